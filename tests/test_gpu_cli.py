"""The CmdStan-grammar executable end to end on the GPU: per-chain Stan CSV
files (file.hpp:189-212 naming), the header/adaptation/draws/footer layout,
init files with the per-chain lookup and base-file fallback
(command_helper.hpp:144-245), and stansummary over the result."""
import json
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import load_stan_csv

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "cmdstan_amd", "bin")
DATA = os.path.join(ROOT, "tests", "golden", "bernoulli.data.json")

pytestmark = pytest.mark.gpu


def _run(args, cwd):
    r = subprocess.run([os.path.join(BIN, "bernoulli")] + args, cwd=cwd, capture_output=True, text=True, timeout=300)
    return r


def test_two_chains_with_base_init_file(tmp_path):
    (tmp_path / "init.json").write_text(json.dumps({"theta": 0.3}))
    r = _run(["sample", "num_chains=2", "data", f"file={DATA}", "init=init.json", "output", "file=out.csv",
              "refresh=500", "random", "seed=12"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Iteration: 2000 / 2000 [100%]  (Sampling)" in r.stdout + r.stderr
    thetas = []
    for k in (1, 2):
        path = tmp_path / f"out_{k}.csv"
        text = path.read_text()
        assert "# Adaptation terminated" in text and "Elapsed Time" in text
        cols, draws, _ = load_stan_csv(str(path))
        assert cols[:8] == ["lp__", "accept_stat__", "stepsize__", "treedepth__", "n_leapfrog__", "divergent__",
                            "energy__", "theta"]
        assert draws.shape == (1000, 8)
        thetas.append(draws[:, 7])
    assert 0.2 < np.mean(thetas) < 0.32
    s = subprocess.run([os.path.join(BIN, "stansummary"), str(tmp_path / "out_1.csv"), str(tmp_path / "out_2.csv")],
                       capture_output=True, text=True, timeout=60)
    assert s.returncode == 0 and "theta" in s.stdout and "2 chains" in s.stdout


def test_per_chain_init_files_and_bad_init(tmp_path):
    for k, v in ((1, 0.2), (2, 0.4)):
        (tmp_path / f"init_{k}.json").write_text(json.dumps({"theta": v}))
    r = _run(["sample", "num_chains=2", "data", f"file={DATA}", "init=init.json", "output", "file=o.csv"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    (tmp_path / "bad.json").write_text(json.dumps({"theta": 1.5}))
    r = _run(["sample", "data", f"file={DATA}", "init=bad.json", "output", "file=b.csv"], tmp_path)
    assert r.returncode != 0 and "theta" in r.stdout + r.stderr
    r = _run(["sample", "num_chains=2", "data", f"file={DATA}", "init=missing.json"], tmp_path)
    assert r.returncode != 0 and "Can't open" in r.stdout + r.stderr


def test_diagnostic_file_per_chain(tmp_path):
    """multi_chain_test.cpp:14-80: id=10, two chains, diagnostic_file=<base>_diag.csv
    gives <base>_diag_10.csv and <base>_diag_11.csv; each row holds the sample
    row's sampler columns, then q, p_<name>, g_<name>."""
    r = _run(["id=10", "sample", "num_warmup=200", "num_samples=50", "num_chains=2", "data", f"file={DATA}",
              "random", "seed=1234", "output", "file=m.csv", "diagnostic_file=m_diag.csv"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    for k in (10, 11):
        cols, draws, _ = load_stan_csv(str(tmp_path / f"m_{k}.csv"))
        dcols, diag, _ = load_stan_csv(str(tmp_path / f"m_diag_{k}.csv"))
        assert dcols == ["lp__", "accept_stat__", "stepsize__", "treedepth__", "n_leapfrog__", "divergent__",
                         "energy__", "theta", "p_theta", "g_theta"]
        assert diag.shape == (50, 10)
        assert np.allclose(diag[:, :7], draws[:, :7], rtol=1e-5)
        # theta = inv_logit(q)
        assert np.allclose(1 / (1 + np.exp(-diag[:, 7])), draws[:, 7], rtol=1e-5)


def test_save_cmdstan_config_json(tmp_path):
    """config_json_test.cpp: save_cmdstan_config=1 writes <output base>_config.json,
    valid JSON holding the argument tree; without it no file is written."""
    r = _run(["sample", "num_samples=20", "num_warmup=20", "data", f"file={DATA}", "output", "file=o.csv",
              "save_cmdstan_config=1"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    cfg = json.loads((tmp_path / "o_config.json").read_text())
    assert cfg["method"]["value"] == "sample"
    s = cfg["method"]["sample"]
    assert s["num_samples"] == 20 and s["num_warmup"] == 20
    assert s["algorithm"]["value"] == "hmc" and s["algorithm"]["hmc"]["engine"]["nuts"]["max_depth"] == 10
    assert cfg["data"]["file"] == DATA and cfg["output"]["file"] == "o.csv"
    assert cfg["model_name"] == "bernoulli_model" and cfg["stan_major_version"] == 2
    r = _run(["sample", "num_samples=20", "num_warmup=20", "data", f"file={DATA}", "output", "file=p.csv"], tmp_path)
    assert r.returncode == 0 and not (tmp_path / "p_config.json").exists()
