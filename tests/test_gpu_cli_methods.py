"""CmdStan CLI behaviour beyond plain sampling, on the GPU, against the
reference's interface tests: method=log_prob (log_prob_test.cpp:58-242),
metric files in JSON and R dump (metric_test.cpp:17-92, the reference's own
metric fixtures), save_metric (save_metric_json_test.cpp:37-108), sig_figs
(output_sig_figs_test.cpp:7-89) and rejected initial values
(command_test.cpp:151-250)."""
import json
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import GOLDEN, load_stan_csv

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "cmdstan_amd", "bin")
BERN_DATA = os.path.join(GOLDEN, "bern.data.json")

pytestmark = pytest.mark.gpu


def _run(family, args, cwd):
    print(f"cli {family} {' '.join(args)}", flush=True)
    return subprocess.run([os.path.join(BIN, family)] + args, cwd=cwd, capture_output=True, text=True, timeout=300)


def _bern_lp(theta, jacobian):
    lp = 2 * np.log(theta) + 8 * np.log1p(-theta)   # Beta(1,1) prior, 2 ones in 10
    g = 2 * (1 - theta) - 8 * theta                   # d/du at u = logit(theta)
    if jacobian:                                      # + log theta + log(1 - theta)
        lp = lp + np.log(theta) + np.log1p(-theta)
        g = g + 1 - 2 * theta
    return lp, g


def _read_log_prob(path):
    names, vals, _ = load_stan_csv(str(path))
    return names, np.atleast_2d(vals)


# ---------------------------------------------------------------- log_prob
def test_log_prob_constrained_csv(tmp_path):
    """log_prob_cparams_csv: the 1000 draws of bern_fitted_params.csv give 1000
    rows of lp__, g_theta (the closed form of the bernoulli model)."""
    r = _run("bernoulli", ["method=log_prob", f"constrained_params={GOLDEN}/bern_fitted_params.csv", "data",
                           f"file={BERN_DATA}", "output", "file=out.csv"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    names, v = _read_log_prob(tmp_path / "out.csv")
    assert names == ["lp__", "g_theta"] and v.shape == (1000, 2)
    fn, fd, _ = load_stan_csv(f"{GOLDEN}/bern_fitted_params.csv")
    theta = fd[:, fn.index("theta")]
    lp, g = _bern_lp(theta, True)
    assert np.allclose(v[:, 0], lp, rtol=2e-5) and np.allclose(v[:, 1], g, rtol=2e-5, atol=1e-5)


@pytest.mark.parametrize("fmt", ["json", "R"])
@pytest.mark.parametrize("jacobian", ["1", "0"])
def test_log_prob_unconstrained_json_and_rdump(tmp_path, fmt, jacobian):
    u = np.array([-1.2, 0.3, 2.0])
    if fmt == "json":
        (tmp_path / "u.json").write_text(json.dumps({"params_r": [[x] for x in u]}))
    else:
        (tmp_path / "u.R").write_text("params_r <- \nstructure(c(%s),\n.Dim = c(3, 1))\n" % ", ".join(map(str, u)))
    r = _run("bernoulli", ["method=log_prob", f"unconstrained_params=u.{fmt}", f"jacobian={jacobian}", "data",
                           f"file={BERN_DATA}", "output", "file=lp.csv", "sig_figs=12"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    names, v = _read_log_prob(tmp_path / "lp.csv")
    assert names == ["lp__", "g_theta"] and v.shape == (3, 2)
    lp, g = _bern_lp(1 / (1 + np.exp(-u)), jacobian == "1")
    assert np.allclose(v[:, 0], lp, rtol=1e-10) and np.allclose(v[:, 1], g, rtol=1e-10, atol=1e-11)
    if fmt == "R":
        assert "is being read as an 'RDump' file" in r.stderr


def test_log_prob_constrained_rdump_and_single_vector(tmp_path):
    (tmp_path / "c.R").write_text("theta <- 0.8\n")
    r = _run("bernoulli", ["method=log_prob", "constrained_params=c.R", "data", f"file={BERN_DATA}", "output",
                           "file=a.csv"], tmp_path)
    assert r.returncode == 0, r.stderr
    names, v = _read_log_prob(tmp_path / "a.csv")
    lp, g = _bern_lp(np.array([0.8]), True)
    assert v.shape == (1, 2) and np.allclose(v[0], [lp[0], g[0]], rtol=1e-5)
    (tmp_path / "u1.json").write_text(json.dumps({"params_r": [0.25]}))
    r = _run("bernoulli", ["method=log_prob", "unconstrained_params=u1.json", "data", f"file={BERN_DATA}", "output",
                           "file=b.csv"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert _read_log_prob(tmp_path / "b.csv")[1].shape == (1, 2)


def test_log_prob_errors(tmp_path):
    base = ["data", f"file={BERN_DATA}", "output", "file=o.csv"]
    r = _run("bernoulli", ["method=log_prob"] + base, tmp_path)        # log_prob_no_params
    assert r.returncode != 0 and "No input parameter files provided" in r.stderr
    r = _run("bernoulli", ["method=log_prob", f"unconstrained_params={GOLDEN}/bern_unconstrained_params.json",
                           f"constrained_params={GOLDEN}/bern_constrained_params.json"] + base, tmp_path)
    assert r.returncode != 0 and "Cannot specify both" in r.stderr
    # ragged params_r (bern_unconstrained_params_short.json): unreadable
    r = _run("bernoulli", ["method=log_prob",
                           f"unconstrained_params={GOLDEN}/bern_unconstrained_params_short.json"] + base, tmp_path)
    assert r.returncode != 0
    # 19 unconstrained values for a model with one parameter
    r = _run("bernoulli", ["method=log_prob", f"unconstrained_params={GOLDEN}/bern_unconstrained_params.R"] + base,
             tmp_path)
    assert r.returncode != 0 and "Incorrect number of unconstrained parameters" in r.stderr
    # log_prob_conflict: constrained_params names the output file
    r = _run("bernoulli", ["method=log_prob", "constrained_params=o.csv"] + base, tmp_path)
    assert r.returncode != 0 and "Filename conflict" in r.stderr


# ------------------------------------------------------------ metric files
# metric_test.cpp: the models "proper" (mu ~ normal(0, 1)) and "test_model"
# (mu1 ~ normal(0, 10), mu2 ~ normal(0, 1)) as gauss densities
MODELS = {"proper": dict(D=1, P=[[1.0]]), "test_model": dict(D=2, P=[[0.01, 0.0], [0.0, 1.0]])}
EXPECTED = {("proper", "diag_e"): "# Diagonal elements of inverse mass matrix:\n# 1.5",
            ("proper", "dense_e"): "# Elements of inverse mass matrix:\n# 1.5",
            ("test_model", "diag_e"): "# Diagonal elements of inverse mass matrix:\n# 1.5, 1.8",
            ("test_model", "dense_e"): "# Elements of inverse mass matrix:\n# 1.5, 1\n# 1, 1.5"}


@pytest.mark.parametrize("model", ["proper", "test_model"])
@pytest.mark.parametrize("metric", ["diag_e", "dense_e"])
def test_metric_file_text(tmp_path, model, metric):
    (tmp_path / "data.json").write_text(json.dumps(MODELS[model]))
    mfile = f"{GOLDEN}/{model}.{metric}_metric.json"
    if not os.path.exists(mfile):
        mfile = f"{GOLDEN}/{model}.{metric}_metric.dat.R"
    for k, (adapt, num_warmup) in enumerate([("0", "0"), ("0", "1"), ("0", "100"), ("1", "0"), ("1", "1"),
                                             ("1", "100")]):
        stepsize = ["0.1", "0.5", "1.25"][k % 3]
        r = _run("gauss", ["sample", "num_samples=100", f"num_warmup={num_warmup}", "adapt", f"engaged={adapt}",
                           "algorithm=hmc", "engine=nuts", f"metric={metric}", f"metric_file={mfile}",
                           f"stepsize={stepsize}", "data", "file=data.json", "output", "file=output.csv"], tmp_path)
        if adapt == "1" and num_warmup == "0":
            assert r.returncode != 0
            continue
        assert r.returncode == 0, r.stdout + r.stderr
        text = (tmp_path / "output.csv").read_text()
        assert text.count(EXPECTED[(model, metric)]) == 1, text[-400:]
        if adapt == "0":
            assert text.count("# Step size = " + stepsize) == 1


@pytest.mark.parametrize("metric,brackets", [("diag_e", 1), ("dense_e", 7), ("unit_e", 1)])
def test_save_metric_json(tmp_path, metric, brackets):
    (tmp_path / "data.json").write_text(json.dumps(dict(D=6)))
    r = _run("gauss", ["random", "seed=1234", "method=sample", "adapt", "save_metric=1", "algorithm=hmc",
                       f"metric={metric}", "data", "file=data.json", "output", "file=output.csv"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    text = (tmp_path / "output_metric.json").read_text()
    obj = json.loads(text)
    assert text.count("stepsize") == 1 and text.count("inv_metric") == 1 and text.count("[") == brackets
    assert obj["metric_type"] == metric and obj["stepsize"] > 0
    r = _run("gauss", ["random", "seed=1234", "method=sample", "adapt", "engaged=0", "save_metric=1", "data",
                       "file=data.json", "output", "file=o2.csv"], tmp_path)
    assert r.returncode != 0


# CmdStan's grammar: hmc's arguments follow algorithm=hmc ("stepsize is either
# mistyped or misplaced" otherwise, as argument_parser reports)
@pytest.mark.parametrize("sig_figs,expect", [("1", 0.1), ("2", 0.12), ("9", 0.123456789)])
def test_sig_figs(tmp_path, sig_figs, expect):
    """The CSV digits follow sig_figs: a step size of 0.123456789123456789
    without adaptation is printed in the stepsize__ column."""
    (tmp_path / "data.json").write_text(json.dumps(dict(D=1)))
    r = _run("gauss", ["sample", "num_warmup=0", "num_samples=1", "adapt", "engaged=0", "algorithm=hmc",
                       "stepsize=0.123456789123456789", "data", "file=data.json", "output", "file=o.csv",
                       f"sig_figs={sig_figs}"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    names, d, _ = load_stan_csv(str(tmp_path / "o.csv"))
    assert abs(np.atleast_2d(d)[0, names.index("stepsize__")] - expect) < 1e-16


# ---------------------------------------------------------- initialisation
@pytest.mark.parametrize("init", ["json", "R"])
def test_user_init_value_fail(tmp_path, init):
    """command_test.cpp user_init_value_fail: one given initial value whose log
    density is not finite is rejected once, and the run fails."""
    (tmp_path / "data.json").write_text(json.dumps(dict(D=1)))
    if init == "json":
        (tmp_path / "init.json").write_text(json.dumps({"x": [1e200]}))
    else:
        (tmp_path / "init.R").write_text("x <- c(1e200)\n")
    r = _run("gauss", ["sample", f"init=init.{init}", "data", "file=data.json", "output", "file=o.csv"], tmp_path)
    assert r.returncode != 0
    assert (r.stdout + r.stderr).count("Rejecting initial value") == 1
    assert "Log probability evaluates to log(0)" in r.stdout


def test_rdump_data_and_per_chain_rdump_inits(tmp_path):
    (tmp_path / "bern.data.R").write_text("N <- 10L\ny <- c(0L, 1L, 0L, 0L, 0L, 0L, 0L, 0L, 0L, 1L)\n")
    for k, v in ((1, 0.1), (2, 0.2)):
        (tmp_path / f"init_{k}.R").write_text(f"theta <- {v}\n")
    r = _run("bernoulli", ["sample", "num_chains=2", "num_warmup=0", "num_samples=1", "adapt", "engaged=0",
                           "algorithm=hmc", "stepsize=1e-9", "init=init.R", "data", "file=bern.data.R", "output",
                           "file=o.csv"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    for k, v in ((1, 0.1), (2, 0.2)):
        names, d, _ = load_stan_csv(str(tmp_path / f"o_{k}.csv"))
        assert abs(np.atleast_2d(d)[0, names.index("theta")] - v) < 1e-6
    r = _run("bernoulli", ["sample", "num_chains=2", "init=init.txt", "data", "file=bern.data.R"], tmp_path)
    assert r.returncode != 0 and "file ending of .txt is not supported by cmdstan" in r.stderr
