"""The diagnose CLI on the reference's example output: its report equals the
golden text of diagnose_test.cpp (tree depth saturation, divergences, E-BFMI,
ESS per draw and R-hat checks; CPU only).  Also: stansummary and diagnose read
only the sampling draws of a file with saved warmup, and the CLI writes the
save_cmdstan_config JSON."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "cmdstan_amd", "bin")
GOLD = os.path.join(ROOT, "tests", "golden")

pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(BIN, "diagnose")), reason="CLI not built")


def _diagnose(*files):
    r = subprocess.run([os.path.join(BIN, "diagnose")] + [os.path.join(GOLD, f) for f in files],
                       capture_output=True, text=True, timeout=120)
    return r


@pytest.mark.parametrize("files,nom", [(["eight_schools_output.csv"], "eight_schools.nom"),
                                       (["corr_gauss_output_depth8.csv"], "corr_gauss_depth8.nom"),
                                       (["mix_output.1.csv", "mix_output.2.csv"], "mix.nom")])
def test_report_matches_golden(files, nom):
    r = _diagnose(*files)
    assert r.returncode == 0, r.stdout + r.stderr
    with open(os.path.join(GOLD, nom)) as f:
        assert r.stdout == f.read()


def test_no_problems_and_missing_file():
    r = _diagnose("bernoulli_chain_1.csv")
    assert r.returncode == 0 and r.stdout.rstrip().endswith("Processing complete, no problems detected.")
    r = _diagnose("no_such_file.csv")
    assert r.returncode != 0 and "File" in r.stdout and "No valid input files, exiting." in r.stdout


def test_saved_warmup_rows_are_skipped(tmp_path):
    """A file whose header says save_warmup = 1 carries num_warmup / thin warmup
    rows before the draws: both tools summarise the draws only."""
    src = open(os.path.join(GOLD, "bernoulli_chain_1.csv")).read().splitlines()
    hdr = next(i for i, l in enumerate(src) if l.startswith("lp__"))
    warm = ["-99,0.5,9,9,1023,1,99,0.999"] * 1000
    text = [l.replace("save_warmup = 0 (Default)", "save_warmup = 1") for l in src[:hdr + 1]] + warm + src[hdr + 1:]
    p = tmp_path / "with_warmup.csv"
    p.write_text("\n".join(text) + "\n")
    a = subprocess.run([os.path.join(BIN, "diagnose"), str(p)], capture_output=True, text=True, timeout=60)
    b = _diagnose("bernoulli_chain_1.csv")
    assert a.stdout == b.stdout
    sa = subprocess.run([os.path.join(BIN, "stansummary"), f"--csv_filename={tmp_path / 'a.csv'}", str(p)],
                        capture_output=True, text=True, timeout=60)
    sb = subprocess.run([os.path.join(BIN, "stansummary"), f"--csv_filename={tmp_path / 'b.csv'}",
                         os.path.join(GOLD, "bernoulli_chain_1.csv")], capture_output=True, text=True, timeout=60)
    assert sa.returncode == 0 and sb.returncode == 0
    assert (tmp_path / "a.csv").read_text() == (tmp_path / "b.csv").read_text()


def test_config_json_is_written_before_sampling(tmp_path):
    """save_cmdstan_config=1 writes <output base>_config.json from the argument
    tree before the model is built (so also where no GPU is present): valid
    JSON with typed values, lists as {"value": choice, choice: {...}}."""
    import json
    exe = os.path.join(BIN, "bernoulli")
    subprocess.run([exe, "sample", "num_samples=33", "adapt", "delta=0.9", "algorithm=hmc", "engine=nuts",
                    "max_depth=7", "output", "file=run.csv", "save_cmdstan_config=1", "random", "seed=5"],
                   cwd=tmp_path, capture_output=True, text=True, timeout=120)
    cfg = json.loads((tmp_path / "run_config.json").read_text())
    s = cfg["method"]["sample"]
    assert cfg["method"]["value"] == "sample" and s["num_samples"] == 33 and s["num_warmup"] == 1000
    assert s["adapt"]["delta"] == 0.9 and s["adapt"]["engaged"] is True
    assert s["algorithm"]["hmc"]["engine"] == {"value": "nuts", "nuts": {"max_depth": 7}}
    assert cfg["random"]["seed"] == 5 and cfg["output"]["file"] == "run.csv"
    assert cfg["data"]["file"] == "" and cfg["mpi_enabled"] is False
