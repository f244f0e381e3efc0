"""The stansummary CLI on the reference's example output: its --csv_filename
rows equal the golden strings of stansummary_test.cpp:481-550 (CPU only)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "cmdstan_amd", "bin", "stansummary")
CSV = os.path.join(ROOT, "tests", "golden", "bernoulli_chain_1.csv")

pytestmark = pytest.mark.skipif(not os.path.exists(BIN), reason="CLI not built")


def _run(args, tmp_path):
    out = tmp_path / "summary.csv"
    r = subprocess.run([BIN, f"--csv_filename={out}"] + args + [CSV], capture_output=True, text=True, timeout=60)
    return r, (out.read_text().splitlines() if out.exists() else [])


def test_csv_output_matches_golden_rows(tmp_path):
    r, lines = _run([], tmp_path)
    assert r.returncode == 0, r.stdout
    assert lines[0] == "name,Mean,MCSE,StdDev,MAD,5%,50%,95%,ESS_bulk,ESS_tail,R_hat"
    assert lines[1] == '"lp__",-7.2719,0.0365168,0.768874,0.303688,-8.98426,-6.97009,-6.75007,519.29,503.309,1.00141'
    assert lines[7] == ('"energy__",7.78428,0.0508815,1.0314,0.745859,6.80565,7.46758,9.8864,489.874,486.438,'
                        '1.00495')
    assert lines[8] == ('"theta",0.256552,0.00610844,0.119654,0.120965,0.0802982,0.24996,0.47034,361.506,395.736,'
                        '1.00186')
    assert "Inference for Stan model: bernoulli_model" in r.stdout


def test_csv_output_without_percentiles(tmp_path):
    r, lines = _run(["-p", ""], tmp_path)
    assert r.returncode == 0
    assert lines[0] == "name,Mean,MCSE,StdDev,MAD,ESS_bulk,ESS_tail,R_hat"
    assert lines[1] == '"lp__",-7.2719,0.0365168,0.768874,0.303688,519.29,503.309,1.00141'


def test_bad_sig_figs_is_an_error(tmp_path):
    r = subprocess.run([BIN, "--sig_figs", "-1", CSV], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "--sig_figs: Value -1 not in range" in r.stdout


def test_autocorr_option(tmp_path):
    """--autocorr n (stansummary_test.cpp:345-375 error cases; report layout of
    stansummary_helper.hpp:700-745): lag 0 is 1, the lags follow numpy's
    correlation of the centred draws, constant columns print nan."""
    import numpy as np
    from tests.conftest import load_stan_csv
    for arg, msg in (("-1", "--autocorr: Number less or equal to 0"), ("0", "--autocorr: Number less or equal to 0"),
                     ("2", "not a valid chain id")):
        r = subprocess.run([BIN, "--autocorr", arg, CSV], capture_output=True, text=True, timeout=60)
        assert r.returncode != 0 and msg in r.stdout
    r = subprocess.run([BIN, "--autocorr", "1", CSV], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    lines = r.stdout.split("Displaying the autocorrelations for chain 1:")[1].strip().splitlines()
    head = lines[0].split()
    assert head[0] == "Lag" and head[-1] == "theta"
    rows = [l.split() for l in lines[1:]]
    names, d, _ = load_stan_csv(CSV)
    x = d[:, names.index("theta")] - d[:, names.index("theta")].mean()
    ref = np.correlate(x, x, mode="full")[len(x) - 1:] / np.dot(x, x)
    got = np.array([float(rw[-1]) for rw in rows])
    assert len(got) == len(x) and got[0] == 1.0
    assert np.allclose(got, ref, rtol=1e-5, atol=1e-6)
    assert rows[1][head.index("stepsize__")] == "nan"


@pytest.mark.parametrize("args,msg", [
    (["--percentiles", "-1"], "Option --percentiles"),           # stansummary_test.cpp:375-406
    (["--percentiles", "101"], "Option --percentiles"),
    (["--percentiles", "2,30,5"], "Option --percentiles 2,30,5: values must be in range (0, 100), inclusive, "
                                  "and strictly increasing."),
    (["--percentiles", "5,abc"], "Option --percentiles"),
    (["-i", "psi"], "--include_param: Unrecognized parameter(s): 'psi' "),  # :408-424
    # lp__ and the sampler columns are not model parameters (stansummary.cpp:158-190)
    (["-i", "psi", "-i", "lp__", "-i", "accept_stat__"],
     "--include_param: Unrecognized parameter(s): 'accept_stat__' 'lp__' 'psi' "),
])
def test_bad_arguments_are_errors(args, msg):
    r = subprocess.run([BIN] + args + [CSV], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and msg in r.stdout


def test_missing_file_and_no_arguments():
    r = subprocess.run([BIN, "no_such.csv"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "File does not exist" in r.stdout   # :290-301
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=60)
    assert "Usage: stansummary" in r.stdout                            # :279-288
    r = subprocess.run([BIN, "--percentiles", "2,50,95", CSV], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "2%" in r.stdout and "95%" in r.stdout


def test_console_rows_match_golden():
    """The console table of stansummary_test.cpp:426-480 (check_console_output):
    column formats chosen separately for the sampler rows and the model rows,
    lp__ printed with the model rows' formats, widths from the wider block."""
    r = subprocess.run([BIN, CSV], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    lines = r.stdout.split("\n")
    for gold in ("lp__            -7.3  3.7e-02     0.77   0.30   -9.0  -7.0  -6.8       519       503    1.0",
                 "accept_stat__   0.90  4.6e-03  1.5e-01  0.064   0.57  0.96   1.0      1284       941   1.00",
                 "energy__         7.8  5.1e-02  1.0e+00   0.75    6.8   7.5   9.9       490       486    1.0",
                 "theta           0.26  6.1e-03     0.12   0.12  0.080  0.25  0.47       362       396    1.0"):
        assert gold in lines, gold


def test_csv_output_sig_figs_golden(tmp_path):
    """stansummary_test.cpp:572-620: with --sig_figs 2 the csv rows use 2
    significant digits."""
    out = tmp_path / "s.csv"
    r = subprocess.run([BIN, f"--csv_filename={out}", "--sig_figs", "2", CSV], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0
    lines = out.read_text().splitlines()
    assert lines[1] == '"lp__",-7.3,0.037,0.77,0.3,-9,-7,-6.8,5.2e+02,5e+02,1'
    assert lines[7] == '"energy__",7.8,0.051,1,0.75,6.8,7.5,9.9,4.9e+02,4.9e+02,1'
    assert lines[8] == '"theta",0.26,0.0061,0.12,0.12,0.08,0.25,0.47,3.6e+02,4e+02,1'


def test_percentile_separators_and_equal_values():
    """The spec splits on commas and spaces alike, runs of them count once, and
    a value may equal the integer part of the one before
    (stansummary.cpp:98-110, stansummary_helper.hpp:300-320)."""
    for spec in ("5 50 95", " 5, 50,,95 ", "5,5,95", "5.5,5.2"):
        r = subprocess.run([BIN, "--percentiles", spec, CSV], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, (spec, r.stdout)
        head = [l for l in r.stdout.splitlines() if "Mean" in l][0].split()
        want = [t for t in spec.replace(",", " ").split()]
        assert head[4:4 + len(want)] == [t + "%" for t in want], (spec, head)


def test_existing_csv_path_is_refused(tmp_path):
    """--csv_filename must name a path that does not exist yet (the CLI11
    NonexistentPath check of stansummary.cpp:62-64)."""
    out = tmp_path / "there.csv"
    out.write_text("keep me\n")
    r = subprocess.run([BIN, f"--csv_filename={out}", CSV], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "Path already exists" in r.stdout
    assert out.read_text() == "keep me\n"


def test_include_param_is_exact_and_autocorr_shows_every_column():
    """A vector's bare name matches none of its elements; the console rows use
    display names; --autocorr prints every column (lp__, sampler, all model
    parameters) under display names whatever --include_param selects
    (stansummary_helper.hpp:700-745)."""
    es = os.path.join(ROOT, "tests", "golden", "eight_schools_output.csv")
    r = subprocess.run([BIN, "-i", "theta", es], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "Unrecognized parameter(s): 'theta' " in r.stdout
    r = subprocess.run([BIN, "-a", "1", "-i", "theta.2", "-i", "mu", es], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    body, ac = r.stdout.split("Displaying the autocorrelations for chain 1:")
    rows = [l.split()[0] for l in body.splitlines() if l and not l.startswith(" ")]
    assert "theta[2]" in rows and "mu" in rows and "theta[1]" not in rows and "tau" not in rows
    head = ac.strip().splitlines()[0].split()
    assert head == ["Lag", "lp__", "accept_stat__", "stepsize__", "treedepth__", "n_leapfrog__", "divergent__",
                    "energy__", "mu", "tau"] + [f"theta[{k}]" for k in range(1, 9)]
    # every name column is as wide as the longest display name plus one
    lines = [l for l in ac.splitlines() if l.strip()]
    lag_w = max(len(str(len(lines) - 1)) + 1, 4)  # lags 0 .. n-1; the Lag column is one wider than n's digits
    assert lines[0] == "Lag".rjust(lag_w) + "".join(h.rjust(len("accept_stat__") + 1) for h in head[1:])


def test_csv_output_include_param_golden(tmp_path):
    """stansummary_test.cpp:622-700 on the eight-schools fixture: selected
    parameters by either naming, then the "# " summary comments."""
    out = tmp_path / "e.csv"
    es = os.path.join(ROOT, "tests", "golden", "eight_schools_output.csv")
    r = subprocess.run([BIN, f"--csv_filename={out}", "--include_param", "theta.6", "-i", "theta[7]", es],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    lines = out.read_text().splitlines()
    assert lines[1] == ('"lp__",-15.5617,0.97319,6.05585,6.3817,-25.3182,-15.7598,-5.47732,41.1897,113.537,'
                        '1.00153')
    assert lines[7] == ('"energy__",20.5888,1.01449,6.43127,6.6161,10.2809,20.8278,30.9921,42.5605,140.171,'
                        '1.00069')
    assert lines[8] == ('"theta[6]",5.001,0.365016,5.76072,5.37947,-4.95375,5.22746,14.1688,230.645,464.978,'
                        '1.00054')
    assert lines[9] == ('"theta[7]",8.54125,0.650098,6.22195,5.35785,-0.814388,8.09342,19.2622,92.3075,241.177,'
                        '1.00244')
    assert lines[10] == "# Inference for Stan model: eight_schools_cp_model"
