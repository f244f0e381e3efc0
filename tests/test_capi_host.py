"""Host-side checks of the C ABI (no GPU needed): the library loads, exports
every symbol the header declares, parses CmdStan JSON data, validates the
argument tree like CmdStan's command layer and reproduces stansummary's golden
statistics."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import cmdstan_amd as ca
from cmdstan_amd._lib import SIGNATURES, nuts_config
from cmdstan_amd.summary import summarize
from tests.conftest import ROOT, load_stan_csv


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "cmdstan_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nuts_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = ca.lib()
    names = _declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
        assert n in SIGNATURES, n


def test_json_data(golden):
    d = ca.Data(json_text=open(f"{golden}/bernoulli.data.json").read())
    L = ca.lib()
    nd = C.c_int()
    dims = (C.c_int64 * 8)()
    vals = C.POINTER(C.c_double)()
    assert L.nuts_data_get(d.h, b"y", C.byref(nd), dims, C.byref(vals)) == 0
    assert nd.value == 1 and dims[0] == 10
    assert [vals[i] for i in range(10)] == [0, 1, 0, 0, 0, 0, 0, 0, 0, 1]
    assert L.nuts_data_get(d.h, b"N", C.byref(nd), dims, C.byref(vals)) == 0
    assert nd.value == 0 and vals[0] == 10
    d2 = ca.Data(json_text='{"a": [[1.5, 2], [3, "inf"]], "b": []}')
    assert L.nuts_data_get(d2.h, b"a", C.byref(nd), dims, C.byref(vals)) == 0
    assert nd.value == 2 and dims[0] == 2 and dims[1] == 2 and vals[3] == np.inf
    with pytest.raises(ca.SamplerError, match="JSON"):
        ca.Data(json_text='{"a": [1, 2')
    with pytest.raises(ca.SamplerError, match="ragged"):
        ca.Data(json_text='{"a": [[1, 2], [3]]}')


def _cfg(**kw):
    c = nuts_config()
    ca.lib().nuts_config_default(C.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def test_config_defaults_match_cmdstan():
    c = _cfg()
    assert (c.num_samples, c.num_warmup, c.save_warmup, c.thin) == (1000, 1000, 0, 1)
    assert (c.adapt_engaged, c.gamma, c.delta, c.kappa, c.t0) == (1, 0.05, 0.8, 0.75, 10)
    assert (c.init_buffer, c.term_buffer, c.window, c.max_depth, c.metric) == (75, 50, 25, 10, 1)
    assert (c.stepsize, c.stepsize_jitter, c.init_radius) == (1.0, 0.0, 2.0)
    assert ca.lib().nuts_output_rows(C.byref(c)) == 1000
    assert ca.lib().nuts_output_rows(C.byref(_cfg(save_warmup=1, thin=3))) == 334 + 334


@pytest.mark.parametrize("kw,msg", [
    (dict(num_warmup=0), "num_warmup"),           # command.hpp:567-574
    (dict(delta=1.0), "delta"),                   # arg_adapt.hpp:23-25
    (dict(stepsize_jitter=1.5), "stepsize_jitter"),
    (dict(max_depth=0), "max_depth"),
    (dict(thin=0), "thin"),
    (dict(stepsize=-1.0), "stepsize"),
])
def test_config_validation(kw, msg):
    c = _cfg(**kw)
    assert ca.lib().nuts_config_validate(C.byref(c)) == 1
    assert msg in ca.last_error()
    assert ca.lib().nuts_config_validate(C.byref(_cfg(num_warmup=0, adapt_engaged=0))) == 0


def test_summary_golden(golden):
    names, d, _ = load_stan_csv(f"{golden}/bernoulli_chain_1.csv")
    s = summarize(d.T[:, None, :])
    gold = {"lp__": [-7.2719, 0.0365168, 0.768874, 0.303688, -8.98426, -6.97009, -6.75007, 519.29, 503.309, 1.00141],
            "energy__": [7.78428, 0.0508815, 1.0314, 0.745859, 6.80565, 7.46758, 9.8864, 489.874, 486.438, 1.00495],
            "theta": [0.256552, 0.00610844, 0.119654, 0.120965, 0.0802982, 0.24996, 0.47034, 361.506, 395.736,
                      1.00186]}
    for k, g in gold.items():
        assert [float(f"{v:.6g}") for v in s[names.index(k)]] == g, k


def test_summary_matches_oracle_multichain():
    from oracle import ess as E
    rng = np.random.default_rng(4)
    x = np.cumsum(rng.standard_normal((3, 4, 400)), axis=2) * 0.1 + rng.standard_normal((3, 4, 400))
    s = summarize(x)
    for p in range(3):
        o = E.summary(x[p].T)
        ref = [o["mean"], o["mcse"], o["sd"], o["mad"], *o["quantiles"], o["ess_bulk"], o["ess_tail"], o["rhat"]]
        assert np.allclose(s[p], ref, rtol=1e-9)


@pytest.mark.gpu
def test_transform_inits_maps_constrained_values():
    """transform_inits: logit / log / identity per parameter, NaN for the
    parameters an init file leaves out, errors for values outside the support."""
    import math
    bern = ca.Model("bernoulli", {"N": 3, "y": [0, 1, 0]})
    assert abs(bern.transform_inits({"theta": 0.25})[0] - math.log(0.25 / 0.75)) < 1e-15
    es = ca.Model("eight_schools_cp", {"J": 3, "y": [1.0, 2.0, 3.0], "sigma": [1.0, 1.0, 2.0]})
    u = es.transform_inits('{"tau": 2.5, "theta": [0.1, 0.2, 0.3]}')
    assert np.isnan(u[0]) and abs(u[1] - math.log(2.5)) < 1e-15 and np.allclose(u[2:], [0.1, 0.2, 0.3])
    with pytest.raises(ca.SamplerError, match="tau"):
        es.transform_inits({"tau": -1.0})
    with pytest.raises(ca.SamplerError, match="theta"):
        bern.transform_inits({"theta": 1.5})


def test_bench_refuses_more_than_one_gpu():
    """bench.py --gpus N > 1 exits with a message: the build is one-GPU (the
    chain-sharded design is documented, not built)."""
    import json
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["n_gpus"] == 2 and "one GPU" in out["error"]


def test_bench_defaults_finish_in_minutes():
    """Without --steps / --warmup the bench measures the documented sizes: C2
    1000 warmup + 1000 draws (20 steps of 50 iterations), the other configs
    their quoted iteration counts."""
    import importlib.util
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("bench_defaults", os.path.join(root, "bench.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    steps, warm = b.default_sizes("c2")
    S = b.step_iterations(steps)
    assert steps * S == 1000 and max(1000, warm * S) == 1000
    assert b.default_sizes("c5") == (100, 150) and b.default_sizes("c3") == (200, 300)
    assert b.default_sizes("c1") == (1000, 1000)


def test_precision_fp32_config_validates():
    """precision 1 is accepted by the config check (FP32 contraction for the
    logistic regression, FP32 leapfrog on the per-chain kernel); 3 is not."""
    assert ca.lib().nuts_config_validate(C.byref(_cfg(precision=1))) == 0
    assert ca.lib().nuts_config_validate(C.byref(_cfg(precision=3))) == 1
    assert "precision" in ca.last_error()


def test_example_plugin_fills_the_plugin_table():
    """The example plugin library loads without a GPU and fills a
    nuts_model_plugin table (dimension, names, callbacks)."""
    from cmdstan_amd._lib import nuts_model_plugin
    path = os.path.join(os.path.dirname(ca.__file__), "plugins", "libfunnel_plugin.so")
    L = C.CDLL(path)
    L.funnel_plugin_create.restype = C.c_void_p
    L.funnel_plugin_create.argtypes = [C.c_int, C.POINTER(nuts_model_plugin)]
    L.funnel_plugin_destroy.argtypes = [C.c_void_p]
    p = nuts_model_plugin()
    h = L.funnel_plugin_create(3, C.byref(p))
    assert h and p.dim == 4 and p.num_outputs == 5
    assert [p.param_names[i].decode() for i in range(5)] == ["v", "sigma", "x.1", "x.2", "x.3"]
    assert [p.unconstrained_names[i].decode() for i in range(4)] == ["v", "x.1", "x.2", "x.3"]
    assert bool(p.log_prob_grad) and bool(p.write_array) and bool(p.transform_inits)
    q = np.array([[0.4, 1.0, -2.0, 0.5]])
    out = np.zeros((1, 5))
    assert p.write_array(p.user, q.ctypes.data_as(C.POINTER(C.c_double)), 1,
                         out.ctypes.data_as(C.POINTER(C.c_double))) == 0
    assert np.allclose(out, [[0.4, np.exp(0.2), 1.0, -2.0, 0.5]])
    L.funnel_plugin_destroy(h)


def test_rdump_reader_matches_json_twins(golden):
    """The reference's R dump fixtures read to the same arrays as their JSON
    twins (column-major R arrays become row-major)."""
    from cmdstan_amd.sampler import data_from_file
    a = data_from_file(f"{golden}/bern_constrained_params.R")
    b = data_from_file(f"{golden}/bern_constrained_params.json")
    for k in ["theta", "mu_vector", "mu_row_vector", "mu_array_1d", "mu_matrix", "mu_matrix_array"]:
        assert np.array_equal(a.get(k), b.get(k)), k
    a = data_from_file(f"{golden}/bern_unconstrained_params.R")
    b = data_from_file(f"{golden}/bern_unconstrained_params.json")
    assert np.array_equal(a.get("params_r"), b.get("params_r"))
    for m in ("proper", "test_model"):
        a = data_from_file(f"{golden}/{m}.dense_e_metric.dat.R")
        b = data_from_file(f"{golden}/{m}.dense_e_metric.json")
        assert np.array_equal(a.get("inv_metric"), b.get("inv_metric"))
    assert data_from_file(f"{golden}/test_model.diag_e_metric.dat.R").get("inv_metric").tolist() == [1.5, 1.8]
    s = ca.Data(rdump_text="n <- 3L\nseq <- 2:5\nz <- integer(0)\nm <- structure(1:6, .Dim = c(2L, 3L))\n")
    assert s.get("seq").tolist() == [2, 3, 4, 5] and s.get("z").shape == (0,)
    assert s.get("m").tolist() == [[1, 3, 5], [2, 4, 6]] and s.get("n") == 3
    with pytest.raises(ca.SamplerError):
        ca.Data(rdump_text="m <- structure(c(1, 2, 3), .Dim = c(2, 2))")
