"""A model outside the registered families through the plugin ABI
(nuts_model_plugin): the example plugin (Neal's funnel, its own gfx950
gradient kernel in cmdstan_amd/plugins/funnel_plugin.hip) against the oracle's
restatement of the same density: log_prob_grad at 1e-12, single-transition
replay on the lockstep drivers, names, write_array, transform_inits and a
posterior check."""
import ctypes as C
import os

import numpy as np
import pytest

import cmdstan_amd as ca
from cmdstan_amd._lib import nuts_model_plugin
from oracle import ess as E
from oracle import oracle as O

pytestmark = pytest.mark.gpu

PLUGIN = os.path.join(os.path.dirname(ca.__file__), "plugins", "libfunnel_plugin.so")


def _funnel(K):
    L = C.CDLL(PLUGIN)
    L.funnel_plugin_create.restype = C.c_void_p
    L.funnel_plugin_create.argtypes = [C.c_int, C.POINTER(nuts_model_plugin)]
    L.funnel_plugin_destroy.argtypes = [C.c_void_p]
    p = nuts_model_plugin()
    h = L.funnel_plugin_create(K, C.byref(p))
    assert h
    return ca.Model.from_plugin(p, keep=(L, h, p)), O.Model("funnel", dict(K=K))


def test_plugin_names_and_log_prob_grad():
    gm, om = _funnel(5)
    assert gm.dims == [6] and gm.num_outputs == [7]
    assert gm.param_names[0] == ["v", "sigma", "x.1", "x.2", "x.3", "x.4", "x.5"]
    assert gm.unconstrained_names[0][:2] == ["v", "x.1"]
    rng = np.random.default_rng(2)
    q = rng.standard_normal((40, 6))
    for jac in (True, False):
        lp, g = gm.log_prob_grad(q, jacobian=jac)
        for i in range(len(q)):
            lp_ref, g_ref = om.log_prob_grad(q[i], jac)
            assert abs(lp[i] - lp_ref) <= 1e-12 * max(1.0, abs(lp_ref)) * 4
            assert np.all(np.abs(g[i] - g_ref) <= 1e-12 * np.maximum(1.0, np.abs(g_ref)) * 4)
    u = gm.transform_inits({"v": 0.5, "x": [1, 2, 3, 4, 5]})
    assert np.allclose(u, [0.5, 1, 2, 3, 4, 5])
    assert np.isnan(gm.transform_inits({"v": 0.1})[1:]).all()


@pytest.mark.parametrize("driver", ["lockstep", "lockstep_block"])
def test_plugin_transition_replay(driver):
    gm, om = _funnel(7)
    rng = np.random.default_rng(4)
    nch, D = 16, 8
    q0 = rng.standard_normal((nch, D)) * 0.5
    u = rng.random((nch, 2048))
    z = rng.standard_normal((nch, D))
    minv = rng.uniform(0.5, 1.5, D)
    for eps in (0.05, 0.2, 0.6):
        qg, infos = gm.transition_replay(q0, u, z, eps, metric="diag_e", inv_metric=minv, max_depth=8, driver=driver)
        for c in range(nch):
            qo, io = om.transition_replay(q0[c], u[c], z[c], eps, metric=1, minv=minv, max_depth=8)
            for k in ("depth", "n_leapfrog", "divergent", "uniforms_used", "normals_used"):
                assert infos[c][k] == io[k], (eps, c, k)
            assert np.allclose(qg[c], qo, rtol=1e-9, atol=1e-11)


def test_plugin_fused_driver_is_refused():
    gm, _ = _funnel(2)
    with pytest.raises(ca.SamplerError, match="lockstep"):
        gm.sample(num_chains=2, fused=1, num_warmup=10, num_samples=10)


def test_plugin_sampling_and_write_array():
    """Draws carry sigma = exp(v / 2) from the plugin's write_array; x is
    symmetric about 0 whatever v does; v follows the oracle's NUTS on the same
    density and settings (NUTS visits a funnel's neck rarely, so both share
    that bias; the comparison is sampler against sampler)."""
    gm, om = _funnel(2)
    kw = dict(num_warmup=600, num_samples=600, delta=0.95)
    fit = gm.sample(num_chains=64, seed=5, **kw)
    assert np.all(fit.status == 0)
    cols = fit.column_names[0]
    arr = np.stack(fit.draws)
    v, sigma = arr[:, :, cols.index("v")], arr[:, :, cols.index("sigma")]
    assert np.allclose(sigma, np.exp(v / 2), rtol=1e-12)
    for k in (1, 2):
        sx = E.summary(arr[:, :, cols.index(f"x.{k}")].T)
        assert abs(sx["mean"]) < 4 * sx["mcse"] + 0.02, sx
    r = om.run(seed=5, num_chains=32, n_threads=8, **kw)
    sg, so = E.summary(v.T), E.summary(r["draws"][:, :, 7].T)
    print(f"plugin funnel v: GPU mean {sg['mean']:.3f} (mcse {sg['mcse']:.3f}), oracle {so['mean']:.3f} "
          f"(mcse {so['mcse']:.3f})")
    assert abs(sg["mean"] - so["mean"]) < 4 * np.hypot(sg["mcse"], so["mcse"])
    assert abs(sg["sd"] - so["sd"]) < 0.15 * so["sd"]
