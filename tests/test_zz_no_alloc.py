"""nuts_log_prob_grad_noalloc (device buffers, scratch in a static device
array) against nuts_log_prob_grad (host buffers, library allocations) on the
same points: two calls in a row, the second with fewer points."""
import ctypes as C

import numpy as np
import pytest

import cmdstan_amd as ca
from cmdstan_amd._lib import last_error, lib
from tests.model_cases import CASES

pytestmark = pytest.mark.gpu

FAMILIES = ["logistic_regression", "hier_logistic_cp", "eight_schools_cp", "hier_normal_nc", "bernoulli", "gauss"]


class _Hip:
    """Device buffers from the HIP runtime the library itself links (torch
    bundles a separate runtime, whose pointers the library cannot use)."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def put(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, shape):
        out = np.empty(shape, dtype=np.float64)
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0  # device to host
        return out


def _noalloc(model, q, jacobian):
    h = _Hip()
    n, D = q.shape
    dq = h.put(q)
    dlp = h.put(np.full(n, np.nan))
    dg = h.put(np.full((n, D), np.nan))
    rc = lib().nuts_log_prob_grad_noalloc(model.h, 0, dq, n, int(jacobian), dlp, dg, None)
    assert rc == 0, last_error()
    assert h.L.hipDeviceSynchronize() == 0
    lp, g = h.get(dlp, (n,)), h.get(dg, (n, D))
    for p in (dq, dlp, dg):
        h.L.hipFree(p)
    return lp, g


@pytest.mark.parametrize("name", FAMILIES)
@pytest.mark.parametrize("jacobian", [True, False])
def test_noalloc_matches_log_prob_grad(name, jacobian):
    gm = ca.Model(name, CASES[name])
    D = gm.dim
    rng = np.random.default_rng(7)
    for n in (300, 37):  # the second call has fewer points and reuses the same scratch
        q = rng.standard_normal((n, D)) * 0.7
        lp_ref, g_ref = gm.log_prob_grad(q, jacobian=jacobian)
        print(f"noalloc {name} jacobian={jacobian} n={n} D={D}", flush=True)
        lp, g = _noalloc(gm, q, jacobian)
        assert np.all(np.isfinite(lp)) and np.all(np.isfinite(g))
        # same kernels and the same work split: equal up to 1e-12 relative
        assert np.allclose(lp, lp_ref, rtol=1e-12, atol=1e-12), np.max(np.abs(lp - lp_ref))
        scale = np.maximum(1.0, np.abs(g_ref))
        assert np.all(np.abs(g - g_ref) <= 1e-12 * scale), np.max(np.abs(g - g_ref) / scale)


def test_noalloc_rejects_too_many_points():
    gm = ca.Model("bernoulli", CASES["bernoulli"])
    rc = lib().nuts_log_prob_grad_noalloc(gm.h, 0, None, 4097, 1, None, None, None)
    assert rc != 0
    assert "4096" in last_error()
