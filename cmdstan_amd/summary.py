"""stansummary statistics through the library's nuts_summary (C++ host code)."""
import os

import numpy as np

from ._lib import DP, lib

STAT_NAMES = ["Mean", "MCSE", "StdDev", "MAD", "5%", "50%", "95%", "ESS_bulk", "ESS_tail", "R_hat"]


def summarize(x, n_threads=None):
    """x: [params][chains][draws] (or [chains][draws] for one parameter) ->
    array [params][10] in STAT_NAMES order."""
    a = np.ascontiguousarray(x, dtype=np.float64)
    if a.ndim == 2:
        a = a[None]
    P, M, N = a.shape
    out = np.zeros((P, 10))
    n_threads = n_threads or min(16, os.cpu_count() or 1)
    rc = lib().nuts_summary(a.ctypes.data_as(DP), N, M, P, n_threads, out.ctypes.data_as(DP))
    if rc != 0:
        raise ValueError("nuts_summary: bad shape")
    return out
