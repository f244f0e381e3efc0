"""ORACLE — test infrastructure only.

ctypes front end of ``liboracle.so`` (the FP64 CPU restatement of Stan's NUTS
path in ``nuts.hpp``/``models.hpp``).  Only ``tests/``, the smoke check and the
bench's ``cpu_baseline`` leg may import this module; the product never does.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

CFG_KEYS = ["num_warmup", "num_samples", "thin", "save_warmup", "adapt", "delta", "gamma", "kappa", "t0",
            "init_buffer", "term_buffer", "window", "stepsize", "jitter", "max_depth", "metric", "init_radius"]
CFG_DEFAULTS = dict(num_warmup=1000, num_samples=1000, thin=1, save_warmup=0, adapt=1, delta=0.8, gamma=0.05,
                    kappa=0.75, t0=10, init_buffer=75, term_buffer=50, window=25, stepsize=1.0, jitter=0.0,
                    max_depth=10, metric=1, init_radius=2.0)
METRICS = {"unit_e": 0, "diag_e": 1, "dense_e": 2}
DP = C.POINTER(C.c_double)


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load(path):
    """A ctypes handle on an oracle library at `path` with every signature set
    (the tests also load a variant built with a planted mistake)."""
    L = C.CDLL(path)
    L.oracle_last_error.restype = C.c_char_p
    L.oracle_model_create.restype = C.c_void_p
    L.oracle_model_create.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int64), C.POINTER(DP)]
    L.oracle_model_destroy.argtypes = [C.c_void_p]
    L.oracle_model_dim.argtypes = [C.c_void_p]
    L.oracle_model_num_out.argtypes = [C.c_void_p]
    L.oracle_log_prob_grad.restype = C.c_double
    L.oracle_log_prob_grad.argtypes = [C.c_void_p, DP, C.c_int, DP]
    L.oracle_log_prob.restype = C.c_double
    L.oracle_log_prob.argtypes = [C.c_void_p, DP, C.c_int]
    L.oracle_log_prob_cs.restype = C.c_double
    L.oracle_log_prob_cs.argtypes = [C.c_void_p, DP, C.c_int, C.c_double, C.c_int]
    L.oracle_write_array.argtypes = [C.c_void_p, DP, DP]
    L.oracle_philox.argtypes = [C.POINTER(C.c_uint32)] * 3
    L.oracle_stream_uniform.restype = C.c_double
    L.oracle_stream_uniform.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64]
    L.oracle_stream_normal_pair.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, DP]
    L.oracle_leapfrog.argtypes = [C.c_void_p, C.c_int, DP, C.c_double, DP, DP, DP, DP]
    L.oracle_transition_replay.argtypes = [C.c_void_p, C.c_int, DP, C.c_double, C.c_double, C.c_int, DP, DP,
                                           C.c_int64, DP, C.c_int64, DP, DP]
    L.oracle_dual_averaging.argtypes = [C.c_double] * 5 + [DP, C.c_int, DP, DP]
    L.oracle_window_schedule.argtypes = [C.c_uint] * 4 + [C.POINTER(C.c_int)] * 2
    L.oracle_run.argtypes = [C.c_void_p, DP, C.c_uint64, C.c_int, C.c_int, C.c_int, DP, DP, DP, DP, DP, DP]
    L.oracle_init_stepsize.argtypes = [C.c_void_p, C.c_int, DP, C.c_double, DP, DP, C.c_int64, DP]
    return L


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = load(path)
    return _LIB


def _dp(a):
    return a.ctypes.data_as(DP)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Model:
    """A model of the oracle's registry bound to a data dict (name -> array)."""

    def __init__(self, name, data, library=None):
        L = self.L = library or lib()
        names, ndims, dims, vals = [], [], [], []
        self._keep = []
        for k, v in data.items():
            a = _f64(np.asarray(v))
            self._keep.append(a)
            names.append(k.encode())
            ndims.append(a.ndim)
            dims.extend(a.shape)
        n = len(names)
        c_names = (C.c_char_p * n)(*names)
        c_ndims = (C.c_int * n)(*ndims)
        c_dims = (C.c_int64 * max(1, len(dims)))(*dims)
        c_vals = (DP * n)(*[_dp(a) for a in self._keep])
        self.h = L.oracle_model_create(name.encode(), n, c_names, c_ndims, c_dims, c_vals)
        if not self.h:
            raise ValueError(L.oracle_last_error().decode())
        self.dim = L.oracle_model_dim(self.h)
        self.num_out = L.oracle_model_num_out(self.h)

    def __del__(self):
        if getattr(self, "h", None) and getattr(self, "L", None) is not None:
            self.L.oracle_model_destroy(self.h)

    def log_prob_grad(self, q, jacobian=True):
        q = _f64(q)
        g = np.zeros(self.dim)
        lp = self.L.oracle_log_prob_grad(self.h, _dp(q), int(jacobian), _dp(g))
        return lp, g

    def log_prob(self, q, jacobian=True):
        q = _f64(q)
        return self.L.oracle_log_prob(self.h, _dp(q), int(jacobian))

    def grad_cs(self, q, h=1e-30, jacobian=True):
        q = _f64(q)
        return np.array([self.L.oracle_log_prob_cs(self.h, _dp(q), k, h, int(jacobian)) for k in range(self.dim)])

    def write_array(self, q):
        q = _f64(q)
        out = np.zeros(self.num_out)
        self.L.oracle_write_array(self.h, _dp(q), _dp(out))
        return out

    def leapfrog(self, q, p, g, V, eps, metric=1, minv=None):
        q, p, g = _f64(q).copy(), _f64(p).copy(), _f64(g).copy()
        Vb = np.array([V], dtype=np.float64)
        minv = _f64(np.ones(self.dim) if minv is None else minv)
        self.L.oracle_leapfrog(self.h, metric, _dp(minv), eps, _dp(q), _dp(p), _dp(g), _dp(Vb))
        return q, p, g, Vb[0]

    def transition_replay(self, q0, uniforms, normals, eps, metric=1, minv=None, jitter=0.0, max_depth=10):
        q0 = _f64(q0)
        u = _f64(uniforms)
        z = _f64(normals)
        if minv is None:
            minv = np.eye(self.dim) if metric == 2 else np.ones(self.dim)
        minv = _f64(minv)
        out_q = np.zeros(self.dim)
        info = np.zeros(9)
        rc = self.L.oracle_transition_replay(self.h, metric, _dp(minv), eps, jitter, max_depth, _dp(q0), _dp(u),
                                            len(u), _dp(z), len(z), _dp(out_q), _dp(info))
        if rc != 0:
            raise RuntimeError(self.L.oracle_last_error().decode())
        keys = ["depth", "n_leapfrog", "divergent", "accept_stat", "energy", "lp", "uniforms_used",
                "normals_used", "stepsize"]
        return out_q, dict(zip(keys, info.tolist()))

    def init_stepsize(self, q0, normals, eps, metric=1, minv=None):
        """base_hmc::init_stepsize from q0 with injected normals; returns
        (final nominal step size, normals used)."""
        q0 = _f64(q0)
        z = _f64(normals)
        if minv is None:
            minv = np.eye(self.dim) if metric == 2 else np.ones(self.dim)
        minv = _f64(minv)
        out = np.zeros(2)
        rc = self.L.oracle_init_stepsize(self.h, metric, _dp(minv), eps, _dp(q0), _dp(z), len(z), _dp(out))
        if rc != 0:
            raise RuntimeError(self.L.oracle_last_error().decode())
        return out[0], int(out[1])

    def run(self, seed, num_chains=4, chain_id0=1, n_threads=None, init_q=None, inv_metric=None, **cfg):
        c = dict(CFG_DEFAULTS)
        for k, v in cfg.items():
            if k not in c:
                raise KeyError(k)
            c[k] = METRICS[v] if k == "metric" and isinstance(v, str) else v
        cfg_arr = _f64([c[k] for k in CFG_KEYS])
        rows = (c["num_warmup"] if c["save_warmup"] else 0)
        rows = (rows + c["thin"] - 1) // c["thin"] if c["save_warmup"] else 0
        rows += (c["num_samples"] + c["thin"] - 1) // c["thin"]
        cols = 7 + self.num_out
        D = self.dim
        msz = D * D if c["metric"] == 2 else D
        draws = np.zeros((num_chains, rows, cols))
        step = np.zeros(num_chains)
        imet = np.zeros((num_chains, msz))
        stats = np.zeros((num_chains, 3))
        iq = None if init_q is None else _f64(np.broadcast_to(init_q, (num_chains, D)))
        im = None if inv_metric is None else _f64(np.broadcast_to(np.ravel(inv_metric), (num_chains, msz)))
        n_threads = n_threads or min(num_chains, os.cpu_count() or 1)
        rc = self.L.oracle_run(self.h, _dp(cfg_arr), seed, chain_id0, num_chains, n_threads,
                              _dp(iq) if iq is not None else None, _dp(im) if im is not None else None,
                              _dp(draws), _dp(step), _dp(imet), _dp(stats))
        if rc != 0:
            raise RuntimeError(self.L.oracle_last_error().decode())
        return dict(draws=draws, stepsize=step, inv_metric=imet, grad_evals=stats[:, 0], leapfrogs=stats[:, 1],
                    seconds=stats[:, 2])


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().oracle_philox(c, k, o)
    return list(o)


def stream_uniform(seed, chain, purpose, idx):
    return lib().oracle_stream_uniform(seed, chain, purpose, idx)


def stream_normal_pair(seed, chain, idx):
    o = np.zeros(2)
    lib().oracle_stream_normal_pair(seed, chain, idx, _dp(o))
    return o


def dual_averaging(accept, mu, delta=0.8, gamma=0.05, kappa=0.75, t0=10):
    a = _f64(accept)
    eps = np.zeros(len(a))
    xb = np.zeros(1)
    lib().oracle_dual_averaging(mu, delta, gamma, kappa, t0, _dp(a), len(a), _dp(eps), _dp(xb))
    return eps, xb[0]


def window_schedule(num_warmup, init_buffer=75, term_buffer=50, window=25):
    col = (C.c_int * num_warmup)()
    upd = (C.c_int * num_warmup)()
    mode = lib().oracle_window_schedule(num_warmup, init_buffer, term_buffer, window, col, upd)
    return mode, np.array(col[:]), np.array(upd[:])
