"""Python mirror of the sampling service interface.

``Model(family, data)`` plays the role of CmdStan's compiled model plus its
``data file=`` context (src/cmdstan/command.hpp:170-171); ``Model.sample``
takes the arguments of ``sample ... algorithm=hmc engine=nuts`` with CmdStan's
names and defaults (src/cmdstan/arguments/arg_sample.hpp, arg_adapt.hpp,
arg_hmc.hpp, arg_nuts.hpp, arg_metric.hpp) and calls ``nuts_sample`` of the C
ABI; ``Model.log_prob_grad`` is the ``log_prob`` method
(src/cmdstan/command_helper.hpp:569-600).
"""
import ctypes as C
import json

import numpy as np

from ._lib import DP, I64P, INTERRUPT_FN, last_error, lib, nuts_config, nuts_output

SAMPLER_COLUMNS = ["lp__", "accept_stat__", "stepsize__", "treedepth__", "n_leapfrog__", "divergent__", "energy__"]
METRICS = {"unit_e": 0, "diag_e": 1, "dense_e": 2}


class SamplerError(RuntimeError):
    pass


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _dp(a):
    return None if a is None else a.ctypes.data_as(DP)


class Data:
    """A var_context: named arrays (from a dict, CmdStan JSON text or R dump text)."""

    def __init__(self, values=None, json_text=None, rdump_text=None):
        L = lib()
        self.h = L.nuts_data_create()
        if json_text is not None:
            if L.nuts_data_parse_json(self.h, json_text.encode()) != 0:
                raise SamplerError(last_error())
        if rdump_text is not None:
            if L.nuts_data_parse_rdump(self.h, rdump_text.encode()) != 0:
                raise SamplerError(last_error())
        for k, v in (values or {}).items():
            a = np.asarray(v)
            is_int = np.issubdtype(a.dtype, np.integer) or a.dtype == bool
            a = _f64(a)
            dims = (C.c_int64 * max(1, a.ndim))(*a.shape)
            if L.nuts_data_add(self.h, k.encode(), a.ndim, dims, _dp(a), int(is_int)) != 0:
                raise SamplerError(last_error())

    def __del__(self):
        if getattr(self, "h", None):
            lib().nuts_data_destroy(self.h)

    def get(self, name):
        """The named array (row-major), reshaped to its dimensions."""
        nd = C.c_int()
        dims = (C.c_int64 * 8)()
        vals = DP()
        if lib().nuts_data_get(self.h, name.encode(), C.byref(nd), dims, C.byref(vals)) != 0:
            raise SamplerError(last_error())
        shape = tuple(dims[i] for i in range(nd.value))
        n = int(np.prod(shape)) if shape else 1
        return np.array([vals[i] for i in range(n)]).reshape(shape)


def data_from_file(path):
    """get_var_context: a .json file is JSON, any other file R dump."""
    text = open(path).read()
    return Data(json_text=text) if path.endswith(".json") else Data(rdump_text=text)


class Fit:
    """Draws and adaptation results of one sampling run."""

    def __init__(self, **kw):
        self.__dict__.update(kw)

    def chain_draws(self, k):
        return self.draws[k]

    def param(self, name, dataset=0):
        cols = self.column_names[dataset]
        j = cols.index(name)
        ks = [k for k in range(len(self.draws)) if self.dataset_of_chain[k] == dataset]
        return np.stack([self.draws[k][:, j] for k in ks], axis=1)  # draws x chains


class Model:
    def __init__(self, family, data):
        datasets = data if isinstance(data, (list, tuple)) else [data]
        self._data = [d if isinstance(d, Data) else (Data(json_text=d) if isinstance(d, str) else Data(d))
                      for d in datasets]
        arr = (C.c_void_p * len(self._data))(*[d.h for d in self._data])
        self.h = lib().nuts_model_create(family.encode(), arr, len(self._data))
        if not self.h:
            raise SamplerError(last_error())
        self.family = family
        self.num_datasets = len(self._data)
        self._load_names()

    @classmethod
    def from_plugin(cls, plugin, keep=None):
        """A model from a nuts_model_plugin structure (a caller-supplied model
        with its own device gradient); `keep` holds whatever must outlive the
        model (the plugin's library handle, its user data)."""
        m = cls.__new__(cls)
        m._data, m._keep = [], keep
        m.h = lib().nuts_model_create_plugin(C.byref(plugin))
        if not m.h:
            raise SamplerError(last_error())
        m.family = "plugin"
        m.num_datasets = 1
        m._load_names()
        return m

    def _load_names(self):
        self.dims, self.num_outputs, self.param_names, self.unconstrained_names = [], [], [], []
        buf = C.create_string_buffer(256)
        for m in range(self.num_datasets):
            d, o = C.c_int(), C.c_int()
            lib().nuts_model_dims(self.h, m, C.byref(d), C.byref(o))
            self.dims.append(d.value)
            self.num_outputs.append(o.value)
            names = []
            for i in range(o.value):
                lib().nuts_model_param_name(self.h, m, i, buf, 256)
                names.append(buf.value.decode())
            self.param_names.append(names)
            un = []
            for i in range(d.value):
                lib().nuts_model_unconstrained_name(self.h, m, i, buf, 256)
                un.append(buf.value.decode())
            self.unconstrained_names.append(un)

    def __del__(self):
        if getattr(self, "h", None):
            lib().nuts_model_destroy(self.h)

    @property
    def dim(self):
        return self.dims[0]

    def transform_inits(self, inits, dataset=0):
        """Constrained values by parameter name (dict or JSON text) to the
        unconstrained vector; parameters left out are NaN (drawn at random)."""
        d = inits if isinstance(inits, Data) else (Data(json_text=inits) if isinstance(inits, str) else Data(inits))
        out = np.zeros(self.dims[dataset])
        if lib().nuts_model_transform_inits(self.h, dataset, d.h, _dp(out)) != 0:
            raise SamplerError(last_error())
        return out

    def log_prob_grad(self, q, jacobian=True, dataset=0, precision=0):
        q = _f64(np.atleast_2d(q))
        n, D = q.shape
        if D != self.dims[dataset]:
            raise ValueError("dimension mismatch")
        lp = np.zeros(n)
        g = np.zeros((n, D))
        if precision in (0, "fp64"):
            rc = lib().nuts_log_prob_grad(self.h, dataset, _dp(q), n, int(jacobian), _dp(lp), _dp(g))
        else:
            p = {"fp32": 1, "bf16": 2}.get(precision, precision)
            rc = lib().nuts_log_prob_grad_precision(self.h, dataset, _dp(q), n, int(jacobian), p, _dp(lp), _dp(g))
        if rc != 0:
            raise SamplerError(last_error())
        return lp, g

    def sample(self, num_chains=1, num_samples=1000, num_warmup=1000, save_warmup=False, thin=1,
               adapt_engaged=True, delta=0.8, gamma=0.05, kappa=0.75, t0=10.0, init_buffer=75, term_buffer=50,
               window=25, max_depth=10, metric="diag_e", stepsize=1.0, stepsize_jitter=0.0, id=1, seed=0,
               init=2.0, inv_metric=None, fused=-1, precision=0, max_saved_params=-1, flow=False, flow_layers=8,
               flow_hidden=128, flow_train_steps=1000, flow_batch=256, flow_lr=1e-3, flow_warmup=300,
               flow_pilot=100, refresh=0, diagnostics=False, energy_error=False, interrupt=None):
        cfg = nuts_config()
        lib().nuts_config_default(C.byref(cfg))
        cfg.num_samples, cfg.num_warmup, cfg.save_warmup, cfg.thin = num_samples, num_warmup, int(save_warmup), thin
        cfg.adapt_engaged, cfg.delta, cfg.gamma, cfg.kappa, cfg.t0 = int(adapt_engaged), delta, gamma, kappa, t0
        cfg.init_buffer, cfg.term_buffer, cfg.window = init_buffer, term_buffer, window
        cfg.max_depth, cfg.metric = max_depth, METRICS[metric] if isinstance(metric, str) else metric
        cfg.stepsize, cfg.stepsize_jitter = stepsize, stepsize_jitter
        cfg.num_chains, cfg.id, cfg.seed, cfg.fused = num_chains, id, seed, fused
        cfg.precision = {"fp64": 0, "fp32": 1, "bf16": 2}.get(precision, precision)
        cfg.max_saved_params = max_saved_params
        cfg.flow = {"off": 0, "auto": 1, "force": 2}.get(flow, int(flow) if not isinstance(flow, str) else -1)
        cfg.flow_layers, cfg.flow_hidden = flow_layers, flow_hidden
        cfg.flow_train_steps, cfg.flow_batch, cfg.flow_lr = flow_train_steps, flow_batch, flow_lr
        cfg.flow_warmup, cfg.flow_pilot = flow_warmup, flow_pilot
        cfg.refresh = refresh
        if interrupt is not None:
            # callbacks::interrupt: a Python callable polled between check intervals; True stops the run
            cfg.interrupt = INTERRUPT_FN(lambda _user: 1 if interrupt() else 0)
        init_arr = None
        if isinstance(init, (int, float)):
            cfg.init_radius = float(init)
        else:
            init_arr = _f64(np.broadcast_to(np.asarray(init, dtype=np.float64),
                                            (num_chains * self.num_datasets, self.dims[0])))
        im = None
        C_tot = num_chains * self.num_datasets
        if inv_metric is not None:
            im = _f64(np.broadcast_to(np.asarray(inv_metric, dtype=np.float64).ravel(),
                                      (C_tot, np.asarray(inv_metric).size)))
        if lib().nuts_config_validate(C.byref(cfg)) != 0:
            raise SamplerError(last_error())
        rows = lib().nuts_output_rows(C.byref(cfg))
        ds_of_chain = [m for m in range(self.num_datasets) for _ in range(num_chains)]
        if max_saved_params < 0:
            cols = [7 + self.num_outputs[m] for m in ds_of_chain]
            names = [SAMPLER_COLUMNS + p for p in self.param_names]
        else:
            cols = [7 + min(self.dims[m], max_saved_params) for m in ds_of_chain]
            names = [SAMPLER_COLUMNS + u[:max_saved_params] for u in self.unconstrained_names]
        total = sum(rows * c for c in cols)
        draws = np.zeros(total)
        step = np.zeros(C_tot)
        dense = cfg.metric == 2
        msz = sum(self.dims[m] ** 2 if dense else self.dims[m] for m in ds_of_chain)
        imet = np.zeros(msz)
        nlf = np.zeros(C_tot, dtype=np.int64)
        nge = np.zeros(C_tot, dtype=np.int64)
        status = np.zeros(C_tot, dtype=np.int32)
        out = nuts_output()
        out.draws, out.stepsize, out.inv_metric = _dp(draws), _dp(step), _dp(imet)
        out.n_leapfrog = nlf.ctypes.data_as(I64P)
        out.grad_evals = nge.ctypes.data_as(I64P)
        out.status = status.ctypes.data_as(C.POINTER(C.c_int))
        dcols = [7 + 3 * self.dims[m] for m in ds_of_chain]
        diag = np.zeros(sum(rows * c for c in dcols)) if diagnostics else None
        out.diagnostics = _dp(diag)
        eerr = np.zeros(C_tot * rows) if energy_error else None
        out.energy_error = _dp(eerr)
        rej = np.zeros(2 * C_tot, dtype=np.int32)
        out.init_rejections = rej.ctypes.data_as(C.POINTER(C.c_int))
        rc = lib().nuts_sample(self.h, C.byref(cfg), _dp(init_arr), _dp(im), C.byref(out))
        err = last_error() if rc != 0 else ""
        per_chain, o, mo, metrics = [], 0, 0, []
        for k in range(C_tot):
            n = rows * cols[k]
            per_chain.append(draws[o:o + n].reshape(rows, cols[k]))
            o += n
            D = self.dims[ds_of_chain[k]]
            if dense:
                metrics.append(imet[mo:mo + D * D].reshape(D, D))
                mo += D * D
            else:
                metrics.append(imet[mo:mo + D])
                mo += D
        diag_rows = None
        if diagnostics:
            diag_rows, o = [], 0
            for k in range(C_tot):
                n = rows * dcols[k]
                diag_rows.append(diag[o:o + n].reshape(rows, dcols[k]))
                o += n
        fit = Fit(draws=per_chain, diagnostics=diag_rows,
                  energy_error=None if eerr is None else [eerr[k * rows:(k + 1) * rows] for k in range(C_tot)],
                  init_rejections=rej.reshape(C_tot, 2), stepsize=step, inv_metric=metrics, n_leapfrog=nlf, grad_evals=nge,
                  status=status, warmup_seconds=out.warmup_seconds, sampling_seconds=out.sampling_seconds,
                  leaf_steps=out.leaf_steps, grad_seconds_per_launch=out.grad_seconds_per_launch,
                  grad_launches_timed=out.grad_launches_timed,
                  grad_chains_per_launch=out.grad_chains_per_launch, flow_used=bool(out.flow_used),
                  flow_train_seconds=out.flow_train_seconds, flow_loss=out.flow_loss,
                  pilot_ess_per_grad_plain=out.pilot_ess_per_grad_plain,
                  pilot_ess_per_grad_flow=out.pilot_ess_per_grad_flow,
                  pilot_ess_per_sec_plain=out.pilot_ess_per_sec_plain,
                  pilot_ess_per_sec_flow=out.pilot_ess_per_sec_flow, pilot_rhat_plain=out.pilot_rhat_plain,
                  pilot_rhat_flow=out.pilot_rhat_flow, tuned_path=out.tuned_path,
                  tuned_blocks_per_cu=out.tuned_blocks_per_cu,
                  tuned_threads_per_chain=out.tuned_threads_per_chain,
                  tuned_launch_seconds=(out.tuned_wave_launch_seconds, out.tuned_block_launch_seconds),
                  column_names=names,
                  dataset_of_chain=ds_of_chain, config=cfg, error=err)
        if rc != 0:
            e = SamplerError(err)
            e.fit = fit  # per-chain status and rejection counts of the failed run
            raise e
        return fit

    def transition_replay(self, q0, uniforms, normals, stepsize, metric="diag_e", inv_metric=None, jitter=0.0,
                          max_depth=10, dataset=0, precision=0, driver="fused"):
        """One transition per chain with injected draws.  driver: "fused" (the
        per-chain kernel), "lockstep" (a wave per chain and the model's batched
        gradient kernel), "lockstep_block" (a 256-thread block per chain) or
        "fused_block" (the per-chain kernel with a block per chain, for models
        whose gradient spreads over a block)."""
        q0 = _f64(np.atleast_2d(q0))
        nch, D = q0.shape
        u = _f64(np.atleast_2d(uniforms))
        z = _f64(np.atleast_2d(normals))
        if u.shape[0] != nch or z.shape[0] != nch:
            raise ValueError("one row of draws per chain")
        m = METRICS[metric] if isinstance(metric, str) else metric
        msz = D * D if m == 2 else D
        base = (np.eye(D).ravel() if m == 2 else np.ones(D)) if inv_metric is None else \
            np.asarray(inv_metric, dtype=np.float64).ravel()
        im = _f64(np.broadcast_to(base, (nch, msz)))
        out_q = np.zeros((nch, D))
        info = np.zeros((nch, 9))
        p = {"fp64": 0, "fp32": 1, "bf16": 2}.get(precision, precision)
        drv = {"fused": 0, "lockstep": 1, "lockstep_block": 2, "fused_block": 3}.get(driver, driver)
        rc = lib().nuts_transition_replay_driver(self.h, dataset, nch, m, stepsize, jitter, max_depth, _dp(im),
                                                 _dp(q0), _dp(u), u.shape[1], _dp(z), z.shape[1], p, drv,
                                                 _dp(out_q), _dp(info))
        if rc != 0:
            raise SamplerError(last_error())
        keys = ["depth", "n_leapfrog", "divergent", "accept_stat", "energy", "lp", "uniforms_used", "normals_used",
                "stepsize"]
        return out_q, [dict(zip(keys, r.tolist())) for r in info]


def read_json_file(path):
    with open(path) as f:
        return f.read()


def data_from_json(path):
    return Data(json_text=read_json_file(path))


__all__ = ["Model", "Data", "Fit", "SamplerError", "SAMPLER_COLUMNS", "json", "data_from_json"]


class Flow:
    """The affine-coupling flow of the learned reparameterisation (GPU), for
    direct use and for its parity tests.  Row arrays are [n][D] float64."""

    def __init__(self, D, n_layers=8, hidden=128, n_glob=0, group_size=1, max_rows=256, seed=1):
        self.h = lib().nuts_flow_create(D, n_layers, hidden, n_glob, group_size, max_rows, seed)
        if not self.h:
            raise SamplerError(last_error())
        self.D, self.max_rows = D, max_rows
        self.n_params = lib().nuts_flow_num_params(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().nuts_flow_destroy(self.h)
            self.h = None

    def _check(self, rc):
        if rc != 0:
            raise SamplerError(last_error())

    def get_params(self):
        p = np.zeros(self.n_params, dtype=np.float32)
        self._check(lib().nuts_flow_get_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._check(lib().nuts_flow_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))))

    def set_standardization(self, loc, scale):
        self._check(lib().nuts_flow_set_standardization(self.h, _dp(_f64(loc)), _dp(_f64(scale))))

    def _rows(self, a):
        a = _f64(np.atleast_2d(a))
        return a, a.shape[0]

    def forward(self, z):
        z, n = self._rows(z)
        x, ld = np.zeros_like(z), np.zeros(n)
        self._check(lib().nuts_flow_forward(self.h, _dp(z), n, _dp(x), _dp(ld)))
        return x, ld

    def inverse(self, x):
        x, n = self._rows(x)
        z, ld = np.zeros_like(x), np.zeros(n)
        self._check(lib().nuts_flow_inverse(self.h, _dp(x), n, _dp(z), _dp(ld)))
        return z, ld

    def pullback_grad(self, z, gx):
        z, n = self._rows(z)
        gx = _f64(np.atleast_2d(gx))
        gz = np.zeros_like(z)
        self._check(lib().nuts_flow_pullback_grad(self.h, _dp(z), _dp(gx), n, _dp(gz)))
        return gz

    def loss_grad(self, x):
        x, n = self._rows(x)
        loss = C.c_double()
        g = np.zeros(self.n_params, dtype=np.float32)
        self._check(lib().nuts_flow_loss_grad(self.h, _dp(x), n, C.byref(loss), g.ctypes.data_as(C.POINTER(C.c_float))))
        return loss.value, g

    def train(self, x, steps, batch=256, lr=1e-3, seed=1):
        x, n = self._rows(x)
        loss = C.c_double()
        self._check(lib().nuts_flow_train(self.h, _dp(x), n, steps, batch, lr, seed, C.byref(loss)))
        return loss.value
