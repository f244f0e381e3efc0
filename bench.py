"""Benchmark: NUTS for Bayesian logistic regression on one MI355X.

Workload (the configuration the headline metric is quoted on): X 100000 x 100
FP64 iid N(0,1); beta* ~ N(0, 1/100); y ~ Bernoulli(logistic(X beta*)); prior
beta ~ N(0, 2.5^2); 1024 chains, diag_e, max_depth 10, 1000 warmup + 1000
draws.

Steps.  A step is a block of S = max(50, ceil(1000 / K)) sampling iterations
of all chains, so the K timed steps always hold at least the workload's 1000
draws per chain (--steps 20: S = 50, 1000 draws).  The adaptation phase runs
max(1000, W * S) iterations (the workload's 1000 warmup iterations, or more
when --warmup W asks for more) and is not timed.  The timed region is the
sampling phase: the library synchronises the GPU stream on both sides of it.

value = min bulk-ESS over the 100 coefficients and lp__ divided by the
sampling seconds, for plain FP64 NUTS.  It is null (with the reason) when the
run has not converged: max R-hat >= 1.01.  Gradient evaluations per second
count every leapfrog of every chain in the sampling phase.

Second leg of the metric ("batched+learned vs plain"): the same workload with
the FP32 contraction and the learned flow in automatic mode (trained on the
warmup draws, kept only if its pilot raises min bulk-ESS per gradient and per
second, else plain NUTS continues); its min bulk-ESS/s over sampling and over
the whole job, and its ratio to the plain FP64 leg.

cpu_baseline: the FP64 CPU oracle (a restatement of Stan's algorithm; CmdStan
itself cannot be built offline) runs a bounded sample of post-warmup
transitions with the GPU-adapted step size and metric on the host cores; its
ESS/s is extrapolated with the GPU run's ESS per gradient evaluation.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
PEAK_FP32_TFLOPS = 157.3  # MI355X dense FP32 matrix
PEAK_BF16_TFLOPS = 2500.0  # MI355X dense BF16 matrix


def reduced_kernel(precision, path, D):
    """(kernel name, matrix-core peak TFLOP/s, matrix products per algorithmic
    multiply-add) of the reduced-precision logistic-regression gradient the
    autotuner picked: path 0 single-pass FP32 cores; path 1 two-pass (BF16
    cores with three-term operands when X is stored in BF16 and D % 8 == 0,
    else FP32 cores); path 2 single-pass BF16 cores with three-term operands
    (six products with X in FP32, three with X stored in BF16)."""
    if path == 2:
        return "glm_mfma_bf16", PEAK_BF16_TFLOPS, 6 if precision == "fp32" else 3
    if path == 1:
        if precision == "bf16" and D % 8 == 0 and os.environ.get("CMDSTAN_AMD_GL_BF16", "1") != "0":
            return "gbl_eta_kernel + gbl_grad_kernel", PEAK_BF16_TFLOPS, 3
        return "gl_eta_kernel + gl_grad_kernel", PEAK_FP32_TFLOPS, 1
    return "glm_mfma_f32", PEAK_FP32_TFLOPS, 1


def measured_traffic(kernel, args):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE + WRITE_SIZE, scaled by the same
    counters on a copy of known size), or None when no summary covers it."""
    path = os.path.join(ROOT, "profiles",
                        f"{kernel}_{args.precision}_logistic_N{args.N}_D{args.D}_C{args.chains}_counters.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def copy_probe(nbytes):
    """A plain device copy of known size: the calibration dispatch for the
    FETCH_SIZE / WRITE_SIZE counters (read nbytes, write nbytes)."""
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    del x, y


def make_data(N, D, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((N, D))
    beta = rng.normal(0.0, np.sqrt(1.0 / D), D)
    eta = X @ beta
    y = (rng.random(N) < 1.0 / (1.0 + np.exp(-eta))).astype(np.int64)
    return dict(N=N, D=D, X=X, y=y, prior_scale=2.5), beta


def host_threads():
    """Host threads this process may use: the CPU affinity set, capped by
    OMP_NUM_THREADS when the host sets it (a host shared by several GPUs gives
    each GPU a CPU share and states it there)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def default_sizes(config):
    """(steps, warmup) when the flags are not given: a run without flags
    finishes in a few minutes and measures the configuration the documented
    figures are quoted on (C2: 20 steps of 50 iterations after 1000 warmup
    iterations)."""
    return {"c1": (1000, 1000), "c2": (20, 5), "c3": (200, 300), "c4": (300, 500),
            "c4f": (30, 0), "c5": (100, 150)}[config]


def step_iterations(steps):
    """Sampling iterations per step: K steps hold at least 1000 draws."""
    return max(50, -(-1000 // max(1, steps)))


def cpu_baseline(data, fit, ess_per_grad, seconds_budget=20.0):
    """Oracle transitions on host threads from GPU-adapted states."""
    from oracle import oracle as O
    om = O.Model("logistic_regression", data)
    threads = host_threads()
    D = data["D"]
    n_trans = 4
    C = len(fit.draws)
    picks = np.linspace(0, C - 1, threads).astype(int)
    grads, secs = 0.0, 0.0
    t_start = time.time()
    while time.time() - t_start < seconds_budget * 0.5:
        init = np.stack([fit.draws[k][-1, 7:7 + D] for k in picks])
        step = float(np.median(fit.stepsize))
        minv = np.median(np.stack(fit.inv_metric), axis=0)
        r = om.run(seed=int(time.time()) & 0xFFFF, num_chains=threads, n_threads=threads, init_q=init,
                   inv_metric=minv, adapt=0, num_warmup=0, num_samples=n_trans, stepsize=step)
        grads += r["grad_evals"].sum()
        secs += r["seconds"].max()
    gps = grads / secs
    return dict(value=ess_per_grad * gps, unit="ESS/s (extrapolated)", cores=threads, kind="port",
                grad_evals_per_sec=gps, cpu_model=cpu_model(), host_nproc=os.cpu_count(),
                sample=f"{threads} chains x {n_trans} post-warmup transitions per round of the FP64 oracle with the "
                       f"GPU-adapted step size and metric, one chain per thread on {threads} host threads (the "
                       f"process's CPU share), {secs:.1f} s CPU wall; ESS/s = GPU ESS per gradient x CPU "
                       f"gradients/s")


def leg_stats(fit, D):
    """min bulk-ESS over the D coefficients and lp__, max R-hat, sampling
    gradient evaluations, and the per-transition energy error when reported."""
    from cmdstan_amd.summary import summarize
    arr = np.stack(fit.draws)  # chains x draws x cols
    params = np.concatenate([arr[:, :, 7:7 + D], arr[:, :, :1]], axis=2).transpose(2, 0, 1)
    st = summarize(params)
    out = dict(min_bulk_ess=float(np.nanmin(st[:, 7])), rhat_max=float(np.nanmax(st[:, 9])),
               grad_evals=float(arr[:, :, 4].sum()))
    ee = getattr(fit, "energy_error", None)
    if ee is not None:
        e = np.concatenate([np.asarray(x).ravel() for x in ee])
        out["energy_error"] = dict(max_abs_dH_per_transition_mean=float(np.mean(e)),
                                   max_abs_dH_per_transition_p99=float(np.quantile(e, 0.99)),
                                   max_abs_dH_per_transition_max=float(np.max(e)))
    return out


def learned_leg(model, args, n_draws, plain, plain_samp_s):
    """Batched FP32 contraction + learned flow (automatic: plain NUTS is the
    fallback), same chains, seed and draw count as the plain FP64 leg."""
    D = args.D
    print("c2 batched fp32 + flow auto", file=sys.stderr, flush=True)
    t0 = time.time()
    fit = model.sample(num_chains=args.chains, num_warmup=1000, num_samples=n_draws, seed=args.seed,
                       metric="diag_e", fused=0, precision="fp32", flow="auto", flow_layers=8, flow_hidden=128,
                       flow_train_steps=args.flow_steps, flow_batch=256, flow_warmup=args.flow_warmup,
                       flow_pilot=args.flow_pilot, energy_error=True)
    wall = time.time() - t0
    st = leg_stats(fit, D)
    ok = st["rhat_max"] < 1.01
    ess_s = st["min_bulk_ess"] / fit.sampling_seconds
    total_s = fit.sampling_seconds + fit.warmup_seconds
    plain_ess_s = plain["min_bulk_ess"] / plain_samp_s
    # error of the FP32 contraction against FP64 at the sampler's own draws
    q = np.stack([fit.draws[k][-1, 7:7 + D] for k in range(0, len(fit.draws), max(1, len(fit.draws) // 64))])
    lp64, g64 = model.log_prob_grad(q)
    lp32, g32 = model.log_prob_grad(q, precision="fp32")
    roof = None
    if fit.grad_seconds_per_launch > 0 and fit.grad_chains_per_launch > 0 and not fit.flow_used:
        t = fit.grad_seconds_per_launch
        ach = 4.0 * args.N * D * fit.grad_chains_per_launch / t / 1e12
        kname, peak, nprod = reduced_kernel("fp32", fit.tuned_path, D)
        roof = dict(bound="mfma", achieved=ach, peak=peak, unit="TFLOP/s", frac=ach / peak, kernel=kname,
                    launch_ms=t * 1e3, mean_active_chains=fit.grad_chains_per_launch,
                    products_per_multiply_add=nprod, frac_of_product_ceiling=ach * nprod / peak)
    return dict(
        path="FP32 contraction (X FP32) + affine-coupling flow (8 layers, MLP 2x128) in automatic mode",
        autotuned_gradient=["single-pass", "two-pass", "single-pass bf16-split"][fit.tuned_path]
        if fit.tuned_path >= 0 else None, roofline=roof,
        flow_used=fit.flow_used, value=ess_s if ok else None,
        value_null_reason=None if ok else f"not converged: max R-hat {st['rhat_max']:.4f} >= 1.01",
        min_bulk_ess=st["min_bulk_ess"], rhat_max=st["rhat_max"],
        ess_per_sec_sampling=ess_s, ess_per_sec_total=st["min_bulk_ess"] / total_s,
        sampling_seconds=fit.sampling_seconds, warmup_seconds=fit.warmup_seconds, wall_seconds=wall,
        grad_evals_per_sec=st["grad_evals"] / fit.sampling_seconds,
        ess_per_grad=st["min_bulk_ess"] / st["grad_evals"],
        pilot_ess_per_grad_plain=fit.pilot_ess_per_grad_plain, pilot_ess_per_grad_flow=fit.pilot_ess_per_grad_flow,
        pilot_ess_per_sec_plain=fit.pilot_ess_per_sec_plain, pilot_ess_per_sec_flow=fit.pilot_ess_per_sec_flow,
        flow_train_seconds=fit.flow_train_seconds, energy_error=st.get("energy_error"),
        ratio_vs_plain_sampling=ess_s / plain_ess_s,
        ratio_vs_plain_total=(st["min_bulk_ess"] / total_s) /
                             (plain["min_bulk_ess"] / (plain_samp_s + args._plain_warm_s)),
        fp32_error_vs_fp64=dict(points=int(len(q)),
                                lp_abs_max=float(np.max(np.abs(lp32 - lp64))),
                                lp_rel_max=float(np.max(np.abs(lp32 - lp64) / np.maximum(1.0, np.abs(lp64)))),
                                grad_rel_max=float(np.max(np.abs(g32 - g64) /
                                                          np.maximum(1.0, np.abs(g64).max(axis=1, keepdims=True))))))


def bench_c1(args):
    """examples/bernoulli shape: 4 chains, diag_e, 1000+1000, FP64; CPU oracle full run beside it."""
    import cmdstan_amd as ca
    from cmdstan_amd.summary import summarize
    from oracle import oracle as O
    data = dict(N=10, y=np.array([0, 1, 0, 0, 0, 0, 0, 0, 0, 1]))
    m = ca.Model("bernoulli", data)
    fit = m.sample(num_chains=4, num_warmup=args.warmup, num_samples=args.steps, seed=args.seed)
    arr = np.stack(fit.draws)
    st = summarize(arr[:, :, 7][None])
    ess = float(st[0, 7])
    t0 = time.time()
    r = O.Model("bernoulli", data).run(seed=args.seed, num_chains=4, n_threads=4, num_warmup=args.warmup,
                                       num_samples=args.steps)
    cpu_s = float(np.max(r["seconds"]))
    from oracle import ess as E
    cpu_ess = E.summary(r["draws"][:, :, 7].T)["ess_bulk"]
    return {"metric": "min bulk-ESS/sec (bernoulli, 4 chains)", "value": ess / fit.sampling_seconds, "unit": "ESS/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": fit.sampling_seconds / args.steps * 1e3, "higher_is_better": True, "dtype": "f64",
            "data": "examples/bernoulli/bernoulli.data.json",
            "config": {"workload": "bernoulli N=10, 4 chains, diag_e, fused per-chain kernel"},
            "grad_evals_per_sec": float(arr[:, :, 4].sum()) / fit.sampling_seconds,
            "sampling_seconds": fit.sampling_seconds, "warmup_seconds": fit.warmup_seconds,
            "theta_mean": float(arr[:, :, 7].mean()),
            "cpu_baseline": {"value": cpu_ess / cpu_s, "unit": "ESS/s over the whole CPU run (warmup+sampling)",
                             "cores": 4, "kind": "port", "seconds": cpu_s,
                             "sample": "full 4-chain run of the FP64 oracle, one thread per chain"}}


def make_c3(n_models, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_models):
        J = int(rng.integers(8, 2001))
        sigma = rng.uniform(5, 20, J)
        mu = rng.normal(0, 5)
        tau = abs(rng.standard_cauchy()) * 5
        theta = mu + tau * rng.standard_normal(J)
        out.append(dict(J=J, y=rng.normal(theta, sigma), sigma=sigma))
    return out


def bench_c3(args):
    """512 non-centered hierarchical normal models (J ~ U{8..2000}), 64 chains each,
    all in one batch, against running the models one at a time."""
    import cmdstan_amd as ca
    from cmdstan_amd.summary import summarize
    nm = args.models
    datasets = make_c3(nm, args.seed)
    keep = 10  # mu, log tau, eta[1..8] per chain
    m = ca.Model("hier_normal_nc", datasets)
    print(f"c3: {nm} models, batched", file=sys.stderr, flush=True)
    # batched: FP64 or the FP32 leapfrog (--precision fp32); model at a time: plain FP64
    fit = m.sample(num_chains=64, num_warmup=args.warmup, num_samples=args.steps, seed=args.seed,
                   max_saved_params=keep, refresh=50, precision=args.precision)
    arr = np.stack(fit.draws)  # chains x draws x (7 + keep)
    C = arr.shape[0]
    ess = []
    for k in range(nm):
        blk = arr[64 * k:64 * (k + 1)]
        st = summarize(blk[:, :, 7:7 + keep].transpose(2, 0, 1))
        ess.append(np.nanmin(st[:, 7]))
    min_ess = float(np.min(ess))
    grads = float(arr[:, :, 4].sum())
    Js = np.array([d["J"] for d in datasets])
    dtype = "f64" if args.precision == "fp64" else "f32 vectors, f64 sums (FP32 leapfrog)"
    rhat = max(float(np.nanmax(summarize(arr[64 * k:64 * (k + 1), :, 7:7 + keep].transpose(2, 0, 1))[:, 9]))
               for k in range(nm))
    if args.c3_batched_only:
        return {"metric": "min bulk-ESS/sec, 512 uneven models batched in one launch",
                "value": min_ess / fit.sampling_seconds, "unit": "ESS/s", "n_gpus": 1, "steps": args.steps,
                "warmup": args.warmup, "dtype": dtype, "grad_evals_per_sec": grads / fit.sampling_seconds,
                "rhat_max": rhat, "sampling_seconds": fit.sampling_seconds, "warmup_seconds": fit.warmup_seconds}
    # model at a time on a spread of sizes, extrapolated to all models
    picks = np.argsort(Js)[np.linspace(0, nm - 1, min(8, nm)).astype(int)]
    t_plain = []
    for k in picks:
        print(f"c3: model J={Js[k]} alone", file=sys.stderr, flush=True)
        f1 = ca.Model("hier_normal_nc", datasets[k]).sample(num_chains=64, num_warmup=args.warmup,
                                                             num_samples=args.steps, seed=args.seed,
                                                             max_saved_params=keep, refresh=100)
        t_plain.append(f1.sampling_seconds)
    # per-model time grows with J: fit time = a + b J over the picks
    A = np.column_stack([np.ones(len(picks)), Js[picks]])
    coef, *_ = np.linalg.lstsq(A, np.array(t_plain), rcond=None)
    plain_total = float(np.sum(coef[0] + coef[1] * Js))
    return {"metric": "min bulk-ESS/sec, 512 uneven models batched in one launch", "value": min_ess / fit.sampling_seconds,
            "unit": "ESS/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": fit.sampling_seconds / args.steps * 1e3, "higher_is_better": True, "dtype": dtype,
            "data": "synthetic",
            "config": {"workload": f"{nm} non-centered hierarchical normal models, J~U{{8..2000}}, 64 chains each "
                                   f"({C} chains), diag_e, fused per-chain kernel; ESS over mu, log tau, eta[1..8]"},
            "grad_evals_per_sec": grads / fit.sampling_seconds, "rhat_max": rhat,
            "sampling_seconds": fit.sampling_seconds, "warmup_seconds": fit.warmup_seconds, "sum_J": int(Js.sum()),
            "plain_model_at_a_time": {"sampling_seconds_extrapolated": plain_total,
                                      "min_bulk_ess_per_sec": min_ess / plain_total,
                                      "measured_models": [int(Js[k]) for k in picks],
                                      "measured_seconds": t_plain},
            "speedup_batched_vs_plain": plain_total / fit.sampling_seconds}


def make_c4(seed, N=50000, G=400, K=10):
    """Centered hierarchical logistic regression: 400 groups of 125 observations,
    10 covariates N(0,1) held in FP32, beta_g ~ N(mu, tau), tau ~ HalfNormal(1)."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((N, K)).astype(np.float32).astype(np.float64)
    grp = np.repeat(np.arange(G), N // G)
    mu = rng.normal(0, 1, K)
    tau = np.abs(rng.normal(0, 1, K))
    beta = mu + tau * rng.standard_normal((G, K))
    eta = np.einsum("nk,nk->n", X, beta[grp])
    y = (rng.random(N) < 1 / (1 + np.exp(-eta))).astype(np.int64)
    return dict(N=N, K=K, G=G, X=X, y=y, group=grp + 1)


def bench_c4(args):
    """Learned flow vs plain NUTS on the centered hierarchical logistic model
    (D = 4020, funnel geometry), 256 chains each."""
    import cmdstan_amd as ca
    from cmdstan_amd.summary import summarize
    data = make_c4(args.seed)
    m = ca.Model("hier_logistic_cp", data)
    D = m.dims[0]

    def min_ess(fit):
        arr = np.stack(fit.draws)  # chains x draws x (7 + D)
        params = np.concatenate([arr[:, :, 7:7 + D], arr[:, :, :1]], axis=2).transpose(2, 0, 1)
        st = summarize(params)
        return float(np.nanmin(st[:, 7])), float(np.nanmax(st[:, 9])), float(arr[:, :, 4].sum())

    out = {}
    fk = dict(flow_layers=8, flow_hidden=128, flow_train_steps=args.flow_steps, flow_batch=256, flow_lr=1e-3,
              flow_warmup=args.flow_warmup, flow_pilot=args.flow_pilot)
    runs = [("plain", {}), ("flow", dict(fk, flow="auto")), ("flow_forced", dict(fk, flow="force"))]
    for name, kw in runs:
        if args.c4_only and name != args.c4_only:
            continue
        t0 = time.time()
        print(f"c4 {name}: sampling", file=sys.stderr, flush=True)
        fit = m.sample(num_chains=256, num_warmup=args.warmup, num_samples=args.steps, seed=args.seed, refresh=50,
                       **kw)
        wall = time.time() - t0
        ess, rhat, grads = min_ess(fit)
        r = dict(min_bulk_ess=ess, rhat_max=rhat, sampling_seconds=fit.sampling_seconds,
                 warmup_seconds=fit.warmup_seconds, wall_seconds=wall,
                 ess_per_sec_sampling=ess / fit.sampling_seconds,
                 ess_per_sec_total=ess / (fit.sampling_seconds + fit.warmup_seconds),
                 ess_per_grad=ess / grads, grad_evals_per_sec=grads / fit.sampling_seconds,
                 divergent=float(np.stack(fit.draws)[:, :, 5].sum()))
        if kw:
            r.update(flow_used=fit.flow_used, flow_train_seconds=fit.flow_train_seconds, flow_loss=fit.flow_loss,
                     pilot_ess_per_grad_plain=fit.pilot_ess_per_grad_plain,
                     pilot_ess_per_grad_flow=fit.pilot_ess_per_grad_flow,
                     pilot_ess_per_sec_plain=fit.pilot_ess_per_sec_plain,
                     pilot_ess_per_sec_flow=fit.pilot_ess_per_sec_flow, pilot_rhat_plain=fit.pilot_rhat_plain,
                     pilot_rhat_flow=fit.pilot_rhat_flow)
        out[name] = r
        print(json.dumps({name: r}), file=sys.stderr, flush=True)
    if args.c4_only:
        return out
    return {"metric": "min bulk-ESS/sec, learned flow vs plain NUTS (centered hierarchical logistic)",
            "value": out["flow"]["ess_per_sec_sampling"], "flow_forced": out["flow_forced"], "unit": "ESS/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "higher_is_better": True, "dtype": "f64 sampler, f32 flow", "data": "synthetic",
            "config": {"workload": "centered hierarchical logistic N=50000, 400 groups x 10 covariates (D=4020), "
                                   "256 chains, diag_e; flow 8 coupling layers, MLP 2x128 FP32"},
            "plain": out["plain"], "flow": out["flow"],
            "speedup_sampling": out["flow"]["ess_per_sec_sampling"] / out["plain"]["ess_per_sec_sampling"],
            "speedup_total": out["flow"]["ess_per_sec_total"] / out["plain"]["ess_per_sec_total"]}


def bench_c4_flow_kernels(args):
    """The flow alone at the C4 shape: forward T(z) and the pullback of a
    gradient for 256 rows, D = 4020, 8 coupling layers, MLP 2 x 128 (the
    kernels a flow leaf adds to a plain one); wall time per forward + pullback
    including the host copies of z and the results (a kernel trace gives the
    GPU time).  CMDSTAN_AMD_FLOW_FUSED=1 selects the fused per-layer kernels."""
    import cmdstan_amd as ca
    D, L, H, rows = 4020, 8, 128, 256
    g = ca.Flow(D, n_layers=L, hidden=H, n_glob=20, group_size=10, max_rows=rows, seed=3)
    rng = np.random.default_rng(args.seed)
    z = rng.standard_normal((rows, D))
    gx = rng.standard_normal((rows, D))
    for _ in range(3):
        g.forward(z)
        g.pullback_grad(z, gx)
    n = max(1, args.steps)
    t0 = time.time()
    for _ in range(n):
        g.forward(z)
        g.pullback_grad(z, gx)
    wall = (time.time() - t0) / n
    return {"metric": "flow forward + pullback wall time (C4 shape)", "value": wall * 1e3, "unit": "ms",
            "higher_is_better": False, "n_gpus": 1, "steps": n, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"affine-coupling flow D={D}, {L} layers, MLP 2x{H}, {rows} rows",
                       "fused_layer_kernels": os.environ.get("CMDSTAN_AMD_FLOW_FUSED", "0") not in ("", "0")}}


def bench_c5(args):
    """Logistic regression N=1,000,000 x D=1000 (X stored BF16 or FP32), 512
    chains, dense_e; the two products on BF16 matrix cores with three-term
    operands (X in BF16) or FP32 matrix cores (X in FP32), FP64 sums."""
    import cmdstan_amd as ca
    from cmdstan_amd.summary import summarize
    N, D, C = args.N5, args.D5, 512
    rng = np.random.default_rng(args.seed)
    print(f"c5: generating X {N}x{D}", file=sys.stderr, flush=True)
    X = rng.standard_normal((N, D), dtype=np.float32).astype(np.float64)
    beta = rng.normal(0, np.sqrt(1.0 / D), D)
    y = (rng.random(N) < 1 / (1 + np.exp(-(X @ beta)))).astype(np.int64)
    m = ca.Model("logistic_regression", dict(N=N, D=D, X=X, y=y))
    del X
    prec = "bf16" if args.precision == "fp64" else args.precision
    print(f"c5: sampling ({prec})", file=sys.stderr, flush=True)
    fit = m.sample(num_chains=C, num_warmup=args.warmup, num_samples=args.steps, seed=args.seed, metric="dense_e",
                   precision=prec, refresh=25)
    arr = np.stack(fit.draws)
    params = np.concatenate([arr[:, :, 7:7 + D], arr[:, :, :1]], axis=2).transpose(2, 0, 1)
    st = summarize(params)
    ess = float(np.nanmin(st[:, 7]))
    grads = float(arr[:, :, 4].sum())
    t = fit.grad_seconds_per_launch
    flops = 4.0 * N * D * fit.grad_chains_per_launch
    # X stored in BF16 (D a multiple of 8) runs on BF16 matrix cores with the
    # other operand as three BF16 terms: three BF16 products per algorithmic
    # multiply-add, so the ceiling at this accuracy is a third of the BF16 peak
    bf16_cores = prec == "bf16" and D % 8 == 0 and os.environ.get("CMDSTAN_AMD_GL_BF16", "1") != "0"
    peak = PEAK_BF16_TFLOPS if bf16_cores else PEAK_FP32_TFLOPS
    roof = dict(bound="mfma", achieved=flops / t / 1e12, peak=peak, unit="TFLOP/s",
                frac=flops / t / 1e12 / peak, traffic=None,
                kernel="gbl_eta_kernel + gbl_grad_kernel" if bf16_cores else "gl_eta_kernel + gl_grad_kernel",
                launch_ms=t * 1e3, mean_active_chains=fit.grad_chains_per_launch) if t > 0 else None
    if roof and bf16_cores:
        roof["products_per_multiply_add"] = 3
        roof["frac_of_three_product_ceiling"] = roof["frac"] * 3
    return {"metric": "min bulk-ESS/sec, large logistic regression (dense_e)", "value": ess / fit.sampling_seconds,
            "unit": "ESS/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": fit.sampling_seconds / args.steps * 1e3, "higher_is_better": True,
            "dtype": f"f32 contraction (X {prec}), f64 sums", "data": "synthetic",
            "config": {"workload": f"logistic regression N={N} D={D}, {C} chains, dense_e, X stored {prec}"},
            "grad_evals_per_sec": grads / fit.sampling_seconds, "min_bulk_ess": ess,
            "rhat_max": float(np.nanmax(st[:, 9])), "sampling_seconds": fit.sampling_seconds,
            "warmup_seconds": fit.warmup_seconds,
            "autotuned_state_machine": {"threads_per_chain": fit.tuned_threads_per_chain,
                                        "median_launch_us_wave": fit.tuned_launch_seconds[0] * 1e6,
                                        "median_launch_us_block": fit.tuned_launch_seconds[1] * 1e6},
            "roofline": roof}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps: c2 steps of at least 50 iterations holding >= 1000 draws (default 20); "
                         "other configs: sampling iterations (defaults c1 1000, c3 200, c4 300, c5 100, c4f 30)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed warmup: c2 steps (default 5; at least 1000 iterations); other configs: "
                         "iterations (defaults c1 1000, c3 300, c4 500, c5 150)")
    ap.add_argument("--chains", type=int, default=1024)
    ap.add_argument("--N", type=int, default=100000)
    ap.add_argument("--D", type=int, default=100)
    ap.add_argument("--seed", type=int, default=20240601)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned-leg", action="store_true",
                    help="c2: skip the batched FP32 + flow leg of the metric")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "bf16"],
                    help="GLM contraction precision (sums, H and tree arithmetic stay FP64)")
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4", "c4f", "c5"],
                    help="c2 (default): the headline workload; c1, c3: secondary measurements")
    ap.add_argument("--models", type=int, default=512, help="c3: number of models")
    ap.add_argument("--c3-batched-only", action="store_true", help="c3: skip the model-at-a-time comparison")
    ap.add_argument("--flow-steps", type=int, default=1000, help="c4: Adam steps")
    ap.add_argument("--N5", type=int, default=1000000, help="c5: observations")
    ap.add_argument("--D5", type=int, default=1000, help="c5: coefficients")
    ap.add_argument("--flow-warmup", type=int, default=300, help="c4: warmup iterations on the flow")
    ap.add_argument("--flow-pilot", type=int, default=100, help="c4: pilot draws per sampler")
    ap.add_argument("--c4-only", default="", choices=["", "plain", "flow", "flow_forced"],
                    help="c4: run one sampler only")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="first run a device copy of BYTES (counter calibration for profiles)")
    args = ap.parse_args()
    steps0, warm0 = default_sizes(args.config)
    if args.steps is None:
        args.steps = steps0
    if args.warmup is None:
        args.warmup = warm0
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the chain-sharded multi-GPU design is documented "
                                   "but not built", "n_gpus": args.gpus}))
        sys.exit(2)

    if args.copy_probe:
        copy_probe(args.copy_probe)
    if args.config == "c1":
        print(json.dumps(bench_c1(args)))
        return
    if args.config == "c3":
        print(json.dumps(bench_c3(args)))
        return
    if args.config == "c4":
        print(json.dumps(bench_c4(args)))
        return
    if args.config == "c4f":
        print(json.dumps(bench_c4_flow_kernels(args)))
        return
    if args.config == "c5":
        print(json.dumps(bench_c5(args)))
        return

    import cmdstan_amd as ca

    data, _ = make_data(args.N, args.D, args.seed)
    model = ca.Model("logistic_regression", data)
    S = step_iterations(args.steps)
    n_draws = args.steps * S
    n_warm = max(1000, args.warmup * S)
    C, D = args.chains, args.D
    print(f"c2 plain {args.precision}: {n_warm} warmup + {n_draws} draws ({args.steps} steps of {S} iterations)",
          file=sys.stderr, flush=True)
    t0 = time.time()
    fit = model.sample(num_chains=C, num_warmup=n_warm, num_samples=n_draws, seed=args.seed,
                       metric="diag_e", fused=0, precision=args.precision, energy_error=True)
    wall = time.time() - t0
    plain = leg_stats(fit, D)
    samp_s = fit.sampling_seconds
    value = plain["min_bulk_ess"] / samp_s
    reason = None
    if not plain["rhat_max"] < 1.01:
        value, reason = None, f"not converged: max R-hat {plain['rhat_max']:.4f} >= 1.01"
    # roofline of the gradient's contraction kernel: 4 N D flops per active chain
    # per launch (X.B and X^T.R), averaged over every lockstep launch of the run
    roof = None
    if fit.grad_seconds_per_launch > 0 and fit.grad_chains_per_launch > 0:
        t = fit.grad_seconds_per_launch
        flops_per_launch = 4.0 * args.N * D * fit.grad_chains_per_launch
        ach = flops_per_launch / t / 1e12
        if args.precision == "fp64":
            # even D runs the double-buffered LDS-copy variant
            kname = "glm_mfma_f64_lds" if D % 2 == 0 else "glm_mfma_f64"
            peak, nprod = PEAK_FP64_TFLOPS, 1
        else:
            kname, peak, nprod = reduced_kernel(args.precision, fit.tuned_path, D)
        roof = dict(bound="mfma", achieved=ach, peak=peak, unit="TFLOP/s", frac=ach / peak,
                    traffic=measured_traffic(kname, args), kernel=kname, launch_ms=t * 1e3,
                    launches=fit.grad_launches_timed, mean_active_chains=fit.grad_chains_per_launch,
                    flops_per_launch=flops_per_launch)
        if nprod > 1:
            roof["products_per_multiply_add"] = nprod
            roof["frac_of_product_ceiling"] = ach * nprod / peak
    learned = None
    if not args.no_learned_leg:
        args._plain_warm_s = fit.warmup_seconds
        learned = learned_leg(model, args, n_draws, plain, samp_s)
    cpu = None
    if not args.no_cpu_baseline and reason is None:
        cpu = cpu_baseline(data, fit, plain["min_bulk_ess"] / plain["grad_evals"])
    line = {
        "metric": "min bulk-ESS/sec at 1 GPU, batched+learned vs plain NUTS; gradient evals/sec",
        "value": value,
        "value_null_reason": reason,
        "unit": "ESS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": samp_s / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else f"f32 contraction (X {args.precision}), f64 sums",
        "data": "synthetic",
        "config": {"workload": f"{'plain FP64' if args.precision == 'fp64' else 'batched ' + args.precision} NUTS, "
                               f"Bayesian logistic regression N={args.N} D={D}, {C} chains, diag_e, "
                               f"{n_warm} warmup + {n_draws} draws",
                   "chains": C, "N": args.N, "D": D, "max_depth": 10, "path": "lockstep + batched gradient kernel",
                   "iterations_per_step": S, "warmup_iterations": n_warm, "draws_per_chain": n_draws},
        "grad_evals_per_sec": plain["grad_evals"] / samp_s,
        "min_bulk_ess": plain["min_bulk_ess"],
        "rhat_max": plain["rhat_max"],
        "sampling_seconds": samp_s,
        "warmup_seconds": fit.warmup_seconds,
        "total_wall_seconds": wall,
        "mean_n_leapfrog": plain["grad_evals"] / (C * n_draws),
        "energy_error": plain.get("energy_error"),
        "autotuned_gradient": {"path": ["single-pass", "two-pass", "single-pass bf16-split"][fit.tuned_path]
                               if fit.tuned_path >= 0 else None,
                               "blocks_per_cu": fit.tuned_blocks_per_cu},
        "autotuned_state_machine": {"threads_per_chain": fit.tuned_threads_per_chain,
                                    "median_launch_us_wave": fit.tuned_launch_seconds[0] * 1e6,
                                    "median_launch_us_block": fit.tuned_launch_seconds[1] * 1e6},
        "roofline": roof,
        "batched_learned": learned,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))

if __name__ == "__main__":
    main()
