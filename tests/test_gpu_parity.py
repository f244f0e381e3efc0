"""Parity of the HIP sampler (through the C ABI) with the FP64 CPU oracle."""
import numpy as np
import pytest

import cmdstan_amd as ca
from oracle import ess as E
from oracle import oracle as O
from tests.model_cases import BERN, CASES, EIGHT, hier_normal, logistic

pytestmark = pytest.mark.gpu


def _grad_scale(name, data, q, g_ref):
    """Conditioning of the gradient sum: sum of |terms| per component."""
    if name == "logistic_regression":
        X = np.asarray(data["X"], dtype=float)
        y = np.asarray(data["y"], dtype=float)
        res = np.abs(1 / (1 + np.exp(-X @ q)) - y)
        return np.abs(X).T @ res + np.abs(q) / 6.25 + 1.0
    return np.maximum(1.0, np.abs(g_ref)) * 8


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("jacobian", [True, False])
def test_log_prob_grad_parity(name, jacobian):
    data = CASES[name]
    gm, om = ca.Model(name, data), O.Model(name, data)
    rng = np.random.default_rng(hash(name) % 1000)
    q = rng.standard_normal((6, gm.dim)) * 0.8
    lp, g = gm.log_prob_grad(q, jacobian=jacobian)
    for i in range(len(q)):
        lp_ref, g_ref = om.log_prob_grad(q[i], jacobian)
        assert abs(lp[i] - lp_ref) <= 1e-12 * max(1.0, abs(lp_ref)) * 10, (name, lp[i], lp_ref)
        scale = _grad_scale(name, data, q[i], g_ref)
        assert np.all(np.abs(g[i] - g_ref) <= 1e-12 * scale), (name, np.max(np.abs(g[i] - g_ref) / scale))


REPLAY = [("gauss", "diag_e"), ("gauss", "unit_e"), ("bernoulli", "diag_e"), ("eight_schools_cp", "diag_e"),
          ("hier_normal_nc", "diag_e"), ("hier_logistic_cp", "diag_e"), ("logistic_regression", "diag_e"),
          ("gauss", "dense_e"), ("logistic_regression", "dense_e"), ("hier_normal_nc", "dense_e"),
          ("bernoulli", "dense_e")]


@pytest.mark.parametrize("name,metric", REPLAY)
def test_transition_replay(name, metric):
    """Same q0 and injected draws: same tree depth, n_leapfrog, divergence and
    selected draw as the oracle, over many chains and step sizes."""
    data = CASES[name]
    gm, om = ca.Model(name, data), O.Model(name, data)
    D = gm.dim
    rng = np.random.default_rng(17)
    nch = 24
    q0 = rng.standard_normal((nch, D)) * 0.5
    u = rng.random((nch, 4096))
    z = rng.standard_normal((nch, D))
    if metric == "dense_e":
        A = rng.standard_normal((D, D))
        minv = np.broadcast_to((A @ A.T / D + 0.5 * np.eye(D)).ravel(), (nch, D * D))
    else:
        minv = rng.uniform(0.3, 2.0, (nch, D)) if metric == "diag_e" else np.ones((nch, D))
    eps_grid = [0.02, 0.1, 0.3, 0.8] if name != "logistic_regression" else [0.01, 0.03, 0.08, 0.2]
    for eps in eps_grid:
        qg, infos = gm.transition_replay(q0, u, z, eps, metric=metric, inv_metric=minv[0], max_depth=8)
        for c in range(nch):
            qo, io = om.transition_replay(q0[c], u[c], z[c], eps, metric=O.METRICS[metric], minv=minv[0],
                                          max_depth=8)
            ig = infos[c]
            for k in ("depth", "n_leapfrog", "divergent", "uniforms_used", "normals_used"):
                assert ig[k] == io[k], (name, eps, c, k, ig[k], io[k])
            tol = 1e-9 if metric != "dense_e" else 1e-8
            assert np.allclose(qg[c], qo, rtol=tol, atol=1e-11), (name, eps, c)
            for k in ("accept_stat", "energy", "lp"):
                assert ig[k] == pytest.approx(io[k], rel=tol, abs=1e-11), (name, eps, c, k)


def test_replay_jitter_and_divergence():
    gm, om = ca.Model("gauss", CASES["gauss"]), O.Model("gauss", CASES["gauss"])
    rng = np.random.default_rng(3)
    q0 = rng.standard_normal((4, gm.dim))
    u = rng.random((4, 512))
    z = rng.standard_normal((4, gm.dim))
    qg, ig = gm.transition_replay(q0, u, z, 0.4, jitter=0.7)
    for c in range(4):
        qo, io = om.transition_replay(q0[c], u[c], z[c], 0.4, jitter=0.7)
        assert ig[c]["stepsize"] == pytest.approx(io["stepsize"], rel=1e-15)
        assert ig[c]["n_leapfrog"] == io["n_leapfrog"]
    qg, ig = gm.transition_replay(q0, u, z, 60.0)
    assert all(i["divergent"] == 1 and i["n_leapfrog"] >= 1 for i in ig)


def test_bernoulli_posterior():
    fit = ca.Model("bernoulli", BERN).sample(num_chains=4, seed=11)
    th = fit.param("theta")
    s = E.summary(th)
    assert abs(s["mean"] - 0.25) < 4 * s["mcse"]
    assert abs(s["sd"] - np.sqrt(27 / (144 * 13))) < 0.01
    assert s["rhat"] < 1.01
    for k in range(4):
        d = fit.draws[k]
        dep, n, div = d[:, 3], d[:, 4], d[:, 5]
        assert np.all((n >= 2 ** dep - 1) & (n <= 2 ** (dep + 1) - 1))
        assert np.all(d[:, 1] >= 0) and np.all(d[:, 1] <= 1)
    assert np.all(fit.status == 0)


def test_bernoulli_same_stream_as_oracle():
    """The GPU uses the same Philox streams as the oracle: with one seed, the
    warmup trajectory agrees until floating-point differences in the
    densities separate them (D=1: for many iterations)."""
    fit = ca.Model("bernoulli", BERN).sample(num_chains=2, seed=99, save_warmup=True, num_warmup=200,
                                               num_samples=50)
    r = O.Model("bernoulli", BERN).run(seed=99, num_chains=2, save_warmup=1, num_warmup=200, num_samples=50)
    for k in range(2):
        g, o = fit.draws[k], r["draws"][k]
        # the step-size heuristic and the first transitions coincide
        assert g[0, 2] == pytest.approx(o[0, 2], rel=1e-12)
        close = np.all(np.isclose(g[:, :8], o[:, :8], rtol=1e-8, atol=1e-10), axis=1)
        first_diff = np.argmin(close) if not close.all() else len(close)
        n = min(first_diff + 1, len(close))
        rel = np.max(np.abs(g[:n, :8] - o[:n, :8]) / (np.abs(o[:n, :8]) + 1e-12), axis=1)
        print(f"chain {k}: agree within 1e-8 for {first_diff} iterations; last rel diffs {rel[-6:]}")
        assert first_diff >= 25, (first_diff, g[min(first_diff, len(g) - 1)], o[min(first_diff, len(o) - 1)])


def test_logistic_lockstep_matches_oracle_posterior():
    data = logistic(seed=5, N=1500, D=6)
    gm = ca.Model("logistic_regression", data)
    fit = gm.sample(num_chains=64, seed=3, fused=0)
    assert fit.leaf_steps > 0
    # 16 oracle chains: its MCSE, not the GPU's, dominates the 4-MCSE band
    r = O.Model("logistic_regression", data).run(seed=3, num_chains=16)
    for j in range(6):
        name = f"beta.{j + 1}"
        xg = fit.param(name)
        xo = r["draws"][:, :, 7 + j].T
        sg, so = E.summary(xg), E.summary(xo)
        assert sg["rhat"] < 1.01
        assert abs(sg["mean"] - so["mean"]) < 4 * np.hypot(sg["mcse"], so["mcse"]), name
        assert abs(sg["sd"] - so["sd"]) < 0.1 * so["sd"], name


def test_lockstep_and_fused_agree_statistically():
    data = logistic(seed=8, N=800, D=4)
    gm = ca.Model("logistic_regression", data)
    a = gm.sample(num_chains=32, seed=4, fused=0)
    b = gm.sample(num_chains=32, seed=4, fused=1)
    for j in range(4):
        sa, sb = E.summary(a.param(f"beta.{j + 1}")), E.summary(b.param(f"beta.{j + 1}"))
        assert abs(sa["mean"] - sb["mean"]) < 4 * np.hypot(sa["mcse"], sb["mcse"])


def test_determinism():
    gm = ca.Model("logistic_regression", logistic(seed=2, N=500, D=3))
    a = gm.sample(num_chains=8, seed=5, num_warmup=150, num_samples=100, fused=0)
    b = gm.sample(num_chains=8, seed=5, num_warmup=150, num_samples=100, fused=0)
    for x, y in zip(a.draws, b.draws):
        assert np.array_equal(x, y)
    em = ca.Model("eight_schools_cp", EIGHT)
    a = em.sample(num_chains=4, seed=5, num_warmup=150, num_samples=100)
    b = em.sample(num_chains=4, seed=5, num_warmup=150, num_samples=100)
    for x, y in zip(a.draws, b.draws):
        assert np.array_equal(x, y)


def test_batched_uneven_hier_normal():
    datasets = [hier_normal(seed=s, J=J) for s, J in [(1, 8), (2, 30), (3, 75)]]
    gm = ca.Model("hier_normal_nc", datasets)
    fit = gm.sample(num_chains=8, seed=21, delta=0.9)
    for m, data in enumerate(datasets):
        r = O.Model("hier_normal_nc", data).run(seed=21, num_chains=4, delta=0.9)
        for name, col in [("mu", 7), ("tau", 8)]:
            xg = fit.param(name, dataset=m)
            xo = r["draws"][:, :, col].T
            sg, so = E.summary(xg), E.summary(xo)
            assert sg["rhat"] < 1.02, (m, name, sg["rhat"])
            assert abs(sg["mean"] - so["mean"]) < 4 * np.hypot(sg["mcse"], so["mcse"]), (m, name)


def test_metric_file_no_adapt_keeps_stepsize_and_metric():
    gm = ca.Model("gauss", dict(D=2, P=np.diag([0.01, 1.0])))
    fit = gm.sample(num_chains=2, adapt_engaged=False, num_warmup=10, num_samples=50, stepsize=0.5,
                    inv_metric=[1.5, 1.8])
    assert np.all(fit.stepsize == 0.5)
    assert all(np.array_equal(m, [1.5, 1.8]) for m in fit.inv_metric)
    assert np.all(fit.draws[0][:, 2] == 0.5)
    fit = gm.sample(num_chains=1, num_warmup=100, num_samples=20, inv_metric=[1.5, 1.8])
    assert np.array_equal(fit.inv_metric[0], [1.5, 1.8])  # 15/75/10 branch: metric kept


def test_adaptation_matches_oracle_schedule():
    """Adapted step size and metric of a gaussian agree with the oracle's in
    distribution (window ends at 100/150/250/450/950 on both)."""
    data = dict(D=3, P=np.diag([1.0, 0.25, 4.0]))
    fit = ca.Model("gauss", data).sample(num_chains=16, seed=8)
    r = O.Model("gauss", data).run(seed=8, num_chains=16)
    mg = np.array(fit.inv_metric)
    mo = r["inv_metric"]
    assert np.allclose(mg.mean(0), mo.mean(0), rtol=0.15)
    assert np.allclose(mg.mean(0), [1.0, 4.0, 0.25], rtol=0.2)
    assert abs(np.mean(fit.stepsize) - np.mean(r["stepsize"])) < 0.15 * np.mean(r["stepsize"])


def test_init_failure_reports_error():
    # theta = logit^-1 with all-ones data still initialises; a zero-radius init of
    # eight schools is valid; an init of inf is not
    gm = ca.Model("gauss", dict(D=2))
    with pytest.raises(ca.SamplerError, match="Initialization failed"):
        gm.sample(num_chains=2, init=np.array([np.inf, 0.0]), num_warmup=10, num_samples=10)


@pytest.mark.parametrize("N,D,C", [(100000, 100, 1100), (777, 1, 70), (1000, 17, 65), (3001, 128, 130),
                                   (64, 4, 3)])
def test_glm_gradient_full_shape(N, D, C):
    """The batched FP64 matrix-core gradient at the benchmark shape and at
    ragged edges, against the closed form X'(y - s(X b)) - b/s^2 in numpy FP64,
    with the error bounded relative to the sum of |terms| (conditioning)."""
    rng = np.random.default_rng(N + D)
    X = rng.standard_normal((N, D))
    beta = rng.normal(0, 1 / np.sqrt(D), D)
    y = (rng.random(N) < 1 / (1 + np.exp(-X @ beta))).astype(np.int64)
    gm = ca.Model("logistic_regression", dict(N=N, D=D, X=X, y=y))
    Q = beta + rng.normal(0, 0.05, (C, D))
    lp, g = gm.log_prob_grad(Q)
    E_ = X @ Q.T
    S = 1 / (1 + np.exp(-E_))
    R = y[:, None] - S
    g_ref = (X.T @ R).T - Q / 6.25
    lp_ref = (y[:, None] * E_ - np.logaddexp(0, E_)).sum(0) - 0.5 * (Q * Q).sum(1) / 6.25
    cond_g = (np.abs(X).T @ np.abs(R)).T + np.abs(Q) / 6.25
    cond_lp = np.abs(y[:, None] * E_).sum(0) + np.logaddexp(0, E_).sum(0)
    assert np.all(np.abs(g - g_ref) <= 1e-12 * cond_g + 1e-300)
    assert np.all(np.abs(lp - lp_ref) <= 1e-12 * cond_lp)


def test_dense_adaptation_recovers_covariance():
    """dense_e warmup on a correlated gaussian: the adapted inverse metric
    approaches the target covariance (regularised), like the oracle's."""
    Sig = np.array([[1.0, 0.8, 0.0], [0.8, 1.0, 0.3], [0.0, 0.3, 2.0]])
    data = dict(D=3, P=np.linalg.inv(Sig))
    try:
        fit = ca.Model("gauss", data).sample(num_chains=16, seed=12, metric="dense_e")
    except ca.SamplerError as e:
        raise AssertionError(f"dense run failed: {e}")
    r = O.Model("gauss", data).run(seed=12, num_chains=8, metric="dense_e")
    mg = np.mean(fit.inv_metric, axis=0)
    mo = r["inv_metric"].reshape(8, 3, 3).mean(0)
    assert np.allclose(mg, Sig, atol=0.15), mg
    assert np.allclose(mg, mo, atol=0.15)
    x = np.stack([fit.param(f"x.{i + 1}") for i in range(3)])
    for i in range(3):
        s = E.summary(x[i])
        assert abs(s["mean"]) < 4 * s["mcse"] and s["rhat"] < 1.01
    assert np.all(fit.status == 0)


def _bf16_round(x):
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("path", [0, 1, 2])
@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_glm_gradient_reduced_precision(precision, path, monkeypatch):
    """FP32 contraction (X in FP32, or stored BF16) with FP64 sums: error of lp
    and gradient against FP64 on the same stored data, relative to the sum of
    |terms|; tolerance 1e-5 (FP32 products, ~2^-24 per term).  Every kernel
    path: 0 single-pass FP32 matrix cores, 1 two-pass, 2 single-pass BF16
    matrix cores with three-term operands."""
    monkeypatch.setenv("CMDSTAN_AMD_GLM_PATH", str(path))
    N, D, C = 20000, 100, 130
    rng = np.random.default_rng(7)
    X = rng.standard_normal((N, D))
    beta = rng.normal(0, 1 / np.sqrt(D), D)
    y = (rng.random(N) < 1 / (1 + np.exp(-X @ beta))).astype(np.int64)
    gm = ca.Model("logistic_regression", dict(N=N, D=D, X=X, y=y))
    Q = beta + rng.normal(0, 0.05, (C, D))
    lp, g = gm.log_prob_grad(Q, precision=precision)
    Xs = X.astype(np.float32).astype(np.float64) if precision == "fp32" else _bf16_round(X)
    E_ = Xs @ Q.T
    R = y[:, None] - 1 / (1 + np.exp(-E_))
    g_ref = (Xs.T @ R).T - Q / 6.25
    lp_ref = (y[:, None] * E_ - np.logaddexp(0, E_)).sum(0) - 0.5 * (Q * Q).sum(1) / 6.25
    cond_g = (np.abs(Xs).T @ np.abs(R)).T + np.abs(Q) / 6.25
    cond_lp = np.abs(y[:, None] * E_).sum(0) + np.logaddexp(0, E_).sum(0)
    err_g = np.max(np.abs(g - g_ref) / cond_g)
    err_lp = np.max(np.abs(lp - lp_ref) / cond_lp)
    print(f"{precision} path {path}: max rel gradient error {err_g:.2e}, lp error {err_lp:.2e} (vs FP64 on stored X)")
    assert err_g < 1e-5 and err_lp < 1e-5
    if precision == "bf16":
        # the storage itself perturbs the data by up to 2^-9 relative
        lp64, g64 = gm.log_prob_grad(Q)
        print("bf16 vs FP64 on the original X: max rel lp change",
              np.max(np.abs(lp - lp64) / cond_lp))


@pytest.mark.parametrize("precision,path", [("fp32", 0), ("fp32", 2), ("bf16", 0), ("bf16", 2)])
def test_reduced_precision_posterior_matches_fp64(precision, path, monkeypatch):
    monkeypatch.setenv("CMDSTAN_AMD_GLM_PATH", str(path))
    data = logistic(seed=9, N=4000, D=8)
    gm = ca.Model("logistic_regression", data)
    a = gm.sample(num_chains=64, seed=6, precision=0)
    b = gm.sample(num_chains=64, seed=6, precision=precision)
    for j in range(8):
        sa, sb = E.summary(a.param(f"beta.{j + 1}")), E.summary(b.param(f"beta.{j + 1}"))
        assert sb["rhat"] < 1.01
        assert abs(sa["mean"] - sb["mean"]) < 4 * np.hypot(sa["mcse"], sb["mcse"])
        assert abs(sa["sd"] - sb["sd"]) < 0.05 * sa["sd"]


def test_state_machine_threads_per_chain_tuned(monkeypatch):
    """Lockstep runs time the wave-per-chain and the block-per-chain state
    machine on their first four check intervals (wave, block, block, wave)
    and report both; the other
    mode replaces the default by dimension (wave below D = 512) only when it is
    more than 20% faster.  CMDSTAN_AMD_BLOCK_PER_CHAIN pins the mode and turns
    the timing off; fused runs report 0 threads."""
    monkeypatch.delenv("CMDSTAN_AMD_BLOCK_PER_CHAIN", raising=False)
    gm = ca.Model("logistic_regression", logistic(seed=4, N=3000, D=20))
    kw = dict(num_chains=64, num_warmup=100, num_samples=50, seed=3, fused=0)
    fit = gm.sample(**kw)
    tw, tb = fit.tuned_launch_seconds
    print(f"state machine launch: wave {tw * 1e6:.1f} us, block {tb * 1e6:.1f} us -> {fit.tuned_threads_per_chain}")
    assert np.all(fit.status == 0) and tw > 0 and tb > 0
    assert fit.tuned_threads_per_chain == (256 if tb < 0.8 * tw else 64)
    monkeypatch.setenv("CMDSTAN_AMD_BLOCK_PER_CHAIN", "1")
    b = gm.sample(**kw)
    assert np.all(b.status == 0)
    assert b.tuned_threads_per_chain == 256 and b.tuned_launch_seconds == (0.0, 0.0)
    f = ca.Model("gauss", CASES["gauss"]).sample(num_chains=4, num_warmup=50, num_samples=20, seed=1, fused=1)
    assert f.tuned_threads_per_chain == 0


def test_dense_deterministic_and_valid_over_seeds():
    """dense_e adaptation over several seeds: every chain finishes, and a
    repeated run is bit-identical (no read of a stale vector across lanes)."""
    Sig = np.array([[1.0, 0.8, 0.0], [0.8, 1.0, 0.3], [0.0, 0.3, 2.0]])
    gm = ca.Model("gauss", dict(D=3, P=np.linalg.inv(Sig)))
    for seed in (1, 2, 3):
        a = gm.sample(num_chains=32, seed=seed, metric="dense_e", num_warmup=300, num_samples=50)
        b = gm.sample(num_chains=32, seed=seed, metric="dense_e", num_warmup=300, num_samples=50)
        assert np.all(a.status == 0)
        for x, y in zip(a.draws, b.draws):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("jacobian", [True, False])
def test_hier_logistic_batched_gradient_large_groups(jacobian):
    """The batched group kernel (lane = point, groups staged through LDS in
    passes of 128 observations) on uneven groups of up to ~900 observations and
    K = 10, against the oracle at 1e-12 of the conditioned scale."""
    from tests.model_cases import hier_logistic
    data = hier_logistic(seed=4, N=5000, K=10, G=9)
    gm, om = ca.Model("hier_logistic_cp", data), O.Model("hier_logistic_cp", data)
    rng = np.random.default_rng(8)
    q = rng.standard_normal((70, gm.dim)) * 0.6
    lp, g = gm.log_prob_grad(q, jacobian=jacobian)
    for i in range(len(q)):
        lp_ref, g_ref = om.log_prob_grad(q[i], jacobian)
        assert abs(lp[i] - lp_ref) <= 1e-11 * max(1.0, abs(lp_ref))
        assert np.all(np.abs(g[i] - g_ref) <= 1e-12 * np.maximum(1.0, np.abs(g_ref)) * 64)


def test_hier_logistic_lockstep_matches_fused_posterior():
    """Batched lockstep kernels and the fused per-chain kernel sample the same
    posterior (means within 4 combined MCSE)."""
    from tests.model_cases import hier_logistic
    data = hier_logistic(seed=2, N=600, K=3, G=5)
    m = ca.Model("hier_logistic_cp", data)
    a = m.sample(num_chains=32, num_warmup=400, num_samples=400, seed=5, fused=0)
    b = m.sample(num_chains=32, num_warmup=400, num_samples=400, seed=6, fused=1)
    assert np.all(a.status == 0) and np.all(b.status == 0)
    for name in ["mu.1", "tau.2", "beta.3.1"]:
        sa, sb = E.summary(a.param(name)), E.summary(b.param(name))
        assert abs(sa["mean"] - sb["mean"]) < 4 * np.hypot(sa["mcse"], sb["mcse"]), (name, sa, sb)


@pytest.mark.parametrize("metric", ["diag_e", "dense_e"])
def test_block_per_chain_state_machine_matches_wave_mode(metric, monkeypatch):
    """The 256-thread-per-chain state machine (used for large D) against the
    one-wave-per-chain one on the same seeds: a short run agrees to rounding,
    a long one has the same posterior."""
    from tests.model_cases import gauss
    m = ca.Model("gauss", gauss(seed=3, D=40))
    kw = dict(num_chains=16, seed=21, metric=metric, fused=0)
    monkeypatch.setenv("CMDSTAN_AMD_BLOCK_PER_CHAIN", "0")
    a = m.sample(num_warmup=20, num_samples=5, **kw)
    la = m.sample(num_warmup=300, num_samples=300, **kw)
    monkeypatch.setenv("CMDSTAN_AMD_BLOCK_PER_CHAIN", "1")
    b = m.sample(num_warmup=20, num_samples=5, **kw)
    lb = m.sample(num_warmup=300, num_samples=300, **dict(kw, seed=22))
    assert np.all(a.status == 0) and np.all(b.status == 0) and np.all(lb.status == 0)
    for x, y in zip(a.draws, b.draws):
        assert np.allclose(x, y, rtol=1e-8, atol=1e-8)
    for i in (1, 7, 40):
        sa, sb = E.summary(la.param(f"x.{i}")), E.summary(lb.param(f"x.{i}"))
        assert abs(sa["mean"] - sb["mean"]) < 4 * np.hypot(sa["mcse"], sb["mcse"])


@pytest.mark.parametrize("family", ["gauss", "hier_logistic_cp", "logistic_regression", "logistic_fp32"])
def test_graph_replay_matches_eager_launches(family, monkeypatch):
    """Lockstep bundles replayed from a captured HIP graph (many replays, with
    the host's checks, synchronisations and eager timed bundles between them)
    give bit-identical draws to the same kernels launched eagerly; the GLM
    families run through the batched matrix-core gradient."""
    from tests.model_cases import gauss, hier_logistic
    data = {"gauss": lambda: gauss(seed=5, D=12), "hier_logistic_cp": lambda: hier_logistic(seed=3, N=400, K=3, G=4),
            "logistic_regression": lambda: logistic(seed=4, N=3000, D=20),
            "logistic_fp32": lambda: logistic(seed=4, N=3000, D=20)}[family]()
    prec = "fp32" if family == "logistic_fp32" else "fp64"
    family = "logistic_regression" if family == "logistic_fp32" else family
    m = ca.Model(family, data)
    kw = dict(num_chains=24, num_warmup=150, num_samples=100, seed=9, fused=0)
    if family == "logistic_regression":
        kw["precision"] = prec
    monkeypatch.setenv("CMDSTAN_AMD_GRAPH", "0")
    a = m.sample(**kw)
    monkeypatch.setenv("CMDSTAN_AMD_GRAPH", "1")
    b = m.sample(**kw)
    assert np.all(a.status == 0) and np.all(b.status == 0)
    for x, y in zip(a.draws, b.draws):
        assert np.array_equal(x, y)
    if family == "logistic_regression":
        assert b.grad_launches_timed > 0 and b.grad_seconds_per_launch > 0


def test_interrupt_callback_stops_the_run():
    """callbacks::interrupt: polled between check intervals; a True return
    stops the run with an error, False lets it finish."""
    from tests.model_cases import gauss
    m = ca.Model("gauss", gauss(seed=1, D=4))
    calls = []

    def stop():
        calls.append(1)
        return len(calls) >= 3

    with pytest.raises(ca.SamplerError, match="interrupted"):
        m.sample(num_chains=8, num_warmup=2000, num_samples=2000, fused=0, interrupt=stop)
    assert len(calls) == 3
    fit = m.sample(num_chains=8, num_warmup=100, num_samples=100, fused=0, interrupt=lambda: False)
    assert np.all(fit.status == 0)


@pytest.mark.parametrize("persistent", ["0", "1"])
def test_graph_replay_matches_eager_with_flow(monkeypatch, persistent):
    """Graph replay of the lockstep bundles with a flow, on the GEMM chain and
    on the persistent flow kernels: draws identical to eager launches."""
    monkeypatch.setenv("CMDSTAN_AMD_FLOW_PERSISTENT", persistent)
    from tests.model_cases import gauss
    m = ca.Model("gauss", gauss(seed=6, D=6))
    kw = dict(num_chains=32, num_warmup=150, num_samples=60, seed=2, flow="force", flow_layers=2, flow_hidden=16,
              flow_train_steps=100, flow_warmup=60, flow_pilot=20)
    monkeypatch.setenv("CMDSTAN_AMD_GRAPH", "0")
    a = m.sample(**kw)
    monkeypatch.setenv("CMDSTAN_AMD_GRAPH", "1")
    b = m.sample(**kw)
    assert a.flow_used == b.flow_used
    for x, y in zip(a.draws, b.draws):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_glm_large_gradient_reduced_precision(precision):
    """Large-D path (two FP32 matrix-core passes, X in FP32 or stored BF16,
    FP64 sums of the row-range partials): error against FP64 on the same stored
    data relative to the sum of |terms|, tolerance 1e-5; ragged N, D, C."""
    N, D, C = 6001, 301, 70
    rng = np.random.default_rng(17)
    X = rng.standard_normal((N, D))
    beta = rng.normal(0, 1 / np.sqrt(D), D)
    y = (rng.random(N) < 1 / (1 + np.exp(-X @ beta))).astype(np.int64)
    gm = ca.Model("logistic_regression", dict(N=N, D=D, X=X, y=y))
    Q = beta + rng.normal(0, 0.05, (C, D))
    lp, g = gm.log_prob_grad(Q, precision=precision)
    Xs = X.astype(np.float32).astype(np.float64) if precision == "fp32" else _bf16_round(X)
    E_ = Xs @ Q.T
    R = y[:, None] - 1 / (1 + np.exp(-E_))
    g_ref = (Xs.T @ R).T - Q / 6.25
    lp_ref = (y[:, None] * E_ - np.logaddexp(0, E_)).sum(0) - 0.5 * (Q * Q).sum(1) / 6.25
    cond_g = (np.abs(Xs).T @ np.abs(R)).T + np.abs(Q) / 6.25
    cond_lp = np.abs(y[:, None] * E_).sum(0) + np.logaddexp(0, E_).sum(0)
    err_g = np.max(np.abs(g - g_ref) / cond_g)
    err_lp = np.max(np.abs(lp - lp_ref) / cond_lp)
    print(f"{precision}: max rel gradient error {err_g:.2e}, lp error {err_lp:.2e}")
    assert err_g < 1e-5 and err_lp < 1e-5


@pytest.mark.parametrize("variant", ["0", "1"])
@pytest.mark.parametrize("N,D,C", [(6001, 304, 70), (1000, 1000, 130), (3000, 8, 300), (2500, 96, 257)])
def test_glm_large_bf16_storage_both_kernels(monkeypatch, variant, N, D, C):
    """X stored in BF16 through the two-pass gradient on FP32 matrix cores
    (variant 0) and on BF16 matrix cores with positions and residuals as three
    BF16 terms and the transposed X (variant 1); D <= 128 forced onto the
    two-pass path.  Error against FP64 on the stored X relative to the sum of
    |terms|, tolerance 1e-5; ragged N and C, several chain and coefficient
    tiles, the smallest D the BF16 kernels take (8)."""
    monkeypatch.setenv("CMDSTAN_AMD_GL_BF16", variant)
    if D <= 128:
        monkeypatch.setenv("CMDSTAN_AMD_GLM_PATH", "1")
    rng = np.random.default_rng(N + D + C)
    X = rng.standard_normal((N, D))
    beta = rng.normal(0, 1 / np.sqrt(D), D)
    y = (rng.random(N) < 1 / (1 + np.exp(-X @ beta))).astype(np.int64)
    gm = ca.Model("logistic_regression", dict(N=N, D=D, X=X, y=y))
    Q = beta + rng.normal(0, 0.05, (C, D))
    print(f"bf16 storage N={N} D={D} C={C} variant={variant}", flush=True)
    lp, g = gm.log_prob_grad(Q, precision="bf16")
    Xs = _bf16_round(X)
    E_ = Xs @ Q.T
    R = y[:, None] - 1 / (1 + np.exp(-E_))
    g_ref = (Xs.T @ R).T - Q / 6.25
    lp_ref = (y[:, None] * E_ - np.logaddexp(0, E_)).sum(0) - 0.5 * (Q * Q).sum(1) / 6.25
    cond_g = (np.abs(Xs).T @ np.abs(R)).T + np.abs(Q) / 6.25
    cond_lp = np.abs(y[:, None] * E_).sum(0) + np.logaddexp(0, E_).sum(0)
    err_g = np.max(np.abs(g - g_ref) / cond_g)
    err_lp = np.max(np.abs(lp - lp_ref) / cond_lp)
    print(f"  max rel gradient error {err_g:.2e}, lp error {err_lp:.2e}")
    assert err_g < 1e-5 and err_lp < 1e-5


def test_glm_large_bf16_kernels_sample_the_fp64_posterior():
    """Sampling through the BF16 matrix-core large-D gradient (X exactly
    representable in BF16, so FP64 and BF16 storage hold the same data) agrees
    with the FP64 per-chain path: means within 4 combined MCSE."""
    N, D = 2000, 136
    rng = np.random.default_rng(8)
    X = _bf16_round(rng.standard_normal((N, D)))
    beta = rng.normal(0, 1 / np.sqrt(D), D)
    y = (rng.random(N) < 1 / (1 + np.exp(-X @ beta))).astype(np.int64)
    m = ca.Model("logistic_regression", dict(N=N, D=D, X=X, y=y))
    a = m.sample(num_chains=32, num_warmup=250, num_samples=150, seed=4, precision="bf16")
    b = m.sample(num_chains=32, num_warmup=250, num_samples=150, seed=5)
    assert np.all(a.status == 0) and np.all(b.status == 0)
    for j in (1, 60, 136):
        sa, sb = E.summary(a.param(f"beta.{j}")), E.summary(b.param(f"beta.{j}"))
        assert abs(sa["mean"] - sb["mean"]) < 4 * np.hypot(sa["mcse"], sb["mcse"]), (j, sa, sb)


def test_glm_large_dense_posterior_matches_fp64():
    """dense_e sampling through the large-D FP32 path agrees with the FP64
    per-chain path (means within 4 combined MCSE)."""
    N, D = 2000, 140
    rng = np.random.default_rng(3)
    X = rng.standard_normal((N, D))
    beta = rng.normal(0, 1 / np.sqrt(D), D)
    y = (rng.random(N) < 1 / (1 + np.exp(-X @ beta))).astype(np.int64)
    m = ca.Model("logistic_regression", dict(N=N, D=D, X=X, y=y))
    a = m.sample(num_chains=32, num_warmup=250, num_samples=150, seed=4, metric="dense_e", precision="fp32")
    b = m.sample(num_chains=32, num_warmup=250, num_samples=150, seed=5, metric="dense_e")
    assert np.all(a.status == 0) and np.all(b.status == 0)
    for j in (1, 50, 140):
        sa, sb = E.summary(a.param(f"beta.{j}")), E.summary(b.param(f"beta.{j}"))
        assert abs(sa["mean"] - sb["mean"]) < 4 * np.hypot(sa["mcse"], sb["mcse"]), (j, sa, sb)


def test_partial_inits_are_used_and_completed():
    """An init giving some parameters: those start where given, the others are
    drawn; the first evaluated state is the given one (replay from it)."""
    data = dict(J=3, y=[1.0, 2.0, 3.0], sigma=[1.0, 1.0, 2.0])
    m = ca.Model("eight_schools_cp", data)
    u = m.transform_inits({"mu": 0.5, "tau": 2.0, "theta": [0.1, 0.2, 0.3]})
    fit = m.sample(num_chains=4, num_warmup=0, num_samples=1, adapt_engaged=False, stepsize=1e-9, init=u, seed=3)
    x = np.stack(fit.draws)[:, 0, 7:]
    assert np.allclose(x[:, 0], 0.5, atol=1e-6) and np.allclose(x[:, 1], 2.0, atol=1e-6)
    part = m.transform_inits({"mu": 0.5})
    fit2 = m.sample(num_chains=8, num_warmup=20, num_samples=10, init=part, seed=4)
    assert np.all(fit2.status == 0)


def test_output_rows_thin_save_warmup_and_zero_samples():
    """Row counts follow generate_transitions: ceil(warmup/thin) saved warmup
    rows when save_warmup, ceil(samples/thin) sampling rows; zero draws is a
    valid request."""
    m = ca.Model("bernoulli", BERN)
    fit = m.sample(num_chains=3, num_warmup=100, num_samples=50, thin=3, save_warmup=True, seed=2)
    assert all(d.shape[0] == 34 + 17 for d in fit.draws)
    fit0 = m.sample(num_chains=2, num_warmup=50, num_samples=0, seed=2)
    assert all(d.shape[0] == 0 for d in fit0.draws) and np.all(fit0.status == 0)


def test_logistic_without_observations_is_the_prior():
    """N = 0 rows: the batched GLM path returns the prior, beta ~ N(0, 2.5^2)."""
    D = 3
    m = ca.Model("logistic_regression", dict(N=0, D=D, X=np.zeros((0, D)), y=np.zeros(0, dtype=np.int64)))
    fit = m.sample(num_chains=64, num_warmup=300, num_samples=300, seed=5)
    assert np.all(fit.status == 0)
    for j in range(1, D + 1):
        s = E.summary(fit.param(f"beta.{j}"))
        assert abs(s["mean"]) < 4 * s["mcse"] and abs(s["sd"] - 2.5) < 0.15


FP32_REPLAY = [("hier_normal_nc", "diag_e"), ("eight_schools_cp", "diag_e"), ("gauss", "unit_e"),
               ("bernoulli", "diag_e"), ("hier_logistic_cp", "diag_e")]


@pytest.mark.parametrize("name,metric", FP32_REPLAY)
def test_fp32_leapfrog_transition_tracks_fp64_oracle(name, metric):
    """FP32 leapfrog (q, p, g, checkpoints and metric stored in FP32; FP64
    arithmetic, H, tree weights and U-turn sums): from the same FP32-exact start
    and metric with the same injected draws, the transition builds the FP64
    oracle's tree in nearly every chain, and there the selected draw agrees to
    1e-4 of the coordinate scale and the energy to 1e-5 relative.  A wrong
    term in the FP32 passes moves every trajectory by far more than that."""
    data = CASES[name]
    gm, om = ca.Model(name, data), O.Model(name, data)
    D = gm.dim
    rng = np.random.default_rng(29)
    nch = 32
    q0 = (rng.standard_normal((nch, D)) * 0.5).astype(np.float32).astype(np.float64)
    u = rng.random((nch, 4096))
    z = rng.standard_normal((nch, D))
    minv = rng.uniform(0.3, 2.0, D) if metric == "diag_e" else np.ones(D)
    minv = minv.astype(np.float32).astype(np.float64)
    agree = total = 0
    worst = 0.0
    for eps in [0.02, 0.1, 0.3]:
        qg, ig = gm.transition_replay(q0, u, z, eps, metric=metric, inv_metric=minv, max_depth=8, precision="fp32")
        for c in range(nch):
            qo, io = om.transition_replay(q0[c], u[c], z[c], eps, metric=O.METRICS[metric], minv=minv, max_depth=8)
            total += 1
            if any(ig[c][k] != io[k] for k in ("depth", "n_leapfrog", "divergent", "uniforms_used")):
                continue
            scale = max(1.0, float(np.max(np.abs(qo))))
            err = float(np.max(np.abs(qg[c] - qo))) / scale
            if err > 1e-4:
                continue  # a knife-edge multinomial choice picked another leaf
            agree += 1
            worst = max(worst, err)
            assert ig[c]["energy"] == pytest.approx(io["energy"], rel=1e-5, abs=1e-5), (name, eps, c)
    assert agree >= 0.9 * total, (name, agree, total)
    assert worst > 0  # the FP32 path really rounds


def test_fp32_leapfrog_batched_hier_normal_matches_oracle_posterior():
    """Batched uneven hierarchical normal models with the FP32 leapfrog: means
    within 4 combined MCSE of the FP64 oracle's, R-hat < 1.02."""
    datasets = [hier_normal(seed=s, J=J) for s, J in [(1, 8), (2, 30), (3, 75)]]
    gm = ca.Model("hier_normal_nc", datasets)
    fit = gm.sample(num_chains=8, seed=23, delta=0.9, precision="fp32")
    assert np.all(fit.status == 0)
    for m, data in enumerate(datasets):
        r = O.Model("hier_normal_nc", data).run(seed=21, num_chains=4, delta=0.9)
        for name, col in [("mu", 7), ("tau", 8)]:
            xg = fit.param(name, dataset=m)
            xo = r["draws"][:, :, col].T
            sg, so = E.summary(xg), E.summary(xo)
            assert sg["rhat"] < 1.02, (m, name, sg["rhat"])
            assert abs(sg["mean"] - so["mean"]) < 4 * np.hypot(sg["mcse"], so["mcse"]), (m, name)


def test_fp32_leapfrog_rejects_dense_and_bf16():
    gm = ca.Model("gauss", CASES["gauss"])
    with pytest.raises(ca.SamplerError):
        gm.sample(num_chains=2, num_warmup=10, num_samples=10, metric="dense_e", precision="fp32", fused=1)
    with pytest.raises(ca.SamplerError):
        gm.sample(num_chains=2, num_warmup=10, num_samples=10, precision="bf16", fused=1)


@pytest.mark.parametrize("metric,precision", [("diag_e", "fp64"), ("unit_e", "fp64"), ("dense_e", "fp64"),
                                              ("diag_e", "fp32")])
def test_diagnostic_rows_hold_q_p_g_of_the_draw(metric, precision):
    """Diagnostic rows (mcmc_writer::write_diagnostic_params): q is the draw's
    unconstrained value (tau = exp(q_tau) in the sample row), g the potential
    gradient at q (log_prob_grad, Jacobian included), lp__ the log density
    there, and p the momentum the energy column was computed with:
    energy__ = -lp__ + 1/2 p' Minv p, also for draws that stayed at z0 with the
    transition's fresh momentum."""
    gm = ca.Model("eight_schools_cp", EIGHT)
    D = gm.dim
    fit = gm.sample(num_chains=4, num_warmup=200, num_samples=100, seed=3, metric=metric, precision=precision,
                    fused=1, diagnostics=True)
    assert np.all(fit.status == 0)
    for k in range(4):
        dr, rows = fit.diagnostics[k], fit.draws[k]
        assert dr.shape == (100, 7 + 3 * D)
        assert np.array_equal(dr[:, :7], rows[:, :7])
        q, p, g = dr[:, 7:7 + D], dr[:, 7 + D:7 + 2 * D], dr[:, 7 + 2 * D:]
        j_tau = fit.column_names[0].index("tau")
        assert np.allclose(np.exp(q[:, 1]), rows[:, j_tau], rtol=1e-12)
        lp, grad = gm.log_prob_grad(q)
        assert np.allclose(lp, dr[:, 0], rtol=1e-12, atol=1e-10)
        assert np.allclose(-grad, g, rtol=1e-6 if precision == "fp32" else 1e-10, atol=1e-6)
        M = fit.inv_metric[k]
        kin = 0.5 * (np.einsum("ri,ij,rj->r", p, M, p) if metric == "dense_e" else
                     np.sum(p * p * (M if metric == "diag_e" else 1.0), axis=1))
        assert np.allclose(dr[:, 6], -dr[:, 0] + kin, rtol=1e-10, atol=1e-9)
    assert np.all(np.isfinite(fit.diagnostics[0]))


def test_diagnostic_rows_rejected_with_flow():
    gm = ca.Model("gauss", CASES["gauss"])
    with pytest.raises(ca.SamplerError):
        gm.sample(num_chains=2, num_warmup=20, num_samples=10, flow="force", diagnostics=True)


def test_fused_block_sampling_matches_wave(monkeypatch):
    """The non-centred hierarchical model sampled by the fused kernel with one
    block per chain (what a run takes once at most two chains per CU remain,
    here from the start) and with one wave per chain
    (CMDSTAN_AMD_FUSED_BLOCK=0): every chain finishes and the posterior means
    of mu, tau and two thetas agree within 4 combined MCSE."""
    from tests.model_cases import hier_normal
    m = ca.Model("hier_normal_nc", hier_normal(seed=8, J=300))
    kw = dict(num_chains=64, num_warmup=300, num_samples=300, fused=1)
    monkeypatch.delenv("CMDSTAN_AMD_FUSED_BLOCK", raising=False)
    a = m.sample(seed=11, **kw)
    monkeypatch.setenv("CMDSTAN_AMD_FUSED_BLOCK", "0")
    b = m.sample(seed=12, **kw)
    assert np.all(a.status == 0) and np.all(b.status == 0)
    for name in ("mu", "tau", "theta.1", "theta.150"):
        sa, sb = E.summary(a.param(name)), E.summary(b.param(name))
        assert abs(sa["mean"] - sb["mean"]) < 4 * np.hypot(sa["mcse"], sb["mcse"]), (name, sa, sb)
