"""The benchmark configurations at their full shapes on the GPU, checked against
the oracle or closed forms: the 512-model uneven batch (C3), the centered
hierarchical logistic model with its flow (C4, D = 4020) and the
1,000,000 x 1000 logistic regression with X stored in FP32 and BF16 (C5)."""
import numpy as np
import pytest

import cmdstan_amd as ca
from oracle import flow as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _c3_datasets(n_models=512, seed=20240601):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_models):
        J = int(rng.integers(8, 2001))
        sigma = rng.uniform(5, 20, J)
        mu = rng.normal(0, 5)
        tau = abs(rng.standard_cauchy()) * 5
        out.append(dict(J=J, y=rng.normal(mu + tau * rng.standard_normal(J), sigma), sigma=sigma))
    return out


def test_c3_full_batch_runs_and_keeps_tree_invariants():
    """512 non-centered hierarchical normal models, J ~ U{8..2000}, 64 chains
    each (32,768 chains) in one batch: every chain finishes and every draw
    satisfies 2^d - 1 <= n_leapfrog <= 2^(d+1) - 1."""
    ds = _c3_datasets()
    m = ca.Model("hier_normal_nc", ds)
    print("c3: 512 models x 64 chains, 60 warmup + 20 draws", flush=True)
    fit = m.sample(num_chains=64, num_warmup=60, num_samples=20, seed=3, max_saved_params=4)
    assert np.all(fit.status == 0)
    arr = np.stack(fit.draws)
    assert arr.shape == (512 * 64, 20, 11)
    d, n = arr[:, :, 3], arr[:, :, 4]
    assert np.all((n >= 2 ** d - 1) & (n <= 2 ** (d + 1) - 1))
    assert np.all((arr[:, :, 1] >= 0) & (arr[:, :, 1] <= 1))
    assert np.all(np.isfinite(arr[:, :, 7:]))


@pytest.mark.parametrize("which", ["smallest", "largest"])
def test_c3_log_prob_grad_parity_extreme_sizes(which):
    ds = _c3_datasets()
    Js = np.array([d["J"] for d in ds])
    k = int(np.argmin(Js) if which == "smallest" else np.argmax(Js))
    m = ca.Model("hier_normal_nc", ds)
    om = O.Model("hier_normal_nc", ds[k])
    rng = np.random.default_rng(k)
    q = rng.standard_normal((16, m.dims[k])) * 0.7
    lp, g = m.log_prob_grad(q, dataset=k)
    for i in range(len(q)):
        lp_ref, g_ref = om.log_prob_grad(q[i], True)
        assert abs(lp[i] - lp_ref) <= 1e-12 * max(1.0, abs(lp_ref)) * 10
        assert np.all(np.abs(g[i] - g_ref) <= 1e-12 * np.maximum(1.0, np.abs(g_ref)) * 8)


def _c4_data(seed=20240601, N=50000, G=400, K=10):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((N, K)).astype(np.float32).astype(np.float64)
    grp = np.repeat(np.arange(G), N // G)
    mu = rng.normal(0, 1, K)
    tau = np.abs(rng.normal(0, 1, K))
    beta = mu + tau * rng.standard_normal((G, K))
    y = (rng.random(N) < 1 / (1 + np.exp(-np.einsum("nk,nk->n", X, beta[grp])))).astype(np.int64)
    return dict(N=N, K=K, G=G, X=X, y=y, group=grp + 1)


def test_c4_batched_gradient_256_points():
    """The batched group kernel at N = 50000, G = 400, K = 10 (D = 4020) for
    256 points against the oracle at 1e-12 of the conditioned scale."""
    data = _c4_data()
    gm, om = ca.Model("hier_logistic_cp", data), O.Model("hier_logistic_cp", data)
    rng = np.random.default_rng(1)
    q = rng.standard_normal((256, gm.dim)) * 0.5
    lp, g = gm.log_prob_grad(q)
    worst = 0.0
    for i in range(256):
        lp_ref, g_ref = om.log_prob_grad(q[i], True)
        assert abs(lp[i] - lp_ref) <= 1e-11 * max(1.0, abs(lp_ref))
        err = np.max(np.abs(g[i] - g_ref) / (np.maximum(1.0, np.abs(g_ref)) * 64))
        worst = max(worst, err)
    print(f"c4 gradient: worst conditioned error {worst * 64:.2e} (bound 64e-12)")
    assert worst <= 1e-12


_FLOW_KERNELS = {"gemm": ("0", "0"), "layers": ("1", "0"), "persistent": ("0", "1")}


def _flow_kernels(monkeypatch, kernels):
    fused, persistent = _FLOW_KERNELS[kernels]
    monkeypatch.setenv("CMDSTAN_AMD_FLOW_FUSED", fused)
    monkeypatch.setenv("CMDSTAN_AMD_FLOW_PERSISTENT", persistent)


@pytest.mark.parametrize("kernels", ["gemm", "layers", "persistent"])
def test_c4_flow_forward_and_pullback_d4020(monkeypatch, kernels):
    """The C4 flow (8 affine-coupling layers, MLP 2 x 128, masks with the 20
    hyperparameters as globals and groups of 10) at D = 4020, 256 rows, against
    the FP64 oracle at the FP32 tolerance 2e-4, on the split-K GEMM chain, on
    the fused per-layer kernels (CMDSTAN_AMD_FLOW_FUSED=1) and on the two
    persistent kernels (CMDSTAN_AMD_FLOW_PERSISTENT=1)."""
    _flow_kernels(monkeypatch, kernels)
    D, L, H, ng, gs = 4020, 8, 128, 20, 10
    g = ca.Flow(D, n_layers=L, hidden=H, n_glob=ng, group_size=gs, max_rows=256, seed=3)
    rng = np.random.default_rng(3)
    p = g.get_params().astype(np.float64)
    p = (p + 0.05 * rng.standard_normal(p.size)).astype(np.float32)
    g.set_params(p)
    loc, scale = rng.standard_normal(D) * 0.3, np.exp(0.2 * rng.standard_normal(D))
    g.set_standardization(loc, scale)
    o = F.Flow(D, F.coupling_masks(D, L, ng, gs), H, p.astype(np.float64), loc, scale)
    z = rng.standard_normal((256, D))
    x, ld = g.forward(z)
    xo, ldo = o.forward(z)
    assert np.max(np.abs(x - xo)) <= 2e-4 * max(1.0, np.max(np.abs(xo)))
    assert np.max(np.abs(ld - ldo)) <= 2e-4 * max(1.0, np.max(np.abs(ldo)))
    gx = rng.standard_normal((256, D))
    gz, gzo = g.pullback_grad(z, gx), o.pullback_grad(z, gx)
    assert np.max(np.abs(gz - gzo)) <= 2e-4 * max(1.0, np.max(np.abs(gzo)))


@pytest.mark.parametrize("rows", [256, 100, 1])
def test_c4_flow_persistent_equals_layer_kernels(monkeypatch, rows):
    """The persistent forward / pullback kernels run the per-layer kernels'
    tiles with the same summation order: bit-identical results at the C4
    shape, for full, ragged and single-row batches."""
    D, L, H, ng, gs = 4020, 8, 128, 20, 10
    rng = np.random.default_rng(5)
    z = rng.standard_normal((rows, D))
    gx = rng.standard_normal((rows, D))
    out = {}
    for kernels in ("layers", "persistent"):
        _flow_kernels(monkeypatch, kernels)
        g = ca.Flow(D, n_layers=L, hidden=H, n_glob=ng, group_size=gs, max_rows=256, seed=3)
        p = g.get_params().astype(np.float64)
        g.set_params((p + 0.05 * np.random.default_rng(6).standard_normal(p.size)).astype(np.float32))
        x, ld = g.forward(z)
        gz = g.pullback_grad(z, gx)
        out[kernels] = (x, ld, gz)
    for a, b in zip(out["layers"], out["persistent"]):
        assert np.array_equal(a, b)


def _bf16_round(x):
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def test_c5_large_gradient_fp32_and_bf16():
    """Logistic regression N = 1,000,000, D = 1000 (X held in FP32 and BF16),
    512 points through the two-pass FP32 matrix-core gradient; 8 of them
    against the closed form on the stored X in FP64, relative to the sum of
    |terms| (tolerance 1e-5: FP32 products)."""
    N, D, C = 1_000_000, 1000, 512
    rng = np.random.default_rng(11)
    print("c5: generating X", flush=True)
    X32 = rng.standard_normal((N, D), dtype=np.float32)
    beta = rng.normal(0, np.sqrt(1.0 / D), D)
    y = (rng.random(N) < 1 / (1 + np.exp(-(X32 @ beta.astype(np.float32)).astype(np.float64)))).astype(np.int64)
    m = ca.Model("logistic_regression", dict(N=N, D=D, X=X32.astype(np.float64), y=y))
    Q = beta + rng.normal(0, 0.01, (C, D))
    check = np.arange(0, C, C // 8)
    for precision in ("fp32", "bf16"):
        print(f"c5: gradient {precision}", flush=True)
        lp, g = m.log_prob_grad(Q, precision=precision)
        assert np.all(np.isfinite(lp)) and np.all(np.isfinite(g))
        Xs = X32 if precision == "fp32" else _bf16_round(X32).astype(np.float32)
        Qc = Q[check]
        E_ = Xs.astype(np.float64) @ Qc.T if precision == "bf16" else X32.astype(np.float64) @ Qc.T
        R = y[:, None] - 1 / (1 + np.exp(-E_))
        Xd = Xs.astype(np.float64)
        g_ref = (Xd.T @ R).T - Qc / 6.25
        lp_ref = (y[:, None] * E_ - np.logaddexp(0, E_)).sum(0) - 0.5 * (Qc * Qc).sum(1) / 6.25
        cond_g = (np.abs(Xd).T @ np.abs(R)).T + np.abs(Qc) / 6.25
        cond_lp = np.abs(y[:, None] * E_).sum(0) + np.logaddexp(0, E_).sum(0)
        err_g = np.max(np.abs(g[check] - g_ref) / cond_g)
        err_lp = np.max(np.abs(lp[check] - lp_ref) / cond_lp)
        print(f"c5 {precision}: max rel gradient error {err_g:.2e}, lp error {err_lp:.2e}", flush=True)
        assert err_g < 1e-5 and err_lp < 1e-5
        del Xd, Xs
