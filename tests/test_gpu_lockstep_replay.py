"""Single-transition replay of the LOCKSTEP drivers against the FP64 oracle:
the state machine in wave-per-chain and block-per-chain mode, with the
gradient computed by the batched kernels the samplers use at scale (the FP64
matrix-core GEMM for logistic regression, the group kernel for the
hierarchical logistic model, the per-chain gradient kernel otherwise).  Same
q0, same injected uniforms and normals: same tree depth, leapfrog count,
divergence and selected draw."""
import numpy as np
import pytest

import cmdstan_amd as ca
from oracle import oracle as O
from tests.model_cases import gauss, hier_logistic, hier_normal, logistic

pytestmark = pytest.mark.gpu


def _replay_and_compare(name, data, driver, metric, eps_grid, nch, max_depth, q_scale=0.5, q_center=None, seed=17,
                        q_tol=1e-9):
    gm, om = ca.Model(name, data), O.Model(name, data)
    D = gm.dim
    rng = np.random.default_rng(seed)
    center = np.zeros(D) if q_center is None else q_center
    q0 = center + rng.standard_normal((nch, D)) * q_scale
    u = rng.random((nch, 4096))
    z = rng.standard_normal((nch, D))
    if metric == "dense_e":
        A = rng.standard_normal((D, D)) / np.sqrt(D)
        minv = A @ A.T * 0.3 + np.eye(D)
        minv_flat = minv.ravel()
    elif metric == "diag_e":
        minv_flat = rng.uniform(0.3, 2.0, D)
    else:
        minv_flat = np.ones(D)
    depths = []
    for eps in eps_grid:
        print(f"replay {name} driver={driver} metric={metric} D={D} eps={eps}", flush=True)
        qg, infos = gm.transition_replay(q0, u, z, eps, metric=metric, inv_metric=minv_flat, max_depth=max_depth,
                                         driver=driver)
        for c in range(nch):
            qo, io = om.transition_replay(q0[c], u[c], z[c], eps, metric=O.METRICS[metric], minv=minv_flat,
                                          max_depth=max_depth)
            ig = infos[c]
            for k in ("depth", "n_leapfrog", "divergent", "uniforms_used", "normals_used"):
                assert ig[k] == io[k], (name, driver, eps, c, k, ig[k], io[k])
            assert np.allclose(qg[c], qo, rtol=q_tol, atol=q_tol * 1e-2), (name, driver, eps, c,
                                                                           np.max(np.abs(qg[c] - qo)))
            for k in ("accept_stat", "energy", "lp"):
                assert ig[k] == pytest.approx(io[k], rel=q_tol, abs=q_tol * 1e-2), (name, eps, c, k)
            depths.append(int(io["depth"]))
    # the grid exercises trees of several depths
    assert len(set(depths)) >= 2, depths


@pytest.mark.parametrize("driver", ["lockstep", "lockstep_block"])
def test_lockstep_replay_logistic_d100_glm_mfma(driver):
    """Logistic regression D = 100: the lockstep gradient is the FP64
    matrix-core GEMM (glm_mfma_f64), the headline path."""
    data = logistic(seed=21, N=3000, D=100)
    beta = np.linalg.lstsq(data["X"], (data["y"] - 0.5) * 4.0, rcond=None)[0]
    _replay_and_compare("logistic_regression", data, driver, "diag_e", [0.005, 0.02, 0.05], nch=12, max_depth=8,
                        q_scale=0.05, q_center=beta)


@pytest.mark.parametrize("name", ["gauss", "hier_normal_nc", "eight_schools_cp", "bernoulli"])
@pytest.mark.parametrize("driver", ["lockstep", "lockstep_block"])
def test_lockstep_replay_per_chain_gradient(name, driver):
    data = {"gauss": gauss(seed=2, D=9), "hier_normal_nc": hier_normal(seed=1, J=30),
            "eight_schools_cp": dict(J=8, y=[28.0, 8, -3, 7, -1, 1, 18, 12],
                                     sigma=[15.0, 10, 16, 11, 9, 11, 10, 18], mu_scale=10.0, tau_scale=10.0),
            "bernoulli": dict(N=10, y=[0, 1, 0, 0, 0, 0, 0, 0, 0, 1])}[name]
    _replay_and_compare(name, data, driver, "diag_e", [0.05, 0.3, 0.9], nch=16, max_depth=8)


@pytest.mark.parametrize("metric", ["dense_e", "unit_e"])
def test_lockstep_replay_gauss_metrics(metric):
    _replay_and_compare("gauss", gauss(seed=4, D=12), "lockstep", metric, [0.05, 0.3, 0.9], nch=16, max_depth=8,
                        q_tol=1e-8)


def test_lockstep_block_replay_hier_logistic_d4020():
    """The C4 model at its full shape (N = 50000, 400 groups x 10 covariates,
    D = 4020), block-per-chain state machine, batched group kernel."""
    rng = np.random.default_rng(5)
    N, G, K = 50000, 400, 10
    X = rng.standard_normal((N, K)).astype(np.float32).astype(np.float64)
    grp = np.repeat(np.arange(G), N // G)
    mu = rng.normal(0, 1, K)
    tau = np.abs(rng.normal(0, 1, K))
    beta = mu + tau * rng.standard_normal((G, K))
    y = (rng.random(N) < 1 / (1 + np.exp(-np.einsum("nk,nk->n", X, beta[grp])))).astype(np.int64)
    data = dict(N=N, K=K, G=G, X=X, y=y, group=grp + 1)
    center = np.concatenate([mu, np.log(tau + 0.1), beta.ravel()])
    _replay_and_compare("hier_logistic_cp", data, "lockstep_block", "diag_e", [0.002, 0.01], nch=6, max_depth=6,
                        q_scale=0.02, q_center=center, q_tol=1e-8)


def test_lockstep_block_replay_logistic_dense_d1000():
    """dense_e at D = 1000 (the C5 metric at its dimension) in block-per-chain
    mode: Cholesky momentum draw, per-leaf matvecs, six U-turn dots per merge."""
    data = logistic(seed=8, N=1500, D=1000)
    _replay_and_compare("logistic_regression", data, "lockstep_block", "dense_e", [0.003, 0.01], nch=4,
                        max_depth=6, q_scale=0.02, q_tol=1e-8)


@pytest.mark.parametrize("metric", ["unit_e", "diag_e"])
def test_fused_block_replay_hier_normal_nc(metric):
    """The fused kernel with one 256-thread block per chain (the driver the
    sampler takes for the last chains of a phase) on the non-centred
    hierarchical normal model, J = 700 (D = 702: each thread owns a few
    elements, the block-wide sums run through LDS), against the oracle."""
    data = hier_normal(seed=31, J=700)
    _replay_and_compare("hier_normal_nc", data, "fused_block", metric, (0.05, 0.12, 0.25), nch=12, max_depth=9,
                        q_scale=0.3)


def test_fused_block_driver_is_refused_for_other_models():
    gm = ca.Model("gauss", gauss(seed=2, D=4))
    q0 = np.zeros((2, 4))
    with pytest.raises(ca.SamplerError, match="block-per-chain"):
        gm.transition_replay(q0, np.full((2, 64), 0.5), np.zeros((2, 4)), 0.1, driver="fused_block")
