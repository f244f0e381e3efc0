"""GPU parity of the affine-coupling flow (FP32 kernels) against the FP64 CPU
oracle (oracle/flow.py) on the same parameters, and the learned
reparameterisation inside the sampler.

Tolerances: the GPU flow runs in FP32 (weights, activations, matrix-core
products with FP32 accumulation); outputs are compared with the FP64 oracle at
a relative tolerance of 2e-4 (5e-4 for parameter gradients, whose entries are
sums over the batch)."""
import numpy as np
import pytest

import cmdstan_amd as ca
from oracle import ess as E
from oracle import flow as F

pytestmark = pytest.mark.gpu


def _pair(D=9, L=4, H=16, n_glob=2, group=1, seed=5, rows=64):
    g = ca.Flow(D, n_layers=L, hidden=H, n_glob=n_glob, group_size=group, max_rows=rows, seed=seed)
    rng = np.random.default_rng(seed)
    p = g.get_params().astype(np.float64)
    p = (p + 0.2 * rng.standard_normal(p.size)).astype(np.float32)
    g.set_params(p)
    loc = rng.standard_normal(D)
    scale = np.exp(0.3 * rng.standard_normal(D))
    g.set_standardization(loc, scale)
    masks = F.coupling_masks(D, L, n_glob, group)
    assert F.param_layout(masks, H)[1] == g.n_params
    o = F.Flow(D, masks, H, p.astype(np.float64), loc, scale)
    return g, o, rng


def _close(a, b, rtol):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b)) <= rtol * max(1.0, np.max(np.abs(b)))


@pytest.mark.parametrize("D,L,H,ng,gs", [(9, 4, 16, 2, 1), (24, 8, 32, 4, 4), (5, 3, 8, 0, 1), (130, 2, 64, 10, 10)])
def test_forward_inverse_match_oracle(D, L, H, ng, gs):
    g, o, rng = _pair(D, L, H, ng, gs)
    z = rng.standard_normal((40, D))
    x, ld = g.forward(z)
    xo, ldo = o.forward(z)
    assert _close(x, xo, 2e-4) and _close(ld, ldo, 2e-4)
    z2, ld2 = g.inverse(xo)
    zo, ldio = o.inverse(xo)
    assert _close(z2, zo, 2e-4) and _close(ld2, ldio, 2e-4)


@pytest.mark.parametrize("D,L,H,ng,gs", [(9, 4, 16, 2, 1), (24, 8, 32, 4, 4)])
def test_pullback_grad_matches_oracle(D, L, H, ng, gs):
    g, o, rng = _pair(D, L, H, ng, gs)
    z = rng.standard_normal((33, D))
    gx = rng.standard_normal((33, D))
    assert _close(g.pullback_grad(z, gx), o.pullback_grad(z, gx), 2e-4)


@pytest.mark.parametrize("n", [1, 17, 64])
def test_loss_grad_matches_oracle(n):
    g, o, rng = _pair(12, 4, 16, 2, 1, rows=64)
    x = rng.standard_normal((n, 12)) * 1.5 + 0.3
    loss, grad = g.loss_grad(x)
    lo, go = o.loss_grad(x)
    assert abs(loss - lo) <= 2e-4 * max(1.0, abs(lo))
    assert _close(grad, go, 5e-4)


def test_adam_training_fits_correlated_gaussian():
    """Maximum likelihood on draws of a strongly correlated gaussian: the loss
    falls towards the gaussian's entropy and the pulled-back draws whiten."""
    D = 6
    rng = np.random.default_rng(0)
    A = np.eye(D) + 0.9 * np.tril(rng.standard_normal((D, D)), -1)
    x = rng.standard_normal((8192, D)) @ A.T
    mu, sd = x.mean(0), x.std(0)
    g = ca.Flow(D, n_layers=6, hidden=32, max_rows=512, seed=3)
    g.set_standardization(mu, sd)
    l0, _ = g.loss_grad(x[:512])
    l1 = g.train(x, steps=1500, batch=512, lr=3e-3, seed=7)
    # exact NLL of the data (without D/2 log 2 pi): 0.5 * log det Sigma + D / 2
    target = 0.5 * np.linalg.slogdet(A @ A.T)[1] + D / 2
    assert l1 < l0 - 0.5
    assert l1 < target + 0.15
    z, _ = g.inverse(x[:512])
    c = np.corrcoef(z.T)
    assert np.max(np.abs(c - np.eye(D))) < 0.2


def test_flow_sampler_on_correlated_gaussian_gains_and_stays_exact():
    """NUTS on the pulled-back density of a correlated gaussian (diag_e cannot
    absorb the correlation): the trained flow raises ESS per gradient over
    plain NUTS, the automatic choice follows its rule, and NUTS through the
    flow keeps the target's moments."""
    D = 8
    rho = 0.95
    Sig = rho * np.ones((D, D)) + (1 - rho) * np.eye(D)
    m = ca.Model("gauss", dict(D=D, P=np.linalg.inv(Sig)))
    kw = dict(num_chains=64, num_warmup=400, num_samples=300, seed=11, flow_layers=4, flow_hidden=32,
              flow_train_steps=800, flow_batch=256, flow_lr=3e-3, flow_warmup=200, flow_pilot=60)
    auto = m.sample(flow="auto", **kw)
    assert np.all(auto.status == 0)
    # the trained network raises ESS per gradient ...
    assert auto.pilot_ess_per_grad_flow > 1.5 * auto.pilot_ess_per_grad_plain
    # ... and the sampler keeps it only if ESS per second rises as well
    stuck = auto.pilot_rhat_plain > 1.05 and auto.pilot_rhat_flow < auto.pilot_rhat_plain
    assert auto.flow_used == (auto.pilot_ess_per_grad_flow > auto.pilot_ess_per_grad_plain and
                              (auto.pilot_ess_per_sec_flow > auto.pilot_ess_per_sec_plain or stuck))
    fit = m.sample(flow="force", **kw)
    assert np.all(fit.status == 0) and fit.flow_used
    x = np.stack([fit.param(f"x.{i + 1}") for i in range(D)])
    for i in range(D):
        s = E.summary(x[i])
        assert abs(s["mean"]) < 4 * s["mcse"], (i, s)
        assert abs(s["sd"] - 1.0) < 0.1, (i, s)
        assert s["rhat"] < 1.01
    # lp__ is the model's log density at x (not at z)
    lp = np.stack([fit.draws[k][:, 0] for k in range(64)]).ravel()
    xs = np.stack([fit.draws[k][:, 7:7 + D] for k in range(64)]).reshape(-1, D)
    ref = -0.5 * np.einsum("ij,jk,ik->i", xs, np.linalg.inv(Sig), xs)
    assert np.allclose(lp, ref, atol=1e-6 * np.max(np.abs(ref)) + 1e-9)


def test_flow_sampler_eight_schools_matches_plain_moments():
    """Centered eight schools (funnel): NUTS through the trained flow agrees
    with a long plain FP64 run within 4 combined MCSE."""
    data = dict(J=8, y=[28, 8, -3, 7, -1, 1, 18, 12], sigma=[15, 10, 16, 11, 9, 11, 10, 18])
    m = ca.Model("eight_schools_cp", data)
    ref = m.sample(num_chains=64, num_warmup=1000, num_samples=1000, seed=3, delta=0.95)
    fit = m.sample(num_chains=64, num_warmup=600, num_samples=600, seed=4, delta=0.95, flow="force", flow_layers=4,
                   flow_hidden=32, flow_train_steps=600, flow_pilot=80, flow_warmup=300)
    assert np.all(fit.status == 0) and fit.flow_used
    for name in ["mu", "tau", "theta.1"]:
        a, b = E.summary(ref.param(name)), E.summary(fit.param(name))
        assert abs(a["mean"] - b["mean"]) < 4 * np.hypot(a["mcse"], b["mcse"]), (name, a, b)
