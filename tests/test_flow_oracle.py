"""Pins of the flow oracle (CPU only): every function of oracle/flow.py is
checked against an independent route (complex-step derivatives, a
finite-difference Jacobian determinant, the inverse map, Adam's closed form)."""
import numpy as np
import pytest

from oracle import flow as F


def _flow(D=7, L=4, H=5, n_glob=2, seed=3, scale=True):
    masks = F.coupling_masks(D, L, n_glob=n_glob, group_size=1)
    rng = np.random.default_rng(seed)
    p = F.init_params(masks, H, seed)
    p += 0.3 * rng.standard_normal(p.size)  # leave the identity
    loc = rng.standard_normal(D) if scale else None
    sc = np.exp(0.3 * rng.standard_normal(D)) if scale else None
    return F.Flow(D, masks, H, p, loc, sc), rng


def test_masks_partition_every_layer():
    for D, ng, gs in [(7, 2, 1), (4020, 20, 10), (10, 0, 1), (3, 0, 1), (2, 0, 1)]:
        for a, b in F.coupling_masks(D, 8, ng, gs):
            assert len(a) > 0 and len(b) > 0
            assert sorted(np.concatenate([a, b]).tolist()) == list(range(D))


def test_group_layers_condition_on_globals():
    masks = F.coupling_masks(4020, 8, n_glob=20, group_size=10)
    a, b = masks[0]
    assert set(range(20)) <= set(a.tolist()) and min(b) >= 20
    a2, b2 = masks[2]
    assert sorted(b2.tolist()) == list(range(20))
    # groups are never split between a and b
    assert all(((i - 20) // 10) % 2 == 1 for i in b)


def test_inverse_undoes_forward():
    fl, rng = _flow()
    z = rng.standard_normal((5, fl.D))
    x, ld = fl.forward(z)
    z2, ld2 = fl.inverse(x)
    assert np.allclose(z2, z, atol=1e-12)
    assert np.allclose(ld2, -ld, atol=1e-12)


def test_logdet_is_log_abs_det_jacobian():
    fl, rng = _flow()
    z = rng.standard_normal(fl.D)
    h = 1e-6
    J = np.zeros((fl.D, fl.D))
    for k in range(fl.D):
        e = np.zeros(fl.D)
        e[k] = h
        J[:, k] = (fl.forward((z + e)[None])[0][0] - fl.forward((z - e)[None])[0][0]) / (2 * h)
    _, ld = fl.forward(z[None])
    assert abs(np.log(abs(np.linalg.det(J))) - ld[0]) < 1e-7


def test_pullback_grad_matches_complex_step():
    fl, rng = _flow()
    A = rng.standard_normal((fl.D, fl.D))
    A = A @ A.T + np.eye(fl.D)

    def U(z):
        x, ld = fl.forward(z[None])
        return 0.5 * x[0] @ A @ x[0] - ld[0]

    z = rng.standard_normal(fl.D)
    x, _ = fl.forward(z[None])
    g = fl.pullback_grad(z[None], (A @ x[0])[None])[0]
    hc = 1e-30
    gcs = np.array([np.imag(U(z + 1j * hc * np.eye(fl.D)[k])) / hc for k in range(fl.D)])
    assert np.allclose(g, gcs, rtol=1e-12, atol=1e-12)


def test_pullback_grad_planted_mistake_is_caught():
    """A sign slip in the log-determinant term moves the gradient visibly."""
    fl, rng = _flow()
    z = rng.standard_normal((1, fl.D))
    x, _ = fl.forward(z)
    g = fl.pullback_grad(z, x)
    g0 = F.Flow(fl.D, fl.masks, fl.H, np.zeros_like(fl.p), fl.loc, fl.scale).pullback_grad(z, x)
    assert np.max(np.abs(g - g0)) > 1e-3


def test_loss_grad_matches_complex_step():
    fl, rng = _flow(D=6, L=3, H=4)
    x = rng.standard_normal((9, fl.D))
    loss, g = fl.loss_grad(x)
    assert np.isclose(loss, fl.loss(x))
    hc = 1e-30
    for k in rng.choice(fl.n_params, 40, replace=False):
        p = fl.p.astype(complex)
        p[k] += 1j * hc
        assert abs(np.imag(fl.loss(x, p)) / hc - g[k]) <= 1e-12 * max(1.0, abs(g[k])), k


def test_loss_is_gaussian_nll_at_identity():
    D = 5
    masks = F.coupling_masks(D, 2)
    fl = F.Flow(D, masks, 3, np.zeros(F.param_layout(masks, 3)[1]))
    x = np.random.default_rng(0).standard_normal((4, D))
    assert np.isclose(fl.loss(x), np.mean(0.5 * np.sum(x * x, 1)))


def test_adam_first_step_closed_form():
    p = np.array([1.0, -2.0, 3.0])
    g = np.array([0.5, -4.0, 1e-3])
    p1, m, v = F.adam_step(p, g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    # bias-corrected m/sqrt(v) = g/|g| on the first step
    assert np.allclose(p1, p - 0.01 * g / (np.abs(g) + 1e-8))
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)
    p2, _, _ = F.adam_step(p1, g, m, v, 2, lr=0.01)
    assert np.allclose(p2, p1 - 0.01 * g / (np.abs(g) + 1e-8), atol=1e-9)
