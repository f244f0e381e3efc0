"""The FP64 CPU oracle, pinned against the reference's fixtures, known-answer
vectors and closed forms (no GPU needed)."""
import os

import numpy as np
import pytest

from tests.conftest import load_stan_csv
from oracle import oracle as O
from oracle import ess as E

BERN = dict(N=10, y=[0, 1, 0, 0, 0, 0, 0, 0, 0, 1])  # examples/bernoulli/bernoulli.data.json


# ----------------------------------------------------------------- Philox
def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32_10
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert O.philox([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert O.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_stream_moments():
    u = np.array([O.stream_uniform(7, 3, 0, i) for i in range(20000)])
    assert np.all((u > 0) & (u < 1))
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1 / 12) < 0.003
    z = np.concatenate([O.stream_normal_pair(7, 3, i) for i in range(10000)])
    assert abs(z.mean()) < 0.03 and abs(z.std() - 1) < 0.02


# ------------------------------------------------------------ log density
def test_bernoulli_lp_matches_fixtures(golden):
    m = O.Model("bernoulli", BERN)
    for f in ["bernoulli_chain_1.csv", "bern_fitted_params.csv", "bern_fitted_params_warmup.csv"]:
        names, d, _ = load_stan_csv(f"{golden}/{f}")
        th = d[:, names.index("theta")]
        q = np.log(th / (1 - th))
        lp = np.array([m.log_prob([x]) for x in q])
        # closed form 3 log(theta) + 9 log(1-theta), fixture printed to 6 digits
        assert np.allclose(lp, 3 * np.log(th) + 9 * np.log1p(-th), rtol=1e-12)
        assert np.max(np.abs(lp - d[:, 0]) / np.abs(d[:, 0])) < 5e-6


def test_eight_schools_lp_matches_fixture(golden):
    names, d, _ = load_stan_csv(f"{golden}/eight_schools_output.csv")
    m = O.Model("eight_schools_cp", dict(J=8, y=[28, 8, -3, 7, -1, 1, 18, 12], sigma=[15, 10, 16, 11, 9, 11, 10, 18],
                                         mu_scale=10, tau_scale=10))
    q = np.column_stack([d[:, 7], np.log(d[:, 8]), d[:, 9:17]])
    lp = np.array([m.log_prob(r) for r in q])
    diff = d[:, 0] - lp
    assert abs(diff.mean()) < 1e-5
    assert np.max(np.abs(diff - diff.mean()) / np.abs(d[:, 0])) < 1e-4
    # a dropped prior is visibly wrong
    m2 = O.Model("eight_schools_cp", dict(J=8, y=[28, 8, -3, 7, -1, 1, 18, 12], sigma=[15, 10, 16, 11, 9, 11, 10, 18]))
    lp2 = np.array([m2.log_prob(r) for r in q])
    assert np.std(d[:, 0] - lp2) > 0.1


def _models(rng):
    N, D = 200, 5
    X = rng.standard_normal((N, D))
    y = (rng.random(N) < 0.4).astype(float)
    J = 7
    hl_N, K, G = 60, 3, 4
    return [
        ("bernoulli", BERN),
        ("logistic_regression", dict(N=N, D=D, X=X, y=y)),
        ("eight_schools_cp", dict(J=8, y=[28, 8, -3, 7, -1, 1, 18, 12], sigma=[15, 10, 16, 11, 9, 11, 10, 18],
                                  mu_scale=10, tau_scale=10)),
        ("hier_normal_nc", dict(J=J, y=rng.standard_normal(J) * 5, sigma=rng.uniform(5, 20, J))),
        ("hier_logistic_cp", dict(N=hl_N, K=K, G=G, X=rng.standard_normal((hl_N, K)),
                                  group=rng.integers(1, G + 1, hl_N), y=(rng.random(hl_N) < 0.5).astype(float))),
        ("gauss", dict(D=3, mu=[1.0, -2.0, 0.5], P=np.array([[2.0, 0.3, 0], [0.3, 1.0, 0.2], [0, 0.2, 0.5]]))),
        ("funnel", dict(K=4)),
    ]


@pytest.mark.parametrize("idx", range(7))
def test_gradient_complex_step(idx):
    rng = np.random.default_rng(11 + idx)
    name, data = _models(rng)[idx]
    m = O.Model(name, data)
    for _ in range(3):
        q = rng.standard_normal(m.dim) * 0.7
        for jac in (True, False):
            lp, g = m.log_prob_grad(q, jac)
            assert lp == pytest.approx(m.log_prob(q, jac), rel=1e-14, abs=1e-12)
            gc = m.grad_cs(q, jacobian=jac)
            scale = max(1.0, np.max(np.abs(gc)))
            assert np.max(np.abs(g - gc)) <= 1e-12 * scale
            # central differences as an independent sanity check
            h = 1e-6
            fd = np.array([(m.log_prob(q + h * e, jac) - m.log_prob(q - h * e, jac)) / (2 * h) for e in np.eye(m.dim)])
            assert np.max(np.abs(fd - g)) <= 1e-5 * scale


def test_funnel_closed_form():
    """Neal's funnel: log p = -v^2/18 - sum_k (x_k^2 e^{-v} + v) / 2; the
    derived output sigma = e^{v/2}."""
    m = O.Model("funnel", dict(K=2))
    q = np.array([0.7, -1.1, 0.4])
    lp, g = m.log_prob_grad(q)
    v, x = q[0], q[1:]
    assert lp == pytest.approx(-v * v / 18 - 0.5 * np.sum(x * x) * np.exp(-v) - v, rel=1e-15)
    assert g == pytest.approx([-v / 9 + 0.5 * np.sum(x * x) * np.exp(-v) - 1.0, *(-x * np.exp(-v))], rel=1e-14)
    assert m.write_array(q) == pytest.approx([v, np.exp(v / 2), *x], rel=1e-15)


def test_logistic_gradient_closed_form():
    rng = np.random.default_rng(5)
    N, D = 300, 4
    X = rng.standard_normal((N, D))
    y = (rng.random(N) < 0.5).astype(float)
    m = O.Model("logistic_regression", dict(N=N, D=D, X=X, y=y, prior_scale=2.5))
    b = rng.standard_normal(D)
    lp, g = m.log_prob_grad(b)
    eta = X @ b
    assert g == pytest.approx(X.T @ (y - 1 / (1 + np.exp(-eta))) - b / 6.25, rel=1e-12, abs=1e-12)
    assert lp == pytest.approx(np.sum(y * eta - np.logaddexp(0, eta)) - 0.5 * b @ b / 6.25, rel=1e-13)


# --------------------------------------------------------------- leapfrog
def test_leapfrog_gaussian_closed_form():
    # V = 1/2 q'Pq: one leapfrog step is the linear map
    # p1 = p - e/2 P q; q' = q + e Minv p1; p' = p1 - e/2 P q'
    P = np.array([[2.0, 0.5], [0.5, 1.0]])
    m = O.Model("gauss", dict(D=2, P=P))
    q = np.array([0.3, -1.2]); p = np.array([0.7, 0.1]); minv = np.array([0.5, 2.0]); e = 0.3
    lp, g = m.log_prob_grad(q)
    q1, p1, g1, V1 = m.leapfrog(q, p, -g, -lp, e, metric=1, minv=minv)
    ph = p - e / 2 * (P @ q)
    qn = q + e * minv * ph
    pn = ph - e / 2 * (P @ qn)
    assert np.allclose(q1, qn, rtol=1e-15, atol=1e-15) and np.allclose(p1, pn, rtol=1e-15, atol=1e-15)
    assert V1 == pytest.approx(0.5 * qn @ P @ qn, rel=1e-14)
    # reversibility: run -e from the end, land back on the start
    q2, p2, g2, V2 = m.leapfrog(q1, p1, g1, V1, -e, metric=1, minv=minv)
    assert np.allclose(q2, q, atol=1e-14) and np.allclose(p2, p, atol=1e-14)


def test_leapfrog_energy_error_second_order():
    m = O.Model("gauss", dict(D=1))
    errs = []
    for e in [0.1, 0.05]:
        q, p = np.array([1.0]), np.array([0.5])
        lp, g = m.log_prob_grad(q)
        H0 = 0.5 * p @ p - lp
        q1, p1, _, V1 = m.leapfrog(q, p, -g, -lp, e, metric=0)
        errs.append(abs(0.5 * p1 @ p1 + V1 - H0))
    assert 6.5 < errs[0] / errs[1] < 9.0  # one step: local energy error O(e^3), ratio ~8


# ------------------------------------------------------------ adaptation
def test_dual_averaging_reproduces_fixture_stepsizes(golden):
    """bern_fitted_params_warmup.csv holds save_warmup draws: outside the
    window resets, stepsize__ at iteration t+1 is exp(x_t) of the dual-averaging
    recurrence fed with the printed accept_stat__ of iteration t."""
    names, d, comments = load_stan_csv(f"{golden}/bern_fitted_params_warmup.csv")
    acc, eps = d[:1000, 1], d[:1000, 2]
    resets = [0, 100, 150, 250, 450, 950]  # 0-based iterations that follow init_stepsize
    for a, b in zip(resets, resets[1:] + [1000]):
        mu = np.log(10 * (1.0 if a == 0 else eps[a]))
        e_pred, xbar = O.dual_averaging(acc[a:b], mu)
        np.testing.assert_allclose(e_pred[:-1], eps[a + 1:b], rtol=2e-4)
        if b == 1000:
            final = float([c for c in comments if c.startswith("# Step size")][0].split("=")[1])
            assert np.exp(xbar) == pytest.approx(final, rel=2e-4)


def test_window_schedule():
    mode, col, upd = O.window_schedule(1000)
    assert mode == 0 and list(np.nonzero(upd)[0] + 1) == [100, 150, 250, 450, 950]
    assert col.sum() == 1000 - 75 - 50 and col[74] == 0 and col[75] == 1 and col[949] == 1 and col[950] == 0
    mode, col, upd = O.window_schedule(100)   # 15/75/10 branch: metric never updated
    assert mode == 2 and upd.sum() == 0
    mode, col, upd = O.window_schedule(19)
    assert mode == 1 and upd.sum() == 0 and col.sum() == 0


# ---------------------------------------------------------- NUTS transition
def test_transition_d1_by_hand():
    """Standard normal, D=1, unit metric: enumerate the first doubling by hand."""
    m = O.Model("gauss", dict(D=1))
    q0, p0, e = 0.4, 1.1, 0.5
    # forward direction (u=0.9 > 0.5), depth-0 tree: one leapfrog
    ph = p0 - e / 2 * q0; q1 = q0 + e * ph; p1 = ph - e / 2 * q1
    H0 = 0.5 * p0 ** 2 + 0.5 * q0 ** 2; h1 = 0.5 * p1 ** 2 + 0.5 * q1 ** 2
    # uniforms: direction, top-level acceptance, then a direction that stops via max_depth=1
    q, info = m.transition_replay([q0], [0.9, 0.0], [p0], e, metric=0, max_depth=1)
    assert info["n_leapfrog"] == 1 and info["depth"] == 1 and info["divergent"] == 0
    assert info["accept_stat"] == pytest.approx(min(1.0, np.exp(H0 - h1)), rel=1e-15)
    w_new = H0 - h1
    if w_new > 0:      # subtree heavier than the old trajectory: taken without a uniform
        assert q[0] == pytest.approx(q1, rel=1e-15) and info["uniforms_used"] == 1
    else:
        assert q[0] == pytest.approx(q1, rel=1e-15) and info["uniforms_used"] == 2


def test_transition_invariants_and_divergence():
    rng = np.random.default_rng(3)
    m = O.Model("gauss", dict(D=4, P=np.diag([1.0, 4.0, 0.25, 9.0])))
    for eps in [0.05, 0.3, 1.0]:
        for _ in range(20):
            u = rng.random(4000); z = rng.standard_normal(4)
            q, info = m.transition_replay(rng.standard_normal(4), u, z, eps, metric=1, max_depth=8)
            d, n = info["depth"], info["n_leapfrog"]
            assert 2 ** d - 1 <= n <= 2 ** (d + 1) - 1
            assert 0 <= info["accept_stat"] <= 1
            if info["divergent"]:
                assert n != 2 ** d - 1
    # a huge step size diverges on the first leaf
    q, info = m.transition_replay(np.ones(4), [0.9] * 10, np.ones(4), 50.0, metric=1)
    assert info["divergent"] == 1 and info["n_leapfrog"] == 1 and info["depth"] == 0


def _gauss_momenta(q0, p0, w, eps, n):
    """Momenta p_0..p_n of n forward leapfrog steps on V = sum_i w_i q_i^2 / 2
    with the unit metric.  Per coordinate one step is the linear map
    (q, p) -> [[1 - e^2 w / 2, e], [-e w (1 - e^2 w / 4), 1 - e^2 w / 2]] (q, p)."""
    q, p = np.array(q0, float), np.array(p0, float)
    w = np.asarray(w, float)
    a = 1 - eps * eps * w / 2
    ps = [p.copy()]
    for _ in range(n):
        q, p = a * q + eps * p, -eps * w * (1 - eps * eps * w / 4) * q + a * p
        ps.append(p.copy())
    return np.array(ps)


# Two trajectories on a D=2 Gaussian (precisions w, unit metric, every
# doubling forward) where Stan's generalized U-turn checks across the two
# halves of a merge stop the tree while the check over the whole merged span
# alone would let it grow.  In each, the merged span [a, b) with halves
# [a, m) and [m, b) has p_a . rho > 0 and p_{b-1} . rho > 0 for rho = sum of
# its momenta, but p_{m} . (rho_[a,m) + p_m) or p_{m-1} . (rho_[m,b) + p_{m-1})
# is negative.
UTURN_TOP = dict(w=[1.0, 9.0], eps=0.5, q0=[1.4, -1.1], p0=[1.3, -0.6], span=(0, 2, 4), n_lf=3, depth=2,
                 n_lf_without=7)   # top-level merge of the second doubling: old {0, 1} + new {2, 3}
UTURN_INNER = dict(w=[1.0, 25.0], eps=0.3, q0=[-0.6, 0.9], p0=[-1.2, 0.3], span=(4, 6, 8), n_lf=7, depth=2,
                   n_lf_without=15)  # root merge of the third doubling's subtree {4, 5} + {6, 7}


def _uturn_dots(case):
    a, m, b = case["span"]
    ps = _gauss_momenta(case["q0"], case["p0"], case["w"], case["eps"], b)
    rho = ps[a:b].sum(0)
    whole = (ps[a] @ rho, ps[b - 1] @ rho)
    left = ps[a:m].sum(0) + ps[m]
    right = ps[m:b].sum(0) + ps[m - 1]
    cross = (ps[a] @ left, ps[m] @ left, ps[m - 1] @ right, ps[b - 1] @ right)
    return whole, cross


@pytest.mark.parametrize("case", [UTURN_TOP, UTURN_INNER], ids=["top_level", "inside_subtree"])
def test_uturn_cross_checks_decide(case):
    whole, cross = _uturn_dots(case)
    assert whole[0] > 0 and whole[1] > 0          # the check over the merged span passes
    assert min(cross) < 0                          # a check across its halves fails
    m = O.Model("gauss", dict(D=2, P=np.diag(case["w"])))
    _, info = m.transition_replay(case["q0"], [0.9] * 64, case["p0"], case["eps"], metric=0, max_depth=6)
    assert info["n_leapfrog"] == case["n_lf"] and info["depth"] == case["depth"]
    assert info["divergent"] == 0


def _oracle_variant(tmp_path, edit):
    """The oracle rebuilt from a copy of its sources with `edit` applied to
    nuts.hpp (a planted mistake, to show a pin test catches it)."""
    import shutil
    import subprocess
    src = os.path.join(os.path.dirname(O.__file__))
    for f in ("oracle_capi.cpp", "nuts.hpp", "models.hpp", "philox.hpp"):
        shutil.copy(os.path.join(src, f), tmp_path / f)
    text = (tmp_path / "nuts.hpp").read_text()
    new = edit(text)
    assert new != text
    (tmp_path / "nuts.hpp").write_text(new)
    so = tmp_path / "liboracle_variant.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-ffp-contract=off", "-pthread", "-shared", "-o", str(so),
                    str(tmp_path / "oracle_capi.cpp")], check=True)
    return O.load(str(so))


def test_uturn_pins_catch_dropped_cross_checks(tmp_path):
    """With the checks across subtree halves deleted from the oracle, both
    trees above grow further: the pins fail on that mistake."""
    import re
    L = _oracle_variant(tmp_path, lambda t: re.sub(r"persist &= criterion\([^;]*\);", ";", t))
    for case in (UTURN_TOP, UTURN_INNER):
        m = O.Model("gauss", dict(D=2, P=np.diag(case["w"])), library=L)
        _, info = m.transition_replay(case["q0"], [0.9] * 64, case["p0"], case["eps"], metric=0, max_depth=6)
        assert info["n_leapfrog"] == case["n_lf_without"] != case["n_lf"]


# base_hmc::init_stepsize on the standard normal from q0 = 0 with the unit
# metric: one leapfrog of size e from (0, p) ends at (e p, p (1 - e^2 / 2)), so
# the probe's H0 - h = -p^2 e^4 / 8, and the step is kept when that exceeds
# log 0.8, i.e. p^2 e^4 < 1.785.  Every probe draws a fresh momentum.
@pytest.mark.parametrize("eps0,normals,expect_eps,expect_used", [
    (1.0, [1.0] * 8, 2.0, 3),       # 1/8 ok -> up; at 1: ok -> 2; at 2: 16/8 too big -> stop at 2
    (1.0, [0.5] * 8, 2.0, 3),       # p^2 = 1/4: e^4 < 7.14 -> 1 ok, 2 (e^4 = 16) too big
    (4.0, [1.0] * 8, 1.0, 4),       # 4: too big -> down; 4 -> 2 -> 1, at 1 the probe is acceptable
    (1.0, [0.1, 3.0, 0.1, 0.1], 1.0, 2),  # direction up from p = 0.1; the second probe (p = 3) is too big
    (0.25, [2.0] * 8, 1.0, 4),      # p^2 e^4: 4/256 ok -> up; at 0.25 ok -> 0.5; 1/4 at 0.5 ok -> 1; 4 at 1 too big
])
def test_init_stepsize_by_hand(eps0, normals, expect_eps, expect_used):
    m = O.Model("gauss", dict(D=1))
    eps, used = m.init_stepsize([0.0], normals, eps0, metric=0)
    assert eps == expect_eps and used == expect_used


def test_init_stepsize_improper_limit():
    # zero momenta: every probe keeps H, the step doubles past 1e7
    m = O.Model("gauss", dict(D=1))
    with pytest.raises(RuntimeError, match="improper"):
        m.init_stepsize([0.0], [0.0] * 64, 1.0, metric=0)


def _two_sample_chi2(a, b, min_count=10):
    from scipy.stats import chi2_contingency
    vals = sorted(set(a.tolist()) | set(b.tolist()))
    ta = np.array([np.sum(a == v) for v in vals])
    tb = np.array([np.sum(b == v) for v in vals])
    keep = (ta + tb) >= min_count
    return chi2_contingency(np.vstack([ta[keep], tb[keep]]))[1]


def test_tree_sizes_match_fixtures_at_their_adapted_step(golden):
    """Adaptation off at each fixture's printed step size and inverse metric:
    the oracle's tree depths (and, for bernoulli, leapfrog counts) have the
    fixture rows' distribution (two-sample chi-square)."""
    names, d, _ = load_stan_csv(f"{golden}/bernoulli_chain_1.csv")   # :40-43 adaptation block
    r = O.Model("bernoulli", BERN).run(seed=5, num_chains=4, adapt=0, num_warmup=100, num_samples=2500,
                                       stepsize=1.02703, inv_metric=[0.526879])
    dr = r["draws"].reshape(-1, r["draws"].shape[2])
    assert _two_sample_chi2(d[:, 3], dr[:, 3]) > 1e-3
    assert _two_sample_chi2(d[:, 4], dr[:, 4]) > 1e-3
    names, d8, _ = load_stan_csv(f"{golden}/eight_schools_output.csv")   # :40-41
    mi = [13.2101, 0.318126, 36.4485, 29.83, 33.6813, 25.6287, 21.8178, 26.9628, 25.8657, 32.7886]
    eight = dict(J=8, y=[28.0, 8, -3, 7, -1, 1, 18, 12], sigma=[15.0, 10, 16, 11, 9, 11, 10, 18], mu_scale=10.0,
                 tau_scale=10.0)
    r = O.Model("eight_schools_cp", eight).run(seed=5, num_chains=4, adapt=0, num_warmup=200, num_samples=2500,
                                                stepsize=0.117114, inv_metric=mi)
    dr = r["draws"].reshape(-1, r["draws"].shape[2])
    assert _two_sample_chi2(d8[:, 3], dr[:, 3]) > 1e-3


def test_fixture_tree_invariant(golden):
    for f in ["bernoulli_chain_1.csv", "eight_schools_output.csv", "bern_fitted_params_warmup.csv"]:
        names, d, _ = load_stan_csv(f"{golden}/{f}")
        dep, n, div = d[:, 3], d[:, 4], d[:, 5]
        assert np.all((n >= 2 ** dep - 1) & (n <= 2 ** (dep + 1) - 1))
        assert not np.any((div == 1) & (n == 2 ** dep - 1))


def test_multinomial_selection_frequencies():
    """Over many replays with fresh uniforms, the selected state of a fixed
    forward depth-2 trajectory follows Stan's scheme: biased progressive
    sampling at the top level, uniform progressive sampling inside subtrees."""
    m = O.Model("gauss", dict(D=1))
    q0, p0, e = 0.2, 0.9, 0.4
    qs, ps = [q0], [p0]
    for _ in range(3):
        ph = ps[-1] - e / 2 * qs[-1]; qn = qs[-1] + e * ph; ps.append(ph - e / 2 * qn); qs.append(qn)
    lw = np.array([0.5 * (a * a + b * b) for a, b in zip(qs, ps)])
    lw = lw[0] - lw
    sub = np.logaddexp(lw[2], lw[3]); tot1 = np.logaddexp(lw[0], lw[1])
    top1_draw = not (lw[1] > lw[0])      # uniform consumed at the first top-level merge
    top2_draw = not (sub > tot1)
    rng = np.random.default_rng(9)
    n = 20000
    picks = []
    for _ in range(n):
        u = [0.9]
        if top1_draw:
            u.append(rng.random())
        u += [0.9, rng.random()]         # direction, inner merge of leaves 2 and 3
        if top2_draw:
            u.append(rng.random())
        q, info = m.transition_replay([q0], u, [p0], e, metric=0, max_depth=2)
        assert info["n_leapfrog"] == 3 and info["uniforms_used"] == len(u)
        picks.append(q[0])
    picks = np.array(picks)
    dist = np.abs(picks[:, None] - np.array(qs)[None, :])
    assert np.all(dist.min(axis=1) < 1e-12)
    counts = np.bincount(dist.argmin(axis=1), minlength=4)
    p_take1 = min(1.0, np.exp(lw[1] - lw[0]))
    p_take_sub = min(1.0, np.exp(sub - tot1))
    p3_in_sub = np.exp(lw[3] - sub)
    probs = np.array([(1 - p_take1) * (1 - p_take_sub), p_take1 * (1 - p_take_sub),
                      p_take_sub * (1 - p3_in_sub), p_take_sub * p3_in_sub])
    keep = probs > 0
    assert np.all(counts[~keep] == 0)
    chi2 = np.sum((counts[keep] - n * probs[keep]) ** 2 / (n * probs[keep]))
    assert chi2 < 20  # at most 3 dof


# -------------------------------------------------------------- end to end
def test_bernoulli_posterior():
    m = O.Model("bernoulli", BERN)
    r = m.run(seed=2024, num_chains=4)
    th = r["draws"][:, :, 7].T
    s = E.summary(th)
    mean, sd = 0.25, np.sqrt(3 * 9 / (12 ** 2 * 13))  # Beta(3, 9)
    assert abs(s["mean"] - mean) < 4 * s["mcse"]
    assert abs(s["sd"] - sd) < 0.01
    assert s["rhat"] < 1.01


def test_metric_file_roundtrip_no_adapt():
    # adapt engaged=0 keeps the given step size and inverse metric
    m = O.Model("gauss", dict(D=2, P=np.diag([1 / 100.0, 1.0])))
    r = m.run(seed=5, num_chains=2, adapt=0, num_warmup=10, num_samples=50, stepsize=0.5,
              inv_metric=[1.5, 1.8])
    assert np.all(r["stepsize"] == 0.5) and np.all(r["inv_metric"] == [1.5, 1.8])
    r = m.run(seed=5, num_chains=1, adapt=1, num_warmup=100, num_samples=50, inv_metric=[1.5, 1.8])
    assert np.all(r["inv_metric"] == [1.5, 1.8])  # 15/75/10 branch never updates the metric


# ----------------------------------------------------------- ESS / R-hat
def test_stansummary_golden(golden):
    names, d, _ = load_stan_csv(f"{golden}/bernoulli_chain_1.csv")
    # src/test/interface/stansummary_test.cpp:481-532
    gold = {"lp__": [-7.2719, 0.0365168, 0.768874, 0.303688, -8.98426, -6.97009, -6.75007, 519.29, 503.309, 1.00141],
            "energy__": [7.78428, 0.0508815, 1.0314, 0.745859, 6.80565, 7.46758, 9.8864, 489.874, 486.438, 1.00495],
            "theta": [0.256552, 0.00610844, 0.119654, 0.120965, 0.0802982, 0.24996, 0.47034, 361.506, 395.736,
                      1.00186]}
    for k, g in gold.items():
        s = E.summary(d[:, names.index(k)])
        got = [s["mean"], s["mcse"], s["sd"], s["mad"], *s["quantiles"], s["ess_bulk"], s["ess_tail"], s["rhat"]]
        assert [float(f"{v:.6g}") for v in got] == g, k
    # console row for accept_stat__ (stansummary_test.cpp:431-433): ESS 1284 / 941
    s = E.summary(d[:, names.index("accept_stat__")])
    assert round(s["ess_bulk"]) == 1284


def test_ess_iid_and_ar1():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((1000, 4))
    assert 3000 < E.ess(x) < 5000
    # AR(1) with phi: ESS/N -> (1 - phi) / (1 + phi)
    phi = 0.8
    y = np.zeros((20000, 2))
    for t in range(1, 20000):
        y[t] = phi * y[t - 1] + rng.standard_normal(2)
    r = E.ess(y) / y.size
    assert abs(r - (1 - phi) / (1 + phi)) < 0.02
