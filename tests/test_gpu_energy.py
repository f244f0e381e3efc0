"""Per-transition energy error (max |H - H0| over a trajectory's leaves) and
the count of rejected initial values, both reported by nuts_sample."""
import numpy as np
import pytest

import cmdstan_amd as ca
from tests.model_cases import gauss

pytestmark = pytest.mark.gpu


def _fixed_step(m, eps, **kw):
    return m.sample(num_chains=32, num_warmup=50, num_samples=300, adapt_engaged=False, stepsize=eps, seed=4,
                    metric="unit_e", energy_error=True, **kw)


@pytest.mark.parametrize("fused", [1, 0])
def test_energy_error_bounds_accept_stat_and_scales_as_eps_squared(fused):
    m = ca.Model("gauss", gauss(seed=1, D=5))
    a = _fixed_step(m, 0.2, fused=fused)
    b = _fixed_step(m, 0.1, fused=fused)
    for fit in (a, b):
        assert np.all(fit.status == 0)
        for d, e in zip(fit.draws, fit.energy_error):
            assert np.all(e >= 0) and np.all(np.isfinite(e))
            # accept_stat is the mean over the leaves of min(1, exp(H0 - h)),
            # so it is at least exp(-max |H - H0|)
            assert np.all(d[:, 1] >= np.exp(-e) * (1 - 1e-12))
    ea = np.mean(np.concatenate(a.energy_error))
    eb = np.mean(np.concatenate(b.energy_error))
    print(f"mean max|dH|: eps 0.2 -> {ea:.3e}, eps 0.1 -> {eb:.3e}, ratio {ea / eb:.2f}")
    assert 2.5 < ea / eb < 6.0  # leapfrog energy error O(eps^2)


def test_energy_error_fp32_leapfrog_close_to_fp64():
    m = ca.Model("hier_normal_nc", dict(J=20, y=np.linspace(-10, 10, 20), sigma=np.full(20, 8.0)))
    kw = dict(num_chains=64, num_warmup=300, num_samples=300, seed=7, energy_error=True)
    a = m.sample(precision="fp64", **kw)
    b = m.sample(precision="fp32", **kw)
    ea, eb = np.concatenate(a.energy_error), np.concatenate(b.energy_error)
    ea, eb = ea[np.isfinite(ea)], eb[np.isfinite(eb)]
    print(f"max|dH| per transition: fp64 median {np.median(ea):.3e}, fp32 median {np.median(eb):.3e}")
    assert np.median(eb) < 2 * np.median(ea) + 1e-3


def test_init_rejection_counted():
    """A given initial value with a non-finite log density is rejected once
    (one try with a user init) and the chain fails with status 1."""
    m = ca.Model("gauss", dict(D=1))
    with pytest.raises(ca.SamplerError) as ei:
        m.sample(num_chains=2, init=[[1e200]], num_warmup=10, num_samples=10)
    fit = ei.value.fit
    assert np.all(fit.status == 1)
    assert fit.init_rejections.tolist() == [[1, 0], [1, 0]]
    fit = m.sample(num_chains=2, init=[[0.5]], num_warmup=10, num_samples=10)
    assert np.all(fit.init_rejections == 0)
