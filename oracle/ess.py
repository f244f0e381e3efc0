"""ORACLE — test infrastructure only.

numpy restatement of the summary statistics CmdStan's ``stansummary`` prints
(src/cmdstan/stansummary_helper.hpp:397-428 ``get_stats``), computed upstream by
``stan::mcmc::chainset`` (not in the reference tree; restated from Vehtari,
Gelman, Simpson, Carpenter, Bürkner 2021, "Rank-normalization, folding, and
localization: an improved R-hat", Bayesian Analysis 16(2)):

* split chains (odd draw counts drop the middle draw),
* rank normalisation with average ranks for ties and Blom offsets
  ``Phi^-1((r - 3/8) / (S + 1/4))``,
* ESS from Geyer's initial positive then initial monotone sequence over the
  multi-chain autocorrelation, capped at ``S log10 S``,
* bulk ESS on rank-normalised split draws; tail ESS = min over the 5% and 95%
  quantile indicators; R-hat = max(bulk, folded tail);
* MCSE(mean) = sd / sqrt(ESS of the raw, unsplit chains), sd with n-1, MAD scaled by
  1.4826, quantiles of type 7.

Pinned by the golden strings in src/test/interface/stansummary_test.cpp:481-532.
"""
import numpy as np
from scipy.special import ndtri


def split_chains(x):
    """x: draws x chains -> (draws//2) x (2*chains)."""
    x = np.asarray(x, dtype=np.float64)
    if x.ndim == 1:
        x = x[:, None]
    n = x.shape[0]
    half = n // 2
    return np.concatenate([x[:half], x[n - half:]], axis=1)


def _autocov(y):
    """Biased autocovariance (1/n sum over overlapping pairs) via FFT."""
    n = len(y)
    yc = y - y.mean()
    m = 1
    while m < 2 * n:
        m *= 2
    f = np.fft.rfft(yc, m)
    ac = np.fft.irfft(f * np.conj(f), m)[:n]
    return ac / n


def ess(x):
    """ESS of a draws x chains matrix (chainset::ess restatement)."""
    x = np.asarray(x, dtype=np.float64)
    n, m = x.shape
    if n < 4:
        return np.nan
    acov = np.stack([_autocov(x[:, j]) for j in range(m)], axis=1)
    chain_mean = x.mean(axis=0)
    chain_var = acov[0] * n / (n - 1.0)
    mean_var = chain_var.mean()
    var_plus = mean_var * (n - 1.0) / n
    if m > 1:
        var_plus += np.var(chain_mean, ddof=1)
    rho = np.zeros(n)
    rho_even = 1.0
    rho[0] = rho_even
    rho_odd = 1.0 - (mean_var - acov[1].mean()) / var_plus
    rho[1] = rho_odd
    t = 1
    while t < n - 4 and (rho_even + rho_odd) > 0:
        rho_even = 1.0 - (mean_var - acov[t + 1].mean()) / var_plus
        rho_odd = 1.0 - (mean_var - acov[t + 2].mean()) / var_plus
        if rho_even + rho_odd >= 0:
            rho[t + 1] = rho_even
            rho[t + 2] = rho_odd
        t += 2
    max_t = t
    if rho_even > 0:
        rho[max_t + 1] = rho_even
    t = 1
    while t <= max_t - 3:
        if rho[t + 1] + rho[t + 2] > rho[t - 1] + rho[t]:
            rho[t + 1] = (rho[t - 1] + rho[t]) / 2.0
            rho[t + 2] = rho[t + 1]
        t += 2
    total = float(n * m)
    tau = -1.0 + 2.0 * rho[:max_t].sum() + rho[max_t + 1]
    return min(total / tau, total * np.log10(total))


def rank_normalize(x):
    x = np.asarray(x, dtype=np.float64)
    flat = x.ravel(order="F")
    S = flat.size
    order = np.argsort(flat, kind="stable")
    sv = flat[order]
    ranks = np.empty(S)
    i = 0
    while i < S:
        j = i + 1
        while j < S and sv[j] == sv[i]:
            j += 1
        ranks[order[i:j]] = (i + 1 + j) / 2.0  # average of ranks i+1..j
        i = j
    z = ndtri((ranks - 0.375) / (S + 0.25))
    return z.reshape(x.shape, order="F")


def quantile7(v, p):
    v = np.sort(np.ravel(np.asarray(v, dtype=np.float64)))
    n = len(v)
    h = (n - 1) * p
    lo = int(np.floor(h))
    hi = min(lo + 1, n - 1)
    return v[lo] + (h - lo) * (v[hi] - v[lo])


def _finite_and_varies(x):
    x = np.asarray(x)
    return np.all(np.isfinite(x)) and np.ptp(x) > 0


def ess_bulk_tail(x):
    s = split_chains(x)
    if s.shape[0] < 4 or not _finite_and_varies(s):
        return np.nan, np.nan
    bulk = ess(rank_normalize(s))
    q05 = quantile7(s, 0.05)
    q95 = quantile7(s, 0.95)
    t05 = ess((s <= q05).astype(np.float64))
    t95 = ess((s <= q95).astype(np.float64))
    return bulk, min(t05, t95)


def rhat(x):
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[0]
    cm = x.mean(axis=0)
    cv = x.var(axis=0, ddof=1)
    var_between = n * np.var(cm, ddof=1)
    var_within = cv.mean()
    return np.sqrt((var_between / var_within + n - 1) / n)


def rhat_bulk_tail(x):
    s = split_chains(x)
    if s.shape[0] < 4 or not _finite_and_varies(s):
        return np.nan, np.nan
    bulk = rhat(rank_normalize(s))
    med = np.median(s)
    tail = rhat(rank_normalize(np.abs(s - med)))
    return bulk, tail


def mcse_mean(x):
    x = np.asarray(x, dtype=np.float64)
    if x.ndim == 1:
        x = x[:, None]
    sd = np.std(x.ravel(), ddof=1)
    return sd / np.sqrt(ess(x))


def summary(x, probs=(0.05, 0.5, 0.95)):
    """stansummary row: Mean, MCSE, StdDev, MAD, quantiles..., ESS_bulk, ESS_tail, R_hat."""
    x = np.asarray(x, dtype=np.float64)
    if x.ndim == 1:
        x = x[:, None]
    v = x.ravel()
    mean = v.mean()
    sd = np.std(v, ddof=1)
    med = np.median(v)
    mad = 1.4826 * np.median(np.abs(v - med))
    qs = [quantile7(v, p) for p in probs]
    eb, et = ess_bulk_tail(x)
    rb, rt = rhat_bulk_tail(x)
    return dict(mean=mean, mcse=mcse_mean(x), sd=sd, mad=mad, quantiles=qs, ess_bulk=eb, ess_tail=et,
                rhat=max(rb, rt))
