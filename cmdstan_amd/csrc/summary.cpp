// Summary statistics of stansummary (src/cmdstan/stansummary_helper.hpp:397-428
// get_stats), computed upstream by stan::mcmc::chainset (not in the reference
// tree).  Restated from Vehtari et al. (2021), "Rank-normalization, folding, and
// localization: an improved R-hat for assessing convergence of MCMC":
// split chains, rank normalisation (average ranks, Blom offsets), Geyer's
// initial positive + monotone sequence, bulk/tail ESS, R-hat = max(bulk, folded
// tail), MCSE(mean) = sd / sqrt(ESS of the unsplit chains).
//
// Autocovariances are accumulated lag by lag only as far as Geyer's truncation
// reads them, which for well-mixing chains is a few tens of lags.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

#include "../../include/cmdstan_amd.h"

namespace {

// Acklam's rational approximation of the standard normal quantile, refined by
// one Halley step on erfc (relative error ~1e-15).
double inv_Phi(double p) {
  static const double a[] = {-3.969683028665376e+01, 2.209460984245205e+02, -2.759285104469687e+02,
                             1.383577518672690e+02, -3.066479806614716e+01, 2.506628277459239e+00};
  static const double b[] = {-5.447609879822406e+01, 1.615858368580409e+02, -1.556989798598866e+02,
                             6.680131188771972e+01, -1.328068155288572e+01};
  static const double c[] = {-7.784894002430293e-03, -3.223964580411365e-01, -2.400758277161838e+00,
                             -2.549732539343734e+00, 4.374664141464968e+00, 2.938163982698783e+00};
  static const double d[] = {7.784695709041462e-03, 3.224671290700398e-01, 2.445134137142996e+00,
                             3.754408661907416e+00};
  if (p <= 0) return -std::numeric_limits<double>::infinity();
  if (p >= 1) return std::numeric_limits<double>::infinity();
  double x;
  if (p < 0.02425) {
    const double q = std::sqrt(-2 * std::log(p));
    x = (((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
        ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1);
  } else if (p > 1 - 0.02425) {
    const double q = std::sqrt(-2 * std::log(1 - p));
    x = -(((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
        ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1);
  } else {
    const double q = p - 0.5, r = q * q;
    x = (((((a[0] * r + a[1]) * r + a[2]) * r + a[3]) * r + a[4]) * r + a[5]) * q /
        (((((b[0] * r + b[1]) * r + b[2]) * r + b[3]) * r + b[4]) * r + 1);
  }
  const double e = 0.5 * std::erfc(-x / std::sqrt(2.0)) - p;
  const double u = e * std::sqrt(2 * M_PI) * std::exp(x * x / 2);
  return x - u / (1 + x * u / 2);
}

// draws x chains, column-major (chain j at x + j*n)
struct Mat {
  int n, m;
  std::vector<double> v;
  double& at(int i, int j) { return v[(size_t)j * n + i]; }
  double at(int i, int j) const { return v[(size_t)j * n + i]; }
};

Mat split(const double* x, int n, int m) {
  const int h = n / 2;
  Mat s{h, 2 * m, std::vector<double>((size_t)h * 2 * m)};
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < h; ++i) {
      s.at(i, 2 * j) = x[(size_t)j * n + i];
      s.at(i, 2 * j + 1) = x[(size_t)j * n + (n - h) + i];
    }
  return s;
}

double ess(const Mat& x) {
  const int n = x.n, m = x.m;
  if (n < 4) return std::numeric_limits<double>::quiet_NaN();
  // constant draws have no effective sample size (NaN, which the tail ESS's
  // min() then passes over when it is the second operand)
  bool varies = false;
  for (double v : x.v) varies = varies || v != x.v[0];
  if (!varies) return std::numeric_limits<double>::quiet_NaN();
  std::vector<double> mean(m), var(m);
  for (int j = 0; j < m; ++j) {
    double s = 0;
    for (int i = 0; i < n; ++i) s += x.at(i, j);
    mean[j] = s / n;
  }
  // mean over chains of the biased autocovariance at lag t
  auto acov_mean = [&](int t) {
    double tot = 0;
    for (int j = 0; j < m; ++j) {
      const double mu = mean[j];
      double s = 0;
      for (int i = 0; i + t < n; ++i) s += (x.at(i, j) - mu) * (x.at(i + t, j) - mu);
      tot += s / n;
    }
    return tot / m;
  };
  double mean_var = 0;
  for (int j = 0; j < m; ++j) {
    const double mu = mean[j];
    double s = 0;
    for (int i = 0; i < n; ++i) s += (x.at(i, j) - mu) * (x.at(i, j) - mu);
    mean_var += (s / n) * n / (n - 1.0);
  }
  mean_var /= m;
  double var_plus = mean_var * (n - 1.0) / n;
  if (m > 1) {
    double mm = 0;
    for (double v : mean) mm += v;
    mm /= m;
    double vb = 0;
    for (double v : mean) vb += (v - mm) * (v - mm);
    var_plus += vb / (m - 1.0);
  }
  std::vector<double> rho(n + 2, 0.0);
  double rho_even = 1.0;
  rho[0] = rho_even;
  double rho_odd = 1.0 - (mean_var - acov_mean(1)) / var_plus;
  rho[1] = rho_odd;
  int t = 1;
  while (t < n - 4 && (rho_even + rho_odd) > 0) {
    rho_even = 1.0 - (mean_var - acov_mean(t + 1)) / var_plus;
    rho_odd = 1.0 - (mean_var - acov_mean(t + 2)) / var_plus;
    if (rho_even + rho_odd >= 0) {
      rho[t + 1] = rho_even;
      rho[t + 2] = rho_odd;
    }
    t += 2;
  }
  const int max_t = t;
  if (rho_even > 0) rho[max_t + 1] = rho_even;
  for (t = 1; t <= max_t - 3; t += 2) {
    if (rho[t + 1] + rho[t + 2] > rho[t - 1] + rho[t]) {
      rho[t + 1] = (rho[t - 1] + rho[t]) / 2;
      rho[t + 2] = rho[t + 1];
    }
  }
  const double total = (double)n * m;
  double sum = 0;
  for (int k = 0; k < max_t; ++k) sum += rho[k];
  const double tau = -1.0 + 2.0 * sum + rho[max_t + 1];
  return std::min(total / tau, total * std::log10(total));
}

Mat rank_normalize(const Mat& x) {
  const size_t S = x.v.size();
  std::vector<std::pair<double, size_t>> vi(S);
  for (size_t i = 0; i < S; ++i) vi[i] = {x.v[i], i};
  std::sort(vi.begin(), vi.end());
  Mat r{x.n, x.m, std::vector<double>(S)};
  for (size_t i = 0; i < S;) {
    size_t j = i + 1;
    while (j < S && vi[j].first == vi[i].first) ++j;
    const double avg = (double)(i + 1 + j) / 2.0;
    const double z = inv_Phi((avg - 0.375) / (S + 0.25));
    for (size_t k = i; k < j; ++k) r.v[vi[k].second] = z;
    i = j;
  }
  return r;
}

double quantile7(std::vector<double> v, double p) {
  std::sort(v.begin(), v.end());
  const size_t n = v.size();
  const double h = (n - 1) * p;
  const size_t lo = (size_t)std::floor(h);
  const size_t hi = std::min(lo + 1, n - 1);
  return v[lo] + (h - lo) * (v[hi] - v[lo]);
}

double rhat(const Mat& x) {
  const int n = x.n, m = x.m;
  std::vector<double> cm(m), cv(m);
  for (int j = 0; j < m; ++j) {
    double s = 0;
    for (int i = 0; i < n; ++i) s += x.at(i, j);
    cm[j] = s / n;
    double q = 0;
    for (int i = 0; i < n; ++i) q += (x.at(i, j) - cm[j]) * (x.at(i, j) - cm[j]);
    cv[j] = q / (n - 1.0);
  }
  double mm = 0, w = 0;
  for (int j = 0; j < m; ++j) { mm += cm[j]; w += cv[j]; }
  mm /= m;
  w /= m;
  double b = 0;
  for (int j = 0; j < m; ++j) b += (cm[j] - mm) * (cm[j] - mm);
  b = n * b / (m - 1.0);
  return std::sqrt((b / w + n - 1) / n);
}

bool finite_and_varies(const Mat& x) {
  double lo = x.v[0], hi = x.v[0];
  for (double v : x.v) {
    if (!std::isfinite(v)) return false;
    lo = std::min(lo, v);
    hi = std::max(hi, v);
  }
  return hi > lo;
}

// out: mean, mcse, sd, mad, q05, q50, q95, ess_bulk, ess_tail, rhat
void param_stats(const double* x, int n, int m, double* out) {
  const double nan = std::numeric_limits<double>::quiet_NaN();
  const size_t S = (size_t)n * m;
  std::vector<double> all(x, x + S);
  double mean = 0;
  for (double v : all) mean += v;
  mean /= S;
  double ss = 0;
  for (double v : all) ss += (v - mean) * (v - mean);
  const double sd = std::sqrt(ss / (S - 1.0));
  const double med = quantile7(all, 0.5);
  std::vector<double> dev(S);
  for (size_t i = 0; i < S; ++i) dev[i] = std::fabs(all[i] - med);
  const double mad = 1.4826 * quantile7(dev, 0.5);
  out[0] = mean;
  out[2] = sd;
  out[3] = mad;
  out[4] = quantile7(all, 0.05);
  out[5] = med;
  out[6] = quantile7(all, 0.95);
  Mat full{n, m, all};
  out[1] = sd / std::sqrt(ess(full));
  Mat s = split(x, n, m);
  if (s.n < 4 || !finite_and_varies(s)) {
    out[7] = out[8] = out[9] = nan;
    return;
  }
  out[7] = ess(rank_normalize(s));
  const double q05 = quantile7(s.v, 0.05), q95 = quantile7(s.v, 0.95);
  Mat i05 = s, i95 = s;
  for (size_t i = 0; i < s.v.size(); ++i) {
    i05.v[i] = s.v[i] <= q05 ? 1.0 : 0.0;
    i95.v[i] = s.v[i] <= q95 ? 1.0 : 0.0;
  }
  out[8] = std::min(ess(i05), ess(i95));
  const double rb = rhat(rank_normalize(s));
  const double sm = quantile7(s.v, 0.5);
  Mat f = s;
  for (double& v : f.v) v = std::fabs(v - sm);
  const double rt = rhat(rank_normalize(f));
  out[9] = std::max(rb, rt);
}

}  // namespace

extern "C" {

/* Summary of n_params parameters; draws[p][chain][draw] (n_draws per chain).
 * out[p][10] = mean, MCSE, sd, MAD, 5%, 50%, 95%, ESS_bulk, ESS_tail, R_hat. */
int nuts_summary(const double* draws, int n_draws, int n_chains, int n_params, int n_threads, double* out) {
  if (n_draws < 1 || n_chains < 1 || n_params < 0) return NUTS_NOT_OK;
  if (n_threads < 1) n_threads = 1;
  std::vector<std::thread> th;
  for (int t = 0; t < n_threads; ++t)
    th.emplace_back([=]() {
      for (int p = t; p < n_params; p += n_threads)
        param_stats(draws + (size_t)p * n_draws * n_chains, n_draws, n_chains, out + (size_t)p * 10);
    });
  for (auto& x : th) x.join();
  return NUTS_OK;
}

}  // extern "C"
