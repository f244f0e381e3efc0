// Per-chain model log densities for the sampler, one wavefront per chain.
//
// Each functor returns V = -log p(q) in Stan's propto + Jacobian convention
// (cmdstan src/cmdstan/command_helper.hpp:569-600 evaluates
// stan::model::log_prob_grad<true, jacobian>) and writes gV = -grad log p.
// Model data are packed per dataset into one double array; `off[c]` gives the
// dataset block of chain c.  write_array maps an unconstrained draw to the
// constrained outputs in Stan's column order.
#pragma once
#include "device_common.h"

namespace nutsgpu {

struct ModelData {
  const double* data;
  const int64_t* off;   // [C] offset of the chain's dataset block in data
  double jac;           // 1: add the log-Jacobian of the constraining transforms
};

// examples/bernoulli/bernoulli.stan: theta ~ beta(1,1); y ~ bernoulli(theta);
// block = [N, sum(y)], q = logit(theta)
struct BernoulliFn {
  static constexpr bool fused = true;
  static constexpr bool block_grad = false;
  static constexpr bool vec32 = true;  // FP32 vector storage supported
  ModelData md;
  template <class T>
  __device__ double operator()(int c, const T* __restrict__ q, T* __restrict__ gV, int, int lane) const {
    const double* d = md.data + md.off[c];
    const double N = d[0], s = d[1];
    const double u = q[0];
    const double t = inv_logit(u);
    const double lt = -log1p_exp(-u), l1t = -log1p_exp(u);
    double lp = s * lt + (N - s) * l1t;
    if (md.jac != 0) lp += lt + l1t;
    const double gr = (s + md.jac) * (1.0 - t) - (N - s + md.jac) * t;
    if (lane == 0) gV[0] = -gr;
    return -lp;
  }
  __device__ int nout(int) const { return 1; }
  __device__ void write_array(int, const double* q, double* out) const { out[0] = inv_logit(q[0]); }
};

// q ~ multi_normal_prec(mu, P); block = [D, mu[D], P[D*D] row-major]
struct GaussFn {
  static constexpr bool fused = true;
  static constexpr bool block_grad = false;
  static constexpr bool vec32 = true;  // FP32 vector storage supported
  ModelData md;
  template <class T>
  __device__ double operator()(int c, const T* __restrict__ q, T* __restrict__ gV, int D, int lane) const {
    const double* d = md.data + md.off[c];
    const double* mu = d + 1;
    const double* P = d + 1 + D;
    double acc = 0;
    for (int i = lane; i < D; i += NUTS_WAVE) {
      double si = 0;
      const double* Pi = P + (int64_t)i * D;
      for (int j = 0; j < D; ++j) si += Pi[j] * (q[j] - mu[j]);
      gV[i] = si;
      acc += (q[i] - mu[i]) * si;
    }
    return 0.5 * wave_sum(acc);
  }
  __device__ int nout(int c) const { return (int)md.data[md.off[c]]; }
  __device__ void write_array(int c, const double* q, double* out) const {
    const int D = (int)md.data[md.off[c]];
    for (int i = 0; i < D; ++i) out[i] = q[i];
  }
};

// Eight schools, centered: (mu, tau, theta[J]); theta ~ normal(mu, tau);
// y ~ normal(theta, sigma); optional mu ~ normal(0, mu_s), tau ~ cauchy(0, tau_s).
// block = [J, mu_s, tau_s, y[J], sigma[J]]
struct EightSchoolsFn {
  static constexpr bool fused = true;
  static constexpr bool block_grad = false;
  static constexpr bool vec32 = true;  // FP32 vector storage supported
  ModelData md;
  template <class T>
  __device__ double operator()(int c, const T* __restrict__ q, T* __restrict__ gV, int, int lane) const {
    const double* d = md.data + md.off[c];
    const int J = (int)d[0];
    const double mu_s = d[1], tau_s = d[2];
    const double* y = d + 3;
    const double* sg = d + 3 + J;
    const double mu = q[0], v = q[1], tau = exp(v);
    double acc = 0, gmu = 0, gtau = 0;
    for (int j = lane; j < J; j += NUTS_WAVE) {
      const double th = q[2 + j];
      const double z = (th - mu) / tau;
      const double e = (y[j] - th) / sg[j];
      acc += (-v - 0.5 * z * z) + (-0.5 * e * e);
      const double dz = (th - mu) / (tau * tau);
      gmu += dz;
      gtau += -1.0 / tau + (th - mu) * (th - mu) / (tau * tau * tau);
      gV[2 + j] = -(-dz + (y[j] - th) / (sg[j] * sg[j]));
    }
    acc = wave_sum(acc); gmu = wave_sum(gmu); gtau = wave_sum(gtau);
    double lp = 0;
    if (mu_s > 0) { lp -= 0.5 * mu * mu / (mu_s * mu_s); gmu -= mu / (mu_s * mu_s); }
    if (tau_s > 0) { lp -= log(1.0 + tau * tau / (tau_s * tau_s)); gtau -= 2.0 * tau / (tau_s * tau_s + tau * tau); }
    lp += md.jac * v + acc;
    if (lane == 0) { gV[0] = -gmu; gV[1] = -(md.jac + tau * gtau); }
    return -lp;
  }
  __device__ int nout(int c) const { return (int)md.data[md.off[c]] + 2; }
  __device__ void write_array(int c, const double* q, double* out) const {
    const int J = (int)md.data[md.off[c]];
    out[0] = q[0]; out[1] = exp(q[1]);
    for (int j = 0; j < J; ++j) out[2 + j] = q[2 + j];
  }
};

// Non-centered hierarchical normal: mu ~ normal(0, mu_s); tau ~ cauchy(0, tau_s), tau>0;
// eta ~ std_normal(); theta = mu + tau*eta; y ~ normal(theta, sigma).
// block = [J, mu_s, tau_s, y[J], sigma[J]]; q = (mu, log tau, eta); outputs mu, tau, eta, theta
struct HierNormalNCFn {
  static constexpr bool fused = true;
  static constexpr bool vec32 = true;       // FP32 vector storage supported
  static constexpr bool block_grad = true;  // gradient over a block per chain (group)
  ModelData md;
  // the gradient with TPC threads per chain (thread tid); sum reduces a
  // per-thread partial over the chain's threads in a fixed order
  template <int TPC, class T, class Sum>
  __device__ double group(int c, const T* __restrict__ q, T* __restrict__ gV, int tid, const Sum& sum) const {
    const double* d = md.data + md.off[c];
    const int J = (int)d[0];
    const double mu_s = d[1], tau_s = d[2];
    const double* y = d + 3;
    const double* sg = d + 3 + J;
    const double mu = q[0], v = q[1], tau = exp(v);
    double acc = 0, gmu = 0, gtau = 0;
#pragma unroll 4
    for (int j = tid; j < J; j += TPC) {
      const double eta = q[2 + j];
      const double th = mu + tau * eta;
      const double e = (y[j] - th) / sg[j];
      const double r = (y[j] - th) / (sg[j] * sg[j]);
      acc += -0.5 * eta * eta - 0.5 * e * e;
      gmu += r;
      gtau += r * eta;
      gV[2 + j] = eta - r * tau;
    }
    acc = sum(acc); gmu = sum(gmu); gtau = sum(gtau);
    double lp = -0.5 * mu * mu / (mu_s * mu_s) - log(1.0 + tau * tau / (tau_s * tau_s));
    lp += md.jac * v;
    lp += acc;
    gmu += -mu / (mu_s * mu_s);
    gtau += -2.0 * tau / (tau_s * tau_s + tau * tau);
    if (tid == 0) { gV[0] = -gmu; gV[1] = -(md.jac + tau * gtau); }
    return -lp;
  }
  template <class T>
  __device__ double operator()(int c, const T* __restrict__ q, T* __restrict__ gV, int, int lane) const {
    return group<NUTS_WAVE>(c, q, gV, lane, [](double x) { return wave_sum(x); });
  }
  __device__ int nout(int c) const { return 2 + 2 * (int)md.data[md.off[c]]; }
  __device__ void write_array(int c, const double* q, double* out) const {
    const int J = (int)md.data[md.off[c]];
    const double tau = exp(q[1]);
    out[0] = q[0]; out[1] = tau;
    for (int j = 0; j < J; ++j) { out[2 + j] = q[2 + j]; out[2 + J + j] = q[0] + tau * q[2 + j]; }
  }
};

// Centered hierarchical logistic: mu[k] ~ normal(0, mu_s); tau[k] ~ normal+(0,1);
// beta[g][k] ~ normal(mu[k], tau[k]); y_n ~ bernoulli_logit(x_n . beta[group_n]).
// Observations are sorted by group on the host.
// block = [N, K, G, mu_s, gstart[G+1], X[N*K], y[N]]; q = (mu[K], log tau[K], beta[G][K])
struct HierLogisticFn {
  static constexpr bool fused = true;
  static constexpr bool block_grad = false;
  static constexpr bool vec32 = true;  // FP32 vector storage supported
  ModelData md;
  template <class T>
  __device__ double operator()(int c, const T* __restrict__ q, T* __restrict__ gV, int, int lane) const {
    const double* d = md.data + md.off[c];
    const int N = (int)d[0], K = (int)d[1], G = (int)d[2];
    const double mu_s = d[3];
    const double* gstart = d + 4;
    const double* X = d + 5 + G;
    const double* y = X + (int64_t)N * K;
    const T* mu = q;
    const T* vv = q + K;
    const T* beta = q + 2 * K;
    T* gb = gV + 2 * K;
    double acc = 0;
    // groups: one lane owns whole groups, so the beta gradient needs no atomics
    for (int gi = lane; gi < G; gi += NUTS_WAVE) {
      const T* b = beta + (int64_t)gi * K;
      double gacc[16];
      for (int k = 0; k < K && k < 16; ++k) {
        const double tk = exp((double)vv[k]);
        const double dd = (double)b[k] - (double)mu[k];
        gacc[k] = -dd / (tk * tk);
        const double z = dd / tk;
        acc += -(double)vv[k] - 0.5 * z * z;
      }
      const int n0 = (int)gstart[gi], n1 = (int)gstart[gi + 1];
      for (int n = n0; n < n1; ++n) {
        const double* x = X + (int64_t)n * K;
        double eta = 0;
        for (int k = 0; k < K; ++k) eta += x[k] * b[k];
        acc += y[n] * eta - log1p_exp(eta);
        const double res = y[n] - inv_logit(eta);
        for (int k = 0; k < K && k < 16; ++k) gacc[k] += res * x[k];
      }
      for (int k = 0; k < K && k < 16; ++k) gb[(int64_t)gi * K + k] = -gacc[k];
    }
    // hyperparameters: lane k owns mu[k], tau[k]
    for (int k = lane; k < K; k += NUTS_WAVE) {
      const double tk = exp((double)vv[k]);
      double gm = -(double)mu[k] / (mu_s * mu_s);
      double gt = -tk;
      for (int gi = 0; gi < G; ++gi) {
        const double dd = (double)beta[(int64_t)gi * K + k] - (double)mu[k];
        gm += dd / (tk * tk);
        gt += -1.0 / tk + dd * dd / (tk * tk * tk);
      }
      acc += -0.5 * (double)mu[k] * (double)mu[k] / (mu_s * mu_s) - 0.5 * tk * tk + md.jac * (double)vv[k];
      gV[k] = -gm;
      gV[K + k] = -(md.jac + tk * gt);
    }
    return -wave_sum(acc);
  }
  __device__ int nout(int c) const {
    const double* d = md.data + md.off[c];
    return 2 * (int)d[1] + (int)d[1] * (int)d[2];
  }
  __device__ void write_array(int c, const double* q, double* out) const {
    const double* d = md.data + md.off[c];
    const int K = (int)d[1], G = (int)d[2];
    for (int k = 0; k < K; ++k) { out[k] = q[k]; out[K + k] = exp(q[K + k]); }
    for (int k = 0; k < K; ++k)
      for (int g = 0; g < G; ++g) out[2 * K + k * G + g] = q[2 * K + g * K + k];
  }
};

// Logistic regression evaluated one chain per wave (used by log_prob_grad and
// as the per-chain reference path); the sampler uses the batched GEMM kernels.
// block = [N, D, 1/s^2, X[N*D], y[N]]
struct LogisticFn {
  static constexpr bool fused = true;
  static constexpr bool block_grad = false;
  static constexpr bool vec32 = false;
  ModelData md;
  __device__ double operator()(int c, const double* __restrict__ q, double* __restrict__ gV, int D, int lane) const {
    const double* d = md.data + md.off[c];
    const int N = (int)d[0];
    const double s2inv = d[2];
    const double* X = d + 3;
    const double* y = X + (int64_t)N * D;
    // each lane accumulates its own partial gradient in gV stripes: first zero
    for (int k = lane; k < D; k += NUTS_WAVE) gV[k] = s2inv * q[k];
    double acc = 0;
    for (int n0 = 0; n0 < N; n0 += NUTS_WAVE) {
      const int n = n0 + lane;
      double res = 0;
      if (n < N) {
        const double* x = X + (int64_t)n * D;
        double eta = 0;
        for (int k = 0; k < D; ++k) eta += x[k] * q[k];
        acc += y[n] * eta - log1p_exp(eta);
        res = inv_logit(eta) - y[n];
      }
      // gV[k] += sum over the rows of this stripe of res * x[k]; the shuffle
      // runs in wave-uniform control flow, the k loop splits over lanes
      const int rmax = N - n0 < NUTS_WAVE ? N - n0 : NUTS_WAVE;
      for (int r = 0; r < rmax; ++r) {
        const double rr = __shfl(res, r, NUTS_WAVE);
        const double* x = X + (int64_t)(n0 + r) * D;
        for (int k = lane; k < D; k += NUTS_WAVE) gV[k] += rr * x[k];
      }
    }
    double pr = 0;
    for (int k = lane; k < D; k += NUTS_WAVE) pr += q[k] * q[k];
    pr = wave_sum(pr);
    return -(wave_sum(acc) - 0.5 * s2inv * pr);
  }
  __device__ int nout(int c) const { return (int)md.data[md.off[c] + 1]; }
  __device__ void write_array(int c, const double* q, double* out) const {
    const int D = (int)md.data[md.off[c] + 1];
    for (int k = 0; k < D; ++k) out[k] = q[k];
  }
};

}  // namespace nutsgpu
