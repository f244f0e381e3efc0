// Batched gradient of the centered hierarchical logistic model for many chains
// (one dataset):
//   mu_k ~ N(0, mu_s), tau_k = exp(v_k) ~ HalfNormal(1) (+ v_k Jacobian),
//   beta[g][k] ~ N(mu_k, tau_k), y_n ~ bernoulli_logit(x_n . beta[group_n]).
// Kernel 1 takes one group and 64 chains per block: lane = chain, the group's
// observations are staged in LDS and read as broadcasts, each lane keeps its
// chain's beta[g] and gradient in registers (K <= 16), the four waves split the
// observations and combine through LDS.  It writes the beta[g] gradient
// directly and per-group partials for the hyperparameters; kernel 2 reduces
// those over groups in a fixed order (deterministic, no atomics).
// Dataset block (host layout): [N, K, G, mu_s, gstart[G+1], X[N*K] sorted by group, y[N]].
#include "device_common.h"
#include "hier_logistic.h"

namespace nutsgpu {

#define HL_TC 64     // chains per block
#define HL_OBS 128   // observations staged per LDS pass
#define HL_KMAX 16

// log(1 + e^e) and sig = 1 / (1 + e^-e) from one exponential and one reciprocal
__device__ __forceinline__ double hl_log1p_exp(double e, double* sig) {
  const double t = exp(-fabs(e));
  double r1;
  const double l1p = log1p_recip_unit(t, r1);
  *sig = e >= 0.0 ? r1 : t * r1;
  return (e > 0.0 ? e : 0.0) + l1p;
}

__global__ void __launch_bounds__(256) hl_group_kernel(HlArgs a) {
  __shared__ double sX[HL_OBS * HL_KMAX];
  __shared__ double sY[HL_OBS];
  __shared__ double sRed[4][HL_TC][HL_KMAX + 1];
  const double* d = a.block;
  const int K = (int)d[1], G = (int)d[2];
  const double* gstart = d + 4;
  const double* X = d + 5 + G;
  const int N = (int)d[0];
  const double* y = X + (int64_t)N * K;
  const int g = blockIdx.x % G;
  const int tile = blockIdx.x / G;
  const int n_act = a.n_dev ? min(a.n_rows, *a.n_dev) : a.n_rows;
  const int p0 = tile * HL_TC;
  if (p0 >= n_act) return;
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const int pos = p0 + lane;
  const bool live = pos < n_act;
  const int64_t row_in = live ? (a.idx_in ? (int64_t)a.idx[pos] : pos) : 0;
  const double* q = a.Q + row_in * a.ldq;
  double b[HL_KMAX], gb[HL_KMAX];
#pragma unroll
  for (int k = 0; k < HL_KMAX; ++k) {
    b[k] = (live && k < K) ? q[2 * K + (int64_t)g * K + k] : 0.0;
    gb[k] = 0.0;
  }
  double lp = 0.0;
  const int n0 = (int)gstart[g], n1 = (int)gstart[g + 1];
  for (int base = n0; base < n1; base += HL_OBS) {
    const int cnt = min(HL_OBS, n1 - base);
    __syncthreads();
    for (int i = t; i < cnt * K; i += 256) sX[i] = X[(int64_t)base * K + i];
    for (int i = t; i < cnt; i += 256) sY[i] = y[base + i];
    __syncthreads();
    // two observations per step: independent exp/log chains in flight
    for (int n = w; n < cnt; n += 8) {
      const int n2 = n + 4;
      const bool two = n2 < cnt;
      const double* xr = sX + n * K;
      const double* xr2 = sX + (two ? n2 : n) * K;
      double e = 0.0, e2 = 0.0;
#pragma unroll
      for (int k = 0; k < HL_KMAX; ++k)
        if (k < K) { e += xr[k] * b[k]; e2 += xr2[k] * b[k]; }
      double sig, sig2;
      const double yy = sY[n], yy2 = two ? sY[n2] : 0.0;
      const double l1 = hl_log1p_exp(e, &sig);
      const double l2 = hl_log1p_exp(e2, &sig2);
      lp += yy * e - l1;
      const double r = yy - sig;
      const double r2 = two ? yy2 - sig2 : 0.0;
      if (two) lp += yy2 * e2 - l2;
#pragma unroll
      for (int k = 0; k < HL_KMAX; ++k)
        if (k < K) gb[k] += r * xr[k] + r2 * xr2[k];
    }
  }
  // combine the four waves' partial sums
#pragma unroll
  for (int k = 0; k < HL_KMAX; ++k)
    if (k < K) sRed[w][lane][k] = gb[k];
  sRed[w][lane][HL_KMAX] = lp;
  __syncthreads();
  if (w == 0 && live) {
    double lpt = 0.0;
    for (int ww = 0; ww < 4; ++ww) lpt += sRed[ww][lane][HL_KMAX];
    const int64_t row_out = a.idx_out ? (int64_t)a.idx[pos] : pos;
    double* gout = a.Gout + row_out * a.ldg;
    double* part = a.work + ((int64_t)g * a.n_rows + pos) * (2 * K + 1);
    double lprior = 0.0;
    for (int k = 0; k < K; ++k) {
      double s = 0.0;
      for (int ww = 0; ww < 4; ++ww) s += sRed[ww][lane][k];
      const double mu = q[k], v = q[K + k];
      const double tau = exp(v);
      const double dd = b[k] - mu;
      const double z = dd / tau;
      // d lp / d beta = likelihood score - (beta - mu) / tau^2
      gout[2 * K + (int64_t)g * K + k] = -(s - dd / (tau * tau));
      part[k] = dd / (tau * tau);                            // d/d mu
      part[K + k] = -1.0 / tau + dd * dd / (tau * tau * tau);  // d/d tau
      lprior += -v - 0.5 * z * z;
    }
    part[2 * K] = lpt + lprior;
  }
}

// hyperparameters, one wave per position: lanes take groups g = lane, lane+64,
// ... and keep the 2K+1 column sums in registers, then combine with butterflies
__global__ void __launch_bounds__(256) hl_hyper_kernel(HlArgs a) {
  const double* d = a.block;
  const int K = (int)d[1], G = (int)d[2];
  const double mu_s = d[3];
  const int n_act = a.n_dev ? min(a.n_rows, *a.n_dev) : a.n_rows;
  const int pos = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (pos >= n_act) return;
  const int64_t row_in = a.idx_in ? (int64_t)a.idx[pos] : pos;
  const int64_t row_out = a.idx_out ? (int64_t)a.idx[pos] : pos;
  const double* q = a.Q + row_in * a.ldq;
  double* gout = a.Gout + row_out * a.ldg;
  const int W = 2 * K + 1;
  double acc[2 * HL_KMAX + 1];
#pragma unroll
  for (int j = 0; j < 2 * HL_KMAX + 1; ++j) acc[j] = 0.0;
  for (int g = lane; g < G; g += NUTS_WAVE) {
    const double* pr = a.work + ((int64_t)g * a.n_rows + pos) * W;
#pragma unroll
    for (int j = 0; j < 2 * HL_KMAX + 1; ++j)
      if (j < W) acc[j] += pr[j];
  }
  double lp = 0.0;
#pragma unroll
  for (int j = 0; j < 2 * HL_KMAX + 1; ++j) {
    if (j >= W) continue;
    const double s = wave_sum(acc[j]);
    if (j < K) {
      const double mu = q[j];
      if (lane == 0) gout[j] = -(s - mu / (mu_s * mu_s));
      lp += -0.5 * mu * mu / (mu_s * mu_s);
    } else if (j < 2 * K) {
      const double v = q[j];
      const double tau = exp(v);
      if (lane == 0) gout[j] = -(a.jac + tau * (s - tau));
      lp += -0.5 * tau * tau + a.jac * v;
    } else {
      lp += s;
    }
  }
  if (lane == 0) a.Vout[row_out] = -lp;
}

void hl_grad_launch(const HlArgs& a, int G, hipStream_t st) {
  const int tiles = (a.n_rows + HL_TC - 1) / HL_TC;
  hipLaunchKernelGGL(hl_group_kernel, dim3(G * tiles), dim3(256), 0, st, a);
  hipLaunchKernelGGL(hl_hyper_kernel, dim3((a.n_rows + 3) / 4), dim3(256), 0, st, a);
}

size_t hl_work_doubles(int G, int K, int n_rows) { return (size_t)G * n_rows * (2 * K + 1); }

}  // namespace nutsgpu
