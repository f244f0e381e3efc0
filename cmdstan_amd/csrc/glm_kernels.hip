// Batched log-density gradient of logistic regression for all chains at once.
//
//   Eta = X · Bᵀ            (N x C; B = positions, C x D, one row per chain)
//   R   = sigmoid(Eta) - y  (never stored in HBM)
//   gV  = Xᵀ R + B / s²     (C x D), V = -(Σ_n y η - log1p(e^η)) + |β|²/(2 s²)
//
// glm_partial_f64: one block owns a chunk of rows and a tile of 16 chains; per
// 64-row sub-tile it stages X in LDS, forms Eta and R for the tile in LDS and
// accumulates the tile's Xᵀ R in registers.  Chunk partials go to a workspace
// that glm_reduce_f64 sums in a fixed order (deterministic, no atomics).
#include "device_common.h"
#include "launch.h"

namespace nutsgpu {

#define GLM_TC 16      // chains per block
#define GLM_TN 64      // rows per sub-tile
#define GLM_DMAX 128   // D capacity of the LDS tiles

__global__ void __launch_bounds__(256) glm_partial_f64(const double* __restrict__ X, const double* __restrict__ y,
                                                        int N, int D, int rows_per_chunk,
                                                        const double* __restrict__ Q, int64_t ldq, int C,
                                                        double* __restrict__ Wg, double* __restrict__ Wlp) {
  __shared__ double sQ[GLM_TC][GLM_DMAX + 1];
  __shared__ double sX[GLM_TN][GLM_DMAX + 1];
  __shared__ double sR[GLM_TN][GLM_TC + 1];
  __shared__ double sLp[16][GLM_TC];
  const int tid = threadIdx.x;
  const int chunk = blockIdx.x;
  const int c0 = blockIdx.y * GLM_TC;
  const int n_begin = chunk * rows_per_chunk;
  const int n_end = min(N, n_begin + rows_per_chunk);

  for (int i = tid; i < GLM_TC * D; i += 256) {
    const int cc = i / D, k = i % D;
    sQ[cc][k] = (c0 + cc < C) ? Q[(int64_t)(c0 + cc) * ldq + k] : 0.0;
  }
  // ownership of the gradient tile: chain gc = tid / 16, k = tid % 16 + 16 j
  const int gc = tid >> 4;
  const int gk0 = tid & 15;
  double gacc[GLM_DMAX / 16];
#pragma unroll
  for (int j = 0; j < GLM_DMAX / 16; ++j) gacc[j] = 0.0;
  double lp_acc = 0.0;  // for (row group, chain) owned in the eta phase
  const int ec = tid & 15;      // eta phase: chain
  const int er0 = tid >> 4;     // rows er0 + 16 i
  __syncthreads();

  for (int nb = n_begin; nb < n_end; nb += GLM_TN) {
    const int nrows = min(GLM_TN, n_end - nb);
    for (int i = tid; i < GLM_TN * D; i += 256) {
      const int r = i / D, k = i % D;
      sX[r][k] = r < nrows ? X[(int64_t)(nb + r) * D + k] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = er0 + 16 * i;
      double eta = 0.0;
      for (int k = 0; k < D; ++k) eta += sX[r][k] * sQ[ec][k];
      double res = 0.0;
      if (r < nrows) {
        const double yy = y[nb + r];
        lp_acc += yy * eta - log1p_exp(eta);
        res = inv_logit(eta) - yy;
      }
      sR[r][ec] = res;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < GLM_DMAX / 16; ++j) {
      const int k = gk0 + 16 * j;
      if (k < D) {
        double a = 0.0;
        for (int r = 0; r < GLM_TN; ++r) a += sR[r][gc] * sX[r][k];
        gacc[j] += a;
      }
    }
    __syncthreads();
  }
  const int cg = c0 + gc;
  if (cg < C) {
#pragma unroll
    for (int j = 0; j < GLM_DMAX / 16; ++j) {
      const int k = gk0 + 16 * j;
      if (k < D) Wg[((int64_t)chunk * C + cg) * D + k] = gacc[j];
    }
  }
  sLp[er0][ec] = lp_acc;
  __syncthreads();
  if (tid < GLM_TC && c0 + tid < C) {
    double s = 0.0;
    for (int i = 0; i < 16; ++i) s += sLp[i][tid];
    Wlp[(int64_t)chunk * C + c0 + tid] = s;
  }
}

__global__ void glm_reduce_f64(const double* __restrict__ Wg, const double* __restrict__ Wlp, int nchunks, int C, int D,
                               double s2inv, const double* __restrict__ Q, int64_t ldq, double* __restrict__ G,
                               double* __restrict__ V) {
  const int c = blockIdx.x;
  const double* q = Q + (int64_t)c * ldq;
  double* g = G + (int64_t)c * ldq;
  for (int k = threadIdx.x; k < D; k += blockDim.x) {
    double s = 0.0;
    for (int ch = 0; ch < nchunks; ++ch) s += Wg[((int64_t)ch * C + c) * D + k];
    g[k] = s + s2inv * q[k];
  }
  if (threadIdx.x < 64) {
    double lp = 0.0, pr = 0.0;
    for (int ch = threadIdx.x; ch < nchunks; ch += 64) lp += Wlp[(int64_t)ch * C + c];
    for (int k = threadIdx.x; k < D; k += 64) pr += q[k] * q[k];
    lp = wave_sum(lp);
    pr = wave_sum(pr);
    if (threadIdx.x == 0) V[c] = -(lp - 0.5 * s2inv * pr);
  }
}

size_t glm_workspace_doubles(int N, int C, int D, int rows_per_chunk) {
  const int nchunks = (N + rows_per_chunk - 1) / rows_per_chunk;
  return (size_t)nchunks * C * D + (size_t)nchunks * C;
}

int glm_grad_f64(const GlmArgs& a, hipStream_t st) {
  if (a.D > GLM_DMAX) return 1;
  const int nchunks = (a.N + a.rows_per_chunk - 1) / a.rows_per_chunk;
  double* Wg = a.work;
  double* Wlp = a.work + (size_t)nchunks * a.C * a.D;
  dim3 grid(nchunks, (a.C + GLM_TC - 1) / GLM_TC);
  hipLaunchKernelGGL(glm_partial_f64, grid, dim3(256), 0, st, a.X, a.y, a.N, a.D, a.rows_per_chunk, a.Q, a.ldq, a.C,
                     Wg, Wlp);
  hipLaunchKernelGGL(glm_reduce_f64, dim3(a.C), dim3(128), 0, st, Wg, Wlp, nchunks, a.C, a.D, a.s2inv, a.Q, a.ldq,
                     a.G, a.V);
  return 0;
}

}  // namespace nutsgpu
