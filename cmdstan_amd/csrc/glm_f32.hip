// Batched logistic-regression gradient with the contraction on FP32 matrix
// cores (v_mfma_f32_16x16x4f32, FP32 accumulation) and X stored in FP32 or
// BF16 (converted to FP32 on its way into LDS).  The per-observation terms are
// formed in FP32 from the FP32 linear predictor and the per-chain log density
// is summed in FP64; chunk partials of the gradient are reduced in FP64, so V
// and ∇V reach the FP64 sampler with an error set by FP32 rounding of the
// products and terms (stated against FP64 in the tests).
//
// Same structure as the FP64 kernel: a block owns (row chunk, 64 chains), four
// waves of 16 chains; phase 1 Eta = X·Qᵀ, pointwise R, phase 2 Gᵀ += Rᵀ·X with
// the K order of phase 2 permuted so its A operand is the phase-1 accumulator.
#include <hip/hip_bf16.h>

#include "device_common.h"
#include "glm_split.h"
#include "launch.h"

namespace nutsgpu {

typedef float f4 __attribute__((ext_vector_type(4)));

#define GF_TN 64
#define GF_TC 64
#define GF_KMAX 128
#define GF_XS (GF_KMAX + 4)  // LDS row stride in floats

// Row (inside a 16-row block) held by register i of lane l in a 16x16 f32 result.
template <int LAYOUT>
__device__ __forceinline__ int f32_row(int l, int i) {
  return LAYOUT == 0 ? 4 * (l >> 4) + i : 4 * i + (l >> 4);
}

__device__ __forceinline__ float to_f32(float x) { return x; }
__device__ __forceinline__ float to_f32(unsigned short b) { return __uint_as_float(((unsigned)b) << 16); }

template <typename XT, int LAYOUT, int KMAX>
__global__ void __launch_bounds__(256)
glm_mfma_f32(const XT* __restrict__ X, const double* __restrict__ y, int N, int D, int total_blocks,
             const double* __restrict__ Q, int64_t ldq, int C, const int* __restrict__ idx,
             const int* __restrict__ n_active, double* __restrict__ Wg, double* __restrict__ Wlp) {
  __shared__ float sX[GF_TN * GF_XS];
  __shared__ double sY[GF_TN];
  const int tid = threadIdx.x;
  const int w = tid >> 6;
  const int l = tid & 63;
  const int lr = l & 15;
  const int lk = l >> 4;

  const int n_act = idx ? *n_active : C;
  const GlmSplit sp = glm_split(n_act, N, total_blocks);
  const int lin = blockIdx.x;
  const int xcd = lin & 7;
  const int j = lin >> 3;
  const int chunk = xcd + 8 * (j / sp.groups);
  const int group = j % sp.groups;
  if (chunk >= sp.chunks || group * GF_TC >= n_act) return;

  const int c0 = group * GF_TC + 16 * w;
  const int KS = (D + 3) >> 2;
  const int DT = (D + 15) >> 4;

  float qreg[KMAX / 4];
  {
    const int pos = c0 + lr;
    const int c = pos < n_act ? (idx ? idx[pos] : pos) : -1;
#pragma unroll
    for (int ks = 0; ks < KMAX / 4; ++ks) {
      const int k = 4 * ks + lk;
      qreg[ks] = (ks < KS && c >= 0 && k < D) ? (float)Q[(int64_t)c * ldq + k] : 0.0f;
    }
  }
  f4 gacc[KMAX / 16];
#pragma unroll
  for (int t = 0; t < KMAX / 16; ++t) gacc[t] = f4{0, 0, 0, 0};
  double lp = 0.0;

  const int n_begin = chunk * sp.rows;
  const int n_end = min(N, n_begin + sp.rows);
  const int ncols = DT * 16;

  for (int nb = n_begin; nb < n_end; nb += GF_TN) {
    __syncthreads();
#pragma unroll
    for (int m = 0; m < GF_TN / 16; ++m) {
      const int r = (tid >> 4) + 16 * m;
      const int n = nb + r;
      const bool rowok = n < n_end;
      const XT* xr = X + (int64_t)n * D;
#pragma unroll
      for (int jj = 0; jj < KMAX / 16; ++jj) {
        const int k = (tid & 15) + 16 * jj;
        if (k < ncols) sX[r * GF_XS + k] = (rowok && k < D) ? to_f32(xr[k]) : 0.0f;
      }
    }
    if (tid < GF_TN) sY[tid] = (nb + tid < n_end) ? y[nb + tid] : 0.0;
    __syncthreads();
    f4 eta[4];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) eta[rb] = f4{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < KMAX / 4; ++ks) {
      if (ks < KS) {
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
          const float a = sX[(16 * rb + lr) * GF_XS + 4 * ks + lk];
          eta[rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, qreg[ks], eta[rb], 0, 0, 0);
        }
      }
    }
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int rr = 16 * rb + f32_row<LAYOUT>(l, i);
        const float e = eta[rb][i];
        float r = 0.0f;
        if (nb + rr < n_end) {
          // terms in FP32 from the FP32 linear predictor, summed in FP64
          const float yy = (float)sY[rr];
          const float t = __expf(-fabsf(e));
          const float l1p = t < 1e-4f ? t * (1.0f - 0.5f * t) : __logf(1.0f + t);
          lp += (double)(yy * e - ((e > 0.0f ? e : 0.0f) + l1p));
          r = (e >= 0.0f ? 1.0f : t) / (1.0f + t) - yy;
        }
        eta[rb][i] = r;
      }
    }
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float a = eta[rb][i];
        const float* xrow = sX + (16 * rb + f32_row<LAYOUT>(l, i)) * GF_XS + lr;
#pragma unroll
        for (int t = 0; t < KMAX / 16; ++t)
          if (t < DT) gacc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, xrow[16 * t], gacc[t], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < KMAX / 16; ++t) {
    if (t < DT) {
      const int d = 16 * t + lr;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int pos = c0 + f32_row<LAYOUT>(l, i);
        if (pos < n_act && d < D) Wg[((int64_t)chunk * sp.stride + pos) * D + d] = (double)gacc[t][i];
      }
    }
  }
  lp += __shfl_xor(lp, 16, 64);
  lp += __shfl_xor(lp, 32, 64);
  if (lk == 0 && c0 + lr < n_act) Wlp[(int64_t)chunk * sp.stride + c0 + lr] = lp;
}

__global__ void glm_to_f32(const double* __restrict__ X, float* __restrict__ Y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) Y[i] = (float)X[i];
}
// round to nearest even bf16
__global__ void glm_to_bf16(const double* __restrict__ X, unsigned short* __restrict__ Y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const unsigned u = __float_as_uint((float)X[i]);
    const unsigned r = u + 0x7FFFu + ((u >> 16) & 1u);
    Y[i] = (unsigned short)(r >> 16);
  }
}

void glm_convert_x(const double* X, void* out, int64_t n, int precision, hipStream_t st) {
  const dim3 grid((unsigned)((n + 255) / 256)), blk(256);
  if (precision == 1) hipLaunchKernelGGL(glm_to_f32, grid, blk, 0, st, X, (float*)out, n);
  else hipLaunchKernelGGL(glm_to_bf16, grid, blk, 0, st, X, (unsigned short*)out, n);
}

// The reduce kernel of the FP64 path sums the chunk partials in FP64.
int glm_grad_mfma_f32(const GlmArgs& a, const GlmPlan& p, const void* Xlow, int precision, int layout,
                      hipStream_t st) {
  if (a.D > GF_KMAX || a.D < 1) return 1;
  double* Wg = a.work;
  double* Wlp = a.work + (size_t)p.total_blocks * GF_TC * a.D;
  (void)layout;  // the 16x16 f32 result layout measured on gfx950 is layout 0
#define GF_LAUNCH(XT, K)                                                                                         \
  hipLaunchKernelGGL((glm_mfma_f32<XT, 0, K>), dim3(p.total_blocks), dim3(256), 0, st, (const XT*)Xlow, a.y, a.N,   \
                     a.D, p.total_blocks, a.Q, a.ldq, a.C, a.idx, a.n_active, Wg, Wlp)
#define GF_BUCKETS(XT)                          \
  if (a.D <= 16) GF_LAUNCH(XT, 16);             \
  else if (a.D <= 32) GF_LAUNCH(XT, 32);        \
  else if (a.D <= 64) GF_LAUNCH(XT, 64);        \
  else if (a.D <= 112) GF_LAUNCH(XT, 112);      \
  else GF_LAUNCH(XT, 128);
  if (a.ev_begin) (void)hipEventRecord(a.ev_begin, st);
  if (precision == 1) { GF_BUCKETS(float) } else { GF_BUCKETS(unsigned short) }
  if (a.ev_end) (void)hipEventRecord(a.ev_end, st);
#undef GF_BUCKETS
#undef GF_LAUNCH
  glm_reduce_launch(a, p, st);
  return 0;
}

}  // namespace nutsgpu
