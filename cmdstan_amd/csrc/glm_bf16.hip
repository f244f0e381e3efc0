// Batched logistic-regression gradient on BF16 matrix cores with FP32-level
// accuracy (v_mfma_f32_16x16x32_bf16, FP32 accumulation): every FP32 operand
// is carried as three BF16 terms whose sum holds its 24-bit significand
// (x = x0 + x1 + x2, each term the BF16 rounding of what the earlier terms
// leave), and each product keeps the six term pairs of weight 2^-16 and above
// (x0 q0, x0 q1, x1 q0, x0 q2, x1 q1, x2 q0); the three dropped pairs weigh
// 2^-24 and below, the size of FP32 rounding.  X held in BF16 (the storage
// variant) is a single exact term, and its products keep the three terms of
// the other operand.
//
// Same structure as the FP64 and FP32 kernels: a block owns (row chunk, 64
// chains), four waves of 16 chains; per 64-row tile
//   phase 1  Eta = X Q^T     M = rows, N = chains, K = D in steps of 32
//   pointwise R = sigmoid(Eta) - y and the log-density terms in FP32 (the
//            per-chain sum in FP64)
//   phase 2  G^T += R^T X    M = chains, N = D, K = rows in steps of 32
// Result layout of the 16x16x32 BF16 product (lane l): A supplies row l % 16,
// k = 8 (l / 16) .. + 7; B supplies k = 8 (l / 16) .. + 7 of column l % 16;
// register i of the result is row 4 (l / 16) + i, column l % 16.  Phase 2
// walks the rows of each 32-row step in the order that puts, in lane group g,
// rows 4g..4g+3 and 16+4g..16+4g+3, which are the phase-1 results the lane
// already holds: R stays in registers.  X is staged in LDS twice, row-major
// for phase 1 and transposed for phase 2, so every operand is one or two
// contiguous LDS reads.
#include "device_common.h"
#include "glm_split.h"
#include "launch.h"

namespace nutsgpu {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf4 __attribute__((ext_vector_type(4)));

#define GB_TC 64             // chains per block
#define GB_KMAX 128          // D capacity
#define GB_XS (GB_KMAX + 8)  // row-major LDS stride (bf16): 16-byte aligned rows, bank offset

__device__ __forceinline__ __bf16 bf_rne(float x) {
  const unsigned u = __float_as_uint(x);
  const unsigned r = u + 0x7FFFu + ((u >> 16) & 1u);
  return __builtin_bit_cast(__bf16, (unsigned short)(r >> 16));
}
__device__ __forceinline__ float bf_f32(__bf16 b) {
  return __uint_as_float(((unsigned)__builtin_bit_cast(unsigned short, b)) << 16);
}
// x = t0 + t1 + t2 (+ a remainder of order 2^-24 |x|)
__device__ __forceinline__ void split3(float x, __bf16& t0, __bf16& t1, __bf16& t2) {
  t0 = bf_rne(x);
  const float r1 = x - bf_f32(t0);
  t1 = bf_rne(r1);
  t2 = bf_rne(r1 - bf_f32(t1));
}

__device__ __forceinline__ f4 mma(const bf8& a, const bf8& b, const f4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// X3: X stored as three BF16 terms (FP32 data), else one exact term (BF16
// storage).  TN rows per tile (32 with three terms: two blocks per CU in LDS).
template <bool X3, int KMAX, int TN>
__global__ void __launch_bounds__(256, 2)
glm_mfma_bf16(const unsigned short* __restrict__ X0, const unsigned short* __restrict__ X1,
              const unsigned short* __restrict__ X2, int ldx, const double* __restrict__ y, int N, int D,
              int total_blocks, const double* __restrict__ Q, int64_t ldq, int C, const int* __restrict__ idx,
              const int* __restrict__ n_active, double* __restrict__ Wg, double* __restrict__ Wlp) {
  constexpr int NT = X3 ? 3 : 1;       // terms of X
  constexpr int KS = KMAX / 32;        // phase-1 k steps
  constexpr int DT = KMAX / 16;        // phase-2 output tiles
  constexpr int TS = TN + 8;           // transposed LDS stride (bf16)
  constexpr int RB = TN / 16;          // row blocks of 16
  __shared__ __attribute__((aligned(16))) __bf16 sX[NT][TN * GB_XS];  // [row][k]
  __shared__ __attribute__((aligned(16))) __bf16 sT[NT][KMAX * TS];   // [k][row]
  __shared__ float sY[TN];
  const int tid = threadIdx.x;
  const int w = tid >> 6;
  const int l = tid & 63;
  const int lr = l & 15;   // row / column inside a 16-wide tile
  const int lg = l >> 4;   // lane group: k = 8 lg .. 8 lg + 7

  const int n_act = idx ? *n_active : C;
  const GlmSplit sp = glm_split(n_act, N, total_blocks);
  const int lin = blockIdx.x;
  const int xcd = lin & 7;
  const int j = lin >> 3;
  const int chunk = xcd + 8 * (j / sp.groups);
  const int group = j % sp.groups;
  if (chunk >= sp.chunks || group * GB_TC >= n_act) return;

  const int c0 = group * GB_TC + 16 * w;  // this wave's 16 chains
  const int ksn = (D + 31) >> 5;          // k steps in use
  const int dtn = (D + 15) >> 4;          // output tiles in use

  // phase-1 B operand: Q[c0 + lr][32 ks + 8 lg + e] as three BF16 terms
  bf8 q0[KS], q1[KS], q2[KS];
  {
    const int pos = c0 + lr;
    const int c = pos < n_act ? (idx ? idx[pos] : pos) : -1;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int k = 32 * ks + 8 * lg + e;
        const float v = (c >= 0 && k < D) ? (float)Q[(int64_t)c * ldq + k] : 0.0f;
        __bf16 a, b, d;
        split3(v, a, b, d);
        q0[ks][e] = a; q1[ks][e] = b; q2[ks][e] = d;
      }
    }
  }
  f4 gacc[DT];
#pragma unroll
  for (int t = 0; t < DT; ++t) gacc[t] = f4{0, 0, 0, 0};
  double lp = 0.0;

  const int n_begin = chunk * sp.rows;
  const int n_end = min(N, n_begin + sp.rows);
  const int kp = ksn * 32;  // staged columns (zero past D)

  for (int nb = n_begin; nb < n_end; nb += TN) {
    __syncthreads();
    // stage the tile: 8 bf16 (16 bytes) per load, row-major and transposed
    for (int v = tid; v < TN * (kp / 8); v += 256) {
      const int r = v / (kp / 8), k = (v % (kp / 8)) * 8;
      const int n = nb + r;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const unsigned short* Xt = t == 0 ? X0 : (t == 1 ? X1 : X2);
        uint4 raw = {0, 0, 0, 0};
        if (n < n_end && k < ldx) raw = *reinterpret_cast<const uint4*>(Xt + (int64_t)n * ldx + k);
        *reinterpret_cast<uint4*>(&sX[t][r * GB_XS + k]) = raw;
        const unsigned wv[4] = {raw.x, raw.y, raw.z, raw.w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          sT[t][(k + 2 * e) * TS + r] = __builtin_bit_cast(__bf16, (unsigned short)(wv[e] & 0xffffu));
          sT[t][(k + 2 * e + 1) * TS + r] = __builtin_bit_cast(__bf16, (unsigned short)(wv[e] >> 16));
        }
      }
    }
    if (tid < TN) sY[tid] = (nb + tid < n_end) ? (float)y[nb + tid] : 0.0f;
    __syncthreads();

    // phase 1: Eta for RB row blocks of 16
    f4 eta[RB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) eta[rb] = f4{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      if (ks < ksn) {
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
          const int off = (16 * rb + lr) * GB_XS + 32 * ks + 8 * lg;
          const bf8 a0 = *reinterpret_cast<const bf8*>(&sX[0][off]);
          eta[rb] = mma(a0, q0[ks], eta[rb]);
          eta[rb] = mma(a0, q1[ks], eta[rb]);
          eta[rb] = mma(a0, q2[ks], eta[rb]);
          if constexpr (X3) {
            const bf8 a1 = *reinterpret_cast<const bf8*>(&sX[1][off]);
            const bf8 a2 = *reinterpret_cast<const bf8*>(&sX[2][off]);
            eta[rb] = mma(a1, q0[ks], eta[rb]);
            eta[rb] = mma(a1, q1[ks], eta[rb]);
            eta[rb] = mma(a2, q0[ks], eta[rb]);
          }
        }
      }
    }
    // pointwise: register i of eta[rb] is row 16 rb + 4 lg + i, chain c0 + lr
    bf8 r0[TN / 32], r1[TN / 32], r2[TN / 32];  // phase-2 A operand per 32-row step, three terms
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int rr = 16 * rb + 4 * lg + i;
        const float e = eta[rb][i];
        float r = 0.0f;
        if (nb + rr < n_end) {
          const float yy = sY[rr];
          const float t = __expf(-fabsf(e));
          const float l1p = t < 1e-4f ? t * (1.0f - 0.5f * t) : __logf(1.0f + t);
          lp += (double)(yy * e - ((e > 0.0f ? e : 0.0f) + l1p));
          r = (e >= 0.0f ? 1.0f : t) / (1.0f + t) - yy;
        }
        // step s = rb / 2 holds rows 32 s + {4 lg + i} (rb even) and 32 s + 16 + {4 lg + i} (rb odd)
        const int s = rb >> 1, e8 = (rb & 1) * 4 + i;
        __bf16 a, b, d;
        split3(r, a, b, d);
        r0[s][e8] = a; r1[s][e8] = b; r2[s][e8] = d;
      }
    }
    // phase 2: G^T[chain][d] += sum over rows R[row][chain] X[row][d]
#pragma unroll
    for (int s = 0; s < TN / 32; ++s) {
      const int rowa = 32 * s + 4 * lg, rowb = rowa + 16;
#pragma unroll
      for (int t = 0; t < DT; ++t) {
        if (t < dtn) {
          const int d = 16 * t + lr;
          bf8 b0;
          {
            const bf4 lo = *reinterpret_cast<const bf4*>(&sT[0][d * TS + rowa]);
            const bf4 hi = *reinterpret_cast<const bf4*>(&sT[0][d * TS + rowb]);
            b0 = bf8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
          }
          gacc[t] = mma(r0[s], b0, gacc[t]);
          gacc[t] = mma(r1[s], b0, gacc[t]);
          gacc[t] = mma(r2[s], b0, gacc[t]);
          if constexpr (X3) {
            const bf4 lo1 = *reinterpret_cast<const bf4*>(&sT[1][d * TS + rowa]);
            const bf4 hi1 = *reinterpret_cast<const bf4*>(&sT[1][d * TS + rowb]);
            const bf8 b1 = bf8{lo1[0], lo1[1], lo1[2], lo1[3], hi1[0], hi1[1], hi1[2], hi1[3]};
            const bf4 lo2 = *reinterpret_cast<const bf4*>(&sT[2][d * TS + rowa]);
            const bf4 hi2 = *reinterpret_cast<const bf4*>(&sT[2][d * TS + rowb]);
            const bf8 b2 = bf8{lo2[0], lo2[1], lo2[2], lo2[3], hi2[0], hi2[1], hi2[2], hi2[3]};
            gacc[t] = mma(r0[s], b1, gacc[t]);
            gacc[t] = mma(r1[s], b1, gacc[t]);
            gacc[t] = mma(r0[s], b2, gacc[t]);
          }
        }
      }
    }
  }
  // gacc[t] register i of lane l = G[chain c0 + 4 lg + i][d 16 t + lr]
#pragma unroll
  for (int t = 0; t < DT; ++t) {
    if (t < dtn) {
      const int d = 16 * t + lr;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int pos = c0 + 4 * lg + i;
        if (pos < n_act && d < D) Wg[((int64_t)chunk * sp.stride + pos) * D + d] = (double)gacc[t][i];
      }
    }
  }
  lp += __shfl_xor(lp, 16, 64);
  lp += __shfl_xor(lp, 32, 64);
  if (lg == 0 && c0 + lr < n_act) Wlp[(int64_t)chunk * sp.stride + c0 + lr] = lp;
}

// X (FP64) to three BF16 terms (or one, the BF16 storage), rows padded to ldx
// (a multiple of 8) with zeros
__global__ void glm_split_x(const double* __restrict__ X, int64_t N, int D, int ldx, int nterms,
                            unsigned short* __restrict__ T0, unsigned short* __restrict__ T1,
                            unsigned short* __restrict__ T2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N * ldx) return;
  const int64_t n = i / ldx;
  const int k = (int)(i % ldx);
  const float v = k < D ? (float)X[n * D + k] : 0.0f;
  __bf16 a, b, c;
  split3(v, a, b, c);
  T0[i] = __builtin_bit_cast(unsigned short, a);
  if (nterms == 3) {
    T1[i] = __builtin_bit_cast(unsigned short, b);
    T2[i] = __builtin_bit_cast(unsigned short, c);
  }
}

int glm_bf16_ldx(int D) { return (D + 31) / 32 * 32; }

void glm_bf16_prepare_x(const double* X, int N, int D, int precision, unsigned short* out, hipStream_t st) {
  const int ldx = glm_bf16_ldx(D);
  const int64_t n = (int64_t)N * ldx;
  const int nt = precision == 1 ? 3 : 1;
  hipLaunchKernelGGL(glm_split_x, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, X, (int64_t)N, D, ldx, nt, out,
                     out + n, out + 2 * n);
}

int glm_grad_mfma_bf16(const GlmArgs& a, const GlmPlan& p, const unsigned short* Xs, int precision, hipStream_t st) {
  if (a.D > GB_KMAX || a.D < 1) return 1;
  double* Wg = a.work;
  double* Wlp = a.work + (size_t)p.total_blocks * GB_TC * a.D;
  const int ldx = glm_bf16_ldx(a.D);
  const int64_t n = (int64_t)a.N * ldx;
  const unsigned short* X0 = Xs;
  const unsigned short* X1 = precision == 1 ? Xs + n : Xs;
  const unsigned short* X2 = precision == 1 ? Xs + 2 * n : Xs;
#define GB_LAUNCH(X3, K)                                                                                     \
  hipLaunchKernelGGL((glm_mfma_bf16<X3, K, X3 ? 32 : 64>), dim3(p.total_blocks), dim3(256), 0, st, X0, X1, X2, \
                     ldx, a.y, a.N, a.D, p.total_blocks, a.Q, a.ldq, a.C, a.idx, a.n_active, Wg, Wlp)
#define GB_BUCKETS(X3)                    \
  if (a.D <= 32) GB_LAUNCH(X3, 32);       \
  else if (a.D <= 64) GB_LAUNCH(X3, 64);  \
  else if (a.D <= 96) GB_LAUNCH(X3, 96);  \
  else GB_LAUNCH(X3, 128);
  if (a.ev_begin) (void)hipEventRecord(a.ev_begin, st);
  if (precision == 1) { GB_BUCKETS(true) } else { GB_BUCKETS(false) }
  if (a.ev_end) (void)hipEventRecord(a.ev_end, st);
#undef GB_BUCKETS
#undef GB_LAUNCH
  glm_reduce_launch(a, p, st);
  return 0;
}

}  // namespace nutsgpu
