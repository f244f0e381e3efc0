// Large-D logistic-regression gradient on BF16 matrix cores for X stored in
// BF16 (v_mfma_f32_16x16x32_bf16, FP32 accumulation).  X is exact in BF16;
// the other operand of each product (chain positions in pass 1, residuals in
// pass 2) is carried as three BF16 terms whose sum holds its 24-bit
// significand (x = x0 + x1 + x2, each term the BF16 rounding of what the
// earlier ones leave), so every product keeps FP32-level accuracy:
//   pass 1  Eta = X . Q^T       M = rows, N = chains, K = D; fused epilogue
//           R = sigmoid(Eta) - y split into three BF16 terms, stored per
//           chain along the rows; lp terms summed per chain in FP64
//   pass 2  G^T = R^T . X       M = chains, N = D, K = rows (split in ranges)
//   reduce  the FP64 row-range sums of the FP32 path (gl_reduce)
// Both products read operands whose K index is contiguous in memory: X row
// major for pass 1, and for pass 2 a transposed copy Xt [D][rows] made once
// per model (HBM is large; the transpose is paid once, not per gradient).
// One 128x128 block tile (four 64x64 waves of 4x4 16x16 products), K in
// slabs of 64 staged through LDS with 16-byte loads and stores; the next
// slab's global loads are in flight while the current slab is multiplied.
// Result layout of the 16x16x32 BF16 product (lane l): A supplies row l % 16,
// k = 8 (l / 16) .. + 7; B supplies column l % 16 over the same k; register i
// of the result is row 4 (l / 16) + i, column l % 16.
#include "device_common.h"
#include "glm_large.h"

namespace nutsgpu {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

#define GBL_BM 128
#define GBL_BK 64
#define GBL_LS (GBL_BK + 8)  // LDS row stride in BF16 elements (144 bytes: 16-byte rows on distinct banks)

__device__ __forceinline__ unsigned short gbl_rne(float x) {
  const unsigned u = __float_as_uint(x);
  return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float gbl_f32(unsigned short b) { return __uint_as_float(((unsigned)b) << 16); }

// x = t0 + t1 + t2 up to 2^-24 |x|; the residues are exact in FP32
__device__ __forceinline__ void gbl_split3(float x, unsigned short& t0, unsigned short& t1, unsigned short& t2) {
  t0 = gbl_rne(x);
  const float r1 = x - gbl_f32(t0);
  t1 = gbl_rne(r1);
  t2 = gbl_rne(r1 - gbl_f32(t1));
}
// the same from FP64 (residues exact in FP64)
__device__ __forceinline__ void gbl_split3(double x, unsigned short& t0, unsigned short& t1, unsigned short& t2) {
  t0 = gbl_rne((float)x);
  const double r1 = x - (double)gbl_f32(t0);
  t1 = gbl_rne((float)r1);
  t2 = gbl_rne((float)(r1 - (double)gbl_f32(t1)));
}

// one operand: T terms of GBL_BM-row tiles, K contiguous; rows at or past
// `rows` and k at or past `kmax` (a multiple of 8) read as zero
struct GblOp {
  const unsigned short* p;
  int64_t ld;     // row stride (elements)
  int64_t plane;  // term stride (elements)
  int rows;
  int kmax;
};

// 16-byte chunks of a T-term 128 x 64 tile: chunk c = term c / 1024, row
// (c % 1024) / 8, k chunk c % 8 (eight threads read one row's 128 bytes)
template <int T>
__device__ __forceinline__ void gbl_fetch(const GblOp& o, int r0, int k0, uint4 (&v)[4 * T]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int u = 0; u < 4 * T; ++u) {
    const int c = tid + 256 * u;
    const int t = c >> 10, row = (c & 1023) >> 3, kc = c & 7;
    const int r = r0 + row, k = k0 + 8 * kc;
    if (r < o.rows && k < o.kmax)
      v[u] = *reinterpret_cast<const uint4*>(o.p + t * o.plane + (int64_t)r * o.ld + k);
    else
      v[u] = make_uint4(0u, 0u, 0u, 0u);
  }
}
template <int T>
__device__ __forceinline__ void gbl_stash(__bf16* s, const uint4 (&v)[4 * T]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int u = 0; u < 4 * T; ++u) {
    const int c = tid + 256 * u;
    const int t = c >> 10, row = (c & 1023) >> 3, kc = c & 7;
    *reinterpret_cast<uint4*>(s + (t * GBL_BM + row) * GBL_LS + 8 * kc) = v[u];
  }
}

// acc[i][j] += sum over the staged slab of A rows (wr + 16 i ..) times B rows
// (wc + 16 j ..), every term pair, smallest terms first
template <int TA, int TB>
__device__ __forceinline__ void gbl_mma(const __bf16* sA, const __bf16* sB, f4 (&acc)[4][4]) {
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, lr = l & 15, lg = l >> 4;
  const int wr = (w >> 1) * 64, wc = (w & 1) * 64;
#pragma unroll
  for (int kk = 0; kk < GBL_BK / 32; ++kk) {
    bf8 a[TA][4], b[TB][4];
#pragma unroll
    for (int t = 0; t < TA; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        a[t][i] = *reinterpret_cast<const bf8*>(sA + (t * GBL_BM + wr + 16 * i + lr) * GBL_LS + 32 * kk + 8 * lg);
#pragma unroll
    for (int t = 0; t < TB; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        b[t][j] = *reinterpret_cast<const bf8*>(sB + (t * GBL_BM + wc + 16 * j + lr) * GBL_LS + 32 * kk + 8 * lg);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int ta = TA - 1; ta >= 0; --ta)
#pragma unroll
          for (int tb = TB - 1; tb >= 0; --tb)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ta][i], b[tb][j], acc[i][j], 0, 0, 0);
  }
}

// C tile (m0, n0) over k in [kb, ke) (ke - kb a multiple of GBL_BK)
template <int TA, int TB>
__device__ __forceinline__ void gbl_gemm(const GblOp& A, const GblOp& B, int m0, int n0, int kb, int ke,
                                         f4 (&acc)[4][4], __bf16* sA, __bf16* sB) {
  if (kb >= ke) return;
  uint4 va[4 * TA], vb[4 * TB];
  gbl_fetch<TA>(A, m0, kb, va);
  gbl_fetch<TB>(B, n0, kb, vb);
  for (int k0 = kb; k0 < ke; k0 += GBL_BK) {
    gbl_stash<TA>(sA, va);
    gbl_stash<TB>(sB, vb);
    __syncthreads();
    if (k0 + GBL_BK < ke) {
      gbl_fetch<TA>(A, m0, k0 + GBL_BK, va);
      gbl_fetch<TB>(B, n0, k0 + GBL_BK, vb);
    }
    gbl_mma<TA, TB>(sA, sB, acc);
    __syncthreads();
  }
}

// pass 1: block (row tile, chain tile)
__global__ void __launch_bounds__(256, 2) gbl_eta_kernel(GlLargeArgs a, GlLargePlan p,
                                                         const unsigned short* __restrict__ X) {
  __shared__ __attribute__((aligned(16))) __bf16 sA[1 * GBL_BM * GBL_LS];
  __shared__ __attribute__((aligned(16))) __bf16 sB[3 * GBL_BM * GBL_LS];
  __shared__ double sLp[2][GBL_BM];
  const int n_act = a.n_active ? *a.n_active : a.C;
  const int m0 = blockIdx.x * GBL_BM, n0 = blockIdx.y * GBL_BM;
  if (n0 >= n_act) return;
  const GblOp A{X, a.D, 0, a.N, a.D};
  const GblOp B{a.Qb, p.Dk, (int64_t)p.Cp * p.Dk, p.Cp, p.Dk};
  f4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4{0, 0, 0, 0};
  gbl_gemm<1, 3>(A, B, m0, n0, 0, p.Dk, acc, sA, sB);

  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, lr = l & 15, lg = l >> 4;
  const int wr = (w >> 1) * 64, wc = (w & 1) * 64;
  const int64_t plane = (int64_t)p.Cp * p.ldn;
  double lpc[4] = {0, 0, 0, 0};
  // this lane's sixteen rows' outcomes, read once (not once per chain tile)
  float yv[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = m0 + wr + 16 * i + 4 * lg + r;
      yv[i][r] = m < a.N ? (float)a.y[m] : 0.0f;
    }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int n = n0 + wc + 16 * j + lr;  // chain position
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int mb = m0 + wr + 16 * i + 4 * lg;  // first of this lane's four rows
      unsigned short t0[4], t1[4], t2[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = mb + r;
        float rr = 0.0f;
        if (m < a.N && n < n_act) {
          const float e = acc[i][j][r];
          const float yy = yv[i][r];
          const float te = __expf(-fabsf(e));
          const float l1p = te < 1e-4f ? te * (1.0f - 0.5f * te) : __logf(1.0f + te);
          lpc[j] += (double)(yy * e - ((e > 0.0f ? e : 0.0f) + l1p));
          // sigmoid(e) - y with the hardware reciprocal of 1 + t in [1, 2]
          rr = (e >= 0.0f ? 1.0f : te) * __builtin_amdgcn_rcpf(1.0f + te) - yy;
        }
        gbl_split3(rr, t0[r], t1[r], t2[r]);
      }
      // rows mb..mb+3 of chain n: 8 bytes per term (rows past N hold zeros)
      unsigned short* dst = a.Rb + (int64_t)n * p.ldn + mb;
      *reinterpret_cast<uint2*>(dst) = make_uint2(t0[0] | ((unsigned)t0[1] << 16), t0[2] | ((unsigned)t0[3] << 16));
      *reinterpret_cast<uint2*>(dst + plane) =
          make_uint2(t1[0] | ((unsigned)t1[1] << 16), t1[2] | ((unsigned)t1[3] << 16));
      *reinterpret_cast<uint2*>(dst + 2 * plane) =
          make_uint2(t2[0] | ((unsigned)t2[1] << 16), t2[2] | ((unsigned)t2[3] << 16));
    }
  }
  // lanes l, l^16, l^32, l^48 hold the same chain
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    lpc[j] += __shfl_xor(lpc[j], 16, 64);
    lpc[j] += __shfl_xor(lpc[j], 32, 64);
  }
  if (l < 16) {
#pragma unroll
    for (int j = 0; j < 4; ++j) sLp[w >> 1][wc + 16 * j + l] = lpc[j];
  }
  __syncthreads();
  if (tid < GBL_BM) {
    const int n = n0 + tid;
    if (n < n_act) a.Wlp[(int64_t)blockIdx.x * a.C + n] = sLp[0][tid] + sLp[1][tid];
  }
}

// pass 2: block (chain tile, coefficient tile, row range)
__global__ void __launch_bounds__(256, 2) gbl_grad_kernel(GlLargeArgs a, GlLargePlan p) {
  __shared__ __attribute__((aligned(16))) __bf16 sA[3 * GBL_BM * GBL_LS];
  __shared__ __attribute__((aligned(16))) __bf16 sB[1 * GBL_BM * GBL_LS];
  const int n_act = a.n_active ? *a.n_active : a.C;
  const int m0 = blockIdx.x * GBL_BM, n0 = blockIdx.y * GBL_BM;
  if (m0 >= n_act) return;
  const int sp = blockIdx.z;
  const int kb = sp * p.rows_per_split;
  const int ke = min((int)p.ldn, kb + p.rows_per_split);
  const GblOp A{a.Rb, p.ldn, (int64_t)p.Cp * p.ldn, p.Cp, (int)p.ldn};
  const GblOp B{a.Xt, p.ldn, 0, p.Dp, (int)p.ldn};
  f4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4{0, 0, 0, 0};
  gbl_gemm<3, 1>(A, B, m0, n0, kb, ke, acc, sA, sB);
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  const int wr = (w >> 1) * 64, wc = (w & 1) * 64;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wr + 16 * i + 4 * (l >> 4) + r;  // chain position
        const int n = n0 + wc + 16 * j + (l & 15);           // coefficient
        if (m < n_act && n < a.D) a.Wg[((int64_t)sp * a.C + m) * a.D + n] = (double)acc[i][j][r];
      }
}

// active chains' positions as three BF16 terms, [term][pos][Dk], zero past D
__global__ void gbl_pack_q(GlLargeArgs a, GlLargePlan p) {
  const int n_act = a.n_active ? *a.n_active : a.C;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int pos = (int)(i / p.Dk), k = (int)(i % p.Dk);
  if (pos >= n_act) return;
  const int c = a.idx ? a.idx[pos] : pos;
  unsigned short t0 = 0, t1 = 0, t2 = 0;
  if (k < a.D) gbl_split3(a.Q[(int64_t)c * a.ldq + k], t0, t1, t2);
  const int64_t plane = (int64_t)p.Cp * p.Dk;
  a.Qb[i] = t0;
  a.Qb[i + plane] = t1;
  a.Qb[i + 2 * plane] = t2;
}

// Xt [Dp][ldn] from the BF16 X [N][D], zero padded (64 x 64 tiles through LDS)
__global__ void __launch_bounds__(256) gbl_transpose_x(const unsigned short* __restrict__ X, int N, int D,
                                                        unsigned short* __restrict__ Xt, int64_t ldn) {
  __shared__ unsigned short t[64][66];
  const int n0 = blockIdx.x * 64, d0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int r = ty; r < 64; r += 4) {
    const int n = n0 + r, d = d0 + tx;
    t[r][tx] = (n < N && d < D) ? X[(int64_t)n * D + d] : (unsigned short)0;
  }
  __syncthreads();
  for (int r = ty; r < 64; r += 4) Xt[(int64_t)(d0 + r) * ldn + n0 + tx] = t[tx][r];
}

bool gl_large_bf16_ok(int D) { return D >= 8 && D % 8 == 0; }

void gl_large_prepare_xt(const unsigned short* Xb, int N, int D, unsigned short* Xt, const GlLargePlan& p,
                         hipStream_t st) {
  const dim3 grid((unsigned)(p.ldn / 64), (unsigned)(p.Dp / 64));
  hipLaunchKernelGGL(gbl_transpose_x, grid, dim3(256), 0, st, Xb, N, D, Xt, p.ldn);
}

void gl_large_grad_bf16(const GlLargeArgs& a, const GlLargePlan& p, const unsigned short* X, hipStream_t st) {
  hipLaunchKernelGGL(gbl_pack_q, dim3((unsigned)(((int64_t)a.C * p.Dk + 255) / 256)), dim3(256), 0, st, a, p);
  if (a.ev_begin) (void)hipEventRecord(a.ev_begin, st);
  const dim3 g1((unsigned)(p.ldn / GBL_BM), (unsigned)(p.Cp / GBL_BM));
  const dim3 g2((unsigned)(p.Cp / GBL_BM), (unsigned)(p.Dp / GBL_BM), (unsigned)p.splits);
  hipLaunchKernelGGL(gbl_eta_kernel, g1, dim3(256), 0, st, a, p, X);
  hipLaunchKernelGGL(gbl_grad_kernel, g2, dim3(256), 0, st, a, p);
  if (a.ev_end) (void)hipEventRecord(a.ev_end, st);
  gl_reduce_launch(a, p, st);
}

}  // namespace nutsgpu
