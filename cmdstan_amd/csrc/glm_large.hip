// Batched logistic-regression gradient for large D (up to thousands of
// coefficients) and large N, on FP32 matrix cores (v_mfma_f32_16x16x4f32),
// with X stored in FP32 or BF16:
//   pass 1  Eta = X . Qf^T     (rows x active chains), fused epilogue:
//           R = sigmoid(Eta) - y stored in FP32, lp terms summed per chain in FP64
//   pass 2  G^T = R^T . X       (active chains x D), split over row ranges
//   reduce  FP64 sum of the row-range partials + prior; V = -(lp - prior)
// Both passes use one 128x128x32 tile kernel (4 waves of 64x64) with the
// operand strides as parameters; BF16 elements are widened to FP32 on their
// way into LDS.
#include "device_common.h"
#include "glm_large.h"

namespace nutsgpu {

typedef float f4 __attribute__((ext_vector_type(4)));

#define GL_BM 128
#define GL_BN 128
#define GL_BK 32

__device__ __forceinline__ float gl_f32(float x) { return x; }
__device__ __forceinline__ float gl_f32(unsigned short b) { return __uint_as_float(((unsigned)b) << 16); }

// eight consecutive elements from a 16-byte aligned address: one 16-byte load
// (BF16) or two (FP32), widened to FP32
__device__ __forceinline__ void gl_load8(const float* p, float (&v)[8]) {
  const float4 a = *reinterpret_cast<const float4*>(p);
  const float4 b = *reinterpret_cast<const float4*>(p + 4);
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}
__device__ __forceinline__ void gl_load8(const unsigned short* p, float (&v)[8]) {
  const uint4 a = *reinterpret_cast<const uint4*>(p);
  const unsigned w[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    v[2 * i] = __uint_as_float(w[i] << 16);
    v[2 * i + 1] = __uint_as_float(w[i] & 0xffff0000u);
  }
}
template <typename T>
__device__ __forceinline__ bool gl_aligned16(const T* p) {
  return (reinterpret_cast<uintptr_t>(p) & 15) == 0;
}

struct GlTile {
  int M, N, K;
  int64_t sam, sak, sbk, sbn;
};

// Accumulates the tile (m0, n0) over k in [kb, ke) into acc (wave layout:
// wave (wr, wc) owns rows wr*64.., cols wc*64.., 4x4 MFMA tiles).
template <typename AT, typename BT>
__device__ __forceinline__ void gl_tile(const AT* __restrict__ A, const BT* __restrict__ B, const GlTile& t, int m0,
                                        int n0, int kb, int ke, f4 (&acc)[4][4], float (*As)[GL_BM + 4],
                                        float (*Bs)[GL_BN + 4]) {
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  const int wr = (w >> 1) * 64, wc = (w & 1) * 64;
  const bool a_k = t.sak == 1;
  const bool b_n = t.sbn == 1;
  for (int k0 = kb; k0 < ke; k0 += GL_BK) {
    // A tile 128 x GL_BK in 128 x 16 slabs: 8 elements per thread per slab,
    // contiguous in memory along k (a_k) or along m; one vector load when all
    // 8 are in range and aligned
#pragma unroll
    for (int slab = 0; slab < GL_BK / 16; ++slab) {
      const int mm = a_k ? (tid >> 1) : (tid & 15) * 8, kk = 16 * slab + (a_k ? (tid & 1) * 8 : (tid >> 4));
      const int m = m0 + mm, k = k0 + kk;
      const AT* src = A + (int64_t)m * t.sam + (int64_t)k * t.sak;
      const bool full = a_k ? (m < t.M && k + 7 < ke) : (m + 7 < t.M && k < ke);
      float v[8];
      if (full && gl_aligned16(src)) {
        gl_load8(src, v);
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int me = a_k ? m : m + e, ke_ = a_k ? k + e : k;
          v[e] = (me < t.M && ke_ < ke) ? gl_f32(A[(int64_t)me * t.sam + (int64_t)ke_ * t.sak]) : 0.0f;
        }
      }
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        if (a_k) As[kk + e][mm] = v[e];
        else As[kk][mm + e] = v[e];
      }
    }
#pragma unroll
    for (int slab = 0; slab < GL_BK / 16; ++slab) {
      const int nn = b_n ? (tid & 15) * 8 : (tid >> 1), kk = 16 * slab + (b_n ? (tid >> 4) : (tid & 1) * 8);
      const int n = n0 + nn, k = k0 + kk;
      const BT* src = B + (int64_t)k * t.sbk + (int64_t)n * t.sbn;
      const bool full = b_n ? (n + 7 < t.N && k < ke) : (n < t.N && k + 7 < ke);
      float v[8];
      if (full && gl_aligned16(src)) {
        gl_load8(src, v);
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int ne = b_n ? n + e : n, ke_ = b_n ? k : k + e;
          v[e] = (ne < t.N && ke_ < ke) ? gl_f32(B[(int64_t)ke_ * t.sbk + (int64_t)ne * t.sbn]) : 0.0f;
        }
      }
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        if (b_n) Bs[kk][nn + e] = v[e];
        else Bs[kk + e][nn] = v[e];
      }
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < GL_BK / 4; ++ks) {
      const int kr = ks * 4 + (l >> 4);
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kr][wr + 16 * i + (l & 15)];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kr][wc + 16 * j + (l & 15)];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
}

// pass 1: rows x chains; block (row tile, chain tile); grid.x = row tiles, grid.y = chain tiles
template <typename XT>
__global__ void __launch_bounds__(256) gl_eta_kernel(GlLargeArgs a, const XT* __restrict__ X) {
  __shared__ float As[GL_BK][GL_BM + 4];
  __shared__ float Bs[GL_BK][GL_BN + 4];
  __shared__ double sLp[2][GL_BN];
  const int n_act = a.n_active ? *a.n_active : a.C;
  const int m0 = blockIdx.x * GL_BM, n0 = blockIdx.y * GL_BN;
  if (n0 >= n_act) return;
  GlTile t{a.N, n_act, a.D, a.D, 1, 1, a.Dq};
  f4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4{0, 0, 0, 0};
  gl_tile<XT, float>(X, a.Qf, t, m0, n0, 0, a.D, acc, As, Bs);
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  const int wr = (w >> 1) * 64, wc = (w & 1) * 64;
  double lpc[4] = {0, 0, 0, 0};  // per n-tile j, this lane's column
  // this lane's sixteen rows' outcomes, read once (not once per chain tile)
  float yv[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = m0 + wr + 16 * i + 4 * (l >> 4) + r;
      yv[i][r] = m < a.N ? (float)a.y[m] : 0.0f;
    }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wr + 16 * i + 4 * (l >> 4) + r;
        const int n = n0 + wc + 16 * j + (l & 15);
        if (m < a.N && n < n_act) {
          const float e = acc[i][j][r];
          const float yy = yv[i][r];
          const float te = __expf(-fabsf(e));
          const float l1p = te < 1e-4f ? te * (1.0f - 0.5f * te) : __logf(1.0f + te);
          lpc[j] += (double)(yy * e - ((e > 0.0f ? e : 0.0f) + l1p));
          // sigmoid(e) - y with the hardware reciprocal of 1 + t in [1, 2]
          a.R[(int64_t)m * a.ldr + n] = (e >= 0.0f ? 1.0f : te) * __builtin_amdgcn_rcpf(1.0f + te) - yy;
        }
      }
  // lanes l, l^16, l^32, l^48 hold the same column
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
  if (tid < GL_BN) {
    const int n = n0 + tid;
    if (n < n_act) a.Wlp[(int64_t)blockIdx.x * a.C + n] = sLp[0][tid] + sLp[1][tid];
  }
}

// pass 2: chains x D, K = rows split into grid.z ranges; FP32 tile sums to FP64 partials
template <typename XT>
__global__ void __launch_bounds__(256) gl_grad_kernel(GlLargeArgs a, const XT* __restrict__ X, int rows_per_split) {
  __shared__ float As[GL_BK][GL_BM + 4];
  __shared__ float Bs[GL_BK][GL_BN + 4];
  const int n_act = a.n_active ? *a.n_active : a.C;
  const int m0 = blockIdx.x * GL_BM, n0 = blockIdx.y * GL_BN;
  if (m0 >= n_act) return;
  const int sp = blockIdx.z;
  const int kb = sp * rows_per_split;
  const int ke = min(a.N, kb + rows_per_split);
  GlTile t{n_act, a.D, a.N, 1, a.ldr, a.D, 1};
  f4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4{0, 0, 0, 0};
  if (kb < ke) gl_tile<float, XT>(a.R, X, t, m0, n0, kb, ke, acc, As, Bs);
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

// active chains' positions to FP32, [pos][Dq] zero padded
__global__ void gl_pack_q(GlLargeArgs a) {
  const int n_act = a.n_active ? *a.n_active : a.C;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int pos = (int)(i / a.Dq), k = (int)(i % a.Dq);
  if (pos >= n_act) return;
  const int c = a.idx ? a.idx[pos] : pos;
  a.Qf[i] = k < a.D ? (float)a.Q[(int64_t)c * a.ldq + k] : 0.0f;
}

// FP64: G = -(sum of partials) + prior, V = -(lp - prior)
__global__ void gl_reduce(GlLargeArgs a, int splits, int row_tiles) {
  const int pos = blockIdx.x;
  const int n_act = a.n_active ? *a.n_active : a.C;
  if (pos >= n_act) return;
  const int c = a.idx ? a.idx[pos] : pos;
  const double* q = a.Q + (int64_t)c * a.ldq;
  double* g = a.G + (int64_t)c * a.ldq;
  for (int k = threadIdx.x; k < a.D; k += blockDim.x) {
    double s = 0.0;
    // split order kept; eight loads in flight instead of one
#pragma unroll 8
    for (int sp = 0; sp < splits; ++sp) s += a.Wg[((int64_t)sp * a.C + pos) * a.D + k];
    g[k] = s + a.s2inv * q[k];
  }
  if (threadIdx.x < 64) {
    double lpv = 0.0, pr = 0.0;
    for (int rt = threadIdx.x; rt < row_tiles; rt += 64) lpv += a.Wlp[(int64_t)rt * a.C + pos];
    for (int k = threadIdx.x; k < a.D; k += 64) pr += q[k] * q[k];
    lpv = wave_sum(lpv);
    pr = wave_sum(pr);
    if (threadIdx.x == 0) a.V[c] = -(lpv - 0.5 * a.s2inv * pr);
  }
}

GlLargePlan gl_large_plan(int N, int C, int D, int num_cus) {
  GlLargePlan p;
  p.row_tiles = (N + GL_BM - 1) / GL_BM;
  p.Dq = (D + 15) / 16 * 16;
  const int tiles2 = ((C + GL_BM - 1) / GL_BM) * ((D + GL_BN - 1) / GL_BN);
  int splits = (2 * (num_cus > 0 ? num_cus : 256) + tiles2 - 1) / tiles2;
  if (splits < 1) splits = 1;
  p.ldn = (int64_t)p.row_tiles * GL_BM;
  // row ranges in whole 64-row slabs (the BF16 variant's K step); a range past
  // N contributes zero partials
  int rps = (int)((p.ldn + splits - 1) / splits);
  rps = (rps + 63) / 64 * 64;
  p.rows_per_split = rps;
  p.splits = (int)((p.ldn + rps - 1) / rps);
  p.ldr = (C + 3) / 4 * 4;
  p.r_floats = (size_t)N * p.ldr;
  p.q_floats = (size_t)C * p.Dq;
  p.wg_doubles = (size_t)p.splits * C * D;
  p.wlp_doubles = (size_t)p.row_tiles * C;
  p.Cp = (C + GL_BM - 1) / GL_BM * GL_BM;
  p.Dk = (D + 63) / 64 * 64;
  p.Dp = (D + GL_BN - 1) / GL_BN * GL_BN;
  p.qb_elems = (size_t)3 * p.Cp * p.Dk;
  p.rb_elems = (size_t)3 * p.Cp * p.ldn;
  p.xt_elems = (size_t)p.Dp * p.ldn;
  return p;
}

void gl_large_grad(const GlLargeArgs& a, const GlLargePlan& p, const void* X, int xtype, hipStream_t st) {
  if (xtype == 2 && a.Qb && a.Rb && a.Xt) {
    gl_large_grad_bf16(a, p, static_cast<const unsigned short*>(X), st);
    return;
  }
  hipLaunchKernelGGL(gl_pack_q, dim3((unsigned)(((int64_t)a.C * p.Dq + 255) / 256)), dim3(256), 0, st, a);
  if (a.ev_begin) (void)hipEventRecord(a.ev_begin, st);
  const dim3 g1(p.row_tiles, (a.C + GL_BN - 1) / GL_BN);
  const dim3 g2((a.C + GL_BM - 1) / GL_BM, (a.D + GL_BN - 1) / GL_BN, p.splits);
  if (xtype == 1) {
    hipLaunchKernelGGL(gl_eta_kernel<float>, g1, dim3(256), 0, st, a, (const float*)X);
    hipLaunchKernelGGL(gl_grad_kernel<float>, g2, dim3(256), 0, st, a, (const float*)X, p.rows_per_split);
  } else {
    hipLaunchKernelGGL(gl_eta_kernel<unsigned short>, g1, dim3(256), 0, st, a, (const unsigned short*)X);
    hipLaunchKernelGGL(gl_grad_kernel<unsigned short>, g2, dim3(256), 0, st, a, (const unsigned short*)X,
                       p.rows_per_split);
  }
  if (a.ev_end) (void)hipEventRecord(a.ev_end, st);
  gl_reduce_launch(a, p, st);
}

void gl_reduce_launch(const GlLargeArgs& a, const GlLargePlan& p, hipStream_t st) {
  hipLaunchKernelGGL(gl_reduce, dim3(a.C), dim3(128), 0, st, a, p.splits, p.row_tiles);
}

}  // namespace nutsgpu
