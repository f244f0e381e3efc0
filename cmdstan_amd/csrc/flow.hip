// Affine-coupling flow kernels: an FP32 matrix-core GEMM (v_mfma_f32_16x16x4f32)
// with split-K and fused bias / tanh / tanh' epilogues, the coupling
// element-wise steps, and Adam.  Host class FlowNet in flow.h.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>

#include "device_common.h"
#include "flow.h"
#include "launch.h"

namespace nutsgpu {

typedef float f4 __attribute__((ext_vector_type(4)));

#define FLOW_S 3.0f  // s = S tanh(s_raw / S)

#define FG_BM 64
#define FG_BN 64
#define FG_BK 64
enum { FG_EPI_NONE = 0, FG_EPI_BIAS = 1, FG_EPI_BIAS_TANH = 2, FG_EPI_DTANH = 3, FG_EPI_ADD = 4 };

struct GemmDesc {
  const float* A; int64_t sam, sak;  // A(m, k) = A[m*sam + k*sak]
  const float* B; int64_t sbk, sbn;  // B(k, n) = B[k*sbk + n*sbn]
  float* C; int64_t ldc;             // C(m, n) = C[m*ldc + n]
  int M, N, K;
  const int* Mdev;                   // optional device row count (<= M)
  int epi;
  const float* bias;                 // [N]
  const float* aux; int64_t ldaux;   // FG_EPI_DTANH: times (1 - aux(m, n)^2)
  float* part; int splits, kper;     // split-K partials [splits][M][N], kper rows of K per split
  const int* aidx_k;                 // optional column gather of A: A(m, k) = A[m*sam + aidx_k[k]*sak]
  const int* aidx_m;                 // optional row gather of A:    A(m, k) = A[aidx_m[m]*sam + k*sak]
  const int* cidx_n;                 // optional column scatter of C: C(m, n) = C[m*ldc + cidx_n[n]]
};

__device__ __forceinline__ int64_t fg_cidx(const GemmDesc& g, int m, int n) {
  return (int64_t)m * g.ldc + (g.cidx_n ? g.cidx_n[n] : n);
}

__device__ __forceinline__ float fg_epi(const GemmDesc& g, float v, int m, int n) {
  if (g.epi == FG_EPI_BIAS) return v + g.bias[n];
  if (g.epi == FG_EPI_BIAS_TANH) return tanhf(v + g.bias[n]);
  if (g.epi == FG_EPI_DTANH) {
    const float h = g.aux[(int64_t)m * g.ldaux + n];
    return v * (1.0f - h * h);
  }
  if (g.epi == FG_EPI_ADD) return g.C[fg_cidx(g, m, n)] + v;
  return v;
}

// 64x64 tile per block of four waves (2x2), each wave 32x32 = 2x2 MFMA tiles;
// K in slabs of 64 (each thread stages 16 elements of A and 16 of B per slab,
// so a slab's loads are in flight together, and the next slab's loads are
// issued before the current slab's products: a tile waits on memory about
// once, however long its K range).
// f32 16x16x4 operands: lane l gives A[l%16][k=l/16], B[k=l/16][l%16]; result
// register i of lane l is D[4(l/16)+i][l%16].
__global__ void __launch_bounds__(256) flow_gemm_kernel(GemmDesc g) {
  __shared__ float As[FG_BK][FG_BM + 4];
  __shared__ float Bs[FG_BK][FG_BN + 4];
  const int M = g.Mdev ? min(g.M, *g.Mdev) : g.M;
  const int m0 = blockIdx.y * FG_BM, n0 = blockIdx.x * FG_BN;
  if (m0 >= M) return;
  const int split = blockIdx.z;
  const int kb = split * g.kper;
  const int ke = min(g.K, kb + g.kper);
  const int t = threadIdx.x, w = t >> 6, l = t & 63;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  f4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f4{0, 0, 0, 0};
  const bool a_k = g.sak == 1;
  const bool b_n = g.sbn == 1;
  float va[16], vb[16];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      // k-contiguous A: 16 consecutive k of one row; else 4 rows x 4 slab rows
      const int mm = a_k ? (t >> 2) : ((t & 15) * 4 + (e & 3));
      const int kk = a_k ? ((t & 3) * 16 + e) : ((t >> 4) + 16 * (e >> 2));
      const int m = m0 + mm, k = k0 + kk;
      float v = 0.0f;
      if (m < M && k < ke) {
        const int64_t mr = g.aidx_m ? g.aidx_m[m] : m;
        const int64_t kc = g.aidx_k ? g.aidx_k[k] : k;
        v = g.A[mr * g.sam + kc * g.sak];
      }
      va[e] = v;
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int nn = b_n ? ((t & 15) * 4 + (e & 3)) : (t >> 2);
      const int kk = b_n ? ((t >> 4) + 16 * (e >> 2)) : ((t & 3) * 16 + e);
      const int n = n0 + nn, k = k0 + kk;
      vb[e] = (n < g.N && k < ke) ? g.B[(int64_t)k * g.sbk + (int64_t)n * g.sbn] : 0.0f;
    }
  };
  // the next slab's loads are issued before the current slab's products
  if (kb < ke) fetch(kb);
  for (int k0 = kb; k0 < ke; k0 += FG_BK) {
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int mm = a_k ? (t >> 2) : ((t & 15) * 4 + (e & 3));
      const int kk = a_k ? ((t & 3) * 16 + e) : ((t >> 4) + 16 * (e >> 2));
      As[kk][mm] = va[e];
      const int nn = b_n ? ((t & 15) * 4 + (e & 3)) : (t >> 2);
      const int kb2 = b_n ? ((t >> 4) + 16 * (e >> 2)) : ((t & 3) * 16 + e);
      Bs[kb2][nn] = vb[e];
    }
    __syncthreads();
    if (k0 + FG_BK < ke) fetch(k0 + FG_BK);
#pragma unroll
    for (int ks = 0; ks < FG_BK / 4; ++ks) {
      const int kr = ks * 4 + (l >> 4);
      const float a0 = As[kr][wm + (l & 15)], a1 = As[kr][wm + 16 + (l & 15)];
      const float b0 = Bs[kr][wn + (l & 15)], b1 = Bs[kr][wn + 16 + (l & 15)];
      acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + 16 * i + 4 * (l >> 4) + r;
        const int n = n0 + wn + 16 * j + (l & 15);
        if (m < M && n < g.N) {
          const float v = acc[i][j][r];
          if (g.splits > 1) g.part[((int64_t)split * g.M + m) * g.N + n] = v;
          else g.C[fg_cidx(g, m, n)] = fg_epi(g, v, m, n);
        }
      }
}

__global__ void flow_gemm_reduce(GemmDesc g) {
  const int M = g.Mdev ? min(g.M, *g.Mdev) : g.M;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)M * g.N) return;
  const int m = (int)(i / g.N), n = (int)(i % g.N);
  float s = 0.0f;
  // split order kept; eight loads in flight instead of one
#pragma unroll 8
  for (int sp = 0; sp < g.splits; ++sp) s += g.part[((int64_t)sp * g.M + m) * g.N + n];
  g.C[fg_cidx(g, m, n)] = fg_epi(g, s, m, n);
}

// ------------------------------------------------------------ element-wise
__device__ __forceinline__ int rows_of(const int* n_dev, int n) { return n_dev ? min(n, *n_dev) : n; }

// ------------------------------------------- fused sampling-path layers
// The sampler's flow gradient (forward T(z) and the pullback) runs each
// coupling layer as three kernels instead of the generic split-K GEMM chain:
//   forward   h1 partials (split over the conditioner columns)
//             h1 = tanh(sum + b1) and h2 = tanh(h1 W2 + b2) per row tile
//             [s, t] = h2 W3 + b3 with the coupling and log-det epilogue
//   pullback  coupling cotangents and GP2 partials (GO W3^T, split over the
//             transformed columns)
//             GP2 = sum (1 - h2^2), GP1 = (GP2 W2^T)(1 - h1^2)
//             GC[:, ia] += GP1 W1^T
// Tiles: 64 rows x 64 columns per block of four waves (2 x 2 of 32 x 32), K in
// slabs of 64 staged through LDS, v_mfma_f32_16x16x4f32.  Every sum has a
// fixed order (no atomics): results are bit-reproducible.
#define FF_LD 68       // LDS row stride of the staged operand slabs (floats)
#define FF_HMAX 256    // widest MLP the fused kernels hold in LDS (64 x H tile)

// acc += A(m, k) B(k, n) over k in [kb, ke) for the block's 64 x 64 tile, K
// in slabs of 64: each thread holds 16 elements of A and 16 of B of a slab in
// registers, and the next slab's loads are issued before the current slab's
// products, so a tile waits on memory about once however long K is.
// A_KC: consecutive k of A are adjacent in memory (a thread loads 16 of one
// row); B_NC: consecutive n of B are adjacent (a thread loads 4 of one k, in
// 4 slab rows).  Products in k order, four at a time (v_mfma_f32_16x16x4f32).
#define FF_BK 64
template <bool A_KC, bool B_NC, class AF, class BF>
__device__ __forceinline__ void ff_mma(const AF& af, const BF& bf, int kb, int ke, f4 (&acc)[2][2],
                                       float (*As)[FF_LD], float (*Bs)[FF_LD]) {
  const int t = threadIdx.x, w = t >> 6, l = t & 63;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  auto a_m = [&](int e) { return A_KC ? (t >> 2) : ((t & 15) * 4 + (e & 3)); };
  auto a_k = [&](int e) { return A_KC ? ((t & 3) * 16 + e) : ((t >> 4) + 16 * (e >> 2)); };
  auto b_n = [&](int e) { return B_NC ? ((t & 15) * 4 + (e & 3)) : (t >> 2); };
  auto b_k = [&](int e) { return B_NC ? ((t >> 4) + 16 * (e >> 2)) : ((t & 3) * 16 + e); };
  float ra[16], rb[16];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int k = k0 + a_k(e);
      ra[e] = k < ke ? af(a_m(e), k) : 0.0f;
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int k = k0 + b_k(e);
      rb[e] = k < ke ? bf(k, b_n(e)) : 0.0f;
    }
  };
  if (kb < ke) fetch(kb);
  for (int k0 = kb; k0 < ke; k0 += FF_BK) {
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      As[a_k(e)][a_m(e)] = ra[e];
      Bs[b_k(e)][b_n(e)] = rb[e];
    }
    __syncthreads();
    if (k0 + FF_BK < ke) fetch(k0 + FF_BK);
#pragma unroll
    for (int ks = 0; ks < FF_BK / 4; ++ks) {  // slab rows past ke hold zeros
      const int kr = ks * 4 + (l >> 4);
      const float a0 = As[kr][wm + (l & 15)], a1 = As[kr][wm + 16 + (l & 15)];
      const float b0 = Bs[kr][wn + (l & 15)], b1 = Bs[kr][wn + 16 + (l & 15)];
      acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
}

__device__ __forceinline__ void ff_zero(f4 (&acc)[2][2]) {
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f4{0, 0, 0, 0};
}

// Calls f(mm, nn, v) for every element of the block's 64 x 64 tile (mm, nn
// local): register r of MFMA tile (i, j) of lane l holds row
// wm + 16 i + 4 (l / 16) + r, column wn + 16 j + l % 16.
template <class F>
__device__ __forceinline__ void ff_store(const f4 (&acc)[2][2], const F& f) {
  const int t = threadIdx.x, w = t >> 6, l = t & 63;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) f(wm + 16 * i + 4 * (l >> 4) + r, wn + 16 * j + (l & 15), acc[i][j][r]);
}

// LDS of one fused tile: the two staged operand slabs, then the tile's own
// buffer (at most 64 x (FF_HMAX + 1) floats).
#define FF_SLABS (2 * FF_BK * FF_LD)
#define FF_POOL (FF_SLABS + 64 * (FF_HMAX + 1))
typedef float (*ff_slab)[FF_LD];
__device__ __forceinline__ ff_slab ff_as(float* pool) { return reinterpret_cast<ff_slab>(pool); }
__device__ __forceinline__ ff_slab ff_bs(float* pool) { return reinterpret_cast<ff_slab>(pool + FF_BK * FF_LD); }

// forward 1, tile (bx, by, bz): h1 pre-activations U[:, ia] W1 over the
// conditioner columns [bz kper, (bz+1) kper); with one split the epilogue
// writes h1 = tanh(. + b1)
__device__ __forceinline__ void ff_h1_tile(int bx, int by, int bz, float* pool, const float* __restrict__ U, int D,
                                           const int* __restrict__ ia, int Da, const float* __restrict__ W1,
                                           const float* __restrict__ b1, int H, const int* __restrict__ n_dev, int n,
                                           int kper, int splits, float* __restrict__ h1, float* __restrict__ part) {
  const int M = rows_of(n_dev, n);
  const int m0 = by * 64, n0 = bx * 64;
  if (m0 >= M) return;
  const int kb = bz * kper, ke = min(Da, kb + kper);
  f4 acc[2][2];
  ff_zero(acc);
  auto af = [&](int mm, int k) { const int m = m0 + mm; return m < M ? U[(int64_t)m * D + ia[k]] : 0.0f; };
  auto bf = [&](int k, int nn) { const int c = n0 + nn; return c < H ? W1[(int64_t)k * H + c] : 0.0f; };
  ff_mma<true, true>(af, bf, kb, ke, acc, ff_as(pool), ff_bs(pool));
  ff_store(acc, [&](int mm, int nn, float v) {
    const int m = m0 + mm, c = n0 + nn;
    if (m >= M || c >= H) return;
    if (splits == 1) h1[(int64_t)m * H + c] = tanhf(v + b1[c]);
    else part[((int64_t)bz * n + m) * H + c] = v;
  });
}

__global__ void __launch_bounds__(256) ff_h1(const float* __restrict__ U, int D, const int* __restrict__ ia, int Da,
                                             const float* __restrict__ W1, const float* __restrict__ b1, int H,
                                             const int* __restrict__ n_dev, int n, int kper, int splits,
                                             float* __restrict__ h1, float* __restrict__ part) {
  __shared__ float pool[FF_SLABS];
  ff_h1_tile(blockIdx.x, blockIdx.y, blockIdx.z, pool, U, D, ia, Da, W1, b1, H, n_dev, n, kper, splits, h1, part);
}

// forward 2, tile (bx, by): h1 = tanh(sum of the partials + b1) for the
// tile's rows (all H columns, kept in LDS; column tile 0 stores it; with one
// split h1 is read as stored), h2 = tanh(h1 W2 + b2)
__device__ __forceinline__ void ff_h2_tile(int bx, int by, float* pool, const float* __restrict__ part, int splits,
                                           const float* __restrict__ b1, float* __restrict__ h1,
                                           const float* __restrict__ W2, const float* __restrict__ b2,
                                           float* __restrict__ h2, int H, const int* __restrict__ n_dev, int n) {
  float* sh = pool + FF_SLABS;  // [64][FF_HMAX + 1]
  const int M = rows_of(n_dev, n);
  const int m0 = by * 64, n0 = bx * 64;
  if (m0 >= M) return;
#pragma unroll 8
  for (int i = threadIdx.x; i < 64 * H; i += 256) {
    const int mm = i / H, c = i % H, m = m0 + mm;
    float v = 0.0f;
    if (m < M) {
      if (splits == 1) {
        v = h1[(int64_t)m * H + c];
      } else {
        float s = 0.0f;
#pragma unroll 8
        for (int z = 0; z < splits; ++z) s += part[((int64_t)z * n + m) * H + c];
        v = tanhf(s + b1[c]);
        if (bx == 0) h1[(int64_t)m * H + c] = v;
      }
    }
    sh[mm * (FF_HMAX + 1) + c] = v;
  }
  __syncthreads();
  f4 acc[2][2];
  ff_zero(acc);
  auto af = [&](int mm, int k) { return sh[mm * (FF_HMAX + 1) + k]; };
  auto bf = [&](int k, int nn) { const int c = n0 + nn; return c < H ? W2[(int64_t)k * H + c] : 0.0f; };
  ff_mma<true, true>(af, bf, 0, H, acc, ff_as(pool), ff_bs(pool));
  ff_store(acc, [&](int mm, int nn, float v) {
    const int m = m0 + mm, c = n0 + nn;
    if (m < M && c < H) h2[(int64_t)m * H + c] = tanhf(v + b2[c]);
  });
  __syncthreads();  // sh is reused by the next tile of a persistent block
}

__global__ void __launch_bounds__(256) ff_h2(const float* __restrict__ part, int splits, const float* __restrict__ b1,
                                             float* __restrict__ h1, const float* __restrict__ W2,
                                             const float* __restrict__ b2, float* __restrict__ h2, int H,
                                             const int* __restrict__ n_dev, int n) {
  __shared__ float pool[FF_POOL];
  ff_h2_tile(blockIdx.x, blockIdx.y, pool, part, splits, b1, h1, W2, b2, h2, H, n_dev, n);
}

// column tiles of the coupling step: the transformed columns' tiles, at
// least one per 64 conditioner columns
__host__ __device__ __forceinline__ int ff_couple_tiles(int Da, int Db) {
  const int tb = (Db + 63) / 64, ta = (Da + 63) / 64;
  return tb > ta ? tb : ta;
}

// forward 3, tile (bx, by) of nt column tiles: [s_raw, t] = h2 W3 + b3 for
// the transformed columns [j0, j0 + 64) (s at column j, t at Db + j), then
// the coupling u_b e^s + t with s = 3 tanh(s_raw / 3), S = s, the row sums of
// s over the tile (ld_part, FP64) and this tile's share of the conditioner
// columns copied through.  nt may exceed the transformed column tiles (a
// layer with few transformed columns spreads its conditioner copy wider):
// the extra tiles only copy.
__device__ __forceinline__ void ff_couple_tile(int bx, int by, int nt, float* pool, const float* __restrict__ h2,
                                               int H, const float* __restrict__ W3, const float* __restrict__ b3,
                                               const int* __restrict__ ia, int Da, const int* __restrict__ ib, int Db,
                                               int D, const float* __restrict__ U, float* __restrict__ Un,
                                               float* __restrict__ S, int64_t lds, double* __restrict__ ld_part,
                                               const int* __restrict__ n_dev, int n) {
  float* ss = pool + FF_SLABS;  // [64][65]
  const int M = rows_of(n_dev, n);
  const int m0 = by * 64, j0 = bx * 64;
  if (m0 >= M) return;
  // conditioner columns pass through unchanged
  const int share = (Da + nt - 1) / nt, ja = bx * share, jz = min(Da, ja + share);
#pragma unroll 4
  for (int i = threadIdx.x; i < 64 * (jz > ja ? jz - ja : 0); i += 256) {
    const int mm = i / (jz - ja), j = ja + i % (jz - ja), m = m0 + mm;
    if (m < M) Un[(int64_t)m * D + ia[j]] = U[(int64_t)m * D + ia[j]];
  }
  if (j0 >= Db) return;  // a tile past the transformed columns only copies
  const int64_t ld3 = 2 * (int64_t)Db;
  auto af = [&](int mm, int k) { const int m = m0 + mm; return m < M ? h2[(int64_t)m * H + k] : 0.0f; };
  f4 as[2][2], at[2][2];
  ff_zero(as);
  ff_zero(at);
  auto bs = [&](int k, int nn) { const int j = j0 + nn; return j < Db ? W3[k * ld3 + j] : 0.0f; };
  auto bt = [&](int k, int nn) { const int j = j0 + nn; return j < Db ? W3[k * ld3 + Db + j] : 0.0f; };
  ff_mma<true, true>(af, bs, 0, H, as, ff_as(pool), ff_bs(pool));
  ff_mma<true, true>(af, bt, 0, H, at, ff_as(pool), ff_bs(pool));
  const int t = threadIdx.x, w = t >> 6, l = t & 63;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int jt = 0; jt < 2; ++jt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int mm = wm + 16 * i + 4 * (l >> 4) + r, nn = wn + 16 * jt + (l & 15);
        const int m = m0 + mm, j = j0 + nn;
        float sv = 0.0f;
        if (m < M && j < Db) {
          sv = FLOW_S * tanhf((as[i][jt][r] + b3[j]) / FLOW_S);
          const int d = ib[j];
          Un[(int64_t)m * D + d] = U[(int64_t)m * D + d] * expf(sv) + (at[i][jt][r] + b3[Db + j]);
          S[(int64_t)m * lds + j] = sv;
        }
        ss[mm * 65 + nn] = sv;
      }
  __syncthreads();
  if (t < 64 && m0 + t < M) {
    double acc = 0.0;
    for (int nn = 0; nn < 64; ++nn) acc += (double)ss[t * 65 + nn];
    ld_part[(int64_t)bx * n + m0 + t] = acc;
  }
  __syncthreads();
}

__global__ void __launch_bounds__(256) ff_couple(const float* __restrict__ h2, int H, const float* __restrict__ W3,
                                                 const float* __restrict__ b3, const int* __restrict__ ia, int Da,
                                                 const int* __restrict__ ib, int Db, int D,
                                                 const float* __restrict__ U, float* __restrict__ Un,
                                                 float* __restrict__ S, int64_t lds, double* __restrict__ ld_part,
                                                 const int* __restrict__ n_dev, int n) {
  __shared__ float pool[FF_SLABS + 64 * 65];
  ff_couple_tile(blockIdx.x, blockIdx.y, gridDim.x, pool, h2, H, W3, b3, ia, Da, ib, Db, D, U, Un, S, lds, ld_part,
                 n_dev, n);
}

// forward end for row r: logdet = sum over layers and tiles of the row sums +
// sum log scale (thread 0); X = loc + scale * U_L over the row
__device__ __forceinline__ void ff_finish_row(int r, const double* __restrict__ ld_part, const int* __restrict__ Db,
                                              int L, int tmax, const float* __restrict__ UL,
                                              const double* __restrict__ loc, const double* __restrict__ scale, int D,
                                              int n, double log_scale_sum, double* __restrict__ logdet,
                                              double* __restrict__ X) {
  // thread t sums the tiles t, t + 256, ... of the layers' tile lists in
  // order; the four waves' sums are added in a fixed order
  __shared__ double red[4];
  double s = 0.0;
  for (int l = 0; l < L; ++l) {
    const int nt = (Db[l] + 63) / 64;
    for (int tt = threadIdx.x; tt < nt; tt += 256) s += ld_part[((int64_t)l * tmax + tt) * n + r];
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
#pragma unroll 4
  for (int j = threadIdx.x; j < D; j += blockDim.x)
    X[(int64_t)r * D + j] = loc[j] + scale[j] * (double)UL[(int64_t)r * D + j];
  __syncthreads();
  if (threadIdx.x == 0) logdet[r] = ((red[0] + red[1]) + (red[2] + red[3])) + log_scale_sum;
  __syncthreads();
}

__global__ void ff_finish(const double* __restrict__ ld_part, const int* __restrict__ Db, int L, int tmax,
                          const float* __restrict__ UL, const double* __restrict__ loc,
                          const double* __restrict__ scale, int D, const int* __restrict__ n_dev, int n,
                          double log_scale_sum, double* __restrict__ logdet, double* __restrict__ X) {
  const int r = blockIdx.x;
  if (r >= rows_of(n_dev, n)) return;
  ff_finish_row(r, ld_part, Db, L, tmax, UL, loc, scale, D, n, log_scale_sum, logdet, X);
}

// pullback 1, tile (bx, by) (one layer, transformed columns [j0, j0 + 64)):
// the coupling cotangents GO = [dU/ds_raw, dU/dt] (dU/ds = gv u e^s - 1, the
// -1 from the log-det), GC_b <- gv e^s in place, and the partial GP2 = GO W3^T
// over these columns for every MLP unit
__device__ __forceinline__ void fb_couple_tile(int bx, int by, float* pool, float* __restrict__ GC,
                                               const float* __restrict__ U, const float* __restrict__ S, int64_t lds,
                                               const int* __restrict__ ib, int Db, int D,
                                               const float* __restrict__ W3, int H, float* __restrict__ part,
                                               const int* __restrict__ n_dev, int n) {
  float* go = pool + FF_SLABS;  // [64][129]
  const int M = rows_of(n_dev, n);
  const int m0 = by * 64, j0 = bx * 64;
  if (m0 >= M) return;
#pragma unroll 8
  for (int i = threadIdx.x; i < 64 * 64; i += 256) {
    const int mm = i >> 6, jj = i & 63, m = m0 + mm, j = j0 + jj;
    float gs = 0.0f, gt = 0.0f;
    if (m < M && j < Db) {
      const int64_t e = (int64_t)m * D + ib[j];
      const float s = S[(int64_t)m * lds + j];
      const float es = expf(s);
      const float gv = GC[e];
      const float q = s / FLOW_S;
      gs = (gv * U[e] * es - 1.0f) * (1.0f - q * q);
      gt = gv;
      GC[e] = gv * es;
    }
    go[mm * 129 + jj] = gs;
    go[mm * 129 + 64 + jj] = gt;
  }
  __syncthreads();
  const int64_t ld3 = 2 * (int64_t)Db;
  auto af = [&](int mm, int k) { return go[mm * 129 + k]; };
  for (int c0 = 0; c0 < H; c0 += 64) {
    f4 acc[2][2];
    ff_zero(acc);
    // B(k, c) = W3[c][j], k < 64: s column j0 + k, else t column Db + j0 + k - 64
    auto bf = [&](int k, int nn) {
      const int c = c0 + nn;
      const int j = j0 + (k & 63);
      if (c >= H || j >= Db) return 0.0f;
      return W3[(int64_t)c * ld3 + (k < 64 ? j : Db + j)];
    };
    ff_mma<true, false>(af, bf, 0, 128, acc, ff_as(pool), ff_bs(pool));
    ff_store(acc, [&](int mm, int nn, float v) {
      const int m = m0 + mm, c = c0 + nn;
      if (m < M && c < H) part[((int64_t)bx * n + m) * H + c] = v;
    });
  }
  __syncthreads();
}

__global__ void __launch_bounds__(256) fb_couple(float* __restrict__ GC, const float* __restrict__ U,
                                                 const float* __restrict__ S, int64_t lds,
                                                 const int* __restrict__ ib, int Db, int D,
                                                 const float* __restrict__ W3, int H, float* __restrict__ part,
                                                 const int* __restrict__ n_dev, int n) {
  __shared__ float pool[FF_SLABS + 64 * 129];
  fb_couple_tile(blockIdx.x, blockIdx.y, pool, GC, U, S, lds, ib, Db, D, W3, H, part, n_dev, n);
}

// pullback 2, tile (bx, by): GP2 = (sum of the partials)(1 - h2^2) for the
// tile's rows (all H in LDS), GP1 = (GP2 W2^T)(1 - h1^2) for the tile's unit
// columns
__device__ __forceinline__ void fb_mlp_tile(int bx, int by, float* pool, const float* __restrict__ part, int splits,
                                            const float* __restrict__ h1, const float* __restrict__ h2,
                                            const float* __restrict__ W2, int H, float* __restrict__ GP1,
                                            const int* __restrict__ n_dev, int n) {
  float* sg = pool + FF_SLABS;  // [64][FF_HMAX + 1]
  const int M = rows_of(n_dev, n);
  const int m0 = by * 64, n0 = bx * 64;
  if (m0 >= M) return;
#pragma unroll 8
  for (int i = threadIdx.x; i < 64 * H; i += 256) {
    const int mm = i / H, c = i % H, m = m0 + mm;
    float v = 0.0f;
    if (m < M) {
      float s = 0.0f;
#pragma unroll 8
      for (int z = 0; z < splits; ++z) s += part[((int64_t)z * n + m) * H + c];
      const float hv = h2[(int64_t)m * H + c];
      v = s * (1.0f - hv * hv);
    }
    sg[mm * (FF_HMAX + 1) + c] = v;
  }
  __syncthreads();
  f4 acc[2][2];
  ff_zero(acc);
  auto af = [&](int mm, int k) { return sg[mm * (FF_HMAX + 1) + k]; };
  // B(k, c) = W2[c][k] (h2 = h1 W2: W2[k_in * H + k_out])
  auto bf = [&](int k, int nn) { const int c = n0 + nn; return c < H ? W2[(int64_t)c * H + k] : 0.0f; };
  ff_mma<true, false>(af, bf, 0, H, acc, ff_as(pool), ff_bs(pool));
  ff_store(acc, [&](int mm, int nn, float v) {
    const int m = m0 + mm, c = n0 + nn;
    if (m < M && c < H) {
      const float hv = h1[(int64_t)m * H + c];
      GP1[(int64_t)m * H + c] = v * (1.0f - hv * hv);
    }
  });
  __syncthreads();
}

__global__ void __launch_bounds__(256) fb_mlp(const float* __restrict__ part, int splits,
                                              const float* __restrict__ h1, const float* __restrict__ h2,
                                              const float* __restrict__ W2, int H, float* __restrict__ GP1,
                                              const int* __restrict__ n_dev, int n) {
  __shared__ float pool[FF_POOL];
  fb_mlp_tile(blockIdx.x, blockIdx.y, pool, part, splits, h1, h2, W2, H, GP1, n_dev, n);
}

// pullback 3, tile (bx, by): GC[:, ia[j]] += GP1 W1^T for the conditioner
// columns [j0, j0 + 64) (each element owned by one tile)
__device__ __forceinline__ void fb_in_tile(int bx, int by, float* pool, const float* __restrict__ GP1, int H,
                                           const float* __restrict__ W1, const int* __restrict__ ia, int Da, int D,
                                           float* __restrict__ GC, const int* __restrict__ n_dev, int n) {
  const int M = rows_of(n_dev, n);
  const int m0 = by * 64, j0 = bx * 64;
  if (m0 >= M) return;
  f4 acc[2][2];
  ff_zero(acc);
  auto af = [&](int mm, int k) { const int m = m0 + mm; return m < M ? GP1[(int64_t)m * H + k] : 0.0f; };
  // B(k, j) = W1[j][k] (W1 [Da][H])
  auto bf = [&](int k, int nn) { const int j = j0 + nn; return j < Da ? W1[(int64_t)j * H + k] : 0.0f; };
  ff_mma<true, false>(af, bf, 0, H, acc, ff_as(pool), ff_bs(pool));
  ff_store(acc, [&](int mm, int nn, float v) {
    const int m = m0 + mm, j = j0 + nn;
    if (m < M && j < Da) GC[(int64_t)m * D + ia[j]] += v;
  });
}

__global__ void __launch_bounds__(256) fb_in(const float* __restrict__ GP1, int H, const float* __restrict__ W1,
                                             const int* __restrict__ ia, int Da, int D, float* __restrict__ GC,
                                             const int* __restrict__ n_dev, int n) {
  __shared__ float pool[FF_SLABS];
  fb_in_tile(blockIdx.x, blockIdx.y, pool, GP1, H, W1, ia, Da, D, GC, n_dev, n);
}

// U0[r][j] = Q[idx[r]][j]
__global__ void flow_gather_z(const double* __restrict__ Q, int64_t ldq, const int* __restrict__ idx,
                              const int* __restrict__ n_dev, int n, int D, float* __restrict__ U0) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int r = (int)(i / D), j = (int)(i % D);
  if (r >= rows_of(n_dev, n)) return;
  const int64_t c = idx ? idx[r] : r;
  U0[i] = (float)Q[c * ldq + j];
}

// forward coupling: Un = U with b-part u_b e^s + t; S[r][j] = s
__global__ void flow_couple_fwd(const float* __restrict__ U, const float* __restrict__ O, const int* __restrict__ ia,
                                const int* __restrict__ ib, int Da, int Db, int D, const int* __restrict__ n_dev, int n,
                                float* __restrict__ S, float* __restrict__ Un) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int r = (int)(i / D), j = (int)(i % D);
  if (r >= rows_of(n_dev, n)) return;
  const int64_t row = (int64_t)r * D;
  if (j < Da) {
    Un[row + ia[j]] = U[row + ia[j]];
  } else {
    const int jb = j - Da, d = ib[jb];
    const float* o = O + (int64_t)r * 2 * Db;
    const float s = FLOW_S * tanhf(o[jb] / FLOW_S);
    S[(int64_t)r * Db + jb] = s;
    Un[row + d] = U[row + d] * expf(s) + o[Db + jb];
  }
}

// inverse coupling (training): Vn = V with b-part (v_b - t) e^{-s}
__global__ void flow_couple_inv(const float* __restrict__ V, const float* __restrict__ O, const int* __restrict__ ia,
                                const int* __restrict__ ib, int Da, int Db, int D, int n, float* __restrict__ S,
                                float* __restrict__ Vn) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int r = (int)(i / D), j = (int)(i % D);
  if (r >= n) return;
  const int64_t row = (int64_t)r * D;
  if (j < Da) {
    Vn[row + ia[j]] = V[row + ia[j]];
  } else {
    const int jb = j - Da, d = ib[jb];
    const float* o = O + (int64_t)r * 2 * Db;
    const float s = FLOW_S * tanhf(o[jb] / FLOW_S);
    S[(int64_t)r * Db + jb] = s;
    Vn[row + d] = (V[row + d] - o[Db + jb]) * expf(-s);
  }
}

// sum over layers of sum_j s, plus sum log scale (FP64), one block per row
__global__ void __launch_bounds__(256) flow_logdet(const float* __restrict__ S, int64_t layer_stride,
                                                   const int* __restrict__ Db, int L, const int* __restrict__ n_dev,
                                                   int n, double log_scale_sum, double* __restrict__ logdet, int sign) {
  __shared__ double red[4];
  const int r = blockIdx.x;
  if (r >= rows_of(n_dev, n)) return;
  double acc = 0;
  for (int l = 0; l < L; ++l) {
    const int db = Db[l];
    const float* s = S + l * layer_stride + (int64_t)r * db;
    for (int j = threadIdx.x; j < db; j += 256) acc += (double)s[j];
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double t = red[0] + red[1] + red[2] + red[3];
    logdet[r] = sign > 0 ? t + log_scale_sum : -t - log_scale_sum;
  }
}

// X = loc + scale * U (FP64 out)
__global__ void flow_to_x(const float* __restrict__ U, const double* __restrict__ loc, const double* __restrict__ scale,
                          int D, const int* __restrict__ n_dev, int n, double* __restrict__ X) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int r = (int)(i / D), j = (int)(i % D);
  if (r >= rows_of(n_dev, n)) return;
  X[i] = loc[j] + scale[j] * (double)U[i];
}

// V0 = (X - loc) / scale (FP32), rows r < n; rows of X picked by `pick` (or r)
__global__ void flow_standardize(const double* __restrict__ X, const int64_t* __restrict__ pick,
                                 const double* __restrict__ loc, const double* __restrict__ scale, int D, int n,
                                 float* __restrict__ V0) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int r = (int)(i / D), j = (int)(i % D);
  if (r >= n) return;
  const int64_t src = pick ? pick[r] : r;
  V0[i] = (float)((X[src * D + j] - loc[j]) / scale[j]);
}

// GC = GX * scale (cotangent of the last layer's output)
__global__ void flow_bwd_init(const double* __restrict__ GX, const double* __restrict__ scale, int D,
                              const int* __restrict__ n_dev, int n, float* __restrict__ GC) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int r = (int)(i / D), j = (int)(i % D);
  if (r >= rows_of(n_dev, n)) return;
  GC[i] = (float)(GX[i] * scale[j]);
}

// sampling VJP through one coupling: GO = [dU/ds_raw, dU/dt]; GC_b <- gv_b e^s
// (U = V(T(z)) - log|det|, so dU/ds = gv_b u_b e^s - 1)
__global__ void flow_couple_bwd_pull(float* __restrict__ GC, const float* __restrict__ U, const float* __restrict__ S,
                                     const int* __restrict__ ib, int Db, int D, const int* __restrict__ n_dev, int n,
                                     float* __restrict__ GO) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int r = (int)(i / Db), j = (int)(i % Db);
  if (r >= rows_of(n_dev, n)) return;
  const int64_t e = (int64_t)r * D + ib[j];
  const float s = S[i];
  const float es = expf(s);
  const float gv = GC[e];
  const float gs = gv * U[e] * es - 1.0f;
  const float q = s / FLOW_S;
  GO[(int64_t)r * 2 * Db + j] = gs * (1.0f - q * q);
  GO[(int64_t)r * 2 * Db + Db + j] = gv;
  GC[e] = gv * es;
}

// training backward through one inverse coupling (loss = mean 0.5|z|^2 + sum s):
// u_b = (v_b - t) e^{-s}: dL/ds = -gu_b u_b + 1/B, dL/dt = -gu_b e^{-s}, gv_b = gu_b e^{-s}
__global__ void flow_couple_bwd_train(float* __restrict__ GC, const float* __restrict__ Uz, const float* __restrict__ S,
                                      const int* __restrict__ ib, int Db, int D, int n, float inv_b,
                                      float* __restrict__ GO) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int r = (int)(i / Db), j = (int)(i % Db);
  if (r >= n) return;
  const int64_t e = (int64_t)r * D + ib[j];
  const float s = S[i];
  const float ems = expf(-s);
  const float gu = GC[e];
  const float q = s / FLOW_S;
  GO[(int64_t)r * 2 * Db + j] = (inv_b - gu * Uz[e]) * (1.0f - q * q);
  GO[(int64_t)r * 2 * Db + Db + j] = -gu * ems;
  GC[e] = gu * ems;
}

// G[idx[r]] = GC[r]; V[idx[r]] = VX[r] - logdet[r]
__global__ void flow_finish(const float* __restrict__ GC, const double* __restrict__ VX,
                            const double* __restrict__ logdet, const int* __restrict__ idx, const int* __restrict__ n_dev,
                            int n, int D, double* __restrict__ G, int64_t ldg, double* __restrict__ V) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int r = (int)(i / D), j = (int)(i % D);
  if (r >= rows_of(n_dev, n)) return;
  const int64_t c = idx ? idx[r] : r;
  G[c * ldg + j] = (double)GC[i];
  if (j == 0) V[c] = VX[r] - logdet[r];
}

// sampler rows: columns 7.. <- X, column 0 -= logdet
__global__ void flow_store_rows(const double* __restrict__ X, const double* __restrict__ logdet, int n, int D,
                                double* __restrict__ rows, int64_t ld) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int r = (int)(i / D), j = (int)(i % D);
  if (r >= n) return;
  rows[r * ld + 7 + j] = X[i];
  if (j == 0) rows[r * ld] -= logdet[r];
}

// z-side rows to FP64
__global__ void flow_rows_out(const float* __restrict__ U, int64_t n, double* __restrict__ Z) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) Z[i] = (double)U[i];
}

// column sums (bias gradients): out[c] = sum_r A[r][c]
__global__ void flow_colsum(const float* __restrict__ A, int rows, int cols, float* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  float s = 0.0f;
  // row order kept; eight loads in flight instead of one
#pragma unroll 8
  for (int r = 0; r < rows; ++r) s += A[(int64_t)r * cols + c];
  out[c] = s;
}

// per-row 0.5|z|^2
__global__ void flow_half_sq(const float* __restrict__ Z, int D, int n, double* __restrict__ out) {
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (r >= n) return;
  double acc = 0;
  for (int j = lane; j < D; j += NUTS_WAVE) {
    const double z = Z[(int64_t)r * D + j];
    acc += z * z;
  }
  acc = wave_sum(acc);
  if (lane == 0) out[r] = 0.5 * acc;
}

__global__ void flow_loss_rows(double* __restrict__ half_sq, const double* __restrict__ logdet_inv, int n) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) half_sq[r] -= logdet_inv[r];
}

__global__ void flow_scale(const float* __restrict__ A, int64_t n, float f, float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = A[i] * f;
}

// minibatch rows: floor(u * n_draws) from a Philox stream keyed by (seed, step)
__global__ void flow_pick_rows(uint64_t seed, uint32_t step, int n, int64_t n_draws, int64_t* __restrict__ pick) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const double u = stream_uniform(seed, step, 3u, (uint64_t)r);
  int64_t k = (int64_t)(u * (double)n_draws);
  pick[r] = k < n_draws ? k : n_draws - 1;
}

// Adam (Kingma & Ba 2015, Algorithm 1), step t >= 1
__global__ void flow_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                          float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ------------------------------------------- persistent sampling path
// The sampler's flow gradient as two launches, forward T(z) and the pullback,
// each one block of 256 threads per CU running every phase of every layer:
//   forward   z rows; per layer h1 partials over the conditioner columns,
//             h1 = tanh(sum + b1), h2 = tanh(h1 W2 + b2), [s, t] = h2 W3 + b3
//             with the coupling and log-det row sums; X and log|det|
//   pullback  GC = GX scale; per layer (last first) coupling cotangents and
//             GP2 partials over the transformed columns, GP2 = their sum,
//             GP1 = (GP2 (1 - h2^2) W2^T)(1 - h1^2), GC[:, ia] += GP1 W1^T;
//             G and V of the chains
// Phases are separated by a grid barrier (fp_sync).  The tiles are the fused
// per-layer kernels' (ff_*_tile, fb_*_tile) and every sum keeps their fixed
// order, so the results equal the per-layer kernels' bit for bit.
#define FP_LMAX 16

struct FpLayer {
  const int* ia;
  const int* ib;
  int Da, Db, kper, splits;  // first product split over the conditioner columns
  const float *W1, *b1, *W2, *b2, *W3, *b3;
};

struct FpArgs {
  FpLayer ly[FP_LMAX];
  int L, D, H, n, tmax;
  const int* n_dev;
  const int* idx;
  int64_t RD, RH, RS;  // per-layer strides of U, H1/H2, S
  float *U, *H1, *H2, *S, *part, *GC, *GP1, *GP2;
  double* ld_part;
  const int* db_dev;
  const double *loc, *scale;
  double log_scale_sum;
  // forward
  const double* Q;
  int64_t ldq;
  double *X, *logdet;
  // pullback
  const double *GX, *VX, *logdet_in;
  double* G;
  int64_t ldg;
  double* V;
  // barrier: bar[0] arrivals, bar[1] departures; err set when a wait times out
  unsigned* bar;
  int* err;
  uint64_t spin_ticks;  // wall-clock ticks a barrier may wait
};

// Grid barrier: every wave waits for its stores to be acknowledged, then one
// thread per block releases at agent scope (an L2 write-back), counts the
// arrival and polls the count with relaxed loads (an acquiring poll would
// invalidate the L2 on every read, under every other block's loads) until all
// blocks have arrived at this barrier, the error flag is set, or the time limit
// passes (which sets it); then it acquires once (L1 / L2 invalidation).  The
// agent-scope write-back and invalidation make the data written before the
// barrier in one XCD visible after it in another.
__device__ __forceinline__ void fp_sync(const FpArgs& a, unsigned& target) {
  target += gridDim.x;
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(a.bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t t0 = wall_clock64();
    while (__hip_atomic_load(a.bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
      if (wall_clock64() - t0 > a.spin_ticks) {
        __hip_atomic_store(a.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// Kernel end: the last block to leave (every block has finished waiting on
// bar[0]) resets both counters for the next launch.
__device__ __forceinline__ void fp_depart(const FpArgs& a) {
  if (threadIdx.x == 0) {
    const unsigned d = __hip_atomic_fetch_add(a.bar + 1, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (d == gridDim.x - 1) {
      __hip_atomic_store(a.bar, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(a.bar + 1, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

__global__ void __launch_bounds__(256) flow_fwd_persistent(FpArgs a) {
  __shared__ float pool[FF_POOL];
  unsigned target = 0;
  const int M = rows_of(a.n_dev, a.n);
  const int G = gridDim.x, D = a.D, H = a.H, n = a.n;
  const int64_t gt = (int64_t)blockIdx.x * 256 + threadIdx.x, gs = (int64_t)G * 256;
#pragma unroll 4
  for (int64_t i = gt; i < (int64_t)M * D; i += gs) {
    const int r = (int)(i / D), j = (int)(i % D);
    const int64_t c = a.idx ? a.idx[r] : r;
    a.U[i] = (float)a.Q[c * a.ldq + j];
  }
  fp_sync(a, target);
  const int tm = (n + 63) / 64, th = (H + 63) / 64;
  for (int l = 0; l < a.L; ++l) {
    const FpLayer& ly = a.ly[l];
    float* U = a.U + l * a.RD;
    float* h1 = a.H1 + l * a.RH;
    float* h2 = a.H2 + l * a.RH;
    const int S = ly.splits;
    for (int it = blockIdx.x; it < th * tm * S; it += G)
      ff_h1_tile(it % th, (it / th) % tm, it / (th * tm), pool, U, D, ly.ia, ly.Da, ly.W1, ly.b1, H, a.n_dev, n,
                 ly.kper, S, h1, a.part);
    fp_sync(a, target);
    if (S > 1) {
      for (int64_t i = gt; i < (int64_t)M * H; i += gs) {
        const int m = (int)(i / H), c = (int)(i % H);
        float s = 0.0f;
#pragma unroll 8
        for (int z = 0; z < S; ++z) s += a.part[((int64_t)z * n + m) * H + c];
        h1[i] = tanhf(s + ly.b1[c]);
      }
      fp_sync(a, target);
    }
    for (int it = blockIdx.x; it < th * tm; it += G)
      ff_h2_tile(it % th, it / th, pool, nullptr, 1, ly.b1, h1, ly.W2, ly.b2, h2, H, a.n_dev, n);
    fp_sync(a, target);
    const int tc = ff_couple_tiles(ly.Da, ly.Db);
    for (int it = blockIdx.x; it < tc * tm; it += G)
      ff_couple_tile(it % tc, it / tc, tc, pool, h2, H, ly.W3, ly.b3, ly.ia, ly.Da, ly.ib, ly.Db, D, U, U + a.RD,
                     a.S + l * a.RS, (int64_t)ly.Db, a.ld_part + (int64_t)l * a.tmax * n, a.n_dev, n);
    fp_sync(a, target);
  }
  for (int r = blockIdx.x; r < M; r += G)
    ff_finish_row(r, a.ld_part, a.db_dev, a.L, a.tmax, a.U + a.L * a.RD, a.loc, a.scale, D, n, a.log_scale_sum,
                  a.logdet, a.X);
  fp_depart(a);
}

__global__ void __launch_bounds__(256) flow_bwd_persistent(FpArgs a) {
  __shared__ float pool[FF_POOL];
  unsigned target = 0;
  const int M = rows_of(a.n_dev, a.n);
  const int G = gridDim.x, D = a.D, H = a.H, n = a.n;
  const int64_t gt = (int64_t)blockIdx.x * 256 + threadIdx.x, gs = (int64_t)G * 256;
#pragma unroll 4
  for (int64_t i = gt; i < (int64_t)M * D; i += gs) a.GC[i] = (float)(a.GX[i] * a.scale[i % D]);
  fp_sync(a, target);
  const int tm = (n + 63) / 64, th = (H + 63) / 64;
  for (int l = a.L - 1; l >= 0; --l) {
    const FpLayer& ly = a.ly[l];
    const float* h1 = a.H1 + l * a.RH;
    const float* h2 = a.H2 + l * a.RH;
    const int tb = (ly.Db + 63) / 64, ta = (ly.Da + 63) / 64;
    for (int it = blockIdx.x; it < tb * tm; it += G)
      fb_couple_tile(it % tb, it / tb, pool, a.GC, a.U + l * a.RD, a.S + l * a.RS, (int64_t)ly.Db, ly.ib, ly.Db, D,
                     ly.W3, H, a.part, a.n_dev, n);
    fp_sync(a, target);
    for (int64_t i = gt; i < (int64_t)M * H; i += gs) {
      const int m = (int)(i / H), c = (int)(i % H);
      float s = 0.0f;
#pragma unroll 8
      for (int z = 0; z < tb; ++z) s += a.part[((int64_t)z * n + m) * H + c];
      a.GP2[i] = s;
    }
    fp_sync(a, target);
    for (int it = blockIdx.x; it < th * tm; it += G)
      fb_mlp_tile(it % th, it / th, pool, a.GP2, 1, h1, h2, ly.W2, H, a.GP1, a.n_dev, n);
    fp_sync(a, target);
    for (int it = blockIdx.x; it < ta * tm; it += G)
      fb_in_tile(it % ta, it / ta, pool, a.GP1, H, ly.W1, ly.ia, ly.Da, D, a.GC, a.n_dev, n);
    fp_sync(a, target);
  }
#pragma unroll 4
  for (int64_t i = gt; i < (int64_t)M * D; i += gs) {
    const int r = (int)(i / D), j = (int)(i % D);
    const int64_t c = a.idx ? a.idx[r] : r;
    a.G[c * a.ldg + j] = (double)a.GC[i];
    if (j == 0) a.V[c] = a.VX[r] - a.logdet_in[r];
  }
  fp_depart(a);
}

static dim3 grid1(int64_t n, int b = 256) { return dim3((unsigned)((n + b - 1) / b)); }

#define HIPF(x)                                                                         \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP: ") + hipGetErrorString(e_)); \
  } while (0)

template <class T>
static T* dalloc(int64_t n) {
  T* p = nullptr;
  if (n > 0) HIPF(hipMalloc(&p, n * sizeof(T)));
  return p;
}

// ------------------------------------------------------------------- masks
FlowSpec flow_default_spec(int D, int L, int H, int n_glob, int group_size) {
  FlowSpec s;
  s.D = D; s.L = L; s.H = H;
  std::vector<int> glob, even, odd, loc;
  for (int i = 0; i < n_glob && i < D; ++i) glob.push_back(i);
  for (int i = n_glob; i < D; ++i) {
    loc.push_back(i);
    (((i - n_glob) / group_size) % 2 == 0 ? even : odd).push_back(i);
  }
  if (odd.empty()) {
    const size_t h = loc.size() / 2;
    even.assign(loc.begin(), loc.begin() + h);
    odd.assign(loc.begin() + h, loc.end());
  }
  auto cat = [](const std::vector<int>& a, const std::vector<int>& b) {
    std::vector<int> r(a);
    r.insert(r.end(), b.begin(), b.end());
    return r;
  };
  for (int l = 0; l < L; ++l) {
    std::vector<int> a, b;
    if (!glob.empty()) {
      const int k = l % 3;
      if (k == 0) { a = cat(glob, even); b = odd; }
      else if (k == 1) { a = cat(glob, odd); b = even; }
      else { a = loc; b = glob; }
    } else {
      a = (l % 2 == 0) ? even : odd;
      b = (l % 2 == 0) ? odd : even;
    }
    s.ia.push_back(a);
    s.ib.push_back(b);
  }
  return s;
}

// ------------------------------------------------------------------ FlowNet
FlowNet::FlowNet(const FlowSpec& spec, int max_rows) : spec_(spec), R_(max_rows) {
  HIPF(hipStreamCreateWithFlags(&own_st_, hipStreamNonBlocking));
  st_ = own_st_;
  const int D = spec.D, L = spec.L, H = spec.H;
  if (D < 2 || L < 1 || H < 1 || max_rows < 1 || (int)spec.ia.size() != L || (int)spec.ib.size() != L)
    throw std::invalid_argument("flow: bad specification");
  int64_t off = 0;
  for (int l = 0; l < L; ++l) {
    Layer ly{};
    ly.Da = (int)spec.ia[l].size();
    ly.Db = (int)spec.ib[l].size();
    if (ly.Da < 1 || ly.Db < 1 || ly.Da + ly.Db != D) throw std::invalid_argument("flow: masks must partition 0..D-1");
    std::vector<int> seen(D, 0);
    for (int v : spec.ia[l]) { if (v < 0 || v >= D || seen[v]++) throw std::invalid_argument("flow: bad mask"); }
    for (int v : spec.ib[l]) { if (v < 0 || v >= D || seen[v]++) throw std::invalid_argument("flow: bad mask"); }
    const int64_t sz[6] = {(int64_t)ly.Da * H, H, (int64_t)H * H, H, (int64_t)H * 2 * ly.Db, 2 * (int64_t)ly.Db};
    for (int k = 0; k < 6; ++k) { ly.off[k] = off; off += sz[k]; }
    ly.ia = dalloc<int>(ly.Da);
    ly.ib = dalloc<int>(ly.Db);
    HIPF(hipMemcpy(ly.ia, spec.ia[l].data(), ly.Da * sizeof(int), hipMemcpyHostToDevice));
    HIPF(hipMemcpy(ly.ib, spec.ib[l].data(), ly.Db * sizeof(int), hipMemcpyHostToDevice));
    Dmax_a_ = std::max(Dmax_a_, ly.Da);
    Dmax_b_ = std::max(Dmax_b_, ly.Db);
    layers_.push_back(ly);
  }
  n_params_ = off;
  params_ = dalloc<float>(off); grad_ = dalloc<float>(off);
  adam_m_ = dalloc<float>(off); adam_v_ = dalloc<float>(off);
  loc_ = dalloc<double>(D); scale_ = dalloc<double>(D);
  const int64_t R = R_;
  U_ = dalloc<float>((L + 1) * R * D);
  H1_ = dalloc<float>(L * R * H); H2_ = dalloc<float>(L * R * H);
  O_ = dalloc<float>(R * 2 * Dmax_b_);
  S_ = dalloc<float>(L * R * Dmax_b_);
  GO_ = dalloc<float>(R * 2 * Dmax_b_);
  GP2_ = dalloc<float>(R * H); GP1_ = dalloc<float>(R * H);
  GC_ = dalloc<float>(R * D);
  rowbuf_ = dalloc<double>(R);
  rowbuf2_ = dalloc<double>(R);
  pick_ = dalloc<int64_t>(R);
  xrows_ = dalloc<double>(R * D);
  // split-K partials: splits * M * N <= (512 + tiles) * 64 * 64 for any GEMM here
  part_n_ = (int64_t)1024 * FG_BM * FG_BN;
  part_ = dalloc<float>(part_n_);
  std::vector<int> dbs(L);
  for (int l = 0; l < L; ++l) dbs[l] = layers_[l].Db;
  ld_part_ = dalloc<double>((int64_t)L * ((Dmax_b_ + 63) / 64) * R);
  db_dev_ = dalloc<int>(L);
  bar_ = dalloc<unsigned>(2);
  bar_err_ = dalloc<int>(1);
  HIPF(hipMemset(bar_, 0, 2 * sizeof(unsigned)));
  HIPF(hipMemset(bar_err_, 0, sizeof(int)));
  {
    int dev = 0, cus = 0, khz = 0;
    HIPF(hipGetDevice(&dev));
    HIPF(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    HIPF(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    n_cu_ = cus;
    spin_ticks_ = (uint64_t)std::max(khz, 1000) * 2000;  // 2 s
  }
  HIPF(hipMemcpy(db_dev_, dbs.data(), L * sizeof(int), hipMemcpyHostToDevice));
  std::vector<double> z(D, 0.0), o(D, 1.0);
  set_standardization(z.data(), o.data());
  init_params(1);
}

FlowNet::~FlowNet() {
  if (own_st_) {
    (void)hipStreamSynchronize(own_st_);
    (void)hipStreamDestroy(own_st_);
  }
  for (auto& l : layers_) { (void)hipFree(l.ia); (void)hipFree(l.ib); }
  for (void* p : {(void*)params_, (void*)grad_, (void*)adam_m_, (void*)adam_v_, (void*)loc_, (void*)scale_, (void*)U_,
                  (void*)H1_, (void*)H2_, (void*)O_, (void*)S_, (void*)GO_, (void*)GP2_, (void*)GP1_,
                  (void*)GC_, (void*)part_, (void*)rowbuf_, (void*)rowbuf2_, (void*)pick_, (void*)xrows_,
                  (void*)db_dev_, (void*)ld_part_, (void*)bar_, (void*)bar_err_})
    if (p) (void)hipFree(p);
}

void FlowNet::set_params(const float* host) {
  HIPF(hipMemcpyAsync(params_, host, n_params_ * sizeof(float), hipMemcpyHostToDevice, st_));
  HIPF(hipStreamSynchronize(st_));
}

void FlowNet::get_params(float* host) const {
  HIPF(hipMemcpyAsync(host, params_, n_params_ * sizeof(float), hipMemcpyDeviceToHost, st_));
  HIPF(hipStreamSynchronize(st_));
}

void FlowNet::init_params(uint64_t seed) {
  std::vector<float> p(n_params_, 0.0f);
  std::mt19937_64 rng(seed);
  const int H = spec_.H;
  for (auto& ly : layers_) {
    const int fan[2][2] = {{ly.Da, H}, {H, H}};
    const int64_t o[2] = {ly.off[0], ly.off[2]};
    for (int k = 0; k < 2; ++k) {
      const double lim = std::sqrt(6.0 / (fan[k][0] + fan[k][1]));
      std::uniform_real_distribution<double> U(-lim, lim);
      for (int64_t i = 0; i < (int64_t)fan[k][0] * fan[k][1]; ++i) p[o[k] + i] = (float)U(rng);
    }
  }
  set_params(p.data());
  HIPF(hipMemsetAsync(adam_m_, 0, n_params_ * sizeof(float), st_));
  HIPF(hipMemsetAsync(adam_v_, 0, n_params_ * sizeof(float), st_));
  HIPF(hipStreamSynchronize(st_));
  adam_t_ = 0;
}

void FlowNet::set_standardization(const double* loc, const double* scale) {
  const int D = spec_.D;
  log_scale_sum_ = 0;
  for (int j = 0; j < D; ++j) {
    if (!(scale[j] > 0) || !std::isfinite(scale[j]) || !std::isfinite(loc[j]))
      throw std::invalid_argument("flow: standardization must be finite with positive scale");
    log_scale_sum_ += std::log(scale[j]);
  }
  HIPF(hipMemcpyAsync(loc_, loc, D * sizeof(double), hipMemcpyHostToDevice, st_));
  HIPF(hipMemcpyAsync(scale_, scale, D * sizeof(double), hipMemcpyHostToDevice, st_));
  HIPF(hipStreamSynchronize(st_));
}

void FlowNet::gemm(const float* A, int64_t sam, int64_t sak, const float* B, int64_t sbk, int64_t sbn, float* C,
                   int64_t ldc, int M, int N, int K, const int* Mdev, int epi, const float* bias, const float* aux,
                   int64_t ldaux, const int* aidx_k, const int* aidx_m, const int* cidx_n) {
  GemmDesc g{};
  g.A = A; g.sam = sam; g.sak = sak; g.B = B; g.sbk = sbk; g.sbn = sbn; g.C = C; g.ldc = ldc;
  g.M = M; g.N = N; g.K = K; g.Mdev = Mdev; g.epi = epi; g.bias = bias; g.aux = aux; g.ldaux = ldaux;
  g.aidx_k = aidx_k; g.aidx_m = aidx_m; g.cidx_n = cidx_n;
  const int tm = (M + FG_BM - 1) / FG_BM, tn = (N + FG_BN - 1) / FG_BN;
  const int kblocks = (K + FG_BK - 1) / FG_BK;
  // split K until about 512 blocks, keeping >= 2 K-slabs (128 of K) per split:
  // the K = H products of the MLP run unsplit
  int splits = std::max(1, std::min(kblocks / 2, (512 + tm * tn - 1) / (tm * tn)));
  g.kper = ((kblocks + splits - 1) / splits) * FG_BK;
  splits = (K + g.kper - 1) / g.kper;
  g.splits = splits;
  if (splits > 1) {
    if ((int64_t)splits * M * N > part_n_) throw std::logic_error("flow: split-K workspace too small");
    g.part = part_;
  }
  hipLaunchKernelGGL(flow_gemm_kernel, dim3(tn, tm, splits), dim3(256), 0, st_, g);
  if (splits > 1) hipLaunchKernelGGL(flow_gemm_reduce, grid1((int64_t)M * N), dim3(256), 0, st_, g);
}

// MLP of layer l on UA (rows x Da): H1, H2 of the layer, O = [s_raw, t]
void FlowNet::mlp(int l, const float* U, int ldu, const int* ia, const int* n_dev, int rows) {
  const Layer& ly = layers_[l];
  const int H = spec_.H;
  const float* P = params_;
  float* h1 = H1_ + (int64_t)l * R_ * H;
  float* h2 = H2_ + (int64_t)l * R_ * H;
  // the conditioner columns ia of U are gathered by the first product's loads
  gemm(U, ldu, 1, P + ly.off[0], H, 1, h1, H, rows, H, ly.Da, n_dev, FG_EPI_BIAS_TANH, P + ly.off[1], nullptr, 0, ia);
  gemm(h1, H, 1, P + ly.off[2], H, 1, h2, H, rows, H, H, n_dev, FG_EPI_BIAS_TANH, P + ly.off[3], nullptr, 0);
  gemm(h2, H, 1, P + ly.off[4], 2 * ly.Db, 1, O_, 2 * ly.Db, rows, 2 * ly.Db, H, n_dev, FG_EPI_BIAS, P + ly.off[5],
       nullptr, 0);
}

// split of the first MLP product over the conditioner columns: about 256
// blocks, at least 32 columns per split, partials within the workspace
static int ff_h1_splits(int Da, int H, int n_rows, int64_t part_n, int* kper) {
  const int tiles = ((H + 63) / 64) * ((n_rows + 63) / 64);
  int splits = std::max(1, std::min((256 + tiles - 1) / tiles, Da / 32));
  while (splits > 1 && (int64_t)splits * n_rows * H > part_n) --splits;
  *kper = ((Da + splits - 1) / splits + 15) / 16 * 16;
  return (Da + *kper - 1) / *kper;
}

// The fused per-layer kernels run only with CMDSTAN_AMD_FLOW_FUSED=1: at 256
// rows and D = 4020 they measured 2.7x slower than the split-K GEMM chain
// (their h1 / GP2 partial sums are reduced by one block per 64 x 64 tile,
// eight blocks in all), so the GEMM chain is the default.
bool FlowNet::fused_ok(int n_rows) const {
  const char* e = std::getenv("CMDSTAN_AMD_FLOW_FUSED");
  if (!e || std::atoi(e) == 0) return false;
  const int64_t rh = (int64_t)n_rows * spec_.H;
  return spec_.H <= FF_HMAX && (int64_t)((Dmax_b_ + 63) / 64) * rh <= part_n_;
}

// The persistent kernels run the sampling path with CMDSTAN_AMD_FLOW_PERSISTENT=1
// (CMDSTAN_AMD_FLOW_FUSED=1, the per-layer kernels, takes precedence) when
// they apply: MLP width within the LDS row buffers, at most FP_LMAX layers,
// the first product's and the pullback's partial sums within the workspace.
// A launch is one block per CU (about 100 KB of LDS each), all resident at
// once, which the grid barriers need.
bool FlowNet::persistent_ok(int n_rows) const {
  const char* e = std::getenv("CMDSTAN_AMD_FLOW_PERSISTENT");
  if (!e || std::atoi(e) == 0) return false;
  if (spec_.H > FF_HMAX || spec_.L > FP_LMAX || n_cu_ < 1) return false;
  const int64_t rh = (int64_t)n_rows * spec_.H;
  for (const Layer& ly : layers_) {
    int kper = 0;
    const int splits = ff_h1_splits(ly.Da, spec_.H, n_rows, part_n_, &kper);
    if ((int64_t)splits * rh > part_n_ || (int64_t)((ly.Db + 63) / 64) * rh > part_n_) return false;
  }
  return true;
}

void FlowNet::persistent_args(FpArgs& a, const int* idx, const int* n_dev, int n_rows) const {
  a = FpArgs{};
  const int D = spec_.D, L = spec_.L, H = spec_.H;
  for (int l = 0; l < L; ++l) {
    const Layer& ly = layers_[l];
    FpLayer& f = a.ly[l];
    f.ia = ly.ia; f.ib = ly.ib; f.Da = ly.Da; f.Db = ly.Db;
    f.splits = ff_h1_splits(ly.Da, H, n_rows, part_n_, &f.kper);
    f.W1 = params_ + ly.off[0]; f.b1 = params_ + ly.off[1]; f.W2 = params_ + ly.off[2];
    f.b2 = params_ + ly.off[3]; f.W3 = params_ + ly.off[4]; f.b3 = params_ + ly.off[5];
  }
  a.L = L; a.D = D; a.H = H; a.n = n_rows; a.tmax = (Dmax_b_ + 63) / 64;
  a.n_dev = n_dev; a.idx = idx;
  a.RD = (int64_t)R_ * D; a.RH = (int64_t)R_ * H; a.RS = (int64_t)R_ * Dmax_b_;
  a.U = U_; a.H1 = H1_; a.H2 = H2_; a.S = S_; a.part = part_; a.GC = GC_; a.GP1 = GP1_; a.GP2 = GP2_;
  a.ld_part = ld_part_; a.db_dev = db_dev_; a.loc = loc_; a.scale = scale_; a.log_scale_sum = log_scale_sum_;
  a.bar = bar_; a.err = bar_err_; a.spin_ticks = spin_ticks_;
}

void FlowNet::check_barriers() {
  if (!ran_persistent_) return;
  int e = 0;
  HIPF(hipMemcpyAsync(&e, bar_err_, sizeof(int), hipMemcpyDeviceToHost, st_));
  HIPF(hipStreamSynchronize(st_));
  if (e) {
    HIPF(hipMemsetAsync(bar_, 0, 2 * sizeof(unsigned), st_));
    HIPF(hipMemsetAsync(bar_err_, 0, sizeof(int), st_));
    HIPF(hipStreamSynchronize(st_));
    throw std::runtime_error("flow: a grid barrier of the persistent sampling kernels timed out");
  }
}

void FlowNet::forward(const double* Q, int64_t ldq, const int* idx, const int* n_dev, int n_rows, double* X,
                      double* logdet) {
  if (n_rows > R_) throw std::invalid_argument("flow: too many rows");
  const int D = spec_.D, L = spec_.L, H = spec_.H;
  const int64_t RD = (int64_t)R_ * D;
  if (!fused_ok(n_rows) && persistent_ok(n_rows)) {
    FpArgs a;
    persistent_args(a, idx, n_dev, n_rows);
    a.Q = Q; a.ldq = ldq; a.X = X; a.logdet = logdet;
    hipLaunchKernelGGL(flow_fwd_persistent, dim3(n_cu_), dim3(256), 0, st_, a);
    ran_persistent_ = true;
    return;
  }
  hipLaunchKernelGGL(flow_gather_z, grid1((int64_t)n_rows * D), dim3(256), 0, st_, Q, ldq, idx, n_dev, n_rows, D, U_);
  if (fused_ok(n_rows)) {
    const float* P = params_;
    const int tm = (n_rows + 63) / 64, th = (H + 63) / 64, tmax = (Dmax_b_ + 63) / 64;
    for (int l = 0; l < L; ++l) {
      const Layer& ly = layers_[l];
      float* U = U_ + l * RD;
      float* h1 = H1_ + (int64_t)l * R_ * H;
      float* h2 = H2_ + (int64_t)l * R_ * H;
      int kper = 0;
      const int splits = ff_h1_splits(ly.Da, H, n_rows, part_n_, &kper);
      hipLaunchKernelGGL(ff_h1, dim3(th, tm, splits), dim3(256), 0, st_, U, D, ly.ia, ly.Da, P + ly.off[0],
                         P + ly.off[1], H, n_dev, n_rows, kper, splits, h1, part_);
      hipLaunchKernelGGL(ff_h2, dim3(th, tm), dim3(256), 0, st_, part_, splits, P + ly.off[1], h1, P + ly.off[2],
                         P + ly.off[3], h2, H, n_dev, n_rows);
      hipLaunchKernelGGL(ff_couple, dim3(ff_couple_tiles(ly.Da, ly.Db), tm), dim3(256), 0, st_, h2, H, P + ly.off[4],
                         P + ly.off[5], ly.ia, ly.Da, ly.ib, ly.Db, D, U, U + RD, S_ + (int64_t)l * R_ * Dmax_b_,
                         (int64_t)ly.Db, ld_part_ + (int64_t)l * tmax * n_rows, n_dev, n_rows);
    }
    hipLaunchKernelGGL(ff_finish, dim3(n_rows), dim3(256), 0, st_, ld_part_, db_dev_, L, tmax, U_ + L * RD, loc_,
                       scale_, D, n_dev, n_rows, log_scale_sum_, logdet, X);
    return;
  }
  for (int l = 0; l < L; ++l) {
    const Layer& ly = layers_[l];
    float* U = U_ + l * RD;
    mlp(l, U, D, ly.ia, n_dev, n_rows);
    hipLaunchKernelGGL(flow_couple_fwd, grid1((int64_t)n_rows * D), dim3(256), 0, st_, U, O_, ly.ia, ly.ib, ly.Da,
                       ly.Db, D, n_dev, n_rows, S_ + (int64_t)l * R_ * Dmax_b_, U + RD);
  }
  hipLaunchKernelGGL(flow_logdet, dim3(n_rows), dim3(256), 0, st_, S_, (int64_t)R_ * Dmax_b_, db_dev_, L, n_dev,
                     n_rows, log_scale_sum_, logdet, 1);
  hipLaunchKernelGGL(flow_to_x, grid1((int64_t)n_rows * D), dim3(256), 0, st_, U_ + L * RD, loc_, scale_, D, n_dev,
                     n_rows, X);
}

void FlowNet::pullback(const double* GX, const double* VX, const double* logdet, const int* idx, const int* n_dev,
                       int n_rows, double* G, int64_t ldg, double* V) {
  const int D = spec_.D, L = spec_.L, H = spec_.H;
  const int64_t RD = (int64_t)R_ * D;
  if (!fused_ok(n_rows) && persistent_ok(n_rows)) {
    FpArgs a;
    persistent_args(a, idx, n_dev, n_rows);
    a.GX = GX; a.VX = VX; a.logdet_in = logdet; a.G = G; a.ldg = ldg; a.V = V;
    hipLaunchKernelGGL(flow_bwd_persistent, dim3(n_cu_), dim3(256), 0, st_, a);
    ran_persistent_ = true;
    return;
  }
  hipLaunchKernelGGL(flow_bwd_init, grid1((int64_t)n_rows * D), dim3(256), 0, st_, GX, scale_, D, n_dev, n_rows, GC_);
  if (fused_ok(n_rows)) {
    const float* P = params_;
    const int tm = (n_rows + 63) / 64, th = (H + 63) / 64;
    for (int l = L - 1; l >= 0; --l) {
      const Layer& ly = layers_[l];
      const float* h1 = H1_ + (int64_t)l * R_ * H;
      const float* h2 = H2_ + (int64_t)l * R_ * H;
      const int tb = (ly.Db + 63) / 64;
      hipLaunchKernelGGL(fb_couple, dim3(tb, tm), dim3(256), 0, st_, GC_, U_ + l * RD, S_ + (int64_t)l * R_ * Dmax_b_,
                         (int64_t)ly.Db, ly.ib, ly.Db, D, P + ly.off[4], H, part_, n_dev, n_rows);
      hipLaunchKernelGGL(fb_mlp, dim3(th, tm), dim3(256), 0, st_, part_, tb, h1, h2, P + ly.off[2], H, GP1_, n_dev,
                         n_rows);
      hipLaunchKernelGGL(fb_in, dim3((ly.Da + 63) / 64, tm), dim3(256), 0, st_, GP1_, H, P + ly.off[0], ly.ia, ly.Da,
                         D, GC_, n_dev, n_rows);
    }
    hipLaunchKernelGGL(flow_finish, grid1((int64_t)n_rows * D), dim3(256), 0, st_, GC_, VX, logdet, idx, n_dev,
                       n_rows, D, G, ldg, V);
    return;
  }
  for (int l = L - 1; l >= 0; --l) {
    const Layer& ly = layers_[l];
    const float* P = params_;
    const float* h1 = H1_ + (int64_t)l * R_ * H;
    const float* h2 = H2_ + (int64_t)l * R_ * H;
    hipLaunchKernelGGL(flow_couple_bwd_pull, grid1((int64_t)n_rows * ly.Db), dim3(256), 0, st_, GC_, U_ + l * RD,
                       S_ + (int64_t)l * R_ * Dmax_b_, ly.ib, ly.Db, D, n_dev, n_rows, GO_);
    // GP2 = (GO W3^T) (1 - h2^2); GP1 = (GP2 W2^T) (1 - h1^2); GUA = GP1 W1^T
    gemm(GO_, 2 * ly.Db, 1, P + ly.off[4], 1, 2 * ly.Db, GP2_, H, n_rows, H, 2 * ly.Db, n_dev, FG_EPI_DTANH, nullptr,
         h2, H);
    gemm(GP2_, H, 1, P + ly.off[2], 1, H, GP1_, H, n_rows, H, H, n_dev, FG_EPI_DTANH, nullptr, h1, H);
    // GC[:, ia] += GP1 W1^T
    gemm(GP1_, H, 1, P + ly.off[0], 1, H, GC_, D, n_rows, ly.Da, H, n_dev, FG_EPI_ADD, nullptr, nullptr, 0, nullptr,
         nullptr, ly.ia);
  }
  hipLaunchKernelGGL(flow_finish, grid1((int64_t)n_rows * D), dim3(256), 0, st_, GC_, VX, logdet, idx, n_dev, n_rows,
                     D, G, ldg, V);
}

// Inverse direction on U_[L] (standardized x, rows): U_[l] = f_l^{-1}(U_[l+1]),
// keeping H1, H2, S of every layer for backward_params().
void FlowNet::inverse_pass(int rows) {
  const int D = spec_.D, L = spec_.L;
  const int64_t RD = (int64_t)R_ * D;
  for (int l = L - 1; l >= 0; --l) {
    const Layer& ly = layers_[l];
    const float* Vin = U_ + (l + 1) * RD;
    mlp(l, Vin, D, ly.ia, nullptr, rows);
    hipLaunchKernelGGL(flow_couple_inv, grid1((int64_t)rows * D), dim3(256), 0, st_, Vin, O_, ly.ia, ly.ib, ly.Da,
                       ly.Db, D, rows, S_ + (int64_t)l * R_ * Dmax_b_, U_ + l * RD);
  }
}

// Row losses 0.5|z|^2 - log|det dT^{-1}/dx| into rowbuf_ (rows) after inverse_pass().
void FlowNet::row_losses(int rows) {
  hipLaunchKernelGGL(flow_logdet, dim3(rows), dim3(256), 0, st_, S_, (int64_t)R_ * Dmax_b_, db_dev_, spec_.L, nullptr,
                     rows, log_scale_sum_, rowbuf2_, -1);
  hipLaunchKernelGGL(flow_half_sq, dim3((rows + 3) / 4), dim3(256), 0, st_, U_, spec_.D, rows, rowbuf_);
  hipLaunchKernelGGL(flow_loss_rows, grid1(rows), dim3(256), 0, st_, rowbuf_, rowbuf2_, rows);
}

// Gradient of the mean row loss with respect to every parameter, into grad_.
void FlowNet::backward_params(int rows) {
  const int D = spec_.D, L = spec_.L, H = spec_.H;
  const int64_t RD = (int64_t)R_ * D;
  float* G = grad_;
  const float* P = params_;
  hipLaunchKernelGGL(flow_scale, grid1((int64_t)rows * D), dim3(256), 0, st_, U_, (int64_t)rows * D, 1.0f / rows, GC_);
  for (int l = 0; l < L; ++l) {
    const Layer& ly = layers_[l];
    const int Db2 = 2 * ly.Db;
    const float* h1 = H1_ + (int64_t)l * R_ * H;
    const float* h2 = H2_ + (int64_t)l * R_ * H;
    hipLaunchKernelGGL(flow_couple_bwd_train, grid1((int64_t)rows * ly.Db), dim3(256), 0, st_, GC_, U_ + l * RD,
                       S_ + (int64_t)l * R_ * Dmax_b_, ly.ib, ly.Db, D, rows, 1.0f / rows, GO_);
    // W3, b3
    gemm(h2, 1, H, GO_, Db2, 1, G + ly.off[4], Db2, H, Db2, rows, nullptr, FG_EPI_NONE, nullptr, nullptr, 0);
    hipLaunchKernelGGL(flow_colsum, grid1(Db2), dim3(256), 0, st_, GO_, rows, Db2, G + ly.off[5]);
    // GP2 = (GO W3^T)(1 - h2^2); W2, b2
    gemm(GO_, Db2, 1, P + ly.off[4], 1, Db2, GP2_, H, rows, H, Db2, nullptr, FG_EPI_DTANH, nullptr, h2, H);
    gemm(h1, 1, H, GP2_, H, 1, G + ly.off[2], H, H, H, rows, nullptr, FG_EPI_NONE, nullptr, nullptr, 0);
    hipLaunchKernelGGL(flow_colsum, grid1(H), dim3(256), 0, st_, GP2_, rows, H, G + ly.off[3]);
    // GP1 = (GP2 W2^T)(1 - h1^2); W1 from the layer's conditioner input, b1
    gemm(GP2_, H, 1, P + ly.off[2], 1, H, GP1_, H, rows, H, H, nullptr, FG_EPI_DTANH, nullptr, h1, H);
    // W1: A(j, r) = U_[l+1][r][ia[j]] (row gather on the conditioner columns)
    gemm(U_ + (l + 1) * RD, 1, D, GP1_, H, 1, G + ly.off[0], H, ly.Da, H, rows, nullptr, FG_EPI_NONE, nullptr, nullptr,
         0, nullptr, ly.ia);
    hipLaunchKernelGGL(flow_colsum, grid1(H), dim3(256), 0, st_, GP1_, rows, H, G + ly.off[1]);
    // cotangent of the conditioner input: GC[:, ia] += GP1 W1^T
    gemm(GP1_, H, 1, P + ly.off[0], 1, H, GC_, D, rows, ly.Da, H, nullptr, FG_EPI_ADD, nullptr, nullptr, 0, nullptr,
         nullptr, ly.ia);
  }
}

void FlowNet::map_rows(double* rows, int64_t ld, int n_rows) {
  if (n_rows > R_) throw std::invalid_argument("flow: too many rows");
  forward(rows + 7, ld, nullptr, nullptr, n_rows, xrows_, rowbuf_);
  hipLaunchKernelGGL(flow_store_rows, grid1((int64_t)n_rows * spec_.D), dim3(256), 0, st_, xrows_, rowbuf_, n_rows,
                     spec_.D, rows, ld);
}

void FlowNet::inverse(const double* X, int n_rows, double* Z, double* logdet) {
  if (n_rows > R_) throw std::invalid_argument("flow: too many rows");
  const int D = spec_.D, L = spec_.L;
  hipLaunchKernelGGL(flow_standardize, grid1((int64_t)n_rows * D), dim3(256), 0, st_, X, nullptr, loc_, scale_, D,
                     n_rows, U_ + (int64_t)L * R_ * D);
  inverse_pass(n_rows);
  hipLaunchKernelGGL(flow_rows_out, grid1((int64_t)n_rows * D), dim3(256), 0, st_, U_, (int64_t)n_rows * D, Z);
  if (logdet)
    hipLaunchKernelGGL(flow_logdet, dim3(n_rows), dim3(256), 0, st_, S_, (int64_t)R_ * Dmax_b_, db_dev_, L, nullptr,
                       n_rows, log_scale_sum_, logdet, -1);
}

double FlowNet::loss_grad(const double* X, int n_rows, float* grad_host) {
  if (n_rows > R_) throw std::invalid_argument("flow: too many rows");
  const int D = spec_.D, L = spec_.L;
  hipLaunchKernelGGL(flow_standardize, grid1((int64_t)n_rows * D), dim3(256), 0, st_, X, nullptr, loc_, scale_, D,
                     n_rows, U_ + (int64_t)L * R_ * D);
  inverse_pass(n_rows);
  row_losses(n_rows);
  backward_params(n_rows);
  std::vector<double> lr(n_rows);
  HIPF(hipMemcpyAsync(lr.data(), rowbuf_, n_rows * sizeof(double), hipMemcpyDeviceToHost, st_));
  if (grad_host) HIPF(hipMemcpyAsync(grad_host, grad_, n_params_ * sizeof(float), hipMemcpyDeviceToHost, st_));
  HIPF(hipStreamSynchronize(st_));
  double s = 0;
  for (double v : lr) s += v;
  return s / n_rows;
}

double FlowNet::train(const double* draws, int64_t n_draws, int steps, int batch, float lr, uint64_t seed,
                      int report) {
  if (batch > R_ || batch < 1 || n_draws < 1) throw std::invalid_argument("flow: bad training batch");
  const int D = spec_.D, L = spec_.L;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  double loss_sum = 0;
  int loss_n = 0;
  std::vector<double> lrow(batch);
  for (int t = 1; t <= steps; ++t) {
    hipLaunchKernelGGL(flow_pick_rows, grid1(batch), dim3(256), 0, st_, seed, (uint32_t)t, batch, n_draws, pick_);
    hipLaunchKernelGGL(flow_standardize, grid1((int64_t)batch * D), dim3(256), 0, st_, draws, pick_, loc_, scale_, D,
                       batch, U_ + (int64_t)L * R_ * D);
    inverse_pass(batch);
    backward_params(batch);
    adam_t_ += 1;
    const float c1 = 1.0f - std::pow(b1, (float)adam_t_), c2 = 1.0f - std::pow(b2, (float)adam_t_);
    hipLaunchKernelGGL(flow_adam, grid1(n_params_), dim3(256), 0, st_, params_, grad_, adam_m_, adam_v_, n_params_, lr,
                       b1, b2, eps, c1, c2);
    if (t > steps - report) {
      row_losses(batch);
      HIPF(hipMemcpyAsync(lrow.data(), rowbuf_, batch * sizeof(double), hipMemcpyDeviceToHost, st_));
      HIPF(hipStreamSynchronize(st_));
      for (double v : lrow) loss_sum += v;
      loss_n += batch;
    }
  }
  HIPF(hipStreamSynchronize(st_));
  return loss_n ? loss_sum / loss_n : 0.0;
}

}  // namespace nutsgpu
