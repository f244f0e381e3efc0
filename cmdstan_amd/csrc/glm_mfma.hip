// Batched logistic-regression gradient on FP64 matrix cores (gfx950).
//
// For a tile of 64 observations and 64 chains (4 waves x 16 chains):
//   phase 1  Eta = X_tile · Qᵀ            v_mfma_f64_16x16x4: M = rows, N = chains, K = D
//   pointwise R = sigmoid(Eta) - y, lp += y·Eta - log1p(e^Eta)   (in registers)
//   phase 2  Gᵀ += Rᵀ · X_tile             M = chains, N = D, K = rows
// The 16x16x4 f64 result layout puts element (4i + l/16, l%16) in register i of
// lane l.  Phase 2 walks the rows of the tile in the order (block rb, register
// i, lane group l/16), so the A operand of each phase-2 step is exactly register
// i of the phase-1 accumulator of block rb: R never leaves the registers.  The
// B operand of phase 2 reads X in that same row order from the LDS tile that
// phase 1 used.
//
// Work split: blocks own (row chunk, 64-chain group); the linear block id is
// remapped so that the chain groups sharing one row chunk run on the same XCD
// and share its L2.  Chunk partials go to a workspace reduced in fixed order
// (deterministic, no atomics).
#include "device_common.h"
#include "glm_split.h"
#include "launch.h"

namespace nutsgpu {

typedef double d4 __attribute__((ext_vector_type(4)));


#define GM_TN 64            // rows per tile
#define GM_TC 64            // chains per block
#define GM_KMAX 128         // D capacity (K of phase 1, N of phase 2)
#define GM_XS (GM_KMAX + 2) // LDS row stride in doubles (== 2 mod 32: conflict-free phase-1 reads)

template <int KMAX>
__global__ void __launch_bounds__(256, 2)
glm_mfma_f64(const double* __restrict__ X, const double* __restrict__ y, int N, int D, int total_blocks,
             const double* __restrict__ Q, int64_t ldq, int C, const int* __restrict__ idx,
             const int* __restrict__ n_active, double* __restrict__ Wg, double* __restrict__ Wlp) {
  __shared__ double sX[GM_TN * GM_XS];
  __shared__ double sY[GM_TN];
  const int tid = threadIdx.x;
  const int w = tid >> 6;
  const int l = tid & 63;
  const int lr = l & 15;   // m/n index inside a 16-wide tile
  const int lk = l >> 4;   // k index inside a 4-deep step

  // XCD-aware mapping: consecutive linear ids go to consecutive XCDs.
  const int n_act = idx ? *n_active : C;
  const GlmSplit sp = glm_split(n_act, N, total_blocks);
  const int lin = blockIdx.x;
  const int xcd = lin & 7;
  const int j = lin >> 3;
  const int chunk = xcd + 8 * (j / sp.groups);
  const int group = j % sp.groups;
  if (chunk >= sp.chunks || group * GM_TC >= n_act) return;

  const int c0 = group * GM_TC + 16 * w;  // this wave's 16 positions
  const int KS = (D + 3) >> 2;            // phase-1 k-steps
  const int DT = (D + 15) >> 4;           // phase-2 output tiles

  // phase-1 B operand for all k-steps: Q[c0 + lr][4 ks + lk]
  double qreg[KMAX / 4];
  {
    const int pos = c0 + lr;
    const int c = pos < n_act ? (idx ? idx[pos] : pos) : -1;
#pragma unroll
    for (int ks = 0; ks < KMAX / 4; ++ks) {
      const int k = 4 * ks + lk;
      qreg[ks] = (ks < KS && c >= 0 && k < D) ? Q[(int64_t)c * ldq + k] : 0.0;
    }
  }
  d4 gacc[KMAX / 16];
#pragma unroll
  for (int t = 0; t < KMAX / 16; ++t) gacc[t] = d4{0, 0, 0, 0};
  double lp = 0.0;

  const int n_begin = chunk * sp.rows;
  const int n_end = min(N, n_begin + sp.rows);
  const int ncols = DT * 16;

  for (int nb = n_begin; nb < n_end; nb += GM_TN) {
    __syncthreads();
    // 16 lanes cover 16 consecutive columns of a row: 128-byte segments
#pragma unroll
    for (int m = 0; m < GM_TN / 16; ++m) {
      const int r = (tid >> 4) + 16 * m;
      const int n = nb + r;
      const bool rowok = n < n_end;
      const double* xr = X + (int64_t)n * D;
#pragma unroll
      for (int jj = 0; jj < KMAX / 16; ++jj) {
        const int k = (tid & 15) + 16 * jj;
        if (k < ncols) sX[r * GM_XS + k] = (rowok && k < D) ? xr[k] : 0.0;
      }
    }
    if (tid < GM_TN) sY[tid] = (nb + tid < n_end) ? y[nb + tid] : 0.0;
    __syncthreads();

    d4 eta[4];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) eta[rb] = d4{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < KMAX / 4; ++ks) {
      if (ks < KS) {
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
          const double a = sX[(16 * rb + lr) * GM_XS + 4 * ks + lk];
          eta[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, qreg[ks], eta[rb], 0, 0, 0);
        }
      }
    }
    // R = sigmoid(eta) - y in place; rows past the chunk contribute nothing
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int rr = 16 * rb + 4 * i + lk;
        const double e = eta[rb][i];
        double r = 0.0;
        if (nb + rr < n_end) {
          // one exp: log1p_exp(e) = max(e,0) + log1p(t), sigmoid(e) from t = e^{-|e|}
          const double yy = sY[rr];
          const double t = exp(-fabs(e));
          double r1;
          const double l1p = log1p_recip_unit(t, r1);
          lp += yy * e - ((e > 0.0 ? e : 0.0) + l1p);
          r = (e >= 0.0 ? r1 : t * r1) - yy;
        }
        eta[rb][i] = r;
      }
    }
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const double a = eta[rb][i];
        const double* xrow = sX + (16 * rb + 4 * i + lk) * GM_XS + lr;
#pragma unroll
        for (int t = 0; t < KMAX / 16; ++t) {
          if (t < DT) gacc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, xrow[16 * t], gacc[t], 0, 0, 0);
        }
      }
    }
  }

  // gacc[t] register i of lane l = G[chain c0 + 4i + lk][d 16t + lr]
#pragma unroll
  for (int t = 0; t < KMAX / 16; ++t) {
    if (t < DT) {
      const int d = 16 * t + lr;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int pos = c0 + 4 * i + lk;
        if (pos < n_act && d < D) Wg[((int64_t)chunk * sp.stride + pos) * D + d] = gacc[t][i];
      }
    }
  }
  // lp: lanes l, l+16, l+32, l+48 hold the same chain
  lp += __shfl_xor(lp, 16, 64);
  lp += __shfl_xor(lp, 32, 64);
  if (lk == 0 && c0 + lr < n_act) Wlp[(int64_t)chunk * sp.stride + c0 + lr] = lp;
}

// Even D: the same products on 32-row tiles double-buffered in LDS.  The next
// tile's rows are copied global -> LDS by global_load_lds_dwordx4 (16 bytes
// per lane, lane l at row base + 16 l, exec-masked lanes leave their slots
// alone -- checked on the device), one instruction per row, issued before the
// current tile's products and waited for (vmcnt, then the barrier) after
// them: the copy needs no registers and overlaps the matrix work.  Columns
// past D stay zero from the initial clear; rows past the chunk hold finite
// data of an earlier tile and contribute nothing (their residual is 0).  Rows
// are summed in the order of the 64-row kernel, so the products match it bit
// for bit; the per-observation terms use two reciprocals instead of divisions.
typedef __attribute__((address_space(3))) void lds_void;
#define GL_TN 32  // rows per double-buffered tile

template <int KMAX>
__global__ void __launch_bounds__(256, 2)
glm_mfma_f64_lds(const double* __restrict__ X, const double* __restrict__ y, int N, int D, int total_blocks,
                 const double* __restrict__ Q, int64_t ldq, int C, const int* __restrict__ idx,
                 const int* __restrict__ n_active, double* __restrict__ Wg, double* __restrict__ Wlp) {
  __shared__ __attribute__((aligned(16))) double sX[2][GL_TN * GM_XS];
  __shared__ double sY[2][GL_TN];
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
  if (chunk >= sp.chunks || group * GM_TC >= n_act) return;

  const int c0 = group * GM_TC + 16 * w;
  const int KS = (D + 3) >> 2;
  const int DT = (D + 15) >> 4;
  double qreg[KMAX / 4];
  {
    const int pos = c0 + lr;
    const int c = pos < n_act ? (idx ? idx[pos] : pos) : -1;
#pragma unroll
    for (int ks = 0; ks < KMAX / 4; ++ks) {
      const int k = 4 * ks + lk;
      qreg[ks] = (ks < KS && c >= 0 && k < D) ? Q[(int64_t)c * ldq + k] : 0.0;
    }
  }
  d4 gacc[KMAX / 16];
#pragma unroll
  for (int t = 0; t < KMAX / 16; ++t) gacc[t] = d4{0, 0, 0, 0};
  double lp = 0.0;

  const int n_begin = chunk * sp.rows;
  const int n_end = min(N, n_begin + sp.rows);
  for (int i = tid; i < 2 * GL_TN * GM_XS; i += 256) (&sX[0][0])[i] = 0.0;
  const int nch = D >> 1;  // 16-byte chunks per row
  // wave w copies rows 8 w .. 8 w + 7 of the tile
  auto issue = [&](int nb, int b) {
#pragma unroll
    for (int rr = 0; rr < GL_TN / 4; ++rr) {
      const int r = w * (GL_TN / 4) + rr;
      const int n = nb + r;
      if (n < n_end && l < nch)
        __builtin_amdgcn_global_load_lds(X + (int64_t)n * D + 2 * l, (lds_void*)&sX[b][r * GM_XS], 16, 0, 0);
    }
  };
  __syncthreads();
  if (tid < GL_TN) sY[0][tid] = (n_begin + tid < n_end) ? y[n_begin + tid] : 0.0;
  issue(n_begin, 0);
  int b = 0;
  for (int nb = n_begin; nb < n_end; nb += GL_TN, b ^= 1) {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    const int nxt = nb + GL_TN;
    double yn = 0.0;
    if (nxt < n_end) {
      issue(nxt, b ^ 1);
      if (tid < GL_TN) yn = (nxt + tid < n_end) ? y[nxt + tid] : 0.0;
    }
    const double* sx = sX[b];
    d4 eta[2];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) eta[rb] = d4{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < KMAX / 4; ++ks) {
      if (ks < KS) {
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) {
          const double a = sx[(16 * rb + lr) * GM_XS + 4 * ks + lk];
          eta[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, qreg[ks], eta[rb], 0, 0, 0);
        }
      }
    }
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int rr = 16 * rb + 4 * i + lk;
        const double e = eta[rb][i];
        double r = 0.0;
        if (nb + rr < n_end) {
          // t = e^{-|e|}: log1p_exp(e) = max(e,0) + log1p(t), sigmoid(e) from 1 / (1 + t)
          const double yy = sY[b][rr];
          const double t = exp(-fabs(e));
          double r1;
          const double l1p = log1p_recip_unit(t, r1);
          lp += yy * e - ((e > 0.0 ? e : 0.0) + l1p);
          r = (e >= 0.0 ? r1 : t * r1) - yy;
        }
        eta[rb][i] = r;
      }
    }
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const double a = eta[rb][i];
        const double* xrow = sx + (16 * rb + 4 * i + lk) * GM_XS + lr;
#pragma unroll
        for (int t = 0; t < KMAX / 16; ++t) {
          if (t < DT) gacc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, xrow[16 * t], gacc[t], 0, 0, 0);
        }
      }
    }
    if (tid < GL_TN) sY[b ^ 1][tid] = yn;
  }

#pragma unroll
  for (int t = 0; t < KMAX / 16; ++t) {
    if (t < DT) {
      const int d = 16 * t + lr;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int pos = c0 + 4 * i + lk;
        if (pos < n_act && d < D) Wg[((int64_t)chunk * sp.stride + pos) * D + d] = gacc[t][i];
      }
    }
  }
  lp += __shfl_xor(lp, 16, 64);
  lp += __shfl_xor(lp, 32, 64);
  if (lk == 0 && c0 + lr < n_act) Wlp[(int64_t)chunk * sp.stride + c0 + lr] = lp;
}

__global__ void glm_mfma_reduce(const double* __restrict__ Wg, const double* __restrict__ Wlp, int N,
                                int total_blocks, int C, int D, double s2inv, const double* __restrict__ Q,
                                int64_t ldq, const int* __restrict__ idx, const int* __restrict__ n_active,
                                double* __restrict__ G, double* __restrict__ V) {
  const int pos = blockIdx.x;
  const int n_act = idx ? *n_active : C;
  if (pos >= n_act) return;
  const GlmSplit sp = glm_split(n_act, N, total_blocks);
  const int c = idx ? idx[pos] : pos;
  const double* q = Q + (int64_t)c * ldq;
  double* g = G + (int64_t)c * ldq;
  for (int k = threadIdx.x; k < D; k += blockDim.x) {
    double s = 0.0;
    // chunk order kept; sixteen loads in flight instead of one
#pragma unroll 16
    for (int ch = 0; ch < sp.chunks; ++ch) s += Wg[((int64_t)ch * sp.stride + pos) * D + k];
    g[k] = s + s2inv * q[k];
  }
  if (threadIdx.x < 64) {
    double lpv = 0.0, pr = 0.0;
    for (int ch = threadIdx.x; ch < sp.chunks; ch += 64) lpv += Wlp[(int64_t)ch * sp.stride + pos];
    for (int k = threadIdx.x; k < D; k += 64) pr += q[k] * q[k];
    lpv = wave_sum(lpv);
    pr = wave_sum(pr);
    if (threadIdx.x == 0) V[c] = -(lpv - 0.5 * s2inv * pr);
  }
}

// Active-chain list for the lockstep gradient: idx[0..n) = chains with need[c]
// set, in chain order (single-block scan, deterministic); clears need[].
__global__ void __launch_bounds__(1024) compact_active(int* __restrict__ need, int C, int* __restrict__ idx,
                                                       int* __restrict__ n_active, long long* __restrict__ tally) {
  __shared__ int wsum[16];
  __shared__ int base;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  if (t == 0) base = 0;
  __syncthreads();
  for (int c0 = 0; c0 < C; c0 += 1024) {
    const int c = c0 + t;
    const int f = (c < C && need[c]) ? 1 : 0;
    if (c < C) need[c] = 0;
    int x = f;  // inclusive wave scan
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    int off = base;
    for (int k = 0; k < wv; ++k) off += wsum[k];
    if (f) idx[off + x - 1] = c;
    __syncthreads();
    if (t == 1023) {
      int tot = 0;
      for (int k = 0; k < 16; ++k) tot += wsum[k];
      base += tot;
    }
    __syncthreads();
  }
  if (t == 0) {
    *n_active = base;
    if (tally) *tally += base;
  }
}

void launch_compact_active(int* need, int C, int* idx, int* n_active, long long* tally, hipStream_t st) {
  hipLaunchKernelGGL(compact_active, dim3(1), dim3(1024), 0, st, need, C, idx, n_active, tally);
}

GlmPlan glm_mfma_plan(int N, int C, int D, int num_cus, int blocks_per_cu) {
  GlmPlan p;
  // blocks_per_cu resident blocks per CU (2 unless tuned), and at least 8 row
  // chunks per chain group
  const int groups = (C + GM_TC - 1) / GM_TC;
  int tb = (blocks_per_cu > 0 ? blocks_per_cu : 2) * (num_cus > 0 ? num_cus : 256);
  if (tb < 8 * groups) tb = 8 * groups;
  p.total_blocks = (tb + 7) & ~7;
  const GlmSplit full = glm_split(C, N, p.total_blocks);
  p.ngroups = full.groups;
  p.nchunks = full.chunks;
  p.rows_per_chunk = full.rows;
  // partials of any launch: chunks * stride <= total_blocks * 64 positions
  p.work_doubles = (size_t)p.total_blocks * GM_TC * D + (size_t)p.total_blocks * GM_TC;
  return p;
}

int glm_grad_mfma_f64(const GlmArgs& a, const GlmPlan& p, hipStream_t st) {
  if (a.D > GM_KMAX || a.D < 1) return 1;
  double* Wg = a.work;
  double* Wlp = a.work + (size_t)p.total_blocks * GM_TC * a.D;
  // even D: rows are whole 16-byte chunks, so the double-buffered LDS copy applies
#define GM_LAUNCH(K)                                                                                                \
  do {                                                                                                              \
    if ((a.D & 1) == 0)                                                                                             \
      hipLaunchKernelGGL(glm_mfma_f64_lds<K>, dim3(p.total_blocks), dim3(256), 0, st, a.X, a.y, a.N, a.D,           \
                         p.total_blocks, a.Q, a.ldq, a.C, a.idx, a.n_active, Wg, Wlp);                              \
    else                                                                                                            \
      hipLaunchKernelGGL(glm_mfma_f64<K>, dim3(p.total_blocks), dim3(256), 0, st, a.X, a.y, a.N, a.D,               \
                         p.total_blocks, a.Q, a.ldq, a.C, a.idx, a.n_active, Wg, Wlp);                              \
  } while (0)
  if (a.ev_begin) (void)hipEventRecord(a.ev_begin, st);
  // operand registers scale with the D bucket (a 64 bucket spilled registers:
  // 33 <= D <= 112 share the 112 instance, whose extra operands are zero)
  if (a.D <= 16) GM_LAUNCH(16);
  else if (a.D <= 32) GM_LAUNCH(32);
  else if (a.D <= 112) GM_LAUNCH(112);
  else GM_LAUNCH(128);
#undef GM_LAUNCH
  if (a.ev_end) (void)hipEventRecord(a.ev_end, st);
  glm_reduce_launch(a, p, st);
  return 0;
}

void glm_reduce_launch(const GlmArgs& a, const GlmPlan& p, hipStream_t st) {
  double* Wg = a.work;
  double* Wlp = a.work + (size_t)p.total_blocks * GM_TC * a.D;
  hipLaunchKernelGGL(glm_mfma_reduce, dim3(a.C), dim3(128), 0, st, Wg, Wlp, a.N, p.total_blocks, a.C, a.D, a.s2inv,
                     a.Q, a.ldq,
                     a.idx, a.n_active, a.G, a.V);
}

}  // namespace nutsgpu
