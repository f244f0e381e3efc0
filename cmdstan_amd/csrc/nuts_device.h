// Batched NUTS as a per-chain state machine, one wavefront per chain.
//
// The upstream recursion (stan::mcmc::base_nuts::transition / build_tree, not in
// the reference tree; entered from cmdstan src/cmdstan/command.hpp:603-697) is
// unrolled into an iteration over the leaves of each doubling.  After leaf i of
// a depth-d subtree, the subtrees that close at that leaf are merged bottom-up:
// level k merges while bit k of i is set, exactly the post-order in which the
// recursion finishes its calls, so the uniforms are consumed in the same order
// and the same three U-turn checks run per merge.  Open left siblings live in
// per-level checkpoints (first/last momentum, rho, log weight, proposal slot).
//
// A chain stops at every point where it needs the potential V(q) and its
// gradient: in the lockstep mode the gradient for all chains is computed by a
// separate kernel (GEMM for GLMs); in the fused mode a model device function is
// called in place and the chain keeps going for a bounded number of gradients.
//
// All threads of a chain execute the same scalar control flow; vectors are
// strided over the chain's threads and reductions give every thread the same
// value, so every thread holds identical scalar state (the staged ChainState
// in LDS is written with identical values by all lanes of a wave).
//
// Metrics (stan::mcmc::{unit,diag,dense}_e_metric): the "sharp" momentum
// Minv·p enters the leapfrog and the U-turn checks.  For unit/diag it is formed
// on the fly; for dense_e every momentum kept for a U-turn check has a stored
// sharp companion, computed once per leaf by a matvec with the per-chain D×D
// inverse metric, and momenta p = L⁻ᵀ z are drawn with the chain's Cholesky
// factor Minv = L Lᵀ, whose buffer also holds L⁻¹ (see dense_factor).
#pragma once
#include "device_common.h"
#include "nuts_types.h"

namespace nutsgpu {

struct ExternalGrad {
  static constexpr bool fused = false;
  template <class T>
  __device__ double operator()(int, const T*, T*, int, int) const { return 0.0; }
};

// dense_e keeps the normal draws of one momentum in an LDS strip of the wave
// while it is formed: D <= 64 * DENSE_RMAX
#define DENSE_RMAX 16

// Makes this wave's earlier global writes visible to later reads by other
// lanes of the same wave (vectors are written lane-strided and then read whole
// by every lane in matvecs, the Welford covariance update, the Cholesky
// columns and the model gradients).  Work-group scope is enough: writer and
// reader run on one CU and share its vector L1, so no L2 write-back or L1
// invalidation is needed (kernels are not built in thread-group split mode).
// Orders LDS writes of this wave before later LDS reads by its other lanes.
__device__ __forceinline__ void lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// TPC threads per chain: 64 (one wave) or a multiple of 64 (one block per
// chain, for large D).  With several waves every wave keeps its own identical
// copy of the scalar state; vectors are strided over the TPC threads, so an
// element is always touched by the same thread, and cross-thread reads of
// vectors (reductions, matvecs, factorisations) go through gsum/gsync, which
// every wave reaches at the same point because the control flow is identical.
// Lane-strided vector passes of a leaf.  The vectors arrive as
// restrict-qualified arguments (pool vectors, Q and G never overlap; two
// read-only arguments may name the same vector), so the unrolled loops issue
// the loads of several elements before their stores instead of waiting one
// memory round trip per element.  Every element keeps the arithmetic and the
// per-thread accumulation order of the step-by-step form.
template <int METRIC, class VT>
__device__ __forceinline__ double sharp(const VT* __restrict__ x, const VT* __restrict__ xs,
                                        const VT* __restrict__ minv, int i) {
  if (METRIC == METRIC_DENSE) return xs[i];
  return METRIC == METRIC_UNIT ? (double)x[i] : (double)minv[i] * (double)x[i];
}

// unit/diag: end_update_p(eps/2), then the kinetic energy's per-thread partial
// of p·Minv p, with the new p also copied to the subtree's rho and, at the
// first leaf of a subtree, to its edge momentum
template <int TPC, bool UNIT, bool EDGE, class VT>
__device__ __forceinline__ double pass_leaf_head(VT* __restrict__ p, const VT* __restrict__ g,
                                                 const VT* __restrict__ minv, VT* __restrict__ rho,
                                                 VT* __restrict__ edge, int D, int tid, double h) {
  double a = 0;
#pragma unroll 4
  for (int i = tid; i < D; i += TPC) {
    const VT pv = (double)p[i] - h * (double)g[i];
    p[i] = pv;
    rho[i] = pv;
    if (EDGE) edge[i] = pv;
    const double pi = pv;
    a += pi * (UNIT ? pi : (double)minv[i] * pi);
  }
  return a;
}

// unit/diag: the closed subtree's checkpoint (first and last momentum, rho),
// the leaf's (q, g) proposal slot when it is kept, then begin_update_p(eps/2)
// and update_q(eps) of the next leaf
template <int TPC, bool UNIT, bool BEG, bool SLOT, class VT>
__device__ __forceinline__ void pass_leaf_tail(VT* __restrict__ p, VT* __restrict__ q, const VT* __restrict__ g,
                                               const VT* __restrict__ minv, const VT* __restrict__ rho,
                                               const VT* __restrict__ beg, VT* __restrict__ lbeg,
                                               VT* __restrict__ lend, VT* __restrict__ lrho, VT* __restrict__ sq,
                                               VT* __restrict__ sg, VT* __restrict__ sp, int D, int tid, double h,
                                               double e) {
#pragma unroll 4
  for (int i = tid; i < D; i += TPC) {
    const VT po = p[i], qo = q[i], gi = g[i];
    lbeg[i] = BEG ? beg[i] : po;
    lend[i] = po;
    lrho[i] = rho[i];
    if (SLOT) {
      sq[i] = qo;
      sg[i] = gi;
      if (sp) sp[i] = po;
    }
    const VT pn = (double)po - h * (double)gi;
    p[i] = pn;
    q[i] = (double)qo + e * ((UNIT ? 1.0 : (double)minv[i]) * (double)pn);
  }
}

// unit/diag: end_update_p(eps/2) of a leaf with the kinetic energy's partial,
// and in the same pass the partial sums of the generalized U-turn checks of
// the first F merges that close at this leaf (levels 0..F-1: begin lb[k], end
// le[k], rho lr[k] of the open left siblings; the closing subtree's begin is
// the leaf for level 0 and lb[k-1] above).  The leaf's momentum goes to dp and
// the merged rho to dr -- where they are read next -- with optional copies to
// db (the leaf as its own subtree's begin) and edge (first leaf of a doubling).
// Every sum follows the separate passes term for term.
template <int TPC, bool UNIT, int F, class VT>
__device__ __forceinline__ void pass_leaf_merge(const VT* __restrict__ p, const VT* __restrict__ g,
                                                const VT* __restrict__ minv, const VT* const (&lb)[3],
                                                const VT* const (&le)[3], const VT* const (&lr)[3],
                                                VT* __restrict__ dp, VT* __restrict__ dr, VT* __restrict__ db,
                                                VT* __restrict__ edge, int D, int tid, double h, double* acc) {
  double kin = 0;
  double a[F > 0 ? F : 1][6];
#pragma unroll
  for (int k = 0; k < F; ++k)
#pragma unroll
    for (int j = 0; j < 6; ++j) a[k][j] = 0;
#pragma unroll 1
  for (int i = tid; i < D; i += TPC) {
    const VT pv = (double)p[i] - h * (double)g[i];
    const double pd = pv;
    const double m = UNIT ? 1.0 : (double)minv[i];
    const double sp = UNIT ? pd : m * pd;
    kin += pd * sp;
    double r = pd;  // rho of the closing subtree, as stored (VT)
    double sbv = pd, sbs = sp;
#pragma unroll
    for (int k = 0; k < F; ++k) {
      const double l = lr[k][i], lbv = lb[k][i], lev = le[k][i];
      const VT rsv = l + r;
      const double rs = rsv;
      const double r2 = l + sbv;
      const double r3 = r + lev;
      const double slb = UNIT ? lbv : m * lbv;
      const double sle = UNIT ? lev : m * lev;
      a[k][0] += slb * rs; a[k][1] += sp * rs;
      a[k][2] += slb * r2; a[k][3] += sbs * r2;
      a[k][4] += sle * r3; a[k][5] += sp * r3;
      r = rs;
      sbv = lbv;
      sbs = slb;
    }
    dp[i] = pv;
    dr[i] = (VT)r;
    if (db) db[i] = pv;
    if (edge) edge[i] = pv;
  }
  acc[0] = kin;
#pragma unroll
  for (int k = 0; k < F; ++k)
#pragma unroll
    for (int j = 0; j < 6; ++j) acc[1 + 6 * k + j] = a[k][j];
}

// unit/diag: the leaf's (q, g, p) proposal slot when it is kept, then
// begin_update_p(eps/2) and update_q(eps) of the next leaf from the closed
// subtree's end momentum po
template <int TPC, bool UNIT, bool SLOT, class VT>
__device__ __forceinline__ void pass_leaf_next(const VT* __restrict__ po_src, VT* __restrict__ p, VT* __restrict__ q,
                                               const VT* __restrict__ g, const VT* __restrict__ minv,
                                               VT* __restrict__ sq, VT* __restrict__ sg, VT* __restrict__ sp, int D,
                                               int tid, double h, double e) {
#pragma unroll 4
  for (int i = tid; i < D; i += TPC) {
    const VT po = po_src[i], qo = q[i], gi = g[i];
    if (SLOT) {
      sq[i] = qo;
      sg[i] = gi;
      if (sp) sp[i] = po;
    }
    const VT pn = (double)po - h * (double)gi;
    p[i] = pn;
    q[i] = (double)qo + e * ((UNIT ? 1.0 : (double)minv[i]) * (double)pn);
  }
}

// the three generalized U-turn checks of a merge inside a subtree (rho of the
// merged subtree written back into srho): partial dot products per thread
template <int TPC, int METRIC, class VT>
__device__ __forceinline__ void pass_merge(const VT* __restrict__ lb, const VT* __restrict__ le,
                                           const VT* __restrict__ lr, const VT* __restrict__ sb,
                                           const VT* __restrict__ p, VT* __restrict__ srho,
                                           const VT* __restrict__ minv, const VT* __restrict__ lbs,
                                           const VT* __restrict__ les, const VT* __restrict__ sbs,
                                           const VT* __restrict__ ps, int D, int tid, double* acc) {
  double a1 = 0, b1 = 0, a2 = 0, b2 = 0, a3 = 0, b3 = 0;
#pragma unroll 4
  for (int i = tid; i < D; i += TPC) {
    const double l = lr[i], r = srho[i];
    const VT rsv = l + r;
    const double rs = rsv;
    const double r2 = l + (double)sb[i];
    const double r3 = r + (double)le[i];
    const double slb = sharp<METRIC>(lb, lbs, minv, i), sp = sharp<METRIC>(p, ps, minv, i);
    a1 += slb * rs; b1 += sp * rs;
    a2 += slb * r2; b2 += sharp<METRIC>(sb, sbs, minv, i) * r2;
    a3 += sharp<METRIC>(le, les, minv, i) * r3; b3 += sp * r3;
    srho[i] = rsv;
  }
  acc[0] = a1; acc[1] = b1; acc[2] = a2; acc[3] = b2; acc[4] = a3; acc[5] = b3;
}

// the same checks for the top-level merge of a doubling (rho of the whole
// trajectory updated in place)
template <int TPC, int METRIC, class VT>
__device__ __forceinline__ void pass_doubling(bool fwd, VT* __restrict__ rho, const VT* __restrict__ srho,
                                              const VT* __restrict__ bb, const VT* __restrict__ ff,
                                              const VT* __restrict__ pfb, const VT* __restrict__ pbf,
                                              const VT* __restrict__ minv, const VT* __restrict__ bbs,
                                              const VT* __restrict__ ffs, const VT* __restrict__ psfb,
                                              const VT* __restrict__ psbf, int D, int tid, double* acc) {
  double a1 = 0, b1 = 0, a2 = 0, b2 = 0, a3 = 0, b3 = 0;
#pragma unroll 4
  for (int i = tid; i < D; i += TPC) {
    const double old = rho[i], nw = srho[i];
    const VT rnv = fwd ? old + nw : nw + old;  // rho_bck + rho_fwd
    const double rn = rnv;
    const double rb = fwd ? old : nw, rf = fwd ? nw : old;
    const double r2 = rb + (double)pfb[i];
    const double r3 = rf + (double)pbf[i];
    const double sbb = sharp<METRIC>(bb, bbs, minv, i), sff = sharp<METRIC>(ff, ffs, minv, i);
    a1 += sbb * rn; b1 += sff * rn;
    a2 += sbb * r2; b2 += sharp<METRIC>(pfb, psfb, minv, i) * r2;
    a3 += sharp<METRIC>(pbf, psbf, minv, i) * r3; b3 += sff * r3;
    rho[i] = rnv;
  }
  acc[0] = a1; acc[1] = b1; acc[2] = a2; acc[3] = b2; acc[4] = a3; acc[5] = b3;
}

// out = M x for the column-major D x D matrix M, summing j = 0..D-1 in order
// like Eigen's row product.  Kept out of line: inside the state machine the
// register pressure made the compiler wait for each column element before
// loading the next; here the loads of four columns issue together.  Rows past
// D read a valid element (row D-1) and are never stored, so the loads carry
// no branches.
template <int TPC, class VT>
__device__ __noinline__ void dense_matvec(const double* __restrict__ M, const VT* __restrict__ x,
                                          VT* __restrict__ out, int D, int tid) {
  constexpr int RB = TPC >= 256 ? 4 : 8;  // TPC * RB rows per pass (dense_e: D <= 1024)
  for (int i0 = 0; i0 < D; i0 += TPC * RB) {
    double acc[RB];
    int ii[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      acc[r] = 0.0;
      const int i = i0 + tid + TPC * r;
      ii[r] = i < D ? i : D - 1;
    }
    int j = 0;
    for (; j + 4 <= D; j += 4) {
      double v[4][RB], xv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        xv[u] = x[j + u];
        const double* col = M + (int64_t)(j + u) * D;
#pragma unroll
        for (int r = 0; r < RB; ++r) v[u][r] = col[ii[r]];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int r = 0; r < RB; ++r) acc[r] += v[u][r] * xv[u];
    }
    for (; j < D; ++j) {
      const double xj = x[j];
      const double* col = M + (int64_t)j * D;
#pragma unroll
      for (int r = 0; r < RB; ++r) acc[r] += col[ii[r]] * xj;
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int i = i0 + tid + TPC * r;
      if (i < D) out[i] = acc[r];
    }
  }
}

// out1 = M x1 and out2 = M x2 in one pass over M (same order of summation as
// dense_matvec): the leaf's sharp momentum and M g share the read of the metric
template <int TPC, class VT>
__device__ __noinline__ void dense_matvec2(const double* __restrict__ M, const VT* __restrict__ x1,
                                           const VT* __restrict__ x2, VT* __restrict__ out1, VT* __restrict__ out2,
                                           int D, int tid) {
  constexpr int RB = TPC >= 256 ? 4 : 8;
  for (int i0 = 0; i0 < D; i0 += TPC * RB) {
    double a1[RB], a2[RB];
    int ii[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      a1[r] = 0.0;
      a2[r] = 0.0;
      const int i = i0 + tid + TPC * r;
      ii[r] = i < D ? i : D - 1;
    }
    int j = 0;
    for (; j + 4 <= D; j += 4) {
      double v[4][RB], y1[4], y2[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        y1[u] = x1[j + u];
        y2[u] = x2[j + u];
        const double* col = M + (int64_t)(j + u) * D;
#pragma unroll
        for (int r = 0; r < RB; ++r) v[u][r] = col[ii[r]];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int r = 0; r < RB; ++r) {
          a1[r] += v[u][r] * y1[u];
          a2[r] += v[u][r] * y2[u];
        }
    }
    for (; j < D; ++j) {
      const double y1j = x1[j], y2j = x2[j];
      const double* col = M + (int64_t)j * D;
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const double m = col[ii[r]];
        a1[r] += m * y1j;
        a2[r] += m * y2j;
      }
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int i = i0 + tid + TPC * r;
      if (i < D) {
        out1[i] = a1[r];
        out2[i] = a2[r];
      }
    }
  }
}

// p = L⁻ᵀ z and ps = Minv p = L Lᵀ L⁻ᵀ z = L z in one pass over the packed
// factor F (column-major D×D: L on and below the diagonal, row r of L⁻¹ in
// column r above it, so F[j][i] (column j, row i < j) = (L⁻¹)[j][i] =
// (L⁻ᵀ)[i][j]).  z is in LDS; the lane owning row i sums column j's element
// into p_i when i < j and into ps_i when i > j; the diagonal terms z_i / L_ii
// and L_ii z_i come last.  Every element of F is read once, columns
// coalesced, with no sequential dependency between rows.  Inlined, not an
// out-of-line call like the matvecs: z lives in the caller's LDS strip, and
// an out-of-line copy that every caller passes the same strip was built with
// that strip's address folded in as LDS offset 0, which is not where the
// calling kernel placed it.
template <int TPC, class VT>
__device__ __forceinline__ void dense_tri_sample(const double* __restrict__ F, const double* __restrict__ z,
                                              VT* __restrict__ p, VT* __restrict__ ps, int D, int tid) {
  constexpr int RB = 4;  // rows per lane per pass: inlined, so few registers
  for (int i0 = 0; i0 < D; i0 += TPC * RB) {
    double a1[RB], a2[RB];
    int ii[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      a1[r] = 0.0;
      a2[r] = 0.0;
      const int i = i0 + tid + TPC * r;
      ii[r] = i < D ? i : D - 1;
    }
    int j = 0;
    for (; j + 4 <= D; j += 4) {
      double v[4][RB], zz[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        zz[u] = z[j + u];
        const double* col = F + (int64_t)(j + u) * D;
#pragma unroll
        for (int r = 0; r < RB; ++r) v[u][r] = col[ii[r]];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int r = 0; r < RB; ++r) {
          const double t = v[u][r] * zz[u];
          a1[r] += ii[r] < j + u ? t : 0.0;
          a2[r] += ii[r] > j + u ? t : 0.0;
        }
    }
    for (; j < D; ++j) {
      const double zj = z[j];
      const double* col = F + (int64_t)j * D;
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const double t = col[ii[r]] * zj;
        a1[r] += ii[r] < j ? t : 0.0;
        a2[r] += ii[r] > j ? t : 0.0;
      }
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int i = i0 + tid + TPC * r;
      if (i < D) {
        const double lii = F[(int64_t)i * D + i];
        p[i] = a1[r] + z[i] / lii;
        ps[i] = a2[r] + lii * z[i];
      }
    }
  }
}

// col += w * d (the dense Welford covariance update, one column per call)
template <int TPC, class VT>
__device__ __forceinline__ void pass_axpy(double* __restrict__ col, const VT* __restrict__ w, double d, int D,
                                          int tid) {
#pragma unroll 4
  for (int i = tid; i < D; i += TPC) col[i] = col[i] + (double)w[i] * d;
}

// X = L⁻¹ (lower triangular) into the strictly upper part of the factor's
// buffer, row r of X in column r: X[r][c] (c < r) at L[r D + c].  Row r by
// forward substitution, X[r][c] = -(Σ_{k=c}^{r-1} L[r][k] X[k][c]) / L[r][r]
// with X[c][c] = 1 / L[c][c]; the lane owning column c reads only its own
// earlier results and the lower triangle, so rows need no barrier between
// them.  k runs in the same order on every lane: L[r][k] is one broadcast
// read and X[k][c] a coalesced one.  Four partial sums per column.  Kept out
// of line: with this loop inlined into the window-end path, the 256-thread
// state machine's single-transition replays (which never reach that path) no
// longer matched the oracle, while this out-of-line form matches; the cause
// was not found.
template <int TPC>
__device__ __noinline__ void dense_inverse(double* L, int D, int tid) {
  for (int r = 1; r < D; ++r) {
    const double lrr = L[(int64_t)r * D + r];
    for (int c0 = 0; c0 < r; c0 += TPC) {
      const int c = c0 + tid;
      const bool own = c < r;
      const int cc = own ? c : 0;
      // k = c term with X[c][c] = 1 / L[c][c]
      double a0 = own ? L[(int64_t)cc * D + r] / L[(int64_t)cc * D + cc] : 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
      int k = c0 + 1;
      for (; k + 4 <= r; k += 4) {
        const double l0 = L[(int64_t)k * D + r], l1 = L[(int64_t)(k + 1) * D + r];
        const double l2 = L[(int64_t)(k + 2) * D + r], l3 = L[(int64_t)(k + 3) * D + r];
        const double x0 = L[(int64_t)k * D + cc], x1 = L[(int64_t)(k + 1) * D + cc];
        const double x2 = L[(int64_t)(k + 2) * D + cc], x3 = L[(int64_t)(k + 3) * D + cc];
        a0 += k > cc ? l0 * x0 : 0.0;
        a1 += k + 1 > cc ? l1 * x1 : 0.0;
        a2 += k + 2 > cc ? l2 * x2 : 0.0;
        a3 += k + 3 > cc ? l3 * x3 : 0.0;
      }
      for (; k < r; ++k) a0 += k > cc ? L[(int64_t)k * D + r] * L[(int64_t)k * D + cc] : 0.0;
      if (own) L[(int64_t)r * D + c] = -((a0 + a1) + (a2 + a3)) / lrr;
    }
  }
}

template <int METRIC, class GradFn, int TPC = NUTS_WAVE, class VT = double>
struct NutsWave {
  static_assert(TPC % NUTS_WAVE == 0 && TPC <= 256, "threads per chain");
  static constexpr bool DENSE = METRIC == METRIC_DENSE;
  static_assert(!DENSE || sizeof(VT) == sizeof(double), "dense_e keeps FP64 vectors");
  const NutsParams& P;
  const NutsBuffers& B;
  ChainState& s;
  const GradFn& grad;
  int c, lane, tid, D, Dp;  // lane in the wave, tid in the chain's thread group
  uint32_t uid;
  // vectors are stored as VT (double, or float for the FP32 leapfrog) and
  // combined in FP64 arithmetic
  VT* q;     // Z end position (dense Q row)
  VT* g;     // Z end potential gradient (dense G row)
  VT* blk;   // pool block of this chain
  VT* minv;  // diagonal inverse metric (unit/diag)
  double* dM;    // dense: inverse metric, column-major D×D
  double* dL;    // dense: lower Cholesky factor, column-major
  double* dW;    // dense: Welford m2, column-major
  // V at Q of the pending gradient: computed in this launch (fused: kept in
  // registers, identical on every lane) or by the gradient kernel of the
  // previous launch (read from memory at the kernel boundary)
  double v_new = 0;
  bool have_v = false;

  __device__ NutsWave(const NutsParams& P_, const NutsBuffers& B_, ChainState& s_, const GradFn& g_, int c_, int lane_,
                      int tid_)
      : P(P_), B(B_), s(s_), grad(g_), c(c_), lane(lane_), tid(tid_) {
    D = B.D[c];
    Dp = B.Dp[c];
    uid = B.uid[c];
    // FP32 vectors use the first half of each buffer at the same element offsets
    q = reinterpret_cast<VT*>(B.Q) + B.off[c];
    g = reinterpret_cast<VT*>(B.G) + B.off[c];
    blk = reinterpret_cast<VT*>(B.pool) + B.off[c] * (int64_t)P.NV;
    minv = reinterpret_cast<VT*>(B.minv) + B.off[c];
    if (DENSE) {
      dM = B.dM + B.moff[c];
      dL = B.dL + B.moff[c];
      dW = B.dW + B.moff[c];
    } else {
      dM = dL = dW = nullptr;
    }
  }

  // group reduction: every thread gets the same sum (fixed order)
  static __device__ double gsum(double v) {
    v = wave_sum(v);
    if constexpr (TPC == NUTS_WAVE) {
      return v;
    } else {
      __shared__ double red[TPC / NUTS_WAVE];
      __syncthreads();
      if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
      __syncthreads();
      double r = 0.0;
#pragma unroll
      for (int w = 0; w < TPC / NUTS_WAVE; ++w) r += red[w];
      return r;
    }
  }
  // global vector writes of the group visible to its other threads
  static __device__ void gsync() {
    if constexpr (TPC == NUTS_WAVE) {
      wave_sync();
    } else {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __syncthreads();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
  }
  static __device__ int gand(int v) {
    v = wave_and(v);
    if constexpr (TPC == NUTS_WAVE) {
      return v;
    } else {
      __shared__ int redi[TPC / NUTS_WAVE];
      __syncthreads();
      if ((threadIdx.x & 63) == 0) redi[threadIdx.x >> 6] = v;
      __syncthreads();
      int r = 1;
#pragma unroll
      for (int w = 0; w < TPC / NUTS_WAVE; ++w) r &= redi[w];
      return r;
    }
  }
  // Reconverges the wave after a lane-strided loop: the scalar state updates
  // that follow run with the full wave, never duplicated into the divergent
  // exits of the loop (where a read-modify-write of the shared state would
  // be applied once per lane group).
  static __device__ void converge() { __builtin_amdgcn_wave_barrier(); }
  static __device__ void lsync() {
    if constexpr (TPC == NUTS_WAVE) lds_sync();
    else __syncthreads();
  }

  __device__ VT* vecp(int id) const { return blk + (int64_t)id * Dp; }
  __device__ VT* Lbeg(int k) const { return vecp(PV_FIXED + s.L_beg_vec[k]); }
  __device__ VT* Lend(int k) const { return vecp(PV_FIXED + P.maxd + k); }
  __device__ VT* Lrho(int k) const { return vecp(PV_FIXED + 2 * P.maxd + k); }
  __device__ VT* Sq(int k) const { return vecp(PV_FIXED + 3 * P.maxd + k); }
  __device__ VT* Sg(int k) const { return vecp(PV_FIXED + 3 * P.maxd + P.nslots + k); }
  __device__ VT* LSbeg(int k) const { return vecp(PV_FIXED + 3 * P.maxd + 2 * P.nslots + k); }
  __device__ VT* LSend(int k) const { return vecp(PV_FIXED + 4 * P.maxd + 2 * P.nslots + k); }
  // momentum of a proposal slot (kept only for the diagnostic rows)
  __device__ VT* Sp(int k) const { return vecp(PV_FIXED + (DENSE ? 5 : 3) * P.maxd + 2 * P.nslots + k); }

  // ------------------------------------------------------------- metric
  __device__ double mi(int i) const { return METRIC == METRIC_UNIT ? 1.0 : (double)minv[i]; }
  // sharp momentum element: x = momentum, xs = its stored companion (dense)
  __device__ double sh(const VT* x, const VT* xs, int i) const {
    if (DENSE) return xs[i];
    return METRIC == METRIC_UNIT ? (double)x[i] : (double)minv[i] * (double)x[i];
  }
  // out = Minv x (dense)
  __device__ void matvec(const VT* x, VT* out) const {
    gsync();
    dense_matvec<TPC, VT>(dM, x, out, D, tid);
    gsync();
  }
  // T = 1/2 p·Minv p; for dense ps must hold Minv p
  __device__ double kinetic(const VT* p, const VT* ps) const {
    const VT* __restrict__ pr = p;
    const VT* __restrict__ sr = DENSE ? ps : (METRIC == METRIC_UNIT ? p : minv);
    double a = 0;
#pragma unroll 4
    for (int i = tid; i < D; i += TPC) {
      const double pv = pr[i];
      a += pv * (DENSE ? (double)sr[i] : (METRIC == METRIC_UNIT ? pv : (double)sr[i] * pv));
    }
    converge();
    return 0.5 * gsum(a);
  }

  // ------------------------------------------------------------- draws
  __device__ double uniform() {
    const uint64_t k = s.uctr++;
    if (B.inj_u) return k < (uint64_t)B.n_inj_u ? B.inj_u[(int64_t)c * B.n_inj_u + k] : 0.5;
    return stream_uniform(P.seed, uid, STREAM_UNIFORM, k);
  }
  __device__ double normal_draw(int i) const {
    if (B.inj_z) {
      const uint64_t k = s.nctr + (uint64_t)i;
      return k < (uint64_t)B.n_inj_z ? B.inj_z[(int64_t)c * B.n_inj_z + k] : 0.0;
    }
    return stream_normal(P.seed, uid, s.nctr + (uint64_t)(i >> 1), i & 1);
  }
  __device__ void advance_normals() { s.nctr += B.inj_z ? (uint64_t)D : (uint64_t)((D + 1) / 2); }

  // {unit,diag,dense}_e_metric::sample_p; for dense also sets ps = Minv p
  __device__ void sample_p(VT* p, VT* ps) {
    if (!DENSE) {
      for (int i = tid; i < D; i += TPC) {
        const double z = normal_draw(i);
        p[i] = METRIC == METRIC_UNIT ? z : z / sqrt((double)minv[i]);
      } converge();
      advance_normals();
      return;
    }
    // the reference solves Lᵀ p = z; p = L⁻ᵀ z with the stored inverse is the
    // same vector without D dependent steps, and Minv p = L z comes with it
    __shared__ double sp[4][NUTS_WAVE * DENSE_RMAX];
    double* zw = sp[TPC == NUTS_WAVE ? (threadIdx.x >> 6) : 0];
    for (int i = tid; i < D; i += TPC) zw[i] = normal_draw(i); converge();
    advance_normals();
    lsync();
    dense_tri_sample<TPC, VT>(dL, zw, p, ps, D, tid);
    converge();
    lsync();  // the strip is free for the next draw
    gsync();
  }

  // ------------------------------------------------------------- vectors
  // the pool vectors never alias each other: restrict lets the unrolled loops
  // issue several independent loads before their stores
  __device__ void copy(VT* __restrict__ dst, const VT* __restrict__ src) const {
#pragma unroll 4
    for (int i = tid; i < D; i += TPC) dst[i] = src[i];
    converge();
  }
  __device__ void swap(VT* a, VT* b) const {
    for (int i = tid; i < D; i += TPC) { const VT t = a[i]; a[i] = b[i]; b[i] = t; } converge();
  }
  // expl_leapfrog: begin_update_p(eps/2), update_q(eps)
  __device__ void leap_begin(double e) const {
    VT* p = vecp(PV_P);
    const double h = 0.5 * e;
    if (DENSE) {
      VT* t = vecp(PV_TMP);
      for (int i = tid; i < D; i += TPC) p[i] = p[i] - h * g[i]; converge();
      matvec(p, t);
      for (int i = tid; i < D; i += TPC) q[i] = q[i] + e * t[i]; converge();
      return;
    }
    VT* __restrict__ pr = p;
    VT* __restrict__ qr = q;
    const VT* __restrict__ gr = g;
#pragma unroll 4
    for (int i = tid; i < D; i += TPC) {
      const VT pi = (double)pr[i] - h * (double)gr[i];
      pr[i] = pi;
      qr[i] = (double)qr[i] + e * (mi(i) * (double)pi);
    }
    converge();
  }
  // end_update_p(eps/2); dense: also the leaf's sharp momentum
  __device__ void leap_end(double e) const {
    VT* __restrict__ p = vecp(PV_P);
    const VT* __restrict__ gr = g;
    const double h = 0.5 * e;
#pragma unroll 4
    for (int i = tid; i < D; i += TPC) p[i] = (double)p[i] - h * (double)gr[i];
    converge();
    if (DENSE) matvec(p, vecp(PV_PS));
  }

  // dense_e leaf: end_update_p(eps/2), then the sharp momentum M p and M g in
  // one pass over the metric; M g stays in PV_TMP for leap_begin_mg
  __device__ void leap_end_mg(double e) const {
    VT* __restrict__ p = vecp(PV_P);
    const VT* __restrict__ gr = g;
    const double h = 0.5 * e;
    for (int i = tid; i < D; i += TPC) p[i] = (double)p[i] - h * (double)gr[i];
    converge();
    gsync();
    dense_matvec2<TPC, VT>(dM, p, g, vecp(PV_PS), vecp(PV_TMP), D, tid);
    gsync();
  }
  // dense_e: begin_update_p(eps/2) and update_q(eps) of the next leaf, with
  // M (p - h g) = M p - h M g from the products leap_end_mg left (g unchanged
  // in between): no second pass over the metric
  __device__ void leap_begin_mg(double e) const {
    VT* __restrict__ p = vecp(PV_P);
    const VT* __restrict__ ps = vecp(PV_PS);
    VT* __restrict__ t = vecp(PV_TMP);
    VT* __restrict__ qr = q;
    const VT* __restrict__ gr = g;
    const double h = 0.5 * e;
    for (int i = tid; i < D; i += TPC) {
      p[i] = (double)p[i] - h * (double)gr[i];
      const double ti = (double)ps[i] - h * (double)t[i];
      qr[i] = (double)qr[i] + e * ti;
    }
    converge();
  }

  // ------------------------------------------------------------- slots
  __device__ int alloc_slot() {
    const uint32_t m = s.free_mask;
    const int k = __builtin_ctz(m);  // a free slot always exists (<= maxd+1 live)
    s.free_mask = m & ~(1u << k);
    return k;
  }
  __device__ void free_slot(int k) { if (k >= 0) s.free_mask |= (1u << k); }
  __device__ int store_leaf_slot() {
    const int k = alloc_slot();
    copy(Sq(k), q);
    copy(Sg(k), g);
    if (P.diag) copy(Sp(k), vecp(PV_P));
    s.slot_V[k] = s.Vz;
    s.slot_H[k] = s.leaf_h;
    return k;
  }

  __device__ double pending_V() const { return have_v ? v_new : B.V[c]; }

  // -------------------------------------------------------- control flow
  __device__ bool request(int next) {
    s.phase = next;
    s.n_grad++;
    return true;
  }
  __device__ void stop(int phase) {
    s.phase = phase;
    if (tid == 0) atomicAdd(B.n_stopped, 1); converge();
  }
  __device__ void fail(int code) {
    s.failure = code;
    stop(PH_FAILED);
  }

  // services::util::initialize
  __device__ bool init_try() {
    const int t = s.init_tries;
    for (int i = tid; i < D; i += TPC) {
      double v;
      // user values where given; NaN (a parameter the init file leaves out)
      // and no user init at all draw uniform(-R, R) on the unconstrained scale
      v = B.user_init ? B.user_init[B.off[c] + i] : __builtin_nan("");
      if (isnan(v)) {
        if (P.init_radius == 0) v = 0.0;
        else v = P.init_radius * (2.0 * stream_uniform(P.seed, uid, STREAM_INIT, (uint64_t)t * (uint64_t)D + i) - 1.0);
      }
      q[i] = v;
    } converge();
    return request(PH_INIT_CHECK);
  }
  __device__ bool init_check() {
    const double V = pending_V();
    int ok = 1;
    for (int i = tid; i < D; i += TPC) ok &= isfinite(g[i]) ? 1 : 0; converge();
    ok = gand(ok) && isfinite(V);
    if (ok) {
      const int k = s.sample_slot;
      copy(Sq(k), q);
      copy(Sg(k), g);
      s.slot_V[k] = V;
      s.Vz = V;
      if (P.adapt) { s.ss_return = SS_RET_TRANS; s.phase = PH_SS_START; }
      else s.phase = PH_TRANS_BEGIN;
    } else {
      if (!isfinite(V)) s.init_rej_lp++;
      else s.init_rej_grad++;
      s.init_tries++;
      if (s.init_tries >= P.max_init_tries) fail(FAIL_INIT);
      else s.phase = PH_INIT_TRY;
    }
    return false;
  }

  // base_hmc::init_stepsize
  __device__ bool ss_probe_begin() {
    const int k = s.sample_slot;
    copy(q, Sq(k));
    copy(g, Sg(k));
    s.Vz = s.slot_V[k];
    VT* p = vecp(PV_P);
    VT* ps = vecp(PV_PS);
    sample_p(p, ps);
    s.ss_H0 = kinetic(p, ps) + s.Vz;
    leap_begin(s.nom_eps);
    return request(PH_SS_PROBE);
  }
  __device__ bool ss_start() {
    const double e = s.nom_eps;
    if (e == 0 || e > 1e7 || isnan(e)) { ss_finish(); return false; }
    s.ss_first = 1;
    return ss_probe_begin();
  }
  __device__ bool ss_probe() {
    s.Vz = pending_V();
    leap_end(s.nom_eps);
    double h = kinetic(vecp(PV_P), vecp(PV_PS)) + s.Vz;
    if (isnan(h)) h = __builtin_inf();
    const double dH = s.ss_H0 - h;
    const double l08 = log(0.8);
    if (s.ss_first) {
      s.ss_dir = dH > l08 ? 1 : -1;
      s.ss_first = 0;
      return ss_probe_begin();
    }
    if ((s.ss_dir == 1) && !(dH > l08)) { ss_finish(); return false; }
    if ((s.ss_dir == -1) && !(dH < l08)) { ss_finish(); return false; }
    s.nom_eps = s.ss_dir == 1 ? s.nom_eps * 2 : s.nom_eps * 0.5;
    if (s.nom_eps > 1e7) { fail(FAIL_STEPSIZE_IMPROPER); return false; }
    if (s.nom_eps == 0) { fail(FAIL_STEPSIZE_ZERO); return false; }
    return ss_probe_begin();
  }
  __device__ void ss_finish() {
    if (s.ss_return == SS_RET_WINDOW) {
      s.sa_mu = log(10 * s.nom_eps);
      s.sa_counter = 0; s.sa_sbar = 0; s.sa_xbar = 0;
      iteration_next();
    } else {
      s.phase = PH_TRANS_BEGIN;
    }
  }

  // base_nuts::transition, set-up part
  __device__ void trans_begin() {
    if (s.iter >= P.iter_target) { stop(PH_PAUSED); return; }
    s.eps = s.nom_eps;
    if (P.jitter != 0) s.eps *= 1.0 + P.jitter * (2.0 * uniform() - 1.0);
    const int k = s.sample_slot;
    // every slot but the current draw's is free at the start of a transition
    // (a transition that ended inside a subtree leaves its checkpoints behind)
    const uint32_t all = P.nslots >= 32 ? 0xFFFFFFFFu : ((1u << P.nslots) - 1u);
    s.free_mask = all & ~(1u << k);
    for (int l = 0; l < P.maxd; ++l) s.L_beg_vec[l] = (uint8_t)l;
    VT* p = vecp(PV_P);
    VT* ps = vecp(PV_PS);
    copy(q, Sq(k));
    copy(g, Sg(k));
    s.Vz = s.slot_V[k];
    sample_p(p, ps);
    // the draw stays at z0 (with the new momentum) unless a leaf replaces it
    if (P.diag) copy(Sp(k), p);
    VT* oq = vecp(PV_OQ); VT* op = vecp(PV_OP); VT* og = vecp(PV_OG);
    VT* pfb = vecp(PV_PFB); VT* pbf = vecp(PV_PBF); VT* rho = vecp(PV_RHO);
    for (int i = tid; i < D; i += TPC) {
      const VT pi = p[i];
      oq[i] = q[i]; op[i] = pi; og[i] = g[i];
      pfb[i] = pi; pbf[i] = pi; rho[i] = pi;
    } converge();
    if (DENSE) {
      VT* ops = vecp(PV_OPS); VT* psfb = vecp(PV_PSFB); VT* psbf = vecp(PV_PSBF);
      for (int i = tid; i < D; i += TPC) { const VT v = ps[i]; ops[i] = v; psfb[i] = v; psbf[i] = v; } converge();
    }
    s.Vo = s.Vz;
    s.z_is_fwd = 1;
    s.log_w = 0;
    s.H0 = kinetic(p, ps) + s.Vz;
    s.slot_H[k] = s.H0;
    s.n_leapfrog = 0;
    s.sum_metro = 0;
    s.depth = 0;
    s.divergent = 0;
    s.max_abs_dH = 0;
    s.phase = PH_DOUBLING_BEGIN;
  }

  __device__ bool doubling_begin() {
    const int fwd = uniform() > 0.5 ? 1 : 0;
    if (fwd != s.z_is_fwd) {
      swap(q, vecp(PV_OQ));
      swap(vecp(PV_P), vecp(PV_OP));
      swap(g, vecp(PV_OG));
      if (DENSE) swap(vecp(PV_PS), vecp(PV_OPS));
      const double t = s.Vz; s.Vz = s.Vo; s.Vo = t;
      s.z_is_fwd = fwd;
    }
    // p_bck_fwd = p_fwd_bck (forward) / p_fwd_bck = p_bck_fwd (backward)
    if (fwd) {
      copy(vecp(PV_PBF), vecp(PV_PFB));
      if (DENSE) copy(vecp(PV_PSBF), vecp(PV_PSFB));
    } else {
      copy(vecp(PV_PFB), vecp(PV_PBF));
      if (DENSE) copy(vecp(PV_PSFB), vecp(PV_PSBF));
    }
    s.dir = fwd ? 1 : -1;
    s.leaf_idx = 0;
    leap_begin(s.dir * s.eps);
    return request(PH_LEAF);
  }

  // build_tree base case plus the merges that close at this leaf (unit/diag).
  // One pass completes the leapfrog, the kinetic energy and the U-turn sums of
  // up to three of the merges that close here, and stores the leaf's momentum
  // and merged rho where they are read next: in the open level K they close
  // into (K = the trailing one bits of the leaf index), or in p and srho at
  // the doubling's last leaf.  When a check fails, the transition ends and
  // those stores are never read.
  __device__ bool leaf() {
    if constexpr (DENSE) {
      return leaf_dense();
    } else {
      constexpr bool UNIT = METRIC == METRIC_UNIT;
      constexpr int FMAX = 3;
      const double e = s.dir * s.eps;
      s.Vz = pending_V();
      s.n_leapfrog++;
      const int K = __builtin_ctz(~(unsigned)s.leaf_idx);
      const bool last = s.leaf_idx == (1 << s.depth) - 1;
      VT* p = vecp(PV_P);
      VT* dp = last ? p : Lend(K);
      VT* dr = last ? vecp(PV_SRHO) : Lrho(K);
      VT* db = (!last && K == 0) ? Lbeg(0) : nullptr;
      VT* edge = s.leaf_idx == 0 ? (s.z_is_fwd ? vecp(PV_PFB) : vecp(PV_PBF)) : nullptr;
      const double hh = 0.5 * e;
      const int F = K < FMAX ? K : FMAX;
      const VT* lb[FMAX];
      const VT* le[FMAX];
      const VT* lr[FMAX];
#pragma unroll
      for (int k = 0; k < FMAX; ++k) {
        const int kk = k < F ? k : 0;
        lb[k] = Lbeg(kk);
        le[k] = Lend(kk);
        lr[k] = Lrho(kk);
      }
      double acc[1 + 6 * FMAX];
      switch (F) {
        case 0: pass_leaf_merge<TPC, UNIT, 0>(p, g, minv, lb, le, lr, dp, dr, db, edge, D, tid, hh, acc); break;
        case 1: pass_leaf_merge<TPC, UNIT, 1>(p, g, minv, lb, le, lr, dp, dr, db, edge, D, tid, hh, acc); break;
        case 2: pass_leaf_merge<TPC, UNIT, 2>(p, g, minv, lb, le, lr, dp, dr, db, edge, D, tid, hh, acc); break;
        default: pass_leaf_merge<TPC, UNIT, 3>(p, g, minv, lb, le, lr, dp, dr, db, edge, D, tid, hh, acc); break;
      }
      converge();
      double h = 0.5 * gsum(acc[0]) + s.Vz;
      if (isnan(h)) h = __builtin_inf();
      if ((h - s.H0) > P.max_deltaH) s.divergent = 1;
      s.max_abs_dH = fmax(s.max_abs_dH, fabs(h - s.H0));
      const double lw = s.H0 - h;
      if (s.H0 - h > 0) s.sum_metro += 1;
      else s.sum_metro += exp(s.H0 - h);
      s.leaf_h = h;
      if (s.divergent) { trans_end(); return false; }
      s.S_logw = lw;
      s.s_prop = PROP_LEAF;
      int k = 0;
      while ((s.leaf_idx >> k) & 1) {
        const double lsw_sub = log_sum_exp2(s.L_logw[k], s.S_logw);
        bool take;
        if (s.S_logw > lsw_sub) take = true;
        else take = uniform() < exp(s.S_logw - lsw_sub);
        if (take) free_slot(s.L_prop[k]);
        else { free_slot(s.s_prop); s.s_prop = s.L_prop[k]; }
        double a1, b1, a2, b2, a3, b3;
        if (k < F) {
          const double* ak = acc + 1 + 6 * k;
          a1 = gsum(ak[0]); b1 = gsum(ak[1]); a2 = gsum(ak[2]);
          b2 = gsum(ak[3]); a3 = gsum(ak[4]); b3 = gsum(ak[5]);
        } else {
          // deeper merges: the separate pass over the stored leaf momentum and
          // the rho merged so far (k >= F >= 1, so the begin is level k-1's)
          double am[6];
          pass_merge<TPC, METRIC, VT>(Lbeg(k), Lend(k), Lrho(k), Lbeg(k - 1), dp, dr, minv, nullptr, nullptr,
                                      nullptr, nullptr, D, tid, am);
          converge();
          a1 = gsum(am[0]); b1 = gsum(am[1]); a2 = gsum(am[2]);
          b2 = gsum(am[3]); a3 = gsum(am[4]); b3 = gsum(am[5]);
        }
        const bool ok = (b1 > 0 && a1 > 0) && (b2 > 0 && a2 > 0) && (b3 > 0 && a3 > 0);
        s.S_logw = lsw_sub;
        if (!ok) { free_slot(s.s_prop); trans_end(); return false; }
        ++k;
      }
      if (last) { doubling_end(); return false; }
      // the closed subtree is the open left sibling at level K: its end and rho
      // are stored there; its begin is level K-1's begin vector (the oldest
      // subtree merged), taken over by a swap, or the leaf itself (K = 0)
      if (K > 0) {
        const uint8_t t = s.L_beg_vec[K];
        s.L_beg_vec[K] = s.L_beg_vec[K - 1];
        s.L_beg_vec[K - 1] = t;
      }
      const int slot = s.s_prop == PROP_LEAF ? alloc_slot() : -1;
      VT* sq = slot >= 0 ? Sq(slot) : nullptr;
      VT* sg = slot >= 0 ? Sg(slot) : nullptr;
      VT* spp = slot >= 0 && P.diag ? Sp(slot) : nullptr;
      if (sq) pass_leaf_next<TPC, UNIT, true>(dp, p, q, g, minv, sq, sg, spp, D, tid, hh, e);
      else pass_leaf_next<TPC, UNIT, false>(dp, p, q, g, minv, sq, sg, spp, D, tid, hh, e);
      converge();
      if (slot >= 0) {
        s.slot_V[slot] = s.Vz;
        s.slot_H[slot] = s.leaf_h;
      }
      s.L_logw[K] = s.S_logw;
      s.L_prop[K] = slot >= 0 ? slot : s.s_prop;
      s.leaf_idx++;
      return request(PH_LEAF);
    }
  }

  // build_tree base case and merges for dense_e (the sharp momenta are kept
  // beside the momenta in every checkpoint)
  __device__ bool leaf_dense() {
    const double e = s.dir * s.eps;
    s.Vz = pending_V();
    VT* p = vecp(PV_P);
    VT* ps = vecp(PV_PS);
    VT* srho = vecp(PV_SRHO);
    s.n_leapfrog++;
    double h;
    if (DENSE) {
      leap_end_mg(e);
      h = kinetic(p, ps) + s.Vz;
    } else {
      // leap_end, kinetic, the edge copy and the srho copy in one pass (srho
      // is written before the divergence check: a diverging transition ends
      // here and never reads it)
      constexpr bool UNIT = METRIC == METRIC_UNIT;
      VT* edge = s.z_is_fwd ? vecp(PV_PFB) : vecp(PV_PBF);
      const double hh = 0.5 * e;
      const double a = s.leaf_idx == 0 ? pass_leaf_head<TPC, UNIT, true>(p, g, minv, srho, edge, D, tid, hh)
                                       : pass_leaf_head<TPC, UNIT, false>(p, g, minv, srho, edge, D, tid, hh);
      converge();
      h = 0.5 * gsum(a) + s.Vz;
    }
    if (isnan(h)) h = __builtin_inf();
    if ((h - s.H0) > P.max_deltaH) s.divergent = 1;
    s.max_abs_dH = fmax(s.max_abs_dH, fabs(h - s.H0));
    const double lw = s.H0 - h;
    if (s.H0 - h > 0) s.sum_metro += 1;
    else s.sum_metro += exp(s.H0 - h);
    s.leaf_h = h;
    if (DENSE && s.leaf_idx == 0) {
      copy(s.z_is_fwd ? vecp(PV_PFB) : vecp(PV_PBF), p);
      copy(s.z_is_fwd ? vecp(PV_PSFB) : vecp(PV_PSBF), ps);
    }
    if (s.divergent) { trans_end(); return false; }

    if (DENSE) copy(srho, p);
    s.S_logw = lw;
    s.s_prop = PROP_LEAF;
    int k = 0;
    while ((s.leaf_idx >> k) & 1) {
      const double lsw_sub = log_sum_exp2(s.L_logw[k], s.S_logw);
      bool take;
      if (s.S_logw > lsw_sub) take = true;
      else take = uniform() < exp(s.S_logw - lsw_sub);
      if (take) free_slot(s.L_prop[k]);
      else { free_slot(s.s_prop); s.s_prop = s.L_prop[k]; }
      const VT* lb = Lbeg(k);
      const VT* le = Lend(k);
      const VT* lr = Lrho(k);
      const VT* sb = k == 0 ? p : Lbeg(k - 1);
      const VT* lbs = DENSE ? LSbeg(k) : nullptr;
      const VT* les = DENSE ? LSend(k) : nullptr;
      const VT* sbs = DENSE ? (k == 0 ? ps : LSbeg(k - 1)) : nullptr;
      double acc[6];
      pass_merge<TPC, METRIC>(lb, le, lr, sb, p, srho, minv, lbs, les, sbs, DENSE ? ps : nullptr, D, tid, acc);
      converge();
      const double a1 = gsum(acc[0]), b1 = gsum(acc[1]), a2 = gsum(acc[2]);
      const double b2 = gsum(acc[3]), a3 = gsum(acc[4]), b3 = gsum(acc[5]);
      const bool ok = (b1 > 0 && a1 > 0) && (b2 > 0 && a2 > 0) && (b3 > 0 && a3 > 0);
      s.S_logw = lsw_sub;
      if (!ok) { free_slot(s.s_prop); trans_end(); return false; }
      ++k;
    }
    if (s.leaf_idx == (1 << s.depth) - 1) { doubling_end(); return false; }
    // the closed subtree becomes the open left sibling at level k
    if (DENSE) {
      copy(Lbeg(k), k == 0 ? p : Lbeg(k - 1));
      copy(Lend(k), p);
      copy(Lrho(k), srho);
      copy(LSbeg(k), k == 0 ? ps : LSbeg(k - 1));
      copy(LSend(k), ps);
      s.L_logw[k] = s.S_logw;
      s.L_prop[k] = s.s_prop == PROP_LEAF ? store_leaf_slot() : s.s_prop;
      s.leaf_idx++;
      leap_begin_mg(e);
      return request(PH_LEAF);
    }
    // unit/diag: the checkpoint copies, the proposal slot and the next leaf's
    // leap_begin in one pass
    constexpr bool UNIT = METRIC == METRIC_UNIT;
    const int slot = s.s_prop == PROP_LEAF ? alloc_slot() : -1;
    VT* lb = Lbeg(k);
    VT* le = Lend(k);
    VT* lr = Lrho(k);
    const VT* beg = k == 0 ? nullptr : Lbeg(k - 1);
    VT* sq = slot >= 0 ? Sq(slot) : nullptr;
    VT* sg = slot >= 0 ? Sg(slot) : nullptr;
    VT* spp = slot >= 0 && P.diag ? Sp(slot) : nullptr;
    const double hh = 0.5 * e;
    if (beg) {
      if (sq) pass_leaf_tail<TPC, UNIT, true, true>(p, q, g, minv, srho, beg, lb, le, lr, sq, sg, spp, D, tid, hh, e);
      else pass_leaf_tail<TPC, UNIT, true, false>(p, q, g, minv, srho, beg, lb, le, lr, sq, sg, spp, D, tid, hh, e);
    } else {
      if (sq) pass_leaf_tail<TPC, UNIT, false, true>(p, q, g, minv, srho, beg, lb, le, lr, sq, sg, spp, D, tid, hh, e);
      else pass_leaf_tail<TPC, UNIT, false, false>(p, q, g, minv, srho, beg, lb, le, lr, sq, sg, spp, D, tid, hh, e);
    }
    converge();
    if (slot >= 0) {
      s.slot_V[slot] = s.Vz;
      s.slot_H[slot] = s.leaf_h;
    }
    s.L_logw[k] = s.S_logw;
    s.L_prop[k] = slot >= 0 ? slot : s.s_prop;
    s.leaf_idx++;
    return request(PH_LEAF);
  }

  // top-level merge of a complete, valid subtree (biased progressive sampling)
  __device__ void doubling_end() {
    s.depth++;
    bool take;
    if (s.S_logw > s.log_w) take = true;
    else take = uniform() < exp(s.S_logw - s.log_w);
    if (take) {
      const int ns = s.s_prop == PROP_LEAF ? store_leaf_slot() : s.s_prop;
      free_slot(s.sample_slot);
      s.sample_slot = ns;
    } else {
      free_slot(s.s_prop);
    }
    s.log_w = log_sum_exp2(s.log_w, s.S_logw);
    const int fwd = s.z_is_fwd;
    const VT* p = vecp(PV_P);
    const VT* op = vecp(PV_OP);
    const VT* pfb = vecp(PV_PFB);
    const VT* pbf = vecp(PV_PBF);
    const VT* srho = vecp(PV_SRHO);
    VT* rho = vecp(PV_RHO);
    const VT* ps = DENSE ? vecp(PV_PS) : nullptr;
    const VT* ops = DENSE ? vecp(PV_OPS) : nullptr;
    const VT* psfb = DENSE ? vecp(PV_PSFB) : nullptr;
    const VT* psbf = DENSE ? vecp(PV_PSBF) : nullptr;
    const VT* bb = fwd ? op : p;   // p_bck_bck
    const VT* ff = fwd ? p : op;   // p_fwd_fwd
    const VT* bbs = fwd ? ops : ps;
    const VT* ffs = fwd ? ps : ops;
    double acc[6];
    pass_doubling<TPC, METRIC>(fwd, rho, srho, bb, ff, pfb, pbf, minv, bbs, ffs, psfb, psbf, D, tid, acc);
    converge();
    const double a1 = gsum(acc[0]), b1 = gsum(acc[1]), a2 = gsum(acc[2]);
    const double b2 = gsum(acc[3]), a3 = gsum(acc[4]), b3 = gsum(acc[5]);
    const bool ok = (b1 > 0 && a1 > 0) && (b2 > 0 && a2 > 0) && (b3 > 0 && a3 > 0);
    if (!ok || s.depth >= P.max_depth) trans_end();
    else s.phase = PH_DOUBLING_BEGIN;
  }

  // window logic of stan::mcmc::windowed_adaptation (unsigned arithmetic)
  __device__ bool in_window() const {
    return (s.w_counter >= P.w_init_buffer) && (s.w_counter < P.w_num_warmup - P.w_term_buffer) &&
           (s.w_counter != P.w_num_warmup);
  }
  __device__ bool end_window() const { return (s.w_counter == s.w_next) && (s.w_counter != P.w_num_warmup); }
  __device__ void next_window() {
    if (s.w_next == P.w_num_warmup - P.w_term_buffer - 1) return;
    s.w_size *= 2;
    s.w_next = s.w_counter + s.w_size;
    if (s.w_next != P.w_num_warmup - P.w_term_buffer - 1) {
      const uint32_t boundary = s.w_next + 2 * s.w_size;
      if (boundary >= P.w_num_warmup - P.w_term_buffer) s.w_next = P.w_num_warmup - P.w_term_buffer - 1;
    }
  }
  // var_adaptation::learn_variance / covar_adaptation::learn_covariance;
  // returns true at a window end
  __device__ bool learn_metric(const VT* qs) {
    VT* wm = vecp(PV_WM);
    VT* wm2 = vecp(PV_WM2);
    if (in_window()) {
      s.welford_n++;
      const double n = (double)s.welford_n;
      if (!DENSE) {
        for (int i = tid; i < D; i += TPC) {
          const double delta = (double)qs[i] - (double)wm[i];
          const double m = wm[i] + delta / n;
          wm[i] = m;
          wm2[i] = wm2[i] + (qs[i] - m) * delta;
        } converge();
      } else {
        // welford_covar_estimator: delta = q - m; m += delta/n; m2 += (q - m) deltaᵀ
        VT* delta = vecp(PV_TMP);
        for (int i = tid; i < D; i += TPC) {
          const double dl = (double)qs[i] - (double)wm[i];
          delta[i] = dl;
          wm[i] = (double)wm[i] + dl / n;
          wm2[i] = (double)qs[i] - (double)wm[i];  // (q - m_new), kept in the diag accumulator slot
        } converge();
        gsync();
        for (int j = 0; j < D; ++j) {
          pass_axpy<TPC>(dW + (int64_t)j * D, wm2, (double)delta[j], D, tid);
          converge();
        }
      }
    }
    if (end_window()) {
      next_window();
      const double n = (double)s.welford_n;
      int ok = 1;
      if (!DENSE) {
        for (int i = tid; i < D; i += TPC) {
          const double var = s.welford_n > 1 ? (double)wm2[i] / (n - 1.0) : (double)minv[i];
          const double v = (n / (n + 5.0)) * var + 1e-3 * (5.0 / (n + 5.0));
          minv[i] = v;
          ok &= isfinite(v) ? 1 : 0;
          wm[i] = 0; wm2[i] = 0;
        } converge();
      } else {
        for (int j = 0; j < D; ++j) {
          double* Mj = dM + (int64_t)j * D;
          double* Wj = dW + (int64_t)j * D;
          for (int i = tid; i < D; i += TPC) {
            const double cov = s.welford_n > 1 ? Wj[i] / (n - 1.0) : Mj[i];
            const double v = (n / (n + 5.0)) * cov + (i == j ? 1e-3 * (5.0 / (n + 5.0)) : 0.0);
            Mj[i] = v;
            ok &= isfinite(v) ? 1 : 0;
            Wj[i] = 0;
          } converge();
        }
        for (int i = tid; i < D; i += TPC) { wm[i] = 0; wm2[i] = 0; } converge();
      }
      s.welford_n = 0;
      s.w_counter++;
      if (!gand(ok)) { fail(FAIL_METRIC_OVERFLOW); return false; }
      if (DENSE) {
        gsync();  // the factorisation reads elements other lanes just wrote
        if (!dense_factor(dM, dL, D, tid)) {
          fail(FAIL_METRIC_NOT_PD);
          return false;
        }
      }
      return true;
    }
    s.w_counter++;
    return false;
  }

  // end of base_nuts::transition + adapt_*_e_nuts::transition + writer
  __device__ void trans_end() {
    const double accept = s.sum_metro / (double)s.n_leapfrog;
    const int k = s.sample_slot;
    s.n_leapfrog_total += s.n_leapfrog;
    const bool warm = s.iter < P.num_warmup;
    const bool save = warm ? (P.save_warmup && (s.iter % P.thin) == 0) : (((s.iter - P.num_warmup) % P.thin) == 0);
    if (save) {
      const int nsave = P.save_params < 0 ? D : (P.save_params < D ? P.save_params : D);
      double* row = B.draws + B.draw_off[c] + (int64_t)s.draw_idx * (7 + nsave);
      if (tid == 0) {
        row[0] = -s.slot_V[k];
        row[1] = accept;
        row[2] = s.eps;
        row[3] = s.depth;
        row[4] = s.n_leapfrog;
        row[5] = s.divergent;
        row[6] = s.slot_H[k];
        if (B.eerr) B.eerr[(int64_t)c * P.out_rows + s.draw_idx] = s.max_abs_dH;
      } converge();
      const VT* sq = Sq(k);
      for (int i = tid; i < nsave; i += TPC) row[7 + i] = sq[i]; converge();
      if (P.diag) {
        // diagnostic_file row: sampler columns, q, p, g of the draw
        double* dr = B.diag + B.diag_off[c] + (int64_t)s.draw_idx * (7 + 3 * D);
        if (tid == 0) {
          for (int j = 0; j < 7; ++j) dr[j] = row[j];
        } converge();
        const VT* sp = Sp(k);
        const VT* sg = Sg(k);
        for (int i = tid; i < D; i += TPC) {
          dr[7 + i] = sq[i];
          dr[7 + D + i] = sp[i];
          dr[7 + 2 * D + i] = sg[i];
        } converge();
      }
      s.draw_idx++;
    }
    if (s.adapt_flag) {
      // stepsize_adaptation::learn_stepsize
      s.sa_counter += 1;
      const double a = accept > 1 ? 1 : accept;
      const double eta = 1.0 / (s.sa_counter + P.t0);
      s.sa_sbar = (1.0 - eta) * s.sa_sbar + eta * (P.delta - a);
      const double x = s.sa_mu - s.sa_sbar * sqrt(s.sa_counter) / P.gamma;
      const double x_eta = pow(s.sa_counter, -P.kappa);
      s.sa_xbar = (1.0 - x_eta) * s.sa_xbar + x_eta * x;
      s.nom_eps = exp(x);
      if (METRIC != METRIC_UNIT && P.metric_adapt) {
        const bool update = learn_metric(Sq(k));
        if (s.phase == PH_FAILED) return;
        if (update) { s.ss_return = SS_RET_WINDOW; s.phase = PH_SS_START; return; }
      }
    }
    iteration_next();
  }

  __device__ void iteration_next() {
    s.iter++;
    if (s.iter == P.num_warmup && s.adapt_flag) {
      s.adapt_flag = 0;
      s.nom_eps = exp(s.sa_xbar);  // stepsize_adaptation::complete_adaptation
    }
    if (s.iter >= P.num_warmup + P.num_samples) stop(PH_DONE);
    else s.phase = PH_TRANS_BEGIN;
  }

  // One step of the machine; true when V, G at Q are needed next.
  __device__ bool dispatch() {
    switch (s.phase) {
      case PH_INIT_TRY: return init_try();
      case PH_INIT_CHECK: return init_check();
      case PH_SS_START: return ss_start();
      case PH_SS_PROBE: return ss_probe();
      case PH_TRANS_BEGIN: trans_begin(); return false;
      case PH_DOUBLING_BEGIN: return doubling_begin();
      case PH_LEAF: return leaf();
      default: return false;
    }
  }

  __device__ void run() {
    int budget = GradFn::fused ? P.fused_budget : 1;
    for (;;) {
      if (s.phase >= PH_PAUSED) return;
      const bool req = dispatch();
      if (!req) continue;
      if constexpr (!GradFn::fused) {
        if (tid == 0) B.need[c] = 1; converge(); return;
      } else {
        gsync();  // the model reads q written lane-strided
        double V;
        if constexpr (TPC == NUTS_WAVE) V = grad(c, q, g, D, lane);
        else V = grad.template group<TPC>(c, q, g, tid, [](double x) { return gsum(x); });  // a block per chain
        v_new = V;
        have_v = true;
        if (tid == 0) B.V[c] = V; converge();  // for the next launch only
        gsync();  // models may write g elements that other lanes read
        if (--budget == 0) return;
      }
    }
  }

  // Lower Cholesky factor of the column-major SPD matrix A into L (column-major),
  // left-looking, one column per step with the wave splitting the rows.
  // Returns false if a pivot is not positive.
  static __device__ bool dense_factor(const double* A, double* L, int D, int tid) {
    for (int j = 0; j < D; ++j) {
      // pivot: A[j][j] - sum_k L[j][k]^2 (row j of L = element j of columns k < j)
      double sq = 0;
      for (int k = tid; k < j; k += TPC) { const double v = L[(int64_t)k * D + j]; sq += v * v; } converge();
      sq = gsum(sq);
      const double piv = A[(int64_t)j * D + j] - sq;
      if (!(piv > 0)) return false;
      const double d = sqrt(piv);
      double* Lj = L + (int64_t)j * D;
      for (int i = j + 1 + tid; i < D; i += TPC) {
        double v = A[(int64_t)j * D + i];  // A[i][j] (lower triangle)
        for (int k = 0; k < j; ++k) v -= L[(int64_t)k * D + i] * L[(int64_t)k * D + j];
        Lj[i] = v / d;
      } converge();
      if (tid == 0) Lj[j] = d; converge();
      gsync();
    }
    dense_inverse<TPC>(L, D, tid);
    gsync();
    return true;
  }
};

// Initial factorisation of every chain's dense inverse metric (one wave per chain).
__global__ void __launch_bounds__(256) dense_factor_kernel(NutsBuffers B, int* failed) {
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= B.C) return;
  const int D = B.D[c];
  const bool ok = NutsWave<METRIC_DENSE, ExternalGrad>::dense_factor(B.dM + B.moff[c], B.dL + B.moff[c], D, lane);
  if (!ok && lane == 0) atomicAdd(failed, 1);
}

// Stage ChainState in LDS, run, write back.  TPC == 64: one wave per chain, 4
// chains per block; TPC == 256: one chain per block, each wave with its own
// identical copy of the state, wave 0 writes it back.
template <int METRIC, class GradFn, int TPC = NUTS_WAVE, class VT = double>
__device__ __forceinline__ void nuts_advance_body(const NutsParams& P, const NutsBuffers& B, const GradFn& grad) {
  __shared__ ChainState sst[4];
  // the wave index is uniform over the wave: keep it (and the chain index and
  // every pointer derived from it) in scalar registers
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int c = TPC == NUTS_WAVE ? blockIdx.x * 4 + w : blockIdx.x;
  const int tid = TPC == NUTS_WAVE ? lane : (int)threadIdx.x;
  if (c >= B.C) return;
  ChainState& s = sst[w];
  {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(B.st + c);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&s);
    for (int i = lane; i < (int)(sizeof(ChainState) / 4); i += 64) dst[i] = src[i];
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  // every wave has its copy of the state before wave 0 may write the new one back
  if (TPC > NUTS_WAVE) __syncthreads();
  if (s.phase < PH_PAUSED) {
    NutsWave<METRIC, GradFn, TPC, VT> wv(P, B, s, grad, c, lane, tid);
    wv.run();
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  if (TPC == NUTS_WAVE || w == 0) {
    uint32_t* dst = reinterpret_cast<uint32_t*>(B.st + c);
    const uint32_t* src = reinterpret_cast<const uint32_t*>(&s);
    for (int i = lane; i < (int)(sizeof(ChainState) / 4); i += 64) dst[i] = src[i];
  }
}

}  // namespace nutsgpu
