// Types shared by the host driver and the NUTS kernels.
//
// Memory layout in HBM (C chains, chain c of dimension D_c, Dp_c = D_c rounded
// up to 16 doubles):
//   Q, G    [sum Dp]  dense position / potential-gradient rows (chain-major);
//                     the gradient kernels read Q and write G, V.  For a GLM
//                     batch with one D these are the X·B operand matrices.
//   V       [C]       potential -log p at Q
//   pool    [sum Dp * NV]  per-chain block of NV vectors (other trajectory end,
//                     U-turn momenta, rho, tree checkpoints, proposal slots,
//                     Welford accumulators), contiguous per chain.
//   minv    [sum Dp]  diagonal inverse metric (diag_e) or [sum D*D] (dense_e)
//   draws   [C][rows][7 + D]  sampler columns + unconstrained q
#pragma once
#include <stdint.h>

#define NUTS_MAX_DEPTH 20
#define NUTS_MAX_SLOTS (NUTS_MAX_DEPTH + 2)

namespace nutsgpu {

enum Phase : int {
  PH_INIT_TRY = 0,
  PH_INIT_CHECK,
  PH_SS_START,
  PH_SS_PROBE,
  PH_TRANS_BEGIN,
  PH_DOUBLING_BEGIN,
  PH_LEAF,
  PH_PAUSED,   // reached the host's iteration target (warmup end)
  PH_DONE,
  PH_FAILED,
};

enum Failure : int {
  FAIL_NONE = 0,
  FAIL_INIT = 1,             // "Initialization failed."
  FAIL_STEPSIZE_IMPROPER = 2,// "Posterior is improper. Please check your model."
  FAIL_STEPSIZE_ZERO = 3,    // "No acceptably small step size could be found..."
  FAIL_METRIC_OVERFLOW = 4,  // "Numerical overflow in metric adaptation."
  FAIL_METRIC_NOT_PD = 5,    // adapted dense metric has no Cholesky factor
};

enum SsReturn : int { SS_RET_TRANS = 0, SS_RET_WINDOW = 1 };

enum Metric : int { METRIC_UNIT = 0, METRIC_DIAG = 1, METRIC_DENSE = 2 };

#define PROP_LEAF (-1)

// Per-chain scalar state (one per chain in HBM; staged in LDS by the wave that
// owns the chain).
struct ChainState {
  int phase, failure, ss_return, iter;
  int draw_idx, depth, leaf_idx, n_leapfrog;
  int divergent, dir, z_is_fwd, init_tries;
  int ss_dir, ss_first, sample_slot, s_prop;
  uint32_t free_mask, w_counter, w_size, w_next;
  int adapt_flag, welford_n;
  int init_rej_lp, init_rej_grad;  // initial values rejected: non-finite log density / gradient
  int L_prop[NUTS_MAX_DEPTH];
  // vector holding the begin momentum of open level k (unit/diag: a subtree
  // that closes into level k takes over level k-1's begin vector by swapping
  // these entries instead of copying the vector)
  uint8_t L_beg_vec[NUTS_MAX_DEPTH];
  double eps, nom_eps, H0, log_w;
  double sum_metro, Vz, Vo, S_logw;
  double leaf_h, ss_H0, sa_counter, sa_sbar;
  double sa_xbar, sa_mu;
  double max_abs_dH, pad3;  // max |H - H0| over the leaves of the current transition
  double L_logw[NUTS_MAX_DEPTH];
  double slot_V[NUTS_MAX_SLOTS];
  double slot_H[NUTS_MAX_SLOTS];
  uint64_t uctr, nctr;
  int64_t n_grad, n_leapfrog_total;
};

// Run-wide parameters (kernel argument by value).
struct NutsParams {
  int num_warmup, num_samples, thin, save_warmup;
  int adapt, max_depth, max_init_tries, iter_target;
  double delta, gamma, kappa, t0;
  double jitter, stepsize, init_radius, max_deltaH;
  uint32_t w_num_warmup, w_init_buffer, w_term_buffer, w_base_window;
  uint32_t w_counter0, w_size0, w_next0, metric_adapt;
  uint64_t seed;
  int nslots, maxd, NV, fused_budget;
  int save_params;  // -1: all D coordinates in each draw row; else the first save_params
  int diag;         // 1: record the diagnostic rows (q, p, g of each saved draw)
  int out_rows, pad;  // saved rows per chain (row stride of eerr)
};

// Device pointers (kernel argument by value).
struct NutsBuffers {
  int C, pad;
  const int* D;             // [C]
  const int* Dp;            // [C]
  const int64_t* off;       // [C] offset into Q/G/minv (diag)
  const int64_t* moff;      // [C] offset into minv for dense (D*D blocks)
  const uint32_t* uid;      // [C] RNG chain number (id + k)
  double* Q;
  double* G;
  double* V;
  double* pool;             // [sum Dp * NV]
  ChainState* st;           // [C]
  double* minv;
  double* draws;            // [C] blocks of rows x (7 + D_c)
  const int64_t* draw_off;  // [C]
  const double* user_init;  // [sum Dp] or null
  int* n_stopped;           // chains in PAUSED/DONE/FAILED
  int* need;                // [C] 1 when the chain waits for V, G at Q
  const double* inj_u;      // injected uniforms [C][n_inj_u] (replay) or null
  const double* inj_z;      // injected normals [C][n_inj_z] or null
  int64_t n_inj_u, n_inj_z;
  // dense_e: per-chain D*D blocks at moff[c], all column-major (element (i, j) at [j*D + i])
  double* dM;               // inverse metric
  double* dL;               // its lower Cholesky factor
  double* dW;               // Welford covariance accumulator m2
  // diagnostic rows (P.diag): [C] blocks of rows x (7 + 3 D_c) at diag_off[c]:
  // the sampler columns, then q, p, g of the draw
  double* diag;
  const int64_t* diag_off;
  // per saved row, the transition's max |H - H0| over its leaves: [C][out_rows] or null
  double* eerr;
};

// Vector ids inside a chain's pool block.
enum PoolVec : int {
  PV_P = 0, PV_OQ, PV_OP, PV_OG, PV_PFB, PV_PBF, PV_RHO, PV_SRHO, PV_WM, PV_WM2,
  // dense_e companions: Minv times the momentum of the same role, and scratch
  PV_PS, PV_OPS, PV_PSFB, PV_PSBF, PV_TMP,
  PV_FIXED  // then L_beg[maxd], L_end[maxd], L_rho[maxd], slot_q[nslots], slot_g[nslots],
            // for dense_e LS_beg[maxd], LS_end[maxd], and with diagnostics slot_p[nslots]
};

}  // namespace nutsgpu
