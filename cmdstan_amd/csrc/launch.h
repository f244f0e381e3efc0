// Host-side launchers for the HIP kernels (implemented in the .hip files).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "nuts_types.h"

namespace nutsgpu {

enum ModelKind : int {
  MK_BERNOULLI = 0,
  MK_GAUSS,
  MK_EIGHT_SCHOOLS,
  MK_HIER_NORMAL_NC,
  MK_HIER_LOGISTIC,
  MK_LOGISTIC,
  MK_PLUGIN,  // caller-supplied model (nuts_model_plugin): lockstep only, gradient by callback
  MK_COUNT
};

struct ModelDataDev {
  const double* data;
  const int64_t* off;  // [C]
  double jac;          // 1 adds the log-Jacobian (sampling), 0 for log_prob jacobian=0
};

// Advance every chain to its next gradient request (lockstep) or through up
// to P.fused_budget gradients computed in place by the model (fused).
void launch_advance_external(int metric, const NutsParams& P, const NutsBuffers& B, hipStream_t st);
// the same with one 256-thread block per chain (large D)
void launch_advance_external_block(int metric, const NutsParams& P, const NutsBuffers& B, hipStream_t st);
// Cholesky factor of every chain's dense inverse metric (B.dM -> B.dL); counts failures.
void launch_dense_factor(const NutsBuffers& B, int* failed, hipStream_t st);
// vec32: FP32 vector storage (unit/diag, models that support it); sums, H, U-turn
// dot products and tree weights stay FP64
void launch_advance_fused(int kind, int metric, bool vec32, const NutsParams& P, const NutsBuffers& B,
                          const ModelDataDev& md, hipStream_t st);
// the same with one 256-thread block per chain (false: the model has no
// block-wide gradient, or dense_e)
bool launch_advance_fused_block(int kind, int metric, bool vec32, const NutsParams& P, const NutsBuffers& B,
                                const ModelDataDev& md, hipStream_t st);
bool model_supports_fused_block(int kind, int metric);
bool model_supports_vec32(int kind);

// Per-chain model gradient: V[c], G row c at Q row c for chains with need[c]
// (need == nullptr: all chains).  Clears need[c].
void launch_model_grad(int kind, int C, const int* D, const int64_t* off, const double* Q, double* G, double* V,
                       int* need, const ModelDataDev& md, hipStream_t st);

// Constrained outputs: in rows [C][rows][7 + D_c] -> out rows [C][rows][7 + nout_c]
void launch_write_array(int kind, int C, int rows, const int* D, const int64_t* in_off, const int64_t* out_off,
                        const int* nout, const double* in, double* out, const ModelDataDev& md, hipStream_t st);

// Batched logistic-regression gradient over a uniform-D chain batch.
struct GlmArgs {
  const double* X; const double* y;
  int N, D, C, rows_per_chunk;
  double s2inv;
  const double* Q; int64_t ldq;
  double* G; double* V;
  double* work;
  const int* idx = nullptr;       // optional compacted chain list (positions -> chains)
  const int* n_active = nullptr;  // its length (device)
  hipEvent_t ev_begin = nullptr;  // optional events recorded around the contraction kernel alone
  hipEvent_t ev_end = nullptr;
};

// FP64 matrix-core version (the sampler's default for logistic regression).
struct GlmPlan {
  int ngroups = 0, nchunks = 0, rows_per_chunk = 0;  // the split with every chain active
  int total_blocks = 0;                               // grid of every launch (the split adapts to the active count)
  size_t work_doubles = 0;
};
GlmPlan glm_mfma_plan(int N, int C, int D, int num_cus, int blocks_per_cu = 2);
// tally (optional, device) accumulates the number of active chains over launches.
void launch_compact_active(int* need, int C, int* idx, int* n_active, long long* tally, hipStream_t st);
void glm_reduce_launch(const GlmArgs& a, const GlmPlan& p, hipStream_t st);

// FP32 matrix-core version; X pre-converted to FP32 (precision 1) or BF16
// (precision 2) by glm_convert_x.  `layout` selects the measured register
// layout of the 16x16 f32 result.
void glm_convert_x(const double* X, void* out, int64_t n, int precision, hipStream_t st);
int glm_grad_mfma_f32(const GlmArgs& a, const GlmPlan& p, const void* Xlow, int precision, int layout,
                      hipStream_t st);
int glm_grad_mfma_f64(const GlmArgs& a, const GlmPlan& p, hipStream_t st);

// BF16 matrix-core version with FP32-level accuracy: every FP32 operand as three
// BF16 terms (X as three terms for precision 1, one exact term when X is stored
// in BF16, precision 2).  glm_bf16_prepare_x writes the terms of X into `out`
// (precision 1: 3 x N x ldx, precision 2: N x ldx values, ldx = glm_bf16_ldx(D)).
int glm_bf16_ldx(int D);
void glm_bf16_prepare_x(const double* X, int N, int D, int precision, unsigned short* out, hipStream_t st);
int glm_grad_mfma_bf16(const GlmArgs& a, const GlmPlan& p, const unsigned short* Xs, int precision, hipStream_t st);

}  // namespace nutsgpu
