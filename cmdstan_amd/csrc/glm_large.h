// Batched logistic-regression gradient for large D on FP32 matrix cores, X
// stored in FP32 or BF16 (two passes: Eta with a fused residual epilogue,
// then R^T X split over row ranges; FP64 sums of the partials).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace nutsgpu {

struct GlLargeArgs {
  const double* y;
  int N, D, C;
  double s2inv;                  // prior precision 1/s^2
  const double* Q; int64_t ldq;  // chain positions (FP64), row stride
  double* G; double* V;          // gradient of V = -lp (stride ldq), V
  const int* idx = nullptr;      // active positions -> chains (or null: all)
  const int* n_active = nullptr;
  float* Qf = nullptr; int Dq = 0;       // packed FP32 positions [C][Dq]
  float* R = nullptr; int64_t ldr = 0;   // residuals [N][ldr]
  double* Wg = nullptr;                  // [splits][C][D]
  double* Wlp = nullptr;                 // [row tiles][C]
  // BF16 matrix-core variant (X stored in BF16): positions and residuals as
  // three BF16 terms, and X transposed [Dp][ldn]
  unsigned short* Qb = nullptr;          // [3][Cp][Dk]
  unsigned short* Rb = nullptr;          // [3][Cp][ldn]
  const unsigned short* Xt = nullptr;    // [Dp][ldn]
  hipEvent_t ev_begin = nullptr, ev_end = nullptr;  // optional, around the two passes
};

struct GlLargePlan {
  int row_tiles = 0, Dq = 0, splits = 0, rows_per_split = 0;
  int64_t ldr = 0;
  size_t r_floats = 0, q_floats = 0, wg_doubles = 0, wlp_doubles = 0;
  // BF16 variant: rows padded to whole 128-row tiles (ldn), chains to whole
  // 128-chain tiles (Cp), coefficients to 64 (Dk, pass-1 K) and 128 (Dp)
  int64_t ldn = 0;
  int Cp = 0, Dk = 0, Dp = 0;
  size_t qb_elems = 0, rb_elems = 0, xt_elems = 0;
};

GlLargePlan gl_large_plan(int N, int C, int D, int num_cus);
// xtype 1: X in FP32, 2: X in BF16 (both on FP32 matrix cores)
void gl_large_grad(const GlLargeArgs& a, const GlLargePlan& p, const void* X, int xtype, hipStream_t st);
void gl_reduce_launch(const GlLargeArgs& a, const GlLargePlan& p, hipStream_t st);

// X in BF16 on BF16 matrix cores (needs D a multiple of 8, a.Qb, a.Rb, a.Xt)
bool gl_large_bf16_ok(int D);
void gl_large_prepare_xt(const unsigned short* Xb, int N, int D, unsigned short* Xt, const GlLargePlan& p,
                         hipStream_t st);
void gl_large_grad_bf16(const GlLargeArgs& a, const GlLargePlan& p, const unsigned short* X, hipStream_t st);

}  // namespace nutsgpu
