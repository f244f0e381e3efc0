// Instantiations of the NUTS state machine (lockstep and fused per model) and
// of the per-chain model kernels.  Compiled with -ffp-contract=off so the
// leapfrog and Hamiltonian arithmetic rounds like the FP64 reference restatement.
#include "launch.h"
#include "flow.h"
#include "models_device.h"
#include "nuts_device.h"

namespace nutsgpu {

template <int METRIC>
__global__ void __launch_bounds__(256) advance_external_kernel(NutsParams P, NutsBuffers B) {
  nuts_advance_body<METRIC>(P, B, ExternalGrad{});
}

// one 256-thread block per chain (large D)
template <int METRIC>
__global__ void __launch_bounds__(256) advance_external_block_kernel(NutsParams P, NutsBuffers B) {
  nuts_advance_body<METRIC, ExternalGrad, 256>(P, B, ExternalGrad{});
}

template <int METRIC, class F, class VT>
__global__ void __launch_bounds__(256) advance_fused_kernel(NutsParams P, NutsBuffers B, F f) {
  nuts_advance_body<METRIC, F, NUTS_WAVE, VT>(P, B, f);
}

// one 256-thread block per chain with the model's gradient spread over the
// block (models with F::block_grad): for few, large chains
template <int METRIC, class F, class VT>
__global__ void __launch_bounds__(256) advance_fused_block_kernel(NutsParams P, NutsBuffers B, F f) {
  nuts_advance_body<METRIC, F, 256, VT>(P, B, f);
}

template <class F>
__global__ void __launch_bounds__(256) model_grad_kernel(F f, int C, const int* D, const int64_t* off, const double* Q,
                                                         double* G, double* V, int* need) {
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= C) return;
  if (need && !need[c]) return;
  const double v = f(c, Q + off[c], G + off[c], D[c], lane);
  if (lane == 0) {
    V[c] = v;
    if (need) need[c] = 0;
  }
}

// Model gradient at packed rows (the flow's x = T(z)): row r belongs to chain idx[r].
template <class F>
__global__ void __launch_bounds__(256) model_grad_rows_kernel(F f, const int* idx, const int* n_dev, int n, int D,
                                                              const double* X, double* GX, double* VX) {
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int nr = n_dev ? min(n, *n_dev) : n;
  if (r >= nr) return;
  const int c = idx ? idx[r] : r;
  const double v = f(c, X + (int64_t)r * D, GX + (int64_t)r * D, D, lane);
  if (lane == 0) VX[r] = v;
}

template <class F>
__global__ void write_array_kernel(F f, int C, int rows, const int* D, const int64_t* in_off, const int64_t* out_off,
                                   const int* nout, const double* in, double* out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)C * rows) return;
  const int c = (int)(t / rows), r = (int)(t % rows);
  const double* src = in + in_off[c] + (int64_t)r * (7 + D[c]);
  double* dst = out + out_off[c] + (int64_t)r * (7 + nout[c]);
  for (int i = 0; i < 7; ++i) dst[i] = src[i];
  f.write_array(c, src + 7, dst + 7);
}

static int blocks_for(int C) { return (C + 3) / 4; }

void launch_advance_external(int metric, const NutsParams& P, const NutsBuffers& B, hipStream_t st) {
  const dim3 grid(blocks_for(B.C)), blk(256);
  if (metric == METRIC_UNIT) hipLaunchKernelGGL(advance_external_kernel<METRIC_UNIT>, grid, blk, 0, st, P, B);
  else if (metric == METRIC_DENSE) hipLaunchKernelGGL(advance_external_kernel<METRIC_DENSE>, grid, blk, 0, st, P, B);
  else hipLaunchKernelGGL(advance_external_kernel<METRIC_DIAG>, grid, blk, 0, st, P, B);
}

void launch_advance_external_block(int metric, const NutsParams& P, const NutsBuffers& B, hipStream_t st) {
  const dim3 grid(B.C), blk(256);
  if (metric == METRIC_UNIT) hipLaunchKernelGGL(advance_external_block_kernel<METRIC_UNIT>, grid, blk, 0, st, P, B);
  else if (metric == METRIC_DENSE)
    hipLaunchKernelGGL(advance_external_block_kernel<METRIC_DENSE>, grid, blk, 0, st, P, B);
  else hipLaunchKernelGGL(advance_external_block_kernel<METRIC_DIAG>, grid, blk, 0, st, P, B);
}

void launch_dense_factor(const NutsBuffers& B, int* failed, hipStream_t st) {
  hipLaunchKernelGGL(dense_factor_kernel, dim3(blocks_for(B.C)), dim3(256), 0, st, B, failed);
}

// vec32: the sampler's vectors in FP32 (unit/diag metrics, models with F::vec32)
template <class F>
static void fused_for_metric(int metric, bool vec32, const NutsParams& P, const NutsBuffers& B, const F& f,
                             hipStream_t st) {
  const dim3 grid(blocks_for(B.C)), blk(256);
  if constexpr (F::vec32) {
    if (vec32) {
      if (metric == METRIC_UNIT)
        hipLaunchKernelGGL((advance_fused_kernel<METRIC_UNIT, F, float>), grid, blk, 0, st, P, B, f);
      else hipLaunchKernelGGL((advance_fused_kernel<METRIC_DIAG, F, float>), grid, blk, 0, st, P, B, f);
      return;
    }
  }
  if (metric == METRIC_UNIT) hipLaunchKernelGGL((advance_fused_kernel<METRIC_UNIT, F, double>), grid, blk, 0, st, P, B, f);
  else if (metric == METRIC_DENSE)
    hipLaunchKernelGGL((advance_fused_kernel<METRIC_DENSE, F, double>), grid, blk, 0, st, P, B, f);
  else hipLaunchKernelGGL((advance_fused_kernel<METRIC_DIAG, F, double>), grid, blk, 0, st, P, B, f);
}

template <class F>
static bool fused_block_for_metric(int metric, bool vec32, const NutsParams& P, const NutsBuffers& B, const F& f,
                                   hipStream_t st) {
  if constexpr (!F::block_grad) {
    return false;
  } else {
    if (metric == METRIC_DENSE) return false;
    const dim3 grid(B.C), blk(256);
    if constexpr (F::vec32) {
      if (vec32) {
        if (metric == METRIC_UNIT)
          hipLaunchKernelGGL((advance_fused_block_kernel<METRIC_UNIT, F, float>), grid, blk, 0, st, P, B, f);
        else hipLaunchKernelGGL((advance_fused_block_kernel<METRIC_DIAG, F, float>), grid, blk, 0, st, P, B, f);
        return true;
      }
    }
    if (metric == METRIC_UNIT)
      hipLaunchKernelGGL((advance_fused_block_kernel<METRIC_UNIT, F, double>), grid, blk, 0, st, P, B, f);
    else hipLaunchKernelGGL((advance_fused_block_kernel<METRIC_DIAG, F, double>), grid, blk, 0, st, P, B, f);
    return true;
  }
}

#define NUTS_MODEL_SWITCH(kind, md, CALL)                       \
  switch (kind) {                                               \
    case MK_BERNOULLI: { BernoulliFn f{{md.data, md.off, md.jac}}; CALL; break; }    \
    case MK_GAUSS: { GaussFn f{{md.data, md.off, md.jac}}; CALL; break; }            \
    case MK_EIGHT_SCHOOLS: { EightSchoolsFn f{{md.data, md.off, md.jac}}; CALL; break; } \
    case MK_HIER_NORMAL_NC: { HierNormalNCFn f{{md.data, md.off, md.jac}}; CALL; break; } \
    case MK_HIER_LOGISTIC: { HierLogisticFn f{{md.data, md.off, md.jac}}; CALL; break; } \
    case MK_LOGISTIC: { LogisticFn f{{md.data, md.off, md.jac}}; CALL; break; }      \
    default: break;                                             \
  }

void launch_advance_fused(int kind, int metric, bool vec32, const NutsParams& P, const NutsBuffers& B,
                          const ModelDataDev& md, hipStream_t st) {
  NUTS_MODEL_SWITCH(kind, md, fused_for_metric(metric, vec32, P, B, f, st));
}

bool launch_advance_fused_block(int kind, int metric, bool vec32, const NutsParams& P, const NutsBuffers& B,
                                const ModelDataDev& md, hipStream_t st) {
  bool ok = false;
  NUTS_MODEL_SWITCH(kind, md, ok = fused_block_for_metric(metric, vec32, P, B, f, st));
  return ok;
}

bool model_supports_fused_block(int kind, int metric) {
  bool r = false;
  ModelDataDev md{nullptr, nullptr, 1.0};
  NUTS_MODEL_SWITCH(kind, md, r = decltype(f)::block_grad);
  return r && metric != METRIC_DENSE;
}

bool model_supports_vec32(int kind) {
  bool r = false;
  ModelDataDev md{nullptr, nullptr, 1.0};
  NUTS_MODEL_SWITCH(kind, md, r = decltype(f)::vec32);
  return r;
}

void launch_model_grad(int kind, int C, const int* D, const int64_t* off, const double* Q, double* G, double* V,
                       int* need, const ModelDataDev& md, hipStream_t st) {
  const dim3 grid(blocks_for(C)), blk(256);
  NUTS_MODEL_SWITCH(kind, md,
                    hipLaunchKernelGGL((model_grad_kernel<decltype(f)>), grid, blk, 0, st, f, C, D, off, Q, G, V, need));
}

void launch_model_grad_rows(int kind, const int* idx, const int* n_dev, int n_rows, int D, const double* X,
                            double* GX, double* VX, const ModelDataDev& md, hipStream_t st) {
  const dim3 grid(blocks_for(n_rows)), blk(256);
  NUTS_MODEL_SWITCH(kind, md,
                    hipLaunchKernelGGL((model_grad_rows_kernel<decltype(f)>), grid, blk, 0, st, f, idx, n_dev, n_rows,
                                       D, X, GX, VX));
}

void launch_write_array(int kind, int C, int rows, const int* D, const int64_t* in_off, const int64_t* out_off,
                        const int* nout, const double* in, double* out, const ModelDataDev& md, hipStream_t st) {
  const int64_t n = (int64_t)C * rows;
  if (n == 0) return;
  const dim3 grid((unsigned)((n + 127) / 128)), blk(128);
  NUTS_MODEL_SWITCH(kind, md,
                    hipLaunchKernelGGL((write_array_kernel<decltype(f)>), grid, blk, 0, st, f, C, rows, D, in_off,
                                       out_off, nout, in, out));
}

}  // namespace nutsgpu
