// Example model plugin (nuts_model_plugin, include/cmdstan_amd.h): Neal's
// funnel, v ~ normal(0, 3), x_k ~ normal(0, exp(v / 2)), k = 1..K, with the
// derived output sigma = exp(v / 2).  It plays the part of a stanc-generated
// model_base (src/cmdstan/command.hpp:71-72): the host callbacks give the
// names, write_array and transform_inits; the gradient is this file's own
// gfx950 kernel, enqueued on the sampler's stream.  Built into
// libfunnel_plugin.so, linked against the library only through its C ABI.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/cmdstan_amd.h"

namespace {

struct Funnel {
  int K;
  std::vector<std::string> names, unames;
  std::vector<const char*> pn, un;
};

// One wavefront per point: V = -log p = v^2 / 18 + sum_k (x_k^2 e^{-v} + v) / 2
// and its gradient.
__global__ void __launch_bounds__(256) funnel_grad_kernel(const double* __restrict__ Q, int64_t ldq,
                                                          const int* __restrict__ idx, const int* __restrict__ d_n,
                                                          int n_rows, int K, double* __restrict__ G,
                                                          double* __restrict__ V) {
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int n = d_n ? min(*d_n, n_rows) : n_rows;
  if (r >= n) return;
  const int64_t c = idx ? idx[r] : r;
  const double* q = Q + c * ldq;
  double* g = G + c * ldq;
  const double v = q[0], iv = exp(-v);
  double s = 0.0;
  for (int k = lane; k < K; k += 64) {
    const double x = q[1 + k];
    s += x * x;
    g[1 + k] = x * iv;
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) {
    g[0] = v / 9.0 - 0.5 * s * iv + 0.5 * K;
    V[c] = v * v / 18.0 + 0.5 * s * iv + 0.5 * K * v;
  }
}

int funnel_log_prob_grad(void* user, const double* Q, int64_t ldq, const int* idx, const int* d_n, int n_rows,
                         int /*jacobian: no constrained parameters*/, double* G, double* V, void* stream) {
  const Funnel* f = static_cast<const Funnel*>(user);
  if (n_rows < 1) return 0;
  hipLaunchKernelGGL(funnel_grad_kernel, dim3((n_rows + 3) / 4), dim3(256), 0, static_cast<hipStream_t>(stream), Q,
                     ldq, idx, d_n, n_rows, f->K, G, V);
  return hipPeekAtLastError() == hipSuccess ? 0 : 1;
}

int funnel_write_array(void* user, const double* q, int n, double* out) {
  const Funnel* f = static_cast<const Funnel*>(user);
  const int D = f->K + 1, O = f->K + 2;
  for (int r = 0; r < n; ++r) {
    out[(int64_t)r * O] = q[(int64_t)r * D];
    out[(int64_t)r * O + 1] = std::exp(0.5 * q[(int64_t)r * D]);
    for (int k = 0; k < f->K; ++k) out[(int64_t)r * O + 2 + k] = q[(int64_t)r * D + 1 + k];
  }
  return 0;
}

int funnel_transform_inits(void* user, const nuts_data* inits, double* out) {
  const Funnel* f = static_cast<const Funnel*>(user);
  int nd = 0;
  int64_t dims[8];
  const double* v = nullptr;
  if (nuts_data_get(inits, "v", &nd, dims, &v) == NUTS_OK) out[0] = v[0];
  if (nuts_data_get(inits, "x", &nd, dims, &v) == NUTS_OK) {
    if (nd != 1 || dims[0] != f->K) return 1;
    for (int k = 0; k < f->K; ++k) out[1 + k] = v[k];
  }
  return 0;
}

}  // namespace

// Fills `out` for a funnel with K scales; returns the handle to pass to
// funnel_plugin_destroy once every model made from `out` is destroyed.
extern "C" void* funnel_plugin_create(int K, nuts_model_plugin* out) {
  if (K < 1 || !out) return nullptr;
  Funnel* f = new Funnel();
  f->K = K;
  f->names = {"v", "sigma"};
  f->unames = {"v"};
  for (int k = 0; k < K; ++k) {
    f->names.push_back("x." + std::to_string(k + 1));
    f->unames.push_back("x." + std::to_string(k + 1));
  }
  for (auto& s : f->names) f->pn.push_back(s.c_str());
  for (auto& s : f->unames) f->un.push_back(s.c_str());
  out->dim = K + 1;
  out->num_outputs = K + 2;
  out->param_names = f->pn.data();
  out->unconstrained_names = f->un.data();
  out->user = f;
  out->log_prob_grad = funnel_log_prob_grad;
  out->write_array = funnel_write_array;
  out->transform_inits = funnel_transform_inits;
  return f;
}

extern "C" void funnel_plugin_destroy(void* h) { delete static_cast<Funnel*>(h); }
