// Batched gradient of the centered hierarchical logistic model (hier_logistic_cp)
// over many chains of one dataset.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace nutsgpu {

struct HlArgs {
  const double* block;   // device dataset block [N, K, G, mu_s, gstart[G+1], X, y]
  const double* Q;       // positions: row idx[pos] (idx_in) or pos, stride ldq
  int64_t ldq;
  const int* idx;        // active positions -> chains
  bool idx_in, idx_out;  // rows of Q / of Gout and Vout are chains (true) or positions
  const int* n_dev;      // active count (device) or null for n_rows
  int n_rows;
  double* Gout;          // gradient of V = -lp, stride ldg
  int64_t ldg;
  double* Vout;
  double* work;          // hl_work_doubles(G, K, n_rows)
  double jac;            // 1: add the log-Jacobian of tau = exp(v)
};

void hl_grad_launch(const HlArgs& a, int G, hipStream_t st);
size_t hl_work_doubles(int G, int K, int n_rows);

}  // namespace nutsgpu
