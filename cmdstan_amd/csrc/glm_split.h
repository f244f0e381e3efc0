// Work split of the batched GLM gradient, recomputed on the device from the
// active-chain count of each launch: chain groups of 64 active chains times
// row chunks (a multiple of the 8 XCDs) fill a fixed grid of total_blocks, so
// a launch with few active chains still spreads its rows over every CU.
#pragma once
#include <hip/hip_runtime.h>

namespace nutsgpu {

struct GlmSplit {
  int groups;  // 64-chain groups with at least one active chain
  int chunks;  // row chunks
  int rows;    // rows per chunk (multiple of 64)
  int stride;  // chain positions per chunk in the partial workspace (groups * 64)
};

__host__ __device__ inline GlmSplit glm_split(int n_act, int N, int total_blocks) {
  GlmSplit s;
  s.groups = n_act > 0 ? (n_act + 63) / 64 : 1;
  int ch = (total_blocks / s.groups) & ~7;
  const int max_ch = (((N + 63) / 64) + 7) & ~7;
  if (ch > max_ch) ch = max_ch;
  if (ch < 8) ch = 8;
  int rpc = (N + ch - 1) / ch;
  rpc = ((rpc + 63) / 64) * 64;
  s.chunks = ch;
  s.rows = rpc;
  s.stride = s.groups * 64;
  return s;
}

}  // namespace nutsgpu
