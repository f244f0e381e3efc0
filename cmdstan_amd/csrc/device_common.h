// Device helpers shared by the sampler and model kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define NUTS_WAVE 64

namespace nutsgpu {

// ------------------------------------------------------------ Philox4x32-10
// Salmon et al., SC'11.  Counter = (idx lo, idx hi, chain, purpose), key = seed.
struct U4 { uint32_t x, y, z, w; };

__device__ __forceinline__ U4 philox10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    U4 o;
    o.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
    o.y = (uint32_t)p1;
    o.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
    o.w = (uint32_t)p0;
    c = o;
  }
  return c;
}

enum : uint32_t { STREAM_UNIFORM = 0u, STREAM_NORMAL = 1u, STREAM_INIT = 2u };

__device__ __forceinline__ U4 stream_block(uint64_t seed, uint32_t chain, uint32_t purpose, uint64_t idx) {
  U4 c{(uint32_t)idx, (uint32_t)(idx >> 32), chain, purpose};
  return philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// (k + 0.5) * 2^-52 with k the top 52 of 64 random bits: exact, strictly in (0,1).
__device__ __forceinline__ double bits_to_unit(uint32_t a, uint32_t b) {
  const uint64_t k = ((uint64_t)a << 20) | (uint64_t)(b >> 12);
  return ((double)k + 0.5) * 2.220446049250313080847e-16;
}

__device__ __forceinline__ double stream_uniform(uint64_t seed, uint32_t chain, uint32_t purpose, uint64_t idx) {
  U4 r = stream_block(seed, chain, purpose, idx);
  return bits_to_unit(r.x, r.y);
}

// Box-Muller; element `which` (0: cos, 1: sin) of pair `idx`.
__device__ __forceinline__ double stream_normal(uint64_t seed, uint32_t chain, uint64_t idx, int which) {
  U4 r = stream_block(seed, chain, STREAM_NORMAL, idx);
  const double u1 = bits_to_unit(r.x, r.y);
  const double u2 = bits_to_unit(r.z, r.w);
  const double rad = sqrt(-2.0 * log(u1));
  const double ang = 6.283185307179586476925286766559 * u2;
  return which ? rad * sin(ang) : rad * cos(ang);
}

// ----------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = NUTS_WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, NUTS_WAVE);
  return v;
}

__device__ __forceinline__ int wave_and(int v) {
#pragma unroll
  for (int off = NUTS_WAVE / 2; off > 0; off >>= 1) v &= __shfl_xor(v, off, NUTS_WAVE);
  return v;
}

// ------------------------------------------------------------ scalar math
// 1/d for d in [1, 8]: the hardware reciprocal estimate refined by two Newton
// steps (no range reduction is needed in this range); within an ulp or two of
// the correctly rounded quotient, at about half the instructions of a division
__device__ __forceinline__ double recip_small(double d) {
  double x = __builtin_amdgcn_rcp(d);
  x = __builtin_fma(x, __builtin_fma(-d, x, 1.0), x);
  return __builtin_fma(x, __builtin_fma(-d, x, 1.0), x);
}

// log(1 + t) and r1 = 1/(1 + t) for t in [0, 1] from one reciprocal.  Above
// t = sqrt(2) - 1 the argument is halved, log(1 + t) = log 2 + 2 atanh(s) with
// s = (t - 1) / (t + 3); below it s = t / (2 + t).  Either way |s| <=
// 3 - 2 sqrt(2) = 0.1716, and 2 (s + s^3/3 + ... + s^21/21) leaves out terms
// below 1e-18 of the result (10 fused multiply-adds; within 2 ulp of log1p
// over [0, 1]).  1 / ((2 + t or 3 + t) (1 + t)) gives both quotients.
__device__ __forceinline__ double log1p_recip_unit(double t, double& r1) {
  const bool hi = t > 0.41421356237309503;
  const double num = hi ? t - 1.0 : t;
  const double den = hi ? t + 3.0 : t + 2.0;
  const double d1 = 1.0 + t;
  const double inv = recip_small(den * d1);
  r1 = den * inv;
  const double s = num * (d1 * inv);
  const double z = s * s;
  double p = 1.0 / 21.0;
#pragma unroll
  for (int k = 19; k >= 1; k -= 2) p = __builtin_fma(p, z, 1.0 / (double)k);
  return __builtin_fma(2.0 * s, p, hi ? 0.69314718055994531 : 0.0);
}

__device__ __forceinline__ double log1p_exp(double a) {
  return a > 0.0 ? a + log1p(exp(-a)) : log1p(exp(a));
}

// stan::math::log_sum_exp(double, double)
__device__ __forceinline__ double log_sum_exp2(double a, double b) {
  const double inf = __builtin_inf();
  if (a == -inf) return b;
  if (a == inf && b == inf) return inf;
  if (a > b) return a + log1p_exp(b - a);
  return b + log1p_exp(a - b);
}

__device__ __forceinline__ double inv_logit(double x) {
  if (x >= 0) return 1.0 / (1.0 + exp(-x));
  const double e = exp(x);
  return e / (1.0 + e);
}

}  // namespace nutsgpu
