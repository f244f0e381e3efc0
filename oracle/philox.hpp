// ORACLE — test infrastructure only. Nothing in the product may include, link or
// execute this file; only tests/, the smoke check and the bench's cpu_baseline leg.
//
// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC'11), restated from the paper, plus the
// project's fixed mapping from (seed, chain, purpose, index) to uniforms and
// normals.  The upstream sampler uses boost's ecuyer1988/mixmax stream
// (stan::services::util::create_rng, called via cmdstan command.hpp:681-688);
// that stream is not reproducible here, so the project defines its own and both
// the oracle and the HIP path implement this definition independently.
//
// Pinned by the Random123 known-answer vectors for philox4x32_10 (tests/test_oracle.py).
#pragma once
#include <cmath>
#include <cstdint>

namespace oracle {

struct P4 { uint32_t v[4]; };

inline void mulhilo32(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

// 10 rounds; key bumped between rounds (9 bumps).
inline P4 philox4x32_10(P4 ctr, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k0 += W0; k1 += W1; }
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo32(M0, ctr.v[0], &hi0, &lo0);
    mulhilo32(M1, ctr.v[2], &hi1, &lo1);
    P4 o;
    o.v[0] = hi1 ^ ctr.v[1] ^ k0;
    o.v[1] = lo1;
    o.v[2] = hi0 ^ ctr.v[3] ^ k1;
    o.v[3] = lo0;
    ctr = o;
  }
  return ctr;
}

// Purposes (4th counter word) keep the three streams of a chain independent.
enum : uint32_t { PURPOSE_UNIFORM = 0u, PURPOSE_NORMAL = 1u, PURPOSE_INIT = 2u };

// Uniform strictly inside (0,1) from 52 random bits: (k + 0.5) * 2^-52, exact
// in binary64 (k + 0.5 needs 53 significant bits), so every platform agrees.
inline double u53(uint32_t a, uint32_t b) {
  uint64_t k = ((uint64_t)a << 20) | (uint64_t)(b >> 12);
  return ((double)k + 0.5) * 2.220446049250313080847e-16;
}

inline P4 stream_block(uint64_t seed, uint32_t chain, uint32_t purpose, uint64_t idx) {
  P4 c;
  c.v[0] = (uint32_t)idx;
  c.v[1] = (uint32_t)(idx >> 32);
  c.v[2] = chain;
  c.v[3] = purpose;
  return philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// One uniform per block (words 0,1).
inline double stream_uniform(uint64_t seed, uint32_t chain, uint32_t purpose, uint64_t idx) {
  P4 r = stream_block(seed, chain, purpose, idx);
  return u53(r.v[0], r.v[1]);
}

// Box-Muller pair from one block: (r cos 2πu2, r sin 2πu2), r = sqrt(-2 log u1).
inline void stream_normal_pair(uint64_t seed, uint32_t chain, uint64_t idx, double* z0, double* z1) {
  P4 r = stream_block(seed, chain, PURPOSE_NORMAL, idx);
  double u1 = u53(r.v[0], r.v[1]);
  double u2 = u53(r.v[2], r.v[3]);
  double rad = std::sqrt(-2.0 * std::log(u1));
  double ang = 6.283185307179586476925286766559 * u2;
  *z0 = rad * std::cos(ang);
  *z1 = rad * std::sin(ang);
}

}  // namespace oracle
