// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror and Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11): ten rounds of
//   (c0, c1, c2, c3) -> (hi(M1 c2) ^ c1 ^ k0, lo(M1 c2), hi(M0 c0) ^ c3 ^ k1, lo(M0 c0))
// with the key bumped by the Weyl constants between rounds. Stateless: the
// draws of a wall interaction are a function of (seed, step, slot, bounce)
// only, so any thread may process any particle.
#pragma once

#include <cstdint>

#ifdef __HIPCC__
#define PX_HD __host__ __device__ __forceinline__
#else
#define PX_HD static inline
#endif

namespace dsmc_rng {

struct U32x4 {
  uint32_t v[4];
};

PX_HD void px_mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
  const uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}

PX_HD U32x4 philox4x32_10(U32x4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#ifdef __HIPCC__
#pragma unroll
#endif
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    px_mulhilo(M0, c.v[0], hi0, lo0);
    px_mulhilo(M1, c.v[2], hi1, lo1);
    U32x4 n;
    n.v[0] = hi1 ^ c.v[1] ^ k0;
    n.v[1] = lo1;
    n.v[2] = hi0 ^ c.v[3] ^ k1;
    n.v[3] = lo0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// (u + 1) / 2^32 in (0, 1], the same mapping as the MT streams' floats.
PX_HD float px_unit(uint32_t u) { return ((float)u + 1.0f) / 4294967296.0f; }

}  // namespace dsmc_rng
