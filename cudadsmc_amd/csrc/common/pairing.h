// Collision-pair selection of the PHYSICS-mode NTC in rounds of disjoint
// pairs, shared by the HIP kernels and the CPU oracle.
//
// A cell with n >= 2 particles and M candidate pairs (Bird's no-time-counter
// count) is processed in rounds of at most P = floor(n / 2) pairs. The pairs
// of a round are disjoint: pair k of round r is (pi(2k), pi(2k + 1)) for a
// random permutation pi of the cell's particles, so every pair of the round
// can be tested and collided at once, and each unordered pair of particles
// is a candidate with probability 1 / (n (n - 1) / 2) per candidate slot,
// as in the sequential NTC loop. A round with a single pair draws it
// directly: i uniform in [0, n), j uniform in [0, n) without i.
//
// The permutation is a keyed pseudo-random permutation of [0, n): a
// balanced six-round Feistel network on the smallest even bit width b with
// 2^b >= n, cycle-walked into [0, n) (Black and Rogaway, "Ciphers with
// arbitrary finite domains", CT-RSA 2002); the round function of round i
// is the murmur3 32-bit finaliser of (half ^ key[i mod 4] ^ i * 0x9E3779B9).
// Four rounds leave pair frequencies measurably uneven for some n (chi-square
// z of 5-30 at n = 5, 8, 64); six are uniform within the test's resolution.
//
// All draws are Philox4x32-10 (philox.h) under key (seed, kNtcKey1) at
// counter (cell, step, round, item):
//   item 0xFFFFFFFF  round keys of the permutation, or (i, j) of a
//                    single-pair round;
//   item k           pair k: acceptance number, cos(theta), azimuth.
#pragma once

#include <cstdint>

#include "philox.h"

namespace dsmc_pairing {

constexpr uint32_t kNtcKey1 = 0x3C6EF372u;
constexpr uint32_t kRoundItem = 0xFFFFFFFFu;

PX_HD uint32_t pp_mix(uint32_t x) {
  x ^= x >> 16;
  x *= 0x85EBCA6Bu;
  x ^= x >> 13;
  x *= 0xC2B2AE35u;
  x ^= x >> 16;
  return x;
}

struct PairPerm {
  uint32_t k[4];
  uint32_t half, mask, n;
};

PX_HD PairPerm pp_make(uint32_t n, const dsmc_rng::U32x4& keys) {
  uint32_t b = 2;
  while (b < 32 && (1ull << b) < (uint64_t)n) ++b;
  b += b & 1u;
  PairPerm p;
  for (int i = 0; i < 4; ++i) p.k[i] = keys.v[i];
  p.half = b / 2u;
  p.mask = (1u << p.half) - 1u;
  p.n = n;
  return p;
}

constexpr int kFeistelRounds = 6;

PX_HD uint32_t pp_feistel(const PairPerm& p, uint32_t x) {
  uint32_t L = x >> p.half, R = x & p.mask;
#ifdef __HIPCC__
#pragma unroll
#endif
  for (int i = 0; i < kFeistelRounds; ++i) {
    const uint32_t t = L ^ (pp_mix(R ^ p.k[i & 3] ^ ((uint32_t)i * 0x9E3779B9u)) & p.mask);
    L = R;
    R = t;
  }
  return (L << p.half) | R;
}

// pi(x) for x in [0, n): the network is a bijection of [0, 2^b), so walking
// its cycle from x reaches an element below n again.
PX_HD uint32_t pp_apply(const PairPerm& p, uint32_t x) {
  x = pp_feistel(p, x);
  while (x >= p.n) x = pp_feistel(p, x);
  return x;
}

// Pairs of round r of a cell with n particles and M candidates (0 when the
// cell has no round r).
PX_HD uint32_t round_pairs(uint32_t n, uint32_t M, uint32_t r) {
  const uint32_t P = n / 2u;
  if (P == 0u) return 0u;
  const uint64_t done = (uint64_t)r * P;
  if (done >= M) return 0u;
  const uint64_t left = (uint64_t)M - done;
  return left < P ? (uint32_t)left : P;
}

PX_HD uint32_t rounds_of(uint32_t n, uint32_t M) {
  const uint32_t P = n / 2u;
  return P == 0u ? 0u : (uint32_t)(((uint64_t)M + P - 1u) / P);
}

// The two particles (cell-local indices) of pair k of a round with m pairs;
// `rk` = Philox at the round's counter (cell, step, r, kRoundItem).
PX_HD void pair_members(uint32_t n, uint32_t m, uint32_t k, const dsmc_rng::U32x4& rk, uint32_t& i, uint32_t& j) {
  if (m == 1u) {
    i = (uint32_t)(((uint64_t)rk.v[0] * n) >> 32);
    j = (uint32_t)(((uint64_t)rk.v[1] * (n - 1u)) >> 32);
    if (j >= i) ++j;
    return;
  }
  const PairPerm p = pp_make(n, rk);
  i = pp_apply(p, 2u * k);
  j = pp_apply(p, 2u * k + 1u);
}

}  // namespace dsmc_pairing
