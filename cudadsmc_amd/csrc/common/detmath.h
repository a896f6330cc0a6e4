// Deterministic single-precision math for the DSMC step.
//
// Every function here is a fixed sequence of IEEE-754 binary32 operations
// (add, sub, mul, correctly rounded div/sqrt, exact int<->float conversions)
// with no fused multiply-add, so the HIP build (gfx950, -ffp-contract=off) and
// the host build (x86-64 SSE, -ffp-contract=off) produce the same bits. This
// is what lets the GPU step be compared bit-for-bit with the CPU oracle even
// where the reference calls transcendental functions (loading: logf/sinf in
// dsmc_kernel.cu:394-415; scattering angles: __cosf/__sinf in
// dsmc_kernel.cu:1065; VHS cross-section powers).
//
// Methods (own implementations, coefficients from the series themselves):
//   sin/cos : Cody-Waite reduction by pi/2 (three-part split), Taylor series
//             of sin to x^9 and cos to x^10 on [-pi/4, pi/4].
//   log     : exponent/mantissa split with m in [sqrt(1/2), sqrt(2)),
//             log(m) = 2 atanh((m-1)/(m+1)) summed to the s^9 term,
//             ln2 split in two parts (Cody-Waite style).
//   exp2    : 2^x = 2^k * e^(f ln2), |f| <= 1/2, Taylor series of e^t to t^7.
//   pow     : x^y = exp2(y * log(x) / ln2) for x > 0.
// Accuracy against libm is checked in the CPU test-suite (a few ulp).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP_DEVICE_COMPILE__)
#define DM_HD __host__ __device__ __forceinline__
#else
#define DM_HD static inline
#endif

DM_HD uint32_t dm_bits(float f) { return __builtin_bit_cast(uint32_t, f); }
DM_HD float dm_from_bits(uint32_t u) { return __builtin_bit_cast(float, u); }

// floor for |x| < 2^31 (exact; the int conversion truncates toward zero).
DM_HD float dm_floorf(float x) {
  float t = (float)(int32_t)x;
  if (t > x) t = t - 1.0f;
  return t;
}

// Round half up; deterministic replacement for rintf on both targets.
DM_HD float dm_roundf(float x) { return dm_floorf(x + 0.5f); }

// float -> uint32 with the saturating semantics of the GPU conversion:
// NaN and negatives give 0, values >= 2^32 give 0xFFFFFFFF. Plain C++ casts
// are undefined for these inputs, so both targets use this helper.
DM_HD uint32_t dm_f2u_sat(float f) {
  if (!(f > 0.0f)) return 0u;
  if (f >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)f;
}

// sin and cos of x (|x| < ~6000; DSMC angles are in (0, 2*pi]).
DM_HD void dm_sincosf(float x, float* s_out, float* c_out) {
  const float TWO_OVER_PI = 0.636619746685028076171875f;
  // pi/2 = P1 + P2 + P3, P1 and P2 with 11 significant bits so k*P1, k*P2
  // are exact for |k| < 2^12.
  const float P1 = 1.5703125f;
  const float P2 = 4.837512969970703125e-4f;
  const float P3 = 7.549790126404332e-08f;
  float kf = dm_roundf(x * TWO_OVER_PI);
  int32_t k = (int32_t)kf;
  float r = x - kf * P1;
  r = r - kf * P2;
  r = r - kf * P3;
  float r2 = r * r;
  // sin r = r - r^3/3! + r^5/5! - r^7/7! + r^9/9!
  float ps = 2.7557319223985893e-06f;          //  1/9!
  ps = ps * r2 + (-1.9841269841269841e-04f);   // -1/7!
  ps = ps * r2 + 8.3333333333333333e-03f;      //  1/5!
  ps = ps * r2 + (-1.6666666666666666e-01f);   // -1/3!
  float sr = r + (r * r2) * ps;
  // cos r = 1 - r^2/2! + r^4/4! - r^6/6! + r^8/8! - r^10/10!
  float pc = -2.7557319223985888e-07f;          // -1/10!
  pc = pc * r2 + 2.4801587301587302e-05f;       //  1/8!
  pc = pc * r2 + (-1.3888888888888889e-03f);    // -1/6!
  pc = pc * r2 + 4.1666666666666664e-02f;       //  1/4!
  pc = pc * r2 + (-0.5f);                       // -1/2!
  float cr = 1.0f + r2 * pc;
  int32_t q = k & 3;
  float s, c;
  if (q == 0) { s = sr; c = cr; }
  else if (q == 1) { s = cr; c = -sr; }
  else if (q == 2) { s = -sr; c = -cr; }
  else { s = -cr; c = sr; }
  *s_out = s;
  *c_out = c;
}

DM_HD float dm_sinf(float x) { float s, c; dm_sincosf(x, &s, &c); return s; }
DM_HD float dm_cosf(float x) { float s, c; dm_sincosf(x, &s, &c); return c; }

// Natural logarithm. x == 0 -> -inf, x < 0 or NaN -> NaN, +inf -> +inf.
DM_HD float dm_logf(float x) {
  if (!(x > 0.0f)) {
    if (x == 0.0f) return dm_from_bits(0xFF800000u);
    return dm_from_bits(0x7FC00000u);
  }
  uint32_t u = dm_bits(x);
  if (u >= 0x7F800000u) return x;  // +inf
  int32_t e = 0;
  if (u < 0x00800000u) {  // subnormal: scale by 2^25 (exact)
    x = x * 33554432.0f;
    u = dm_bits(x);
    e = -25;
  }
  e += (int32_t)((u >> 23) & 0xFFu) - 127;
  float m = dm_from_bits((u & 0x007FFFFFu) | 0x3F800000u);  // [1, 2)
  if (m > 1.41421353816986083984375f) {
    m = m * 0.5f;
    e += 1;
  }
  float s = (m - 1.0f) / (m + 1.0f);
  float s2 = s * s;
  // 2 atanh(s) = 2s + 2s^3/3 + 2s^5/5 + 2s^7/7 + 2s^9/9
  float p = 2.2222222222222222e-01f;   // 2/9
  p = p * s2 + 2.8571428571428571e-01f;  // 2/7
  p = p * s2 + 4.0000000000000000e-01f;  // 2/5
  p = p * s2 + 6.6666666666666667e-01f;  // 2/3
  float lm = (s + s) + (s * s2) * p;
  const float LN2_HI = 0.693145751953125f;        // 16 significant bits
  const float LN2_LO = 1.428606765330187e-06f;    // ln2 - LN2_HI
  float ef = (float)e;
  return ef * LN2_HI + (ef * LN2_LO + lm);
}

// 2^x. Results below 2^-126 flush to zero (never reached by the DSMC paths).
DM_HD float dm_exp2f(float x) {
  if (x != x) return x;
  if (x >= 128.0f) return dm_from_bits(0x7F800000u);
  if (x < -126.0f) return 0.0f;
  float kf = dm_roundf(x);
  float f = x - kf;  // exact, |f| <= 1/2
  const float LN2 = 0.693147182464599609375f;
  float t = f * LN2;
  float p = 1.9841269841269841e-04f;     // 1/7!
  p = p * t + 1.3888888888888889e-03f;   // 1/6!
  p = p * t + 8.3333333333333333e-03f;   // 1/5!
  p = p * t + 4.1666666666666664e-02f;   // 1/4!
  p = p * t + 1.6666666666666666e-01f;   // 1/3!
  p = p * t + 0.5f;                      // 1/2!
  p = p * t + 1.0f;
  p = p * t + 1.0f;
  int32_t k = (int32_t)kf;
  if (k < -126) {  // p * 2^k with k in [-126-1, ...]: split the scaling
    p = p * dm_from_bits((uint32_t)(k + 127 + 1) << 23);
    return p * 0.5f;
  }
  return p * dm_from_bits((uint32_t)(k + 127) << 23);
}

DM_HD float dm_expf(float x) { return dm_exp2f(x * 1.44269502162933349609375f); }

// x^y for x > 0 (x == 0 gives 0 for y > 0).
DM_HD float dm_powf(float x, float y) {
  if (x == 0.0f) return y > 0.0f ? 0.0f : dm_from_bits(0x7F800000u);
  return dm_exp2f(y * (dm_logf(x) * 1.44269502162933349609375f));
}
