// ojmi_device.h -- device primitives shared by the annealing kernels (gfx950).
//
//  * Philox4x32-10 counter-based uniforms (Salmon et al., SC'11).  The word
//    used by spin i of replica r, Trotter slice m, global sweep t is
//      Philox(key = (seed_lo, seed_hi), ctr = (i >> 2, t, r, m))[i & 3].
//  * A deterministic log on (0,1) built only from IEEE +,-,*,/ and fma, so a
//    CPU replay computes the same bits.
//  * The Metropolis threshold: spin i flips iff dE_i < theta_i with
//    theta = log(U) * (-1/beta), U = (u + 1/2) 2^-32 (fp64) or
//    U = (u>>9 + 1/2) 2^-23 (fp32).  dE <= 0 always flips (theta > 0), and the
//    flip probability of dE > 0 is exp(-beta dE): the reference's rule
//    `dE <= 0 || exp(-beta dE) > U` (include/openjij/updater/single_spin_flip.hpp:80-82)
//    written as a threshold so that it can be drawn before dE is known.
//  * Wave64 helpers: ballot, quad transpose through DPP.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ojmi {

__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t& o0,
                                              uint32_t& o1, uint32_t& o2, uint32_t& o3) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
    const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  o0 = c0;
  o1 = c1;
  o2 = c2;
  o3 = c3;
}

__device__ __forceinline__ double det_log_d(double x) {
  uint64_t b = (uint64_t)__double_as_longlong(x);
  int e = (int)((b >> 52) & 0x7ff) - 1023;
  b = (b & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
  double m = __longlong_as_double((long long)b);
  if (m > 0x1.6a09e667f3bcdp+0) {
    m = m * 0.5;
    e += 1;
  }
  const double f = m - 1.0;
  const double t = f / (2.0 + f);
  const double t2 = t * t;
  double p = 0x1.642c8590b2164p-5;
  p = __builtin_fma(t2, p, 0x1.8618618618618p-5);
  p = __builtin_fma(t2, p, 0x1.af286bca1af28p-5);
  p = __builtin_fma(t2, p, 0x1.e1e1e1e1e1e1ep-5);
  p = __builtin_fma(t2, p, 0x1.1111111111111p-4);
  p = __builtin_fma(t2, p, 0x1.3b13b13b13b14p-4);
  p = __builtin_fma(t2, p, 0x1.745d1745d1746p-4);
  p = __builtin_fma(t2, p, 0x1.c71c71c71c71cp-4);
  p = __builtin_fma(t2, p, 0x1.2492492492492p-3);
  p = __builtin_fma(t2, p, 0x1.999999999999ap-3);
  p = __builtin_fma(t2, p, 0x1.5555555555555p-2);
  const double s = t2 * p;
  const double tt = t + t;
  const double lm = __builtin_fma(tt, s, tt);
  return __builtin_fma((double)e, 0x1.62e42fefa39efp-1, lm);
}

__device__ __forceinline__ float det_log_f(float x) {
  uint32_t b = __float_as_uint(x);
  int e = (int)((b >> 23) & 0xff) - 127;
  b = (b & 0x007fffffu) | 0x3f800000u;
  float m = __uint_as_float(b);
  if (m > 0x1.6a09e6p+0f) {
    m = m * 0.5f;
    e += 1;
  }
  const float f = m - 1.0f;
  const float t = f / (2.0f + f);
  const float t2 = t * t;
  float p = 0x1.745d18p-4f;
  p = __builtin_fmaf(t2, p, 0x1.c71c72p-4f);
  p = __builtin_fmaf(t2, p, 0x1.24924ap-3f);
  p = __builtin_fmaf(t2, p, 0x1.99999ap-3f);
  p = __builtin_fmaf(t2, p, 0x1.555556p-2f);
  const float s = t2 * p;
  const float tt = t + t;
  const float lm = __builtin_fmaf(tt, s, tt);
  return __builtin_fmaf((float)e, 0x1.62e430p-1f, lm);
}

__device__ __forceinline__ double uniform_d(uint32_t u) { return ((double)u + 0.5) * 0x1p-32; }
__device__ __forceinline__ float uniform_f(uint32_t u) {
  return ((float)(u >> 9) + 0.5f) * 0x1p-23f;
}

// log(U): the negative of the Trotter threshold, or theta * beta.
template <typename TT>
__device__ __forceinline__ TT log_uniform(uint32_t u);
template <>
__device__ __forceinline__ double log_uniform<double>(uint32_t u) {
  return det_log_d(uniform_d(u));
}
template <>
__device__ __forceinline__ float log_uniform<float>(uint32_t u) {
  return det_log_f(uniform_f(u));
}

// log of the acceptance threshold's random factor for rule 0 (Metropolis:
// flip iff dE < log(U) * (-1/beta), single_spin_flip.hpp:80-82) or rule 1
// (heat bath: flip iff 1 / (1 + exp(beta dE)) > U, single_spin_flip.hpp:365-371,
// i.e. dE < log(U / (1 - U)) * (-1/beta)).  1 - U is exactly the uniform of
// ~u in both precisions (U = (u + 1/2) 2^-32, or ((u >> 9) + 1/2) 2^-23), so
// log(1 - U) is the same deterministic logarithm of another word.
// A per-sweep parameter read from memory by every lane, moved to scalar
// registers: the load is waited for once, where it is read, instead of at
// every later use inside the sweep (where the wait would also drain
// prefetches in flight).  All lanes hold the same value, so this is exact.
__device__ __forceinline__ float sweep_scalar(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
__device__ __forceinline__ double sweep_scalar(double x) {
  const long long b = __double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// The value lane `i` (wave-uniform) holds, in scalar registers.
__device__ __forceinline__ int32_t lane_value(int32_t x, int i) {
  return __builtin_amdgcn_readlane(x, i);
}
__device__ __forceinline__ float lane_value(float x, int i) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), i));
}
__device__ __forceinline__ double lane_value(double x, int i) {
  const long long b = __double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, i);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), i);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

constexpr int RULE_METROPOLIS = 0;
constexpr int RULE_HEAT_BATH = 1;
template <typename TT>
__device__ __forceinline__ TT accept_log(uint32_t u, int rule) {
  TT x = log_uniform<TT>(u);
  if (rule == RULE_HEAT_BATH) x = x - log_uniform<TT>(~u);
  return x;
}

// Integer fields: the flip test dE < theta (theta = accept_log(u) * nb in
// double) equals dE < ceil(theta) for an integer dE, so a sweep can hold
// T = ceil(theta), clamped to int32 (dE = +-2F is even, never INT32_MAX), as
// an int32.  T comes from a hardware base-2 logarithm in single precision
// (v_log_f32, within 1 ulp; the single-precision uniform is within 2^-23 of
// U) and single-precision products whenever the bracket th -+ eps,
// eps = |c| (m 2^-16 + 2^-18) with c = ln2 * nb and m the sum of the
// logarithms' magnitudes, holds one value of ceil below 2^24: eps is at least
// four times the error of th and far above its rounding, so theta lies
// strictly inside the same unit interval.  Otherwise (about 1 word in 10^4 at
// |theta| ~ 10, and every |theta| >= 2^24) the double logarithm decides, so T
// is ceil(theta) exactly in every case.
__device__ __forceinline__ int32_t clamp_ceil_i32(double x) {
  const double c = __builtin_ceil(x);
  return c >= 2147483647.0 ? INT32_MAX : (c <= -2147483648.0 ? INT32_MIN : (int32_t)c);
}
__device__ __forceinline__ int32_t int_threshold(uint32_t u, int rule, double nb) {
  const float c = (float)(0x1.62e42fefa39efp-1 * nb);
  const float l2a = __builtin_amdgcn_logf(__builtin_fmaf((float)u, 0x1p-32f, 0x1p-33f));
  float x = l2a, mag = __builtin_fabsf(l2a);
  if (rule == RULE_HEAT_BATH) {
    const float l2b = __builtin_amdgcn_logf(__builtin_fmaf((float)~u, 0x1p-32f, 0x1p-33f));
    x -= l2b;
    mag += __builtin_fabsf(l2b);
  }
  const float th = x * c;
  const float eps = __builtin_fabsf(c) * __builtin_fmaf(mag, 0x1p-16f, 0x1p-18f);
  const float lo = __builtin_ceilf(th - eps), hi = __builtin_ceilf(th + eps);
  if (lo == hi && __builtin_fabsf(lo) <= 16777216.0f) return (int32_t)lo;
  return clamp_ceil_i32(accept_log<double>(u, rule) * nb);
}

// ---------------------------------------------------------------------------
// wave64 helpers
__device__ __forceinline__ uint64_t ballot64(bool p) { return __ballot(p); }

__device__ __forceinline__ uint32_t dpp_xor1(uint32_t v) {
  // quad_perm [1,0,3,2]
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);
}
__device__ __forceinline__ uint32_t dpp_xor2(uint32_t v) {
  // quad_perm [2,3,0,1]
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);
}

// In each quad of lanes, lane q holds words A[q][0..3]; afterwards it holds
// B[q][j] = A[j][q] in w[j].
__device__ __forceinline__ void quad_transpose(uint32_t& w0, uint32_t& w1, uint32_t& w2,
                                               uint32_t& w3, int lane) {
  const bool b1 = (lane & 2) != 0;
  {
    const uint32_t s0 = b1 ? w0 : w2;
    const uint32_t s1 = b1 ? w1 : w3;
    const uint32_t r0 = dpp_xor2(s0);
    const uint32_t r1 = dpp_xor2(s1);
    if (b1) {
      w0 = r0;
      w1 = r1;
    } else {
      w2 = r0;
      w3 = r1;
    }
  }
  const bool b0 = (lane & 1) != 0;
  {
    const uint32_t s0 = b0 ? w0 : w1;
    const uint32_t s1 = b0 ? w2 : w3;
    const uint32_t r0 = dpp_xor1(s0);
    const uint32_t r1 = dpp_xor1(s1);
    if (b0) {
      w0 = r0;
      w2 = r1;
    } else {
      w1 = r0;
      w3 = r1;
    }
  }
}

// Lanes strictly above `i0` (i0 in 0..63).
__device__ __forceinline__ uint64_t lanes_above(int i0) {
  return i0 >= 63 ? 0ull : (~0ull << (i0 + 1));
}

template <typename T>
__device__ __forceinline__ T neg_inf();
template <>
__device__ __forceinline__ double neg_inf<double>() {
  return -__builtin_huge_val();
}
template <>
__device__ __forceinline__ float neg_inf<float>() {
  return -__builtin_huge_valf();
}

}  // namespace ojmi
