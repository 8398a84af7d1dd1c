// Elementary functions and the counter-based generator of the emission path.
//
// Every discrete decision of the emission samplers (accept/reject, escape
// test, Poisson count) must come out identical on the GPU and on the CPU
// restatement used to check it, so the library does not use the vendor math
// library there.  The algorithms, stated from the mathematics:
//   log(x): x = 2^k m with m in [sqrt(1/2), sqrt(2)); f = m-1, s = f/(2+f),
//           log(m) = 2 atanh(s) = f - s (f - R(s^2)), R(z) = sum_{k>=1} 2 z^k/(2k+1)
//           truncated after z^11; log(x) = k ln2_hi + (log(m) + k ln2_lo).
//   exp(x): k = round(x/ln2), r = (x - k ln2_hi) - k ln2_lo, |r| <= ln2/2,
//           Taylor series of exp(r) to r^13, scaled by 2^k in two exact steps.
//   Philox4x64-10: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
//           easy as 1, 2, 3", SC'11 (round multipliers and Weyl key
//           increments as published there).
// All code here is compiled with floating-point contraction disabled.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define RD_HD __host__ __device__ __forceinline__

namespace rd {

RD_HD uint64_t dbits(double x) { return (uint64_t)__builtin_bit_cast(unsigned long long, x); }
RD_HD double bitsd(uint64_t u) { return __builtin_bit_cast(double, (unsigned long long)u); }

static constexpr double kLn2Hi = 0x1.62e42p-1;
static constexpr double kLn2Lo = 0x1.fdf473de6af28p-22;
static constexpr double kInvLn2 = 0x1.71547652b82fep+0;
static constexpr double kSqrtHalf = 0.70710678118654752440;

RD_HD double rd_log(double x) {
  if (x != x) return x;
  if (x < 0.0) return __builtin_nan("");
  if (x == 0.0) return -__builtin_inf();
  if (x == __builtin_inf()) return x;
  int k = 0;
  uint64_t u = dbits(x);
  if ((u >> 52) == 0) {
    x = x * 0x1p54;
    k -= 54;
    u = dbits(x);
  }
  k += (int)((u >> 52) & 0x7ff) - 1023;
  double mm = bitsd((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
  if (mm > 2.0 * kSqrtHalf) {
    mm = mm * 0.5;
    k += 1;
  }
  double f = mm - 1.0;
  double s = f / (2.0 + f);
  double z = s * s;
  double R = 2.0 / 23.0;
  R = 2.0 / 21.0 + z * R;
  R = 2.0 / 19.0 + z * R;
  R = 2.0 / 17.0 + z * R;
  R = 2.0 / 15.0 + z * R;
  R = 2.0 / 13.0 + z * R;
  R = 2.0 / 11.0 + z * R;
  R = 2.0 / 9.0 + z * R;
  R = 2.0 / 7.0 + z * R;
  R = 2.0 / 5.0 + z * R;
  R = 2.0 / 3.0 + z * R;
  R = z * R;
  double lm = f - s * (f - R);
  double dk = (double)k;
  return dk * kLn2Hi + (lm + dk * kLn2Lo);
}

RD_HD double rd_exp(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return __builtin_inf();
  if (x < -745.1332191019412) return 0.0;
  double kd = __builtin_floor(x * kInvLn2 + 0.5);
  int k = (int)kd;
  double r = (x - kd * kLn2Hi) - kd * kLn2Lo;
  double p = 1.0 / 6227020800.0;
  p = 1.0 / 479001600.0 + r * p;
  p = 1.0 / 39916800.0 + r * p;
  p = 1.0 / 3628800.0 + r * p;
  p = 1.0 / 362880.0 + r * p;
  p = 1.0 / 40320.0 + r * p;
  p = 1.0 / 5040.0 + r * p;
  p = 1.0 / 720.0 + r * p;
  p = 1.0 / 120.0 + r * p;
  p = 1.0 / 24.0 + r * p;
  p = 1.0 / 6.0 + r * p;
  p = 0.5 + r * p;
  p = 1.0 + r * p;
  p = 1.0 + r * p;
  int k1 = k / 2, k2 = k - k / 2;
  double s1 = bitsd((uint64_t)(int64_t)(k1 + 1023) << 52);
  double s2 = bitsd((uint64_t)(int64_t)(k2 + 1023) << 52);
  return (p * s1) * s2;
}

struct u64x4 {
  uint64_t v[4];
};

RD_HD void mulhilo64(uint64_t a, uint64_t b, uint64_t& hi, uint64_t& lo) {
#if defined(__HIP_DEVICE_COMPILE__)
  hi = __umul64hi(a, b);
  lo = a * b;
#else
  unsigned __int128 p = (unsigned __int128)a * (unsigned __int128)b;
  hi = (uint64_t)(p >> 64);
  lo = (uint64_t)p;
#endif
}

RD_HD u64x4 philox4x64(uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3, uint64_t k0,
                       uint64_t k1) {
  const uint64_t M0 = 0xD2E7470EE14C6C93ULL, M1 = 0xCA5A826395121157ULL;
  const uint64_t W0 = 0x9E3779B97F4A7C15ULL, W1 = 0xBB67AE8584CAA73BULL;
  uint64_t x0 = c0, x1 = c1, x2 = c2, x3 = c3;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += W0;
      k1 += W1;
    }
    uint64_t hi0, lo0, hi1, lo1;
    mulhilo64(M0, x0, hi0, lo0);
    mulhilo64(M1, x2, hi1, lo1);
    uint64_t y0 = hi1 ^ x1 ^ k0;
    uint64_t y2 = hi0 ^ x3 ^ k1;
    x0 = y0;
    x1 = lo1;
    x2 = y2;
    x3 = lo0;
  }
  u64x4 o;
  o.v[0] = x0;
  o.v[1] = x1;
  o.v[2] = x2;
  o.v[3] = x3;
  return o;
}

// Uniform double in (0,1) from 64 random bits: ((u >> 11) + 0.5) * 2^-53.
RD_HD double u01(uint64_t u) { return ((double)(u >> 11) + 0.5) * 0x1p-53; }

}  // namespace rd
