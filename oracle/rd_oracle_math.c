/*
 * ORACLE -- test infrastructure only (see rd_oracle.h).
 *
 * Elementary functions and the counter-based generator used by the emission
 * restatement.  The reference uses the compiler's libm and the Fortran
 * RANDOM_NUMBER intrinsic (src/main.F90:747-793); neither is reproducible
 * across compilers or devices, so this build defines its own, and the product
 * implements the same algorithms so that every discrete emission decision is
 * bit-identical between the HIP path and this restatement.
 *
 *  log(x): x = 2^k m, m in [sqrt(1/2), sqrt(2)); with f = m-1 and s = f/(2+f),
 *          log(m) = 2 atanh(s) = f - s (f - R(s^2)), R(z) = sum_{k>=1} 2 z^k/(2k+1)
 *          (the atanh Taylor series), truncated after z^11; result
 *          k*ln2_hi + (log(m) + k*ln2_lo).
 *  exp(x): k = round(x/ln2), r = (x - k ln2_hi) - k ln2_lo, |r| <= ln2/2,
 *          exp(r) by its Taylor series to r^13 (Horner), scaled by 2^k.
 *  Philox4x64-10: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 *          easy as 1, 2, 3", SC'11 -- round multipliers and Weyl key
 *          increments as published there.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "rd_oracle.h"

static const double LN2_HI = 0x1.62e42p-1;            /* ln2 with 20 significant bits */
static const double LN2_LO = 0x1.fdf473de6af28p-22;   /* ln2 - LN2_HI */
static const double INV_LN2 = 0x1.71547652b82fep+0;   /* 1/ln2 */
static const double SQRT_HALF = 0.70710678118654752440;

static inline uint64_t dbits(double x) {
  uint64_t u;
  memcpy(&u, &x, 8);
  return u;
}
static inline double bitsd(uint64_t u) {
  double x;
  memcpy(&x, &u, 8);
  return x;
}

double rdo_log(double x) {
  if (x != x) return x;
  if (x < 0.0) return NAN;
  if (x == 0.0) return -INFINITY;
  if (x == INFINITY) return x;
  int k = 0;
  uint64_t u = dbits(x);
  if ((u >> 52) == 0) { /* subnormal */
    x = x * 0x1p54;
    k -= 54;
    u = dbits(x);
  }
  k += (int)((u >> 52) & 0x7ff) - 1023;
  double mm = bitsd((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL); /* [1,2) */
  if (mm > 2.0 * SQRT_HALF) {
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
  return dk * LN2_HI + (lm + dk * LN2_LO);
}

double rdo_exp(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return INFINITY;
  if (x < -745.1332191019412) return 0.0;
  double kd = floor(x * INV_LN2 + 0.5);
  int k = (int)kd;
  double r = (x - kd * LN2_HI) - kd * LN2_LO;
  double p = 1.0 / 6227020800.0; /* 1/13! */
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
  /* scale by 2^k in two exact steps so that k outside the normal range works */
  int k1 = k / 2, k2 = k - k / 2;
  double s1 = bitsd((uint64_t)(k1 + 1023) << 52);
  double s2 = bitsd((uint64_t)(k2 + 1023) << 52);
  return (p * s1) * s2;
}

/* ------------------------------------------------------------------------ */
static inline void mulhilo64(uint64_t a, uint64_t b, uint64_t* hi, uint64_t* lo) {
  unsigned __int128 p = (unsigned __int128)a * (unsigned __int128)b;
  *hi = (uint64_t)(p >> 64);
  *lo = (uint64_t)p;
}

void rdo_philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]) {
  const uint64_t M0 = 0xD2E7470EE14C6C93ULL, M1 = 0xCA5A826395121157ULL;
  const uint64_t W0 = 0x9E3779B97F4A7C15ULL, W1 = 0xBB67AE8584CAA73BULL;
  uint64_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
  uint64_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += W0;
      k1 += W1;
    }
    uint64_t hi0, lo0, hi1, lo1;
    mulhilo64(M0, x0, &hi0, &lo0);
    mulhilo64(M1, x2, &hi1, &lo1);
    uint64_t y0 = hi1 ^ x1 ^ k0;
    uint64_t y1 = lo1;
    uint64_t y2 = hi0 ^ x3 ^ k1;
    uint64_t y3 = lo0;
    x0 = y0;
    x1 = y1;
    x2 = y2;
    x3 = y3;
  }
  out[0] = x0;
  out[1] = x1;
  out[2] = x2;
  out[3] = x3;
}
