// Pair and field arithmetic shared by the pair kernels and the emission
// engine: the FP64 forms of the reference expressions (planar image series,
// hyperboloid tip vacuum field and sphere image) and the mixed-precision
// packed records and image weights.
#pragma once

#include <hip/hip_runtime.h>

#include "rd_internal.h"

namespace rd {

// FP64 device forms of the reference expressions.
__device__ __forceinline__ void ic_planar64(double z_a, double z_b, double dx, double dy,
                                            double dxy2, int nic, double dd, double es,
                                            double& icx, double& icy, double& icz) {
  double z_ic = -1.0 * z_b;
  double diff_z = z_a - z_ic;
  double r = __builtin_sqrt(dxy2 + diff_z * diff_z) + es;
  double inv_r3 = 1.0 / (r * r * r);
  double ix = -dx * inv_r3, iy = -dy * inv_r3, iz = -diff_z * inv_r3;
  for (int n = 1; n <= nic; ++n) {
    double tnd = 2.0 * n * dd;
    z_ic = tnd - z_b;
    diff_z = z_a - z_ic;
    r = __builtin_sqrt(dxy2 + diff_z * diff_z) + es;
    inv_r3 = 1.0 / (r * r * r);
    ix = ix - dx * inv_r3;
    iy = iy - dy * inv_r3;
    iz = iz - diff_z * inv_r3;
    z_ic = -2.0 * n * dd - z_b;
    diff_z = z_a - z_ic;
    r = __builtin_sqrt(dxy2 + diff_z * diff_z) + es;
    inv_r3 = 1.0 / (r * r * r);
    ix = ix - dx * inv_r3;
    iy = iy - dy * inv_r3;
    iz = iz - diff_z * inv_r3;
    z_ic = tnd + z_b;
    diff_z = z_a - z_ic;
    r = __builtin_sqrt(dxy2 + diff_z * diff_z) + es;
    inv_r3 = 1.0 / (r * r * r);
    ix = ix + dx * inv_r3;
    iy = iy + dy * inv_r3;
    iz = iz + diff_z * inv_r3;
    z_ic = -2.0 * n * dd + z_b;
    diff_z = z_a - z_ic;
    r = __builtin_sqrt(dxy2 + diff_z * diff_z) + es;
    inv_r3 = 1.0 / (r * r * r);
    ix = ix + dx * inv_r3;
    iy = iy + dy * inv_r3;
    iz = iz + diff_z * inv_r3;
  }
  icx = ix;
  icy = iy;
  icz = iz;
}

// Tip vacuum field (acc_tip_field_E.inc:13-34); cos/sin of atan2(y,x) are
// taken as x/rho, y/rho.
__device__ __host__ __forceinline__ void tip_vacuum64(const DevGeom& g, double x, double y,
                                                      double z, double& fx, double& fy,
                                                      double& fz) {
  double af = g.a_foci, sz = g.shift_z;
  double r_p = __builtin_sqrt(x * x + y * y + (z + af - sz) * (z + af - sz));
  double r_m = __builtin_sqrt(x * x + y * y + (z - af - sz) * (z - af - sz));
  double xi = (r_p + r_m) / (2.0 * af);
  double eta = (r_p - r_m) / (2.0 * af);
  double c, s;
  if ((__builtin_fabs(x) < 1.0e-18) && (__builtin_fabs(y) < 1.0e-18)) {
    c = 1.0;
    s = 0.0;
  } else {
    double rho = __builtin_sqrt(x * x + y * y);
    c = x / rho;
    s = y / rho;
  }
  double pre = g.pre_E * 1.0 / (xi * xi - eta * eta);
  double fac_xy;
  if (__builtin_fabs(xi - 1.0) < 1.0e-6)
    fac_xy = 0.0;
  else
    fac_xy = eta * __builtin_sqrt((xi * xi - 1.0) / (1.0 - eta * eta));
  fx = -1.0 * pre * fac_xy * c;
  fy = -1.0 * pre * fac_xy * s;
  fz = pre * xi;
}

// Vacuum field of either geometry.
__device__ __forceinline__ void vacuum64(const DevGeom& g, double x, double y, double z,
                                         double& fx, double& fy, double& fz) {
  if (g.kind == RD_GEOM_TIP) {
    tip_vacuum64(g, x, y, z, fx, fy, fz);
  } else {
    fx = 0.0;
    fy = 0.0;
    fz = g.Ez;
  }
}

// Sphere image point of the particle at a (acc_tip_image_point.inc:14-18).
__device__ __forceinline__ void tip_image_point64(double x_a, double y_a, double z_a, double z0,
                                                  double R, double& dis_a, double& x_im,
                                                  double& y_im, double& z_im) {
  double dz = z_a - z0;
  double da = __builtin_sqrt(x_a * x_a + y_a * y_a + dz * dz);
  double zi =
      z0 + R * R / (__builtin_sqrt(1 + x_a * x_a / (dz * dz) + y_a * y_a / (dz * dz)) * da);
  z_im = zi;
  x_im = (zi - z0) * x_a / dz;
  y_im = (zi - z0) * y_a / dz;
  dis_a = da;
}

// Sphere image force (acc_tip_ic_force.inc:17-22), t**1.5 = t*sqrt(t).
__device__ __forceinline__ void tip_ic_force64(double x_a, double y_a, double z_a, double x_b,
                                               double y_b, double z_b, double dis_a,
                                               double x_im, double y_im, double z_im, double R,
                                               double kt, double& icx, double& icy,
                                               double& icz) {
  double ta = (x_b - x_a) * (x_b - x_a) + (y_b - y_a) * (y_b - y_a) + (z_b - z_a) * (z_b - z_a);
  double tb = (x_b - x_im) * (x_b - x_im) + (y_b - y_im) * (y_b - y_im) +
              (z_b - z_im) * (z_b - z_im);
  double tda = ta * __builtin_sqrt(ta);
  double tdb = tb * __builtin_sqrt(tb);
  icx = kt * ((x_a - x_b) / tda - (R * (x_im - x_b)) / (dis_a * tdb));
  icy = kt * ((y_a - y_b) / tda - (R * (y_im - y_b)) / (dis_a * tdb));
  icz = kt * ((z_a - z_b) / tda - (R * (z_im - z_b)) / (dis_a * tdb));
}

// ------------------------------------------------------------------------
// Mixed precision.
//
// Packed record (nm): hi = (x_hi, y_hi, z_hi, q/q_0), lo = (x_lo, y_lo, z_lo,
// d - z) with x = x_hi + x_lo to ~2^-48 relative.
__device__ __forceinline__ void pack1(double xm, double ym, double zm, double qv, double d,
                                      float4& hi, float4& lo) {
  const double s = 1.0e9;
  double xn = xm * s, yn = ym * s, zn = zm * s;
  float xh = (float)xn, yh = (float)yn, zh = (float)zn;
  hi = make_float4(xh, yh, zh, (float)(qv / Consts::q_0));
  lo = make_float4((float)(xn - (double)xh), (float)(yn - (double)yh), (float)(zn - (double)zh),
                   (float)(d * s - zn));
}

__device__ __forceinline__ float rsq(float v) { return __builtin_amdgcn_rsqf(v); }

// Guard for squared distances that can be exactly zero (a pair with itself
// or coincident particles, a particle exactly on an electrode): keeps every
// FP32 term finite; the reference's r + 1e-18 m softening gives the same 0
// contribution for zero offsets.
constexpr float kR2Min = 1.0e-20f;

// One image partner of a series at dz = DZ: add 1/r^3 to the suffix sum S.
#define RD_PARTNER_SUM(DZ, S)                   \
  {                                             \
    float _dz = (DZ);                           \
    float _r2 = __builtin_fmaf(_dz, _dz, dxy2); \
    float _ri = rsq(_r2);                       \
    S = __builtin_fmaf(_ri * _ri, _ri, S);      \
  }

// Image-charge weights of one pair: returns (sum_same - sum_opp) of 1/r^3 in
// wxy and of dz/r^3 in wz.  zsum = z_a + z_b, wsum = (d - z_a) + (d - z_b),
// dzab = z_a - z_b.
//
// The partners n >= 1 of a series lie at dz = a + t_n and a - t_n, t_n = 2nd
// (a = zsum for the opposite charges, dzab for the same charges), so
//   sum dz/r^3 = a sum 1/r^3 + 2d sum_n n (1/r+^3 - 1/r-^3),
// and the n-weighted sum is the sum over k of the suffix sums
// sum_{n >= k} (1/r+^3 - 1/r-^3), built with n descending.  A partner then
// costs one fma into its suffix sum in place of forming dz/r^3 and adding it
// twice.  Every partner in that form has |dz| >= t_n / 2, so a r^-3 and
// t_n r^-3 do not cancel.  The opposite partner across the anode, at
// zsum - 2d = -wsum, keeps its own term: it is close to the target for
// particles near the anode.
// NEAR: the caller guarantees nic <= 1 (no partners n >= 2), so the
// suffix-sum loop is compiled out.
template <bool NEAR = false>
__device__ __forceinline__ void ic_planar32(float zsum, float wsum, float dzab, float dxy2,
                                            int nic, float two_d, float& wxy, float& wz) {
  float wo, wzo;
  {
    float r2 = fmaxf(__builtin_fmaf(zsum, zsum, dxy2), kR2Min);
    float ri = rsq(r2);
    wo = ri * ri * ri;
    wzo = zsum * wo;
  }
  if (nic < 1) {
    wxy = -wo;
    wz = -wzo;
    return;
  }
  {
    float dz1 = -wsum;
    float r2 = fmaxf(__builtin_fmaf(dz1, dz1, dxy2), kR2Min);
    float ri = rsq(r2);
    float r3 = ri * ri * ri;
    wo += r3;
    wzo = __builtin_fmaf(dz1, r3, wzo);
  }
  // suffix sums of 1/r^3: p (a + t_n) and m (a - t_n), o(pposite) and s(ame)
  float po = 0.f, mo = 0.f, ps = 0.f, ms = 0.f, ro = 0.f, rs = 0.f;
  if (NEAR) {
    // n = 1 alone: the sums start at its terms (the same values as adding
    // them to zero, without the additions: every term is positive)
    auto r3 = [dxy2](float dz) {
      const float ri = rsq(__builtin_fmaf(dz, dz, dxy2));
      return ri * ri * ri;
    };
    po = r3(zsum + two_d);
    ps = r3(dzab + two_d);
    ms = r3(dzab - two_d);
    ro = po;
    rs = ps - ms;
  } else {
    float tn = (float)nic * two_d;
    for (int n = nic; n >= 2; --n) {
      RD_PARTNER_SUM(zsum + tn, po);
      RD_PARTNER_SUM(zsum - tn, mo);
      RD_PARTNER_SUM(dzab + tn, ps);
      RD_PARTNER_SUM(dzab - tn, ms);
      ro = (ro + po) - mo;
      rs = (rs + ps) - ms;
      tn -= two_d;
    }
    RD_PARTNER_SUM(zsum + two_d, po);
    RD_PARTNER_SUM(dzab + two_d, ps);
    RD_PARTNER_SUM(dzab - two_d, ms);
    ro = (ro + po) - mo;
    rs = (rs + ps) - ms;
  }
  const float so = NEAR ? po : po + mo, ss = ps + ms;
  wxy = ss - (wo + so);
  wz = __builtin_fmaf(dzab, ss, two_d * rs) - (__builtin_fmaf(zsum, so, wzo) + two_d * ro);
}

// Direct plus image-charge weight of one pair whose target lies on the
// cathode plane z = 0 (the emission surface: surface-field grids and
// Metropolis-Hastings sampling points).  There every partner has a twin at
// the same distance -- the n = 0 image with the source itself, the opposite
// partner at 2nd - z_b with the same-charge partner at z_b - 2nd, the one at
// -2nd - z_b with the one at 2nd + z_b -- whose lateral terms cancel and whose
// normal terms are equal, so the pair adds
//   cz = -2 sum_{k=-nic..nic} u_k / (dxy2 + u_k^2)^(3/2),  u_k = z_b + 2kd,
//   cxy = 0,
// the same series as ic_planar32 plus the direct term, with 2 nic + 1 distances
// instead of 4 nic + 2.  surface_sum32 returns the sum without the factor -2,
// which the callers apply once to their FP64 total (exactly: a power of two).
// zb = z_b, wb = d - z_b and dnm = d (nm); the partners below the cathode are
// formed from wb, as ic_planar32 forms the anode image.
// surface_far32: w plus the partners k != 0 of that series, one by one.
__device__ __forceinline__ float surface_far32(float zb, float wb, float dxy2, int nic, float dnm,
                                               float w) {
  const float two_d = dnm + dnm;
  float up = zb, um = -(wb + dnm);
  for (int n = 1; n <= nic; ++n) {
    up += two_d;
    float r2 = __builtin_fmaf(up, up, dxy2);
    float ri = rsq(r2);
    w = __builtin_fmaf(up, ri * ri * ri, w);
    r2 = __builtin_fmaf(um, um, dxy2);
    ri = rsq(r2);
    w = __builtin_fmaf(um, ri * ri * ri, w);
    um -= two_d;
  }
  return w;
}

__device__ __forceinline__ float surface_sum32(float zb, float wb, float dxy2, int nic,
                                               float dnm) {
  float r2 = fmaxf(__builtin_fmaf(zb, zb, dxy2), kR2Min);
  float ri = rsq(r2);
  return surface_far32(zb, wb, dxy2, nic, dnm, zb * (ri * ri * ri));
}

// The partners k != 0 of the series above as a power series in s = dxy2.
// They lie at |u_k| >= d, so
//   far(s) = sum_{k != 0} u_k / (s + u_k^2)^(3/2) = sum_m c_m s^m,
//   c_m = binom(-3/2, m) sum_{k != 0} sgn(u_k) (u_k^2)^-(m+1),
// and each partner's series in s / u_k^2 <= s / d^2 alternates with
// decreasing terms.  With kFarTerms terms and s <= kFarS d^2 the truncation
// stays below |binom(-3/2, 8)| 0.1^8 = 3.4e-8 of each partner's term.  The
// series is evaluated in t = s / d^2 <= kFarS with the coefficients
// a_m = c_m d^(2m): FP64 sums over the partners, scaled in FP64 and rounded
// to FP32 once (surface_far_coeffs).  c_m itself falls as d^-(2m+2) and
// would leave FP32's range from d ~ 1 um; a_m ~ d^-2 stays in it for every
// gap.  Pairs beyond kFarS d^2 take surface_far32.
constexpr int kFarTerms = 8;
constexpr float kFarS = 0.1f;

// a_0 .. a_7 of one source: zb = z_b, wb = d - z_b, dnm = d (nm).
__device__ __forceinline__ void surface_far_coeffs(double zb, double wb, double dnm, int nic,
                                                   float4& c0, float4& c1) {
  double S[kFarTerms];
  for (int m = 0; m < kFarTerms; ++m) S[m] = 0.0;
  for (int k = 1; k <= nic; ++k) {
    const double up = zb + 2.0 * k * dnm;          // u_k
    const double um = wb + (2.0 * k - 1.0) * dnm;  // |u_-k| = 2kd - z_b
    const double ia = 1.0 / (up * up), ib = 1.0 / (um * um);
    double pa = ia, pb = ib;
    for (int m = 0; m < kFarTerms; ++m) {
      S[m] += pa - pb;
      pa *= ia;
      pb *= ib;
    }
  }
  float c[kFarTerms];
  double b = 1.0;  // binom(-3/2, m)
  double d2m = 1.0;  // d^(2m)
  for (int m = 0; m < kFarTerms; ++m) {
    c[m] = (float)(b * S[m] * d2m);
    b *= (-1.5 - m) / (m + 1.0);
    d2m *= dnm * dnm;
  }
  c0 = make_float4(c[0], c[1], c[2], c[3]);
  c1 = make_float4(c[4], c[5], c[6], c[7]);
}

// sum_m a_m t^m, t = s / d^2.
__device__ __forceinline__ float surface_far_eval(float t, float4 c0, float4 c1) {
  float v = __builtin_fmaf(c1.w, t, c1.z);
  v = __builtin_fmaf(v, t, c1.y);
  v = __builtin_fmaf(v, t, c1.x);
  v = __builtin_fmaf(v, t, c0.w);
  v = __builtin_fmaf(v, t, c0.z);
  v = __builtin_fmaf(v, t, c0.y);
  return __builtin_fmaf(v, t, c0.x);
}

// q_j/q_0 times surface_sum32 for the packed target (h1, l1) on the cathode
// and the packed source (h2, l2); N0: the series stops at n = 0 (nic = 0).
// The source height is its high part alone (the float sum of both parts
// rounds to it but for ties).
template <bool N0>
__device__ __forceinline__ float surface_term32(const float4 h1, const float4 l1, const float4 h2,
                                                const float4 l2, int nic, float dnm) {
  const float dx = (h1.x - h2.x) + (l1.x - l2.x);
  const float dy = (h1.y - h2.y) + (l1.y - l2.y);
  const float dxy2 = __builtin_fmaf(dy, dy, dx * dx);
  if (N0) {
    const float ri = rsq(fmaxf(__builtin_fmaf(h2.z, h2.z, dxy2), kR2Min));
    return h2.w * (h2.z * (ri * ri * ri));
  }
  return h2.w * surface_sum32(h2.z, l2.w, dxy2, nic, dnm);
}

// With N_ic_max = 0: acc + q_j/q_0 z_b / r^3 for one source, in FP32.  The
// callers add blocks of at most kSurfBlock such terms in FP32 and each block
// sum to their FP64 total.  Every term of a block has the sign of the
// charges (z_b >= 0, one species), so a block's FP32 sum is within
// (kSurfBlock - 1) u = 15 * 2^-24 (< 1e-6) of its exact value; per source it
// saves the conversion and the FP64 add.  The same holds for the paired-image
// sum with N_ic_max > 0: sum_{k=-nic..nic} u_k / (rho^2 + u_k^2)^(3/2),
// u_k = z_b + 2kd, is positive for every z_b in (0, d), rho and N_ic_max
// (the terms n and -n-1 pair off towards (2 nic + 1) z_b / rho^3 at large
// rho; checked on a grid of z_b and rho for N_ic_max up to 20), so the
// cathode field's far-series loop adds its terms in the same blocks.
constexpr int kSurfBlock = 16;
__device__ __forceinline__ float surface_acc32(const float4 h1, const float4 l1, const float4 h2,
                                               const float4 l2, float acc) {
  const float dx = (h1.x - h2.x) + (l1.x - l2.x);
  const float dy = (h1.y - h2.y) + (l1.y - l2.y);
  const float dxy2 = __builtin_fmaf(dy, dy, dx * dx);
  const float ri = rsq(fmaxf(__builtin_fmaf(h2.z, h2.z, dxy2), kR2Min));
  return __builtin_fmaf(h2.w, h2.z * (ri * ri * ri), acc);
}

// surface_acc32 for a source s2 = (x, y, z, q_j/q_0 z) of high parts
// (k_pack_particles' srf) and its low parts l2.
__device__ __forceinline__ float surface_accq32(const float4 h1, const float4 l1, const float4 s2,
                                                const float4 l2, float acc) {
  const float dx = (h1.x - s2.x) + (l1.x - l2.x);
  const float dy = (h1.y - s2.y) + (l1.y - l2.y);
  const float dxy2 = __builtin_fmaf(dy, dy, dx * dx);
  const float ri = rsq(fmaxf(__builtin_fmaf(s2.z, s2.z, dxy2), kR2Min));
  return __builtin_fmaf(s2.w, ri * ri * ri, acc);
}

// True when every active target of the wavefront lies on the plane z = 0 and
// the image series is on: the wavefront may then use surface_term32.
__device__ __forceinline__ bool wave_on_surface(int ic, bool act, float4 h1, float4 l1) {
  return ic && __all(!act || (h1.z == 0.0f && l1.z == 0.0f));
}

// ------------------------------------------------------------------------
// Field at one target per lane from the sources of chunks [c0, c1).  Every
// lane of the workgroup takes part in the LDS staging whether or not it has
// a target (`act`), so the calls sit in uniform control flow.

// s_q += P[q * S + k * ld] for q = 0, 1, 2 and k = 0 .. n-1, in k order (the
// same additions as a plain loop); the loads are issued eight chunks at a
// time, so that their latencies overlap instead of adding up.
__device__ __forceinline__ void sum3_in_order(const double* __restrict__ P, size_t S, size_t ld,
                                              int n, double& s0, double& s1, double& s2) {
  // batches of 16 partials (48 loads in flight), the last one predicated:
  // the sums are latency-bound, so the number of round trips is what counts
  constexpr int B = 16;
  for (int k = 0; k < n; k += B) {
    double a[B], b[B], c[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const bool in = k + u < n;
      const size_t o = in ? (size_t)(k + u) * ld : 0;
      a[u] = in ? P[0 * S + o] : 0.0;
      b[u] = in ? P[1 * S + o] : 0.0;
      c[u] = in ? P[2 * S + o] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
      if (k + u < n) {
        s0 = s0 + a[u];
        s1 = s1 + b[u];
        s2 = s2 + c[u];
      }
    }
  }
}

// FP64: one partial per 256-source chunk, P[k * S + c * m + p] for the
// components k = 0, 1, 2 (the canonical order: sources within a chunk in
// index order, chunk partials added in chunk order by the finishing step).
template <int KIND>
__device__ __forceinline__ void field_chunks64(const DevGeom& g, bool act, double x1, double y1,
                                               double z1, const double* __restrict__ x,
                                               const double* __restrict__ y,
                                               const double* __restrict__ z,
                                               const double* __restrict__ q, int n, int c0,
                                               int c1, double* sx, double* sy, double* sz,
                                               double* sq, double* __restrict__ P, size_t S,
                                               int m, int p, int cfirst = 0) {
  const int t = threadIdx.x;
  double da = 0, xim = 0, yim = 0, zim = 0;
  if (KIND == RD_GEOM_TIP && act && g.ic) tip_image_point64(x1, y1, z1, g.z0, g.R, da, xim, yim, zim);
  for (int c = c0; c < c1; ++c) {
    const int jb = c * kChunk;
    __syncthreads();
    if (jb + t < n) {
      sx[t] = x[jb + t];
      sy[t] = y[jb + t];
      sz[t] = z[jb + t];
      sq[t] = q[jb + t];
    }
    __syncthreads();
    const int cnt = min(kChunk, n - jb);
    double fxs = 0.0, fys = 0.0, fzs = 0.0;
    if (act) {
      for (int jj = 0; jj < cnt; ++jj) {
        double x2 = sx[jj], y2 = sy[jj], z2 = sz[jj];
        double pre = sq[jj] * g.dfc;
        double dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
        double ax, ay, az;
        if (KIND == RD_GEOM_PLANAR) {
          double dxy2 = dx * dx + dy * dy;
          double r = __builtin_sqrt(dxy2 + dz * dz) + g.es;
          double inv_r3 = 1.0 / (r * r * r);
          ax = pre * inv_r3 * dx;
          ay = pre * inv_r3 * dy;
          az = pre * inv_r3 * dz;
          if (g.ic) {
            double icx, icy, icz;
            ic_planar64(z1, z2, dx, dy, dxy2, g.nic, g.d, g.es, icx, icy, icz);
            ax = ax + pre * icx;
            ay = ay + pre * icy;
            az = az + pre * icz;
          }
        } else {
          double r = __builtin_sqrt(dx * dx + dy * dy + dz * dz) + g.es;
          double inv_r3 = 1.0 / (r * r * r);
          ax = pre * inv_r3 * dx;
          ay = pre * inv_r3 * dy;
          az = pre * inv_r3 * dz;
          if (g.ic) {
            double icx, icy, icz;
            tip_ic_force64(x1, y1, z1, x2, y2, z2, da, xim, yim, zim, g.R, g.ktip, icx, icy, icz);
            ax = ax + pre * icx;
            ay = ay + pre * icy;
            az = az + pre * icz;
          }
        }
        fxs = fxs + ax;
        fys = fys + ay;
        fzs = fzs + az;
      }
      P[0 * S + (size_t)(c - cfirst) * m + p] = fxs;
      P[1 * S + (size_t)(c - cfirst) * m + p] = fys;
      P[2 * S + (size_t)(c - cfirst) * m + p] = fzs;
    }
  }
}

// Field = E_vac + the chunk partials in chunk order.
__device__ __forceinline__ void field_finish64(const DevGeom& g, const double* __restrict__ P,
                                               size_t S, int m, int p, int nchunks, double x,
                                               double y, double z, double& fx, double& fy,
                                               double& fz) {
  double sx = 0.0, sy = 0.0, sz = 0.0;
  sum3_in_order(P + p, S, m, nchunks, sx, sy, sz);
  double ex, ey, ez;
  vacuum64(g, x, y, z, ex, ey, ez);
  fx = ex + sx;
  fy = ey + sy;
  fz = ez + sz;
}

// p[j] for a wavefront-uniform j of memory that no launch in flight writes:
// read through the constant address space, so the load goes through the
// scalar cache into SGPRs.
__device__ __forceinline__ float4 load_uniform(const float4* p, int j) {
  typedef const __attribute__((address_space(4))) float cfloat;
  const cfloat* q = (const cfloat*)(const float*)(p + j);
  return make_float4(q[0], q[1], q[2], q[3]);
}

// v[q] = the minimum of v[q] over the wavefront's 64 lanes (butterfly).
__device__ __forceinline__ void wave_min4(float v[4]) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = fminf(v[q], __shfl_xor(v[q], o));
}

// Mixed (planar): FP64 sums, in nm^-2 per unit charge, of the packed sources
// of chunks [c0, c1) at the packed target (h1, l1).
__device__ __forceinline__ void field_seg32(const DevGeom& g, bool act, float4 h1, float4 l1,
                                            const float4* __restrict__ hi,
                                            const float4* __restrict__ lo,
                                            const float4* __restrict__ far, int n, int c0,
                                            int c1, float4* shi, float4* slo, float4* sfar,
                                            double& fx, double& fy, double& fz) {
  const int ic = g.ic, nic = g.nic;
  const float two_d = (float)(2.0 * g.d * 1e9);
  const float dnm = (float)(g.d * 1e9);
  const float slim = kFarS * dnm * dnm;
  const float inv_d2 = (float)(1.0 / ((g.d * 1e9) * (g.d * 1e9)));
  const bool surf = wave_on_surface(ic, act, h1, l1);
  const int t = threadIdx.x;
  // Lateral box of the wavefront's targets and of each chunk's sources (high
  // parts): when every pair of a chunk lies within the far series' range,
  // the wavefront takes the chunk without the per-pair test (the same choice
  // the test makes for each pair, so the same result).
  const bool boxes = far && nic > 0;
  float tb[4] = {act ? h1.x : 3e38f, act ? -h1.x : 3e38f, act ? h1.y : 3e38f,
                 act ? -h1.y : 3e38f};  // min x, -max x, min y, -max y
  if (boxes) wave_min4(tb);
  __shared__ float sbox[4][kBlock / 64];
  fx = 0.0;
  fy = 0.0;
  fz = 0.0;
  for (int c = c0; c < c1; ++c) {
    const int jb = c * kChunk;
    __syncthreads();
    float sb[4] = {3e38f, 3e38f, 3e38f, 3e38f};
    if (jb + t < n) {
      const float4 hv = hi[jb + t];
      shi[t] = hv;
      slo[t] = lo[jb + t];
      if (far) {
        sfar[2 * t] = far[2 * (jb + t)];
        sfar[2 * t + 1] = far[2 * (jb + t) + 1];
      }
      sb[0] = hv.x;
      sb[1] = -hv.x;
      sb[2] = hv.y;
      sb[3] = -hv.y;
    }
    if (boxes) {
      wave_min4(sb);
      if ((t & 63) == 0)
        for (int q = 0; q < 4; ++q) sbox[q][t >> 6] = sb[q];
    }
    __syncthreads();
    const int cnt = min(kChunk, n - jb);
    bool inside = false;
    if (boxes) {
      for (int q = 0; q < 4; ++q) {
        sb[q] = sbox[q][0];
        for (int u = 1; u < (int)(blockDim.x >> 6); ++u) sb[q] = fminf(sb[q], sbox[q][u]);
      }
      // largest |dx|, |dy| over the chunk's pairs, with a margin of 1e-3 nm for
      // the low parts and the rounding of the differences
      const float ex = fmaxf(-tb[1] - sb[0], -sb[1] - tb[0]) + 1e-3f;
      const float ey = fmaxf(-tb[3] - sb[2], -sb[3] - tb[2]) + 1e-3f;
      inside = fmaxf(ex, 0.f) * fmaxf(ex, 0.f) + fmaxf(ey, 0.f) * fmaxf(ey, 0.f) <=
               0.999f * slim;
    }
    if (act && surf) {  // fz holds the sum without the factor -2 until the end
      if (nic > 0 && far && inside)
        for (int jb2 = 0; jb2 < cnt; jb2 += kSurfBlock) {
          const int je2 = min(cnt, jb2 + kSurfBlock);
          float acc = 0.f;
#pragma unroll 4
          for (int jj = jb2; jj < je2; ++jj) {
            const float4 h2 = shi[jj], l2 = slo[jj];
            const float dx = (h1.x - h2.x) + (l1.x - l2.x);
            const float dy = (h1.y - h2.y) + (l1.y - l2.y);
            const float s = __builtin_fmaf(dy, dy, dx * dx);
            const float ri = rsq(fmaxf(__builtin_fmaf(h2.z, h2.z, s), kR2Min));
            float w = h2.z * (ri * ri * ri);
            // the series coefficients through the scalar cache: the LDS
            // reads of the sources are the kernel's limit
            w += surface_far_eval(s * inv_d2, load_uniform(far, 2 * (jb + jj)),
                                  load_uniform(far, 2 * (jb + jj) + 1));
            acc = __builtin_fmaf(h2.w, w, acc);
          }
          fz += (double)acc;
        }
      else if (nic > 0 && far)
        for (int jb2 = 0; jb2 < cnt; jb2 += kSurfBlock) {
          const int je2 = min(cnt, jb2 + kSurfBlock);
          float acc = 0.f;
          for (int jj = jb2; jj < je2; ++jj) {
            const float4 h2 = shi[jj], l2 = slo[jj];
            const float dx = (h1.x - h2.x) + (l1.x - l2.x);
            const float dy = (h1.y - h2.y) + (l1.y - l2.y);
            const float s = __builtin_fmaf(dy, dy, dx * dx);
            const float ri = rsq(fmaxf(__builtin_fmaf(h2.z, h2.z, s), kR2Min));
            float w = h2.z * (ri * ri * ri);
            if (s <= slim)
              w += surface_far_eval(s * inv_d2, sfar[2 * jj], sfar[2 * jj + 1]);
            else
              w = surface_far32(h2.z, l2.w, s, nic, dnm, w);
            acc = __builtin_fmaf(h2.w, w, acc);
          }
          fz += (double)acc;
        }
      else if (nic == 0)
        for (int jb2 = 0; jb2 < cnt; jb2 += kSurfBlock) {
          const int je2 = min(cnt, jb2 + kSurfBlock);
          float acc = 0.f;
          for (int jj = jb2; jj < je2; ++jj) acc = surface_acc32(h1, l1, shi[jj], slo[jj], acc);
          fz += (double)acc;
        }
      else
        for (int jj = 0; jj < cnt; ++jj)
          fz += (double)surface_term32<false>(h1, l1, shi[jj], slo[jj], nic, dnm);
    } else if (act) {
      for (int jj = 0; jj < cnt; ++jj) {
        const float4 h2 = shi[jj], l2 = slo[jj];
        float dx = (h1.x - h2.x) + (l1.x - l2.x);
        float dy = (h1.y - h2.y) + (l1.y - l2.y);
        float dz = (h1.z - h2.z) + (l1.z - l2.z);
        float dxy2 = __builtin_fmaf(dy, dy, dx * dx);
        float r2 = fmaxf(__builtin_fmaf(dz, dz, dxy2), kR2Min);
        float ri = rsq(r2);
        float r3 = ri * ri * ri;
        float cxy = r3, cz = dz * r3;
        if (ic) {
          float wxy, wz;
          ic_planar32(h1.z + h2.z, l1.w + l2.w, dz, dxy2, nic, two_d, wxy, wz);
          cxy += wxy;
          cz += wz;
        }
        cxy *= h2.w;
        cz *= h2.w;
        fx += (double)(cxy * dx);
        fy += (double)(cxy * dy);
        fz += (double)cz;
      }
    }
  }
  if (surf) fz = -2.0 * fz;
}

// Field (V/m) of the planar geometry from the particle sums (sx, sy, sz) of a
// mixed evaluation (nm^-2 per unit charge).
__device__ __forceinline__ void field_from_sums32(const DevGeom& g, double sx, double sy,
                                                  double sz, double& fx, double& fy, double& fz) {
  double pre = g.dfc * Consts::q_0 * 1.0e18;
  fx = pre * sx;
  fy = pre * sy;
  fz = g.Ez + pre * sz;
}

// Field (V/m) from the segment sums of a mixed evaluation, in segment order.
__device__ __forceinline__ void field_finish32(const DevGeom& g, const double* __restrict__ P,
                                               size_t S, int m, int p, int segs, double& fx,
                                               double& fy, double& fz) {
  double sx = 0.0, sy = 0.0, sz = 0.0;
  sum3_in_order(P + p, S, m, segs, sx, sy, sz);
  field_from_sums32(g, sx, sy, sz, fx, fy, fz);
}

// Field (V/m) of the tip geometry from the segment sums of a mixed
// evaluation: the FP64 vacuum field at the point plus the particle sums.
__device__ __forceinline__ void field_finish32_tip(const DevGeom& g, const double* __restrict__ P,
                                                   size_t S, int m, int p, int segs, double x,
                                                   double y, double z, double& fx, double& fy,
                                                   double& fz) {
  double sx = 0.0, sy = 0.0, sz = 0.0;
  sum3_in_order(P + p, S, m, segs, sx, sy, sz);
  double ex, ey, ez;
  tip_vacuum64(g, x, y, z, ex, ey, ez);
  const double pre = g.dfc * Consts::q_0 * 1.0e18;
  fx = ex + pre * sx;
  fy = ey + pre * sy;
  fz = ez + pre * sz;
}

// ------------------------------------------------------------------------
// Mixed precision, hyperboloid tip.  The sphere image of a point (FP64,
// tip_image_point64) is packed like a position: ihi = (x_im, y_im, z_im in nm
// as float, distance of the point from the sphere centre in nm), ilo = the
// low parts of the coordinates.  Pair terms are FP32 in nm; the image term
// keeps the reference's factor k_tip (acc_tip_ic_force.inc:17-22), so a
// pair adds  q_j/q_0 * [ d/r^3 + k_tip (d/r^3 + R e/(D t^3)) ]  (nm^-2,
// scaled by 1e18 * dfc * q_i * q_0 at the end), with d = target - source,
// e = source - image, t = |e|, D the image owner's distance from the centre.
__device__ __forceinline__ void pack_image(double xm, double ym, double zm, double z0, double R,
                                           float4& ihi, float4& ilo) {
  double da, xi, yi, zi;
  tip_image_point64(xm, ym, zm, z0, R, da, xi, yi, zi);
  const double s = 1.0e9;
  double xn = xi * s, yn = yi * s, zn = zi * s;
  float xh = (float)xn, yh = (float)yn, zh = (float)zn;
  ihi = make_float4(xh, yh, zh, (float)(da * s));
  ilo = make_float4((float)(xn - (double)xh), (float)(yn - (double)yh), (float)(zn - (double)zh),
                    0.0f);
}

// Target-image pair (the field at points, and accelerations with j > i):
// the image is the target's.  Returns the FP32 terms scaled by q_j/q_0.
__device__ __forceinline__ void tip_pair32_target(const float4 h1, const float4 l1,
                                                  const float4 ih1, const float4 il1,
                                                  const float4 h2, const float4 l2, float kt,
                                                  float Rn, float& cx, float& cy, float& cz) {
  const float dx = (h1.x - h2.x) + (l1.x - l2.x);
  const float dy = (h1.y - h2.y) + (l1.y - l2.y);
  const float dz = (h1.z - h2.z) + (l1.z - l2.z);
  const float r2 = fmaxf(__builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)), kR2Min);
  const float ri = rsq(r2);
  const float r3 = ri * ri * ri;
  if (kt == 0.0f) {  // no image charge
    cx = h2.w * (dx * r3);
    cy = h2.w * (dy * r3);
    cz = h2.w * (dz * r3);
    return;
  }
  // e = source - image of the target
  const float ex = (h2.x - ih1.x) + (l2.x - il1.x);
  const float ey = (h2.y - ih1.y) + (l2.y - il1.y);
  const float ez = (h2.z - ih1.z) + (l2.z - il1.z);
  const float t2 = fmaxf(__builtin_fmaf(ez, ez, __builtin_fmaf(ey, ey, ex * ex)), kR2Min);
  const float ti = rsq(t2);
  const float w = Rn * (ti * ti * ti) / ih1.w;
  const float a = 1.0f + kt;
  cx = h2.w * (dx * r3 * a + kt * (w * ex));
  cy = h2.w * (dy * r3 * a + kt * (w * ey));
  cz = h2.w * (dz * r3 * a + kt * (w * ez));
}

// Source-image pair (accelerations with j < i): the image is the source's;
// the reference negates the x and y image components (acc_tip_ic_force.inc
// with the operands exchanged).  f = target - image of the source.
__device__ __forceinline__ void tip_pair32_source(const float4 h1, const float4 l1,
                                                  const float4 h2, const float4 l2,
                                                  const float4 ih2, const float4 il2, float kt,
                                                  float Rn, float& cx, float& cy, float& cz) {
  const float dx = (h1.x - h2.x) + (l1.x - l2.x);
  const float dy = (h1.y - h2.y) + (l1.y - l2.y);
  const float dz = (h1.z - h2.z) + (l1.z - l2.z);
  const float r2 = fmaxf(__builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)), kR2Min);
  const float ri = rsq(r2);
  const float r3 = ri * ri * ri;
  if (kt == 0.0f) {  // no image charge
    cx = h2.w * (dx * r3);
    cy = h2.w * (dy * r3);
    cz = h2.w * (dz * r3);
    return;
  }
  const float fx = (h1.x - ih2.x) + (l1.x - il2.x);
  const float fy = (h1.y - ih2.y) + (l1.y - il2.y);
  const float fz = (h1.z - ih2.z) + (l1.z - il2.z);
  const float t2 = fmaxf(__builtin_fmaf(fz, fz, __builtin_fmaf(fy, fy, fx * fx)), kR2Min);
  const float ti = rsq(t2);
  const float w = Rn * (ti * ti * ti) / ih2.w;
  // x, y: d/r^3 + k_tip (d/r^3 - R f/(D t^3));  z: d/r^3 + k_tip (-d/r^3 + R f/(D t^3))
  cx = h2.w * (dx * r3 * (1.0f + kt) - kt * (w * fx));
  cy = h2.w * (dy * r3 * (1.0f + kt) - kt * (w * fy));
  cz = h2.w * (dz * r3 * (1.0f - kt) + kt * (w * fz));
}

// Mixed (tip): FP64 sums of the packed sources of chunks [c0, c1) at the
// packed target (h1, l1) with its packed image (ih1, il1).
__device__ __forceinline__ void field_seg32_tip(const DevGeom& g, bool act, float4 h1, float4 l1,
                                                float4 ih1, float4 il1,
                                                const float4* __restrict__ hi,
                                                const float4* __restrict__ lo, int n, int c0,
                                                int c1, float4* shi, float4* slo, double& fx,
                                                double& fy, double& fz) {
  const float kt = g.ic ? (float)g.ktip : 0.0f;
  const float Rn = (float)(g.R * 1e9);
  const int t = threadIdx.x;
  fx = 0.0;
  fy = 0.0;
  fz = 0.0;
  for (int c = c0; c < c1; ++c) {
    const int jb = c * kChunk;
    __syncthreads();
    if (jb + t < n) {
      shi[t] = hi[jb + t];
      slo[t] = lo[jb + t];
    }
    __syncthreads();
    const int cnt = min(kChunk, n - jb);
    if (act) {
      for (int jj = 0; jj < cnt; ++jj) {
        float cx, cy, cz;
        tip_pair32_target(h1, l1, ih1, il1, shi[jj], slo[jj], kt, Rn, cx, cy, cz);
        fx += (double)cx;
        fy += (double)cy;
        fz += (double)cz;
      }
    }
  }
}

}  // namespace rd
