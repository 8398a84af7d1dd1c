// All-pairs Coulomb force, image charges and surface field for gfx950.
//
// Reference: Calculate_Acceleration_Particles_ACC (src/mod_verlet.F90:1217-1429),
// Calc_Field_at / Calc_Field_at_Batch (:1466-1529, :1635-1911), the planar image
// series (src/acc_ic_planar_series.inc:19-63) and the hyperboloid tip
// (src/acc_tip_field_E.inc, src/acc_tip_image_point.inc, src/acc_tip_ic_force.inc).
//
// Shape: one lane per target (particle or field point), 256-lane workgroups;
// sources are staged through LDS 256 at a time and read as wavefront-uniform
// broadcasts.  The source range is split over gridDim.y segments so that a
// launch has enough workgroups to fill 256 CUs; every (segment, target)
// writes FP64 partial sums and a finishing kernel adds them in a fixed order.
// The pair work is vector-ALU arithmetic (square roots, reciprocals, FMAs),
// not a contraction, so nothing here uses MFMA.
//
// FP64 kernels: operand-for-operand the reference expressions, with the
//   partials formed per 256-source chunk and added chunk by chunk, so the
//   result is bit-identical to the CPU restatement that uses the same order.
// Mixed kernels: positions packed as float hi + float lo (in nm); pair
//   offsets are formed from both parts, so they are as accurate as FP32 allows
//   even for close pairs; image offsets are formed from sums of non-negative
//   heights (z_a + z_b from the cathode, (d - z_a) + (d - z_b) from the anode),
//   which never cancel; the pair arithmetic is FP32 with v_rsq_f32 and the
//   per-target sums are FP64.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "rd_internal.h"

namespace rd {

// ------------------------------------------------------------------------
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

// Sphere image point of the particle at a (acc_tip_image_point.inc:14-18).
__device__ __forceinline__ void tip_image_point64(double x_a, double y_a, double z_a, double z0,
                                                  double R, double& dis_a, double& x_im,
                                                  double& y_im, double& z_im) {
  double dz = z_a - z0;
  double da = __builtin_sqrt(x_a * x_a + y_a * y_a + dz * dz);
  double zi = z0 + R * R / (__builtin_sqrt(1 + x_a * x_a / (dz * dz) + y_a * y_a / (dz * dz)) * da);
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

// Image points of every particle (they depend on the particle alone).
__global__ void k_tip_images64(int n, const double* __restrict__ x, const double* __restrict__ y,
                               const double* __restrict__ z, DevGeom g, double* __restrict__ img) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double da, xi, yi, zi;
  tip_image_point64(x[i], y[i], z[i], g.z0, g.R, da, xi, yi, zi);
  img[i] = da;
  img[n + i] = xi;
  img[2 * n + i] = yi;
  img[3 * n + i] = zi;
}

// ------------------------------------------------------------------------
// FP64 accelerations, planar.  part layout: [3][nchunks][n].
__global__ __launch_bounds__(kBlock) void k_accel_planar64(
    int n, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ q, DevGeom g, int cps, int nchunks,
    double* __restrict__ part) {
  __shared__ double sx[kChunk], sy[kChunk], sz[kChunk], sq[kChunk];
  const int t = threadIdx.x;
  const int i = blockIdx.x * kBlock + t;
  const bool act = i < n;
  double x1 = 0, y1 = 0, z1 = 0, qd1 = 0;
  if (act) {
    x1 = x[i];
    y1 = y[i];
    z1 = z[i];
    qd1 = q[i] * g.dfc;
  }
  const int c0 = blockIdx.y * cps;
  const int c1 = min(nchunks, c0 + cps);
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
    double px = 0.0, py = 0.0, pz = 0.0;
    if (act) {
      for (int jj = 0; jj < cnt; ++jj) {
        const int j = jb + jj;
        if (j == i) continue;
        double x2 = sx[jj], y2 = sy[jj], z2 = sz[jj];
        double pre = qd1 * sq[jj];
        double dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
        double dxy2 = dx * dx + dy * dy;
        double r = __builtin_sqrt(dxy2 + dz * dz) + g.es;
        double inv_r3 = 1.0 / (r * r * r);
        double axx = pre * inv_r3 * dx;
        double ayy = pre * inv_r3 * dy;
        double azz = pre * inv_r3 * dz;
        if (g.ic) {
          double za = j > i ? z1 : z2;
          double zb = j > i ? z2 : z1;
          double icx, icy, icz;
          ic_planar64(za, zb, dx, dy, dxy2, g.nic, g.d, g.es, icx, icy, icz);
          axx = axx + pre * icx;
          ayy = ayy + pre * icy;
          azz = azz + pre * icz;
        }
        px = px + axx;
        py = py + ayy;
        pz = pz + azz;
      }
      part[(size_t)(0 * nchunks + c) * n + i] = px;
      part[(size_t)(1 * nchunks + c) * n + i] = py;
      part[(size_t)(2 * nchunks + c) * n + i] = pz;
    }
  }
}

// FP64 accelerations, hyperboloid tip (image data from k_tip_images64).
__global__ __launch_bounds__(kBlock) void k_accel_tip64(
    int n, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ q, const double* __restrict__ img,
    DevGeom g, int cps, int nchunks, double* __restrict__ part) {
  __shared__ double sx[kChunk], sy[kChunk], sz[kChunk], sq[kChunk];
  __shared__ double sd[kChunk], sxi[kChunk], syi[kChunk], szi[kChunk];
  const int t = threadIdx.x;
  const int i = blockIdx.x * kBlock + t;
  const bool act = i < n;
  double x1 = 0, y1 = 0, z1 = 0, qd1 = 0, d1 = 0, xi1 = 0, yi1 = 0, zi1 = 0;
  if (act) {
    x1 = x[i];
    y1 = y[i];
    z1 = z[i];
    qd1 = q[i] * g.dfc;
    if (g.ic) {
      d1 = img[i];
      xi1 = img[n + i];
      yi1 = img[2 * n + i];
      zi1 = img[3 * n + i];
    }
  }
  const int c0 = blockIdx.y * cps;
  const int c1 = min(nchunks, c0 + cps);
  for (int c = c0; c < c1; ++c) {
    const int jb = c * kChunk;
    __syncthreads();
    if (jb + t < n) {
      sx[t] = x[jb + t];
      sy[t] = y[jb + t];
      sz[t] = z[jb + t];
      sq[t] = q[jb + t];
      if (g.ic) {
        sd[t] = img[jb + t];
        sxi[t] = img[n + jb + t];
        syi[t] = img[2 * n + jb + t];
        szi[t] = img[3 * n + jb + t];
      }
    }
    __syncthreads();
    const int cnt = min(kChunk, n - jb);
    double px = 0.0, py = 0.0, pz = 0.0;
    if (act) {
      for (int jj = 0; jj < cnt; ++jj) {
        const int j = jb + jj;
        if (j == i) continue;
        double x2 = sx[jj], y2 = sy[jj], z2 = sz[jj];
        double pre = qd1 * sq[jj];
        double dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
        double r = __builtin_sqrt(dx * dx + dy * dy + dz * dz) + g.es;
        double inv_r3 = 1.0 / (r * r * r);
        double axx = pre * inv_r3 * dx;
        double ayy = pre * inv_r3 * dy;
        double azz = pre * inv_r3 * dz;
        if (g.ic) {
          double icx, icy, icz;
          if (j > i) {
            tip_ic_force64(x1, y1, z1, x2, y2, z2, d1, xi1, yi1, zi1, g.R, g.ktip, icx, icy, icz);
            axx = axx + pre * 1.0 * icx;
            ayy = ayy + pre * 1.0 * icy;
          } else {
            tip_ic_force64(x2, y2, z2, x1, y1, z1, sd[jj], sxi[jj], syi[jj], szi[jj], g.R,
                           g.ktip, icx, icy, icz);
            axx = axx + pre * -1.0 * icx;
            ayy = ayy + pre * -1.0 * icy;
          }
          azz = azz + pre * icz;
        }
        px = px + axx;
        py = py + ayy;
        pz = pz + azz;
      }
      part[(size_t)(0 * nchunks + c) * n + i] = px;
      part[(size_t)(1 * nchunks + c) * n + i] = py;
      part[(size_t)(2 * nchunks + c) * n + i] = pz;
    }
  }
}

// Sum the chunk partials in order and form a = (F + q E_vac)/m.
__global__ void k_accel_finish64(int n, int nchunks, const double* __restrict__ part,
                                 const double* __restrict__ x, const double* __restrict__ y,
                                 const double* __restrict__ z, const double* __restrict__ q,
                                 const double* __restrict__ m, DevGeom g, double* __restrict__ ax,
                                 double* __restrict__ ay, double* __restrict__ az) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int c = 0; c < nchunks; ++c) {
    sx = sx + part[(size_t)(0 * nchunks + c) * n + i];
    sy = sy + part[(size_t)(1 * nchunks + c) * n + i];
    sz = sz + part[(size_t)(2 * nchunks + c) * n + i];
  }
  double q1 = q[i];
  double im = 1.0 / m[i];
  if (g.kind == RD_GEOM_PLANAR) {
    ax[i] = sx * im;
    ay[i] = sy * im;
    az[i] = (sz + q1 * g.Ez) * im;
  } else {
    double fx, fy, fz;
    tip_vacuum64(g, x[i], y[i], z[i], fx, fy, fz);
    ax[i] = (sx + q1 * fx) * im;
    ay[i] = (sy + q1 * fy) * im;
    az[i] = (sz + q1 * fz) * im;
  }
}

// ------------------------------------------------------------------------
// FP64 field at points (SoA px,py,pz), planar and tip.
template <int KIND>
__global__ __launch_bounds__(kBlock) void k_field64(
    int m, const double* __restrict__ ptx, const double* __restrict__ pty,
    const double* __restrict__ ptz, int n, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z, const double* __restrict__ q,
    DevGeom g, int cps, int nchunks, double* __restrict__ part) {
  __shared__ double sx[kChunk], sy[kChunk], sz[kChunk], sq[kChunk];
  const int t = threadIdx.x;
  const int p = blockIdx.x * kBlock + t;
  const bool act = p < m;
  double x1 = 0, y1 = 0, z1 = 0, da = 0, xim = 0, yim = 0, zim = 0;
  if (act) {
    x1 = ptx[p];
    y1 = pty[p];
    z1 = ptz[p];
    if (KIND == RD_GEOM_TIP && g.ic) tip_image_point64(x1, y1, z1, g.z0, g.R, da, xim, yim, zim);
  }
  const int c0 = blockIdx.y * cps;
  const int c1 = min(nchunks, c0 + cps);
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
      part[(size_t)(0 * nchunks + c) * m + p] = fxs;
      part[(size_t)(1 * nchunks + c) * m + p] = fys;
      part[(size_t)(2 * nchunks + c) * m + p] = fzs;
    }
  }
}

// field = E_vac + sum of chunk partials (in order).  nchunks may be 0.
__global__ void k_field_finish64(int m, int nchunks, const double* __restrict__ part,
                                 const double* __restrict__ ptx, const double* __restrict__ pty,
                                 const double* __restrict__ ptz, DevGeom g,
                                 double* __restrict__ fx, double* __restrict__ fy,
                                 double* __restrict__ fz) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= m) return;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int c = 0; c < nchunks; ++c) {
    sx = sx + part[(size_t)(0 * nchunks + c) * m + p];
    sy = sy + part[(size_t)(1 * nchunks + c) * m + p];
    sz = sz + part[(size_t)(2 * nchunks + c) * m + p];
  }
  double ex = 0.0, ey = 0.0, ez = g.Ez;
  if (g.kind == RD_GEOM_TIP) tip_vacuum64(g, ptx[p], pty[p], ptz[p], ex, ey, ez);
  fx[p] = ex + sx;
  fy[p] = ey + sy;
  fz[p] = ez + sz;
}

// ------------------------------------------------------------------------
// Mixed precision.
//
// Packed source/target record (nm): hi = (x_hi, y_hi, z_hi, q/q_0),
// lo = (x_lo, y_lo, z_lo, d - z) with x = x_hi + x_lo to ~2^-48 relative.
__global__ void k_pack_mixed(int n, const double* __restrict__ x, const double* __restrict__ y,
                             const double* __restrict__ z, const double* __restrict__ q,
                             double d, float4* __restrict__ hi, float4* __restrict__ lo) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double s = 1.0e9;
  double xn = x[i] * s, yn = y[i] * s, zn = z[i] * s;
  float xh = (float)xn, yh = (float)yn, zh = (float)zn;
  float qr = q ? (float)(q[i] / Consts::q_0) : 0.0f;
  hi[i] = make_float4(xh, yh, zh, qr);
  lo[i] = make_float4((float)(xn - (double)xh), (float)(yn - (double)yh),
                      (float)(zn - (double)zh), (float)(d * s - zn));
}

__device__ __forceinline__ float rsq(float v) { return __builtin_amdgcn_rsqf(v); }

// Guard for squared distances that can be exactly zero (a pair with itself
// or coincident particles, a particle exactly on an electrode): keeps every
// FP32 term finite; the reference's r + 1e-18 m softening gives the same 0
// contribution for zero offsets.
constexpr float kR2Min = 1.0e-20f;

// One image partner: accumulate 1/r^3 and dz/r^3.
#define RD_PARTNER(DZ, W, WZ)                 \
  {                                           \
    float _dz = (DZ);                         \
    float _r2 = __builtin_fmaf(_dz, _dz, dxy2); \
    float _ri = rsq(_r2);                     \
    float _r3 = _ri * _ri * _ri;              \
    W += _r3;                                 \
    WZ = __builtin_fmaf(_dz, _r3, WZ);        \
  }

// Image-charge weights of one pair: returns (sum_same - sum_opp) of 1/r^3 in
// wxy and of dz/r^3 in wz.  zsum = z_a + z_b, wsum = (d - z_a) + (d - z_b),
// dzab = z_a - z_b.
__device__ __forceinline__ void ic_planar32(float zsum, float wsum, float dzab, float dxy2,
                                            int nic, float two_d, float& wxy, float& wz) {
  float wo = 0.f, wzo = 0.f, ws = 0.f, wzs = 0.f;
  {
    float dz0 = zsum;
    float r2 = fmaxf(__builtin_fmaf(dz0, dz0, dxy2), kR2Min);
    float ri = rsq(r2);
    float r3 = ri * ri * ri;
    wo += r3;
    wzo = __builtin_fmaf(dz0, r3, wzo);
  }
  if (nic >= 1) {
    // n = 1: the opposite partner across the anode is at distance -(wsum)
    float dz1 = -wsum;
    float r2 = fmaxf(__builtin_fmaf(dz1, dz1, dxy2), kR2Min);
    float ri = rsq(r2);
    float r3 = ri * ri * ri;
    wo += r3;
    wzo = __builtin_fmaf(dz1, r3, wzo);
    RD_PARTNER(zsum + two_d, wo, wzo);
    RD_PARTNER(dzab - two_d, ws, wzs);
    RD_PARTNER(dzab + two_d, ws, wzs);
  }
  float tn = two_d;
  for (int n = 2; n <= nic; ++n) {
    tn += two_d;
    RD_PARTNER(zsum - tn, wo, wzo);
    RD_PARTNER(zsum + tn, wo, wzo);
    RD_PARTNER(dzab - tn, ws, wzs);
    RD_PARTNER(dzab + tn, ws, wzs);
  }
  wxy = ws - wo;
  wz = wzs - wzo;
}

// Mixed accelerations (planar).  part layout: [3][segs][n] FP64 sums of
// q_j/q_0 * (offset / r^3 + image terms) in nm^-2.
__global__ __launch_bounds__(kBlock) void k_accel_planar32(
    int n, const float4* __restrict__ hi, const float4* __restrict__ lo, int ic, int nic,
    float two_d, int cps, int nchunks, double* __restrict__ part) {
  __shared__ float4 shi[kChunk], slo[kChunk];
  const int t = threadIdx.x;
  const int i = blockIdx.x * kBlock + t;
  const bool act = i < n;
  float4 h1 = make_float4(0, 0, 0, 0), l1 = make_float4(0, 0, 0, 0);
  if (act) {
    h1 = hi[i];
    l1 = lo[i];
  }
  double fx = 0.0, fy = 0.0, fz = 0.0;
  const int c0 = blockIdx.y * cps;
  const int c1 = min(nchunks, c0 + cps);
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
        const int j = jb + jj;
        const float4 h2 = shi[jj], l2 = slo[jj];
        float dx = (h1.x - h2.x) + (l1.x - l2.x);
        float dy = (h1.y - h2.y) + (l1.y - l2.y);
        float dz = (h1.z - h2.z) + (l1.z - l2.z);
        float qj = (j == i) ? 0.0f : h2.w;
        float dxy2 = __builtin_fmaf(dy, dy, dx * dx);
        float r2 = fmaxf(__builtin_fmaf(dz, dz, dxy2), kR2Min);
        float ri = rsq(r2);
        float r3 = ri * ri * ri;
        float cxy = r3, cz = dz * r3;
        if (ic) {
          float wxy, wz;
          float dzab = (j > i) ? dz : -dz;
          ic_planar32(h1.z + h2.z, l1.w + l2.w, dzab, dxy2, nic, two_d, wxy, wz);
          cxy += wxy;
          cz += wz;
        }
        cxy *= qj;
        cz *= qj;
        fx += (double)(cxy * dx);
        fy += (double)(cxy * dy);
        fz += (double)cz;
      }
    }
  }
  if (act) {
    const int segs = gridDim.y;
    part[(size_t)(0 * segs + blockIdx.y) * n + i] = fx;
    part[(size_t)(1 * segs + blockIdx.y) * n + i] = fy;
    part[(size_t)(2 * segs + blockIdx.y) * n + i] = fz;
  }
}

__global__ void k_accel_finish32(int n, int segs, const double* __restrict__ part,
                                 const double* __restrict__ q, const double* __restrict__ m,
                                 DevGeom g, double* __restrict__ ax, double* __restrict__ ay,
                                 double* __restrict__ az) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int s = 0; s < segs; ++s) {
    sx += part[(size_t)(0 * segs + s) * n + i];
    sy += part[(size_t)(1 * segs + s) * n + i];
    sz += part[(size_t)(2 * segs + s) * n + i];
  }
  double q1 = q[i];
  double im = 1.0 / m[i];
  double pre = g.dfc * q1 * Consts::q_0 * 1.0e18 * im;
  ax[i] = pre * sx;
  ay[i] = pre * sy;
  az[i] = pre * sz + q1 * g.Ez * im;
}

// Mixed field at points (planar).
__global__ __launch_bounds__(kBlock) void k_field_planar32(
    int m, const float4* __restrict__ phi, const float4* __restrict__ plo, int n,
    const float4* __restrict__ hi, const float4* __restrict__ lo, int ic, int nic, float two_d,
    int cps, int nchunks, double* __restrict__ part) {
  __shared__ float4 shi[kChunk], slo[kChunk];
  const int t = threadIdx.x;
  const int p = blockIdx.x * kBlock + t;
  const bool act = p < m;
  float4 h1 = make_float4(0, 0, 0, 0), l1 = make_float4(0, 0, 0, 0);
  if (act) {
    h1 = phi[p];
    l1 = plo[p];
  }
  double fx = 0.0, fy = 0.0, fz = 0.0;
  const int c0 = blockIdx.y * cps;
  const int c1 = min(nchunks, c0 + cps);
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
  if (act) {
    const int segs = gridDim.y;
    part[(size_t)(0 * segs + blockIdx.y) * m + p] = fx;
    part[(size_t)(1 * segs + blockIdx.y) * m + p] = fy;
    part[(size_t)(2 * segs + blockIdx.y) * m + p] = fz;
  }
}

__global__ void k_field_finish32(int m, int segs, const double* __restrict__ part, DevGeom g,
                                 double* __restrict__ fx, double* __restrict__ fy,
                                 double* __restrict__ fz) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= m) return;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  for (int s = 0; s < segs; ++s) {
    sx += part[(size_t)(0 * segs + s) * m + p];
    sy += part[(size_t)(1 * segs + s) * m + p];
    sz += part[(size_t)(2 * segs + s) * m + p];
  }
  double pre = g.dfc * Consts::q_0 * 1.0e18;
  fx[p] = pre * sx;
  fy[p] = pre * sy;
  fz[p] = g.Ez + pre * sz;
}

__global__ void k_grid_points(int nx, int ny, double x0, double y0, double lx, double ly,
                              double zz, double* __restrict__ ptx, double* __restrict__ pty,
                              double* __restrict__ ptz) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nx * ny) return;
  int ix = p % nx, iy = p / nx;
  ptx[p] = x0 + (ix + 0.5) * (lx / nx);
  pty[p] = y0 + (iy + 0.5) * (ly / ny);
  ptz[p] = zz;
}

// ------------------------------------------------------------------------
// Launch helpers.

// Segments of the source range: enough workgroups for ~8 per CU.
static int pick_segments(int target_blocks, int nchunks) {
  const int want = 2048;
  int segs = (want + target_blocks - 1) / target_blocks;
  segs = std::max(1, std::min(segs, nchunks));
  return segs;
}

int launch_accel(rd_ctx* c, int precision, hipEvent_t k0, hipEvent_t k1) {
  const int n = c->n;
  if (n == 0) return RD_OK;
  const int tb = (n + kBlock - 1) / kBlock;
  const int nchunks = (n + kChunk - 1) / kChunk;
  const int fin = (n + 255) / 256;
  if (precision == RD_PREC_FP64) {
    int r = ensure(c, c->part, sizeof(double) * 3 * (size_t)nchunks * n);
    if (r) return r;
    int segs = pick_segments(tb, nchunks);
    int cps = (nchunks + segs - 1) / segs;
    segs = (nchunks + cps - 1) / cps;
    dim3 grid(tb, segs);
    if (c->g.kind == RD_GEOM_PLANAR) {
      if (k0) RD_HIP(hipEventRecord(k0, c->stream));
      k_accel_planar64<<<grid, kBlock, 0, c->stream>>>(n, c->x, c->y, c->z, c->q, c->g, cps,
                                                        nchunks, (double*)c->part.p);
      if (k1) RD_HIP(hipEventRecord(k1, c->stream));
    } else {
      r = ensure(c, c->img64, sizeof(double) * 4 * (size_t)n);
      if (r) return r;
      if (c->g.ic)
        k_tip_images64<<<fin, 256, 0, c->stream>>>(n, c->x, c->y, c->z, c->g,
                                                    (double*)c->img64.p);
      if (k0) RD_HIP(hipEventRecord(k0, c->stream));
      k_accel_tip64<<<grid, kBlock, 0, c->stream>>>(n, c->x, c->y, c->z, c->q,
                                                     (const double*)c->img64.p, c->g, cps,
                                                     nchunks, (double*)c->part.p);
      if (k1) RD_HIP(hipEventRecord(k1, c->stream));
    }
    k_accel_finish64<<<fin, 256, 0, c->stream>>>(n, nchunks, (const double*)c->part.p, c->x,
                                                  c->y, c->z, c->q, c->m, c->g, c->ax, c->ay,
                                                  c->az);
    RD_HIP(hipGetLastError());
    return RD_OK;
  }
  if (precision != RD_PREC_MIXED) return fail(RD_ERR_ARG, "unknown precision");
  if (c->g.kind != RD_GEOM_PLANAR)
    return fail(RD_ERR_ARG, "mixed precision is implemented for the planar geometry only");
  int r = ensure(c, c->pack_hi, sizeof(float4) * (size_t)n);
  if (!r) r = ensure(c, c->pack_lo, sizeof(float4) * (size_t)n);
  if (r) return r;
  int segs = pick_segments(tb, nchunks);
  int cps = (nchunks + segs - 1) / segs;
  segs = (nchunks + cps - 1) / cps;
  r = ensure(c, c->part, sizeof(double) * 3 * (size_t)segs * n);
  if (r) return r;
  k_pack_mixed<<<fin, 256, 0, c->stream>>>(n, c->x, c->y, c->z, c->q, c->g.d,
                                            (float4*)c->pack_hi.p, (float4*)c->pack_lo.p);
  if (k0) RD_HIP(hipEventRecord(k0, c->stream));
  k_accel_planar32<<<dim3(tb, segs), kBlock, 0, c->stream>>>(
      n, (const float4*)c->pack_hi.p, (const float4*)c->pack_lo.p, c->g.ic, c->g.nic,
      (float)(2.0 * c->g.d * 1e9), cps, nchunks, (double*)c->part.p);
  if (k1) RD_HIP(hipEventRecord(k1, c->stream));
  k_accel_finish32<<<fin, 256, 0, c->stream>>>(n, segs, (const double*)c->part.p, c->q, c->m,
                                                c->g, c->ax, c->ay, c->az);
  RD_HIP(hipGetLastError());
  return RD_OK;
}

// Field at the m points held in c->pts (SoA); result SoA in c->fld.
int launch_field(rd_ctx* c, int m, int precision, hipEvent_t k0, hipEvent_t k1) {
  if (m == 0) return RD_OK;
  const int n = c->n;
  double* ptx = (double*)c->pts.p;
  double* pty = ptx + m;
  double* ptz = pty + m;
  double* fx = (double*)c->fld.p;
  double* fy = fx + m;
  double* fz = fy + m;
  const int tb = (m + kBlock - 1) / kBlock;
  const int nchunks = (n + kChunk - 1) / kChunk;
  const int fin = (m + 255) / 256;
  if (precision == RD_PREC_FP64) {
    int segs = 1, cps = 1;
    if (nchunks > 0) {
      int r = ensure(c, c->part, sizeof(double) * 3 * (size_t)nchunks * m);
      if (r) return r;
      segs = pick_segments(tb, nchunks);
      cps = (nchunks + segs - 1) / segs;
      segs = (nchunks + cps - 1) / cps;
      if (k0) RD_HIP(hipEventRecord(k0, c->stream));
      if (c->g.kind == RD_GEOM_PLANAR)
        k_field64<RD_GEOM_PLANAR><<<dim3(tb, segs), kBlock, 0, c->stream>>>(
            m, ptx, pty, ptz, n, c->x, c->y, c->z, c->q, c->g, cps, nchunks,
            (double*)c->part.p);
      else
        k_field64<RD_GEOM_TIP><<<dim3(tb, segs), kBlock, 0, c->stream>>>(
            m, ptx, pty, ptz, n, c->x, c->y, c->z, c->q, c->g, cps, nchunks,
            (double*)c->part.p);
      if (k1) RD_HIP(hipEventRecord(k1, c->stream));
    }
    k_field_finish64<<<fin, 256, 0, c->stream>>>(m, nchunks, (const double*)c->part.p, ptx, pty,
                                                  ptz, c->g, fx, fy, fz);
    RD_HIP(hipGetLastError());
    return RD_OK;
  }
  if (precision != RD_PREC_MIXED) return fail(RD_ERR_ARG, "unknown precision");
  if (c->g.kind != RD_GEOM_PLANAR)
    return fail(RD_ERR_ARG, "mixed precision is implemented for the planar geometry only");
  int r = ensure(c, c->pts_hi, sizeof(float4) * (size_t)m);
  if (!r) r = ensure(c, c->pts_lo, sizeof(float4) * (size_t)m);
  if (!r) r = ensure(c, c->pack_hi, sizeof(float4) * (size_t)std::max(n, 1));
  if (!r) r = ensure(c, c->pack_lo, sizeof(float4) * (size_t)std::max(n, 1));
  if (r) return r;
  int segs = 1, cps = 1;
  if (nchunks > 0) {
    segs = pick_segments(tb, nchunks);
    cps = (nchunks + segs - 1) / segs;
    segs = (nchunks + cps - 1) / cps;
  }
  r = ensure(c, c->part, sizeof(double) * 3 * (size_t)segs * m);
  if (r) return r;
  k_pack_mixed<<<fin, 256, 0, c->stream>>>(m, ptx, pty, ptz, nullptr, c->g.d,
                                            (float4*)c->pts_hi.p, (float4*)c->pts_lo.p);
  if (n > 0)
    k_pack_mixed<<<(n + 255) / 256, 256, 0, c->stream>>>(n, c->x, c->y, c->z, c->q, c->g.d,
                                                          (float4*)c->pack_hi.p,
                                                          (float4*)c->pack_lo.p);
  if (k0) RD_HIP(hipEventRecord(k0, c->stream));
  k_field_planar32<<<dim3(tb, segs), kBlock, 0, c->stream>>>(
      m, (const float4*)c->pts_hi.p, (const float4*)c->pts_lo.p, n, (const float4*)c->pack_hi.p,
      (const float4*)c->pack_lo.p, c->g.ic, c->g.nic, (float)(2.0 * c->g.d * 1e9), cps, nchunks,
      (double*)c->part.p);
  if (k1) RD_HIP(hipEventRecord(k1, c->stream));
  k_field_finish32<<<fin, 256, 0, c->stream>>>(m, segs, (const double*)c->part.p, c->g, fx, fy,
                                                fz);
  RD_HIP(hipGetLastError());
  return RD_OK;
}

int launch_grid_points(rd_ctx* c, double x0, double y0, double lx, double ly, int nx, int ny,
                       double zz) {
  int m = nx * ny;
  int r = ensure(c, c->pts, sizeof(double) * 3 * (size_t)m);
  if (!r) r = ensure(c, c->fld, sizeof(double) * 3 * (size_t)m);
  if (r) return r;
  double* ptx = (double*)c->pts.p;
  k_grid_points<<<(m + 255) / 256, 256, 0, c->stream>>>(nx, ny, x0, y0, lx, ly, zz, ptx, ptx + m,
                                                         ptx + 2 * m);
  RD_HIP(hipGetLastError());
  return RD_OK;
}

}  // namespace rd
