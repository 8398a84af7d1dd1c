/*
 * ORACLE -- test infrastructure only (see rd_oracle.h).
 *
 * CPU restatement of the RUMDEED hot path.  Sources followed (relative to the
 * RUMDEED tree):
 *   constants ............ src/mod_global.F90:26-76, src/mod_field_emission_v2.F90:36-47
 *   planar image series .. src/acc_ic_planar_series.inc:19-63 (= Force_Image_charges_v2,
 *                          src/mod_verlet.F90:1924-1979)
 *   tip vacuum field ..... src/acc_tip_field_E.inc:13-34 (= field_E_Hyperboloid,
 *                          src/mod_hyperboloid_tip.f90:115-154)
 *   tip sphere image ..... src/acc_tip_image_point.inc:14-18, src/acc_tip_ic_force.inc:17-22
 *                          (= Sphere_IC_field, src/mod_hyperboloid_tip.f90:168-210)
 *   accelerations ........ Calculate_Acceleration_Particles_ACC, src/mod_verlet.F90:1217-1429
 *   field at points ...... Calc_Field_at / Calc_Field_at_Batch, src/mod_verlet.F90:1466-1529,
 *                          :1635-1911
 *
 * Canonical FP64 summation order.  The reference sums in an order that
 * depends on its OpenMP/OpenACC schedule.  This restatement fixes one order,
 * which the product's FP64 kernels reproduce bit for bit:
 *   for every target, the sources are split into consecutive chunks of
 *   RDO_CHUNK; each chunk is summed sequentially into a zeroed partial and the
 *   partials are added sequentially, chunk 0 first, into a zeroed total.
 * Per-pair terms follow the reference expressions operand for operand, with
 * floating-point contraction disabled (-ffp-contract=off).
 *
 * Deliberate restatements (documented, pinned by tests against the reference
 * build): in the tip vacuum field cos(atan2(y,x)) and sin(atan2(y,x)) are
 * written x/rho and y/rho, and t**(3/2) is written t*sqrt(t); both agree with
 * the reference to a few ulp and are what the product computes.
 */
#include "rd_oracle.h"

#include <math.h>
#include <stdint.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define RDO_CHUNK 256

/* ------------------------------------------------------------------------ */
/* Constants, built as src/mod_global.F90:26-76 builds them.                 */
static const double PI = 3.141592653589793238462643383279502884197;
static const double H_PLANCK = 6.62607015e-34;
static const double K_B = 1.380649e-23;
static const double C_LIGHT = 299792458.0;
static const double MU_0 = 1.25663706212e-6;
static const double M_0 = 9.1093837015e-31;
static const double Q_0 = 1.602176634e-19;
static const double M_U = 1.66053906660e-27;
static const double LENGTH_SCALE = 1.0e-9;
static const double TIME_SCALE = 1.0e-12;

static double eps0(void) { return 1.0 / (MU_0 * (C_LIGHT * C_LIGHT)); }
static double div_fac_c(void) { return 1.0 / (4.0 * PI * eps0() * 1.0); }
static double eps_soft(void) { return LENGTH_SCALE * LENGTH_SCALE; }
static double h_bar(void) { return H_PLANCK / (2.0 * PI); }
/* 1*q_0/(4*pi*epsilon_0) inside the sphere image force, acc_tip_ic_force.inc:20-22 */
static double k_tip(void) { return 1.0 * Q_0 / (4.0 * PI * eps0()); }

void rdo_get_consts(rdo_consts* o) {
  o->pi = PI;
  o->h = H_PLANCK;
  o->k_b = K_B;
  o->c = C_LIGHT;
  o->mu_0 = MU_0;
  o->epsilon_0 = eps0();
  o->h_bar = h_bar();
  o->m_0 = M_0;
  o->q_0 = Q_0;
  o->q_02 = Q_0 * Q_0;
  o->m_u = M_U;
  o->m_N2 = 28.0134 * M_U;
  o->m_N2p = 28.0134 * M_U - M_0;
  o->div_fac_c = div_fac_c();
  o->length_scale = LENGTH_SCALE;
  o->time_scale = TIME_SCALE;
  o->eps_soft = eps_soft();
  /* src/mod_field_emission_v2.F90:36-47 */
  o->a_FN = (Q_0 * Q_0) / (16.0 * (PI * PI) * h_bar());
  o->b_FN = -(4.0 / (3.0 * h_bar()) * sqrt(2.0 * M_0 * Q_0));
  o->l_const = Q_0 / (4.0 * PI * eps0());
}

int rdo_sum_chunk(void) { return RDO_CHUNK; }

/* ------------------------------------------------------------------------ */
/* Geometry set-up.                                                          */
void rdo_geom_planar(rdo_geom* g, double d, double V, int image_charge, int n_ic_max) {
  memset(g, 0, sizeof(*g));
  g->kind = RDO_GEOM_PLANAR;
  g->image_charge = image_charge;
  g->n_ic_max = n_ic_max;
  g->d = d;
  g->V = V;
  g->E_z = -1.0 * V / d; /* src/mod_verlet.F90:2051 */
}

/* Init_Emission_Tip, src/mod_emission_tip.f90:105-125. */
void rdo_geom_tip(rdo_geom* g, double V, double d_tip, double R_base, double h_tip,
                  int image_charge) {
  memset(g, 0, sizeof(*g));
  g->kind = RDO_GEOM_TIP;
  g->image_charge = image_charge;
  g->V = V;
  g->d_tip = d_tip;
  g->R_base = R_base;
  g->h_tip = h_tip;
  g->d = d_tip + h_tip;
  g->max_xi = h_tip / d_tip + 1.0;
  g->a_foci = sqrt(d_tip * d_tip * (R_base * R_base) / (h_tip * h_tip + 2 * d_tip * h_tip) +
                   d_tip * d_tip);
  g->eta_1 = -1.0 * d_tip / g->a_foci;
  g->theta_tip = acos(d_tip / g->a_foci);
  g->r_tip = g->a_foci * sin(g->theta_tip) * tan(g->theta_tip);
  g->shift_z = fabs(g->a_foci * g->eta_1 * g->max_xi);
  double eta_2 = 0.0;
  double pf = log((1.0 + g->eta_1) / (1.0 - g->eta_1) * (1.0 - eta_2) / (1.0 + eta_2));
  g->pre_fac_E_tip_unit_voltage = 2.0 * 1.0 / (g->a_foci * pf);
  g->pre_fac_E_tip = 2.0 * V / (g->a_foci * pf);
  g->E_z = 0.0;
}

/* ------------------------------------------------------------------------ */
/* Planar image-charge series, acc_ic_planar_series.inc:19-63.  Returns the  */
/* sum without the charge prefactor.                                         */
static inline void ic_planar(double z_a, double z_b, double diff_x, double diff_y,
                             double dxy2, int nic, double dd, double es, double* icx,
                             double* icy, double* icz) {
  double z_ic = -1.0 * z_b;
  double diff_z = z_a - z_ic;
  double r = sqrt(dxy2 + diff_z * diff_z) + es;
  double inv_r3 = 1.0 / (r * r * r);
  double ix = -diff_x * inv_r3;
  double iy = -diff_y * inv_r3;
  double iz = -diff_z * inv_r3;
  for (int n = 1; n <= nic; ++n) {
    double tnd = 2.0 * n * dd;
    /* opposite charge, +n and -n */
    z_ic = tnd - z_b;
    diff_z = z_a - z_ic;
    r = sqrt(dxy2 + diff_z * diff_z) + es;
    inv_r3 = 1.0 / (r * r * r);
    ix = ix - diff_x * inv_r3;
    iy = iy - diff_y * inv_r3;
    iz = iz - diff_z * inv_r3;
    z_ic = -2.0 * n * dd - z_b;
    diff_z = z_a - z_ic;
    r = sqrt(dxy2 + diff_z * diff_z) + es;
    inv_r3 = 1.0 / (r * r * r);
    ix = ix - diff_x * inv_r3;
    iy = iy - diff_y * inv_r3;
    iz = iz - diff_z * inv_r3;
    /* same charge, +n and -n */
    z_ic = tnd + z_b;
    diff_z = z_a - z_ic;
    r = sqrt(dxy2 + diff_z * diff_z) + es;
    inv_r3 = 1.0 / (r * r * r);
    ix = ix + diff_x * inv_r3;
    iy = iy + diff_y * inv_r3;
    iz = iz + diff_z * inv_r3;
    z_ic = -2.0 * n * dd + z_b;
    diff_z = z_a - z_ic;
    r = sqrt(dxy2 + diff_z * diff_z) + es;
    inv_r3 = 1.0 / (r * r * r);
    ix = ix + diff_x * inv_r3;
    iy = iy + diff_y * inv_r3;
    iz = iz + diff_z * inv_r3;
  }
  *icx = ix;
  *icy = iy;
  *icz = iz;
}

/* Tip vacuum field, acc_tip_field_E.inc:13-34 (cos/sin of atan2 written as
 * x/rho, y/rho; t**1.5 not used here). */
static inline void tip_vacuum(const rdo_geom* g, double x, double y, double z, double* fx,
                              double* fy, double* fz) {
  double af = g->a_foci, sz = g->shift_z;
  double r_p = sqrt(x * x + y * y + (z + af - sz) * (z + af - sz));
  double r_m = sqrt(x * x + y * y + (z - af - sz) * (z - af - sz));
  double xi = (r_p + r_m) / (2.0 * af);
  double eta = (r_p - r_m) / (2.0 * af);
  double c, s;
  if ((fabs(x) < 1.0e-18) && (fabs(y) < 1.0e-18)) {
    c = 1.0;
    s = 0.0;
  } else {
    double rho = sqrt(x * x + y * y);
    c = x / rho;
    s = y / rho;
  }
  double pre = g->pre_fac_E_tip * 1.0 / (xi * xi - eta * eta);
  double fac_xy;
  if (fabs(xi - 1.0) < 1.0e-6)
    fac_xy = 0.0;
  else
    fac_xy = eta * sqrt((xi * xi - 1.0) / (1.0 - eta * eta));
  *fx = -1.0 * pre * fac_xy * c;
  *fy = -1.0 * pre * fac_xy * s;
  *fz = pre * xi;
}

/* Sphere image point of the particle at a, acc_tip_image_point.inc:14-18. */
static inline void tip_image_point(double x_a, double y_a, double z_a, double z0, double R,
                                   double* dis_a, double* x_im, double* y_im, double* z_im) {
  double dz = z_a - z0;
  double da = sqrt(x_a * x_a + y_a * y_a + dz * dz);
  double zi = z0 + R * R / (sqrt(1 + x_a * x_a / (dz * dz) + y_a * y_a / (dz * dz)) * da);
  *z_im = zi;
  *x_im = (zi - z0) * x_a / dz;
  *y_im = (zi - z0) * y_a / dz;
  *dis_a = da;
}

/* Sphere image force, acc_tip_ic_force.inc:17-22, with t**1.5 = t*sqrt(t). */
static inline void tip_ic_force(double x_a, double y_a, double z_a, double x_b, double y_b,
                                double z_b, double dis_a, double x_im, double y_im,
                                double z_im, double R, double kt, double* icx, double* icy,
                                double* icz) {
  double ta = (x_b - x_a) * (x_b - x_a) + (y_b - y_a) * (y_b - y_a) + (z_b - z_a) * (z_b - z_a);
  double tb = (x_b - x_im) * (x_b - x_im) + (y_b - y_im) * (y_b - y_im) +
              (z_b - z_im) * (z_b - z_im);
  double tda = ta * sqrt(ta);
  double tdb = tb * sqrt(tb);
  *icx = kt * ((x_a - x_b) / tda - (R * (x_im - x_b)) / (dis_a * tdb));
  *icy = kt * ((y_a - y_b) / tda - (R * (y_im - y_b)) / (dis_a * tdb));
  *icz = kt * ((z_a - z_b) / tda - (R * (z_im - z_b)) / (dis_a * tdb));
}

void rdo_vacuum_field(const rdo_geom* g, const double* p, double* out) {
  if (g->kind == RDO_GEOM_TIP) {
    tip_vacuum(g, p[0], p[1], p[2], &out[0], &out[1], &out[2]);
  } else {
    out[0] = 0.0;
    out[1] = 0.0;
    out[2] = g->E_z;
  }
}

/* ------------------------------------------------------------------------ */
/* Pair term of the gather kernel for target i and source j != i.           */
/* Planar: src/mod_verlet.F90:1281-1330.  Tip: :1355-1413.                   */
static inline void pair_term(const rdo_geom* g, int first, int j, double x1, double y1,
                             double z1, double qd1, const double* pos, const double* q,
                             double es, double kt, double z0, double R, double* fx,
                             double* fy, double* fz) {
  double x2 = pos[3 * j], y2 = pos[3 * j + 1], z2 = pos[3 * j + 2];
  double pre = qd1 * q[j];
  double dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
  double ax, ay, az;
  if (g->kind == RDO_GEOM_PLANAR) {
    double dxy2 = dx * dx + dy * dy;
    double r = sqrt(dxy2 + dz * dz) + es;
    double inv_r3 = 1.0 / (r * r * r);
    ax = pre * inv_r3 * dx;
    ay = pre * inv_r3 * dy;
    az = pre * inv_r3 * dz;
    if (g->image_charge) {
      double z_a, z_b, icx, icy, icz;
      if (first) {
        z_a = z1;
        z_b = z2;
      } else {
        z_a = z2;
        z_b = z1;
      }
      ic_planar(z_a, z_b, dx, dy, dxy2, g->n_ic_max, g->d, es, &icx, &icy, &icz);
      ax = ax + pre * icx;
      ay = ay + pre * icy;
      az = az + pre * icz;
    }
  } else {
    double r = sqrt(dx * dx + dy * dy + dz * dz) + es;
    double inv_r3 = 1.0 / (r * r * r);
    ax = pre * inv_r3 * dx;
    ay = pre * inv_r3 * dy;
    az = pre * inv_r3 * dz;
    if (g->image_charge) {
      double xa, ya, za, xb, yb, zb, sgn;
      if (first) {
        xa = x1; ya = y1; za = z1; xb = x2; yb = y2; zb = z2; sgn = 1.0;
      } else {
        xa = x2; ya = y2; za = z2; xb = x1; yb = y1; zb = z1; sgn = -1.0;
      }
      double da, xi, yi, zi, icx, icy, icz;
      tip_image_point(xa, ya, za, z0, R, &da, &xi, &yi, &zi);
      tip_ic_force(xa, ya, za, xb, yb, zb, da, xi, yi, zi, R, kt, &icx, &icy, &icz);
      ax = ax + pre * sgn * icx;
      ay = ay + pre * sgn * icy;
      az = az + pre * icz;
    }
  }
  *fx = ax;
  *fy = ay;
  *fz = az;
}

/* Acceleration of target i over all n sources in the canonical order. */
/* Acceleration of target i.  The target takes the first role of a pair
 * (the image series evaluated at it) when it has the lower index
 * (Calculate_Acceleration_Particles_ACC, src/mod_verlet.F90:1217-1429).
 * With sp != NULL and K > 0, the two-time-step accelerations of step `step`
 * (Update_Elec_Acceleration / Update_Ion_Acceleration,
 * src/mod_verlet.F90:899-1151): between an electron and an ion the electron
 * takes the first role; an ion between its update steps (step % K != 0)
 * gets the electrons' reactions only, at its update steps also the other
 * ions and the vacuum field; an ion adds onto the acceleration it holds. */
static void accel_target_k(const rdo_geom* g, int n, const double* pos, const double* q,
                           const double* m, const int* sp, int K, int step, int i,
                           double* out) {
  const double es = eps_soft(), dfc = div_fac_c(), kt = k_tip();
  const double z0 = g->h_tip - g->r_tip, R = g->r_tip;
  {
    double* acc = out - 3 * i;
    double x1 = pos[3 * i], y1 = pos[3 * i + 1], z1 = pos[3 * i + 2];
    double q1 = q[i];
    double qd1 = q1 * dfc;
    const int tts = sp != NULL && K > 0;
    const int sp1 = tts ? sp[i] : 0;
    const int ion_only_e = tts && sp1 == 2 && step % K != 0;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int c0 = 0; c0 < n; c0 += RDO_CHUNK) {
      int c1 = c0 + RDO_CHUNK < n ? c0 + RDO_CHUNK : n;
      double px = 0.0, py = 0.0, pz = 0.0;
      for (int j = c0; j < c1; ++j) {
        if (j == i) continue;
        int first = j > i;
        if (tts) {
          if (ion_only_e && sp[j] == 2) continue;
          if (sp1 != sp[j]) first = sp1 == 1;
        }
        double fx, fy, fz;
        pair_term(g, first, j, x1, y1, z1, qd1, pos, q, es, kt, z0, R, &fx, &fy, &fz);
        px = px + fx;
        py = py + fy;
        pz = pz + fz;
      }
      sx = sx + px;
      sy = sy + py;
      sz = sz + pz;
    }
    double im = 1.0 / m[i];
    const int add = tts && sp1 == 2;
    const int vac = !add || step % K == 0;
    double a[3];
    if (g->kind == RDO_GEOM_PLANAR) {
      a[0] = sx * im;
      a[1] = sy * im;
      a[2] = (vac ? sz + q1 * g->E_z : sz) * im;
    } else if (vac) {
      double fx, fy, fz;
      tip_vacuum(g, x1, y1, z1, &fx, &fy, &fz);
      a[0] = (sx + q1 * fx) * im;
      a[1] = (sy + q1 * fy) * im;
      a[2] = (sz + q1 * fz) * im;
    } else {
      a[0] = sx * im;
      a[1] = sy * im;
      a[2] = sz * im;
    }
    for (int k = 0; k < 3; ++k) acc[3 * i + k] = add ? acc[3 * i + k] + a[k] : a[k];
  }
}
static void accel_target(const rdo_geom* g, int n, const double* pos, const double* q,
                         const double* m, int i, double* out) {
  accel_target_k(g, n, pos, q, m, NULL, 0, 0, i, out);
}

void rdo_accel_tts(const rdo_geom* g, int n, const double* pos, const double* q,
                   const double* m, const int* species, int K, int step, double* acc,
                   int nthreads) {
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
#endif
  for (int i = 0; i < n; ++i) accel_target_k(g, n, pos, q, m, species, K, step, i, &acc[3 * i]);
}

void rdo_accel(const rdo_geom* g, int n, const double* pos, const double* q, const double* m,
               double* acc, int nthreads) {
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
#endif
  for (int i = 0; i < n; ++i) accel_target(g, n, pos, q, m, i, &acc[3 * i]);
}

void rdo_accel_subset(const rdo_geom* g, int n, const double* pos, const double* q,
                      const double* m, int nt, const int* targets, double* acc, int nthreads) {
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
#endif
  for (int k = 0; k < nt; ++k) accel_target(g, n, pos, q, m, targets[k], &acc[3 * k]);
}

/* Field term of source j at point p (Calc_Field_at, src/mod_verlet.F90:1497-1524,
 * in the expression form of the batch kernels, :1761-1791 / :1841-1875). */
static inline void field_term(const rdo_geom* g, double x1, double y1, double z1, int j,
                              const double* pos, const double* q, double dfc, double es,
                              double kt, double R, double da, double xim, double yim,
                              double zim, double* fx, double* fy, double* fz) {
  double x2 = pos[3 * j], y2 = pos[3 * j + 1], z2 = pos[3 * j + 2];
  double pre = q[j] * dfc;
  double dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
  double ax, ay, az;
  if (g->kind == RDO_GEOM_PLANAR) {
    double dxy2 = dx * dx + dy * dy;
    double r = sqrt(dxy2 + dz * dz) + es;
    double inv_r3 = 1.0 / (r * r * r);
    ax = pre * inv_r3 * dx;
    ay = pre * inv_r3 * dy;
    az = pre * inv_r3 * dz;
    if (g->image_charge) {
      double icx, icy, icz;
      ic_planar(z1, z2, dx, dy, dxy2, g->n_ic_max, g->d, es, &icx, &icy, &icz);
      ax = ax + pre * icx;
      ay = ay + pre * icy;
      az = az + pre * icz;
    }
  } else {
    double r = sqrt(dx * dx + dy * dy + dz * dz) + es;
    double inv_r3 = 1.0 / (r * r * r);
    ax = pre * inv_r3 * dx;
    ay = pre * inv_r3 * dy;
    az = pre * inv_r3 * dz;
    if (g->image_charge) {
      double icx, icy, icz;
      tip_ic_force(x1, y1, z1, x2, y2, z2, da, xim, yim, zim, R, kt, &icx, &icy, &icz);
      ax = ax + pre * icx;
      ay = ay + pre * icy;
      az = az + pre * icz;
    }
  }
  *fx = ax;
  *fy = ay;
  *fz = az;
}

void rdo_field(const rdo_geom* g, int n, const double* pos, const double* q, int mpts,
               const double* pts, double* out, int nthreads) {
  const double es = eps_soft(), dfc = div_fac_c(), kt = k_tip();
  const double z0 = g->h_tip - g->r_tip, R = g->r_tip;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
#endif
  for (int p = 0; p < mpts; ++p) {
    double x1 = pts[3 * p], y1 = pts[3 * p + 1], z1 = pts[3 * p + 2];
    double da = 0, xim = 0, yim = 0, zim = 0;
    if (g->kind == RDO_GEOM_TIP && g->image_charge)
      tip_image_point(x1, y1, z1, z0, R, &da, &xim, &yim, &zim);
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int c0 = 0; c0 < n; c0 += RDO_CHUNK) {
      int c1 = c0 + RDO_CHUNK < n ? c0 + RDO_CHUNK : n;
      double px = 0.0, py = 0.0, pz = 0.0;
      for (int j = c0; j < c1; ++j) {
        double fx, fy, fz;
        field_term(g, x1, y1, z1, j, pos, q, dfc, es, kt, R, da, xim, yim, zim, &fx, &fy, &fz);
        px = px + fx;
        py = py + fy;
        pz = pz + fz;
      }
      sx = sx + px;
      sy = sy + py;
      sz = sz + pz;
    }
    double e[3];
    rdo_vacuum_field(g, &pts[3 * p], e);
    out[3 * p] = e[0] + sx;
    out[3 * p + 1] = e[1] + sy;
    out[3 * p + 2] = e[2] + sz;
  }
}

/* Conditioning of the sums above: for every point, the sum over the sources
 * of the norm of each source's term plus the norm of the vacuum field (the
 * scale against which a sum of rounded terms is accurate), and the same for
 * listed targets of the accelerations (times q_i / m_i like the sum itself). */
void rdo_field_scale(const rdo_geom* g, int n, const double* pos, const double* q, int mpts,
                     const double* pts, double* out, int nthreads) {
  const double es = eps_soft(), dfc = div_fac_c(), kt = k_tip();
  const double z0 = g->h_tip - g->r_tip, R = g->r_tip;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
#endif
  for (int p = 0; p < mpts; ++p) {
    double x1 = pts[3 * p], y1 = pts[3 * p + 1], z1 = pts[3 * p + 2];
    double da = 0, xim = 0, yim = 0, zim = 0;
    if (g->kind == RDO_GEOM_TIP && g->image_charge)
      tip_image_point(x1, y1, z1, z0, R, &da, &xim, &yim, &zim);
    double sc = 0.0;
    for (int j = 0; j < n; ++j) {
      double fx, fy, fz;
      field_term(g, x1, y1, z1, j, pos, q, dfc, es, kt, R, da, xim, yim, zim, &fx, &fy, &fz);
      sc += sqrt(fx * fx + fy * fy + fz * fz);
    }
    double e[3];
    rdo_vacuum_field(g, &pts[3 * p], e);
    out[p] = sc + sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
  }
}

void rdo_accel_scale(const rdo_geom* g, int n, const double* pos, const double* q,
                     const double* m, int nt, const int* targets, double* out, int nthreads) {
  const double es = eps_soft(), dfc = div_fac_c(), kt = k_tip();
  const double z0 = g->h_tip - g->r_tip, R = g->r_tip;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
#endif
  for (int k = 0; k < nt; ++k) {
    int i = targets[k];
    double x1 = pos[3 * i], y1 = pos[3 * i + 1], z1 = pos[3 * i + 2];
    double qd1 = q[i] * dfc;
    double sc = 0.0;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      double fx, fy, fz;
      pair_term(g, j > i, j, x1, y1, z1, qd1, pos, q, es, kt, z0, R, &fx, &fy, &fz);
      sc += sqrt(fx * fx + fy * fy + fz * fz);
    }
    double e[3];
    rdo_vacuum_field(g, &pos[3 * i], e);
    sc += fabs(q[i]) * sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    out[k] = sc / m[i];
  }
}

/* ------------------------------------------------------------------------ */
/* Long-double brute force: same reference semantics, natural order,        */
/* libm-style transcendental functions (atan2l/cosl/sinl/powl) as written.  */
typedef long double LD;

static inline void ic_planar_ld(LD z_a, LD z_b, LD dx, LD dy, LD dxy2, int nic, LD dd, LD es,
                                LD* icx, LD* icy, LD* icz) {
  LD ix = 0, iy = 0, iz = 0;
  for (int n = 0; n <= nic; ++n) {
    LD zs[4];
    LD sg[4];
    int cnt;
    if (n == 0) {
      zs[0] = -z_b;
      sg[0] = -1;
      cnt = 1;
    } else {
      zs[0] = 2.0L * n * dd - z_b;
      sg[0] = -1;
      zs[1] = -2.0L * n * dd - z_b;
      sg[1] = -1;
      zs[2] = 2.0L * n * dd + z_b;
      sg[2] = 1;
      zs[3] = -2.0L * n * dd + z_b;
      sg[3] = 1;
      cnt = 4;
    }
    for (int k = 0; k < cnt; ++k) {
      LD dz = z_a - zs[k];
      LD r = sqrtl(dxy2 + dz * dz) + es;
      LD inv = 1.0L / (r * r * r);
      ix += sg[k] * dx * inv;
      iy += sg[k] * dy * inv;
      iz += sg[k] * dz * inv;
    }
  }
  *icx = ix;
  *icy = iy;
  *icz = iz;
}

static void tip_vacuum_ld(const rdo_geom* g, LD x, LD y, LD z, LD* fx, LD* fy, LD* fz) {
  LD af = g->a_foci, sz = g->shift_z;
  LD r_p = sqrtl(x * x + y * y + (z + af - sz) * (z + af - sz));
  LD r_m = sqrtl(x * x + y * y + (z - af - sz) * (z - af - sz));
  LD xi = (r_p + r_m) / (2.0L * af);
  LD eta = (r_p - r_m) / (2.0L * af);
  LD phi = ((fabsl(x) < 1.0e-18L) && (fabsl(y) < 1.0e-18L)) ? 0.0L : atan2l(y, x);
  LD pre = (LD)g->pre_fac_E_tip / (xi * xi - eta * eta);
  LD fac = (fabsl(xi - 1.0L) < 1.0e-6L) ? 0.0L : eta * sqrtl((xi * xi - 1.0L) / (1.0L - eta * eta));
  *fx = -pre * fac * cosl(phi);
  *fy = -pre * fac * sinl(phi);
  *fz = pre * xi;
}

static void tip_ic_ld(LD xa, LD ya, LD za, LD xb, LD yb, LD zb, LD z0, LD R, LD kt, LD* icx,
                      LD* icy, LD* icz) {
  LD dz = za - z0;
  LD da = sqrtl(xa * xa + ya * ya + dz * dz);
  LD zi = z0 + R * R / (sqrtl(1 + xa * xa / (dz * dz) + ya * ya / (dz * dz)) * da);
  LD xi = (zi - z0) * xa / dz;
  LD yi = (zi - z0) * ya / dz;
  LD ta = powl((xb - xa) * (xb - xa) + (yb - ya) * (yb - ya) + (zb - za) * (zb - za), 1.5L);
  LD tb = powl((xb - xi) * (xb - xi) + (yb - yi) * (yb - yi) + (zb - zi) * (zb - zi), 1.5L);
  *icx = kt * ((xa - xb) / ta - (R * (xi - xb)) / (da * tb));
  *icy = kt * ((ya - yb) / ta - (R * (yi - yb)) / (da * tb));
  *icz = kt * ((za - zb) / ta - (R * (zi - zb)) / (da * tb));
}

static void accel_target_ld(const rdo_geom* g, int n, const double* pos, const double* q,
                            const double* m, int i, double* out) {
  const LD es = (LD)eps_soft();
  const LD dfc = 1.0L / (4.0L * 3.141592653589793238462643383279502884L *
                         (1.0L / ((LD)MU_0 * (LD)C_LIGHT * (LD)C_LIGHT)));
  const LD kt = (LD)Q_0 * dfc;
  const LD z0 = (LD)g->h_tip - (LD)g->r_tip, R = g->r_tip;
  {
    double* acc = out - 3 * i;
    LD x1 = pos[3 * i], y1 = pos[3 * i + 1], z1 = pos[3 * i + 2];
    LD q1 = q[i];
    LD ax = 0, ay = 0, az = 0;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      LD x2 = pos[3 * j], y2 = pos[3 * j + 1], z2 = pos[3 * j + 2];
      LD pre = q1 * dfc * (LD)q[j];
      LD dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
      LD r = sqrtl(dx * dx + dy * dy + dz * dz) + es;
      LD inv = 1.0L / (r * r * r);
      ax += pre * inv * dx;
      ay += pre * inv * dy;
      az += pre * inv * dz;
      if (!g->image_charge) continue;
      LD icx, icy, icz;
      if (g->kind == RDO_GEOM_PLANAR) {
        LD za = j > i ? z1 : z2, zb = j > i ? z2 : z1;
        ic_planar_ld(za, zb, dx, dy, dx * dx + dy * dy, g->n_ic_max, (LD)g->d, es, &icx, &icy,
                     &icz);
        ax += pre * icx;
        ay += pre * icy;
        az += pre * icz;
      } else {
        if (j > i) {
          tip_ic_ld(x1, y1, z1, x2, y2, z2, z0, R, kt, &icx, &icy, &icz);
          ax += pre * icx;
          ay += pre * icy;
        } else {
          tip_ic_ld(x2, y2, z2, x1, y1, z1, z0, R, kt, &icx, &icy, &icz);
          ax -= pre * icx;
          ay -= pre * icy;
        }
        az += pre * icz;
      }
    }
    LD ex = 0, ey = 0, ez = g->E_z;
    if (g->kind == RDO_GEOM_TIP) tip_vacuum_ld(g, x1, y1, z1, &ex, &ey, &ez);
    LD im = 1.0L / (LD)m[i];
    acc[3 * i] = (double)((ax + q1 * ex) * im);
    acc[3 * i + 1] = (double)((ay + q1 * ey) * im);
    acc[3 * i + 2] = (double)((az + q1 * ez) * im);
  }
}

void rdo_accel_ld(const rdo_geom* g, int n, const double* pos, const double* q, const double* m,
                  double* acc, int nthreads) {
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
#endif
  for (int i = 0; i < n; ++i) accel_target_ld(g, n, pos, q, m, i, &acc[3 * i]);
}

void rdo_accel_subset_ld(const rdo_geom* g, int n, const double* pos, const double* q,
                         const double* m, int nt, const int* targets, double* acc,
                         int nthreads) {
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
#endif
  for (int k = 0; k < nt; ++k) accel_target_ld(g, n, pos, q, m, targets[k], &acc[3 * k]);
}

void rdo_field_ld(const rdo_geom* g, int n, const double* pos, const double* q, int mpts,
                  const double* pts, double* out, int nthreads) {
  const LD es = (LD)eps_soft();
  const LD dfc = 1.0L / (4.0L * 3.141592653589793238462643383279502884L *
                         (1.0L / ((LD)MU_0 * (LD)C_LIGHT * (LD)C_LIGHT)));
  const LD kt = (LD)Q_0 * dfc;
  const LD z0 = (LD)g->h_tip - (LD)g->r_tip, R = g->r_tip;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
#endif
  for (int p = 0; p < mpts; ++p) {
    LD x1 = pts[3 * p], y1 = pts[3 * p + 1], z1 = pts[3 * p + 2];
    LD fx = 0, fy = 0, fz = 0;
    for (int j = 0; j < n; ++j) {
      LD x2 = pos[3 * j], y2 = pos[3 * j + 1], z2 = pos[3 * j + 2];
      LD pre = (LD)q[j] * dfc;
      LD dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
      LD r = sqrtl(dx * dx + dy * dy + dz * dz) + es;
      LD inv = 1.0L / (r * r * r);
      fx += pre * inv * dx;
      fy += pre * inv * dy;
      fz += pre * inv * dz;
      if (!g->image_charge) continue;
      LD icx, icy, icz;
      if (g->kind == RDO_GEOM_PLANAR)
        ic_planar_ld(z1, z2, dx, dy, dx * dx + dy * dy, g->n_ic_max, (LD)g->d, es, &icx, &icy,
                     &icz);
      else
        tip_ic_ld(x1, y1, z1, x2, y2, z2, z0, R, kt, &icx, &icy, &icz);
      fx += pre * icx;
      fy += pre * icy;
      fz += pre * icz;
    }
    LD ex = 0, ey = 0, ez = g->E_z;
    if (g->kind == RDO_GEOM_TIP) tip_vacuum_ld(g, x1, y1, z1, &ex, &ey, &ez);
    out[3 * p] = (double)(ex + fx);
    out[3 * p + 1] = (double)(ey + fy);
    out[3 * p + 2] = (double)(ez + fz);
  }
}
