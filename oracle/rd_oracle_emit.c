/*
 * ORACLE -- test infrastructure only (see rd_oracle.h).
 *
 * CPU restatement of one RUMDEED simulation step, one simulation at a time:
 *   emission ........ Do_Field_Emission_Planar_rectangle + Metropolis_Hastings_rectangle_J_batch
 *                     (src/mod_field_emission_v2.F90:261-395, :1284-1516),
 *                     Do_Field_Thermo_Emission_Planar_simple + its MH
 *                     (src/mod_field_thermo_emission.F90:136-446),
 *                     Do_Field_Emission_Tip_OLDCODE + Metro_algo_tip_v3
 *                     (src/mod_emission_tip.f90:417-534, :1213-1390)
 *   FN / GTF ........ src/mod_field_emission_v2.F90:515-625, src/mod_emission_tip.f90:1657-1758,
 *                     src/mod_kevin_rjgtf_v2.f90:56-175
 *   work function ... w_theta_checkerboard, src/mod_work_function.F90:389-487
 *   Beeman .......... Update_ElecIon_Position / _Velocity, src/mod_verlet.F90:197-232, :449-509;
 *                     two-time-step Update_Species_Position / _Velocity and the species
 *                     accelerations, src/mod_verlet.F90:182-193, :234-290, :511-580, :899-1151
 *   boundaries ...... Check_Boundary_Planar src/mod_verlet.F90:325-338,
 *                     Check_Boundary_Tip src/mod_emission_tip.f90:1627-1647
 *   removal ......... Mark_Particles_Remove, Remove_Particles, record_lifetime,
 *                     src/mod_pair.F90:169-562, :1135-1160
 *   Add_Particle .... src/mod_pair.F90:29-159
 *
 * Where the reference leaves a choice to its runtime (RNG stream, summation
 * order of the supply integral, chains sampled in lockstep) this restatement
 * fixes the same choice as the device engine:
 *   - draws come from Philox4x64-10 with key (seed, stream) and counter
 *     (step, purpose<<40 | sub, chain, block);
 *   - the planar supply integral (Divonne in the reference) is the midpoint
 *     rule on a grid of the emitter whose cells follow the work function's
 *     sections (at least quad points per side), refined by doubling the
 *     points per side until two levels agree within the reference's Divonne
 *     tolerances; the tip keeps the reference's 100 x 100 (xi, phi) cells;
 *     each level is summed in tiles of 1024 points, pairwise within a tile,
 *     tiles in order;
 *   - all chains of a step see the particles of the start of the step; the
 *     planar FE width is updated per jump from the batch acceptance rate, the
 *     thermal-field and tip widths per chain in chain order after the step.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "rd_oracle.h"

#define LS 1.0e-9
#define KTINY 2.2250738585072014e-308
#define KHUGE 1.7976931348623157e308
#define PI_D 3.141592653589793238462643383279502884197
#define NLIFE 1000

enum { EM_NONE = 0, EM_TIP = 3, EM_GTF = 9, EM_FE = 10 };

/* ---- own sin/cos (same algorithm as the device library) ---------------- */
static const double PIO2_HI = 0x1.921fb54400000p+0;
static const double PIO2_MID = 0x1.0b4611a600000p-34;
static const double PIO2_LO = 0x1.3198a2e037073p-69;
static const double INV_PIO2 = 0x1.45f306dc9c883p-1;

static double sin_poly(double r) {
  double z = r * r;
  double p = 1.0 / 355687428096000.0;
  p = -1.0 / 1307674368000.0 + z * p;
  p = 1.0 / 6227020800.0 + z * p;
  p = -1.0 / 39916800.0 + z * p;
  p = 1.0 / 362880.0 + z * p;
  p = -1.0 / 5040.0 + z * p;
  p = 1.0 / 120.0 + z * p;
  p = -1.0 / 6.0 + z * p;
  return r + (r * z) * p;
}
static double cos_poly(double r) {
  double z = r * r;
  double p = 1.0 / 20922789888000.0;
  p = -1.0 / 87178291200.0 + z * p;
  p = 1.0 / 479001600.0 + z * p;
  p = -1.0 / 3628800.0 + z * p;
  p = 1.0 / 40320.0 + z * p;
  p = -1.0 / 720.0 + z * p;
  p = 1.0 / 24.0 + z * p;
  p = -0.5 + z * p;
  return 1.0 + z * p;
}
void rdo_sincos(double x, double* s, double* c) {
  double kd = floor(x * INV_PIO2 + 0.5);
  double r = ((x - kd * PIO2_HI) - kd * PIO2_MID) - kd * PIO2_LO;
  double sp = sin_poly(r), cp = cos_poly(r);
  int q = ((int)kd) & 3;
  if (q == 0) { *s = sp; *c = cp; }
  else if (q == 1) { *s = cp; *c = -sp; }
  else if (q == 2) { *s = -sp; *c = -cp; }
  else { *s = -cp; *c = sp; }
}

static int fnint(double x) {
  double t = trunc(x);
  double f = x - t;
  if (f >= 0.5) t = t + 1.0;
  else if (f <= -0.5) t = t - 1.0;
  return (int)t;
}
static double fmodulo(double a, double p) { return a - floor(a / p) * p; }

/* ---- counter-based draws ------------------------------------------------ */
enum { T_START = 1, T_JUMP = 2, T_ACCEPT = 3, T_EMIT = 4, T_POISSON = 5, T_VEL = 6 };
static uint64_t tagv(uint64_t kind, uint64_t sub) { return (kind << 40) | sub; }
static double u01(uint64_t u) { return ((double)(u >> 11) + 0.5) * 0x1p-53; }

static void ph(uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3, uint64_t k0, uint64_t k1,
               uint64_t out[4]) {
  uint64_t ctr[4] = {c0, c1, c2, c3}, key[2] = {k0, k1};
  rdo_philox4x64(ctr, key, out);
}

/* box_muller (src/mod_global.F90:578-595), polar form */
static void box_muller(uint64_t seed, uint64_t sim, uint64_t step, uint64_t tg, uint64_t chain,
                       double m1, double m2, double s1, double s2, double* o1, double* o2) {
  for (uint64_t blk = 0; blk < 4096; ++blk) {
    uint64_t r[4];
    ph(step, tg, chain, blk, seed, sim, r);
    for (int a = 0; a < 2; ++a) {
      double x1 = 2.0 * u01(r[2 * a]) - 1.0;
      double x2 = 2.0 * u01(r[2 * a + 1]) - 1.0;
      double w = x1 * x1 + x2 * x2;
      if ((w < 1.0) && (w > 0.0)) {
        double f = sqrt((-2.0 * rdo_log(w)) / w);
        *o1 = x1 * f * s1 + m1;
        *o2 = x2 * f * s2 + m2;
        return;
      }
    }
  }
  *o1 = m1;
  *o2 = m2;
}
static double uniform(uint64_t seed, uint64_t sim, uint64_t step, uint64_t tg, uint64_t chain,
                      int which) {
  uint64_t r[4];
  ph(step, tg, chain, 0, seed, sim, r);
  return u01(r[which & 3]);
}
/* Rand_Poisson (src/mod_global.F90:600-643) */
static int poisson(uint64_t seed, uint64_t sim, uint64_t step, double lambda) {
  double lambda_left = lambda;
  int k = 0;
  double p = 1.0;
  const double STEP = 500.0;
  uint64_t cnt = 0;
  do {
    k = k + 1;
    uint64_t r[4];
    ph(step, tagv(T_POISSON, 0), cnt >> 2, 0, seed, sim, r);
    double u = u01(r[cnt & 3]);
    ++cnt;
    p = p * u;
    while ((p < 1.0) && (lambda_left > 0.0)) {
      if (lambda_left > STEP) {
        p = p * rdo_exp(STEP);
        lambda_left = lambda_left - STEP;
      } else {
        p = p * rdo_exp(lambda_left);
        lambda_left = 0.0;
      }
    }
  } while (p >= 1.0 && k < (1 << 30));
  return k - 1;
}

/* ---- FN functions (src/mod_field_emission_v2.F90:515-625) --------------- */
static double a_FN_(void) { rdo_consts c; rdo_get_consts(&c); return c.a_FN; }
static double b_FN_(void) { rdo_consts c; rdo_get_consts(&c); return c.b_FN; }
static double l_const_(void) { rdo_consts c; rdo_get_consts(&c); return c.l_const; }

static double fn_v(int ic, double lc, double F, double w) {
  if (!ic) return 1.0;
  double l = lc * (-1.0 * F) / (w * w);
  if (l > 1.0) l = 1.0;
  return 1.0 - l + 1.0 / 6.0 * l * rdo_log(l);
}
static double fn_t(int ic, double lc, double F, double w) {
  if (!ic) return 1.0;
  double l = lc * (-1.0 * F) / (w * w);
  if (l > 1.0) l = 1.0;
  return 1.0 + l * (1.0 / 9.0 - 1.0 / 18.0 * rdo_log(l));
}
double rdo_fn_v(int ic, double F, double w) { return fn_v(ic, l_const_(), F, w); }
double rdo_fn_t(int ic, double F, double w) { return fn_t(ic, l_const_(), F, w); }
double rdo_escape_log(int ic, double F, double w) {
  double sw = sqrt(w);
  return b_FN_() * (sw * sw * sw) * fn_v(ic, l_const_(), F, w) / (-1.0 * F);
}
double rdo_supply_log(int ic, double F, double w) {
  return 2.0 * rdo_log(-1.0 * F) - 2.0 * rdo_log(fn_t(ic, l_const_(), F, w)) - rdo_log(w);
}
double rdo_supply(int ic, double dt_q0, double F, double w) {
  double t = fn_t(ic, l_const_(), F, w);
  return dt_q0 * a_FN_() / (t * t * w) * (F * F);
}

/* ---- GTF (src/mod_kevin_rjgtf_v2.f90) ------------------------------------ */
static double gtf_Nns(double n, double s) {
  if (n == 1.0) return (s + 1.0) * rdo_exp(-s);
  double x = n * n;
  double y = 1.0 / x;
  double z = (n - 1.0) * s;
  double sng;
  if (fabs(z) > 1.0e-5)
    sng = (x + 1.0) * (x * rdo_exp(-s) - rdo_exp(-n * s)) / (x - 1.0);
  else
    sng = (0.5 * (x + 1.0) * rdo_exp(-s) / (n + 1.0)) *
          ((1.0 - n) * (s * s) + 2.0 * (1.0 + n) + 2.0 * s);
  double sn = -x * (0.10593434 * x + 0.35506593);
  double sd = -y * (0.10593434 * y + 0.35506593);
  double sfn = x * rdo_exp(-s);
  double v = sng + sn * rdo_exp(-n * s) + x * sd * rdo_exp(-s);
  return v > sfn ? v : sfn;
}
double rdo_gtf(double F, double To, double Phi) {
  const double pi = 3.14159265358979324, kb = 1.0 / 11604.50635, hbar = 0.6582119571;
  const double c = 299.7924580, mo = 5.685630103, afs = 1.0 / 137.035999084;
  const double Qo = afs * hbar * c / 4.0, cm = 1.0e7, Amp = 6.241509074e3;
  const double Arld = (mo * (kb * kb) / (2.0 * (pi * pi) * (hbar * hbar * hbar))) * (cm * cm) / Amp;
  double Fo = fabs(F) * 1.0e-9;
  if (Fo < 1.0e-9) return 0.0;
  const double chem = 7.0;
  double s4 = sqrt(4.0 * Qo * Fo);
  double yo = s4 / Phi;
  double phix = Phi - s4;
  double ly = rdo_log(yo);
  double ty = 1.0 + (yo * yo) * (1.0 - ly) / 9.0;
  double vy = 1.0 - (yo * yo) * (3.0 - ly) / 3.0;
  double Tmin = (hbar * Fo / (4.0 * kb * ty)) * sqrt(2.0 / (mo * Phi));
  double Tmax = hbar * Fo / (kb * pi * sqrt(mo * Phi * yo));
  double betaT = 1.0 / (kb * To);
  double betau = (2.0 / (hbar * Fo)) * sqrt(2.0 * mo * Phi) * ty;
  double betap = (pi / (hbar * Fo)) * sqrt(mo * Phi * yo);
  double theto = (4.0 * sqrt(2.0 * mo * (Phi * Phi * Phi)) / (3.0 * hbar * Fo)) * vy;
  double nft, sft;
  if (To < Tmin) {
    nft = betaT / betau;
    sft = theto;
  } else if (To > Tmax) {
    nft = betaT / betap;
    sft = betap * phix;
  } else {
    double Ap = 3.0 * (betap + betau) - 6.0 * theto / phix;
    double Bp = -2.0 * (betap + 2.0 * betau) + 6.0 * theto / phix;
    double Cp = betau - betaT;
    double po = (-Bp - sqrt(Bp * Bp - 4.0 * Ap * Cp)) / (2.0 * Ap);
    double Em = chem + po * phix;
    double theta = ((1.0 - po) * (1.0 - po)) * (2.0 * po + 1.0) * theto -
                   phix * po * (1.0 - po) * ((1.0 - po) * betau - po * betap);
    nft = 1.0;
    sft = theta + betaT * (Em - chem);
  }
  double J = Arld * gtf_Nns(nft, sft) * (To * To);
  return J * 1.0e4;
}

/* ---- checkerboard (w_theta_checkerboard) ---------------------------------- */
double rdo_work_checker(const double* table, int nx, int ny, double ex0, double ey0, double elx,
                        double ely, double x, double y, int* sec) {
  double xs = (x - ex0) / elx;
  double ys = (y - ey0) / ely;
  double x_len = 1.0 / nx, y_len = 1.0 / ny;
  int xi = (int)floor(xs / x_len) + 1;
  int yi = (int)floor(ys / y_len) + 1;
  if (xi > nx) xi = nx;
  else if (xi < 1) xi = 1;
  if (yi > ny) yi = ny;
  else if (yi < 1) yi = 1;
  *sec = nx * (yi - 1) + xi;
  int yr = ny - yi + 1;
  return table[(yr - 1) * nx + (xi - 1)];
}

/* ---- tip surface (mod_hyperboloid_tip.f90:25-163) ------------------------- */
static void tip_xyz(const rdo_geom* g, double xi, double eta, double phi, double* x, double* y,
                    double* z) {
  double xy = g->a_foci * sqrt((xi * xi - 1.0) * (1.0 - eta * eta));
  double s, c;
  rdo_sincos(phi, &s, &c);
  *x = xy * c;
  *y = xy * s;
  *z = g->a_foci * xi * eta + g->shift_z;
}
static void tip_normal(const rdo_geom* g, double x, double y, double* nx, double* ny, double* nz) {
  double e1 = g->eta_1;
  double eta_fac = e1 / sqrt(1 - e1 * e1);
  double div_fac = -1.0 / sqrt(x * x + y * y + g->a_foci * g->a_foci * (1 - e1 * e1));
  double ux = eta_fac * x * div_fac, uy = eta_fac * y * div_fac, uz = 1.0;
  double nrm = sqrt(ux * ux + uy * uy + uz * uz);
  *nx = ux / nrm;
  *ny = uy / nrm;
  *nz = uz / nrm;
}
static double tip_area(const rdo_geom* g, double xi1, double xi2, double phi1, double phi2) {
  double e2 = g->eta_1 * g->eta_1;
  double s1 = sqrt(xi1 * xi1 - e2), s2 = sqrt(xi2 * xi2 - e2);
  double f1 = xi1 * s1 - e2 * rdo_log(xi1 + s1);
  double f2 = xi2 * s2 - e2 * rdo_log(xi2 + s2);
  return 0.5 * (g->a_foci * g->a_foci) * sqrt(1.0 - e2) * (phi2 - phi1) * (f2 - f1);
}
static double tip_eta(const rdo_geom* g, double x, double y, double z) {
  double a = g->a_foci, sz = g->shift_z;
  return 1.0 / (2.0 * a) * (sqrt(x * x + y * y + (z + a - sz) * (z + a - sz)) -
                            sqrt(x * x + y * y + (z - a - sz) * (z - a - sz)));
}
static double field_normal(const rdo_geom* g, double x, double y, const double* E) {
  double nx, ny, nz;
  tip_normal(g, x, y, &nx, &ny, &nz);
  return nx * E[0] + ny * E[1] + nz * E[2];
}

/* Public entry points of the tip surface functions, for the tests that pin
 * them against the reference's own (xyz_corr, surface_normal, Tip_Area,
 * eta_coor, Field_normal). */
void rdo_tip_xyz(const rdo_geom* g, double xi, double eta, double phi, double* out) {
  tip_xyz(g, xi, eta, phi, &out[0], &out[1], &out[2]);
}
void rdo_tip_normal(const rdo_geom* g, const double* p, double* out) {
  tip_normal(g, p[0], p[1], &out[0], &out[1], &out[2]);
}
double rdo_tip_area(const rdo_geom* g, double xi1, double xi2, double phi1, double phi2) {
  return tip_area(g, xi1, xi2, phi1, phi2);
}
double rdo_tip_eta(const rdo_geom* g, const double* p) { return tip_eta(g, p[0], p[1], p[2]); }
double rdo_field_normal(const rdo_geom* g, const double* p, const double* E) {
  return field_normal(g, p[0], p[1], E);
}

/* Get_MB_Velocity (src/mod_velocity.f90:53-68): two Box-Muller pairs of
 * standard deviation sqrt(k_b T / m_0); x from the first pair, y and z from
 * the second (the reference overwrites element 2), z made positive.  The
 * draws of candidate k of a step. */
static void mb_velocity(uint64_t seed, uint64_t stream, int step, int k, double T, double* v) {
  rdo_consts cc;
  rdo_get_consts(&cc);
  double sd = sqrt(cc.k_b * T / cc.m_0);
  double b1, b2, b3, b4;
  box_muller(seed, stream, (uint64_t)step, tagv(T_VEL, 0), (uint64_t)k, 0, 0, sd, sd, &b1, &b2);
  box_muller(seed, stream, (uint64_t)step, tagv(T_VEL, 1), (uint64_t)k, 0, 0, sd, sd, &b3, &b4);
  v[0] = b1;
  v[1] = b3;
  v[2] = fabs(b4);
}
void rdo_mb_velocity(uint64_t seed, uint64_t stream, int step, double T, int n, double* out) {
  for (int k = 0; k < n; ++k) mb_velocity(seed, stream, step, k, T, &out[3 * k]);
}

/* ==================================================================== */
typedef struct {
  rdo_geom g;
  double box_top, dt, dt2, dt_q0;
  int mode, delay;
  double ex0, ey0, elx, ely, T, w_tip;
  int wf_nx, wf_ny;
  double wf[96 * 96];
  int quad, ndim;
  int qgx, qgy;  /* level-0 points per side of the planar supply (x, y) */
  int quad_lmax; /* refinement doublings of the planar supply integral */
  int tts;       /* two-time-step atom_time_interval; 0: the standard integrator */
  uint64_t seed, stream;
  int cap, n, nrID;
  double *pos, *prev, *vel, *acc, *a1, *a2, *q, *m;
  int *species, *pstep, *emitter, *section, *id, *alive;
  double residual, MH_std, a_rate, N_sup, df_avg, F_avg[3], ramo[3], vsum[9];
  int emitted, n_round, rem_total, rem_top, rem_bot, rem_elec, rem_ion, rem_top_emit, dropped;
  int start_rounds, n_elec;
  int supply_level, supply_cells, supply_fail;
  double supply_err;
  int life[2][NLIFE];
  int nthreads;
  int serial; /* 1: the reference's serial samplers (emit_serial) */
  int emit_step; /* step of the last emission phase */
  /* particle records of the last step (the density_emit / density_absorb
   * writes of Add_Particle and Mark_Particles_Remove) */
  struct rdo_emit_rec* elog;
  struct rdo_abs_rec* alog;
  int n_elog, n_alog;
  /* recording planes (planes_N, planes_z) and the last step's crossings,
   * cap records per plane */
  int nplanes;
  double planes[10];
  struct rdo_plane_rec* plog;
  int n_plog[10];
} rdo_sim;

/* Same layouts as rd_emit_record / rd_absorb_record (include/rumdeed.h). */
typedef struct rdo_emit_rec {
  double x, y, z;
  int32_t emitter, section, id, species;
} rdo_emit_rec;
typedef struct rdo_abs_rec {
  double x, y, z, vx, vy, vz;
  int32_t emitter, section, id, where;
} rdo_abs_rec;
typedef struct rdo_plane_rec {
  double x, y, vx, vy, vz;
  int32_t emitter, section, id, plane;
} rdo_plane_rec;

/* Per-step record, same layout as rd_step_record (include/rumdeed.h). */
typedef struct {
  int32_t step, n, emitted, n_round;
  int32_t absorbed, absorbed_top, absorbed_bot, dropped;
  int32_t start_rounds, supply_level, supply_cells, supply_fail;
  double n_sup, ramo_current, mh_std, a_rate, df_avg;
  double f_avg[3];
  double ramo[3];
  double avg_elec_vel[3];
  double supply_err;
} rdo_record;

int rdo_quad_lmax(int g0, int levels);
int rdo_quad_side(int quad, int sections);

rdo_sim* rdo_sim_create(int cap, int nthreads) {
  rdo_sim* s = (rdo_sim*)calloc(1, sizeof(rdo_sim));
  s->emit_step = -1;
  s->cap = cap;
  s->nthreads = nthreads < 1 ? 1 : nthreads;
  s->pos = calloc(3 * (size_t)cap, 8);
  s->prev = calloc(3 * (size_t)cap, 8);
  s->vel = calloc(3 * (size_t)cap, 8);
  s->acc = calloc(3 * (size_t)cap, 8);
  s->a1 = calloc(3 * (size_t)cap, 8);
  s->a2 = calloc(3 * (size_t)cap, 8);
  s->q = calloc((size_t)cap, 8);
  s->m = calloc((size_t)cap, 8);
  s->species = calloc((size_t)cap, 4);
  s->pstep = calloc((size_t)cap, 4);
  s->emitter = calloc((size_t)cap, 4);
  s->section = calloc((size_t)cap, 4);
  s->id = calloc((size_t)cap, 4);
  s->alive = calloc((size_t)cap, 4);
  s->elog = calloc((size_t)cap, sizeof(rdo_emit_rec));
  s->alog = calloc((size_t)cap, sizeof(rdo_abs_rec));
  s->plog = calloc(10 * (size_t)cap, sizeof(rdo_plane_rec));
  s->MH_std = 0.0125;
  s->a_rate = 1.0;
  s->quad = 32;
  s->qgx = s->qgy = 32;
  s->quad_lmax = rdo_quad_lmax(32, 0);
  s->ndim = 200;
  s->w_tip = 4.7;
  return s;
}

void rdo_sim_destroy(rdo_sim* s) {
  if (!s) return;
  free(s->pos); free(s->prev); free(s->vel); free(s->acc); free(s->a1); free(s->a2);
  free(s->q); free(s->m); free(s->species); free(s->pstep); free(s->emitter);
  free(s->section); free(s->id); free(s->alive); free(s->elog); free(s->alog);
  free(s->plog);
  free(s);
}

void rdo_sim_set_geometry(rdo_sim* s, const rdo_geom* g, double box_top) {
  s->g = *g;
  s->box_top = box_top > 0.0 ? box_top : g->d;
}
void rdo_sim_set_time_step(rdo_sim* s, double dt) {
  rdo_consts c;
  rdo_get_consts(&c);
  s->dt = dt;
  s->dt2 = dt * dt;
  s->dt_q0 = dt / c.q_0;
}
void rdo_sim_set_seed(rdo_sim* s, uint64_t seed, uint64_t stream) {
  s->seed = seed;
  s->stream = stream;
}
void rdo_sim_set_emitter(rdo_sim* s, int mode, int delay, double x0, double y0, double lx,
                         double ly, double T, double w_tip, int nx, int ny, const double* table,
                         int quad, int ndim) {
  s->mode = mode;
  s->delay = delay;
  s->ex0 = x0; s->ey0 = y0; s->elx = lx; s->ely = ly;
  s->T = T;
  s->w_tip = w_tip;
  s->wf_nx = nx;
  s->wf_ny = ny;
  if (table && nx > 0 && ny > 0) memcpy(s->wf, table, sizeof(double) * nx * ny);
  s->quad = quad > 0 ? quad : 32;
  s->qgx = rdo_quad_side(s->quad, nx);
  s->qgy = rdo_quad_side(s->quad, ny);
  s->quad_lmax = rdo_quad_lmax(s->qgx > s->qgy ? s->qgx : s->qgy, 0);
  int ndef = mode == EM_GTF ? 25 : (mode == EM_TIP ? 80 : 200);
  s->ndim = ndim > 0 ? ndim : ndef;
  if (mode == EM_GTF) { s->MH_std = 0.125; s->a_rate = 1.0; }
  else if (mode == EM_TIP) { s->MH_std = 1.0; s->a_rate = 0.5; }
  else { s->MH_std = 0.0125; s->a_rate = 1.0; }
  s->residual = 0.0;
}

/* Level-0 points per side of the planar supply along an axis the work
 * function splits into `sections` equal parts: the least multiple of
 * `sections` not below quad, so that every cell of the grid (and of each
 * refinement) lies in one section and the integrand is smooth on it. */
int rdo_quad_side(int quad, int sections) {
  if (sections < 1) sections = 1;
  return sections * ((quad + sections - 1) / sections);
}
/* Doublings of the planar supply rule from g0 points on the longer side:
 * levels < 0 none, 0 four, else `levels`; never beyond 1024 points per side. */
int rdo_quad_lmax(int g0, int levels) {
  int l = levels < 0 ? 0 : (levels == 0 ? 4 : levels);
  while (l > 0 && (g0 << l) > 1024) --l;
  return l;
}
void rdo_sim_set_two_time_step(rdo_sim* s, int K) { s->tts = K > 0 ? K : 0; }
void rdo_sim_set_quad_levels(rdo_sim* s, int levels) {
  s->quad_lmax = rdo_quad_lmax(s->qgx > s->qgy ? s->qgx : s->qgy, levels);
}

static void vacuum(const rdo_sim* s, const double* p, double* e) { rdo_vacuum_field(&s->g, p, e); }

/* Add_Particle (src/mod_pair.F90:29-159): append with the Beeman seed. */
static void add_particle(rdo_sim* s, const double* p, const double* v, int spc, int step,
                         int emit, int sec) {
  rdo_consts c;
  rdo_get_consts(&c);
  if (s->n + 1 > s->cap) {
    s->dropped++;
    return;
  }
  int i = s->n;
  for (int k = 0; k < 3; ++k) {
    s->pos[3 * i + k] = p[k];
    s->prev[3 * i + k] = -1.0 * LS;
    s->vel[3 * i + k] = v[k];
  }
  double qq = spc == 2 ? 1.0 * c.q_0 : -1.0 * c.q_0;
  double mm = spc == 2 ? c.m_N2p : 1.0 * c.m_0;
  double e[3];
  vacuum(s, p, e);
  double qm = qq / mm;
  for (int k = 0; k < 3; ++k) s->acc[3 * i + k] = s->a1[3 * i + k] = s->a2[3 * i + k] = qm * e[k];
  s->q[i] = qq;
  s->m[i] = mm;
  s->species[i] = spc;
  s->pstep[i] = step;
  s->emitter[i] = emit;
  s->section[i] = sec;
  s->id[i] = s->nrID;
  s->alive[i] = 1;
  s->n++;
  s->nrID++;
}

void rdo_sim_add(rdo_sim* s, int n, const double* pos, const double* vel, const int* species,
                 int step, int emit, const int* sec) {
  for (int k = 0; k < n; ++k) {
    double z[3] = {0, 0, 0};
    add_particle(s, &pos[3 * k], vel ? &vel[3 * k] : z, species ? species[k] : 1, step, emit,
                 sec ? sec[k] : 1);
  }
}

void rdo_sim_put(rdo_sim* s, int n, const double* pos, const double* vel, const double* q,
                 const double* m, const int* species) {
  s->n = n;
  s->nrID = n;
  memcpy(s->pos, pos, sizeof(double) * 3 * n);
  if (vel) memcpy(s->vel, vel, sizeof(double) * 3 * n);
  else memset(s->vel, 0, sizeof(double) * 3 * n);
  memset(s->prev, 0, sizeof(double) * 3 * n);
  memset(s->acc, 0, sizeof(double) * 3 * n);
  memset(s->a1, 0, sizeof(double) * 3 * n);
  memset(s->a2, 0, sizeof(double) * 3 * n);
  memcpy(s->q, q, sizeof(double) * n);
  memcpy(s->m, m, sizeof(double) * n);
  for (int i = 0; i < n; ++i) {
    s->species[i] = species ? species[i] : 1;
    s->pstep[i] = 0;
    s->emitter[i] = 0;
    s->section[i] = 1;
    s->id[i] = i;
    s->alive[i] = 1;
  }
}

/* ---- emission ------------------------------------------------------------ */
static double work_at(const rdo_sim* s, double x, double y, int* sec) {
  return rdo_work_checker(s->wf, s->wf_nx, s->wf_ny, s->ex0, s->ey0, s->elx, s->ely, x, y, sec);
}

static double chain_target(const rdo_sim* s, double F, double a, double b) {
  if (s->mode == EM_FE) {
    int sec;
    return rdo_supply_log(s->g.image_charge, F, work_at(s, a, b, &sec));
  }
  if (s->mode == EM_GTF) {
    int sec;
    double J = rdo_gtf(F, s->T, work_at(s, a, b, &sec));
    return rdo_log(J > KTINY ? J : KTINY);
  }
  double t = fn_t(s->g.image_charge, l_const_(), F, s->w_tip);
  double sup = s->dt_q0 * a_FN_() / (t * t * s->w_tip) * (F * F);
  return rdo_log(sup > KTINY ? sup : KTINY) + 0.5 * rdo_log(a * a - s->g.eta_1 * s->g.eta_1);
}

static void field_at(const rdo_sim* s, int m, const double* pts, double* out) {
  rdo_field(&s->g, s->n, s->pos, s->q, m, pts, out, s->nthreads);
}

static void add_particle(rdo_sim* s, const double* p, const double* v, int spc, int step,
                         int emit, int sec);
static void emit_serial(rdo_sim* s, int step, int nr);

/* One level of the supply integral: the integrand at the Gx x Gy cell
 * centres of the planar emitter, x fastest (Gx = 0: the tip's 100 x 100
 * (xi, phi) cells) and its sums -- value, f_x, f_y, f_z (tip: the normal
 * field) -- in the device's order: tiles of 1024 points, a pairwise tree
 * within a tile (zeros beyond the last point), tile sums added in tile
 * order. */
static void supply_level(rdo_sim* s, int Gx, int Gy, double tot[4]) {
  const int tip = Gx == 0;
  int M = tip ? 100 * 100 : Gx * Gy;
  double* pts = calloc((size_t)3 * M, sizeof(double));
  double* fld = malloc(sizeof(double) * 3 * M);
  for (int k = 0; k < M; ++k) {
    if (tip) {
      const int nr = 100;
      int i = k / nr + 1, j = k % nr + 1;
      double len_xi = (s->g.max_xi - 1.0) / nr, len_phi = 2.0 * PI_D / nr;
      tip_xyz(&s->g, 1.0 + (i - 0.5) * len_xi, s->g.eta_1, (j - 0.5) * len_phi, &pts[3 * k],
              &pts[3 * k + 1], &pts[3 * k + 2]);
    } else {
      int ix = k % Gx, iy = k / Gx;
      double u = (ix + 0.5) / Gx, v = (iy + 0.5) / Gy;
      pts[3 * k] = s->ex0 + u * s->elx;
      pts[3 * k + 1] = s->ey0 + v * s->ely;
      pts[3 * k + 2] = 0.0;
    }
  }
  field_at(s, M, pts, fld);
  double* rv = malloc(sizeof(double) * 4 * M); /* val, fx, fy, F (or fz) per point */
  const int ic = s->g.image_charge;
  for (int k = 0; k < M; ++k) {
    double val = 0.0;
    if (tip) {
      double F = field_normal(&s->g, pts[3 * k], pts[3 * k + 1], &fld[3 * k]);
      if (F < 0.0) {
        const int nr = 100;
        int i = k / nr + 1, j = k % nr + 1;
        double len_xi = (s->g.max_xi - 1.0) / nr, len_phi = 2.0 * PI_D / nr;
        double A_f = tip_area(&s->g, 1.0 + (i - 1.0) * len_xi, 1.0 + (i + 0.0) * len_xi,
                              (j - 1.0) * len_phi, (j + 0.0) * len_phi);
        double t = fn_t(ic, l_const_(), F, s->w_tip);
        rdo_consts c;
        rdo_get_consts(&c);
        val = A_f * a_FN_() * (F * F) * s->dt / (c.q_0 * s->w_tip * (t * t));
      }
      rv[4 * k + 1] = 0.0;
      rv[4 * k + 2] = 0.0;
      rv[4 * k + 3] = F;
    } else {
      double F = fld[3 * k + 2];
      if (F < 0.0) {
        int sec;
        double w = work_at(s, pts[3 * k], pts[3 * k + 1], &sec);
        if (s->mode == EM_FE) val = rdo_supply(ic, s->dt_q0, F, w);
        else val = rdo_gtf(F, s->T, w) * s->dt_q0;
      }
      val = (s->elx * s->ely) * val;
      rv[4 * k + 1] = fld[3 * k];
      rv[4 * k + 2] = fld[3 * k + 1];
      rv[4 * k + 3] = fld[3 * k + 2];
    }
    rv[4 * k] = val;
  }
  /* the device's order: tiles of 1024 points, a pairwise tree within a tile
     (zeros beyond the last point), tile sums added in tile order */
  tot[0] = tot[1] = tot[2] = tot[3] = 0.0;
  for (int b = 0; b < M; b += 1024) {
    double t[4][1024];
    for (int q = 0; q < 4; ++q)
      for (int j = 0; j < 1024; ++j) t[q][j] = (b + j < M) ? rv[4 * (b + j) + q] : 0.0;
    for (int w = 512; w > 0; w >>= 1)
      for (int q = 0; q < 4; ++q)
        for (int j = 0; j < w; ++j) t[q][j] = t[q][j] + t[q][j + w];
    for (int q = 0; q < 4; ++q) tot[q] = tot[q] + t[q][0];
  }
  free(rv);
  free(pts);
  free(fld);
}

/* The planar supply: the midpoint rule on qgx x qgy points (cells within the
 * work function's sections), then on twice as many per side, and so on up to
 * quad_lmax doublings, until two
 * successive levels agree within the reference's Divonne tolerances
 * (src/mod_global.F90:89-90, the test of src/mod_cuba_integration.F90:87:
 * |difference| <= max(0.5 electrons, 1e-3 |integral|)).  The last
 * level is the result; *M receives its number of points. */
static void supply_planar(rdo_sim* s, double tot[4], int* Mout) {
  s->supply_level = 0;
  s->supply_fail = 0;
  s->supply_err = 0.0;
  int Gx = s->qgx, Gy = s->qgy;
  supply_level(s, Gx, Gy, tot);
  double prev = tot[0] / ((double)Gx * (double)Gy);
  for (int L = 1; L <= s->quad_lmax; ++L) {
    const int G2x = s->qgx << L, G2y = s->qgy << L;
    double t2[4];
    supply_level(s, G2x, G2y, t2);
    const double cur = t2[0] / ((double)G2x * (double)G2y);
    double tol = 1.0e-3 * fabs(cur);
    if (tol < 0.5) tol = 0.5;
    const double err = fabs(cur - prev);
    const int conv = err <= tol;
    for (int k = 0; k < 4; ++k) tot[k] = t2[k];
    Gx = G2x;
    Gy = G2y;
    prev = cur;
    s->supply_level = L;
    s->supply_err = err;
    s->supply_fail = !conv; /* final unless converged earlier: only the last level's counts */
    if (conv) break;
  }
  s->supply_cells = Gx * Gy;
  *Mout = Gx * Gy;
}

static void do_emission(rdo_sim* s, int step) {
  if (s->mode == EM_NONE || !(s->delay < step)) return;
  const int tip = s->mode == EM_TIP;
  /* supply integral: the tip's 100 x 100 cells; the planar emitter's
     midpoint rule refined (supply_planar) */
  double tot[4];
  int M;
  if (tip) {
    M = 100 * 100;
    supply_level(s, 0, 0, tot);
    s->supply_cells = M;
  } else {
    supply_planar(s, tot, &M);
  }
  const double sum = tot[0], f0 = tot[1], f1 = tot[2], f2 = tot[3];
  const int ic = s->g.image_charge;
  int nr;
  if (tip) {
    s->N_sup = sum;
    nr = fnint(sum);
    if (nr < 0) nr = 0;
    s->F_avg[0] = 0.0; s->F_avg[1] = 0.0; s->F_avg[2] = f2 / M;
  } else {
    double nsup = sum / (double)M;
    s->N_sup = nsup;
    if (s->mode == EM_FE) {
      nr = fnint(nsup + s->residual);
      s->residual = nsup - nr;
      if (nr < 0) nr = 0;
    } else {
      nr = poisson(s->seed, s->stream, (uint64_t)step, nsup);
    }
    s->F_avg[0] = f0 / M; s->F_avg[1] = f1 / M; s->F_avg[2] = f2 / M;
  }
  s->n_round = nr;
  if (nr <= 0) return;
  if (s->serial) {
    emit_serial(s, step, nr);
    return;
  }

  /* chains */
  const int C = nr;
  double *cx = malloc(8 * C), *cy = malloc(8 * C), *px = malloc(8 * C), *py = malloc(8 * C);
  double *F = malloc(8 * C), *sup = malloc(8 * C), *D = malloc(8 * C);
  double *o = malloc(24 * (size_t)C);
  int *ok = calloc(C, 4), *acc = calloc(C, 4), *rej = calloc(C, 4), *pv = malloc(4 * C);
  double* cp = malloc(24 * (size_t)C);
  double* cf = malloc(24 * (size_t)C);
  for (int k = 0; k < C; ++k) pv[k] = 1;
  if (tip) {
    if (s->MH_std > 0.125) s->MH_std = 0.125;
    else if (s->MH_std < 0.0005) s->MH_std = 0.0005;
  }
  for (int round = 1;; ++round) {
    const int last = round >= 10001;
    for (int k = 0; k < C; ++k) {
      if (ok[k]) {
        cp[3 * k] = o[3 * k]; cp[3 * k + 1] = o[3 * k + 1]; cp[3 * k + 2] = o[3 * k + 2];
        continue;
      }
      uint64_t r[4];
      ph((uint64_t)step, tagv(T_START, (uint64_t)round), (uint64_t)k, 0, s->seed, s->stream, r);
      double u1 = u01(r[0]), u2 = u01(r[1]);
      if (tip) {
        px[k] = 1.0 + (s->g.max_xi - 1.0) * u1;
        py[k] = 2.0 * PI_D * u2;
        tip_xyz(&s->g, px[k], s->g.eta_1, py[k], &cp[3 * k], &cp[3 * k + 1], &cp[3 * k + 2]);
      } else {
        px[k] = u1 * s->elx + s->ex0;
        py[k] = u2 * s->ely + s->ey0;
        cp[3 * k] = px[k]; cp[3 * k + 1] = py[k]; cp[3 * k + 2] = 0.0;
      }
    }
    field_at(s, C, cp, cf);
    int pend = 0;
    for (int k = 0; k < C; ++k) {
      if (ok[k]) continue;
      double Fk = tip ? field_normal(&s->g, cp[3 * k], cp[3 * k + 1], &cf[3 * k]) : cf[3 * k + 2];
      if (Fk < 0.0) {
        ok[k] = 1;
        cx[k] = px[k];
        cy[k] = py[k];
        F[k] = Fk;
        sup[k] = chain_target(s, Fk, px[k], py[k]);
        o[3 * k] = cp[3 * k]; o[3 * k + 1] = cp[3 * k + 1]; o[3 * k + 2] = cp[3 * k + 2];
      } else if (last) {
        ok[k] = -1;
        F[k] = 1.0;
        sup[k] = -KHUGE;
        if (tip) {
          cx[k] = 1.0; cy[k] = 0.0;
          tip_xyz(&s->g, 1.0, s->g.eta_1, 0.0, &o[3 * k], &o[3 * k + 1], &o[3 * k + 2]);
        } else {
          cx[k] = s->ex0; cy[k] = s->ey0;
          o[3 * k] = s->ex0; o[3 * k + 1] = s->ey0; o[3 * k + 2] = 0.0;
        }
      } else {
        ++pend;
      }
    }
    s->start_rounds = round;
    if (pend == 0 || last) break;
  }
  const int first = fnint(s->ndim * 0.25);
  for (int it = 1; it <= s->ndim; ++it) {
    const double MH = s->MH_std;
    const uint64_t tg = tagv(T_JUMP, (uint64_t)it);
    for (int k = 0; k < C; ++k) {
      if (ok[k] != 1) {
        cp[3 * k] = o[3 * k]; cp[3 * k + 1] = o[3 * k + 1]; cp[3 * k + 2] = o[3 * k + 2];
        continue;
      }
      if (tip) {
        double s1, s2;
        if (it > first) { s1 = (s->g.max_xi - 1.0) * MH; s2 = 2.0 * PI_D * MH; }
        else { s1 = (s->g.max_xi - 1.0) * 0.10; s2 = 2.0 * PI_D * 0.10; }
        double d1, d2;
        box_muller(s->seed, s->stream, (uint64_t)step, tg, (uint64_t)k, 0.0, 0.0, s1, s2, &d1, &d2);
        double nxi = cx[k] + d1;
        double nphi = fmodulo(cy[k] + d2, 2.0 * PI_D);
        if (nxi > s->g.max_xi) nxi = 2.0 * s->g.max_xi - nxi;
        if (nxi < 1.0) nxi = 2.0 - nxi;
        pv[k] = !((nxi < 1.0) || (nxi > s->g.max_xi));
        px[k] = nxi;
        py[k] = nphi;
        if (pv[k]) tip_xyz(&s->g, nxi, s->g.eta_1, nphi, &cp[3 * k], &cp[3 * k + 1], &cp[3 * k + 2]);
        else { cp[3 * k] = o[3 * k]; cp[3 * k + 1] = o[3 * k + 1]; cp[3 * k + 2] = o[3 * k + 2]; }
        continue;
      }
      double sx, sy;
      if (s->mode == EM_FE) {
        double f = (it > first) ? MH : 0.10;
        sx = s->elx * f;
        sy = s->ely * f;
      } else {
        sx = s->elx * MH;
        sy = s->ely * MH;
      }
      double nx, ny;
      box_muller(s->seed, s->stream, (uint64_t)step, tg, (uint64_t)k, cx[k], cy[k], sx, sy, &nx, &ny);
      if (s->mode == EM_FE) {
        double x_max = s->ex0 + s->elx, x_min = s->ex0, y_max = s->ey0 + s->ely, y_min = s->ey0;
        if (nx > x_max) nx = x_max - (nx - x_max);
        else if (nx < x_min) nx = (x_min - nx) + x_min;
        if (ny > y_max) ny = y_max - (ny - y_max);
        else if (ny < y_min) ny = (y_min - ny) + y_min;
      } else {
        double sxs = (nx - s->ex0) / s->elx, sys = (ny - s->ey0) / s->ely;
        if ((sxs > 1.0) || (sxs < 0.0)) sxs = 1.0 - (sxs - floor(sxs));
        if ((sys > 1.0) || (sys < 0.0)) sys = 1.0 - (sys - floor(sys));
        nx = sxs * s->elx + s->ex0;
        ny = sys * s->ely + s->ey0;
      }
      px[k] = nx;
      py[k] = ny;
      cp[3 * k] = nx; cp[3 * k + 1] = ny; cp[3 * k + 2] = 0.0;
    }
    field_at(s, C, cp, cf);
    int na = 0, nrj = 0;
    for (int k = 0; k < C; ++k) {
      if (ok[k] != 1) continue;
      int accd = 0, rejd = 0;
      if (tip && !pv[k]) {
        rejd = 1;
      } else {
        double Fk = tip ? field_normal(&s->g, cp[3 * k], cp[3 * k + 1], &cf[3 * k]) : cf[3 * k + 2];
        int bad = (s->mode == EM_GTF) ? (Fk > 0.0) : (Fk >= 0.0);
        if (bad) {
          rejd = 1;
        } else {
          double sn = chain_target(s, Fk, px[k], py[k]);
          double alpha = sn - sup[k];
          if (sn >= sup[k]) accd = 1;
          else {
            double rnd = uniform(s->seed, s->stream, (uint64_t)step, tagv(T_ACCEPT, (uint64_t)it),
                                 (uint64_t)k, 0);
            accd = rdo_log(rnd) <= alpha;
          }
          if (accd) {
            cx[k] = px[k]; cy[k] = py[k]; sup[k] = sn; F[k] = Fk;
            o[3 * k] = cp[3 * k]; o[3 * k + 1] = cp[3 * k + 1]; o[3 * k + 2] = cp[3 * k + 2];
          } else {
            rejd = 1;
          }
        }
      }
      na += accd;
      nrj += rejd;
      if (s->mode == EM_GTF || (tip && it > first)) {
        acc[k] += accd;
        rej[k] += rejd;
      }
    }
    if (s->mode == EM_FE && it > first && na + nrj > 0) {
      s->a_rate = (double)na / (double)(na + nrj);
      s->MH_std = s->MH_std * rdo_exp(0.025 * (s->a_rate - 0.35));
      if (s->MH_std > 0.1250) s->MH_std = 0.1250;
      else if (s->MH_std < 0.00005) s->MH_std = 0.00005;
    }
  }
  if (s->mode != EM_FE) {
    const double lo = tip ? 0.0005 : 0.005;
    for (int k = 0; k < C; ++k) {
      if (ok[k] != 1) continue;
      if (acc[k] + rej[k] > 0) {
        s->a_rate = (double)acc[k] / (double)(rej[k] + acc[k]);
        s->MH_std = s->MH_std * rdo_exp(0.025 * (s->a_rate - 0.35));
        if (s->MH_std > 0.1250) s->MH_std = 0.1250;
        else if (s->MH_std < lo) s->MH_std = lo;
      }
    }
  }
  for (int k = 0; k < C; ++k) {
    if (s->mode == EM_FE) {
      if (ok[k] == 1) {
        int sec;
        D[k] = rdo_escape_log(ic, F[k], work_at(s, cx[k], cy[k], &sec));
      } else {
        D[k] = -KHUGE;
      }
    } else if (tip) {
      if (ok[k] == 1) {
        double sw = sqrt(s->w_tip);
        D[k] = rdo_exp(b_FN_() * (sw * sw * sw) * fn_v(ic, l_const_(), F[k], s->w_tip) / fabs(F[k]));
      } else {
        D[k] = 0.0;
      }
    } else {
      D[k] = 0.0;
    }
  }
  /* escape test and Add_Particle in chain order */
  double dsum_blocks[256];
  for (int t = 0; t < 256; ++t) dsum_blocks[t] = 0.0;
  int emitted = 0;
  rdo_consts cc;
  rdo_get_consts(&cc);
  for (int k = 0; k < C; ++k) {
    double p[3], v[3] = {0, 0, 0};
    int sec = 1, flag = 0;
    if (s->mode == EM_FE) {
      double Df = D[k];
      if (F[k] >= 0.0) Df = -KHUGE;
      dsum_blocks[k % 256] = dsum_blocks[k % 256] + rdo_exp(Df);
      double rnd = uniform(s->seed, s->stream, (uint64_t)step, tagv(T_EMIT, 0), (uint64_t)k, 0);
      if (rdo_log(rnd) <= Df) {
        flag = 1;
        p[0] = cx[k]; p[1] = cy[k]; p[2] = 1.0 * LS;
        work_at(s, p[0], p[1], &sec);
      }
    } else if (s->mode == EM_GTF) {
      if (ok[k] == 1) {
        flag = 1;
        p[0] = cx[k]; p[1] = cy[k]; p[2] = 1.0 * LS;
        mb_velocity(s->seed, s->stream, step, k, s->T, v);
        work_at(s, p[0], p[1], &sec);
      }
    } else {
      double rnd = uniform(s->seed, s->stream, (uint64_t)step, tagv(T_EMIT, 0), (uint64_t)k, 0);
      if ((F[k] < 0.0) && (rnd <= D[k])) {
        flag = 1;
        double nx, ny, nz;
        tip_normal(&s->g, o[3 * k], o[3 * k + 1], &nx, &ny, &nz);
        p[0] = o[3 * k] + nx * LS;
        p[1] = o[3 * k + 1] + ny * LS;
        p[2] = o[3 * k + 2] + nz * LS;
        sec = 1;
      }
    }
    if (flag) {
      int before = s->n;
      add_particle(s, p, v, 1, step, 1, sec);
      if (s->n > before) {
        emitted++;
        /* Add_Particle's density_emit record (src/mod_pair.F90:63-95) */
        rdo_emit_rec* r = &s->elog[s->n_elog++];
        r->x = p[0];
        r->y = p[1];
        r->z = p[2];
        r->emitter = 1;
        r->section = sec;
        r->id = s->id[before];
        r->species = 1;
      }
    }
  }
  s->emitted = emitted;
  if (s->mode == EM_FE) {
    /* the device reduces the per-lane sums with a fixed tree of 256 lanes */
    for (int w = 128; w > 0; w >>= 1)
      for (int t = 0; t < w; ++t) dsum_blocks[t] = dsum_blocks[t] + dsum_blocks[t + w];
    s->df_avg = dsum_blocks[0] / s->N_sup;
  }
  free(cx); free(cy); free(px); free(py); free(F); free(sup); free(D); free(o);
  free(ok); free(acc); free(rej); free(pv); free(cp); free(cf);
}


/* ---- the reference's default serial samplers --------------------------------
 * For statistical comparison with the lockstep sampling of do_emission (the
 * reference's mh_batch semantics, which the device runs).  Each of the step's
 * nr chains runs to completion in turn against the particles present at that
 * moment -- so it sees the electrons emitted by the chains before it -- and
 * the candidate is tested and added before the next chain starts:
 *   FE:  Metropolis_Hastings_rectangle_J (src/mod_field_emission_v2.F90:1122-1265)
 *        and the emission loop of Do_Field_Emission_Planar_rectangle (:307-385);
 *   GTF: Metropolis_Hastings_rectangle_J and Do_Field_Thermo_Emission_Planar_simple
 *        (src/mod_field_thermo_emission.F90:136-364);
 *   tip: Metro_algo_tip_v3 and Do_Field_Emission_Tip_OLDCODE
 *        (src/mod_emission_tip.f90:1241-1390, :488-530).
 * The width MH_std is fixed within a chain and updated once per chain from the
 * chain's acceptance rate.  Draws use the same counters as the lockstep form
 * (chain k, jump it), so the two differ only by their semantics. */
static double field_one(const rdo_sim* s, const double* p, double* e) {
  rdo_field(&s->g, s->n, s->pos, s->q, 1, p, e, 1);
  return s->mode == EM_TIP ? field_normal(&s->g, p[0], p[1], e) : e[2];
}

static void emit_serial(rdo_sim* s, int step, int nr) {
  const int tip = s->mode == EM_TIP, gtf = s->mode == EM_GTF;
  const int ic = s->g.image_charge;
  const int first = gtf ? 0 : fnint(s->ndim * 0.25);
  const double lo = tip ? 0.0005 : (gtf ? 0.005 : 0.00005);
  rdo_consts cc;
  rdo_get_consts(&cc);
  int emitted = 0;
  double dsum = 0.0;
  for (int k = 0; k < nr; ++k) {
    if (tip) { /* Metro_algo_tip_v3 clamps the shared width before use */
      if (s->MH_std > 0.125) s->MH_std = 0.125;
      else if (s->MH_std < 0.0005) s->MH_std = 0.0005;
    }
    double cx = 0.0, cy = 0.0, p[3], e[3], F = 1.0;
    int ok = 0;
    for (int round = 1; round <= 10001 && !ok; ++round) {
      uint64_t r[4];
      ph((uint64_t)step, tagv(T_START, (uint64_t)round), (uint64_t)k, 0, s->seed, s->stream, r);
      double u1 = u01(r[0]), u2 = u01(r[1]);
      if (tip) {
        cx = 1.0 + (s->g.max_xi - 1.0) * u1;
        cy = 2.0 * PI_D * u2;
        tip_xyz(&s->g, cx, s->g.eta_1, cy, &p[0], &p[1], &p[2]);
      } else {
        cx = u1 * s->elx + s->ex0;
        cy = u2 * s->ely + s->ey0;
        p[0] = cx; p[1] = cy; p[2] = 0.0;
      }
      F = field_one(s, p, e);
      ok = F < 0.0;
    }
    if (!ok) {  /* a failed chain emits nothing */
      if (s->mode == EM_FE) dsum += 0.0;
      continue;
    }
    double sup = chain_target(s, F, cx, cy);
    const double MH = s->MH_std;
    int acc = 0, rej = 0;
    for (int it = 1; it <= s->ndim; ++it) {
      const int counted = gtf || it > first;
      const uint64_t tg = tagv(T_JUMP, (uint64_t)it);
      double nx, ny, q[3];
      if (tip) {
        double f = it > first ? MH : 0.10;
        double d1, d2;
        box_muller(s->seed, s->stream, (uint64_t)step, tg, (uint64_t)k, 0.0, 0.0,
                   (s->g.max_xi - 1.0) * f, 2.0 * PI_D * f, &d1, &d2);
        nx = cx + d1;
        ny = fmodulo(cy + d2, 2.0 * PI_D);
        if (nx > s->g.max_xi) nx = 2.0 * s->g.max_xi - nx;
        if (nx < 1.0) nx = 2.0 - nx;
        if ((nx < 1.0) || (nx > s->g.max_xi)) {
          if (counted) rej++;
          continue;
        }
        tip_xyz(&s->g, nx, s->g.eta_1, ny, &q[0], &q[1], &q[2]);
      } else {
        double f = gtf ? MH : (it > first ? MH : 0.10);
        box_muller(s->seed, s->stream, (uint64_t)step, tg, (uint64_t)k, cx, cy, s->elx * f,
                   s->ely * f, &nx, &ny);
        if (gtf) {
          double sxs = (nx - s->ex0) / s->elx, sys = (ny - s->ey0) / s->ely;
          if ((sxs > 1.0) || (sxs < 0.0)) sxs = 1.0 - (sxs - floor(sxs));
          if ((sys > 1.0) || (sys < 0.0)) sys = 1.0 - (sys - floor(sys));
          nx = sxs * s->elx + s->ex0;
          ny = sys * s->ely + s->ey0;
        } else {
          double x_max = s->ex0 + s->elx, x_min = s->ex0, y_max = s->ey0 + s->ely, y_min = s->ey0;
          if (nx > x_max) nx = x_max - (nx - x_max);
          else if (nx < x_min) nx = (x_min - nx) + x_min;
          if (ny > y_max) ny = y_max - (ny - y_max);
          else if (ny < y_min) ny = (y_min - ny) + y_min;
        }
        q[0] = nx; q[1] = ny; q[2] = 0.0;
      }
      double Fn = field_one(s, q, e);
      if (gtf ? (Fn > 0.0) : (Fn >= 0.0)) {
        if (counted) rej++;
        continue;
      }
      double sn = chain_target(s, Fn, nx, ny);
      int accd = sn >= sup;
      if (!accd) {
        double rnd = uniform(s->seed, s->stream, (uint64_t)step, tagv(T_ACCEPT, (uint64_t)it),
                             (uint64_t)k, 0);
        accd = rdo_log(rnd) <= sn - sup;
      }
      if (accd) {
        cx = nx; cy = ny; sup = sn; F = Fn;
        p[0] = q[0]; p[1] = q[1]; p[2] = q[2];
        if (counted) acc++;
      } else if (counted) {
        rej++;
      }
    }
    if (acc + rej > 0) {
      s->a_rate = (double)acc / (double)(acc + rej);
      s->MH_std = s->MH_std * rdo_exp(0.025 * (s->a_rate - 0.35));
      if (s->MH_std > 0.1250) s->MH_std = 0.1250;
      else if (s->MH_std < lo) s->MH_std = lo;
    }
    double pe[3], v[3] = {0.0, 0.0, 0.0};
    int sec = 1, flag = 0;
    if (s->mode == EM_FE) {
      double Df = rdo_escape_log(ic, F, work_at(s, cx, cy, &sec));
      dsum += rdo_exp(Df);
      double rnd = uniform(s->seed, s->stream, (uint64_t)step, tagv(T_EMIT, 0), (uint64_t)k, 0);
      if (rdo_log(rnd) <= Df) {
        flag = 1;
        pe[0] = cx; pe[1] = cy; pe[2] = 1.0 * LS;
        work_at(s, cx, cy, &sec);
      }
    } else if (gtf) {
      flag = 1;
      pe[0] = cx; pe[1] = cy; pe[2] = 1.0 * LS;
      mb_velocity(s->seed, s->stream, step, k, s->T, v);
      work_at(s, cx, cy, &sec);
    } else {
      double sw = sqrt(s->w_tip);
      double D = rdo_exp(b_FN_() * (sw * sw * sw) * fn_v(ic, l_const_(), F, s->w_tip) / fabs(F));
      double rnd = uniform(s->seed, s->stream, (uint64_t)step, tagv(T_EMIT, 0), (uint64_t)k, 0);
      if (rnd <= D) {
        flag = 1;
        double nx, ny, nz;
        tip_normal(&s->g, p[0], p[1], &nx, &ny, &nz);
        pe[0] = p[0] + nx * LS;
        pe[1] = p[1] + ny * LS;
        pe[2] = p[2] + nz * LS;
      }
    }
    if (flag) {
      int before = s->n;
      add_particle(s, pe, v, 1, step, 1, sec);
      if (s->n > before) {
        emitted++;
        rdo_emit_rec* r = &s->elog[s->n_elog++];
        r->x = pe[0]; r->y = pe[1]; r->z = pe[2];
        r->emitter = 1; r->section = sec; r->id = s->id[before]; r->species = 1;
      }
    }
  }
  s->emitted = emitted;
  if (s->mode == EM_FE) s->df_avg = dsum / s->N_sup;
}

void rdo_sim_set_serial(rdo_sim* s, int on) { s->serial = on; }

/* ---- Beeman update, boundaries, Ramo ---------------------------------------- */
static void mark_remove(rdo_sim* s, int i, int reason) {
  if (!s->alive[i]) return;
  s->alive[i] = 0;
  s->q[i] = 0.0;
  {
    /* Mark_Particles_Remove's record (src/mod_pair.F90:240-256): position and
     * the velocity at marking; particles are marked in index order */
    rdo_abs_rec* a = &s->alog[s->n_alog++];
    a->x = s->pos[3 * i];
    a->y = s->pos[3 * i + 1];
    a->z = s->pos[3 * i + 2];
    a->vx = s->vel[3 * i];
    a->vy = s->vel[3 * i + 1];
    a->vz = s->vel[3 * i + 2];
    a->emitter = s->emitter[i];
    a->section = s->section[i];
    a->id = s->id[i];
    a->where = reason;
  }
  s->rem_total++;
  if (s->species[i] == 1) {
    s->rem_elec++;
    if (reason == 1) {
      s->rem_top++;
      if (s->emitter[i] == 1) s->rem_top_emit++;
    } else {
      s->rem_bot++;
    }
  } else {
    s->rem_ion++;
    if (reason == 1) s->rem_top++;
    else s->rem_bot++;
  }
}

/* Check_Planes (src/mod_verlet.F90:343-367): a crossing from below of a
 * plane with height > 0, recorded with the position after the move and the
 * velocity before the velocity update; particles in index order. */
static void check_planes(rdo_sim* s, int i) {
  for (int k = 0; k < s->nplanes; ++k) {
    const double zp = s->planes[k];
    if (!(zp > 0.0)) continue;
    if (s->pos[3 * i + 2] > zp && s->prev[3 * i + 2] < zp) {
      rdo_plane_rec* r = &s->plog[(size_t)k * s->cap + s->n_plog[k]++];
      r->x = s->pos[3 * i];
      r->y = s->pos[3 * i + 1];
      r->vx = s->vel[3 * i];
      r->vy = s->vel[3 * i + 1];
      r->vz = s->vel[3 * i + 2];
      r->emitter = s->emitter[i];
      r->section = s->section[i];
      r->id = s->id[i];
      r->plane = k + 1;
    }
  }
}

/* Velocity_Verlet (src/mod_verlet.F90:123-162): positions, accelerations,
 * velocities.  Two-time-step mode (s->tts = atom_time_interval K > 0,
 * Update_Species_Position / _Velocity, :234-290, :511-580): ions move only
 * at steps that are multiples of K, over K time steps; between them nothing
 * of an ion changes, no boundary is checked and it adds nothing to the Ramo
 * current or the velocity sums. */
static int ion_rests(const rdo_sim* s, int i, int step) {
  return s->tts > 0 && s->species[i] == 2 && step % s->tts != 0;
}
static void update_position(rdo_sim* s, int step) {
  const double c16 = 1.0 / 6.0;
  for (int i = 0; i < s->n; ++i) {
    if (ion_rests(s, i, step)) continue;
    const double ks = (s->tts > 0 && s->species[i] == 2) ? (double)s->tts : 1.0;
    for (int k = 0; k < 3; ++k) {
      double x = s->pos[3 * i + k];
      s->prev[3 * i + k] = x;
      double a = s->acc[3 * i + k], a1 = s->a1[3 * i + k];
      s->pos[3 * i + k] =
          x + s->vel[3 * i + k] * s->dt * ks + c16 * (4.0 * a - a1) * s->dt2 * (ks * ks);
      s->a2[3 * i + k] = a1;
      s->a1[3 * i + k] = a;
      s->acc[3 * i + k] = 0.0;
    }
    double z = s->pos[3 * i + 2];
    if (z < 0.0) mark_remove(s, i, 2);
    else if (z > s->box_top) mark_remove(s, i, 1);
    if (s->g.kind == RDO_GEOM_TIP) {
      double eta = tip_eta(&s->g, s->pos[3 * i], s->pos[3 * i + 1], z);
      if (eta < s->g.eta_1) mark_remove(s, i, 2);
    }
    check_planes(s, i);
  }
  if (s->tts > 0)
    rdo_accel_tts(&s->g, s->n, s->pos, s->q, s->m, s->species, s->tts, step, s->acc, s->nthreads);
  else
    rdo_accel(&s->g, s->n, s->pos, s->q, s->m, s->acc, s->nthreads);
  /* velocity + Ramo; the device sums 256-particle blocks with a fixed tree
   * and then the block sums in order: reproduced here */
  const int nb = s->n > 0 ? (s->n + 255) / 256 : 1;
  double tot[13];
  for (int k = 0; k < 13; ++k) tot[k] = 0.0;
  for (int b = 0; b < nb; ++b) {
    double lane[13][256];
    for (int t = 0; t < 256; ++t) {
      for (int k = 0; k < 13; ++k) lane[k][t] = 0.0;
      int i = b * 256 + t;
      if (i >= s->n || ion_rests(s, i, step)) continue;
      const double ks = (s->tts > 0 && s->species[i] == 2) ? (double)s->tts : 1.0;
      double v[3];
      for (int k = 0; k < 3; ++k)
        v[k] = s->vel[3 * i + k] +
               c16 * (2.0 * s->acc[3 * i + k] + 5.0 * s->a1[3 * i + k] - s->a2[3 * i + k]) *
                   s->dt * ks;
      for (int k = 0; k < 3; ++k) s->vel[3 * i + k] = v[k];
      double u[3];
      if (s->g.kind == RDO_GEOM_TIP) {
        double e[3];
        vacuum(s, &s->pos[3 * i], e);
        for (int k = 0; k < 3; ++k) u[k] = e[k] * s->g.pre_fac_E_tip_unit_voltage / s->g.pre_fac_E_tip;
      } else {
        u[0] = 0.0; u[1] = 0.0; u[2] = -1.0 / s->g.d;
      }
      double EzV = v[0] * u[0] + v[1] * u[1] + v[2] * u[2];
      int spc = s->species[i];
      if (spc >= 1 && spc <= 3) lane[spc - 1][t] = s->q[i] * EzV;
      lane[3][t] = v[0]; lane[4][t] = v[1]; lane[5][t] = v[2];
      if (spc == 1) { lane[6][t] = v[0]; lane[7][t] = v[1]; lane[8][t] = v[2]; lane[12][t] = 1.0; }
      else if (spc == 2) { lane[9][t] = v[0]; lane[10][t] = v[1]; lane[11][t] = v[2]; }
    }
    for (int w = 128; w > 0; w >>= 1)
      for (int t = 0; t < w; ++t)
        for (int k = 0; k < 13; ++k) lane[k][t] = lane[k][t] + lane[k][t + w];
    for (int k = 0; k < 13; ++k) tot[k] = tot[k] + lane[k][0];
  }
  for (int k = 0; k < 3; ++k) s->ramo[k] = tot[k];
  for (int k = 0; k < 9; ++k) s->vsum[k] = tot[3 + k];
  s->n_elec = (int)tot[12];
}

/* Remove_Particles: stable compaction + record_lifetime */
static void remove_particles(rdo_sim* s, int step) {
  if (s->rem_total == 0) return;
  int j = 0;
  for (int i = 0; i < s->n; ++i) {
    if (s->alive[i]) {
      if (j != i) {
        for (int k = 0; k < 3; ++k) {
          s->pos[3 * j + k] = s->pos[3 * i + k];
          s->prev[3 * j + k] = s->prev[3 * i + k];
          s->vel[3 * j + k] = s->vel[3 * i + k];
          s->acc[3 * j + k] = s->acc[3 * i + k];
          s->a1[3 * j + k] = s->a1[3 * i + k];
          s->a2[3 * j + k] = s->a2[3 * i + k];
        }
        s->q[j] = s->q[i]; s->m[j] = s->m[i];
        s->species[j] = s->species[i]; s->pstep[j] = s->pstep[i];
        s->emitter[j] = s->emitter[i]; s->section[j] = s->section[i]; s->id[j] = s->id[i];
        s->alive[j] = 1;
      }
      ++j;
    } else {
      int lt = step - s->pstep[i];
      if (lt <= 0) lt = 1;
      if (lt > NLIFE) lt = NLIFE;
      int spc = s->species[i];
      if (spc >= 1 && spc <= 2) s->life[spc - 1][lt - 1]++;
    }
  }
  s->n = j;
}

static void reset_counters(rdo_sim* s) {
  s->n_elog = 0;
  s->n_alog = 0;
  for (int k = 0; k < 10; ++k) s->n_plog[k] = 0;
  s->emitted = 0;
  s->n_round = 0;
  s->rem_total = s->rem_top = s->rem_bot = s->rem_elec = s->rem_ion = s->rem_top_emit = 0;
  s->start_rounds = 0;
  s->supply_level = 0;
  s->supply_cells = 0;
  s->supply_fail = 0;
  s->supply_err = 0.0;
  s->N_sup = 0.0;
  s->df_avg = 0.0;
  s->F_avg[0] = s->F_avg[1] = s->F_avg[2] = 0.0;
}

/* One main-loop step (main.F90:175-219); phases as the device engine:
 * bit 0 emission, bit 1 position/force/velocity, bit 2 removal. */
void rdo_sim_step(rdo_sim* s, int step, int phases, rdo_record* rec) {
  if (phases & 1) {
    reset_counters(s);
    do_emission(s, step);
    s->emit_step = step;
  }
  if (phases & 2) {
    /* a step's counters start at zero with its first phase: the emission
       phase of the same step when it ran on its own, else this update */
    if (!(phases & 1) && s->emit_step != step) reset_counters(s);
    update_position(s, step);
  }
  if (phases & 4) remove_particles(s, step);
  if (rec) {
    memset(rec, 0, sizeof(*rec));
    rec->step = step;
    rec->n = s->n;
    rec->emitted = s->emitted;
    rec->n_round = s->n_round;
    rec->absorbed = s->rem_total;
    rec->absorbed_top = s->rem_top;
    rec->absorbed_bot = s->rem_bot;
    rec->dropped = s->dropped;
    rec->start_rounds = s->start_rounds;
    rec->supply_level = s->supply_level;
    rec->supply_cells = s->supply_cells;
    rec->supply_fail = s->supply_fail;
    rec->supply_err = s->supply_err;
    rec->n_sup = s->N_sup;
    rec->ramo_current = s->ramo[0] + s->ramo[1] + s->ramo[2];
    rec->mh_std = s->MH_std;
    rec->a_rate = s->a_rate;
    rec->df_avg = s->df_avg;
    for (int k = 0; k < 3; ++k) {
      rec->f_avg[k] = s->F_avg[k];
      rec->ramo[k] = s->ramo[k];
      rec->avg_elec_vel[k] = s->n_elec > 0 ? s->vsum[3 + k] / s->n_elec : 0.0;
    }
  }
}

int rdo_sim_count(const rdo_sim* s) { return s->n; }

/* The last step's particle records; returns how many (copied when out is
 * not NULL). */
int rdo_sim_emit_log(const rdo_sim* s, rdo_emit_rec* out) {
  if (out) memcpy(out, s->elog, sizeof(rdo_emit_rec) * (size_t)s->n_elog);
  return s->n_elog;
}
void rdo_sim_set_planes(rdo_sim* s, int n, const double* z) {
  s->nplanes = n < 0 ? 0 : (n > 10 ? 10 : n);
  for (int k = 0; k < 10; ++k) s->planes[k] = k < s->nplanes ? z[k] : 0.0;
}
/* All planes' crossings of the last step, plane 1 first. */
int rdo_sim_plane_log(const rdo_sim* s, rdo_plane_rec* out) {
  int at = 0;
  for (int k = 0; k < 10; ++k) {
    if (out)
      memcpy(out + at, s->plog + (size_t)k * s->cap, sizeof(rdo_plane_rec) * (size_t)s->n_plog[k]);
    at += s->n_plog[k];
  }
  return at;
}
int rdo_sim_absorb_log(const rdo_sim* s, rdo_abs_rec* out) {
  if (out) memcpy(out, s->alog, sizeof(rdo_abs_rec) * (size_t)s->n_alog);
  return s->n_alog;
}

void rdo_sim_get(const rdo_sim* s, double* pos, double* prev, double* vel, double* acc,
                 double* a1, double* a2, int* ints, double* q) {
  int n = s->n;
  if (pos) memcpy(pos, s->pos, 24 * (size_t)n);
  if (prev) memcpy(prev, s->prev, 24 * (size_t)n);
  if (vel) memcpy(vel, s->vel, 24 * (size_t)n);
  if (acc) memcpy(acc, s->acc, 24 * (size_t)n);
  if (a1) memcpy(a1, s->a1, 24 * (size_t)n);
  if (a2) memcpy(a2, s->a2, 24 * (size_t)n);
  if (q) memcpy(q, s->q, 8 * (size_t)n);
  if (ints)
    for (int i = 0; i < n; ++i) {
      ints[5 * i] = s->species[i];
      ints[5 * i + 1] = s->pstep[i];
      ints[5 * i + 2] = s->emitter[i];
      ints[5 * i + 3] = s->section[i];
      ints[5 * i + 4] = s->id[i];
    }
}

void rdo_sim_life(const rdo_sim* s, int* hist) {
  for (int sp = 0; sp < 2; ++sp)
    for (int l = 0; l < NLIFE; ++l) hist[sp * NLIFE + l] = s->life[sp][l];
}
