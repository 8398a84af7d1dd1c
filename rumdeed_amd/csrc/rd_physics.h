// Emission physics on the device: Fowler-Nordheim functions, Jensen's
// general thermal-field current density, the checkerboard work function and
// the hyperboloid-tip surface parametrisation.  All FP64, with the library's
// own log/exp/sin/cos (rd_math.h) so that every decision is reproducible.
//
// Sources (RUMDEED tree):
//   FN functions ....... src/mod_field_emission_v2.F90:515-625 (planar),
//                        src/mod_emission_tip.f90:1657-1758 (tip)
//   GTF ................ src/mod_kevin_rjgtf_v2.f90:56-175
//   work function ...... src/mod_work_function.F90:389-487 (checkerboard)
//   tip surface ........ src/mod_hyperboloid_tip.f90:25-163
#pragma once

#include "rd_math.h"

namespace rd {

// ---- own sine/cosine: reduction by pi/2 (three-part constant: pi/2 from
// Machin's formula, cut after 33, 66 and 119 bits), Taylor polynomials on
// |r| <= pi/4 to r^17 / r^16.
static constexpr double kPio2Hi = 0x1.921fb54400000p+0;
static constexpr double kPio2Mid = 0x1.0b4611a600000p-34;
static constexpr double kPio2Lo = 0x1.3198a2e037073p-69;
static constexpr double kInvPio2 = 0x1.45f306dc9c883p-1;

RD_HD double sin_poly(double r) {
  double z = r * r;
  double p = 1.0 / 355687428096000.0;  // 1/17!
  p = -1.0 / 1307674368000.0 + z * p;  // -1/15!
  p = 1.0 / 6227020800.0 + z * p;      // 1/13!
  p = -1.0 / 39916800.0 + z * p;       // -1/11!
  p = 1.0 / 362880.0 + z * p;          // 1/9!
  p = -1.0 / 5040.0 + z * p;           // -1/7!
  p = 1.0 / 120.0 + z * p;             // 1/5!
  p = -1.0 / 6.0 + z * p;              // -1/3!
  return r + (r * z) * p;
}
RD_HD double cos_poly(double r) {
  double z = r * r;
  double p = 1.0 / 20922789888000.0;  // 1/16!
  p = -1.0 / 87178291200.0 + z * p;   // -1/14!
  p = 1.0 / 479001600.0 + z * p;      // 1/12!
  p = -1.0 / 3628800.0 + z * p;       // -1/10!
  p = 1.0 / 40320.0 + z * p;          // 1/8!
  p = -1.0 / 720.0 + z * p;           // -1/6!
  p = 1.0 / 24.0 + z * p;             // 1/4!
  p = -0.5 + z * p;                   // -1/2!
  return 1.0 + z * p;
}
// sin and cos together, for |x| < 2^20.
RD_HD void rd_sincos(double x, double& s, double& c) {
  double kd = __builtin_floor(x * kInvPio2 + 0.5);
  double r = ((x - kd * kPio2Hi) - kd * kPio2Mid) - kd * kPio2Lo;
  double sp = sin_poly(r), cp = cos_poly(r);
  int q = ((int)kd) & 3;
  if (q == 0) {
    s = sp;
    c = cp;
  } else if (q == 1) {
    s = cp;
    c = -sp;
  } else if (q == 2) {
    s = -sp;
    c = -cp;
  } else {
    s = -cp;
    c = sp;
  }
}

// ---- constants of the emission formulas (src/mod_field_emission_v2.F90:36-47)
struct EmitConsts {
  double a_FN, b_FN, l_const, q0, m0, k_b;
};
RD_HD EmitConsts emit_consts() {
  const double pi = 3.141592653589793238462643383279502884197;
  const double q0 = 1.602176634e-19, m0 = 9.1093837015e-31;
  const double mu0 = 1.25663706212e-6, c = 299792458.0, h = 6.62607015e-34;
  const double eps0 = 1.0 / (mu0 * (c * c));
  const double hbar = h / (2.0 * pi);
  EmitConsts e;
  e.a_FN = (q0 * q0) / (16.0 * (pi * pi) * hbar);
  e.b_FN = -(4.0 / (3.0 * hbar) * __builtin_sqrt(2.0 * m0 * q0));
  e.l_const = q0 / (4.0 * pi * eps0);
  e.q0 = q0;
  e.m0 = m0;
  e.k_b = 1.380649e-23;
  return e;
}

// l = y^2 of the Forbes-Deane approximation, clamped to 1.
RD_HD double fn_l(double l_const, double F, double w) { return l_const * (-1.0 * F) / (w * w); }
RD_HD double fn_v(int ic, double l_const, double F, double w) {
  if (!ic) return 1.0;
  double l = fn_l(l_const, F, w);
  if (l > 1.0) l = 1.0;
  return 1.0 - l + 1.0 / 6.0 * l * rd_log(l);
}
RD_HD double fn_t(int ic, double l_const, double F, double w) {
  if (!ic) return 1.0;
  double l = fn_l(l_const, F, w);
  if (l > 1.0) l = 1.0;
  return 1.0 + l * (1.0 / 9.0 - 1.0 / 18.0 * rd_log(l));
}
// Escape_Prob_log: b_FN w^(3/2) v_y / (-F)
RD_HD double fn_escape_log(const EmitConsts& e, int ic, double F, double w) {
  double sw = __builtin_sqrt(w);
  return e.b_FN * (sw * sw * sw) * fn_v(ic, e.l_const, F, w) / (-1.0 * F);
}
// Elec_Supply_log: 2 log(-F) - 2 log(t_y) - log(w)
RD_HD double fn_supply_log(const EmitConsts& e, int ic, double F, double w) {
  return 2.0 * rd_log(-1.0 * F) - 2.0 * rd_log(fn_t(ic, e.l_const, F, w)) - rd_log(w);
}
// Elec_Supply_V2: (dt/q0) a_FN / (t_y^2 w) F^2
RD_HD double fn_supply(const EmitConsts& e, int ic, double dt_q0, double F, double w) {
  double t = fn_t(ic, e.l_const, F, w);
  return dt_q0 * e.a_FN / (t * t * w) * (F * F);
}

// ---- Jensen oGTF (src/mod_kevin_rjgtf_v2.f90).  Units eV, nm, fs, q = 1.
struct GtfConsts {
  double pi, kb, hbar, c, mo, afs, Qo, cm, Amp, Arld;
};
RD_HD GtfConsts gtf_consts() {
  GtfConsts k;
  k.pi = 3.14159265358979324;
  k.kb = 1.0 / 11604.50635;
  k.hbar = 0.6582119571;
  k.c = 299.7924580;
  k.mo = 5.685630103;
  k.afs = 1.0 / 137.035999084;
  k.Qo = k.afs * k.hbar * k.c / 4.0;
  k.cm = 1.0e7;
  k.Amp = 6.241509074e3;
  k.Arld = (k.mo * (k.kb * k.kb) / (2.0 * (k.pi * k.pi) * (k.hbar * k.hbar * k.hbar))) *
           (k.cm * k.cm) / k.Amp;
  return k;
}

RD_HD double gtf_Nns(double n, double s) {
  if (n == 1.0) return (s + 1.0) * rd_exp(-s);
  double x = n * n;
  double y = 1.0 / x;
  double z = (n - 1.0) * s;
  double sng;
  if (__builtin_fabs(z) > 1.0e-5)
    sng = (x + 1.0) * (x * rd_exp(-s) - rd_exp(-n * s)) / (x - 1.0);
  else
    sng = (0.5 * (x + 1.0) * rd_exp(-s) / (n + 1.0)) *
          ((1.0 - n) * (s * s) + 2.0 * (1.0 + n) + 2.0 * s);
  double sn = -x * (0.10593434 * x + 0.35506593);
  double sd = -y * (0.10593434 * y + 0.35506593);
  double sfn = x * rd_exp(-s);
  double v = sng + sn * rd_exp(-n * s) + x * sd * rd_exp(-s);
  return v > sfn ? v : sfn;
}

// Get_Kevin_Jgtf_v2: current density (A/m^2) for field F (V/m), T (K), phi (eV).
RD_HD double gtf_current(double F, double To, double Phi) {
  const GtfConsts k = gtf_consts();
  double Fo = __builtin_fabs(F) * 1.0e-9;
  if (Fo < 1.0e-9) return 0.0;
  const double chem = 7.0;
  double s4 = __builtin_sqrt(4.0 * k.Qo * Fo);
  double yo = s4 / Phi;
  double phix = Phi - s4;
  double ly = rd_log(yo);
  double ty = 1.0 + (yo * yo) * (1.0 - ly) / 9.0;
  double vy = 1.0 - (yo * yo) * (3.0 - ly) / 3.0;
  double Tmin = (k.hbar * Fo / (4.0 * k.kb * ty)) * __builtin_sqrt(2.0 / (k.mo * Phi));
  double Tmax = k.hbar * Fo / (k.kb * k.pi * __builtin_sqrt(k.mo * Phi * yo));
  double betaT = 1.0 / (k.kb * To);
  double betau = (2.0 / (k.hbar * Fo)) * __builtin_sqrt(2.0 * k.mo * Phi) * ty;
  double betap = (k.pi / (k.hbar * Fo)) * __builtin_sqrt(k.mo * Phi * yo);
  double theto = (4.0 * __builtin_sqrt(2.0 * k.mo * (Phi * Phi * Phi)) / (3.0 * k.hbar * Fo)) * vy;
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
    double po = (-Bp - __builtin_sqrt(Bp * Bp - 4.0 * Ap * Cp)) / (2.0 * Ap);
    double Em = chem + po * phix;
    double theta = ((1.0 - po) * (1.0 - po)) * (2.0 * po + 1.0) * theto -
                   phix * po * (1.0 - po) * ((1.0 - po) * betau - po * betap);
    nft = 1.0;
    sft = theta + betaT * (Em - chem);
  }
  double J = k.Arld * gtf_Nns(nft, sft) * (To * To);
  return J * 1.0e4;
}

// ---- checkerboard work function (w_theta_checkerboard).  table: ny rows as
// read from the file (first row = top), nx columns.  Returns w, sets sec.
RD_HD double work_checker(const double* table, int nx, int ny, double ex0, double ey0,
                          double elx, double ely, double x, double y, int& sec) {
  double xs = (x - ex0) / elx;
  double ys = (y - ey0) / ely;
  double x_len = 1.0 / nx, y_len = 1.0 / ny;
  int xi = (int)__builtin_floor(xs / x_len) + 1;
  int yi = (int)__builtin_floor(ys / y_len) + 1;
  if (xi > nx)
    xi = nx;
  else if (xi < 1)
    xi = 1;
  if (yi > ny)
    yi = ny;
  else if (yi < 1)
    yi = 1;
  sec = nx * (yi - 1) + xi;
  int yr = ny - yi + 1;
  return table[(yr - 1) * nx + (xi - 1)];
}

// ---- hyperboloid tip surface (mod_hyperboloid_tip.f90:25-163)
struct TipSurf {
  double a, eta1, shift_z, max_xi;
};
RD_HD void tip_xyz(const TipSurf& t, double xi, double eta, double phi, double& x, double& y,
                   double& z) {
  double xy = t.a * __builtin_sqrt((xi * xi - 1.0) * (1.0 - eta * eta));
  double s, c;
  rd_sincos(phi, s, c);
  x = xy * c;
  y = xy * s;
  z = t.a * xi * eta + t.shift_z;
}
RD_HD void tip_normal(const TipSurf& t, double x, double y, double& nx, double& ny, double& nz) {
  double eta_fac = t.eta1 / __builtin_sqrt(1 - t.eta1 * t.eta1);
  double div_fac = -1.0 / __builtin_sqrt(x * x + y * y + t.a * t.a * (1 - t.eta1 * t.eta1));
  double ux = eta_fac * x * div_fac;
  double uy = eta_fac * y * div_fac;
  double uz = 1.0;
  double nrm = __builtin_sqrt(ux * ux + uy * uy + uz * uz);
  nx = ux / nrm;
  ny = uy / nrm;
  nz = uz / nrm;
}
RD_HD double tip_area(const TipSurf& t, double xi1, double xi2, double phi1, double phi2) {
  double e2 = t.eta1 * t.eta1;
  double s1 = __builtin_sqrt(xi1 * xi1 - e2), s2 = __builtin_sqrt(xi2 * xi2 - e2);
  double f1 = xi1 * s1 - e2 * rd_log(xi1 + s1);
  double f2 = xi2 * s2 - e2 * rd_log(xi2 + s2);
  return 0.5 * (t.a * t.a) * __builtin_sqrt(1.0 - e2) * (phi2 - phi1) * (f2 - f1);
}
RD_HD double tip_eta(const TipSurf& t, double x, double y, double z) {
  return 1.0 / (2.0 * t.a) *
         (__builtin_sqrt(x * x + y * y + (z + t.a - t.shift_z) * (z + t.a - t.shift_z)) -
          __builtin_sqrt(x * x + y * y + (z - t.a - t.shift_z) * (z - t.a - t.shift_z)));
}

// Fortran modulo(a, p) = a - floor(a/p) p
RD_HD double fmodulo(double a, double p) { return a - __builtin_floor(a / p) * p; }

// ---- counter-based draws.  Every stream is keyed by the simulation seed;
// the counter holds (step, tag, chain, block).
enum : uint64_t {
  kTagStart = 1,    // uniform start positions, per start round
  kTagJump = 2,     // Gaussian proposal, per jump
  kTagAccept = 3,   // Metropolis accept draw, per jump
  kTagEmit = 4,     // escape test, per chain
  kTagPoisson = 5,  // Poisson count, per step
  kTagVel = 6,      // Maxwell-Boltzmann velocity, per chain
};
RD_HD uint64_t tag(uint64_t kind, uint64_t sub) { return (kind << 40) | sub; }

// Polar Box-Muller (box_muller, src/mod_global.F90:578-595) on the stream
// (seed; step, tg, chain, block = 0, 1, ...): two draws per attempt, two
// attempts per Philox block, until 0 < w < 1.
// The unit pair (x1 f, x2 f) of the polar method; box_muller scales it as
// (x1 f) s + m.  Returns false (pair 0, 0) if no attempt of the bounded
// stream was accepted.
RD_HD bool box_muller_unit(uint64_t seed, uint64_t sim, uint64_t step, uint64_t tg,
                           uint64_t chain, double& u1, double& u2) {
  for (uint64_t blk = 0; blk < 4096; ++blk) {
    u64x4 r = philox4x64(step, tg, chain, blk, seed, sim);
    for (int a = 0; a < 2; ++a) {
      double x1 = 2.0 * u01(r.v[2 * a]) - 1.0;
      double x2 = 2.0 * u01(r.v[2 * a + 1]) - 1.0;
      double w = x1 * x1 + x2 * x2;
      if ((w < 1.0) && (w > 0.0)) {
        double f = __builtin_sqrt((-2.0 * rd_log(w)) / w);
        u1 = x1 * f;
        u2 = x2 * f;
        return true;
      }
    }
  }
  u1 = 0.0;
  u2 = 0.0;
  return false;
}

RD_HD void box_muller(uint64_t seed, uint64_t sim, uint64_t step, uint64_t tg, uint64_t chain,
                      double m1, double m2, double s1, double s2, double& o1, double& o2) {
  double u1, u2;
  if (box_muller_unit(seed, sim, step, tg, chain, u1, u2)) {
    o1 = u1 * s1 + m1;
    o2 = u2 * s2 + m2;
  } else {
    o1 = m1;
    o2 = m2;
  }
}

RD_HD double uniform(uint64_t seed, uint64_t sim, uint64_t step, uint64_t tg, uint64_t chain,
                     int which) {
  u64x4 r = philox4x64(step, tg, chain, 0, seed, sim);
  return u01(r.v[which & 3]);
}


// Fortran nint: round half away from zero (x - trunc(x) is exact).
RD_HD int nint(double x) {
  double t = __builtin_trunc(x);
  double f = x - t;
  if (f >= 0.5)
    t = t + 1.0;
  else if (f <= -0.5)
    t = t - 1.0;
  return (int)t;
}

}  // namespace rd
