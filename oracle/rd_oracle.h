/*
 * ORACLE -- test infrastructure only.
 *
 * CPU restatement (plain C, FP64 / long double) of the RUMDEED per-step
 * particle hot path.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker.  The
 * product (rumdeed_amd/librumdeed.so) never links or calls it.
 *
 * Every function cites the reference routine it restates (paths relative to
 * the RUMDEED source tree).
 */
#ifndef RD_ORACLE_H
#define RD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { RDO_GEOM_PLANAR = 1, RDO_GEOM_TIP = 2 };

/* Geometry of the diode; SI units (m, V).  Mirrors the module variables the
 * reference reads: d, E_z, image_charge, N_ic_max (src/mod_global.F90:191-195,
 * :313-314) and the hyperboloid parameters set by Init_Emission_Tip
 * (src/mod_emission_tip.f90:105-125). */
typedef struct {
  int32_t kind;          /* RDO_GEOM_PLANAR | RDO_GEOM_TIP */
  int32_t image_charge;  /* 0 / 1 */
  int32_t n_ic_max;      /* planar image series length */
  int32_t pad0;
  double d;              /* gap spacing (m); tip: d_tip + h_tip */
  double V;              /* voltage (V) */
  double E_z;            /* planar vacuum field -V/d (V/m) */
  double a_foci, shift_z, pre_fac_E_tip, pre_fac_E_tip_unit_voltage;
  double h_tip, r_tip, eta_1, max_xi, d_tip, R_base, theta_tip;
} rdo_geom;

/* Physical constants exactly as src/mod_global.F90:26-76 builds them. */
typedef struct {
  double pi, h, k_b, c, mu_0, epsilon_0, h_bar, m_0, q_0, q_02, m_u, m_N2, m_N2p;
  double div_fac_c, length_scale, time_scale, eps_soft;
  double a_FN, b_FN, l_const;
} rdo_consts;

void rdo_get_consts(rdo_consts* out);
void rdo_geom_planar(rdo_geom* g, double d, double V, int image_charge, int n_ic_max);
void rdo_geom_tip(rdo_geom* g, double V, double d_tip, double R_base, double h_tip, int image_charge);

/* Summation chunk of the canonical FP64 order (see rd_oracle.c). */
int rdo_sum_chunk(void);

/* Accelerations in the gather form of src/mod_verlet.F90:1217-1429 with the
 * canonical FP64 summation order.  pos/acc are (3,n) column-major. */
void rdo_accel(const rdo_geom* g, int n, const double* pos, const double* q,
               const double* m, double* acc, int nthreads);
/* The two-time-step accelerations of step `step` with atom_time_interval
 * K (Update_Elec_Acceleration / Update_Ion_Acceleration,
 * src/mod_verlet.F90:899-1151), species 1 electron, 2 ion: electrons as
 * rdo_accel but with the electron in the first role of an electron-ion
 * pair; an ion adds onto acc the electrons' reactions, and at steps that are
 * multiples of K also the other ions' forces and the vacuum field. */
void rdo_accel_tts(const rdo_geom* g, int n, const double* pos, const double* q,
                   const double* m, const int* species, int K, int step, double* acc,
                   int nthreads);
/* Same physics, long double, natural j order (brute-force error reference). */
void rdo_accel_ld(const rdo_geom* g, int n, const double* pos, const double* q,
                  const double* m, double* acc, int nthreads);
/* The accelerations of the nt listed targets (0-based indices) over all n
 * sources: rows k of acc (3,nt) equal rows targets[k] of rdo_accel /
 * rdo_accel_ld.  For sampled checks of large snapshots. */
void rdo_accel_subset(const rdo_geom* g, int n, const double* pos, const double* q,
                      const double* m, int nt, const int* targets, double* acc, int nthreads);
void rdo_accel_subset_ld(const rdo_geom* g, int n, const double* pos, const double* q,
                         const double* m, int nt, const int* targets, double* acc,
                         int nthreads);

/* Field at M points (src/mod_verlet.F90:1466-1529 / :1635-1911), canonical
 * FP64 order, and its long-double brute-force counterpart. */
void rdo_field(const rdo_geom* g, int n, const double* pos, const double* q,
               int mpts, const double* pts, double* out, int nthreads);
void rdo_field_ld(const rdo_geom* g, int n, const double* pos, const double* q,
                  int mpts, const double* pts, double* out, int nthreads);

/* Scale of those sums: per point (target) the sum of the norms of the
 * sources' terms plus the norm of the vacuum term.  A sum of terms rounded
 * to relative u is accurate to u times this scale. */
void rdo_field_scale(const rdo_geom* g, int n, const double* pos, const double* q, int mpts,
                     const double* pts, double* out, int nthreads);
void rdo_accel_scale(const rdo_geom* g, int n, const double* pos, const double* q,
                     const double* m, int nt, const int* targets, double* out, int nthreads);

/* Vacuum field at one point (field_E_planar / field_E_Hyperboloid). */
void rdo_vacuum_field(const rdo_geom* g, const double* p, double* out);

/* Own elementary functions (identical algorithm to the product's), exposed
 * so tests can pin them against libm. */
double rdo_log(double x);
double rdo_exp(double x);
void rdo_philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]);

#ifdef __cplusplus
}
#endif
#endif
