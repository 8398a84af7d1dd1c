/*
 * rumdeed.h -- C ABI of librumdeed.so, the MI355X (gfx950) implementation of
 * the RUMDEED per-step particle hot path.
 *
 * Conventions (all entry points):
 *   - return int status: RD_OK (0) or an RD_ERR_* code; rd_last_error()
 *     gives the message of the last failure on the calling thread.  Nothing
 *     crosses the ABI as an exception.
 *   - SI units (m, s, V, C, kg), FP64, host buffers owned by the caller.
 *     Vectors are (3,n) column-major like the Fortran particle arrays
 *     (particles_cur_pos(1:3, 1:n), src/mod_global.F90:128-135).
 *   - Particle indices in this ABI are 0-based; the Fortran layer adds 1.
 *   - A context owns its particle state in HBM and one HIP stream; calls on
 *     one context are serialised by the caller (the reference calls the path
 *     from its master thread, outside OpenMP regions).
 *
 * The reference interface each entry point replaces is cited as file:line in
 * the RUMDEED source tree.
 */
#ifndef RUMDEED_H
#define RUMDEED_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RD_OK 0
#define RD_ERR_ARG 1      /* invalid argument */
#define RD_ERR_HIP 2      /* HIP runtime failure */
#define RD_ERR_CAPACITY 3 /* particle or batch capacity exceeded */
#define RD_ERR_STATE 4    /* call not valid in the current state */

/* Geometries implemented on the device (ACC_Geometry, src/mod_verlet.F90:1168-1180). */
#define RD_GEOM_PLANAR 1 /* field_E_planar + Force_Image_charges_v2 */
#define RD_GEOM_TIP 2    /* field_E_Hyperboloid + Sphere_IC_field */

/* Arithmetic of the pair kernels. */
#define RD_PREC_FP64 0  /* plain FP64, fixed summation order (bit-reproducible) */
#define RD_PREC_MIXED 1 /* FP32 pair arithmetic, FP64 per-particle accumulation */

/* Particle species (src/mod_global.F90:103-107). */
#define RD_SPECIES_ELEC 1
#define RD_SPECIES_ION 2

typedef struct rd_ctx rd_ctx;

/* Diode geometry.  Planar: gap d (box_dim(3)) and voltage V_s, image charge
 * switch and series length N_ic_max (src/mod_global.F90:191-195, :313-314).
 * Tip: emitters_dim(1:3,1) = d_tip, R_base, h_tip (src/mod_emission_tip.f90:105-108);
 * the prolate-spheroidal parameters are derived as Init_Emission_Tip does. */
typedef struct {
  int32_t kind;         /* RD_GEOM_PLANAR | RD_GEOM_TIP */
  int32_t image_charge; /* 0 / 1 */
  int32_t n_ic_max;     /* planar image series, 0..64 */
  int32_t reserved;
  double V;             /* V_s (V) */
  double d;             /* planar gap (m); ignored for the tip */
  double d_tip, R_base, h_tip; /* tip (m); ignored for planar */
} rd_geometry;

/* ---- library and context ------------------------------------------------ */
const char* rd_version(void);
const char* rd_last_error(void);
int rd_device_count(int32_t* n);

/* Allocate a context with room for max_particles in HBM on `device`
 * (replaces the particle-array allocation of Init, src/main.F90:422-456, and
 * the device residency of Ensure_Device_Particles, src/mod_verlet.F90:77-84). */
int rd_ctx_create(int32_t device, int64_t max_particles, rd_ctx** out);
int rd_ctx_destroy(rd_ctx* ctx);

/* Geometry and time step (Read_Input_Variables, src/main.F90:350-379;
 * Init_Field_Emission_v2 / Init_Emission_Tip pointer binding). */
int rd_set_geometry(rd_ctx* ctx, const rd_geometry* geom);
int rd_set_time_step(rd_ctx* ctx, double dt);

/* ---- particle store ------------------------------------------------------ */
/* Replace the whole particle set (snapshot put; acceleration histories are
 * zeroed).  pos, vel: (3,n); q, m: (n); species: (n) or NULL for electrons. */
int rd_put_particles(rd_ctx* ctx, int32_t n, const double* pos, const double* vel,
                     const double* q, const double* m, const int32_t* species);
int rd_particle_count(rd_ctx* ctx, int32_t* n);
/* Copy the live particles back (any pointer may be NULL). */
int rd_get_particles(rd_ctx* ctx, double* pos, double* vel, double* q, double* m,
                     int32_t* species);

/* ---- forces and fields ---------------------------------------------------- */
/* Accelerations of all particles into the current-acceleration slot
 * (Calculate_Acceleration_Particles, src/mod_verlet.F90:597-620, gather form
 * of Calculate_Acceleration_Particles_ACC, :1217-1429). */
int rd_accel(rd_ctx* ctx, int32_t precision);
int rd_get_accel(rd_ctx* ctx, double* acc);

/* Electric field at m points (Calc_Field_at_Batch, src/mod_verlet.F90:1635-1911;
 * each point as Calc_Field_at, :1466-1529).  pts, field: (3,m) host buffers. */
int rd_field_batch(rd_ctx* ctx, int32_t m, const double* pts, double* field, int32_t precision);

/* Field on the nx*ny cell centres of the rectangle [x0,x0+lx]x[y0,y0+ly] at
 * height z (points ordered x fastest), kept in HBM; copied to `field` (3,nx*ny)
 * when it is not NULL.  The surface grid of the force/field snapshot. */
int rd_field_grid(rd_ctx* ctx, double x0, double y0, double lx, double ly, int32_t nx,
                  int32_t ny, double z, int32_t precision, double* field);

/* ---- measurement ----------------------------------------------------------- */
/* Average device time (ms) of the accelerations / grid field over `reps`
 * launches, from HIP events on the context's stream; `kernel_ms` receives
 * the average of the pair kernel alone (may be NULL). */
int rd_time_accel(rd_ctx* ctx, int32_t precision, int32_t reps, double* total_ms,
                  double* kernel_ms);
int rd_time_field_grid(rd_ctx* ctx, double x0, double y0, double lx, double ly, int32_t nx,
                       int32_t ny, double z, int32_t precision, int32_t reps, double* total_ms,
                       double* kernel_ms);

#ifdef __cplusplus
}
#endif
#endif
