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
  double box_top;       /* removal height box_dim(3) (m); 0 = the gap (d, or d_tip + h_tip) */
} rd_geometry;

/* Emission models (EMISSION_MODE, src/mod_global.F90:292-307). */
#define RD_EMIT_NONE 0
#define RD_EMIT_TIP_FE 3      /* mod_emission_tip, EMITTERS_TYPE = 1 */
#define RD_EMIT_PLANAR_GTF 9  /* mod_field_thermo_emission */
#define RD_EMIT_PLANAR_FE 10  /* mod_field_emission_v2, rectangle */

/* One emitter (emitters_pos/dim/delay, the `work` file, T_temp). */
typedef struct {
  int32_t mode;           /* RD_EMIT_* */
  int32_t delay;          /* emits on steps with delay < step */
  double x0, y0;          /* emitters_pos(1:2,1) (m) */
  double lx, ly;          /* emitters_dim(1:2,1) (m) */
  double T;               /* T_temp (K) of the cathode, thermal-field emission */
  double w_tip;           /* tip work function (eV) */
  int32_t wf_nx, wf_ny;   /* checkerboard work function; 1 x 1 = constant */
  const double* wf_table; /* wf_ny rows of wf_nx values, first row = top, as in `work` */
  int32_t quad_n;         /* supply-integral points per side, planar (0 = 32) */
  int32_t ndim;           /* Metropolis-Hastings jumps per chain (0 = 200 FE, 25 GTF, 80 tip) */
  int32_t quad_levels;    /* planar supply: doublings of quad_n until two levels agree within
                             max(0.5 electrons, 1e-3 relative), the reference's Divonne
                             tolerances; -1 none (the fixed rule), 0 up to 4, k in 1..8 up to
                             k; never beyond 1024 points per side */
  int32_t reserved;
} rd_emitter;

/* Per-step record of one simulation (the figures of emitted.dt, absorbed.dt,
 * ramo_current.dt, field.dt: src/mod_pair.F90:790-837, mod_field_emission_v2.F90:389). */
typedef struct {
  int32_t step, n, emitted, n_round;
  int32_t absorbed, absorbed_top, absorbed_bot, dropped;
  int32_t start_rounds;
  int32_t supply_level;  /* planar supply: the level used (quad_n << level points per side) */
  int32_t supply_cells;  /* supply integral: cells of the level used (tip: 100 x 100); 0 when
                            the step has no emission phase */
  int32_t supply_fail;   /* planar supply: 1 when the last allowed level still differs from the
                            one before by more than the tolerance, else 0 */
  double n_sup, ramo_current, mh_std, a_rate, df_avg;
  double f_avg[3];
  double ramo[3];
  double avg_elec_vel[3];
  double supply_err;     /* planar supply: |I_L - I_(L-1)| of the last two levels (electrons);
                            0 for the single level of the fixed rule.  With supply_cells and
                            supply_fail: the line of integration.dt
                            (src/mod_field_emission_v2.F90:654-655) */
} rd_step_record;

/* Particle records of one step for RUMDEED's density files (as
 * `rd_set_logs` collects them; rd_get_logs).  rd_emit_record: every
 * electron the device emitted, in id order, at its emission point (the
 * density_emit*.bin writes of Add_Particle, src/mod_pair.F90:63-95).
 * rd_absorb_record: every particle absorbed, in index order, with its
 * position and its velocity when it was marked (Mark_Particles_Remove,
 * src/mod_pair.F90:240-256); where = 1 top (anode), 2 bottom (cathode or
 * tip).  rd_plane_record: every crossing of a recording plane from below,
 * by plane and then in index order, with the position after the move and
 * the velocity before this step's velocity update (Check_Planes,
 * src/mod_verlet.F90:343-367); plane = 1..RD_MAX_PLANES, the planes-<plane>.bin
 * file.  Positions in m, velocities in m/s. */
typedef struct {
  double x, y, z;
  int32_t emitter, section, id, species;
} rd_emit_record;

typedef struct {
  double x, y, z, vx, vy, vz;
  int32_t emitter, section, id, where;
} rd_absorb_record;

typedef struct {
  double x, y, vx, vy, vz;
  int32_t emitter, section, id, plane;
} rd_plane_record;

/* planes_N_max (src/mod_global.F90:337) */
#define RD_MAX_PLANES 10

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

/* Two-time-step mode: two_time_step with atom_time_interval K
 * (src/main.F90:391-395; Update_Particle_Position / _Acceleration /
 * _Velocity, src/mod_verlet.F90:182-193, 396-419, 886-897; the species
 * routines :234-290, 511-580, 899-1151).  Electrons move every step; ions
 * move, and their velocities change, only at steps that are multiples of K,
 * over K time steps; ions receive the electrons' forces every step and the
 * other ions' forces and the vacuum field at their steps, added to the
 * acceleration they hold.  K = 0 (the default) is the standard integrator.
 * FP64 on both geometries, mixed precision on the planar one (the tip
 * returns RD_ERR_ARG there).  rd_accel keeps the standard accelerations. */
int rd_set_two_time_step(rd_ctx* ctx, int32_t atom_time_interval);

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

/* Allocation-free variant of rd_field_batch.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, errors and results as rd_field_batch, bit for
 * bit.  Every device buffer it needs beyond the context's particle store --
 * host staging, points, fields, partial sums, launch table, packed copies --
 * comes from one file-scope static device array of RD_STATIC_SCRATCH_BYTES
 * per device, held from the library's load on that device until the process
 * ends; the call makes no device or pinned-host allocation.  It needs about
 * 72 B per point plus 24 B per point per 256 particles (FP64 partial sums;
 * mixed: 32 B per point and 32-64 B per particle more), and returns
 * RD_ERR_CAPACITY beyond the array.  Every context on a device shares that
 * device's array, which keeps the last call's contents: calls that overlap on
 * one device race. */
#define RD_STATIC_SCRATCH_BYTES (64ull << 20)
int rd_field_batch_static(rd_ctx* ctx, int32_t m, const double* pts, double* field,
                          int32_t precision);

/* Field on the nx*ny cell centres of the rectangle [x0,x0+lx]x[y0,y0+ly] at
 * height z (points ordered x fastest), kept in HBM; copied to `field` (3,nx*ny)
 * when it is not NULL.  The surface grid of the force/field snapshot. */
int rd_field_grid(rd_ctx* ctx, double x0, double y0, double lx, double ly, int32_t nx,
                  int32_t ny, double z, int32_t precision, double* field);

/* ---- time step (the main loop, src/main.F90:175-219) ---------------------- */
/* Emitter of the current simulation (Init_Field_Emission_v2 / Init_Emission_Tip /
 * Init_Field_Thermo_Emission with Read_work_function). */
int rd_set_emitter(rd_ctx* ctx, const rd_emitter* em);
/* Seed of the counter-based generator: every random draw of the simulation
 * is a function of (seed, stream, step, purpose, chain), so results do not
 * depend on what else runs on the device. */
int rd_set_seed(rd_ctx* ctx, uint64_t seed, uint64_t stream);

/* Add_Particle (src/mod_pair.F90:29-159) for n particles at once: appends
 * them in order with the Beeman history seeded from the vacuum field.
 * pos, vel: (3,n); species: (n) or NULL for electrons; sec: (n) or NULL (1). */
int rd_add_particles(rd_ctx* ctx, int32_t n, const double* pos, const double* vel,
                     const int32_t* species, int32_t step, int32_t emit, const int32_t* sec);

/* Update_Position (src/mod_verlet.F90:115-162): Beeman position update with
 * boundary marking (Update_ElecIon_Position, :197-232), accelerations, Beeman
 * velocity update and Ramo current (Update_ElecIon_Velocity, :449-509). */
int rd_update_position(rd_ctx* ctx, int32_t step, int32_t precision, rd_step_record* rec);
/* Remove_Particles (src/mod_pair.F90:352-562): order-preserving compaction of
 * the particles marked in the last position update, lifetime histogram. */
int rd_remove_particles(rd_ctx* ctx, int32_t step, rd_step_record* rec);
/* ptr_Do_Emission: surface supply integral and Metropolis-Hastings site
 * sampling of the current emitter, then the escape test and Add_Particle
 * (Do_Field_Emission_Planar_rectangle, src/mod_field_emission_v2.F90:261-395;
 * Do_Field_Thermo_Emission_Planar_simple, src/mod_field_thermo_emission.F90:136-193;
 * Do_Field_Emission_Tip_OLDCODE, src/mod_emission_tip.f90:417-534). */
int rd_do_emission(rd_ctx* ctx, int32_t step, int32_t precision, rd_step_record* rec);
/* nsteps iterations of the main loop (emission, update, removal) after the
 * last completed step; records: nsteps entries (may be NULL). */
int rd_run(rd_ctx* ctx, int32_t nsteps, int32_t precision, rd_step_record* records);

/* Full particle state of the current simulation (any pointer may be NULL).
 * vec3 arrays (3,n); ints (5,n): species, birth step, emitter, section, id. */
int rd_get_state(rd_ctx* ctx, double* pos, double* prev_pos, double* vel, double* acc,
                 double* acc_prev, double* acc_prev2, int32_t* ints);
/* life_time histogram (src/mod_pair.F90:1135-1160): hist (1000, 2) for
 * electrons then ions, lifetime in steps 1..1000. */
int rd_get_life_time(rd_ctx* ctx, int32_t* hist);

/* ---- many simulations on one device (batched serving) --------------------- */
/* A pool: one particle store of total_particles shared by up to max_slots
 * independent simulations; every step of rd_pool_step advances all open
 * simulations with one launch per kernel.  rd_ctx_create is a pool of one
 * slot that spans the whole store. */
int rd_pool_create(int32_t device, int64_t total_particles, int32_t max_slots, rd_ctx** out);
/* HBM bytes a simulation of `capacity` particles holds in a pool. */
int rd_slot_footprint(int64_t capacity, int64_t* bytes);
/* Open a simulation with room for `capacity` particles if a contiguous range
 * of the store is free (RD_ERR_CAPACITY otherwise); it becomes current. */
int rd_slot_open(rd_ctx* ctx, int64_t capacity, int32_t* slot);
int rd_slot_close(rd_ctx* ctx, int32_t slot);
/* Make `slot` the simulation the single-simulation entry points act on. */
int rd_slot_select(rd_ctx* ctx, int32_t slot);
/* Last completed step of a simulation. */
int rd_slot_step(rd_ctx* ctx, int32_t slot, int32_t* step);
/* One main-loop step of every open simulation; records: one per slot index
 * (max_slots entries, may be NULL); slots that are not open are left untouched. */
int rd_pool_step(rd_ctx* ctx, int32_t precision, rd_step_record* records);

/* Launch the Metropolis-Hastings jump loop of each emission step (about
 * 400 kernels at 200 jumps) as one HIP graph (on != 0) or as separate
 * launches (0, the default).  The graph is captured when the loop's launch
 * parameters change and replayed while they do not; results are identical
 * either way.  Library control, no reference counterpart. */
int rd_set_graphs(rd_ctx* ctx, int32_t on);

/* Pool tuning: the mixed-precision jump launches of a step give each
 * workgroup four partial sums of the chain fields once they hold at least
 * min_workgroups workgroups of one partial each (2048 by default; 0: always).
 * The partials keep their size and order, so results are identical for every
 * threshold.  Library control, no reference counterpart. */
int rd_set_jump_pairing(rd_ctx* ctx, int64_t min_workgroups);

/* Mixed-precision jump loops of small simulations (on != 0, the default): in
 * a context made for one simulation (rd_ctx_create), a planar simulation whose
 * chains x electrons per jump are few (by its own sizes) runs all its
 * Metropolis-Hastings jumps in one workgroup of one launch instead of two
 * launches per jump; pools keep the two launches, which serve all their
 * simulations at once.  The chain fields add the same FP32 pair terms in
 * another fixed order, so the fields agree to rounding and the chains take
 * the same sites unless an accept test falls within that rounding.  Library
 * control, no reference counterpart. */
int rd_set_solo_jumps(rd_ctx* ctx, int32_t on);

/* Forms of the mixed-precision planar accelerations (flags, 0 = both on,
 * the default).  RD_FORCE_NO_SYMMETRIC: every simulation takes the gather
 * form (each ordered pair evaluated at its target) instead of the
 * pair-symmetric form from 16384 particles.  RD_FORCE_NO_FAR_EXPANSION: the
 * image partners n >= 2 of the planar series are summed pair by pair in
 * FP32 instead of as one FP64 expansion of their field over all particles.
 * Library control for measurement and testing, no reference counterpart;
 * FP64 accelerations are not affected. */
#define RD_FORCE_NO_SYMMETRIC 1
#define RD_FORCE_NO_FAR_EXPANSION 2
int rd_set_force_options(rd_ctx* ctx, int32_t flags);

/* Bound on the partial sums of one pair launch (bytes; 0 = the default,
 * 512 MiB).  The FP64 forces and fields keep one partial per target and
 * source chunk of 256, the pair-symmetric mixed forces one per particle and
 * block of 1024; a launch whose partials exceed the bound runs in bands whose
 * partials are added, in the same order, to running sums, so the memory of
 * the sums is O(N) and every result is the same for every bound.  Library
 * control, no reference counterpart. */
int rd_set_partial_budget(rd_ctx* ctx, int64_t bytes);

/* Device bytes the context's work buffers hold (partial sums, packed copies,
 * launch tables, chains; grown on demand, kept until rd_ctx_destroy).
 * Library query, no reference counterpart. */
int rd_work_bytes(rd_ctx* ctx, int64_t* bytes);

/* Collect the particle records of each step (on != 0) or not (0, the
 * default).  Library control, no reference counterpart: the reference
 * writes these records to files inside Add_Particle and
 * Mark_Particles_Remove. */
int rd_set_logs(rd_ctx* ctx, int32_t on);
/* The records of the last step of the selected simulation: up to max_emit
 * emission and max_absorb absorption records; n_emit / n_absorb receive the
 * step's counts (RD_ERR_CAPACITY when they exceed the maxima). */
int rd_get_logs(rd_ctx* ctx, int32_t max_emit, rd_emit_record* emit, int32_t* n_emit,
                int32_t max_absorb, rd_absorb_record* absorb, int32_t* n_absorb);
/* Recording planes of the selected simulation (planes_N and planes_z of the
 * input namelist, src/mod_global.F90:337-342, scaled to m as
 * src/main.F90:375 does): n <= RD_MAX_PLANES heights; a height <= 0 keeps
 * its number but records nothing, as Check_Planes skips it.  n = 0 (the
 * default) records none.  Crossings are collected while rd_set_logs is on. */
int rd_set_planes(rd_ctx* ctx, int32_t n, const double* z);
/* The plane crossings of the last step of the selected simulation (the
 * planes-k.bin writes of Check_Planes): up to max_records records, plane 1
 * first; n receives the count (RD_ERR_CAPACITY when it exceeds max_records). */
int rd_get_plane_logs(rd_ctx* ctx, int32_t max_records, rd_plane_record* records, int32_t* n);

/* ---- measurement ----------------------------------------------------------- */
/* Average device time (ms) of the accelerations / grid field over `reps`
 * launches, from HIP events on the context's stream; `kernel_ms` receives
 * the average of the pair kernel alone (may be NULL). */
int rd_time_accel(rd_ctx* ctx, int32_t precision, int32_t reps, double* total_ms,
                  double* kernel_ms);
int rd_time_field_grid(rd_ctx* ctx, double x0, double y0, double lx, double ly, int32_t nx,
                       int32_t ny, double z, int32_t precision, int32_t reps, double* total_ms,
                       double* kernel_ms);
/* Average device time (ms) of a plain copy of `bytes` bytes (multiple of 16,
 * at most 16 GiB) between two scratch buffers allocated for the call: the
 * streaming-bandwidth probe and the known-size kernel against which HBM
 * traffic counters are calibrated.  No reference counterpart. */
int rd_time_copy(rd_ctx* ctx, int64_t bytes, int32_t reps, double* kernel_ms);

#ifdef __cplusplus
}
#endif
#endif
