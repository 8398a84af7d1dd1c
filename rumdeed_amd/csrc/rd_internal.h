// Internal definitions of librumdeed.so: simulation slots, device state and
// the launch tables shared by the translation units.
//
// A context owns one set of particle arrays in HBM (structure of arrays,
// capacity `cap`) and up to kMaxSlots independent simulations ("slots"),
// each holding a contiguous index range [base, base + slot_cap).  Every
// kernel of the step is launched once for all active slots: blockIdx.z (or
// the block index) selects a launch-table entry, so many small simulations
// share each launch.  A single simulation is a context with one slot.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/rumdeed.h"

namespace rd {

// Physical constants as src/mod_global.F90:26-76 builds them (FP64).
struct Consts {
  static constexpr double pi = 3.141592653589793238462643383279502884197;
  static constexpr double h = 6.62607015e-34;
  static constexpr double k_b = 1.380649e-23;
  static constexpr double c = 299792458.0;
  static constexpr double mu_0 = 1.25663706212e-6;
  static constexpr double m_0 = 9.1093837015e-31;
  static constexpr double q_0 = 1.602176634e-19;
  static constexpr double m_u = 1.66053906660e-27;
  static constexpr double length_scale = 1.0e-9;
  static constexpr double time_scale = 1.0e-12;
};
double epsilon_0();
double div_fac_c();  // 1/(4 pi epsilon_0 epsilon_r)
double eps_soft();   // length_scale**2, the r softening of the pair sums
double k_tip();      // 1*q_0/(4 pi epsilon_0), inside the sphere image force
double m_N2p();

// Geometry in the form the kernels use.
struct DevGeom {
  int kind, ic, nic, pad;
  double d, Ez, es, dfc, ktip;
  double box_top;  // removal height (box_dim(3))
  // hyperboloid tip (Init_Emission_Tip, src/mod_emission_tip.f90:105-125)
  double a_foci, shift_z, pre_E, pre_E_unit, z0, R, eta_1, h_tip, d_tip, R_base, max_xi, theta;
};

// Read-only description of one simulation slot.
struct SimDesc {
  DevGeom g;
  double dt, dt2, dt_q0;
  int mode;   // RD_EMIT_*
  int delay;  // emitters_delay
  double ex0, ey0, elx, ely;  // planar emitter rectangle
  double T;                   // cathode temperature (GTF)
  double w_tip;               // tip work function
  int wf_nx, wf_ny, wf_off;   // checkerboard table in ctx->wf
  int quad;                   // requested supply points per side (planar)
  int qgx, qgy;               // level-0 supply points per side (x, y): multiples
                              // of the work function's sections (quad_side)
  int quad_lmax;              // refinement doublings allowed (planar)
  int tts;                    // atom_time_interval of the two-time-step mode; 0: off
  int base, cap;
  int ndim;                   // Metropolis-Hastings jumps per chain
  int nplanes;                // recording planes (Check_Planes)
  uint64_t seed, stream;
  double planes[RD_MAX_PLANES];  // plane heights (m); <= 0: unused
};

// Mutable per-slot state, resident in HBM.
struct SimState {
  int n;          // live particles
  int nrID;       // next particle id
  int n_round;    // chains of the current step
  int emitted;    // electrons emitted this step
  int rem_total, rem_top, rem_bot, rem_elec, rem_ion, rem_top_emit;
  int dropped;    // refused for capacity (cumulative)
  int pending;    // chains still without a start position
  int it_a, it_r; // accept / reject count of the current jump
  int start_rounds;
  int n_elec;     // electrons in the last velocity update
  double residual;  // rounding residual of the supply (planar FE)
  double MH_std, a_rate;
  double N_sup, df_avg;
  double F_avg[3];
  double ramo[3];   // Ramo current per species (A)
  double vsum[9];   // velocity sums: all, electrons, ions
  int plane_n[RD_MAX_PLANES];  // plane crossings of the last step
  int first_dead;     // lowest index marked for removal since the counters' reset
  double qsum[2][4];  // supply sums (value, f_x, f_y, f_z) of levels of parity 0 / 1
  int q_final;        // the supply's last computed level is its result
  int q_level;        // that level
  int q_cells;        // its cells
  int q_fail;         // the last allowed level did not meet the tolerance
  double q_err;       // |difference| of the last two levels (electrons)
};

// One launch-table entry (per active slot and launch).
struct Ent {
  int slot, base, n_ub, nchunks_ub;
  int cps, segs;
  long long part_off;
  int pbase, m;   // field points [pbase, pbase + m)
  int cbase, nch; // MH chains [cbase, cbase + nch)
  int step, nblk; // time step; per-slot block count (reductions)
  int wg0, wg1;   // first workgroup of the entry in a compact launch (two layouts)
  int nic_pair;   // image partners n <= nic_pair in the mixed pair kernels (the
                  // rest, when fewer than N_ic_max, come from rd_tail.hip)
  int lev;        // refinement level of the supply points (quad << lev per side)
  int tts;        // accelerations of the two-time-step update: atom_time_interval;
                  // 0: the standard accelerations (Calculate_Acceleration_Particles)
  // Banded launches (partial sums bounded by rd_ctx::part_budget).  FP64: the
  // launch takes source chunks [band0, band1), its partial slot k holding
  // chunk band0 + k; pair-symmetric: the launch takes block rows
  // [band0, band1), the reaction partials of row X in slot X - band0 at
  // part_off, the direct partials of the band's particles at d_off.
  int band0, band1;
  int pslots;        // partial slots per component at part_off
  int banded;        // the sums come from acc (k_fold64 / k_fold_sym), not one partial set
  int d_ld;          // leading dimension of the direct partials (pair-symmetric)
  long long acc_off; // running sums [3][targets] at rd_ctx::acc + acc_off
  long long d_off;   // direct partials [3][segs][d_ld] (pair-symmetric, banded)
  // Compaction by chunks (k_compact_count / _move / _back): the entry's header
  // and chunk counts from int2 slot c_off of rd_ctx::ctmp, its survivors staged
  // from particle t_off of the staging arrays behind them.  Supply levels
  // (k_quad_tiles / _tsum): c_off is the entry's first tile-sum slot.
  long long c_off, t_off;
};

// Entry of a compact launch whose workgroups [wg0, next wg0) belong to one
// entry: the last entry with wg0 <= b (wg1 for the second layout).  The
// offsets are non-decreasing, so the entries that pass form a prefix: the
// first wavefront tests 64 entries per round with independent loads and
// broadcasts the answer through LDS.  Every thread of the block calls it, in
// uniform control flow (it may hold a barrier), and the block has >= 64
// lanes.
__device__ __forceinline__ int find_entry(const Ent* tab, int ne, int b, bool second = false) {
  __shared__ int s_entry;
  if (ne == 1) return 0;  // uniform over the launch
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    int best = 0;
    for (int k0 = 0; k0 < ne; k0 += 64) {
      const int k = k0 + lane;
      const bool ok = k < ne && (second ? tab[k].wg1 : tab[k].wg0) <= b;
      const unsigned long long m = __ballot(ok);
      if (m) best = k0 + 63 - __clzll((long long)m);
      if (m != ~0ull) break;
    }
    if (lane == 0) s_entry = best;
  }
  __syncthreads();
  return s_entry;
}

constexpr int kChunk = 256;  // canonical FP64 summation chunk
constexpr int kJumpStreams = 2;  // streams of a pool's jump loop (3: the same makespan; 4: slower)
constexpr int kBlock = 256;
constexpr int kMaxSlots = 4096;
constexpr int kLife = 1000;  // MAX_LIFE_TIME

struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct Slot {
  bool active = false;
  SimDesc desc{};
  int n = 0;         // host mirror of state.n after the last sync
  int step = 0;      // last completed step
  int emit_step = -1;  // step of the last emission phase (its counters stand)
  rd_geometry geom{};
  bool geom_set = false;
};

}  // namespace rd

// Jump workgroups (one partial each) from which a pool's mixed jump launches
// give every workgroup kJumpPairs partials, paying the proposal and set-up
// once for all of them (measured on the 256-request serving stream: 66.2 s
// with 3 chunks per workgroup, 57.6 s with 6, 54.5 s with 12; below this size
// the pool is latency-bound and more, smaller workgroups are faster).
constexpr int kJumpPairWgs = 2048;
constexpr int kJumpPairs = 4;

// Partial sums of one pair launch are bounded by this many bytes (512 MiB);
// larger launches run in bands whose partials are folded, in order, into
// running sums, with the results of the unbanded order.
constexpr size_t kPartBudget = size_t(512) << 20;

struct rd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int64_t cap = 0;
  int nslots_max = 1;
  std::vector<rd::Slot> slots;

  // particle state, structure of arrays in HBM
  double *x = nullptr, *y = nullptr, *z = nullptr;
  double *px = nullptr, *py = nullptr, *pz = nullptr;  // previous position
  double *vx = nullptr, *vy = nullptr, *vz = nullptr;
  double *ax = nullptr, *ay = nullptr, *az = nullptr;     // current acceleration
  double *a1x = nullptr, *a1y = nullptr, *a1z = nullptr;  // previous
  double *a2x = nullptr, *a2y = nullptr, *a2z = nullptr;  // previous-previous
  double *q = nullptr, *m = nullptr;
  int32_t *species = nullptr, *pstep = nullptr, *emitter = nullptr, *section = nullptr,
          *id = nullptr, *alive = nullptr;

  rd::SimDesc* d_desc = nullptr;
  rd::SimState* d_state = nullptr;
  rd::SimState* h_state = nullptr;  // pinned mirror
  int32_t* d_life = nullptr;        // [slot][species 1..2][kLife]
  double* d_wf = nullptr;           // work-function tables
  int wf_used = 0, wf_cap = 0;

  // work buffers (grown on demand, never shrunk)
  rd::Buf part, pack_hi, pack_lo, pts, fld, pts_hi, pts_lo, stage, img64, tab, red, chains;
  rd::Buf pack_far;  // far-partner series of the particles (surface_far_coeffs)
  rd::Buf pack_srf;  // (x, y, z, q z) high parts of the particles: cathode-surface sums
  rd::Buf rnd;  // random numbers of the Metropolis-Hastings jumps of a step
  rd::Buf cnt;  // accept / reject counts per accept workgroup of a jump
  rd::Buf img_hi, img_lo;  // packed sphere images of the particles (mixed, tip)
  // far image partners as an expansion (rd_tail.hip): the plan of the last
  // (d, N_ic_max), work arrays, lateral boxes; force options (rd_set_force_options)
  rd::Buf tail_plan, tail_work, tail_box;
  // running sums of banded launches and the partial-sum bound of one band
  // (rd_set_partial_budget)
  rd::Buf acc;
  rd::Buf ctmp;  // survivors of the simulations compacted by chunks (kCompactSplit)
  size_t part_budget = kPartBudget;
  size_t work_bytes = 0;  // device bytes held by the work buffers (ensure)
  double tail_d = -1.0;
  int tail_nic = -1, tail_nnz = 0;
  int force_opts = 0;
  // particle records of the step (rd_set_logs): velocity and reason at
  // marking (4 doubles a particle), emission and absorption records
  int logs = 0;
  rd::Buf mark, elog, alog, plog;  // plog: RD_MAX_PLANES regions of cap records
  rd::Ent* h_tab = nullptr;  // pinned staging of launch tables (ring)
  int tab_used = 0;          // ring entries used since the last stream synchronisation
  int tab_last = -1;         // ring index of the last upload
  size_t tab_last_n = 0;
  int cur = 0;               // slot the single-simulation entry points act on

  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;

  // Metropolis-Hastings jump loop as a HIP graph (rd_set_graphs): the last
  // captured loop and the launch parameters it was captured with; replayed
  // while they are unchanged.
  int graphs = 0;
  hipGraphExec_t jump_exec = nullptr;
  // Jump workgroups from which a pool's mixed jump launches give each
  // workgroup kJumpPairs partials (rd_set_jump_pairing; kJumpPairWgs by default).
  long long jump_pair_wgs = kJumpPairWgs;
  std::vector<unsigned char> jump_key;

  // further streams of the pool jump loop (default priority) and the events
  // that fork them from and join them to `stream`
  hipStream_t jstream[rd::kJumpStreams - 1] = {};
  hipEvent_t ev_fork = nullptr, ev_join[rd::kJumpStreams - 1] = {};
  bool solo_jumps = true;          // small simulations' jumps in one workgroup (mixed)
  hipStream_t sstream = nullptr;  // one-workgroup jump loops of small simulations
  hipEvent_t ev_solo = nullptr;
};

namespace rd {
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);
int ensure(rd_ctx* c, Buf& b, size_t bytes);
int build_geom(const rd_geometry* in, DevGeom* g);

// Pair kernels (rd_pair.hip).  `tab` is a device launch table of `ne`
// entries; all launches go to c->stream.
// Accelerations of the particles of every entry (n from the slot state).
int launch_accel(rd_ctx* c, const std::vector<Ent>& ents, int precision, hipEvent_t k0,
                 hipEvent_t k1);
// Device memory given to a launch instead of the context's work buffers:
// points and results SoA (stride `mtot`), and [next, end) for everything else
// the launch needs (launch table, partial sums, packed copies), handed out in
// 256-byte aligned pieces.  The launch allocates nothing and fails with
// RD_ERR_CAPACITY when [next, end) is too small.
struct FixedBufs {
  double* pts;
  double* fld;
  char* next;
  char* end;
  void* take(size_t bytes) {
    const size_t a = (bytes + 255) & ~(size_t)255;
    if ((size_t)(end - next) < a) return nullptr;
    void* p = next;
    next += a;
    return p;
  }
};
// Field at the points [pbase, pbase + m) of every entry from its particles,
// points and results SoA in c->pts / c->fld (stride `mtot`), or in fixed's
// buffers when `fixed` is given.
int launch_field(rd_ctx* c, const std::vector<Ent>& ents, int mtot, int precision, hipEvent_t k0,
                 hipEvent_t k1, FixedBufs* fixed = nullptr);
int upload_tab(rd_ctx* c, const std::vector<Ent>& ents, Ent** dtab);

// Level-0 points per side of the planar supply along an axis the work
// function splits into `sections` equal parts: the least multiple of
// `sections` not below quad, so that every cell of the grid (and of each
// refinement) lies in one section and the integrand is smooth on it.
inline int quad_side(int quad, int sections) {
  if (sections < 1) sections = 1;
  return sections * ((quad + sections - 1) / sections);
}
// Doublings of the planar supply rule from g0 points on the longer side:
// levels < 0 none, 0 four, else `levels`; never beyond 1024 points per side.
inline int quad_lmax(int g0, int levels) {
  int l = levels < 0 ? 0 : (levels == 0 ? 4 : levels);
  while (l > 0 && (g0 << l) > 1024) --l;
  return l;
}
// Packed mixed-precision copies (c->pack_hi / pack_lo, and c->pack_srf when
// `surface`) of every entry's particles.
int launch_pack(rd_ctx* c, const std::vector<Ent>& ents, bool surface = false);
// Split nchunks source chunks into `segs` (at most max_segs) segments of
// `cps` chunks each.
void segment(int nchunks, int max_segs, int& segs, int& cps);
// Segments for a simulation whose targets fill `target_blocks` workgroups:
// about `want` workgroups for that simulation.
inline int seg_limit(int target_blocks, int want) {
  const int tb = target_blocks > 0 ? target_blocks : 1;
  return (want + tb - 1) / tb;
}
constexpr int kWantPair = 8192;  // accelerations, field at points
constexpr int kWantJump = 512;   // fields of the Metropolis-Hastings chains
constexpr int kPoolJumpChunks = 3;  // least source chunks per jump partial in a pool
int launch_grid_points(rd_ctx* c, double x0, double y0, double lx, double ly, int nx, int ny,
                       double zz);

// Far image partners (n >= 2) of the mixed planar accelerations (rd_tail.hip).
struct TailUse {
  int ent;             // index into the entries of the launch
  double cx, cy, cz;   // expansion centre (nm)
};
bool tail_wanted(const rd_ctx* c, const DevGeom& g, int precision);
int tail_decide(rd_ctx* c, std::vector<Ent>& ents, int precision, std::vector<TailUse>& use);
int tail_apply(rd_ctx* c, const std::vector<Ent>& ents, const std::vector<TailUse>& use);

// Engine (rd_engine.hip).
int engine_step(rd_ctx* c, int precision, rd_step_record* rec_per_slot);
int engine_add(rd_ctx* c, int slot, int n, const double* pos, const double* vel,
               const int32_t* species, int step, int emit, const int32_t* sec);
int engine_position(rd_ctx* c, const std::vector<Ent>& ents);
int engine_velocity(rd_ctx* c, const std::vector<Ent>& ents);
int engine_compact(rd_ctx* c, const std::vector<Ent>& ents);
int sync_state(rd_ctx* c);
}  // namespace rd

#define RD_HIP(call)                                      \
  do {                                                    \
    hipError_t _e = (call);                               \
    if (_e != hipSuccess) return rd::hip_fail(_e, #call); \
  } while (0)
