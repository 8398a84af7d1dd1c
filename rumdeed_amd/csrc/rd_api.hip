// C ABI of librumdeed.so: pools and simulation slots, particle store, force
// and field entry points, the time step, timing.  See include/rumdeed.h for
// the contract of each function.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "rd_internal.h"

namespace rd {

static thread_local std::string g_err;

double epsilon_0() { return 1.0 / (Consts::mu_0 * (Consts::c * Consts::c)); }
double div_fac_c() { return 1.0 / (4.0 * Consts::pi * epsilon_0() * 1.0); }
double eps_soft() { return Consts::length_scale * Consts::length_scale; }
double k_tip() { return 1.0 * Consts::q_0 / (4.0 * Consts::pi * epsilon_0()); }
double m_N2p() { return 28.0134 * Consts::m_u - Consts::m_0; }

void set_error(const std::string& msg) { g_err = msg; }
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
int hip_fail(hipError_t e, const char* what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return RD_ERR_HIP;
}

int ensure(rd_ctx* c, Buf& b, size_t bytes) {
  if (bytes <= b.bytes) return RD_OK;
  if (b.p) {
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    hipFree(b.p);
    c->work_bytes -= b.bytes;
    b.p = nullptr;
    b.bytes = 0;
  }
  size_t want = bytes + bytes / 4 + 4096;
  hipError_t e = hipMalloc(&b.p, want);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc(work buffer)");
  b.bytes = want;
  c->work_bytes += want;
  return RD_OK;
}

static void free_buf(Buf& b) {
  if (b.p) hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

// Device-side transposes between the (3,n) host layout and the SoA store.
__global__ void k_aos_to_soa(int n, const double* __restrict__ a, double* __restrict__ x,
                             double* __restrict__ y, double* __restrict__ z) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = a[3 * i];
  y[i] = a[3 * i + 1];
  z[i] = a[3 * i + 2];
}
__global__ void k_soa_to_aos(int n, const double* __restrict__ x, const double* __restrict__ y,
                             const double* __restrict__ z, double* __restrict__ a) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  a[3 * i] = x[i];
  a[3 * i + 1] = y[i];
  a[3 * i + 2] = z[i];
}
__global__ void k_fill(int n, double v, double* __restrict__ a) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}
__global__ void k_fill_i(int n, int32_t v, int32_t* __restrict__ a) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}
__global__ void k_iota(int n, int32_t v0, int32_t* __restrict__ a) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v0 + i;
}

static int upload_vec3(rd_ctx* c, int n, const double* host, double* x, double* y, double* z) {
  if (n == 0) return RD_OK;
  int r = ensure(c, c->stage, sizeof(double) * 3 * (size_t)n);
  if (r) return r;
  RD_HIP(hipMemcpyAsync(c->stage.p, host, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
                        c->stream));
  k_aos_to_soa<<<(n + 255) / 256, 256, 0, c->stream>>>(n, (const double*)c->stage.p, x, y, z);
  RD_HIP(hipGetLastError());
  RD_HIP(hipStreamSynchronize(c->stream));
  return RD_OK;
}

static int download_vec3(rd_ctx* c, int n, const double* x, const double* y, const double* z,
                         double* host) {
  if (n == 0) return RD_OK;
  int r = ensure(c, c->stage, sizeof(double) * 3 * (size_t)n);
  if (r) return r;
  k_soa_to_aos<<<(n + 255) / 256, 256, 0, c->stream>>>(n, x, y, z, (double*)c->stage.p);
  RD_HIP(hipGetLastError());
  RD_HIP(hipMemcpyAsync(host, c->stage.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost,
                        c->stream));
  RD_HIP(hipStreamSynchronize(c->stream));
  return RD_OK;
}

static int fill(rd_ctx* c, int n, double v, double* a) {
  if (n == 0) return RD_OK;
  k_fill<<<(n + 255) / 256, 256, 0, c->stream>>>(n, v, a);
  RD_HIP(hipGetLastError());
  return RD_OK;
}

// Derived geometry, as Read_Input_Variables (src/main.F90:350-356) and
// Init_Emission_Tip (src/mod_emission_tip.f90:105-125) compute it.
int build_geom(const rd_geometry* in, DevGeom* g) {
  std::memset(g, 0, sizeof(*g));
  g->kind = in->kind;
  g->ic = in->image_charge ? 1 : 0;
  g->nic = in->n_ic_max;
  g->es = eps_soft();
  g->dfc = div_fac_c();
  g->ktip = k_tip();
  if (in->kind == RD_GEOM_PLANAR) {
    if (!(in->d > 0.0)) return fail(RD_ERR_ARG, "planar gap d must be > 0");
    if (in->n_ic_max < 0 || in->n_ic_max > 64) return fail(RD_ERR_ARG, "n_ic_max outside 0..64");
    g->d = in->d;
    g->Ez = -1.0 * in->V / in->d;
  } else if (in->kind == RD_GEOM_TIP) {
    double dtip = in->d_tip, rb = in->R_base, ht = in->h_tip;
    if (!(dtip > 0.0 && rb > 0.0 && ht > 0.0))
      return fail(RD_ERR_ARG, "tip d_tip, R_base, h_tip must be > 0");
    g->d_tip = dtip;
    g->R_base = rb;
    g->h_tip = ht;
    g->d = dtip + ht;
    g->max_xi = ht / dtip + 1.0;
    g->a_foci = std::sqrt(dtip * dtip * (rb * rb) / (ht * ht + 2 * dtip * ht) + dtip * dtip);
    g->eta_1 = -1.0 * dtip / g->a_foci;
    g->theta = std::acos(dtip / g->a_foci);
    double r_tip = g->a_foci * std::sin(g->theta) * std::tan(g->theta);
    g->shift_z = std::fabs(g->a_foci * g->eta_1 * g->max_xi);
    double eta_2 = 0.0;
    double pf = std::log((1.0 + g->eta_1) / (1.0 - g->eta_1) * (1.0 - eta_2) / (1.0 + eta_2));
    g->pre_E_unit = 2.0 * 1.0 / (g->a_foci * pf);
    g->pre_E = 2.0 * in->V / (g->a_foci * pf);
    g->R = r_tip;
    g->z0 = ht - r_tip;
    g->Ez = 0.0;
  } else {
    return fail(RD_ERR_ARG, "unknown geometry kind");
  }
  g->box_top = in->box_top > 0.0 ? in->box_top : g->d;
  return RD_OK;
}

static int put_desc(rd_ctx* c, int s) {
  RD_HIP(hipMemcpyAsync(c->d_desc + s, &c->slots[s].desc, sizeof(SimDesc), hipMemcpyHostToDevice,
                        c->stream));
  RD_HIP(hipStreamSynchronize(c->stream));
  return RD_OK;
}

// Write the host copy of one slot's state to the device (after a sync).
static int put_state(rd_ctx* c, int s) {
  RD_HIP(hipMemcpyAsync(c->d_state + s, c->h_state + s, sizeof(SimState), hipMemcpyHostToDevice,
                        c->stream));
  RD_HIP(hipStreamSynchronize(c->stream));
  return RD_OK;
}

static void default_state(SimState* S) {
  std::memset(S, 0, sizeof(*S));
  S->MH_std = 0.0125;  // mod_field_emission_v2.F90:56
  S->a_rate = 1.0;
}

static int slot_range_free(rd_ctx* c, int64_t cap, int* base) {
  // first fit over the ranges of the open slots, sorted by base
  std::vector<std::pair<int, int>> used;
  for (const Slot& s : c->slots)
    if (s.active) used.push_back({s.desc.base, s.desc.cap});
  std::sort(used.begin(), used.end());
  int64_t pos = 0;
  for (auto& u : used) {
    if (u.first - pos >= cap) break;
    pos = std::max<int64_t>(pos, (int64_t)u.first + u.second);
  }
  if (pos + cap > c->cap) return RD_ERR_CAPACITY;
  *base = (int)pos;
  return RD_OK;
}

static int open_slot(rd_ctx* c, int64_t cap, int32_t* out) {
  if (cap < 1) return fail(RD_ERR_ARG, "slot capacity must be >= 1");
  int idx = -1;
  for (int s = 0; s < c->nslots_max; ++s)
    if (!c->slots[s].active) {
      idx = s;
      break;
    }
  if (idx < 0) return fail(RD_ERR_CAPACITY, "all simulation slots are in use");
  int base = 0;
  if (slot_range_free(c, cap, &base))
    return fail(RD_ERR_CAPACITY, "no contiguous free range of the particle store");
  int r = sync_state(c);
  if (r) return r;
  Slot& sl = c->slots[idx];
  sl = Slot();
  sl.active = true;
  std::memset(&sl.desc, 0, sizeof(SimDesc));
  sl.desc.base = base;
  sl.desc.cap = (int)cap;
  sl.desc.quad = 32;
  sl.desc.qgx = sl.desc.qgy = 32;
  sl.desc.quad_lmax = quad_lmax(32, 0);
  sl.desc.ndim = 200;
  sl.desc.w_tip = 4.7;
  sl.desc.wf_nx = 0;
  sl.n = 0;
  sl.step = 0;
  default_state(&c->h_state[idx]);
  r = put_state(c, idx);
  if (!r) r = put_desc(c, idx);
  if (r) return r;
  // zero the lifetime histogram of the slot
  int32_t* life = c->d_life + (size_t)idx * 2 * kLife;
  k_fill_i<<<(2 * kLife + 255) / 256, 256, 0, c->stream>>>(2 * kLife, 0, life);
  RD_HIP(hipStreamSynchronize(c->stream));
  c->cur = idx;
  *out = idx;
  return RD_OK;
}

static rd::Slot* cur_slot(rd_ctx* c) {
  if (c->cur < 0 || c->cur >= c->nslots_max || !c->slots[c->cur].active) return nullptr;
  return &c->slots[c->cur];
}

static Ent slot_ent(rd_ctx* c, int s) {
  Ent e{};
  e.slot = s;
  e.base = c->slots[s].desc.base;
  e.n_ub = c->slots[s].n;
  e.step = c->slots[s].step;
  return e;
}

int engine_phases(rd_ctx* c, const std::vector<int>& slots, int precision, int phases,
                  const int* steps, rd_step_record* recs);

}  // namespace rd

using namespace rd;

#define RD_CUR(c)                                                        \
  if (!(c)) return fail(RD_ERR_ARG, "null context");                     \
  rd::Slot* sl = cur_slot(c);                                            \
  if (!sl) return fail(RD_ERR_STATE, "no open simulation slot selected"); \
  RD_HIP(hipSetDevice((c)->device))

extern "C" {

const char* rd_version(void) { return "rumdeed-mi355x 0.2 (gfx950)"; }
const char* rd_last_error(void) { return g_err.c_str(); }

int rd_device_count(int32_t* n) {
  if (!n) return fail(RD_ERR_ARG, "null pointer");
  int d = 0;
  hipError_t e = hipGetDeviceCount(&d);
  if (e != hipSuccess) {
    *n = 0;
    return hip_fail(e, "hipGetDeviceCount");
  }
  *n = d;
  return RD_OK;
}

int rd_slot_footprint(int64_t capacity, int64_t* bytes) {
  if (!bytes || capacity < 0) return fail(RD_ERR_ARG, "bad argument");
  // 20 FP64 + 6 int32 per particle, packed copies (32 B), tip image data
  // (32 B), and the FP64 chunk partials 3 * ceil(n/256) * n * 8 B.
  int64_t per = 20 * 8 + 6 * 4 + 32 + 32;
  int64_t nch = (capacity + kChunk - 1) / kChunk;
  *bytes = capacity * per + 3 * nch * capacity * 8 + (int64_t)sizeof(SimState) + 2 * kLife * 4;
  return RD_OK;
}

int rd_pool_create(int32_t device, int64_t total, int32_t max_slots, rd_ctx** out) {
  if (!out) return fail(RD_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (total < 1 || total > (int64_t)1 << 30) return fail(RD_ERR_ARG, "max_particles outside 1..2^30");
  if (max_slots < 1 || max_slots > kMaxSlots) return fail(RD_ERR_ARG, "max_slots outside 1..4096");
  int nd = 0;
  hipError_t e = hipGetDeviceCount(&nd);
  if (e != hipSuccess || nd == 0) return fail(RD_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= nd) return fail(RD_ERR_ARG, "device index out of range");
  RD_HIP(hipSetDevice(device));
  rd_ctx* c = new rd_ctx();
  c->device = device;
  c->cap = total;
  c->nslots_max = max_slots;
  c->slots.assign(max_slots, Slot());
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return hip_fail(e, "hipStreamCreate");
  }
  double** d_arrays[] = {&c->x,   &c->y,   &c->z,   &c->px,  &c->py,  &c->pz,  &c->vx,
                         &c->vy,  &c->vz,  &c->ax,  &c->ay,  &c->az,  &c->a1x, &c->a1y,
                         &c->a1z, &c->a2x, &c->a2y, &c->a2z, &c->q,   &c->m};
  int32_t** i_arrays[] = {&c->species, &c->pstep, &c->emitter, &c->section, &c->id, &c->alive};
  size_t cap = (size_t)total;
  for (double** p : d_arrays) {
    e = hipMalloc((void**)p, sizeof(double) * cap);
    if (e != hipSuccess) {
      rd_ctx_destroy(c);
      return hip_fail(e, "hipMalloc(particle store)");
    }
  }
  for (int32_t** p : i_arrays) {
    e = hipMalloc((void**)p, sizeof(int32_t) * cap);
    if (e != hipSuccess) {
      rd_ctx_destroy(c);
      return hip_fail(e, "hipMalloc(particle store)");
    }
  }
  e = hipMalloc((void**)&c->d_desc, sizeof(SimDesc) * max_slots);
  if (e == hipSuccess) e = hipMalloc((void**)&c->d_state, sizeof(SimState) * max_slots);
  if (e == hipSuccess) e = hipHostMalloc((void**)&c->h_state, sizeof(SimState) * max_slots);
  if (e == hipSuccess) e = hipMalloc((void**)&c->d_life, sizeof(int32_t) * 2 * kLife * max_slots);
  if (e != hipSuccess) {
    rd_ctx_destroy(c);
    return hip_fail(e, "hipMalloc(slot state)");
  }
  std::memset(c->h_state, 0, sizeof(SimState) * max_slots);
  hipEventCreate(&c->ev0);
  hipEventCreate(&c->ev1);
  hipEventCreate(&c->ev2);
  c->cur = -1;
  *out = c;
  return RD_OK;
}

int rd_ctx_create(int32_t device, int64_t max_particles, rd_ctx** out) {
  int r = rd_pool_create(device, max_particles, 1, out);
  if (r) return r;
  int32_t s;
  r = open_slot(*out, max_particles, &s);
  if (r) {
    rd_ctx_destroy(*out);
    *out = nullptr;
  }
  return r;
}

int rd_ctx_destroy(rd_ctx* c) {
  if (!c) return RD_OK;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  double* d_arrays[] = {c->x,   c->y,   c->z,   c->px,  c->py,  c->pz,  c->vx,
                        c->vy,  c->vz,  c->ax,  c->ay,  c->az,  c->a1x, c->a1y,
                        c->a1z, c->a2x, c->a2y, c->a2z, c->q,   c->m};
  for (double* p : d_arrays)
    if (p) hipFree(p);
  int32_t* i_arrays[] = {c->species, c->pstep, c->emitter, c->section, c->id, c->alive};
  for (int32_t* p : i_arrays)
    if (p) hipFree(p);
  if (c->d_desc) hipFree(c->d_desc);
  if (c->d_state) hipFree(c->d_state);
  if (c->d_life) hipFree(c->d_life);
  if (c->d_wf) hipFree(c->d_wf);
  if (c->jump_exec) hipGraphExecDestroy(c->jump_exec);
  if (c->h_state) hipHostFree(c->h_state);
  if (c->h_tab) hipHostFree(c->h_tab);
  Buf* bufs[] = {&c->part, &c->pack_hi, &c->pack_lo, &c->pack_far, &c->pack_srf, &c->pts, &c->fld, &c->pts_hi, &c->pts_lo,
                 &c->stage, &c->img64, &c->tab, &c->red, &c->chains, &c->rnd, &c->img_hi, &c->img_lo, &c->cnt,
                 &c->mark, &c->elog, &c->alog, &c->plog, &c->tail_plan, &c->tail_work,
                 &c->tail_box, &c->acc, &c->ctmp};
  for (Buf* b : bufs) free_buf(*b);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  if (c->ev2) hipEventDestroy(c->ev2);
  if (c->ev_fork) hipEventDestroy(c->ev_fork);
  for (int h = 0; h < kJumpStreams - 1; ++h) {
    if (c->ev_join[h]) hipEventDestroy(c->ev_join[h]);
    if (c->jstream[h]) hipStreamDestroy(c->jstream[h]);
  }
  if (c->ev_solo) hipEventDestroy(c->ev_solo);
  if (c->sstream) hipStreamDestroy(c->sstream);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return RD_OK;
}

int rd_slot_open(rd_ctx* c, int64_t capacity, int32_t* slot) {
  if (!c || !slot) return fail(RD_ERR_ARG, "null pointer");
  RD_HIP(hipSetDevice(c->device));
  return open_slot(c, capacity, slot);
}

int rd_slot_close(rd_ctx* c, int32_t s) {
  if (!c) return fail(RD_ERR_ARG, "null context");
  if (s < 0 || s >= c->nslots_max || !c->slots[s].active) return fail(RD_ERR_ARG, "slot not open");
  RD_HIP(hipSetDevice(c->device));
  RD_HIP(hipStreamSynchronize(c->stream));
  c->slots[s].active = false;
  if (c->cur == s) c->cur = -1;
  return RD_OK;
}

int rd_slot_select(rd_ctx* c, int32_t s) {
  if (!c) return fail(RD_ERR_ARG, "null context");
  if (s < 0 || s >= c->nslots_max || !c->slots[s].active) return fail(RD_ERR_ARG, "slot not open");
  c->cur = s;
  return RD_OK;
}

int rd_slot_step(rd_ctx* c, int32_t s, int32_t* step) {
  if (!c || !step) return fail(RD_ERR_ARG, "null pointer");
  if (s < 0 || s >= c->nslots_max || !c->slots[s].active) return fail(RD_ERR_ARG, "slot not open");
  *step = c->slots[s].step;
  return RD_OK;
}

int rd_set_geometry(rd_ctx* c, const rd_geometry* geom) {
  RD_CUR(c);
  if (!geom) return fail(RD_ERR_ARG, "null geometry");
  DevGeom g;
  int r = build_geom(geom, &g);
  if (r) return r;
  sl->desc.g = g;
  sl->geom = *geom;
  sl->geom_set = true;
  return put_desc(c, c->cur);
}

int rd_set_time_step(rd_ctx* c, double dt) {
  RD_CUR(c);
  if (!(dt > 0.0)) return fail(RD_ERR_ARG, "time step must be > 0");
  sl->desc.dt = dt;
  sl->desc.dt2 = dt * dt;
  sl->desc.dt_q0 = dt / Consts::q_0;
  return put_desc(c, c->cur);
}

int rd_set_two_time_step(rd_ctx* c, int32_t atom_time_interval) {
  RD_CUR(c);
  if (atom_time_interval < 0 || atom_time_interval > (1 << 20))
    return fail(RD_ERR_ARG, "atom_time_interval outside 0..2^20");
  sl->desc.tts = atom_time_interval;
  return put_desc(c, c->cur);
}

int rd_set_seed(rd_ctx* c, uint64_t seed, uint64_t stream) {
  RD_CUR(c);
  sl->desc.seed = seed;
  sl->desc.stream = stream;
  return put_desc(c, c->cur);
}

int rd_set_emitter(rd_ctx* c, const rd_emitter* em) {
  RD_CUR(c);
  if (!em) return fail(RD_ERR_ARG, "null emitter");
  SimDesc& d = sl->desc;
  int mode = em->mode;
  if (mode != RD_EMIT_NONE && mode != RD_EMIT_PLANAR_FE && mode != RD_EMIT_PLANAR_GTF &&
      mode != RD_EMIT_TIP_FE)
    return fail(RD_ERR_ARG, "unknown emission mode");
  if (mode != RD_EMIT_NONE && !sl->geom_set) return fail(RD_ERR_STATE, "set the geometry first");
  if (mode == RD_EMIT_TIP_FE && d.g.kind != RD_GEOM_TIP)
    return fail(RD_ERR_ARG, "tip emission needs the tip geometry");
  if ((mode == RD_EMIT_PLANAR_FE || mode == RD_EMIT_PLANAR_GTF) && d.g.kind != RD_GEOM_PLANAR)
    return fail(RD_ERR_ARG, "planar emission needs the planar geometry");
  if (mode == RD_EMIT_PLANAR_FE || mode == RD_EMIT_PLANAR_GTF) {
    if (!(em->lx > 0.0 && em->ly > 0.0)) return fail(RD_ERR_ARG, "emitter size must be > 0");
    int nx = em->wf_nx, ny = em->wf_ny;
    if (nx < 1 || ny < 1 || nx * ny > 96 * 96 || !em->wf_table)
      return fail(RD_ERR_ARG, "work-function table missing or larger than 96 x 96");
    // append the table to the work-function store
    int need = c->wf_used + nx * ny;
    if (need > c->wf_cap) {
      int ncap = std::max(need, 2 * c->wf_cap + 256);
      double* nd = nullptr;
      RD_HIP(hipMalloc((void**)&nd, sizeof(double) * ncap));
      if (c->d_wf) {
        RD_HIP(hipMemcpy(nd, c->d_wf, sizeof(double) * c->wf_used, hipMemcpyDeviceToDevice));
        hipFree(c->d_wf);
      }
      c->d_wf = nd;
      c->wf_cap = ncap;
    }
    RD_HIP(hipMemcpy(c->d_wf + c->wf_used, em->wf_table, sizeof(double) * nx * ny,
                     hipMemcpyHostToDevice));
    d.wf_off = c->wf_used;
    d.wf_nx = nx;
    d.wf_ny = ny;
    c->wf_used = need;
  }
  if (mode == RD_EMIT_TIP_FE && !(em->w_tip > 0.0))
    return fail(RD_ERR_ARG, "tip work function must be > 0");
  if (em->quad_n < 0 || em->quad_n > 1024) return fail(RD_ERR_ARG, "quad_n outside 0..1024");
  if (em->ndim < 0 || em->ndim > 100000) return fail(RD_ERR_ARG, "ndim outside 0..100000");
  if (em->quad_levels < -1 || em->quad_levels > 8)
    return fail(RD_ERR_ARG, "quad_levels outside -1..8");
  d.mode = mode;
  d.delay = em->delay;
  d.ex0 = em->x0;
  d.ey0 = em->y0;
  d.elx = em->lx;
  d.ely = em->ly;
  d.T = em->T;
  d.w_tip = em->w_tip;
  d.quad = em->quad_n > 0 ? em->quad_n : 32;
  d.qgx = quad_side(d.quad, d.wf_nx);
  d.qgy = quad_side(d.quad, d.wf_ny);
  if (d.qgx > 2048 || d.qgy > 2048)
    return fail(RD_ERR_ARG, "supply grid beyond 2048 points per side (quad_n and the work-function sections)");
  d.quad_lmax = quad_lmax(std::max(d.qgx, d.qgy), em->quad_levels);
  int ndef = mode == RD_EMIT_PLANAR_GTF ? 25 : (mode == RD_EMIT_TIP_FE ? 80 : 200);
  d.ndim = em->ndim > 0 ? em->ndim : ndef;
  int r = sync_state(c);
  if (r) return r;
  SimState& S = c->h_state[c->cur];
  // module defaults of the samplers (src/mod_field_emission_v2.F90:55-56,
  // src/mod_field_thermo_emission.F90:32-33, src/mod_emission_tip.f90:50)
  if (mode == RD_EMIT_PLANAR_GTF) {
    S.MH_std = 0.125;
    S.a_rate = 1.0;
  } else if (mode == RD_EMIT_TIP_FE) {
    S.MH_std = 1.0;
    S.a_rate = 0.5;
  } else {
    S.MH_std = 0.0125;
    S.a_rate = 1.0;
  }
  S.residual = 0.0;
  r = put_state(c, c->cur);
  if (!r) r = put_desc(c, c->cur);
  return r;
}

int rd_put_particles(rd_ctx* c, int32_t n, const double* pos, const double* vel, const double* q,
                     const double* m, const int32_t* species) {
  RD_CUR(c);
  if (n < 0) return fail(RD_ERR_ARG, "negative particle count");
  if (n > sl->desc.cap) return fail(RD_ERR_CAPACITY, "particle count exceeds the slot capacity");
  if (n > 0 && (!pos || !q || !m)) return fail(RD_ERR_ARG, "pos, q and m are required");
  const int b = sl->desc.base;
  int r = upload_vec3(c, n, pos, c->x + b, c->y + b, c->z + b);
  if (r) return r;
  if (vel) {
    r = upload_vec3(c, n, vel, c->vx + b, c->vy + b, c->vz + b);
    if (r) return r;
  } else {
    fill(c, n, 0.0, c->vx + b);
    fill(c, n, 0.0, c->vy + b);
    fill(c, n, 0.0, c->vz + b);
  }
  double* zero[] = {c->px, c->py, c->pz, c->ax, c->ay, c->az, c->a1x, c->a1y, c->a1z, c->a2x,
                    c->a2y, c->a2z};
  for (double* p : zero) fill(c, n, 0.0, p + b);
  if (n > 0) {
    RD_HIP(hipMemcpyAsync(c->q + b, q, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    RD_HIP(hipMemcpyAsync(c->m + b, m, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    if (species)
      RD_HIP(hipMemcpyAsync(c->species + b, species, sizeof(int32_t) * n, hipMemcpyHostToDevice,
                            c->stream));
    else
      k_fill_i<<<(n + 255) / 256, 256, 0, c->stream>>>(n, RD_SPECIES_ELEC, c->species + b);
    int32_t* izero[] = {c->pstep, c->emitter};
    for (int32_t* p : izero) k_fill_i<<<(n + 255) / 256, 256, 0, c->stream>>>(n, 0, p + b);
    k_fill_i<<<(n + 255) / 256, 256, 0, c->stream>>>(n, 1, c->section + b);
    k_fill_i<<<(n + 255) / 256, 256, 0, c->stream>>>(n, 1, c->alive + b);
    k_iota<<<(n + 255) / 256, 256, 0, c->stream>>>(n, 0, c->id + b);
    RD_HIP(hipGetLastError());
  }
  r = sync_state(c);
  if (r) return r;
  SimState& S = c->h_state[c->cur];
  S.n = n;
  S.nrID = n;
  r = put_state(c, c->cur);
  if (r) return r;
  sl->n = n;
  return RD_OK;
}

int rd_particle_count(rd_ctx* c, int32_t* n) {
  RD_CUR(c);
  if (!n) return fail(RD_ERR_ARG, "null pointer");
  *n = sl->n;
  return RD_OK;
}

int rd_get_particles(rd_ctx* c, double* pos, double* vel, double* q, double* m,
                     int32_t* species) {
  RD_CUR(c);
  const int n = sl->n, b = sl->desc.base;
  int r = RD_OK;
  if (pos) r = download_vec3(c, n, c->x + b, c->y + b, c->z + b, pos);
  if (!r && vel) r = download_vec3(c, n, c->vx + b, c->vy + b, c->vz + b, vel);
  if (r) return r;
  if (n > 0) {
    if (q)
      RD_HIP(hipMemcpyAsync(q, c->q + b, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    if (m)
      RD_HIP(hipMemcpyAsync(m, c->m + b, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    if (species)
      RD_HIP(hipMemcpyAsync(species, c->species + b, sizeof(int32_t) * n, hipMemcpyDeviceToHost,
                            c->stream));
  }
  RD_HIP(hipStreamSynchronize(c->stream));
  return RD_OK;
}

int rd_get_state(rd_ctx* c, double* pos, double* prev, double* vel, double* acc, double* a1,
                 double* a2, int32_t* ints) {
  RD_CUR(c);
  const int n = sl->n, b = sl->desc.base;
  int r = RD_OK;
  if (pos) r = download_vec3(c, n, c->x + b, c->y + b, c->z + b, pos);
  if (!r && prev) r = download_vec3(c, n, c->px + b, c->py + b, c->pz + b, prev);
  if (!r && vel) r = download_vec3(c, n, c->vx + b, c->vy + b, c->vz + b, vel);
  if (!r && acc) r = download_vec3(c, n, c->ax + b, c->ay + b, c->az + b, acc);
  if (!r && a1) r = download_vec3(c, n, c->a1x + b, c->a1y + b, c->a1z + b, a1);
  if (!r && a2) r = download_vec3(c, n, c->a2x + b, c->a2y + b, c->a2z + b, a2);
  if (r) return r;
  if (ints && n > 0) {
    std::vector<int32_t> tmp((size_t)n);
    int32_t* cols[] = {c->species, c->pstep, c->emitter, c->section, c->id};
    for (int k = 0; k < 5; ++k) {
      RD_HIP(hipMemcpy(tmp.data(), cols[k] + b, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
      for (int i = 0; i < n; ++i) ints[5 * (size_t)i + k] = tmp[i];
    }
  }
  return RD_OK;
}

int rd_get_life_time(rd_ctx* c, int32_t* hist) {
  RD_CUR(c);
  if (!hist) return fail(RD_ERR_ARG, "null pointer");
  std::vector<int32_t> tmp(2 * kLife);
  RD_HIP(hipStreamSynchronize(c->stream));
  RD_HIP(hipMemcpy(tmp.data(), c->d_life + (size_t)c->cur * 2 * kLife, sizeof(int32_t) * 2 * kLife,
                   hipMemcpyDeviceToHost));
  // (1000, 2) column-major like life_time(1:MAX_LIFE_TIME, 1:nrSpecies)
  for (int s = 0; s < 2; ++s)
    for (int l = 0; l < kLife; ++l) hist[(size_t)s * kLife + l] = tmp[(size_t)s * kLife + l];
  return RD_OK;
}

int rd_set_logs(rd_ctx* c, int32_t on) {
  if (!c) return fail(RD_ERR_ARG, "null context");
  RD_HIP(hipSetDevice(c->device));
  if (on) {
    int r = ensure(c, c->mark, sizeof(double) * 4 * (size_t)c->cap);
    if (!r) r = ensure(c, c->elog, sizeof(rd_emit_record) * (size_t)c->cap);
    if (!r) r = ensure(c, c->alog, sizeof(rd_absorb_record) * (size_t)c->cap);
    if (!r) r = ensure(c, c->plog, sizeof(rd_plane_record) * RD_MAX_PLANES * (size_t)c->cap);
    if (r) return r;
  }
  c->logs = on != 0;
  return RD_OK;
}

int rd_get_logs(rd_ctx* c, int32_t max_emit, rd_emit_record* emit, int32_t* n_emit,
                int32_t max_absorb, rd_absorb_record* absorb, int32_t* n_absorb) {
  RD_CUR(c);
  if (!n_emit || !n_absorb) return fail(RD_ERR_ARG, "null pointer");
  if (!c->logs) return fail(RD_ERR_STATE, "particle records are off (rd_set_logs)");
  RD_HIP(hipStreamSynchronize(c->stream));
  const SimState& S = c->h_state[c->cur];
  *n_emit = S.emitted;
  *n_absorb = S.rem_total;
  if (S.emitted > max_emit || S.rem_total > max_absorb)
    return fail(RD_ERR_CAPACITY, "record buffers smaller than the step's counts");
  if ((S.emitted > 0 && !emit) || (S.rem_total > 0 && !absorb))
    return fail(RD_ERR_ARG, "null record buffer");
  const size_t base = (size_t)sl->desc.base;
  if (S.emitted > 0)
    RD_HIP(hipMemcpy(emit, (rd_emit_record*)c->elog.p + base, sizeof(rd_emit_record) * S.emitted,
                     hipMemcpyDeviceToHost));
  if (S.rem_total > 0)
    RD_HIP(hipMemcpy(absorb, (rd_absorb_record*)c->alog.p + base,
                     sizeof(rd_absorb_record) * S.rem_total, hipMemcpyDeviceToHost));
  return RD_OK;
}

int rd_set_planes(rd_ctx* c, int32_t n, const double* z) {
  RD_CUR(c);
  if (n < 0 || n > RD_MAX_PLANES) return fail(RD_ERR_ARG, "plane count outside 0..RD_MAX_PLANES");
  if (n > 0 && !z) return fail(RD_ERR_ARG, "null plane heights");
  for (int k = 0; k < n; ++k)
    if (!std::isfinite(z[k])) return fail(RD_ERR_ARG, "plane height not finite");
  SimDesc& d = sl->desc;
  d.nplanes = n;
  for (int k = 0; k < RD_MAX_PLANES; ++k) d.planes[k] = k < n ? z[k] : 0.0;
  return put_desc(c, c->cur);
}

int rd_get_plane_logs(rd_ctx* c, int32_t max_records, rd_plane_record* records, int32_t* n) {
  RD_CUR(c);
  if (!n) return fail(RD_ERR_ARG, "null pointer");
  if (!c->logs) return fail(RD_ERR_STATE, "particle records are off (rd_set_logs)");
  RD_HIP(hipStreamSynchronize(c->stream));
  const SimState& S = c->h_state[c->cur];
  long long total = 0;
  for (int k = 0; k < RD_MAX_PLANES; ++k) total += S.plane_n[k];
  *n = (int32_t)total;
  if (total > max_records) return fail(RD_ERR_CAPACITY, "record buffer smaller than the count");
  if (total > 0 && !records) return fail(RD_ERR_ARG, "null record buffer");
  const size_t base = (size_t)sl->desc.base;
  long long at = 0;
  for (int k = 0; k < RD_MAX_PLANES; ++k) {
    const int m = S.plane_n[k];
    if (m <= 0) continue;
    RD_HIP(hipMemcpy(records + at, (rd_plane_record*)c->plog.p + (size_t)k * c->cap + base,
                     sizeof(rd_plane_record) * m, hipMemcpyDeviceToHost));
    at += m;
  }
  return RD_OK;
}

int rd_accel(rd_ctx* c, int32_t precision) {
  RD_CUR(c);
  if (!sl->geom_set) return fail(RD_ERR_STATE, "geometry not set");
  std::vector<Ent> ents{slot_ent(c, c->cur)};
  int r = launch_accel(c, ents, precision, nullptr, nullptr);
  if (r) return r;
  RD_HIP(hipStreamSynchronize(c->stream));
  return RD_OK;
}

int rd_get_accel(rd_ctx* c, double* acc) {
  RD_CUR(c);
  if (!acc) return fail(RD_ERR_ARG, "null pointer");
  const int b = sl->desc.base;
  return download_vec3(c, sl->n, c->ax + b, c->ay + b, c->az + b, acc);
}

static int field_points(rd_ctx* c, rd::Slot* sl, int m, int precision, hipEvent_t k0,
                        hipEvent_t k1) {
  Ent e = slot_ent(c, c->cur);
  e.pbase = 0;
  e.m = m;
  (void)sl;
  std::vector<Ent> ents{e};
  return launch_field(c, ents, m, precision, k0, k1);
}

int rd_field_batch(rd_ctx* c, int32_t m, const double* pts, double* field, int32_t precision) {
  RD_CUR(c);
  if (!sl->geom_set) return fail(RD_ERR_STATE, "geometry not set");
  if (m < 0) return fail(RD_ERR_ARG, "negative point count");
  if (m == 0) return RD_OK;
  if (!pts || !field) return fail(RD_ERR_ARG, "null point/field buffer");
  int r = ensure(c, c->pts, sizeof(double) * 3 * (size_t)m);
  if (!r) r = ensure(c, c->fld, sizeof(double) * 3 * (size_t)m);
  if (r) return r;
  double* ptx = (double*)c->pts.p;
  r = upload_vec3(c, m, pts, ptx, ptx + m, ptx + 2 * m);
  if (!r) r = field_points(c, sl, m, precision, nullptr, nullptr);
  if (r) return r;
  double* f = (double*)c->fld.p;
  return download_vec3(c, m, f, f + m, f + 2 * m, field);
}

}  // extern "C"

// ---- allocation-free field at points -------------------------------------
// Work memory of rd_field_batch_static: one array per device, reserved when
// the library's code object is loaded on that device and held until the
// process ends.  Host code never uses the array's name as an address: its
// device address comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) unsigned char scratch[RD_STATIC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
std::map<int, char*> g_scratch;  // device -> address of `scratch` on it

// The device address of `scratch` on the current device `dev`.
int scratch_on(int dev, char** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch.find(dev);
  if (it != g_scratch.end()) {
    *out = it->second;
    return RD_OK;
  }
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
  if (!p) return fail(RD_ERR_HIP, "no device address for the static work array");
  g_scratch[dev] = (char*)p;
  *out = (char*)p;
  return RD_OK;
}
}  // namespace

extern "C" {

int rd_field_batch_static(rd_ctx* c, int32_t m, const double* pts, double* field,
                          int32_t precision) {
  RD_CUR(c);
  if (!sl->geom_set) return fail(RD_ERR_STATE, "geometry not set");
  if (m < 0) return fail(RD_ERR_ARG, "negative point count");
  if (m == 0) return RD_OK;
  if (!pts || !field) return fail(RD_ERR_ARG, "null point/field buffer");
  if (precision != RD_PREC_FP64 && precision != RD_PREC_MIXED)
    return fail(RD_ERR_ARG, "unknown precision");
  char* base = nullptr;
  int r = scratch_on(c->device, &base);
  if (r) return r;
  // staging (AoS), points and fields (SoA), then the launch's own pieces
  FixedBufs fb;
  fb.next = base;
  fb.end = base + RD_STATIC_SCRATCH_BYTES;
  const size_t v3 = sizeof(double) * 3 * (size_t)m;
  double* stage = (double*)fb.take(v3);
  fb.pts = (double*)fb.take(v3);
  fb.fld = (double*)fb.take(v3);
  if (!stage || !fb.pts || !fb.fld)
    return fail(RD_ERR_CAPACITY, "rd_field_batch_static: " + std::to_string(m) +
                                     " points exceed the static work array of " +
                                     std::to_string(RD_STATIC_SCRATCH_BYTES >> 20) + " MiB");
  RD_HIP(hipMemcpyAsync(stage, pts, v3, hipMemcpyHostToDevice, c->stream));
  k_aos_to_soa<<<(m + 255) / 256, 256, 0, c->stream>>>(m, stage, fb.pts, fb.pts + m,
                                                        fb.pts + 2 * m);
  RD_HIP(hipGetLastError());
  Ent e = slot_ent(c, c->cur);
  e.pbase = 0;
  e.m = m;
  std::vector<Ent> ents{e};
  r = launch_field(c, ents, m, precision, nullptr, nullptr, &fb);
  if (r) {
    if (r == RD_ERR_CAPACITY)
      return fail(RD_ERR_CAPACITY,
                  "rd_field_batch_static: " + std::to_string(m) + " points over " +
                      std::to_string(sl->n) + " particles exceed the static work array of " +
                      std::to_string(RD_STATIC_SCRATCH_BYTES >> 20) + " MiB");
    return r;
  }
  k_soa_to_aos<<<(m + 255) / 256, 256, 0, c->stream>>>(m, fb.fld, fb.fld + m, fb.fld + 2 * m,
                                                        stage);
  RD_HIP(hipGetLastError());
  RD_HIP(hipMemcpyAsync(field, stage, v3, hipMemcpyDeviceToHost, c->stream));
  RD_HIP(hipStreamSynchronize(c->stream));
  return RD_OK;
}

int rd_field_grid(rd_ctx* c, double x0, double y0, double lx, double ly, int32_t nx, int32_t ny,
                  double zz, int32_t precision, double* field) {
  RD_CUR(c);
  if (!sl->geom_set) return fail(RD_ERR_STATE, "geometry not set");
  if (nx < 1 || ny < 1 || (int64_t)nx * ny > (1 << 26)) return fail(RD_ERR_ARG, "bad grid size");
  int m = nx * ny;
  int r = launch_grid_points(c, x0, y0, lx, ly, nx, ny, zz);
  if (!r) r = field_points(c, sl, m, precision, nullptr, nullptr);
  if (r) return r;
  if (field) {
    double* f = (double*)c->fld.p;
    return download_vec3(c, m, f, f + m, f + 2 * m, field);
  }
  RD_HIP(hipStreamSynchronize(c->stream));
  return RD_OK;
}

// ---- time step -------------------------------------------------------------
static int check_step_ready(rd_ctx* c, rd::Slot* sl) {
  (void)c;
  if (!sl->geom_set) return fail(RD_ERR_STATE, "geometry not set");
  if (!(sl->desc.dt > 0.0)) return fail(RD_ERR_STATE, "time step not set");
  return RD_OK;
}

int rd_add_particles(rd_ctx* c, int32_t n, const double* pos, const double* vel,
                     const int32_t* species, int32_t step, int32_t emit, const int32_t* sec) {
  RD_CUR(c);
  if (!sl->geom_set) return fail(RD_ERR_STATE, "geometry not set");
  if (n < 0) return fail(RD_ERR_ARG, "negative particle count");
  if (n > 0 && !pos) return fail(RD_ERR_ARG, "null positions");
  if (species)
    for (int k = 0; k < n; ++k)
      if (species[k] != RD_SPECIES_ELEC && species[k] != RD_SPECIES_ION)
        return fail(RD_ERR_ARG, "unknown particle species");
  return engine_add(c, c->cur, n, pos, vel, species, step, emit, sec);
}

static int run_phases(rd_ctx* c, int32_t step, int32_t precision, int phases, rd_step_record* rec) {
  std::vector<int> slots{c->cur};
  int steps[1] = {step};
  return engine_phases(c, slots, precision, phases, steps, rec);
}

int rd_do_emission(rd_ctx* c, int32_t step, int32_t precision, rd_step_record* rec) {
  RD_CUR(c);
  int r = check_step_ready(c, sl);
  if (r) return r;
  return run_phases(c, step, precision, 1, rec);
}

int rd_update_position(rd_ctx* c, int32_t step, int32_t precision, rd_step_record* rec) {
  RD_CUR(c);
  int r = check_step_ready(c, sl);
  if (r) return r;
  return run_phases(c, step, precision, 2, rec);
}

int rd_remove_particles(rd_ctx* c, int32_t step, rd_step_record* rec) {
  RD_CUR(c);
  int r = run_phases(c, step, RD_PREC_FP64, 4, rec);
  if (r) return r;
  sl->step = step;
  return RD_OK;
}

int rd_run(rd_ctx* c, int32_t nsteps, int32_t precision, rd_step_record* records) {
  RD_CUR(c);
  int r = check_step_ready(c, sl);
  if (r) return r;
  if (nsteps < 0) return fail(RD_ERR_ARG, "negative step count");
  for (int k = 0; k < nsteps; ++k) {
    int step = sl->step + 1;
    r = run_phases(c, step, precision, 7, records ? &records[k] : nullptr);
    if (r) return r;
    sl->step = step;
  }
  return RD_OK;
}

int rd_pool_step(rd_ctx* c, int32_t precision, rd_step_record* records) {
  if (!c) return fail(RD_ERR_ARG, "null context");
  RD_HIP(hipSetDevice(c->device));
  std::vector<int> slots;
  std::vector<int> steps;
  for (int s = 0; s < c->nslots_max; ++s) {
    Slot& sl = c->slots[s];
    if (!sl.active) continue;
    if (!sl.geom_set || !(sl.desc.dt > 0.0)) return fail(RD_ERR_STATE, "slot not configured");
    slots.push_back(s);
    steps.push_back(sl.step + 1);
  }
  if (slots.empty()) return RD_OK;
  std::vector<rd_step_record> recs(slots.size());
  int r = engine_phases(c, slots, precision, 7, steps.data(), recs.data());
  if (r) return r;
  for (size_t k = 0; k < slots.size(); ++k) {
    c->slots[slots[k]].step = steps[k];
    if (records) records[slots[k]] = recs[k];
  }
  return RD_OK;
}

// ---- measurement -------------------------------------------------------------
static int time_loop(rd_ctx* c, int reps, double* total_ms, double* kernel_ms,
                     int (*body)(rd_ctx*, void*, hipEvent_t, hipEvent_t), void* arg) {
  if (reps < 1) return fail(RD_ERR_ARG, "reps must be >= 1");
  int r = body(c, arg, nullptr, nullptr);  // warm-up
  if (r) return r;
  RD_HIP(hipStreamSynchronize(c->stream));
  double ktot = 0.0;
  RD_HIP(hipEventRecord(c->ev2, c->stream));
  for (int k = 0; k < reps; ++k) {
    r = body(c, arg, c->ev0, c->ev1);
    if (r) return r;
    RD_HIP(hipEventSynchronize(c->ev1));
    float ms = 0.f;
    RD_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    ktot += ms;
  }
  hipEvent_t end;
  RD_HIP(hipEventCreate(&end));
  RD_HIP(hipEventRecord(end, c->stream));
  RD_HIP(hipEventSynchronize(end));
  float tot = 0.f;
  RD_HIP(hipEventElapsedTime(&tot, c->ev2, end));
  hipEventDestroy(end);
  if (total_ms) *total_ms = tot / reps;
  if (kernel_ms) *kernel_ms = ktot / reps;
  return RD_OK;
}

struct AccelArg {
  int precision;
};
static int accel_body(rd_ctx* c, void* a, hipEvent_t k0, hipEvent_t k1) {
  std::vector<Ent> ents{slot_ent(c, c->cur)};
  return launch_accel(c, ents, ((AccelArg*)a)->precision, k0, k1);
}

int rd_time_accel(rd_ctx* c, int32_t precision, int32_t reps, double* total_ms,
                  double* kernel_ms) {
  RD_CUR(c);
  if (!sl->geom_set) return fail(RD_ERR_STATE, "geometry not set");
  if (sl->n < 1) return fail(RD_ERR_STATE, "no particles");
  AccelArg a{precision};
  return time_loop(c, reps, total_ms, kernel_ms, accel_body, &a);
}

struct GridArg {
  int m, precision;
};
static int grid_body(rd_ctx* c, void* a, hipEvent_t k0, hipEvent_t k1) {
  GridArg* g = (GridArg*)a;
  return field_points(c, &c->slots[c->cur], g->m, g->precision, k0, k1);
}

int rd_time_field_grid(rd_ctx* c, double x0, double y0, double lx, double ly, int32_t nx,
                       int32_t ny, double zz, int32_t precision, int32_t reps, double* total_ms,
                       double* kernel_ms) {
  RD_CUR(c);
  if (!sl->geom_set) return fail(RD_ERR_STATE, "geometry not set");
  if (sl->n < 1) return fail(RD_ERR_STATE, "no particles");
  if (nx < 1 || ny < 1 || (int64_t)nx * ny > (1 << 26)) return fail(RD_ERR_ARG, "bad grid size");
  int r = launch_grid_points(c, x0, y0, lx, ly, nx, ny, zz);
  if (r) return r;
  GridArg a{nx * ny, precision};
  return time_loop(c, reps, total_ms, kernel_ms, grid_body, &a);
}

}  // extern "C"

// Plain device-to-device copy: 16 B per lane and iteration, grid-stride.
__global__ __launch_bounds__(256) void k_copy16(const float4* __restrict__ src,
                                                float4* __restrict__ dst, int64_t n16) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
    dst[i] = src[i];
}

struct CopyArg {
  const float4* src;
  float4* dst;
  int64_t n16;
};
static int copy_body(rd_ctx* c, void* a, hipEvent_t k0, hipEvent_t k1) {
  CopyArg* g = (CopyArg*)a;
  if (k0) RD_HIP(hipEventRecord(k0, c->stream));
  k_copy16<<<4096, 256, 0, c->stream>>>(g->src, g->dst, g->n16);
  if (k1) RD_HIP(hipEventRecord(k1, c->stream));
  RD_HIP(hipGetLastError());
  return RD_OK;
}

extern "C" {

int rd_set_graphs(rd_ctx* c, int32_t on) {
  if (!c) return fail(RD_ERR_ARG, "null context");
  RD_HIP(hipSetDevice(c->device));
  RD_HIP(hipStreamSynchronize(c->stream));
  if (c->jump_exec) RD_HIP(hipGraphExecDestroy(c->jump_exec));
  c->jump_exec = nullptr;
  c->jump_key.clear();
  c->graphs = on != 0;
  return RD_OK;
}

int rd_set_force_options(rd_ctx* c, int32_t flags) {
  if (!c) return fail(RD_ERR_ARG, "null context");
  if (flags & ~(RD_FORCE_NO_SYMMETRIC | RD_FORCE_NO_FAR_EXPANSION))
    return fail(RD_ERR_ARG, "unknown force option");
  c->force_opts = flags;
  return RD_OK;
}

int rd_work_bytes(rd_ctx* c, int64_t* bytes) {
  if (!c || !bytes) return fail(RD_ERR_ARG, "null pointer");
  *bytes = (int64_t)c->work_bytes;
  return RD_OK;
}

int rd_set_partial_budget(rd_ctx* c, int64_t bytes) {
  if (!c) return fail(RD_ERR_ARG, "null context");
  if (bytes < 0) return fail(RD_ERR_ARG, "negative partial-sum budget");
  c->part_budget = bytes == 0 ? kPartBudget : (size_t)bytes;
  return RD_OK;
}

int rd_set_jump_pairing(rd_ctx* c, int64_t min_workgroups) {
  if (!c) return fail(RD_ERR_ARG, "null context");
  if (min_workgroups < 0) return fail(RD_ERR_ARG, "negative workgroup threshold");
  c->jump_pair_wgs = min_workgroups;
  return RD_OK;
}

int rd_set_solo_jumps(rd_ctx* c, int32_t on) {
  if (!c) return fail(RD_ERR_ARG, "null context");
  c->solo_jumps = on != 0;
  return RD_OK;
}

int rd_time_copy(rd_ctx* c, int64_t bytes, int32_t reps, double* kernel_ms) {
  if (!c) return fail(RD_ERR_ARG, "null context");
  if (bytes < 16 || bytes % 16 != 0 || bytes > (int64_t(1) << 34))
    return fail(RD_ERR_ARG, "bytes must be a multiple of 16 in [16, 16 GiB]");
  RD_HIP(hipSetDevice(c->device));
  void *src = nullptr, *dst = nullptr;
  hipError_t e = hipMalloc(&src, bytes);
  if (e == hipSuccess) e = hipMalloc(&dst, bytes);
  if (e == hipSuccess) e = hipMemsetAsync(src, 0, bytes, c->stream);
  int r = RD_OK;
  if (e != hipSuccess) {
    r = hip_fail(e, "copy buffers");
  } else {
    CopyArg a{(const float4*)src, (float4*)dst, bytes / 16};
    r = time_loop(c, reps, nullptr, kernel_ms, copy_body, &a);
    hipStreamSynchronize(c->stream);
  }
  if (src) hipFree(src);
  if (dst) hipFree(dst);
  return r;
}

}  // extern "C"
