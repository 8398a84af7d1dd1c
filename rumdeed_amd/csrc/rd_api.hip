// C ABI of librumdeed.so: contexts, particle store, force and field entry
// points, timing.  See include/rumdeed.h for the contract of each function.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
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
    b.p = nullptr;
    b.bytes = 0;
  }
  size_t want = bytes + bytes / 4 + 4096;
  hipError_t e = hipMalloc(&b.p, want);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc(work buffer)");
  b.bytes = want;
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

int upload_vec3(rd_ctx* c, int n, const double* host, double* x, double* y, double* z) {
  if (n == 0) return RD_OK;
  int r = ensure(c, c->stage, sizeof(double) * 3 * (size_t)n);
  if (r) return r;
  RD_HIP(hipMemcpyAsync(c->stage.p, host, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
                        c->stream));
  k_aos_to_soa<<<(n + 255) / 256, 256, 0, c->stream>>>(n, (const double*)c->stage.p, x, y, z);
  RD_HIP(hipGetLastError());
  return RD_OK;
}

int download_vec3(rd_ctx* c, int n, const double* x, const double* y, const double* z,
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
  g->d_nm = (float)(g->d * 1e9);
  return RD_OK;
}

}  // namespace rd

using namespace rd;

extern "C" {

const char* rd_version(void) { return "rumdeed-mi355x 0.1 (gfx950)"; }
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

int rd_ctx_create(int32_t device, int64_t max_particles, rd_ctx** out) {
  if (!out) return fail(RD_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (max_particles < 1 || max_particles > (int64_t)1 << 30)
    return fail(RD_ERR_ARG, "max_particles outside 1..2^30");
  int nd = 0;
  hipError_t e = hipGetDeviceCount(&nd);
  if (e != hipSuccess || nd == 0) return fail(RD_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= nd) return fail(RD_ERR_ARG, "device index out of range");
  RD_HIP(hipSetDevice(device));
  rd_ctx* c = new rd_ctx();
  c->device = device;
  c->cap = max_particles;
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return hip_fail(e, "hipStreamCreate");
  }
  double** d_arrays[] = {&c->x, &c->y, &c->z, &c->px, &c->py, &c->pz, &c->vx, &c->vy,
                         &c->vz, &c->ax, &c->ay, &c->az, &c->a1x, &c->a1y, &c->a1z, &c->a2x,
                         &c->a2y, &c->a2z, &c->q, &c->m};
  int32_t** i_arrays[] = {&c->species, &c->step, &c->emitter, &c->section, &c->id};
  size_t cap = (size_t)max_particles;
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
  hipEventCreate(&c->ev0);
  hipEventCreate(&c->ev1);
  hipEventCreate(&c->ev2);
  *out = c;
  return RD_OK;
}

int rd_ctx_destroy(rd_ctx* c) {
  if (!c) return RD_OK;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  double* d_arrays[] = {c->x,  c->y,  c->z,  c->px,  c->py,  c->pz,  c->vx,
                        c->vy, c->vz, c->ax, c->ay,  c->az,  c->a1x, c->a1y,
                        c->a1z, c->a2x, c->a2y, c->a2z, c->q, c->m};
  for (double* p : d_arrays)
    if (p) hipFree(p);
  int32_t* i_arrays[] = {c->species, c->step, c->emitter, c->section, c->id};
  for (int32_t* p : i_arrays)
    if (p) hipFree(p);
  Buf* bufs[] = {&c->part, &c->pack_hi, &c->pack_lo, &c->pack_im, &c->pts, &c->fld,
                 &c->pts_hi, &c->pts_lo, &c->stage, &c->img64};
  for (Buf* b : bufs) free_buf(*b);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  if (c->ev2) hipEventDestroy(c->ev2);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return RD_OK;
}

int rd_set_geometry(rd_ctx* c, const rd_geometry* geom) {
  if (!c || !geom) return fail(RD_ERR_ARG, "null pointer");
  DevGeom g;
  int r = build_geom(geom, &g);
  if (r) return r;
  c->g = g;
  c->geom_in = *geom;
  c->geom_set = true;
  return RD_OK;
}

int rd_set_time_step(rd_ctx* c, double dt) {
  if (!c) return fail(RD_ERR_ARG, "null pointer");
  if (!(dt > 0.0)) return fail(RD_ERR_ARG, "time step must be > 0");
  c->dt = dt;
  return RD_OK;
}

int rd_put_particles(rd_ctx* c, int32_t n, const double* pos, const double* vel, const double* q,
                     const double* m, const int32_t* species) {
  if (!c) return fail(RD_ERR_ARG, "null context");
  if (n < 0) return fail(RD_ERR_ARG, "negative particle count");
  if (n > c->cap) return fail(RD_ERR_CAPACITY, "particle count exceeds context capacity");
  if (n > 0 && (!pos || !q || !m)) return fail(RD_ERR_ARG, "pos, q and m are required");
  RD_HIP(hipSetDevice(c->device));
  int r = upload_vec3(c, n, pos, c->x, c->y, c->z);
  if (r) return r;
  RD_HIP(hipStreamSynchronize(c->stream));
  if (vel) {
    r = upload_vec3(c, n, vel, c->vx, c->vy, c->vz);
    if (r) return r;
    RD_HIP(hipStreamSynchronize(c->stream));
  } else {
    fill(c, n, 0.0, c->vx);
    fill(c, n, 0.0, c->vy);
    fill(c, n, 0.0, c->vz);
  }
  double* zero[] = {c->px, c->py, c->pz, c->ax, c->ay, c->az, c->a1x,
                    c->a1y, c->a1z, c->a2x, c->a2y, c->a2z};
  for (double* p : zero) fill(c, n, 0.0, p);
  if (n > 0) {
    RD_HIP(hipMemcpyAsync(c->q, q, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    RD_HIP(hipMemcpyAsync(c->m, m, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    if (species)
      RD_HIP(hipMemcpyAsync(c->species, species, sizeof(int32_t) * n, hipMemcpyHostToDevice,
                            c->stream));
    else
      k_fill_i<<<(n + 255) / 256, 256, 0, c->stream>>>(n, RD_SPECIES_ELEC, c->species);
    int32_t* izero[] = {c->step, c->emitter, c->section};
    for (int32_t* p : izero) k_fill_i<<<(n + 255) / 256, 256, 0, c->stream>>>(n, 0, p);
    RD_HIP(hipGetLastError());
  }
  RD_HIP(hipStreamSynchronize(c->stream));
  c->n = n;
  return RD_OK;
}

int rd_particle_count(rd_ctx* c, int32_t* n) {
  if (!c || !n) return fail(RD_ERR_ARG, "null pointer");
  *n = c->n;
  return RD_OK;
}

int rd_get_particles(rd_ctx* c, double* pos, double* vel, double* q, double* m,
                     int32_t* species) {
  if (!c) return fail(RD_ERR_ARG, "null context");
  RD_HIP(hipSetDevice(c->device));
  int n = c->n;
  int r = RD_OK;
  if (pos) r = download_vec3(c, n, c->x, c->y, c->z, pos);
  if (!r && vel) r = download_vec3(c, n, c->vx, c->vy, c->vz, vel);
  if (r) return r;
  if (n > 0) {
    if (q) RD_HIP(hipMemcpyAsync(q, c->q, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    if (m) RD_HIP(hipMemcpyAsync(m, c->m, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    if (species)
      RD_HIP(hipMemcpyAsync(species, c->species, sizeof(int32_t) * n, hipMemcpyDeviceToHost,
                            c->stream));
  }
  RD_HIP(hipStreamSynchronize(c->stream));
  return RD_OK;
}

int rd_accel(rd_ctx* c, int32_t precision) {
  if (!c) return fail(RD_ERR_ARG, "null context");
  if (!c->geom_set) return fail(RD_ERR_STATE, "geometry not set");
  RD_HIP(hipSetDevice(c->device));
  int r = launch_accel(c, precision, nullptr, nullptr);
  if (r) return r;
  RD_HIP(hipStreamSynchronize(c->stream));
  return RD_OK;
}

int rd_get_accel(rd_ctx* c, double* acc) {
  if (!c || !acc) return fail(RD_ERR_ARG, "null pointer");
  RD_HIP(hipSetDevice(c->device));
  return download_vec3(c, c->n, c->ax, c->ay, c->az, acc);
}

int rd_field_batch(rd_ctx* c, int32_t m, const double* pts, double* field, int32_t precision) {
  if (!c) return fail(RD_ERR_ARG, "null context");
  if (!c->geom_set) return fail(RD_ERR_STATE, "geometry not set");
  if (m < 0) return fail(RD_ERR_ARG, "negative point count");
  if (m == 0) return RD_OK;
  if (!pts || !field) return fail(RD_ERR_ARG, "null point/field buffer");
  RD_HIP(hipSetDevice(c->device));
  int r = ensure(c, c->pts, sizeof(double) * 3 * (size_t)m);
  if (!r) r = ensure(c, c->fld, sizeof(double) * 3 * (size_t)m);
  if (r) return r;
  double* ptx = (double*)c->pts.p;
  r = upload_vec3(c, m, pts, ptx, ptx + m, ptx + 2 * m);
  if (!r) r = launch_field(c, m, precision, nullptr, nullptr);
  if (r) return r;
  double* f = (double*)c->fld.p;
  return download_vec3(c, m, f, f + m, f + 2 * m, field);
}

int rd_field_grid(rd_ctx* c, double x0, double y0, double lx, double ly, int32_t nx, int32_t ny,
                  double zz, int32_t precision, double* field) {
  if (!c) return fail(RD_ERR_ARG, "null context");
  if (!c->geom_set) return fail(RD_ERR_STATE, "geometry not set");
  if (nx < 1 || ny < 1 || (int64_t)nx * ny > (1 << 26)) return fail(RD_ERR_ARG, "bad grid size");
  RD_HIP(hipSetDevice(c->device));
  int m = nx * ny;
  int r = launch_grid_points(c, x0, y0, lx, ly, nx, ny, zz);
  if (!r) r = launch_field(c, m, precision, nullptr, nullptr);
  if (r) return r;
  if (field) {
    double* f = (double*)c->fld.p;
    return download_vec3(c, m, f, f + m, f + 2 * m, field);
  }
  RD_HIP(hipStreamSynchronize(c->stream));
  return RD_OK;
}

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
  return launch_accel(c, ((AccelArg*)a)->precision, k0, k1);
}

int rd_time_accel(rd_ctx* c, int32_t precision, int32_t reps, double* total_ms,
                  double* kernel_ms) {
  if (!c) return fail(RD_ERR_ARG, "null context");
  if (!c->geom_set) return fail(RD_ERR_STATE, "geometry not set");
  if (c->n < 1) return fail(RD_ERR_STATE, "no particles");
  RD_HIP(hipSetDevice(c->device));
  AccelArg a{precision};
  return time_loop(c, reps, total_ms, kernel_ms, accel_body, &a);
}

struct GridArg {
  int m, precision;
};
static int grid_body(rd_ctx* c, void* a, hipEvent_t k0, hipEvent_t k1) {
  GridArg* g = (GridArg*)a;
  return launch_field(c, g->m, g->precision, k0, k1);
}

int rd_time_field_grid(rd_ctx* c, double x0, double y0, double lx, double ly, int32_t nx,
                       int32_t ny, double zz, int32_t precision, int32_t reps, double* total_ms,
                       double* kernel_ms) {
  if (!c) return fail(RD_ERR_ARG, "null context");
  if (!c->geom_set) return fail(RD_ERR_STATE, "geometry not set");
  if (c->n < 1) return fail(RD_ERR_STATE, "no particles");
  if (nx < 1 || ny < 1 || (int64_t)nx * ny > (1 << 26)) return fail(RD_ERR_ARG, "bad grid size");
  RD_HIP(hipSetDevice(c->device));
  int r = launch_grid_points(c, x0, y0, lx, ly, nx, ny, zz);
  if (r) return r;
  GridArg a{nx * ny, precision};
  return time_loop(c, reps, total_ms, kernel_ms, grid_body, &a);
}

}  // extern "C"
