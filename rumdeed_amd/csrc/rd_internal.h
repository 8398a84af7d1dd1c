// Internal definitions of librumdeed.so: the per-context device state and
// the kernel launch helpers shared by the translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

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
  // hyperboloid tip (Init_Emission_Tip, src/mod_emission_tip.f90:105-125)
  double a_foci, shift_z, pre_E, pre_E_unit, z0, R, eta_1, h_tip, d_tip, R_base, max_xi, theta;
  // mixed-precision kernels work in nanometres
  float d_nm, pad2;
};

// Canonical FP64 summation: chunks of kChunk sources summed sequentially,
// chunk partials summed sequentially (the CPU restatement uses the same).
constexpr int kChunk = 256;
constexpr int kBlock = 256;

struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
};

}  // namespace rd

struct rd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int64_t cap = 0;
  int32_t n = 0;
  bool geom_set = false;
  rd_geometry geom_in{};
  rd::DevGeom g{};
  double dt = 0.0;

  // particle state, structure of arrays in HBM
  double *x = nullptr, *y = nullptr, *z = nullptr;
  double *px = nullptr, *py = nullptr, *pz = nullptr;  // previous position
  double *vx = nullptr, *vy = nullptr, *vz = nullptr;
  double *ax = nullptr, *ay = nullptr, *az = nullptr;     // current acceleration
  double *a1x = nullptr, *a1y = nullptr, *a1z = nullptr;  // previous
  double *a2x = nullptr, *a2y = nullptr, *a2z = nullptr;  // previous-previous
  double *q = nullptr, *m = nullptr;
  int32_t *species = nullptr, *step = nullptr, *emitter = nullptr, *section = nullptr,
          *id = nullptr;

  // work buffers (grown on demand, never shrunk)
  rd::Buf part;      // partial sums of the pair kernels
  rd::Buf pack_hi;   // mixed-precision packed sources
  rd::Buf pack_lo;
  rd::Buf pack_im;   // tip image points (FP64 or packed)
  rd::Buf pts;       // field points (SoA x,y,z)
  rd::Buf fld;       // field results (SoA)
  rd::Buf pts_hi, pts_lo;
  rd::Buf stage;     // host<->device staging (AoS)
  rd::Buf img64;     // tip image data FP64: dis, xim, yim, zim

  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
};

namespace rd {
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);
int ensure(rd_ctx* c, Buf& b, size_t bytes);

// Pair kernels (rd_pair.hip).  All launches go to c->stream.
int launch_accel(rd_ctx* c, int precision, hipEvent_t k0, hipEvent_t k1);
int launch_field(rd_ctx* c, int m, int precision, hipEvent_t k0, hipEvent_t k1);
int launch_grid_points(rd_ctx* c, double x0, double y0, double lx, double ly, int nx, int ny,
                       double zz);
}  // namespace rd

#define RD_HIP(call)                                    \
  do {                                                  \
    hipError_t _e = (call);                             \
    if (_e != hipSuccess) return rd::hip_fail(_e, #call); \
  } while (0)
