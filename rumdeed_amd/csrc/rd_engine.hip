// The device-resident time step: emission, Beeman update with boundary
// marking and Ramo current, order-preserving compaction, batched Add_Particle.
//
// Reference (RUMDEED tree):
//   main loop ................. src/main.F90:175-219
//   Beeman position ........... Update_ElecIon_Position, src/mod_verlet.F90:197-232
//   boundaries ................ Check_Boundary_Planar :325-338, Check_Boundary_Tip
//                               src/mod_emission_tip.f90:1627-1647
//   Beeman velocity, Ramo ..... Update_ElecIon_Velocity, src/mod_verlet.F90:449-509
//   marking / compaction ...... Mark_Particles_Remove, Remove_Particles, record_lifetime,
//                               src/mod_pair.F90:169-562, :1135-1223
//   Add_Particle .............. src/mod_pair.F90:29-159
//   planar FE emission ........ Do_Field_Emission_Planar_rectangle,
//                               Metropolis_Hastings_rectangle_J_batch,
//                               src/mod_field_emission_v2.F90:261-395, :1284-1516
//   thermal-field emission .... src/mod_field_thermo_emission.F90:136-446
//   tip emission .............. Do_Field_Emission_Tip_OLDCODE, Metro_algo_tip_v3,
//                               src/mod_emission_tip.f90:417-534, :1213-1390
//
// Every kernel takes a launch table: one entry per simulation slot, so a
// pool of independent simulations advances with one launch per kernel.
//
// Emission semantics on the device (all simulations, all modes): the
// Metropolis-Hastings chains of a step run in lockstep against the particle
// configuration of the start of the step (the reference's mh_batch
// semantics, src/mod_global.F90:316-328).  The planar FE sampler updates the
// proposal width once per jump from the batch acceptance rate, as the
// reference's batch sampler does; the thermal-field and tip samplers, whose
// reference updates the width once per chain, apply the per-chain updates in
// chain order after the step's chains (all chains of a step use the width of
// the start of the step).  The supply integral replaces the Cuba adaptive
// integrator (not available) with a midpoint rule on a fixed grid, summed in
// tiles of 1024 points (a fixed pairwise tree within a tile, tiles in order).
// Random numbers come from Philox4x64-10 streams keyed by the
// simulation seed and indexed by (step, purpose, chain).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <type_traits>

#include "rd_field.h"
#include "rd_internal.h"
#include "rd_physics.h"

namespace rd {

static constexpr double kTiny = 2.2250738585072014e-308;  // tiny(1.0d0)
static constexpr double kHuge = 1.7976931348623157e308;   // huge(1.0d0)

// ------------------------------------------------------------------------
// Chain arrays (structure of arrays over all chains of a step).
struct Chains {
  double *cx, *cy;    // current position (x, y) planar, (xi, phi) tip
  double *px, *py;    // proposal
  double *F, *sup, *D;
  double *ox, *oy, *oz;  // final emission site (tip: xyz on the surface)
  int *ok, *acc, *rej, *pv;  // start found; per-chain counts; proposal valid
};

static Chains chain_view(rd_ctx* c, int C) {
  Chains h;
  double* d = (double*)c->chains.p;
  h.cx = d;
  h.cy = d + C;
  h.px = d + 2 * (size_t)C;
  h.py = d + 3 * (size_t)C;
  h.F = d + 4 * (size_t)C;
  h.sup = d + 5 * (size_t)C;
  h.D = d + 6 * (size_t)C;
  h.ox = d + 7 * (size_t)C;
  h.oy = d + 8 * (size_t)C;
  h.oz = d + 9 * (size_t)C;
  int* ip = (int*)(d + 10 * (size_t)C);
  h.ok = ip;
  h.acc = ip + C;
  h.rej = ip + 2 * (size_t)C;
  h.pv = ip + 3 * (size_t)C;
  return h;
}
static size_t chain_bytes(int C) { return (size_t)C * (10 * sizeof(double) + 4 * sizeof(int)); }

// Work function at a surface point of a planar emitter.
__device__ __forceinline__ double work_at(const SimDesc& s, const double* wf, double x, double y,
                                          int& sec) {
  return work_checker(wf + s.wf_off, s.wf_nx, s.wf_ny, s.ex0, s.ey0, s.elx, s.ely, x, y, sec);
}

__device__ __forceinline__ TipSurf tip_surf(const DevGeom& g) {
  TipSurf t;
  t.a = g.a_foci;
  t.eta1 = g.eta_1;
  t.shift_z = g.shift_z;
  t.max_xi = g.max_xi;
  return t;
}

// Normal component of the field on the tip surface (Field_normal).
__device__ __forceinline__ double field_normal(const DevGeom& g, double x, double y, double ex,
                                              double ey, double ez) {
  double nx, ny, nz;
  tip_normal(tip_surf(g), x, y, nx, ny, nz);
  return nx * ex + ny * ey + nz * ez;
}

// MH chain target at a surface point with field component F (< 0).
__device__ __forceinline__ double chain_target(const SimDesc& s, const double* wf,
                                               const EmitConsts& ec, double F, double a, double b,
                                               double xs, double ys, double zs) {
  if (s.mode == RD_EMIT_PLANAR_FE) {
    int sec;
    double w = work_at(s, wf, a, b, sec);
    return fn_supply_log(ec, s.g.ic, F, w);  // Elec_Supply_log
  }
  if (s.mode == RD_EMIT_PLANAR_GTF) {
    int sec;
    double w = work_at(s, wf, a, b, sec);
    double J = gtf_current(F, s.T, w);
    return rd_log(J > kTiny ? J : kTiny);
  }
  // tip: Tip_fe_target_log(eta_f, xi, pos); a = xi
  double t = fn_t(s.g.ic, ec.l_const, F, s.w_tip);
  double sup = s.dt_q0 * ec.a_FN / (t * t * s.w_tip) * (F * F);
  (void)xs;
  (void)ys;
  (void)zs;
  return rd_log(sup > kTiny ? sup : kTiny) + 0.5 * rd_log(a * a - s.g.eta_1 * s.g.eta_1);
}

// ------------------------------------------------------------------------
__global__ void k_reset_counters(const Ent* tab, SimState* st) {
  const Ent e = tab[blockIdx.x];
  if (threadIdx.x != 0) return;
  SimState& s = st[e.slot];
  s.emitted = 0;
  s.n_round = 0;
  s.rem_total = s.rem_top = s.rem_bot = s.rem_elec = s.rem_ion = s.rem_top_emit = 0;
  s.first_dead = 0x7fffffff;
  s.pending = 0;
  s.start_rounds = 0;
  s.q_level = 0;
  s.q_cells = 0;
  s.q_fail = 0;
  s.q_err = 0.0;
  s.N_sup = 0.0;
  s.df_avg = 0.0;
  for (int k = 0; k < 3; ++k) s.F_avg[k] = 0.0;
}

// Quadrature points of the supply integral: the tip's 100 x 100 (xi, phi)
// cells; the planar emitter's cell centres at level e.lev (qgx << lev by
// qgy << lev, x fastest).
__global__ void k_quad_points(const Ent* tab, const SimDesc* desc, double* __restrict__ ptx,
                              double* __restrict__ pty, double* __restrict__ ptz) {
  const Ent e = tab[blockIdx.z];
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= e.m) return;
  const SimDesc& s = desc[e.slot];
  double x, y, z;
  if (s.mode == RD_EMIT_TIP_FE) {
    const int nr = 100;
    int i = k / nr + 1, j = k % nr + 1;
    double len_xi = (s.g.max_xi - 1.0) / nr;
    double len_phi = 2.0 * 3.141592653589793238462643383279502884197 / nr;
    double xi_c = 1.0 + (i - 0.5) * len_xi;
    double phi_c = (j - 0.5) * len_phi;
    tip_xyz(tip_surf(s.g), xi_c, s.g.eta_1, phi_c, x, y, z);
  } else {
    const int Gx = s.qgx << e.lev, Gy = s.qgy << e.lev;
    int ix = k % Gx, iy = k / Gx;
    double u = (ix + 0.5) / Gx, v = (iy + 0.5) / Gy;
    x = s.ex0 + u * s.elx;
    y = s.ey0 + v * s.ely;
    z = 0.0;
  }
  ptx[e.pbase + k] = x;
  pty[e.pbase + k] = y;
  ptz[e.pbase + k] = z;
}

// Integrand per point and the sums of one level (value, f_x, f_y, f_z; the
// tip: value, 0, 0, normal field): tiles of 1024 points, a fixed pairwise
// tree within a tile (zeros beyond the last point), tile sums added in tile
// order.  k_quad_tiles: a workgroup per tile (its integrand values and its
// tree; the tile sums at qt + 4 (e.c_off + tile)); k_quad_tsum: the tile sums
// in tile order into S.qsum[lev & 1].
__global__ __launch_bounds__(1024) void k_quad_tiles(const Ent* tab, const SimDesc* desc,
                                                    const double* wf,
                                                    const double* __restrict__ ptx,
                                                    const double* __restrict__ pty,
                                                    const double* __restrict__ fx,
                                                    const double* __restrict__ fy,
                                                    const double* __restrict__ fz,
                                                    double* __restrict__ qt) {
  constexpr int kTile = 1024;
  __shared__ double lds[4 * kTile];
  double *t0 = lds, *t1 = lds + kTile, *t2 = lds + 2 * kTile, *t3 = lds + 3 * kTile;
  const Ent e = tab[blockIdx.y];
  const int b = blockIdx.x * kTile;
  if (b >= e.m) return;  // uniform
  const SimDesc& s = desc[e.slot];
  const EmitConsts ec = emit_consts();
  const bool tip = s.mode == RD_EMIT_TIP_FE;
  const int k = threadIdx.x;  // blockDim.x == kTile
  double val = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
  if (b + k < e.m) {
    const int kk = b + k;
    const int p = e.pbase + kk;
    if (tip) {
      const double F = field_normal(s.g, ptx[p], pty[p], fx[p], fy[p], fz[p]);
      g2 = F;
      if (F < 0.0) {
        const int nr = 100;
        int i = kk / nr + 1, j = kk % nr + 1;
        double len_xi = (s.g.max_xi - 1.0) / nr;
        double len_phi = 2.0 * 3.141592653589793238462643383279502884197 / nr;
        double xi_1 = 1.0 + (i - 1.0) * len_xi, phi_1 = (j - 1.0) * len_phi;
        double xi_2 = 1.0 + (i + 0.0) * len_xi, phi_2 = (j + 0.0) * len_phi;
        double A_f = tip_area(tip_surf(s.g), xi_1, xi_2, phi_1, phi_2);
        double t = fn_t(s.g.ic, ec.l_const, F, s.w_tip);
        // Elec_Supply (src/mod_emission_tip.f90:1711-1718)
        val = A_f * ec.a_FN * (F * F) * s.dt / (ec.q0 * s.w_tip * (t * t));
      }
    } else {
      const double F = fz[p];
      if (F < 0.0) {
        int sec;
        double w = work_at(s, wf, ptx[p], pty[p], sec);
        if (s.mode == RD_EMIT_PLANAR_FE)
          val = fn_supply(ec, s.g.ic, s.dt_q0, F, w);
        else
          val = gtf_current(F, s.T, w) * s.dt_q0;
      }
      val = (s.elx * s.ely) * val;
      g0 = fx[p];
      g1 = fy[p];
      g2 = F;
    }
  }
  t0[k] = val;
  t1[k] = g0;
  t2[k] = g1;
  t3[k] = g2;
  for (int w = kTile / 2; w > 0; w >>= 1) {
    __syncthreads();
    if (k < w) {
      t0[k] = t0[k] + t0[k + w];
      t1[k] = t1[k] + t1[k + w];
      t2[k] = t2[k] + t2[k + w];
      t3[k] = t3[k] + t3[k + w];
    }
  }
  __syncthreads();
  if (k == 0) {
    double* o = qt + 4 * (e.c_off + blockIdx.x);
    o[0] = t0[0];
    o[1] = t1[0];
    o[2] = t2[0];
    o[3] = t3[0];
  }
}

__global__ void k_quad_tsum(const Ent* tab, SimState* st, const double* __restrict__ qt) {
  const Ent e = tab[blockIdx.x];
  if (threadIdx.x != 0) return;
  double sum = 0.0, f0 = 0.0, f1 = 0.0, f2 = 0.0;
  const int nt = (e.m + 1023) / 1024;
  for (int b = 0; b < nt; ++b) {
    const double* o = qt + 4 * (e.c_off + b);
    sum = sum + o[0];
    f0 = f0 + o[1];
    f1 = f1 + o[2];
    f2 = f2 + o[3];
  }
  double* q = st[e.slot].qsum[e.lev & 1];
  q[0] = sum;
  q[1] = f0;
  q[2] = f1;
  q[3] = f2;
}

// The supply's level e.lev is final when it is the tip's rule, the last level
// allowed, or within max(0.5, 1e-3 |I|) of the level before (the reference's
// Divonne tolerances, src/mod_global.F90:89-90 with the test of
// src/mod_cuba_integration.F90:87); I = sum / (Gx Gy) electrons.  A final level
// gives N_sup, the number of chains and F_avg; otherwise S.q_final = 0 and the
// host computes the next level.
__global__ __launch_bounds__(1024) void k_quad_close(const Ent* tab, const SimDesc* desc,
                                                    SimState* st) {
  constexpr int kTile = 1024;
  __shared__ double lds[4 * kTile];
  __shared__ int s_final;
  const Ent e = tab[blockIdx.x];
  const SimDesc& s = desc[e.slot];
  SimState& S = st[e.slot];
  const bool tip = s.mode == RD_EMIT_TIP_FE;
  const int Gx = s.qgx << e.lev, Gy = s.qgy << e.lev;
  const double* cur = S.qsum[e.lev & 1];
  const double sum = cur[0];
  const double lambda = tip ? sum : sum / ((double)Gx * (double)Gy);
  if (threadIdx.x == 0) {
    int fin = 1, failed = 0;
    double err = 0.0;
    if (!tip && e.lev > 0) {
      const int Gpx = s.qgx << (e.lev - 1), Gpy = s.qgy << (e.lev - 1);
      const double prev = S.qsum[(e.lev - 1) & 1][0] / ((double)Gpx * (double)Gpy);
      double tol = 1.0e-3 * fabs(lambda);
      if (tol < 0.5) tol = 0.5;
      err = fabs(lambda - prev);
      const int conv = err <= tol;
      if (e.lev < s.quad_lmax)
        fin = conv;
      else
        failed = !conv;
    }
    S.q_final = fin;
    S.q_level = e.lev;
    S.q_cells = tip ? 100 * 100 : Gx * Gy;
    S.q_fail = failed;
    S.q_err = err;
    s_final = fin;
  }
  __syncthreads();
  if (!s_final) return;
  const double f0 = cur[1], f1 = cur[2], f2 = cur[3];
  const double M = tip ? 100.0 * 100.0 : (double)Gx * (double)Gy;
  if (s.mode == RD_EMIT_PLANAR_GTF) {
    // Rand_Poisson: the product of uniforms runs in one lane; the uniforms,
    // in stream order, are drawn kPois at a time by the whole block.
    constexpr int kPois = 4 * kTile;
    double* ut = lds;
    __shared__ int s_more;
    double lambda_left = lambda, pp = 1.0;
    int kk = 0;
    for (uint64_t b0 = 0;; b0 += kPois) {
      for (int j = threadIdx.x; j < kPois / 4; j += blockDim.x) {
        const u64x4 r =
            philox4x64((uint64_t)e.step, tag(kTagPoisson, 0), b0 / 4 + j, 0, s.seed, s.stream);
        for (int q = 0; q < 4; ++q) ut[4 * j + q] = u01(r.v[q]);
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        bool more = true;
        for (int q = 0; q < kPois && more; ++q) {
          kk = kk + 1;
          pp = pp * ut[q];
          while ((pp < 1.0) && (lambda_left > 0.0)) {
            if (lambda_left > 500.0) {
              pp = pp * rd_exp(500.0);
              lambda_left = lambda_left - 500.0;
            } else {
              pp = pp * rd_exp(lambda_left);
              lambda_left = 0.0;
            }
          }
          more = pp >= 1.0 && kk < (1 << 30);
        }
        s_more = more;
      }
      __syncthreads();
      if (!s_more) break;
    }
    if (threadIdx.x == 0) {
      S.N_sup = lambda;
      S.n_round = kk - 1;
      S.F_avg[0] = f0 / M;
      S.F_avg[1] = f1 / M;
      S.F_avg[2] = f2 / M;
    }
    return;
  }
  if (threadIdx.x != 0) return;
  int nr;
  if (tip) {
    S.N_sup = sum;
    nr = nint(sum);
    if (nr < 0) nr = 0;
    S.F_avg[0] = 0.0;
    S.F_avg[1] = 0.0;
    S.F_avg[2] = f2 / M;
  } else {
    S.N_sup = lambda;
    nr = nint(lambda + S.residual);
    S.residual = lambda - nr;
    if (nr < 0) nr = 0;
    S.F_avg[0] = f0 / M;
    S.F_avg[1] = f1 / M;
    S.F_avg[2] = f2 / M;
  }
  S.n_round = nr;
}

// ---- Metropolis-Hastings ------------------------------------------------
__global__ void k_mh_init(const Ent* tab, const SimDesc* desc, SimState* st, Chains ch) {
  const Ent e = tab[blockIdx.z];
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k == 0) {
    SimState& S = st[e.slot];
    S.pending = e.nch;
    if (desc[e.slot].mode == RD_EMIT_TIP_FE) {
      // Metro_algo_tip_v3 brings the shared width into its limits first
      if (S.MH_std > 0.125)
        S.MH_std = 0.125;
      else if (S.MH_std < 0.0005)
        S.MH_std = 0.0005;
    }
  }
  if (k >= e.nch) return;
  const int c = e.cbase + k;
  ch.ok[c] = 0;
  ch.acc[c] = 0;
  ch.rej[c] = 0;
  ch.pv[c] = 1;
}

// Uniform start proposals for the chains without a start site.
__global__ void k_mh_start(const Ent* tab, const SimDesc* desc, Chains ch, int round,
                           double* __restrict__ ptx, double* __restrict__ pty,
                           double* __restrict__ ptz) {
  const Ent e = tab[blockIdx.z];
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= e.nch) return;
  const int c = e.cbase + k;
  if (ch.ok[c]) return;
  const SimDesc& s = desc[e.slot];
  u64x4 r = philox4x64((uint64_t)e.step, tag(kTagStart, (uint64_t)round), (uint64_t)k, 0, s.seed,
                       s.stream);
  double u1 = u01(r.v[0]), u2 = u01(r.v[1]);
  double a, b, x, y, z;
  if (s.mode == RD_EMIT_TIP_FE) {
    a = 1.0 + (s.g.max_xi - 1.0) * u1;
    b = 2.0 * 3.141592653589793238462643383279502884197 * u2;
    tip_xyz(tip_surf(s.g), a, s.g.eta_1, b, x, y, z);
  } else {
    a = u1 * s.elx + s.ex0;
    b = u2 * s.ely + s.ey0;
    x = a;
    y = b;
    z = 0.0;
  }
  ch.px[c] = a;
  ch.py[c] = b;
  ptx[e.pbase + k] = x;
  pty[e.pbase + k] = y;
  ptz[e.pbase + k] = z;
}

__global__ __launch_bounds__(256) void k_mh_start_check(const Ent* tab, const SimDesc* desc,
                                                        SimState* st, const double* wf,
                                                        Chains ch, int round, int last,
                                                        const double* __restrict__ ptx,
                                                        const double* __restrict__ pty,
                                                        const double* __restrict__ ptz,
                                                        const double* __restrict__ fx,
                                                        const double* __restrict__ fy,
                                                        const double* __restrict__ fz) {
  __shared__ int red[256];
  const Ent e = tab[blockIdx.x];
  const SimDesc& s = desc[e.slot];
  const EmitConsts ec = emit_consts();
  int pend = 0;
  for (int k = threadIdx.x; k < e.nch; k += blockDim.x) {
    const int c = e.cbase + k;
    if (ch.ok[c]) continue;
    const int p = e.pbase + k;
    double F = (s.mode == RD_EMIT_TIP_FE)
                   ? field_normal(s.g, ptx[p], pty[p], fx[p], fy[p], fz[p])
                   : fz[p];
    if (F < 0.0) {
      ch.ok[c] = 1;
      ch.cx[c] = ch.px[c];
      ch.cy[c] = ch.py[c];
      ch.F[c] = F;
      ch.sup[c] = chain_target(s, wf, ec, F, ch.px[c], ch.py[c], ptx[p], pty[p], ptz[p]);
      ch.ox[c] = ptx[p];
      ch.oy[c] = pty[p];
      ch.oz[c] = ptz[p];
    } else if (last) {
      // failed chain: unfavourable field, no emission
      ch.ok[c] = -1;
      ch.F[c] = 1.0;
      ch.sup[c] = -kHuge;
      if (s.mode == RD_EMIT_TIP_FE) {
        double x, y, z;
        tip_xyz(tip_surf(s.g), 1.0, s.g.eta_1, 0.0, x, y, z);
        ch.cx[c] = 1.0;
        ch.cy[c] = 0.0;
        ch.ox[c] = x;
        ch.oy[c] = y;
        ch.oz[c] = z;
      } else {
        ch.cx[c] = s.ex0;
        ch.cy[c] = s.ey0;
        ch.ox[c] = s.ex0;
        ch.oy[c] = s.ey0;
        ch.oz[c] = 0.0;
      }
    } else {
      ++pend;
    }
  }
  red[threadIdx.x] = pend;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    st[e.slot].pending = red[0];
    st[e.slot].start_rounds = round;
  }
}

// Gaussian proposal of jump `it` (1-based) for chain k of entry e, with the
// width MH of the start of the jump and the unit normal pair (u1, u2) of the
// chain's (step, jump) stream (box_muller_unit).  (a, b) are the chain coordinates
// (planar x, y; tip xi, phi) and (x, y, z) the surface point whose field the
// accept step needs.  valid: 1 proposal inside the patch, 0 tip proposal
// outside it, -1 chain without a start site (the point is its final site and
// nothing is recorded).
struct Proposal {
  double a, b, x, y, z;
  int valid;
};

__device__ __forceinline__ Proposal mh_proposal(const Ent& e, const SimDesc& s, double MH,
                                                const Chains& ch, int it, int k, double u1,
                                                double u2) {
  Proposal P;
  const int c = e.cbase + k;
  if (ch.ok[c] != 1) {
    P.a = P.b = 0.0;
    P.x = ch.ox[c];
    P.y = ch.oy[c];
    P.z = ch.oz[c];
    P.valid = -1;
    return P;
  }
  const double PI2 = 2.0 * 3.141592653589793238462643383279502884197;
  if (s.mode == RD_EMIT_TIP_FE) {
    const int first = nint(s.ndim * 0.25);
    double s1, s2;
    if (it > first) {
      s1 = (s.g.max_xi - 1.0) * MH;
      s2 = PI2 * MH;
    } else {
      s1 = (s.g.max_xi - 1.0) * 0.10;
      s2 = PI2 * 0.10;
    }
    const double d1 = u1 * s1 + 0.0, d2 = u2 * s2 + 0.0;
    double nxi = ch.cx[c] + d1;
    double nphi = fmodulo(ch.cy[c] + d2, PI2);
    if (nxi > s.g.max_xi) nxi = 2.0 * s.g.max_xi - nxi;
    if (nxi < 1.0) nxi = 2.0 - nxi;
    P.valid = !((nxi < 1.0) || (nxi > s.g.max_xi));
    P.a = nxi;
    P.b = nphi;
    if (P.valid) {
      tip_xyz(tip_surf(s.g), nxi, s.g.eta_1, nphi, P.x, P.y, P.z);
    } else {
      P.x = ch.ox[c];
      P.y = ch.oy[c];
      P.z = ch.oz[c];
    }
    return P;
  }
  double sx, sy;
  if (s.mode == RD_EMIT_PLANAR_FE) {
    const int first = nint(s.ndim * 0.25);
    double f = (it > first) ? MH : 0.10;
    sx = s.elx * f;
    sy = s.ely * f;
  } else {
    sx = s.elx * MH;
    sy = s.ely * MH;
  }
  double nx = u1 * sx + ch.cx[c], ny = u2 * sy + ch.cy[c];
  if (s.mode == RD_EMIT_PLANAR_FE) {
    // check_limits_metro_rec (src/mod_field_emission_v2.F90:1466-1516): reflect
    double x_max = s.ex0 + s.elx, x_min = s.ex0;
    double y_max = s.ey0 + s.ely, y_min = s.ey0;
    if (nx > x_max)
      nx = x_max - (nx - x_max);
    else if (nx < x_min)
      nx = (x_min - nx) + x_min;
    if (ny > y_max)
      ny = y_max - (ny - y_max);
    else if (ny < y_min)
      ny = (y_min - ny) + y_min;
  } else {
    // check_limits_metro_rec (src/mod_field_thermo_emission.F90:369-389): wrap
    double sxs = (nx - s.ex0) / s.elx, sys = (ny - s.ey0) / s.ely;
    if ((sxs > 1.0) || (sxs < 0.0)) sxs = 1.0 - (sxs - __builtin_floor(sxs));
    if ((sys > 1.0) || (sys < 0.0)) sys = 1.0 - (sys - __builtin_floor(sys));
    nx = sxs * s.elx + s.ex0;
    ny = sys * s.ely + s.ey0;
  }
  P.a = nx;
  P.b = ny;
  P.x = nx;
  P.y = ny;
  P.z = 0.0;
  P.valid = 1;
  return P;
}

// Record a proposal: the chain's proposed coordinates and its field point.
__device__ __forceinline__ void mh_record(const Ent& e, const SimDesc& s, const Chains& ch, int k,
                                          const Proposal& P, double* __restrict__ ptx,
                                          double* __restrict__ pty, double* __restrict__ ptz) {
  const int c = e.cbase + k;
  const int p = e.pbase + k;
  if (P.valid >= 0) {
    if (s.mode == RD_EMIT_TIP_FE) ch.pv[c] = P.valid;
    ch.px[c] = P.a;
    ch.py[c] = P.b;
  }
  ptx[p] = P.x;
  pty[p] = P.y;
  ptz[p] = P.z;
}

// The planar FE width update of one jump from the batch counts
// (MH_std_update, src/mod_field_emission_v2.F90:603-612).
__device__ __forceinline__ double fe_width_after(double mh, int it, int first, int na, int nr) {
  if ((it > first) && (na + nr > 0)) {
    mh = mh * rd_exp(0.025 * ((double)na / (double)(na + nr) - 0.35));
    if (mh > 0.1250)
      mh = 0.1250;
    else if (mh < 0.00005)
      mh = 0.00005;
  }
  return mh;
}
__device__ __forceinline__ void fe_width_update(SimState& S, int it, int first, int na, int nr) {
  S.it_a = na;
  S.it_r = nr;
  if ((it > first) && (na + nr > 0)) {
    S.a_rate = (double)na / (double)(na + nr);
    S.MH_std = S.MH_std * rd_exp(0.025 * (S.a_rate - 0.35));
    if (S.MH_std > 0.1250)
      S.MH_std = 0.1250;
    else if (S.MH_std < 0.00005)
      S.MH_std = 0.00005;
  }
}

// What the accept step of one chain reads, loaded together ahead of the
// field sums so that the loads overlap.
struct AcceptIn {
  int ok, pv;            // start found; proposal valid (tip)
  double px, py, sup;    // proposal (chain coordinates), current log target
  double lr;             // log-uniform of the jump
  double x, y, z;        // proposal point
};
__device__ __forceinline__ AcceptIn accept_inputs(const Ent& e, const SimDesc& s,
                                                  const Chains& ch, const double* lr,
                                                  const double* ptx, const double* pty,
                                                  const double* ptz, int k) {
  AcceptIn in{};
  if (k < 0) return in;
  const int c = e.cbase + k, p = e.pbase + k;
  in.ok = ch.ok[c];
  in.pv = s.mode == RD_EMIT_TIP_FE ? ch.pv[c] : 1;
  in.px = ch.px[c];
  in.py = ch.py[c];
  in.sup = ch.sup[c];
  in.lr = lr[c];
  in.x = ptx[p];
  in.y = pty[p];
  in.z = ptz[p];
  return in;
}

// Accept/reject of jump `it` for chain k of entry e (k < 0: no chain), one
// chain per lane of a block (a multiple of 64 lanes; ra, rr: LDS arrays of
// blockDim.x / 64 entries for the batch acceptance counts; in: the chain's inputs;
// (fxk, fyk, fzk): the field at the chain's proposal).  Planar FE: with
// counts == nullptr the block holds every chain and updates the width
// itself; otherwise it stores its (accepted, rejected) counts there (see
// split_width).
__device__ void mh_accept_block(const Ent& e, const SimDesc& s, SimState* st, const double* wf,
                                const Chains& ch, int it, const AcceptIn& in, double fxk,
                                double fyk, double fzk, int* ra, int* rr, int k, int* counts) {
  const EmitConsts ec = emit_consts();
  const int first = nint(s.ndim * 0.25);
  const bool tip = s.mode == RD_EMIT_TIP_FE;
  const bool counted = (s.mode == RD_EMIT_PLANAR_GTF) || (it > first);
  int na = 0, nrj = 0;
  const int c = e.cbase + k;
  if (k >= 0 && in.ok == 1) {
    bool acc = false;
    bool rejected = false;
    double F = 0.0;
    if (tip && !in.pv) {
      rejected = true;
    } else {
      F = tip ? field_normal(s.g, in.x, in.y, fxk, fyk, fzk) : fzk;
      bool bad = (s.mode == RD_EMIT_PLANAR_GTF) ? (F > 0.0) : (F >= 0.0);
      if (bad) {
        rejected = true;
      } else {
        double sup_new = chain_target(s, wf, ec, F, in.px, in.py, in.x, in.y, in.z);
        double alpha = sup_new - in.sup;
        if (sup_new >= in.sup) {
          acc = true;
        } else {
          acc = in.lr <= alpha;
        }
        if (acc) {
          ch.cx[c] = in.px;
          ch.cy[c] = in.py;
          ch.sup[c] = sup_new;
          ch.F[c] = F;
          ch.ox[c] = in.x;
          ch.oy[c] = in.y;
          ch.oz[c] = in.z;
        } else {
          rejected = true;
        }
      }
    }
    if (acc) ++na;
    if (rejected) ++nrj;
    if (counted && !tip && s.mode == RD_EMIT_PLANAR_GTF) {
      if (acc) ch.acc[c] += 1;
      if (rejected) ch.rej[c] += 1;
    }
    if (tip && it > first) {
      if (acc) ch.acc[c] += 1;
      if (rejected) ch.rej[c] += 1;
    }
  }
  if (s.mode != RD_EMIT_PLANAR_FE) return;  // per-chain schedules: no batch update
  // batch counts: one chain per lane, so a ballot per wavefront, then the
  // wavefront counts (integers: any order gives the same sums)
  const unsigned long long ba = __ballot(na), br = __ballot(nrj);
  if ((threadIdx.x & 63) == 0) {
    ra[threadIdx.x >> 6] = __popcll(ba);
    rr[threadIdx.x >> 6] = __popcll(br);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int ta = 0, tr = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      ta += ra[w];
      tr += rr[w];
    }
    if (counts) {
      counts[0] = ta;
      counts[1] = tr;
    } else {
      fe_width_update(st[e.slot], it, first, ta, tr);
    }
  }
}

// ------------------------------------------------------------------------
// One Metropolis-Hastings jump in two launches.
//
// k_mh_field64 / k_mh_field32: every workgroup proposes the jump for its own
// chains (the proposals are counter-based, so every workgroup of a chain
// computes the same one; the first source segment's workgroup records it)
// and writes the field partial sums of one source segment.  k_mh_settle: one
// workgroup per kSettle chains of a simulation adds the partials in a fixed
// order and runs the accept step for those chains (the planar FE width of
// simulations whose chains take several workgroups: see split_width).
// Kernel boundaries are the only synchronisation: no counters, no
// device-wide fences.
struct JumpArgs {
  double *ptx, *pty, *ptz, *fx, *fy, *fz;
  double* part;
  const double *X, *Y, *Z, *Q;    // FP64 sources
  const float4 *HI, *LO;          // packed sources (mixed)
  const float4* SRF;              // (x, y, z, q z) high parts (mixed, cathode surface)
  const double *ru1, *ru2, *rlr;  // per jump and chain: unit normals, log-uniform ([it-1][C])
  int* counts;                    // [jump parity][accept workgroup][accepted, rejected]
  int C;                          // chains in all
  int nsettle;                    // accept workgroups of a jump
};

constexpr int kSettle = 256;  // lanes of a finish-and-accept workgroup

// Workgroups of a simulation's finish-and-accept: one per kSettle chains.
__host__ __device__ inline int settle_groups(int m) { return max(1, (m + kSettle - 1) / kSettle); }

// Planar FE simulations whose accept step takes several workgroups ("split"):
// the workgroups of jump j store their batch counts in counts[j & 1]; the
// width after jump j, f(width after j - 1, counts of j), is formed by every
// proposal of jump j + 1 from the stored width and those counts, and stored
// into the state by the first accept workgroup of jump j + 1 (nothing else
// reads the stored width during that launch, and nothing writes it during a
// proposal launch).  The width after the last jump is stored by
// k_mh_width_last.  No launch is added per jump.
__device__ __forceinline__ bool split_width(const SimDesc& s, const Ent& e) {
  return s.mode == RD_EMIT_PLANAR_FE && settle_groups(e.nch) > 1;
}
__device__ __forceinline__ void split_counts(const Ent& e, const int* counts, int nsettle, int j,
                                             int& na, int& nr) {
  const int* cn = counts + (size_t)(j & 1) * 2 * nsettle + 2 * (size_t)e.wg1;
  na = 0;
  nr = 0;
  for (int g = 0; g < settle_groups(e.nch); ++g) {
    na += cn[2 * g];
    nr += cn[2 * g + 1];
  }
}

// The random numbers of every jump of every chain, drawn before the jumps so
// that none of them is on a jump's critical path: the unit normal pair of the
// proposal and the log of the accept uniform, from the chain's (step, jump)
// streams.
__global__ void k_mh_random(const Ent* tab, int ne, const SimDesc* desc, int C,
                            double* __restrict__ u1, double* __restrict__ u2,
                            double* __restrict__ lr) {
  const int cg = blockIdx.x * blockDim.x + threadIdx.x;  // chain over all entries
  if (cg >= C) return;
  int lo = 0, hi = ne - 1;  // entry holding chain cg (chains are laid out in entry order)
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tab[mid].cbase <= cg)
      lo = mid;
    else
      hi = mid - 1;
  }
  const Ent e = tab[lo];
  const SimDesc& s = desc[e.slot];
  const int k = cg - e.cbase;
  const int it = blockIdx.y + 1;
  if (k >= e.nch || it > s.ndim) return;
  const size_t o = (size_t)(it - 1) * C + e.cbase + k;
  double a, b;
  box_muller_unit(s.seed, s.stream, (uint64_t)e.step, tag(kTagJump, (uint64_t)it), (uint64_t)k, a,
                  b);
  u1[o] = a;
  u2[o] = b;
  lr[o] = rd_log(
      uniform(s.seed, s.stream, (uint64_t)e.step, tag(kTagAccept, (uint64_t)it), (uint64_t)k, 0));
}

__device__ __forceinline__ Proposal jump_proposal(const Ent& e, const SimDesc& s,
                                                  const SimState* st, const Chains& ch, int it,
                                                  int k, const JumpArgs& a) {
  const size_t o = (size_t)(it - 1) * a.C + e.cbase + k;
  double mh = st[e.slot].MH_std;
  if (it >= 2 && split_width(s, e)) {
    int na, nr;
    split_counts(e, a.counts, a.nsettle, it - 1, na, nr);
    mh = fe_width_after(mh, it - 1, nint(s.ndim * 0.25), na, nr);
  }
  return mh_proposal(e, s, mh, ch, it, k, a.ru1[o], a.ru2[o]);
}

// FP64: one lane per chain, 256 chains per workgroup, sources in 256-source
// chunks (the canonical order of the field at points: partials per chunk,
// bit-identical to the oracle).
template <int KIND>
__global__ __launch_bounds__(kBlock) void k_mh_field64(const Ent* tab, int ne,
                                                       const SimDesc* desc, const SimState* st,
                                                       Chains ch, int it, JumpArgs a) {
  __shared__ double sx[kChunk], sy[kChunk], sz[kChunk], sq[kChunk];
  const Ent e = tab[find_entry(tab, ne, blockIdx.x)];
  const SimDesc& s = desc[e.slot];
  if (s.g.kind != KIND || it > s.ndim) return;
  const int m = e.nch;
  const int gx = (m + kBlock - 1) / kBlock;
  const int local = blockIdx.x - e.wg0;
  const int bx = local % gx, by = local / gx;  // chain group, source segment
  const DevGeom g = s.g;
  const int n = st[e.slot].n;
  const int nchunks = (n + kChunk - 1) / kChunk;
  const int k = bx * kBlock + threadIdx.x;
  const bool act = k < m;
  Proposal P{};
  if (act) {
    P = jump_proposal(e, s, st, ch, it, k, a);
    if (by == 0) mh_record(e, s, ch, k, P, a.ptx, a.pty, a.ptz);
  }
  const int c0 = by * e.cps;
  const int c1 = min(nchunks, c0 + e.cps);
  if (c0 < nchunks)
    field_chunks64<KIND>(g, act, P.x, P.y, P.z, a.X + e.base, a.Y + e.base, a.Z + e.base,
                         a.Q + e.base, n, c0, c1, sx, sy, sz, sq, a.part + e.part_off,
                         (size_t)e.nchunks_ub * m, m, k);
}

// Mixed: 64 chains per workgroup, one per lane; the kJumpWaves wavefronts
// take an equal slice of every 256-source chunk of the segment each (sources
// staged in LDS and read as wavefront-wide broadcasts, or read through the
// scalar cache), and their FP64 lane sums are added in wavefront order.  Many
// short slices keep the per-wavefront dependency chain short: a lone jump is
// latency-bound (4 wavefronts per workgroup measured slower in pools too).
// A workgroup computes PPS consecutive segment partials (the set-up above the
// source loop is paid once for all of them); the partials are the same for
// every PPS.  At most 64 VGPRs (eight waves per SIMD: two workgroups per CU).
constexpr int kJumpTargets = 64;
constexpr int kJumpWaves = 16;
constexpr int kJumpSlice = kChunk / kJumpWaves;

// One mixed pair term (general series).
__device__ __forceinline__ void mixed_pair(const float4 h1, const float4 l1, const float4 h2,
                                           const float4 l2, int ic, int nic, float two_d,
                                           double& fx, double& fy, double& fz) {
  float dx = (h1.x - h2.x) + (l1.x - l2.x);
  float dy = (h1.y - h2.y) + (l1.y - l2.y);
  float dxy2 = __builtin_fmaf(dy, dy, dx * dx);
  float dz = (h1.z - h2.z) + (l1.z - l2.z);
  float r2 = fmaxf(__builtin_fmaf(dz, dz, dxy2), kR2Min);
  float ri = rsq(r2);
  float r3 = ri * ri * ri;
  float cxy = r3, cz = dz * r3;
  if (ic) {
    float wxy, wz;
    ic_planar32(h1.z + h2.z, l1.w + l2.w, dz, dxy2, nic, two_d, wxy, wz);
    cxy += wxy;
    cz += wz;
  }
  cxy *= h2.w;
  cz *= h2.w;
  fx += (double)(cxy * dx);
  fy += (double)(cxy * dy);
  fz += (double)cz;
}

// Layout of a simulation's chains in the mixed jump field, chosen from its
// own chain count only: with few chains one wavefront per chain (its lanes
// stride over the sources), otherwise one lane per chain (the wavefronts
// take slices of every chunk).
constexpr int kWaveChains = 32;  // at most this many chains: a wavefront per chain
__host__ __device__ inline bool jump_by_wave(int m) { return m <= kWaveChains; }
__host__ __device__ inline int jump_groups32(int m) {
  return jump_by_wave(m) ? (m + kJumpWaves - 1) / kJumpWaves
                         : (m + kJumpTargets - 1) / kJumpTargets;
}

template <int KIND, int PPS>
__global__ __launch_bounds__(64 * kJumpWaves, 8) void k_mh_field32(const Ent* tab, int ne,
                                                                const SimDesc* desc,
                                                                const SimState* st, Chains ch,
                                                                int it, JumpArgs a) {
  __shared__ float4 shi[kChunk], slo[kChunk];
  // the partials' sums alternate between two buffers: the barrier of partial
  // q + 1 orders every read of partial q - 1's buffer before its reuse
  __shared__ double red[PPS > 1 ? 2 : 1][3][kJumpWaves][kJumpTargets];
  __shared__ float4 tgt[4][kJumpTargets];  // wavefront 0's packed targets
  const Ent e = tab[find_entry(tab, ne, blockIdx.x)];
  const SimDesc& s = desc[e.slot];
  if (s.g.kind != KIND || it > s.ndim) return;
  const int m = e.nch;
  const bool by_wave = jump_by_wave(m);
  const int gx = jump_groups32(m);
  const int local = blockIdx.x - e.wg0;
  const int bx = local % gx, by = local / gx;  // chain group, workgroup segment
  const DevGeom& g = s.g;
  const int n = st[e.slot].n;
  const int nchunks = (n + kChunk - 1) / kChunk;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  // the chain of this lane (uniform across a wavefront when by_wave)
  const int k = by_wave ? bx * kJumpWaves + w : bx * kJumpTargets + lane;
  const bool act = k < m;
  float4 h1 = make_float4(0, 0, 0, 0), l1 = make_float4(0, 0, 0, 0);
  float4 ih1 = make_float4(0, 0, 0, 1), il1 = make_float4(0, 0, 0, 0);
  // One lane per chain: every wavefront holds the same kJumpTargets chains, so
  // wavefront 0 forms and packs their proposals and the others read the packed
  // targets from LDS (by_wave depends on m alone: uniform over the workgroup).
  if (act && (by_wave || w == 0)) {
    const Proposal P = jump_proposal(e, s, st, ch, it, k, a);
    if (by == 0 && (!by_wave || lane == 0)) mh_record(e, s, ch, k, P, a.ptx, a.pty, a.ptz);
    pack1(P.x, P.y, P.z, 0.0, g.d, h1, l1);
    if (KIND == RD_GEOM_TIP) pack_image(P.x, P.y, P.z, g.z0, g.R, ih1, il1);
  }
  if (!by_wave) {
    if (w == 0) {
      tgt[0][lane] = h1;
      tgt[1][lane] = l1;
      tgt[2][lane] = ih1;
      tgt[3][lane] = il1;
    }
    __syncthreads();
    if (w != 0) {
      h1 = tgt[0][lane];
      l1 = tgt[1][lane];
      ih1 = tgt[2][lane];
      il1 = tgt[3][lane];
    }
  }
  const float4* hi = a.HI + e.base;
  const float4* lo = a.LO + e.base;
  const float4* srf = a.SRF + e.base;
  const int ic = g.ic, nic = g.nic;
  const float two_d = (float)(2.0 * g.d * 1e9);
  const float kt = g.ic ? (float)g.ktip : 0.0f;
  const float Rn = (float)(g.R * 1e9);
  const bool surf = KIND == RD_GEOM_PLANAR && wave_on_surface(ic, act, h1, l1);
  // Planar chains on the cathode, one per lane: a wavefront's sources are
  // the same for all its lanes, so they are read through the scalar cache
  // into SGPRs (no LDS traffic) in the same order as the staged loop below.
  // Every wavefront of the workgroup holds the same chains, so the branch is
  // uniform over the workgroup.
  const bool scalar_src = KIND == RD_GEOM_PLANAR && !by_wave &&
                          __syncthreads_or(surf ? 1 : 0) && __syncthreads_and(surf ? 1 : 0);
  double* part = a.part + e.part_off;
  const size_t S = (size_t)e.segs * m;
#pragma unroll
  for (int q = 0; q < PPS; ++q) {
    const int sg = by * PPS + q;  // the partial
    if (sg >= e.segs) break;
    const int c0 = sg * e.cps;
    const int c1 = min(nchunks, c0 + e.cps);
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (scalar_src) {
      const int wu = __builtin_amdgcn_readfirstlane(w);
      const float dnm = 0.5f * two_d;
      for (int c = c0; c < c1; ++c) {
        const int jb = c * kChunk + wu * kJumpSlice;
        const int je = min(n, jb + kJumpSlice);
        if (act) {  // fz holds the sum without the factor -2 until the end
          static_assert(kJumpSlice <= kSurfBlock, "a slice is one FP32 block");
          auto sweep = [&](auto n0) {
            constexpr bool N0 = decltype(n0)::value;
            float acc = 0.f;  // N0: the slice's block sum
            int j = jb;
            for (; j + 4 <= je; j += 4) {  // four sources' loads in flight together
              float4 hh[4], ll[4];
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                hh[u] = load_uniform(N0 ? srf : hi, j + u);
                ll[u] = load_uniform(lo, j + u);
              }
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                if (N0)
                  acc = surface_accq32(h1, l1, hh[u], ll[u], acc);
                else
                  fz += (double)surface_term32<false>(h1, l1, hh[u], ll[u], nic, dnm);
              }
            }
            for (; j < je; ++j) {
              if (N0)
                acc = surface_accq32(h1, l1, load_uniform(srf, j), load_uniform(lo, j), acc);
              else
                fz += (double)surface_term32<false>(h1, l1, load_uniform(hi, j),
                                                    load_uniform(lo, j), nic, dnm);
            }
            if (N0) fz += (double)acc;
          };
          if (nic == 0)
            sweep(std::true_type{});
          else
            sweep(std::false_type{});
        }
      }
    }
    for (int c = scalar_src ? c1 : c0; c < c1; ++c) {
      const int jb = c * kChunk;
      __syncthreads();
      if (t < kChunk && jb + t < n) {
        shi[t] = hi[jb + t];
        slo[t] = lo[jb + t];
      }
      __syncthreads();
      const int cnt = min(kChunk, n - jb);
      // sources of this lane in the chunk: a slice per wavefront, or a stride
      // of 64 over the chunk per lane
      const int j0 = by_wave ? lane : w * kJumpSlice;
      const int j1 = by_wave ? cnt : min(cnt, j0 + kJumpSlice);
      const int dj = by_wave ? 64 : 1;
      if (act) {
        if (KIND == RD_GEOM_PLANAR) {
          // one loop per form, so that no branch sits in the source loop
          if (surf) {  // fz holds the sum without the factor -2 until the end
            const float dnm = 0.5f * two_d;
            if (nic == 0) {
              // a lane's sources of the chunk: a slice of kJumpSlice, or at most
              // kChunk / 64 by wavefront: one FP32 block either way
              static_assert(kChunk / 64 <= kSurfBlock, "one FP32 block per chunk");
              float acc = 0.f;
              for (int jj = j0; jj < j1; jj += dj)
                acc = surface_acc32(h1, l1, shi[jj], slo[jj], acc);
              fz += (double)acc;
            } else
              for (int jj = j0; jj < j1; jj += dj)
                fz += (double)surface_term32<false>(h1, l1, shi[jj], slo[jj], nic, dnm);
          } else if (ic) {
            for (int jj = j0; jj < j1; jj += dj)
              mixed_pair(h1, l1, shi[jj], slo[jj], 1, nic, two_d, fx, fy, fz);
          } else {
            for (int jj = j0; jj < j1; jj += dj)
              mixed_pair(h1, l1, shi[jj], slo[jj], 0, 0, two_d, fx, fy, fz);
          }
        } else {
          for (int jj = j0; jj < j1; jj += dj) {
            float cx, cy, cz;
            tip_pair32_target(h1, l1, ih1, il1, shi[jj], slo[jj], kt, Rn, cx, cy, cz);
            fx += (double)cx;
            fy += (double)cy;
            fz += (double)cz;
          }
        }
      }
    }
    if (surf) fz = -2.0 * fz;
    const int qb = q & 1;
    red[qb][0][w][lane] = fx;
    red[qb][1][w][lane] = fy;
    red[qb][2][w][lane] = fz;
    __syncthreads();
    if (by_wave) {
      // lane q of wavefront w adds component q over the 64 lanes, in lane order
      if (lane < 3 && act) {
        double v = red[qb][lane][w][0];
        for (int u = 1; u < 64; ++u) v += red[qb][lane][w][u];
        part[lane * S + (size_t)sg * m + k] = v;
      }
    } else if (w < 3 && act) {
      // wavefront q adds component q over the wavefronts, in wavefront order
      double v = red[qb][w][0][lane];
      for (int u = 1; u < kJumpWaves; ++u) v += red[qb][w][u][lane];
      part[w * S + (size_t)sg * m + k] = v;
    }
  }
}

// Finish and accept: one workgroup per kSettle chains of a simulation.  FP64
// adds the chunk partials of every chain in chunk order (plus the vacuum
// field); mixed adds the segment partials in segment order.
template <int PREC>
__global__ __launch_bounds__(kSettle) void k_mh_settle(const Ent* tab, int ne,
                                                       const SimDesc* desc, SimState* st,
                                                       const double* wf, Chains ch, int it,
                                                       JumpArgs a) {
  __shared__ int ra[kSettle / 64], rr[kSettle / 64];
  const Ent e = tab[find_entry(tab, ne, blockIdx.x, true)];
  const SimDesc& s = desc[e.slot];
  if (it > s.ndim) return;
  const int m = e.nch;
  const int grp = blockIdx.x - e.wg1;
  const int k0 = grp * kSettle, k1 = min(m, k0 + kSettle);
  const int n = st[e.slot].n;
  const int nchunks = (n + kChunk - 1) / kChunk;
  const double* part = a.part + e.part_off;
  // one chain per lane: k1 - k0 <= kSettle
  const int k = k0 + (int)threadIdx.x < k1 ? k0 + (int)threadIdx.x : -1;
  const AcceptIn in =
      accept_inputs(e, s, ch, a.rlr + (size_t)(it - 1) * a.C, a.ptx, a.pty, a.ptz, k);
  double fxk = 0.0, fyk = 0.0, fzk = 0.0;
  if (k >= 0) {
    const int p = e.pbase + k;
    if (PREC == RD_PREC_FP64) {
      field_finish64(s.g, part, (size_t)e.nchunks_ub * m, m, k, nchunks, in.x, in.y, in.z, fxk,
                     fyk, fzk);
    } else {
      const int segs = nchunks == 0 ? 1 : min(e.segs, (nchunks + e.cps - 1) / e.cps);
      if (s.g.kind == RD_GEOM_TIP)
        field_finish32_tip(s.g, part, (size_t)e.segs * m, m, k, segs, in.x, in.y, in.z, fxk,
                           fyk, fzk);
      else
        field_finish32(s.g, part, (size_t)e.segs * m, m, k, segs, fxk, fyk, fzk);
    }
    a.fx[p] = fxk;
    a.fy[p] = fyk;
    a.fz[p] = fzk;
  }
  const bool split = split_width(s, e);
  if (split && grp == 0 && threadIdx.x == 0 && it >= 2) {
    // store the width after the previous jump (its proposals formed it too)
    int na, nr;
    split_counts(e, a.counts, a.nsettle, it - 1, na, nr);
    fe_width_update(st[e.slot], it - 1, nint(s.ndim * 0.25), na, nr);
  }
  mh_accept_block(e, s, st, wf, ch, it, in, fxk, fyk, fzk, ra, rr, k,
                  split ? a.counts + (size_t)(it & 1) * 2 * a.nsettle + 2 * (size_t)blockIdx.x
                        : nullptr);
}

// Split planar FE simulations (split_width): the width after their last jump.
__global__ void k_mh_width_last(const Ent* tab, const SimDesc* desc, SimState* st,
                                const int* __restrict__ counts, int nsettle) {
  const Ent e = tab[blockIdx.x];
  const SimDesc& s = desc[e.slot];
  if (threadIdx.x != 0 || s.ndim < 1 || !split_width(s, e)) return;
  int na, nr;
  split_counts(e, counts, nsettle, s.ndim, na, nr);
  fe_width_update(st[e.slot], s.ndim, nint(s.ndim * 0.25), na, nr);
}

// ------------------------------------------------------------------------
// Every jump of a small planar simulation in one workgroup (mixed precision).
//
// A lone small simulation (a few chains over a few hundred electrons) spends
// two launches per jump on a few microseconds of pair terms: the launches,
// not the terms, set its step time.  Here one workgroup of kSoloThreads lanes
// takes all jumps 1..ndim of one simulation, with barriers between the
// phases of a jump:
//   1. lane k < m proposes for chain k (jump_proposal: the same random
//      numbers, width and reflection as the two-launch path), records the
//      proposal and packs its point into LDS;
//   2. the lanes form mp columns (mp: the power of two >= m) and
//      kSoloThreads / mp rows, lane t taking chain t % mp and the sources
//      j = row, row + rows, ... of each staged block, with the FP32 pair terms
//      of the two-launch path added in FP64 (on the cathode with N_ic_max = 0,
//      FP32 blocks of at most kSurfBlock of the lane's own terms); the rows'
//      sums are then added pairwise in a fixed tree;
//   3. lane k < m forms the field (field_from_sums32) and runs the accept
//      step with the batch width update of mh_accept_block.
// The sources do not move during the jumps: up to kSoloStage of them stay in
// LDS for all jumps, more are staged per jump.  Which simulations run here
// depends on the kind of context (a lone simulation) and on the simulation's
// own chain count, electron count and image series (solo_jumps).  Tip
// simulations start hundreds of chains or more and keep the two-launch path.
constexpr int kSoloThreads = 1024;
constexpr int kSoloStage = 2048;     // sources held in LDS
constexpr double kSoloWork = 2.0e6;  // most estimated VALU work per jump, in lane-ops

// Estimated lane-ops of one jump's pair terms: chains x electrons x the
// instructions of one term (surface form with N_ic_max = 0: 12, each image
// partner pair about 10 more; without images: 25).  At 64 lane-ops per cycle
// on one CU, kSoloWork is about 13 us per jump, against about 15 us for the
// two launches of a lone simulation.
__host__ __device__ inline double solo_work(int ic, int nic, int m, int n) {
  const double per = ic ? 12.0 + 10.0 * nic : 25.0;
  return per * (double)m * (double)(n > 1 ? n : 1);
}
__host__ __device__ inline bool solo_jumps(int kind, int ic, int nic, int m, int n) {
  return kind == RD_GEOM_PLANAR && m >= 1 && m <= kSettle && solo_work(ic, nic, m, n) <= kSoloWork;
}

__global__ __launch_bounds__(kSoloThreads) void k_mh_solo32(const Ent* tab, const SimDesc* desc,
                                                            SimState* st, const double* wf,
                                                            Chains ch, JumpArgs a) {
  __shared__ float4 shi[kSoloStage], slo[kSoloStage];
  __shared__ float4 tgt[2][kSettle];
  __shared__ double red[3][kSoloThreads];
  __shared__ int ra[kSoloThreads / 64], rr[kSoloThreads / 64];
  const Ent e = tab[blockIdx.x];
  const SimDesc& s = desc[e.slot];
  const DevGeom& g = s.g;
  const int m = e.nch;
  const int n = st[e.slot].n;
  const int t = threadIdx.x;
  int lmp = 0;
  while ((1 << lmp) < m) ++lmp;
  const int mp = 1 << lmp, rows = kSoloThreads >> lmp;
  const int col = t & (mp - 1), row = t >> lmp;
  const bool act = col < m;
  const float4* hi = a.HI + e.base;
  const float4* lo = a.LO + e.base;
  const int ic = g.ic, nic = g.nic;
  const float two_d = (float)(2.0 * g.d * 1e9);
  const float dnm = 0.5f * two_d;
  const bool resident = n <= kSoloStage;
  if (resident) {
    for (int j = t; j < n; j += kSoloThreads) {
      shi[j] = hi[j];
      slo[j] = lo[j];
    }
  }
  for (int it = 1; it <= s.ndim; ++it) {
    // 1. proposals
    AcceptIn in{};
    if (t < m) {
      const Proposal P = jump_proposal(e, s, st, ch, it, t, a);
      mh_record(e, s, ch, t, P, a.ptx, a.pty, a.ptz);
      float4 h1, l1;
      pack1(P.x, P.y, P.z, 0.0, g.d, h1, l1);
      tgt[0][t] = h1;
      tgt[1][t] = l1;
      in = accept_inputs(e, s, ch, a.rlr + (size_t)(it - 1) * a.C, a.ptx, a.pty, a.ptz, t);
    }
    __syncthreads();
    // 2. pair terms
    float4 h1 = make_float4(0, 0, 0, 0), l1 = make_float4(0, 0, 0, 0);
    if (act) {
      h1 = tgt[0][col];
      l1 = tgt[1][col];
    }
    const bool surf = wave_on_surface(ic, act, h1, l1);
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int jb = 0; jb < n; jb += kSoloStage) {
      if (!resident) {
        __syncthreads();
        for (int j = t; j < kSoloStage && jb + j < n; j += kSoloThreads) {
          shi[j] = hi[jb + j];
          slo[j] = lo[jb + j];
        }
        __syncthreads();
      }
      const int cnt = min(kSoloStage, n - jb);
      if (!act) continue;
      if (surf) {  // fz holds the sum without the factor -2 until the end
        if (nic == 0) {
          for (int j0 = row; j0 < cnt; j0 += rows * kSurfBlock) {
            const int je = min(cnt, j0 + rows * kSurfBlock);
            float acc = 0.f;
            for (int jj = j0; jj < je; jj += rows)
              acc = surface_acc32(h1, l1, shi[jj], slo[jj], acc);
            fz += (double)acc;
          }
        } else {
          for (int jj = row; jj < cnt; jj += rows)
            fz += (double)surface_term32<false>(h1, l1, shi[jj], slo[jj], nic, dnm);
        }
      } else if (ic) {
        for (int jj = row; jj < cnt; jj += rows)
          mixed_pair(h1, l1, shi[jj], slo[jj], 1, nic, two_d, fx, fy, fz);
      } else {
        for (int jj = row; jj < cnt; jj += rows)
          mixed_pair(h1, l1, shi[jj], slo[jj], 0, 0, two_d, fx, fy, fz);
      }
    }
    if (surf) fz = -2.0 * fz;
    red[0][t] = fx;
    red[1][t] = fy;
    red[2][t] = fz;
    // rows added pairwise: row r takes row r + h for h = rows/2, rows/4, ..., 1
    for (int h = rows >> 1; h > 0; h >>= 1) {
      __syncthreads();
      if (row < h) {
        const int u = t + (h << lmp);
        red[0][t] += red[0][u];
        red[1][t] += red[1][u];
        red[2][t] += red[2][u];
      }
    }
    __syncthreads();
    // 3. field and accept (lane t < m holds chain t: row 0)
    double fxk = 0.0, fyk = 0.0, fzk = 0.0;
    if (t < m) {
      field_from_sums32(g, red[0][t], red[1][t], red[2][t], fxk, fyk, fzk);
      const int p = e.pbase + t;
      a.fx[p] = fxk;
      a.fy[p] = fyk;
      a.fz[p] = fzk;
    }
    mh_accept_block(e, s, st, wf, ch, it, in, fxk, fyk, fzk, ra, rr, t < m ? t : -1, nullptr);
    // the width and the chains' states are read by the next jump's proposals
    __syncthreads();
  }
}

// Per-chain width updates (thermal-field, tip) in chain order and the escape
// quantity of every chain.
__global__ __launch_bounds__(256) void k_mh_finish(const Ent* tab, const SimDesc* desc,
                                                   SimState* st, const double* wf, Chains ch) {
  constexpr int kTile = 1024;
  __shared__ double fac[kTile], rate[kTile];
  const Ent e = tab[blockIdx.x];
  const SimDesc& s = desc[e.slot];
  const EmitConsts ec = emit_consts();
  if (s.mode != RD_EMIT_PLANAR_FE) {
    // The width factor of every chain depends on its own counts only, so the
    // exponentials are formed in parallel; the products and clamps, which
    // chain the chains together, follow in chain order in one lane.
    SimState& S = st[e.slot];
    const double lo = (s.mode == RD_EMIT_TIP_FE) ? 0.0005 : 0.005;
    double mh = S.MH_std, ar = S.a_rate;
    for (int b = 0; b < e.nch; b += kTile) {
      const int cnt = min(kTile, e.nch - b);
      __syncthreads();
      for (int k = threadIdx.x; k < cnt; k += blockDim.x) {
        const int c = e.cbase + b + k;
        const int a = ch.acc[c], r = ch.rej[c];
        if (ch.ok[c] == 1 && a + r > 0) {
          const double ak = (double)a / (double)(r + a);
          rate[k] = ak;
          fac[k] = rd_exp(0.025 * (ak - 0.35));
        } else {
          rate[k] = -1.0;  // no update
        }
      }
      __syncthreads();
      if (threadIdx.x == 0)
        for (int k0 = 0; k0 < cnt; k0 += 8) {
          // eight loads issued together, updates in chain order
          double rk[8], fk[8];
          const int nk = min(8, cnt - k0);
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            rk[j] = j < nk ? rate[k0 + j] : -1.0;
            fk[j] = j < nk ? fac[k0 + j] : 1.0;
          }
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            if (rk[j] < 0.0) continue;
            ar = rk[j];
            // the clamp to [lo, 0.125] (lo < 0.125) as max then min: the same
            // values as the two comparisons, with a shorter dependent chain
            mh = __builtin_fmin(__builtin_fmax(mh * fk[j], lo), 0.1250);
          }
        }
    }
    if (threadIdx.x == 0) {
      S.MH_std = mh;
      S.a_rate = ar;
    }
  }
  for (int k = threadIdx.x; k < e.nch; k += blockDim.x) {
    const int c = e.cbase + k;
    double D;
    if (s.mode == RD_EMIT_PLANAR_FE) {
      if (ch.ok[c] == 1) {
        int sec;
        double w = work_at(s, wf, ch.cx[c], ch.cy[c], sec);
        D = fn_escape_log(ec, s.g.ic, ch.F[c], w);
      } else {
        D = -kHuge;
      }
    } else if (s.mode == RD_EMIT_TIP_FE) {
      // Escape_Prob_Tip: exp(b_FN w^(3/2) v_y / |F|); failed chains 0
      if (ch.ok[c] == 1) {
        double sw = __builtin_sqrt(s.w_tip);
        D = rd_exp(ec.b_FN * (sw * sw * sw) * fn_v(s.g.ic, ec.l_const, ch.F[c], s.w_tip) /
                   __builtin_fabs(ch.F[c]));
      } else {
        D = 0.0;
      }
    } else {
      D = 0.0;
    }
    ch.D[c] = D;
  }
}

// Escape test and append (Add_Particle for the emitted electrons, in chain
// order).  One block of 256 per simulation.
__global__ __launch_bounds__(256) void k_emit(const Ent* tab, const SimDesc* desc, SimState* st,
                                              const double* wf, Chains ch, double* X, double* Y,
                                              double* Z, double* PX, double* PY, double* PZ,
                                              double* VX, double* VY, double* VZ, double* AX,
                                              double* AY, double* AZ, double* A1X, double* A1Y,
                                              double* A1Z, double* A2X, double* A2Y, double* A2Z,
                                              double* Q, double* M, int32_t* SP, int32_t* STEP,
                                              int32_t* EMIT, int32_t* SEC, int32_t* ID,
                                              int32_t* ALIVE, rd_emit_record* elog) {
  __shared__ int scan[256];
  __shared__ double dsum[256];
  __shared__ int s_off, s_drop;
  const Ent e = tab[blockIdx.x];
  const SimDesc& s = desc[e.slot];
  SimState& S = st[e.slot];
  const EmitConsts ec = emit_consts();
  const int n0 = S.n, id0 = S.nrID;
  const double qe = -1.0 * ec.q0, me = 1.0 * ec.m0;
  if (threadIdx.x == 0) {
    s_off = 0;
    s_drop = 0;
  }
  double dacc = 0.0;
  __syncthreads();
  for (int kb = 0; kb < e.nch; kb += 256) {
    const int k = kb + threadIdx.x;
    int flag = 0;
    double x = 0, y = 0, z = 0, vx = 0, vy = 0, vz = 0;
    int sec = 1;
    if (k < e.nch) {
      const int c = e.cbase + k;
      if (s.mode == RD_EMIT_PLANAR_FE) {
        double Df = ch.D[c];
        if (ch.F[c] >= 0.0) Df = -kHuge;
        dacc = dacc + rd_exp(Df);
        double rnd = uniform(s.seed, s.stream, (uint64_t)e.step, tag(kTagEmit, 0), (uint64_t)k, 0);
        if (rd_log(rnd) <= Df) {
          flag = 1;
          x = ch.cx[c];
          y = ch.cy[c];
          z = 1.0 * Consts::length_scale;
          work_at(s, wf, x, y, sec);
        }
      } else if (s.mode == RD_EMIT_PLANAR_GTF) {
        if (ch.ok[c] == 1) {
          flag = 1;
          x = ch.cx[c];
          y = ch.cy[c];
          z = 1.0 * Consts::length_scale;
          // Get_MB_Velocity (src/mod_velocity.f90:53-68)
          double sd = __builtin_sqrt(ec.k_b * s.T / ec.m0);
          double b1, b2, b3, b4;
          box_muller(s.seed, s.stream, (uint64_t)e.step, tag(kTagVel, 0), (uint64_t)k, 0.0, 0.0,
                     sd, sd, b1, b2);
          box_muller(s.seed, s.stream, (uint64_t)e.step, tag(kTagVel, 1), (uint64_t)k, 0.0, 0.0,
                     sd, sd, b3, b4);
          vx = b1;
          vy = b3;
          vz = __builtin_fabs(b4);
          work_at(s, wf, x, y, sec);
        }
      } else {
        double rnd = uniform(s.seed, s.stream, (uint64_t)e.step, tag(kTagEmit, 0), (uint64_t)k, 0);
        if ((ch.F[c] < 0.0) && (rnd <= ch.D[c])) {
          flag = 1;
          double nx, ny, nz;
          tip_normal(tip_surf(s.g), ch.ox[c], ch.oy[c], nx, ny, nz);
          x = ch.ox[c] + nx * Consts::length_scale;
          y = ch.oy[c] + ny * Consts::length_scale;
          z = ch.oz[c] + nz * Consts::length_scale;
          sec = 1;
        }
      }
    }
    scan[threadIdx.x] = flag;
    __syncthreads();
    // inclusive Hillis-Steele scan
    for (int w = 1; w < 256; w <<= 1) {
      int v = (int)threadIdx.x >= w ? scan[threadIdx.x - w] : 0;
      __syncthreads();
      scan[threadIdx.x] += v;
      __syncthreads();
    }
    const int total = scan[255];
    const int rank = scan[threadIdx.x] - flag;
    if (flag) {
      const int slot_idx = n0 + s_off + rank;  // index within the simulation
      if (slot_idx >= s.cap) {
        atomicAdd(&s_drop, 1);
      } else {
        const int gi = s.base + slot_idx;
        X[gi] = x;
        Y[gi] = y;
        Z[gi] = z;
        PX[gi] = -1.0 * Consts::length_scale;
        PY[gi] = -1.0 * Consts::length_scale;
        PZ[gi] = -1.0 * Consts::length_scale;
        VX[gi] = vx;
        VY[gi] = vy;
        VZ[gi] = vz;
        double ex, ey, ez;
        if (s.g.kind == RD_GEOM_TIP) {
          // field_E_Hyperboloid at the emission point
          double af = s.g.a_foci, sz = s.g.shift_z;
          double r_p = __builtin_sqrt(x * x + y * y + (z + af - sz) * (z + af - sz));
          double r_m = __builtin_sqrt(x * x + y * y + (z - af - sz) * (z - af - sz));
          double xi = (r_p + r_m) / (2.0 * af);
          double eta = (r_p - r_m) / (2.0 * af);
          double cc, ss;
          if ((__builtin_fabs(x) < 1.0e-18) && (__builtin_fabs(y) < 1.0e-18)) {
            cc = 1.0;
            ss = 0.0;
          } else {
            double rho = __builtin_sqrt(x * x + y * y);
            cc = x / rho;
            ss = y / rho;
          }
          double pre = s.g.pre_E * 1.0 / (xi * xi - eta * eta);
          double fac = (__builtin_fabs(xi - 1.0) < 1.0e-6)
                           ? 0.0
                           : eta * __builtin_sqrt((xi * xi - 1.0) / (1.0 - eta * eta));
          ex = -1.0 * pre * fac * cc;
          ey = -1.0 * pre * fac * ss;
          ez = pre * xi;
        } else {
          ex = 0.0;
          ey = 0.0;
          ez = s.g.Ez;
        }
        const double qm = qe / me;
        AX[gi] = A1X[gi] = A2X[gi] = qm * ex;
        AY[gi] = A1Y[gi] = A2Y[gi] = qm * ey;
        AZ[gi] = A1Z[gi] = A2Z[gi] = qm * ez;
        Q[gi] = qe;
        M[gi] = me;
        SP[gi] = RD_SPECIES_ELEC;
        STEP[gi] = e.step;
        EMIT[gi] = 1;
        SEC[gi] = sec;
        ID[gi] = id0 + s_off + rank;
        ALIVE[gi] = 1;
        if (elog) {
          rd_emit_record r;
          r.x = x;
          r.y = y;
          r.z = z;
          r.emitter = 1;
          r.section = sec;
          r.id = id0 + s_off + rank;
          r.species = RD_SPECIES_ELEC;
          elog[s.base + (slot_idx - n0)] = r;
        }
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) s_off += total;
    __syncthreads();
  }
  dsum[threadIdx.x] = dacc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) dsum[threadIdx.x] = dsum[threadIdx.x] + dsum[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int added = s_off - s_drop;
    S.n = n0 + added;
    S.nrID = id0 + added;
    S.emitted = added;
    S.dropped += s_drop;
    if (s.mode == RD_EMIT_PLANAR_FE) S.df_avg = dsum[0] / S.N_sup;
  }
}

// ---- Add_Particle from the host ------------------------------------------
__global__ void k_add(int slot, const SimDesc* desc, SimState* st, int n, const double* __restrict__ in,
                      int step, int emit, double* X, double* Y, double* Z, double* PX, double* PY,
                      double* PZ, double* VX, double* VY, double* VZ, double* AX, double* AY,
                      double* AZ, double* A1X, double* A1Y, double* A1Z, double* A2X, double* A2Y,
                      double* A2Z, double* Q, double* M, int32_t* SP, int32_t* STEP, int32_t* EMIT,
                      int32_t* SEC, int32_t* ID, int32_t* ALIVE) {
  // in: n rows of (x, y, z, vx, vy, vz, species, section)
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  const SimDesc& s = desc[slot];
  const SimState& S = st[slot];
  if (k >= n) return;
  const int idx = S.n + k;
  if (idx >= s.cap) return;
  const int gi = s.base + idx;
  const double* r = in + 8 * (size_t)k;
  double x = r[0], y = r[1], z = r[2];
  int spc = (int)r[6];
  double q = (spc == RD_SPECIES_ION) ? 1.0 * Consts::q_0 : -1.0 * Consts::q_0;
  double m = (spc == RD_SPECIES_ION) ? (28.0134 * Consts::m_u - Consts::m_0) : 1.0 * Consts::m_0;
  X[gi] = x;
  Y[gi] = y;
  Z[gi] = z;
  PX[gi] = -1.0 * Consts::length_scale;
  PY[gi] = -1.0 * Consts::length_scale;
  PZ[gi] = -1.0 * Consts::length_scale;
  VX[gi] = r[3];
  VY[gi] = r[4];
  VZ[gi] = r[5];
  double ex, ey, ez;
  if (s.g.kind == RD_GEOM_TIP) {
    double af = s.g.a_foci, sz = s.g.shift_z;
    double r_p = __builtin_sqrt(x * x + y * y + (z + af - sz) * (z + af - sz));
    double r_m = __builtin_sqrt(x * x + y * y + (z - af - sz) * (z - af - sz));
    double xi = (r_p + r_m) / (2.0 * af);
    double eta = (r_p - r_m) / (2.0 * af);
    double cc, ss;
    if ((__builtin_fabs(x) < 1.0e-18) && (__builtin_fabs(y) < 1.0e-18)) {
      cc = 1.0;
      ss = 0.0;
    } else {
      double rho = __builtin_sqrt(x * x + y * y);
      cc = x / rho;
      ss = y / rho;
    }
    double pre = s.g.pre_E * 1.0 / (xi * xi - eta * eta);
    double fac = (__builtin_fabs(xi - 1.0) < 1.0e-6)
                     ? 0.0
                     : eta * __builtin_sqrt((xi * xi - 1.0) / (1.0 - eta * eta));
    ex = -1.0 * pre * fac * cc;
    ey = -1.0 * pre * fac * ss;
    ez = pre * xi;
  } else {
    ex = 0.0;
    ey = 0.0;
    ez = s.g.Ez;
  }
  const double qm = q / m;
  AX[gi] = A1X[gi] = A2X[gi] = qm * ex;
  AY[gi] = A1Y[gi] = A2Y[gi] = qm * ey;
  AZ[gi] = A1Z[gi] = A2Z[gi] = qm * ez;
  Q[gi] = q;
  M[gi] = m;
  SP[gi] = spc;
  STEP[gi] = step;
  EMIT[gi] = emit;
  SEC[gi] = (int)r[7];
  ID[gi] = S.nrID + k;
  ALIVE[gi] = 1;
}

__global__ void k_add_commit(int slot, const SimDesc* desc, SimState* st, int n) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  SimState& S = st[slot];
  int room = desc[slot].cap - S.n;
  int added = n < room ? n : room;
  S.n += added;
  S.nrID += added;
  S.dropped += n - added;
}

// ---- Beeman position update + boundary marking ---------------------------
__device__ __forceinline__ void mark_remove(SimState& S, int32_t* ALIVE, double* Q,
                                            const int32_t* SP, const int32_t* EMIT, int gi,
                                            int li, int reason, double* mark, double vx,
                                            double vy, double vz) {
  if (!ALIVE[gi]) return;
  ALIVE[gi] = 0;
  Q[gi] = 0.0;
  atomicMin(&S.first_dead, li);
  if (mark) {  // the velocity at marking, for the absorption record
    mark[4 * (size_t)gi + 0] = vx;
    mark[4 * (size_t)gi + 1] = vy;
    mark[4 * (size_t)gi + 2] = vz;
    mark[4 * (size_t)gi + 3] = reason;
  }
  atomicAdd(&S.rem_total, 1);
  const int spc = SP[gi];
  if (spc == RD_SPECIES_ELEC) {
    atomicAdd(&S.rem_elec, 1);
    if (reason == 1) {
      atomicAdd(&S.rem_top, 1);
      if (EMIT[gi] == 1) atomicAdd(&S.rem_top_emit, 1);
    } else {
      atomicAdd(&S.rem_bot, 1);
    }
  } else {
    atomicAdd(&S.rem_ion, 1);
    if (reason == 1)
      atomicAdd(&S.rem_top, 1);
    else
      atomicAdd(&S.rem_bot, 1);
  }
}

__global__ void k_position(const Ent* tab, const SimDesc* desc, SimState* st, double* X,
                           double* Y, double* Z, double* PX, double* PY, double* PZ,
                           const double* VX, const double* VY, const double* VZ, double* AX,
                           double* AY, double* AZ, double* A1X, double* A1Y, double* A1Z,
                           double* A2X, double* A2Y, double* A2Z, double* Q, const int32_t* SP,
                           const int32_t* EMIT, int32_t* ALIVE, double* mark) {
  const Ent e = tab[blockIdx.z];
  SimState& S = st[e.slot];
  const int n = S.n;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const SimDesc& s = desc[e.slot];
  const int gi = e.base + i;
  // two-time-step mode (Update_Species_Position, src/mod_verlet.F90:234-290):
  // an ion moves only at steps that are multiples of K, over K time steps;
  // between them nothing of it changes and no boundary is checked
  const bool ion_k = s.tts > 0 && SP[gi] == RD_SPECIES_ION;
  if (ion_k && (e.step % s.tts) != 0) return;
  const double ks = ion_k ? (double)s.tts : 1.0;
  const double dt = s.dt, dt2 = s.dt2;
  const double c16 = 1.0 / 6.0;
  double x = X[gi], y = Y[gi], z = Z[gi];
  const double zprev = z;
  PX[gi] = x;
  PY[gi] = y;
  PZ[gi] = z;
  double ax = AX[gi], ay = AY[gi], az = AZ[gi];
  double a1x = A1X[gi], a1y = A1Y[gi], a1z = A1Z[gi];
  x = x + VX[gi] * dt * ks + c16 * (4.0 * ax - a1x) * dt2 * (ks * ks);
  y = y + VY[gi] * dt * ks + c16 * (4.0 * ay - a1y) * dt2 * (ks * ks);
  z = z + VZ[gi] * dt * ks + c16 * (4.0 * az - a1z) * dt2 * (ks * ks);
  X[gi] = x;
  Y[gi] = y;
  Z[gi] = z;
  A2X[gi] = a1x;
  A2Y[gi] = a1y;
  A2Z[gi] = a1z;
  A1X[gi] = ax;
  A1Y[gi] = ay;
  A1Z[gi] = az;
  AX[gi] = 0.0;
  AY[gi] = 0.0;
  AZ[gi] = 0.0;
  // ptr_Check_Boundary
  const double vx = VX[gi], vy = VY[gi], vz = VZ[gi];
  if (mark) {  // Check_Planes: the velocity of a crossing, for its record
    bool cross = false;
    for (int p = 0; p < s.nplanes; ++p) {
      const double zp = s.planes[p];
      cross |= zp > 0.0 && z > zp && zprev < zp;
    }
    if (cross) {
      mark[4 * (size_t)gi + 0] = vx;
      mark[4 * (size_t)gi + 1] = vy;
      mark[4 * (size_t)gi + 2] = vz;
    }
  }
  if (z < 0.0)
    mark_remove(S, ALIVE, Q, SP, EMIT, gi, i, 2, mark, vx, vy, vz);
  else if (z > s.g.box_top)
    mark_remove(S, ALIVE, Q, SP, EMIT, gi, i, 1, mark, vx, vy, vz);
  if (s.g.kind == RD_GEOM_TIP) {
    double eta = tip_eta(tip_surf(s.g), x, y, z);
    if (eta < s.g.eta_1) mark_remove(S, ALIVE, Q, SP, EMIT, gi, i, 2, mark, vx, vy, vz);
  }
}

// ---- Beeman velocity update + Ramo current --------------------------------
// Per block: fixed-order tree sums of 13 values (Ramo per species 3, velocity
// sums 9, electron count 1) into red[entry][block][16].
__global__ __launch_bounds__(256) void k_velocity(const Ent* tab, const SimDesc* desc,
                                                  const SimState* st, const double* X,
                                                  const double* Y, const double* Z, double* VX,
                                                  double* VY, double* VZ, const double* AX,
                                                  const double* AY, const double* AZ,
                                                  const double* A1X, const double* A1Y,
                                                  const double* A1Z, const double* A2X,
                                                  const double* A2Y, const double* A2Z,
                                                  const double* Q, const int32_t* SP,
                                                  double* red) {
  __shared__ double sm[13][256];
  const Ent e = tab[blockIdx.z];
  const int n = st[e.slot].n;
  if ((int)(blockIdx.x * blockDim.x) >= max(n, 1) && blockIdx.x > 0) return;
  const SimDesc& s = desc[e.slot];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v[13];
  for (int k = 0; k < 13; ++k) v[k] = 0.0;
  // two-time-step mode (Update_Species_Velocity, src/mod_verlet.F90:511-580):
  // ions move only at steps that are multiples of K, over K time steps
  const bool ion_k = s.tts > 0 && i < n && SP[e.base + i] == RD_SPECIES_ION;
  if (i < n && !(ion_k && (e.step % s.tts) != 0)) {
    const int gi = e.base + i;
    const double dt = s.dt;
    const double c16 = 1.0 / 6.0;
    const double ks = ion_k ? (double)s.tts : 1.0;
    double vx = VX[gi] + c16 * (2.0 * AX[gi] + 5.0 * A1X[gi] - A2X[gi]) * dt * ks;
    double vy = VY[gi] + c16 * (2.0 * AY[gi] + 5.0 * A1Y[gi] - A2Y[gi]) * dt * ks;
    double vz = VZ[gi] + c16 * (2.0 * AZ[gi] + 5.0 * A1Z[gi] - A2Z[gi]) * dt * ks;
    VX[gi] = vx;
    VY[gi] = vy;
    VZ[gi] = vz;
    double ux, uy, uz;
    if (s.g.kind == RD_GEOM_TIP) {
      double ex, ey, ez;
      double x = X[gi], y = Y[gi], z = Z[gi];
      double af = s.g.a_foci, sz = s.g.shift_z;
      double r_p = __builtin_sqrt(x * x + y * y + (z + af - sz) * (z + af - sz));
      double r_m = __builtin_sqrt(x * x + y * y + (z - af - sz) * (z - af - sz));
      double xi = (r_p + r_m) / (2.0 * af);
      double eta = (r_p - r_m) / (2.0 * af);
      double cc, ss;
      if ((__builtin_fabs(x) < 1.0e-18) && (__builtin_fabs(y) < 1.0e-18)) {
        cc = 1.0;
        ss = 0.0;
      } else {
        double rho = __builtin_sqrt(x * x + y * y);
        cc = x / rho;
        ss = y / rho;
      }
      double pre = s.g.pre_E * 1.0 / (xi * xi - eta * eta);
      double fac = (__builtin_fabs(xi - 1.0) < 1.0e-6)
                       ? 0.0
                       : eta * __builtin_sqrt((xi * xi - 1.0) / (1.0 - eta * eta));
      ex = -1.0 * pre * fac * cc;
      ey = -1.0 * pre * fac * ss;
      ez = pre * xi;
      ux = ex * s.g.pre_E_unit / s.g.pre_E;
      uy = ey * s.g.pre_E_unit / s.g.pre_E;
      uz = ez * s.g.pre_E_unit / s.g.pre_E;
    } else {
      ux = 0.0;
      uy = 0.0;
      uz = -1.0 / s.g.d;
    }
    double EzV = vx * ux + vy * uy + vz * uz;
    const int spc = SP[gi];
    const double cur = Q[gi] * EzV;
    if (spc >= 1 && spc <= 3) v[spc - 1] = cur;
    v[3] = vx;
    v[4] = vy;
    v[5] = vz;
    if (spc == RD_SPECIES_ELEC) {
      v[6] = vx;
      v[7] = vy;
      v[8] = vz;
      v[12] = 1.0;
    } else if (spc == RD_SPECIES_ION) {
      v[9] = vx;
      v[10] = vy;
      v[11] = vz;
    }
  }
  for (int k = 0; k < 13; ++k) sm[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 13; ++k) sm[k][threadIdx.x] = sm[k][threadIdx.x] + sm[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 13)
    red[((size_t)blockIdx.z * e.nblk + blockIdx.x) * 16 + threadIdx.x] = sm[threadIdx.x][0];
}

// Block sums of k_velocity added in block order: lane k < 13 takes sum k,
// with its loads issued eight blocks at a time.
__global__ void k_velocity_finish(const Ent* tab, SimState* st, const double* red) {
  const Ent e = tab[blockIdx.x];
  const int k = threadIdx.x;
  if (k >= 13) return;
  SimState& S = st[e.slot];
  const int nb = min(max(1, (S.n + 255) / 256), e.nblk);
  const double* r = red + (size_t)blockIdx.x * e.nblk * 16 + k;
  double acc = 0.0;
  int b = 0;
  for (; b + 8 <= nb; b += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = r[(size_t)(b + u) * 16];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc = acc + v[u];
  }
  for (; b < nb; ++b) acc = acc + r[(size_t)b * 16];
  if (k < 3)
    S.ramo[k] = acc;
  else if (k < 12)
    S.vsum[k - 3] = acc;
  else
    S.n_elec = (int)acc;  // electrons updated this step
}

// ---- order-preserving compaction (one block of 1024 per simulation) ------
// Chunks of 1024 particles in index order; survivors move down to the running
// offset, keeping their order (Remove_Particles, src/mod_pair.F90:352-562).
__global__ __launch_bounds__(1024) void k_compact(
    const Ent* tab, const SimDesc* desc, SimState* st, int32_t* life, double* X, double* Y,
    double* Z, double* PX, double* PY, double* PZ, double* VX, double* VY, double* VZ, double* AX,
    double* AY, double* AZ, double* A1X, double* A1Y, double* A1Z, double* A2X, double* A2Y,
    double* A2Z, double* Q, double* M, int32_t* SP, int32_t* STEP, int32_t* EMIT, int32_t* SEC,
    int32_t* ID, int32_t* ALIVE, const double* mark, rd_absorb_record* alog) {
  __shared__ int wsum[1024 / 64], rsum[1024 / 64];
  __shared__ int s_off, s_roff;
  const Ent e = tab[blockIdx.x];
  SimState& S = st[e.slot];
  if (S.rem_total == 0) return;
  const int n = S.n;
  const int base = e.base;
  // everything before the chunk of the first marked particle stays in place
  // (First_Dead_Index, src/mod_pair.F90:570-582)
  const int start = (min(S.first_dead, n) / 1024) * 1024;
  if (threadIdx.x == 0) {
    s_off = start;
    s_roff = 0;
  }
  __syncthreads();
  for (int cb = start; cb < n; cb += 1024) {
    const int i = cb + threadIdx.x;
    const int flag = (i < n && ALIVE[base + i]) ? 1 : 0;
    // rank among the chunk's survivors: a ballot per wavefront, then the
    // survivors of the lower wavefronts
    const unsigned long long mask = __ballot(flag);
    const bool gone = i < n && !flag;  // removed: an absorption record
    const unsigned long long gmask = __ballot(gone);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
      wsum[wv] = __popcll(mask);
      rsum[wv] = __popcll(gmask);
    }
    __syncthreads();
    const unsigned long long below = (1ull << lane) - 1ull;
    int rank = __popcll(mask & below), total = 0;
    int rrank = __popcll(gmask & below), rtotal = 0;
    for (int u = 0; u < 1024 / 64; ++u) {
      if (u < wv) {
        rank += wsum[u];
        rrank += rsum[u];
      }
      total += wsum[u];
      rtotal += rsum[u];
    }
    const int off = s_off, roff = s_roff;
    // Up to the first removed particle nothing moves (First_Dead_Index,
    // src/mod_pair.F90:570-582): such a chunk is only counted.  Uniform over
    // the workgroup, like the barriers inside.
    const bool moves = !(off == cb && total == min(1024, n - cb));
    double r[21];
    int ri[5];
    if (moves && flag) {
      const int gi = base + i;
      r[0] = X[gi]; r[1] = Y[gi]; r[2] = Z[gi];
      r[3] = PX[gi]; r[4] = PY[gi]; r[5] = PZ[gi];
      r[6] = VX[gi]; r[7] = VY[gi]; r[8] = VZ[gi];
      r[9] = AX[gi]; r[10] = AY[gi]; r[11] = AZ[gi];
      r[12] = A1X[gi]; r[13] = A1Y[gi]; r[14] = A1Z[gi];
      r[15] = A2X[gi]; r[16] = A2Y[gi]; r[17] = A2Z[gi];
      r[18] = Q[gi]; r[19] = M[gi];
      ri[0] = SP[gi]; ri[1] = STEP[gi]; ri[2] = EMIT[gi]; ri[3] = SEC[gi]; ri[4] = ID[gi];
    }
    if (gone) {
      // record_lifetime (src/mod_pair.F90:1135-1160)
      const int gi = base + i;
      int lt = e.step - STEP[gi];
      if (lt <= 0) lt = 1;
      if (lt > kLife) lt = kLife;
      int spc = SP[gi];
      if (spc >= 1 && spc <= 2) atomicAdd(&life[((size_t)e.slot * 2 + (spc - 1)) * kLife + (lt - 1)], 1);
    }
    if (gone && alog) {  // in index order, before any particle moves
      const int gi = base + i;
      rd_absorb_record a;
      a.x = X[gi];
      a.y = Y[gi];
      a.z = Z[gi];
      a.vx = mark[4 * (size_t)gi + 0];
      a.vy = mark[4 * (size_t)gi + 1];
      a.vz = mark[4 * (size_t)gi + 2];
      a.emitter = EMIT[gi];
      a.section = SEC[gi];
      a.id = ID[gi];
      a.where = (int)mark[4 * (size_t)gi + 3];
      alog[base + roff + rrank] = a;
    }
    __syncthreads();  // every read of this chunk is done before any write
    if (moves && flag) {
      const int d = base + off + rank;
      X[d] = r[0]; Y[d] = r[1]; Z[d] = r[2];
      PX[d] = r[3]; PY[d] = r[4]; PZ[d] = r[5];
      VX[d] = r[6]; VY[d] = r[7]; VZ[d] = r[8];
      AX[d] = r[9]; AY[d] = r[10]; AZ[d] = r[11];
      A1X[d] = r[12]; A1Y[d] = r[13]; A1Z[d] = r[14];
      A2X[d] = r[15]; A2Y[d] = r[16]; A2Z[d] = r[17];
      Q[d] = r[18]; M[d] = r[19];
      SP[d] = ri[0]; STEP[d] = ri[1]; EMIT[d] = ri[2]; SEC[d] = ri[3]; ID[d] = ri[4];
      ALIVE[d] = 1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      s_off = off + total;
      s_roff = roff + rtotal;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) S.n = s_off;
}

// ------------------------------------------------------------------------
// Compaction of a large simulation by chunks of 1024 particles, in three
// launches with a workgroup per chunk (k_compact takes one workgroup per
// simulation and walks its chunks in turn: about 5 us per chunk).  The result
// is the same data movement: survivors in index order, absorption records in
// index order, the same lifetime counts.
//   k_compact_count: each chunk from the one holding the first marked
//     particle counts its survivors and its removed particles;
//   k_compact_move:  each chunk adds the counts of the chunks before it (its
//     destination and record offsets), writes its absorption records and
//     lifetimes, and stages its survivors in the staging arrays;
//   k_compact_back:  the staged survivors return to the particle arrays from
//     the first chunk's start on, and the new count is stored.
// The staging keeps the moves free of overlaps between workgroups.  Header
// slot (c_off): the particle count before compaction, written by pass 1 and
// read by pass 3 (which stores the new count into the state).
constexpr int kCompactSplit = 8 * 1024;  // particles from which a simulation is compacted by chunks

struct CompactStage {
  int2* cnt;        // headers and chunk counts (survivors, removed)
  double* d;        // 20 staged double arrays of ts particles each
  int32_t* i;       // 5 staged int arrays
  long long ts;     // staged particles per array
};

// Survivors and removed particles of the block's chunk, and the ranks of this
// thread's particle among each (every thread of the block calls it).
__device__ __forceinline__ void chunk_ranks(int flag, int gone, int& rank, int& total, int& rrank,
                                            int& rtotal) {
  __shared__ int wsum[1024 / 64], rsum[1024 / 64];
  const unsigned long long mask = __ballot(flag), gmask = __ballot(gone);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    wsum[wv] = __popcll(mask);
    rsum[wv] = __popcll(gmask);
  }
  __syncthreads();
  const unsigned long long below = (1ull << lane) - 1ull;
  rank = __popcll(mask & below);
  rrank = __popcll(gmask & below);
  total = 0;
  rtotal = 0;
  for (int u = 0; u < 1024 / 64; ++u) {
    if (u < wv) {
      rank += wsum[u];
      rrank += rsum[u];
    }
    total += wsum[u];
    rtotal += rsum[u];
  }
  __syncthreads();
}

// Sum of the (survivors, removed) counts of chunks [c0, c1) over the block.
__device__ __forceinline__ int2 chunk_prefix(const int2* cnt, int c0, int c1) {
  __shared__ int pa[1024 / 64], pr[1024 / 64];
  int a = 0, r = 0;
  for (int c = c0 + (int)threadIdx.x; c < c1; c += blockDim.x) {
    a += cnt[c].x;
    r += cnt[c].y;
  }
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o);
    r += __shfl_xor(r, o);
  }
  if ((threadIdx.x & 63) == 0) {
    pa[threadIdx.x >> 6] = a;
    pr[threadIdx.x >> 6] = r;
  }
  __syncthreads();
  a = 0;
  r = 0;
  for (int u = 0; u < (int)(blockDim.x >> 6); ++u) {
    a += pa[u];
    r += pr[u];
  }
  __syncthreads();
  return make_int2(a, r);
}

__global__ __launch_bounds__(1024) void k_compact_count(const Ent* tab, const SimState* st,
                                                       const int32_t* ALIVE, CompactStage cs) {
  const Ent e = tab[blockIdx.y];
  const SimState& S = st[e.slot];
  const int n = S.n;
  const int c = blockIdx.x, cb = c * 1024;
  if (S.rem_total == 0 || cb >= n) return;
  const int c0 = min(S.first_dead, n) / 1024;
  if (c < c0) return;
  const int i = cb + threadIdx.x;
  const int flag = (i < n && ALIVE[e.base + i]) ? 1 : 0;
  const int gone = (i < n && !flag) ? 1 : 0;
  int rank, total, rrank, rtotal;
  chunk_ranks(flag, gone, rank, total, rrank, rtotal);
  if (threadIdx.x == 0) {
    cs.cnt[e.c_off + 1 + c] = make_int2(total, rtotal);
    if (c == c0) cs.cnt[e.c_off] = make_int2(n, 0);
  }
}

__global__ __launch_bounds__(1024) void k_compact_move(
    const Ent* tab, const SimState* st, int32_t* life, const double* X, const double* Y,
    const double* Z, const double* PX, const double* PY, const double* PZ, const double* VX,
    const double* VY, const double* VZ, const double* AX, const double* AY, const double* AZ,
    const double* A1X, const double* A1Y, const double* A1Z, const double* A2X, const double* A2Y,
    const double* A2Z, const double* Q, const double* M, const int32_t* SP, const int32_t* STEP,
    const int32_t* EMIT, const int32_t* SEC, const int32_t* ID, const int32_t* ALIVE,
    const double* mark, rd_absorb_record* alog, CompactStage cs) {
  const Ent e = tab[blockIdx.y];
  const SimState& S = st[e.slot];
  const int n = S.n;
  const int c = blockIdx.x, cb = c * 1024;
  if (S.rem_total == 0 || cb >= n) return;
  const int c0 = min(S.first_dead, n) / 1024;
  if (c < c0) return;
  const int2* cnt = cs.cnt + e.c_off + 1;
  const int2 pre = chunk_prefix(cnt, c0, c);  // survivors and records before this chunk
  const int i = cb + threadIdx.x;
  const int base = e.base;
  const int flag = (i < n && ALIVE[base + i]) ? 1 : 0;
  const int gone = (i < n && !flag) ? 1 : 0;
  int rank, total, rrank, rtotal;
  chunk_ranks(flag, gone, rank, total, rrank, rtotal);
  const int gi = base + i;
  if (gone) {
    // record_lifetime (src/mod_pair.F90:1135-1160)
    int lt = e.step - STEP[gi];
    if (lt <= 0) lt = 1;
    if (lt > kLife) lt = kLife;
    const int spc = SP[gi];
    if (spc >= 1 && spc <= 2) atomicAdd(&life[((size_t)e.slot * 2 + (spc - 1)) * kLife + (lt - 1)], 1);
    if (alog) {
      rd_absorb_record a;
      a.x = X[gi];
      a.y = Y[gi];
      a.z = Z[gi];
      a.vx = mark[4 * (size_t)gi + 0];
      a.vy = mark[4 * (size_t)gi + 1];
      a.vz = mark[4 * (size_t)gi + 2];
      a.emitter = EMIT[gi];
      a.section = SEC[gi];
      a.id = ID[gi];
      a.where = (int)mark[4 * (size_t)gi + 3];
      alog[base + pre.y + rrank] = a;
    }
  }
  if (flag) {
    const long long k = e.t_off + pre.x + rank;  // staged position
    const long long T = cs.ts;
    const double v[20] = {X[gi],   Y[gi],   Z[gi],   PX[gi],  PY[gi],  PZ[gi],  VX[gi],
                          VY[gi],  VZ[gi],  AX[gi],  AY[gi],  AZ[gi],  A1X[gi], A1Y[gi],
                          A1Z[gi], A2X[gi], A2Y[gi], A2Z[gi], Q[gi],   M[gi]};
#pragma unroll
    for (int a = 0; a < 20; ++a) cs.d[a * T + k] = v[a];
    const int32_t w[5] = {SP[gi], STEP[gi], EMIT[gi], SEC[gi], ID[gi]};
#pragma unroll
    for (int a = 0; a < 5; ++a) cs.i[a * T + k] = w[a];
  }
}

__global__ __launch_bounds__(1024) void k_compact_back(
    const Ent* tab, SimState* st, double* X, double* Y, double* Z, double* PX, double* PY,
    double* PZ, double* VX, double* VY, double* VZ, double* AX, double* AY, double* AZ,
    double* A1X, double* A1Y, double* A1Z, double* A2X, double* A2Y, double* A2Z, double* Q,
    double* M, int32_t* SP, int32_t* STEP, int32_t* EMIT, int32_t* SEC, int32_t* ID,
    int32_t* ALIVE, CompactStage cs) {
  const Ent e = tab[blockIdx.y];
  SimState& S = st[e.slot];
  if (S.rem_total == 0) return;
  const int n = cs.cnt[e.c_off].x;  // the count before compaction (pass 1)
  const int c = blockIdx.x, cb = c * 1024;
  if (cb >= n) return;
  const int c0 = min(S.first_dead, n) / 1024;
  if (c < c0) return;
  const int2 all = chunk_prefix(cs.cnt + e.c_off + 1, c0, (n + 1023) / 1024);
  const int start = c0 * 1024, nn = start + all.x;  // the new count
  if (c == c0 && threadIdx.x == 0) S.n = nn;
  const int d = cb + threadIdx.x;  // destination index
  if (d >= nn) return;
  const long long k = e.t_off + (d - start);
  const long long T = cs.ts;
  const int gd = e.base + d;
  double* dst[20] = {X,  Y,  Z,  PX,  PY,  PZ,  VX,  VY,  VZ,  AX,
                     AY, AZ, A1X, A1Y, A1Z, A2X, A2Y, A2Z, Q, M};
#pragma unroll
  for (int a = 0; a < 20; ++a) dst[a][gd] = cs.d[a * T + k];
  int32_t* dsti[5] = {SP, STEP, EMIT, SEC, ID};
#pragma unroll
  for (int a = 0; a < 5; ++a) dsti[a][gd] = cs.i[a * T + k];
  ALIVE[gd] = 1;
}

// ---- Check_Planes records ------------------------------------------------
// One workgroup per simulation walks its particles in index order after the
// move (removed ones included, as Check_Planes runs before removal) and lists
// each crossing from below, z_prev < z_p < z, per plane in index order: a
// ballot per wavefront and plane ranks the crossings.  Plane p's records of
// the simulation at `base` go to plog[p * cap + base + rank].
__global__ __launch_bounds__(1024) void k_planes(const Ent* tab, const SimDesc* desc,
                                                 SimState* st, const double* X, const double* Y,
                                                 const double* Z, const double* PZ,
                                                 const int32_t* EMIT, const int32_t* SEC,
                                                 const int32_t* ID, const double* mark,
                                                 rd_plane_record* plog, long long cap) {
  __shared__ int wsum[RD_MAX_PLANES][1024 / 64];
  __shared__ int s_off[RD_MAX_PLANES];
  const Ent e = tab[blockIdx.x];
  SimState& S = st[e.slot];
  const SimDesc& s = desc[e.slot];
  const int n = S.n, np = s.nplanes;
  if (threadIdx.x < RD_MAX_PLANES) s_off[threadIdx.x] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int cb = 0; cb < n; cb += 1024) {
    const int i = cb + threadIdx.x;
    const int gi = e.base + i;
    double z = 0.0, zprev = 0.0;
    if (i < n) {
      z = Z[gi];
      zprev = PZ[gi];
    }
    unsigned long long bm[RD_MAX_PLANES];
#pragma unroll
    for (int p = 0; p < RD_MAX_PLANES; ++p) {
      const double zp = p < np ? s.planes[p] : 0.0;
      bm[p] = __ballot(i < n && zp > 0.0 && z > zp && zprev < zp);
      if (lane == 0) wsum[p][wv] = __popcll(bm[p]);
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < RD_MAX_PLANES; ++p) {
      if ((bm[p] >> lane) & 1ull) {
        int rank = s_off[p] + __popcll(bm[p] & below);
        for (int u = 0; u < wv; ++u) rank += wsum[p][u];
        rd_plane_record r;
        r.x = X[gi];
        r.y = Y[gi];
        r.vx = mark[4 * (size_t)gi + 0];
        r.vy = mark[4 * (size_t)gi + 1];
        r.vz = mark[4 * (size_t)gi + 2];
        r.emitter = EMIT[gi];
        r.section = SEC[gi];
        r.id = ID[gi];
        r.plane = p + 1;
        plog[(long long)p * cap + e.base + rank] = r;
      }
    }
    __syncthreads();
    if (threadIdx.x < RD_MAX_PLANES) {
      int t = 0;
      for (int u = 0; u < 1024 / 64; ++u) t += wsum[threadIdx.x][u];
      s_off[threadIdx.x] += t;
    }
    __syncthreads();
  }
  if (threadIdx.x < RD_MAX_PLANES) S.plane_n[threadIdx.x] = s_off[threadIdx.x];
}

// ------------------------------------------------------------------------
// Host orchestration.
#define RD_ARRAYS                                                                          \
  c->x, c->y, c->z, c->px, c->py, c->pz, c->vx, c->vy, c->vz, c->ax, c->ay, c->az, c->a1x, \
      c->a1y, c->a1z, c->a2x, c->a2y, c->a2z, c->q, c->m, c->species, c->pstep, c->emitter, \
      c->section, c->id, c->alive

int sync_state(rd_ctx* c) {
  int hi = 0;
  for (int s = 0; s < c->nslots_max; ++s)
    if (c->slots[s].active) hi = s + 1;
  if (hi > 0)
    RD_HIP(hipMemcpyAsync(c->h_state, c->d_state, sizeof(SimState) * hi, hipMemcpyDeviceToHost,
                          c->stream));
  RD_HIP(hipStreamSynchronize(c->stream));
  c->tab_used = 0;  // every table copy has completed
  for (int s = 0; s < c->nslots_max; ++s)
    if (c->slots[s].active) c->slots[s].n = c->h_state[s].n;
  return RD_OK;
}

int engine_add(rd_ctx* c, int slot, int n, const double* pos, const double* vel,
               const int32_t* species, int step, int emit, const int32_t* sec) {
  if (n <= 0) return RD_OK;
  std::vector<double> rows((size_t)n * 8);
  for (int k = 0; k < n; ++k) {
    double* r = &rows[(size_t)k * 8];
    r[0] = pos[3 * k];
    r[1] = pos[3 * k + 1];
    r[2] = pos[3 * k + 2];
    r[3] = vel ? vel[3 * k] : 0.0;
    r[4] = vel ? vel[3 * k + 1] : 0.0;
    r[5] = vel ? vel[3 * k + 2] : 0.0;
    r[6] = species ? species[k] : RD_SPECIES_ELEC;
    r[7] = sec ? sec[k] : 1;
  }
  int r = ensure(c, c->stage, sizeof(double) * rows.size());
  if (r) return r;
  RD_HIP(hipMemcpyAsync(c->stage.p, rows.data(), sizeof(double) * rows.size(),
                        hipMemcpyHostToDevice, c->stream));
  k_add<<<(n + 255) / 256, 256, 0, c->stream>>>(slot, c->d_desc, c->d_state, n,
                                                 (const double*)c->stage.p, step, emit, RD_ARRAYS);
  k_add_commit<<<1, 1, 0, c->stream>>>(slot, c->d_desc, c->d_state, n);
  RD_HIP(hipGetLastError());
  return sync_state(c);
}

static int max_n_ub(const std::vector<Ent>& ents) {
  int m = 0;
  for (const Ent& e : ents) m = std::max(m, e.n_ub);
  return m;
}

int engine_position(rd_ctx* c, const std::vector<Ent>& ents) {
  if (ents.empty()) return RD_OK;
  int mx = max_n_ub(ents);
  if (mx == 0) return RD_OK;
  Ent* dt;
  int r = upload_tab(c, ents, &dt);
  if (r) return r;
  k_position<<<dim3((mx + 255) / 256, 1, ents.size()), 256, 0, c->stream>>>(
      dt, c->d_desc, c->d_state, c->x, c->y, c->z, c->px, c->py, c->pz, c->vx, c->vy, c->vz,
      c->ax, c->ay, c->az, c->a1x, c->a1y, c->a1z, c->a2x, c->a2y, c->a2z, c->q, c->species,
      c->emitter, c->alive, c->logs ? (double*)c->mark.p : nullptr);
  RD_HIP(hipGetLastError());
  return RD_OK;
}

int engine_velocity(rd_ctx* c, const std::vector<Ent>& in) {
  if (in.empty()) return RD_OK;
  std::vector<Ent> ents = in;
  int mx = std::max(1, max_n_ub(ents));
  int nb = (mx + 255) / 256;
  for (Ent& e : ents) e.nblk = nb;
  int r = ensure(c, c->red, sizeof(double) * 16 * (size_t)nb * ents.size());
  if (r) return r;
  Ent* dt;
  r = upload_tab(c, ents, &dt);
  if (r) return r;
  k_velocity<<<dim3(nb, 1, ents.size()), 256, 0, c->stream>>>(
      dt, c->d_desc, c->d_state, c->x, c->y, c->z, c->vx, c->vy, c->vz, c->ax, c->ay, c->az,
      c->a1x, c->a1y, c->a1z, c->a2x, c->a2y, c->a2z, c->q, c->species, (double*)c->red.p);
  k_velocity_finish<<<ents.size(), 64, 0, c->stream>>>(dt, c->d_state, (const double*)c->red.p);
  RD_HIP(hipGetLastError());
  return RD_OK;
}

int engine_compact(rd_ctx* c, const std::vector<Ent>& ents) {
  if (ents.empty()) return RD_OK;
  Ent* dt;
  int r = upload_tab(c, ents, &dt);
  if (r) return r;
  if (c->logs)  // plane crossings, before removal reorders the particles
    k_planes<<<ents.size(), 1024, 0, c->stream>>>(
        dt, c->d_desc, c->d_state, c->x, c->y, c->z, c->pz, c->emitter, c->section, c->id,
        (const double*)c->mark.p, (rd_plane_record*)c->plog.p, (long long)c->cap);
  // simulations of up to kCompactSplit particles: one workgroup each; larger
  // ones: a workgroup per chunk (the same data movement either way)
  std::vector<Ent> small, large;
  long long cslots = 0, ts = 0;
  int maxch = 0;
  for (Ent e : ents) {
    if (e.n_ub < kCompactSplit) {
      small.push_back(e);
      continue;
    }
    const int nch = (e.n_ub + 1023) / 1024;
    e.c_off = cslots;
    e.t_off = ts;
    cslots += 1 + nch;
    ts += (long long)nch * 1024;
    maxch = std::max(maxch, nch);
    large.push_back(e);
  }
  const double* mark = c->logs ? (const double*)c->mark.p : nullptr;
  rd_absorb_record* alog = c->logs ? (rd_absorb_record*)c->alog.p : nullptr;
  if (!small.empty()) {
    Ent* ds = dt;
    if (!large.empty()) {
      r = upload_tab(c, small, &ds);
      if (r) return r;
    }
    k_compact<<<small.size(), 1024, 0, c->stream>>>(ds, c->d_desc, c->d_state, c->d_life,
                                                     RD_ARRAYS, mark, alog);
  }
  if (!large.empty()) {
    const size_t cnt_bytes = ((sizeof(int2) * (size_t)cslots + 255) / 256) * 256;
    r = ensure(c, c->ctmp, cnt_bytes + (20 * sizeof(double) + 5 * sizeof(int32_t)) * (size_t)ts);
    if (r) return r;
    CompactStage cs;
    cs.cnt = (int2*)c->ctmp.p;
    cs.d = (double*)((char*)c->ctmp.p + cnt_bytes);
    cs.i = (int32_t*)(cs.d + 20 * (size_t)ts);
    cs.ts = ts;
    Ent* dl;
    r = upload_tab(c, large, &dl);
    if (r) return r;
    const dim3 g(maxch, large.size());
    k_compact_count<<<g, 1024, 0, c->stream>>>(dl, c->d_state, c->alive, cs);
    k_compact_move<<<g, 1024, 0, c->stream>>>(dl, c->d_state, c->d_life, RD_ARRAYS, mark, alog,
                                               cs);
    k_compact_back<<<g, 1024, 0, c->stream>>>(dl, c->d_state, c->x, c->y, c->z, c->px, c->py,
                                               c->pz, c->vx, c->vy, c->vz, c->ax, c->ay, c->az,
                                               c->a1x, c->a1y, c->a1z, c->a2x, c->a2y, c->a2z,
                                               c->q, c->m, c->species, c->pstep, c->emitter,
                                               c->section, c->id, c->alive, cs);
  }
  RD_HIP(hipGetLastError());
  return RD_OK;
}

static int reset_counters(rd_ctx* c, const std::vector<Ent>& ents) {
  Ent* dt;
  int r = upload_tab(c, ents, &dt);
  if (r) return r;
  k_reset_counters<<<ents.size(), 64, 0, c->stream>>>(dt, c->d_state);
  RD_HIP(hipGetLastError());
  return RD_OK;
}

// The Metropolis-Hastings jumps 1..maxndim of the chain entries `in` (C
// chains in all: the stride of the point and field arrays): per jump, one
// field launch per geometry present, then the finish-and-accept launch; after
// the last jump, the width of split planar FE simulations (split_width).
//
// In a pool the simulations are dealt into up to kJumpStreams parts of about
// equal work, each with its own launch table on its own stream.  Nothing of
// one part depends on another during the jumps, so the latency-bound accept
// step of one part overlaps the field launches of the others.  Every launch
// still sees each simulation exactly as before, so results do not change.
static int run_jumps(rd_ctx* c, const std::vector<Ent>& in, int C, int precision,
                     const Chains& V, int maxndim) {
  const bool fp64 = precision == RD_PREC_FP64;
  const int per = fp64 ? kBlock : kJumpTargets;  // chains per workgroup
  std::vector<Ent> all;
  for (const Ent& e : in)
    if (e.nch > 0) all.push_back(e);
  if (all.empty()) return RD_OK;
  JumpArgs a;
  std::memset(&a, 0, sizeof(a));  // its bytes form part of the graph key
  double* ptx = (double*)c->pts.p;
  double* f = (double*)c->fld.p;
  a.ptx = ptx;
  a.pty = ptx + C;
  a.ptz = ptx + 2 * (size_t)C;
  a.fx = f;
  a.fy = f + C;
  a.fz = f + 2 * (size_t)C;
  a.X = c->x;
  a.Y = c->y;
  a.Z = c->z;
  a.Q = c->q;
  int r = RD_OK;
  if (!fp64) {
    r = launch_pack(c, all, true);
    if (r) return r;
  }
  // A lone small simulation takes all its jumps in one workgroup
  // (k_mh_solo32); the others, and every simulation of a pool, the two
  // launches per jump below.  In a pool those launches serve every
  // simulation at once, and a solo workgroup would hold a CU (and most of its
  // LDS) for the whole jump phase: measured 7% fewer simulations per hour on
  // the serving stream.
  std::vector<Ent> ents, solo;
  const bool lone = c->nslots_max == 1;
  for (const Ent& e : all) {
    const DevGeom& g = c->slots[e.slot].desc.g;
    if (!fp64 && lone && c->solo_jumps && solo_jumps(g.kind, g.ic, g.nic, e.nch, e.n_ub))
      solo.push_back(e);
    else
      ents.push_back(e);
  }
  a.HI = (const float4*)c->pack_hi.p;
  a.LO = (const float4*)c->pack_lo.p;
  a.SRF = (const float4*)c->pack_srf.p;

  // parts: consecutive runs of the table of about equal work
  struct Half {
    std::vector<Ent> ents;
    int nwg = 0, nsettle = 0;
    bool any_split = false, any_planar = false, any_tip = false;
    Ent* dt = nullptr;
    JumpArgs a;
    hipStream_t stream = nullptr;
  };
  const int nparts = (c->nslots_max > 1 && !c->graphs)
                         ? std::min<int>(kJumpStreams, (int)ents.size())
                         : 1;
  Half hv[kJumpStreams];
  int nh = 0;
  {
    double total = 0.0, acc = 0.0;
    for (const Ent& e : ents) total += (double)e.nch * std::max(e.n_ub, 1);
    size_t k = 0;
    for (int h = 0; h < nparts; ++h) {
      // entries until this part reaches its share; leave one per later part
      const size_t keep = (size_t)(nparts - 1 - h);
      while (k < ents.size() - keep) {
        hv[h].ents.push_back(ents[k]);
        acc += (double)ents[k].nch * std::max(ents[k].n_ub, 1);
        ++k;
        if (h + 1 < nparts && acc >= total * (h + 1) / nparts) break;
      }
      if (!hv[h].ents.empty()) nh = h + 1;
    }
  }
  // A context made for one simulation spreads each jump's sums over up to
  // kWantJump workgroups (latency); in a pool, whose launches serve many
  // simulations, every mixed partial covers at least kPoolJumpChunks chunks
  // of sources (throughput; the FP64 kernel, bound by its arithmetic, keeps
  // the finer split).  Either way the split depends on the simulation's own
  // size and the kind of context only.
  const int min_cps = (c->nslots_max > 1 && !fp64) ? kPoolJumpChunks : 1;
  long long off = 0;
  int nsettle_all = 0;
  for (int h = 0; h < nh; ++h) {
    Half& H = hv[h];
    for (Ent& e : H.ents) {
      if (c->slots[e.slot].desc.g.kind == RD_GEOM_TIP)
        H.any_tip = true;
      else
        H.any_planar = true;
      e.wg1 = H.nsettle;
      H.nsettle += settle_groups(e.nch);
      if (settle_groups(e.nch) > 1 && c->slots[e.slot].desc.mode == RD_EMIT_PLANAR_FE)
        H.any_split = true;
      const int nch = (e.n_ub + kChunk - 1) / kChunk;
      const int gx = fp64 ? (e.nch + per - 1) / per : jump_groups32(e.nch);
      e.nchunks_ub = nch;
      segment(nch, seg_limit(gx, kWantJump), e.segs, e.cps);
      if (e.cps < min_cps) {
        e.cps = std::min(min_cps, std::max(nch, 1));
        e.segs = (std::max(nch, 1) + e.cps - 1) / e.cps;
      }
      e.part_off = off;
      off += 3LL * (fp64 ? std::max(nch, 1) : e.segs) * e.nch;
    }
    nsettle_all += H.nsettle;
  }
  // Partials per mixed workgroup: one, or kJumpPairs when the launches are
  // large enough to keep every CU busy with that many times fewer workgroups
  // (each then pays the proposal and set-up once for all its sources).  The
  // partials themselves keep their size, so the sums do not depend on this
  // choice.
  int pps = 1;
  if (!fp64 && c->nslots_max > 1) {
    long long wg_all = 0;
    for (int h = 0; h < nh; ++h)
      for (const Ent& e : hv[h].ents) wg_all += (long long)jump_groups32(e.nch) * e.segs;
    if (wg_all >= c->jump_pair_wgs) pps = kJumpPairs;
  }
  for (int h = 0; h < nh; ++h) {
    Half& H = hv[h];
    for (Ent& e : H.ents) {
      const int gx = fp64 ? (e.nch + per - 1) / per : jump_groups32(e.nch);
      e.wg0 = H.nwg;
      H.nwg += gx * ((e.segs + pps - 1) / pps);
    }
  }
  r = ensure(c, c->part, sizeof(double) * (size_t)std::max(off, 1LL));
  if (r) return r;
  const size_t nr = (size_t)maxndim * C;
  r = ensure(c, c->rnd, sizeof(double) * 3 * nr);
  if (r) return r;
  r = ensure(c, c->cnt, sizeof(int) * 4 * (size_t)std::max(nsettle_all, 1));
  if (r) return r;
  if (nh > 1 && !c->jstream[0]) {
    for (int h = 0; h < kJumpStreams - 1; ++h) {
      RD_HIP(hipStreamCreateWithFlags(&c->jstream[h], hipStreamNonBlocking));
      RD_HIP(hipEventCreateWithFlags(&c->ev_join[h], hipEventDisableTiming));
    }
  }
  // the solo launch overlaps the two-launch jumps on a stream of its own
  // (one stream when the loop is captured)
  const bool solo_own = !solo.empty() && nh > 0 && !c->graphs;
  if (solo_own && !c->sstream) {
    RD_HIP(hipStreamCreateWithFlags(&c->sstream, hipStreamNonBlocking));
    RD_HIP(hipEventCreateWithFlags(&c->ev_solo, hipEventDisableTiming));
  }
  if ((nh > 1 || solo_own) && !c->ev_fork)
    RD_HIP(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
  double* rb = (double*)c->rnd.p;
  a.ru1 = rb;
  a.ru2 = rb + nr;
  a.rlr = rb + 2 * nr;
  a.C = C;
  a.part = (double*)c->part.p;
  Ent* dall = nullptr;  // every entry, for the random numbers of all chains
  Ent* dsolo = nullptr;
  if (nh > 1 || !solo.empty()) {
    r = upload_tab(c, all, &dall);
    if (r) return r;
  }
  if (!solo.empty()) {
    r = upload_tab(c, solo, &dsolo);
    if (r) return r;
  }
  int cnt_off = 0;
  for (int h = 0; h < nh; ++h) {
    Half& H = hv[h];
    r = upload_tab(c, H.ents, &H.dt);
    if (r) return r;
    H.a = a;
    H.a.counts = (int*)c->cnt.p + cnt_off;
    H.a.nsettle = H.nsettle;
    cnt_off += 4 * H.nsettle;
    H.stream = h == 0 ? c->stream : c->jstream[h - 1];
  }
  if (nh == 1 && solo.empty()) dall = hv[0].dt;
  const int ne_all = (int)all.size();
  // Everything the launches below depend on besides device data.  With graphs
  // on, an unchanged key replays the captured loop (the kernels read the
  // launch table and the chains from device memory, so new contents under
  // the same pointers are picked up).  Graphs use one stream (nh == 1).
  const Half& G = hv[0];
  std::vector<unsigned char> key;
  auto put = [&key](const void* v, size_t bytes) {
    const unsigned char* b = (const unsigned char*)v;
    key.insert(key.end(), b, b + bytes);
  };
  const int ints[] = {(int)G.ents.size(), G.nwg, G.nsettle, (int)G.any_split,
                      (int)G.any_planar, (int)G.any_tip, (int)fp64, maxndim, C,
                      (int)solo.size(), nh};
  const void* ptrs[] = {G.dt, dall, dsolo, c->d_desc, c->d_state, c->d_wf};
  put(ints, sizeof(ints));
  put(ptrs, sizeof(ptrs));
  put(&G.a, sizeof(G.a));
  put(&V, sizeof(V));
  if (c->graphs && c->jump_exec && key == c->jump_key) {
    RD_HIP(hipGraphLaunch(c->jump_exec, c->stream));
    return RD_OK;
  }
  if (c->graphs) {
    // the stream was synchronised after the previous launch of the old graph
    if (c->jump_exec) RD_HIP(hipGraphExecDestroy(c->jump_exec));
    c->jump_exec = nullptr;
    c->jump_key.clear();
    RD_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  }
  k_mh_random<<<dim3((C + 255) / 256, maxndim), 256, 0, c->stream>>>(dall, ne_all, c->d_desc, C,
                                                                       rb, rb + nr, rb + 2 * nr);
  if (nh > 1 || solo_own) {
    RD_HIP(hipEventRecord(c->ev_fork, c->stream));
    for (int h = 1; h < nh; ++h) RD_HIP(hipStreamWaitEvent(hv[h].stream, c->ev_fork, 0));
    if (solo_own) RD_HIP(hipStreamWaitEvent(c->sstream, c->ev_fork, 0));
  }
  if (!solo.empty()) {
    hipStream_t ss = solo_own ? c->sstream : c->stream;
    k_mh_solo32<<<(int)solo.size(), kSoloThreads, 0, ss>>>(dsolo, c->d_desc, c->d_state,
                                                          c->d_wf, V, a);
    if (solo_own) RD_HIP(hipEventRecord(c->ev_solo, c->sstream));
  }
  for (int it = 1; it <= maxndim; ++it) {
    for (int h = 0; h < nh; ++h) {
      const Half& H = hv[h];
      const int ne = (int)H.ents.size();
      if (fp64) {
        if (H.any_planar)
          k_mh_field64<RD_GEOM_PLANAR><<<H.nwg, kBlock, 0, H.stream>>>(
              H.dt, ne, c->d_desc, c->d_state, V, it, H.a);
        if (H.any_tip)
          k_mh_field64<RD_GEOM_TIP><<<H.nwg, kBlock, 0, H.stream>>>(H.dt, ne, c->d_desc,
                                                                     c->d_state, V, it, H.a);
        k_mh_settle<RD_PREC_FP64><<<H.nsettle, kSettle, 0, H.stream>>>(
            H.dt, ne, c->d_desc, c->d_state, c->d_wf, V, it, H.a);
      } else {
        if (H.any_planar) {
          if (pps == kJumpPairs)
            k_mh_field32<RD_GEOM_PLANAR, kJumpPairs><<<H.nwg, 64 * kJumpWaves, 0, H.stream>>>(
                H.dt, ne, c->d_desc, c->d_state, V, it, H.a);
          else
            k_mh_field32<RD_GEOM_PLANAR, 1><<<H.nwg, 64 * kJumpWaves, 0, H.stream>>>(
                H.dt, ne, c->d_desc, c->d_state, V, it, H.a);
        }
        if (H.any_tip) {
          if (pps == kJumpPairs)
            k_mh_field32<RD_GEOM_TIP, kJumpPairs><<<H.nwg, 64 * kJumpWaves, 0, H.stream>>>(
                H.dt, ne, c->d_desc, c->d_state, V, it, H.a);
          else
            k_mh_field32<RD_GEOM_TIP, 1><<<H.nwg, 64 * kJumpWaves, 0, H.stream>>>(
                H.dt, ne, c->d_desc, c->d_state, V, it, H.a);
        }
        k_mh_settle<RD_PREC_MIXED><<<H.nsettle, kSettle, 0, H.stream>>>(
            H.dt, ne, c->d_desc, c->d_state, c->d_wf, V, it, H.a);
      }
    }
  }
  for (int h = 0; h < nh; ++h) {
    const Half& H = hv[h];
    if (H.any_split)
      k_mh_width_last<<<(int)H.ents.size(), 64, 0, H.stream>>>(H.dt, c->d_desc, c->d_state,
                                                               H.a.counts, H.nsettle);
  }
  for (int h = 1; h < nh; ++h) {
    RD_HIP(hipEventRecord(c->ev_join[h - 1], hv[h].stream));
    RD_HIP(hipStreamWaitEvent(c->stream, c->ev_join[h - 1], 0));
  }
  if (solo_own) RD_HIP(hipStreamWaitEvent(c->stream, c->ev_solo, 0));
  if (c->graphs) {
    hipGraph_t g = nullptr;
    const hipError_t le = hipGetLastError();
    const hipError_t ce = hipStreamEndCapture(c->stream, &g);
    if (le != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return hip_fail(le, "jump loop launch (capture)");
    }
    RD_HIP(ce);
    const hipError_t ie = hipGraphInstantiate(&c->jump_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
      c->jump_exec = nullptr;
      return hip_fail(ie, "hipGraphInstantiate (jump loop)");
    }
    c->jump_key = key;
    RD_HIP(hipGraphLaunch(c->jump_exec, c->stream));
    return RD_OK;
  }
  RD_HIP(hipGetLastError());
  return RD_OK;
}

// Emission for the entries whose simulation emits at this step.  On return
// the device state holds the new particle counts; n_round[k] receives the
// number of chains of entry k (an upper bound of the electrons added).
static int engine_emission(rd_ctx* c, const std::vector<Ent>& all, int precision,
                           std::vector<int>& n_round) {
  n_round.assign(all.size(), 0);
  // Supply integral.  The first batch holds level 0 of every emitting entry
  // and, where refinement is allowed, level 1; each later batch holds the next
  // level of the entries whose last two levels disagree.
  std::vector<Ent> q, top;
  std::vector<int> qidx;
  for (size_t k = 0; k < all.size(); ++k) {
    const Ent& e0 = all[k];
    const SimDesc& d = c->slots[e0.slot].desc;
    if (d.mode == RD_EMIT_NONE || !(d.delay < e0.step)) continue;
    Ent e = e0;
    e.lev = 0;
    q.push_back(e);
    if (d.mode != RD_EMIT_TIP_FE && d.quad_lmax >= 1) {
      e.lev = 1;
      q.push_back(e);
    }
    top.push_back(e);
    qidx.push_back((int)k);
  }
  if (q.empty()) return RD_OK;
  int r;
  while (!q.empty()) {
    int mtot = 0, maxm = 0, ntiles = 0, maxt = 0;
    for (Ent& e : q) {
      const SimDesc& d = c->slots[e.slot].desc;
      e.pbase = mtot;
      e.m = (d.mode == RD_EMIT_TIP_FE) ? 100 * 100 : (d.qgx << e.lev) * (d.qgy << e.lev);
      mtot += e.m;
      maxm = std::max(maxm, e.m);
      const int nt = (e.m + 1023) / 1024;  // tiles of the level's sums
      e.c_off = ntiles;
      ntiles += nt;
      maxt = std::max(maxt, nt);
    }
    r = ensure(c, c->pts, sizeof(double) * 3 * (size_t)mtot);
    if (!r) r = ensure(c, c->fld, sizeof(double) * 3 * (size_t)mtot);
    if (!r) r = ensure(c, c->stage, sizeof(double) * 4 * (size_t)ntiles);
    if (r) return r;
    double* ptx = (double*)c->pts.p;
    Ent* dq;
    r = upload_tab(c, q, &dq);
    if (r) return r;
    k_quad_points<<<dim3((maxm + 255) / 256, 1, q.size()), 256, 0, c->stream>>>(
        dq, c->d_desc, ptx, ptx + mtot, ptx + 2 * mtot);
    r = launch_field(c, q, mtot, precision, nullptr, nullptr);
    if (r) return r;
    double* f = (double*)c->fld.p;
    k_quad_tiles<<<dim3(maxt, q.size()), 1024, 0, c->stream>>>(
        dq, c->d_desc, c->d_wf, ptx, ptx + mtot, f, f + mtot, f + 2 * mtot, (double*)c->stage.p);
    k_quad_tsum<<<q.size(), 64, 0, c->stream>>>(dq, c->d_state, (const double*)c->stage.p);
    Ent* dtop;
    r = upload_tab(c, top, &dtop);
    if (r) return r;
    k_quad_close<<<top.size(), 1024, 0, c->stream>>>(dtop, c->d_desc, c->d_state);
    RD_HIP(hipGetLastError());
    r = sync_state(c);
    if (r) return r;
    q.clear();
    std::vector<Ent> next;
    for (Ent e : top)
      if (!c->h_state[e.slot].q_final) {
        e.lev += 1;
        q.push_back(e);
        next.push_back(e);
      }
    top.swap(next);
  }
  // chains
  std::vector<Ent> ch;
  int C = 0, maxc = 0, maxndim = 0;
  for (size_t k = 0; k < qidx.size(); ++k) {
    Ent e = all[qidx[k]];
    int nr = c->h_state[e.slot].n_round;
    n_round[qidx[k]] = nr;
    if (nr <= 0) continue;
    e.cbase = C;
    e.nch = nr;
    e.pbase = C;
    e.m = nr;
    C += nr;
    maxc = std::max(maxc, nr);
    maxndim = std::max(maxndim, c->slots[e.slot].desc.ndim);
    ch.push_back(e);
  }
  if (C == 0) return RD_OK;
  r = ensure(c, c->chains, chain_bytes(C));
  if (!r) r = ensure(c, c->pts, sizeof(double) * 3 * (size_t)C);
  if (!r) r = ensure(c, c->fld, sizeof(double) * 3 * (size_t)C);
  if (r) return r;
  double* ptx = (double*)c->pts.p;
  double* f = (double*)c->fld.p;
  double *pty = ptx + C, *ptz = ptx + 2 * (size_t)C;
  double *fx = f, *fy = f + C, *fz = f + 2 * (size_t)C;
  Chains V = chain_view(c, C);
  Ent* dc;
  r = upload_tab(c, ch, &dc);
  if (r) return r;
  dim3 cg((maxc + 255) / 256, 1, ch.size());
  k_mh_init<<<cg, 256, 0, c->stream>>>(dc, c->d_desc, c->d_state, V);
  // start sites: rounds until every chain has F < 0 (at most 10000)
  for (int round = 1;; ++round) {
    const int last = round >= 10001;
    k_mh_start<<<cg, 256, 0, c->stream>>>(dc, c->d_desc, V, round, ptx, pty, ptz);
    r = launch_field(c, ch, C, precision, nullptr, nullptr);
    if (r) return r;
    k_mh_start_check<<<ch.size(), 256, 0, c->stream>>>(dc, c->d_desc, c->d_state, c->d_wf, V,
                                                        round, last, ptx, pty, ptz, fx, fy, fz);
    RD_HIP(hipGetLastError());
    r = sync_state(c);
    if (r) return r;
    int pend = 0;
    for (const Ent& e : ch) pend += c->h_state[e.slot].pending;
    if (pend == 0 || last) break;
  }
  r = run_jumps(c, ch, C, precision, V, maxndim);
  if (r) return r;
  r = upload_tab(c, ch, &dc);
  if (r) return r;
  k_mh_finish<<<ch.size(), 256, 0, c->stream>>>(dc, c->d_desc, c->d_state, c->d_wf, V);
  k_emit<<<ch.size(), 256, 0, c->stream>>>(dc, c->d_desc, c->d_state, c->d_wf, V, RD_ARRAYS,
                                            c->logs ? (rd_emit_record*)c->elog.p : nullptr);
  RD_HIP(hipGetLastError());
  return RD_OK;
}

static void fill_record(rd_ctx* c, int slot, int step, rd_step_record* rec) {
  const SimState& S = c->h_state[slot];
  std::memset(rec, 0, sizeof(*rec));
  rec->step = step;
  rec->n = S.n;
  rec->emitted = S.emitted;
  rec->n_round = S.n_round;
  rec->absorbed = S.rem_total;
  rec->absorbed_top = S.rem_top;
  rec->absorbed_bot = S.rem_bot;
  rec->dropped = S.dropped;
  rec->start_rounds = S.start_rounds;
  rec->supply_level = S.q_level;
  rec->supply_cells = S.q_cells;
  rec->supply_fail = S.q_fail;
  rec->supply_err = S.q_err;
  rec->n_sup = S.N_sup;
  rec->ramo_current = S.ramo[0] + S.ramo[1] + S.ramo[2];
  rec->mh_std = S.MH_std;
  rec->a_rate = S.a_rate;
  rec->df_avg = S.df_avg;
  for (int k = 0; k < 3; ++k) {
    rec->f_avg[k] = S.F_avg[k];
    rec->ramo[k] = S.ramo[k];
    rec->avg_elec_vel[k] = S.n_elec > 0 ? S.vsum[3 + k] / S.n_elec : 0.0;
  }
}

// One main-loop step (emission, position/force/velocity, removal) of the
// given slots.  phases: bit 0 emission, bit 1 update, bit 2 removal.
int engine_phases(rd_ctx* c, const std::vector<int>& slots, int precision, int phases,
                  const int* steps, rd_step_record* recs) {
  std::vector<Ent> ents;
  for (size_t k = 0; k < slots.size(); ++k) {
    Ent e{};
    e.slot = slots[k];
    e.base = c->slots[e.slot].desc.base;
    e.n_ub = c->slots[e.slot].n;
    e.step = steps[k];
    ents.push_back(e);
  }
  if (ents.empty()) return RD_OK;
  if ((phases & 2) && precision == RD_PREC_MIXED)
    for (const Ent& e : ents) {
      const SimDesc& d = c->slots[e.slot].desc;
      if (d.tts > 0 && d.g.kind == RD_GEOM_TIP)
        return fail(RD_ERR_ARG, "two-time-step mode on the tip runs in FP64 only");
    }
  int r;
  if (phases & 1) {
    r = reset_counters(c, ents);
    if (r) return r;
    for (const Ent& e : ents) c->slots[e.slot].emit_step = e.step;
  }
  std::vector<int> nr;
  if (phases & 1) {
    r = engine_emission(c, ents, precision, nr);
    if (r) return r;
    for (size_t k = 0; k < ents.size(); ++k) {
      const SimDesc& d = c->slots[ents[k].slot].desc;
      ents[k].n_ub = std::min(d.cap, ents[k].n_ub + nr[k]);
    }
  }
  if (phases & 2) {
    if (!(phases & 1)) {
      // A step's counters start at zero with its first phase: the emission
      // phase of the same step when the caller ran it on its own, else this
      // update.
      std::vector<Ent> fresh;
      for (const Ent& e : ents)
        if (c->slots[e.slot].emit_step != e.step) fresh.push_back(e);
      if (!fresh.empty()) {
        r = reset_counters(c, fresh);
        if (r) return r;
      }
    }
    r = engine_position(c, ents);
    if (!r) {
      // Update_Particle_Acceleration: the two-time-step accelerations where
      // the simulation runs that mode
      std::vector<Ent> acc = ents;
      for (Ent& e : acc) e.tts = c->slots[e.slot].desc.tts;
      r = launch_accel(c, acc, precision, nullptr, nullptr);
    }
    if (!r) r = engine_velocity(c, ents);
    if (r) return r;
  }
  if (phases & 4) {
    r = engine_compact(c, ents);
    if (r) return r;
  }
  r = sync_state(c);
  if (r) return r;
  if (recs)
    for (size_t k = 0; k < ents.size(); ++k) fill_record(c, ents[k].slot, ents[k].step, &recs[k]);
  return RD_OK;
}

}  // namespace rd
