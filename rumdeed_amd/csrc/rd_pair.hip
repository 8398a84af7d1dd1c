// All-pairs Coulomb force, image charges and field at points for gfx950.
//
// Reference: Calculate_Acceleration_Particles_ACC (src/mod_verlet.F90:1217-1429),
// Calc_Field_at / Calc_Field_at_Batch (:1466-1529, :1635-1911), the planar image
// series (src/acc_ic_planar_series.inc:19-63) and the hyperboloid tip
// (src/acc_tip_field_E.inc, src/acc_tip_image_point.inc, src/acc_tip_ic_force.inc).
//
// Shape: one lane per target (particle or field point), 256-lane workgroups;
// sources are staged through LDS 256 at a time and read as wavefront-uniform
// broadcasts.  blockIdx.z selects a simulation (launch-table entry), so one
// launch serves every simulation of a pool; the source range of a simulation
// is split over gridDim.y segments so that a launch has enough workgroups to
// fill 256 CUs.  Every (segment, target) writes FP64 partial sums and a
// finishing kernel adds them in a fixed order.  The pair work is vector-ALU
// arithmetic (square roots, reciprocals, FMAs), not a contraction, so
// nothing here uses MFMA.
//
// FP64 kernels: operand-for-operand the reference expressions, with the
//   partials formed per 256-source chunk and added chunk by chunk, so the
//   result is bit-identical to the CPU restatement that uses the same order.
// Mixed kernels: positions packed as float hi + float lo (in nm); pair
//   offsets are formed from both parts, so they are as accurate as FP32 allows
//   even for close pairs; image offsets are formed from sums of non-negative
//   heights (z_a + z_b from the cathode, (d - z_a) + (d - z_b) from the anode),
//   which never cancel; the pair arithmetic is FP32 with v_rsq_f32 and the
//   per-target sums are FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "rd_field.h"
#include "rd_internal.h"

namespace rd {

// ------------------------------------------------------------------------

// Workgroup of a compact pair launch: the launch-table entry it belongs to
// (workgroups [wg0, next wg0) of the 1-D grid), its target block bx and its
// source segment by.  The host lays the entry out as gx target blocks by
// segs segments with gx = max(1, ceil(targets / kBlock)).
struct Tile {
  Ent e;
  int bx, by;
};
__device__ __forceinline__ Tile tile_of(const Ent* tab, int ne, bool points) {
  Tile t;
  t.e = tab[find_entry(tab, ne, blockIdx.x)];
  const int targets = points ? t.e.m : t.e.n_ub;
  const int gx = max(1, (targets + kBlock - 1) / kBlock);
  const int local = blockIdx.x - t.e.wg0;
  t.bx = local % gx;
  t.by = local / gx;
  return t;
}

// Image points of every particle (they depend on the particle alone).
__global__ void k_tip_images64(const Ent* tab, const SimDesc* desc, const SimState* st,
                               const double* __restrict__ X, const double* __restrict__ Y,
                               const double* __restrict__ Z, double* __restrict__ img) {
  const Ent e = tab[blockIdx.z];
  const int n = st[e.slot].n;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const DevGeom& g = desc[e.slot].g;
  if (g.kind != RD_GEOM_TIP || !g.ic) return;
  const int gi = e.base + i;
  double da, xi, yi, zi;
  tip_image_point64(X[gi], Y[gi], Z[gi], g.z0, g.R, da, xi, yi, zi);
  img[4 * (size_t)gi] = da;
  img[4 * (size_t)gi + 1] = xi;
  img[4 * (size_t)gi + 2] = yi;
  img[4 * (size_t)gi + 3] = zi;
}

// ------------------------------------------------------------------------
// Two-time-step accelerations (Update_Elec_Acceleration and
// Update_Ion_Acceleration, src/mod_verlet.F90:899-1151, with atom_time_interval
// K = Ent::tts > 0).  Every step an electron gets the force of every particle
// and the vacuum field; an ion gets the reaction of every electron, and at
// steps that are multiples of K also the force of the other ions and the
// vacuum field, all added to the acceleration it holds (an ion's is cleared
// only by its own position update, every K steps).  Between an electron and
// an ion the electron takes the first role of the pair (its image series is
// evaluated at the electron; the ion gets the mirrored reaction); within a
// species the lower index does, as in the standard loops.
struct TtsRole {
  int sp1 = 0;            // the target's species
  bool ion_only_e = false;  // ion target between its updates: electron sources only
};
__device__ __forceinline__ TtsRole tts_role(const Ent& e, const int32_t* sp, int gi, bool act) {
  TtsRole r;
  if (e.tts > 0 && act) {
    r.sp1 = sp[gi];
    r.ion_only_e = r.sp1 == RD_SPECIES_ION && (e.step % e.tts) != 0;
  }
  return r;
}
// Whether the target takes the first role of the pair with source j (species
// sp2), and whether the pair contributes at all.
__device__ __forceinline__ bool tts_pair(const TtsRole& r, int i, int j, int sp2, bool& first) {
  first = j > i;
  if (r.sp1 != sp2) first = r.sp1 == RD_SPECIES_ELEC;
  return !(r.ion_only_e && sp2 == RD_SPECIES_ION);
}

// FP64 accelerations.  part layout per entry: [3][nchunks_ub][n_ub].
template <int KIND>
__global__ __launch_bounds__(kBlock) void k_accel64(const Ent* tab, int ne, const SimDesc* desc,
                                                    const SimState* st,
                                                    const double* __restrict__ X,
                                                    const double* __restrict__ Y,
                                                    const double* __restrict__ Z,
                                                    const double* __restrict__ Q,
                                                    const int32_t* __restrict__ SP,
                                                    const double* __restrict__ img,
                                                    double* __restrict__ part) {
  __shared__ double sx[kChunk], sy[kChunk], sz[kChunk], sq[kChunk];
  __shared__ int ssp[kChunk];
  __shared__ double sd[KIND == RD_GEOM_TIP ? kChunk : 1], sxi[KIND == RD_GEOM_TIP ? kChunk : 1],
      syi[KIND == RD_GEOM_TIP ? kChunk : 1], szi[KIND == RD_GEOM_TIP ? kChunk : 1];
  const Tile tl = tile_of(tab, ne, false);
  const int n = st[tl.e.slot].n;
  if (tl.bx * kBlock >= n) return;
  const int nchunks = min((n + kChunk - 1) / kChunk, tl.e.band1);
  const int c0 = tl.e.band0 + tl.by * tl.e.cps;
  if (c0 >= nchunks) return;
  const int c1 = min(nchunks, c0 + tl.e.cps);
  const DevGeom g = desc[tl.e.slot].g;
  if (g.kind != KIND) return;
  const int base = tl.e.base;
  const double* x = X + base;
  const double* y = Y + base;
  const double* z = Z + base;
  const double* q = Q + base;
  double* P = part + tl.e.part_off;
  const int t = threadIdx.x;
  const int i = tl.bx * kBlock + t;
  const bool act = i < n;
  const bool tts = tl.e.tts > 0;
  const TtsRole role = tts_role(tl.e, SP, base + i, act);
  double x1 = 0, y1 = 0, z1 = 0, qd1 = 0, d1 = 0, xi1 = 0, yi1 = 0, zi1 = 0;
  if (act) {
    x1 = x[i];
    y1 = y[i];
    z1 = z[i];
    qd1 = q[i] * g.dfc;
    if (KIND == RD_GEOM_TIP && g.ic) {
      d1 = img[4 * (size_t)(base + i)];
      xi1 = img[4 * (size_t)(base + i) + 1];
      yi1 = img[4 * (size_t)(base + i) + 2];
      zi1 = img[4 * (size_t)(base + i) + 3];
    }
  }
  for (int c = c0; c < c1; ++c) {
    const int jb = c * kChunk;
    __syncthreads();
    if (jb + t < n) {
      sx[t] = x[jb + t];
      sy[t] = y[jb + t];
      sz[t] = z[jb + t];
      sq[t] = q[jb + t];
      if (tts) ssp[t] = SP[base + jb + t];
      if (KIND == RD_GEOM_TIP && g.ic) {
        const size_t gj = 4 * (size_t)(base + jb + t);
        sd[t] = img[gj];
        sxi[t] = img[gj + 1];
        syi[t] = img[gj + 2];
        szi[t] = img[gj + 3];
      }
    }
    __syncthreads();
    const int cnt = min(kChunk, n - jb);
    double px = 0.0, py = 0.0, pz = 0.0;
    if (act) {
      for (int jj = 0; jj < cnt; ++jj) {
        const int j = jb + jj;
        if (j == i) continue;
        // the target takes the first role of the pair (the lower index; see
        // TtsRole for the two-time-step update)
        bool first = j > i;
        if (tts && !tts_pair(role, i, j, ssp[jj], first)) continue;
        double x2 = sx[jj], y2 = sy[jj], z2 = sz[jj];
        double pre = qd1 * sq[jj];
        double dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
        double axx, ayy, azz;
        if (KIND == RD_GEOM_PLANAR) {
          double dxy2 = dx * dx + dy * dy;
          double r = __builtin_sqrt(dxy2 + dz * dz) + g.es;
          double inv_r3 = 1.0 / (r * r * r);
          axx = pre * inv_r3 * dx;
          ayy = pre * inv_r3 * dy;
          azz = pre * inv_r3 * dz;
          if (g.ic) {
            double za = first ? z1 : z2;
            double zb = first ? z2 : z1;
            double icx, icy, icz;
            ic_planar64(za, zb, dx, dy, dxy2, g.nic, g.d, g.es, icx, icy, icz);
            axx = axx + pre * icx;
            ayy = ayy + pre * icy;
            azz = azz + pre * icz;
          }
        } else {
          double r = __builtin_sqrt(dx * dx + dy * dy + dz * dz) + g.es;
          double inv_r3 = 1.0 / (r * r * r);
          axx = pre * inv_r3 * dx;
          ayy = pre * inv_r3 * dy;
          azz = pre * inv_r3 * dz;
          if (g.ic) {
            double icx, icy, icz;
            if (first) {
              tip_ic_force64(x1, y1, z1, x2, y2, z2, d1, xi1, yi1, zi1, g.R, g.ktip, icx, icy,
                             icz);
              axx = axx + pre * 1.0 * icx;
              ayy = ayy + pre * 1.0 * icy;
            } else {
              tip_ic_force64(x2, y2, z2, x1, y1, z1, sd[jj], sxi[jj], syi[jj], szi[jj], g.R,
                             g.ktip, icx, icy, icz);
              axx = axx + pre * -1.0 * icx;
              ayy = ayy + pre * -1.0 * icy;
            }
            azz = azz + pre * icz;
          }
        }
        px = px + axx;
        py = py + ayy;
        pz = pz + azz;
      }
      const size_t S = (size_t)tl.e.pslots * tl.e.n_ub;
      const size_t o = (size_t)(c - tl.e.band0) * tl.e.n_ub + i;
      P[0 * S + o] = px;
      P[1 * S + o] = py;
      P[2 * S + o] = pz;
    }
  }
}

// Running sums of a banded FP64 launch: acc = acc + the band's chunk
// partials, in chunk order (the first band starts from zero), so that the
// bands together add every chunk in the order of one unbanded launch.
// points: the targets are the entry's m points, else its particles.
__global__ void k_fold64(const Ent* tab, const SimState* st, const double* __restrict__ part,
                         double* __restrict__ acc, int points) {
  const Ent e = tab[blockIdx.z];
  if (!e.banded) return;
  const int n = st[e.slot].n;
  const int ld = points ? e.m : e.n_ub;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (points ? e.m : n)) return;
  const int nchunks = (n + kChunk - 1) / kChunk;
  const int cl = min(nchunks, e.band1);
  double* A = acc + e.acc_off;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  if (e.band0 > 0) {
    if (cl <= e.band0) return;
    sx = A[p];
    sy = A[(size_t)ld + p];
    sz = A[2 * (size_t)ld + p];
  }
  if (cl > e.band0)
    sum3_in_order(part + e.part_off + p, (size_t)e.pslots * ld, ld, cl - e.band0, sx, sy, sz);
  A[p] = sx;
  A[(size_t)ld + p] = sy;
  A[2 * (size_t)ld + p] = sz;
}

// Sum the chunk partials in order (banded: the running sums) and form
// a = (F + q E_vac)/m.
__global__ void k_accel_finish64(const Ent* tab, const SimDesc* desc, const SimState* st,
                                 const double* __restrict__ part, const double* __restrict__ acc,
                                 const double* __restrict__ X,
                                 const double* __restrict__ Y, const double* __restrict__ Z,
                                 const double* __restrict__ Q, const double* __restrict__ M,
                                 const int32_t* __restrict__ SP,
                                 double* __restrict__ AX, double* __restrict__ AY,
                                 double* __restrict__ AZ) {
  const Ent e = tab[blockIdx.z];
  const int n = st[e.slot].n;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const DevGeom& g = desc[e.slot].g;
  const int nchunks = (n + kChunk - 1) / kChunk;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  if (e.banded) {
    const double* A = acc + e.acc_off;
    sx = A[i];
    sy = A[(size_t)e.n_ub + i];
    sz = A[2 * (size_t)e.n_ub + i];
  } else {
    sum3_in_order(part + e.part_off + i, (size_t)e.pslots * e.n_ub, e.n_ub, nchunks, sx, sy, sz);
  }
  const int gi = e.base + i;
  double q1 = Q[gi];
  double im = 1.0 / M[gi];
  // two-time-step ions add onto the acceleration they hold, with the vacuum
  // field only at their update steps
  const bool add = e.tts > 0 && SP[gi] == RD_SPECIES_ION;
  const bool vac = !add || (e.step % e.tts) == 0;
  double ax, ay, az;
  if (g.kind == RD_GEOM_PLANAR) {
    ax = sx * im;
    ay = sy * im;
    az = (vac ? sz + q1 * g.Ez : sz) * im;
  } else if (vac) {
    double fx, fy, fz;
    tip_vacuum64(g, X[gi], Y[gi], Z[gi], fx, fy, fz);
    ax = (sx + q1 * fx) * im;
    ay = (sy + q1 * fy) * im;
    az = (sz + q1 * fz) * im;
  } else {
    ax = sx * im;
    ay = sy * im;
    az = sz * im;
  }
  if (add) {
    ax = AX[gi] + ax;
    ay = AY[gi] + ay;
    az = AZ[gi] + az;
  }
  AX[gi] = ax;
  AY[gi] = ay;
  AZ[gi] = az;
}

// ------------------------------------------------------------------------
// FP64 field at points.  Points of an entry: [pbase, pbase + m) of the SoA
// point arrays; part layout per entry: [3][nchunks_ub][m].
template <int KIND>
__global__ __launch_bounds__(kBlock) void k_field64(const Ent* tab, int ne, const SimDesc* desc,
                                                    const SimState* st,
                                                    const double* __restrict__ ptx,
                                                    const double* __restrict__ pty,
                                                    const double* __restrict__ ptz,
                                                    const double* __restrict__ X,
                                                    const double* __restrict__ Y,
                                                    const double* __restrict__ Z,
                                                    const double* __restrict__ Q,
                                                    double* __restrict__ part) {
  __shared__ double sx[kChunk], sy[kChunk], sz[kChunk], sq[kChunk];
  const Tile tl = tile_of(tab, ne, true);
  const Ent& e = tl.e;
  const int m = e.m;
  if (tl.bx * kBlock >= m) return;
  const int n = st[e.slot].n;
  const int nchunks = min((n + kChunk - 1) / kChunk, e.band1);
  const int c0 = e.band0 + tl.by * e.cps;
  if (c0 >= nchunks) return;
  const int c1 = min(nchunks, c0 + e.cps);
  const DevGeom g = desc[e.slot].g;
  if (g.kind != KIND) return;
  const int p = tl.bx * kBlock + threadIdx.x;
  const bool act = p < m;
  double x1 = 0, y1 = 0, z1 = 0;
  if (act) {
    x1 = ptx[e.pbase + p];
    y1 = pty[e.pbase + p];
    z1 = ptz[e.pbase + p];
  }
  field_chunks64<KIND>(g, act, x1, y1, z1, X + e.base, Y + e.base, Z + e.base, Q + e.base, n, c0,
                       c1, sx, sy, sz, sq, part + e.part_off, (size_t)e.pslots * m, m, p,
                       e.band0);
}

// field = E_vac + sum of chunk partials (in order).
__global__ void k_field_finish64(const Ent* tab, const SimDesc* desc, const SimState* st,
                                 const double* __restrict__ part, const double* __restrict__ acc,
                                 const double* __restrict__ ptx,
                                 const double* __restrict__ pty, const double* __restrict__ ptz,
                                 double* __restrict__ fx, double* __restrict__ fy,
                                 double* __restrict__ fz) {
  const Ent e = tab[blockIdx.z];
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= e.m) return;
  const int n = st[e.slot].n;
  const int nchunks = (n + kChunk - 1) / kChunk;
  const int gp = e.pbase + p;
  if (e.banded) {
    const double* A = acc + e.acc_off;
    double ex, ey, ez;
    vacuum64(desc[e.slot].g, ptx[gp], pty[gp], ptz[gp], ex, ey, ez);
    fx[gp] = ex + A[p];
    fy[gp] = ey + A[(size_t)e.m + p];
    fz[gp] = ez + A[2 * (size_t)e.m + p];
    return;
  }
  field_finish64(desc[e.slot].g, part + e.part_off, (size_t)e.pslots * e.m, e.m, p, nchunks,
                 ptx[gp], pty[gp], ptz[gp], fx[gp], fy[gp], fz[gp]);
}
// ------------------------------------------------------------------------
// Mixed precision (packed records and image weights: rd_field.h).

// Pack the particles of every entry (targets = sources).
__global__ void k_pack_particles(const Ent* tab, const SimDesc* desc, const SimState* st,
                                 const double* __restrict__ X, const double* __restrict__ Y,
                                 const double* __restrict__ Z, const double* __restrict__ Q,
                                 float4* __restrict__ hi, float4* __restrict__ lo,
                                 float4* __restrict__ srf) {
  const Ent e = tab[blockIdx.z];
  const int n = st[e.slot].n;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int gi = e.base + i;
  float4 h, l;
  pack1(X[gi], Y[gi], Z[gi], Q[gi], desc[e.slot].g.d, h, l);
  hi[gi] = h;
  lo[gi] = l;
  if (srf) srf[gi] = make_float4(h.x, h.y, h.z, h.w * h.z);
}

// Far-partner coefficients (surface_far_coeffs) of the particles of every
// planar entry with image partners beyond n = 0: far[2 i], far[2 i + 1].
__global__ void k_pack_far(const Ent* tab, const SimDesc* desc, const SimState* st,
                           const double* __restrict__ Z, float4* __restrict__ far) {
  const Ent e = tab[blockIdx.z];
  const DevGeom& g = desc[e.slot].g;
  if (g.kind != RD_GEOM_PLANAR || !g.ic || g.nic < 1) return;
  const int n = st[e.slot].n;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t gi = (size_t)e.base + i;
  const double zn = Z[gi] * 1.0e9, dnm = g.d * 1.0e9;
  float4 c0, c1;
  surface_far_coeffs(zn, dnm - zn, dnm, g.nic, c0, c1);
  far[2 * gi] = c0;
  far[2 * gi + 1] = c1;
}

// Pack the field points of every entry.
__global__ void k_pack_points(const Ent* tab, const SimDesc* desc, const double* __restrict__ ptx,
                              const double* __restrict__ pty, const double* __restrict__ ptz,
                              float4* __restrict__ hi, float4* __restrict__ lo) {
  const Ent e = tab[blockIdx.z];
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= e.m) return;
  const int gp = e.pbase + p;
  float4 h, l;
  pack1(ptx[gp], pty[gp], ptz[gp], 0.0, desc[e.slot].g.d, h, l);
  hi[gp] = h;
  lo[gp] = l;
}


// Mixed accelerations (planar).  part layout per entry: [3][segs][n_ub] FP64
// sums of q_j/q_0 * (offset / r^3 + image terms) in nm^-2.
template <bool NEAR, bool TTS>
__global__ __launch_bounds__(kBlock) void k_accel_planar32(const Ent* tab, int ne, const SimDesc* desc,
                                                           const SimState* st,
                                                           const float4* __restrict__ HI,
                                                           const float4* __restrict__ LO,
                                                           const int32_t* __restrict__ SP,
                                                           double* __restrict__ part) {
  __shared__ float4 shi[kChunk], slo[kChunk];
  __shared__ int ssp[TTS ? kChunk : 1];
  const Tile tl = tile_of(tab, ne, false);
  const Ent& e = tl.e;
  const int n = st[e.slot].n;
  if (tl.bx * kBlock >= n) return;
  const int nchunks = (n + kChunk - 1) / kChunk;
  const int c0 = tl.by * e.cps;
  if (c0 >= nchunks) return;
  const int c1 = min(nchunks, c0 + e.cps);
  const DevGeom& gg = desc[e.slot].g;
  if (gg.kind != RD_GEOM_PLANAR) return;
  const int ic = gg.ic, nic = e.nic_pair;
  const float two_d = (float)(2.0 * gg.d * 1e9);
  const float4* hi = HI + e.base;
  const float4* lo = LO + e.base;
  const int t = threadIdx.x;
  const int i = tl.bx * kBlock + t;
  const bool act = i < n;
  float4 h1 = make_float4(0, 0, 0, 0), l1 = make_float4(0, 0, 0, 0);
  if (act) {
    h1 = hi[i];
    l1 = lo[i];
  }
  TtsRole role;
  if (TTS) role = tts_role(e, SP, e.base + i, act);
  double fx = 0.0, fy = 0.0, fz = 0.0;
  for (int c = c0; c < c1; ++c) {
    const int jb = c * kChunk;
    __syncthreads();
    if (jb + t < n) {
      shi[t] = hi[jb + t];
      slo[t] = lo[jb + t];
      if (TTS) ssp[t] = e.tts > 0 ? SP[e.base + jb + t] : 0;
    }
    __syncthreads();
    const int cnt = min(kChunk, n - jb);
    if (act) {
      for (int jj = 0; jj < cnt; ++jj) {
        const int j = jb + jj;
        const float4 h2 = shi[jj], l2 = slo[jj];
        float dx = (h1.x - h2.x) + (l1.x - l2.x);
        float dy = (h1.y - h2.y) + (l1.y - l2.y);
        float dz = (h1.z - h2.z) + (l1.z - l2.z);
        float qj = (j == i) ? 0.0f : h2.w;
        bool first = j > i;  // the target takes the first role of the pair
        if (TTS && e.tts > 0 && !tts_pair(role, i, j, ssp[jj], first)) qj = 0.0f;
        float dxy2 = __builtin_fmaf(dy, dy, dx * dx);
        float r2 = fmaxf(__builtin_fmaf(dz, dz, dxy2), kR2Min);
        float ri = rsq(r2);
        float r3 = ri * ri * ri;
        float cxy = r3, cz = dz * r3;
        if (ic) {
          float wxy, wz;
          float dzab = first ? dz : -dz;
          ic_planar32<NEAR>(h1.z + h2.z, l1.w + l2.w, dzab, dxy2, nic, two_d, wxy, wz);
          cxy += wxy;
          cz += wz;
        }
        cxy *= qj;
        cz *= qj;
        fx += (double)(cxy * dx);
        fy += (double)(cxy * dy);
        fz += (double)cz;
      }
    }
  }
  if (act) {
    double* P = part + e.part_off;
    const size_t S = (size_t)e.segs * e.n_ub;
    P[0 * S + (size_t)tl.by * e.n_ub + i] = fx;
    P[1 * S + (size_t)tl.by * e.n_ub + i] = fy;
    P[2 * S + (size_t)tl.by * e.n_ub + i] = fz;
  }
}

// ------------------------------------------------------------------------
// Mixed accelerations (planar), pair-symmetric form for large simulations.
//
// Calculate_Acceleration_Particles_Planar (src/mod_verlet.F90:805-872)
// evaluates every unordered pair once: the image series of the pair (i < j)
// at z_i from j's partners, +F on i, and on j the Coulomb part reversed and
// the image part with x and y reversed and z kept.  This kernel does the
// same on the GPU.  Particles form blocks of kSymBlock; workgroup (X, Y),
// X <= Y, takes the block pair.
//   X < Y: wavefront w holds 64 targets of block X (one per lane) and, in
//     phase p, the 64 sources of tile (w + p) mod kSymWaves of block Y.  Over
//     64 steps lane l pairs its target with every source of the tile; the
//     sources' FP64 reaction sums ride along the wavefront (rotated one lane
//     per step), so every source's sum is added in a fixed order without
//     atomics.  Each tile belongs to one wavefront per phase, and the phases
//     are separated by barriers.
//   X = Y: the block's pairs by gather, each lane over the whole block,
//     with the reference's lower-index roles (as k_accel_planar32).
// Partials: for a particle of block X, slot Y holds its sum from the block
// pair (X, Y) for Y > X, from (Y, X) for Y < X, and the diagonal for Y = X:
// the [3][segs][n_ub] layout of the gather kernels with segs = blocks, so
// k_accel_finish32 adds them in block order.  A banded launch takes the rows
// X in [band0, band1): slot X of the particles of later blocks (the
// reaction sums) at part_off, [3][pslots][n_ub] from slot band0; the slots
// Y >= X of the rows' own particles at d_off, [3][segs][d_ld] from particle
// band0 * kSymBlock; k_fold_sym adds them to the running sums in slot order.
constexpr int kSymWaves = 16;
constexpr int kSymBlock = 64 * kSymWaves;
constexpr int kSymMin = 16384;

// One lane to the next: lane l takes lane l + 1's value (wave_rol:1).  Every
// lane receives a value, so the old-value operand is v itself: the move can
// write in place, with no register cleared for it first.
__device__ __forceinline__ int wave_rotate(int v) {
  return __builtin_amdgcn_update_dpp(v, v, 0x134, 0xF, 0xF, false);
}
__device__ __forceinline__ double wave_rotate(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const int lo = wave_rotate((int)(unsigned)b);
  const int hi = wave_rotate((int)(unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

template <bool NEAR>
__global__ __launch_bounds__(kSymBlock) void k_accel_sym32(const Ent* tab, int ne,
                                                           const SimDesc* desc, const SimState* st,
                                                           const float4* __restrict__ HI,
                                                           const float4* __restrict__ LO,
                                                           double* __restrict__ part) {
  __shared__ float4 shi[kSymBlock], slo[kSymBlock];
  __shared__ double jsum[3][kSymBlock];
  const Ent e = tab[find_entry(tab, ne, blockIdx.x, true)];
  const int nb = e.segs;
  int local = blockIdx.x - e.wg1;
  int X = e.band0;
  while (local >= nb - X) {
    local -= nb - X;
    ++X;
  }
  const int Y = X + local;
  const int n = st[e.slot].n;
  if (Y * kSymBlock >= n) return;  // X <= Y: both blocks hold particles
  const DevGeom& gg = desc[e.slot].g;
  const int ic = gg.ic, nic = e.nic_pair;
  const float two_d = (float)(2.0 * gg.d * 1e9);
  const float dnm = (float)(gg.d * 1e9);
  const float4* hi = HI + e.base;
  const float4* lo = LO + e.base;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  // padding (beyond n): charge 0 at the origin, so every term it enters is 0
  const float4 h0 = make_float4(0, 0, 0, 0), l0 = make_float4(0, 0, 0, dnm);
  const int i = X * kSymBlock + t;
  // (values, not a select of addresses: selecting between a global element
  // and a local would put the local in scratch memory)
  float4 h1 = h0, l1 = l0;
  if (i < n) {
    h1 = hi[i];
    l1 = lo[i];
  }
  const int js = Y * kSymBlock + t;
  float4 hs = h0, ls = l0;
  if (js < n) {
    hs = hi[js];
    ls = lo[js];
  }
  shi[t] = hs;
  slo[t] = ls;
  jsum[0][t] = 0.0;
  jsum[1][t] = 0.0;
  jsum[2][t] = 0.0;
  __syncthreads();
  // direct sums of the row's particles: D[c][slot][i - i0]; reaction sums of
  // the column's particles: R[c][X - band0][js]
  double* D = part + (e.banded ? e.d_off : e.part_off);
  const size_t DS = (size_t)e.segs * e.d_ld;
  const int i0 = e.banded ? e.band0 * kSymBlock : 0;
  double* R = part + e.part_off;
  const size_t RS = (size_t)e.pslots * e.n_ub;
  double fx = 0.0, fy = 0.0, fz = 0.0;
  if (X == Y) {
    const int cnt = min(kSymBlock, n - Y * kSymBlock);
    for (int jj = 0; jj < cnt; ++jj) {
      const int j = Y * kSymBlock + jj;
      const float4 h2 = shi[jj], l2 = slo[jj];
      float dx = (h1.x - h2.x) + (l1.x - l2.x);
      float dy = (h1.y - h2.y) + (l1.y - l2.y);
      float dz = (h1.z - h2.z) + (l1.z - l2.z);
      float qj = (j == i) ? 0.0f : h2.w;
      float dxy2 = __builtin_fmaf(dy, dy, dx * dx);
      float r2 = fmaxf(__builtin_fmaf(dz, dz, dxy2), kR2Min);
      float ri = rsq(r2);
      float r3 = ri * ri * ri;
      float cxy = r3, cz = dz * r3;
      if (ic) {
        float wxy, wz;
        float dzab = (j > i) ? dz : -dz;
        ic_planar32<NEAR>(h1.z + h2.z, l1.w + l2.w, dzab, dxy2, nic, two_d, wxy, wz);
        cxy += wxy;
        cz += wz;
      }
      cxy *= qj;
      cz *= qj;
      fx += (double)(cxy * dx);
      fy += (double)(cxy * dy);
      fz += (double)cz;
    }
    if (i < n) {
      const size_t o = (size_t)X * e.d_ld + (i - i0);
      D[0 * DS + o] = fx;
      D[1 * DS + o] = fy;
      D[2 * DS + o] = fz;
    }
    return;
  }
  // Electrons only (every q/q_0 of both blocks exactly -1, uniform over the
  // workgroup): a product with -1 is a sign, so each term is formed once for
  // the pair and added to the target's sum and subtracted from the source's
  // (or the reverse) -- the same roundings as multiplying by the charges.
  // The image series' switches (ic; with NEAR, nic_pair in {0, 1}) are
  // compile-time constants of each copy of the loop, so the loop holds no
  // branch.
  const bool uni = __syncthreads_and(h1.w == -1.0f && hs.w == -1.0f);
  auto tiles = [&](auto uni_t, auto ic_t, auto n1_t) {
    constexpr bool UNI = decltype(uni_t)::value;
    constexpr bool IC = decltype(ic_t)::value;
    constexpr bool N1 = decltype(n1_t)::value;  // NEAR: nic_pair == 1
    const int nicv = NEAR ? (N1 ? 1 : 0) : nic;
    for (int ph = 0; ph < kSymWaves; ++ph) {
      const int tb = ((w + ph) & (kSymWaves - 1)) * 64;  // the tile of this wavefront
      int jl = lane;  // the source whose reaction sum this lane holds
      double gx = jsum[0][tb + lane], gy = jsum[1][tb + lane], gz = jsum[2][tb + lane];
      for (int step = 0; step < 64; ++step) {
        const float4 h2 = shi[tb + jl], l2 = slo[tb + jl];
        float dx = (h1.x - h2.x) + (l1.x - l2.x);
        float dy = (h1.y - h2.y) + (l1.y - l2.y);
        float dz = (h1.z - h2.z) + (l1.z - l2.z);
        float dxy2 = __builtin_fmaf(dy, dy, dx * dx);
        float r2 = fmaxf(__builtin_fmaf(dz, dz, dxy2), kR2Min);
        float ri = rsq(r2);
        float r3 = ri * ri * ri;
        float cxy = r3, czi = dz * r3, czj = -czi;
        if (IC) {
          float wxy, wz;
          ic_planar32<NEAR>(h1.z + h2.z, l1.w + l2.w, dz, dxy2, nicv, two_d, wxy, wz);
          cxy += wxy;
          czi += wz;
          czj += wz;
        }
        if (UNI) {
          const double tx = (double)(cxy * dx), ty = (double)(cxy * dy);
          fx -= tx;
          fy -= ty;
          fz -= (double)czi;
          gx += tx;
          gy += ty;
          gz -= (double)czj;
        } else {
          const float ci = cxy * h2.w, cj = cxy * h1.w;
          fx += (double)(ci * dx);
          fy += (double)(ci * dy);
          fz += (double)(czi * h2.w);
          gx -= (double)(cj * dx);
          gy -= (double)(cj * dy);
          gz += (double)(czj * h1.w);
        }
        gx = wave_rotate(gx);
        gy = wave_rotate(gy);
        gz = wave_rotate(gz);
        jl = wave_rotate(jl);
      }
      // 64 rotations: every lane holds its own source's sum again
      jsum[0][tb + lane] = gx;
      jsum[1][tb + lane] = gy;
      jsum[2][tb + lane] = gz;
      __syncthreads();
    }
  };
  auto by_series = [&](auto uni_t) {
    if (!ic)
      tiles(uni_t, std::false_type{}, std::false_type{});
    else if (NEAR && nic >= 1)
      tiles(uni_t, std::true_type{}, std::true_type{});
    else
      tiles(uni_t, std::true_type{}, std::false_type{});
  };
  if (uni)
    by_series(std::true_type{});
  else
    by_series(std::false_type{});
  if (i < n) {
    const size_t o = (size_t)Y * e.d_ld + (i - i0);
    D[0 * DS + o] = fx;
    D[1 * DS + o] = fy;
    D[2 * DS + o] = fz;
  }
  if (js < n) {
    const size_t o = (size_t)(X - e.band0) * e.n_ub + js;
    R[0 * RS + o] = jsum[0][t];
    R[1 * RS + o] = jsum[1][t];
    R[2 * RS + o] = jsum[2][t];
  }
}

// Running sums of a banded pair-symmetric launch over rows [band0, band1):
// for a particle of a later block the slots band0..band1-1 (reaction sums),
// for a particle of the band's blocks every slot from band0 on (reaction
// sums below its own block, then its direct sums), added in slot order to
// the sums of the earlier bands (the first band starts from zero).
__global__ void k_fold_sym(const Ent* tab, const SimState* st, const double* __restrict__ part,
                           double* __restrict__ acc) {
  const Ent e = tab[blockIdx.z];
  if (!e.banded) return;
  const int n = st[e.slot].n;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int B = i / kSymBlock;
  const int x0 = e.band0, x1 = e.band1;
  if (B < x0) return;
  const int nbk = min(e.segs, (n + kSymBlock - 1) / kSymBlock);
  const int last = B >= x1 ? x1 : nbk;
  const double* R = part + e.part_off + i;
  const size_t RS = (size_t)e.pslots * e.n_ub;
  const double* D = part + e.d_off + (i - x0 * kSymBlock);
  const size_t DS = (size_t)e.segs * e.d_ld;
  double* A = acc + e.acc_off;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  if (x0 > 0) {
    sx = A[i];
    sy = A[(size_t)e.n_ub + i];
    sz = A[2 * (size_t)e.n_ub + i];
  }
  for (int s = x0; s < last; ++s) {
    double vx, vy, vz;
    if (s < B) {
      const size_t o = (size_t)(s - x0) * e.n_ub;
      vx = R[o];
      vy = R[RS + o];
      vz = R[2 * RS + o];
    } else {
      const size_t o = (size_t)s * e.d_ld;
      vx = D[o];
      vy = D[DS + o];
      vz = D[2 * DS + o];
    }
    sx = sx + vx;
    sy = sy + vy;
    sz = sz + vz;
  }
  A[i] = sx;
  A[(size_t)e.n_ub + i] = sy;
  A[2 * (size_t)e.n_ub + i] = sz;
}

__global__ void k_accel_finish32(const Ent* tab, const SimDesc* desc, const SimState* st,
                                 const double* __restrict__ part, const double* __restrict__ acc,
                                 const int32_t* __restrict__ SP,
                                 const double* __restrict__ X,
                                 const double* __restrict__ Y, const double* __restrict__ Z,
                                 const double* __restrict__ Q, const double* __restrict__ M,
                                 double* __restrict__ AX, double* __restrict__ AY,
                                 double* __restrict__ AZ) {
  const Ent e = tab[blockIdx.z];
  const int n = st[e.slot].n;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const DevGeom& g = desc[e.slot].g;
  const int nchunks = (n + kChunk - 1) / kChunk;
  const int segs = min(e.segs, (nchunks + e.cps - 1) / e.cps);
  double sx = 0.0, sy = 0.0, sz = 0.0;
  if (e.banded) {
    const double* A = acc + e.acc_off;
    sx = A[i];
    sy = A[(size_t)e.n_ub + i];
    sz = A[2 * (size_t)e.n_ub + i];
  } else {
    sum3_in_order(part + e.part_off + i, (size_t)e.segs * e.n_ub, e.n_ub, segs, sx, sy, sz);
  }
  const int gi = e.base + i;
  double q1 = Q[gi];
  double im = 1.0 / M[gi];
  double pre = g.dfc * q1 * Consts::q_0 * 1.0e18 * im;
  if (g.kind == RD_GEOM_TIP) {
    double ex, ey, ez;
    tip_vacuum64(g, X[gi], Y[gi], Z[gi], ex, ey, ez);
    AX[gi] = pre * sx + q1 * ex * im;
    AY[gi] = pre * sy + q1 * ey * im;
    AZ[gi] = pre * sz + q1 * ez * im;
    return;
  }
  // two-time-step ions (planar only) add onto the acceleration they hold,
  // with the vacuum field only at their update steps
  const bool add = e.tts > 0 && SP[gi] == RD_SPECIES_ION;
  const double ax = pre * sx, ay = pre * sy;
  const double az = (!add || (e.step % e.tts) == 0) ? pre * sz + q1 * g.Ez * im : pre * sz;
  AX[gi] = add ? AX[gi] + ax : ax;
  AY[gi] = add ? AY[gi] + ay : ay;
  AZ[gi] = add ? AZ[gi] + az : az;
}

// Packed sphere images of every particle of the tip entries (mixed).
__global__ void k_pack_images(const Ent* tab, const SimDesc* desc, const SimState* st,
                              const double* __restrict__ X, const double* __restrict__ Y,
                              const double* __restrict__ Z, float4* __restrict__ ihi,
                              float4* __restrict__ ilo) {
  const Ent e = tab[blockIdx.z];
  const int n = st[e.slot].n;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const DevGeom& g = desc[e.slot].g;
  if (g.kind != RD_GEOM_TIP) return;
  const int gi = e.base + i;
  float4 h, l;
  pack_image(X[gi], Y[gi], Z[gi], g.z0, g.R, h, l);
  ihi[gi] = h;
  ilo[gi] = l;
}

// Mixed accelerations (tip).  Same layout as the planar kernel; a source
// with j > i uses the target's sphere image, j < i the source's (the lower
// index owns the image, as in Calculate_Acceleration_Particles_ACC).
__global__ __launch_bounds__(kBlock) void k_accel_tip32(const Ent* tab, int ne, const SimDesc* desc,
                                                        const SimState* st,
                                                        const float4* __restrict__ HI,
                                                        const float4* __restrict__ LO,
                                                        const float4* __restrict__ IHI,
                                                        const float4* __restrict__ ILO,
                                                        double* __restrict__ part) {
  __shared__ float4 shi[kChunk], slo[kChunk], sih[kChunk], sil[kChunk];
  const Tile tl = tile_of(tab, ne, false);
  const Ent& e = tl.e;
  const int n = st[e.slot].n;
  if (tl.bx * kBlock >= n) return;
  const int nchunks = (n + kChunk - 1) / kChunk;
  const int c0 = tl.by * e.cps;
  if (c0 >= nchunks) return;
  const int c1 = min(nchunks, c0 + e.cps);
  const DevGeom& gg = desc[e.slot].g;
  if (gg.kind != RD_GEOM_TIP) return;
  const float kt = gg.ic ? (float)gg.ktip : 0.0f;
  const float Rn = (float)(gg.R * 1e9);
  const float4* hi = HI + e.base;
  const float4* lo = LO + e.base;
  const float4* ih = IHI + e.base;
  const float4* il = ILO + e.base;
  const int t = threadIdx.x;
  const int i = tl.bx * kBlock + t;
  const bool act = i < n;
  float4 h1 = make_float4(0, 0, 0, 0), l1 = h1, ih1 = make_float4(0, 0, 0, 1), il1 = h1;
  if (act) {
    h1 = hi[i];
    l1 = lo[i];
    ih1 = ih[i];
    il1 = il[i];
  }
  double fx = 0.0, fy = 0.0, fz = 0.0;
  for (int c = c0; c < c1; ++c) {
    const int jb = c * kChunk;
    __syncthreads();
    if (jb + t < n) {
      shi[t] = hi[jb + t];
      slo[t] = lo[jb + t];
      sih[t] = ih[jb + t];
      sil[t] = il[jb + t];
    }
    __syncthreads();
    const int cnt = min(kChunk, n - jb);
    if (act) {
      for (int jj = 0; jj < cnt; ++jj) {
        const int j = jb + jj;
        float4 h2 = shi[jj];
        if (j == i) h2.w = 0.0f;
        float cx, cy, cz;
        if (j > i)
          tip_pair32_target(h1, l1, ih1, il1, h2, slo[jj], kt, Rn, cx, cy, cz);
        else
          tip_pair32_source(h1, l1, h2, slo[jj], sih[jj], sil[jj], kt, Rn, cx, cy, cz);
        fx += (double)cx;
        fy += (double)cy;
        fz += (double)cz;
      }
    }
  }
  if (act) {
    double* P = part + e.part_off;
    const size_t S = (size_t)e.segs * e.n_ub;
    P[0 * S + (size_t)tl.by * e.n_ub + i] = fx;
    P[1 * S + (size_t)tl.by * e.n_ub + i] = fy;
    P[2 * S + (size_t)tl.by * e.n_ub + i] = fz;
  }
}

// Mixed field at points (tip): the image of every point is its own.
__global__ __launch_bounds__(kBlock) void k_field_tip32(const Ent* tab, int ne, const SimDesc* desc,
                                                        const SimState* st,
                                                        const double* __restrict__ ptx,
                                                        const double* __restrict__ pty,
                                                        const double* __restrict__ ptz,
                                                        const float4* __restrict__ HI,
                                                        const float4* __restrict__ LO,
                                                        double* __restrict__ part) {
  __shared__ float4 shi[kChunk], slo[kChunk];
  const Tile tl = tile_of(tab, ne, true);
  const Ent& e = tl.e;
  const int m = e.m;
  if (tl.bx * kBlock >= m) return;
  const int n = st[e.slot].n;
  const int nchunks = (n + kChunk - 1) / kChunk;
  const int c0 = tl.by * e.cps;
  if (c0 >= max(nchunks, 1)) return;
  const int c1 = min(nchunks, c0 + e.cps);
  const DevGeom& g = desc[e.slot].g;
  if (g.kind != RD_GEOM_TIP) return;
  const int p = tl.bx * kBlock + threadIdx.x;
  const bool act = p < m;
  float4 h1 = make_float4(0, 0, 0, 0), l1 = h1, ih1 = make_float4(0, 0, 0, 1), il1 = h1;
  if (act) {
    const double x = ptx[e.pbase + p], y = pty[e.pbase + p], z = ptz[e.pbase + p];
    pack1(x, y, z, 0.0, g.d, h1, l1);
    pack_image(x, y, z, g.z0, g.R, ih1, il1);
  }
  double fx, fy, fz;
  field_seg32_tip(g, act, h1, l1, ih1, il1, HI + e.base, LO + e.base, n, c0, c1, shi, slo, fx,
                  fy, fz);
  if (act) {
    double* P = part + e.part_off;
    const size_t S = (size_t)e.segs * m;
    P[0 * S + (size_t)tl.by * m + p] = fx;
    P[1 * S + (size_t)tl.by * m + p] = fy;
    P[2 * S + (size_t)tl.by * m + p] = fz;
  }
}

// Mixed field at points (planar).
__global__ __launch_bounds__(kBlock) void k_field_planar32(const Ent* tab, int ne, const SimDesc* desc,
                                                           const SimState* st,
                                                           const float4* __restrict__ PHI,
                                                           const float4* __restrict__ PLO,
                                                           const float4* __restrict__ HI,
                                                           const float4* __restrict__ LO,
                                                           const float4* __restrict__ FAR,
                                                           double* __restrict__ part) {
  __shared__ float4 shi[kChunk], slo[kChunk], sfar[2 * kChunk];
  const Tile tl = tile_of(tab, ne, true);
  const Ent& e = tl.e;
  const int m = e.m;
  if (tl.bx * kBlock >= m) return;
  const int n = st[e.slot].n;
  const int nchunks = (n + kChunk - 1) / kChunk;
  const int c0 = tl.by * e.cps;
  if (c0 >= max(nchunks, 1)) return;
  const int c1 = min(nchunks, c0 + e.cps);
  const DevGeom& g = desc[e.slot].g;
  if (g.kind != RD_GEOM_PLANAR) return;
  const int p = tl.bx * kBlock + threadIdx.x;
  const bool act = p < m;
  float4 h1 = make_float4(0, 0, 0, 0), l1 = make_float4(0, 0, 0, 0);
  if (act) {
    h1 = PHI[e.pbase + p];
    l1 = PLO[e.pbase + p];
  }
  double fx, fy, fz;
  field_seg32(g, act, h1, l1, HI + e.base, LO + e.base, FAR ? FAR + 2 * (size_t)e.base : nullptr,
              n, c0, c1, shi, slo, sfar, fx, fy, fz);
  if (act) {
    double* P = part + e.part_off;
    const size_t S = (size_t)e.segs * m;
    P[0 * S + (size_t)tl.by * m + p] = fx;
    P[1 * S + (size_t)tl.by * m + p] = fy;
    P[2 * S + (size_t)tl.by * m + p] = fz;
  }
}

__global__ void k_field_finish32(const Ent* tab, const SimDesc* desc, const SimState* st,
                                 const double* __restrict__ part, const double* __restrict__ ptx,
                                 const double* __restrict__ pty, const double* __restrict__ ptz,
                                 double* __restrict__ fx, double* __restrict__ fy,
                                 double* __restrict__ fz) {
  const Ent e = tab[blockIdx.z];
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= e.m) return;
  const int n = st[e.slot].n;
  const int nchunks = (n + kChunk - 1) / kChunk;
  const int segs = nchunks == 0 ? 1 : min(e.segs, (nchunks + e.cps - 1) / e.cps);
  const int gp = e.pbase + p;
  const DevGeom& g = desc[e.slot].g;
  if (g.kind == RD_GEOM_TIP)
    field_finish32_tip(g, part + e.part_off, (size_t)e.segs * e.m, e.m, p, segs, ptx[gp],
                       pty[gp], ptz[gp], fx[gp], fy[gp], fz[gp]);
  else
    field_finish32(g, part + e.part_off, (size_t)e.segs * e.m, e.m, p, segs, fx[gp], fy[gp],
                   fz[gp]);
}

__global__ void k_grid_points(int nx, int ny, double x0, double y0, double lx, double ly,
                              double zz, double* __restrict__ ptx, double* __restrict__ pty,
                              double* __restrict__ ptz) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nx * ny) return;
  int ix = p % nx, iy = p / nx;
  ptx[p] = x0 + (ix + 0.5) * (lx / nx);
  pty[p] = y0 + (iy + 0.5) * (ly / ny);
  ptz[p] = zz;
}

// ------------------------------------------------------------------------
// Launch helpers.

// Segments of a simulation's source range: at most max_segs segments of
// whole chunks.  The split depends on the simulation's own size only, so the
// order of every mixed-precision sum, and with it every result, is the same
// whether the simulation runs alone or in a pool.
void segment(int nchunks, int max_segs, int& segs, int& cps) {
  const int nc = std::max(nchunks, 1);
  cps = (nc + max_segs - 1) / max_segs;
  segs = (nc + cps - 1) / cps;
}

// Launch tables travel through a ring of pinned host slots and device slots.
// A ring slot is rewritten only after the stream has been synchronised since
// its last use (the host waits when the ring is exhausted), and a table equal
// to the previous upload reuses it, so the Metropolis-Hastings loop, which
// launches the same table for every jump, uploads it once.
int upload_tab(rd_ctx* c, const std::vector<Ent>& ents, Ent** dtab) {
  const int kRing = 256;
  const int per = 2 * c->nslots_max;  // the supply's first two levels share a table
  if (ents.size() > (size_t)per) return fail(RD_ERR_ARG, "launch table larger than the pool");
  int r = ensure(c, c->tab, sizeof(Ent) * (size_t)per * kRing);
  if (r) return r;
  if (!c->h_tab) {
    RD_HIP(hipHostMalloc((void**)&c->h_tab, sizeof(Ent) * (size_t)per * kRing));
    c->tab_used = 0;
    c->tab_last = -1;
  }
  if (c->tab_last >= 0 && c->tab_last_n == ents.size() &&
      std::memcmp(c->h_tab + (size_t)c->tab_last * per, ents.data(), sizeof(Ent) * ents.size()) ==
          0) {
    *dtab = (Ent*)c->tab.p + (size_t)c->tab_last * per;
    return RD_OK;
  }
  if (c->tab_used >= kRing) {
    RD_HIP(hipStreamSynchronize(c->stream));
    c->tab_used = 0;
  }
  int k = c->tab_used++;
  Ent* h = c->h_tab + (size_t)k * per;
  Ent* d = (Ent*)c->tab.p + (size_t)k * per;
  std::memcpy(h, ents.data(), sizeof(Ent) * ents.size());
  RD_HIP(hipMemcpyAsync(d, h, sizeof(Ent) * ents.size(), hipMemcpyHostToDevice, c->stream));
  c->tab_last = k;
  c->tab_last_n = ents.size();
  *dtab = d;
  return RD_OK;
}

// Whether a simulation's mixed planar accelerations take the pair-symmetric
// form: from its own particle bound alone, so that its result does not depend
// on the other simulations of a pool.  Below kSymMin particles the gather
// form, whose source segments spread one simulation over more workgroups,
// finishes sooner.
static bool sym_form(const rd_ctx* c, const Ent& e) {
  return c->slots[e.slot].desc.g.kind == RD_GEOM_PLANAR && e.n_ub >= kSymMin && e.tts == 0 &&
         !(c->force_opts & RD_FORCE_NO_SYMMETRIC);
}

// FP64 pair launches in bands: the fewest bands B for which one band's
// partials, ceil(nchunks / B) source chunks of every entry's targets, fit
// rd_ctx::part_budget.  Band b takes chunks [b K, (b + 1) K) of an entry
// (K = ceil(nchunks / B)), split into segments for the grid; with B > 1,
// k_fold64 adds each band's partials to the running sums in chunk order.
// points: targets are the entries' field points (k_field64), else their
// particles (k_accel64).  *dlast receives the last band's table, which the
// finish kernel reads (banded, acc_off, part_off, pslots).
static int launch_fp64_bands(rd_ctx* c, std::vector<Ent>& ents, bool points, hipEvent_t k0,
                             bool any_planar, bool any_tip, Ent** dlast, int mtot = 0) {
  auto targets = [&](const Ent& e) { return points ? e.m : e.n_ub; };
  int maxc = 1, maxt = 0;
  long long total = 0;
  for (const Ent& e : ents) {
    maxc = std::max(maxc, e.nchunks_ub);
    maxt = std::max(maxt, targets(e));
    total += 24LL * std::max(e.nchunks_ub, 1) * targets(e);
  }
  int B = (int)std::max(1LL, std::min<long long>(maxc, total / (long long)c->part_budget));
  for (;; ++B) {
    long long bytes = 0;
    for (const Ent& e : ents)
      bytes += 24LL * ((std::max(e.nchunks_ub, 1) + B - 1) / B) * targets(e);
    if ((size_t)bytes <= c->part_budget || B >= maxc) break;
  }
  long long off = 0, acc_n = 0;
  for (Ent& e : ents) {
    const int K = (std::max(e.nchunks_ub, 1) + B - 1) / B;
    e.pslots = K;
    e.banded = B > 1;
    e.part_off = off;
    off += 3LL * K * targets(e);
    e.acc_off = acc_n;
    if (B > 1) acc_n += 3LL * targets(e);
  }
  int r = ensure(c, c->part, sizeof(double) * (size_t)std::max(off, 1LL));
  if (!r && acc_n > 0) r = ensure(c, c->acc, sizeof(double) * (size_t)acc_n);
  if (r) return r;
  double* part = (double*)c->part.p;
  const int ne = (int)ents.size();
  dim3 fin((maxt + 255) / 256, 1, ne);
  double* ptx = (double*)c->pts.p;
  if (k0) RD_HIP(hipEventRecord(k0, c->stream));
  for (int b = 0; b < B; ++b) {
    int nwg = 0;
    for (Ent& e : ents) {
      const int nch = std::max(e.nchunks_ub, 1);
      e.band0 = std::min(nch, b * e.pslots);
      e.band1 = std::min(nch, (b + 1) * e.pslots);
      e.wg0 = nwg;
      e.segs = 0;
      if (e.band1 > e.band0 && targets(e) > 0) {
        const int gx = std::max(1, (targets(e) + kBlock - 1) / kBlock);
        segment(e.band1 - e.band0, seg_limit(gx, kWantPair), e.segs, e.cps);
        nwg += gx * e.segs;
      }
    }
    r = upload_tab(c, ents, dlast);
    if (r) return r;
    if (nwg > 0) {
      if (points) {
        if (any_planar)
          k_field64<RD_GEOM_PLANAR><<<nwg, kBlock, 0, c->stream>>>(
              *dlast, ne, c->d_desc, c->d_state, ptx, ptx + mtot, ptx + 2 * (size_t)mtot, c->x,
              c->y, c->z, c->q, part);
        if (any_tip)
          k_field64<RD_GEOM_TIP><<<nwg, kBlock, 0, c->stream>>>(
              *dlast, ne, c->d_desc, c->d_state, ptx, ptx + mtot, ptx + 2 * (size_t)mtot, c->x,
              c->y, c->z, c->q, part);
      } else {
        if (any_planar)
          k_accel64<RD_GEOM_PLANAR><<<nwg, kBlock, 0, c->stream>>>(
              *dlast, ne, c->d_desc, c->d_state, c->x, c->y, c->z, c->q, c->species,
              (const double*)c->img64.p, part);
        if (any_tip)
          k_accel64<RD_GEOM_TIP><<<nwg, kBlock, 0, c->stream>>>(
              *dlast, ne, c->d_desc, c->d_state, c->x, c->y, c->z, c->q, c->species,
              (const double*)c->img64.p, part);
      }
    }
    if (B > 1) k_fold64<<<fin, 256, 0, c->stream>>>(*dlast, c->d_state, part, (double*)c->acc.p, points);
  }
  RD_HIP(hipGetLastError());
  return RD_OK;
}

int launch_accel(rd_ctx* c, const std::vector<Ent>& in, int precision, hipEvent_t k0,
                 hipEvent_t k1) {
  std::vector<Ent> ents;
  int maxn = 0;
  bool any_tip = false, any_planar = false;
  for (const Ent& e : in)
    if (e.n_ub > 0) {
      ents.push_back(e);
      maxn = std::max(maxn, e.n_ub);
      if (c->slots[e.slot].desc.g.kind == RD_GEOM_TIP)
        any_tip = true;
      else
        any_planar = true;
    }
  if (ents.empty()) return RD_OK;
  if (precision != RD_PREC_FP64 && precision != RD_PREC_MIXED)
    return fail(RD_ERR_ARG, "unknown precision");
  if (precision == RD_PREC_MIXED)
    for (const Ent& e : ents)
      if (e.tts > 0 && c->slots[e.slot].desc.g.kind == RD_GEOM_TIP)
        return fail(RD_ERR_ARG, "two-time-step mode on the tip runs in FP64 only");
  // far image partners by expansion where it converges (sets nic_pair)
  std::vector<TailUse> tail;
  int r = tail_decide(c, ents, precision, tail);
  if (r) return r;
  for (Ent& e : ents) {
    e.nchunks_ub = (e.n_ub + kChunk - 1) / kChunk;
    e.banded = 0;
    e.acc_off = 0;
    e.d_off = 0;
  }
  dim3 fin((maxn + 255) / 256, 1, ents.size());
  if (precision == RD_PREC_FP64) {
    if (any_tip) {
      r = ensure(c, c->img64, sizeof(double) * 4 * (size_t)c->cap);
      if (r) return r;
      Ent* dt0 = nullptr;
      r = upload_tab(c, ents, &dt0);
      if (r) return r;
      k_tip_images64<<<fin, 256, 0, c->stream>>>(dt0, c->d_desc, c->d_state, c->x, c->y, c->z,
                                                  (double*)c->img64.p);
    }
    Ent* dt = nullptr;
    r = launch_fp64_bands(c, ents, false, k0, any_planar, any_tip, &dt);
    if (r) return r;
    if (k1) RD_HIP(hipEventRecord(k1, c->stream));
    k_accel_finish64<<<fin, 256, 0, c->stream>>>(dt, c->d_desc, c->d_state, (double*)c->part.p,
                                                  (const double*)c->acc.p, c->x, c->y, c->z,
                                                  c->q, c->m, c->species, c->ax, c->ay, c->az);
    RD_HIP(hipGetLastError());
    return RD_OK;
  }
  // Mixed.  Gather-form entries: one partial set each, kept to the finish;
  // pair-symmetric entries: nb = ceil(n_ub / kSymBlock) slots, in B bands of
  // ceil(nb / B) block rows whose partials (reaction and direct) fit the
  // budget, after the gather partials.
  long long off = 0;
  int nwg = 0;  // workgroups of the compact gather launch
  std::vector<int> sym;
  for (size_t k = 0; k < ents.size(); ++k) {
    Ent& e = ents[k];
    const int gx = std::max(1, (e.n_ub + kBlock - 1) / kBlock);
    e.wg0 = nwg;
    e.band0 = 0;
    if (sym_form(c, e)) {
      e.segs = (e.n_ub + kSymBlock - 1) / kSymBlock;
      e.cps = kSymBlock / kChunk;
      e.band1 = e.segs;
      sym.push_back((int)k);
    } else {
      segment(e.nchunks_ub, seg_limit(gx, kWantPair), e.segs, e.cps);
      nwg += gx * e.segs;
      e.band1 = e.segs;
      e.pslots = e.segs;
      e.d_ld = e.n_ub;
      e.part_off = off;
      off += 3LL * e.segs * e.n_ub;
    }
  }
  int B = 1;
  if (!sym.empty()) {
    auto band_bytes = [&](int b) {
      long long bytes = 0;
      for (int k : sym) {
        const Ent& e = ents[k];
        const long long K = (e.segs + b - 1) / b;
        bytes += 24LL * K * (e.n_ub + (b > 1 ? (long long)e.segs * kSymBlock : 0));
      }
      return bytes;
    };
    int maxnb = 1;
    for (int k : sym) maxnb = std::max(maxnb, ents[k].segs);
    while (B < maxnb && (size_t)band_bytes(B) > c->part_budget) ++B;
  }
  long long acc_n = 0;
  for (int k : sym) {
    Ent& e = ents[k];
    const int K = (e.segs + B - 1) / B;
    e.pslots = K;
    e.part_off = off;
    off += 3LL * K * e.n_ub;
    if (B > 1) {
      e.banded = 1;
      e.d_ld = K * kSymBlock;
      e.d_off = off;
      off += 3LL * e.segs * e.d_ld;
      e.acc_off = acc_n;
      acc_n += 3LL * e.n_ub;
    } else {
      e.d_ld = e.n_ub;
    }
  }
  r = ensure(c, c->part, sizeof(double) * (size_t)std::max(off, 1LL));
  if (!r && acc_n > 0) r = ensure(c, c->acc, sizeof(double) * (size_t)acc_n);
  if (r) return r;
  double* part = (double*)c->part.p;
  // the table of band b: the symmetric entries' rows and workgroup offsets
  auto band_table = [&](int b, int& nsym) {
    nsym = 0;
    std::vector<Ent> t = ents;
    for (Ent& e : t) {
      e.wg1 = nsym;
      if (!sym_form(c, e)) continue;
      const int K = e.pslots;
      e.band0 = std::min(e.segs, b * K);
      e.band1 = std::min(e.segs, (b + 1) * K);
      for (int X = e.band0; X < e.band1; ++X) nsym += e.segs - X;
    }
    return t;
  };
  int nsym = 0;
  std::vector<Ent> t0 = band_table(0, nsym);
  Ent* dt = nullptr;
  r = upload_tab(c, t0, &dt);
  if (r) return r;
  const int ne = (int)ents.size();
  const int grid = nwg;
  r = ensure(c, c->pack_hi, sizeof(float4) * (size_t)c->cap);
  if (!r) r = ensure(c, c->pack_lo, sizeof(float4) * (size_t)c->cap);
  if (r) return r;
  k_pack_particles<<<fin, 256, 0, c->stream>>>(dt, c->d_desc, c->d_state, c->x, c->y, c->z,
                                                c->q, (float4*)c->pack_hi.p,
                                                (float4*)c->pack_lo.p, nullptr);
  if (any_tip) {
    r = ensure(c, c->img_hi, sizeof(float4) * (size_t)c->cap);
    if (!r) r = ensure(c, c->img_lo, sizeof(float4) * (size_t)c->cap);
    if (r) return r;
    k_pack_images<<<fin, 256, 0, c->stream>>>(dt, c->d_desc, c->d_state, c->x, c->y, c->z,
                                               (float4*)c->img_hi.p, (float4*)c->img_lo.p);
  }
  // every entry within partners n <= 1: the pair kernels without the series loop
  bool near = true, tts = false;
  for (const Ent& e : ents) {
    near &= e.nic_pair <= 1;
    tts |= e.tts > 0;
  }
  if (k0) RD_HIP(hipEventRecord(k0, c->stream));
  if (any_planar && grid > 0) {
    const float4* ph = (const float4*)c->pack_hi.p;
    const float4* pl = (const float4*)c->pack_lo.p;
    if (tts) {
      if (near)
        k_accel_planar32<true, true><<<grid, kBlock, 0, c->stream>>>(
            dt, ne, c->d_desc, c->d_state, ph, pl, c->species, part);
      else
        k_accel_planar32<false, true><<<grid, kBlock, 0, c->stream>>>(
            dt, ne, c->d_desc, c->d_state, ph, pl, c->species, part);
    } else if (near) {
      k_accel_planar32<true, false><<<grid, kBlock, 0, c->stream>>>(
          dt, ne, c->d_desc, c->d_state, ph, pl, nullptr, part);
    } else {
      k_accel_planar32<false, false><<<grid, kBlock, 0, c->stream>>>(
          dt, ne, c->d_desc, c->d_state, ph, pl, nullptr, part);
    }
  }
  if (any_tip)
    k_accel_tip32<<<grid, kBlock, 0, c->stream>>>(
        dt, ne, c->d_desc, c->d_state, (const float4*)c->pack_hi.p, (const float4*)c->pack_lo.p,
        (const float4*)c->img_hi.p, (const float4*)c->img_lo.p, part);
  for (int b = 0; b < B; ++b) {
    if (b > 0) {
      std::vector<Ent> tb = band_table(b, nsym);
      r = upload_tab(c, tb, &dt);
      if (r) return r;
    }
    if (nsym > 0) {
      if (near)
        k_accel_sym32<true><<<nsym, kSymBlock, 0, c->stream>>>(
            dt, ne, c->d_desc, c->d_state, (const float4*)c->pack_hi.p,
            (const float4*)c->pack_lo.p, part);
      else
        k_accel_sym32<false><<<nsym, kSymBlock, 0, c->stream>>>(
            dt, ne, c->d_desc, c->d_state, (const float4*)c->pack_hi.p,
            (const float4*)c->pack_lo.p, part);
    }
    if (B > 1)
      k_fold_sym<<<fin, 256, 0, c->stream>>>(dt, c->d_state, part, (double*)c->acc.p);
  }
  if (k1) RD_HIP(hipEventRecord(k1, c->stream));
  k_accel_finish32<<<fin, 256, 0, c->stream>>>(dt, c->d_desc, c->d_state, part,
                                                (const double*)c->acc.p, c->species, c->x, c->y,
                                                c->z, c->q, c->m, c->ax, c->ay, c->az);
  RD_HIP(hipGetLastError());
  if (!tail.empty()) return tail_apply(c, ents, tail);
  return RD_OK;
}

int launch_field(rd_ctx* c, const std::vector<Ent>& in, int mtot, int precision, hipEvent_t k0,
                 hipEvent_t k1, FixedBufs* fixed) {
  std::vector<Ent> ents;
  int maxm = 0;
  bool any_tip = false, any_planar = false;
  for (const Ent& e : in)
    if (e.m > 0) {
      ents.push_back(e);
      maxm = std::max(maxm, e.m);
      if (c->slots[e.slot].desc.g.kind == RD_GEOM_TIP)
        any_tip = true;
      else
        any_planar = true;
    }
  if (ents.empty()) return RD_OK;
  if (precision != RD_PREC_FP64 && precision != RD_PREC_MIXED)
    return fail(RD_ERR_ARG, "unknown precision");
  long long off = 0;
  int maxn = 0, nwg = 0;  // nwg: workgroups of the compact field launch
  for (Ent& e : ents) {
    int nch = (e.n_ub + kChunk - 1) / kChunk;
    const int gx = std::max(1, (e.m + kBlock - 1) / kBlock);
    e.nchunks_ub = nch;
    segment(nch, seg_limit(gx, kWantPair), e.segs, e.cps);
    maxn = std::max(maxn, e.n_ub);
    e.part_off = off;
    e.band0 = 0;
    e.band1 = std::max(nch, 1);
    e.pslots = precision == RD_PREC_FP64 ? std::max(nch, 1) : e.segs;
    e.banded = 0;
    e.acc_off = 0;
    off += 3LL * e.pslots * e.m;
    e.wg0 = nwg;
    nwg += gx * e.segs;
  }
  dim3 fin((maxm + 255) / 256, 1, ents.size());
  if (precision == RD_PREC_FP64 && !fixed) {
    // FP64 in bands of bounded partials (the fixed buffers take one band)
    double* ptx = (double*)c->pts.p;
    double* fx = (double*)c->fld.p;
    Ent* dt = nullptr;
    int r = launch_fp64_bands(c, ents, true, k0, any_planar, any_tip, &dt, mtot);
    if (r) return r;
    if (k1) RD_HIP(hipEventRecord(k1, c->stream));
    k_field_finish64<<<fin, 256, 0, c->stream>>>(dt, c->d_desc, c->d_state, (double*)c->part.p,
                                                  (const double*)c->acc.p, ptx, ptx + mtot,
                                                  ptx + 2 * (size_t)mtot, fx, fx + mtot,
                                                  fx + 2 * (size_t)mtot);
    RD_HIP(hipGetLastError());
    return RD_OK;
  }
  const size_t part_bytes = sizeof(double) * (size_t)std::max(off, 1LL);
  double* part = nullptr;
  Ent* dt = nullptr;
  int r;
  if (fixed) {
    part = (double*)fixed->take(part_bytes);
    dt = (Ent*)fixed->take(sizeof(Ent) * ents.size());
    if (!part || !dt) return fail(RD_ERR_CAPACITY, "field launch exceeds the fixed buffer");
    RD_HIP(hipMemcpyAsync(dt, ents.data(), sizeof(Ent) * ents.size(), hipMemcpyHostToDevice,
                          c->stream));
  } else {
    r = ensure(c, c->part, part_bytes);
    if (r) return r;
    part = (double*)c->part.p;
    r = upload_tab(c, ents, &dt);
    if (r) return r;
  }
  const int ne = (int)ents.size();
  double* ptx = fixed ? fixed->pts : (double*)c->pts.p;
  double* pty = ptx + mtot;
  double* ptz = pty + mtot;
  double* fx = fixed ? fixed->fld : (double*)c->fld.p;
  double* fy = fx + mtot;
  double* fz = fy + mtot;
  const int grid = nwg;
  if (precision == RD_PREC_FP64) {
    if (k0) RD_HIP(hipEventRecord(k0, c->stream));
    if (maxn > 0) {
      if (any_planar)
        k_field64<RD_GEOM_PLANAR><<<grid, kBlock, 0, c->stream>>>(
            dt, ne, c->d_desc, c->d_state, ptx, pty, ptz, c->x, c->y, c->z, c->q, part);
      if (any_tip)
        k_field64<RD_GEOM_TIP><<<grid, kBlock, 0, c->stream>>>(
            dt, ne, c->d_desc, c->d_state, ptx, pty, ptz, c->x, c->y, c->z, c->q, part);
    }
    if (k1) RD_HIP(hipEventRecord(k1, c->stream));
    k_field_finish64<<<fin, 256, 0, c->stream>>>(dt, c->d_desc, c->d_state, part, nullptr, ptx,
                                                  pty, ptz, fx, fy, fz);
    RD_HIP(hipGetLastError());
    return RD_OK;
  }
  // far-partner series of the sources, for planar entries with image partners
  // beyond n = 0 (used where a wavefront's points lie on the cathode)
  bool any_far = false;
  for (const Ent& e : ents) {
    const DevGeom& g = c->slots[e.slot].desc.g;
    any_far |= g.kind == RD_GEOM_PLANAR && g.ic && g.nic >= 1;
  }
  // Packed copies.  The kernels index the particle copies by store index
  // (base + i); fixed buffers hold only the single entry's particles, so
  // their pointers are shifted down by that entry's base.
  float4 *phi = nullptr, *plo = nullptr, *ahi = nullptr, *alo = nullptr, *afar = nullptr;
  if (fixed) {
    if (ne != 1) return fail(RD_ERR_ARG, "fixed buffers serve one simulation");
    const size_t np = (size_t)std::max(maxn, 1);
    phi = (float4*)fixed->take(sizeof(float4) * (size_t)mtot);
    plo = (float4*)fixed->take(sizeof(float4) * (size_t)mtot);
    float4* h = (float4*)fixed->take(sizeof(float4) * np);
    float4* l = (float4*)fixed->take(sizeof(float4) * np);
    float4* f = any_far ? (float4*)fixed->take(2 * sizeof(float4) * np) : h;
    if (!phi || !plo || !h || !l || !f)
      return fail(RD_ERR_CAPACITY, "field launch exceeds the fixed buffer");
    const uintptr_t shift = sizeof(float4) * (uintptr_t)ents[0].base;
    ahi = (float4*)((uintptr_t)h - shift);
    alo = (float4*)((uintptr_t)l - shift);
    afar = (float4*)((uintptr_t)f - 2 * shift);
  } else {
    r = ensure(c, c->pts_hi, sizeof(float4) * (size_t)mtot);
    if (!r) r = ensure(c, c->pts_lo, sizeof(float4) * (size_t)mtot);
    if (!r) r = ensure(c, c->pack_hi, sizeof(float4) * (size_t)c->cap);
    if (!r) r = ensure(c, c->pack_lo, sizeof(float4) * (size_t)c->cap);
    if (!r && any_far && maxn > 0) r = ensure(c, c->pack_far, 2 * sizeof(float4) * (size_t)c->cap);
    if (r) return r;
    phi = (float4*)c->pts_hi.p;
    plo = (float4*)c->pts_lo.p;
    ahi = (float4*)c->pack_hi.p;
    alo = (float4*)c->pack_lo.p;
    afar = (float4*)c->pack_far.p;
  }
  k_pack_points<<<fin, 256, 0, c->stream>>>(dt, c->d_desc, ptx, pty, ptz, phi, plo);
  if (maxn > 0)
    k_pack_particles<<<dim3((maxn + 255) / 256, 1, ne), 256, 0, c->stream>>>(
        dt, c->d_desc, c->d_state, c->x, c->y, c->z, c->q, ahi, alo, nullptr);
  if (any_far && maxn > 0)
    k_pack_far<<<dim3((maxn + 255) / 256, 1, ne), 256, 0, c->stream>>>(dt, c->d_desc, c->d_state,
                                                                         c->z, afar);
  if (k0) RD_HIP(hipEventRecord(k0, c->stream));
  if (any_planar)
    k_field_planar32<<<grid, kBlock, 0, c->stream>>>(dt, ne, c->d_desc, c->d_state, phi, plo,
                                                      ahi, alo, any_far ? afar : nullptr, part);
  if (any_tip)
    k_field_tip32<<<grid, kBlock, 0, c->stream>>>(dt, ne, c->d_desc, c->d_state, ptx, pty, ptz,
                                                   ahi, alo, part);
  if (k1) RD_HIP(hipEventRecord(k1, c->stream));
  k_field_finish32<<<fin, 256, 0, c->stream>>>(dt, c->d_desc, c->d_state, part, ptx, pty, ptz,
                                                fx, fy, fz);
  RD_HIP(hipGetLastError());
  return RD_OK;
}

int launch_pack(rd_ctx* c, const std::vector<Ent>& ents, bool surface) {
  int maxn = 0;
  for (const Ent& e : ents) maxn = std::max(maxn, e.n_ub);
  if (ents.empty() || maxn == 0) return RD_OK;
  int r = ensure(c, c->pack_hi, sizeof(float4) * (size_t)c->cap);
  if (!r) r = ensure(c, c->pack_lo, sizeof(float4) * (size_t)c->cap);
  if (!r && surface) r = ensure(c, c->pack_srf, sizeof(float4) * (size_t)c->cap);
  if (r) return r;
  Ent* dt = nullptr;
  r = upload_tab(c, ents, &dt);
  if (r) return r;
  k_pack_particles<<<dim3((maxn + 255) / 256, 1, ents.size()), 256, 0, c->stream>>>(
      dt, c->d_desc, c->d_state, c->x, c->y, c->z, c->q, (float4*)c->pack_hi.p,
      (float4*)c->pack_lo.p, surface ? (float4*)c->pack_srf.p : nullptr);
  RD_HIP(hipGetLastError());
  return RD_OK;
}

int launch_grid_points(rd_ctx* c, double x0, double y0, double lx, double ly, int nx, int ny,
                       double zz) {
  int m = nx * ny;
  int r = ensure(c, c->pts, sizeof(double) * 3 * (size_t)m);
  if (!r) r = ensure(c, c->fld, sizeof(double) * 3 * (size_t)m);
  if (r) return r;
  double* ptx = (double*)c->pts.p;
  k_grid_points<<<(m + 255) / 256, 256, 0, c->stream>>>(nx, ny, x0, y0, lx, ly, zz, ptx, ptx + m,
                                                         ptx + 2 * m);
  RD_HIP(hipGetLastError());
  return RD_OK;
}

}  // namespace rd
