// rd_tail.hip -- the far image partners of the planar image series in the
// mixed-precision accelerations, as one FP64 expansion for all particles.
//
// The planar image series of a pair (acc_ic_planar_series.inc:19-63) has,
// for source b, the opposite-charge partners at +-2nd - z_b and the
// same-charge partners at +-2nd + z_b.  Those with n >= 2 ("far": at least
// 3d from any particle in the gap) make, summed over all sources, a smooth
// field in the gap.  With K(D) = sum_{n=2..N} sum_+- E(D -+ 2nd z^), E(D) =
// D/|D|^3, the far part of the acceleration sums of particle k is
//
//   T_k = - sum_j q_j K(x_k - x_j, y_k - y_j, z_k + z_j)   opposite partners
//         + sum_j q_j K(r_k - r_j)                         same-charge partners
//         + q_k K(0, 0, 2 z_k)                             no self image (j != k)
//         - 2 z^ sum_{j<k} q_j K_z(r_k - r_j)              lower-index roles
//
// The last line is the reference's pair convention (src/mod_verlet.F90:
// 736-742, :867-871): the series of the pair (i < j) is evaluated at z_i and
// applied to j with x and y reversed and z kept, which for the same-charge
// partners reverses their z term on the higher-index particle.
//
// K is the field of point charges at +-2nd z^, n >= 2, so both sums are
// Taylor series of 1/|D| about fixed centres (the sources' centre for the
// same-charge images, its mirror for the opposite ones), written as local
// expansions at the particles' centre from the sources' moments
// M_delta = sum_j q_j (-w_j)^delta, w_j = r_j - centre.  The index-ordered sum
// takes the moments of the blocks of 64 particles before a particle's block
// (an ordered prefix) and the earlier particles of its own block directly.
// Everything here is FP64; the pair kernels then run the partners n <= 1
// only.  The truncation at degree kTailDeg leaves about rho^(kTailDeg+1) of
// the far field, rho = 2 R / (3 d) for particles within R of the centre; the
// expansion is used for rho <= kTailRhoMax, otherwise the pair kernels keep
// every partner.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "rd_internal.h"

namespace rd {

constexpr int kTailDeg = 13;
constexpr int kTailTerms = (kTailDeg + 1) * (kTailDeg + 2) * (kTailDeg + 3) / 6;
constexpr int kTailBlk = 64;
constexpr double kTailRhoMax = 0.42;

// Terms in the order a = 0..D, b = 0..D-a, c = 0..D-a-b (x^a y^b z^c).
static void tail_terms(std::vector<int>& a, std::vector<int>& b, std::vector<int>& c) {
  for (int i = 0; i <= kTailDeg; ++i)
    for (int j = 0; i + j <= kTailDeg; ++j)
      for (int k = 0; i + j + k <= kTailDeg; ++k) {
        a.push_back(i);
        b.push_back(j);
        c.push_back(k);
      }
}

static int term_index(int i, int j, int k) {
  // position of (i, j, k) in the order of tail_terms
  int t = 0;
  for (int ii = 0; ii < i; ++ii) {
    const int m = kTailDeg - ii;
    t += (m + 1) * (m + 2) / 2;
  }
  for (int jj = 0; jj < j; ++jj) t += kTailDeg - i - jj + 1;
  return t + k;
}

// Taylor coefficients a_t(x) of 1/|x - v| = sum_t a_t(x) v^t, |t| <= kTailDeg,
// by the recurrence |t| |x|^2 a_t = (2|t| - 1) sum_i x_i a_{t-e_i}
// - (|t| - 1) sum_i a_{t-2e_i} (Duan and Krasny's treecode coefficients).
static void inv_r_taylor(const double x[3], const std::vector<int>& A, const std::vector<int>& B,
                         const std::vector<int>& C, std::vector<double>& out) {
  const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
  out.assign(A.size(), 0.0);
  // terms are visited in an order where every t - e_i and t - 2 e_i comes
  // first: by total degree
  for (int deg = 0; deg <= kTailDeg; ++deg)
    for (size_t t = 0; t < A.size(); ++t) {
      const int k[3] = {A[t], B[t], C[t]};
      if (k[0] + k[1] + k[2] != deg) continue;
      if (deg == 0) {
        out[t] = 1.0 / std::sqrt(r2);
        continue;
      }
      double s1 = 0.0, s2 = 0.0;
      for (int d = 0; d < 3; ++d) {
        int m1[3] = {k[0], k[1], k[2]};
        m1[d] -= 1;
        if (m1[d] >= 0) s1 += x[d] * out[term_index(m1[0], m1[1], m1[2])];
        int m2[3] = {k[0], k[1], k[2]};
        m2[d] -= 2;
        if (m2[d] >= 0) s2 += out[term_index(m2[0], m2[1], m2[2])];
      }
      out[t] = ((2.0 * deg - 1.0) * s1 - (deg - 1.0) * s2) / (deg * r2);
    }
}

static double binom(int n, int k) {
  double r = 1.0;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

// Host plan of one (d, N_ic_max): the terms and, per term beta, the list of
// (delta, v0, v1) with L_beta = sum (v0 - v1) M_delta for the field of all
// far images and sum v0 M_delta for the same-charge ones; lengths in nm.
static int build_plan(rd_ctx* c, double d_m, int nic) {
  if (c->tail_d == d_m && c->tail_nic == nic && c->tail_plan.p) return RD_OK;
  std::vector<int> A, B, C;
  tail_terms(A, B, C);
  const int T = (int)A.size();
  const double dnm = d_m * 1.0e9;
  std::vector<double> b0(T, 0.0), b1(T, 0.0), tmp;
  // centre of the same-charge kernel: 0; of the opposite one: (0, 0, d)
  for (int n = 2; n <= nic; ++n)
    for (int s = -1; s <= 1; s += 2) {
      const double x0[3] = {0.0, 0.0, s * 2.0 * n * dnm};
      inv_r_taylor(x0, A, B, C, tmp);
      for (int t = 0; t < T; ++t) b0[t] += tmp[t];
      const double x1[3] = {0.0, 0.0, s * 2.0 * n * dnm - dnm};
      inv_r_taylor(x1, A, B, C, tmp);
      for (int t = 0; t < T; ++t) b1[t] += tmp[t];
    }
  std::vector<int> rowp(T + 1, 0), col;
  std::vector<double> v0, v1;
  for (int bt = 0; bt < T; ++bt) {
    rowp[bt] = (int)col.size();
    for (int dt = 0; dt < T; ++dt) {
      const int g0 = A[bt] + A[dt], g1 = B[bt] + B[dt], g2 = C[bt] + C[dt];
      if (g0 + g1 + g2 > kTailDeg) continue;
      const int gt = term_index(g0, g1, g2);
      const double cb = binom(g0, A[bt]) * binom(g1, B[bt]) * binom(g2, C[bt]);
      col.push_back(dt);
      v0.push_back(cb * b0[gt]);
      v1.push_back(cb * b1[gt] * ((C[dt] & 1) ? -1.0 : 1.0));
    }
  }
  rowp[T] = (int)col.size();
  const int nnz = (int)col.size();
  std::vector<int> packed(T);
  for (int t = 0; t < T; ++t) packed[t] = A[t] | (B[t] << 8) | (C[t] << 16);
  const size_t bytes = sizeof(int) * (size_t)(T + T + 1 + nnz) + sizeof(double) * 2 * (size_t)nnz;
  int r = ensure(c, c->tail_plan, bytes + 64);
  if (r) return r;
  std::vector<char> h(bytes + 64, 0);
  char* p = h.data();
  std::memcpy(p, v0.data(), sizeof(double) * nnz);
  std::memcpy(p + sizeof(double) * nnz, v1.data(), sizeof(double) * nnz);
  char* q = p + sizeof(double) * 2 * nnz;
  std::memcpy(q, packed.data(), sizeof(int) * T);
  std::memcpy(q + sizeof(int) * T, rowp.data(), sizeof(int) * (T + 1));
  std::memcpy(q + sizeof(int) * (2 * T + 1), col.data(), sizeof(int) * nnz);
  RD_HIP(hipMemcpy(c->tail_plan.p, h.data(), bytes, hipMemcpyHostToDevice));
  c->tail_d = d_m;
  c->tail_nic = nic;
  c->tail_nnz = nnz;
  return RD_OK;
}

// Lateral bounding box of one simulation's particles: out = (xmin, xmax,
// ymin, ymax) in m.  One workgroup of 1024 lanes, eight independent loads in
// flight per lane (min and max are exact, so the order does not matter).
__global__ __launch_bounds__(1024) void k_tail_box(const double* __restrict__ X,
                                                    const double* __restrict__ Y, int n,
                                                    double* __restrict__ out) {
  __shared__ double s[4][1024];
  double v[4] = {INFINITY, -INFINITY, INFINITY, -INFINITY};
  constexpr int kU = 4;
  for (int i0 = threadIdx.x; i0 < n; i0 += 1024 * kU) {
    double xs[kU], ys[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int i = i0 + 1024 * u;
      xs[u] = i < n ? X[i] : X[i0];
      ys[u] = i < n ? Y[i] : Y[i0];
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      v[0] = fmin(v[0], xs[u]);
      v[1] = fmax(v[1], xs[u]);
      v[2] = fmin(v[2], ys[u]);
      v[3] = fmax(v[3], ys[u]);
    }
  }
  for (int k = 0; k < 4; ++k) s[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s[0][threadIdx.x] = fmin(s[0][threadIdx.x], s[0][threadIdx.x + w]);
      s[1][threadIdx.x] = fmax(s[1][threadIdx.x], s[1][threadIdx.x + w]);
      s[2][threadIdx.x] = fmin(s[2][threadIdx.x], s[2][threadIdx.x + w]);
      s[3][threadIdx.x] = fmax(s[3][threadIdx.x], s[3][threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) out[threadIdx.x] = s[threadIdx.x][0];
}

// Powers u^0 .. u^kTailDeg of the three coordinates of the block's particles,
// column `lane` of pw (the block's first wavefront fills it).
__device__ __forceinline__ void tail_powers(double (*pw)[kTailDeg + 1][kTailBlk], int lane,
                                            double ux, double uy, double uz) {
  double px = 1.0, py = 1.0, pz = 1.0;
  for (int k = 0; k <= kTailDeg; ++k) {
    pw[0][k][lane] = px;
    pw[1][k][lane] = py;
    pw[2][k][lane] = pz;
    px *= ux;
    py *= uy;
    pz *= uz;
  }
}

// Moments of every block of kTailBlk particles: Mblk[blk][t] =
// sum_j q_j/q_0 (-w_j)^t, w_j = r_j - centre (nm), j in particle order.  The
// block's powers sit in LDS; padding particles carry charge 0.
__global__ __launch_bounds__(256) void k_tail_moments(
    const double* __restrict__ X, const double* __restrict__ Y, const double* __restrict__ Z,
    const double* __restrict__ Q, int n, double cx, double cy, double cz,
    const int* __restrict__ terms, double* __restrict__ Mblk) {
  __shared__ double pw[3][kTailDeg + 1][kTailBlk];
  __shared__ double sq[kTailBlk];
  const int j0 = blockIdx.x * kTailBlk;
  if (threadIdx.x < kTailBlk) {
    const int j = j0 + threadIdx.x;
    double wx = 0.0, wy = 0.0, wz = 0.0, qq = 0.0;
    if (j < n) {
      wx = cx - X[j] * 1.0e9;
      wy = cy - Y[j] * 1.0e9;
      wz = cz - Z[j] * 1.0e9;
      qq = Q[j] / Consts::q_0;
    }
    tail_powers(pw, threadIdx.x, wx, wy, wz);
    sq[threadIdx.x] = qq;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < kTailTerms; t += 256) {
    const int pk = terms[t];
    const int a = pk & 255, b = (pk >> 8) & 255, cc = (pk >> 16) & 255;
    double s = 0.0;
    for (int jj = 0; jj < kTailBlk; ++jj) s += sq[jj] * pw[0][a][jj] * pw[1][b][jj] * pw[2][cc][jj];
    Mblk[(size_t)blockIdx.x * kTailTerms + t] = s;
  }
}

// Ordered prefix over the blocks, two levels: within groups of kScanGroup
// blocks (Mpre[blk] = sum of the group's earlier blocks, Gsum[g] = the
// group's sum) and over the groups (Gpre[g], Mtot); a block's prefix is
// Gpre[g] + Mpre[blk], a fixed order.
constexpr int kScanGroup = 32;
__global__ __launch_bounds__(256) void k_tail_scan_groups(const double* __restrict__ Mblk,
                                                          int nblk, double* __restrict__ Mpre,
                                                          double* __restrict__ Gsum) {
  const int t = blockIdx.y * 256 + threadIdx.x;
  if (t >= kTailTerms) return;
  const int g = blockIdx.x;
  const int b0 = g * kScanGroup, b1 = min(nblk, b0 + kScanGroup);
  double v[kScanGroup];
#pragma unroll
  for (int k = 0; k < kScanGroup; ++k)
    v[k] = b0 + k < b1 ? Mblk[(size_t)(b0 + k) * kTailTerms + t] : 0.0;
  double run = 0.0;
#pragma unroll
  for (int k = 0; k < kScanGroup; ++k) {
    if (b0 + k < b1) Mpre[(size_t)(b0 + k) * kTailTerms + t] = run;
    run += v[k];
  }
  Gsum[(size_t)g * kTailTerms + t] = run;
}
__global__ __launch_bounds__(256) void k_tail_scan_top(const double* __restrict__ Gsum,
                                                       int ngroups, double* __restrict__ Gpre,
                                                       double* __restrict__ Mtot) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= kTailTerms) return;
  double run = 0.0;
  for (int g = 0; g < ngroups; ++g) {
    const double v = Gsum[(size_t)g * kTailTerms + t];
    Gpre[(size_t)g * kTailTerms + t] = run;
    run += v;
  }
  Mtot[t] = run;
}

// Local expansions: Lpre[blk][beta] from the block's prefix moments
// Gpre[blk / kScanGroup] + Mpre[blk] (same-charge kernel), L[beta] from all
// moments (both kernels; column blk = nblk).  A wavefront takes one row beta
// for 64 blocks, one per lane, so the operator's entries (v0, v1, col) are
// the same for every lane of the row.
__global__ __launch_bounds__(256) void k_tail_local(
    const double* __restrict__ v0, const double* __restrict__ v1, const int* __restrict__ rowp,
    const int* __restrict__ col, const double* __restrict__ Mpre, const double* __restrict__ Gpre,
    const double* __restrict__ Mtot, int nblk, double* __restrict__ Lpre, double* __restrict__ L) {
  const int lane = threadIdx.x & 63;
  const int bt = blockIdx.x * 4 + (threadIdx.x >> 6);  // the row of this wavefront
  const int blk = blockIdx.y * 64 + lane;               // the block of this lane
  if (bt >= kTailTerms || blk > nblk) return;
  const int k0 = rowp[bt], k1 = rowp[bt + 1];
  double s = 0.0;
  if (blk < nblk) {
    const double* Mi = Mpre + (size_t)blk * kTailTerms;
    const double* Gi = Gpre + (size_t)(blk / kScanGroup) * kTailTerms;
    for (int k = k0; k < k1; ++k) {
      const int cc = col[k];
      s += v0[k] * (Gi[cc] + Mi[cc]);
    }
    Lpre[(size_t)blk * kTailTerms + bt] = s;
  } else {
    for (int k = k0; k < k1; ++k) s += (v0[k] - v1[k]) * Mtot[col[k]];
    L[bt] = s;
  }
}

// z component of K(D) (nm): the field of unit charges at +-2nd z^, n = 2..nic.
__device__ __forceinline__ double far_kz(double dx, double dy, double dz, double two_d, int nic) {
  const double r2xy = dx * dx + dy * dy;
  double s = 0.0;
  for (int k = 2; k <= nic; ++k) {
    const double tn = k * two_d;
    const double zp = dz - tn, zm = dz + tn;
    const double rp = sqrt(r2xy + zp * zp), rm = sqrt(r2xy + zm * zm);
    s += zp / (rp * rp * rp) + zm / (rm * rm * rm);
  }
  return s;
}

// The far part of every particle's acceleration, added to AX, AY, AZ: one
// block of kTailBlk particles per workgroup of kTailWaves wavefronts; lane l
// of every wavefront takes particle l, wavefront w the terms t = w mod
// kTailWaves and the block's earlier particles j = w mod kTailWaves (the
// reference's index-ordered part), and the wavefronts' sums are added in
// wavefront order.
constexpr int kTailWaves = 8;
__global__ __launch_bounds__(kTailWaves * kTailBlk) void k_tail_apply(
    const double* __restrict__ X, const double* __restrict__ Y, const double* __restrict__ Z,
    const double* __restrict__ Q, const double* __restrict__ Ms, int n, double cx, double cy,
    double cz, double dnm, int nic, double dfc, const int* __restrict__ terms,
    const double* __restrict__ L, const double* __restrict__ Lpre, double* __restrict__ AX,
    double* __restrict__ AY, double* __restrict__ AZ) {
  __shared__ double pw[3][kTailDeg + 1][kTailBlk];
  __shared__ double red[5][kTailWaves][kTailBlk];
  const int lane = threadIdx.x & (kTailBlk - 1);
  const int w = threadIdx.x / kTailBlk;
  const int blk = blockIdx.x;
  const int i = blk * kTailBlk + lane;
  const bool act = i < n;
  double xi = 0.0, yi = 0.0, zi = 0.0;
  if (act) {
    xi = X[i] * 1.0e9;
    yi = Y[i] * 1.0e9;
    zi = Z[i] * 1.0e9;
  }
  if (w == 0) tail_powers(pw, lane, xi - cx, yi - cy, zi - cz);
  __syncthreads();
  // -grad of sum_t L_t u^t, and -d/dz of sum_t Lpre_t u^t
  double gx = 0.0, gy = 0.0, gz = 0.0, qz = 0.0;
  const double* Lp = Lpre + (size_t)blk * kTailTerms;
  for (int t = w; t < kTailTerms; t += kTailWaves) {
    const int pk = terms[t];
    const int a = pk & 255, b = (pk >> 8) & 255, c = (pk >> 16) & 255;
    const double xa = pw[0][a][lane], yb = pw[1][b][lane], zc = pw[2][c][lane];
    const double lt = L[t];
    if (a > 0) gx += lt * a * pw[0][a - 1][lane] * yb * zc;
    if (b > 0) gy += lt * b * xa * pw[1][b - 1][lane] * zc;
    if (c > 0) {
      const double dzt = c * xa * yb * pw[2][c - 1][lane];
      gz += lt * dzt;
      qz += Lp[t] * dzt;
    }
  }
  // earlier particles of this block, directly
  const double two_d = 2.0 * dnm;
  double qd = 0.0;
  if (act)
    for (int j = blk * kTailBlk + w; j < i; j += kTailWaves)
      qd += Q[j] / Consts::q_0 *
            far_kz(xi - X[j] * 1.0e9, yi - Y[j] * 1.0e9, zi - Z[j] * 1.0e9, two_d, nic);
  red[0][w][lane] = gx;
  red[1][w][lane] = gy;
  red[2][w][lane] = gz;
  red[3][w][lane] = qz;
  red[4][w][lane] = qd;
  __syncthreads();
  if (w != 0 || !act) return;
  double sum[5];
  for (int k = 0; k < 5; ++k) {
    double v = red[k][0][lane];
    for (int ww = 1; ww < kTailWaves; ++ww) v += red[k][ww][lane];
    sum[k] = v;
  }
  const double qi = Q[i] / Consts::q_0;
  const double self_z = qi * far_kz(0.0, 0.0, 2.0 * zi, two_d, nic);
  const double tx = -sum[0], ty = -sum[1];
  const double tz = -sum[2] + self_z - 2.0 * (-sum[3] + sum[4]);
  // acceleration: dfc q_i q_0 / m_i times T (in (q/q_0) / nm^2 -> 1e18 / m^2)
  const double pre = dfc * Q[i] * Consts::q_0 * 1.0e18 / Ms[i];
  AX[i] += pre * tx;
  AY[i] += pre * ty;
  AZ[i] += pre * tz;
}

bool tail_wanted(const rd_ctx* c, const DevGeom& g, int precision) {
  return precision == RD_PREC_MIXED && !(c->force_opts & RD_FORCE_NO_FAR_EXPANSION) &&
         g.kind == RD_GEOM_PLANAR && g.ic && g.nic >= 2;
}

// Which of the candidate entries (tail_wanted) take the expansion: their
// particles' lateral box keeps rho <= kTailRhoMax.  Sets e.nic_pair (1 for
// those, the full series otherwise) and the expansion centre.
int tail_decide(rd_ctx* c, std::vector<Ent>& ents, int precision,
                std::vector<TailUse>& use) {
  use.clear();
  std::vector<int> cand;
  for (size_t k = 0; k < ents.size(); ++k) {
    const DevGeom& g = c->slots[ents[k].slot].desc.g;
    ents[k].nic_pair = g.nic;
    if (tail_wanted(c, g, precision) && ents[k].n_ub > 0 && ents[k].tts == 0)
      cand.push_back((int)k);
  }
  if (cand.empty()) return RD_OK;
  // exact particle counts (inside a step the host's may predate the emission)
  int r = sync_state(c);
  if (r) return r;
  r = ensure(c, c->tail_box, sizeof(double) * 4 * cand.size());
  if (r) return r;
  double* box = (double*)c->tail_box.p;
  for (size_t k = 0; k < cand.size(); ++k) {
    const Ent& e = ents[cand[k]];
    const int n = c->slots[e.slot].n;
    if (n < 1) continue;
    k_tail_box<<<1, 1024, 0, c->stream>>>(c->x + e.base, c->y + e.base, n, box + 4 * k);
  }
  RD_HIP(hipGetLastError());
  std::vector<double> h(4 * cand.size(), 0.0);
  RD_HIP(hipMemcpyAsync(h.data(), box, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                        c->stream));
  RD_HIP(hipStreamSynchronize(c->stream));
  for (size_t k = 0; k < cand.size(); ++k) {
    Ent& e = ents[cand[k]];
    const int n = c->slots[e.slot].n;
    if (n < 1) continue;
    const DevGeom& g = c->slots[e.slot].desc.g;
    const double dnm = g.d * 1.0e9;
    const double hx = 0.5 * (h[4 * k + 1] - h[4 * k]) * 1.0e9;
    const double hy = 0.5 * (h[4 * k + 3] - h[4 * k + 2]) * 1.0e9;
    const double R = std::sqrt(hx * hx + hy * hy + 0.25 * dnm * dnm);
    const double rho = 2.0 * R / (3.0 * dnm);
    if (!(rho <= kTailRhoMax)) continue;
    e.nic_pair = 1;
    TailUse u;
    u.ent = cand[k];
    u.cx = 0.5 * (h[4 * k] + h[4 * k + 1]) * 1.0e9;
    u.cy = 0.5 * (h[4 * k + 2] + h[4 * k + 3]) * 1.0e9;
    u.cz = 0.5 * dnm;
    use.push_back(u);
  }
  return RD_OK;
}

// The far part of the accelerations of the entries in `use` (after the pair
// kernels' finish has written AX, AY, AZ).
int tail_apply(rd_ctx* c, const std::vector<Ent>& ents, const std::vector<TailUse>& use) {
  for (const TailUse& u : use) {
    const Ent& e = ents[u.ent];
    const int n = c->slots[e.slot].n;
    const DevGeom& g = c->slots[e.slot].desc.g;
    int r = build_plan(c, g.d, g.nic);
    if (r) return r;
    const int T = kTailTerms;
    const int nnz = c->tail_nnz;
    const double* v0 = (const double*)c->tail_plan.p;
    const double* v1 = v0 + nnz;
    const int* terms = (const int*)(v1 + nnz);
    const int* rowp = terms + T;
    const int* col = rowp + T + 1;
    const int nblk = (n + kTailBlk - 1) / kTailBlk;
    const int ngroups = (nblk + kScanGroup - 1) / kScanGroup;
    r = ensure(c, c->tail_work,
               sizeof(double) * (size_t)T * (3 * (size_t)nblk + 2 * (size_t)ngroups + 2));
    if (r) return r;
    double* Mblk = (double*)c->tail_work.p;
    double* Mpre = Mblk + (size_t)nblk * T;
    double* Lpre = Mpre + (size_t)nblk * T;
    double* Gsum = Lpre + (size_t)nblk * T;
    double* Gpre = Gsum + (size_t)ngroups * T;
    double* Mtot = Gpre + (size_t)ngroups * T;
    double* L = Mtot + T;
    const int tb = (T + 255) / 256;
    const double* x = c->x + e.base;
    const double* y = c->y + e.base;
    const double* z = c->z + e.base;
    const double* q = c->q + e.base;
    k_tail_moments<<<nblk, 256, 0, c->stream>>>(x, y, z, q, n, u.cx, u.cy, u.cz, terms, Mblk);
    k_tail_scan_groups<<<dim3(ngroups, tb), 256, 0, c->stream>>>(Mblk, nblk, Mpre, Gsum);
    k_tail_scan_top<<<tb, 256, 0, c->stream>>>(Gsum, ngroups, Gpre, Mtot);
    k_tail_local<<<dim3((T + 3) / 4, (nblk + 1 + 63) / 64), 256, 0, c->stream>>>(
        v0, v1, rowp, col, Mpre, Gpre, Mtot, nblk, Lpre, L);
    k_tail_apply<<<nblk, kTailWaves * kTailBlk, 0, c->stream>>>(
        x, y, z, q, c->m + e.base, n, u.cx, u.cy, u.cz, g.d * 1.0e9, g.nic, g.dfc, terms, L, Lpre,
        c->ax + e.base, c->ay + e.base, c->az + e.base);
    RD_HIP(hipGetLastError());
  }
  return RD_OK;
}

}  // namespace rd
