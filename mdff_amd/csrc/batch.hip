// batch.hip -- many configurations of uneven atom counts evaluated together
// (force matching over a set of cells), and the adjoint of the force-matching
// loss with respect to the BMHFTD parameters, the charges and the
// polarizabilities.
//
// Atoms of all configurations are concatenated (configuration c owns atoms
// off_c .. off_c + n_c - 1); every launch covers the whole batch.  The
// real-space kernels are those of the single-cell path instantiated with
// BATCH = true (cell, Ewald alpha and cutoffs looked up per atom); the
// reciprocal-space sums are batched here: one workgroup per (configuration,
// block of k vectors) for the structure factors, lanes per atom over the
// configuration's k vectors for the field and force.  In FP32 the
// structure factors and the per-atom k sums are accumulated with
// compensated (Kahan-Babuska) summation; loss and gradient reductions are
// FP64 in either precision.
//
// Gradient of L = sum |F - F_t|^2 / (3 N_tot):
//   g_i = dL/dF_i = 2 (F_i - F_t,i) / (3 N_tot)
//   BMHFTD parameter p:  dL/dp = sum over directed pairs (r_ij . g_i) dw_ij/dp,
//     with F_i = sum_j w_ij r_ij; the pair terms are recomputed (k_b_sr_adj).
//   charge q_a, polarizability alpha_t: the electrostatic force is a
//     quadratic form F(x) of the sources x = (q, mu), and mu is the SCF fixed
//     point, so dF/dtheta = 2 B(x, x') = [F(x + x') - F(x - x')] / 2 with the
//     tangent x' = (dq/dtheta, dmu/dtheta); dmu solves the tangent SCF
//       (1 - alpha T) dmu = alpha E_static(dq)     (charges)
//       (1 - alpha T) dmu = P_t mu / alpha_t       (polarizability of type t)
//     with the SCF's own dipole-field operator T.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/mdffgpu.h"
#include "kernels.h"
#include "mixed.h"
#include "rtypes.h"
#include "tables.h"

struct mdffgpu_batch;

namespace mdff {
namespace batch {

constexpr int BK = 256;        // k vectors per structure-factor workgroup
constexpr int BTA = 32;        // atoms per LDS tile of the structure factor
constexpr int NTAB_MAX = 75;   // phase-table entries per atom (kmax <= 12)
constexpr int NPAIR_MAX = 3;   // unordered type pairs (two species)
constexpr int NSR = 5;         // A, B, C6, C8, BD per pair
constexpr int NMO = 3;         // Morse eps, sigma, rho per pair
constexpr int NDAMP_MAX = 2;   // species with a Tang-Toennies damping parameter b

// ---------------------------------------------------------------------------
// compensated accumulation (Kahan-Babuska / Neumaier) in type R
template <typename R>
struct KSum {
  R s = R(0), c = R(0);
  __device__ __forceinline__ void add(R x) {
    const R t = s + x;
    if (fabs(s) >= fabs(x))
      c += (s - t) + x;
    else
      c += (x - t) + s;
    s = t;
  }
  __device__ __forceinline__ R val() const { return s + c; }
};

template <typename R>
struct C2 {
  R x, y;
};
template <typename R>
__device__ __forceinline__ C2<R> cmul(C2<R> a, C2<R> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// ---------------------------------------------------------------------------
// fractional coordinates and the per-atom phase tables e^{2 pi i m s_a},
// layout [entry][atom]: m1 = 0..KX, then m2 = -KY..KY, then m3 = -KZ..KZ
__global__ __launch_bounds__(MDFF_BLOCK) void k_b_prepare(int N, const double* __restrict__ pos,
                                                           const int* __restrict__ cfg,
                                                           const BatchCfg* __restrict__ bc,
                                                           int KX, int KY, int KZ,
                                                           double* __restrict__ s,
                                                           double4* __restrict__ s4,
                                                           double2* __restrict__ tab) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double* B = bc[cfg[i]].cell.B;
  const double x = pos[i], y = pos[i + N], z = pos[i + 2 * N];
  const double s1 = x * B[0] + y * B[1] + z * B[2];
  const double s2 = x * B[3] + y * B[4] + z * B[5];
  const double s3 = x * B[6] + y * B[7] + z * B[8];
  s[i] = s1;
  s[i + N] = s2;
  s[i + 2 * N] = s3;
  s4[i] = make_double4(s1, s2, s3, 0.0);
  int e = 0;
  for (int m = 0; m <= KX; ++m, ++e) {
    double sn, cs;
    sincospi(2.0 * m * s1, &sn, &cs);
    tab[(size_t)e * N + i] = make_double2(cs, sn);
  }
  for (int m = -KY; m <= KY; ++m, ++e) {
    double sn, cs;
    sincospi(2.0 * m * s2, &sn, &cs);
    tab[(size_t)e * N + i] = make_double2(cs, sn);
  }
  for (int m = -KZ; m <= KZ; ++m, ++e) {
    double sn, cs;
    sincospi(2.0 * m * s3, &sn, &cs);
    tab[(size_t)e * N + i] = make_double2(cs, sn);
  }
}

// Full neighbour list of every atom within its own configuration, j in
// increasing order (row-major nbr[i * maxnb + k] as the single-cell path).
__global__ __launch_bounds__(MDFF_BLOCK) void k_b_nlist(int N, const int* __restrict__ cfg,
                                                         const BatchCfg* __restrict__ bc,
                                                         const double4* __restrict__ s4,
                                                         int maxnb, int* __restrict__ cnt,
                                                         int* __restrict__ nbr,
                                                         int* __restrict__ ovf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const BatchCfg& b = bc[cfg[i]];
  const double cut2 = fmax(b.cut2, b.rc2);
  const double4 si = s4[i];
  int m = 0;
  for (int j = b.off; j < b.off + b.n; ++j) {
    if (j == i) continue;
    const double4 sj = s4[j];
    double rx, ry, rz;
    const double d2 = min_image(b.cell, si.x - sj.x, si.y - sj.y, si.z - sj.z, rx, ry, rz);
    if (d2 <= cut2) {
      if (m < maxnb) nbr[(size_t)i * maxnb + m] = j;
      ++m;
    }
  }
  cnt[i] = min(m, maxnb);
  if (m > maxnb) atomicMax(ovf, m);
}

// ---------------------------------------------------------------------------
// Structure factors rho(k) = sum_j (q_j + i k.mu_j) e^{i k.r_j} of one block
// of BK k vectors of one configuration; the atoms' phase tables are staged
// through LDS in tiles of BTA atoms.  kn = (e1, e2, e3) table entries of k,
// kv = (kx, ky, kz, A_k).
template <typename R>
__global__ __launch_bounds__(BK) void k_b_rho(int N, int ntab, const BatchCfg* __restrict__ bc,
                                              const int2* __restrict__ items,
                                              const int4* __restrict__ kn,
                                              const double4* __restrict__ kv,
                                              const double2* __restrict__ tab,
                                              const double* __restrict__ q,
                                              const double* __restrict__ mu, int chg, int dip,
                                              double2* __restrict__ rho) {
  __shared__ C2<R> sT[NTAB_MAX][BTA + 1];
  __shared__ R sW[4][BTA];
  const int2 it = items[blockIdx.x];
  const BatchCfg& b = bc[it.x];
  const int kk = it.y + threadIdx.x;
  const bool act = kk < b.nk;
  int4 e = make_int4(0, 0, 0, 0);
  R kx = 0, ky = 0, kz = 0;
  if (act) {
    e = kn[b.koff + kk];
    const double4 v = kv[b.koff + kk];
    kx = R(v.x);
    ky = R(v.y);
    kz = R(v.z);
  }
  KSum<R> aR, aI;
  const int j1 = b.off + b.n;
  for (int j0 = b.off; j0 < j1; j0 += BTA) {
    const int m = min(BTA, j1 - j0);
    __syncthreads();
    for (int t = threadIdx.x; t < ntab * BTA; t += blockDim.x) {
      const int en = t / BTA, a = t % BTA;
      C2<R> v = {R(0), R(0)};
      if (a < m) {
        const double2 d = tab[(size_t)en * N + j0 + a];
        v = {R(d.x), R(d.y)};
      }
      sT[en][a] = v;
    }
    for (int t = threadIdx.x; t < 4 * BTA; t += blockDim.x) {
      const int w = t / BTA, a = t % BTA;
      R v = R(0);
      if (a < m) {
        const int j = j0 + a;
        if (w == 0)
          v = chg ? R(q[j]) : R(0);
        else
          v = dip ? R(mu[j + (size_t)(w - 1) * N]) : R(0);
      }
      sW[w][a] = v;
    }
    __syncthreads();
    if (act) {
      for (int a = 0; a < m; ++a) {
        const C2<R> P = cmul(cmul(sT[e.x][a], sT[e.y][a]), sT[e.z][a]);
        const R wr = sW[0][a];
        const R wi = kx * sW[1][a] + ky * sW[2][a] + kz * sW[3][a];
        aR.add(wr * P.x - wi * P.y);
        aI.add(wr * P.y + wi * P.x);
      }
    }
  }
  if (act) rho[b.koff + kk] = make_double2((double)aR.val(), (double)aI.val());
}

// ---------------------------------------------------------------------------
// Reciprocal-space field and force of every atom over its configuration's k
// vectors, LK lanes per atom (lanes l and l + 64/LK share an atom... the
// atoms of a wave are consecutive for each lane group):
//   s_k = Re(conj(P_i) g), t_k = Im(conj(P_i) g),  g = A_k rho(k)
//   E_i = -2 (4 pi/V) sum_k k t_k
//   F_i = q_i E_i + 2 (4 pi/V) sum_k k (k.mu_i) s_k        (FORCE)
// (multipole_ES_rec, src/field.f90:2746-2801, x2 for the half mesh).
template <typename R, bool FORCE, int LK>
__global__ __launch_bounds__(MDFF_BLOCK) void k_b_apply(int N, const int* __restrict__ cfg,
                                                        const BatchCfg* __restrict__ bc,
                                                        const int4* __restrict__ kn,
                                                        const double4* __restrict__ kv,
                                                        const double2* __restrict__ rho,
                                                        const double2* __restrict__ tab,
                                                        const double* __restrict__ q,
                                                        const double* __restrict__ mu,
                                                        double* __restrict__ Erec,
                                                        double* __restrict__ Frec) {
  constexpr int APW = 64 / LK;  // atoms per wave
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int sub = lane / APW;
  const int i = wave * APW + (lane % APW);
  const bool act = i < N;
  KSum<R> U[3], W[3];
  R m0 = 0, m1 = 0, m2 = 0;
  if (act) {
    const BatchCfg& b = bc[cfg[i]];
    if (FORCE) {
      m0 = R(mu[i]);
      m1 = R(mu[i + N]);
      m2 = R(mu[i + 2 * N]);
    }
    for (int k = sub; k < b.nk; k += LK) {
      const int4 e = kn[b.koff + k];
      const double4 v = kv[b.koff + k];
      const double2 r = rho[b.koff + k];
      const R gr = R(v.w * r.x), gi = R(v.w * r.y);
      const double2 t1 = tab[(size_t)e.x * N + i];
      const double2 t2 = tab[(size_t)e.y * N + i];
      const double2 t3 = tab[(size_t)e.z * N + i];
      const C2<R> P = cmul(cmul(C2<R>{R(t1.x), R(t1.y)}, C2<R>{R(t2.x), R(t2.y)}),
                           C2<R>{R(t3.x), R(t3.y)});
      const R sk = P.x * gr + P.y * gi;
      const R tk = P.x * gi - P.y * gr;
      const R kx = R(v.x), ky = R(v.y), kz = R(v.z);
      U[0].add(kx * tk);
      U[1].add(ky * tk);
      U[2].add(kz * tk);
      if (FORCE) {
        const R km = (kx * m0 + ky * m1 + kz * m2) * sk;
        W[0].add(kx * km);
        W[1].add(ky * km);
        W[2].add(kz * km);
      }
    }
  }
  double u[3], w[3];
  for (int c = 0; c < 3; ++c) {
    u[c] = (double)U[c].val();
    w[c] = FORCE ? (double)W[c].val() : 0.0;
  }
  for (int off = APW; off < 64; off <<= 1)
    for (int c = 0; c < 3; ++c) {
      u[c] += __shfl_xor(u[c], off, 64);
      if (FORCE) w[c] += __shfl_xor(w[c], off, 64);
    }
  if (act && sub == 0) {
    const double f = 2.0 * bc[cfg[i]].fpi_V;
    for (int c = 0; c < 3; ++c) Erec[i + (size_t)c * N] = -f * u[c];
    if (FORCE) {
      const double qi = q[i];
      for (int c = 0; c < 3; ++c) Frec[i + (size_t)c * N] = qi * (-f * u[c]) + f * w[c];
    }
  }
}

// Sum of the parts of one multipole_ES call (src/field.f90:1817-2054,
// without surface term):
//   STATIC  E_out = E_dir + E_rec + 2 selfa2 mu
//   IND     E_out = S + E_dir + E_rec + 2 selfa2 mu
//   FINAL   F_out = (F_dir + F_rec) coul_unit
template <int MODE>
__global__ __launch_bounds__(MDFF_BLOCK) void k_b_combine(
    int N, const int* __restrict__ cfg, const BatchCfg* __restrict__ bc, double coul_unit,
    const double* __restrict__ dir, const double* __restrict__ rec,
    const double* __restrict__ mu, const double* __restrict__ S, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double sa2 = bc[cfg[i]].selfa2;
  for (int c = 0; c < 3; ++c) {
    const size_t x = i + (size_t)c * N;
    if (MODE == ES_FINAL) {
      out[x] = (dir[x] + rec[x]) * coul_unit;
    } else {
      double v = dir[x] + rec[x] + 2.0 * sa2 * mu[x];
      if (MODE == ES_IND) v += S[x];
      out[x] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// small vector kernels (graph-safe replacements of memset / memcpy)
__global__ __launch_bounds__(MDFF_BLOCK) void k_b_fill(size_t n, double v, double* __restrict__ x) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// z = x + s y
__global__ __launch_bounds__(MDFF_BLOCK) void k_b_axpy(size_t n, const double* __restrict__ x,
                                                        double s, const double* __restrict__ y,
                                                        double* __restrict__ z) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = x[i] + s * y[i];
}

// unit charges on the atoms of type a: src = [typ == a], qp = q + src,
// qm = q - src
__global__ __launch_bounds__(MDFF_BLOCK) void k_b_qsrc(int N, const int* __restrict__ typ, int a,
                                                        const double* __restrict__ q,
                                                        double* __restrict__ src,
                                                        double* __restrict__ qp,
                                                        double* __restrict__ qm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double v = (typ[i] == a) ? 1.0 : 0.0;
  src[i] = v;
  qp[i] = q[i] + v;
  qm[i] = q[i] - v;
}

// source of the polarizability tangent: S = P_t mu / alpha_t^2 (so that
// alpha (S + T dmu) = P_t mu / alpha_t + alpha T dmu)
__global__ __launch_bounds__(MDFF_BLOCK) void k_b_asrc(int N, const int* __restrict__ typ, int t,
                                                        double alpha_t,
                                                        const double* __restrict__ mu,
                                                        double* __restrict__ S) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double f = (typ[i] == t) ? 1.0 / (alpha_t * alpha_t) : 0.0;
  for (int c = 0; c < 3; ++c) S[i + (size_t)c * N] = f * mu[i + (size_t)c * N];
}

// loss rows: 0 sum (F - F_t)^2; g = 2 (F - F_t) / (3 N_tot), F = F_sr + F_el
__global__ __launch_bounds__(MDFF_BLOCK) void k_b_loss(int N, double inv3n,
                                                        const double* __restrict__ Fsr,
                                                        const double* __restrict__ Fel,
                                                        const double* __restrict__ Ft,
                                                        double* __restrict__ F,
                                                        double* __restrict__ g,
                                                        double* __restrict__ part) {
  __shared__ double red[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[1] = {0.0};
  if (i < N) {
    for (int c = 0; c < 3; ++c) {
      const size_t x = i + (size_t)c * N;
      const double f = Fsr[x] + Fel[x];
      const double d = f - Ft[x];
      F[x] = f;
      if (g) g[x] = 2.0 * d * inv3n;
      acc[0] += d * d;
    }
  }
  block_sum<1>(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// rows: sum_i g_i . (Fp_i - Fm_i)
__global__ __launch_bounds__(MDFF_BLOCK) void k_b_contract(int N, const double* __restrict__ g,
                                                            const double* __restrict__ Fp,
                                                            const double* __restrict__ Fm,
                                                            double* __restrict__ part) {
  __shared__ double red[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[1] = {0.0};
  if (i < N)
    for (int c = 0; c < 3; ++c) {
      const size_t x = i + (size_t)c * N;
      acc[0] += g[x] * (Fp[x] - Fm[x]);
    }
  block_sum<1>(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// ---------------------------------------------------------------------------
// Adjoint of the BMHFTD pair forces (src/field.f90:2952-2976): for every
// directed pair (i, j) inside the short-range cutoff, with r = r_i - r_j,
// d = |r|, s = r . g_i and the pair weight
//   w = [B A e^{-Bd} - 6 C f6 / d^7 - 8 D f8 / d^9 + C fd6 / d^6 + D fd8 / d^8] / d
// (f_n, fd_n Tang-Toennies with b = BD, c = 1), accumulate s dw/dp:
//   dw/dA  = B e^{-Bd} / d
//   dw/dB  = A e^{-Bd} (1 - B d) / d
//   dw/dC  = (-6 f6 / d + fd6) / d^7
//   dw/dD  = (-8 f8 / d + fd8) / d^9
//   dw/dBD = (C fd6 / d^6 + D fd8 / d^8) (1 / BD - d) / d
// (df/db = (d/b) fd, dfd/db = fd ((n+1)/b - d)).  Pair terms are recomputed
// here rather than stored by the forward pass.  Rows: NSR * pair index.
template <typename R>
__global__ __launch_bounds__(MDFF_BLOCK) void k_b_sr_adj(SrArgs a, const int* __restrict__ upair,
                                                          const double* __restrict__ g,
                                                          double* __restrict__ part) {
  __shared__ double red[4 * NSR * NPAIR_MAX];
  const int n = a.n;
  const PairTables* T = a.T;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = tid / LPA, sub = tid % LPA;
  double acc[NSR * NPAIR_MAX];
#pragma unroll
  for (int k = 0; k < NSR * NPAIR_MAX; ++k) acc[k] = 0.0;
  if (i < n) {
    const double si1 = a.s[i], si2 = a.s[i + n], si3 = a.s[i + 2 * n];
    const int ti = a.typ[i];
    const int c = a.cnt[i];
    const BatchCfg& bcf = a.bc[a.cfg[i]];
    const R rc2 = R(bcf.rc2);
    const R g0 = R(g[i]), g1 = R(g[i + n]), g2 = R(g[i + 2 * n]);
    const int* nb = a.nbr + (size_t)i * a.maxnb;
    for (int k = sub; k < c; k += LPA) {
      const int j = nb[k];
      R rx, ry, rz;
      const R rsq = min_image_t<R>(bcf.cell, si1 - a.s4[j].x, si2 - a.s4[j].y, si3 - a.s4[j].z,
                                   rx, ry, rz);
      if (!(rsq < rc2)) continue;
      const int t = ti + MDFF_NT * a.typ[j];
      if (T->bm_A[t] == 0.0) continue;
      const R A = R(T->bm_A[t]), B = R(T->bm_B[t]), C = R(T->bm_C[t]), D = R(T->bm_D[t]),
              BD = R(T->bm_BD[t]);
      const R d = sqrt(rsq);
      const R id = R(1) / d;
      const R ir2 = R(1) / rsq;
      const R ir6 = ir2 * ir2 * ir2;
      const R ir8 = ir6 * ir2;
      const R eb = rexp<R>(-B * d);
      R f6 = R(1), fd6 = R(0), f8 = R(1), fd8 = R(0);
      if (a.flags & SR_BMD) {
        tt_damp_t<R>(a.tt, BD, R(1), d, 6, f6, fd6);
        tt_damp_t<R>(a.tt, BD, R(1), d, 8, f8, fd8);
      }
      const R s = rx * g0 + ry * g1 + rz * g2;
      const int p = upair[t];
      acc[NSR * p + 0] += (double)(s * B * eb * id);
      acc[NSR * p + 1] += (double)(s * A * eb * (R(1) - B * d) * id);
      acc[NSR * p + 2] += (double)(s * (R(-6) * f6 * id + fd6) * ir6 * id);
      acc[NSR * p + 3] += (double)(s * (R(-8) * f8 * id + fd8) * ir8 * id);
      if (a.flags & SR_BMD)
        acc[NSR * p + 4] += (double)(s * (C * fd6 * ir6 + D * fd8 * ir8) * (R(1) / BD - d) * id);
    }
  }
  block_sum<NSR * NPAIR_MAX>(acc, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < NSR * NPAIR_MAX; ++k) part[(size_t)k * gridDim.x + blockIdx.x] = acc[k];
}

// Adjoint of the Morse pair forces of k_shortrange: X = e^{-rho (d - sigma)},
// w = 2 eps rho X (X - 1) / d, so that for s = r . g_i
//   dw/deps   = 2 rho X (X - 1) / d
//   dw/dsigma = 2 eps rho^2 X (2 X - 1) / d
//   dw/drho   = 2 eps X [(X - 1) - rho (d - sigma) (2 X - 1)] / d
// Rows: NMO * pair index.
template <typename R>
__global__ __launch_bounds__(MDFF_BLOCK) void k_b_mo_adj(SrArgs a, const int* __restrict__ upair,
                                                          const double* __restrict__ g,
                                                          double* __restrict__ part) {
  __shared__ double red[4 * NMO * NPAIR_MAX];
  const int n = a.n;
  const PairTables* T = a.T;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = tid / LPA, sub = tid % LPA;
  double acc[NMO * NPAIR_MAX];
#pragma unroll
  for (int k = 0; k < NMO * NPAIR_MAX; ++k) acc[k] = 0.0;
  if (i < n) {
    const double si1 = a.s[i], si2 = a.s[i + n], si3 = a.s[i + 2 * n];
    const int ti = a.typ[i];
    const int c = a.cnt[i];
    const BatchCfg& bcf = a.bc[a.cfg[i]];
    const R rc2 = R(bcf.rc2);
    const R g0 = R(g[i]), g1 = R(g[i + n]), g2 = R(g[i + 2 * n]);
    const int* nb = a.nbr + (size_t)i * a.maxnb;
    for (int k = sub; k < c; k += LPA) {
      const int j = nb[k];
      R rx, ry, rz;
      const R rsq = min_image_t<R>(bcf.cell, si1 - a.s4[j].x, si2 - a.s4[j].y, si3 - a.s4[j].z,
                                   rx, ry, rz);
      if (!(rsq < rc2)) continue;
      const int t = ti + MDFF_NT * a.typ[j];
      const R eps = R(T->mo_eps[t]), rho = R(T->mo_rho[t]), sig = R(T->mo_sig[t]);
      const R d = sqrt(rsq);
      const R X = rexp<R>(-d * rho) * R(T->mo_rs[t]);
      const R s = (rx * g0 + ry * g1 + rz * g2) / d;
      const int p = upair[t];
      acc[NMO * p + 0] += (double)(s * R(2) * rho * X * (X - R(1)));
      acc[NMO * p + 1] += (double)(s * R(2) * eps * rho * rho * X * (R(2) * X - R(1)));
      acc[NMO * p + 2] +=
          (double)(s * R(2) * eps * X * ((X - R(1)) - rho * (d - sig) * (R(2) * X - R(1))));
    }
  }
  block_sum<NMO * NPAIR_MAX>(acc, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < NMO * NPAIR_MAX; ++k) part[(size_t)k * gridDim.x + blockIdx.x] = acc[k];
}

// Species of the damping parameters: slot[t] >= 0 for a species whose
// Tang-Toennies b (dip_pol_damp_b(t, t, :)) is a fitted parameter.
struct DampSlots {
  int slot[MDFF_NT];
};

// Explicit derivatives with respect to the damping b of the Tang-Toennies
// dipole damping (src/field.f90:2205-2231, 2468-2538), FP64, pair terms
// recomputed.  The damping of centre i's own terms uses the parameters of
// its species (dmp slot o of k_es_dir), so b_t enters
//   the static field   E_i += q_j (F1 - F1d_i) r / d^3
//   the final force    F_i += [q_i (T2 - T2d_j) mu_j - q_j (T2 - T2d_i) mu_i] coul_unit
// with F1d = 1 - f, F2d = F1d + (d / 3) fd and, for the Tang-Toennies f, fd
// of order n, df/db = (d / b) fd, dfd/db = fd ((n + 1) / b - d).  Writes
// dEs[m] = dE_static / db_m (3N each, every atom) and the block partials of
// sum_i g_i . dF_i / db_m (rows m).
__global__ __launch_bounds__(MDFF_BLOCK) void k_b_damp_adj(EsDirArgs a, DampSlots ds,
                                                            const double* __restrict__ g,
                                                            double coul_unit, int ndamp,
                                                            double* __restrict__ dEs,
                                                            double* __restrict__ part) {
  __shared__ double red[4 * NDAMP_MAX];
  const int n = a.n;
  const PairTables* T = a.T;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = tid / LPA, sub = tid % LPA;
  double de[NDAMP_MAX][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
  double df[NDAMP_MAX][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
  if (i < n) {
    const double si1 = a.s[i], si2 = a.s[i + n], si3 = a.s[i + 2 * n];
    const int ti = a.typ[i];
    const BatchCfg& bcf = a.bc[a.cfg[i]];
    const double qi = a.q[i];
    const double mi[3] = {a.mu[i], a.mu[i + n], a.mu[i + 2 * n]};
    const int mi_slot = ds.slot[ti];
    const int c = a.cnt[i];
    const int* nb = a.nbr + (size_t)i * a.maxnb;
    for (int k = sub; k < c; k += LPA) {
      const int j = nb[k];
      const double4 sj = a.s4[j];
      double rx, ry, rz;
      const double d2 = min_image(bcf.cell, si1 - sj.x, si2 - sj.y, si3 - sj.z, rx, ry, rz);
      if (d2 > bcf.cut2) continue;
      const int tj = a.typ[j];
      const int o = (i < j) ? 0 : 1;
      const int pij = ti + MDFF_NT * tj, pji = tj + MDFF_NT * ti;
      if (!T->dmp_on[o][pij]) continue;
      const int mj_slot = ds.slot[tj];
      if (mi_slot < 0 && mj_slot < 0) continue;
      const double d = sqrt(d2);
      const double dm3 = 1.0 / (d2 * d), dm5 = dm3 / d2;
      const double qj = a.q[j];
      const double mj[3] = {a.mu[j], a.mu[j + n], a.mu[j + 2 * n]};
      const double r3[3] = {rx, ry, rz};
      if (mi_slot >= 0 && T->dmp_b[o][pij] != 0.0) {
        const double b = T->dmp_b[o][pij];
        const int kk = T->dmp_k[o][pij];
        double f, fd;
        tt_damp_t<double>(a.tt, b, T->dmp_c[o][pij], d, kk, f, fd);
        const double dF1 = -(d / b) * fd;
        const double dF2 = dF1 + (d / 3.0) * fd * ((kk + 1) / b - d);
        const double rmi = rx * mi[0] + ry * mi[1] + rz * mi[2];
        for (int x = 0; x < 3; ++x) {
          de[mi_slot][x] += -qj * dF1 * dm3 * r3[x];
          df[mi_slot][x] += qj * (r3[x] * rmi * 3.0 * dF2 * dm5 - mi[x] * dF1 * dm3) * coul_unit;
        }
      }
      if (mj_slot >= 0 && T->dmp_b[1 - o][pji] != 0.0) {
        const double b = T->dmp_b[1 - o][pji];
        const int kk = T->dmp_k[1 - o][pji];
        double f, fd;
        tt_damp_t<double>(a.tt, b, T->dmp_c[1 - o][pji], d, kk, f, fd);
        const double dF1 = -(d / b) * fd;
        const double dF2 = dF1 + (d / 3.0) * fd * ((kk + 1) / b - d);
        const double rmj = rx * mj[0] + ry * mj[1] + rz * mj[2];
        for (int x = 0; x < 3; ++x)
          df[mj_slot][x] -= qi * (r3[x] * rmj * 3.0 * dF2 * dm5 - mj[x] * dF1 * dm3) * coul_unit;
      }
    }
  }
  for (int off = LPA / 2; off > 0; off >>= 1)
    for (int m = 0; m < NDAMP_MAX; ++m)
      for (int x = 0; x < 3; ++x) {
        de[m][x] += __shfl_xor(de[m][x], off, 64);
        df[m][x] += __shfl_xor(df[m][x], off, 64);
      }
  double acc[NDAMP_MAX] = {0.0, 0.0};
  if (i < n && sub == 0) {
    for (int m = 0; m < ndamp; ++m)
      for (int x = 0; x < 3; ++x) {
        dEs[(size_t)m * 3 * n + i + (size_t)x * n] = de[m][x];
        acc[m] += g[i + (size_t)x * n] * df[m][x];
      }
  }
  block_sum<NDAMP_MAX>(acc, red);
  if (threadIdx.x == 0)
    for (int m = 0; m < NDAMP_MAX; ++m) part[(size_t)m * gridDim.x + blockIdx.x] = acc[m];
}

}  // namespace batch
}  // namespace mdff

// ===========================================================================
// host side
// ===========================================================================
namespace mdff {
namespace batch {
namespace {

struct BFail {
  int code;
};

inline int nblk(size_t n) { return (int)((n + MDFF_BLOCK - 1) / MDFF_BLOCK); }

}  // namespace
}  // namespace batch
}  // namespace mdff

struct mdffgpu_batch {
  mdffgpu_params P;
  int nconf = 0, N = 0, KX = 0, KY = 0, KZ = 0, ntab = 0, NK = 0, nitems = 0, maxnb = 0;
  int ntype = 0, npair = 0, npol = 0, nparam = 0;
  // Morse parameters in theta; species with a fitted damping b
  bool lmo = false;
  int ndamp = 0;
  int damp_type[mdff::batch::NDAMP_MAX];
  mdff::batch::DampSlots dslots;
  double* d_dEs = nullptr;
  bool mixed = false, pim = false;
  // FP32 path: reciprocal sums on the matrix cores (configurations with a
  // common kmax); positions' version for its phase tables
  mdff::mixed::MixedState* mx = nullptr;
  bool mx_tried = false;
  long long epoch = 0;
  int pol_type[MDFF_NT];
  double pol_alpha[MDFF_NT];
  // SCF pair table: the polarizable species (bit per species), the atoms of
  // those species in index order and the entries per lane of k_dd_tab2's
  // pipeline
  unsigned polmask = 0;
  int npolrows = 0, dd_u = 8;
  int* d_polrows = nullptr;
  int upair_h[256];
  std::vector<mdff::BatchCfg> bc_h;
  mdff::BatchCfg* d_bc = nullptr;
  int *d_cfg = nullptr, *d_typ = nullptr, *d_upair = nullptr, *d_cnt = nullptr,
      *d_nbr = nullptr, *d_ovf = nullptr;
  int4* d_kn = nullptr;
  double4* d_kv = nullptr;
  int2* d_items = nullptr;
  mdff::PairTables pt_h;
  mdff::TypeTables ty_h;
  mdff::TTTable tt;
  mdff::PairTables* d_pt = nullptr;
  mdff::TypeTables* d_ty = nullptr;
  double *d_pos = nullptr, *d_s = nullptr;
  double4* d_s4 = nullptr;
  double2* d_tab = nullptr;
  void* d_t2 = nullptr;
  size_t t2_cap = 0;
  double4* d_mu4 = nullptr;
  double *d_q = nullptr, *d_zero = nullptr;
  double *d_mu = nullptr, *d_E = nullptr, *d_Es = nullptr, *d_dir = nullptr, *d_rec = nullptr,
         *d_frec = nullptr, *d_Fel = nullptr, *d_Fsr = nullptr, *d_F = nullptr, *d_g = nullptr,
         *d_Ft = nullptr;
  double *d_dmu = nullptr, *d_Et = nullptr, *d_S = nullptr, *d_qsrc = nullptr, *d_qp = nullptr,
         *d_qm = nullptr, *d_mup = nullptr, *d_mum = nullptr, *d_Fp = nullptr, *d_Fm = nullptr;
  double2* d_rho = nullptr;
  double* d_part = nullptr;
  size_t part_cap = 0;
  double* d_out = nullptr;
  double* h_out = nullptr;
  hipStream_t own = nullptr, st = nullptr;
  bool positions_set = false;
  int apply_lanes = 4;
  // graph of loss_grad with a fixed SCF iteration count
  bool graph_on = false;
  hipGraphExec_t gexec = nullptr;
  int graph_iters = -1;
  std::vector<void*> allocs;
  std::string err;
};

namespace mdff {
namespace batch {
namespace {

[[noreturn]] void bfail(mdffgpu_batch* b, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (b) b->err = buf;
  throw BFail{code};
}

thread_local std::string g_batch_create_error;

#define BCK(b, x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess)                                                           \
      bfail(b, MDFFGPU_ERR_DEVICE, "%s: %s (%s:%d)", #x, hipGetErrorString(e_),    \
            __FILE__, __LINE__);                                                    \
  } while (0)

void launch_check(mdffgpu_batch* b, const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) bfail(b, MDFFGPU_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

template <typename T>
void balloc(mdffgpu_batch* b, T** p, size_t count) {
  void* q = nullptr;
  BCK(b, hipMalloc(&q, (count ? count : 1) * sizeof(T)));
  b->allocs.push_back(q);
  *p = (T*)q;
}

void fill(mdffgpu_batch* b, double* x, size_t n, double v) {
  k_b_fill<<<nblk(n), MDFF_BLOCK, 0, b->st>>>(n, v, x);
}

void copy(mdffgpu_batch* b, double* dst, const double* src, size_t n) {
  k_b_axpy<<<nblk(n), MDFF_BLOCK, 0, b->st>>>(n, src, 0.0, src, dst);
}

void ensure_part(mdffgpu_batch* b, size_t n) {
  if (n <= b->part_cap) return;
  void* q = nullptr;
  BCK(b, hipMalloc(&q, n * sizeof(double)));
  if (b->d_part) {
    (void)hipFree(b->d_part);
    b->allocs.erase(std::find(b->allocs.begin(), b->allocs.end(), (void*)b->d_part));
  }
  b->allocs.push_back(q);
  b->d_part = (double*)q;
  b->part_cap = n;
}

// reduce rows [r0, r0 + nrows) of the block partials into d_out[slot ...]
void reduce(mdffgpu_batch* b, int nrows, int nb, int slot) {
  k_reduce_rows<<<nrows, MDFF_BLOCK, 0, b->st>>>(b->d_part, nb, b->d_out + slot);
}

EsDirArgs es_args(mdffgpu_batch* b, const double* q, const double* mu) {
  EsDirArgs da;
  std::memset(&da, 0, sizeof(da));
  da.n = b->N;
  da.maxnb = b->maxnb;
  da.s = b->d_s;
  da.s4 = b->d_s4;
  da.typ = b->d_typ;
  da.cnt = b->d_cnt;
  da.nbr = b->d_nbr;
  da.T = b->d_pt;
  da.tt = b->tt;
  da.q = q;
  da.mu = mu;
  da.part = b->d_part;
  da.bc = b->d_bc;
  da.cfg = b->d_cfg;
  da.polmask = b->polmask;
  return da;
}

SrArgs sr_args(mdffgpu_batch* b) {
  SrArgs sa;
  std::memset(&sa, 0, sizeof(sa));
  const mdffgpu_params& P = b->P;
  sa.n = b->N;
  sa.maxnb = b->maxnb;
  sa.s = b->d_s;
  sa.s4 = b->d_s4;
  sa.typ = b->d_typ;
  sa.cnt = b->d_cnt;
  sa.nbr = b->d_nbr;
  sa.T = b->d_pt;
  sa.tt = b->tt;
  sa.flags = (P.lnmlj ? SR_LJ : 0) | ((P.lbmhft || P.lbmhftd) ? SR_BM : 0) |
             (P.lbmhftd ? SR_BMD : 0) | (P.lmorse ? SR_MO : 0) |
             (P.trunc == 1 ? SR_TRUNC1 : 0) | (P.trunc == 2 ? SR_TRUNC2 : 0);
  sa.f = b->d_Fsr;
  sa.part = b->d_part;
  sa.bc = b->d_bc;
  sa.cfg = b->d_cfg;
  return sa;
}

template <typename R, bool FORCE>
void launch_apply(mdffgpu_batch* b, const double* q, const double* mu, double* Erec,
                  double* Frec) {
  const int LK = b->apply_lanes;
  const size_t threads = (size_t)b->N * LK;
  const int grid = nblk(threads);
  switch (LK) {
    case 2:
      k_b_apply<R, FORCE, 2><<<grid, MDFF_BLOCK, 0, b->st>>>(b->N, b->d_cfg, b->d_bc, b->d_kn,
                                                              b->d_kv, b->d_rho, b->d_tab, q,
                                                              mu, Erec, Frec);
      break;
    case 8:
      k_b_apply<R, FORCE, 8><<<grid, MDFF_BLOCK, 0, b->st>>>(b->N, b->d_cfg, b->d_bc, b->d_kn,
                                                              b->d_kv, b->d_rho, b->d_tab, q,
                                                              mu, Erec, Frec);
      break;
    case 16:
      k_b_apply<R, FORCE, 16><<<grid, MDFF_BLOCK, 0, b->st>>>(b->N, b->d_cfg, b->d_bc, b->d_kn,
                                                               b->d_kv, b->d_rho, b->d_tab, q,
                                                               mu, Erec, Frec);
      break;
    default:
      k_b_apply<R, FORCE, 4><<<grid, MDFF_BLOCK, 0, b->st>>>(b->N, b->d_cfg, b->d_bc, b->d_kn,
                                                              b->d_kv, b->d_rho, b->d_tab, q,
                                                              mu, Erec, Frec);
      break;
  }
}

// Matrix-core reciprocal sums of the FP32 path, created on first use when
// every configuration has the same kmax (else the per-k kernels serve).
mdff::mixed::MixedState* mx_state(mdffgpu_batch* b) {
  if (b->mx || b->mx_tried) return b->mx;
  b->mx_tried = true;
  const BatchCfg& x0 = b->bc_h[0];
  for (const BatchCfg& x : b->bc_h)
    for (int a = 0; a < 3; ++a)
      if (x.kmax[a] != x0.kmax[a]) return nullptr;
  std::vector<mdff::mixed::MxSeg> segs(b->bc_h.size());
  for (size_t c = 0; c < segs.size(); ++c) {
    std::memset(&segs[c], 0, sizeof(segs[c]));
    segs[c].cell = b->bc_h[c].cell;
    segs[c].alpha = b->bc_h[c].alpha;
    segs[c].off = b->bc_h[c].off;
    segs[c].n = b->bc_h[c].n;
  }
  b->mx = mdff::mixed::create_batch(b->N, x0.kmax, segs);
  return b->mx;
}

// One multipole_ES call over the batch with sources (q, mu):
//   STATIC  out = static field (charges only)
//   IND     out = S + dipole field of mu (the SCF field operator)
//   FINAL   out = electrostatic forces * coul_unit
template <typename R>
void es_call(mdffgpu_batch* b, int mode, const double* q, const double* mu, const double* S,
             double* out) {
  using R4 = typename std::conditional<sizeof(R) == sizeof(double), double4, float4>::type;
  const int N = b->N;
  const int gbl = nblk((size_t)N * LPA);
  EsDirArgs da = es_args(b, q, mu);
  if (mode == ES_STATIC) {
    da.task = TASK_QQ;
    da.ef = b->d_dir;
    k_es_dir<ES_STATIC, R, true><<<gbl, MDFF_BLOCK, 0, b->st>>>(da);
  } else if (mode == ES_IND) {
    da.ef = b->d_dir;
    R4* m4 = reinterpret_cast<R4*>(b->d_mu4);
    k_pack_mu<R4><<<nblk(N), MDFF_BLOCK, 0, b->st>>>(N, mu, m4);
    if constexpr (sizeof(R) == sizeof(double)) {
      k_dd_field<<<nblk((size_t)N * DD_LANES), MDFF_BLOCK, 0, b->st>>>(da, (const double*)b->d_t2,
                                                                     m4);
    } else {
      const size_t cap = (size_t)b->maxnb * N;
      const uint4* u4 = reinterpret_cast<const uint4*>(b->d_t2);
      const float2* tab = reinterpret_cast<const float2*>(u4 + N);
      const int* tj = reinterpret_cast<const int*>(tab + cap);
      // the rows of the polarizable atoms only: the field elsewhere is unused
      auto launch = [&](auto ut, auto* tjp) {
        constexpr int U = decltype(ut)::value;
        using IT = typename std::remove_const<std::remove_pointer_t<decltype(tjp)>>::type;
        k_dd_tab2<IT, U><<<dd_tab_blocks(b->npolrows), MDFF_BLOCK, 0, b->st>>>(
            N, b->maxnb, tab, tjp, tj + cap, m4, u4, Cell{}, b->d_bc, b->d_cfg, nullptr,
            b->d_polrows, b->npolrows, b->d_dir);
      };
      auto launch_u = [&](auto* tjp) {
        if (b->dd_u == 4)
          launch(std::integral_constant<int, 4>{}, tjp);
        else if (b->dd_u == 6)
          launch(std::integral_constant<int, 6>{}, tjp);
        else
          launch(std::integral_constant<int, 8>{}, tjp);
      };
      if (b->npolrows > 0 && dd_idx16(N))
        launch_u(reinterpret_cast<const unsigned short*>(tj));
      else if (b->npolrows > 0)
        launch_u(tj);
    }
  } else {
    da.task = TASK_QQ | (b->pim ? (TASK_QD | TASK_DD) : 0);
    da.f = b->d_dir;
    k_es_dir<ES_FINAL, R, true><<<gbl, MDFF_BLOCK, 0, b->st>>>(da);
  }
  launch_check(b, "batch real space");
  const int chg = (mode != ES_IND) ? 1 : 0;
  const int dip = (mode != ES_STATIC && b->pim) ? 1 : 0;
  if (sizeof(R) == sizeof(float) && mx_state(b)) {
    mdff::mixed::kspace_batch(b->mx, b->st, b->d_s, b->epoch, b->d_cfg, q, mu, chg, dip,
                              mode == ES_FINAL, b->d_rec, b->d_frec);
  } else {
    k_b_rho<R><<<b->nitems, BK, 0, b->st>>>(N, b->ntab, b->d_bc, b->d_items, b->d_kn, b->d_kv,
                                            b->d_tab, q, mu, chg, dip, b->d_rho);
    if (mode == ES_FINAL)
      launch_apply<R, true>(b, q, mu, b->d_rec, b->d_frec);
    else
      launch_apply<R, false>(b, q, mu, b->d_rec, nullptr);
  }
  launch_check(b, "batch k space");
  const int gb = nblk(N);
  const double cu = b->P.coul_unit;
  if (mode == ES_STATIC)
    k_b_combine<ES_STATIC><<<gb, MDFF_BLOCK, 0, b->st>>>(N, b->d_cfg, b->d_bc, cu, b->d_dir,
                                                         b->d_rec, mu, S, out);
  else if (mode == ES_IND)
    k_b_combine<ES_IND><<<gb, MDFF_BLOCK, 0, b->st>>>(N, b->d_cfg, b->d_bc, cu, b->d_dir,
                                                      b->d_rec, mu, S, out);
  else
    k_b_combine<ES_FINAL><<<gb, MDFF_BLOCK, 0, b->st>>>(N, b->d_cfg, b->d_bc, cu, b->d_dir,
                                                        b->d_frec, mu, S, out);
  launch_check(b, "batch combine");
}

// Jacobi iteration x <- alpha (S + T x) from x = 0, with E the field
// buffer; iters > 0: exactly iters iterations; iters == 0: until the SCF
// residual (src/field.f90:3378-3400, over all polarizable atoms of the
// batch) is below conv_tol_ind, within min/max_scf_pol_iter.  Returns the
// number of iterations.
template <typename R>
int jacobi(mdffgpu_batch* b, const double* S, double* x, double* E, int iters, double* rmsd) {
  const int N = b->N;
  const int gb = nblk(N);
  const mdffgpu_params& P = b->P;
  copy(b, E, S, 3 * (size_t)N);
  int it = 0;
  *rmsd = 0.0;
  const int nmax = iters > 0 ? iters : std::max(P.max_scf_pol_iter, 1);
  while (it < nmax) {
    ++it;
    k_induced<<<gb, MDFF_BLOCK, 0, b->st>>>(N, b->d_typ, b->d_ty, E, x, b->d_part, nullptr,
                                            nullptr, nullptr);
    es_call<R>(b, ES_IND, b->d_zero, x, S, E);
    if (iters == 0) {
      k_scf_resid<<<gb, MDFF_BLOCK, 0, b->st>>>(N, b->d_typ, b->d_ty, x, E, b->d_part);
      reduce(b, 2, gb, 0);
      BCK(b, hipMemcpyAsync(b->h_out, b->d_out, 2 * sizeof(double), hipMemcpyDeviceToHost,
                            b->st));
      BCK(b, hipStreamSynchronize(b->st));
      *rmsd = b->h_out[1] > 0 ? std::sqrt(b->h_out[0] / b->h_out[1]) : 0.0;
      if (it >= P.min_scf_pol_iter && *rmsd <= P.conv_tol_ind) break;
    }
  }
  return it;
}

// Forward pass: F_sr, the SCF dipoles, F_el.  Returns the SCF iterations.
template <typename R>
int forward(mdffgpu_batch* b, int iters, double* rmsd) {
  const int N = b->N;
  const int gbl = nblk((size_t)N * LPA);
  const mdffgpu_params& P = b->P;
  const bool sr = P.lnmlj || P.lbmhft || P.lbmhftd || P.lmorse;
  if (sr) {
    SrArgs sa = sr_args(b);
    k_shortrange<R, true><<<gbl, MDFF_BLOCK, 0, b->st>>>(sa);
    launch_check(b, "batch short range");
  } else {
    fill(b, b->d_Fsr, 3 * (size_t)N, 0.0);
  }
  int n_scf = 0;
  *rmsd = 0.0;
  if (b->pim) {
    es_call<R>(b, ES_STATIC, b->d_q, b->d_zero, nullptr, b->d_Es);
    n_scf = jacobi<R>(b, b->d_Es, b->d_mu, b->d_E, iters, rmsd);
  } else {
    fill(b, b->d_mu, 3 * (size_t)N, 0.0);
  }
  if (P.lcoulomb)
    es_call<R>(b, ES_FINAL, b->d_q, b->d_mu, nullptr, b->d_Fel);
  else
    fill(b, b->d_Fel, 3 * (size_t)N, 0.0);
  return n_scf;
}

// Backward pass after forward(): loss and gradient into d_out[0 ..].
//   d_out[0]                   loss
//   d_out[1 + NSR p + k]       BMHFTD parameter k of unordered pair p
//   d_out[1 + NSR npair + a]   2 x gradient of q_a (halved on the host)
//   d_out[1 + NSR npair + ntype + m]  2 x gradient of alpha of the m-th
//                                     polarizable type
template <typename R>
void backward(mdffgpu_batch* b, int n_scf) {
  const int N = b->N;
  const size_t n3 = 3 * (size_t)N;
  const int gb = nblk(N);
  const int gbl = nblk((size_t)N * LPA);
  k_b_loss<<<gb, MDFF_BLOCK, 0, b->st>>>(N, 1.0 / (3.0 * N), b->d_Fsr, b->d_Fel, b->d_Ft,
                                         b->d_F, b->d_g, b->d_part);
  reduce(b, 1, gb, 0);
  const mdffgpu_params& P = b->P;
  if (P.lbmhftd || P.lbmhft) {
    SrArgs sa = sr_args(b);
    k_b_sr_adj<R><<<gbl, MDFF_BLOCK, 0, b->st>>>(sa, b->d_upair, b->d_g, b->d_part);
    launch_check(b, "batch short-range adjoint");
    reduce(b, NSR * b->npair, gbl, 1);
  } else {
    fill(b, b->d_out + 1, NSR * b->npair, 0.0);
  }
  int slot = 1 + NSR * b->npair;
  if (b->lmo) {
    SrArgs sa = sr_args(b);
    k_b_mo_adj<R><<<gbl, MDFF_BLOCK, 0, b->st>>>(sa, b->d_upair, b->d_g, b->d_part);
    launch_check(b, "batch Morse adjoint");
    reduce(b, NMO * b->npair, gbl, slot);
    slot += NMO * b->npair;
  }
  double rms;
  if (!P.lcoulomb) {
    fill(b, b->d_out + slot, b->ntype + b->npol + 2 * b->ndamp, 0.0);
    return;
  }
  for (int a = 0; a < b->ntype; ++a, ++slot) {
    k_b_qsrc<<<gb, MDFF_BLOCK, 0, b->st>>>(N, b->d_typ, a, b->d_q, b->d_qsrc, b->d_qp,
                                           b->d_qm);
    const double* mup = b->d_zero;
    const double* mum = b->d_zero;
    if (b->pim) {
      es_call<R>(b, ES_STATIC, b->d_qsrc, b->d_zero, nullptr, b->d_S);
      jacobi<R>(b, b->d_S, b->d_dmu, b->d_Et, n_scf, &rms);
      k_b_axpy<<<nblk(n3), MDFF_BLOCK, 0, b->st>>>(n3, b->d_mu, 1.0, b->d_dmu, b->d_mup);
      k_b_axpy<<<nblk(n3), MDFF_BLOCK, 0, b->st>>>(n3, b->d_mu, -1.0, b->d_dmu, b->d_mum);
      mup = b->d_mup;
      mum = b->d_mum;
    }
    es_call<R>(b, ES_FINAL, b->d_qp, mup, nullptr, b->d_Fp);
    es_call<R>(b, ES_FINAL, b->d_qm, mum, nullptr, b->d_Fm);
    k_b_contract<<<gb, MDFF_BLOCK, 0, b->st>>>(N, b->d_g, b->d_Fp, b->d_Fm, b->d_part);
    reduce(b, 1, gb, slot);
  }
  for (int m = 0; m < b->npol; ++m, ++slot) {
    k_b_asrc<<<gb, MDFF_BLOCK, 0, b->st>>>(N, b->d_typ, b->pol_type[m], b->pol_alpha[m],
                                           b->d_mu, b->d_S);
    jacobi<R>(b, b->d_S, b->d_dmu, b->d_Et, n_scf, &rms);
    k_b_axpy<<<nblk(n3), MDFF_BLOCK, 0, b->st>>>(n3, b->d_mu, 1.0, b->d_dmu, b->d_mup);
    k_b_axpy<<<nblk(n3), MDFF_BLOCK, 0, b->st>>>(n3, b->d_mu, -1.0, b->d_dmu, b->d_mum);
    es_call<R>(b, ES_FINAL, b->d_q, b->d_mup, nullptr, b->d_Fp);
    es_call<R>(b, ES_FINAL, b->d_q, b->d_mum, nullptr, b->d_Fm);
    k_b_contract<<<gb, MDFF_BLOCK, 0, b->st>>>(N, b->d_g, b->d_Fp, b->d_Fm, b->d_part);
    reduce(b, 1, gb, slot);
  }
  if (b->ndamp > 0) {
    // damping b: explicit derivatives of the static field and the final
    // force (after the parameter slots, 1 + nparam ..), then the tangent
    // dmu = alpha (dE_static/db + T dmu) through the quadratic form
    k_b_damp_adj<<<gbl, MDFF_BLOCK, 0, b->st>>>(es_args(b, b->d_q, b->d_mu), b->dslots, b->d_g,
                                                P.coul_unit, b->ndamp, b->d_dEs, b->d_part);
    launch_check(b, "batch damping adjoint");
    reduce(b, b->ndamp, gbl, 1 + b->nparam);
    for (int m = 0; m < b->ndamp; ++m, ++slot) {
      jacobi<R>(b, b->d_dEs + (size_t)m * n3, b->d_dmu, b->d_Et, n_scf, &rms);
      k_b_axpy<<<nblk(n3), MDFF_BLOCK, 0, b->st>>>(n3, b->d_mu, 1.0, b->d_dmu, b->d_mup);
      k_b_axpy<<<nblk(n3), MDFF_BLOCK, 0, b->st>>>(n3, b->d_mu, -1.0, b->d_dmu, b->d_mum);
      es_call<R>(b, ES_FINAL, b->d_q, b->d_mup, nullptr, b->d_Fp);
      es_call<R>(b, ES_FINAL, b->d_q, b->d_mum, nullptr, b->d_Fm);
      k_b_contract<<<gb, MDFF_BLOCK, 0, b->st>>>(N, b->d_g, b->d_Fp, b->d_Fm, b->d_part);
      reduce(b, 1, gb, slot);
    }
  }
}

void set_model(mdffgpu_batch* b, const mdffgpu_params& P, bool first) {
  if (P.ntype < 1 || P.ntype > 2)
    bfail(b, MDFFGPU_ERR_UNSUPPORTED, "the batched path supports one or two species");
  for (int k = 0; k < 48; ++k)
    if (P.dip[k] != 0.0) bfail(b, MDFFGPU_ERR_UNSUPPORTED, "permanent dipoles in a batch");
  if (P.lsurf) bfail(b, MDFFGPU_ERR_UNSUPPORTED, "surface term in a batch");
  for (int k = 0; k < 144; ++k)
    if (P.quad[k] != 0.0) bfail(b, MDFFGPU_ERR_UNSUPPORTED, "quadrupoles in a batch");
  for (int t = 0; t < 16; ++t)
    if (P.lquad_polar[t] || P.lwfc[t] != 0)
      bfail(b, MDFFGPU_ERR_UNSUPPORTED, "quadrupole polarizabilities or Wannier centres in a batch");
  if (P.task_coul_quad[0] || P.task_coul_quad[1] || P.task_coul_quad[2])
    bfail(b, MDFFGPU_ERR_UNSUPPORTED, "quadrupole tasks in a batch");
  if (P.precision != MDFFGPU_PREC_FP64 && P.precision != MDFFGPU_PREC_MIXED)
    bfail(b, MDFFGPU_ERR_ARG, "unknown precision");
  if (!first) {
    if (P.ntype != b->P.ntype || P.precision != b->P.precision ||
        P.lcoulomb != b->P.lcoulomb || P.lbmhftd != b->P.lbmhftd)
      bfail(b, MDFFGPU_ERR_ARG, "set_model may change parameter values only");
    for (int t = 0; t < 16; ++t)
      if ((P.ldip_polar[t] != 0) != (b->P.ldip_polar[t] != 0))
        bfail(b, MDFFGPU_ERR_ARG, "set_model may change parameter values only");
  }
  const int bad = mdff::make_tables(P, b->pt_h, b->ty_h, b->tt);
  if (bad >= 0)
    bfail(b, MDFFGPU_ERR_TT_ORDER, "Tang-Toennies order %d > %d", bad, MDFFGPU_TT_MAX_ORDER);
  b->P = P;
  b->ntype = P.ntype;
  b->npair = P.ntype * (P.ntype + 1) / 2;
  b->mixed = (P.precision == MDFFGPU_PREC_MIXED);
  b->npol = 0;
  b->pim = false;
  for (int t = 0; t < P.ntype; ++t)
    if (b->ty_h.polar[t]) {
      const double* pm = b->ty_h.pol[t];
      if (pm[0] <= 0.0 || pm[0] != pm[4] || pm[0] != pm[8] || pm[1] != 0.0 || pm[2] != 0.0 ||
          pm[3] != 0.0 || pm[5] != 0.0 || pm[6] != 0.0 || pm[7] != 0.0)
        bfail(b, MDFFGPU_ERR_UNSUPPORTED, "batched polarizabilities must be isotropic");
      b->pol_type[b->npol] = t;
      b->pol_alpha[b->npol] = pm[0];
      ++b->npol;
      b->pim = true;
    }
  if (b->pim && !P.lcoulomb) bfail(b, MDFFGPU_ERR_ARG, "polarizable ions need lcoulomb");
  // the damping b of each species whose own field terms are damped: the
  // entries its centres use, (t, t, u) when i < j and (t, u, t) otherwise
  // (src/field.f90:2207-2211), on damped pairs with c != 0 -- all with one
  // value of b per species
  int ndamp = 0;
  int dtype[NDAMP_MAX] = {0, 0};
  DampSlots ds;
  for (int t = 0; t < MDFF_NT; ++t) ds.slot[t] = -1;
  auto L3 = [&](int x, int y, int z) { return P.ldip_damping[x + 16 * (y + 16 * z)] != 0; };
  auto I3 = [](int x, int y, int z) { return x + 16 * (y + 16 * z); };
  if (b->pim)
    for (int t = 0; t < P.ntype; ++t) {
      bool on = false;
      double b0 = 0.0;
      for (int u = 0; u < P.ntype; ++u) {
        const int e[2] = {I3(t, t, u), I3(t, u, t)};
        const bool act[2] = {L3(t, t, u) || L3(u, t, u), L3(u, u, t) || L3(t, u, t)};
        for (int o = 0; o < 2; ++o) {
          if (!act[o] || P.dip_pol_damp_c[e[o]] == 0.0) continue;
          if (!on) b0 = P.dip_pol_damp_b[e[o]];
          if (P.dip_pol_damp_b[e[o]] != b0)
            bfail(b, MDFFGPU_ERR_UNSUPPORTED,
                  "the batch fits one damping b per species: dip_pol_damp_b(%d,:,:) differ", t + 1);
          on = true;
        }
      }
      if (!on) continue;
      ds.slot[t] = ndamp;
      dtype[ndamp++] = t;
    }
  if (!first && (ndamp != b->ndamp || (P.lmorse != 0) != b->lmo))
    bfail(b, MDFFGPU_ERR_ARG, "set_model may change parameter values only");
  b->ndamp = ndamp;
  b->dslots = ds;
  for (int m = 0; m < NDAMP_MAX; ++m) b->damp_type[m] = dtype[m];
  b->lmo = P.lmorse != 0;
  b->nparam = (NSR + (b->lmo ? NMO : 0)) * b->npair + b->ntype + b->npol + b->ndamp;
  int p = 0;
  for (int x = 0; x < 256; ++x) b->upair_h[x] = 0;
  for (int ti = 0; ti < P.ntype; ++ti)
    for (int tj = ti; tj < P.ntype; ++tj, ++p) {
      b->upair_h[ti + 16 * tj] = p;
      b->upair_h[tj + 16 * ti] = p;
    }
  if (first) {
    balloc(b, &b->d_pt, 1);
    balloc(b, &b->d_ty, 1);
    balloc(b, &b->d_upair, 256);
  }
  BCK(b, hipMemcpy(b->d_pt, &b->pt_h, sizeof(b->pt_h), hipMemcpyHostToDevice));
  BCK(b, hipMemcpy(b->d_ty, &b->ty_h, sizeof(b->ty_h), hipMemcpyHostToDevice));
  BCK(b, hipMemcpy(b->d_upair, b->upair_h, sizeof(b->upair_h), hipMemcpyHostToDevice));
  if (!first) {
    // charges per atom follow the new qch
    std::vector<double> q(b->N);
    std::vector<int> typ(b->N);
    BCK(b, hipMemcpy(typ.data(), b->d_typ, b->N * sizeof(int), hipMemcpyDeviceToHost));
    for (int i = 0; i < b->N; ++i) q[i] = P.qch[typ[i]];
    BCK(b, hipMemcpy(b->d_q, q.data(), b->N * sizeof(double), hipMemcpyHostToDevice));
  }
  if (b->gexec) {
    (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
    b->graph_iters = -1;
  }
}

void create_impl(const mdffgpu_params* model, int nconf, const mdffgpu_batch_config* cfgs,
                 const int32_t* itype, mdffgpu_batch* b) {
  if (!model || !cfgs || !itype || nconf < 1) bfail(b, MDFFGPU_ERR_ARG, "bad batch arguments");
  set_model(b, *model, true);
  b->nconf = nconf;
  b->bc_h.resize(nconf);
  int off = 0;
  for (int c = 0; c < nconf; ++c) {
    const mdffgpu_batch_config& g = cfgs[c];
    BatchCfg& x = b->bc_h[c];
    std::memset(&x, 0, sizeof(x));
    if (g.natm < 1) bfail(b, MDFFGPU_ERR_ARG, "configuration %d has no atoms", c);
    double W[3];
    if (!make_cell(g.A, x.cell, W)) bfail(b, MDFFGPU_ERR_ARG, "configuration %d: bad cell", c);
    const double wmin = std::min(W[0], std::min(W[1], W[2]));
    if (g.cutlongrange > 0.5 * wmin || g.cutshortrange > 0.5 * wmin)
      bfail(b, MDFFGPU_ERR_ARG, "configuration %d: cutoff above half the cell width", c);
    x.alpha = g.alphaES;
    x.cut2 = g.cutlongrange * g.cutlongrange;
    x.rc2 = g.cutshortrange * g.cutshortrange;
    x.fpi_V = 4.0 * MDFF_PI / x.cell.omega;
    x.selfa2 = 2.0 * g.alphaES * g.alphaES * g.alphaES / (3.0 * std::sqrt(MDFF_PI));
    x.n = g.natm;
    x.off = off;
    for (int a = 0; a < 3; ++a) {
      if (g.kmax[a] < 1 || g.kmax[a] > 12)
        bfail(b, MDFFGPU_ERR_ARG, "configuration %d: kmax must be 1..12", c);
      x.kmax[a] = g.kmax[a];
    }
    b->KX = std::max(b->KX, g.kmax[0]);
    b->KY = std::max(b->KY, g.kmax[1]);
    b->KZ = std::max(b->KZ, g.kmax[2]);
    off += g.natm;
  }
  b->N = off;
  const int N = b->N;
  b->ntab = (b->KX + 1) + (2 * b->KY + 1) + (2 * b->KZ + 1);
  // k lists (half space of kpoint_sum_init, src/kspace.f90:162-191)
  std::vector<int4> kn;
  std::vector<double4> kv;
  std::vector<int2> items;
  for (int c = 0; c < nconf; ++c) {
    BatchCfg& x = b->bc_h[c];
    x.koff = (int)kn.size();
    const double* B = x.cell.B;
    const double a2 = x.alpha * x.alpha;
    for (int n1 = 0; n1 <= x.kmax[0]; ++n1)
      for (int n2 = -x.kmax[1]; n2 <= x.kmax[1]; ++n2)
        for (int n3 = -x.kmax[2]; n3 <= x.kmax[2]; ++n3) {
          const bool half = n1 > 0 || (n1 == 0 && n2 > 0) || (n1 == 0 && n2 == 0 && n3 > 0);
          if (!half) continue;
          const double tpi = 2.0 * MDFF_PI;
          const double kx = tpi * (n1 * B[0] + n2 * B[3] + n3 * B[6]);
          const double ky = tpi * (n1 * B[1] + n2 * B[4] + n3 * B[7]);
          const double kz = tpi * (n1 * B[2] + n2 * B[5] + n3 * B[8]);
          const double kk = kx * kx + ky * ky + kz * kz;
          kn.push_back(make_int4(n1, (b->KX + 1) + (n2 + b->KY),
                                 (b->KX + 1) + (2 * b->KY + 1) + (n3 + b->KZ), 0));
          kv.push_back(make_double4(kx, ky, kz, std::exp(-kk * 0.25 / a2) / kk));
        }
    x.nk = (int)kn.size() - x.koff;
    for (int k0 = 0; k0 < x.nk; k0 += BK) items.push_back(make_int2(c, k0));
  }
  b->NK = (int)kn.size();
  b->nitems = (int)items.size();
  std::vector<int> cfg(N), typ(N);
  std::vector<double> q(N);
  for (int c = 0; c < nconf; ++c)
    for (int i = 0; i < b->bc_h[c].n; ++i) cfg[b->bc_h[c].off + i] = c;
  for (int i = 0; i < N; ++i) {
    if (itype[i] < 1 || itype[i] > model->ntype)
      bfail(b, MDFFGPU_ERR_ARG, "itype(%d) = %d outside 1..ntype", i + 1, itype[i]);
    typ[i] = itype[i] - 1;
    q[i] = model->qch[typ[i]];
  }
  BCK(b, hipStreamCreateWithFlags(&b->own, hipStreamNonBlocking));
  b->st = b->own;
  BCK(b, hipHostMalloc((void**)&b->h_out, 64 * sizeof(double), hipHostMallocDefault));
  balloc(b, &b->d_bc, nconf);
  balloc(b, &b->d_cfg, N);
  balloc(b, &b->d_typ, N);
  balloc(b, &b->d_kn, b->NK);
  balloc(b, &b->d_kv, b->NK);
  balloc(b, &b->d_items, b->nitems);
  balloc(b, &b->d_rho, b->NK);
  BCK(b, hipMemcpy(b->d_bc, b->bc_h.data(), nconf * sizeof(BatchCfg), hipMemcpyHostToDevice));
  BCK(b, hipMemcpy(b->d_cfg, cfg.data(), N * sizeof(int), hipMemcpyHostToDevice));
  BCK(b, hipMemcpy(b->d_typ, typ.data(), N * sizeof(int), hipMemcpyHostToDevice));
  // the polarizable atoms (rows of the SCF pair table) and the pipeline
  // depth of k_dd_tab2 for the longest expected row: the configuration with
  // the most polarizable neighbours inside its cutoff, 15 % to spare
  b->polmask = 0;
  for (int t = 0; t < b->ntype; ++t)
    if (b->ty_h.polar[t]) b->polmask |= 1u << t;
  std::vector<int> prow;
  std::vector<int> npc(nconf, 0);
  for (int i = 0; i < N; ++i)
    if ((b->polmask >> typ[i]) & 1u) {
      prow.push_back(i);
      ++npc[cfg[i]];
    }
  b->npolrows = (int)prow.size();
  double row_max = 0.0;
  for (int c = 0; c < nconf; ++c) {
    const double rc3 = std::pow(b->bc_h[c].cut2, 1.5);
    row_max = std::max(row_max, npc[c] / b->bc_h[c].cell.omega * (4.0 / 3.0) * MDFF_PI * rc3);
  }
  const double need = 1.15 * row_max / 64.0;
  b->dd_u = need <= 4.0 ? 4 : need <= 6.0 ? 6 : 8;
  balloc(b, &b->d_polrows, std::max(1, b->npolrows));
  if (b->npolrows > 0)
    BCK(b, hipMemcpy(b->d_polrows, prow.data(), prow.size() * sizeof(int), hipMemcpyHostToDevice));
  BCK(b, hipMemcpy(b->d_kn, kn.data(), kn.size() * sizeof(int4), hipMemcpyHostToDevice));
  BCK(b, hipMemcpy(b->d_kv, kv.data(), kv.size() * sizeof(double4), hipMemcpyHostToDevice));
  BCK(b, hipMemcpy(b->d_items, items.data(), items.size() * sizeof(int2),
                   hipMemcpyHostToDevice));
  const size_t n3 = 3 * (size_t)N;
  balloc(b, &b->d_pos, n3);
  balloc(b, &b->d_s, n3);
  balloc(b, &b->d_s4, N);
  balloc(b, &b->d_mu4, N);
  balloc(b, &b->d_tab, (size_t)b->ntab * N);
  balloc(b, &b->d_cnt, N);
  balloc(b, &b->d_ovf, 1);
  balloc(b, &b->d_q, N);
  balloc(b, &b->d_zero, n3);
  double** vec3[] = {&b->d_mu, &b->d_E, &b->d_Es, &b->d_dir, &b->d_rec, &b->d_frec,
                     &b->d_Fel, &b->d_Fsr, &b->d_F, &b->d_g, &b->d_Ft, &b->d_dmu,
                     &b->d_Et, &b->d_S, &b->d_mup, &b->d_mum, &b->d_Fp, &b->d_Fm};
  for (double** p : vec3) balloc(b, p, n3);
  balloc(b, &b->d_qsrc, N);
  balloc(b, &b->d_qp, N);
  balloc(b, &b->d_qm, N);
  balloc(b, &b->d_out, 64);
  if (b->ndamp > 0) balloc(b, &b->d_dEs, (size_t)b->ndamp * n3);
  BCK(b, hipMemcpy(b->d_q, q.data(), N * sizeof(double), hipMemcpyHostToDevice));
  BCK(b, hipMemset(b->d_zero, 0, n3 * sizeof(double)));
  BCK(b, hipMemset(b->d_mu, 0, n3 * sizeof(double)));
  ensure_part(b, (size_t)(NSR * NPAIR_MAX + 8) * nblk((size_t)N * LPA) + 4096);
  // neighbour capacity from the densest configuration
  double est = 16.0;
  for (const BatchCfg& x : b->bc_h) {
    const double rl = std::sqrt(std::max(x.cut2, x.rc2));
    est = std::max(est, 4.0 / 3.0 * MDFF_PI * rl * rl * rl * x.n / x.cell.omega);
  }
  b->maxnb = std::min((int)(1.4 * est) + 32, N);
}

template <typename R>
void build_pairs(mdffgpu_batch* b) {
  EsDirArgs da = es_args(b, b->d_q, b->d_zero);
  // FP64: five doubles per list entry; FP32: the fixed-point positions
  // (uint4 per atom), compacted (A2, A1) float2 + index per entry, then the
  // per-atom counts
  const size_t cap = (size_t)b->maxnb * b->N;
  const size_t need = sizeof(R) == sizeof(double)
                          ? 5 * cap * sizeof(double)
                          : (size_t)b->N * sizeof(uint4) + cap * (sizeof(float2) + sizeof(int)) +
                                (size_t)b->N * sizeof(int);
  if (need > b->t2_cap) {
    if (b->d_t2) {
      (void)hipFree(b->d_t2);
      b->allocs.erase(std::find(b->allocs.begin(), b->allocs.end(), b->d_t2));
    }
    b->d_t2 = nullptr;
    BCK(b, hipMalloc(&b->d_t2, need));
    b->allocs.push_back(b->d_t2);
    b->t2_cap = need;
  }
  if constexpr (sizeof(R) == sizeof(double)) {
    k_pairs_t2<true><<<nblk((size_t)b->N * LPA), MDFF_BLOCK, 0, b->st>>>(da, (double*)b->d_t2);
  } else {
    uint4* u4 = reinterpret_cast<uint4*>(b->d_t2);
    float2* tab = reinterpret_cast<float2*>(u4 + b->N);
    int* tj = reinterpret_cast<int*>(tab + cap);
    da.dd16 = dd_idx16(b->N);
    // fractional coordinates in each configuration's own cell: the fixed-
    // point difference is that cell's minimum image
    k_pack_u4<<<nblk(b->N), MDFF_BLOCK, 0, b->st>>>(b->N, b->d_s, nullptr, u4);
    k_dd_table<true><<<nblk((size_t)b->N * 64), MDFF_BLOCK, 0, b->st>>>(da, tab, tj,
                                                                                tj + cap);
  }
  launch_check(b, "batch pair table");
}

void set_positions_impl(mdffgpu_batch* b, const double* d_pos) {
  const int N = b->N;
  BCK(b, hipMemcpyAsync(b->d_pos, d_pos, 3 * (size_t)N * sizeof(double),
                        hipMemcpyDeviceToDevice, b->st));
  k_b_prepare<<<nblk(N), MDFF_BLOCK, 0, b->st>>>(N, b->d_pos, b->d_cfg, b->d_bc, b->KX, b->KY,
                                                 b->KZ, b->d_s, b->d_s4, b->d_tab);
  launch_check(b, "batch prepare");
  for (int attempt = 0; attempt < 2; ++attempt) {
    if (b->d_nbr) {
      (void)hipFree(b->d_nbr);
      b->allocs.erase(std::find(b->allocs.begin(), b->allocs.end(), (void*)b->d_nbr));
      b->d_nbr = nullptr;
    }
    balloc(b, &b->d_nbr, (size_t)b->maxnb * N);
    BCK(b, hipMemsetAsync(b->d_ovf, 0, sizeof(int), b->st));
    k_b_nlist<<<nblk(N), MDFF_BLOCK, 0, b->st>>>(N, b->d_cfg, b->d_bc, b->d_s4, b->maxnb,
                                                 b->d_cnt, b->d_nbr, b->d_ovf);
    launch_check(b, "batch neighbour list");
    int ovf = 0;
    BCK(b, hipMemcpyAsync(&ovf, b->d_ovf, sizeof(int), hipMemcpyDeviceToHost, b->st));
    BCK(b, hipStreamSynchronize(b->st));
    if (ovf == 0) break;
    if (attempt == 1) bfail(b, MDFFGPU_ERR_NLIST, "neighbour list overflow (%d)", ovf);
    b->maxnb = ovf + 16;
  }
  if (b->pim) {
    if (b->mixed)
      build_pairs<float>(b);
    else
      build_pairs<double>(b);
  }
  b->positions_set = true;
  ++b->epoch;
  if (b->mixed) mx_state(b);  // allocates: never first inside a graph capture
  if (b->gexec) {
    (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
    b->graph_iters = -1;
  }
}

// forward + backward on b->st; results in d_out, copied to h_out
void loss_grad_seq(mdffgpu_batch* b, int iters, int* n_scf, double* rmsd) {
  if (b->mixed) {
    *n_scf = forward<float>(b, iters, rmsd);
    backward<float>(b, *n_scf);
  } else {
    *n_scf = forward<double>(b, iters, rmsd);
    backward<double>(b, *n_scf);
  }
  BCK(b, hipMemcpyAsync(b->h_out, b->d_out, (1 + b->nparam + b->ndamp) * sizeof(double),
                        hipMemcpyDeviceToHost, b->st));
}

}  // namespace
}  // namespace batch
}  // namespace mdff

// ===========================================================================
// C ABI
// ===========================================================================
using namespace mdff::batch;

extern "C" {

int mdffgpu_batch_create(const mdffgpu_params* model, int32_t nconf,
                         const mdffgpu_batch_config* cfgs, const int32_t* itype,
                         mdffgpu_batch** out) {
  if (!out) return MDFFGPU_ERR_ARG;
  *out = nullptr;
  mdffgpu_batch* b = new mdffgpu_batch();
  try {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
      bfail(b, MDFFGPU_ERR_DEVICE, "no HIP device visible");
    create_impl(model, nconf, cfgs, itype, b);
    *out = b;
    return MDFFGPU_OK;
  } catch (BFail& f) {
    g_batch_create_error = b->err;
    mdffgpu_batch_destroy(b);
    return f.code;
  } catch (std::exception& e) {
    g_batch_create_error = e.what();
    mdffgpu_batch_destroy(b);
    return MDFFGPU_ERR_ALLOC;
  }
}

void mdffgpu_batch_destroy(mdffgpu_batch* b) {
  if (!b) return;
  if (b->st) (void)hipStreamSynchronize(b->st);
  if (b->mx) mdff::mixed::destroy(b->mx);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  for (void* p : b->allocs) (void)hipFree(p);
  if (b->h_out) (void)hipHostFree(b->h_out);
  if (b->own) (void)hipStreamDestroy(b->own);
  delete b;
}

int mdffgpu_batch_set_model(mdffgpu_batch* b, const mdffgpu_params* model) {
  if (!b || !model) return MDFFGPU_ERR_ARG;
  try {
    set_model(b, *model, false);
    return MDFFGPU_OK;
  } catch (BFail& f) {
    return f.code;
  }
}

int mdffgpu_batch_set_positions(mdffgpu_batch* b, const double* d_pos, void* stream) {
  if (!b || !d_pos) return MDFFGPU_ERR_ARG;
  try {
    if (stream) BCK(b, hipStreamSynchronize((hipStream_t)stream));
    set_positions_impl(b, d_pos);
    return MDFFGPU_OK;
  } catch (BFail& f) {
    return f.code;
  }
}

int mdffgpu_batch_forces(mdffgpu_batch* b, double* d_force, int32_t scf_iters,
                         int32_t* n_scf, double* rmsd, void* stream) {
  if (!b || !d_force || scf_iters < 0) return MDFFGPU_ERR_ARG;
  try {
    if (!b->positions_set) bfail(b, MDFFGPU_ERR_ARG, "positions not set");
    if (stream) BCK(b, hipStreamSynchronize((hipStream_t)stream));
    double r = 0.0;
    const int ns = b->mixed ? forward<float>(b, scf_iters, &r) : forward<double>(b, scf_iters, &r);
    k_b_axpy<<<nblk(3 * (size_t)b->N), MDFF_BLOCK, 0, b->st>>>(3 * (size_t)b->N, b->d_Fsr, 1.0,
                                                               b->d_Fel, d_force);
    launch_check(b, "batch forces");
    BCK(b, hipStreamSynchronize(b->st));
    if (n_scf) *n_scf = ns;
    if (rmsd) *rmsd = r;
    return MDFFGPU_OK;
  } catch (BFail& f) {
    return f.code;
  }
}

int32_t mdffgpu_batch_num_params(const mdffgpu_batch* b) { return b ? b->nparam : 0; }

int64_t mdffgpu_batch_num_atoms(const mdffgpu_batch* b) { return b ? b->N : 0; }

int mdffgpu_batch_loss_grad(mdffgpu_batch* b, const double* d_target, double* d_force,
                            double* loss, double* grad, int32_t scf_iters, int32_t* n_scf,
                            double* rmsd, void* stream) {
  if (!b || !d_target || !loss || !grad || scf_iters < 0) return MDFFGPU_ERR_ARG;
  try {
    if (!b->positions_set) bfail(b, MDFFGPU_ERR_ARG, "positions not set");
    if (b->graph_on && scf_iters == 0 && b->pim)
      bfail(b, MDFFGPU_ERR_ARG, "graph replay needs a fixed SCF iteration count");
    if (stream) BCK(b, hipStreamSynchronize((hipStream_t)stream));
    const size_t n3 = 3 * (size_t)b->N;
    BCK(b, hipMemcpyAsync(b->d_Ft, d_target, n3 * sizeof(double), hipMemcpyDeviceToDevice,
                          b->st));
    int ns = 0;
    double r = 0.0;
    if (b->graph_on) {
      if (!b->gexec || b->graph_iters != scf_iters) {
        if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
        b->gexec = nullptr;
        hipGraph_t g = nullptr;
        BCK(b, hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
        try {
          loss_grad_seq(b, scf_iters, &ns, &r);
        } catch (...) {
          (void)hipStreamEndCapture(b->st, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        BCK(b, hipStreamEndCapture(b->st, &g));
        hipError_t e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess)
          bfail(b, MDFFGPU_ERR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
        b->graph_iters = scf_iters;
      }
      BCK(b, hipGraphLaunch(b->gexec, b->st));
      ns = scf_iters;
    } else {
      loss_grad_seq(b, scf_iters, &ns, &r);
    }
    if (d_force)
      BCK(b, hipMemcpyAsync(d_force, b->d_F, n3 * sizeof(double), hipMemcpyDeviceToDevice,
                            b->st));
    BCK(b, hipStreamSynchronize(b->st));
    *loss = b->h_out[0] / (3.0 * b->N);
    // pair parameters directly; charges, polarizabilities and the dipole
    // tangent of the damping b as 2 B(x, x') (halved), plus b's explicit part
    const int nsr = (NSR + (b->lmo ? NMO : 0)) * b->npair;
    for (int k = 0; k < nsr; ++k) grad[k] = b->h_out[1 + k];
    for (int k = nsr; k < b->nparam; ++k) grad[k] = 0.5 * b->h_out[1 + k];
    for (int m = 0; m < b->ndamp; ++m)
      grad[b->nparam - b->ndamp + m] += b->h_out[1 + b->nparam + m];
    if (n_scf) *n_scf = ns;
    if (rmsd) *rmsd = r;
    return MDFFGPU_OK;
  } catch (BFail& f) {
    return f.code;
  }
}

int mdffgpu_batch_graph(mdffgpu_batch* b, int32_t enable) {
  if (!b) return MDFFGPU_ERR_ARG;
  b->graph_on = enable != 0;
  if (!b->graph_on && b->gexec) {
    (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
    b->graph_iters = -1;
  }
  return MDFFGPU_OK;
}

int mdffgpu_batch_set_tuning(mdffgpu_batch* b, int32_t apply_lanes) {
  if (!b) return MDFFGPU_ERR_ARG;
  if (apply_lanes != 2 && apply_lanes != 4 && apply_lanes != 8 && apply_lanes != 16)
    return MDFFGPU_ERR_ARG;
  if (apply_lanes != b->apply_lanes && b->gexec) {
    (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
    b->graph_iters = -1;
  }
  b->apply_lanes = apply_lanes;
  return MDFFGPU_OK;
}

int mdffgpu_batch_set_mixed_tuning(mdffgpu_batch* b, int32_t apply_slices, int32_t split_atoms) {
  if (!b) return MDFFGPU_ERR_ARG;
  if (!b->mx) return MDFFGPU_ERR_UNSUPPORTED;
  if (b->gexec) {
    (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
    b->graph_iters = -1;
  }
  if (b->st) (void)hipStreamSynchronize(b->st);  // buffers may be replaced
  return mdff::mixed::set_tuning(b->mx, apply_slices, split_atoms) ? MDFFGPU_OK
                                                                   : MDFFGPU_ERR_ARG;
}

const char* mdffgpu_batch_last_error(const mdffgpu_batch* b) {
  if (!b) return g_batch_create_error.c_str();
  return b->err.c_str();
}

}  // extern "C"
