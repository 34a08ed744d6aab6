// quad.hip -- quadrupole terms of the real-space Ewald sum and the
// Wannier-centre dipoles, FP64.
//
// multipole_ES_dir (src/field.f90:2542-2601) adds, per pair (ia, ja) of its
// half list with r = r_ia - r_ja and the screened tensors T2, T3, T4 of
// :2248-2318 (F1..F4 of the erfc recursion :2219-2223):
//   q-Q   u += q_i T2:Q_j / 3 + q_j T2:Q_i / 3
//   d-Q   u += mu_i.T3:Q_j / 3 - mu_j.T3:Q_i / 3
//   Q-Q   u += Q_i:T4:Q_j / 9
//         E(ia) -= T3:Q_j,  E(ja) += T3:Q_i
//         EFG(ia) += T4:Q_j, EFG(ja) -= T4:Q_i
// and no force or stress.  Every energy term is symmetric under the swap
// of the pair, so each centre atom takes half of it; the fields follow from
// T3(-r) = -T3(r); the quadrupole-quadrupole EFG takes the sign of the
// pair's order, + at the smaller index ('ia' of the all-pairs loop i < j) in
// the caller's numbering (s4.w).
// With symmetric Q the contractions reduce to u = Q r, s = r.Q.r and
// tr Q:
//   T2:Q      = s F2' - tr F1'
//   (T3:Q)_a  = -r_a s F3' + (r_a tr + 2 u_a) F2'
//   (T4:Q)_ab = r_a r_b s F4' - (r_a r_b tr + 2 r_a u_b + 2 r_b u_a
//               + d_ab s) F3' + (d_ab tr + 2 Q_ab) F2'
//   Q_i:T4:Q_j = s_i s_j F4' - (s_i tr_j + tr_i s_j + 4 u_i.u_j) F3'
//               + (tr_i tr_j + 2 Q_i:Q_j) F2'
// with F1' = F1/d^3, F2' = 3 F2/d^5, F3' = 15 F3/d^7, F4' = 105 F4/d^9.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace mdff {

namespace {
// symmetric part of Q as xx yy zz xy xz yz, from the (n,3,3) array
__device__ __forceinline__ void load_q6(const double* th, int n, int i, double (&q)[6]) {
  q[0] = th[i];
  q[1] = th[i + 4 * (size_t)n];
  q[2] = th[i + 8 * (size_t)n];
  q[3] = 0.5 * (th[i + 3 * (size_t)n] + th[i + 1 * (size_t)n]);
  q[4] = 0.5 * (th[i + 6 * (size_t)n] + th[i + 2 * (size_t)n]);
  q[5] = 0.5 * (th[i + 7 * (size_t)n] + th[i + 5 * (size_t)n]);
}

__device__ __forceinline__ void q_dot_r(const double (&q)[6], double rx, double ry, double rz,
                                        double (&u)[3]) {
  u[0] = q[0] * rx + q[3] * ry + q[4] * rz;
  u[1] = q[3] * rx + q[1] * ry + q[5] * rz;
  u[2] = q[4] * rx + q[5] * ry + q[2] * rz;
}
}  // namespace

// One group of LPA lanes per centre atom (atoms in spatial order), the
// neighbours of the full list strided over the lanes.  Fields and EFGs are
// added to a.ef / a.efg (Fortran (n,3) and (n,3,3)); energy block partials
// go to a.part.
__global__ __launch_bounds__(MDFF_BLOCK) void k_es_quad(QuadArgs a) {
  __shared__ double red[4];
  const int n = a.n;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int i0 = tid / LPA, sub = tid % LPA;
  const int i = (a.perm && i0 < n) ? a.perm[i0] : i0;
  const bool QQU = a.do_energy && (a.task & TASK_QQUAD), DQU = a.do_energy && (a.task & TASK_DQUAD);
  const bool QUQ = a.task & TASK_QUADQ;
  const bool EF = QUQ && a.do_efield, EFG = QUQ && a.do_efg;
  const double alpha = a.alpha, alpha2 = alpha * alpha, alpha3 = alpha2 * alpha,
               alpha5 = alpha3 * alpha2, alpha7 = alpha5 * alpha2;
  const double rpiroot = 1.0 / sqrt(MDFF_PI);
  double uacc = 0.0;
  double e[3] = {0.0, 0.0, 0.0};
  double g[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (i < n) {
    const double si1 = a.s[i], si2 = a.s[i + n], si3 = a.s[i + 2 * n];
    const double wi = a.s4[i].w;  // the caller's index of atom i
    const double qi = a.q[i];
    const double mi[3] = {a.mu[i], a.mu[i + n], a.mu[i + 2 * n]};
    double Qi[6];
    load_q6(a.th, n, i, Qi);
    const double tri = Qi[0] + Qi[1] + Qi[2];
    const int c = a.cnt[i];
    const int* nb = a.nbr + (size_t)i * a.maxnb;
    for (int k = sub; k < c; k += LPA) {
      const int j = nb[k];
      const double4 sj = a.s4[j];
      double rx, ry, rz;
      const double d2 = min_image(a.cell, si1 - sj.x, si2 - sj.y, si3 - sj.z, rx, ry, rz);
      if (d2 > a.cut2) continue;
      const double d = sqrt(d2);
      const double d3 = d2 * d, d5 = d3 * d2, d7 = d5 * d2;
      const double dm1 = 1.0 / d, dm3 = dm1 / d2, dm5 = dm3 / d2, dm7 = dm5 / d2,
                   dm9 = dm7 / d2;
      const double expon = exp(-alpha2 * d2) * rpiroot;
      const double F0 = errfc_nr(alpha * d);
      const double F1 = F0 + 2.0 * alpha * d * expon;
      const double F2 = F1 + 4.0 * alpha3 * d3 * expon / 3.0;
      const double F3 = F2 + 8.0 * alpha5 * d5 * expon / 15.0;
      const double F4 = F3 + 16.0 * alpha7 * d7 * expon / 105.0;
      const double f1 = F1 * dm3, f2 = 3.0 * F2 * dm5, f3 = 15.0 * F3 * dm7,
                   f4 = 105.0 * F4 * dm9;
      double Qj[6];
      load_q6(a.th, n, j, Qj);
      const double trj = Qj[0] + Qj[1] + Qj[2];
      double ui[3], uj[3];
      q_dot_r(Qi, rx, ry, rz, ui);
      q_dot_r(Qj, rx, ry, rz, uj);
      const double r3[3] = {rx, ry, rz};
      const double s_i = rx * ui[0] + ry * ui[1] + rz * ui[2];
      const double s_j = rx * uj[0] + ry * uj[1] + rz * uj[2];
      if (QQU) {
        const double qj = a.q[j];
        uacc += 0.5 * (qi * (s_j * f2 - trj * f1) + qj * (s_i * f2 - tri * f1)) / 3.0;
      }
      // (T3:Q_j)_a
      double v3[3];
      for (int x = 0; x < 3; ++x) v3[x] = -r3[x] * s_j * f3 + (r3[x] * trj + 2.0 * uj[x]) * f2;
      if (DQU) {
        const double mj[3] = {a.mu[j], a.mu[j + n], a.mu[j + 2 * n]};
        double w3[3];
        for (int x = 0; x < 3; ++x) w3[x] = -r3[x] * s_i * f3 + (r3[x] * tri + 2.0 * ui[x]) * f2;
        const double t1 = mi[0] * v3[0] + mi[1] * v3[1] + mi[2] * v3[2];
        const double t2 = mj[0] * w3[0] + mj[1] * w3[1] + mj[2] * w3[2];
        uacc += 0.5 * (t1 - t2) / 3.0;
      }
      if (QUQ) {
        const double uu = ui[0] * uj[0] + ui[1] * uj[1] + ui[2] * uj[2];
        const double qq = Qi[0] * Qj[0] + Qi[1] * Qj[1] + Qi[2] * Qj[2] +
                          2.0 * (Qi[3] * Qj[3] + Qi[4] * Qj[4] + Qi[5] * Qj[5]);
        if (a.do_energy)
          uacc += 0.5 * (s_i * s_j * f4 - (s_i * trj + tri * s_j + 4.0 * uu) * f3 +
                         (tri * trj + 2.0 * qq) * f2) / 9.0;
        if (EF)
          for (int x = 0; x < 3; ++x) e[x] -= v3[x];
        if (EFG) {
          const double sg = (wi < sj.w) ? 1.0 : -1.0;
          // (T4:Q_j)_ab for ab = xx yy zz xy xz yz
          const int ia_[6] = {0, 1, 2, 0, 0, 1}, ib_[6] = {0, 1, 2, 1, 2, 2};
          for (int m = 0; m < 6; ++m) {
            const int x = ia_[m], y = ib_[m];
            const double dxy = (x == y) ? 1.0 : 0.0;
            const double v = r3[x] * r3[y] * s_j * f4 -
                             (r3[x] * r3[y] * trj + 2.0 * r3[x] * uj[y] + 2.0 * r3[y] * uj[x] +
                              dxy * s_j) * f3 +
                             (dxy * trj + 2.0 * Qj[m]) * f2;
            g[m] += sg * v;
          }
        }
      }
    }
  }
  for (int off = LPA / 2; off > 0; off >>= 1) {
    if (EF)
      for (int x = 0; x < 3; ++x) e[x] += __shfl_xor(e[x], off, 64);
    if (EFG)
      for (int m = 0; m < 6; ++m) g[m] += __shfl_xor(g[m], off, 64);
  }
  if (i < n && sub == 0) {
    if (EF)
      for (int x = 0; x < 3; ++x) a.ef[i + (size_t)x * n] += e[x];
    if (EFG) {
      const int map[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};  // Fortran (a,b) -> xx yy zz xy xz yz
      for (int k = 0; k < 9; ++k) a.efg[i + (size_t)k * n] += g[map[k]];
    }
  }
  if (a.do_energy) {
    double acc[1] = {uacc};
    block_sum<1>(acc, red);
    if (threadIdx.x == 0) a.part[blockIdx.x] = acc[0];
  }
}

// moment_from_WFc (src/field.f90:3492-3672): for an ion of a species with
// lwfc > 0, the centres (atoms of species with lwfc < 0) whose squared
// minimum-image distance is below rcut (the reference compares the square
// with rcut_wfc, :3572) give mu = -2 sum (r_c - r_ion) (:3606-3629); a count
// other than lwfc is the reference's STOP (:3594-3599), reported through
// bad.  Ions of species with lwfc = 0 and the centres get zeros.
__global__ __launch_bounds__(MDFF_BLOCK) void k_wfc(WfcArgs a) {
  const int n = a.n;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int t = a.typ[i];
  double acc[3] = {0.0, 0.0, 0.0};
  if (a.lwfc[t] > 0) {
    const double si1 = a.s[i], si2 = a.s[i + n], si3 = a.s[i + 2 * n];
    const int cnt = a.nbr ? a.cnt[i] : n;
    int found = 0;
    for (int k = 0; k < cnt; ++k) {
      const int j = a.nbr ? a.nbr[(size_t)i * a.maxnb + k] : k;
      if (a.lwfc[a.typ[j]] >= 0) continue;
      const double sj1 = a.s[j], sj2 = a.s[j + n], sj3 = a.s[j + 2 * n];
      double rx, ry, rz;
      const double d2 = min_image(a.cell, si1 - sj1, si2 - sj2, si3 - sj3, rx, ry, rz);
      if (!(d2 < a.rcut)) continue;
      ++found;
      min_image(a.cell, sj1 - si1, sj2 - si2, sj3 - si3, rx, ry, rz);
      acc[0] += rx;
      acc[1] += ry;
      acc[2] += rz;
    }
    if (found != a.lwfc[t]) atomicMax(a.bad, (a.ext ? a.ext[i] : i) + 1);
  }
  for (int x = 0; x < 3; ++x) a.mu[i + (size_t)x * n] = -2.0 * acc[x];
}

}  // namespace mdff
