// efg.hip -- standalone electric field gradients (calc = 'efg') and the
// per-site NMR parameters of their statistics, FP64 on gfx950.
//
//   k_efg_mask     moments of the contributing atoms (lefg_it_contrib)
//   k_efg_es_dir   real-space part of multipole_efg_ES (src/efg.f90:1401-1499):
//                  every pair j != i under the minimum image, no cutoff
//   k_efg_ds       multipole_efg_DS (src/efg.f90:1151-1237): bare tensors
//                  over image cells, pairs closer than the cutoff
//   k_efg_combine  real + reciprocal + self + charged-cell correction,
//                  symmetrised, units and sign (src/efg.f90:1560-1612, 484-488)
//   k_efg_nmr      Czjzek components, principal values in the NMR order and
//                  the asymmetry (efg_stat :2233-2269, nmr_convention
//                  :2546-2648)
//
// The all-pairs sum is compute-bound FP64 VALU work (two exponentials, one
// reciprocal, one square root and about 120 further operations per pair):
// each thread owns one site i and walks a slice of the j atoms staged through
// LDS in tiles of the block size; the j range is split over grid.y so that
// even a few thousand sites fill the 256 CUs, and the slices' partial
// tensors are summed in a fixed order by the combine.
#include "kernels.h"

namespace mdff {

__global__ __launch_bounds__(MDFF_BLOCK) void k_efg_mask(int n, const int* __restrict__ typ,
                                                         int it_efg, const double* __restrict__ q,
                                                         const double* __restrict__ mu,
                                                         double* __restrict__ qe,
                                                         double* __restrict__ mue) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool on = typ[i] == it_efg;
  qe[i] = on ? q[i] : 0.0;
  for (int a = 0; a < 3; ++a) mue[i + a * n] = on ? mu[i + a * n] : 0.0;
}

// Rank-2 and rank-3 screened tensors of one pair (src/efg.f90:1434-1463) and
// their contraction with the source's charge and dipole into V.
template <bool DIP>
__device__ __forceinline__ void efg_pair_es(double x, double y, double z, double d2, double alpha,
                                            double qj, double mx, double my, double mz,
                                            double (&V)[6]) {
  const double rpi = 0.56418958354775628694807945156077;  // 1/sqrt(pi)
  const double a2 = alpha * alpha, a3 = a2 * alpha, a5 = a3 * a2;
  const double d = sqrt(d2);
  const double d3 = d2 * d, d5 = d3 * d2;
  const double dm5 = 1.0 / (d2 * d3);
  const double expon = exp(-a2 * d2) * rpi;
  const double F1 = errfc_nr(alpha * d) + 2.0 * alpha * d * expon;
  const double F2 = F1 + 4.0 * a3 * d3 * expon / 3.0;
  const double t2 = 3.0 * F2 * dm5, t1 = d2 * F1 * dm5;
  V[0] -= qj * (x * x * t2 - t1);
  V[1] -= qj * (y * y * t2 - t1);
  V[2] -= qj * (z * z * t2 - t1);
  V[3] -= qj * (x * y * t2);
  V[4] -= qj * (x * z * t2);
  V[5] -= qj * (y * z * t2);
  if (DIP) {
    const double F3 = F2 + 8.0 * a5 * d5 * expon / 15.0;
    const double dm7 = dm5 / d2 * 3.0;
    const double c3 = -5.0 * F3 * dm7, c1 = d2 * F2 * dm7;
    const double Txxx = x * (x * x * c3 + 3.0 * c1), Tyyy = y * (y * y * c3 + 3.0 * c1),
                 Tzzz = z * (z * z * c3 + 3.0 * c1);
    const double Txxy = y * (x * x * c3 + c1), Txxz = z * (x * x * c3 + c1);
    const double Tyyx = x * (y * y * c3 + c1), Tyyz = z * (y * y * c3 + c1);
    const double Tzzx = x * (z * z * c3 + c1), Tzzy = y * (z * z * c3 + c1);
    const double Txyz = x * y * z * c3;
    V[0] += Txxx * mx + Txxy * my + Txxz * mz;
    V[1] += Tyyx * mx + Tyyy * my + Tyyz * mz;
    V[2] += Tzzx * mx + Tzzy * my + Tzzz * mz;
    V[3] += Txxy * mx + Tyyx * my + Txyz * mz;
    V[4] += Txxz * mx + Txyz * my + Tzzx * mz;
    V[5] += Txyz * mx + Tyyz * my + Tzzy * mz;
  }
}

// grid (ceil(n / MDFF_BLOCK), nsplit); slice y covers j in [y*jlen, (y+1)*jlen).
// part rows: [(y*6 + c)*n + i].
template <bool DIP>
__global__ __launch_bounds__(MDFF_BLOCK) void k_efg_es_dir(EfgDirArgs a) {
  __shared__ double ss[MDFF_BLOCK][8];
  const int n = a.n;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = i < n;
  double si0 = 0, si1 = 0, si2 = 0;
  if (act) {
    si0 = a.s[i];
    si1 = a.s[i + n];
    si2 = a.s[i + 2 * n];
  }
  double V[6] = {0, 0, 0, 0, 0, 0};
  const int j0 = blockIdx.y * a.jlen;
  const int j1 = min(n, j0 + a.jlen);
  for (int jb = j0; jb < j1; jb += MDFF_BLOCK) {
    const int m = min(MDFF_BLOCK, j1 - jb);
    __syncthreads();
    if (threadIdx.x < m) {
      const int j = jb + threadIdx.x;
      ss[threadIdx.x][0] = a.s[j];
      ss[threadIdx.x][1] = a.s[j + n];
      ss[threadIdx.x][2] = a.s[j + 2 * n];
      ss[threadIdx.x][3] = a.q[j];
      if (DIP) {
        ss[threadIdx.x][4] = a.mu[j];
        ss[threadIdx.x][5] = a.mu[j + n];
        ss[threadIdx.x][6] = a.mu[j + 2 * n];
      }
    }
    __syncthreads();
    if (!act) continue;
    for (int t = 0; t < m; ++t) {
      if (jb + t == i) continue;
      const double qj = ss[t][3];
      double mx = 0, my = 0, mz = 0;
      if (DIP) {
        mx = ss[t][4];
        my = ss[t][5];
        mz = ss[t][6];
        if (qj == 0.0 && mx == 0.0 && my == 0.0 && mz == 0.0) continue;
      } else if (qj == 0.0) {
        continue;
      }
      double x, y, z;
      const double d2 = min_image(a.cell, si0 - ss[t][0], si1 - ss[t][1], si2 - ss[t][2], x, y, z);
      efg_pair_es<DIP>(x, y, z, d2, a.alpha, qj, mx, my, mz, V);
    }
  }
  if (act) {
    double* out = a.part + (size_t)blockIdx.y * 6 * n;
    for (int c = 0; c < 6; ++c) out[(size_t)c * n + i] = V[c];
  }
}

template __global__ void k_efg_es_dir<false>(EfgDirArgs);
template __global__ void k_efg_es_dir<true>(EfgDirArgs);

// One block per site i: the block's threads stride over (cell, j) pairs;
// the block sum goes to part rows [c*n + i].  Positions are wrapped into
// [-1/2, 1/2] as efgcalc leaves them (src/efg.f90:424-426).
template <bool DIP>
__global__ __launch_bounds__(MDFF_BLOCK) void k_efg_ds(EfgDsArgs a) {
  __shared__ double red[4 * 6];
  const int n = a.n;
  const int i = blockIdx.x;
  const double* A = a.cell.A;
  double si[3];
  for (int k = 0; k < 3; ++k) {
    const double s = a.s[i + k * n];
    si[k] = s - nint_d(s);
  }
  double V[6] = {0, 0, 0, 0, 0, 0};
  const long long total = (long long)a.ncell * n;
  for (long long p = threadIdx.x; p < total; p += blockDim.x) {
    const int cidx = (int)(p / n);
    const int j = (int)(p - (long long)cidx * n);
    const double bx = a.cells[3 * cidx], by = a.cells[3 * cidx + 1], bz = a.cells[3 * cidx + 2];
    if (bx == 0.0 && by == 0.0 && bz == 0.0 && j == i) continue;
    const double qj = a.q[j];
    double mx = 0, my = 0, mz = 0;
    if (DIP) {
      mx = a.mu[j];
      my = a.mu[j + n];
      mz = a.mu[j + 2 * n];
    }
    double sj[3];
    for (int k = 0; k < 3; ++k) {
      const double s = a.s[j + k * n];
      sj[k] = s - nint_d(s);
    }
    const double sx = si[0] - (sj[0] + bx), sy = si[1] - (sj[1] + by), sz = si[2] - (sj[2] + bz);
    const double x = sx * A[0] + sy * A[3] + sz * A[6];
    const double y = sx * A[1] + sy * A[4] + sz * A[7];
    const double z = sx * A[2] + sy * A[5] + sz * A[8];
    const double d2 = x * x + y * y + z * z;
    if (!(d2 < a.cut2 && d2 != 0.0)) continue;
    const double d = sqrt(d2);
    const double dm5 = 1.0 / (d2 * d2 * d);
    const double t2 = 3.0 * dm5;
    V[0] -= qj * (x * x * t2 - d2 * dm5);
    V[1] -= qj * (y * y * t2 - d2 * dm5);
    V[2] -= qj * (z * z * t2 - d2 * dm5);
    V[3] -= qj * (x * y * t2);
    V[4] -= qj * (x * z * t2);
    V[5] -= qj * (y * z * t2);
    if (DIP) {
      const double dm7 = dm5 / d2 * 3.0;
      const double c3 = 5.0 * dm7, c1 = d2 * dm7;
      const double Txxx = x * (x * x * c3 - 3.0 * c1), Tyyy = y * (y * y * c3 - 3.0 * c1),
                   Tzzz = z * (z * z * c3 - 3.0 * c1);
      const double Txxy = y * (x * x * c3 - c1), Txxz = z * (x * x * c3 - c1);
      const double Tyyx = x * (y * y * c3 - c1), Tyyz = z * (y * y * c3 - c1);
      const double Tzzx = x * (z * z * c3 - c1), Tzzy = y * (z * z * c3 - c1);
      const double Txyz = x * y * z * c3;
      V[0] -= Txxx * mx + Txxy * my + Txxz * mz;
      V[1] -= Tyyx * mx + Tyyy * my + Tyyz * mz;
      V[2] -= Tzzx * mx + Tzzy * my + Tzzz * mz;
      V[3] -= Txxy * mx + Tyyx * my + Txyz * mz;
      V[4] -= Txxz * mx + Txyz * my + Tzzx * mz;
      V[5] -= Txyz * mx + Tyyz * my + Tzzy * mz;
    }
  }
  block_sum<6>(V, red);
  if (threadIdx.x == 0)
    for (int c = 0; c < 6; ++c) a.part[(size_t)c * n + i] = V[c];
}

template __global__ void k_efg_ds<false>(EfgDsArgs);
template __global__ void k_efg_ds<true>(EfgDsArgs);

__global__ __launch_bounds__(MDFF_BLOCK) void k_efg_combine(EfgCombineArgs a) {
  const int n = a.n;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double V[6] = {0, 0, 0, 0, 0, 0};
  for (int sp = 0; sp < a.ndir; ++sp) {
    const double* p = a.dir + (size_t)sp * 6 * n;
    for (int c = 0; c < 6; ++c) V[c] += p[(size_t)c * n + i];
  }
  if (a.nrec > 0) {
    // reciprocal sums over (n1, n2, n3): 2 (4 pi / V) (2 pi)^2 B W B^T
    double W[6] = {0, 0, 0, 0, 0, 0};
    for (int sp = 0; sp < a.nrec; ++sp) {
      const double* p = a.rec + (size_t)sp * 9 * n;
      for (int c = 0; c < 6; ++c) W[c] += p[(size_t)(3 + c) * n + i];
    }
    const double Wm[3][3] = {{W[0], W[3], W[4]}, {W[3], W[1], W[5]}, {W[4], W[5], W[2]}};
    const double* B = a.cell.B;
    const double tpi = 2.0 * MDFF_PI;
    const double fg = 2.0 * a.fpi_V * tpi * tpi;
    const int rr[6] = {0, 1, 2, 0, 0, 1}, cc[6] = {0, 1, 2, 1, 2, 2};
    for (int k = 0; k < 6; ++k) {
      double acc = 0.0;
      for (int x = 0; x < 3; ++x)
        for (int y = 0; y < 3; ++y) acc += B[rr[k] + 3 * x] * Wm[x][y] * B[cc[k] + 3 * y];
      V[k] += fg * acc;
    }
  }
  // self term (zero for the direct sum) and charged-cell correction
  const double diag = a.selfa * a.q[i] + a.corr;
  V[0] += diag;
  V[1] += diag;
  V[2] += diag;
  const double s = a.scale;
  const int rr[6] = {0, 1, 2, 0, 0, 1}, cc[6] = {0, 1, 2, 1, 2, 2};
  for (int k = 0; k < 6; ++k) {
    const double v = s * V[k];
    a.out[i + (size_t)n * (rr[k] + 3 * cc[k])] = v;
    if (rr[k] != cc[k]) a.out[i + (size_t)n * (cc[k] + 3 * rr[k])] = v;
  }
}

// ---------------------------------------------------------------------------
// Eigenvalues of a symmetric 3x3 matrix by cyclic Jacobi rotations (the
// classical method, e.g. Golub & Van Loan, Matrix Computations, 8.5), in
// ascending order as LAPACK's DSYEV returns them.
__device__ void sym3_eigvals(double a00, double a11, double a22, double a01, double a02,
                             double a12, double (&w)[3]) {
  double m[3][3] = {{a00, a01, a02}, {a01, a11, a12}, {a02, a12, a22}};
  for (int sweep = 0; sweep < 32; ++sweep) {
    const double off = m[0][1] * m[0][1] + m[0][2] * m[0][2] + m[1][2] * m[1][2];
    const double dia = m[0][0] * m[0][0] + m[1][1] * m[1][1] + m[2][2] * m[2][2];
    if (off == 0.0 || off <= 1e-36 * dia) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        const double apq = m[p][q];
        if (apq == 0.0) continue;
        const double theta = (m[q][q] - m[p][p]) / (2.0 * apq);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; ++k) {
          const double mkp = m[k][p], mkq = m[k][q];
          m[k][p] = c * mkp - s * mkq;
          m[k][q] = s * mkp + c * mkq;
        }
        for (int k = 0; k < 3; ++k) {
          const double mpk = m[p][k], mqk = m[q][k];
          m[p][k] = c * mpk - s * mqk;
          m[q][k] = s * mpk + c * mqk;
        }
        m[p][q] = m[q][p] = 0.0;
      }
  }
  double e0 = m[0][0], e1 = m[1][1], e2 = m[2][2], t;
  if (e0 > e1) { t = e0; e0 = e1; e1 = t; }
  if (e1 > e2) { t = e1; e1 = e2; e2 = t; }
  if (e0 > e1) { t = e0; e0 = e1; e1 = t; }
  w[0] = e0;
  w[1] = e1;
  w[2] = e2;
}

// efg6 (n,6): vxx vyy vzz vxy vxz vyz; nmr (n,4); U (n,5); S (n); bad[0]
// counts the sites whose ordering fails the reference's checks.
__global__ __launch_bounds__(MDFF_BLOCK) void k_efg_nmr(int n, const double* __restrict__ efg6,
                                                        double* __restrict__ nmr,
                                                        double* __restrict__ U,
                                                        double* __restrict__ S,
                                                        int* __restrict__ bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double vxx = efg6[i], vyy = efg6[i + n], vzz = efg6[i + 2 * n];
  const double vxy = efg6[i + 3 * n], vxz = efg6[i + 4 * n], vyz = efg6[i + 5 * n];
  // Czjzek components (src/efg.f90:2235-2244)
  const double sq3 = 1.0 / sqrt(3.0), sq32 = sq3 * 0.5;
  const double u[5] = {vzz * 0.5, vxz * sq3, vyz * sq3, vxy * sq3, (vxx - vyy) * sq32};
  double ss = 0.0;
  for (int k = 0; k < 5; ++k) {
    U[i + (size_t)k * n] = u[k];
    ss = ss + u[k] * u[k];
  }
  S[i] = sqrt(ss);
  double w[3];
  sym3_eigvals(vxx, vyy, vzz, vxy, vxz, vyz, w);
  // nmr_convention: |Vzz - iso| >= |Vyy - iso| >= |Vxx - iso|
  double diso = (w[0] + w[1] + w[2]) / 3.0;
  if (fabs(diso) != 0.0 && fabs(diso) > 1e-3) diso = 0.0;
  const double a0 = fabs(w[0] - diso), a1 = fabs(w[1] - diso), a2 = fabs(w[2] - diso);
  double d11, d22, d33;
  if (a0 >= a1 && a0 >= a2) {
    d11 = w[0];
    if (a1 >= a2) { d22 = w[1]; d33 = w[2]; } else { d22 = w[2]; d33 = w[1]; }
  } else if (a1 >= a0 && a1 >= a2) {
    d11 = w[1];
    if (a0 >= a2) { d22 = w[0]; d33 = w[2]; } else { d22 = w[2]; d33 = w[0]; }
  } else {
    d11 = w[2];
    if (a0 >= a1) { d22 = w[0]; d33 = w[1]; } else { d22 = w[1]; d33 = w[0]; }
  }
  const double xx = d33, yy = d22, zz = d11;
  double eta = (zz != 0.0) ? (xx - yy) / zz : 0.0;
  // the reference's small-value test compares with a default-REAL 1.0E-05
  const double tiny = (double)1.0e-5f;
  if (xx == yy || zz == 0.0 || (fabs(xx) < tiny && fabs(yy) < tiny && fabs(zz) < tiny)) eta = 0.0;
  nmr[i] = xx;
  nmr[i + n] = yy;
  nmr[i + 2 * n] = zz;
  nmr[i + 3 * n] = eta;
  const bool ok = !(fabs(zz - diso) < fabs(xx - diso) || fabs(zz - diso) < fabs(yy - diso) ||
                    fabs(yy - diso) < fabs(xx - diso));
  if (!ok) atomicAdd(bad, 1);
}

}  // namespace mdff
