// common.h -- shared device-side definitions of the MI355X force/field
// library (gfx950, wave64).
//
// Data layout in HBM (one context = one configuration):
//   positions   Cartesian SoA  x[N] y[N] z[N]                 (double)
//   fractional  SoA            s1[N] s2[N] s3[N]              (double)
//   moments     SoA            q[N], mu = mx[N] my[N] mz[N]
//   EFG         Fortran (N,3,3): efg[i + N*(a + 3*b)]
//   neighbours  ELL, column-major: nbr[k*N + i] (k < cnt[i]), full list
//               (each pair appears in both atoms' rows), so every per-atom
//               output is written by exactly one thread: no atomics.
//   k vectors   chunks of KC consecutive n3 at fixed (n1,n2)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define MDFF_NT 16
#define MDFF_PI 3.14159265358979323846264338327950
#define MDFF_BLOCK 256

namespace mdff {

// Cell: A (direct, columns = lattice vectors), B (reciprocal / Omega,
// columns), both column-major like simu_cell%A and %B.
struct Cell {
  double A[9];
  double B[9];
  double omega;
};

// Per ordered type pair (ti, tj) tables, index ti + 16*tj.
struct PairTables {
  // n-m Lennard-Jones, derived as initialize_param_non_bonded
  // (src/field.f90:1021-1064)
  double lj_epsp[256], lj_sigsq[256], lj_fc[256], lj_ptwo[256],
      lj_qtwo[256], lj_p[256], lj_q[256], lj_uc[256], lj_uc1[256],
      lj_uc2[256];
  // BMHFTD (src/field.f90:2952-2976)
  double bm_A[256], bm_B[256], bm_C[256], bm_D[256], bm_BD[256];
  // Morse
  double mo_eps[256], mo_rho[256], mo_rs[256], mo_sig[256];
  // Tang-Toennies dipole damping seen from the centre atom ti of a pair
  // with tj.  Slot o = 0 when ti's atom has the smaller index (it is 'ia'
  // of src/field.f90:2207-2211), o = 1 otherwise.
  //   dmp_on[o]    ldamp of the pair
  //   dmp_*[o]     parameters for ti's own field terms
  int32_t dmp_on[2][256];
  double dmp_b[2][256], dmp_c[2][256];
  int32_t dmp_k[2][256];
};

struct TypeTables {
  double qch[MDFF_NT];
  double dip[MDFF_NT][3];
  double pol[MDFF_NT][9];     // poldip(it,a,b) at [a + 3*b]
  double invpol[MDFF_NT][9];  // 3x3 inverse (src/read_pos.f90:201-210)
  int32_t polar[MDFF_NT];     // ldip_polar
  // quadrupole polarizability polquad(it,g,a,b) at [g + 3*(a + 3*b)], and
  // lquad_polar (src/field.f90:138, 154)
  double polq[MDFF_NT][27];
  int32_t qpolar[MDFF_NT];
};

// E_TT(k) = 1/k! (src/tt_damp.f90:13-29)
struct TTTable {
  double e[9];
};

__device__ __forceinline__ double nint_d(double x) {
  // Fortran NINT: round half away from zero
  return round(x);
}

// Minimum image of s_i - s_j in fractional coordinates mapped to Cartesian
// (src/field.f90:1262-1273); returns |r|^2.
__device__ __forceinline__ double min_image(const Cell& c, double dsx,
                                            double dsy, double dsz,
                                            double& rx, double& ry,
                                            double& rz) {
  dsx = dsx - nint_d(dsx);
  dsy = dsy - nint_d(dsy);
  dsz = dsz - nint_d(dsz);
  rx = dsx * c.A[0] + dsy * c.A[3] + dsz * c.A[6];
  ry = dsx * c.A[1] + dsy * c.A[4] + dsz * c.A[7];
  rz = dsx * c.A[2] + dsy * c.A[5] + dsz * c.A[8];
  return rx * rx + ry * ry + rz * rz;
}

// errfc of src/functions.f90:28-40 (Numerical Recipes erfcc).  The FP64
// path evaluates this formula, not erfc(), so that it matches the
// reference to rounding.
__device__ __forceinline__ double errfc_nr(double x) {
  double z = fabs(x);
  double t = 1.0 / (1.0 + 0.5 * z);
  double p =
      t * (1.00002368 +
           t * (.37409196 +
                t * (.09678418 +
                     t * (-.18628806 +
                          t * (.27886807 +
                               t * (-1.13520398 +
                                    t * (1.48851587 +
                                         t * (-.82215223 + t * .17087277))))))));
  double r = t * exp(-z * z - 1.26551223 + p);
  return (x < 0.0) ? 2.0 - r : r;
}

__device__ __forceinline__ float errfc_nr_f(float x) {
  float z = fabsf(x);
  float t = 1.0f / (1.0f + 0.5f * z);
  float p =
      t * (1.00002368f +
           t * (.37409196f +
                t * (.09678418f +
                     t * (-.18628806f +
                          t * (.27886807f +
                               t * (-1.13520398f +
                                    t * (1.48851587f +
                                         t * (-.82215223f + t * .17087277f))))))));
  float r = t * __expf(-z * z - 1.26551223f + p);
  return (x < 0.0f) ? 2.0f - r : r;
}

// Tang-Toennies f and f' (src/field.f90:3775-3805)
__device__ __forceinline__ void tt_damp(const TTTable& E, double b, double c,
                                        double r, int order, double& f,
                                        double& fd) {
  double br = b * r;
  double ex = exp(-br) * c;
  double ff = E.e[order];
  for (int k = order - 1; k >= 1; --k) ff = ff * br + E.e[k];
  ff = ff * br + E.e[0];
  f = 1.0 - ff * ex;
  double brn = 1.0;
  for (int k = 0; k < order; ++k) brn *= br;
  fd = E.e[order] * brn * ex * b;
}

__device__ __forceinline__ void tt_damp_f(const TTTable& E, float b, float c,
                                          float r, int order, float& f,
                                          float& fd) {
  float br = b * r;
  float ex = __expf(-br) * c;
  float ff = (float)E.e[order];
  for (int k = order - 1; k >= 1; --k) ff = ff * br + (float)E.e[k];
  ff = ff * br + 1.0f;
  f = 1.0f - ff * ex;
  float brn = 1.0f;
  for (int k = 0; k < order; ++k) brn *= br;
  fd = (float)E.e[order] * brn * ex * b;
}

// Deterministic block reduction of NV doubles per thread into one value per
// slot; thread 0 holds the results in v.  blockDim.x == MDFF_BLOCK.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds) {
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double x = v[k];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) lds[wid * NV + k] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int w = 0; w < MDFF_BLOCK / 64; ++w) s += lds[w * NV + k];
      v[k] = s;
    }
  }
  __syncthreads();
}

}  // namespace mdff
