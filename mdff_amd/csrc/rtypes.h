// rtypes.h -- arithmetic of the real-space pair terms in type R (double for
// the FP64 path, float for the mixed path).
#pragma once

#include "common.h"

namespace mdff {

// four-vector of type R (rx, ry, rz, radial factor of the SCF pair table)
template <typename R>
struct R4Sel;
template <>
struct R4Sel<double> {
  typedef double4 type;
};
template <>
struct R4Sel<float> {
  typedef float4 type;
};
template <typename R>
using R4t = typename R4Sel<R>::type;

template <typename R>
__device__ __forceinline__ R rexp(R x);
template <>
__device__ __forceinline__ double rexp<double>(double x) {
  return exp(x);
}
template <>
__device__ __forceinline__ float rexp<float>(float x) {
  return __expf(x);
}

template <typename R>
__device__ __forceinline__ R rpow(R x, R y);
template <>
__device__ __forceinline__ double rpow<double>(double x, double y) {
  return pow(x, y);
}
template <>
__device__ __forceinline__ float rpow<float>(float x, float y) {
  return powf(x, y);
}

// errfc of src/functions.f90:28-40 in type R
template <typename R>
__device__ __forceinline__ R errfc_t(R x) {
  const R z = fabs(x);
  const R t = R(1) / (R(1) + R(0.5) * z);
  const R p =
      t * (R(1.00002368) +
           t * (R(.37409196) +
                t * (R(.09678418) +
                     t * (R(-.18628806) +
                          t * (R(.27886807) +
                               t * (R(-1.13520398) +
                                    t * (R(1.48851587) + t * (R(-.82215223) + t * R(.17087277)))))))));
  const R r = t * rexp<R>(-z * z - R(1.26551223) + p);
  return (x < R(0)) ? R(2) - r : r;
}

// Tang-Toennies f and f' (src/field.f90:3775-3805) in type R
template <typename R>
__device__ __forceinline__ void tt_damp_t(const TTTable& E, R b, R c, R r, int order, R& f,
                                          R& fd) {
  const R br = b * r;
  const R ex = rexp<R>(-br) * c;
  R ff = R(E.e[order]);
  for (int k = order - 1; k >= 1; --k) ff = ff * br + R(E.e[k]);
  ff = ff * br + R(E.e[0]);
  f = R(1) - ff * ex;
  R brn = R(1);
  for (int k = 0; k < order; ++k) brn *= br;
  fd = R(E.e[order]) * brn * ex * b;
}

// Minimum image in type R: the fractional difference is formed in double
// (positions are stored in double), the rest in R.
template <typename R>
__device__ __forceinline__ R min_image_t(const Cell& c, double dsx, double dsy, double dsz,
                                         R& rx, R& ry, R& rz) {
  const R sx = R(dsx - nint_d(dsx));
  const R sy = R(dsy - nint_d(dsy));
  const R sz = R(dsz - nint_d(dsz));
  rx = sx * R(c.A[0]) + sy * R(c.A[3]) + sz * R(c.A[6]);
  ry = sx * R(c.A[1]) + sy * R(c.A[4]) + sz * R(c.A[7]);
  rz = sx * R(c.A[2]) + sy * R(c.A[5]) + sz * R(c.A[8]);
  return rx * rx + ry * ry + rz * rz;
}

}  // namespace mdff
