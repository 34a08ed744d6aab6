// md.hip -- integrator kernels of the device-resident MD loop: the particle
// updates of prop_velocity_verlet (src/integration.f90:131-194) and
// prop_pos_vel_verlet (:333-385), the velocity scaling and kinetic energy
// of the Nose-Hoover chains, and the chains themselves (chain_nhc2
// :261-324, chain_nhcn :564-713), which are scalar recurrences run by one
// thread so that the thermostat state never leaves the device.
//
// Every particle update keeps the reference's operation order, e.g.
// r + v dt + (f dt^2/2) / m, so that the FP64 trajectory follows it to
// rounding.  Kinetic energies are block partials of sum m |v|^2 (the
// caller reduces them in a fixed order and halves the sum, as calc_temp,
// src/kinetic.f90:531-556).
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace mdff {

__device__ __forceinline__ void kin_partial(double v, double* part) {
  __shared__ double red[MDFF_BLOCK / 64];
  double acc[1] = {v};
  block_sum<1>(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// prop_velocity_verlet, first half: fs = f; r += v dt + (f dtsq2) / m
__global__ __launch_bounds__(MDFF_BLOCK) void k_vv_pos(MdArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const double m = a.mass[i];
  for (int c = 0; c < 3; ++c) {
    const size_t k = i + (size_t)c * a.n;
    const double f = a.f[k];
    a.fs[k] = f;
    a.pos[k] = a.pos[k] + a.vel[k] * a.dt + (f * a.dtsq2) / m;
  }
}

// prop_velocity_verlet, second half: v += (fs + f) dt2 / m, with the
// kinetic-energy partials of the new velocities
__global__ __launch_bounds__(MDFF_BLOCK) void k_vv_vel(MdArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double mv2 = 0.0;
  if (i < a.n) {
    const double m = a.mass[i];
    double v[3];
    for (int c = 0; c < 3; ++c) {
      const size_t k = i + (size_t)c * a.n;
      v[c] = a.vel[k] + (a.fs[k] + a.f[k]) * a.dt2 / m;
      a.vel[k] = v[c];
    }
    mv2 = (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * m;
  }
  kin_partial(mv2, a.part);
}

// prop_pos_vel_verlet, first half: v += f dt2 / m; r += v dt
__global__ __launch_bounds__(MDFF_BLOCK) void k_kick_drift(MdArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const double m = a.mass[i];
  for (int c = 0; c < 3; ++c) {
    const size_t k = i + (size_t)c * a.n;
    const double v = a.vel[k] + a.f[k] * a.dt2 / m;
    a.vel[k] = v;
    a.pos[k] = a.pos[k] + v * a.dt;
  }
}

// prop_pos_vel_verlet, second half: v += f dt2 / m, kinetic partials
__global__ __launch_bounds__(MDFF_BLOCK) void k_kick(MdArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double mv2 = 0.0;
  if (i < a.n) {
    const double m = a.mass[i];
    double v[3];
    for (int c = 0; c < 3; ++c) {
      const size_t k = i + (size_t)c * a.n;
      v[c] = a.vel[k] + a.f[k] * a.dt2 / m;
      a.vel[k] = v[c];
    }
    mv2 = (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * m;
  }
  kin_partial(mv2, a.part);
}

// v *= s (the chain's scale factor, on the device), kinetic partials of
// the scaled velocities; with s == nullptr only the partials
__global__ __launch_bounds__(MDFF_BLOCK) void k_vscale(MdArgs a, const NhcState* s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double mv2 = 0.0;
  if (i < a.n) {
    const double m = a.mass[i];
    const double sc = s ? s->s : 1.0;
    double v[3];
    for (int c = 0; c < 3; ++c) {
      const size_t k = i + (size_t)c * a.n;
      v[c] = s ? sc * a.vel[k] : a.vel[k];
      if (s) a.vel[k] = v[c];
    }
    mv2 = (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * m;
  }
  kin_partial(mv2, a.part);
}

// Per-atom masses from the species.
__global__ __launch_bounds__(MDFF_BLOCK) void k_massia(int n, const int* typ, NhcMass m,
                                                       double* mass) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) mass[i] = m.m[typ[i]];
}

// chain_nhcn (src/integration.f90:564-713), trotter_order -1, on one
// thread.  kin = sum_mv2 / 2 enters; the scale factor goes to st->s.  G
// holds zeros at entry as in the reference (allocated and zeroed per
// call), so that on the first inner step the descending sweep uses
// G(2..nhc_n-1) = 0 (only G(nhc_n) and G(1) are set before it).
__global__ void k_nhcn_chain(NhcState* st, const double* sum_mv2, int nhc_n, int yosh, int mults,
                             double dt, double temp, double L, int final) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double* vxi = st->vxi;
  double* xi = st->xi;
  const double* Q = st->Q;
  double G[MDFF_NHC_MAX];
  for (int k = 0; k < nhc_n; ++k) G[k] = 0.0;
  const double kin = 0.5 * (*sum_mv2);
  double s = 1.0;
  G[0] = 2.0 * kin - L * temp;
  const int last = nhc_n - 1;
  for (int k = 0; k < mults; ++k) {
    for (int j = 0; j < yosh; ++j) {
      const double dts = st->yw[j] * dt / (double)mults;
      const double dts2 = dts * 0.5;
      const double dts4 = dts2 * 0.5;
      const double dts8 = dts4 * 0.5;
      G[last] = vxi[last - 1] * vxi[last - 1] / Q[last - 1] - temp;
      vxi[last] = vxi[last] + G[last] * dts4;
      for (int h = last - 1; h >= 0; --h) {
        vxi[h] = vxi[h] * exp(-vxi[h + 1] * dts8 / Q[h + 1]);
        vxi[h] = vxi[h] + G[h] * dts4;
        vxi[h] = vxi[h] * exp(-vxi[h + 1] * dts8 / Q[h + 1]);
      }
      s = s * exp(-vxi[0] * dts2 / Q[0]);
      for (int h = 0; h < nhc_n; ++h) xi[h] = xi[h] + vxi[h] * dts2 / Q[h];
      G[0] = 2.0 * kin * s * s - L * temp;
      for (int h = 0; h < last; ++h) {
        vxi[h] = vxi[h] * exp(-vxi[h + 1] * dts8 / Q[h + 1]);
        vxi[h] = vxi[h] + G[h] * dts4;
        vxi[h] = vxi[h] * exp(-vxi[h + 1] * dts8 / Q[h + 1]);
        G[h + 1] = vxi[h] * vxi[h] / Q[h] - temp;
      }
      vxi[last] = vxi[last] + G[last] * dts4;
    }
  }
  st->s = s;
  if (final) {
    // conserved-energy terms of nhcn (src/integration.f90:524-541)
    const double nvt1 = L * temp * xi[0];
    const double nvt2 = vxi[0] * vxi[0] * 0.5 / Q[0];
    double nvt3 = 0.0, nvt4 = 0.0;
    for (int h = 1; h < nhc_n; ++h) {
      nvt3 = nvt3 + vxi[h] * vxi[h] * 0.5 / Q[h];
      nvt4 = nvt4 + temp * xi[h];
    }
    st->e_nvt = nvt1 + nvt2 + nvt3 + nvt4;
  }
}

// chain_nhc2 (src/integration.f90:261-324) on one thread; kin = sum_mv2 / 2.
__global__ void k_nhc2_chain(NhcState* st, const double* sum_mv2, double dt, double temp,
                             double L, int final) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double* vxi = st->vxi;
  double* xi = st->xi;
  const double* Q = st->Q;
  double kin = 0.5 * (*sum_mv2);
  double s = 1.0;
  const double dt2 = dt * 0.5, dt4 = dt2 * 0.5, dt8 = dt4 * 0.5;
  double G1 = (2.0 * kin - L * temp) / Q[0];
  double G2 = (Q[0] * vxi[0] * vxi[0] - temp) / Q[1];
  vxi[1] = vxi[1] + G2 * dt4;
  vxi[0] = vxi[0] * exp(-vxi[1] * dt8);
  vxi[0] = vxi[0] + G1 * dt4;
  vxi[0] = vxi[0] * exp(-vxi[1] * dt8);
  s = s * exp(-vxi[0] * dt2);
  kin = s * s * kin;
  xi[0] = xi[0] + vxi[0] * dt2;
  xi[1] = xi[1] + vxi[1] * dt2;
  G1 = (2.0 * kin - L * temp) / Q[0];
  vxi[0] = vxi[0] * exp(-vxi[1] * dt8);
  vxi[0] = vxi[0] + G1 * dt4;
  vxi[0] = vxi[0] * exp(-vxi[1] * dt8);
  G2 = (Q[0] * vxi[0] * vxi[0] - temp) / Q[1];
  vxi[1] = vxi[1] + G2 * dt4;
  st->s = s;
  if (final) {
    // nhc2 (src/integration.f90:237-243)
    double e = 0.0;
    e = e + L * temp * xi[0];
    e = e + vxi[0] * vxi[0] * 0.5 / Q[0];
    e = e + vxi[1] * vxi[1] * 0.5 / Q[1];
    e = e + temp * xi[1];
    st->e_nvt = e;
  }
}

}  // namespace mdff
