// kernels64.hip -- FP64 kernels of the force/field evaluation (gfx950).
//
// Every kernel maps one output row to one thread, so nothing is written by
// two threads and no atomics are needed; block partial sums go to a partial
// buffer that k_reduce_rows sums in a fixed order, which makes every result
// bitwise reproducible from run to run.
#include "kernels.h"

namespace mdff {

// ---------------------------------------------------------------------------
// Cartesian -> fractional (kardir with BASIS = B, src/lattice.f90:183-213)
// and the per-atom phase bases e^{2 pi i s_a} used by the k-space kernels.
__global__ __launch_bounds__(MDFF_BLOCK) void k_prepare(
    int n, const double* __restrict__ pos, Cell cell, const int* __restrict__ ext,
    const int* __restrict__ typ, double* __restrict__ s, double4* __restrict__ s4,
    double2* __restrict__ eph) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x = pos[i], y = pos[i + n], z = pos[i + 2 * n];
  const double* B = cell.B;
  double s1 = x * B[0] + y * B[1] + z * B[2];
  double s2 = x * B[3] + y * B[4] + z * B[5];
  double s3 = x * B[6] + y * B[7] + z * B[8];
  s[i] = s1;
  s[i + n] = s2;
  s[i + 2 * n] = s3;
  // w: 16 x the caller's index of the atom (the pair-order conventions
  // compare it) + its species (s4_type)
  const long long w = 16ll * (ext ? ext[i] : i) + typ[i];
  s4[i] = make_double4(s1, s2, s3, (double)w);
  double sn, cs;
  sincospi(2.0 * s1, &sn, &cs);
  eph[i] = make_double2(cs, sn);
  sincospi(2.0 * s2, &sn, &cs);
  eph[i + n] = make_double2(cs, sn);
  sincospi(2.0 * s3, &sn, &cs);
  eph[i + 2 * n] = make_double2(cs, sn);
}

// ---------------------------------------------------------------------------
// Full neighbour list (replaces vnlist_pbc, src/tools.f90:274-370): every j
// whose minimum image lies within cut (list cutoff incl. skin), in index
// order; row-major rows nbr[i*maxnb + k] so that the lanes sharing an atom
// read contiguously.  Brute force over LDS tiles of fractional positions,
// one thread per atom i, with the candidates j split into gridDim.y ranges
// of jlen atoms so that the grid fills the device: a COUNT pass stores each
// range's count per atom (scnt[range][i]), k_nlist_offsets turns them into
// offsets and the row count, and a FILL pass writes each range's neighbours
// at its offset, so that every row stays in index order.
template <bool FILL>
__global__ __launch_bounds__(MDFF_BLOCK) void k_nlist(int n, const double* __restrict__ s,
                                                      Cell cell, double cut2, int maxnb, int jlen,
                                                      int* __restrict__ scnt,
                                                      int* __restrict__ nbr) {
  __shared__ double ts[3][MDFF_BLOCK];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int sp = blockIdx.y;
  const int j0 = sp * jlen, j1 = min(n, j0 + jlen);
  const bool act = i < n;
  double si1 = 0, si2 = 0, si3 = 0;
  if (act) {
    si1 = s[i];
    si2 = s[i + n];
    si3 = s[i + 2 * n];
  }
  const int off = (FILL && act) ? scnt[(size_t)sp * n + i] : 0;
  // orthorhombic cell: r_a = A_aa (ds_a - round(ds_a)); rounding half to
  // even instead of away from zero gives the same |r| (a difference of
  // exactly half the cell maps to -0.5 or 0.5)
  const bool ortho = cell.A[1] == 0.0 && cell.A[2] == 0.0 && cell.A[3] == 0.0 &&
                     cell.A[5] == 0.0 && cell.A[6] == 0.0 && cell.A[7] == 0.0;
  const double a0 = cell.A[0], a4 = cell.A[4], a8 = cell.A[8];
  int c = 0;
  for (int base = j0; base < j1; base += MDFF_BLOCK) {
    const int r = base + threadIdx.x;
    __syncthreads();
    if (r < j1) {
      ts[0][threadIdx.x] = s[r];
      ts[1][threadIdx.x] = s[r + n];
      ts[2][threadIdx.x] = s[r + 2 * n];
    }
    __syncthreads();
    const int m = min(MDFF_BLOCK, j1 - base);
    if (act) {
      for (int jj = 0; jj < m; ++jj) {
        double rx, ry, rz, d2;
        if (ortho) {
          const double dx = si1 - ts[0][jj], dy = si2 - ts[1][jj], dz = si3 - ts[2][jj];
          rx = (dx - rint(dx)) * a0;
          ry = (dy - rint(dy)) * a4;
          rz = (dz - rint(dz)) * a8;
          d2 = rx * rx + ry * ry + rz * rz;
        } else {
          d2 = min_image(cell, si1 - ts[0][jj], si2 - ts[1][jj], si3 - ts[2][jj], rx, ry, rz);
        }
        const int jg = base + jj;
        if (jg != i && d2 <= cut2) {
          if (FILL && off + c < maxnb) nbr[(size_t)i * maxnb + off + c] = jg;
          ++c;
        }
      }
    }
  }
  if (!FILL && act) scnt[(size_t)sp * n + i] = c;
}

// per atom: the range counts of k_nlist<false> become exclusive offsets (in
// place), their sum the row count; the largest row count above maxnb goes to
// overflow
__global__ __launch_bounds__(MDFF_BLOCK) void k_nlist_offsets(int n, int nsp, int maxnb,
                                                              int* __restrict__ scnt,
                                                              int* __restrict__ cnt,
                                                              int* __restrict__ overflow) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int acc = 0;
  for (int sp = 0; sp < nsp; ++sp) {
    const int v = scnt[(size_t)sp * n + i];
    scnt[(size_t)sp * n + i] = acc;
    acc += v;
  }
  cnt[i] = acc;
  if (acc > maxnb) atomicMax(overflow, acc);
}

template __global__ void k_nlist<false>(int, const double*, Cell, double, int, int, int*, int*);
template __global__ void k_nlist<true>(int, const double*, Cell, double, int, int, int*, int*);

// ---------------------------------------------------------------------------
// Structure factor (multipole_ES_rec, src/field.f90:2697-2742) for all
// half-space k vectors, FP64.  Thread = one k chunk (fixed n1, n2, KC
// consecutive n3); blockIdx.y = atom split.  Per LDS atom tile the block
// builds E_a[t][m] = e^{2 pi i m s_a}, m = 0..kmax_a, by recurrence from the
// per-atom bases; e^{-i x} = conj(e^{i x}) gives the negative m.  Along the
// chunk the phase advances by e3.
//   rho(k) = sum_j (q_j + i k.mu_j - k.Q_j.k / 3) e^{i k.r_j},
//   k.mu_j = 2 pi n.(B^T mu_j),  k.Q_j.k = 4 pi^2 n.(B^T Q_j B).n
// (the quadrupole term with QUAD, src/field.f90:2723-2740)
template <bool CHG, bool DIP, bool QUAD>
__global__ __launch_bounds__(MDFF_BLOCK) void k_rho64(RhoArgs a) {
  extern __shared__ double2 sh[];
  const int TA = RHO_TILE;
  const int L1 = a.kmax[0] + 1, L2 = a.kmax[1] + 1, L3 = a.kmax[2] + 1;
  double2* E1 = sh;                 // [TA][L1]
  double2* E2 = E1 + TA * L1;       // [TA][L2]
  double2* E3 = E2 + TA * L2;       // [TA][L3]
  double* aq = (double*)(E3 + TA * L3);  // [TA]
  double* am = aq + TA;                  // [3][TA]  2 pi B^T mu
  // QUAD: [6][TA] 4 pi^2 / 3 (B^T Q B) as xx yy zz xy xz yz (2 x off-diagonal)
  double* aqd = am + 3 * TA;
  const int ck = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = ck < a.nchunk;
  KChunk C = act ? a.chunks[ck] : a.chunks[0];
  const int KC = a.KC;
  double accR[KC_MAX], accI[KC_MAX];
#pragma unroll
  for (int c = 0; c < KC_MAX; ++c) accR[c] = accI[c] = 0.0;
  const int n = a.n;
  const int j0 = a.j_begin + blockIdx.y * a.atoms_per_split;
  const int j1 = min(a.j_end, j0 + a.atoms_per_split);
  const double tpi = 2.0 * MDFF_PI;
  for (int tb = j0; tb < j1; tb += TA) {
    const int m = min(TA, j1 - tb);
    __syncthreads();
    // tables by recurrence: thread -> (atom, axis)
    for (int w = threadIdx.x; w < 3 * TA; w += blockDim.x) {
      const int t = w / 3, ax = w % 3;
      if (t < m) {
        const double2 e = a.eph[tb + t + ax * n];
        const int L = (ax == 0) ? L1 : (ax == 1 ? L2 : L3);
        double2* E = (ax == 0) ? E1 + t * L1 : (ax == 1 ? E2 + t * L2 : E3 + t * L3);
        double2 p = make_double2(1.0, 0.0);
        E[0] = p;
        for (int l = 1; l < L; ++l) {
          p = make_double2(p.x * e.x - p.y * e.y, p.x * e.y + p.y * e.x);
          E[l] = p;
        }
      }
    }
    for (int t = threadIdx.x; t < m; t += blockDim.x) {
      const int j = tb + t;
      aq[t] = CHG ? a.q[j] : 0.0;
      if (DIP) {
        const double mx = a.mu[j], my = a.mu[j + n], mz = a.mu[j + 2 * n];
        const double* B = a.cell.B;
        am[t] = tpi * (B[0] * mx + B[1] * my + B[2] * mz);
        am[TA + t] = tpi * (B[3] * mx + B[4] * my + B[5] * mz);
        am[2 * TA + t] = tpi * (B[6] * mx + B[7] * my + B[8] * mz);
      }
      if (QUAD) {
        // k.Q.k with the reference's upper-triangle components, Q_xy used
        // for both xy and yx (src/field.f90:2729-2732)
        double Q[3][3];
        const double* th = a.th;
        Q[0][0] = th[j];
        Q[1][1] = th[j + 4 * n];
        Q[2][2] = th[j + 8 * n];
        Q[0][1] = Q[1][0] = th[j + 3 * n];
        Q[0][2] = Q[2][0] = th[j + 6 * n];
        Q[1][2] = Q[2][1] = th[j + 7 * n];
        const double* B = a.cell.B;  // column c = reciprocal vector c
        double P[3][3];
        for (int u = 0; u < 3; ++u)
          for (int v = 0; v < 3; ++v) {
            double acc = 0.0;
            for (int x = 0; x < 3; ++x)
              for (int y = 0; y < 3; ++y) acc += B[x + 3 * u] * Q[x][y] * B[y + 3 * v];
            P[u][v] = acc;
          }
        const double f = tpi * tpi / 3.0;
        aqd[t] = f * P[0][0];
        aqd[TA + t] = f * P[1][1];
        aqd[2 * TA + t] = f * P[2][2];
        aqd[3 * TA + t] = 2.0 * f * P[0][1];
        aqd[4 * TA + t] = 2.0 * f * P[0][2];
        aqd[5 * TA + t] = 2.0 * f * P[1][2];
      }
    }
    __syncthreads();
    if (act) {
      for (int t = 0; t < m; ++t) {
        double2 e1 = E1[t * L1 + C.n1];
        double2 e2 = E2[t * L2 + abs(C.n2)];
        if (C.n2 < 0) e2.y = -e2.y;
        double2 e3 = E3[t * L3 + abs(C.n3s)];
        if (C.n3s < 0) e3.y = -e3.y;
        const double2 st = E3[t * L3 + 1];
        double2 p12 = make_double2(e1.x * e2.x - e1.y * e2.y, e1.x * e2.y + e1.y * e2.x);
        double pr = p12.x * e3.x - p12.y * e3.y;
        double pi = p12.x * e3.y + p12.y * e3.x;
        const double qj = aq[t];
        double tt = 0.0, dt = 0.0;
        if (DIP) {
          tt = C.n1 * am[t] + C.n2 * am[TA + t] + C.n3s * am[2 * TA + t];
          dt = am[2 * TA + t];
        }
        // QUAD: kq(n3) = c0 + n3 (c1 + n3 c2) along the chunk
        double c0 = 0.0, c1 = 0.0, c2 = 0.0;
        if (QUAD) {
          const double n1 = C.n1, n2 = C.n2;
          c0 = n1 * n1 * aqd[t] + n2 * n2 * aqd[TA + t] + n1 * n2 * aqd[3 * TA + t];
          c1 = n1 * aqd[4 * TA + t] + n2 * aqd[5 * TA + t];
          c2 = aqd[2 * TA + t];
        }
#pragma unroll
        for (int c = 0; c < KC_MAX; ++c) {
          if (c < KC) {
            if (CHG) {
              accR[c] = fma(pr, qj, accR[c]);
              accI[c] = fma(pi, qj, accI[c]);
            }
            if (DIP) {
              accR[c] = fma(-pi, tt, accR[c]);
              accI[c] = fma(pr, tt, accI[c]);
              tt += dt;
            }
            if (QUAD) {
              const double n3 = (double)(C.n3s + c);
              const double kq = c0 + n3 * (c1 + n3 * c2);
              accR[c] = fma(-pr, kq, accR[c]);
              accI[c] = fma(-pi, kq, accI[c]);
            }
            const double nr = pr * st.x - pi * st.y;
            pi = pr * st.y + pi * st.x;
            pr = nr;
          }
        }
      }
    }
  }
  if (act) {
    double2* out = a.part + (size_t)blockIdx.y * a.kpad + (size_t)ck * KC;
    for (int c = 0; c < KC; ++c) out[c] = make_double2(accR[c], accI[c]);
  }
}

template __global__ void k_rho64<true, false, false>(RhoArgs);
template __global__ void k_rho64<false, true, false>(RhoArgs);
template __global__ void k_rho64<true, true, false>(RhoArgs);
template __global__ void k_rho64<false, true, true>(RhoArgs);
template __global__ void k_rho64<true, true, true>(RhoArgs);

// Sum of the atom-split partials, g = A_k rho, energy and stress per k
// (src/field.f90:2742-2745, 2803-2815).  Partial rows: 0 u, 1..6 tau.
__global__ __launch_bounds__(MDFF_BLOCK) void k_rho_reduce64(RhoRedArgs a) {
  __shared__ double red[4 * 7];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[7];
#pragma unroll
  for (int c = 0; c < 7; ++c) acc[c] = 0.0;
  if (k < a.kpad) {
    double rR = 0.0, rI = 0.0;
    for (int sp = 0; sp < a.nsplit; ++sp) {
      const double2 v = a.part[(size_t)sp * a.kpad + k];
      rR += v.x;
      rI += v.y;
    }
    const double Ak = a.Ak[k];
    a.g[k] = make_double2(Ak * rR, Ak * rI);
    if (a.do_energy) {
      const double str = (rR * rR + rI * rI) * Ak;
      const double kx = a.kvec[k], ky = a.kvec[k + a.kpad], kz = a.kvec[k + 2 * a.kpad];
      const double kc = a.kcoe[k];
      acc[0] = str;
      acc[1] = (1.0 - kc * kx * kx) * str;
      acc[2] = (1.0 - kc * ky * ky) * str;
      acc[3] = (1.0 - kc * kz * kz) * str;
      acc[4] = -kc * kx * ky * str;
      acc[5] = -kc * kx * kz * str;
      acc[6] = -kc * ky * kz * str;
    }
  }
  if (a.do_energy) {
    block_sum<7>(acc, red);
    if (threadIdx.x == 0)
      for (int c = 0; c < 7; ++c) a.pout[(size_t)c * gridDim.x + blockIdx.x] = acc[c];
  }
}

// ---------------------------------------------------------------------------
// k-space field / EFG apply per atom (multipole_ES_rec, src/field.f90:
// 2746-2801), FP64.  Thread = atom i; blockIdx.y = k split (a contiguous
// range of chunks).  With P = e^{i k.r_i} and g = A_k rho(k):
//   recarg  = Im(g conj(P))... = g_I P_R - g_R P_I   (field, charge force)
//   recarg2 = g_R P_R + g_I P_I                      (EFG, dipole force)
// accumulated in the fractional basis: U_c = sum n_c recarg,
// V_cd = sum n_c n_d recarg2, c,d = 1..3.  The phase walks each (n1,n2) run
// by recurrence; runs and n1 changes advance by e2 and e1.
// Output rows (split-major): [split][comp][i], comp 0..2 U, 3..8 V (11 22 33
// 12 13 23).
template <bool EFG>
__global__ __launch_bounds__(MDFF_BLOCK) void k_apply64(ApplyArgs a) {
  __shared__ KChunk sc[APPLY_TILE];
  __shared__ double2 sg[APPLY_TILE * KC_MAX];
  const int n = a.n;
  const int i = a.i_begin + blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = i < a.i_end;
  const int KC = a.KC;
  const int c0 = blockIdx.y * a.chunks_per_split;
  const int c1 = min(a.nchunk, c0 + a.chunks_per_split);
  double U1 = 0, U2 = 0, U3 = 0;
  double V11 = 0, V22 = 0, V33 = 0, V12 = 0, V13 = 0, V23 = 0;
  double2 e1 = make_double2(1, 0), e2 = make_double2(1, 0), e3 = make_double2(1, 0);
  double s1 = 0, s2 = 0, s3 = 0;
  if (act) {
    e1 = a.eph[i];
    e2 = a.eph[i + n];
    e3 = a.eph[i + 2 * n];
    s1 = a.s[i];
    s2 = a.s[i + n];
    s3 = a.s[i + 2 * n];
  }
  // running phases: P12 = e1^n1 e2^n2, ph = P12 e3^n3
  double2 P12 = make_double2(1, 0), ph = make_double2(1, 0);
  int cur_n1 = -1000000, cur_n2 = -1000000, next_n3 = -1000000;
  for (int cb = c0; cb < c1; cb += APPLY_TILE) {
    const int m = min(APPLY_TILE, c1 - cb);
    __syncthreads();
    for (int w = threadIdx.x; w < m; w += blockDim.x) sc[w] = a.chunks[cb + w];
    for (int w = threadIdx.x; w < m * KC; w += blockDim.x)
      sg[w] = a.g[(size_t)cb * KC + w];
    __syncthreads();
    if (!act) continue;
    for (int t = 0; t < m; ++t) {
      const KChunk C = sc[t];
      if (C.n1 != cur_n1) {
        double sn, cs;
        sincospi(2.0 * (C.n1 * s1 + C.n2 * s2), &sn, &cs);
        P12 = make_double2(cs, sn);
        cur_n1 = C.n1;
        cur_n2 = C.n2;
        sincospi(2.0 * C.n3s * s3, &sn, &cs);
        ph = make_double2(P12.x * cs - P12.y * sn, P12.x * sn + P12.y * cs);
      } else if (C.n2 != cur_n2) {
        // next run of the same n1: advance P12 by e2^(dn2) (dn2 = 1)
        P12 = make_double2(P12.x * e2.x - P12.y * e2.y, P12.x * e2.y + P12.y * e2.x);
        cur_n2 = C.n2;
        double sn, cs;
        sincospi(2.0 * C.n3s * s3, &sn, &cs);
        ph = make_double2(P12.x * cs - P12.y * sn, P12.x * sn + P12.y * cs);
      } else if (C.n3s != next_n3) {
        double sn, cs;
        sincospi(2.0 * C.n3s * s3, &sn, &cs);
        ph = make_double2(P12.x * cs - P12.y * sn, P12.x * sn + P12.y * cs);
      }
      double pr = ph.x, pi = ph.y;
      double R0 = 0, R3 = 0, Q0 = 0, Q3 = 0, Q33 = 0;
      double n3 = (double)C.n3s;
      const double2* g = sg + t * KC;
#pragma unroll
      for (int c = 0; c < KC_MAX; ++c) {
        if (c < KC) {
          const double2 gk = g[c];
          const double ra = gk.y * pr - gk.x * pi;
          R0 += ra;
          R3 = fma(n3, ra, R3);
          if (EFG) {
            const double rb = gk.x * pr + gk.y * pi;
            Q0 += rb;
            Q3 = fma(n3, rb, Q3);
            Q33 = fma(n3 * n3, rb, Q33);
          }
          const double nr = pr * e3.x - pi * e3.y;
          pi = pr * e3.y + pi * e3.x;
          pr = nr;
          n3 += 1.0;
        }
      }
      ph = make_double2(pr, pi);
      next_n3 = C.n3s + KC;
      const double dn1 = C.n1, dn2 = C.n2;
      U1 = fma(dn1, R0, U1);
      U2 = fma(dn2, R0, U2);
      U3 += R3;
      if (EFG) {
        V11 = fma(dn1 * dn1, Q0, V11);
        V22 = fma(dn2 * dn2, Q0, V22);
        V33 += Q33;
        V12 = fma(dn1 * dn2, Q0, V12);
        V13 = fma(dn1, Q3, V13);
        V23 = fma(dn2, Q3, V23);
      }
    }
  }
  if (act) {
    double* out = a.part + (size_t)blockIdx.y * (EFG ? 9 : 3) * n;
    out[i] = U1;
    out[i + n] = U2;
    out[i + 2 * n] = U3;
    if (EFG) {
      out[i + 3 * n] = V11;
      out[i + 4 * n] = V22;
      out[i + 5 * n] = V33;
      out[i + 6 * n] = V12;
      out[i + 7 * n] = V13;
      out[i + 8 * n] = V23;
    }
  }
}

template __global__ void k_apply64<false>(ApplyArgs);
template __global__ void k_apply64<true>(ApplyArgs);

// ---------------------------------------------------------------------------
// Per-call moment sums of multipole_ES (src/field.f90:1864-1880): rows
// 0..2 qtot = sum q r, 3..5 mutot, 6 qsq, 7 musq.
__global__ __launch_bounds__(MDFF_BLOCK) void k_moments(
    int n, const double* __restrict__ pos, const double* __restrict__ q,
    const double* __restrict__ mu, double* __restrict__ part) {
  __shared__ double red[4 * 8];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) acc[k] = 0.0;
  if (i < n) {
    const double qi = q[i];
    const double m0 = mu[i], m1 = mu[i + n], m2 = mu[i + 2 * n];
    acc[0] = qi * pos[i];
    acc[1] = qi * pos[i + n];
    acc[2] = qi * pos[i + 2 * n];
    acc[3] = m0;
    acc[4] = m1;
    acc[5] = m2;
    acc[6] = qi * qi;
    acc[7] = m0 * m0 + m1 * m1 + m2 * m2;
  }
  block_sum<8>(acc, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < 8; ++k) part[(size_t)k * gridDim.x + blockIdx.x] = acc[k];
}

// ---------------------------------------------------------------------------
// Per-atom total of one multipole_ES call (src/field.f90:1933-1990): sums
// the k-split partials, maps U, V back to Cartesian with k = 2 pi B n, adds
// the self (and surface) terms.  In ES_IND mode the result is the induced
// field and E = E_stat + E_ind is formed; in ES_FINAL mode the forces are
// added with coul_unit and ef_t / efg_t are written.
template <int MODE>
__device__ __forceinline__ void combine_atom(const CombineArgs& a, int i, double* pap = nullptr) {
  const int n = a.n;
  const bool EFG = (MODE == ES_FINAL) && a.do_efg;
  const bool EFIELD = (MODE != ES_FINAL) || a.do_efield;
  const int ncomp = (MODE == ES_FINAL && (a.do_efg || a.dip_force)) ? 9 : 3;
  double U[3] = {0, 0, 0}, V[6] = {0, 0, 0, 0, 0, 0};
  for (int sp = 0; sp < a.nsplit; ++sp) {
    const double* p = a.apart + (size_t)sp * ncomp * n;
    U[0] += p[i];
    U[1] += p[i + n];
    U[2] += p[i + 2 * n];
    if (ncomp == 9)
      for (int c = 0; c < 6; ++c) V[c] += p[i + (3 + c) * n];
  }
  const double* B = a.cell.B;
  const double tpi = 2.0 * MDFF_PI;
  // E_rec = -2 (4 pi / V) sum_k k recarg
  const double fe = -2.0 * a.fpi_V * tpi;
  double er[3];
  for (int r = 0; r < 3; ++r)
    er[r] = fe * (B[r] * U[0] + B[r + 3] * U[1] + B[r + 6] * U[2]);
  // raw k-space EFG, 2 (4 pi/V) (2 pi)^2 B V B^T (full symmetric)
  double G[3][3] = {{0}};
  if (ncomp == 9) {
    const double Vm[3][3] = {{V[0], V[3], V[4]}, {V[3], V[1], V[5]}, {V[4], V[5], V[2]}};
    const double fg = 2.0 * a.fpi_V * tpi * tpi;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        double acc = 0.0;
        for (int x = 0; x < 3; ++x)
          for (int y = 0; y < 3; ++y) acc += B[r + 3 * x] * Vm[x][y] * B[c + 3 * y];
        G[r][c] = fg * acc;
      }
  }
  const double m0 = a.mu[i], m1 = a.mu[i + n], m2 = a.mu[i + 2 * n];
  const double mu3[3] = {m0, m1, m2};
  double ef[3];
  for (int r = 0; r < 3; ++r) {
    double v = 0.0;
    if (EFIELD) v = a.ef_dir[i + r * n] + er[r];
    if (a.lsurf) v = v + (a.qm[r] + a.qm[3 + r]);
    ef[r] = v + 2.0 * a.selfa2 * mu3[r];
  }
  if (MODE == ES_STATIC) {
    for (int r = 0; r < 3; ++r) a.E_stat[i + r * n] = ef[r];
  } else if (MODE == ES_IND) {
    if (a.cg_Ap) {
      // conjugate-gradient SCF: Ap = alpha^-1 p - T p, T p = ef
      const int t = a.typ[i];
      double ap[3] = {0, 0, 0};
      if (a.T->polar[t]) {
        const double* ip = a.T->invpol[t];
        for (int r = 0; r < 3; ++r) {
          ap[r] = ip[r] * m0 + ip[r + 3] * m1 + ip[r + 6] * m2 - ef[r];
          *pap += mu3[r] * ap[r];
        }
      }
      for (int r = 0; r < 3; ++r) a.cg_Ap[i + r * n] = ap[r];
    } else {
      for (int r = 0; r < 3; ++r) a.E[i + r * n] = a.E_stat[i + r * n] + ef[r];
    }
  } else {
    if (a.ef_out)
      for (int r = 0; r < 3; ++r) a.ef_out[i + r * n] = ef[r];
    if (a.efg_out) {
      const double qi = a.q[i];
      for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
          double v = 0.0;
          if (EFG) {
            v = a.efg_dir[i + n * (r + 3 * c)];
            if (r <= c) v += G[r][c];  // k-space EFG has the upper 6 only
          }
          if (r == c) v += -2.0 * a.selfa2 * qi;
          a.efg_out[i + n * (r + 3 * c)] = v * a.coul_unit;
        }
    }
    // forces: f += (f_rec + f_dir [+ f_surf]) coul_unit
    const double qi = a.q[i];
    for (int r = 0; r < 3; ++r) {
      double fr = qi * er[r];
      if (a.dip_force) fr += G[r][0] * m0 + G[r][1] * m1 + G[r][2] * m2;
      double v = fr + a.f_dir[i + r * n];
      if (a.lsurf) v = v + (-(qi * (a.qm[r] + a.qm[3 + r])) * a.fpi_3V);
      a.f[i + r * n] = a.f[i + r * n] + v * a.coul_unit;
    }
  }
}

// SCF stopping criterion of atom i (src/field.f90:3378-3400): squared
// deviations of mu / alpha from the new field, and the polarizable count.
// the rows of the SCF stopping criterion of atom i (src/field.f90:3378-3400):
// dipole terms on ldip_polar species, quadrupole terms (theta_aa /
// polquad_aaa - E_a) on lquad_polar species, each species kind counted in
// npol
__device__ __forceinline__ void resid_terms(const TypeTables* T, int t, int n, int i,
                                            const double* mu, const double* th,
                                            const double* E, double (&acc)[2]) {
  if (T->polar[t]) {
    acc[1] += 1.0;
    for (int r = 0; r < 3; ++r) {
      const double pa = T->pol[t][r + 3 * r];
      if (pa == 0.0) continue;
      const double d = mu[i + r * n] / pa - E[i + r * n];
      acc[0] += d * d;
    }
  }
  if (th && T->qpolar[t]) {
    acc[1] += 1.0;
    for (int r = 0; r < 3; ++r) {
      const double pq = T->polq[t][r + 3 * (r + 3 * r)];
      if (pq == 0.0) continue;
      const double d = th[i + (size_t)(r + 3 * r) * n] / pq - E[i + r * n];
      acc[0] += d * d;
    }
  }
}

__device__ __forceinline__ void resid_atom(const CombineArgs& a, int i, double (&acc)[2]) {
  resid_terms(a.T, a.typ[i], a.n, i, a.mu, a.th, a.E, acc);
}

template <int MODE>
__global__ __launch_bounds__(MDFF_BLOCK) void k_combine64(CombineArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (MODE != ES_IND || !a.rpart) {
    if (i < a.n) combine_atom<MODE>(a, i);
    return;
  }
  // ES_IND with the SCF residual fused: block partials rows 0, 1 of rpart
  __shared__ double red[4 * 2];
  double acc[2] = {0.0, 0.0};
  if (i < a.n) {
    if (a.cg_Ap) {
      combine_atom<MODE>(a, i, &acc[0]);
    } else {
      combine_atom<MODE>(a, i);
      resid_atom(a, i, acc);
    }
  }
  block_sum<2>(acc, red);
  if (threadIdx.x == 0) {
    a.rpart[blockIdx.x] = acc[0];
    a.rpart[gridDim.x + blockIdx.x] = acc[1];
  }
}

template __global__ void k_combine64<ES_STATIC>(CombineArgs);
template __global__ void k_combine64<ES_IND>(CombineArgs);
template __global__ void k_combine64<ES_FINAL>(CombineArgs);

// ---------------------------------------------------------------------------
// induced_moment (src/field.f90:1716-1803): mu = alpha E on polarizable
// atoms; partial row 0 = sum mu alpha^-1 mu (u_pol before the 1/2).
// With mu4 (the FP32 SCF field of the mixed path), the moments are also
// stored as float4 at the atom's spatial rank (k_pack_mu_ranked fused).
// induced_moment (src/field.f90:1716-1803) of one atom: mu = alpha E on
// ldip_polar or lquad_polar species, theta = polquad E (when th is given),
// and its u_pol term: mu alpha^-1 mu / 2 on ldip_polar species,
// sum theta_ab^2 / (6 polquad_111) on lquad_polar ones.
__device__ __forceinline__ double induced_atom(const TypeTables* T, int t, int n, int i,
                                               const double* E, double (&m)[3], double* th) {
  double u = 0.0;
  m[0] = m[1] = m[2] = 0.0;
  if (!T->polar[t] && !T->qpolar[t]) {
    if (th)
      for (int k = 0; k < 9; ++k) th[i + (size_t)k * n] = 0.0;
    return 0.0;
  }
  const double e[3] = {E[i], E[i + n], E[i + 2 * n]};
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) m[r] = m[r] + T->pol[t][r + 3 * c] * e[c];
  if (T->polar[t]) {
    double v = 0.0;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) v = v + m[r] * T->invpol[t][r + 3 * c] * m[c];
    u = 0.5 * v;
  }
  if (th) {
    double uq = 0.0;
    for (int k = 0; k < 9; ++k) {
      double v = 0.0;
      for (int g = 0; g < 3; ++g) v = v + T->polq[t][g + 3 * k] * e[g];
      th[i + (size_t)k * n] = v;
      if (T->qpolar[t]) uq = uq + v * v / T->polq[t][0];
    }
    u = u + uq / 6.0;
  }
  return u;
}

__global__ __launch_bounds__(MDFF_BLOCK) void k_induced(
    int n, const int* __restrict__ typ, const TypeTables* __restrict__ T,
    const double* __restrict__ E, double* __restrict__ mu, double* __restrict__ part,
    const int* __restrict__ rank, float4* __restrict__ mu4, double* __restrict__ th) {
  __shared__ double red[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[1] = {0.0};
  if (i < n) {
    double m[3];
    acc[0] = 2.0 * induced_atom(T, typ[i], n, i, E, m, th);  // halved with the sum
    mu[i] = m[0];
    mu[i + n] = m[1];
    mu[i + 2 * n] = m[2];
    if (mu4) mu4[rank[i]] = make_float4((float)m[0], (float)m[1], (float)m[2], 0.f);
  }
  block_sum<1>(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// SCF stopping criterion (src/field.f90:3378-3400): row 0 = sum of squared
// residuals mu/alpha_aa - E_a, row 1 = number of polarizable atoms.
__global__ __launch_bounds__(MDFF_BLOCK) void k_scf_resid(
    int n, const int* __restrict__ typ, const TypeTables* __restrict__ T,
    const double* __restrict__ mu, const double* __restrict__ E,
    double* __restrict__ part) {
  __shared__ double red[4 * 2];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[2] = {0.0, 0.0};
  if (i < n) {
    const int t = typ[i];
    if (T->polar[t]) {
      acc[1] = 1.0;
      for (int r = 0; r < 3; ++r) {
        const double pa = T->pol[t][r + 3 * r];
        if (pa == 0.0) continue;
        const double d = mu[i + r * n] / pa - E[i + r * n];
        acc[0] = acc[0] + d * d;
      }
    }
  }
  block_sum<2>(acc, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = acc[0];
    part[gridDim.x + blockIdx.x] = acc[1];
  }
}

// Dipole extrapolation and history updates (src/field.f90:3807-3992).
//   op 0: mu = h0                         (poly order 0)
//   op 1: mu = 2 h0 - h1, h1 = h0         (poly order 1, itime > 1)
//   op 2: mu = h0, h1 = h0                (poly order 1, itime <= 1)
//   op 3: ASPC predictor, mu = sum_k b_k h_k, shift h_k <- h_{k-1}
//   op 4: h0 = mu                         (store after the SCF)
//   op 5: ASPC corrector, mu = w alpha E + (1 - w) mu_p (+ u_pol row, and
//         with rpart the SCF residual rows of the corrected moments)
// hist layout: Fortran dipia_ind_t(nh, n, 3): [t + nh*(i + n*a)].
__global__ __launch_bounds__(MDFF_BLOCK) void k_extrap(ExtrapArgs a) {
  __shared__ double red[4];
  __shared__ double red2[4 * 2];
  const int n = a.n;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[1] = {0.0};
  double rs[2] = {0.0, 0.0};
  if (i < n) {
    const int nh = a.nh;
    for (int r = 0; r < 3; ++r) {
      double* h = a.hist + (size_t)nh * (i + (size_t)n * r);
      double* m = a.mu + i + (size_t)r * n;
      switch (a.op) {
        case 0: *m = h[0]; break;
        case 1: *m = 2.0 * h[0] - h[1]; h[1] = h[0]; break;
        case 2: *m = h[0]; h[1] = h[0]; break;
        case 3: {
          double v = 0.0;
          for (int k = 0; k <= a.ord; ++k) v = v + a.b[k] * h[k];
          *m = v;
          for (int k = a.ord; k >= 1; --k) h[k] = h[k - 1];
          a.mup[i + (size_t)r * n] = v;
          break;
        }
        case 4: h[0] = *m; break;
        default: break;
      }
    }
    if (a.op == 5) {
      // induced_moment of the new field (theta_ind as it leaves it), then
      // the dipoles mixed with the predictor's (src/field.f90:3982-3987)
      const int t = a.typ[i];
      double mm[3];
      acc[0] = 2.0 * induced_atom(a.T, t, n, i, a.E, mm, a.th);  // halved with the sum
      for (int r = 0; r < 3; ++r)
        a.mu[i + r * n] = a.w * mm[r] + (1.0 - a.w) * a.mup[i + r * n];
      // the stopping criterion of this iteration is taken on the corrected
      // moments (src/field.f90:3368-3400: the corrector runs before it)
      if (a.rpart) resid_terms(a.T, t, n, i, a.mu, a.th, a.E, rs);
    }
  }
  if (a.op == 5) {
    block_sum<1>(acc, red);
    if (threadIdx.x == 0) a.part[blockIdx.x] = acc[0];
    if (a.rpart) {
      block_sum<2>(rs, red2);
      if (threadIdx.x == 0) {
        a.rpart[blockIdx.x] = rs[0];
        a.rpart[gridDim.x + blockIdx.x] = rs[1];
      }
    }
  }
}

// Largest squared displacement since the list build: block maxima.
__global__ __launch_bounds__(MDFF_BLOCK) void k_disp(int n, const double* __restrict__ pos,
                                                     const double* __restrict__ ref,
                                                     double* __restrict__ part) {
  __shared__ double red[MDFF_BLOCK / 64];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double d = 0.0;
  if (i < n) {
    const double dx = pos[i] - ref[i], dy = pos[i + n] - ref[i + n],
                 dz = pos[i + 2 * n] - ref[i + 2 * n];
    d = dx * dx + dy * dy + dz * dz;
  }
  for (int off = 32; off > 0; off >>= 1) d = fmax(d, __shfl_down(d, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = d;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int w = 0; w < MDFF_BLOCK / 64; ++w) m = fmax(m, red[w]);
    part[blockIdx.x] = m;
  }
}

__global__ __launch_bounds__(MDFF_BLOCK) void k_reduce_max(const double* __restrict__ part,
                                                           int nb, double* __restrict__ out) {
  __shared__ double red[MDFF_BLOCK / 64];
  double m = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) m = fmax(m, part[b]);
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int w = 0; w < MDFF_BLOCK / 64; ++w) r = fmax(r, red[w]);
    *out = r;
  }
}

// Row sums of block partials in a fixed order: out[r] = sum_b part[r*nb+b].
__global__ __launch_bounds__(MDFF_BLOCK) void k_reduce_rows(
    const double* __restrict__ part, int nb, double* __restrict__ out) {
  __shared__ double red[4];
  const int r = blockIdx.x;
  double acc[1] = {0.0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x) acc[0] += part[(size_t)r * nb + b];
  block_sum<1>(acc, red);
  if (threadIdx.x == 0) out[r] = acc[0];
}

// k_reduce_rows into two places: out and out2 (the device address of pinned
// host memory, read by the host after a fenced event: no copy launch)
__global__ __launch_bounds__(MDFF_BLOCK) void k_reduce_rows2(
    const double* __restrict__ part, int nb, double* __restrict__ out, double* out2) {
  __shared__ double red[4];
  const int r = blockIdx.x;
  double acc[1] = {0.0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x) acc[0] += part[(size_t)r * nb + b];
  block_sum<1>(acc, red);
  if (threadIdx.x == 0) {
    out[r] = acc[0];
    out2[r] = acc[0];
  }
}

// ---------------------------------------------------------------------------
// Preconditioned conjugate gradient of the induced dipoles (an alternative
// to the Jacobi iteration of moment_from_pola_scf, src/field.f90:3200-3458,
// for the same fixed point mu = alpha (E_stat + T mu)).  Per-atom 3x3 blocks,
// FP64; block partials in rows of a.part for k_reduce_rows.
__device__ __forceinline__ void mat3(const double* m, const double* v, int i, int n,
                                     double (&o)[3]) {
  for (int r = 0; r < 3; ++r)
    o[r] = m[r] * v[i] + m[r + 3] * v[i + n] + m[r + 6] * v[i + 2 * n];
}

// r = E - alpha^-1 x, z = alpha r, p = z; rows: r.z, |r|^2, polarizable count
__global__ __launch_bounds__(MDFF_BLOCK) void k_cg_init(CgArgs a) {
  __shared__ double red[4 * 3];
  const int n = a.n;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[3] = {0.0, 0.0, 0.0};
  if (i < n) {
    const int t = a.typ[i];
    double r[3] = {0, 0, 0}, z[3] = {0, 0, 0};
    if (a.T->polar[t]) {
      double ax[3];
      mat3(a.T->invpol[t], a.x, i, n, ax);
      for (int c = 0; c < 3; ++c) r[c] = a.E[i + c * n] - ax[c];
      for (int c = 0; c < 3; ++c)
        z[c] = a.T->pol[t][c] * r[0] + a.T->pol[t][c + 3] * r[1] + a.T->pol[t][c + 6] * r[2];
      for (int c = 0; c < 3; ++c) {
        acc[0] += r[c] * z[c];
        acc[1] += r[c] * r[c];
      }
      acc[2] = 1.0;
    }
    for (int c = 0; c < 3; ++c) {
      a.r[i + c * n] = r[c];
      a.z[i + c * n] = z[c];
      a.p[i + c * n] = z[c];
    }
    if (a.p4) a.p4[a.rank[i]] = make_float4((float)z[0], (float)z[1], (float)z[2], 0.f);
  }
  block_sum<3>(acc, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) a.part[k * gridDim.x + blockIdx.x] = acc[k];
}

// step = rz_old / pAp: x += step p, r -= step Ap, z = alpha r; rows: r.z (new),
// |r|^2, polarizable count.  p.Ap is summed from the combine's block
// partials by every block, in the fixed order of k_reduce_rows.
__global__ __launch_bounds__(MDFF_BLOCK) void k_cg_update(CgArgs a) {
  __shared__ double red[4 * 3];
  __shared__ double s_pap;
  const int n = a.n;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  // the atom's vectors first: their loads run under the reduction of p.Ap
  const int t = i < n ? a.typ[i] : 0;
  const bool pol = i < n && a.T->polar[t];
  double xv[3] = {0, 0, 0}, pv[3] = {0, 0, 0}, rv[3] = {0, 0, 0}, av[3] = {0, 0, 0};
  if (pol)
    for (int c = 0; c < 3; ++c) {
      xv[c] = a.x[i + c * n];
      pv[c] = a.p[i + c * n];
      rv[c] = a.r[i + c * n];
      av[c] = a.Ap[i + c * n];
    }
  const double rz = a.sc[a.s_rz];
  {
    double v[1] = {0.0};
    for (int b = threadIdx.x; b < a.npap; b += blockDim.x) v[0] += a.pap_part[b];
    block_sum<1>(v, red);
    if (threadIdx.x == 0) s_pap = v[0];
    __syncthreads();
  }
  const double pap = s_pap;
  const double step = pap != 0.0 ? rz / pap : 0.0;
  double acc[3] = {0.0, 0.0, 0.0};
  if (i < n) {
    if (pol) {
      double r[3];
      for (int c = 0; c < 3; ++c) {
        a.x[i + c * n] = xv[c] + step * pv[c];
        r[c] = rv[c] - step * av[c];
        a.r[i + c * n] = r[c];
      }
      for (int c = 0; c < 3; ++c) {
        const double z =
            a.T->pol[t][c] * r[0] + a.T->pol[t][c + 3] * r[1] + a.T->pol[t][c + 6] * r[2];
        a.z[i + c * n] = z;
        acc[0] += r[c] * z;
        acc[1] += r[c] * r[c];
      }
      acc[2] = 1.0;
    }
  }
  block_sum<3>(acc, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) a.part[k * gridDim.x + blockIdx.x] = acc[k];
}

// p = z + (rz_new / rz_old) p, and its FP32 rank-ordered copy.  rz_new and
// the other sums of the update (rows r.z, |r|^2, polarizable count of
// a.part, one partial per block of the same grid) are summed by every block
// in the fixed order of k_reduce_rows; block 0 stores the three sums at
// a.sc_out and a.sc_host.
__global__ __launch_bounds__(MDFF_BLOCK) void k_cg_dir(CgArgs a) {
  __shared__ double red[4 * 3];
  __shared__ double s_rzn;
  const int n = a.n;
  // the atom's vectors first: their loads run under the reduction
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double zv[3] = {0, 0, 0}, pv[3] = {0, 0, 0};
  if (i < n)
    for (int c = 0; c < 3; ++c) {
      zv[c] = a.z[i + c * n];
      pv[c] = a.p[i + c * n];
    }
  const double rz = a.sc[a.s_rz];
  {
    const int nb = gridDim.x;
    double v[3] = {0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
      for (int k = 0; k < 3; ++k) v[k] += a.part[(size_t)k * nb + b];
    block_sum<3>(v, red);
    if (threadIdx.x == 0) {
      s_rzn = v[0];
      if (blockIdx.x == 0)
        for (int k = 0; k < 3; ++k) {
          a.sc_out[k] = v[k];
          a.sc_host[k] = v[k];
        }
    }
    __syncthreads();
  }
  if (i >= n) return;
  const double beta = rz != 0.0 ? s_rzn / rz : 0.0;
  double p[3];
  for (int c = 0; c < 3; ++c) {
    p[c] = zv[c] + beta * pv[c];
    a.p[i + c * n] = p[c];
  }
  if (a.p4) a.p4[a.rank[i]] = make_float4((float)p[0], (float)p[1], (float)p[2], 0.f);
  if (a.wt)
    for (int c = 0; c < 3; ++c) {
      const double m = a.wB[3 * c] * p[0] + a.wB[3 * c + 1] * p[1] + a.wB[3 * c + 2] * p[2];
      a.wt[(size_t)c * a.wt_npad + i] = (float)(m * 6.283185307179586) * a.wscale;
    }
}

// u_pol partials of x (the quantity k_induced accumulates): x alpha^-1 x
__global__ __launch_bounds__(MDFF_BLOCK) void k_cg_upol(CgArgs a) {
  __shared__ double red[4];
  const int n = a.n;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[1] = {0.0};
  if (i < n) {
    const int t = a.typ[i];
    if (a.T->polar[t]) {
      double ix[3];
      mat3(a.T->invpol[t], a.x, i, n, ix);
      for (int c = 0; c < 3; ++c) acc[0] += a.x[i + c * n] * ix[c];
    }
  }
  block_sum<1>(acc, red);
  if (threadIdx.x == 0) a.part[blockIdx.x] = acc[0];
}

// out = a + b elementwise (total dipoles mu = dipia + mu_ind)
__global__ __launch_bounds__(MDFF_BLOCK) void k_add(int m, const double* __restrict__ a,
                                                    const double* __restrict__ b,
                                                    double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) out[i] = a[i] + b[i];
}

}  // namespace mdff
