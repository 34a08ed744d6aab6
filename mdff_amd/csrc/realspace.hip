// realspace.hip -- real-space kernels (short-range pairs and the Ewald
// real-space multipole sum), templated on the arithmetic type R: double for
// the plain FP64 path, float for the mixed path.  Block partial sums of
// energies, virials and stress are always reduced in double.
#include <type_traits>

#include "kernels.h"
#include "rtypes.h"

namespace mdff {

// ---------------------------------------------------------------------------
// Short-range pair forces: n-m LJ (engforce_nmlj_pbc, src/field.f90:1181),
// BMHFTD (engforce_bmhftd_pbc, :2876) and Morse, per centre atom over its
// full list, energies halved.  The first short-range term zeroes the
// forces; with BMHFTD on, the LJ forces and stress are dropped as the
// reference's routine overwrites them (:2902-2907).  Partial rows:
//   0 u_lj 1 vir_lj 2 u_bm 3 vir_bm 4 u_mo 5 vir_mo 6..11 tau (xx yy zz xy
//   xz yz) of the terms that contribute forces.
template <typename R, bool BATCH>
__global__ __launch_bounds__(MDFF_BLOCK) void k_shortrange(SrArgs a) {
  __shared__ double red[4 * 12];
  const int n = a.n;
  const bool LJ = a.flags & SR_LJ, BM = a.flags & SR_BM, BMD = a.flags & SR_BMD,
             MO = a.flags & SR_MO;
  const PairTables* T = a.T;
  // FP32: the BMHFTD parameters of the type pairs in LDS, (A, B, C, D) and
  // BD (the FP32 kernels convert them to float anyway)
  constexpr bool F32 = sizeof(R) == sizeof(float);
  __shared__ float4 sbm[F32 ? 256 : 1];
  __shared__ float sbd[F32 ? 256 : 1];
  if constexpr (F32) {
    if (BM) {
      for (int t = threadIdx.x; t < 256; t += blockDim.x) {
        sbm[t] = make_float4((float)T->bm_A[t], (float)T->bm_B[t], (float)T->bm_C[t],
                             (float)T->bm_D[t]);
        sbd[t] = (float)T->bm_BD[t];
      }
      __syncthreads();
    }
  }
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int i0 = tid / LPA, sub = tid % LPA;
  const int i = (a.perm && i0 < n) ? a.perm[i0] : i0;  // atoms in spatial order
  R acc[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) acc[k] = R(0);
  R fx = 0, fy = 0, fz = 0;
  if (i < n) {
    const double si1 = a.s[i], si2 = a.s[i + n], si3 = a.s[i + 2 * n];
    const int ti = a.typ[i];
    const int c = a.cnt[i];
    const int cf = BATCH ? a.cfg[i] : 0;
    const Cell& cell = BATCH ? a.bc[cf].cell : a.cell;
    const R rc2 = R(BATCH ? a.bc[cf].rc2 : a.rc2);
    const int* nb = a.nbr + (size_t)i * a.maxnb;
    // two-stage pipeline over the neighbours: a neighbour's position and
    // type are loaded one iteration ahead, its index two ahead; the loads are
    // unconditional (slots past the count read the centre atom), so that the
    // count of loads in flight stays exact
    auto nbj = [&](int k) {
      const int v = nb[min(k, c - 1)];
      return k < c ? v : i;
    };
    int jn = 0, jnn = 0, tjn = 0;
    double4 sjn = make_double4(0.0, 0.0, 0.0, 0.0);
    if (c > 0) {
      jn = nbj(sub);
      sjn = a.s4[jn];
      tjn = BATCH ? a.typ[jn] : 0;
      jnn = nbj(sub + LPA);
    }
    for (int k = sub; k < c; k += LPA) {
      const int tj = BATCH ? tjn : s4_type(sjn);
      const double4 sj = sjn;
      jn = jnn;
      sjn = a.s4[jn];
      if (BATCH) tjn = a.typ[jn];
      jnn = nbj(k + 2 * LPA);
      R rx, ry, rz;
      const R rsq = min_image_t<R>(cell, si1 - sj.x, si2 - sj.y, si3 - sj.z, rx, ry, rz);
      if (!(rsq < rc2)) continue;
      const int t = ti + MDFF_NT * tj;
      R w = R(0);
      if (LJ) {
        const R sr2 = R(T->lj_sigsq[t]) / rsq;
        const R srp = rpow<R>(sr2, R(T->lj_ptwo[t]));
        const R srq = rpow<R>(sr2, R(T->lj_qtwo[t]));
        R u = R(T->lj_epsp[t]) * (R(T->lj_p[t]) * srq - R(T->lj_q[t]) * srp);
        if (a.flags & SR_TRUNC1) u = u - R(T->lj_uc[t]);
        if (a.flags & SR_TRUNC2) u = u + R(T->lj_uc1[t]) * rsq - R(T->lj_uc2[t]);
        const R wl = R(T->lj_fc[t]) * (srq - srp) * sr2;
        acc[0] += R(0.5) * u;
        acc[1] += R(0.5) * wl * rsq;
        if (!BM) w += wl;
      }
      float4 bmf = make_float4(0.f, 0.f, 0.f, 0.f);
      if constexpr (F32)
        if (BM) bmf = sbm[t];
      const R bA = F32 ? R(bmf.x) : R(T->bm_A[t]);
      if (BM && (F32 ? bmf.x != 0.f : T->bm_A[t] != 0.0)) {
        const R bB = F32 ? R(bmf.y) : R(T->bm_B[t]);
        const R bC = F32 ? R(bmf.z) : R(T->bm_C[t]);
        const R bD = F32 ? R(bmf.w) : R(T->bm_D[t]);
        const R ir2 = R(1) / rsq;
        const R d = sqrt(rsq);
        const R erh = bA * rexp<R>(-bB * d);
        R ir6 = ir2 * ir2 * ir2;
        R ir8 = ir6 * ir2;
        ir6 = ir6 * bC;
        ir8 = ir8 * bD;
        R ir6d = ir6, ir8d = ir8, fd6 = 0, fd8 = 0;
        if (BMD) {
          R f6, f8;
          const R bBD = F32 ? R(sbd[t]) : R(T->bm_BD[t]);
          tt_damp_t<R>(a.tt, bBD, R(1), d, 6, f6, fd6);
          tt_damp_t<R>(a.tt, bBD, R(1), d, 8, f8, fd8);
          ir6d = ir6 * f6;
          ir8d = ir8 * f8;
        }
        const R ir7 = R(6) * ir6d / d;
        const R ir9 = R(8) * ir8d / d;
        acc[2] += R(0.5) * (erh - ir6d - ir8d);
        const R wb = (bB * erh - ir7 - ir9 + ir6 * fd6 + ir8 * fd8) / d;
        acc[3] += R(0.5) * wb * rsq;
        w += wb;
      }
      if (MO) {
        const R d = sqrt(rsq);
        const R X = rexp<R>(-d * R(T->mo_rho[t])) * R(T->mo_rs[t]);
        acc[4] += R(0.5) * R(T->mo_eps[t]) * ((X - R(1)) * (X - R(1)) - R(1));
        const R wm = R(2) * R(T->mo_eps[t]) * R(T->mo_rho[t]) * X * (X - R(1)) / d;
        acc[5] += R(0.5) * wm * rsq;
        w += wm;
      }
      const R gx = w * rx, gy = w * ry, gz = w * rz;
      fx += gx;
      fy += gy;
      fz += gz;
      acc[6] += R(0.5) * rx * gx;
      acc[7] += R(0.5) * ry * gy;
      acc[8] += R(0.5) * rz * gz;
      acc[9] += R(0.5) * rx * gy;
      acc[10] += R(0.5) * rx * gz;
      acc[11] += R(0.5) * ry * gz;
    }
  }
  for (int off = LPA / 2; off > 0; off >>= 1) {
    fx += __shfl_xor(fx, off, 64);
    fy += __shfl_xor(fy, off, 64);
    fz += __shfl_xor(fz, off, 64);
  }
  if (i < n && sub == 0) {
    a.f[i] = fx;
    a.f[i + n] = fy;
    a.f[i + 2 * n] = fz;
  }
  double dacc[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) dacc[k] = (double)acc[k];
  block_sum<12>(dacc, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < 12; ++k) a.part[(size_t)k * gridDim.x + blockIdx.x] = dacc[k];
}

template __global__ void k_shortrange<double, false>(SrArgs);
template __global__ void k_shortrange<float, false>(SrArgs);
template __global__ void k_shortrange<double, true>(SrArgs);
template __global__ void k_shortrange<float, true>(SrArgs);

// ---------------------------------------------------------------------------
// Ewald real space, multipole_ES_dir (src/field.f90:2057-2643), charges and
// dipoles, per centre atom i over its full list.  The reference visits each
// pair once as (ia<ja) and updates both atoms; here each directed pair
// (i,j) adds i's share, with r = r_i - r_j:
//   field      qq  q_j r (F1 - F1d(p_i)) / d^3        dd  T2 mu_j
//   EFG        qq  -q_j T2                            dd  T3 mu_j
//   force      qq  q_i q_j r F1 / d^3                 dd  mu_i T3 mu_j
//              qd  q_i (T2 - T2d(p_j)) mu_j - q_j (T2 - T2d(p_i)) mu_i
//   energy     qq  q_i q_j T0 / 2                     dd  -mu_i T2 mu_j / 2
//              qd  q_j (T1 - T1d(p_i)) . mu_i
// where p_i are the damping parameters the reference uses for atom i of the
// pair (dmp_* slot 0 if i < j in the caller's numbering, s4.w, slot 1
// otherwise).
// Partial rows: 0 u, 1..6 tau (xx yy zz xy xz yz), before the 1/(2 Omega).
template <int MODE, typename R, bool BATCH>
__global__ __launch_bounds__(MDFF_BLOCK) __attribute__((amdgpu_waves_per_eu(
    (MODE == ES_STATIC && sizeof(R) == sizeof(float) && !BATCH) ? 7 : 1))) void k_es_dir(EsDirArgs a) {
  __shared__ double red[4 * 7];
  // FP32: the Tang-Toennies damping table of the pair types in LDS, one
  // float4 (on, b, c, k) per (slot o, type pair) (the FP32 kernels convert
  // b and c to float anyway)
  constexpr bool F32 = sizeof(R) == sizeof(float);
  using R4 = typename std::conditional<F32, float4, double4>::type;
  __shared__ float4 sdmp[F32 ? 2 * 256 : 1];
  if constexpr (F32) {
    if (MODE != ES_IND) {
      for (int t = threadIdx.x; t < 2 * 256; t += blockDim.x) {
        const int o = t >> 8, pr = t & 255;
        sdmp[t] = make_float4((float)a.T->dmp_on[o][pr], (float)a.T->dmp_b[o][pr],
                              (float)a.T->dmp_c[o][pr], (float)a.T->dmp_k[o][pr]);
      }
      __syncthreads();
    }
  }
  const int n = a.n;
  const bool QQ = a.task & 1, QD = a.task & 2, DD = a.task & 4;
  const bool EFIELD = (MODE != ES_FINAL) || a.do_efield;
  const bool EFG = (MODE == ES_FINAL) && a.do_efg;
  const bool FORCE = (MODE == ES_FINAL);
  const bool DAMP = (MODE != ES_IND);
  // the static FP32 pass of one configuration may write the SCF pair table
  constexpr bool TAB = (MODE == ES_STATIC) && F32 && !BATCH;
  static_assert(LPA <= 32 && 64 % LPA == 0, "lane groups inside one wave");
  int tbase = 0;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int row_v = tid / LPA, sub = tid % LPA;
  // processing row: the v-th listed rank (rows past the list inactive) or v
  const int i0 = a.rows ? (row_v < a.nrows ? a.rows[row_v] : n) : row_v;
  const int i = (a.perm && i0 < n) ? a.perm[i0] : i0;  // atoms in spatial order
  const int cf = (BATCH && i < n) ? a.cfg[i] : 0;
  const Cell& cell = BATCH ? a.bc[cf].cell : a.cell;
  const R alpha = R(BATCH ? a.bc[cf].alpha : a.alpha), alpha2 = alpha * alpha,
          alpha3 = alpha2 * alpha, alpha5 = alpha3 * alpha2;
  const R rpiroot = R(1.0 / sqrt(MDFF_PI));
  const R cut2 = R(BATCH ? a.bc[cf].cut2 : a.cut2);
  const R c43a3 = R(4) * alpha3 / R(3), c815a5 = R(8) * alpha5 / R(15);  // FP32 constants
  R acc[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) acc[k] = R(0);
  R e0 = 0, e1 = 0, e2 = 0;
  R g00 = 0, g11 = 0, g22 = 0, g01 = 0, g02 = 0, g12 = 0;
  R f0 = 0, f1 = 0, f2 = 0;
  if (i < n) {
    const double si1 = a.s[i], si2 = a.s[i + n], si3 = a.s[i + 2 * n];
    // the caller's index of atom i (s4.w), for the pair-order convention
    const double wi = BATCH ? 0.0 : a.s4[i].w;
    const int ti = a.typ[i];
    const R qi = R(a.q[i]);
    const R mi0 = R(a.mu[i]), mi1 = R(a.mu[i + n]), mi2 = R(a.mu[i + 2 * n]);
    const bool dip_i = (mi0 != R(0) || mi1 != R(0) || mi2 != R(0));
    const int c = a.cnt[i];
    const int* nb = a.nbr + (size_t)i * a.maxnb;
    // the next neighbour's index one iteration ahead, and all of a
    // neighbour's data loaded together, before its cutoff test
    int jn = sub < c ? nb[sub] : 0;
    for (int k = sub; k < c; k += LPA) {
      const int j = jn;
      if (k + LPA < c) jn = nb[k + LPA];
      const double4 sj = a.s4[j];
      // charge and dipole: one gather of the packed (q, mu) when given
      R qj, mj0, mj1, mj2;
      if (a.qm) {
        const R4 v = reinterpret_cast<const R4*>(a.qm)[j];
        qj = v.x;
        mj0 = v.y;
        mj1 = v.z;
        mj2 = v.w;
      } else {
        qj = R(a.q[j]);
        mj0 = R(a.mu[j]);
        mj1 = R(a.mu[j + n]);
        mj2 = R(a.mu[j + 2 * n]);
      }
      const int tj = BATCH ? a.typ[j] : s4_type(sj);
      R rx, ry, rz;
      const R d2 = min_image_t<R>(cell, si1 - sj.x, si2 - sj.y, si3 - sj.z, rx, ry, rz);
      // SCF pair table: the entry's slot among the tabled pairs (inside the
      // cutoff, both species in polmask) of the atom's lane group, in the
      // group's visiting order
      int tpos = 0;
      bool tin = false;
      if constexpr (TAB) {
        if (a.ddtab) {
          tin = !(d2 > cut2) && ((a.polmask >> ti) & (a.polmask >> tj) & 1u);
          const unsigned gm = (unsigned)(__ballot(tin) >> (threadIdx.x & (64 - LPA))) &
                              (unsigned)((1ull << LPA) - 1ull);
          tpos = tbase + __popc(gm & ((1u << sub) - 1u));
          tbase += __popc(gm);
        }
      }
      if (d2 > cut2) continue;
      const bool dip_j = (mj0 != R(0) || mj1 != R(0) || mj2 != R(0));
      const bool dd = DD && (dip_i || dip_j);
      const bool qd = QD && (dip_i || dip_j);
      const int o = (BATCH ? (i < j) : (wi < sj.w)) ? 0 : 1;
      const int pij = ti + MDFF_NT * tj;
      const int pji = tj + MDFF_NT * ti;
      float4 dmij = make_float4(0.f, 0.f, 0.f, 0.f);
      if constexpr (F32)
        if (DAMP) dmij = sdmp[o * 256 + pij];
      const bool ldamp = DAMP && (F32 ? dmij.x != 0.f : a.T->dmp_on[o][pij] != 0);
      R d, dm1, dm3, dm5, dm7;
      if constexpr (sizeof(R) == sizeof(float)) {
        // FP32: one reciprocal square root, powers by products
        dm1 = rsqrtf(d2);
        d = d2 * dm1;
        const R dm2 = dm1 * dm1;
        dm3 = dm2 * dm1;
        dm5 = dm3 * dm2;
        dm7 = dm5 * dm2;
      } else {
        d = sqrt(d2);
        dm1 = R(1) / d;
        dm3 = dm1 / d2;
        dm5 = dm3 / d2;
        dm7 = dm5 / d2;
      }
      const R d3 = d2 * d;
      const R d5 = d3 * d2;
      const R expon = rexp<R>(-alpha2 * d2) * rpiroot;
      const R F0 = errfc_t<R>(alpha * d);
      const R F1 = F0 + R(2) * alpha * d * expon;
      const R F2 = F1 + (sizeof(R) == sizeof(float) ? c43a3 * d3 * expon
                                                    : R(4) * alpha3 * d3 * expon / R(3));
      const R F1_dm3 = F1 * dm3;
      const R F2_dm5 = R(3) * F2 * dm5;
      if constexpr (TAB) {
        if (tin) {
          // (A2, A1), A2 = 3 F2 / d^5, A1 = F1 / d^3 (undamped)
          const size_t e = (size_t)i0 * a.maxnb + tpos;
          a.ddtab[e] = make_float2(F2_dm5, F1_dm3);
          const int js = a.rank ? a.rank[j] : j;
          if (a.dd16)
            reinterpret_cast<unsigned short*>(a.ddtj)[e] = (unsigned short)js;
          else
            a.ddtj[e] = js;
        }
      }
      R F1d_i = 0, F2d_i = 0, F1d_j = 0, F2d_j = 0;
      if (ldamp) {
        R fdmp, fdd;
        if constexpr (F32)
          tt_damp_t<R>(a.tt, R(dmij.y), R(dmij.z), d, (int)dmij.w, fdmp, fdd);
        else
          tt_damp_t<R>(a.tt, R(a.T->dmp_b[o][pij]), R(a.T->dmp_c[o][pij]), d,
                       a.T->dmp_k[o][pij], fdmp, fdd);
        F1d_i = -fdmp + R(1);
        F2d_i = F1d_i + (d / R(3)) * fdd;
        if (FORCE && qd) {
          if constexpr (F32) {
            const float4 dmji = sdmp[(1 - o) * 256 + pji];
            tt_damp_t<R>(a.tt, R(dmji.y), R(dmji.z), d, (int)dmji.w, fdmp, fdd);
          } else {
            tt_damp_t<R>(a.tt, R(a.T->dmp_b[1 - o][pji]), R(a.T->dmp_c[1 - o][pji]), d,
                         a.T->dmp_k[1 - o][pji], fdmp, fdd);
          }
          F1d_j = -fdmp + R(1);
          F2d_j = F1d_j + (d / R(3)) * fdd;
        }
      }
      const R rmi = rx * mi0 + ry * mi1 + rz * mi2;
      const R rmj = rx * mj0 + ry * mj1 + rz * mj2;
      if (QQ && EFIELD) {
        const R w = qj * (F1 - F1d_i) * dm3;
        e0 += w * rx;
        e1 += w * ry;
        e2 += w * rz;
      }
      if (dd && EFIELD) {
        // T2 mu_j = r (r.mu_j) 3F2/d^5 - mu_j F1/d^3
        e0 += rx * rmj * F2_dm5 - mj0 * F1_dm3;
        e1 += ry * rmj * F2_dm5 - mj1 * F1_dm3;
        e2 += rz * rmj * F2_dm5 - mj2 * F1_dm3;
      }
      if (EFG || FORCE) {
        const R F3 = F2 + (sizeof(R) == sizeof(float) ? c815a5 * d5 * expon
                                                      : R(8) * alpha5 * d5 * expon / R(15));
        const R F3_dm7 = F3 * dm7 * R(15);
        if (EFG) {
          if (QQ) {
            const R w = -qj * F2_dm5;
            g00 += w * rx * rx + qj * F1_dm3;
            g11 += w * ry * ry + qj * F1_dm3;
            g22 += w * rz * rz + qj * F1_dm3;
            g01 += w * rx * ry;
            g02 += w * rx * rz;
            g12 += w * ry * rz;
          }
          if (dd) {
            // (T3 mu_j)_ab = -r_a r_b (r.mu_j) 15F3/d^7
            //               + (delta_ab r.mu_j + r_a mu_jb + r_b mu_ja) 3F2/d^5
            const R w3 = -rmj * F3_dm7;
            g00 += rx * rx * w3 + (rmj + R(2) * rx * mj0) * F2_dm5;
            g11 += ry * ry * w3 + (rmj + R(2) * ry * mj1) * F2_dm5;
            g22 += rz * rz * w3 + (rmj + R(2) * rz * mj2) * F2_dm5;
            g01 += rx * ry * w3 + (rx * mj1 + ry * mj0) * F2_dm5;
            g02 += rx * rz * w3 + (rx * mj2 + rz * mj0) * F2_dm5;
            g12 += ry * rz * w3 + (ry * mj2 + rz * mj1) * F2_dm5;
          }
        }
        if (FORCE) {
          R px = 0, py = 0, pz = 0, u = 0;
          if (QQ) {
            const R w = qi * qj * F1_dm3;
            px += w * rx;
            py += w * ry;
            pz += w * rz;
            u += R(0.5) * qi * qj * dm1 * F0;
          }
          if (dd) {
            const R mm = mi0 * mj0 + mi1 * mj1 + mi2 * mj2;
            const R w3 = -rmi * rmj * F3_dm7;
            px += rx * w3 + (mi0 * rmj + rmi * mj0 + rx * mm) * F2_dm5;
            py += ry * w3 + (mi1 * rmj + rmi * mj1 + ry * mm) * F2_dm5;
            pz += rz * w3 + (mi2 * rmj + rmi * mj2 + rz * mm) * F2_dm5;
            u -= R(0.5) * (rmi * rmj * F2_dm5 - mm * F1_dm3);
          }
          if (qd) {
            // q_i (T2 - T2d_j) mu_j - q_j (T2 - T2d_i) mu_i
            const R a2j = qi * (F2_dm5 - R(3) * F2d_j * dm5);
            const R a1j = qi * (F1_dm3 - F1d_j * dm3);
            const R a2i = qj * (F2_dm5 - R(3) * F2d_i * dm5);
            const R a1i = qj * (F1_dm3 - F1d_i * dm3);
            px += rx * rmj * a2j - mj0 * a1j - (rx * rmi * a2i - mi0 * a1i);
            py += ry * rmj * a2j - mj1 * a1j - (ry * rmi * a2i - mi1 * a1i);
            pz += rz * rmj * a2j - mj2 * a1j - (rz * rmi * a2i - mi2 * a1i);
            // q_j (T1 - T1d_i).mu_i, T1 = -r F1/d^3
            u += -qj * rmi * (F1 - F1d_i) * dm3;
          }
          f0 += px;
          f1 += py;
          f2 += pz;
          acc[0] += u;
          acc[1] += rx * px;
          acc[2] += ry * py;
          acc[3] += rz * pz;
          acc[4] += R(0.5) * (rx * py + ry * px);
          acc[5] += R(0.5) * (rx * pz + rz * px);
          acc[6] += R(0.5) * (ry * pz + rz * py);
        }
      }
    }
  }
  for (int off = LPA / 2; off > 0; off >>= 1) {
    if (EFIELD) {
      e0 += __shfl_xor(e0, off, 64);
      e1 += __shfl_xor(e1, off, 64);
      e2 += __shfl_xor(e2, off, 64);
    }
    if (EFG) {
      g00 += __shfl_xor(g00, off, 64);
      g11 += __shfl_xor(g11, off, 64);
      g22 += __shfl_xor(g22, off, 64);
      g01 += __shfl_xor(g01, off, 64);
      g02 += __shfl_xor(g02, off, 64);
      g12 += __shfl_xor(g12, off, 64);
    }
    if (FORCE) {
      f0 += __shfl_xor(f0, off, 64);
      f1 += __shfl_xor(f1, off, 64);
      f2 += __shfl_xor(f2, off, 64);
    }
  }
  if constexpr (TAB)
    if (a.ddtab && i < n && sub == 0) {  // lane 0 visits every slot
      a.ddcnt[i0] = tbase;
      if (tbase == 0) {  // a row without pairs holds one zero entry (k_dd_tab2 reads it)
        a.ddtab[(size_t)i0 * a.maxnb] = make_float2(0.f, 0.f);
        if (a.dd16)
          reinterpret_cast<unsigned short*>(a.ddtj)[(size_t)i0 * a.maxnb] = 0;
        else
          a.ddtj[(size_t)i0 * a.maxnb] = 0;
      }
    }
  if (i < n && sub == 0) {
    if (EFIELD) {
      a.ef[i] = e0;
      a.ef[i + n] = e1;
      a.ef[i + 2 * n] = e2;
    }
    if (EFG) {
      a.efg[i] = g00;
      a.efg[i + n] = g01;
      a.efg[i + 2 * n] = g02;
      a.efg[i + 3 * n] = g01;
      a.efg[i + 4 * n] = g11;
      a.efg[i + 5 * n] = g12;
      a.efg[i + 6 * n] = g02;
      a.efg[i + 7 * n] = g12;
      a.efg[i + 8 * n] = g22;
    }
    if (FORCE) {
      a.f[i] = f0;
      a.f[i + n] = f1;
      a.f[i + 2 * n] = f2;
    }
  }
  if (FORCE) {
    double dacc[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) dacc[k] = (double)acc[k];
    block_sum<7>(dacc, red);
    if (threadIdx.x == 0)
      for (int k = 0; k < 7; ++k) a.part[(size_t)k * gridDim.x + blockIdx.x] = dacc[k];
  }
}

template __global__ void k_es_dir<ES_STATIC, double, false>(EsDirArgs);
template __global__ void k_es_dir<ES_IND, double, false>(EsDirArgs);
template __global__ void k_es_dir<ES_FINAL, double, false>(EsDirArgs);
template __global__ void k_es_dir<ES_STATIC, float, false>(EsDirArgs);
template __global__ void k_es_dir<ES_FINAL, float, false>(EsDirArgs);
template __global__ void k_es_dir<ES_STATIC, double, true>(EsDirArgs);
template __global__ void k_es_dir<ES_FINAL, double, true>(EsDirArgs);
template __global__ void k_es_dir<ES_STATIC, float, true>(EsDirArgs);
template __global__ void k_es_dir<ES_FINAL, float, true>(EsDirArgs);

// ---------------------------------------------------------------------------
// Dipole field tensor of the SCF (multipole_ES_dir with task d-d and no
// damping, :2407-2421), FP64 path.  The pair geometry is fixed during the
// SCF, so k_pairs_t2 stores for every directed pair inside the cutoff, in the
// row layout of nbr, the four-vector (r, 3F2/d^5) and then, in a plane of
// its own, F1/d^3; pairs outside the cutoff get zero factors.  k_dd_field
// applies E_i = sum_j [r (r.mu_j) 3F2/d^5 - mu_j F1/d^3].  With polmask set,
// both run over the atoms of its species only (the field is used at the
// polarizable atoms only; elsewhere it is not written).
template <bool BATCH>
__global__ __launch_bounds__(MDFF_BLOCK) void k_pairs_t2(EsDirArgs a, double* __restrict__ tab) {
  const int n = a.n;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int i0 = tid / LPA, sub = tid % LPA;
  if (i0 >= n) return;
  const int i = a.perm ? a.perm[i0] : i0;  // atoms in spatial order
  const int cf = BATCH ? a.cfg[i] : 0;
  const Cell& cell = BATCH ? a.bc[cf].cell : a.cell;
  const double cut2 = BATCH ? a.bc[cf].cut2 : a.cut2;
  const double alpha = BATCH ? a.bc[cf].alpha : a.alpha;
  const double alpha2 = alpha * alpha, alpha3 = alpha2 * alpha;
  const double piroot = sqrt(MDFF_PI);
  const double si1 = a.s[i], si2 = a.s[i + n], si3 = a.s[i + 2 * n];
  // rows of the species in polmask only (all with 0): k_dd_field skips
  // the others
  if (a.polmask && !((a.polmask >> a.typ[i]) & 1u)) return;
  const int c = a.cnt[i];
  const size_t plane = (size_t)a.maxnb * n;
  for (int k = sub; k < c; k += LPA) {
    const size_t e = (size_t)i * a.maxnb + k;
    const int j = a.nbr[e];
    double rx, ry, rz, A2 = 0.0, A1 = 0.0;
    const double d2 =
        min_image(cell, si1 - a.s4[j].x, si2 - a.s4[j].y, si3 - a.s4[j].z, rx, ry, rz);
    if (!(d2 > cut2)) {
      const double d = sqrt(d2);
      const double d3 = d2 * d;
      const double dm1 = 1.0 / d;
      const double dm3 = dm1 / d2;
      const double dm5 = dm3 / d2;
      const double expon = exp(-alpha2 * d2) / piroot;
      const double F0 = errfc_nr(alpha * d);
      const double F1 = F0 + 2.0 * alpha * d * expon;
      const double F2 = F1 + 4.0 * alpha3 * d3 * expon / 3.0;
      A2 = 3.0 * F2 * dm5;
      A1 = F1 * dm3;
    }
    reinterpret_cast<double4*>(tab)[e] = make_double4(rx, ry, rz, A2);
    tab[4 * plane + e] = A1;
  }
}

// Dipoles of the SCF iteration in FP32, stored at the atom's spatial rank:
// a row of the compacted pair table lists its neighbours by rank, so the
// lanes of one gather read neighbouring 16-byte slots.
__global__ __launch_bounds__(MDFF_BLOCK) void k_pack_mu_ranked(int n, const double* __restrict__ mu,
                                                               const int* __restrict__ rank,
                                                               float4* __restrict__ mu4) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  mu4[rank[i]] = make_float4((float)mu[i], (float)mu[i + n], (float)mu[i + 2 * n], 0.f);
}

// Charges and dipoles packed as (q, mu_x, mu_y, mu_z) per atom for the
// static and final real-space passes: a neighbour's moments are one gather.
template <typename R4>
__global__ __launch_bounds__(MDFF_BLOCK) void k_pack_qmu(int n, const double* __restrict__ q,
                                                         const double* __restrict__ mu,
                                                         R4* __restrict__ qm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  R4 v;
  v.x = q[i];
  v.y = mu[i];
  v.z = mu[i + n];
  v.w = mu[i + 2 * n];
  qm[i] = v;
}
template __global__ void k_pack_qmu<float4>(int, const double*, const double*, float4*);
template __global__ void k_pack_qmu<double4>(int, const double*, const double*, double4*);

// Dipoles of the SCF iteration packed as (x, y, z, 0) per atom, so that a
// neighbour's dipole is one gather.
template <typename R4>
__global__ __launch_bounds__(MDFF_BLOCK) void k_pack_mu(int n, const double* __restrict__ mu,
                                                        R4* __restrict__ mu4) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  R4 v;
  v.x = mu[i];
  v.y = mu[i + n];
  v.z = mu[i + 2 * n];
  v.w = 0;
  mu4[i] = v;
}

// DD_LANES lanes per atom (a whole wave), pairs taken 2 at a time per lane
// (loads in flight together), accumulated in list order per lane.  A
// neighbour without a dipole contributes exact zeros.
__global__ __launch_bounds__(MDFF_BLOCK) void k_dd_field(EsDirArgs a,
                                                         const double* __restrict__ tab,
                                                         const double4* __restrict__ mu4) {
  using R = double;
  using R4 = double4;
  const int n = a.n;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int i0 = tid / DD_LANES, sub = tid % DD_LANES;
  const int i = (a.perm && i0 < n) ? a.perm[i0] : i0;  // atoms in spatial order
  R e0 = 0, e1 = 0, e2 = 0;
  const size_t plane = (size_t)a.maxnb * n;
  const R4* t4 = reinterpret_cast<const R4*>(tab);
  const R* t1 = tab + 4 * plane;
  const bool act = i < n && (!a.polmask || ((a.polmask >> a.typ[i]) & 1u));
  const int c = act ? a.cnt[i] : 0;
  const size_t row = (size_t)(act ? i : 0) * a.maxnb;
  constexpr int U = 2;
  for (int k0 = sub; k0 < c; k0 += U * DD_LANES) {
    int jj[U];
    R4 g[U], m[U];
    R A1[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + u * DD_LANES;
      const size_t e = row + (k < c ? k : 0);
      jj[u] = a.nbr[e];
      g[u] = t4[e];
      A1[u] = t1[e];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) m[u] = mu4[jj[u]];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (k0 + u * DD_LANES >= c) break;
      const R rmj = g[u].x * m[u].x + g[u].y * m[u].y + g[u].z * m[u].z;
      e0 += g[u].x * rmj * g[u].w - m[u].x * A1[u];
      e1 += g[u].y * rmj * g[u].w - m[u].y * A1[u];
      e2 += g[u].z * rmj * g[u].w - m[u].z * A1[u];
    }
  }
  for (int off = DD_LANES / 2; off > 0; off >>= 1) {
    e0 += __shfl_xor(e0, off, 64);
    e1 += __shfl_xor(e1, off, 64);
    e2 += __shfl_xor(e2, off, 64);
  }
  if (act && sub == 0) {
    a.ef[i] = e0;
    a.ef[i + n] = e1;
    a.ef[i + 2 * n] = e2;
  }
}

// ---------------------------------------------------------------------------
// SCF dipole field of the mixed path, from a compacted FP32 pair table.  Once
// per evaluation k_dd_table (or the static FP32 real-space pass) keeps, per
// atom and in list order, only the list entries inside the cutoff, with
// A2 = 3F2/d^5, A1 = F1/d^3 (the undamped d-d terms of
// src/field.f90:2407-2421, formed in FP32 from an FP64 fractional
// difference) and the neighbour's slot in the dipole array (its spatial
// rank when a.rank is given, else its index).  Rows and counts are in the
// processing order i0 (atom perm[i0] when perm is given).  Each SCF
// iteration then applies E_i = sum_j [A2 r (r.mu_j) - A1 mu_j] (k_dd_tab2),
// forming r from 32-bit fixed-point positions: 10 bytes of table and one
// 32-byte gather per pair.
template <bool BATCH>
__global__ __launch_bounds__(MDFF_BLOCK) void k_dd_table(EsDirArgs a, float2* __restrict__ tab,
                                                         int* __restrict__ tj,
                                                         int* __restrict__ tcnt) {
  const int n = a.n;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int i0 = tid >> 6, lane = tid & 63;
  if (i0 >= n) return;  // whole waves
  const int i = a.perm ? a.perm[i0] : i0;  // atoms in spatial order
  const int cf = BATCH ? a.cfg[i] : 0;
  const Cell& cell = BATCH ? a.bc[cf].cell : a.cell;
  const float cut2 = (float)(BATCH ? a.bc[cf].cut2 : a.cut2);
  const double alphad = BATCH ? a.bc[cf].alpha : a.alpha;
  const float alpha = (float)alphad, alpha2 = alpha * alpha;
  const float c43a3 = (float)(4.0 / 3.0 * alphad * alphad * alphad);
  const double4 si = a.s4[i];
  // only pairs of two species in polmask (all with 0)
  const unsigned pm = a.polmask ? a.polmask : ~0u;
  const bool pol_i = (pm >> (BATCH ? a.typ[i] : s4_type(si))) & 1u;
  const int c = pol_i ? a.cnt[i] : 0;
  const size_t lrow = (size_t)i * a.maxnb;  // the atom's list row
  const size_t row = (size_t)i0 * a.maxnb;  // its table row: processing order
  const unsigned long long below = (1ull << lane) - 1ull;
  int base = 0;
  for (int k0 = 0; k0 < c; k0 += 64) {
    const int k = k0 + lane;
    float2 v = make_float2(0.f, 0.f);
    int j = 0;
    bool in = false;
    if (k < c) {
      j = a.nbr[lrow + k];
      const double4 sj = a.s4[j];
      float rx, ry, rz;
      const float d2 = min_image_t<float>(cell, si.x - sj.x, si.y - sj.y, si.z - sj.z, rx, ry, rz);
      in = !(d2 > cut2) && ((pm >> (BATCH ? a.typ[j] : s4_type(sj))) & 1u);
      if (in) {
        const float d = sqrtf(d2);
        const float dm1 = 1.f / d;
        const float dm3 = dm1 * dm1 * dm1;
        const float expon = __expf(-alpha2 * d2) * 0.56418958354775628f;  // 1/sqrt(pi)
        const float F1 = errfc_t<float>(alpha * d) + 2.f * alpha * d * expon;
        const float F2 = F1 + c43a3 * d2 * d * expon;
        v = make_float2(3.f * F2 * dm3 * dm1 * dm1, F1 * dm3);
      }
    }
    const unsigned long long m = __ballot(in);
    if (in) {
      const size_t e = row + base + __popcll(m & below);
      tab[e] = v;
      const int js = a.rank ? a.rank[j] : j;
      if (a.dd16)
        reinterpret_cast<unsigned short*>(tj)[e] = (unsigned short)js;
      else
        tj[e] = js;
    }
    base += __popcll(m);
  }
  if (lane == 0) {
    tcnt[i0] = base;
    if (base == 0) {  // a row without pairs holds one zero entry (k_dd_tab2 reads it)
      tab[row] = make_float2(0.f, 0.f);
      if (a.dd16)
        reinterpret_cast<unsigned short*>(tj)[row] = 0;
      else
        tj[row] = 0;
    }
  }
}

// Fractional coordinates as 32-bit fixed point in [0, 1), rank ordered:
// the difference of two, read as a signed integer, is the minimum-image
// fractional difference s_i - s_j - nint(s_i - s_j) to 2^-32.
__global__ __launch_bounds__(MDFF_BLOCK) void k_pack_u4(int n, const double* __restrict__ s,
                                                        const int* __restrict__ rank,
                                                        uint4* __restrict__ u4) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  unsigned int u[3];
  for (int a = 0; a < 3; ++a) {
    const double x = s[i + (size_t)a * n];
    const double f = x - floor(x);  // [0, 1]
    u[a] = (unsigned int)(unsigned long long)(f * 4294967296.0);  // 1 wraps to 0
  }
  u4[rank ? rank[i] : i] = make_uint4(u[0], u[1], u[2], 0u);
}

// The SCF dipole field from the (A2, A1) table of k_dd_table (or the static
// FP32 pass): r = A (u_i - u_j) 2^-32 from the fixed-point positions,
// E_i = sum_j [A2 r (r.mu_j) - A1 mu_j], with A the cell of the row's
// configuration (bc, cfg: a batch) or cell; a neighbour's position and
// dipole are two 16-byte gathers from the same slot.  One wave per atom,
// persistent: the waves of the grid take the visits v = wave, wave + waves,
// ... of the nrows rows listed in rows (all n rows in order when rows is
// null; the rows of atoms outside the table's species are empty and need
// not be listed, their field is not written); while a wave gathers and sums
// one row, the loads of its next row are already in flight, so that the
// table stream does not pause.  The first 64 U entries of a row are
// pipelined (U per lane, chosen for the expected row length); longer rows
// finish in a plain loop.  The pipelined loads are issued unconditionally: indexes are
// clamped to the row's last entry (a row without pairs holds one zero
// entry) and to the last row, and the entries past the count are masked by
// a zero factor, so that no branch splits the loads in flight.
template <typename IT, int U>
__global__ __launch_bounds__(MDFF_BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_dd_tab2(
    int n, int maxnb, const float2* __restrict__ tab, const IT* __restrict__ tj,
    const int* __restrict__ tcnt, const float4* __restrict__ mu4, const uint4* __restrict__ u4,
    Cell cell, const BatchCfg* __restrict__ bc, const int* __restrict__ cfg,
    const int* __restrict__ perm, const int* __restrict__ rows, int nrows,
    double* __restrict__ ef) {
  constexpr float Q = 2.3283064365386963e-10f;  // 2^-32
  float Am[9];
  auto set_cell = [&](const Cell& cl) {
#pragma unroll
    for (int k = 0; k < 9; ++k) Am[k] = (float)cl.A[k] * Q;
  };
  set_cell(cell);
  const int lane = threadIdx.x & 63;
  const int wid =
      __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  const int nw = (int)((gridDim.x * blockDim.x) >> 6);
  // the row of visit v (clamped to the last visit: loads past it are
  // issued and never used)
  auto rowof = [&](int v) { return rows ? rows[min(v, nrows - 1)] : min(v, n - 1); };
  auto load = [&](int r, int c, float2 (&g)[U], int (&jj)[U]) {
    const size_t row = (size_t)r * maxnb;
    const float2* trow = tab + row;
    const IT* jrow = tj + row;
    const int last = max(c - 1, 0);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = lane + u * 64;
      const int kc = min(k, last);
      const float2 v = trow[kc];
      jj[u] = jrow[kc];
      const float mk = k < c ? 1.f : 0.f;
      g[u] = make_float2(v.x * mk, v.y * mk);
    }
  };
  // the field of one pair: r from the fixed-point difference
  auto pair = [&](const uint4 ui, const uint4 uj, const float2 t, const float4 m, float& e0,
                  float& e1, float& e2) {
    const float dx = (float)(int)(ui.x - uj.x), dy = (float)(int)(ui.y - uj.y),
                dz = (float)(int)(ui.z - uj.z);
    const float rx = dx * Am[0] + dy * Am[3] + dz * Am[6];
    const float ry = dx * Am[1] + dy * Am[4] + dz * Am[7];
    const float rz = dx * Am[2] + dy * Am[5] + dz * Am[8];
    const float rm = (rx * m.x + ry * m.y + rz * m.z) * t.x;
    e0 += rx * rm - t.y * m.x;
    e1 += ry * rm - t.y * m.y;
    e2 += rz * rm - t.y * m.z;
  };
  float2 g[U];
  int jj[U];
  int r = rowof(wid), rn = rowof(wid + nw);
  int c = tcnt[r];
  int cn = tcnt[rn];
  load(r, c, g, jj);
  for (int v = wid; v < nrows; v += nw) {
    const int iout = perm ? perm[r] : r;
    if (bc) set_cell(bc[cfg[iout]].cell);
    const uint4 ui = u4[r];
    float4 m[U];
    uint4 q[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      m[u] = mu4[jj[u]];
      q[u] = u4[jj[u]];
    }
    // the next row and the count of the one after it (rows past the last
    // are loaded clamped and never used)
    float2 gn[U];
    int jn[U];
    load(rn, cn, gn, jn);
    const int rnn = rowof(v + 2 * nw);
    const int cnn = tcnt[rnn];
    float e0 = 0.f, e1 = 0.f, e2 = 0.f;
#pragma unroll
    for (int u = 0; u < U; ++u) pair(ui, q[u], g[u], m[u], e0, e1, e2);
    // entries past the pipelined 512
    const size_t row = (size_t)r * maxnb;
    for (int k = U * 64 + lane; k < c; k += 64) {
      const int j = tj[row + k];
      pair(ui, u4[j], tab[row + k], mu4[j], e0, e1, e2);
    }
    for (int off = 32; off > 0; off >>= 1) {
      e0 += __shfl_xor(e0, off, 64);
      e1 += __shfl_xor(e1, off, 64);
      e2 += __shfl_xor(e2, off, 64);
    }
    if (lane == 0) {
      ef[iout] = e0;
      ef[iout + n] = e1;
      ef[iout + 2 * n] = e2;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      g[u] = gn[u];
      jj[u] = jn[u];
    }
    r = rn;
    rn = rnn;
    c = cn;
    cn = cnn;
  }
}

#define MDFF_DD_TAB2(IT, U)                                                                  \
  template __global__ void k_dd_tab2<IT, U>(int, int, const float2*, const IT*, const int*,        \
                                            const float4*, const uint4*, Cell, const BatchCfg*,    \
                                            const int*, const int*, const int*, int, double*);
MDFF_DD_TAB2(int, 4)
MDFF_DD_TAB2(int, 6)
MDFF_DD_TAB2(int, 8)
MDFF_DD_TAB2(unsigned short, 4)
MDFF_DD_TAB2(unsigned short, 6)
MDFF_DD_TAB2(unsigned short, 8)
#undef MDFF_DD_TAB2
template __global__ void k_pairs_t2<false>(EsDirArgs, double*);
template __global__ void k_pairs_t2<true>(EsDirArgs, double*);
template __global__ void k_dd_table<false>(EsDirArgs, float2*, int*, int*);
template __global__ void k_dd_table<true>(EsDirArgs, float2*, int*, int*);
template __global__ void k_pack_mu<double4>(int, const double*, double4*);
template __global__ void k_pack_mu<float4>(int, const double*, float4*);

}  // namespace mdff
