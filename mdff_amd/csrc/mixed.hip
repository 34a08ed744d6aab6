// mixed.hip -- reciprocal-space part of the mixed-precision path on the
// matrix cores (v_mfma_f32_16x16x32_f16, gfx950).
//
// With separable phases e^{i k.r_j} = E1_j[n1] E2_j[n2] E3_j[n3]
// (E_a[n] = e^{2 pi i n s_a}), the weighted structure factors of
// multipole_ES_rec (src/field.f90:2697-2742) are a GEMM
//   M_w[n1][(n2,n3)] = sum_j (w_j E1_j[n1]) (E2_j[n2] E3_j[n3])
// with rows (weight, n1), columns (n2, n3) and the atoms as the reduction
// dimension, and the field / EFG apply (:2746-2801) is the transposed
// contraction
//   X_w[i][n1] = sum_{(n2,n3)} conj(E2_i[n2] E3_i[n3]) (w g)[n1][(n2,n3)]
// with the atoms as rows.  Complex products are real GEMMs over interleaved
// (re, im) K; every operand is split into fp16 hi + lo parts and three
// products (hi.hi + hi.lo + lo.hi) are accumulated in FP32 (22-bit
// operands).  Structure factors are accumulated in FP32 over 32-atom
// stages and compensated in FP64 across stages and atom splits; energy,
// stress and g = A_k rho are formed in FP64.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "mixed.h"

namespace mdff {
namespace mixed {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));  // 16 raw bytes

constexpr int RHO_ATOMS_PER_SPLIT = 2048;
constexpr float DIP_SCALE = 256.0f;     // dipole weights are scaled by 2^8
// The apply operand w g of each n1 plane is normalised to |.| <= APPLY_SCALE
// (a power of two): most elements of a plane are far below its maximum, and
// this keeps their fp16 lo parts above the subnormal range (fp16 normals
// start at 2^-14; |lo| <= 2^-11 |x|).  2^12 leaves the largest hi part at
// 4096, well inside the fp16 range.
constexpr float APPLY_SCALE = 4096.0f;

struct MixedState {
  mdffgpu_ctx* ctx;
  int N1 = 0, N2 = 0, N3 = 0, C23 = 0, C23p = 0, npad = 0;
  int kx = 0, ky = 0, kz = 0;
  int nsplit = 0;
  // phase tables, float2 (re, im)
  float2* tj1 = nullptr;  // [n1][j]   (N1 x npad)
  float2* tj2 = nullptr;  // [n2+ky][j]
  float2* tj3 = nullptr;  // [n3+kz][j]
  double* rpart = nullptr;      // [split][row][col]  FP64 partials
  size_t rpart_cap = 0;
  double2* g = nullptr;         // [n1][c]
  unsigned int* gmax = nullptr; // max |g| per n1 plane, as float bits
  _Float16* bfrag = nullptr;    // apply B fragments, hi then lo
  size_t bfrag_cap = 0;
  float* X = nullptr;           // [slice][i][col]
  float* Wt = nullptr;          // [w][j] structure-factor weights
  u32x4* afrag = nullptr;       // structure-factor A fragments of one call
  size_t X_cap = 0;
};

// ---------------------------------------------------------------------------
// phase tables from the FP64 fractional coordinates
__global__ __launch_bounds__(MDFF_BLOCK) void mx_tables(int n, int npad, const double* __restrict__ s,
                                                         int kx, int ky, int kz, float2* tj1,
                                                         float2* tj2, float2* tj3) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= npad) return;
  const bool act = j < n;
  const double s1 = act ? s[j] : 0.0, s2 = act ? s[j + n] : 0.0, s3 = act ? s[j + 2 * n] : 0.0;
  for (int m = 0; m <= kx; ++m) {
    double sn, cs;
    sincospi(2.0 * m * s1, &sn, &cs);
    tj1[(size_t)m * npad + j] = act ? make_float2((float)cs, (float)sn) : make_float2(0.f, 0.f);
  }
  for (int m = -ky; m <= ky; ++m) {
    double sn, cs;
    sincospi(2.0 * m * s2, &sn, &cs);
    tj2[(size_t)(m + ky) * npad + j] = act ? make_float2((float)cs, (float)sn) : make_float2(0.f, 0.f);
  }
  for (int m = -kz; m <= kz; ++m) {
    double sn, cs;
    sincospi(2.0 * m * s3, &sn, &cs);
    tj3[(size_t)(m + kz) * npad + j] = act ? make_float2((float)cs, (float)sn) : make_float2(0.f, 0.f);
  }
}

// hi + lo = x to 22 bits.  x is first pinned as a rounded FP32 value:
// otherwise the compiler may fuse the producing multiply into both
// conversions (v_fma_mix*_f16 from the exact product for lo, v_cvt_pk_f16_f32
// of the FP32-rounded product for hi), and at an fp16 rounding tie the two
// roundings of "hi" differ by one ulp, so hi + lo misses x by an ulp.
__device__ __forceinline__ void split16(float x, _Float16& hi, _Float16& lo) {
  __asm__ volatile("" : "+v"(x));
  hi = (_Float16)x;
  lo = (_Float16)(x - (float)hi);
}

// ---------------------------------------------------------------------------
// Weights of the structure-factor GEMM, FP32, [w][j] (zero past n):
// w0 = q (when CHG), then 2^8 * 2 pi B^T mu (when DIP).  Also resets the
// per-n1 maxima of |g| that mx_rho_reduce accumulates.
__global__ __launch_bounds__(MDFF_BLOCK) void mx_weights(int n, int npad, int chg, int dip,
                                                          const double* __restrict__ q,
                                                          const double* __restrict__ mu, Cell cell,
                                                          float* __restrict__ Wt,
                                                          unsigned int* __restrict__ gmax,
                                                          int nreset) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.x == 0 && threadIdx.x < nreset) gmax[threadIdx.x] = 0u;
  if (j >= npad) return;
  int w = 0;
  if (chg) Wt[(size_t)(w++) * npad + j] = j < n ? (float)q[j] : 0.f;
  if (dip) {
    const double* B = cell.B;
    for (int a = 0; a < 3; ++a) {
      float v = 0.f;
      if (j < n) {
        const double m = B[3 * a] * mu[j] + B[3 * a + 1] * mu[j + n] + B[3 * a + 2] * mu[j + 2 * n];
        v = (float)(m * 6.283185307179586) * DIP_SCALE;
      }
      Wt[(size_t)(w++) * npad + j] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// A fragments of the structure-factor GEMM for the whole call, in the
// fragment order the GEMM stages through LDS: per K step g (16 atoms), row
// tile rt, hi|lo, lane l: atoms 16 g + 4 (l / 16) + (0..3), row
// r = rt*16 + l%16 = w*N1 + n1, values (Re, Im) of W_w[j] E1_j[n1] split
// into fp16 hi + lo.  Rows past NW*N1 are zero.
template <int NRT>
__global__ __launch_bounds__(MDFF_BLOCK) void mx_rho_afrag(int npad, int N1, int NW,
                                                            const float* __restrict__ Wt,
                                                            const float2* __restrict__ tj1,
                                                            u32x4* __restrict__ afrag) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)(npad / 16) * NRT * 64;
  if (t >= total) return;
  const int l = (int)(t & 63);
  const int rt = (int)((t >> 6) % NRT);
  const int g = (int)((t >> 6) / NRT);
  const int r = rt * 16 + (l & 15);
  const int j = g * 16 + 4 * (l >> 4);
  half8 h, lo;
  if (r < NW * N1) {
    const int w = r / N1, n1 = r % N1;
    const float4 wv = *(const float4*)(Wt + (size_t)w * npad + j);
    const float4 e01 = *(const float4*)(tj1 + (size_t)n1 * npad + j);
    const float4 e23 = *(const float4*)(tj1 + (size_t)n1 * npad + j + 2);
    const float ww[4] = {wv.x, wv.y, wv.z, wv.w};
    const float er[4] = {e01.x, e01.z, e23.x, e23.z};
    const float ei[4] = {e01.y, e01.w, e23.y, e23.w};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      _Float16 a, bb;
      split16(ww[e] * er[e], a, bb);
      h[2 * e] = a;
      lo[2 * e] = bb;
      split16(ww[e] * ei[e], a, bb);
      h[2 * e + 1] = a;
      lo[2 * e + 1] = bb;
    }
  } else {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      h[e] = (_Float16)0.f;
      lo[e] = (_Float16)0.f;
    }
  }
  u32x4* o = afrag + (((size_t)g * NRT + rt) * 2) * 64 + l;
  o[0] = __builtin_bit_cast(u32x4, h);
  o[64] = __builtin_bit_cast(u32x4, lo);
}

// ---------------------------------------------------------------------------
// Structure-factor GEMM.  Rows r = w*N1 + n1 (weight, n1), NRT tiles;
// columns: Re and Im planes of the (n2, n3) mesh, col = part*C23p + c;
// K = 2 x atoms (re, im interleaved).
//   A[r][2j] = Re(W_w[j] E1_j[n1]),  A[r][2j+1] = Im(W_w[j] E1_j[n1])
//   Re plane: B[2j][c] = Re P_j[c], B[2j+1][c] = -Im P_j[c]
//   Im plane: B[2j][c] = Im P_j[c], B[2j+1][c] =  Re P_j[c]     (P = E2 E3)
// The A fragments of a 32-atom stage are formed once per workgroup into LDS
// (fragment order, double buffered, one barrier per stage); each wave forms
// the B fragments of its CT column groups (16 c each, both planes from one
// set of products) in registers, the loads of the next K step in flight.
// Grid (column blocks of NWV*CT groups, atom splits).  Output:
// part[split][r][col] (rows NRT*16, 2*C23p columns).
constexpr int RHO_STAGE = 32;  // atoms per LDS stage (two K steps)

template <int NRT, int CT, int NWV>
__global__ __launch_bounds__(64 * NWV) void mx_rho_gemm(
    int npad, int N3, int C23, int C23p, const u32x4* __restrict__ afrag,
    const float2* __restrict__ tj2, const float2* __restrict__ tj3, int atoms_per_split,
    double* __restrict__ out) {
  __shared__ u32x4 sA[2][2][NRT][2][64];  // [buf][kstep][rt][hi|lo][lane], fp16 x8
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int kgrp = lane >> 4;
  int col[CT], n2i[CT], n3i[CT];
  bool colok[CT], grpok[CT];
#pragma unroll
  for (int t = 0; t < CT; ++t) {
    const int grp = (blockIdx.x * NWV + wave) * CT + t;
    grpok[t] = grp * 16 < C23p;
    col[t] = grp * 16 + (lane & 15);
    colok[t] = col[t] < C23;
    n2i[t] = colok[t] ? col[t] / N3 : 0;
    n3i[t] = colok[t] ? col[t] % N3 : 0;
  }
  // FP32 MFMA accumulators flushed into FP64 every 64 atoms: the structure
  // factors of small |k| are far smaller than the partial sums of a long
  // FP32 chain (charge neutrality), so the chain is kept short and the sum
  // over chains is compensated in FP64.
  floatx4 acc[CT][2][NRT];
  double accd[CT][2][NRT][4];
#pragma unroll
  for (int t = 0; t < CT; ++t)
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int r = 0; r < NRT; ++r) {
        acc[t][p][r] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int e = 0; e < 4; ++e) accd[t][p][r][e] = 0.0;
      }
  const int j0 = blockIdx.y * atoms_per_split;
  const int j1 = min(npad, j0 + atoms_per_split);

  // A fragments of a stage (32 atoms): a contiguous block of ABLK 16-byte
  // words of afrag, copied through registers into LDS
  constexpr int ABLK = 2 * NRT * 2 * 64;
  constexpr int NTH = 64 * NWV;
  constexpr int APER = (ABLK + NTH - 1) / NTH;
  u32x4 apre[APER];
#define MX_FETCH_A(jb)                                                                   \
  {                                                                                      \
    const u32x4* src_ = afrag + (size_t)((jb) / 16) * NRT * 2 * 64;                      \
    _Pragma("unroll") for (int u = 0; u < APER; ++u) {                                   \
      const int idx_ = threadIdx.x + u * NTH;                                            \
      if (ABLK % NTH == 0 || idx_ < ABLK) apre[u] = src_[idx_];                          \
    }                                                                                    \
  }
#define MX_STAGE_A(b)                                                                    \
  {                                                                                      \
    u32x4* dst_ = &sA[b][0][0][0][0];                                                    \
    _Pragma("unroll") for (int u = 0; u < APER; ++u) {                                   \
      const int idx_ = threadIdx.x + u * NTH;                                            \
      if (ABLK % NTH == 0 || idx_ < ABLK) dst_[idx_] = apre[u];                          \
    }                                                                                    \
  }
  // phase products of 4 atoms for column group t
  float4 p2a[CT], p2b[CT], p3a[CT], p3b[CT];
  auto load_p = [&](int j) {
#pragma unroll
    for (int t = 0; t < CT; ++t) {
      const float2* a2 = tj2 + (size_t)n2i[t] * npad + j;
      const float2* a3 = tj3 + (size_t)n3i[t] * npad + j;
      p2a[t] = *(const float4*)a2;
      p2b[t] = *(const float4*)(a2 + 2);
      p3a[t] = *(const float4*)a3;
      p3b[t] = *(const float4*)(a3 + 2);
    }
  };

  // B fragments of both planes from the loaded phases
  half8 bre_h[CT], bre_l[CT], bim_h[CT], bim_l[CT];
  auto make_b = [&]() {
#pragma unroll
    for (int t = 0; t < CT; ++t) {
      const float e2r[4] = {p2a[t].x, p2a[t].z, p2b[t].x, p2b[t].z};
      const float e2i[4] = {p2a[t].y, p2a[t].w, p2b[t].y, p2b[t].w};
      const float e3r[4] = {p3a[t].x, p3a[t].z, p3b[t].x, p3b[t].z};
      const float e3i[4] = {p3a[t].y, p3a[t].w, p3b[t].y, p3b[t].w};
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        float pr = e2r[e] * e3r[e] - e2i[e] * e3i[e];
        float pi = e2r[e] * e3i[e] + e2i[e] * e3r[e];
        if (!colok[t]) {
          pr = 0.f;
          pi = 0.f;
        }
        _Float16 rh, rl, ih, il;
        split16(pr, rh, rl);
        split16(pi, ih, il);
        bre_h[t][2 * e] = rh;
        bre_h[t][2 * e + 1] = -ih;
        bre_l[t][2 * e] = rl;
        bre_l[t][2 * e + 1] = -il;
        bim_h[t][2 * e] = ih;
        bim_h[t][2 * e + 1] = rh;
        bim_l[t][2 * e] = il;
        bim_l[t][2 * e + 1] = rl;
      }
    }
  };

  int buf = 0;
  if (j0 < j1) {
    MX_FETCH_A(j0)
    MX_STAGE_A(0)
    load_p(j0 + 4 * kgrp);
    make_b();
    if (j0 + 16 < j1) load_p(j0 + 16 + 4 * kgrp);
  }
  __syncthreads();
  int stage = 0;
  for (int jb = j0; jb < j1; jb += RHO_STAGE, ++stage) {
    const bool more = jb + RHO_STAGE < j1;
    if (more) MX_FETCH_A(jb + RHO_STAGE)
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      // row tiles two at a time: four independent accumulators (2 tiles x
      // Re/Im planes) between dependent MFMAs; the next pair's A fragments
      // are read from LDS ahead
      u32x4 a0h = sA[buf][ks][0][0][lane], a0l = sA[buf][ks][0][1][lane];
      u32x4 a1h = a0h, a1l = a0l;
      if (NRT > 1) {
        a1h = sA[buf][ks][1][0][lane];
        a1l = sA[buf][ks][1][1][lane];
      }
#pragma unroll
      for (int r0 = 0; r0 < NRT; r0 += 2) {
        u32x4 n0h = a0h, n0l = a0l, n1h = a1h, n1l = a1l;
        if (r0 + 2 < NRT) {
          n0h = sA[buf][ks][r0 + 2][0][lane];
          n0l = sA[buf][ks][r0 + 2][1][lane];
        }
        if (r0 + 3 < NRT) {
          n1h = sA[buf][ks][r0 + 3][0][lane];
          n1l = sA[buf][ks][r0 + 3][1][lane];
        }
        const bool two = r0 + 1 < NRT;
        const half8 h0 = __builtin_bit_cast(half8, a0h), l0 = __builtin_bit_cast(half8, a0l);
        const half8 h1 = __builtin_bit_cast(half8, a1h), l1 = __builtin_bit_cast(half8, a1l);
#pragma unroll
        for (int t = 0; t < CT; ++t) {
          if (!grpok[t]) continue;
          acc[t][0][r0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(h0, bre_h[t], acc[t][0][r0], 0, 0, 0);
          acc[t][1][r0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(h0, bim_h[t], acc[t][1][r0], 0, 0, 0);
          if (two) {
            acc[t][0][r0 + 1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(h1, bre_h[t], acc[t][0][r0 + 1], 0, 0, 0);
            acc[t][1][r0 + 1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(h1, bim_h[t], acc[t][1][r0 + 1], 0, 0, 0);
          }
          acc[t][0][r0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(h0, bre_l[t], acc[t][0][r0], 0, 0, 0);
          acc[t][1][r0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(h0, bim_l[t], acc[t][1][r0], 0, 0, 0);
          if (two) {
            acc[t][0][r0 + 1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(h1, bre_l[t], acc[t][0][r0 + 1], 0, 0, 0);
            acc[t][1][r0 + 1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(h1, bim_l[t], acc[t][1][r0 + 1], 0, 0, 0);
          }
          acc[t][0][r0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(l0, bre_h[t], acc[t][0][r0], 0, 0, 0);
          acc[t][1][r0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(l0, bim_h[t], acc[t][1][r0], 0, 0, 0);
          if (two) {
            acc[t][0][r0 + 1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(l1, bre_h[t], acc[t][0][r0 + 1], 0, 0, 0);
            acc[t][1][r0 + 1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(l1, bim_h[t], acc[t][1][r0 + 1], 0, 0, 0);
          }
        }
        a0h = n0h;
        a0l = n0l;
        a1h = n1h;
        a1l = n1l;
      }
      // the next K step's B fragments (phases already loaded), and the loads
      // of the one after
      const int jn = jb + (ks + 1) * 16;
      if (jn < j1) {
        make_b();
        if (jn + 16 < j1) load_p(jn + 16 + 4 * kgrp);
      }
    }
    // FP64 flush every second stage (64 atoms) and after the last
    if ((stage & 1) || !more) {
#pragma unroll
      for (int t = 0; t < CT; ++t)
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
          for (int r = 0; r < NRT; ++r) {
#pragma unroll
            for (int e = 0; e < 4; ++e) accd[t][p][r][e] += (double)acc[t][p][r][e];
            acc[t][p][r] = floatx4{0.f, 0.f, 0.f, 0.f};
          }
    }
    if (more) MX_STAGE_A(buf ^ 1)
    __syncthreads();
    buf ^= 1;
  }
  // D layout: acc[..][rt][e] = D[rt*16 + 4*kgrp + e][col]
  double* o = out + (size_t)blockIdx.y * (NRT * 16) * (2 * C23p);
#pragma unroll
  for (int t = 0; t < CT; ++t) {
    if (!grpok[t]) continue;
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
        for (int e = 0; e < 4; ++e)
          o[(size_t)(rt * 16 + 4 * kgrp + e) * (2 * C23p) + p * C23p + col[t]] = accd[t][p][rt][e];
  }
}
#undef MX_FETCH_A
#undef MX_STAGE_A

// ---------------------------------------------------------------------------
// Sum of the splits in FP64, rho(k) = M_q + i sum_c n_c M_c, g = A_k rho on
// the half-space mesh (zero elsewhere), energy and stress rows
// (src/field.f90:2742-2745, 2803-2815), and max |g|.
// Grid (column blocks, N1): one n1 plane per block row.
__global__ __launch_bounds__(MDFF_BLOCK) void mx_rho_reduce(
    int N1, int N2, int N3, int C23, int C23p, int NW, int chg, int dip, int nsplit, int NRT16,
    const double* __restrict__ part, Cell cell, double alpha, int do_energy, double2* __restrict__ g,
    unsigned int* __restrict__ gmax, double* __restrict__ pout) {
  __shared__ double red[4 * 7];
  __shared__ float wmax[MDFF_BLOCK / 64];
  const int n1 = blockIdx.y;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int idx = n1 * C23 + c;
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  float gabs = 0.f;
  if (c < C23) {
    const int ky = (N2 - 1) / 2, kz = (N3 - 1) / 2;
    const int n2 = c / N3 - ky, n3 = c % N3 - kz;
    double Mr[4] = {0, 0, 0, 0}, Mi[4] = {0, 0, 0, 0};
    const size_t sstr = (size_t)NRT16 * (2 * C23p);
    const double* p0 = part + (size_t)n1 * (2 * C23p) + c;
    for (int w = 0; w < NW; ++w) {
      const double* p = p0 + (size_t)w * N1 * (2 * C23p);
      double vr[4], vi[4];
      int sp = 0;
      for (; sp + 4 <= nsplit; sp += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          vr[u] = p[(sp + u) * sstr];
          vi[u] = p[(sp + u) * sstr + C23p];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          Mr[w] += vr[u];
          Mi[w] += vi[u];
        }
      }
      for (; sp < nsplit; ++sp) {
        Mr[w] += p[sp * sstr];
        Mi[w] += p[sp * sstr + C23p];
      }
    }
    double rR = 0.0, rI = 0.0;
    int w = 0;
    if (chg) {
      rR = Mr[0];
      rI = Mi[0];
      w = 1;
    }
    if (dip) {
      const double sc = 1.0 / DIP_SCALE;
      const double tR = (n1 * Mr[w] + n2 * Mr[w + 1] + n3 * Mr[w + 2]) * sc;
      const double tI = (n1 * Mi[w] + n2 * Mi[w + 1] + n3 * Mi[w + 2]) * sc;
      rR -= tI;  // i * (tR + i tI)
      rI += tR;
    }
    const bool half = (n1 > 0) || (n1 == 0 && n2 > 0) || (n1 == 0 && n2 == 0 && n3 > 0);
    double2 gv = make_double2(0.0, 0.0);
    if (half) {
      const double tpi = 2.0 * MDFF_PI;
      const double* B = cell.B;
      const double kx = tpi * (n1 * B[0] + n2 * B[3] + n3 * B[6]);
      const double ky2 = tpi * (n1 * B[1] + n2 * B[4] + n3 * B[7]);
      const double kz2 = tpi * (n1 * B[2] + n2 * B[5] + n3 * B[8]);
      const double kk = kx * kx + ky2 * ky2 + kz2 * kz2;
      const double a2 = alpha * alpha;
      const double Ak = exp(-kk * 0.25 / a2) / kk;
      gv = make_double2(Ak * rR, Ak * rI);
      gabs = (float)fmax(fabs(gv.x), fabs(gv.y));
      if (do_energy) {
        const double kc = 2.0 * (1.0 / kk + 1.0 / a2 / 4.0);
        const double str = (rR * rR + rI * rI) * Ak;
        acc[0] = str;
        acc[1] = (1.0 - kc * kx * kx) * str;
        acc[2] = (1.0 - kc * ky2 * ky2) * str;
        acc[3] = (1.0 - kc * kz2 * kz2) * str;
        acc[4] = -kc * kx * ky2 * str;
        acc[5] = -kc * kx * kz2 * str;
        acc[6] = -kc * ky2 * kz2 * str;
      }
    }
    g[idx] = gv;
  }
  // max |g| of the n1 plane: block maximum, one atomic per block
  for (int off = 32; off > 0; off >>= 1) gabs = fmaxf(gabs, __shfl_xor(gabs, off, 64));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = gabs;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = 0.f;
    for (int w = 0; w < MDFF_BLOCK / 64; ++w) m = fmaxf(m, wmax[w]);
    if (m > 0.f) atomicMax(gmax + n1, __float_as_uint(m));
  }
  if (do_energy) {
    block_sum<7>(acc, red);
    const int nb = gridDim.x * gridDim.y;
    const int b = blockIdx.y * gridDim.x + blockIdx.x;
    if (threadIdx.x == 0)
      for (int k = 0; k < 7; ++k) pout[(size_t)k * nb + b] = acc[k];
  }
}

// Scale of apply column (w, n1): values of w(n) g(n1, n) are at most
// APPLY_SCALE after multiplying by APPLY_SCALE 2^-e; gmaxbits = max |g| over
// the n1 plane, wmax = largest |w| (1, kmax2, kmax3, ...).  A power of two
// keeps the scaling exact.
__device__ __forceinline__ float class_scale(unsigned int gmaxbits, float wmax) {
  const float gm = __uint_as_float(gmaxbits) * wmax;
  if (!(gm > 0.f)) return APPLY_SCALE;
  return exp2f(-ceilf(log2f(gm))) * APPLY_SCALE;
}

// weights of the apply: 0 -> 1, 1 -> n2, 2 -> n3, 3 -> n2^2, 4 -> n2 n3,
// 5 -> n3^2
__device__ __forceinline__ float apply_weight(int w, int n2, int n3) {
  switch (w) {
    case 0: return 1.f;
    case 1: return (float)n2;
    case 2: return (float)n3;
    case 3: return (float)(n2 * n2);
    case 4: return (float)(n2 * n3);
    default: return (float)(n3 * n3);
  }
}

__device__ __forceinline__ float apply_wmax(int w, int ky, int kz) {
  switch (w) {
    case 0: return 1.f;
    case 1: return (float)ky;
    case 2: return (float)kz;
    case 3: return (float)(ky * ky);
    case 4: return (float)(ky * kz);
    default: return (float)(kz * kz);
  }
}

// ---------------------------------------------------------------------------
// B fragments of the apply GEMM.  cols = (part, w, n1): col = part*(W*N1) +
// w*N1 + n1, NCT tiles of 16; K: k' = 2c + t over C23p columns c.
//   Re col: B[2c] = Gr, B[2c+1] = Gi;  Im col: B[2c] = Gi, B[2c+1] = -Gr
// with G = APPLY_SCALE 2^-e(w,n1) w(n) g(n).  Layout [ks][hi|lo][ct][lane] of half8, so
// that the fragments of one K step are contiguous.
__global__ __launch_bounds__(MDFF_BLOCK) void mx_gprep(int N1, int N3, int C23, int C23p, int W,
                                                        int NCT, const double2* __restrict__ g,
                                                        const unsigned int* __restrict__ gmax,
                                                        int ky, int kz, _Float16* __restrict__ bf) {
  const int KS = (2 * C23p) / 32;
  const size_t total = (size_t)NCT * KS * 64;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int lane = (int)(t % 64);
  const int ct = (int)((t / 64) % NCT);
  const int ks = (int)(t / (64 * (size_t)NCT));
  const int colg = ct * 16 + (lane & 15);
  const int ncols = 2 * W * N1;
  int part = 0, w = 0, n1 = 0;
  float cs = 0.f;
  if (colg < ncols) {
    part = colg / (W * N1);
    const int wn = colg % (W * N1);
    w = wn / N1;
    n1 = wn % N1;
    cs = class_scale(gmax[n1], apply_wmax(w, ky, kz));
  }
  half8 h, l;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int kp = ks * 32 + 8 * (lane >> 4) + e;
    const int c = kp >> 1, tt = kp & 1;
    float v = 0.f;
    if (colg < ncols && c < C23) {
      const int n2 = c / N3 - ky, n3 = c % N3 - kz;
      const double2 gv = g[(size_t)n1 * C23 + c];
      const float sc = cs * apply_weight(w, n2, n3);
      const float gr = (float)gv.x * sc, gi = (float)gv.y * sc;
      if (part == 0)
        v = tt ? gi : gr;
      else
        v = tt ? -gr : gi;
    }
    _Float16 hh, ll;
    split16(v, hh, ll);
    h[e] = hh;
    l[e] = ll;
  }
  half8* o = (half8*)bf + (size_t)ks * 2 * NCT * 64 + ct * 64 + lane;
  o[0] = h;
  o[NCT * 64] = l;
}

// ---------------------------------------------------------------------------
// Apply GEMM: rows = atoms, K = 2 C23p, NCT column tiles.  A workgroup of
// APP_WAVES waves owns APP_WAVES*RT*16 atoms and one of S slices of K; wave
// w owns row tiles RT*w .. RT*w+RT-1 and all column tiles.  A fragments
// (A[i][2c] = Re P_i[c], A[i][2c+1] = Im P_i[c], P = E2 E3) are formed in
// registers from the atom-contiguous phase tables; the B fragments of a K
// step are staged in LDS (double buffered) and shared by the waves.
// Block -> (row block, K slice) is XCD aware: consecutive block ids land on
// consecutive XCDs, and the blocks of one XCD share a K slice, so that each
// XCD's L2 holds only its slice of the B fragments.
// Output Xp[s][col][i] (FP32, NCT*16 columns, atoms contiguous).
constexpr int APP_WAVES = 4;
constexpr int NXCD = 8;

template <int NCT, int RT>
__global__ __launch_bounds__(64 * APP_WAVES) void mx_apply_gemm(
    int npad, int N3, int C23, int KS, int S, int ksper, int mblocks,
    const float2* __restrict__ tj2, const float2* __restrict__ tj3,
    const _Float16* __restrict__ bf, float* __restrict__ Xp) {
  __shared__ u32x4 sB[2][2 * NCT * 64];
  const int per = NXCD / S;
  const int xcd = blockIdx.x % NXCD;
  const int sl = xcd / per;
  const int mb = (blockIdx.x / NXCD) * per + (xcd % per);
  if (mb >= mblocks) return;  // whole block, before any barrier
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int kgrp = lane >> 4;
  const int ks0 = sl * ksper;
  const int ks1 = min(KS, ks0 + ksper);
  const int ibase = mb * (APP_WAVES * RT * 16) + wave * RT * 16 + (lane & 15);
  floatx4 acc[RT][NCT];
#pragma unroll
  for (int r = 0; r < RT; ++r)
#pragma unroll
    for (int t = 0; t < NCT; ++t) acc[r][t] = floatx4{0.f, 0.f, 0.f, 0.f};
  // B fragments of one K step: NB 16-byte words, staged through LDS
  constexpr int NB = 2 * NCT * 64;
  constexpr int NT = 64 * APP_WAVES;
  constexpr int PER_FULL = NB / NT;
  constexpr int PER_T = (NB + NT - 1) / NT;
  const u32x4* bsrc = reinterpret_cast<const u32x4*>(bf);
  u32x4 pre[PER_T];
#define MX_FETCH_B(ks)                                                                   \
  {                                                                                      \
    _Pragma("unroll") for (int u = 0; u < PER_FULL; ++u) pre[u] =                       \
        bsrc[(size_t)(ks) * NB + threadIdx.x + u * NT];                                  \
    if (PER_T > PER_FULL) {                                                              \
      const int idx_ = threadIdx.x + PER_FULL * NT;                                      \
      pre[PER_T - 1] = idx_ < NB ? bsrc[(size_t)(ks) * NB + idx_] : u32x4{0u, 0u, 0u, 0u}; \
    }                                                                                    \
  }
#define MX_STAGE_B(buf)                                                                  \
  {                                                                                      \
    _Pragma("unroll") for (int u = 0; u < PER_FULL; ++u) sB[buf][threadIdx.x + u * NT] = \
        pre[u];                                                                          \
    if (PER_T > PER_FULL) {                                                              \
      const int idx_ = threadIdx.x + PER_FULL * NT;                                      \
      if (idx_ < NB) sB[buf][idx_] = pre[PER_T - 1];                                     \
    }                                                                                    \
  }
  if (ks0 < ks1) {
    MX_FETCH_B(ks0)
    MX_STAGE_B(0)
  }
  __syncthreads();
  // phases E2, E3 of the K step's columns c = 16 ks + 4 kgrp + e for the RT
  // atoms of this lane; columns past C23 read column 0 and are zeroed
  float2 pe2[RT][4], pe3[RT][4];
  unsigned okm = 0;
  auto load_e = [&](int ks) {
    const int cb = ks * 16 + 4 * kgrp;
    int n2 = cb / N3, n3 = cb - (cb / N3) * N3;
    okm = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const bool ok = cb + e < C23;
      okm |= ok ? (1u << e) : 0u;
      const int m2 = ok ? n2 : 0, m3 = ok ? n3 : 0;
#pragma unroll
      for (int r = 0; r < RT; ++r) {
        pe2[r][e] = tj2[(size_t)m2 * npad + ibase + 16 * r];
        pe3[r][e] = tj3[(size_t)m3 * npad + ibase + 16 * r];
      }
      if (++n3 == N3) {
        n3 = 0;
        ++n2;
      }
    }
  };
  // A fragments: A[i][2c] = Re P_i[c], A[i][2c+1] = Im P_i[c]
  auto make_frags = [&](half8 (&h8)[RT], half8 (&l8)[RT]) {
#pragma unroll
    for (int e = 0; e < 4; ++e)
#pragma unroll
      for (int r = 0; r < RT; ++r) {
        const float2 e2 = pe2[r][e], e3 = pe3[r][e];
        const bool ok = (okm >> e) & 1u;
        const float pr = ok ? e2.x * e3.x - e2.y * e3.y : 0.f;
        const float pi = ok ? e2.x * e3.y + e2.y * e3.x : 0.f;
        _Float16 h, l;
        split16(pr, h, l);
        h8[r][2 * e] = h;
        l8[r][2 * e] = l;
        split16(pi, h, l);
        h8[r][2 * e + 1] = h;
        l8[r][2 * e + 1] = l;
      }
  };
  half8 ah[RT], al[RT];
  if (ks0 < ks1) {
    load_e(ks0);
    make_frags(ah, al);
    if (ks0 + 1 < ks1) load_e(ks0 + 1);
  }
  int cur = 0;
  for (int ks = ks0; ks < ks1; ++ks) {
    const bool more = ks + 1 < ks1;
    if (more) MX_FETCH_B(ks + 1)
    const u32x4* B = sB[cur];
    // column tiles two at a time: four independent accumulators between
    // dependent MFMAs; the next pair's fragments are read from LDS ahead
    u32x4 h0 = B[lane], l0 = B[NCT * 64 + lane];
    u32x4 h1 = u32x4{0u, 0u, 0u, 0u}, l1 = u32x4{0u, 0u, 0u, 0u};
    if (NCT > 1) {
      h1 = B[64 + lane];
      l1 = B[NCT * 64 + 64 + lane];
    }
#pragma unroll
    for (int c0 = 0; c0 < NCT; c0 += 2) {
      u32x4 nh0 = h0, nl0 = l0, nh1 = h1, nl1 = l1;
      if (c0 + 2 < NCT) {
        nh0 = B[(c0 + 2) * 64 + lane];
        nl0 = B[NCT * 64 + (c0 + 2) * 64 + lane];
      }
      if (c0 + 3 < NCT) {
        nh1 = B[(c0 + 3) * 64 + lane];
        nl1 = B[NCT * 64 + (c0 + 3) * 64 + lane];
      }
      const bool two = c0 + 1 < NCT;
      const half8 b10 = __builtin_bit_cast(half8, h0), b20 = __builtin_bit_cast(half8, l0);
      const half8 b11 = __builtin_bit_cast(half8, h1), b21 = __builtin_bit_cast(half8, l1);
#pragma unroll
      for (int r = 0; r < RT; ++r) {
        acc[r][c0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[r], b10, acc[r][c0], 0, 0, 0);
        if (two)
          acc[r][c0 + 1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[r], b11, acc[r][c0 + 1], 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < RT; ++r) {
        acc[r][c0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[r], b20, acc[r][c0], 0, 0, 0);
        if (two)
          acc[r][c0 + 1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[r], b21, acc[r][c0 + 1], 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < RT; ++r) {
        acc[r][c0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[r], b10, acc[r][c0], 0, 0, 0);
        if (two)
          acc[r][c0 + 1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[r], b11, acc[r][c0 + 1], 0, 0, 0);
      }
      h0 = nh0;
      l0 = nl0;
      h1 = nh1;
      l1 = nl1;
    }
    // the next K step's fragments (independent of the MFMAs above) and the
    // phases of the one after
    if (more) {
      make_frags(ah, al);
      if (ks + 2 < ks1) load_e(ks + 2);
      MX_STAGE_B(cur ^ 1)
    }
    __syncthreads();
    cur ^= 1;
  }
  // D layout: acc[r][ct][e] = D[4*kgrp + e][ct*16 + lane%16]: rows = atoms.
  // X is column major (atoms contiguous): 4 consecutive atoms per store.
  float* X = Xp + (size_t)sl * npad * (NCT * 16);
  const int i0 = mb * (APP_WAVES * RT * 16) + wave * RT * 16 + 4 * kgrp;
#pragma unroll
  for (int r = 0; r < RT; ++r)
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) {
      const int colx = ct * 16 + (lane & 15);
      *(floatx4*)(X + (size_t)colx * npad + i0 + 16 * r) = acc[r][ct];
    }
}
#undef MX_FETCH_B
#undef MX_STAGE_B

// ---------------------------------------------------------------------------
// Sum of the K slices and fold over n1 with conj(E1_i[n1]) (FP64), into the
// U, V sums of k_apply64:
//   U1 = Im S'_1, U2 = Im S_n2, U3 = Im S_n3,
//   V11 = Re S''_1, V22 = Re S_n2n2, V33 = Re S_n3n3, V12 = Re S'_n2,
//   V13 = Re S'_n3, V23 = Re S_n2n3,
// S_w = sum_n1 conj(E1[n1]) X_w[n1], S' with an extra n1, S'' with n1^2.
__global__ __launch_bounds__(MDFF_BLOCK) void mx_apply_epi(int n, int npad, int N1, int W,
                                                            int ncolp, int S, int ky, int kz,
                                                            const float* __restrict__ Xp,
                                                            const float2* __restrict__ tj1,
                                                            const unsigned int* __restrict__ gmax,
                                                            double* __restrict__ out) {
  // grid (atom blocks, W): one weight class per thread
  const int w = blockIdx.y;
  __shared__ double sinv[25];
  if (threadIdx.x < N1)
    sinv[threadIdx.x] = 1.0 / (double)class_scale(gmax[threadIdx.x], apply_wmax(w, ky, kz));
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t sstride = (size_t)npad * ncolp;
  const float* xre = Xp + (size_t)(w * N1) * npad + i;
  const float* xim = Xp + (size_t)(W * N1 + w * N1) * npad + i;
  double sr = 0, si = 0, pr = 0, pi = 0, qr = 0;
  for (int m = 0; m < N1; ++m) {
    double xr = 0.0, xi = 0.0;
    for (int sl = 0; sl < S; ++sl) {
      xr += xre[sl * sstride + (size_t)m * npad];
      xi += xim[sl * sstride + (size_t)m * npad];
    }
    xr *= sinv[m];
    xi *= sinv[m];
    const float2 e = tj1[(size_t)m * npad + i];
    const double er = e.x, ei = -(double)e.y;  // conj(E1)
    const double vr = er * xr - ei * xi, vi = er * xi + ei * xr;
    sr += vr;
    si += vi;
    pr += m * vr;
    pi += m * vi;
    qr += (double)m * m * vr;
  }
  switch (w) {
    case 0:
      out[i] = pi;                     // U1 = Im S'_0
      if (W == 6) out[i + 3 * n] = qr; // V11 = Re S''_0
      break;
    case 1:
      out[i + n] = si;                 // U2 = Im S_1
      if (W == 6) out[i + 6 * n] = pr; // V12 = Re S'_1
      break;
    case 2:
      out[i + 2 * n] = si;             // U3 = Im S_2
      if (W == 6) out[i + 7 * n] = pr; // V13 = Re S'_2
      break;
    case 3: out[i + 4 * n] = sr; break;  // V22 = Re S_3
    case 4: out[i + 8 * n] = sr; break;  // V23 = Re S_4
    default: out[i + 5 * n] = sr; break; // V33 = Re S_5
  }
}

// ---------------------------------------------------------------------------
// host side

template <typename T>
static void mx_alloc(T** p, size_t count) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  if (hipMalloc((void**)p, (count ? count : 1) * sizeof(T)) != hipSuccess) *p = nullptr;
}

// apply launch shapes: (NCT, RT) and the number of K slices
constexpr int APP_NCT_F = 10, APP_RT_F = 2;  // field calls (W = 3)
constexpr int APP_NCT_E = 19, APP_RT_E = 1;  // final call with EFG (W = 6)
constexpr int APP_ROWS_MAX = APP_WAVES * 2 * 16;
constexpr int APP_SLICES = 2;
constexpr int RHO_NRT_MAX = 7;  // 4 weights x 25 n1 rows
constexpr int RHO_CT = 1, RHO_NWV = 4;

static bool sizes(MixedState* m) {
  CtxView v = view(m->ctx);
  m->kx = v.kmax[0];
  m->ky = v.kmax[1];
  m->kz = v.kmax[2];
  m->N1 = m->kx + 1;
  m->N2 = 2 * m->ky + 1;
  m->N3 = 2 * m->kz + 1;
  m->C23 = m->N2 * m->N3;
  m->C23p = (m->C23 + 63) / 64 * 64;  // whole column blocks of the rho GEMM
  // whole row blocks of the apply GEMM (tables are zero past n)
  m->npad = (v.n + APP_ROWS_MAX - 1) / APP_ROWS_MAX * APP_ROWS_MAX;
  m->nsplit = (m->npad + RHO_ATOMS_PER_SPLIT - 1) / RHO_ATOMS_PER_SPLIT;
  if (m->N1 > 25) return false;  // 4 N1 rows fit the 7 row tiles of mx_rho_gemm
  const size_t np = m->npad;
  mx_alloc(&m->tj1, (size_t)m->N1 * np);
  mx_alloc(&m->tj2, (size_t)m->N2 * np);
  mx_alloc(&m->tj3, (size_t)m->N3 * np);
  mx_alloc(&m->g, (size_t)m->N1 * m->C23);
  mx_alloc(&m->gmax, 32);
  m->rpart_cap = (size_t)m->nsplit * RHO_NRT_MAX * 16 * 2 * m->C23p;
  mx_alloc(&m->rpart, m->rpart_cap);
  mx_alloc(&m->Wt, (size_t)4 * np);
  mx_alloc(&m->afrag, (np / 16) * RHO_NRT_MAX * 2 * 64);
  const int KS = 2 * m->C23p / 32;
  m->bfrag_cap = (size_t)2 * APP_NCT_E * KS * 64 * 8;
  mx_alloc(&m->bfrag, m->bfrag_cap);
  m->X_cap = (size_t)APP_SLICES * np * APP_NCT_E * 16;
  mx_alloc(&m->X, m->X_cap);
  return m->tj1 && m->tj2 && m->tj3 && m->g && m->gmax && m->rpart && m->bfrag && m->X && m->Wt &&
         m->afrag;
}

MixedState* create(mdffgpu_ctx* c) {
  MixedState* m = new MixedState();
  m->ctx = c;
  CtxView v = view(c);
  if (v.P->lcoulomb && !sizes(m)) {
    destroy(m);
    return nullptr;
  }
  return m;
}

void destroy(MixedState* m) {
  if (!m) return;
  void* p[] = {m->tj1, m->tj2, m->tj3, m->rpart, m->g, m->gmax, m->bfrag, m->X, m->Wt, m->afrag};
  for (void* x : p)
    if (x) (void)hipFree(x);
  delete m;
}

void on_cell(MixedState* m) {
  if (view(m->ctx).P->lcoulomb) sizes(m);
}

template <int NRT>
static void launch_rho(MixedState* m, const CtxView& v) {
  const int groups = m->C23p / 16;
  dim3 grid((groups + RHO_CT * RHO_NWV - 1) / (RHO_CT * RHO_NWV), m->nsplit);
  mx_rho_gemm<NRT, RHO_CT, RHO_NWV><<<grid, 64 * RHO_NWV, 0, v.st>>>(
      m->npad, m->N3, m->C23, m->C23p, m->afrag, m->tj2, m->tj3, RHO_ATOMS_PER_SPLIT, m->rpart);
}

template <int NCT, int RT>
static void launch_apply(MixedState* m, const CtxView& v) {
  const int KS = 2 * m->C23p / 32;
  const int S = APP_SLICES;
  const int ksper = (KS + S - 1) / S;
  const int rows = APP_WAVES * RT * 16;
  const int mblocks = m->npad / rows;
  const int per = NXCD / S;
  const int grid = (mblocks + per - 1) / per * NXCD;
  mx_apply_gemm<NCT, RT><<<grid, 64 * APP_WAVES, 0, v.st>>>(m->npad, m->N3, m->C23, KS, S, ksper,
                                                           mblocks, m->tj2, m->tj3, m->bfrag,
                                                           m->X);
}

void kspace(MixedState* m, int mode, const double* q, const double* mu, bool chg, bool dip,
            bool efg, double* apart) {
  CtxView v = view(m->ctx);
  hipStream_t st = v.st;
  const int n = v.n;
  // phase tables: positions are fixed within an evaluation, rebuild on the
  // first reciprocal call of each evaluation (the static or final call)
  if (mode != ES_IND) {
    ProfScope _ps(m->ctx, K_MX_PHASE);
    mx_tables<<<(m->npad + MDFF_BLOCK - 1) / MDFF_BLOCK, MDFF_BLOCK, 0, st>>>(
        n, m->npad, v.s, m->kx, m->ky, m->kz, m->tj1, m->tj2, m->tj3);
  }
  const int NW = (chg ? 1 : 0) + (dip ? 3 : 0);
  const int NRT = (NW * m->N1 + 15) / 16;
  const int NRTl = (NRT <= 2) ? 2 : (NRT <= 4) ? 4 : (NRT <= 5) ? 5 : 7;
  {
    ProfScope _ps(m->ctx, K_MX_WEIGHTS);
    mx_weights<<<(m->npad + MDFF_BLOCK - 1) / MDFF_BLOCK, MDFF_BLOCK, 0, st>>>(
        n, m->npad, chg, dip, q, mu, v.cell, m->Wt, m->gmax, m->N1);
    const size_t at = (size_t)(m->npad / 16) * NRTl * 64;
    const unsigned ga = (unsigned)((at + MDFF_BLOCK - 1) / MDFF_BLOCK);
    switch (NRTl) {
      case 2: mx_rho_afrag<2><<<ga, MDFF_BLOCK, 0, st>>>(m->npad, m->N1, NW, m->Wt, m->tj1, m->afrag); break;
      case 4: mx_rho_afrag<4><<<ga, MDFF_BLOCK, 0, st>>>(m->npad, m->N1, NW, m->Wt, m->tj1, m->afrag); break;
      case 5: mx_rho_afrag<5><<<ga, MDFF_BLOCK, 0, st>>>(m->npad, m->N1, NW, m->Wt, m->tj1, m->afrag); break;
      default: mx_rho_afrag<7><<<ga, MDFF_BLOCK, 0, st>>>(m->npad, m->N1, NW, m->Wt, m->tj1, m->afrag); break;
    }
  }
  {
    ProfScope _ps(m->ctx, K_MX_RHO);
    switch (NRTl) {
      case 2: launch_rho<2>(m, v); break;
      case 4: launch_rho<4>(m, v); break;
      case 5: launch_rho<5>(m, v); break;
      default: launch_rho<7>(m, v); break;
    }
  }
  const dim3 gred((m->C23 + MDFF_BLOCK - 1) / MDFF_BLOCK, m->N1);
  const int gk = gred.x * gred.y;
  {
    ProfScope _ps(m->ctx, K_MX_RHO_RED);
    mx_rho_reduce<<<gred, MDFF_BLOCK, 0, st>>>(m->N1, m->N2, m->N3, m->C23, m->C23p, NW, chg, dip,
                                             m->nsplit, NRTl * 16, m->rpart, v.cell,
                                             v.P->alphaES, mode == ES_FINAL, m->g, m->gmax,
                                             v.part);
  }
  if (mode == ES_FINAL)
    k_reduce_rows<<<7, MDFF_BLOCK, 0, st>>>(v.part, gk, v.sc + v.sc_rec);
  const int W = efg ? 6 : 3;
  const int ncols = 2 * W * m->N1;
  const int NCT = (ncols + 15) / 16;
  const int NCTl = (NCT <= APP_NCT_F) ? APP_NCT_F : APP_NCT_E;
  const int KS = 2 * m->C23p / 32;
  {
    ProfScope _ps(m->ctx, K_MX_GPREP);
    const size_t total = (size_t)NCTl * KS * 64;
    mx_gprep<<<(unsigned)((total + MDFF_BLOCK - 1) / MDFF_BLOCK), MDFF_BLOCK, 0, st>>>(
        m->N1, m->N3, m->C23, m->C23p, W, NCTl, m->g, m->gmax, m->ky, m->kz, m->bfrag);
  }
  {
    ProfScope _ps(m->ctx, K_MX_APPLY);
    if (NCTl == APP_NCT_F)
      launch_apply<APP_NCT_F, APP_RT_F>(m, v);
    else
      launch_apply<APP_NCT_E, APP_RT_E>(m, v);
  }
  {
    ProfScope _ps(m->ctx, K_MX_COMBINE);
    mx_apply_epi<<<dim3((n + MDFF_BLOCK - 1) / MDFF_BLOCK, W), MDFF_BLOCK, 0, st>>>(
        n, m->npad, m->N1, W, NCTl * 16, APP_SLICES, m->ky, m->kz, m->X, m->tj1, m->gmax, apart);
  }
}

}  // namespace mixed
}  // namespace mdff
