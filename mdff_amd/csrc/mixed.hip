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
#include <type_traits>
#include <vector>

#include "mixed.h"

namespace mdff {
namespace mixed {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));  // 16 raw bytes

constexpr int RHO_ATOMS_PER_SPLIT = 1024;  // default atoms per structure-factor split
constexpr float DIP_SCALE = DIP_WEIGHT_SCALE;  // dipole weights are scaled by 2^8
// The apply operand w g of each n1 plane is normalised to |.| <= APPLY_SCALE
// (a power of two): most elements of a plane are far below its maximum, and
// this keeps their fp16 lo parts above the subnormal range (fp16 normals
// start at 2^-14; |lo| <= 2^-11 |x|).  2^12 leaves the largest hi part at
// 4096, well inside the fp16 range.
constexpr float APPLY_SCALE = 4096.0f;

struct MixedState {
  mdffgpu_ctx* ctx;
  int N1 = 0, N2 = 0, N3 = 0, C23 = 0, C23p = 0, npad = 0;
  int CH = 0, CHp = 0;  // conjugate-paired half of the (n2, n3) columns
  int kx = 0, ky = 0, kz = 0;
  int nsplit = 0;
  // one configuration: the atoms that can carry a dipole (the dipole-only
  // structure factors sum over their slots, in splits of their own) and the
  // atoms whose field the charge- and dipole-field calls must give (the
  // rows of their apply GEMM and combine), as whole 128-slot blocks
  // [dj0, dj1) and [fj0, fj1) of the padded axis; all slots by default
  int dip0 = 0, dip1 = -1, fld0 = 0, fld1 = -1;
  int dj0 = 0, dj1 = 0, fj0 = 0, fj1 = 0;
  int2* d_split_dip = nullptr;
  int nsplit_dip = 0, nsplit_dip_cap = 0;
  int cg_parts = 0;  // p.Ap block partials of the last fused combine
  // phase tables, float2 (re, im)
  float2* tj1 = nullptr;  // [n1][j]   (N1 x npad)
  float2* tj2 = nullptr;  // [n2+ky][j]
  float2* tj3 = nullptr;  // [n3+kz][j]
  double* rpart = nullptr;      // [split][row][col]  FP64 (or FP32) partials
  size_t rpart_cap = 0;
  double2* g = nullptr;         // [n1][c]
  unsigned int* gmax = nullptr; // max |g| per n1 plane, as float bits
  _Float16* bfrag = nullptr;    // apply B fragments, hi then lo
  size_t bfrag_cap = 0;
  float* X = nullptr;           // [slice][i][col]
  float* Wt = nullptr;          // [w][j] structure-factor weights
  u32x4* afrag = nullptr;       // structure-factor A fragments of one call
  // charge structure factors M_q[seg][n1][c] (FP64) of the last charge-only
  // call, for the evaluation (epoch) and charge array they belong to
  double2* mq = nullptr;
  long long mq_epoch = -1;
  const double* mq_q = nullptr;
  long long tables_eval = -1;   // evaluation the phase tables belong to
  size_t X_cap = 0;
  // configurations on the padded atom axis: none (one configuration, the
  // context's) or a segment table of the batched path
  int N = 0;                    // atoms of the caller's arrays
  int nseg = 1;
  MxSeg* d_seg = nullptr;
  int* d_kseg = nullptr;        // segment of each 32-slot group
  int2* d_split = nullptr;      // [j0, j1) of each structure-factor split
  int2* d_colofs = nullptr;     // apply: phase-table row offsets of each column
  double* sinv = nullptr;       // [seg][w][n1] apply scales
  double* apart = nullptr;      // batched path: U, V sums [9][N]
  // launch shape (autotuned for batches): K slices of the apply GEMM and
  // atoms per structure-factor split; the segment table as laid out
  int app_slices = 2, split_atoms = RHO_ATOMS_PER_SPLIT, x_slices = 0;
  std::vector<MxSeg> h_seg;
};

// Atom of padded slot jp (-1 for padding) and its segment z.  Without a
// segment table the slots are the atoms 0 .. n-1 of one configuration.
__device__ __forceinline__ int slot_atom(int jp, int n, const MxSeg* __restrict__ seg,
                                         const int* __restrict__ kseg, int& z) {
  if (!seg) {
    z = 0;
    return jp < n ? jp : -1;
  }
  z = kseg[jp >> 5];
  const int l = jp - seg[z].poff;
  return l < seg[z].n ? seg[z].off + l : -1;
}

// ---------------------------------------------------------------------------
// phase tables from the FP64 fractional coordinates: one thread per (slot,
// table row), rows 0..kx of E1, then the 2 ky + 1 of E2 and 2 kz + 1 of E3
__global__ __launch_bounds__(MDFF_BLOCK) void mx_tables(int n, int npad, const double* __restrict__ s,
                                                         const MxSeg* __restrict__ seg,
                                                         const int* __restrict__ kseg,
                                                         int kx, int ky, int kz, float2* tj1,
                                                         float2* tj2, float2* tj3) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= npad) return;
  int row = blockIdx.y, axis = 0, m = row;
  float2* out = tj1;
  if (row > kx) {
    row -= kx + 1;
    if (row < 2 * ky + 1) {
      axis = 1;
      m = row - ky;
      out = tj2;
    } else {
      row -= 2 * ky + 1;
      axis = 2;
      m = row - kz;
      out = tj3;
    }
  }
  int z;
  const int i = slot_atom(j, n, seg, kseg, z);
  float2 v = make_float2(0.f, 0.f);
  if (i >= 0) {
    double sn, cs;
    sincospi(2.0 * m * s[i + (size_t)axis * n], &sn, &cs);
    v = make_float2((float)cs, (float)sn);
  }
  out[(size_t)row * npad + j] = v;
}

// hi + lo = x to 22 bits.  x is first pinned as a rounded FP32 value:
// otherwise the compiler may fuse the producing multiply into both
// conversions (v_fma_mix*_f16 from the exact product for lo, v_cvt_pk_f16_f32
// of the FP32-rounded product for hi), and at an fp16 rounding tie the two
// roundings of "hi" differ by one ulp, so hi + lo misses x by an ulp.  The
// pin is not volatile: it has no side effects, so the scheduler may move
// memory operations across it.
__device__ __forceinline__ void split16(float x, _Float16& hi, _Float16& lo) {
  __asm__("" : "+v"(x));
  hi = (_Float16)x;
  lo = (_Float16)(x - (float)hi);
}

// split16 of two values at once: one packed conversion for both hi parts
// (v_cvt_pk_f16_f32), their FP32 values back, two residuals and one packed
// conversion for the lo parts; the same roundings as split16.
typedef _Float16 half2v __attribute__((ext_vector_type(2)));
typedef float float2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void split16x2(float x0, float x1, half2v& hi, half2v& lo) {
  __asm__("" : "+v"(x0), "+v"(x1));
  hi = __builtin_convertvector((float2v){x0, x1}, half2v);
  const float2v hf = __builtin_convertvector(hi, float2v);
  lo = __builtin_convertvector((float2v){x0 - hf.x, x1 - hf.y}, half2v);
}

// ---------------------------------------------------------------------------
// Weights of the structure-factor GEMM, FP32, [w][slot] (zero on padding):
// w0 = q (when CHG), then 2^8 * 2 pi B^T mu (when DIP), B of the slot's
// configuration.
__global__ __launch_bounds__(MDFF_BLOCK) void mx_weights(int n, int npad, int chg, int dip,
                                                          const double* __restrict__ q,
                                                          const double* __restrict__ mu, Cell cell,
                                                          const MxSeg* __restrict__ seg,
                                                          const int* __restrict__ kseg,
                                                          float* __restrict__ Wt) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= npad) return;
  int z;
  const int i = slot_atom(j, n, seg, kseg, z);
  int w = 0;
  if (chg) Wt[(size_t)(w++) * npad + j] = i >= 0 ? (float)q[i] : 0.f;
  if (dip) {
    const double* B = seg ? seg[z].cell.B : cell.B;
    for (int a = 0; a < 3; ++a) {
      float v = 0.f;
      if (i >= 0) {
        const double m = B[3 * a] * mu[i] + B[3 * a + 1] * mu[i + n] + B[3 * a + 2] * mu[i + 2 * n];
        v = (float)(m * 6.283185307179586) * DIP_SCALE;
      }
      Wt[(size_t)(w++) * npad + j] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// A fragments of the structure-factor GEMM for the whole call, in the
// fragment order the GEMM stages through LDS: per K step g (32 atoms), row
// tile rt, hi|lo, lane l: row r = rt*16 + l%16 = 2 p + x for the pair
// p = w*N1 + n1 and x = re|im, atoms 32 g + 8 (l / 16) + (0..7), values
// x-part of W_w[j] E1_j[n1] split into fp16 hi + lo.  Rows past 2 NW N1 are
// zero.  Also resets the per-(configuration, n1) maxima of |g| that
// mx_rho_reduce accumulates.
template <int NRT2>
__global__ __launch_bounds__(MDFF_BLOCK) void mx_rho_afrag(int npad, int N1, int NW,
                                                            const float* __restrict__ Wt,
                                                            const float2* __restrict__ tj1,
                                                            u32x4* __restrict__ afrag,
                                                            unsigned int* __restrict__ gmax,
                                                            int nreset) {
  if (blockIdx.x == 0)
    for (int t = threadIdx.x; t < nreset; t += blockDim.x) gmax[t] = 0u;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)(npad / 32) * NRT2 * 64;
  if (t >= total) return;
  const int l = (int)(t & 63);
  const int rt = (int)((t >> 6) % NRT2);
  const int g = (int)((t >> 6) / NRT2);
  const int r = rt * 16 + (l & 15);
  const int p = r >> 1, x = r & 1;
  const int j = g * 32 + 8 * (l >> 4);
  half8 h, lo;
  if (p < NW * N1) {
    const int w = p / N1, n1 = p % N1;
    const float4* wp = reinterpret_cast<const float4*>(Wt + (size_t)w * npad + j);
    const float4* ep = reinterpret_cast<const float4*>(tj1 + (size_t)n1 * npad + j);
    const float4 w0 = wp[0], w1 = wp[1];
    const float ww[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float4 e = ep[q];  // atoms j + 2q, j + 2q + 1 as (re, im, re, im)
      _Float16 a, b;
      split16(ww[2 * q] * (x ? e.y : e.x), a, b);
      h[2 * q] = a;
      lo[2 * q] = b;
      split16(ww[2 * q + 1] * (x ? e.w : e.z), a, b);
      h[2 * q + 1] = a;
      lo[2 * q + 1] = b;
    }
  } else {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      h[e] = (_Float16)0.f;
      lo[e] = (_Float16)0.f;
    }
  }
  u32x4* o = afrag + (((size_t)g * NRT2 + rt) * 2) * 64 + l;
  o[0] = __builtin_bit_cast(u32x4, h);
  o[64] = __builtin_bit_cast(u32x4, lo);
}

// ---------------------------------------------------------------------------
// Structure-factor GEMM with conjugate-paired columns.  The (n2, n3) mesh is
// closed under negation: column c' = C23 - 1 - c has the phases
// B_c' = E2[-n2] E3[-n3] = conj(B_c).  For the half c < CH, the four real
// products
//   P_xy[p][c] = sum_j A_x[p][j] B_y[j][c],  A = W_w E1[n1],  x, y = re|im
// give both structure factors of the pair,
//   S_c  = (P_rr - P_ii) + i (P_ri + P_ir),
//   S_c' = (P_rr + P_ii) + i (P_ir - P_ri),
// at half the MFMA work and half the B operand of complex products over all
// columns.  GEMM rows (p, x) as row 2p + x (NRT2 tiles), columns (c, y) as
// two tiles (y = re, im) per 16 columns c, K = atoms (32 per MFMA step).
// The B operand is never stored: B[j][c] = E2_j[n2] E3_j[n3] (c = n2 N3 +
// n3) is combined and split into fp16 hi + lo in registers from the
// separable phase tables, each lane its column and 8 atoms (mx_rho_bfrag).
// Two kernels share this scheme:
//  - mx_rho_gen (FP32 chains per split, and the 4-tile calls): a wave owns
//    one column tile and all row tiles, one wave per SIMD, A fragments read
//    from LDS one step ahead into a second register set;
//  - mx_rho_gemm (FP64 flushes, up to 13 row tiles): a wave owns one column
//    tile and RTW row tiles of NRT2.
// With charge weights (F64) the FP32 MFMA chains are flushed into FP64
// every 128 atoms: neutral sums cancel at small |k|; dipole weights alone
// keep one FP32 chain per split.  Blocks map to (atom split, column block)
// XCD-aware: the blocks of one XCD take consecutive indices of the
// split-major order, so each XCD's L2 holds the A fragments and phases of
// about one split; a split is an atom range of one configuration.  Every
// K step issues the same loads, unconditionally (addresses past the split
// are clamped to its last step and the data left unused), so that the
// compiler's count of loads in flight is exact around the loop and a wait
// for one step's data does not drain the prefetches behind it.  Columns
// past CH load the last column's phases: a B column only reaches the same
// D column, and those are not stored.  Output: part[split][p][col] with S
// at Re plane col and Im plane C23p + col, for both c and c'.
constexpr int NXCD = 8;

// element type of the per-split structure-factor sums: FP64 where the FP32
// chains are flushed into FP64, FP32 where a split is one FP32 chain
template <bool F64>
using RhoPart = typename std::conditional<F64, double, float>::type;

// B fragments (Re hi, Re lo, Im hi, Im lo) of one K step and column from
// the phases of 8 atoms, e = E2 rows (4 float4 = 8 atoms as re, im pairs)
// then E3 rows.
__device__ __forceinline__ void mx_rho_bfrag(half8 (&d)[4], const float4 (&e)[8]) {
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const float4 a = e[u], b = e[4 + u];  // atoms 2u, 2u + 1 as (re, im, re, im)
    half2v h, lo;
    split16x2(a.x * b.x - a.y * b.y, a.z * b.z - a.w * b.w, h, lo);
    d[0][2 * u] = h.x;
    d[0][2 * u + 1] = h.y;
    d[1][2 * u] = lo.x;
    d[1][2 * u + 1] = lo.y;
    split16x2(a.x * b.y + a.y * b.x, a.z * b.w + a.w * b.z, h, lo);
    d[2][2 * u] = h.x;
    d[2][2 * u + 1] = h.y;
    d[3][2 * u] = lo.x;
    d[3][2 * u + 1] = lo.y;
  }
}

// one quarter (atoms 2u, 2u + 1) of mx_rho_bfrag
__device__ __forceinline__ void mx_rho_bfrag_part(half8 (&d)[4], const float4 (&e)[8], int u) {
  const float4 a = e[u], b = e[4 + u];
  half2v h, lo;
  split16x2(a.x * b.x - a.y * b.y, a.z * b.z - a.w * b.w, h, lo);
  d[0][2 * u] = h.x;
  d[0][2 * u + 1] = h.y;
  d[1][2 * u] = lo.x;
  d[1][2 * u + 1] = lo.y;
  split16x2(a.x * b.y + a.y * b.x, a.z * b.w + a.w * b.z, h, lo);
  d[2][2 * u] = h.x;
  d[2][2 * u + 1] = h.y;
  d[3][2 * u] = lo.x;
  d[3][2 * u + 1] = lo.y;
}

// the E2, E3 phases of a lane's column for the 8 atoms jb + 8 kgrp + (0..7)
__device__ __forceinline__ void mx_rho_phases(float4 (&d)[8], const float4* e2row,
                                              const float4* e3row, int jb, int kgrp) {
  const int q = (jb + 8 * kgrp) >> 1;  // float4 = two atoms
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    d[u] = e2row[q + u];
    d[4 + u] = e3row[q + u];
  }
}

template <int NRT2, int RTW, bool F64>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(
    (RTW > 5 || F64) ? 1 : 2))) void mx_rho_gemm(int C23, int CH, int C23p, int nrt_used, int ncb,
                                                 int nblk_used, const int2* __restrict__ splits,
                                                 const u32x4* __restrict__ afrag, int npad, int N3,
                                                 const float2* __restrict__ tj2,
                                                 const float2* __restrict__ tj3,
                                                 double* __restrict__ out) {
  constexpr int RG = (NRT2 + RTW - 1) / RTW;  // row groups
  constexpr int CPB = 4 / RG;                 // column tiles per block
  constexpr int NRTP = RG * RTW;              // row tiles incl. padding
  static_assert(RG * CPB == 4, "four waves per block");
  __shared__ u32x4 sA[2][NRTP][2][64];  // [buf][rt][hi|lo][lane], fp16 x8
  const int per = gridDim.x / NXCD;
  const int gid = (blockIdx.x % NXCD) * per + blockIdx.x / NXCD;
  if (gid >= nblk_used) return;  // whole block, before any barrier
  const int split = gid / ncb, cblk = gid % ncb;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int kgrp = lane >> 4;
  const int rg = wave % RG;
  const int ctile = cblk * CPB + wave / RG;
  const int rt0 = rg * RTW;
  const bool tile_ok = ctile * 16 < CH && rt0 < nrt_used;  // wave-uniform
  const int c = ctile * 16 + (lane & 15);
  const bool cok = c < CH;
  // padding row tiles (past NRT2) multiply zeros
  if constexpr (NRTP > NRT2) {
    constexpr int PADW = (NRTP - NRT2) * 2 * 64;
    for (int idx = threadIdx.x; idx < 2 * PADW; idx += 256)
      (&sA[idx / PADW][NRT2][0][0])[idx % PADW] = u32x4{0u, 0u, 0u, 0u};
  }
  floatx4 acc[2][RTW];
  double accd[2][RTW][4];
#pragma unroll
  for (int y = 0; y < 2; ++y)
#pragma unroll
    for (int r = 0; r < RTW; ++r) {
      acc[y][r] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int e = 0; e < 4; ++e) accd[y][r][e] = 0.0;
    }
  const int j0 = splits[split].x;
  const int j1 = splits[split].y;  // j1 - j0: a multiple of 128
  const int jlast = j1 - 32;

  // A fragments of one K step: ABLK contiguous 16-byte words of afrag,
  // fetched two K steps before they are staged, in a ring of two register
  // sets (slot = step parity)
  constexpr int ABLK = NRT2 * 2 * 64;
  constexpr int APER = (ABLK + 255) / 256;
  u32x4 apre[2][APER];
  auto fetch_a = [&](int jb, int slot) {
    const u32x4* src = afrag + (size_t)(jb / 32) * ABLK;
#pragma unroll
    for (int u = 0; u < APER; ++u) {
      const int idx = threadIdx.x + u * 256;
      if (ABLK % 256 == 0 || idx < ABLK) apre[slot][u] = src[idx];
    }
  };
  auto stage_a = [&](int buf, int slot) {
    u32x4* dst = &sA[buf][0][0][0];
#pragma unroll
    for (int u = 0; u < APER; ++u) {
      const int idx = threadIdx.x + u * 256;
      if (ABLK % 256 == 0 || idx < ABLK) dst[idx] = apre[slot][u];
    }
  };
  // B fragments one K step ahead (ring of two), made from phases loaded
  // EPF steps ahead (EPF divides the 4 unrolled steps: ring index = step
  // index mod EPF)
  const int cl = min(c, CH - 1);
  const int n2c = cl / N3;
  const float4* e2row = reinterpret_cast<const float4*>(tj2 + (size_t)n2c * npad);
  const float4* e3row = reinterpret_cast<const float4*>(tj3 + (size_t)(cl - n2c * N3) * npad);
  constexpr int EPF = RTW > 5 ? 4 : 2;
  float4 raw[EPF][8];
  half8 bq[2][4];

  mx_rho_phases(raw[0], e2row, e3row, j0, kgrp);
  mx_rho_bfrag(bq[0], raw[0]);
#pragma unroll
  for (int k = 1; k <= EPF; ++k) mx_rho_phases(raw[k % EPF], e2row, e3row, min(j0 + 32 * k, jlast), kgrp);
  fetch_a(j0, 0);
  stage_a(0, 0);
  fetch_a(min(j0 + 32, jlast), 1);
  fetch_a(min(j0 + 64, jlast), 0);
  __syncthreads();
  const floatx4 zero = floatx4{0.f, 0.f, 0.f, 0.f};
  // 128 atoms (4 K steps) per FP32 chain, then an FP64 flush; the first K
  // step of a chain starts from zero
  for (int jb = j0; jb < j1; jb += 128) {
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const int js = jb + 32 * ks;
      const int cs = ks & 1, ns = cs ^ 1;
      if (tile_ok) {
        // three passes (lo.hi, then hi.hi and hi.lo), 2 RTW independent
        // accumulators between dependent MFMAs
        half8 al[RTW], ah[RTW];
#pragma unroll
        for (int r = 0; r < RTW; ++r) {
          al[r] = __builtin_bit_cast(half8, sA[ks & 1][rt0 + r][1][lane]);
          ah[r] = __builtin_bit_cast(half8, sA[ks & 1][rt0 + r][0][lane]);
        }
        // all LDS reads issued before the first MFMA waits on one
        __builtin_amdgcn_sched_barrier(0);
        // the next step's B fragments, between this step's MFMAs
        mx_rho_bfrag(bq[ns], raw[(ks + 1) % EPF]);
#pragma unroll
        for (int r = 0; r < RTW; ++r) {
          acc[0][r] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[r], bq[cs][0], (ks || !F64) ? acc[0][r] : zero, 0, 0, 0);
          acc[1][r] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[r], bq[cs][2], (ks || !F64) ? acc[1][r] : zero, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < RTW; ++r) {
          acc[0][r] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[r], bq[cs][0], acc[0][r], 0, 0, 0);
          acc[1][r] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[r], bq[cs][2], acc[1][r], 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < RTW; ++r) {
          acc[0][r] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[r], bq[cs][1], acc[0][r], 0, 0, 0);
          acc[1][r] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[r], bq[cs][3], acc[1][r], 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < 6 * RTW; ++i) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
          __builtin_amdgcn_sched_group_barrier(0x002, 3, 0);  // VALU
        }
      }
      // phases EPF steps ahead, the next step's A fragments into the other
      // LDS buffer, and the A fragments three steps ahead
      static_assert(4 % EPF == 0, "ring index by the unrolled step");
      mx_rho_phases(raw[(ks + 1) % EPF], e2row, e3row, min(js + 32 * (1 + EPF), jlast), kgrp);
      stage_a(ns, ns);
      fetch_a(min(js + 96, jlast), ns);
      __syncthreads();
    }
    if (F64 && tile_ok) {
#pragma unroll
      for (int y = 0; y < 2; ++y)
#pragma unroll
        for (int r = 0; r < RTW; ++r)
#pragma unroll
          for (int e = 0; e < 4; ++e) accd[y][r][e] += (double)acc[y][r][e];
    }
  }
  // FP32 chains are widened at the store
  auto P = [&](int y, int r, int e) -> double {
    if constexpr (F64) return accd[y][r][e];
    else return (double)acc[y][r][e];
  };
  if (!tile_ok || !cok) return;
  // D layout: acc[y][r][e] = P_xy of row (rt0 + r)*16 + 4 kgrp + e: e = 0, 1
  // are (re, im) of pair p0, e = 2, 3 of pair p0 + 1
  double* o = out + (size_t)split * (NRT2 * 8) * (2 * C23p);
  const int cp = C23 - 1 - c;
#pragma unroll
  for (int r = 0; r < RTW; ++r) {
    if (rt0 + r >= nrt_used) continue;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int p = (rt0 + r) * 8 + 2 * kgrp + h;
      const double Prr = P(0, r, 2 * h), Pir = P(0, r, 2 * h + 1);
      const double Pri = P(1, r, 2 * h), Pii = P(1, r, 2 * h + 1);
      double* op = o + (size_t)p * (2 * C23p);
      op[c] = Prr - Pii;
      op[C23p + c] = Pri + Pir;
      if (cp != c) {  // the centre column (0, 0) pairs with itself
        op[cp] = Prr + Pii;
        op[C23p + cp] = Pir - Pri;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Structure-factor GEMM, one wave per SIMD: every wave owns one column
// tile and all NRT2 row tiles (four column tiles per block, one block per
// CU).  A fragments: fetched from afrag three K steps ahead, staged into a
// ring of three LDS buffers two steps ahead, and read into a second
// register set one step ahead, under the MFMAs of the current step, with
// the next step's B fragments.  One barrier per K step; the MFMAs run
// unconditionally (waves past CH work on clamped phases and store
// nothing), so that a step is one scheduling region.
template <int NRT2, bool F64>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void mx_rho_gen(
    int C23, int CH, int C23p, int nrt_used, int ncb, int nblk_used, const int2* __restrict__ splits,
    const u32x4* __restrict__ afrag, int npad, int N3, const float2* __restrict__ tj2,
    const float2* __restrict__ tj3, RhoPart<F64>* __restrict__ out) {
  constexpr int NBUF = 3;
  __shared__ u32x4 sA[NBUF][NRT2][2][64];  // [buf][rt][hi|lo][lane], fp16 x8
  const int per = gridDim.x / NXCD;
  const int gid = (blockIdx.x % NXCD) * per + blockIdx.x / NXCD;
  if (gid >= nblk_used) return;  // whole block, before any barrier
  const int split = gid / ncb, cblk = gid % ncb;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int kgrp = lane >> 4;
  const int ctile = cblk * 4 + wave;
  const int c = ctile * 16 + (lane & 15);
  const int cl = min(c, CH - 1);
  const int n2c = cl / N3;
  const float4* e2row = reinterpret_cast<const float4*>(tj2 + (size_t)n2c * npad);
  const float4* e3row = reinterpret_cast<const float4*>(tj3 + (size_t)(cl - n2c * N3) * npad);
  const int j0 = splits[split].x;
  const int j1 = splits[split].y;  // j1 - j0: a multiple of 128
  const int jlast = j1 - 32;

  constexpr int ABLK = NRT2 * 2 * 64;
  constexpr int APER = (ABLK + 255) / 256;
  u32x4 apre[APER];
  auto fetch_a = [&](int jb) {
    const u32x4* src = afrag + (size_t)(jb / 32) * ABLK;
#pragma unroll
    for (int u = 0; u < APER; ++u) {
      const int idx = threadIdx.x + u * 256;
      if (ABLK % 256 == 0 || idx < ABLK) apre[u] = src[idx];
    }
  };
  auto stage_a = [&](int buf) {
    u32x4* dst = &sA[buf][0][0][0];
#pragma unroll
    for (int u = 0; u < APER; ++u) {
      const int idx = threadIdx.x + u * 256;
      if (ABLK % 256 == 0 || idx < ABLK) dst[idx] = apre[u];
    }
  };
  half8 ah[2][NRT2], al[2][NRT2];
  auto read_a = [&](int buf, int set) {
#pragma unroll
    for (int r = 0; r < NRT2; ++r) {
      al[set][r] = __builtin_bit_cast(half8, sA[buf][r][1][lane]);
      ah[set][r] = __builtin_bit_cast(half8, sA[buf][r][0][lane]);
    }
  };
  // phases EPF steps ahead (EPF divides the 4 unrolled steps), B fragments
  // one step ahead
  constexpr int EPF = NRT2 > 4 ? 2 : 4;
  float4 raw[EPF][8];
  half8 bq[2][4];  // Re hi, Re lo, Im hi, Im lo

  floatx4 acc[2][NRT2];
  double accd[2][NRT2][4];
#pragma unroll
  for (int y = 0; y < 2; ++y)
#pragma unroll
    for (int r = 0; r < NRT2; ++r) {
      acc[y][r] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int e = 0; e < 4; ++e) accd[y][r][e] = 0.0;
    }
  // prologue: A(0) in registers, A(1) in LDS buffer 1, A(2) fetched, B(0)
  // made, phases of steps 1..EPF in flight
  fetch_a(j0);
  stage_a(0);
  fetch_a(min(j0 + 32, jlast));
  stage_a(1);
  fetch_a(min(j0 + 64, jlast));
  mx_rho_phases(raw[0], e2row, e3row, j0, kgrp);
  mx_rho_bfrag(bq[0], raw[0]);
#pragma unroll
  for (int k = 1; k <= EPF; ++k) mx_rho_phases(raw[k % EPF], e2row, e3row, min(j0 + 32 * k, jlast), kgrp);
  __syncthreads();
  read_a(0, 0);
  const floatx4 zero = floatx4{0.f, 0.f, 0.f, 0.f};
  int bnext = 1;  // LDS buffer of step t + 1
  for (int jb = j0; jb < j1; jb += 128) {
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const int js = jb + 32 * ks;
      const int cs = ks & 1, ns = cs ^ 1;
      // A(t+1) from LDS and B(t+1) from the phases, under the MFMAs of t
      read_a(bnext, ns);
      mx_rho_bfrag(bq[ns], raw[(ks + 1) % EPF]);
#pragma unroll
      for (int r = 0; r < NRT2; ++r) {
        acc[0][r] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[cs][r], bq[cs][0], (ks || !F64) ? acc[0][r] : zero, 0, 0, 0);
        acc[1][r] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[cs][r], bq[cs][2], (ks || !F64) ? acc[1][r] : zero, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < NRT2; ++r) {
        acc[0][r] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[cs][r], bq[cs][0], acc[0][r], 0, 0, 0);
        acc[1][r] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[cs][r], bq[cs][2], acc[1][r], 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < NRT2; ++r) {
        acc[0][r] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[cs][r], bq[cs][1], acc[0][r], 0, 0, 0);
        acc[1][r] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[cs][r], bq[cs][3], acc[1][r], 0, 0, 0);
      }
#pragma unroll
      for (int i = 0; i < 6 * NRT2; ++i) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
        __builtin_amdgcn_sched_group_barrier(0x002, 2, 0);  // VALU
        if (i < 2 * NRT2) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // LDS read
      }
      // phases EPF steps ahead; A(t+2) into the LDS buffer A(t-1) left
      // (read before the last barrier), then A(t+3) fetched
      static_assert(4 % EPF == 0, "ring index by the unrolled step");
      mx_rho_phases(raw[(ks + 1) % EPF], e2row, e3row, min(js + 32 * (1 + EPF), jlast), kgrp);
      const int bstage = bnext == NBUF - 1 ? 0 : bnext + 1;
      stage_a(bstage);
      fetch_a(min(js + 96, jlast));
      bnext = bstage;
      __syncthreads();
    }
    if (F64) {
#pragma unroll
      for (int y = 0; y < 2; ++y)
#pragma unroll
        for (int r = 0; r < NRT2; ++r)
#pragma unroll
          for (int e = 0; e < 4; ++e) accd[y][r][e] += (double)acc[y][r][e];
    }
  }
  auto P = [&](int y, int r, int e) -> double {
    if constexpr (F64) return accd[y][r][e];
    else return (double)acc[y][r][e];
  };
  if (ctile * 16 >= CH || c >= CH) return;
  // FP32 chains: the split's sums are FP32 values, stored as such
  using PT = RhoPart<F64>;
  PT* o = out + (size_t)split * (NRT2 * 8) * (2 * C23p);
  const int cp = C23 - 1 - c;
#pragma unroll
  for (int r = 0; r < NRT2; ++r) {
    if (r >= nrt_used) continue;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int p = r * 8 + 2 * kgrp + h;
      const double Prr = P(0, r, 2 * h), Pir = P(0, r, 2 * h + 1);
      const double Pri = P(1, r, 2 * h), Pii = P(1, r, 2 * h + 1);
      PT* op = o + (size_t)p * (2 * C23p);
      op[c] = (PT)(Prr - Pii);
      op[C23p + c] = (PT)(Pri + Pir);
      if (cp != c) {  // the centre column (0, 0) pairs with itself
        op[cp] = (PT)(Prr + Pii);
        op[C23p + cp] = (PT)(Pir - Pri);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Structure factors of the dipole-only calls (one FP32 chain per split) on
// v_mfma_f32_32x32x16_f16.  A wave owns 32 columns c of the conjugate-paired
// half and all rows (NRT32 tiles of 32 (pair, re|im) rows, the row layout of
// mx_rho_gen); per 16-atom K step it reads one A fragment pair (hi, lo) per
// row tile from LDS, one row tile ahead of its MFMAs, and makes the B
// fragments (Re hi, Re lo, Im hi, Im lo) of its column for its 8 atoms from
// the E2, E3 phases, one half step ahead; six MFMAs per row tile.  Against
// mx_rho_gen's 16-column waves this halves the LDS bytes of A per flop, and
// a 32x32x16 product reads half the operand registers of a 16x16x32 one per
// flop.  A fragments (mx_rho_afrag32): [32-atom step g][half s][rt][hi|lo]
// [lane] of 8 halves, row rt*32 + lane%32, atoms 32 g + 16 s + 8 (lane/32) +
// 0..7.  A fragments of a 32-atom step go through a ring of three LDS
// buffers as in mx_rho_gen (staged two steps ahead, one barrier per step).
// Output: the FP32 partials of mx_rho_gen<NRT2 = 2 NRT32, false>.
template <int NRT32>
__global__ __launch_bounds__(MDFF_BLOCK) void mx_rho_afrag32(int npad, int N1, int NW,
                                                              const float* __restrict__ Wt,
                                                              const float2* __restrict__ tj1,
                                                              u32x4* __restrict__ afrag,
                                                              unsigned int* __restrict__ gmax,
                                                              int nreset, int jb0, int jb1) {
  if (blockIdx.x == 0)
    for (int t = threadIdx.x; t < nreset; t += blockDim.x) gmax[t] = 0u;
  // the fragments of slots jb0 .. jb1 - 1 (multiples of 32)
  const size_t t = (size_t)(jb0 / 32) * 2 * NRT32 * 64 + (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t total = (size_t)(jb1 / 32) * 2 * NRT32 * 64;
  if (t >= total) return;
  const int l = (int)(t & 63);
  const size_t f = t >> 6;  // (g * 2 + s) * NRT32 + rt
  const int rt = (int)(f % NRT32);
  const int hs = (int)(f / NRT32);  // g * 2 + s
  const int r = rt * 32 + (l & 31);
  const int p = r >> 1, x = r & 1;
  const int j = hs * 16 + 8 * (l >> 5);
  half8 h, lo;
  if (p < NW * N1) {
    const int w = p / N1, n1 = p % N1;
    const float4* wp = reinterpret_cast<const float4*>(Wt + (size_t)w * npad + j);
    const float4* ep = reinterpret_cast<const float4*>(tj1 + (size_t)n1 * npad + j);
    const float4 w0 = wp[0], w1 = wp[1];
    const float ww[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float4 e = ep[q];  // atoms j + 2q, j + 2q + 1 as (re, im, re, im)
      _Float16 a, b;
      split16(ww[2 * q] * (x ? e.y : e.x), a, b);
      h[2 * q] = a;
      lo[2 * q] = b;
      split16(ww[2 * q + 1] * (x ? e.w : e.z), a, b);
      h[2 * q + 1] = a;
      lo[2 * q + 1] = b;
    }
  } else {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      h[e] = (_Float16)0.f;
      lo[e] = (_Float16)0.f;
    }
  }
  u32x4* o = afrag + (f * 2) * 64 + l;
  o[0] = __builtin_bit_cast(u32x4, h);
  o[64] = __builtin_bit_cast(u32x4, lo);
}

typedef float floatx16 __attribute__((ext_vector_type(16)));

template <int NRT32>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void mx_rho_w(
    int C23, int CH, int C23p, int ncb, int nblk_used, const int2* __restrict__ splits,
    const u32x4* __restrict__ afrag, int npad, int N3, const float2* __restrict__ tj2,
    const float2* __restrict__ tj3, float* __restrict__ out) {
  constexpr int NBUF = 3;
  constexpr int ABLK = 2 * NRT32 * 2 * 64;  // 16-byte words of A per 32-atom step
  constexpr int APER = (ABLK + 255) / 256;
  constexpr int R = 4;  // phase ring: half steps loaded ahead
  __shared__ u32x4 sA[NBUF][ABLK];
  const int per = gridDim.x / NXCD;
  const int gid = (blockIdx.x % NXCD) * per + blockIdx.x / NXCD;
  if (gid >= nblk_used) return;  // whole block, before any barrier
  const int split = gid / ncb, cblk = gid % ncb;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int half = lane >> 5;  // atoms 8 half .. 8 half + 7 of a 16-atom step
  const int ctile = cblk * 4 + wave;
  const int c = ctile * 32 + (lane & 31);
  const int cl = min(c, CH - 1);
  const int n2c = cl / N3;
  const float4* e2row = reinterpret_cast<const float4*>(tj2 + (size_t)n2c * npad);
  const float4* e3row = reinterpret_cast<const float4*>(tj3 + (size_t)(cl - n2c * N3) * npad);
  const int j0 = splits[split].x;
  const int j1 = splits[split].y;  // j1 - j0: a multiple of 128
  const int jlast = j1 - 32;

  u32x4 apre[APER];
  auto fetch_a = [&](int jb) {
    const u32x4* src = afrag + (size_t)(jb / 32) * ABLK;
#pragma unroll
    for (int u = 0; u < APER; ++u) {
      const int idx = threadIdx.x + u * 256;
      if (ABLK % 256 == 0 || idx < ABLK) apre[u] = src[idx];
    }
  };
  auto stage_a = [&](int buf) {
#pragma unroll
    for (int u = 0; u < APER; ++u) {
      const int idx = threadIdx.x + u * 256;
      if (ABLK % 256 == 0 || idx < ABLK) sA[buf][idx] = apre[u];
    }
  };
  // the E2, E3 phases of the lane's column for its 8 atoms of half step s
  // of the 32-atom step at jb (clamped to the split's last step)
  auto phases = [&](float4 (&d)[8], int jb, int s) {
    const int q = (jb + 16 * s + 8 * half) >> 1;  // float4 = two atoms
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      d[u] = e2row[q + u];
      d[4 + u] = e3row[q + u];
    }
  };
  floatx16 accr[NRT32], acci[NRT32];
#pragma unroll
  for (int r = 0; r < NRT32; ++r)
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      accr[r][e] = 0.f;
      acci[r][e] = 0.f;
    }
  float4 raw[R][8];
  half8 bq[2][4];   // B of the current and the next half step
  half8 ah[2], al[2];  // A of one row tile, by row-tile parity
  // prologue: A of steps 0, 1 in LDS, step 2 fetched; phases of half steps
  // 0 .. R-1; B of half step 0
  fetch_a(j0);
  stage_a(0);
  fetch_a(min(j0 + 32, jlast));
  stage_a(1);
  fetch_a(min(j0 + 64, jlast));
#pragma unroll
  for (int h = 0; h < R; ++h) phases(raw[h], min(j0 + 32 * (h / 2), jlast), h & 1);
  mx_rho_bfrag(bq[0], raw[0]);
  __syncthreads();
  ah[0] = __builtin_bit_cast(half8, sA[0][lane]);
  al[0] = __builtin_bit_cast(half8, sA[0][64 + lane]);
  int bcur = 0;  // LDS buffer of the current 32-atom step
  for (int jb = j0; jb < j1; jb += 128) {
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const int js = jb + 32 * ks;
      const int bnext = bcur == NBUF - 1 ? 0 : bcur + 1;
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int h = 2 * ks + s;  // half step within the unrolled block
        const int cb = h & 1;
        // per row tile, in program order (the sections are scheduling
        // fences): the next row tile's A from LDS, a quarter of the next
        // half step's B fragments (from phases loaded R - 1 half steps
        // ago), then this row tile's six MFMAs with that VALU work between
        // them
#pragma unroll
        for (int rt = 0; rt < NRT32; ++rt) {
          const int q = h * NRT32 + rt;  // row-tile index of the unrolled block
          const int cs = q & 1;
          {
            const int nrt = rt + 1 < NRT32 ? rt + 1 : 0;
            const int ns = rt + 1 < NRT32 ? s : s ^ 1;
            const int nb = (rt + 1 < NRT32 || s == 0) ? bcur : bnext;
            const u32x4* src = &sA[nb][(ns * NRT32 + nrt) * 128];
            ah[cs ^ 1] = __builtin_bit_cast(half8, src[lane]);
            al[cs ^ 1] = __builtin_bit_cast(half8, src[64 + lane]);
          }
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if ((u < NRT32 ? u : NRT32 - 1) == rt) mx_rho_bfrag_part(bq[cb ^ 1], raw[(h + 1) % R], u);
          accr[rt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[cs], bq[cb][0], accr[rt], 0, 0, 0);
          acci[rt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[cs], bq[cb][2], acci[rt], 0, 0, 0);
          accr[rt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[cs], bq[cb][0], accr[rt], 0, 0, 0);
          acci[rt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[cs], bq[cb][2], acci[rt], 0, 0, 0);
          accr[rt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[cs], bq[cb][1], accr[rt], 0, 0, 0);
          acci[rt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[cs], bq[cb][3], acci[rt], 0, 0, 0);
          __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);  // LDS read
#pragma unroll
          for (int i = 0; i < 6; ++i) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
            __builtin_amdgcn_sched_group_barrier(0x002, 4, 0);  // VALU
          }
          __builtin_amdgcn_sched_barrier(0);
        }
        // phases R half steps ahead, into the ring slot this half step's B
        // was made from
        {
          const int hn = h + R;  // half step index within the block, may pass 7
          phases(raw[h % R], min(jb + 32 * (hn / 2), jlast), hn & 1);
        }
      }
      // A of step t + 2 into the buffer step t - 1 used (read before the
      // last barrier), then A of step t + 3 fetched
      const int bstage = bnext == NBUF - 1 ? 0 : bnext + 1;
      stage_a(bstage);
      fetch_a(min(js + 96, jlast));
      bcur = bnext;
      __syncthreads();
    }
  }
  if (c >= CH) return;
  // D layout of v_mfma_f32_32x32x16_f16: element e of a lane is row
  // 8 (e/4) + 4 (lane/32) + e%4, column lane%32; rows 2p + x, x = row
  // parity = e%2, so elements e, e + 1 (e even) are (re, im) of pair p
  float* o = out + (size_t)split * (NRT32 * 16) * (2 * C23p);
  const int cp = C23 - 1 - c;
#pragma unroll
  for (int r = 0; r < NRT32; ++r)
#pragma unroll
    for (int e = 0; e < 16; e += 2) {
      const int p = r * 16 + 4 * (e >> 2) + 2 * half + ((e & 3) >> 1);
      const float Prr = accr[r][e], Pir = accr[r][e + 1];
      const float Pri = acci[r][e], Pii = acci[r][e + 1];
      float* op = o + (size_t)p * (2 * C23p);
      op[c] = Prr - Pii;
      op[C23p + c] = Pri + Pir;
      if (cp != c) {  // the centre column (0, 0) pairs with itself
        op[cp] = Prr + Pii;
        op[C23p + cp] = Pir - Pri;
      }
    }
}

// ---------------------------------------------------------------------------
// Sum of the splits in FP64, rho(k) = M_q + i sum_c n_c M_c, g = A_k rho on
// the half-space mesh (zero elsewhere), energy and stress rows
// (src/field.f90:2742-2745, 2803-2815), and max |g|.
// Grid (column blocks, N1, configurations): one n1 plane of one
// configuration per block row; g and the maxima are per configuration.
template <typename PT, int NW>
__global__ __launch_bounds__(MDFF_BLOCK) void mx_rho_reduce(
    int N1, int N2, int N3, int C23, int C23p, int chg, int dip, int nsplit, int split_rows,
    const PT* __restrict__ part, Cell cell_, double alpha_, const MxSeg* __restrict__ seg,
    int do_energy, double2* __restrict__ g, unsigned int* __restrict__ gmax,
    double* __restrict__ pout, const double2* __restrict__ mq, double2* __restrict__ mq_out) {
  __shared__ double red[4 * 7];
  __shared__ float wmax[MDFF_BLOCK / 64];
  const int z = blockIdx.z;
  const Cell& cell = seg ? seg[z].cell : cell_;
  const double alpha = seg ? seg[z].alpha : alpha_;
  const int sp0 = seg ? seg[z].split0 : 0;
  if (seg) nsplit = seg[z].nsplit;
  g += (size_t)z * N1 * C23;
  gmax += (size_t)z * N1;
  if (mq) mq += (size_t)z * N1 * C23;
  if (mq_out) mq_out += (size_t)z * N1 * C23;
  const int n1 = blockIdx.y;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int idx = n1 * C23 + c;
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  float gabs = 0.f;
  if (c < C23) {
    const int ky = (N2 - 1) / 2, kz = (N3 - 1) / 2;
    const int n2 = c / N3 - ky, n3 = c % N3 - kz;
    double Mr[4] = {0, 0, 0, 0}, Mi[4] = {0, 0, 0, 0};
    const size_t sstr = (size_t)split_rows * (2 * C23p);
    const PT* p0 = part + (size_t)sp0 * sstr + (size_t)n1 * (2 * C23p) + c;
    const size_t wstr = (size_t)N1 * (2 * C23p);
    // per weight class, the splits in order; the loads of all classes of
    // four splits issue together
    int sp = 0;
    for (; sp + 4 <= nsplit; sp += 4) {
      PT vr[NW][4], vi[NW][4];
#pragma unroll
      for (int w = 0; w < NW; ++w)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const PT* p = p0 + (size_t)w * wstr + (size_t)(sp + u) * sstr;
          vr[w][u] = p[0];
          vi[w][u] = p[C23p];
        }
#pragma unroll
      for (int w = 0; w < NW; ++w)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          Mr[w] += vr[w][u];
          Mi[w] += vi[w][u];
        }
    }
    for (; sp < nsplit; ++sp)
#pragma unroll
      for (int w = 0; w < NW; ++w) {
        const PT* p = p0 + (size_t)w * wstr + (size_t)sp * sstr;
        Mr[w] += p[0];
        Mi[w] += p[C23p];
      }
    double rR = 0.0, rI = 0.0;
    int w = 0;
    if (chg) {
      rR = Mr[0];
      rI = Mi[0];
      w = 1;
      if (mq_out) mq_out[idx] = make_double2(rR, rI);
    } else if (mq) {
      // the charge sums of this evaluation, cached by its static call
      const double2 v = mq[idx];
      rR = v.x;
      rI = v.y;
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
// B fragments of the apply GEMM over the conjugate-paired half of the
// columns.  With c' = C23 - 1 - c the partner of c (n2, n3 negated),
// P_c' = conj(P_c) for the phases P = E2 E3, so the sum over all columns of
// conj(P_c) G_c equals the sum over c < CH of
//   Pr (G_c + G_c') + i Pi (G_c' - G_c) = Pr U + i Pi V
// (the self-paired centre column (0, 0), where Pi = 0, takes U = G_c).
// cols = (part, w, n1): col = part*(W*N1) + w*N1 + n1, NCT tiles of 16;
// K: k' = 2c + t over CHp columns c:
//   Re col: B[2c] = Ur, B[2c+1] = -Vi;  Im col: B[2c] = Ui, B[2c+1] = Vr
// with G = APPLY_SCALE 2^-e(w,n1) w(n) g(n).  Layout [ks][hi|lo][ct][lane] of
// half8, so that the fragments of one K step are contiguous; one such block
// per configuration (grid.y), and sinv[w][n1] = 2^e(w,n1) / APPLY_SCALE.
__global__ __launch_bounds__(MDFF_BLOCK) void mx_gprep(int N1, int N3, int C23, int CH, int CHp,
                                                        int W, int NCT, const double2* __restrict__ g,
                                                        const unsigned int* __restrict__ gmax,
                                                        int ky, int kz, _Float16* __restrict__ bf,
                                                        double* __restrict__ sinv) {
  const int KS = (2 * CHp) / 32;
  const size_t total = (size_t)NCT * KS * 64;
  // two threads per fragment lane, 4 of its 8 elements each
  const size_t t2 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int hh = (int)(t2 & 1);
  const size_t t = t2 >> 1;
  if (t >= total) return;
  // configuration blockIdx.y: its g planes, maxima, fragments and scales
  const int z = blockIdx.y;
  g += (size_t)z * N1 * C23;
  gmax += (size_t)z * N1;
  bf += (size_t)z * total * 16;
  sinv += (size_t)z * W * N1;
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
    if (part == 0 && ks == 0 && lane < 16 && hh == 0) sinv[w * N1 + n1] = 1.0 / (double)cs;
  }
  half4 h, l;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int kp = ks * 32 + 8 * (lane >> 4) + 4 * hh + e;
    const int c = kp >> 1, tt = kp & 1;
    float v = 0.f;
    if (colg < ncols && c < CH) {
      const int n2 = c / N3 - ky, n3 = c % N3 - kz;
      const double2 gv = g[(size_t)n1 * C23 + c];
      const double sc = (double)cs * apply_weight(w, n2, n3);
      double ur = gv.x * sc, ui = gv.y * sc, vr = -ur, vi = -ui;
      if (2 * c + 1 != C23) {
        // partner column: n2, n3 negated (odd weights change sign)
        const double2 gp = g[(size_t)n1 * C23 + (C23 - 1 - c)];
        const double scp = (double)cs * apply_weight(w, -n2, -n3);
        ur += gp.x * scp;
        ui += gp.y * scp;
        vr += gp.x * scp;
        vi += gp.y * scp;
      }
      if (part == 0)
        v = (float)(tt ? -vi : ur);
      else
        v = (float)(tt ? vr : ui);
    }
    _Float16 hh, ll;
    split16(v, hh, ll);
    h[e] = hh;
    l[e] = ll;
  }
  half4* o = (half4*)bf + 2 * ((size_t)ks * 2 * NCT * 64 + ct * 64 + lane) + hh;
  o[0] = h;
  o[2 * NCT * 64] = l;
}

// ---------------------------------------------------------------------------
// Apply GEMM: rows = atoms (the row blocks mb0 .. mb0 + mblocks - 1), K =
// 2 CHp (the conjugate-paired half of the columns, operand of mx_gprep), NCT
// column tiles.  A workgroup of
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

template <int NCT, int RT>
__global__ __launch_bounds__(64 * APP_WAVES) void mx_apply_gemm(
    int npad, int KS, int S, int ksper, int mb0, int mblocks, const int2* __restrict__ colofs,
    const float2* __restrict__ tj2, const float2* __restrict__ tj3,
    const _Float16* __restrict__ bf, const int* __restrict__ kseg, float* __restrict__ Xp) {
  __shared__ u32x4 sB[2][2 * NCT * 64];
  const int per = NXCD / S;
  const int xcd = blockIdx.x % NXCD;
  const int sl = xcd / per;
  const int mbl = (blockIdx.x / NXCD) * per + (xcd % per);
  if (mbl >= mblocks) return;  // whole block, before any barrier
  const int mb = mb0 + mbl;    // row blocks mb0 .. mb0 + mblocks - 1
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
  // B fragments of the row block's configuration (rows never straddle one)
  const int zc = kseg ? kseg[(mb * APP_WAVES * RT * 16) >> 5] : 0;
  const u32x4* bsrc = reinterpret_cast<const u32x4*>(bf) + (size_t)zc * KS * NB;
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
  // atoms of this lane, through the column table (columns past CH read the
  // last column: the operand rows they meet are zero)
  float2 pe2[RT][4], pe3[RT][4];
  auto load_e = [&](int ks) {
    const int4* co = reinterpret_cast<const int4*>(colofs + ks * 16 + 4 * kgrp);
    const int4 c01 = co[0], c23 = co[1];
    const int o2[4] = {c01.x, c01.z, c23.x, c23.z};
    const int o3[4] = {c01.y, c01.w, c23.y, c23.w};
#pragma unroll
    for (int e = 0; e < 4; ++e)
#pragma unroll
      for (int r = 0; r < RT; ++r) {
        pe2[r][e] = tj2[o2[e] + ibase + 16 * r];
        pe3[r][e] = tj3[o3[e] + ibase + 16 * r];
      }
  };
  // A fragments: A[i][2c] = Re P_i[c], A[i][2c+1] = Im P_i[c]
  auto make_frags = [&](half8 (&h8)[RT], half8 (&l8)[RT]) {
#pragma unroll
    for (int e = 0; e < 4; ++e)
#pragma unroll
      for (int r = 0; r < RT; ++r) {
        const float2 e2 = pe2[r][e], e3 = pe3[r][e];
        half2v h, l;
        split16x2(e2.x * e3.x - e2.y * e3.y, e2.x * e3.y + e2.y * e3.x, h, l);
        h8[r][2 * e] = h.x;
        h8[r][2 * e + 1] = h.y;
        l8[r][2 * e] = l.x;
        l8[r][2 * e + 1] = l.y;
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
      // the next pair's LDS reads in flight under this pair's MFMAs
      __builtin_amdgcn_sched_barrier(0);
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
constexpr int EPI_CHUNKS = 4;  // n1 ranges per slot, one wave each

template <int S>
__global__ __launch_bounds__(MDFF_BLOCK) void mx_apply_epi(int n, int npad, int j0, int N1, int W,
                                                            int ncolp,
                                                            const MxSeg* __restrict__ seg,
                                                            const int* __restrict__ kseg,
                                                            const float* __restrict__ Xp,
                                                            const float2* __restrict__ tj1,
                                                            const double* __restrict__ sinv,
                                                            double* __restrict__ out) {
  // grid (64-slot blocks, W): wave c of the block folds the n1 range c of
  // 64 consecutive slots for weight class w (more loads in flight than one
  // thread per slot), then the ranges are summed in order
  static_assert(MDFF_BLOCK == 64 * EPI_CHUNKS, "one wave per n1 range");
  __shared__ double part[EPI_CHUNKS - 1][5][64];
  const int w = blockIdx.y;
  const int lane = threadIdx.x & 63, ch = threadIdx.x >> 6;
  const int jp = j0 + blockIdx.x * 64 + lane;  // slots j0 .. of the grid
  int z = 0, i = -1;
  if (jp < npad) i = slot_atom(jp, n, seg, kseg, z);
  double sr = 0, si = 0, pr = 0, pi = 0, qr = 0;
  if (i >= 0) {
    const int per = (N1 + EPI_CHUNKS - 1) / EPI_CHUNKS;
    const int m0 = ch * per, m1 = min(N1, m0 + per);
    const double* sv = sinv + (size_t)(z * W + w) * N1;
    const size_t sstride = (size_t)npad * ncolp;
    const float* xre = Xp + (size_t)(w * N1) * npad + jp;
    const float* xim = Xp + (size_t)(W * N1 + w * N1) * npad + jp;
#pragma unroll 7
    for (int m = m0; m < m1; ++m) {
      double xr = 0.0, xi = 0.0;
#pragma unroll
      for (int sl = 0; sl < S; ++sl) {
        xr += xre[sl * sstride + (size_t)m * npad];
        xi += xim[sl * sstride + (size_t)m * npad];
      }
      xr *= sv[m];
      xi *= sv[m];
      const float2 e = tj1[(size_t)m * npad + jp];
      const double er = e.x, ei = -(double)e.y;  // conj(E1)
      const double vr = er * xr - ei * xi, vi = er * xi + ei * xr;
      sr += vr;
      si += vi;
      pr += m * vr;
      pi += m * vi;
      qr += (double)m * m * vr;
    }
  }
  if (ch > 0) {
    part[ch - 1][0][lane] = sr;
    part[ch - 1][1][lane] = si;
    part[ch - 1][2][lane] = pr;
    part[ch - 1][3][lane] = pi;
    part[ch - 1][4][lane] = qr;
  }
  __syncthreads();
  if (ch > 0 || i < 0) return;
  for (int c = 0; c < EPI_CHUNKS - 1; ++c) {
    sr += part[c][0][lane];
    si += part[c][1][lane];
    pr += part[c][2][lane];
    pi += part[c][3][lane];
    qr += part[c][4][lane];
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
// ES_IND call of the conjugate-gradient SCF, one configuration, field only:
// mx_apply_epi's fold of the three field classes (U1 = Im S'_0, U2 = Im S_1,
// U3 = Im S_2) for the 64 atoms of a block, then, for those atoms, the
// combine of k_combine64<ES_IND> in conjugate-gradient mode (no surface
// term): E = E_dir + E_rec(U) + 2 selfa2 p, Ap = alpha^-1 p - E, and the
// block partial of p.Ap at ca.rpart[block].  The operations and their order
// are those of the two kernels it replaces.
template <int S>
__global__ __launch_bounds__(MDFF_BLOCK) void mx_apply_epi_cg(int n, int npad, int j0, int N1,
                                                               int ncolp,
                                                               const float* __restrict__ Xp,
                                                               const float2* __restrict__ tj1,
                                                               const double* __restrict__ sinv,
                                                               CombineArgs ca) {
  static_assert(MDFF_BLOCK == 64 * EPI_CHUNKS, "one wave per n1 range");
  constexpr int W = 3;
  __shared__ double part[EPI_CHUNKS - 1][W][64];
  __shared__ double red[4];
  const int lane = threadIdx.x & 63, ch = threadIdx.x >> 6;
  const int jp = j0 + blockIdx.x * 64 + lane;  // slots j0 .. of the grid
  const int i = jp < n ? jp : -1;              // one configuration: slot = atom
  double u[W] = {0.0, 0.0, 0.0};
  if (i >= 0) {
    const int per = (N1 + EPI_CHUNKS - 1) / EPI_CHUNKS;
    const int m0 = ch * per, m1 = min(N1, m0 + per);
    const size_t sstride = (size_t)npad * ncolp;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const double* sv = sinv + (size_t)w * N1;
      const float* xre = Xp + (size_t)(w * N1) * npad + jp;
      const float* xim = Xp + (size_t)(W * N1 + w * N1) * npad + jp;
#pragma unroll 7
      for (int m = m0; m < m1; ++m) {
        double xr = 0.0, xi = 0.0;
#pragma unroll
        for (int sl = 0; sl < S; ++sl) {
          xr += xre[sl * sstride + (size_t)m * npad];
          xi += xim[sl * sstride + (size_t)m * npad];
        }
        xr *= sv[m];
        xi *= sv[m];
        const float2 e = tj1[(size_t)m * npad + jp];
        const double er = e.x, ei = -(double)e.y;  // conj(E1)
        const double vi = er * xi + ei * xr;
        u[w] += (w == 0) ? m * vi : vi;
      }
    }
  }
  if (ch > 0)
    for (int w = 0; w < W; ++w) part[ch - 1][w][lane] = u[w];
  __syncthreads();
  double pap[1] = {0.0};
  if (ch == 0 && i >= 0) {
    for (int c = 0; c < EPI_CHUNKS - 1; ++c)
      for (int w = 0; w < W; ++w) u[w] += part[c][w][lane];
    const double* B = ca.cell.B;
    const double fe = -2.0 * ca.fpi_V * (2.0 * MDFF_PI);
    const double m3[3] = {ca.mu[i], ca.mu[i + n], ca.mu[i + 2 * n]};
    double ef[3];
    for (int r = 0; r < 3; ++r) {
      const double er = fe * (B[r] * u[0] + B[r + 3] * u[1] + B[r + 6] * u[2]);
      ef[r] = (ca.ef_dir[i + r * n] + er) + 2.0 * ca.selfa2 * m3[r];
    }
    const int t = ca.typ[i];
    double ap[3] = {0, 0, 0};
    if (ca.T->polar[t]) {
      const double* ip = ca.T->invpol[t];
      for (int r = 0; r < 3; ++r) {
        ap[r] = ip[r] * m3[0] + ip[r + 3] * m3[1] + ip[r + 6] * m3[2] - ef[r];
        pap[0] += m3[r] * ap[r];
      }
    }
    for (int r = 0; r < 3; ++r) ca.cg_Ap[i + r * n] = ap[r];
  }
  block_sum<1>(pap, red);
  if (threadIdx.x == 0) ca.rpart[blockIdx.x] = pap[0];
}

// ---------------------------------------------------------------------------
// Batched path: reciprocal field and force of atom i from the U, V sums of
// mx_apply_epi with its configuration's lattice (as k_combine64 forms them):
//   Erec = -2 (4 pi/V) (2 pi) B U,
//   Frec = q Erec + 2 (4 pi/V) (2 pi)^2 B V B^T mu.
__global__ __launch_bounds__(MDFF_BLOCK) void mx_batch_fields(int N, const int* __restrict__ cfg,
                                                               const MxSeg* __restrict__ seg,
                                                               const double* __restrict__ ap,
                                                               const double* __restrict__ q,
                                                               const double* __restrict__ mu,
                                                               int force, double* __restrict__ Erec,
                                                               double* __restrict__ Frec) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const Cell& cell = seg[cfg[i]].cell;
  const double* B = cell.B;
  const double tpi = 2.0 * MDFF_PI;
  const double fpi_V = 4.0 * MDFF_PI / cell.omega;
  const double fe = -2.0 * fpi_V * tpi;
  const double U[3] = {ap[i], ap[i + N], ap[i + 2 * (size_t)N]};
  double er[3];
  for (int r = 0; r < 3; ++r) {
    er[r] = fe * (B[r] * U[0] + B[r + 3] * U[1] + B[r + 6] * U[2]);
    Erec[i + (size_t)r * N] = er[r];
  }
  if (!force) return;
  double V[6];
  for (int c = 0; c < 6; ++c) V[c] = ap[i + (size_t)(3 + c) * N];
  const double Vm[3][3] = {{V[0], V[3], V[4]}, {V[3], V[1], V[5]}, {V[4], V[5], V[2]}};
  const double fg = 2.0 * fpi_V * tpi * tpi;
  const double m[3] = {mu[i], mu[i + N], mu[i + 2 * (size_t)N]};
  const double qi = q[i];
  for (int r = 0; r < 3; ++r) {
    double fr = qi * er[r];
    for (int c = 0; c < 3; ++c) {
      double G = 0.0;
      for (int x = 0; x < 3; ++x)
        for (int y = 0; y < 3; ++y) G += B[r + 3 * x] * Vm[x][y] * B[c + 3 * y];
      fr += fg * G * m[c];
    }
    Frec[i + (size_t)r * N] = fr;
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
constexpr int APP_NCT_S = 4, APP_RT_S = 1;   // small meshes: 2 W N1 <= 64
constexpr int APP_NCT_F = 10, APP_RT_F = 1;  // field calls (W = 3)
constexpr int APP_NCT_E = 19, APP_RT_E = 1;  // final call with EFG (W = 6)
constexpr int APP_ROWS_MAX = APP_WAVES * 2 * 16;
// structure-factor GEMM shapes (row tiles of (pair, re|im) rows, row tiles
// per wave): charges only, dipoles only, charges and dipoles, N1 <= 25
constexpr int RHO_NRT2_MAX = 13;  // 2 x 4 weights x 25 n1 rows

static void mesh(MixedState* m, const int kmax[3]) {
  m->kx = kmax[0];
  m->ky = kmax[1];
  m->kz = kmax[2];
  m->N1 = m->kx + 1;
  m->N2 = 2 * m->ky + 1;
  m->N3 = 2 * m->kz + 1;
  m->C23 = m->N2 * m->N3;
  m->C23p = (m->C23 + 63) / 64 * 64;  // row stride of the structure-factor partials
  m->CH = (m->C23 + 1) / 2;
  m->CHp = (m->CH + 15) / 16 * 16;
}

// buffers for npad slots, nseg configurations and nsplit splits
static bool allocate(MixedState* m) {
  const size_t np = m->npad, ns = m->nseg;
  mx_alloc(&m->tj1, (size_t)m->N1 * np);
  mx_alloc(&m->tj2, (size_t)m->N2 * np);
  mx_alloc(&m->tj3, (size_t)m->N3 * np);
  mx_alloc(&m->g, ns * m->N1 * m->C23);
  mx_alloc(&m->mq, ns * m->N1 * m->C23);
  m->mq_epoch = -1;
  mx_alloc(&m->gmax, ns * m->N1);
  mx_alloc(&m->sinv, ns * 6 * m->N1);
  m->rpart_cap = (size_t)m->nsplit * RHO_NRT2_MAX * 8 * 2 * m->C23p;
  mx_alloc(&m->rpart, m->rpart_cap);
  mx_alloc(&m->Wt, (size_t)4 * np);
  mx_alloc(&m->afrag, (np / 32) * RHO_NRT2_MAX * 2 * 64);
  m->tables_eval = -1;
  const int KS = 2 * m->CHp / 32;
  m->bfrag_cap = ns * 2 * APP_NCT_E * KS * 64 * 8;
  mx_alloc(&m->bfrag, m->bfrag_cap);
  m->x_slices = m->app_slices;
  m->X_cap = (size_t)m->x_slices * np * APP_NCT_E * 16;
  mx_alloc(&m->X, m->X_cap);
  // rows a partial call leaves alone hold finite values
  if (m->X && hipMemset(m->X, 0, m->X_cap * sizeof(float)) != hipSuccess) return false;
  mx_alloc(&m->d_split, (size_t)m->nsplit);
  // per column c < CHp of the paired half: element offsets of its E2 and E3
  // rows, (n2 npad, n3 npad), with c clamped to CH - 1 (padding columns
  // multiply the zero rows of the apply operand)
  {
    std::vector<int2> co(m->CHp);
    for (int c = 0; c < m->CHp; ++c) {
      const int cc = std::min(c, m->CH - 1);
      co[c] = make_int2((cc / m->N3) * (int)np, (cc % m->N3) * (int)np);
    }
    mx_alloc(&m->d_colofs, co.size());
    if (!m->d_colofs ||
        hipMemcpy(m->d_colofs, co.data(), co.size() * sizeof(int2), hipMemcpyHostToDevice) !=
            hipSuccess)
      return false;
  }
  return m->tj1 && m->tj2 && m->tj3 && m->g && m->mq && m->gmax && m->sinv && m->rpart && m->bfrag &&
         m->X && m->Wt && m->afrag && m->d_split;
}

// structure-factor splits of split_atoms slots over the segments (their
// split0 / nsplit are set), or over [0, npad) without a segment table
static std::vector<int2> make_splits(MixedState* m) {
  std::vector<int2> sp;
  const int sa = m->split_atoms;
  if (m->h_seg.empty()) {
    for (int j = 0; j < m->npad; j += sa) sp.push_back(make_int2(j, std::min(m->npad, j + sa)));
    return sp;
  }
  for (MxSeg& sg : m->h_seg) {
    sg.split0 = (int)sp.size();
    for (int j = 0; j < sg.np; j += sa)
      sp.push_back(make_int2(sg.poff + j, sg.poff + std::min(sg.np, j + sa)));
    sg.nsplit = (int)sp.size() - sg.split0;
  }
  return sp;
}

// The slot ranges of the dipole-only structure factors and of the field
// calls' rows (whole 128-slot blocks of one configuration) and the splits of
// the first, from dj0.
static bool make_ranges(MixedState* m) {
  auto blk = [&](int a, int b, int& j0, int& j1) {
    if (b < 0 || !m->h_seg.empty()) {
      j0 = 0;
      j1 = m->npad;
      return;
    }
    j0 = std::min(m->npad, a / 128 * 128);
    j1 = std::min(m->npad, (b + 127) / 128 * 128);
    if (j1 <= j0) j1 = std::min(m->npad, j0 + 128);
    if (j1 <= j0) j0 = j1 - 128;
  };
  blk(m->dip0, m->dip1, m->dj0, m->dj1);
  blk(m->fld0, m->fld1, m->fj0, m->fj1);
  // as many splits as give one workgroup of mx_rho_w (ncb column blocks per
  // split) to each CU in one round, in whole 128-slot steps
  static int cus = 0;
  if (cus <= 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
  }
  const int ncb = ((m->CH + 31) / 32 + 3) / 4;
  const int want = std::max(1, cus / std::max(1, ncb));
  int sa = std::max(128, ((m->dj1 - m->dj0 + want - 1) / want + 127) / 128 * 128);
  // no more splits than the partials buffer holds
  const size_t cap_splits = m->rpart_cap / ((size_t)RHO_NRT2_MAX * 8 * 2 * m->C23p);
  if ((size_t)((m->dj1 - m->dj0 + sa - 1) / sa) > cap_splits) sa = std::max(sa, m->split_atoms);
  std::vector<int2> sp;
  for (int j = m->dj0; j < m->dj1; j += sa) sp.push_back(make_int2(j, std::min(m->dj1, j + sa)));
  if ((int)sp.size() > m->nsplit_dip_cap) {
    mx_alloc(&m->d_split_dip, sp.size());
    if (!m->d_split_dip) return false;
    m->nsplit_dip_cap = (int)sp.size();
  }
  m->nsplit_dip = (int)sp.size();
  return hipMemcpy(m->d_split_dip, sp.data(), sp.size() * sizeof(int2),
                   hipMemcpyHostToDevice) == hipSuccess;
}

bool set_active(MixedState* m, int dip0, int dip1, int fld0, int fld1) {
  m->dip0 = dip0;
  m->dip1 = dip1;
  m->fld0 = fld0;
  m->fld1 = fld1;
  return make_ranges(m);
}

static bool sizes(MixedState* m) {
  CtxView v = view(m->ctx);
  mesh(m, v.kmax);
  if (m->N1 > 25) return false;  // 8 N1 rows fit the 13 row tiles of mx_rho_gemm
  m->N = v.n;
  m->nseg = 1;
  // whole row blocks of the apply GEMM (tables are zero past n)
  m->npad = (v.n + APP_ROWS_MAX - 1) / APP_ROWS_MAX * APP_ROWS_MAX;
  const std::vector<int2> sp = make_splits(m);
  m->nsplit = (int)sp.size();
  if (!allocate(m)) return false;
  return hipMemcpy(m->d_split, sp.data(), sp.size() * sizeof(int2), hipMemcpyHostToDevice) ==
             hipSuccess &&
         make_ranges(m);
}

MixedState* create_batch(int N, const int kmax[3], const std::vector<MxSeg>& segs_in) {
  MixedState* m = new MixedState();
  m->ctx = nullptr;
  mesh(m, kmax);
  m->N = N;
  m->nseg = (int)segs_in.size();
  m->h_seg = segs_in;
  std::vector<MxSeg>& segs = m->h_seg;
  int poff = 0;
  for (MxSeg& sg : segs) {
    sg.poff = poff;
    sg.np = std::max(APP_ROWS_MAX, (sg.n + APP_ROWS_MAX - 1) / APP_ROWS_MAX * APP_ROWS_MAX);
    poff += sg.np;
  }
  m->npad = poff;
  const std::vector<int2> sp = make_splits(m);
  m->nsplit = (int)sp.size();
  std::vector<int> kseg(m->npad / 32);
  for (int z = 0; z < m->nseg; ++z)
    for (int k = segs[z].poff / 32; k < (segs[z].poff + segs[z].np) / 32; ++k) kseg[k] = z;
  bool ok = m->N1 <= 25 && allocate(m);
  mx_alloc(&m->d_seg, segs.size());
  mx_alloc(&m->d_kseg, kseg.size());
  mx_alloc(&m->apart, (size_t)9 * N);
  ok = ok && m->d_seg && m->d_kseg && m->apart;
  ok = ok && hipMemcpy(m->d_seg, segs.data(), segs.size() * sizeof(MxSeg),
                       hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMemcpy(m->d_kseg, kseg.data(), kseg.size() * sizeof(int),
                       hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMemcpy(m->d_split, sp.data(), sp.size() * sizeof(int2),
                       hipMemcpyHostToDevice) == hipSuccess;
  if (!ok) {
    destroy(m);
    return nullptr;
  }
  return m;
}

bool set_tuning(MixedState* m, int app_slices, int split_atoms) {
  if (app_slices != 1 && app_slices != 2 && app_slices != 4) return false;
  if (split_atoms < 128 || split_atoms % 128 != 0) return false;
  if (app_slices > m->x_slices) {
    m->x_slices = app_slices;
    m->X_cap = (size_t)m->x_slices * m->npad * APP_NCT_E * 16;
    mx_alloc(&m->X, m->X_cap);
    if (!m->X || hipMemset(m->X, 0, m->X_cap * sizeof(float)) != hipSuccess) return false;
  }
  m->app_slices = app_slices;
  if (split_atoms == m->split_atoms) return true;
  m->split_atoms = split_atoms;
  const std::vector<int2> sp = make_splits(m);
  if ((int)sp.size() > m->nsplit) {
    mx_alloc(&m->d_split, sp.size());
    m->rpart_cap = sp.size() * RHO_NRT2_MAX * 8 * 2 * m->C23p;
    mx_alloc(&m->rpart, m->rpart_cap);
    if (!m->d_split || !m->rpart) return false;
  }
  m->nsplit = (int)sp.size();
  if (hipMemcpy(m->d_split, sp.data(), sp.size() * sizeof(int2), hipMemcpyHostToDevice) !=
      hipSuccess)
    return false;
  if (m->h_seg.empty() && !make_ranges(m)) return false;
  return m->h_seg.empty() ||
         hipMemcpy(m->d_seg, m->h_seg.data(), m->h_seg.size() * sizeof(MxSeg),
                   hipMemcpyHostToDevice) == hipSuccess;
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
  void* p[] = {m->d_colofs, m->tj1,   m->tj2,  m->tj3,  m->rpart,  m->g,       m->mq,
               m->gmax,     m->bfrag, m->X,    m->Wt,   m->afrag,  m->sinv,    m->d_seg,
               m->d_kseg,   m->d_split, m->apart, m->d_split_dip};
  for (void* x : p)
    if (x) (void)hipFree(x);
  delete m;
}

void on_cell(MixedState* m) {
  if (view(m->ctx).P->lcoulomb) sizes(m);
}

// the splits of a structure-factor call
struct Splits {
  const int2* d;
  int n;
};

template <int NRT2, int RTW, bool F64>
static void launch_rho(MixedState* m, hipStream_t st, int nrt_used, Splits sp) {
  constexpr int CPB = 4 / ((NRT2 + RTW - 1) / RTW);
  const int ctiles = (m->CH + 15) / 16;
  const int ncb = (ctiles + CPB - 1) / CPB;
  const int used = ncb * sp.n;
  const int grid = (used + NXCD - 1) / NXCD * NXCD;
  mx_rho_gemm<NRT2, RTW, F64><<<grid, 256, 0, st>>>(m->C23, m->CH, m->C23p, nrt_used, ncb, used,
                                                    sp.d, m->afrag, m->npad, m->N3, m->tj2,
                                                    m->tj3, m->rpart);
}

template <int NRT2, bool F64>
static void launch_rho_gen(MixedState* m, hipStream_t st, int nrt_used, Splits sp) {
  const int ctiles = (m->CH + 15) / 16;
  const int ncb = (ctiles + 3) / 4;
  const int used = ncb * sp.n;
  const int grid = (used + NXCD - 1) / NXCD * NXCD;
  mx_rho_gen<NRT2, F64><<<grid, 256, 0, st>>>(m->C23, m->CH, m->C23p, nrt_used, ncb, used, sp.d,
                                               m->afrag, m->npad, m->N3, m->tj2, m->tj3,
                                               reinterpret_cast<RhoPart<F64>*>(m->rpart));
}

template <int NRT32>
static void launch_rho_w(MixedState* m, hipStream_t st, Splits sp) {
  const int ctiles = (m->CH + 31) / 32;
  const int ncb = (ctiles + 3) / 4;
  const int used = ncb * sp.n;
  const int grid = (used + NXCD - 1) / NXCD * NXCD;
  mx_rho_w<NRT32><<<grid, 256, 0, st>>>(m->C23, m->CH, m->C23p, ncb, used, sp.d, m->afrag,
                                        m->npad, m->N3, m->tj2, m->tj3,
                                        reinterpret_cast<float*>(m->rpart));
}

// rows j0 .. j1 - 1 (whole row blocks)
template <int NCT, int RT>
static void launch_apply(MixedState* m, hipStream_t st, int j0, int j1) {
  const int KS = 2 * m->CHp / 32;
  const int S = m->app_slices;
  const int ksper = (KS + S - 1) / S;
  const int rows = APP_WAVES * RT * 16;
  const int mb0 = j0 / rows;
  const int mblocks = (j1 - j0) / rows;
  const int per = NXCD / S;
  const int grid = (mblocks + per - 1) / per * NXCD;
  mx_apply_gemm<NCT, RT><<<grid, 64 * APP_WAVES, 0, st>>>(m->npad, KS, S, ksper, mb0, mblocks,
                                                         m->d_colofs, m->tj2, m->tj3, m->bfrag,
                                                         m->d_kseg, m->X);
}

// One reciprocal-space call: the context's configuration (energy and stress
// rows when `energy`) or the configurations of a batch (segment table).
struct KsCall {
  hipStream_t st;
  const double* s;      // fractional coordinates, SoA over the caller's atoms
  long long epoch;      // positions' version: the phase tables follow it
  Cell cell;            // one configuration: its lattice and alpha
  double alpha;
  bool energy;
  bool weights_ready = false;  // dipole weights already in Wt (one configuration)
  double* part;         // energy / stress partials and their slots
  double* sc_rec;
};

static bool run_kspace(MixedState* m, const KsCall& a, const double* q, const double* mu, bool chg,
                       bool dip, bool efg, double* apart, const CombineArgs* cg = nullptr) {
  hipStream_t st = a.st;
  // A call with charges and dipoles at positions whose charge-only call has
  // already run (the static call of the same evaluation, same charges) takes
  // the charge structure factors from that call and runs the dipole-only
  // GEMM: the charge sums do not change within an evaluation.
  const bool mq_in = chg && dip && m->mq_epoch == a.epoch && m->mq_q == q;
  const bool mq_save = chg && !dip;
  if (mq_in) chg = false;
  const int n = m->N;
  const MxSeg* seg = m->d_seg;
  // one configuration: the dipole-only structure factors sum over the slots
  // that can carry a dipole, and a call without energy or EFG (the static
  // and SCF field calls) gives the field at the rows that need it only
  const Splits spl = (!seg && !chg && dip) ? Splits{m->d_split_dip, m->nsplit_dip}
                                           : Splits{m->d_split, m->nsplit};
  const bool rows_part = !seg && !a.energy && !efg;
  const int rj0 = rows_part ? m->fj0 : 0, rj1 = rows_part ? m->fj1 : m->npad;
  // phase tables: positions are fixed within an evaluation, built on its
  // first reciprocal call
  if (m->tables_eval != a.epoch) {
    ProfScope _ps(m->ctx, K_MX_PHASE);
    const dim3 gt((m->npad + MDFF_BLOCK - 1) / MDFF_BLOCK,
                  m->kx + 1 + (2 * m->ky + 1) + (2 * m->kz + 1));
    mx_tables<<<gt, MDFF_BLOCK, 0, st>>>(
        n, m->npad, a.s, seg, m->d_kseg, m->kx, m->ky, m->kz, m->tj1, m->tj2, m->tj3);
    m->tables_eval = a.epoch;
  }
  const int NW = (chg ? 1 : 0) + (dip ? 3 : 0);
  const int NRT2 = (2 * NW * m->N1 + 15) / 16;  // row tiles with data
  int NRT2l = (NRT2 <= 4) ? 4 : (NRT2 <= 10) ? 10 : 13;
  // dipole-only calls: 32x32 tiles (mx_rho_w), NRT32 tiles of 32 rows
  const int NRT32 = (2 * NW * m->N1 + 31) / 32;
  const bool wide = !chg && dip && NRT32 <= 5;
  if (wide) NRT2l = 2 * NRT32;  // rows per split in the partials: NRT2l * 8 pairs
  {
    ProfScope _ps(m->ctx, K_MX_WEIGHTS);
    if (!(a.weights_ready && !chg && dip && !seg))
      mx_weights<<<(m->npad + MDFF_BLOCK - 1) / MDFF_BLOCK, MDFF_BLOCK, 0, st>>>(
          n, m->npad, chg, dip, q, mu, a.cell, seg, m->d_kseg, m->Wt);
    const size_t at = (size_t)(m->npad / 32) * NRT2l * 64;
    const unsigned ga = (unsigned)((at + MDFF_BLOCK - 1) / MDFF_BLOCK);
    const int nreset = m->nseg * m->N1;
    if (wide) {
      // the fragments of the slots the structure-factor splits cover
      const int jb0 = (spl.d == m->d_split_dip) ? m->dj0 : 0;
      const int jb1 = (spl.d == m->d_split_dip) ? m->dj1 : m->npad;
      const size_t atw = (size_t)((jb1 - jb0) / 32) * NRT2l * 64;
      const unsigned gw = (unsigned)std::max<size_t>(1, (atw + MDFF_BLOCK - 1) / MDFF_BLOCK);
#define MX_AF32(K)                                                                           \
  mx_rho_afrag32<K><<<gw, MDFF_BLOCK, 0, st>>>(m->npad, m->N1, NW, m->Wt, m->tj1, m->afrag, \
                                               m->gmax, nreset, jb0, jb1)
      switch (NRT32) {
        case 1: MX_AF32(1); break;
        case 2: MX_AF32(2); break;
        case 3: MX_AF32(3); break;
        case 4: MX_AF32(4); break;
        default: MX_AF32(5); break;
      }
#undef MX_AF32
    } else if (NRT2l == 4)
      mx_rho_afrag<4><<<ga, MDFF_BLOCK, 0, st>>>(m->npad, m->N1, NW, m->Wt, m->tj1, m->afrag,
                                                 m->gmax, nreset);
    else if (NRT2l == 10)
      mx_rho_afrag<10><<<ga, MDFF_BLOCK, 0, st>>>(m->npad, m->N1, NW, m->Wt, m->tj1, m->afrag,
                                                  m->gmax, nreset);
    else
      mx_rho_afrag<13><<<ga, MDFF_BLOCK, 0, st>>>(m->npad, m->N1, NW, m->Wt, m->tj1, m->afrag,
                                                  m->gmax, nreset);
  }
  {
    ProfScope _ps(m->ctx, K_MX_RHO);
    // FP32 chains over a whole split for dipole weights alone; with charges
    // (neutral sums that cancel at small |k|) the chains are flushed to FP64
    // one wave per SIMD over all row tiles where the registers allow (FP32
    // chains, or 4 row tiles), row groups with FP64 flushes otherwise
    if (wide) {
      switch (NRT32) {
        case 1: launch_rho_w<1>(m, st, spl); break;
        case 2: launch_rho_w<2>(m, st, spl); break;
        case 3: launch_rho_w<3>(m, st, spl); break;
        case 4: launch_rho_w<4>(m, st, spl); break;
        default: launch_rho_w<5>(m, st, spl); break;
      }
    } else if (NRT2l == 4)
      chg ? launch_rho_gen<4, true>(m, st, NRT2, spl) : launch_rho_gen<4, false>(m, st, NRT2, spl);
    else if (NRT2l == 10)
      chg ? launch_rho<10, 5, true>(m, st, NRT2, spl)
          : launch_rho_gen<10, false>(m, st, NRT2, spl);
    else
      launch_rho<13, 7, true>(m, st, NRT2, spl);
  }
  const dim3 gred((m->C23 + MDFF_BLOCK - 1) / MDFF_BLOCK, m->N1, m->nseg);
  const int gk = gred.x * gred.y;
  {
    ProfScope _ps(m->ctx, K_MX_RHO_RED);
    // FP32 split sums from the FP32-chain kernels (dipole weights alone)
    const double* rp = m->rpart;
    const float* rpf = reinterpret_cast<const float*>(m->rpart);
#define MX_RED(PT, NWC, P)                                                                     \
  mx_rho_reduce<PT, NWC><<<gred, MDFF_BLOCK, 0, st>>>(                                          \
      m->N1, m->N2, m->N3, m->C23, m->C23p, chg, dip, spl.n, NRT2l * 8, P, a.cell, a.alpha,     \
      seg, a.energy, m->g, m->gmax, a.part, mq_in ? m->mq : nullptr, mq_save ? m->mq : nullptr)
    if (!chg && (wide || NRT2l != 13))
      MX_RED(float, 3, rpf);
    else if (NW == 1)
      MX_RED(double, 1, rp);
    else if (NW == 3)
      MX_RED(double, 3, rp);
    else
      MX_RED(double, 4, rp);
#undef MX_RED
  }
  if (mq_save) {
    m->mq_epoch = a.epoch;
    m->mq_q = q;
  }
  if (a.energy) k_reduce_rows<<<7, MDFF_BLOCK, 0, st>>>(a.part, gk, a.sc_rec);
  const int W = efg ? 6 : 3;
  const int ncols = 2 * W * m->N1;
  const int NCT = (ncols + 15) / 16;
  const int NCTl = (NCT <= APP_NCT_S) ? APP_NCT_S : (NCT <= APP_NCT_F) ? APP_NCT_F : APP_NCT_E;
  const int KS = 2 * m->CHp / 32;
  {
    ProfScope _ps(m->ctx, K_MX_GPREP);
    const size_t total = (size_t)NCTl * KS * 64 * 2;  // two threads per fragment lane
    const dim3 gg((unsigned)((total + MDFF_BLOCK - 1) / MDFF_BLOCK), m->nseg);
    mx_gprep<<<gg, MDFF_BLOCK, 0, st>>>(m->N1, m->N3, m->C23, m->CH, m->CHp, W, NCTl, m->g,
                                        m->gmax, m->ky, m->kz, m->bfrag, m->sinv);
  }
  {
    ProfScope _ps(m->ctx, K_MX_APPLY);
    if (NCTl == APP_NCT_S)
      launch_apply<APP_NCT_S, APP_RT_S>(m, st, rj0, rj1);
    else if (NCTl == APP_NCT_F)
      launch_apply<APP_NCT_F, APP_RT_F>(m, st, rj0, rj1);
    else
      launch_apply<APP_NCT_E, APP_RT_E>(m, st, rj0, rj1);
  }
  if (cg && W == 3 && !seg) {
    // ES_IND of the conjugate-gradient SCF: fold and combine in one kernel
    ProfScope _ps(m->ctx, K_MX_COMBINE);
    const unsigned gb = (unsigned)((rj1 - rj0) / 64);
    if (m->app_slices == 1)
      mx_apply_epi_cg<1><<<gb, MDFF_BLOCK, 0, st>>>(n, m->npad, rj0, m->N1, NCTl * 16, m->X,
                                                    m->tj1, m->sinv, *cg);
    else if (m->app_slices == 2)
      mx_apply_epi_cg<2><<<gb, MDFF_BLOCK, 0, st>>>(n, m->npad, rj0, m->N1, NCTl * 16, m->X,
                                                    m->tj1, m->sinv, *cg);
    else
      mx_apply_epi_cg<4><<<gb, MDFF_BLOCK, 0, st>>>(n, m->npad, rj0, m->N1, NCTl * 16, m->X,
                                                    m->tj1, m->sinv, *cg);
    m->cg_parts = (int)gb;
    return true;
  }
  {
    ProfScope _ps(m->ctx, K_MX_COMBINE);
    // the K slices a compile-time count, so that all of a thread's loads
    // issue together
    const dim3 ge((rj1 - rj0) / 64, W);
    if (m->app_slices == 1)
      mx_apply_epi<1><<<ge, MDFF_BLOCK, 0, st>>>(n, m->npad, rj0, m->N1, W, NCTl * 16, seg,
                                                 m->d_kseg, m->X, m->tj1, m->sinv, apart);
    else if (m->app_slices == 2)
      mx_apply_epi<2><<<ge, MDFF_BLOCK, 0, st>>>(n, m->npad, rj0, m->N1, W, NCTl * 16, seg,
                                                 m->d_kseg, m->X, m->tj1, m->sinv, apart);
    else
      mx_apply_epi<4><<<ge, MDFF_BLOCK, 0, st>>>(n, m->npad, rj0, m->N1, W, NCTl * 16, seg,
                                                 m->d_kseg, m->X, m->tj1, m->sinv, apart);
  }
  return false;
}

int kspace(MixedState* m, int mode, const double* q, const double* mu, bool chg, bool dip,
           bool efg, double* apart, bool weights_ready, const CombineArgs* cg) {
  CtxView v = view(m->ctx);
  KsCall a;
  a.weights_ready = weights_ready;
  a.st = v.st;
  a.s = v.s;
  a.epoch = v.evals;
  a.cell = v.cell;
  a.alpha = v.P->alphaES;
  a.energy = (mode == ES_FINAL);
  a.part = v.part;
  a.sc_rec = v.sc + v.sc_rec;
  return run_kspace(m, a, q, mu, chg, dip, efg, apart, cg) ? m->cg_parts : 0;
}

float* dipole_weights(MixedState* m, int* npad) {
  *npad = m->npad;
  return m->d_seg ? nullptr : m->Wt;  // a batch has its own slot layout
}

void kspace_batch(MixedState* m, hipStream_t st, const double* s, long long epoch,
                  const int* cfg, const double* q, const double* mu, bool chg, bool dip,
                  bool force, double* Erec, double* Frec) {
  KsCall a;
  a.st = st;
  a.s = s;
  a.epoch = epoch;
  std::memset(&a.cell, 0, sizeof(a.cell));
  a.alpha = 0.0;
  a.energy = false;
  a.part = nullptr;
  a.sc_rec = nullptr;
  run_kspace(m, a, q, mu, chg, dip, force, m->apart);
  mx_batch_fields<<<(m->N + MDFF_BLOCK - 1) / MDFF_BLOCK, MDFF_BLOCK, 0, st>>>(
      m->N, cfg, m->d_seg, m->apart, q, mu, force ? 1 : 0, Erec, Frec);
}

}  // namespace mixed
}  // namespace mdff
