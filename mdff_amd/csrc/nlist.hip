// nlist.hip -- O(N) neighbour list on the device: a cell grid over the
// fractional coordinates, the atoms sorted into it, and a list search over
// the stencil of neighbouring cells.  Replaces vnlist_pbc
// (src/tools.f90:274-370), which tests every pair; the pair set is the same
// (every j != i whose minimum image, s - nint(s) as src/field.f90:1268-1273,
// lies within the list cutoff), so the brute-force k_nlist remains only for
// cells too small for a stencil without repeated cells.
//
// The grid: nc_a cells along lattice vector a, of perpendicular width
// W_a / nc_a.  A pair within the cutoff lc differs in fractional coordinate
// a by at most lc / W_a after the minimum image, so its cells differ by at
// most m_a = ceil(lc nc_a / W_a) (cyclically); with 2 m_a + 1 <= nc_a the
// stencil visits each cell once.
//
// The cells are row-major (a = 0 slowest), so the 2 m_2 + 1 cells of a
// stencil row along the third axis are one contiguous range of the sorted
// atoms (two when the row wraps).  The search gives one wave to each centre
// atom: the lanes take the candidates of all stencil rows as one sequence,
// a ballot compacts the hits, and the row is sorted by atom index in LDS,
// the order of the all-pairs build.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace mdff {

__device__ __forceinline__ int cell_coord(double s, int nc) {
  double f = s - floor(s);
  return min(nc - 1, (int)(f * nc));
}

// Zero the cell counts (ncell + 1 entries).
__global__ __launch_bounds__(MDFF_BLOCK) void k_cell_zero(int m, int* __restrict__ cnt) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m) cnt[k] = 0;
}

// Cell of every atom, counts per cell (cnt[cell + 1]).
__global__ __launch_bounds__(MDFF_BLOCK) void k_cell_bin(int n, const double4* __restrict__ s4,
                                                         int nc0, int nc1, int nc2,
                                                         int* __restrict__ cellof,
                                                         int* __restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 s = s4[i];
  const int c = (cell_coord(s.x, nc0) * nc1 + cell_coord(s.y, nc1)) * nc2 + cell_coord(s.z, nc2);
  cellof[i] = c;
  atomicAdd(&cnt[c + 1], 1);
}

// Inclusive scan of cnt[1..m] in place (cnt[0] = 0), so that cnt[c] is the
// first slot of cell c; one block.  cur[c] = cnt[c], the fill cursors.
__global__ __launch_bounds__(1024) void k_cell_scan(int m, int* __restrict__ cnt,
                                                     int* __restrict__ cur) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int per = (m + 1023) / 1024;
  const int b = 1 + t * per, e = min(m + 1, b + per);
  int acc = 0;
  for (int k = b; k < e; ++k) acc += cnt[k];
  part[t] = acc;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int v = (t >= off) ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = (t > 0) ? part[t - 1] : 0;
  for (int k = b; k < e; ++k) {
    run += cnt[k];
    cnt[k] = run;
  }
  __syncthreads();
  for (int k = t; k < m; k += 1024) cur[k] = cnt[k];
}

// Atoms into their cells (slot order within a cell is arbitrary here and
// fixed by k_cell_order).
__global__ __launch_bounds__(MDFF_BLOCK) void k_cell_fill(int n, const int* __restrict__ cellof,
                                                          int* __restrict__ cur,
                                                          int* __restrict__ perm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  perm[atomicAdd(&cur[cellof[i]], 1)] = i;
}

// Atoms of each cell in index order; rank[perm[r]] = r.
__global__ __launch_bounds__(MDFF_BLOCK) void k_cell_order(int m, const int* __restrict__ cnt,
                                                           int* __restrict__ perm,
                                                           int* __restrict__ rank) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m) return;
  const int b = cnt[c], e = cnt[c + 1];
  for (int r = b + 1; r < e; ++r) {
    const int v = perm[r];
    int q = r - 1;
    while (q >= b && perm[q] > v) {
      perm[q + 1] = perm[q];
      --q;
    }
    perm[q + 1] = v;
  }
  for (int r = b; r < e; ++r) rank[perm[r]] = r;
}

// Ranks r whose atom perm[r] has a species with its bit in polmask, in rank
// order: one workgroup takes the ranks in chunks of its size, and the
// threads of a hit find its slot from the wave ballots of the chunk.
__global__ __launch_bounds__(1024) void k_pol_rows(int n, const int* __restrict__ perm,
                                                   const int* __restrict__ typ,
                                                   unsigned polmask, int* __restrict__ rows) {
  __shared__ int wsum[1024 / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nw = (int)(blockDim.x >> 6);
  int base = 0;
  for (int r0 = 0; r0 < n; r0 += blockDim.x) {
    const int r = r0 + (int)threadIdx.x;
    const bool hit = r < n && ((polmask >> typ[perm[r]]) & 1u);
    const unsigned long long m = __ballot(hit);
    if (lane == 0) wsum[w] = __popcll(m);
    __syncthreads();
    int off = base, tot = 0;
    for (int k = 0; k < nw; ++k) {
      if (k < w) off += wsum[k];
      tot += wsum[k];
    }
    if (hit) rows[off + __popcll(m & ((1ull << lane) - 1ull))] = r;
    base += tot;
    __syncthreads();
  }
}

// Neighbour rows from the stencil: one wave (and workgroup) per centre
// atom, the centres in cell order.  Rows nbr[i * maxnb + k]; cnt[i] the row
// length; the longest row above maxnb goes to overflow (the caller grows
// the list and runs again).  The stencil rows live in the lanes (row r in
// lane r, at most 64 rows); the lanes find the row of a candidate by binary
// lifting over the row bases with wave shuffles.  The hits are gathered in
// LDS and sorted by atom index (bitonic, NL_SORT entries at most) before
// they are written, so that a row lists its neighbours in index order as
// the all-pairs build does: the pair kernels then meet the species in long
// runs (the atoms of a species are contiguous in index), which keeps their
// species-dependent work converged across a lane group.  Longer rows are
// written unsorted.
constexpr int NL_SORT = 1024;
__global__ __launch_bounds__(64) void k_nlist_cells(NlCellArgs a) {
  __shared__ int srow[NL_SORT];
  const int lane = threadIdx.x;
  const int wv = blockIdx.x;
  if (wv >= a.n) return;  // the whole (one-wave) workgroup
  const int i = a.perm[wv];
  const double4 si = a.s4[i];
  const int c = a.cellof[i];
  const int c2 = c % a.nc[2], c1 = (c / a.nc[2]) % a.nc[1], c0 = c / (a.nc[2] * a.nc[1]);
  const int w1 = 2 * a.m[1] + 1;
  const int nrow = (2 * a.m[0] + 1) * w1;
  // stencil rows (d0, d1): the third-axis run c2 - m2 .. c2 + m2 as one or
  // two ranges of sorted atoms
  int s0 = 0, l0 = 0, s1 = 0, l1 = 0;
  if (lane < nrow) {
    int x = c0 + lane / w1 - a.m[0], y = c1 + lane % w1 - a.m[1];
    x += (x < 0) ? a.nc[0] : ((x >= a.nc[0]) ? -a.nc[0] : 0);
    y += (y < 0) ? a.nc[1] : ((y >= a.nc[1]) ? -a.nc[1] : 0);
    const int rowc = (x * a.nc[1] + y) * a.nc[2];
    const int lo = c2 - a.m[2], hi = c2 + a.m[2];
    if (lo < 0) {
      s0 = a.cstart[rowc + lo + a.nc[2]];
      l0 = a.cstart[rowc + a.nc[2]] - s0;
      s1 = a.cstart[rowc];
      l1 = a.cstart[rowc + hi + 1] - s1;
    } else if (hi >= a.nc[2]) {
      s0 = a.cstart[rowc + lo];
      l0 = a.cstart[rowc + a.nc[2]] - s0;
      s1 = a.cstart[rowc];
      l1 = a.cstart[rowc + hi - a.nc[2] + 1] - s1;
    } else {
      s0 = a.cstart[rowc + lo];
      l0 = a.cstart[rowc + hi + 1] - s0;
    }
  }
  // exclusive prefix of the row lengths across the wave
  const int tot = l0 + l1;
  int inc = tot;
  for (int off = 1; off < 64; off <<= 1) {
    const int v = __shfl_up(inc, off, 64);
    if (lane >= off) inc += v;
  }
  const int total = __shfl(inc, 63, 64);
  const int base = inc - tot;
  int count = 0;
  int* row = a.nbr + (size_t)i * a.maxnb;
  for (int v0 = 0; v0 < total; v0 += 64) {
    const int v = v0 + lane;
    // the last row r < nrow with base[r] <= v (bases do not decrease)
    int r = 0;
    for (int step = 32; step > 0; step >>= 1) {
      const int cand = r + step;
      const int b = __shfl(base, min(cand, 63), 64);
      if (cand < nrow && b <= v) r = cand;
    }
    const int off = v - __shfl(base, r, 64);
    const int rl = __shfl(l0, r, 64);
    const int r0 = __shfl(s0, r, 64), r1 = __shfl(s1, r, 64);
    const int slot = (off < rl) ? r0 + off : r1 + off - rl;
    bool hit = false;
    int j = 0;
    if (v < total) {
      j = a.perm[slot];
      const double4 sj = a.s4[j];
      double rx, ry, rz;
      const double d2 = min_image(a.cell, si.x - sj.x, si.y - sj.y, si.z - sj.z, rx, ry, rz);
      hit = (j != i) && (d2 <= a.cut2);
    }
    const unsigned long long mask = __ballot(hit);
    if (hit) {
      const int k = count + __popcll(mask & ((1ull << lane) - 1ull));
      if (k < NL_SORT)
        srow[k] = j;
      else if (k < a.maxnb)
        row[k] = j;
    }
    count += __popcll(mask);
  }
  __syncthreads();
  if (count <= NL_SORT) {
    int P = 1;
    while (P < count) P <<= 1;
    for (int k = count + lane; k < P; k += 64) srow[k] = 0x7fffffff;
    __syncthreads();
    for (int size = 2; size <= P; size <<= 1)
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        for (int t = lane; t < (P >> 1); t += 64) {
          // compare-exchange pair (x, x + stride) of the t-th comparator
          const int x = 2 * t - (t & (stride - 1));
          const int y = x + stride;
          const bool up = (x & size) == 0;
          const int vx = srow[x], vy = srow[y];
          if ((vx > vy) == up) {
            srow[x] = vy;
            srow[y] = vx;
          }
        }
        __syncthreads();
      }
  }
  // the first min(count, NL_SORT) entries, sorted when the row fits the LDS
  const int nw = min(min(count, NL_SORT), a.maxnb);
  for (int k = lane; k < nw; k += 64) row[k] = srow[k];
  if (lane == 0) {
    a.cnt[i] = count;
    if (count > a.maxnb) atomicMax(a.overflow, count);
  }
}

// Per-atom arrays across the context's internal numbering, laid out
// (inner, n, planes) with the inner index fastest: gather dst(k, i, p) =
// src(k, idx[i], p), scatter dst(k, idx[i], p) = src(k, i, p), with idx[i]
// the caller's index of internal atom i.
__global__ __launch_bounds__(MDFF_BLOCK) void k_perm_gather(int n, int inner, int planes,
                                                            const int* __restrict__ idx,
                                                            const double* __restrict__ src,
                                                            double* __restrict__ dst) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t tot = (size_t)n * inner * planes;
  if (t >= tot) return;
  const size_t k = t % inner, r = t / inner;
  const size_t i = r % n, p = r / n;
  dst[t] = src[k + (size_t)inner * (idx[i] + (size_t)n * p)];
}

__global__ __launch_bounds__(MDFF_BLOCK) void k_perm_scatter(int n, int inner, int planes,
                                                             const int* __restrict__ idx,
                                                             const double* __restrict__ src,
                                                             double* __restrict__ dst) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t tot = (size_t)n * inner * planes;
  if (t >= tot) return;
  const size_t k = t % inner, r = t / inner;
  const size_t i = r % n, p = r / n;
  dst[k + (size_t)inner * (idx[i] + (size_t)n * p)] = src[t];
}

}  // namespace mdff
