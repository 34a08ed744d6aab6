// mixed.h -- mixed-precision evaluation path (FP32 pair, erfc and k-space
// arithmetic; FP64 energy / virial reductions and SCF residual).
#pragma once

#include "../../include/mdffgpu.h"
#include "kernels.h"

#include <vector>

namespace mdff {
namespace mixed {

// Read-only view of the context buffers the mixed path shares with the
// FP64 path (neighbour list, k mesh, per-atom arrays, scalar slots).
struct CtxView {
  int n;
  hipStream_t st;
  Cell cell;
  const mdffgpu_params* P;
  int kmax[3];
  int KC, nchunk, kpad;
  const KChunk* chunks;
  const double* Ak;
  const double* kcoe;
  const double* kvec;
  long long evals;  // evaluation counter: phase tables are per evaluation
  const double* pos;
  const double* s;
  const double2* eph;
  const int* typ;
  const int* cnt;
  const int* nbr;
  int maxnb;
  const PairTables* pt;
  TTTable tt;
  const TypeTables* ttab;
  double* efdir;
  double* efgdir;
  double* fdir;
  double* force;
  double* Estat;
  double* E;
  double* ef;
  double* efg;
  double* part;
  double* sc;
  int sc_sr, sc_dir, sc_rec, sc_mom;
};

struct MixedState;

// One configuration on the padded atom axis of the reciprocal-space GEMMs:
// atoms off .. off+n-1 of the caller's arrays occupy padded slots
// poff .. poff+np-1 (np a multiple of 128, slots past n are zero), and
// structure-factor splits split0 .. split0+nsplit-1.
struct MxSeg {
  Cell cell;
  double alpha;
  int off, n, poff, np, split0, nsplit;
};

// defined in mdffgpu.hip
CtxView view(mdffgpu_ctx* c);

MixedState* create(mdffgpu_ctx* c);
void destroy(MixedState* m);
void on_cell(MixedState* m);
// One configuration: atoms dip0 .. dip1 - 1 hold every atom that can carry
// a dipole (the dipole-only structure factors sum over them) and atoms
// fld0 .. fld1 - 1 every atom whose field the static and SCF calls must
// give (their apply rows); dip1 or fld1 < 0: all atoms.  The field of the
// other atoms in those calls is left as the buffers held it.
bool set_active(MixedState* m, int dip0, int dip1, int fld0, int fld1);
// Reciprocal-space part of one multipole_ES call: writes the per-atom U, V
// sums (apply-partials layout of k_apply64 with one split) to apart and the
// energy / stress rows to the scalar slots.  With weights_ready (dipole
// weights alone) the caller has already written the FP32 weights of mu
// (dipole_weights).
// With cg (an ES_IND call of the conjugate-gradient SCF, one configuration,
// no surface term) the combine of k_combine64<ES_IND> is done here too;
// returns the number of p.Ap block partials written to cg->rpart, 0 if not.
int kspace(MixedState* m, int mode, const double* q, const double* mu, bool chg, bool dip,
           bool efg, double* apart, bool weights_ready = false, const CombineArgs* cg = nullptr);
// Structure-factor weights of dipoles alone, one configuration (slot = atom):
// Wt[c * npad + i] = (float)(2 pi (B^T mu_i)_c) * DIP_WEIGHT_SCALE, c = 0..2;
// the slots past the atoms stay zero.
constexpr float DIP_WEIGHT_SCALE = 256.0f;
float* dipole_weights(MixedState* m, int* npad);

// Batched path: configurations with a common kmax, one segment each (cell,
// alpha, atom range of the batch arrays; the padding is laid out here).
MixedState* create_batch(int N, const int kmax[3], const std::vector<MxSeg>& segs);
// Launch shape: K slices of the apply GEMM (1, 2 or 4) and atoms per
// structure-factor split (a multiple of 128); false if rejected.
bool set_tuning(MixedState* m, int app_slices, int split_atoms);
// Reciprocal-space field (and, with force, force) of every atom of the
// batch for sources (q, mu): Erec = -2 (4 pi/V) sum_k k t_k,
// Frec = q Erec + 2 (4 pi/V) sum_k k (k.mu) s_k (src/field.f90:2746-2801).
// epoch changes whenever the positions do (the phase tables follow it);
// cfg is the configuration of each atom.
void kspace_batch(MixedState* m, hipStream_t st, const double* s, long long epoch,
                  const int* cfg, const double* q, const double* mu, bool chg, bool dip,
                  bool force, double* Erec, double* Frec);

}  // namespace mixed
}  // namespace mdff
