// kernels.h -- kernel argument blocks and launch declarations shared by the
// kernel files and the host orchestration (mdffgpu.hip).
#pragma once

#include <algorithm>

#include "common.h"

struct mdffgpu_ctx;

namespace mdff {

constexpr int KC_MAX = 16;      // k vectors per chunk (consecutive n3)
constexpr int RHO_TILE = 32;    // atoms per LDS tile in the structure factor
constexpr int APPLY_TILE = 64;  // k chunks per LDS tile in the apply
constexpr int DD_LANES = 64;    // lanes per atom of the SCF dipole field

enum { ES_STATIC = 0, ES_IND = 1, ES_FINAL = 2 };
enum { SR_LJ = 1, SR_BM = 2, SR_BMD = 4, SR_MO = 8, SR_TRUNC1 = 16, SR_TRUNC2 = 32 };
enum { TASK_QQ = 1, TASK_QD = 2, TASK_DD = 4, TASK_QQUAD = 8, TASK_DQUAD = 16, TASK_QUADQ = 32 };
constexpr int TASK_QUAD_ANY = TASK_QQUAD | TASK_DQUAD | TASK_QUADQ;

// Kernel identifiers of the launch instrumentation (mdffgpu_profile_*).
enum KernelId {
  K_PREPARE, K_NLIST, K_SR, K_DIR_STATIC, K_DIR_IND, K_DIR_FINAL, K_RHO_CHG, K_RHO_DIP,
  K_RHO_BOTH, K_RHO_RED, K_APPLY_F, K_APPLY_EFG, K_MOMENTS, K_COMBINE, K_INDUCED,
  K_RESID, K_EXTRAP, K_REDUCE, K_ADD, K_PAIRS_T2, K_DD_FIELD, K_MX_SR, K_MX_DIR, K_MX_PHASE, K_MX_RHO,
  K_MX_RHO_RED, K_MX_APPLY, K_MX_COMBINE, K_MX_WEIGHTS, K_MX_GPREP, K_BATCH, K_ADJ,
  K_EFG_DIR, K_EFG_DS, K_EFG_COMBINE, K_MD, K_QUAD, K_WFC, K_COUNT
};
extern const char* const kKernelNames[K_COUNT];

// RAII bracket of one launch with HIP events on the context's stream when
// instrumentation is enabled (implemented in mdffgpu.hip).
struct ProfScope {
  mdffgpu_ctx* c;
  int id;
  hipEvent_t a, b;
  ProfScope(mdffgpu_ctx* ctx, int kid);
  ~ProfScope() noexcept(false);
};

// Per-configuration constants of a batch of configurations (the batched
// path evaluates many cells at once; kernels instantiated with BATCH read
// the cell, Ewald parameter and cutoffs of atom i's configuration here).
struct BatchCfg {
  Cell cell;
  double alpha;   // Ewald alpha
  double cut2;    // real-space Ewald cutoff^2
  double rc2;     // short-range cutoff^2
  double fpi_V;   // 4 pi / Omega
  double selfa2;  // 2 alpha^3 / (3 sqrt(pi))
  int n, off;     // atoms, first atom in the batch arrays
  int koff, nk;   // half-space k vectors of this configuration
  int kmax[3];
  int pad;
};

// A chunk of KC consecutive k vectors (n1, n2, n3s .. n3s+KC-1) of the
// half-space mesh of kpoint_sum_init (src/kspace.f90:162-191).
struct KChunk {
  int n1, n2, n3s, pad;
};

struct SrArgs {
  int n, maxnb;
  const double* s;
  const double4* s4;  // the same, one (s1, s2, s3, w) per atom (w: see s4_type)
  const int* typ;
  const int* cnt;
  const int* nbr;
  Cell cell;
  const PairTables* T;
  TTTable tt;
  double rc2;
  int flags;
  double* f;
  double* part;
  const BatchCfg* bc;  // batched path: per-configuration constants
  const int* cfg;      // batched path: configuration of each atom
  const int* perm;     // atoms in this order (spatial locality), or null
};

struct EsDirArgs {
  int n, maxnb;
  const double* s;
  const double4* s4;  // the same, one (s1, s2, s3, w) per atom (w: see s4_type)
  const int* typ;
  const int* cnt;
  const int* nbr;
  Cell cell;
  const PairTables* T;
  TTTable tt;
  const double* q;
  const double* mu;
  double cut2, alpha;
  int task, do_efield, do_efg;
  double* ef;
  double* efg;
  double* f;
  double* part;
  const BatchCfg* bc;  // batched path: per-configuration constants
  const int* cfg;      // batched path: configuration of each atom
  const int* rank;     // spatial rank of each atom (SCF pair table), or null
  const int* perm;     // atoms in this order (spatial locality), or null
  // FP32 static pass of a single configuration: the compacted SCF pair
  // table (k_dd_table's layout) written on the way, or null
  // neighbours' charge and dipole packed per atom as (q, mu_x, mu_y, mu_z),
  // float4 for the FP32 kernels and double4 for the FP64 ones, or null
  // (q and mu read separately)
  const void* qm;
  float2* ddtab;
  int* ddtj;
  int* ddcnt;
  int dd16;  // pair table: neighbour slots as 16-bit indexes (n <= 65536)
  // pair table: only pairs of two species whose bits are set (the
  // polarizable ones; the dipole field is needed at them only and only
  // their dipoles are nonzero)
  unsigned polmask;
  // centre rows visited (the ranks listed, nrows of them), or all rows
  const int* rows;
  int nrows;
};

struct RhoArgs {
  int n, nchunk, KC, kmax[3];
  const KChunk* chunks;
  const double2* eph;
  const double* q;
  const double* mu;
  const double* th;  // quadrupoles (n,3,3) Fortran order, for the QUAD variant
  Cell cell;
  int atoms_per_split, kpad;
  double2* part;
  int j_begin, j_end;  // the atoms summed (splits of atoms_per_split from j_begin)
};

struct RhoRedArgs {
  int kpad, nsplit, do_energy;
  const double2* part;
  const double* Ak;
  const double* kcoe;
  const double* kvec;
  double2* g;
  double* pout;
};

struct ApplyArgs {
  int n, nchunk, KC, chunks_per_split;
  const KChunk* chunks;
  const double2* g;
  const double2* eph;
  const double* s;
  const double4* s4;  // the same, one (s1, s2, s3, w) per atom (w: see s4_type)
  double* part;
  int i_begin, i_end;  // the atoms whose sums are formed (others left as they are)
};

struct CombineArgs {
  int n, nsplit, lsurf, do_efield, do_efg, dip_force;
  const double* apart;
  Cell cell;
  double fpi_V, fpi_3V, selfa2, coul_unit;
  const double* mu;
  const double* q;
  const double* ef_dir;
  const double* efg_dir;
  const double* f_dir;
  const double* qm;  // device: qtot[3], mutot[3]
  double* E_stat;
  double* E;
  double* ef_out;
  double* efg_out;
  double* f;
  // ES_IND: SCF residual partials (2 rows) when non-null
  const int* typ;
  const TypeTables* T;
  double* rpart;
  // ES_IND of the conjugate-gradient SCF, with rpart: mu is the search
  // direction p; writes Ap = alpha^-1 p - E_ind(p) instead of E and the block
  // partials of p.Ap into row 0 of rpart
  double* cg_Ap;
  // ES_IND with quadrupole polarizabilities: the induced quadrupoles of the
  // residual (n,3,3), or null
  const double* th;
};

struct ExtrapArgs {
  int n, nh, op, ord;
  double b[6];
  double w;
  double* hist;
  double* mu;
  double* mup;
  const double* E;
  const int* typ;
  const TypeTables* T;
  double* part;
  double* rpart;  // op 5: SCF residual rows 0, 1 of the corrected moments, or null
  double* th;     // op 5 with quadrupole polarizabilities: theta_ind = polquad E, or null
};

__global__ void k_prepare(int n, const double* pos, Cell cell, const int* ext, const int* typ,
                          double* s, double4* s4, double2* eph);
// s4.w of a single configuration: 16 x (the caller's index of the atom) +
// its species, so that one gather gives position, species and the index
// the pair-order conventions compare (the batch writes 0)
__device__ __forceinline__ int s4_type(const double4& s) { return ((int)s.w) & (MDFF_NT - 1); }
template <bool FILL>
__global__ void k_nlist(int n, const double* s, Cell cell, double cut2, int maxnb, int jlen,
                        int* scnt, int* nbr);
__global__ void k_nlist_offsets(int n, int nsp, int maxnb, int* scnt, int* cnt, int* overflow);
constexpr int NLIST_RANGES = 16;  // candidate ranges of the list build (grid.y)
// O(N) list build over a cell grid (nlist.hip)
struct NlCellArgs {
  int n, maxnb;
  int nc[3];  // cells along each lattice vector
  int m[3];   // stencil half-widths, 2 m + 1 <= nc, (2 m0 + 1)(2 m1 + 1) <= 64
  const double4* s4;
  const int* cellof;
  const int* cstart;  // first sorted slot of each cell (ncell + 1)
  const int* perm;    // atoms sorted by cell
  Cell cell;
  double cut2;
  int* cnt;
  int* nbr;
  int* overflow;
};
__global__ void k_cell_zero(int m, int* cnt);
__global__ void k_cell_bin(int n, const double4* s4, int nc0, int nc1, int nc2, int* cellof,
                           int* cnt);
__global__ void k_cell_scan(int m, int* cnt, int* cur);
__global__ void k_cell_fill(int n, const int* cellof, int* cur, int* perm);
__global__ void k_cell_order(int m, const int* cnt, int* perm, int* rank);
__global__ void k_pol_rows(int n, const int* perm, const int* typ, unsigned polmask, int* rows);
__global__ void k_nlist_cells(NlCellArgs a);
__global__ void k_perm_gather(int n, int inner, int planes, const int* idx, const double* src,
                              double* dst);
__global__ void k_perm_scatter(int n, int inner, int planes, const int* idx, const double* src,
                               double* dst);

// device-resident MD (md.hip)
constexpr int MDFF_NHC_MAX = 16;
struct MdArgs {
  int n;
  const double* mass;  // per atom
  double* pos;
  double* vel;
  const double* f;
  double* fs;  // forces of the previous step (velocity Verlet)
  double dt, dt2, dtsq2;
  double* part;  // block partials of sum m |v|^2
};
// Thermostat state of the Nose-Hoover chains on the device.
struct NhcState {
  double vxi[MDFF_NHC_MAX], xi[MDFF_NHC_MAX], Q[MDFF_NHC_MAX];
  double yw[9];  // Yoshida weights
  double s;      // scale factor of the last chain call
  double e_nvt;
};
struct NhcMass {
  double m[MDFF_NT];
};
__global__ void k_vv_pos(MdArgs a);
__global__ void k_vv_vel(MdArgs a);
__global__ void k_kick_drift(MdArgs a);
__global__ void k_kick(MdArgs a);
__global__ void k_vscale(MdArgs a, const NhcState* s);
__global__ void k_massia(int n, const int* typ, NhcMass m, double* mass);
__global__ void k_nhcn_chain(NhcState* st, const double* sum_mv2, int nhc_n, int yosh, int mults,
                             double dt, double temp, double L, int final);
__global__ void k_nhc2_chain(NhcState* st, const double* sum_mv2, double dt, double temp,
                             double L, int final);
template <typename R, bool BATCH = false>
__global__ void k_shortrange(SrArgs a);
template <int MODE, typename R, bool BATCH = false>
__global__ void k_es_dir(EsDirArgs a);
template <bool BATCH = false>
__global__ void k_pairs_t2(EsDirArgs a, double* tab);
__global__ void k_dd_field(EsDirArgs a, const double* tab, const double4* mu4);
template <typename R4>
__global__ void k_pack_mu(int n, const double* mu, R4* mu4);
__global__ void k_pack_mu_ranked(int n, const double* mu, const int* rank, float4* mu4);
template <typename R4>
__global__ void k_pack_qmu(int n, const double* q, const double* mu, R4* qm);
template <bool BATCH>
__global__ void k_dd_table(EsDirArgs a, float2* tab, int* tj, int* tcnt);
__global__ void k_pack_u4(int n, const double* s, const int* rank, uint4* u4);
template <typename IT, int U>
__global__ void k_dd_tab2(int n, int maxnb, const float2* tab, const IT* tj, const int* tcnt,
                          const float4* mu4, const uint4* u4, Cell cell, const BatchCfg* bc,
                          const int* cfg, const int* perm, const int* rows, int nrows,
                          double* ef);
// neighbour slots of the FP32 pair table fit 16 bits
inline bool dd_idx16(long long n) { return n <= 65536; }
// Grid of k_dd_tab2 (persistent waves): four per SIMD of the device, at most
// one per atom.
inline int dd_tab_blocks(int n) {
  static int cus = 0;
  if (cus <= 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
  }
  const long long waves = std::min<long long>(n, (long long)cus * 16);
  return (int)((waves * 64 + MDFF_BLOCK - 1) / MDFF_BLOCK);
}
template <bool CHG, bool DIP, bool QUAD = false>
__global__ void k_rho64(RhoArgs a);
__global__ void k_rho_reduce64(RhoRedArgs a);
template <bool EFG>
__global__ void k_apply64(ApplyArgs a);
__global__ void k_moments(int n, const double* pos, const double* q, const double* mu,
                          double* part);
template <int MODE>
__global__ void k_combine64(CombineArgs a);
// Preconditioned conjugate gradient for the induced dipoles,
// (alpha^-1 - T) x = E_stat with the per-type polarizability alpha as the
// preconditioner; atoms without polarizability keep zeros.  The operator's
// T p is read as E - E_stat after an ES_IND call on p.
struct CgArgs {
  int n;
  const int* typ;
  const TypeTables* T;
  double* x;        // induced dipoles (the SCF's mu_ind)
  double* r;        // residual E_stat + T x - alpha^-1 x
  double* z;        // alpha r
  double* p;        // search direction
  double* Ap;       // (alpha^-1 - T) p
  const double* E;  // E_stat + T v of the last ES_IND call (v = x or p)
  const double* Estat;
  const double* sc;  // device scalars: r.z of the current iterate at s_rz
  int s_rz;
  const double* pap_part;  // k_cg_update: block partials of p.Ap (npap of them)
  int npap;
  double* part;      // block partials, rows of nblk
  const int* rank;   // with p4: FP32 rank-ordered copy of p for k_dd_tab
  float4* p4;
  double* sc_out;    // k_cg_dir: the update's three sums, device scalars
  double* sc_host;   //   and the same in mapped host memory
  // k_cg_dir, when wt is set: FP32 structure-factor weights of p,
  // wt[c * wt_npad + i] = (float)(2 pi (B^T p_i)_c) * wscale
  float* wt;
  int wt_npad;
  float wscale;
  double wB[9];
};
__global__ void k_cg_init(CgArgs a);
__global__ void k_cg_update(CgArgs a);
__global__ void k_cg_dir(CgArgs a);
__global__ void k_cg_upol(CgArgs a);
__global__ void k_induced(int n, const int* typ, const TypeTables* T, const double* E,
                          double* mu, double* part, const int* rank, float4* mu4, double* th);
__global__ void k_scf_resid(int n, const int* typ, const TypeTables* T, const double* mu,
                            const double* E, double* part);

// Quadrupole terms of the real-space Ewald sum (realspace.hip): charge-,
// dipole- and quadrupole-quadrupole energies, the quadrupole-quadrupole
// field and EFG (src/field.f90:2542-2601), FP64.
struct QuadArgs {
  int n, maxnb;
  const double* s;
  const double4* s4;
  const int* cnt;
  const int* nbr;
  const int* perm;
  Cell cell;
  double cut2, alpha;
  const double* q;
  const double* mu;
  const double* th;  // (n,3,3) Fortran order
  int task;          // TASK_QQUAD / TASK_DQUAD / TASK_QUADQ bits
  int do_energy, do_efield, do_efg;
  double* ef;    // added to
  double* efg;   // added to
  double* part;  // block partials of the energy (one row)
};
__global__ void k_es_quad(QuadArgs a);

// Wannier-centre dipoles (moment_from_WFc, src/field.f90:3492-3672) from the
// neighbour list, or from all atoms when nbr is null.
struct WfcArgs {
  int n, maxnb;
  const double* s;
  const int* typ;
  const int* cnt;
  const int* nbr;
  int lwfc[MDFF_NT];  // per species
  Cell cell;
  double rcut;      // compared with the squared distance
  double* mu;       // (n,3), written for every atom
  int* bad;         // set to 1 + an ion without its lwfc centres
  const int* ext;   // the caller's index of each atom (null: the identity)
};
__global__ void k_wfc(WfcArgs a);
__global__ void k_extrap(ExtrapArgs a);
__global__ void k_reduce_rows(const double* part, int nb, double* out);
__global__ void k_reduce_rows2(const double* part, int nb, double* out, double* out2);
__global__ void k_disp(int n, const double* pos, const double* ref, double* part);
__global__ void k_reduce_max(const double* part, int nb, double* out);

// lanes that share one centre atom in the real-space kernels
constexpr int LPA = 32;
__global__ void k_add(int m, const double* a, const double* b, double* out);

// ---------------------------------------------------------------------------
// standalone electric field gradients (efg.hip)
struct EfgDirArgs {
  int n, jlen;
  const double* s;   // fractional coordinates SoA (3n)
  const double* q;   // contributing charges
  const double* mu;  // contributing dipoles SoA (3n)
  Cell cell;
  double alpha;
  double* part;      // (grid.y, 6, n)
};
struct EfgDsArgs {
  int n, ncell;
  const double* s;
  const double* q;
  const double* mu;
  const double* cells;  // image cells (n1, n2, n3) as doubles, 3 per cell
  Cell cell;
  double cut2;
  double* part;         // (6, n)
};
struct EfgCombineArgs {
  int n, ndir, nrec;
  const double* dir;  // (ndir, 6, n) real-space partials
  const double* rec;  // (nrec, 9, n) reciprocal partials of k_apply64<true>
  Cell cell;
  double fpi_V;       // 4 pi / V
  double selfa;       // -4 alpha^3 / (3 sqrt(pi))
  const double* q;    // contributing charges (self term)
  double corr;        // charged-cell correction on the diagonal
  double scale;       // coul_unit, negated with lefg_vasp_sign
  double* out;        // (n, 3, 3) Fortran order
};
__global__ void k_efg_mask(int n, const int* typ, int it_efg, const double* q, const double* mu,
                           double* qe, double* mue);
template <bool DIP>
__global__ void k_efg_es_dir(EfgDirArgs a);
template <bool DIP>
__global__ void k_efg_ds(EfgDsArgs a);
__global__ void k_efg_combine(EfgCombineArgs a);
__global__ void k_efg_nmr(int n, const double* efg6, double* nmr, double* U, double* S, int* bad);

}  // namespace mdff
