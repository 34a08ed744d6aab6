// mdffgpu.hip -- C ABI (include/mdffgpu.h) and host orchestration of the
// force/field evaluation: engforce_driver (src/field.f90:1111-1166) with
// the polarizable-ion SCF (moment_from_pola_scf, :3200-3458) and the Ewald
// calls (multipole_ES, :1817-2054) on one MI355X.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mdffgpu.h"
#include "kernels.h"
#include "mixed.h"
#include "tables.h"

using namespace mdff;

namespace {

// scalar slots of ctx->d_sc
enum {
  SC_SR = 0,     // 12: u_lj vir_lj u_bm vir_bm u_mo vir_mo tau6
  SC_DIR = 12,   // 7: u tau6
  SC_REC = 19,   // 7: u tau6
  SC_MOM = 26,   // 8: qtot3 mutot3 qsq musq
  SC_UPOL = 34,  // 1
  SC_RMSD = 35,  // 2: sum, npol
  SC_DISP = 37,  // max squared displacement since the list build
  SC_CG = 40,    // 6: conjugate-gradient SCF, two (r.z, |r|^2, npol) triples
  SC_ESTAT = 46, // 2: sum of |E_stat|^2 over polarizable atoms, their count
  SC_KIN = 48,   // sum m |v|^2 of the MD loop, then the thermostat energy
  SC_QUAD = 50,  // real-space quadrupole energy of the final call
  SC_N = 52
};

thread_local std::string g_create_error;

const double B_ASPC[6][6] = {
    {1.0, 0, 0, 0, 0, 0},
    {2.0, -1.0, 0, 0, 0, 0},
    {2.5, -2.0, 0.5, 0, 0, 0},
    {2.8, -2.8, 1.2, -0.2, 0, 0},
    {3.0, -24.0 / 7.0, 27.0 / 14.0, -4.0 / 7.0, 1.0 / 14.0, 0},
    {22.0 / 7.0, -55.0 / 14.0, 55.0 / 21.0, -22.0 / 21.0, 5.0 / 21.0, -1.0 / 42.0}};
const double W_ASPC[6] = {0.0, 2.0 / 3.0, 0.6, 4.0 / 7.0, 5.0 / 9.0, 6.0 / 11.0};

struct Fail {
  int code;
};

}  // namespace

struct mdffgpu_ctx {
  mdffgpu_params P;
  int n = 0;
  int device = 0;
  std::string err;
  std::vector<int> typ0;
  Cell cell;
  double W[3];
  // k-space mesh
  int kmax[3] = {0, 0, 0};
  int KC = 1, nchunk = 0, kpad = 0, nk = 0;
  std::vector<KChunk> chunks_h;
  int rho_split = 1, rho_atoms = 1, app_split = 1, app_chunks = 1;
  // neighbour list
  int maxnb = 0;
  double list_cut = 0.0;
  // device buffers
  double *d_pos = nullptr, *d_s = nullptr;
  double2* d_eph = nullptr;
  int *d_typ = nullptr, *d_cnt = nullptr, *d_nbr = nullptr, *d_ovf = nullptr;
  int* d_nlr = nullptr;  // per-range neighbour counts, then offsets, of a list build
  double *d_q = nullptr, *d_qzero = nullptr, *d_dip = nullptr, *d_mu = nullptr,
         *d_muind = nullptr, *d_mup = nullptr;
  double4* d_s4 = nullptr;   // fractional coordinates, one double4 per atom
  // spatial order of the atoms (cells of about 4 A, row-major, atom index
  // within a cell), refreshed with the neighbour list: perm[rank] = atom
  int* d_perm = nullptr;
  int* d_rank = nullptr;
  // FP32 SCF pair table: its species (the dipole-polarizable ones, bit per
  // species) and the ranks of their atoms in rank order (npolrows of them,
  // k_pol_rows at every list build)
  unsigned polmask = 0;
  int npolrows = 0;
  // FP64 SCF field: the species whose field the SCF reads (dipole or
  // quadrupole polarizable)
  unsigned fieldmask = 0;
  // FP64 reciprocal sums: the atoms kr0 .. kr1 - 1 (context numbering) hold
  // every atom of those species (all atoms by default)
  int kr0 = 0, kr1 = 0;
  int* d_polrows = nullptr;
  // internal numbering of the atoms, fixed at the first evaluation: species
  // by species, by cell of about 4 A (row-major) and caller index within a
  // species, so that a neighbour row meets the species in long runs and
  // neighbours sit close in memory.  int2ext[k] is the caller's index of
  // internal atom k; every entry point gathers its per-atom inputs and
  // scatters its outputs across it (renum false: the identity).
  bool renum_decided = false, renum = false;
  int* d_int2ext = nullptr;
  double* d_xbuf = nullptr;  // staging of the boundary permutation
  // cell grid of the list build (nlist.hip): cells per axis, stencil
  // half-widths, whether the stencil search applies; cell of each atom,
  // first sorted slot of each cell, fill cursors
  int gnc[3] = {1, 1, 1}, gm[3] = {0, 0, 0};
  bool gstencil = false;
  int ncell_cap = 0;
  int *d_cellof = nullptr, *d_cstart = nullptr, *d_ccur = nullptr;
  // device-resident MD (mdffgpu_md_run): velocities, previous forces,
  // per-atom masses, thermostat state
  double *d_vel = nullptr, *d_fs = nullptr, *d_massia = nullptr;
  NhcState* d_nhc = nullptr;
  // quadrupoles (FP64 path): permanent ones (quadia), induced ones of the
  // SCF, their sum; any static quadrupole, any quadrupole polarizability,
  // any quadrupole term at all
  double *d_thstat = nullptr, *d_thind = nullptr, *d_theta = nullptr;
  bool lqstat = false, lqpol = false, lquad = false;
  bool quad_final = false;  // the last final call had quadrupole terms (SC_QUAD)
  // Wannier-centre dipoles and the index + 1 of an ion missing centres
  double* d_muwfc = nullptr;
  int* d_wfcbad = nullptr;
  bool lwannier = false;
  double mx_floor = 0.0;     // SCF residual floor of the mixed path (0 = unknown)
  // outcome of the last mixed-precision solve, folded into mx_floor once the
  // evaluation's scalars are on the host (update_floor)
  bool scf_stalled = false, scf_done = false;
  double scf_best = 0.0, scf_last = 0.0;
  double4* d_mu4 = nullptr;  // packed dipoles of the SCF field (double4 or float4)
  double *d_Estat = nullptr, *d_E = nullptr, *d_efdir = nullptr, *d_efgdir = nullptr,
         *d_fdir = nullptr, *d_force = nullptr, *d_ef = nullptr, *d_efg = nullptr;
  double* d_apart = nullptr;
  double2* d_rpart = nullptr;
  double2* d_g = nullptr;
  double *d_Ak = nullptr, *d_kcoe = nullptr, *d_kvec = nullptr;
  KChunk* d_chunks = nullptr;
  double *d_part = nullptr, *d_sc = nullptr, *d_hist = nullptr;
  size_t part_cap = 0;
  void* d_t2 = nullptr;  // SCF pair tensor table (double or float)
  size_t t2_cap = 0;
  bool t2_in_static = false;  // the next static FP32 pass writes the FP32 table
  bool wt_ready = false;      // the next ES_IND call's dipole weights are written
  // device-resident outputs of the current call, copied before its final
  // synchronisation (null: none)
  double *out_force = nullptr, *out_mu = nullptr, *out_ef = nullptr, *out_efg = nullptr;
  // standalone EFG (mdffgpu_efg): contributing moments, real-space
  // partials, image cells of the direct sum, the (natm,3,3) result
  double *d_efgq = nullptr, *d_efgmu = nullptr, *d_efgpart = nullptr, *d_efgcells = nullptr,
         *d_efgout = nullptr;
  size_t efgpart_cap = 0, efgcells_cap = 0;
  double* d_rref = nullptr;  // positions at the last list build
  bool list_valid = false;
  int list_builds = 0;
  PairTables* d_pt = nullptr;
  TypeTables* d_tt = nullptr;
  PairTables pt_h;
  TypeTables tt_h;
  TTTable tt;
  double u_pol = 0.0;
  double* d_upart = nullptr;  // block partials of u_pol (reduced once per SCF)
  double* d_cg = nullptr;     // conjugate-gradient SCF: r, z, p, Ap (3N each)
  double* cg_Ap = nullptr;    // set while ES_IND calls are the CG's T p products
  int cg_npap = 0;            // block partials of p.Ap the last such call wrote
  bool upol_pending = false;
  double* h_sc = nullptr;  // pinned
  double* h_sc_dev = nullptr;  // device address of h_sc
  hipEvent_t ev_sc = nullptr;  // the SCF residual has landed in h_sc
  int* h_ovf = nullptr;    // pinned
  // allocation-free mode (mdffgpu_engforce_noalloc): the device buffers
  // are carved from the library's static scratch array, the host scalars
  // live in pageable memory, the stream is the caller's
  bool noalloc = false;
  char* arena = nullptr;
  size_t arena_cap = 0, arena_off = 0;
  std::vector<double> h_sc_buf;
  std::vector<int> h_ovf_buf;
  hipStream_t own_stream = nullptr;
  hipStream_t st = nullptr;
  int linduced = 0, ldip_static = 0;
  mixed::MixedState* mx = nullptr;
  // kernel instrumentation: HIP events around each launch of the own stream
  int prof = 0;
  long long evals = 0;  // engforce evaluations run (positions may change between)
  int prof_only = -1;  // with prof: time only this kernel id (-1: all)
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> prof_pending;
  std::vector<hipEvent_t> ev_pool;
  double prof_ms[mdff::K_COUNT] = {0};
  long long prof_n[mdff::K_COUNT] = {0};
};

namespace mdff {
const char* const kKernelNames[K_COUNT] = {
    "k_prepare", "k_nlist", "k_shortrange", "k_es_dir<static>", "k_es_dir<ind>",
    "k_es_dir<final>", "k_rho64<chg>", "k_rho64<dip>", "k_rho64<chg+dip>", "k_rho_reduce64",
    "k_apply64<field>", "k_apply64<efg>", "k_moments", "k_combine64", "k_induced",
    "k_scf_resid", "k_extrap", "k_reduce_rows", "k_add", "k_pairs_t2", "k_dd_field",
    "mx_shortrange", "mx_es_dir",
    "mx_phase", "mx_rho", "mx_rho_reduce", "mx_apply", "mx_combine", "mx_weights", "mx_gprep",
    "batch", "adjoint", "k_efg_es_dir", "k_efg_ds", "k_efg_combine", "md", "k_es_quad", "k_wfc"};

ProfScope::ProfScope(mdffgpu_ctx* ctx, int kid) : c(ctx), id(kid), a(nullptr), b(nullptr) {
  if (!c || !c->prof || (c->prof_only != -1 && c->prof_only != kid)) return;
  if (c->ev_pool.size() < 2) {
    // timing only, read after a stream synchronisation: no system-scope
    // fence at record (it writes back and invalidates the caches, a gap of
    // several microseconds before the next launch)
    for (int k = 0; k < 2; ++k) {
      hipEvent_t e;
      if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) == hipSuccess)
        c->ev_pool.push_back(e);
    }
  }
  if (c->ev_pool.size() < 2) return;
  a = c->ev_pool.back();
  c->ev_pool.pop_back();
  b = c->ev_pool.back();
  c->ev_pool.pop_back();
  (void)hipEventRecord(a, c->st);
}

ProfScope::~ProfScope() noexcept(false) {
  if (!a) return;
  (void)hipEventRecord(b, c->st);
  c->prof_pending.push_back({id, {a, b}});
}
}  // namespace mdff

namespace {

void fail(mdffgpu_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c)
    c->err = buf;
  else
    g_create_error = buf;
  throw Fail{code};
}

void hipck(mdffgpu_ctx* c, hipError_t e, const char* what) {
  if (e != hipSuccess) {
    fail(c, e == hipErrorOutOfMemory ? MDFFGPU_ERR_ALLOC : MDFFGPU_ERR_DEVICE,
         "%s: %s", what, hipGetErrorString(e));
  }
}

#define HCK(c, x) hipck((c), (x), #x)

// Next 256-byte aligned block of the allocation-free arena.  Blocks are
// never returned: the arena belongs to one evaluation.
void* arena_take(mdffgpu_ctx* c, size_t bytes) {
  const size_t off = (c->arena_off + 255) & ~(size_t)255;
  if (off + bytes > c->arena_cap)
    fail(c, MDFFGPU_ERR_ALLOC,
         "the allocation-free evaluation needs more than the %zu MiB of static scratch "
         "(natm, neighbour list or k mesh too large for it)",
         c->arena_cap >> 20);
  c->arena_off = off + bytes;
  return c->arena + off;
}

template <typename T>
void dalloc(mdffgpu_ctx* c, T** p, size_t count) {
  if (count == 0) count = 1;
  if (c->noalloc) {
    *p = static_cast<T*>(arena_take(c, count * sizeof(T)));
    return;
  }
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  hipck(c, hipMalloc((void**)p, count * sizeof(T)), "hipMalloc");
}

inline int nblk(int n) { return (n + MDFF_BLOCK - 1) / MDFF_BLOCK; }

void check_launch(mdffgpu_ctx* c, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) fail(c, MDFFGPU_ERR_DEVICE, "launch %s: %s", what, hipGetErrorString(e));
}

// After a stream synchronisation: fold the recorded intervals in.
void prof_collect(mdffgpu_ctx* c) {
  for (auto& p : c->prof_pending) {
    float ms = 0.f;
    HCK(c, hipEventElapsedTime(&ms, p.second.first, p.second.second));
    c->prof_ms[p.first] += ms;
    c->prof_n[p.first] += 1;
    c->ev_pool.push_back(p.second.first);
    c->ev_pool.push_back(p.second.second);
  }
  c->prof_pending.clear();
}

// 3x3 inverse, Gauss-Jordan with partial pivoting (DGETRF/DGETRI of
// src/read_pos.f90:201-210 for the polarizability); singular -> zeros.

// lattice (src/lattice.f90:60-168)
void set_lattice(mdffgpu_ctx* c, const double* A) {
  if (!make_cell(A, c->cell, c->W)) fail(c, MDFFGPU_ERR_ARG, "cell volume is not positive");
}

// Half-space k mesh of kpoint_sum_init (src/kspace.f90:142-207) arranged
// in chunks of KC consecutive n3; padding slots carry A_k = 0.
void build_kmesh(mdffgpu_ctx* c) {
  const mdffgpu_params& P = c->P;
  for (int a = 0; a < 3; ++a) {
    c->kmax[a] = P.kES[a];
    if (P.kES[a] < 0 || P.kES[a] > 512) fail(c, MDFFGPU_ERR_ARG, "kES out of range");
  }
  const int kx = c->kmax[0], ky = c->kmax[1], kz = c->kmax[2];
  const int L3 = 2 * kz + 1;
  const int nch3 = (L3 + KC_MAX - 1) / KC_MAX;
  c->KC = (L3 + nch3 - 1) / nch3;
  const int KC = c->KC;
  std::vector<KChunk>& ch = c->chunks_h;
  ch.clear();
  std::vector<double> Ak, kcoe, kv[3];
  const double tpi = 2.0 * MDFF_PI;
  const double alpha2 = P.alphaES * P.alphaES;
  const double* B = c->cell.B;
  int nk = 0;
  for (int n1 = 0; n1 <= kx; ++n1) {
    int n2min = (n1 == 0) ? 0 : -ky;
    for (int n2 = n2min; n2 <= ky; ++n2) {
      int lo = (n1 == 0 && n2 == 0) ? 1 : -kz;
      if (lo > kz) continue;
      for (int s = lo; s <= kz; s += KC) {
        KChunk k;
        k.n1 = n1;
        k.n2 = n2;
        k.n3s = s;
        k.pad = 0;
        ch.push_back(k);
        for (int cc = 0; cc < KC; ++cc) {
          int n3 = s + cc;
          if (n3 <= kz) {
            double x = n1, y = n2, z = n3;
            double vx = tpi * (x * B[0] + y * B[3] + z * B[6]);
            double vy = tpi * (x * B[1] + y * B[4] + z * B[7]);
            double vz = tpi * (x * B[2] + y * B[5] + z * B[8]);
            double kk = vx * vx + vy * vy + vz * vz;
            Ak.push_back(std::exp(-kk * 0.25 / alpha2) / kk);
            kcoe.push_back(2.0 * (1.0 / kk + 1.0 / alpha2 / 4.0));
            kv[0].push_back(vx);
            kv[1].push_back(vy);
            kv[2].push_back(vz);
            ++nk;
          } else {
            Ak.push_back(0.0);
            kcoe.push_back(0.0);
            kv[0].push_back(0.0);
            kv[1].push_back(0.0);
            kv[2].push_back(0.0);
          }
        }
      }
    }
  }
  c->nk = nk;
  c->nchunk = (int)ch.size();
  c->kpad = c->nchunk * KC;
  if (c->nchunk == 0) return;
  std::vector<double> kvec(3 * (size_t)c->kpad);
  for (int a = 0; a < 3; ++a)
    std::memcpy(kvec.data() + (size_t)a * c->kpad, kv[a].data(), c->kpad * sizeof(double));
  dalloc(c, &c->d_chunks, c->nchunk);
  dalloc(c, &c->d_Ak, c->kpad);
  dalloc(c, &c->d_kcoe, c->kpad);
  dalloc(c, &c->d_kvec, 3 * (size_t)c->kpad);
  dalloc(c, &c->d_g, c->kpad);
  HCK(c, hipMemcpy(c->d_chunks, ch.data(), ch.size() * sizeof(KChunk), hipMemcpyHostToDevice));
  HCK(c, hipMemcpy(c->d_Ak, Ak.data(), Ak.size() * sizeof(double), hipMemcpyHostToDevice));
  HCK(c, hipMemcpy(c->d_kcoe, kcoe.data(), kcoe.size() * sizeof(double), hipMemcpyHostToDevice));
  HCK(c, hipMemcpy(c->d_kvec, kvec.data(), kvec.size() * sizeof(double), hipMemcpyHostToDevice));
  // work split so that a launch has about 1024 blocks
  const int n = c->n;
  int gx = (c->nchunk + MDFF_BLOCK - 1) / MDFF_BLOCK;
  c->rho_split = std::max(1, std::min((1024 + gx - 1) / gx, (n + RHO_TILE - 1) / RHO_TILE));
  c->rho_atoms = (n + c->rho_split - 1) / c->rho_split;
  c->rho_split = (n + c->rho_atoms - 1) / c->rho_atoms;
  int ax = nblk(n);
  c->app_split = std::max(1, std::min((1024 + ax - 1) / ax, c->nchunk));
  c->app_chunks = (c->nchunk + c->app_split - 1) / c->app_split;
  c->app_split = (c->nchunk + c->app_chunks - 1) / c->app_chunks;
  dalloc(c, &c->d_rpart, (size_t)c->rho_split * c->kpad);
  dalloc(c, &c->d_apart, (size_t)c->app_split * 9 * n);
  // rows a partial call leaves alone hold finite values
  HCK(c, hipMemsetAsync(c->d_apart, 0, (size_t)c->app_split * 9 * n * sizeof(double), c->st));
}

void build_tables(mdffgpu_ctx* c) {
  const int bad = make_tables(c->P, c->pt_h, c->tt_h, c->tt);
  if (bad >= 0)
    fail(c, MDFFGPU_ERR_TT_ORDER, "Tang-Toennies order %d > %d", bad, MDFFGPU_TT_MAX_ORDER);
  c->linduced = 0;
  c->ldip_static = 0;
  for (int k = 0; k < 48; ++k)
    if (c->P.dip[k] != 0.0) c->ldip_static = 1;
  c->polmask = 0;
  c->fieldmask = 0;
  for (int t = 0; t < c->P.ntype; ++t) {
    if (c->tt_h.polar[t] || c->tt_h.qpolar[t]) {
      c->linduced = 1;
      c->fieldmask |= 1u << t;
    }
    if (c->tt_h.polar[t]) c->polmask |= 1u << t;
  }
  HCK(c, hipMemcpy(c->d_pt, &c->pt_h, sizeof(c->pt_h), hipMemcpyHostToDevice));
  HCK(c, hipMemcpy(c->d_tt, &c->tt_h, sizeof(c->tt_h), hipMemcpyHostToDevice));
}

void validate(mdffgpu_ctx* c, const mdffgpu_params* p) {
  if (!p) fail(c, MDFFGPU_ERR_ARG, "params is NULL");
  if (p->natm < 1) fail(c, MDFFGPU_ERR_ARG, "natm must be positive");
  if (p->ntype < 1 || p->ntype > MDFFGPU_NTYPEMAX) fail(c, MDFFGPU_ERR_ARG, "ntype out of range");
  if ((p->lnmlj || p->lbmhft || p->lbmhftd || p->lmorse) && !(p->cutshortrange > 0.0))
    fail(c, MDFFGPU_ERR_ARG, "cutshortrange is null");
  if (p->lcoulomb && !(p->cutlongrange > 0.0)) fail(c, MDFFGPU_ERR_ARG, "cutlongrange is null");
  if (p->lcoulomb && !(p->alphaES > 0.0)) fail(c, MDFFGPU_ERR_ARG, "alphaES must be positive");
  if (p->lcoulomb && (p->kES[0] < 1 || p->kES[1] < 1 || p->kES[2] < 1))
    fail(c, MDFFGPU_ERR_ARG, "kES must be at least 1 along every axis");
  if (p->algo_ext_dipole == MDFFGPU_EXT_POLY && (p->extrapolate_order < 0 || p->extrapolate_order > 1))
    fail(c, MDFFGPU_ERR_UNSUPPORTED,
         "polynomial dipole extrapolation of order %d is not supported (0 or 1)",
         p->extrapolate_order);
  if (p->algo_ext_dipole == MDFFGPU_EXT_ASPC && (p->extrapolate_order < 0 || p->extrapolate_order > 5))
    fail(c, MDFFGPU_ERR_ARG, "ASPC extrapolation order must be 0..5");
  if (p->algo_ext_dipole != MDFFGPU_EXT_POLY && p->algo_ext_dipole != MDFFGPU_EXT_ASPC)
    fail(c, MDFFGPU_ERR_ARG, "unknown algo_ext_dipole");
  if (p->precision != MDFFGPU_PREC_FP64 && p->precision != MDFFGPU_PREC_MIXED)
    fail(c, MDFFGPU_ERR_ARG, "unknown precision");
  if (p->max_scf_pol_iter < 0 || p->min_scf_pol_iter < 0)
    fail(c, MDFFGPU_ERR_ARG, "SCF iteration bounds must be non-negative");
  if (p->scf_solver != MDFFGPU_SCF_JACOBI && p->scf_solver != MDFFGPU_SCF_PCG)
    fail(c, MDFFGPU_ERR_ARG, "unknown scf_solver");
  if (p->lcoulomb) {
    bool anyq = p->task_coul_quad[0] || p->task_coul_quad[1] || p->task_coul_quad[2];
    bool qpol = false;
    for (int t = 0; t < p->ntype; ++t) {
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          const double v = p->quad[t + 16 * (a + 3 * b)];
          if (v != 0.0) anyq = true;
          if (v != p->quad[t + 16 * (b + 3 * a)])
            fail(c, MDFFGPU_ERR_ARG, "quad(%d,:,:) is not symmetric", t + 1);
          for (int g = 0; g < 3; ++g)
            if (p->polquad[t + 16 * (g + 3 * (a + 3 * b))] !=
                p->polquad[t + 16 * (g + 3 * (b + 3 * a))])
              fail(c, MDFFGPU_ERR_ARG, "polquad(%d,g,:,:) is not symmetric", t + 1);
        }
      if (p->lquad_polar[t]) {
        qpol = anyq = true;
        // u_pol divides by polquad(it,1,1,1) (src/field.f90:1793)
        if (p->polquad[t] == 0.0)
          fail(c, MDFFGPU_ERR_ARG, "lquad_polar(%d) with polquad(%d,1,1,1) = 0", t + 1, t + 1);
      }
    }
    if (anyq && p->precision != MDFFGPU_PREC_FP64)
      fail(c, MDFFGPU_ERR_UNSUPPORTED, "quadrupoles run on the FP64 path only");
    if (qpol && p->scf_solver != MDFFGPU_SCF_JACOBI)
      fail(c, MDFFGPU_ERR_UNSUPPORTED,
           "quadrupole polarizabilities need the Jacobi SCF (scf_solver = 0)");
    bool wfc = false;
    for (int t = 0; t < p->ntype; ++t)
      if (p->lwfc[t] > 0) wfc = true;
    if (wfc && !(p->rcut_wfc > 0.0)) fail(c, MDFFGPU_ERR_ARG, "rcut_wfc must be positive");
  }
  // Tang-Toennies orders (src/field.f90:355-358), before any device work
  std::unique_ptr<PairTables> pt(new PairTables());
  std::unique_ptr<TypeTables> ty(new TypeTables());
  TTTable t3;
  const int bad = make_tables(*p, *pt, *ty, t3);
  if (bad >= 0)
    fail(c, MDFFGPU_ERR_TT_ORDER, "Tang-Toennies order %d > %d", bad, MDFFGPU_TT_MAX_ORDER);
}

void ensure_part(mdffgpu_ctx* c, size_t count) {
  if (count > c->part_cap) {
    dalloc(c, &c->d_part, count);
    c->part_cap = count;
  }
}

void reduce(mdffgpu_ctx* c, int rows, int nb, int slot) {
  {
    ProfScope _ps(c, K_REDUCE);
    k_reduce_rows<<<rows, MDFF_BLOCK, 0, c->st>>>(c->d_part, nb, c->d_sc + slot);
  }
  check_launch(c, "k_reduce_rows");
}

void reduce_upol(mdffgpu_ctx* c) {
  k_reduce_rows<<<1, MDFF_BLOCK, 0, c->st>>>(c->d_upart, nblk(c->n), c->d_sc + SC_UPOL);
  check_launch(c, "k_reduce_rows");
}

void alloc_atoms(mdffgpu_ctx* c) {
  const size_t n = c->n;
  dalloc(c, &c->d_pos, 3 * n);
  dalloc(c, &c->d_s, 3 * n);
  dalloc(c, &c->d_eph, 3 * n);
  dalloc(c, &c->d_typ, n);
  dalloc(c, &c->d_cnt, n);
  dalloc(c, &c->d_ovf, 1);
  dalloc(c, &c->d_nlr, (size_t)NLIST_RANGES * n);
  dalloc(c, &c->d_q, n);
  dalloc(c, &c->d_qzero, n);
  dalloc(c, &c->d_dip, 3 * n);
  dalloc(c, &c->d_mu, 3 * n);
  dalloc(c, &c->d_muind, 3 * n);
  dalloc(c, &c->d_mu4, n);
  dalloc(c, &c->d_s4, n);
  dalloc(c, &c->d_upart, nblk(c->n));
  if (c->P.scf_solver == MDFFGPU_SCF_PCG) dalloc(c, &c->d_cg, 12 * n);
  dalloc(c, &c->d_perm, n);
  dalloc(c, &c->d_rank, n);
  dalloc(c, &c->d_polrows, n);
  dalloc(c, &c->d_cellof, n);
  dalloc(c, &c->d_mup, 3 * n);
  dalloc(c, &c->d_Estat, 3 * n);
  dalloc(c, &c->d_E, 3 * n);
  dalloc(c, &c->d_efdir, 3 * n);
  dalloc(c, &c->d_efgdir, 9 * n);
  dalloc(c, &c->d_fdir, 3 * n);
  dalloc(c, &c->d_force, 3 * n);
  dalloc(c, &c->d_ef, 3 * n);
  dalloc(c, &c->d_efg, 9 * n);
  dalloc(c, &c->d_sc, SC_N);
  dalloc(c, &c->d_rref, 3 * n);
  dalloc(c, &c->d_hist, (size_t)(c->P.extrapolate_order + 1) * 3 * n);
  dalloc(c, &c->d_pt, 1);
  dalloc(c, &c->d_tt, 1);
  HCK(c, hipMemsetAsync(c->d_qzero, 0, n * sizeof(double), c->st));
  HCK(c, hipMemsetAsync(c->d_hist, 0, (size_t)(c->P.extrapolate_order + 1) * 3 * n * sizeof(double),
                        c->st));
  HCK(c, hipMemsetAsync(c->d_sc, 0, SC_N * sizeof(double), c->st));
  ensure_part(c, 64 * (size_t)nblk(c->n) + 4096);
}

void set_list_capacity(mdffgpu_ctx* c, int maxnb) {
  c->maxnb = std::max(1, std::min(maxnb, c->n));
  dalloc(c, &c->d_nbr, (size_t)c->maxnb * c->n);
}

int initial_capacity(mdffgpu_ctx* c) {
  // 1.5 x the mean number of neighbours inside the list cutoff, + 64
  const double rc = c->list_cut + c->P.skindiff;
  const double nmean = 4.0 / 3.0 * MDFF_PI * rc * rc * rc * c->n / c->cell.omega;
  return (int)std::min<double>(c->n, 1.5 * nmean + 64.0);
}

void update_cutoffs(mdffgpu_ctx* c) {
  const mdffgpu_params& P = c->P;
  double rc = 0.0;
  if (P.lnmlj || P.lbmhft || P.lbmhftd || P.lmorse) rc = std::max(rc, P.cutshortrange);
  if (P.lcoulomb) rc = std::max(rc, P.cutlongrange);
  c->list_cut = rc;
}

// ---------------------------------------------------------------------------
// One multipole_ES call on device buffers.
// FP32 SCF dipole field of the mixed path (packed rank-ordered dipoles)
bool scf_field_fp32(const mdffgpu_ctx* c) {
  return (c->P.precision == MDFFGPU_PREC_MIXED) &&
         !(c->P.mixed_fp64_parts & MDFFGPU_MIXED_FP64_SCF_FIELD);
}

// Real-space part of the SCF dipole field (ES_IND): E_dir of the dipoles mu
// into d_efdir, from the pair table of the evaluation.  With packed, the
// FP32 rank-ordered copy of mu is already in d_mu4 (k_induced wrote it).
// FP32 SCF pair table of the mixed path in d_t2: the evaluation's
// fixed-point positions (uint4 per rank), the (A2, A1) entries and the
// neighbour slots (maxnb per row), the row counts.
uint4* dd_u4(mdffgpu_ctx* c) { return reinterpret_cast<uint4*>(c->d_t2); }
float2* dd_tab(mdffgpu_ctx* c) { return reinterpret_cast<float2*>(dd_u4(c) + c->n); }
int* dd_tj(mdffgpu_ctx* c) {
  return reinterpret_cast<int*>(dd_tab(c) + (size_t)c->maxnb * c->n);
}

// Entries per lane of k_dd_tab2's pipeline: the expected row of the table
// (the polarizable neighbours inside the cutoff, from their density) with
// 15 % to spare, as 4, 6 or 8.
int dd_unroll(const mdffgpu_ctx* c) {
  const double rc = c->P.cutlongrange;
  const double mean = c->npolrows / c->cell.omega * (4.0 / 3.0) * MDFF_PI * rc * rc * rc;
  const double need = 1.15 * mean / 64.0;
  return need <= 4.0 ? 4 : need <= 6.0 ? 6 : 8;
}

template <typename IT, int U>
void launch_dd_tab2(mdffgpu_ctx* c, const float2* tab, const IT* tj, const int* tcnt,
                    const float4* m4) {
  const int nr = c->npolrows;
  k_dd_tab2<IT, U><<<dd_tab_blocks(nr), MDFF_BLOCK, 0, c->st>>>(
      c->n, c->maxnb, tab, tj, tcnt, m4, dd_u4(c), c->cell, nullptr, nullptr, c->d_perm,
      c->d_polrows, nr, c->d_efdir);
}

template <typename IT>
void launch_dd_tab2(mdffgpu_ctx* c, const float2* tab, const IT* tj, const int* tcnt,
                    const float4* m4) {
  switch (dd_unroll(c)) {
    case 4: launch_dd_tab2<IT, 4>(c, tab, tj, tcnt, m4); break;
    case 6: launch_dd_tab2<IT, 6>(c, tab, tj, tcnt, m4); break;
    default: launch_dd_tab2<IT, 8>(c, tab, tj, tcnt, m4); break;
  }
}

void es_ind_dir(mdffgpu_ctx* c, const double* mu, bool packed = false) {
  const int n = c->n;
  hipStream_t st = c->st;
  const int gb = nblk(n);
  ProfScope _ps(c, K_DD_FIELD);
  if (scf_field_fp32(c)) {
    float4* m4 = reinterpret_cast<float4*>(c->d_mu4);
    const size_t cap = (size_t)c->maxnb * n;
    const float2* tab = dd_tab(c);
    const int* tj = dd_tj(c);
    if (!packed) k_pack_mu_ranked<<<gb, MDFF_BLOCK, 0, st>>>(n, mu, c->d_rank, m4);
    // the rows of the polarizable atoms only: the field elsewhere is unused
    if (c->npolrows > 0 && dd_idx16(n))
      launch_dd_tab2(c, tab, reinterpret_cast<const unsigned short*>(tj), tj + cap, m4);
    else if (c->npolrows > 0)
      launch_dd_tab2(c, tab, tj, tj + cap, m4);
  } else {
    EsDirArgs da;
    std::memset(&da, 0, sizeof(da));
    da.n = n;
    da.maxnb = c->maxnb;
    da.cnt = c->d_cnt;
    da.nbr = c->d_nbr;
    da.perm = c->d_perm;
    da.ef = c->d_efdir;
    da.typ = c->d_typ;
    da.polmask = c->fieldmask;
    k_pack_mu<double4><<<gb, MDFF_BLOCK, 0, st>>>(n, mu, c->d_mu4);
    k_dd_field<<<nblk(n * DD_LANES), MDFF_BLOCK, 0, st>>>(da, (const double*)c->d_t2, c->d_mu4);
  }
  check_launch(c, "k_dd_field");
}

// One Ewald call; with dir_done (ES_IND only) the real-space part has
// already been launched (es_ind_dir) for these dipoles.
// th: quadrupoles (n,3,3) of the call, or null; their terms run with the
// TASK_Q*QUAD bits of task (FP64 path only).
void es_call(mdffgpu_ctx* c, int mode, const double* q, const double* mu, int task,
             int do_efield, int do_efg, bool dir_done = false, const double* th = nullptr) {
  const int n = c->n;
  hipStream_t st = c->st;
  const mdffgpu_params& P = c->P;
  const int gb = nblk(n);
  const int gbl = nblk(n * LPA);
  // real space
  EsDirArgs da;
  std::memset(&da, 0, sizeof(da));
  da.n = n;
  da.maxnb = c->maxnb;
  da.s = c->d_s;
  da.s4 = c->d_s4;
  da.typ = c->d_typ;
  da.cnt = c->d_cnt;
  da.nbr = c->d_nbr;
  da.cell = c->cell;
  da.T = c->d_pt;
  da.tt = c->tt;
  da.q = q;
  da.mu = mu;
  da.cut2 = P.cutlongrange * P.cutlongrange;
  da.alpha = P.alphaES;
  da.perm = c->d_perm;
  da.task = task;
  da.do_efield = do_efield;
  da.do_efg = do_efg;
  da.ef = c->d_efdir;
  da.efg = c->d_efgdir;
  da.f = c->d_fdir;
  da.part = c->d_part;
  const bool mixed = (P.precision == MDFFGPU_PREC_MIXED);
  const bool mx = mixed && !(P.mixed_fp64_parts & MDFFGPU_MIXED_FP64_DIRECT);
  const bool mxk = mixed && !(P.mixed_fp64_parts & MDFFGPU_MIXED_FP64_KSPACE);
  if (mode == ES_STATIC && mx && c->t2_in_static) {
    // the evaluation's FP32 SCF pair table, from the same pairs and terms
    const size_t cap = (size_t)c->maxnb * n;
    da.ddtab = dd_tab(c);
    da.ddtj = dd_tj(c);
    da.ddcnt = da.ddtj + cap;
    da.rank = c->d_rank;
    da.dd16 = dd_idx16(n);
    da.polmask = c->polmask;
    c->t2_in_static = false;
  }
  if (mode != ES_IND) {
    // the neighbours' (q, mu) in one gather, in the SCF dipole buffer
    // (free outside the SCF iterations)
    if (mx)
      k_pack_qmu<<<gb, MDFF_BLOCK, 0, st>>>(n, q, mu, reinterpret_cast<float4*>(c->d_mu4));
    else
      k_pack_qmu<<<gb, MDFF_BLOCK, 0, st>>>(n, q, mu, c->d_mu4);
    check_launch(c, "k_pack_qmu");
    da.qm = c->d_mu4;
  }
  if (mode == ES_STATIC) {
    ProfScope _ps(c, mx ? K_MX_DIR : K_DIR_STATIC);
    // FP32 static pass with the FP32 SCF field: the rows of the polarizable
    // atoms only (the static field is read there only, and their rows hold
    // the SCF pair table)
    const bool prow = mx && c->linduced && scf_field_fp32(c) && c->npolrows > 0;
    if (prow) {
      da.rows = c->d_polrows;
      da.nrows = c->npolrows;
    }
    if (mx)
      k_es_dir<ES_STATIC, float><<<prow ? nblk(c->npolrows * LPA) : gbl, MDFF_BLOCK, 0, st>>>(da);
    else
      k_es_dir<ES_STATIC, double><<<gbl, MDFF_BLOCK, 0, st>>>(da);
  } else if (mode == ES_IND) {
    if (!dir_done) es_ind_dir(c, mu);
  } else {
    ProfScope _ps(c, mx ? K_MX_DIR : K_DIR_FINAL);
    if (mx)
      k_es_dir<ES_FINAL, float><<<gbl, MDFF_BLOCK, 0, st>>>(da);
    else
      k_es_dir<ES_FINAL, double><<<gbl, MDFF_BLOCK, 0, st>>>(da);
  }
  check_launch(c, "k_es_dir");
  if (mode == ES_FINAL) reduce(c, 7, gbl, SC_DIR);
  const bool quad = th && (task & TASK_QUAD_ANY);
  if (mode == ES_FINAL) c->quad_final = quad;
  if (quad) {
    // quadrupole terms of the real-space sum, added to its field and EFG
    QuadArgs qa;
    qa.n = n;
    qa.maxnb = c->maxnb;
    qa.s = c->d_s;
    qa.s4 = c->d_s4;
    qa.cnt = c->d_cnt;
    qa.nbr = c->d_nbr;
    qa.perm = c->d_perm;
    qa.cell = c->cell;
    qa.cut2 = P.cutlongrange * P.cutlongrange;
    qa.alpha = P.alphaES;
    qa.q = q;
    qa.mu = mu;
    qa.th = th;
    qa.task = task;
    qa.do_energy = (mode == ES_FINAL);
    qa.do_efield = (mode != ES_FINAL) || do_efield;
    qa.do_efg = (mode == ES_FINAL) && do_efg;
    qa.ef = c->d_efdir;
    qa.efg = c->d_efgdir;
    qa.part = c->d_part;
    {
      ProfScope _ps(c, K_QUAD);
      k_es_quad<<<gbl, MDFF_BLOCK, 0, st>>>(qa);
    }
    check_launch(c, "k_es_quad");
    if (qa.do_energy) reduce(c, 1, gbl, SC_QUAD);
  }
  CombineArgs ca;
  std::memset(&ca, 0, sizeof(ca));
  ca.n = n;
  ca.lsurf = P.lsurf;
  ca.do_efield = do_efield;
  ca.do_efg = do_efg;
  ca.apart = c->d_apart;
  ca.cell = c->cell;
  ca.fpi_V = 2.0 * (2.0 * MDFF_PI / c->cell.omega);
  ca.fpi_3V = 2.0 * (2.0 * MDFF_PI / c->cell.omega / 3.0);
  const double selfa = P.alphaES / std::sqrt(MDFF_PI);
  ca.selfa2 = 2.0 * selfa * P.alphaES * P.alphaES / 3.0;
  ca.coul_unit = P.coul_unit;
  ca.mu = mu;
  ca.q = q;
  ca.ef_dir = c->d_efdir;
  ca.efg_dir = c->d_efgdir;
  ca.f_dir = c->d_fdir;
  ca.qm = c->d_sc + SC_MOM;
  ca.E_stat = c->d_Estat;
  ca.E = c->d_E;
  ca.ef_out = c->d_ef;
  ca.efg_out = c->d_efg;
  ca.f = c->d_force;
  if (mode == ES_IND) {
    // the SCF residual of the new field, fused (rows for SC_RMSD)
    ca.typ = c->d_typ;
    ca.T = c->d_tt;
    ca.rpart = c->d_part;
    ca.cg_Ap = c->cg_Ap;
    ca.th = c->lqpol ? c->d_thind : nullptr;
  }
  // reciprocal space
  const bool chg = (mode != ES_IND);
  const bool dip = (task & (TASK_QD | TASK_DD)) != 0;
  const bool efgv = (mode == ES_FINAL) && (do_efg || dip);
  int nsplit_apply = c->app_split;
  const bool wt_ready = c->wt_ready && mode == ES_IND;
  c->wt_ready = false;
  // the conjugate-gradient T p products: the mixed reciprocal space also
  // does the combine (no surface term)
  int fused_parts = 0;
  c->cg_npap = gb;
  if (mxk && c->nchunk > 0) {
    const bool fuse = (mode == ES_IND) && c->cg_Ap && !P.lsurf;
    fused_parts = mixed::kspace(c->mx, mode, q, mu, chg, dip, efgv, c->d_apart, wt_ready,
                                fuse ? &ca : nullptr);
    nsplit_apply = 1;
  } else if (c->nchunk > 0) {
    RhoArgs ra;
    ra.n = n;
    ra.nchunk = c->nchunk;
    ra.KC = c->KC;
    for (int a = 0; a < 3; ++a) ra.kmax[a] = c->kmax[a];
    ra.chunks = c->d_chunks;
    ra.eph = c->d_eph;
    ra.q = q;
    ra.mu = mu;
    ra.th = th;
    ra.cell = c->cell;
    ra.atoms_per_split = c->rho_atoms;
    ra.kpad = c->kpad;
    ra.part = c->d_rpart;
    // the SCF calls' moments sit on the atoms of the polarizable species
    ra.j_begin = (mode == ES_IND) ? c->kr0 : 0;
    ra.j_end = (mode == ES_IND) ? c->kr1 : n;
    const int nsp = std::max(1, (ra.j_end - ra.j_begin + c->rho_atoms - 1) / c->rho_atoms);
    dim3 g((c->nchunk + MDFF_BLOCK - 1) / MDFF_BLOCK, nsp);
    size_t lds = (size_t)RHO_TILE * (c->kmax[0] + c->kmax[1] + c->kmax[2] + 3) * sizeof(double2) +
                 10 * RHO_TILE * sizeof(double);
    // the quadrupoles enter rho only together with the dipoles
    // (src/field.f90:2723-2740 sit behind the ldip test)
    if (quad && dip) {
      ProfScope _ps(c, chg ? K_RHO_BOTH : K_RHO_DIP);
      if (chg)
        k_rho64<true, true, true><<<g, MDFF_BLOCK, lds, st>>>(ra);
      else
        k_rho64<false, true, true><<<g, MDFF_BLOCK, lds, st>>>(ra);
    } else if (chg && dip)
      {
        ProfScope _ps(c, K_RHO_BOTH);
        k_rho64<true, true><<<g, MDFF_BLOCK, lds, st>>>(ra);
      }
    else if (dip)
      {
        ProfScope _ps(c, K_RHO_DIP);
        k_rho64<false, true><<<g, MDFF_BLOCK, lds, st>>>(ra);
      }
    else
      {
        ProfScope _ps(c, K_RHO_CHG);
        k_rho64<true, false><<<g, MDFF_BLOCK, lds, st>>>(ra);
      }
    check_launch(c, "k_rho64");
    RhoRedArgs rr;
    rr.kpad = c->kpad;
    rr.nsplit = nsp;
    rr.do_energy = (mode == ES_FINAL);
    rr.part = c->d_rpart;
    rr.Ak = c->d_Ak;
    rr.kcoe = c->d_kcoe;
    rr.kvec = c->d_kvec;
    rr.g = c->d_g;
    rr.pout = c->d_part;
    const int gk = nblk(c->kpad);
    {
      ProfScope _ps(c, K_RHO_RED);
      k_rho_reduce64<<<gk, MDFF_BLOCK, 0, st>>>(rr);
    }
    check_launch(c, "k_rho_reduce64");
    if (mode == ES_FINAL) reduce(c, 7, gk, SC_REC);
    ApplyArgs aa;
    aa.n = n;
    aa.nchunk = c->nchunk;
    aa.KC = c->KC;
    aa.chunks_per_split = c->app_chunks;
    aa.chunks = c->d_chunks;
    aa.g = c->d_g;
    aa.eph = c->d_eph;
    aa.s = c->d_s;
    aa.s4 = c->d_s4;
    aa.part = c->d_apart;
    // the static and SCF calls give the field where the SCF reads it only
    aa.i_begin = (mode != ES_FINAL) ? c->kr0 : 0;
    aa.i_end = (mode != ES_FINAL) ? c->kr1 : n;
    dim3 ga(std::max(1, nblk(aa.i_end - aa.i_begin)), c->app_split);
    if (efgv)
      {
        ProfScope _ps(c, K_APPLY_EFG);
        k_apply64<true><<<ga, MDFF_BLOCK, 0, st>>>(aa);
      }
    else
      {
        ProfScope _ps(c, K_APPLY_F);
        k_apply64<false><<<ga, MDFF_BLOCK, 0, st>>>(aa);
      }
    check_launch(c, "k_apply64");
  } else {
    HCK(c, hipMemsetAsync(c->d_apart, 0, (size_t)c->app_split * 9 * n * sizeof(double), st));
    if (mode == ES_FINAL) HCK(c, hipMemsetAsync(c->d_sc + SC_REC, 0, 7 * sizeof(double), st));
  }
  // moment sums (self and surface terms)
  if (mode == ES_FINAL || P.lsurf) {
    {
      ProfScope _ps(c, K_MOMENTS);
      k_moments<<<gb, MDFF_BLOCK, 0, st>>>(n, c->d_pos, q, mu, c->d_part);
    }
    check_launch(c, "k_moments");
    reduce(c, 8, gb, SC_MOM);
  }
  if (fused_parts > 0) {
    c->cg_npap = fused_parts;
    check_launch(c, "mx_apply_epi_cg");
    return;
  }
  ca.nsplit = (c->nchunk > 0) ? nsplit_apply : 1;
  ca.dip_force = dip;
  if (mode == ES_STATIC)
    {
      ProfScope _ps(c, K_COMBINE);
      k_combine64<ES_STATIC><<<gb, MDFF_BLOCK, 0, st>>>(ca);
    }
  else if (mode == ES_IND)
    {
      ProfScope _ps(c, K_COMBINE);
      k_combine64<ES_IND><<<gb, MDFF_BLOCK, 0, st>>>(ca);
    }
  else
    {
      ProfScope _ps(c, K_COMBINE);
      k_combine64<ES_FINAL><<<gb, MDFF_BLOCK, 0, st>>>(ca);
    }
  check_launch(c, "k_combine64");
}

// Pair table of the SCF dipole field (k_pairs_t2): once per evaluation,
// positions are fixed during the SCF.
void build_t2(mdffgpu_ctx* c) {
  const int n = c->n;
  const bool mx = (c->P.precision == MDFFGPU_PREC_MIXED) &&
                  !(c->P.mixed_fp64_parts & MDFFGPU_MIXED_FP64_SCF_FIELD);
  // FP64: five doubles per list entry; mixed: the fixed-point positions,
  // (A2, A1) float2 + index per entry, the per-atom counts
  const size_t cap = (size_t)c->maxnb * n;
  const size_t need = mx ? (size_t)n * sizeof(uint4) + cap * (sizeof(float2) + sizeof(int)) +
                               (size_t)n * sizeof(int)
                         : 5 * cap * sizeof(double);
  if (need > c->t2_cap) {
    char* t2 = static_cast<char*>(c->d_t2);
    dalloc(c, &t2, need);
    c->d_t2 = t2;
    c->t2_cap = need;
  }
  EsDirArgs da;
  std::memset(&da, 0, sizeof(da));
  da.n = n;
  da.maxnb = c->maxnb;
  da.s = c->d_s;
  da.s4 = c->d_s4;
  da.cnt = c->d_cnt;
  da.nbr = c->d_nbr;
  da.cell = c->cell;
  da.cut2 = c->P.cutlongrange * c->P.cutlongrange;
  da.alpha = c->P.alphaES;
  da.rank = c->d_rank;
  da.perm = c->d_perm;
  da.dd16 = dd_idx16(n);
  da.polmask = c->polmask;
  // the FP32 table with an FP32 static real-space pass: written by that
  // pass (the SCF's first Ewald call, after this one in every evaluation)
  if (mx) {
    // positions are fixed during the SCF: the fixed-point copy of this
    // evaluation, in the rank order of the dipole slots
    k_pack_u4<<<nblk(n), MDFF_BLOCK, 0, c->st>>>(n, c->d_s, c->d_rank, dd_u4(c));
    check_launch(c, "k_pack_u4");
  }
  c->t2_in_static = mx && !(c->P.mixed_fp64_parts & MDFFGPU_MIXED_FP64_DIRECT);
  if (c->t2_in_static) return;
  ProfScope _ps(c, K_PAIRS_T2);
  if (mx) {
    int* tj = dd_tj(c);
    k_dd_table<false><<<nblk(n * 64), MDFF_BLOCK, 0, c->st>>>(da, dd_tab(c), tj, tj + cap);
  } else {
    da.typ = c->d_typ;
    da.polmask = c->fieldmask;
    k_pairs_t2<<<nblk(n * LPA), MDFF_BLOCK, 0, c->st>>>(da, (double*)c->d_t2);
  }
  check_launch(c, "k_pairs_t2");
}

void sync_scalars(mdffgpu_ctx* c) {
  HCK(c, hipMemcpyAsync(c->h_sc, c->d_sc, SC_N * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HCK(c, hipStreamSynchronize(c->st));
  prof_collect(c);
}

void extrap(mdffgpu_ctx* c, int op, int ord, double w, double* rpart = nullptr) {
  ExtrapArgs ea;
  ea.n = c->n;
  ea.nh = c->P.extrapolate_order + 1;
  ea.op = op;
  ea.ord = ord;
  for (int k = 0; k < 6; ++k) ea.b[k] = (ord >= 0 && ord < 6) ? B_ASPC[ord][k] : 0.0;
  ea.w = w;
  ea.hist = c->d_hist;
  ea.mu = c->d_muind;
  ea.mup = c->d_mup;
  ea.E = c->d_E;
  ea.typ = c->d_typ;
  ea.T = c->d_tt;
  ea.part = c->d_upart;
  ea.rpart = rpart;
  ea.th = c->lqpol ? c->d_thind : nullptr;
  {
    ProfScope _ps(c, K_EXTRAP);
    k_extrap<<<nblk(c->n), MDFF_BLOCK, 0, c->st>>>(ea);
  }
  check_launch(c, "k_extrap");
}

int aspc_order(const mdffgpu_params& P, int itime) {
  if (itime >= P.extrapolate_order + 1) return P.extrapolate_order;
  if (itime == 1 || itime == 0) return 0;
  return itime - 1;
}

// moment_from_pola_scf (src/field.f90:3200-3458) on device; returns the
// iteration count and final rmsd.
void scf(mdffgpu_ctx* c, int itime, int* n_scf, double* rmsd_out, int* n_es,
         double* tol_out) {
  const mdffgpu_params& P = c->P;
  const int n = c->n;
  const int gb = nblk(n);
  hipStream_t st = c->st;
  c->cg_Ap = nullptr;
  HCK(c, hipMemsetAsync(c->d_muind, 0, 3 * (size_t)n * sizeof(double), st));
  if (c->lqpol) HCK(c, hipMemsetAsync(c->d_thind, 0, 9 * (size_t)n * sizeof(double), st));
  *n_scf = 0;
  *rmsd_out = 0.0;
  *tol_out = P.conv_tol_ind;
  if (!c->linduced) return;
  // static moments: every task when there are static quadrupoles
  // (src/field.f90:3255-3268)
  int task_static = TASK_QQ | ((c->ldip_static || c->lqstat) ? (TASK_QD | TASK_DD) : 0) |
                    (c->lqstat ? TASK_QUAD_ANY : 0);
  es_call(c, ES_STATIC, c->d_q, c->d_dip, task_static, 1, 0, false,
          c->lqstat ? c->d_thstat : nullptr);
  ++*n_es;
  HCK(c, hipMemcpyAsync(c->d_E, c->d_Estat, 3 * (size_t)n * sizeof(double),
                        hipMemcpyDeviceToDevice, st));
  int iscf = 0;
  double rmsd = 1.79769313486231570e308;
  bool upol_new = false;
  const bool mixed_prec = (P.precision == MDFFGPU_PREC_MIXED);
  // Mixed precision: the FP64 residual of FP32 fields has a noise floor
  // that may lie above the tolerance.  The context keeps an estimate of
  // that floor (the best residual of the last stalled solve) and stops at
  // 1.5x the estimate; without an estimate, or well above it, a solve is
  // taken as stalled when two successive iterations fail to improve on the
  // best residual while the residual is below 0.1 of the first one (the
  // Jacobi residual need not fall monotonically; a diverging iteration
  // stays above the first residual).
  const double tol = mixed_prec ? std::max(P.conv_tol_ind, 1.5 * c->mx_floor) : P.conv_tol_ind;
  *tol_out = tol;
  bool stalled = false;
  if (mixed_prec) {
    // scale of the static field (mu_ind is still zero here): caps the floor
    {
      ProfScope _ps(c, K_RESID);
      k_scf_resid<<<gb, MDFF_BLOCK, 0, st>>>(n, c->d_typ, c->d_tt, c->d_muind, c->d_Estat,
                                             c->d_part);
    }
    check_launch(c, "k_scf_resid");
    reduce(c, 2, gb, SC_ESTAT);
  }
  double rmsd_best = rmsd, rmsd_first = 0.0;
  int no_progress = 0;
  // Mixed precision: while the host waits for an iteration's residual, the
  // next iteration's first part is already queued behind it (new moments
  // mu = alpha E and their real-space field), so that the GPU does not idle
  // through the round trip.  A solve that stops there keeps those moments:
  // one Jacobi step past the accepted iterate, with a change below the
  // residual just accepted.  The FP64 path follows the reference's
  // sequence exactly.
  const bool spec = mixed_prec;
  if (spec && !c->ev_sc) {
    HCK(c, hipEventCreateWithFlags(&c->ev_sc, hipEventDisableTiming));
    HCK(c, hipHostGetDevicePointer((void**)&c->h_sc_dev, c->h_sc, 0));
  }
  // with the FP32 SCF field, k_induced also packs the moments for k_dd_tab2
  auto induced = [&](bool pack) {
    {
      ProfScope _ps(c, K_INDUCED);
      k_induced<<<gb, MDFF_BLOCK, 0, st>>>(n, c->d_typ, c->d_tt, c->d_E, c->d_muind, c->d_upart,
                                           c->d_rank,
                                           pack ? reinterpret_cast<float4*>(c->d_mu4) : nullptr,
                                           c->lqpol ? c->d_thind : nullptr);
    }
    check_launch(c, "k_induced");
    upol_new = true;
  };
  const bool pack = spec && scf_field_fp32(c);
  bool queued = false;  // this iteration's moments and real-space field are queued
  if (P.scf_solver == MDFFGPU_SCF_PCG) {
    // Preconditioned conjugate gradient on (alpha^-1 - T) mu = E_stat from
    // the same initial guess as the Jacobi iteration; one ES_IND call per
    // iteration (the first on the guess, then on the search directions).
    // The residual r = E_stat + T mu - alpha^-1 mu is kept by the recursion;
    // for a diagonal alpha its squares are the reference's criterion
    // (mu_a / alpha_aa - E_a)^2.  While the host reads an iteration's
    // residual, the next direction and its real-space field are queued.
    const bool pk = scf_field_fp32(c);
    CgArgs g;
    std::memset(&g, 0, sizeof(g));
    g.n = n;
    g.typ = c->d_typ;
    g.T = c->d_tt;
    g.x = c->d_muind;
    g.r = c->d_cg;
    g.z = g.r + 3 * (size_t)n;
    g.p = g.z + 3 * (size_t)n;
    g.Ap = g.p + 3 * (size_t)n;
    g.E = c->d_E;
    g.Estat = c->d_Estat;
    g.sc = c->d_sc;
    g.part = c->d_part;
    g.rank = c->d_rank;
    g.p4 = pk ? reinterpret_cast<float4*>(c->d_mu4) : nullptr;
    // with the mixed reciprocal space, k_cg_dir also writes the FP32
    // structure-factor weights of the direction
    if (mixed_prec && !(P.mixed_fp64_parts & MDFFGPU_MIXED_FP64_KSPACE) && c->mx &&
        c->nchunk > 0) {
      g.wt = mixed::dipole_weights(c->mx, &g.wt_npad);
      g.wscale = mixed::DIP_WEIGHT_SCALE;
      for (int k = 0; k < 9; ++k) g.wB[k] = c->cell.B[k];
    }
    if (!c->ev_sc) {
      HCK(c, hipEventCreateWithFlags(&c->ev_sc, hipEventDisableTiming));
      HCK(c, hipHostGetDevicePointer((void**)&c->h_sc_dev, c->h_sc, 0));
    }
    ++iscf;
    if (P.calc_mode == 0)
      induced(false);
    else if (P.algo_ext_dipole == MDFFGPU_EXT_POLY)
      extrap(c, (P.extrapolate_order == 0) ? 0 : (itime > P.extrapolate_order ? 1 : 2), 0, 0.0);
    else
      extrap(c, 3, aspc_order(P, itime), 0.0);
    es_call(c, ES_IND, c->d_qzero, c->d_muind, TASK_DD, 1, 0);
    ++*n_es;
    k_cg_init<<<gb, MDFF_BLOCK, 0, st>>>(g);
    check_launch(c, "k_cg_init");
    int cur = 0;  // triple of the current r.z, |r|^2, npol
    auto base = [](int k) { return SC_CG + 3 * k; };
    // the initial sums into the device scalars and mapped host memory; the
    // first T p product is queued before the host reads them (unused if the
    // initial iterate already meets the tolerance)
    {
      ProfScope _ps(c, K_REDUCE);
      k_reduce_rows2<<<3, MDFF_BLOCK, 0, st>>>(c->d_part, gb, c->d_sc + base(cur),
                                               c->h_sc_dev + base(cur));
    }
    check_launch(c, "k_reduce_rows2");
    HCK(c, hipEventRecord(c->ev_sc, st));
    bool first_queued = iscf < P.max_scf_pol_iter;
    if (first_queued) {
      c->cg_Ap = g.Ap;
      es_call(c, ES_IND, c->d_qzero, g.p, TASK_DD, 1, 0);
      c->cg_Ap = nullptr;
    }
    HCK(c, hipEventSynchronize(c->ev_sc));
    rmsd = std::sqrt(c->h_sc[base(cur) + 1] / c->h_sc[base(cur) + 2]);
    rmsd_first = rmsd;
    rmsd_best = rmsd;
    bool queued_p = false;  // the direction and its real-space field are queued
    double rmsd_prev = 0.0;  // the residual before the current one
    while (((iscf < P.max_scf_pol_iter && rmsd > tol) || iscf < P.min_scf_pol_iter) && !stalled) {
      ++iscf;
      // T p by an ES_IND call whose combine forms Ap = alpha^-1 p - T p and
      // the partials of p.Ap (row 0); the update sums those itself and
      // writes its own partials to rows 1..3
      if (first_queued) {
        first_queued = false;  // the product queued after the initial sums
      } else {
        c->cg_Ap = g.Ap;
        es_call(c, ES_IND, c->d_qzero, g.p, TASK_DD, 1, 0, queued_p && pk);
        c->cg_Ap = nullptr;
      }
      queued_p = false;
      ++*n_es;
      g.s_rz = base(cur);
      g.pap_part = c->d_part;
      g.npap = c->cg_npap;
      g.part = c->d_part + std::max(gb, c->cg_npap);
      k_cg_update<<<gb, MDFF_BLOCK, 0, st>>>(g);
      check_launch(c, "k_cg_update");
      // the new direction from the update's sums, which the same kernel
      // stores for the next update and for the host's convergence test
      const int nxt = cur ^ 1;
      g.sc_out = c->d_sc + base(nxt);
      g.sc_host = c->h_sc_dev + base(nxt);
      k_cg_dir<<<gb, MDFF_BLOCK, 0, st>>>(g);
      check_launch(c, "k_cg_dir");
      c->wt_ready = (g.wt != nullptr);
      HCK(c, hipEventRecord(c->ev_sc, st));
      // the new direction's real-space field is queued behind the residual
      // unless the last two residuals, extrapolated geometrically, put this
      // one below the tolerance (then the solve most likely stops here)
      const bool last_likely =
          iscf >= P.min_scf_pol_iter &&
          (iscf >= P.max_scf_pol_iter ||
           (iscf >= 2 && rmsd_prev > 0.0 && rmsd * (rmsd / rmsd_prev) < tol));
      queued_p = pk && !last_likely;
      if (queued_p) es_ind_dir(c, g.p, true);
      HCK(c, hipEventSynchronize(c->ev_sc));
      cur = nxt;
      rmsd_prev = rmsd;
      rmsd = std::sqrt(c->h_sc[base(cur) + 1] / c->h_sc[base(cur) + 2]);
      if (rmsd < rmsd_best) {
        rmsd_best = rmsd;
        no_progress = 0;
      } else {
        ++no_progress;
      }
      if (mixed_prec && iscf >= std::max(P.min_scf_pol_iter, 3) && no_progress >= 2 &&
          rmsd < 0.1 * rmsd_first && (c->mx_floor == 0.0 || rmsd < 5.0 * c->mx_floor))
        stalled = true;
    }
    g.part = c->d_upart;
    k_cg_upol<<<gb, MDFF_BLOCK, 0, st>>>(g);
    check_launch(c, "k_cg_upol");
    upol_new = true;
  }
  while (P.scf_solver != MDFFGPU_SCF_PCG &&
         ((iscf < P.max_scf_pol_iter && rmsd > tol) || iscf < P.min_scf_pol_iter) && !stalled) {
    ++iscf;
    if (queued) {
      // k_induced and the real-space field were queued by the last iteration
    } else if (iscf != 1 || P.calc_mode == 0) {
      induced(false);
    } else if (P.algo_ext_dipole == MDFFGPU_EXT_POLY) {
      int op = (P.extrapolate_order == 0) ? 0 : (itime > P.extrapolate_order ? 1 : 2);
      extrap(c, op, 0, 0.0);
    } else {
      extrap(c, 3, aspc_order(P, itime), 0.0);
    }
    // induced moments: dipole-dipole and quadrupole-quadrupole tasks
    // (src/field.f90:3304-3306)
    es_call(c, ES_IND, c->d_qzero, c->d_muind, TASK_DD | (c->lqpol ? TASK_QUADQ : 0), 1, 0,
            queued, c->lqpol ? c->d_thind : nullptr);
    queued = false;
    ++*n_es;
    if (iscf == 1 && P.calc_mode == 1 && P.algo_ext_dipole == MDFFGPU_EXT_ASPC) {
      // the corrector, then the residual rows again from the corrected
      // moments (the reference tests convergence after the corrector)
      const int eo = aspc_order(P, itime);
      extrap(c, 5, eo, W_ASPC[eo], c->d_part);
      upol_new = true;
    }
    if (spec) {
      // residual rows of the ES_IND combine, straight into pinned memory
      {
        ProfScope _ps(c, K_REDUCE);
        k_reduce_rows2<<<2, MDFF_BLOCK, 0, st>>>(c->d_part, gb, c->d_sc + SC_RMSD,
                                                 c->h_sc_dev + SC_RMSD);
      }
      check_launch(c, "k_reduce_rows2");
      HCK(c, hipEventRecord(c->ev_sc, st));
      induced(pack);
      es_ind_dir(c, c->d_muind, pack);
      queued = true;
      // the profiling events of the queued launches are folded in at the
      // next stream synchronisation
      HCK(c, hipEventSynchronize(c->ev_sc));
    } else {
      reduce(c, 2, gb, SC_RMSD);  // residual rows of the ES_IND combine
      sync_scalars(c);
    }
    rmsd = std::sqrt(c->h_sc[SC_RMSD] / c->h_sc[SC_RMSD + 1]);
    if (iscf == 1) rmsd_first = rmsd;
    if (rmsd < rmsd_best) {
      rmsd_best = rmsd;
      no_progress = 0;
    } else {
      ++no_progress;
    }
    if (mixed_prec && iscf >= std::max(P.min_scf_pol_iter, 3) && no_progress >= 2 &&
        rmsd < 0.1 * rmsd_first && (c->mx_floor == 0.0 || rmsd < 5.0 * c->mx_floor))
      stalled = true;
  }
  if (mixed_prec) {
    c->scf_done = true;
    c->scf_stalled = stalled;
    c->scf_best = rmsd_best;
    c->scf_last = rmsd;
  }
  extrap(c, 4, 0, 0.0);
  // u_pol of the last induced moments: the partials of the last k_induced /
  // ASPC corrector, read with the evaluation's final scalars
  if (upol_new) {
    reduce_upol(c);
    c->upol_pending = true;
  }
  *n_scf = iscf;
  *rmsd_out = rmsd;
}

void sym9(const double* t6, double scale, double* out) {
  // t6: xx yy zz xy xz yz -> Fortran (3,3)
  out[0] = t6[0] * scale;
  out[4] = t6[1] * scale;
  out[8] = t6[2] * scale;
  out[1] = out[3] = t6[3] * scale;
  out[2] = out[6] = t6[4] * scale;
  out[5] = out[7] = t6[5] * scale;
}

// Cell grid of the list build: cells of at least 4 A and at least a third of
// the list cutoff across (at most 64 per axis), so that the stencil
// half-width m = ceil(lc nc / W) is at most 3 and its rows fit one wave.
// The stencil search needs 2 m + 1 <= nc on every axis (no cell visited
// twice); smaller cells fall back to the brute-force build.
void grid_setup(mdffgpu_ctx* c) {
  const double lc = c->list_cut + c->P.skindiff;
  const double w = std::max(4.0, lc / 3.0);
  bool ok = lc > 0.0;
  for (int a = 0; a < 3; ++a) {
    c->gnc[a] = std::max(1, std::min(64, (int)(c->W[a] / w)));
    c->gm[a] = (int)std::ceil(lc * c->gnc[a] / c->W[a]);
    ok = ok && (2 * c->gm[a] + 1 <= c->gnc[a]);
  }
  c->gstencil = ok && (2 * c->gm[0] + 1) * (2 * c->gm[1] + 1) <= 64;
  const int ncell = c->gnc[0] * c->gnc[1] * c->gnc[2];
  if (ncell > c->ncell_cap) {
    dalloc(c, &c->d_cstart, (size_t)ncell + 1);
    dalloc(c, &c->d_ccur, (size_t)ncell);
    c->ncell_cap = ncell;
  }
}

// Spatial order of the atoms on the device: atoms sorted by cell of the
// list grid (row-major), by index within a cell; perm[rank] = atom.  The
// pair kernels visit the atoms in this order.
void cell_sort(mdffgpu_ctx* c) {
  const int n = c->n;
  const int ncell = c->gnc[0] * c->gnc[1] * c->gnc[2];
  hipStream_t st = c->st;
  k_cell_zero<<<nblk(ncell + 1), MDFF_BLOCK, 0, st>>>(ncell + 1, c->d_cstart);
  k_cell_bin<<<nblk(n), MDFF_BLOCK, 0, st>>>(n, c->d_s4, c->gnc[0], c->gnc[1], c->gnc[2],
                                             c->d_cellof, c->d_cstart);
  k_cell_scan<<<1, 1024, 0, st>>>(ncell, c->d_cstart, c->d_ccur);
  k_cell_fill<<<nblk(n), MDFF_BLOCK, 0, st>>>(n, c->d_cellof, c->d_ccur, c->d_perm);
  k_cell_order<<<nblk(ncell), MDFF_BLOCK, 0, st>>>(ncell, c->d_cstart, c->d_perm, c->d_rank);
  if (c->linduced && scf_field_fp32(c))
    k_pol_rows<<<1, 1024, 0, st>>>(n, c->d_perm, c->d_typ, c->polmask, c->d_polrows);
  check_launch(c, "cell sort");
}

// Atom ranges of the mixed reciprocal space in the context's numbering: the
// atoms of species that can carry a dipole (polarizable, a static dipole or
// Wannier centres) and of the polarizable species (the SCF reads the field
// there only).  The allocation-free context keeps all atoms in both.
void mx_active(mdffgpu_ctx* c) {
  if (!c->mx || c->noalloc) return;
  const mdffgpu_params& P = c->P;
  unsigned dipm = c->polmask;
  for (int t = 0; t < P.ntype; ++t) {
    for (int a = 0; a < 3; ++a)
      if (P.dip[t + 16 * a] != 0.0) dipm |= 1u << t;
    if (P.lwfc[t] > 0) dipm |= 1u << t;
  }
  auto range = [&](unsigned mask, int& a0, int& a1) {
    a0 = 0;
    a1 = -1;  // none: all atoms
    for (int i = 0; i < c->n; ++i)
      if ((mask >> c->typ0[i]) & 1u) {
        if (a1 < 0) a0 = i;
        a1 = i + 1;
      }
  };
  int d0, d1, f0, f1;
  range(dipm, d0, d1);
  range(c->polmask, f0, f1);
  if (!mixed::set_active(c->mx, d0, d1, f0, f1))
    fail(c, MDFFGPU_ERR_ALLOC, "mixed reciprocal-space atom ranges");
}

// The same for the FP64 reciprocal sums: the SCF calls' moments (induced
// dipoles and quadrupoles) and the field the SCF reads sit on the atoms of
// the dipole- or quadrupole-polarizable species, atoms kr0 .. kr1 - 1 (all
// atoms in the allocation-free context or without such species).
void k_ranges(mdffgpu_ctx* c) {
  c->kr0 = 0;
  c->kr1 = c->n;
  if (c->noalloc) return;
  int a0 = -1, a1 = -1;
  for (int i = 0; i < c->n; ++i)
    if ((c->fieldmask >> c->typ0[i]) & 1u) {
      if (a0 < 0) a0 = i;
      a1 = i + 1;
    }
  if (a0 >= 0) {
    c->kr0 = a0;
    c->kr1 = a1;
  }
}

// Per-atom species data of typeinfo_init (src/read_pos.f90:175-199) in the
// context's numbering (typ0): species, charge, static dipole and, with
// quadrupoles, the permanent quadrupole (quadia, :197).
void load_species(mdffgpu_ctx* c) {
  const mdffgpu_params& P = c->P;
  const int n = c->n;
  HCK(c, hipMemcpy(c->d_typ, c->typ0.data(), n * sizeof(int), hipMemcpyHostToDevice));
  c->npolrows = 0;
  for (int i = 0; i < n; ++i) c->npolrows += (int)((c->polmask >> c->typ0[i]) & 1u);
  std::vector<double> q(n), dip(3 * (size_t)n);
  for (int i = 0; i < n; ++i) {
    const int t = c->typ0[i];
    q[i] = P.qch[t];
    for (int a = 0; a < 3; ++a) dip[i + (size_t)n * a] = P.dip[t + 16 * a];
  }
  HCK(c, hipMemcpy(c->d_q, q.data(), q.size() * sizeof(double), hipMemcpyHostToDevice));
  HCK(c, hipMemcpy(c->d_dip, dip.data(), dip.size() * sizeof(double), hipMemcpyHostToDevice));
  if (c->d_thstat) {
    std::vector<double> th(9 * (size_t)n);
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < 9; ++k) th[i + (size_t)n * k] = P.quad[c->typ0[i] + 16 * k];
    HCK(c, hipMemcpy(c->d_thstat, th.data(), th.size() * sizeof(double), hipMemcpyHostToDevice));
  }
}

// Internal numbering, decided once, at the first evaluation, from the
// caller's positions (host arrays): key (species, cell of about 4 A on the
// lattice, caller index).  The allocation-free context keeps the caller's
// order.  The dipole history already set is carried over.
void ensure_renum(mdffgpu_ctx* c, const double* rx, const double* ry, const double* rz) {
  if (c->renum_decided) return;
  c->renum_decided = true;
  const int n = c->n;
  if (c->noalloc || n < 2) return;
  int nc[3];
  for (int a = 0; a < 3; ++a) nc[a] = std::max(1, std::min(64, (int)(c->W[a] / 4.0)));
  const long long ncell = (long long)nc[0] * nc[1] * nc[2];
  const double* B = c->cell.B;
  std::vector<long long> key(n);
  for (int i = 0; i < n; ++i) {
    const double x = rx[i], y = ry[i], z = rz[i];
    int cc[3];
    for (int a = 0; a < 3; ++a) {
      double s = x * B[3 * a] + y * B[3 * a + 1] + z * B[3 * a + 2];
      s -= std::floor(s);
      int k = std::isfinite(s) ? (int)(s * nc[a]) : 0;
      cc[a] = std::max(0, std::min(nc[a] - 1, k));
    }
    key[i] = (long long)c->typ0[i] * ncell + ((long long)cc[0] * nc[1] + cc[1]) * nc[2] + cc[2];
  }
  std::vector<int> ord(n);
  for (int i = 0; i < n; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return key[a] < key[b]; });
  bool ident = true;
  for (int i = 0; i < n && ident; ++i) ident = ord[i] == i;
  if (ident) return;
  const int K = c->P.extrapolate_order + 1;
  dalloc(c, &c->d_int2ext, (size_t)n);
  dalloc(c, &c->d_xbuf, (size_t)std::max(9, 3 * K) * n);
  HCK(c, hipMemcpy(c->d_int2ext, ord.data(), n * sizeof(int), hipMemcpyHostToDevice));
  std::vector<int> t(n);
  for (int i = 0; i < n; ++i) t[i] = c->typ0[ord[i]];
  c->typ0.swap(t);
  load_species(c);
  mx_active(c);
  k_ranges(c);
  c->renum = true;
  const size_t nh = (size_t)3 * K * n;
  k_perm_gather<<<(int)((nh + MDFF_BLOCK - 1) / MDFF_BLOCK), MDFF_BLOCK, 0, c->st>>>(
      n, K, 3, c->d_int2ext, c->d_hist, c->d_xbuf);
  check_launch(c, "k_perm_gather");
  HCK(c, hipMemcpyAsync(c->d_hist, c->d_xbuf, nh * sizeof(double), hipMemcpyDeviceToDevice, c->st));
  HCK(c, hipStreamSynchronize(c->st));
}

// The same from positions on the device (SoA [x|y|z], caller's order).
void ensure_renum_dev(mdffgpu_ctx* c, const double* d_pos) {
  if (c->renum_decided) return;
  if (c->noalloc || c->n < 2) {
    c->renum_decided = true;
    return;
  }
  const size_t n = c->n;
  std::vector<double> h(3 * n);
  HCK(c, hipMemcpyAsync(h.data(), d_pos, 3 * n * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HCK(c, hipStreamSynchronize(c->st));
  ensure_renum(c, h.data(), h.data() + n, h.data() + 2 * n);
}

// Boundary copies across the numbering, (inner, n, planes) doubles with the
// inner index fastest: gather_dev fills dst (internal numbering) from the
// caller's device array src, scatter_dev the caller's dst from src.
void gather_dev(mdffgpu_ctx* c, double* dst, const double* src, int planes, int inner = 1) {
  const size_t cnt = (size_t)planes * inner * c->n;
  if (!c->renum) {
    HCK(c, hipMemcpyAsync(dst, src, cnt * sizeof(double), hipMemcpyDeviceToDevice, c->st));
    return;
  }
  k_perm_gather<<<(int)((cnt + MDFF_BLOCK - 1) / MDFF_BLOCK), MDFF_BLOCK, 0, c->st>>>(
      c->n, inner, planes, c->d_int2ext, src, dst);
  check_launch(c, "k_perm_gather");
}

void scatter_dev(mdffgpu_ctx* c, double* dst, const double* src, int planes, int inner = 1) {
  const size_t cnt = (size_t)planes * inner * c->n;
  if (!c->renum) {
    HCK(c, hipMemcpyAsync(dst, src, cnt * sizeof(double), hipMemcpyDeviceToDevice, c->st));
    return;
  }
  k_perm_scatter<<<(int)((cnt + MDFF_BLOCK - 1) / MDFF_BLOCK), MDFF_BLOCK, 0, c->st>>>(
      c->n, inner, planes, c->d_int2ext, src, dst);
  check_launch(c, "k_perm_scatter");
}

// The same for host arrays of n doubles, one pointer per plane.
void gather_host(mdffgpu_ctx* c, double* dst, const double* const* src, int planes) {
  const size_t n = c->n, b = n * sizeof(double);
  double* stage = c->renum ? c->d_xbuf : dst;
  for (int p = 0; p < planes; ++p)
    HCK(c, hipMemcpyAsync(stage + p * n, src[p], b, hipMemcpyHostToDevice, c->st));
  if (c->renum) gather_dev(c, dst, stage, planes);
}

void scatter_host(mdffgpu_ctx* c, double* const* dst, const double* src, int planes) {
  const size_t n = c->n, b = n * sizeof(double);
  if (c->renum) {
    scatter_dev(c, c->d_xbuf, src, planes);
    src = c->d_xbuf;
  }
  for (int p = 0; p < planes; ++p)
    HCK(c, hipMemcpyAsync(dst[p], src + p * n, b, hipMemcpyDeviceToHost, c->st));
}

// A contiguous host array of planes x n doubles.
void scatter_host(mdffgpu_ctx* c, double* dst, const double* src, int planes) {
  double* d[9];
  for (int p = 0; p < planes; ++p) d[p] = dst + (size_t)p * c->n;
  scatter_host(c, d, src, planes);
}

// Noise-floor estimate of the mixed SCF after a solve (scalars on the
// host).  Only a stalled solve raises it, and never above FLOOR_CAP FP32
// epsilons of the RMS static field; a solve that exits at the iteration
// limit leaves it as it is; one that reaches conv_tol_ind drops it, and one
// that converges below it lowers it.
constexpr double FLOOR_CAP = 64.0;
void update_floor(mdffgpu_ctx* c) {
  if (!c->scf_done) return;
  c->scf_done = false;
  const double* h = c->h_sc;
  const double estat = (h[SC_ESTAT + 1] > 0.0) ? std::sqrt(h[SC_ESTAT] / h[SC_ESTAT + 1]) : 0.0;
  const double cap = FLOOR_CAP * (double)FLT_EPSILON * estat;
  if (c->scf_stalled)
    c->mx_floor = std::min(c->scf_best, cap);
  else if (c->scf_last <= c->P.conv_tol_ind)
    c->mx_floor = 0.0;
  else if (c->mx_floor > 0.0 && c->scf_last <= 1.5 * c->mx_floor)
    c->mx_floor = std::min(c->mx_floor, c->scf_best);
}

// engforce_driver on the device buffers of ctx; d_pos and d_force are
// loaded by the caller.
// Verlet list of the current positions: rebuilt when invalid or when the
// skin test fails, grown until it holds every row.
void update_list(mdffgpu_ctx* c) {
  const mdffgpu_params& P = c->P;
  const int n = c->n;
  const int gb = nblk(n);
  hipStream_t st = c->st;
  bool rebuild = !c->list_valid;
  if (c->list_cut > 0.0 && !rebuild) {
    // Verlet skin (vnlistcheck, src/tools.f90:446-535): the list holds every
    // pair within cutoff + skin at the last build, so it stays complete
    // while no atom has moved more than skin/2
    {
      ProfScope _ps(c, K_NLIST);
      k_disp<<<gb, MDFF_BLOCK, 0, st>>>(n, c->d_pos, c->d_rref, c->d_part);
      k_reduce_max<<<1, MDFF_BLOCK, 0, st>>>(c->d_part, gb, c->d_sc + SC_DISP);
    }
    check_launch(c, "k_disp");
    HCK(c, hipMemcpyAsync(c->h_sc + SC_DISP, c->d_sc + SC_DISP, sizeof(double),
                          hipMemcpyDeviceToHost, st));
    HCK(c, hipStreamSynchronize(st));
    prof_collect(c);
    rebuild = 2.0 * std::sqrt(c->h_sc[SC_DISP]) > P.skindiff;
  }
  if (c->list_cut > 0.0 && rebuild) {
    HCK(c, hipMemcpyAsync(c->d_rref, c->d_pos, 3 * (size_t)n * sizeof(double),
                          hipMemcpyDeviceToDevice, st));
    c->list_valid = true;
    ++c->list_builds;
    {
      ProfScope _ps(c, K_NLIST);
      grid_setup(c);
      cell_sort(c);
    }
    for (;;) {
      HCK(c, hipMemsetAsync(c->d_ovf, 0, sizeof(int), st));
      {
        ProfScope _ps(c, K_NLIST);
        const double lc = c->list_cut + P.skindiff;
        if (c->gstencil) {
          NlCellArgs na;
          na.n = n;
          na.maxnb = c->maxnb;
          for (int a = 0; a < 3; ++a) {
            na.nc[a] = c->gnc[a];
            na.m[a] = c->gm[a];
          }
          na.s4 = c->d_s4;
          na.cellof = c->d_cellof;
          na.cstart = c->d_cstart;
          na.perm = c->d_perm;
          na.cell = c->cell;
          na.cut2 = lc * lc;
          na.cnt = c->d_cnt;
          na.nbr = c->d_nbr;
          na.overflow = c->d_ovf;
          k_nlist_cells<<<n, 64, 0, st>>>(na);
        } else {
          // candidate ranges of whole tiles
          const int jlen =
              (n + NLIST_RANGES * MDFF_BLOCK - 1) / (NLIST_RANGES * MDFF_BLOCK) * MDFF_BLOCK;
          const dim3 gl(gb, (n + jlen - 1) / jlen);
          k_nlist<false><<<gl, MDFF_BLOCK, 0, st>>>(n, c->d_s, c->cell, lc * lc, c->maxnb, jlen,
                                                    c->d_nlr, c->d_nbr);
          k_nlist_offsets<<<gb, MDFF_BLOCK, 0, st>>>(n, (int)gl.y, c->maxnb, c->d_nlr, c->d_cnt,
                                                     c->d_ovf);
          k_nlist<true><<<gl, MDFF_BLOCK, 0, st>>>(n, c->d_s, c->cell, lc * lc, c->maxnb, jlen,
                                                   c->d_nlr, c->d_nbr);
        }
      }
      check_launch(c, "k_nlist");
      HCK(c, hipMemcpyAsync(c->h_ovf, c->d_ovf, sizeof(int), hipMemcpyDeviceToHost, st));
      HCK(c, hipStreamSynchronize(st));
      prof_collect(c);
      if (*c->h_ovf <= c->maxnb) break;
      set_list_capacity(c, (int)(*c->h_ovf * 1.2) + 16);
    }
  }
}

void run(mdffgpu_ctx* c, int itime, mdffgpu_energies* en) {
  ++c->evals;
  const mdffgpu_params& P = c->P;
  const int n = c->n;
  const int gb = nblk(n);
  hipStream_t st = c->st;
  std::memset(en, 0, sizeof(*en));
  {
    ProfScope _ps(c, K_PREPARE);
    k_prepare<<<gb, MDFF_BLOCK, 0, st>>>(n, c->d_pos, c->cell, c->renum ? c->d_int2ext : nullptr,
                                         c->d_typ, c->d_s, c->d_s4, c->d_eph);
  }
  check_launch(c, "k_prepare");
  update_list(c);
  if (P.lcoulomb && c->linduced) build_t2(c);
  const bool sr = P.lnmlj || P.lbmhft || P.lbmhftd || P.lmorse;
  const bool mxs = (P.precision == MDFFGPU_PREC_MIXED) &&
                   !(P.mixed_fp64_parts & MDFFGPU_MIXED_FP64_SHORTRANGE);
  if (sr) {
    int flags = (P.lnmlj ? SR_LJ : 0) | ((P.lbmhft || P.lbmhftd) ? SR_BM : 0) |
                (P.lbmhftd ? SR_BMD : 0) | (P.lmorse ? SR_MO : 0) |
                (P.trunc == 1 ? SR_TRUNC1 : 0) | (P.trunc == 2 ? SR_TRUNC2 : 0);
    {
        ProfScope _ps(c, mxs ? K_MX_SR : K_SR);
        SrArgs sa;
        std::memset(&sa, 0, sizeof(sa));
        sa.n = n;
        sa.maxnb = c->maxnb;
        sa.perm = c->d_perm;
        sa.s = c->d_s;
        sa.s4 = c->d_s4;
        sa.typ = c->d_typ;
        sa.cnt = c->d_cnt;
        sa.nbr = c->d_nbr;
        sa.cell = c->cell;
        sa.T = c->d_pt;
        sa.tt = c->tt;
        sa.rc2 = P.cutshortrange * P.cutshortrange;
        sa.flags = flags;
        sa.f = c->d_force;
        sa.part = c->d_part;
        if (mxs)
          k_shortrange<float><<<nblk(n * LPA), MDFF_BLOCK, 0, st>>>(sa);
        else
          k_shortrange<double><<<nblk(n * LPA), MDFF_BLOCK, 0, st>>>(sa);
    }
    check_launch(c, "k_shortrange");
    reduce(c, 12, nblk(n * LPA), SC_SR);
  }
  int n_es = 0;
  if (P.lcoulomb) {
    scf(c, itime, &en->n_scf, &en->rmsd, &n_es, &en->scf_tol);
    {
      ProfScope _ps(c, K_ADD);
      k_add<<<nblk(3 * n), MDFF_BLOCK, 0, st>>>(3 * n, c->d_dip, c->d_muind, c->d_mu);
    }
    check_launch(c, "k_add");
    if (c->lquad) {
      // theta = quadia + theta_ind (get_moments, src/field.f90:3744)
      ProfScope _ps(c, K_ADD);
      k_add<<<nblk(9 * n), MDFF_BLOCK, 0, st>>>(9 * n, c->d_thstat, c->d_thind, c->d_theta);
    }
    if (c->lwannier) {
      // moment_from_WFc after the SCF (src/field.f90:3723-3737)
      HCK(c, hipMemsetAsync(c->d_wfcbad, 0, sizeof(int), st));
      WfcArgs wa;
      wa.n = n;
      wa.maxnb = c->maxnb;
      wa.s = c->d_s;
      wa.typ = c->d_typ;
      // the list holds every pair within list_cut; centres farther out
      // need the scan of all atoms
      const bool use_list = c->list_cut > 0.0 && P.rcut_wfc <= c->list_cut * c->list_cut;
      wa.cnt = use_list ? c->d_cnt : nullptr;
      wa.nbr = use_list ? c->d_nbr : nullptr;
      for (int t = 0; t < MDFF_NT; ++t) wa.lwfc[t] = P.lwfc[t];
      wa.cell = c->cell;
      wa.rcut = P.rcut_wfc;
      wa.mu = c->d_muwfc;
      wa.bad = c->d_wfcbad;
      wa.ext = c->renum ? c->d_int2ext : nullptr;
      {
        ProfScope _ps(c, K_WFC);
        k_wfc<<<gb, MDFF_BLOCK, 0, st>>>(wa);
      }
      check_launch(c, "k_wfc");
      HCK(c, hipMemcpyAsync(c->h_ovf, c->d_wfcbad, sizeof(int), hipMemcpyDeviceToHost, st));
      if (P.ldip_wfc) k_add<<<nblk(3 * n), MDFF_BLOCK, 0, st>>>(3 * n, c->d_mu, c->d_muwfc, c->d_mu);
    }
    int task = (P.task_coul[0] ? TASK_QQ : 0) | (P.task_coul[1] ? TASK_QD : 0) |
               (P.task_coul[2] ? TASK_DD : 0) | (P.task_coul_quad[0] ? TASK_QQUAD : 0) |
               (P.task_coul_quad[1] ? TASK_DQUAD : 0) | (P.task_coul_quad[2] ? TASK_QUADQ : 0);
    es_call(c, ES_FINAL, c->d_q, c->d_mu, task, P.doefield, P.doefg, false,
            c->lquad ? c->d_theta : nullptr);
    ++n_es;
  }
  // device outputs queued ahead of the one synchronisation of the call
  if (c->out_force) scatter_dev(c, c->out_force, c->d_force, 3);
  if (P.lcoulomb) {
    if (c->out_mu) scatter_dev(c, c->out_mu, c->d_mu, 3);
    if (c->out_ef) scatter_dev(c, c->out_ef, c->d_ef, 3);
    if (c->out_efg) scatter_dev(c, c->out_efg, c->d_efg, 9);
  }
  sync_scalars(c);
  update_floor(c);
  if (c->upol_pending) {
    c->u_pol = 0.5 * c->h_sc[SC_UPOL];
    c->upol_pending = false;
  }
  if (P.lcoulomb && c->lwannier && *c->h_ovf != 0)
    fail(c, MDFFGPU_ERR_WFC,
         "moment_from_WFc: wannier center is missing for atom %d (fewer or more than lwfc "
         "centres with r^2 < rcut_wfc)",
         *c->h_ovf);
  const double* h = c->h_sc;
  const double omega = c->cell.omega;
  if (sr) {
    if (P.lnmlj) {
      en->u_lj = h[SC_SR + 0];
      en->vir_lj = h[SC_SR + 1] / 3.0;
    }
    if (P.lbmhft || P.lbmhftd) {
      en->u_bmhft = h[SC_SR + 2];
      en->vir_bmhft = h[SC_SR + 3] / 3.0;
    }
    if (P.lmorse) {
      en->u_morse = h[SC_SR + 4];
      en->vir_morse = h[SC_SR + 5] / 3.0;
    }
    sym9(h + SC_SR + 6, 1.0 / omega / P.press_unit, en->tau_nonb);
  }
  if (P.lcoulomb) {
    const double tpi_V = 2.0 * MDFF_PI / omega;
    const double alpha = P.alphaES, alpha2 = alpha * alpha;
    const double selfa = alpha / std::sqrt(MDFF_PI);
    const double selfa2 = 2.0 * selfa * alpha2 / 3.0;
    const double u_dir = h[SC_DIR] + (c->quad_final ? h[SC_QUAD] : 0.0);
    const double u_rec = h[SC_REC] * 2.0 * tpi_V;
    const double* m = h + SC_MOM;
    const double u_self = -selfa * m[6] - selfa2 * m[7];
    double u_surf = (m[0] * m[0] + m[1] * m[1] + m[2] * m[2]) +
                    2.0 * (m[0] * m[3] + m[1] * m[4] + m[2] * m[5]) +
                    (m[3] * m[3] + m[4] * m[4] + m[5] * m[5]);
    u_surf = u_surf * (tpi_V / 3.0);
    const double cu = P.coul_unit;
    if (P.lsurf)
      en->u_coul = (u_dir + u_rec + u_surf + u_self + c->u_pol) * cu;
    else
      en->u_coul = (u_dir + u_rec + u_self + c->u_pol) * cu;
    en->u_pol = c->u_pol;
    double td[9], tr[9];
    sym9(h + SC_DIR + 1, 0.5 / omega, td);
    sym9(h + SC_REC + 1, 2.0 * tpi_V / omega, tr);
    for (int k = 0; k < 9; ++k) en->tau_coul[k] = (td[k] + tr[k]) * cu / P.press_unit;
    en->pvirial_coul =
        1.0 / 3.0 * (en->tau_coul[0] + en->tau_coul[4] + en->tau_coul[8]) * P.press_unit;
  }
  en->n_es_calls = n_es;
}

// Device buffer of at least count doubles, grown (contents dropped) when
// too small.
void grow(mdffgpu_ctx* c, double** p, size_t* cap, size_t count) {
  if (*p && *cap >= count) return;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  dalloc(c, p, count);
  *cap = count;
}

// Image cells of multipole_efg_DS that can hold a pair closer than cut:
// cells n with |A n| < cut + |a1| + |a2| + |a3| (wrapped fractional
// differences lie in [-1, 1]^3), in the order of direct_sum_init /
// reorder_rpt (src/rspace.f90:64-114: x outer, z inner, stably sorted by
// |n|^2).
std::vector<double> ds_cells(const Cell& cl, int nc, double cut) {
  double reach = cut;
  for (int k = 0; k < 3; ++k)
    reach += std::sqrt(cl.A[3 * k] * cl.A[3 * k] + cl.A[3 * k + 1] * cl.A[3 * k + 1] +
                       cl.A[3 * k + 2] * cl.A[3 * k + 2]);
  struct Cn {
    int x, y, z, rr;
  };
  std::vector<Cn> keep;
  for (int x = -nc; x <= nc; ++x)
    for (int y = -nc; y <= nc; ++y)
      for (int z = -nc; z <= nc; ++z) {
        double v[3];
        for (int a = 0; a < 3; ++a) v[a] = x * cl.A[a] + y * cl.A[3 + a] + z * cl.A[6 + a];
        if (std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) >= reach) continue;
        keep.push_back({x, y, z, x * x + y * y + z * z});
      }
  std::stable_sort(keep.begin(), keep.end(), [](const Cn& a, const Cn& b) { return a.rr < b.rr; });
  std::vector<double> out;
  out.reserve(3 * keep.size());
  for (const Cn& k : keep) {
    out.push_back(k.x);
    out.push_back(k.y);
    out.push_back(k.z);
  }
  return out;
}

// One configuration of efgcalc (src/efg.f90:404-488) on the positions in
// d_pos: moments, then the EFG of every site into d_efgout.
void efg_frame(mdffgpu_ctx* c, const mdffgpu_efg_opts& o, mdffgpu_energies* en) {
  const int n = c->n;
  const int gb = nblk(n);
  hipStream_t st = c->st;
  const bool ewald = (o.longrange == 0);
  ++c->evals;
  {
    ProfScope _ps(c, K_PREPARE);
    k_prepare<<<gb, MDFF_BLOCK, 0, st>>>(n, c->d_pos, c->cell, c->renum ? c->d_int2ext : nullptr,
                                         c->d_typ, c->d_s, c->d_s4, c->d_eph);
  }
  check_launch(c, "k_prepare");
  // get_moments (src/field.f90:3686-3765) with calc = 'efg': the SCF starts
  // from the static field, no extrapolation
  std::memset(en, 0, sizeof(*en));
  if (c->linduced) {
    const int calc_mode = c->P.calc_mode;
    c->P.calc_mode = 0;
    try {
      update_list(c);
      build_t2(c);
      int n_es = 0;
      scf(c, 0, &en->n_scf, &en->rmsd, &n_es, &en->scf_tol);
      en->n_es_calls = n_es;
    } catch (...) {
      c->P.calc_mode = calc_mode;
      throw;
    }
    c->P.calc_mode = calc_mode;
    {
      ProfScope _ps(c, K_ADD);
      k_add<<<nblk(3 * n), MDFF_BLOCK, 0, st>>>(3 * n, c->d_dip, c->d_muind, c->d_mu);
    }
    check_launch(c, "k_add");
  } else {
    HCK(c, hipMemcpyAsync(c->d_mu, c->d_dip, 3 * (size_t)n * sizeof(double),
                          hipMemcpyDeviceToDevice, st));
  }
  const bool dip = c->linduced || c->ldip_static;
  // moments of the contributing atoms
  const double* q = c->d_q;
  const double* mu = c->d_mu;
  if (o.lefg_it_contrib) {
    if (!c->d_efgq) {
      dalloc(c, &c->d_efgq, n);
      dalloc(c, &c->d_efgmu, 3 * (size_t)n);
    }
    k_efg_mask<<<gb, MDFF_BLOCK, 0, st>>>(n, c->d_typ, o.it_efg - 1, c->d_q, c->d_mu, c->d_efgq,
                                          c->d_efgmu);
    check_launch(c, "k_efg_mask");
    q = c->d_efgq;
    mu = c->d_efgmu;
  }
  if (!c->d_efgout) dalloc(c, &c->d_efgout, 9 * (size_t)n);
  EfgCombineArgs ca;
  std::memset(&ca, 0, sizeof(ca));
  ca.n = n;
  ca.cell = c->cell;
  ca.fpi_V = 4.0 * MDFF_PI / c->cell.omega;
  ca.q = q;
  ca.out = c->d_efgout;
  ca.scale = c->P.coul_unit * (o.lefg_vasp_sign ? -1.0 : 1.0);
  if (ewald) {
    // real space: j slices so that a launch has about 2048 blocks
    const int ib = gb;
    int js = std::max(1, std::min((2048 + ib - 1) / ib, (n + MDFF_BLOCK - 1) / MDFF_BLOCK));
    int jlen = (n + js - 1) / js;
    jlen = (jlen + MDFF_BLOCK - 1) / MDFF_BLOCK * MDFF_BLOCK;
    js = (n + jlen - 1) / jlen;
    grow(c, &c->d_efgpart, &c->efgpart_cap, (size_t)js * 6 * n);
    EfgDirArgs da;
    da.n = n;
    da.jlen = jlen;
    da.s = c->d_s;
    da.q = q;
    da.mu = mu;
    da.cell = c->cell;
    da.alpha = c->P.alphaES;
    da.part = c->d_efgpart;
    {
      ProfScope _ps(c, K_EFG_DIR);
      const dim3 g(ib, js);
      if (dip)
        k_efg_es_dir<true><<<g, MDFF_BLOCK, 0, st>>>(da);
      else
        k_efg_es_dir<false><<<g, MDFF_BLOCK, 0, st>>>(da);
    }
    check_launch(c, "k_efg_es_dir");
    ca.dir = c->d_efgpart;
    ca.ndir = js;
    // reciprocal space: charge_density_k (src/efg.f90:2888-2930) and the
    // EFG of every k (:1507-1556), on the FP64 structure-factor kernels
    if (c->nchunk > 0) {
      RhoArgs ra;
      ra.n = n;
      ra.nchunk = c->nchunk;
      ra.KC = c->KC;
      for (int a = 0; a < 3; ++a) ra.kmax[a] = c->kmax[a];
      ra.chunks = c->d_chunks;
      ra.eph = c->d_eph;
      ra.q = q;
      ra.mu = mu;
      ra.cell = c->cell;
      ra.atoms_per_split = c->rho_atoms;
      ra.kpad = c->kpad;
      ra.part = c->d_rpart;
      ra.j_begin = 0;
      ra.j_end = n;
      dim3 g((c->nchunk + MDFF_BLOCK - 1) / MDFF_BLOCK, c->rho_split);
      size_t lds = (size_t)RHO_TILE * (c->kmax[0] + c->kmax[1] + c->kmax[2] + 3) *
                       sizeof(double2) + 4 * RHO_TILE * sizeof(double);
      if (dip) {
        ProfScope _ps(c, K_RHO_BOTH);
        k_rho64<true, true><<<g, MDFF_BLOCK, lds, st>>>(ra);
      } else {
        ProfScope _ps(c, K_RHO_CHG);
        k_rho64<true, false><<<g, MDFF_BLOCK, lds, st>>>(ra);
      }
      check_launch(c, "k_rho64");
      RhoRedArgs rr;
      rr.kpad = c->kpad;
      rr.nsplit = c->rho_split;
      rr.do_energy = 0;
      rr.part = c->d_rpart;
      rr.Ak = c->d_Ak;
      rr.kcoe = c->d_kcoe;
      rr.kvec = c->d_kvec;
      rr.g = c->d_g;
      rr.pout = c->d_part;
      {
        ProfScope _ps(c, K_RHO_RED);
        k_rho_reduce64<<<nblk(c->kpad), MDFF_BLOCK, 0, st>>>(rr);
      }
      check_launch(c, "k_rho_reduce64");
      ApplyArgs aa;
      aa.n = n;
      aa.nchunk = c->nchunk;
      aa.KC = c->KC;
      aa.chunks_per_split = c->app_chunks;
      aa.chunks = c->d_chunks;
      aa.g = c->d_g;
      aa.eph = c->d_eph;
      aa.s = c->d_s;
      aa.s4 = c->d_s4;
      aa.part = c->d_apart;
      aa.i_begin = 0;
      aa.i_end = n;
      {
        ProfScope _ps(c, K_APPLY_EFG);
        k_apply64<true><<<dim3(gb, c->app_split), MDFF_BLOCK, 0, st>>>(aa);
      }
      check_launch(c, "k_apply64");
      ca.rec = c->d_apart;
      ca.nrec = c->app_split;
    }
    const double alpha = c->P.alphaES;
    ca.selfa = -4.0 * alpha * alpha * alpha / 3.0 / std::sqrt(MDFF_PI);
    if (o.lmp_correction) {
      // total charge of the contributing atoms (src/efg.f90:1596-1608)
      double Q = 0.0;
      for (int i = 0; i < n; ++i) {
        const int t = c->typ0[i];
        if (o.lefg_it_contrib && t != o.it_efg - 1) continue;
        Q = Q + c->P.qch[t];
      }
      ca.corr = Q * ca.fpi_V / 3.0;
    }
  } else {
    std::vector<double> cells = ds_cells(c->cell, o.ncelldirect, o.cutdirect);
    const int ncell = (int)(cells.size() / 3);
    if (ncell == 0) fail(c, MDFFGPU_ERR_ARG, "no image cell within the direct-sum cutoff");
    grow(c, &c->d_efgcells, &c->efgcells_cap, cells.size());
    grow(c, &c->d_efgpart, &c->efgpart_cap, 6 * (size_t)n);
    HCK(c, hipMemcpyAsync(c->d_efgcells, cells.data(), cells.size() * sizeof(double),
                          hipMemcpyHostToDevice, st));
    EfgDsArgs da;
    da.n = n;
    da.ncell = ncell;
    da.s = c->d_s;
    da.q = q;
    da.mu = mu;
    da.cells = c->d_efgcells;
    da.cell = c->cell;
    da.cut2 = o.cutdirect * o.cutdirect;
    da.part = c->d_efgpart;
    {
      ProfScope _ps(c, K_EFG_DS);
      if (dip)
        k_efg_ds<true><<<n, MDFF_BLOCK, 0, st>>>(da);
      else
        k_efg_ds<false><<<n, MDFF_BLOCK, 0, st>>>(da);
    }
    check_launch(c, "k_efg_ds");
    ca.dir = c->d_efgpart;
    ca.ndir = 1;
  }
  {
    ProfScope _ps(c, K_EFG_COMBINE);
    k_efg_combine<<<gb, MDFF_BLOCK, 0, st>>>(ca);
  }
  check_launch(c, "k_efg_combine");
  sync_scalars(c);
  if (c->linduced) update_floor(c);
  if (c->upol_pending) {
    c->u_pol = 0.5 * c->h_sc[SC_UPOL];
    c->upol_pending = false;
  }
}

// The context of mdffgpu_create; with arena (the allocation-free entry
// point) every device buffer comes from that memory instead of hipMalloc,
// the host scalars are pageable and the work goes to the caller's stream.
// Quadrupole and Wannier-centre state (FP64 path): buffers of the
// permanent (filled by load_species), induced and total quadrupoles, the
// Wannier-centre dipoles.
void setup_quad_wfc(mdffgpu_ctx* c) {
  const mdffgpu_params& P = c->P;
  const int n = c->n;
  c->lqstat = c->lqpol = c->lwannier = false;
  for (int k = 0; k < 144; ++k)
    if (P.quad[k] != 0.0) c->lqstat = true;
  for (int t = 0; t < P.ntype; ++t) {
    if (P.lquad_polar[t]) c->lqpol = true;
    if (P.lwfc[t] > 0) c->lwannier = true;
  }
  c->lquad = c->lqstat || c->lqpol || P.task_coul_quad[0] || P.task_coul_quad[1] ||
             P.task_coul_quad[2];
  if (c->lquad) {
    dalloc(c, &c->d_thstat, 9 * (size_t)n);
    dalloc(c, &c->d_thind, 9 * (size_t)n);
    dalloc(c, &c->d_theta, 9 * (size_t)n);
    HCK(c, hipMemsetAsync(c->d_thind, 0, 9 * (size_t)n * sizeof(double), c->st));
    HCK(c, hipMemsetAsync(c->d_theta, 0, 9 * (size_t)n * sizeof(double), c->st));
  }
  if (c->lwannier) {
    dalloc(c, &c->d_muwfc, 3 * (size_t)n);
    dalloc(c, &c->d_wfcbad, 1);
  }
}

void create_impl(const mdffgpu_params* params, const int32_t* itype, const double* A,
                 mdffgpu_ctx** out, char* arena = nullptr, size_t arena_cap = 0,
                 hipStream_t stream = nullptr) {
  validate(nullptr, params);
  if (!itype || !A || !out) fail(nullptr, MDFFGPU_ERR_ARG, "NULL argument");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    fail(nullptr, MDFFGPU_ERR_DEVICE, "no HIP device visible");
  mdffgpu_ctx* c = new mdffgpu_ctx();
  try {
    c->P = *params;
    c->n = params->natm;
    HCK(c, hipGetDevice(&c->device));
    c->typ0.resize(c->n);
    for (int i = 0; i < c->n; ++i) {
      if (itype[i] < 1 || itype[i] > params->ntype)
        fail(c, MDFFGPU_ERR_ARG, "itype(%d) = %d outside 1..ntype", i + 1, itype[i]);
      c->typ0[i] = itype[i] - 1;
    }
    if (arena) {
      c->noalloc = true;
      c->arena = arena;
      c->arena_cap = arena_cap;
      c->st = stream;
      c->h_sc_buf.assign(SC_N, 0.0);
      c->h_ovf_buf.assign(1, 0);
      c->h_sc = c->h_sc_buf.data();
      c->h_ovf = c->h_ovf_buf.data();
    } else {
      HCK(c, hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking));
      c->st = c->own_stream;
      HCK(c, hipHostMalloc((void**)&c->h_sc, SC_N * sizeof(double), hipHostMallocDefault));
      HCK(c, hipHostMalloc((void**)&c->h_ovf, sizeof(int), hipHostMallocDefault));
    }
    set_lattice(c, A);
    alloc_atoms(c);
    build_tables(c);
    if (c->P.lcoulomb) setup_quad_wfc(c);
    load_species(c);
    k_ranges(c);
    update_cutoffs(c);
    if (c->list_cut > 0.0) set_list_capacity(c, initial_capacity(c));
    if (c->P.lcoulomb) build_kmesh(c);
    size_t need = 12 * (size_t)std::max(nblk(c->kpad), nblk(c->n * LPA)) + 4096;
    ensure_part(c, need);
    if (c->P.precision == MDFFGPU_PREC_MIXED) {
      c->mx = mixed::create(c);
      if (!c->mx) fail(c, MDFFGPU_ERR_UNSUPPORTED, "mixed-precision path unavailable");
      mx_active(c);
    }
    // the clears of alloc_atoms are done before the context is handed out
    HCK(c, hipStreamSynchronize(c->st));
    *out = c;
  } catch (Fail&) {
    g_create_error = c->err;
    mdffgpu_destroy(c);
    throw;
  }
}

void copy_in_host(mdffgpu_ctx* c, const double* rx, const double* ry, const double* rz,
                  const double* fx, const double* fy, const double* fz) {
  const double* r[3] = {rx, ry, rz};
  const double* f[3] = {fx, fy, fz};
  gather_host(c, c->d_pos, r, 3);
  gather_host(c, c->d_force, f, 3);
}

}  // namespace

// ===========================================================================
// mixed-precision hooks need access to the context internals
namespace mdff {

void inv3(const double* m, double* out) {
  double a[3][6];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      a[i][j] = m[i + 3 * j];
      a[i][j + 3] = (i == j) ? 1.0 : 0.0;
    }
  for (int col = 0; col < 3; ++col) {
    int piv = col;
    for (int r = col + 1; r < 3; ++r)
      if (std::fabs(a[r][col]) > std::fabs(a[piv][col])) piv = r;
    if (a[piv][col] == 0.0) {
      for (int k = 0; k < 9; ++k) out[k] = 0.0;
      return;
    }
    if (piv != col)
      for (int k = 0; k < 6; ++k) std::swap(a[col][k], a[piv][k]);
    double d = a[col][col];
    for (int k = 0; k < 6; ++k) a[col][k] = a[col][k] / d;
    for (int r = 0; r < 3; ++r)
      if (r != col) {
        double f = a[r][col];
        for (int k = 0; k < 6; ++k) a[r][k] = a[r][k] - f * a[col][k];
      }
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) out[i + 3 * j] = a[i][j + 3];
}

bool make_cell(const double* A, Cell& cl, double W[3]) {
  std::memcpy(cl.A, A, 9 * sizeof(double));
  const double* a1 = A;
  const double* a2 = A + 3;
  const double* a3 = A + 6;
  double* B = cl.B;
  B[0] = a2[1] * a3[2] - a2[2] * a3[1];
  B[1] = a2[2] * a3[0] - a2[0] * a3[2];
  B[2] = a2[0] * a3[1] - a2[1] * a3[0];
  B[3] = a3[1] * a1[2] - a3[2] * a1[1];
  B[4] = a3[2] * a1[0] - a3[0] * a1[2];
  B[5] = a3[0] * a1[1] - a3[1] * a1[0];
  B[6] = a1[1] * a2[2] - a1[2] * a2[1];
  B[7] = a1[2] * a2[0] - a1[0] * a2[2];
  B[8] = a1[0] * a2[1] - a1[1] * a2[0];
  const double omega = B[0] * A[0] + B[1] * A[1] + B[2] * A[2];
  if (!(omega > 0.0)) return false;
  cl.omega = omega;
  for (int i = 0; i < 3; ++i) {
    const double* b = B + 3 * i;
    W[i] = omega / std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  }
  for (int i = 0; i < 9; ++i) B[i] = B[i] / omega;
  return true;
}

int make_tables(const mdffgpu_params& P, PairTables& T, TypeTables& Y, TTTable& tt) {
  int bad = -1;
  std::memset(&T, 0, sizeof(T));
  const double two16 = std::pow(2.0, 1.0 / 6.0);
  const double rc = P.cutshortrange, rcsq = rc * rc;
  const int nt = P.ntype;
  auto L = [&](int a, int b, int cc) { return P.ldip_damping[a + 16 * (b + 16 * cc)] != 0; };
  auto D = [&](const double* tab, int a, int b, int cc) { return tab[a + 16 * (b + 16 * cc)]; };
  for (int ti = 0; ti < nt; ++ti)
    for (int tj = 0; tj < nt; ++tj) {
      const int t = ti + 16 * tj;
      // initialize_param_non_bonded, src/field.f90:1021-1064
      const double pp = P.plj[t], qq = P.qlj[t];
      const double epsp = P.epslj[t] / (qq - pp);
      const double sigsq = two16 * two16 * P.sigmalj[t] * P.sigmalj[t];
      const double sr2 = sigsq / rcsq;
      const double sr = std::sqrt(sr2);
      const double srp = std::pow(sr, pp), srq = std::pow(sr, qq);
      const double ppqq = pp * qq;
      T.lj_epsp[t] = epsp;
      T.lj_sigsq[t] = sigsq;
      T.lj_fc[t] = ppqq * epsp / sigsq;
      T.lj_ptwo[t] = pp * 0.5;
      T.lj_qtwo[t] = qq * 0.5;
      T.lj_p[t] = pp;
      T.lj_q[t] = qq;
      T.lj_uc[t] = epsp * (pp * srq - qq * srp);
      T.lj_uc1[t] = epsp * ppqq / (2.0 * rcsq) * (srq - srp);
      T.lj_uc2[t] = 0.5 * epsp * ((2.0 * pp + ppqq) * srq - (2.0 * qq + ppqq) * srp);
      T.bm_A[t] = P.Abmhftd[t];
      T.bm_B[t] = P.Bbmhftd[t];
      T.bm_C[t] = P.Cbmhftd[t];
      T.bm_D[t] = P.Dbmhftd[t];
      T.bm_BD[t] = P.BDbmhftd[t];
      T.mo_eps[t] = P.epsmor[t];
      T.mo_rho[t] = P.rhomor[t];
      T.mo_rs[t] = std::exp(P.rhomor[t] * P.sigmamor[t]);
      T.mo_sig[t] = P.sigmamor[t];
      // damping as seen from ti (src/field.f90:2207-2211)
      T.dmp_on[0][t] = L(ti, ti, tj) || L(tj, ti, tj);
      T.dmp_on[1][t] = L(tj, tj, ti) || L(ti, tj, ti);
      T.dmp_b[0][t] = D(P.dip_pol_damp_b, ti, ti, tj);
      T.dmp_c[0][t] = D(P.dip_pol_damp_c, ti, ti, tj);
      T.dmp_k[0][t] = P.dip_pol_damp_k[ti + 16 * (ti + 16 * tj)];
      T.dmp_b[1][t] = D(P.dip_pol_damp_b, ti, tj, ti);
      T.dmp_c[1][t] = D(P.dip_pol_damp_c, ti, tj, ti);
      T.dmp_k[1][t] = P.dip_pol_damp_k[ti + 16 * (tj + 16 * ti)];
      for (int o = 0; o < 2; ++o)
        if ((T.dmp_k[o][t] < 0 || T.dmp_k[o][t] > MDFFGPU_TT_MAX_ORDER) && bad < 0)
          bad = T.dmp_k[o][t];
    }
  std::memset(&Y, 0, sizeof(Y));
  for (int t = 0; t < 16; ++t) {
    Y.qch[t] = P.qch[t];
    for (int a = 0; a < 3; ++a) Y.dip[t][a] = P.dip[t + 16 * a];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) Y.pol[t][a + 3 * b] = P.poldip[t + 16 * (a + 3 * b)];
    Y.polar[t] = P.ldip_polar[t] ? 1 : 0;
    if (Y.polar[t]) inv3(Y.pol[t], Y.invpol[t]);
    for (int k = 0; k < 27; ++k) Y.polq[t][k] = P.polquad[t + 16 * k];
    Y.qpolar[t] = P.lquad_polar[t] ? 1 : 0;
  }
  tt.e[0] = 1.0;
  for (int k = 1; k <= 8; ++k) tt.e[k] = tt.e[k - 1] / (double)k;
  return bad;
}

}  // namespace mdff

namespace mdff {
namespace mixed {
CtxView view(mdffgpu_ctx* c) {
  CtxView v;
  v.n = c->n;
  v.st = c->st;
  v.cell = c->cell;
  v.P = &c->P;
  v.kmax[0] = c->kmax[0];
  v.kmax[1] = c->kmax[1];
  v.kmax[2] = c->kmax[2];
  v.KC = c->KC;
  v.nchunk = c->nchunk;
  v.kpad = c->kpad;
  v.chunks = c->d_chunks;
  v.Ak = c->d_Ak;
  v.kcoe = c->d_kcoe;
  v.kvec = c->d_kvec;
  v.evals = c->evals;
  v.pos = c->d_pos;
  v.s = c->d_s;
  v.eph = c->d_eph;
  v.typ = c->d_typ;
  v.cnt = c->d_cnt;
  v.nbr = c->d_nbr;
  v.maxnb = c->maxnb;
  v.pt = c->d_pt;
  v.tt = c->tt;
  v.ttab = c->d_tt;
  v.efdir = c->d_efdir;
  v.efgdir = c->d_efgdir;
  v.fdir = c->d_fdir;
  v.force = c->d_force;
  v.Estat = c->d_Estat;
  v.E = c->d_E;
  v.ef = c->d_ef;
  v.efg = c->d_efg;
  v.part = c->d_part;
  v.sc = c->d_sc;
  v.sc_sr = SC_SR;
  v.sc_dir = SC_DIR;
  v.sc_rec = SC_REC;
  v.sc_mom = SC_MOM;
  return v;
}
}  // namespace mixed
}  // namespace mdff

// ===========================================================================
// Static scratch of the allocation-free entry point (mdffgpu_engforce_noalloc)
namespace {

constexpr size_t kScratchBytes = (size_t)MDFFGPU_NOALLOC_SCRATCH_MIB << 20;

// Every device the library runs on holds one copy until the process ends.
// Host code never uses the name as an address: the device address of each
// device's copy comes from hipGetSymbolAddress (scratch_ptr).
static __device__ __align__(256) char scratch[kScratchBytes];

std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
char* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch, looked up once per device
// (with that device current) and kept.
char* scratch_ptr(int dev, std::string* err) {
  if (dev < 0 || dev >= kMaxDevices) {
    *err = "device index out of range of the scratch table";
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return nullptr;
    }
    g_scratch_ptr[dev] = static_cast<char*>(ptr);
  }
  return g_scratch_ptr[dev];
}

}  // namespace

// ===========================================================================
// C ABI
extern "C" {

int mdffgpu_create(const mdffgpu_params* params, const int32_t* itype, const double* A,
                   mdffgpu_ctx** out) {
  if (out) *out = nullptr;
  try {
    create_impl(params, itype, A, out);
    return MDFFGPU_OK;
  } catch (Fail& f) {
    return f.code;
  } catch (...) {
    g_create_error = "unexpected host exception";
    return MDFFGPU_ERR_DEVICE;
  }
}

void mdffgpu_destroy(mdffgpu_ctx* c) {
  if (!c) return;
  if (c->noalloc) {
    // buffers of the static scratch and pageable host memory: nothing to free
    delete c;
    return;
  }
  if (c->mx) mixed::destroy(c->mx);
  double* dp[] = {c->d_rref, c->d_pos, c->d_s, c->d_q, c->d_qzero, c->d_dip, c->d_mu, c->d_muind,
                  c->d_mup, c->d_Estat, c->d_E, c->d_efdir, c->d_efgdir, c->d_fdir,
                  c->d_force, c->d_ef, c->d_efg, c->d_apart, c->d_Ak, c->d_kcoe, c->d_kvec,
                  c->d_part, c->d_sc, c->d_hist, c->d_upart, c->d_cg, c->d_efgq,
                  c->d_efgmu, c->d_efgpart, c->d_efgcells, c->d_efgout};
  for (double* p : dp)
    if (p) (void)hipFree(p);
  if (c->d_t2) (void)hipFree(c->d_t2);
  void* vp[] = {c->d_eph, c->d_typ, c->d_cnt, c->d_nbr, c->d_ovf, c->d_rpart, c->d_g,
                c->d_chunks, c->d_pt, c->d_tt, c->d_mu4, c->d_s4, c->d_perm, c->d_rank,
                c->d_nlr, c->d_cellof, c->d_cstart, c->d_ccur, c->d_vel, c->d_fs,
                c->d_massia, c->d_nhc, c->d_thstat, c->d_thind, c->d_theta, c->d_muwfc,
                c->d_wfcbad, c->d_int2ext, c->d_xbuf};
  for (void* p : vp)
    if (p) (void)hipFree(p);
  for (auto& p : c->prof_pending) {
    (void)hipEventDestroy(p.second.first);
    (void)hipEventDestroy(p.second.second);
  }
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->ev_sc) (void)hipEventDestroy(c->ev_sc);
  if (c->h_sc) (void)hipHostFree(c->h_sc);
  if (c->h_ovf) (void)hipHostFree(c->h_ovf);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
}

int mdffgpu_set_cell(mdffgpu_ctx* c, const double* A) {
  if (!c || !A) return MDFFGPU_ERR_ARG;
  try {
    set_lattice(c, A);
    update_cutoffs(c);
    c->list_valid = false;
    c->mx_floor = 0.0;
    if (c->list_cut > 0.0) set_list_capacity(c, std::max(c->maxnb, initial_capacity(c)));
    if (c->P.lcoulomb) build_kmesh(c);
    if (c->mx) mixed::on_cell(c->mx);
    return MDFFGPU_OK;
  } catch (Fail& f) {
    return f.code;
  }
}

int mdffgpu_engforce(mdffgpu_ctx* c, const double* rx, const double* ry, const double* rz,
                     int32_t itime, double* fx, double* fy, double* fz, double* mu_t,
                     double* theta_t, double* ef_t, double* efg_t, mdffgpu_energies* en) {
  if (!c) return MDFFGPU_ERR_ARG;
  if (!rx || !ry || !rz || !fx || !fy || !fz || !en) {
    c->err = "NULL position, force or energy argument";
    return MDFFGPU_ERR_ARG;
  }
  try {
    c->st = c->own_stream;
    ensure_renum(c, rx, ry, rz);
    copy_in_host(c, rx, ry, rz, fx, fy, fz);
    run(c, itime, en);
    const size_t b = (size_t)c->n * sizeof(double);
    double* f[3] = {fx, fy, fz};
    scatter_host(c, f, c->d_force, 3);
    if (c->P.lcoulomb) {
      if (mu_t) scatter_host(c, mu_t, c->d_mu, 3);
      if (ef_t) scatter_host(c, ef_t, c->d_ef, 3);
      if (efg_t) scatter_host(c, efg_t, c->d_efg, 9);
    }
    if (theta_t) {
      if (c->P.lcoulomb && c->lquad)
        scatter_host(c, theta_t, c->d_theta, 9);
      else
        std::memset(theta_t, 0, 9 * b);
    }
    HCK(c, hipStreamSynchronize(c->st));
    return MDFFGPU_OK;
  } catch (Fail& f) {
    (void)hipStreamSynchronize(c->st);
    return f.code;
  }
}

int mdffgpu_engforce_device(mdffgpu_ctx* c, const double* d_pos, int32_t itime,
                            double* d_force, double* d_mu, double* d_ef, double* d_efg,
                            mdffgpu_energies* en, void* stream) {
  if (!c) return MDFFGPU_ERR_ARG;
  if (!d_pos || !d_force || !en) {
    c->err = "NULL position, force or energy argument";
    return MDFFGPU_ERR_ARG;
  }
  try {
    c->st = stream ? (hipStream_t)stream : c->own_stream;
    ensure_renum_dev(c, d_pos);
    gather_dev(c, c->d_pos, d_pos, 3);
    gather_dev(c, c->d_force, d_force, 3);
    // the outputs are copied by run() ahead of its final synchronisation
    c->out_force = d_force;
    c->out_mu = d_mu;
    c->out_ef = d_ef;
    c->out_efg = d_efg;
    run(c, itime, en);
    c->out_force = c->out_mu = c->out_ef = c->out_efg = nullptr;
    c->st = c->own_stream;
    return MDFFGPU_OK;
  } catch (Fail& f) {
    c->out_force = c->out_mu = c->out_ef = c->out_efg = nullptr;
    c->st = c->own_stream;
    return f.code;
  }
}

// Yoshida weights of chain_nhcn (src/integration.f90:596-632): the
// published Yoshida/Suzuki factorisation weights of orders 1 to 9.
static bool yoshida_weights(int order, double* w) {
  switch (order) {
    case 1:
      w[0] = 1.0;
      return true;
    case 3:
      w[0] = 1.0 / (2.0 - std::pow(2.0, 1.0 / 3.0));
      w[1] = 1.0 - 2.0 * w[0];
      w[2] = w[0];
      return true;
    case 5:
      w[0] = 1.0 / (4.0 - std::pow(4.0, 1.0 / 3.0));
      w[1] = w[2] = w[3] = w[0];
      w[4] = 1.0 - 4.0 * w[0];
      return true;
    case 7:
      w[0] = .78451361047756;
      w[1] = .235573213359357;
      w[2] = -1.17767998417887;
      w[3] = 1.0 - 2.0 * (w[0] + w[1] + w[2]);
      w[4] = w[2];
      w[5] = w[1];
      w[6] = w[0];
      return true;
    case 9:
      w[0] = 0.192;
      w[1] = 0.554910818409783619692725006662999;
      w[2] = 0.124659619941888644216504240951585;
      w[3] = -0.843182063596933505315033808282941;
      w[4] = 1.0 - 2.0 * (w[0] + w[1] + w[2] + w[3]);
      w[5] = w[3];
      w[6] = w[2];
      w[7] = w[1];
      w[8] = w[0];
      return true;
    default:
      return false;
  }
}

// Kinetic sum m |v|^2 into d_sc[SC_KIN]: s = nullptr only sums, otherwise
// scales the velocities by the chain's factor first.
static void md_kin(mdffgpu_ctx* c, MdArgs& a, const NhcState* s) {
  const int gb = nblk(c->n);
  {
    ProfScope _ps(c, K_MD);
    k_vscale<<<gb, MDFF_BLOCK, 0, c->st>>>(a, s);
    k_reduce_rows<<<1, MDFF_BLOCK, 0, c->st>>>(c->d_part, gb, c->d_sc + SC_KIN);
  }
  check_launch(c, "k_vscale");
}

static void md_chain(mdffgpu_ctx* c, const mdffgpu_md_params& mp, double L, int final) {
  ProfScope _ps(c, K_MD);
  if (mp.integrator == MDFFGPU_NVT_NHC2)
    k_nhc2_chain<<<1, 64, 0, c->st>>>(c->d_nhc, c->d_sc + SC_KIN, mp.dt, mp.temp, L, final);
  else
    k_nhcn_chain<<<1, 64, 0, c->st>>>(c->d_nhc, c->d_sc + SC_KIN, mp.nhc_n, mp.nhc_yosh_order,
                                      mp.nhc_mults, mp.dt, mp.temp, L, final);
}

// The md_run time loop (src/runmd.f90:233-330) on device buffers; see the
// header.
static void md_loop(mdffgpu_ctx* c, const mdffgpu_md_params& mp, int itime0, int nsteps, double* xi,
             double* vxi, mdffgpu_md_step* steps) {
  const int n = c->n;
  const int gb = nblk(n);
  hipStream_t st = c->st;
  const bool nvt = mp.integrator != MDFFGPU_NVE_VV;
  const int nhc = (mp.integrator == MDFFGPU_NVT_NHC2) ? 2 : mp.nhc_n;
  // degrees of freedom: nhc2 3(N-1) (src/integration.f90:220), nhcn 3N (:505)
  const double L = (mp.integrator == MDFFGPU_NVT_NHC2) ? 3.0 * ((double)n - 1.0) : 3.0 * (double)n;
  NhcMass nm;
  for (int t = 0; t < MDFF_NT; ++t) nm.m[t] = mp.mass[t];
  k_massia<<<gb, MDFF_BLOCK, 0, st>>>(n, c->d_typ, nm, c->d_massia);
  check_launch(c, "k_massia");
  NhcState hs;
  std::memset(&hs, 0, sizeof(hs));
  if (nvt) {
    // thermostat masses: nhcn Q = timesca^2 temp, Q(1) *= L (:508-510);
    // nhc2 Q(1) = timesca^2 temp L, Q(2) = timesca^2 temp (:223-224)
    const double tq = std::pow(mp.timesca_thermo, 2.0) * mp.temp;
    for (int h = 0; h < nhc; ++h) {
      hs.Q[h] = tq;
      hs.xi[h] = xi[h];
      hs.vxi[h] = vxi[h];
    }
    hs.Q[0] = hs.Q[0] * L;
    yoshida_weights(mp.nhc_yosh_order, hs.yw);
    HCK(c, hipMemcpyAsync(c->d_nhc, &hs, sizeof(hs), hipMemcpyHostToDevice, st));
  }
  MdArgs a;
  a.n = n;
  a.mass = c->d_massia;
  a.pos = c->d_pos;
  a.vel = c->d_vel;
  a.f = c->d_force;
  a.fs = c->d_fs;
  a.dt = mp.dt;
  a.dt2 = mp.dt * 0.5;
  a.dtsq2 = mp.dt * mp.dt * 0.5;
  a.part = c->d_part;
  for (int k = 0; k < nsteps; ++k) {
    const int itime = itime0 + k;
    mdffgpu_energies en;
    if (!nvt) {
      // prop_velocity_verlet (src/integration.f90:131-194)
      {
        ProfScope _ps(c, K_MD);
        k_vv_pos<<<gb, MDFF_BLOCK, 0, st>>>(a);
      }
      check_launch(c, "k_vv_pos");
      run(c, itime, &en);
      {
        ProfScope _ps(c, K_MD);
        k_vv_vel<<<gb, MDFF_BLOCK, 0, st>>>(a);
        k_reduce_rows<<<1, MDFF_BLOCK, 0, st>>>(c->d_part, gb, c->d_sc + SC_KIN);
      }
      check_launch(c, "k_vv_vel");
    } else {
      // nhc2 / nhcn: chain(kin of calc_temp), prop_pos_vel_verlet, chain
      // (kin of the second kick), calc_temp of the scaled velocities
      md_kin(c, a, nullptr);
      md_chain(c, mp, L, 0);
      {
        ProfScope _ps(c, K_MD);
        k_vscale<<<gb, MDFF_BLOCK, 0, st>>>(a, c->d_nhc);
        k_kick_drift<<<gb, MDFF_BLOCK, 0, st>>>(a);
      }
      check_launch(c, "k_kick_drift");
      run(c, itime, &en);
      {
        ProfScope _ps(c, K_MD);
        k_kick<<<gb, MDFF_BLOCK, 0, st>>>(a);
        k_reduce_rows<<<1, MDFF_BLOCK, 0, st>>>(c->d_part, gb, c->d_sc + SC_KIN);
      }
      check_launch(c, "k_kick");
      md_chain(c, mp, L, 1);
      md_kin(c, a, c->d_nhc);
      HCK(c, hipMemcpyAsync(c->h_sc + SC_KIN + 1, &c->d_nhc->e_nvt, sizeof(double),
                            hipMemcpyDeviceToHost, st));
    }
    HCK(c, hipMemcpyAsync(c->h_sc + SC_KIN, c->d_sc + SC_KIN, sizeof(double),
                          hipMemcpyDeviceToHost, st));
    HCK(c, hipStreamSynchronize(st));
    prof_collect(c);
    if (steps) {
      mdffgpu_md_step& o = steps[k];
      o.en = en;
      o.e_kin = 0.5 * c->h_sc[SC_KIN];
      // calc_temp: T = 2 ekin / boltz_unit / L, L = 3 natm (src/kinetic.f90:541-553)
      o.temp_r = 2.0 * o.e_kin / mp.boltz_unit;
      o.temp_r = o.temp_r / (3.0 * (double)n);
      o.e_nvt = nvt ? c->h_sc[SC_KIN + 1] : 0.0;
    }
  }
  if (nvt) {
    HCK(c, hipMemcpyAsync(&hs, c->d_nhc, sizeof(hs), hipMemcpyDeviceToHost, st));
    HCK(c, hipStreamSynchronize(st));
    for (int h = 0; h < nhc; ++h) {
      xi[h] = hs.xi[h];
      vxi[h] = hs.vxi[h];
    }
  }
}

int mdffgpu_md_run(mdffgpu_ctx* c, const mdffgpu_md_params* mp, int32_t itime0, int32_t nsteps,
                   double* d_pos, double* d_vel, double* d_force, double* xi, double* vxi,
                   mdffgpu_md_step* steps, void* stream) {
  if (!c) return MDFFGPU_ERR_ARG;
  if (!mp || !d_pos || !d_vel || !d_force || nsteps < 0) {
    c->err = "NULL MD parameter, position, velocity or force argument, or nsteps < 0";
    return MDFFGPU_ERR_ARG;
  }
  if (c->noalloc) {
    c->err = "the MD loop needs a context of mdffgpu_create";
    return MDFFGPU_ERR_ARG;
  }
  const bool nvt = mp->integrator != MDFFGPU_NVE_VV;
  if (mp->integrator < MDFFGPU_NVE_VV || mp->integrator > MDFFGPU_NVT_NHCN) {
    c->err = "integrator must be nve-vv, nvt-nhc2 or nvt-nhcn";
    return MDFFGPU_ERR_UNSUPPORTED;
  }
  double yw[9];
  if (nvt && (!xi || !vxi)) {
    c->err = "the Nose-Hoover integrators need the thermostat coordinates xi, vxi";
    return MDFFGPU_ERR_ARG;
  }
  if (mp->integrator == MDFFGPU_NVT_NHCN &&
      (mp->nhc_n < 2 || mp->nhc_n > MDFFGPU_NHC_MAX || mp->nhc_mults < 1 ||
       !yoshida_weights(mp->nhc_yosh_order, yw))) {
    // nhc_n = 1 is refused by md_check_tag (src/md.f90:302-305); the
    // Yoshida order by chain_nhcn (src/integration.f90:596-599)
    c->err = "nvt-nhcn needs 2 <= nhc_n <= 16, nhc_mults >= 1 and a Yoshida order of 1, 3, 5, 7 "
             "or 9";
    return MDFFGPU_ERR_ARG;
  }
  if (!(mp->dt > 0.0) || !(mp->boltz_unit > 0.0) || (nvt && !(mp->temp > 0.0)) ||
      (nvt && !(mp->timesca_thermo > 0.0))) {
    c->err = "dt and boltz_unit must be positive; temp and timesca_thermo too with a thermostat";
    return MDFFGPU_ERR_ARG;
  }
  for (int t = 0; t < c->P.ntype; ++t)
    if (!(mp->mass[t] > 0.0)) {
      c->err = "every species needs a positive mass";
      return MDFFGPU_ERR_ARG;
    }
  try {
    c->st = stream ? (hipStream_t)stream : c->own_stream;
    if (!c->d_vel) {
      dalloc(c, &c->d_vel, 3 * (size_t)c->n);
      dalloc(c, &c->d_fs, 3 * (size_t)c->n);
      dalloc(c, &c->d_massia, (size_t)c->n);
      dalloc(c, &c->d_nhc, 1);
    }
    ensure_renum_dev(c, d_pos);
    gather_dev(c, c->d_pos, d_pos, 3);
    gather_dev(c, c->d_vel, d_vel, 3);
    gather_dev(c, c->d_force, d_force, 3);
    md_loop(c, *mp, itime0, nsteps, xi, vxi, steps);
    scatter_dev(c, d_pos, c->d_pos, 3);
    scatter_dev(c, d_vel, c->d_vel, 3);
    scatter_dev(c, d_force, c->d_force, 3);
    HCK(c, hipStreamSynchronize(c->st));
    c->st = c->own_stream;
    return MDFFGPU_OK;
  } catch (Fail& f) {
    c->st = c->own_stream;
    return f.code;
  }
}

int mdffgpu_list_info(const mdffgpu_ctx* c, int32_t* builds, int32_t* cell_search) {
  if (!c) return MDFFGPU_ERR_ARG;
  if (builds) *builds = c->list_builds;
  if (cell_search) *cell_search = c->gstencil ? 1 : 0;
  return MDFFGPU_OK;
}

int mdffgpu_engforce_noalloc(const mdffgpu_params* params, const int32_t* itype,
                             const double* A, const double* d_pos, int32_t itime,
                             double* d_hist, double* d_force, double* d_mu, double* d_ef,
                             double* d_efg, mdffgpu_energies* en, void* stream) {
  if (params && (params->precision != MDFFGPU_PREC_FP64 ||
                 params->scf_solver != MDFFGPU_SCF_JACOBI)) {
    g_create_error = "the allocation-free evaluation runs the FP64 path with the Jacobi SCF";
    return MDFFGPU_ERR_UNSUPPORTED;
  }
  if (!d_pos || !d_force || !en) {
    g_create_error = "NULL position, force or energy argument";
    return MDFFGPU_ERR_ARG;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    g_create_error = "no current HIP device";
    return MDFFGPU_ERR_DEVICE;
  }
  std::string err;
  char* arena = scratch_ptr(dev, &err);
  if (!arena) {
    g_create_error = err;
    return MDFFGPU_ERR_DEVICE;
  }
  hipStream_t st = (hipStream_t)stream;
  mdffgpu_ctx* c = nullptr;
  try {
    // every call on a device shares its scratch: calls that overlap on one
    // device race (see the header)
    create_impl(params, itype, A, &c, arena, kScratchBytes, st);
    const size_t n = c->n, b = n * sizeof(double);
    const size_t nh = (size_t)(c->P.extrapolate_order + 1) * 3 * n * sizeof(double);
    if (d_hist) HCK(c, hipMemcpyAsync(c->d_hist, d_hist, nh, hipMemcpyDeviceToDevice, st));
    HCK(c, hipMemcpyAsync(c->d_pos, d_pos, 3 * b, hipMemcpyDeviceToDevice, st));
    HCK(c, hipMemcpyAsync(c->d_force, d_force, 3 * b, hipMemcpyDeviceToDevice, st));
    c->out_force = d_force;
    c->out_mu = d_mu;
    c->out_ef = d_ef;
    c->out_efg = d_efg;
    run(c, itime, en);
    if (d_hist) {
      HCK(c, hipMemcpyAsync(d_hist, c->d_hist, nh, hipMemcpyDeviceToDevice, st));
      HCK(c, hipStreamSynchronize(st));
    }
    mdffgpu_destroy(c);
    return MDFFGPU_OK;
  } catch (Fail& f) {
    if (c) {
      g_create_error = c->err;
      (void)hipStreamSynchronize(st);
      mdffgpu_destroy(c);
    }
    return f.code;
  } catch (...) {
    g_create_error = "unexpected host exception";
    if (c) mdffgpu_destroy(c);
    return MDFFGPU_ERR_DEVICE;
  }
}

int mdffgpu_efg(mdffgpu_ctx* c, const mdffgpu_efg_opts* o, const double* rx, const double* ry,
                const double* rz, double* mu_t, double* efg_t, mdffgpu_energies* en) {
  if (!c) return MDFFGPU_ERR_ARG;
  if (!o || !rx || !ry || !rz || !efg_t) {
    c->err = "NULL options, position or EFG argument";
    return MDFFGPU_ERR_ARG;
  }
  if (c->noalloc) {
    c->err = "not available on an allocation-free context";
    return MDFFGPU_ERR_UNSUPPORTED;
  }
  if (o->lefg_old) {
    c->err = "lefg_old (charge-only efg_ES / efg_DS) is not supported";
    return MDFFGPU_ERR_UNSUPPORTED;
  }
  if (c->lquad || c->lwannier) {
    // multipole_efg_ES / _DS take charges and dipoles (src/efg.f90:1278-1625)
    c->err = "the standalone EFG takes charges and dipoles, not quadrupoles or Wannier centres";
    return MDFFGPU_ERR_UNSUPPORTED;
  }
  if (o->longrange != 0 && o->longrange != 1) {
    c->err = "longrange must be 0 (ewald) or 1 (direct)";
    return MDFFGPU_ERR_ARG;
  }
  if (!c->P.lcoulomb) {
    c->err = "the EFG evaluation needs lcoulomb";
    return MDFFGPU_ERR_ARG;
  }
  if (o->longrange == 1 && c->linduced) {
    c->err = "induced dipoles need the Ewald sum (longrange = 'ewald')";
    return MDFFGPU_ERR_UNSUPPORTED;
  }
  if (o->longrange == 1 && (o->ncelldirect < 0 || o->ncelldirect > 256 || !(o->cutdirect > 0.0))) {
    c->err = "direct sum: ncelldirect must be 0..256 and the cutoff positive";
    return MDFFGPU_ERR_ARG;
  }
  if (o->longrange == 0 && !(c->P.alphaES > 0.0)) {
    c->err = "Ewald sum: alphaES must be positive";
    return MDFFGPU_ERR_ARG;
  }
  if (o->lefg_it_contrib && (o->it_efg < 1 || o->it_efg > c->P.ntype)) {
    c->err = "it_efg must be a species index 1..ntype";
    return MDFFGPU_ERR_ARG;
  }
  try {
    c->st = c->own_stream;
    ensure_renum(c, rx, ry, rz);
    const double* r[3] = {rx, ry, rz};
    gather_host(c, c->d_pos, r, 3);
    mdffgpu_energies tmp;
    efg_frame(c, *o, en ? en : &tmp);
    scatter_host(c, efg_t, c->d_efgout, 9);
    if (mu_t) scatter_host(c, mu_t, c->d_mu, 3);
    HCK(c, hipStreamSynchronize(c->st));
    return MDFFGPU_OK;
  } catch (Fail& f) {
    (void)hipStreamSynchronize(c->st);
    return f.code;
  }
}

int mdffgpu_efg_nmr(int32_t n, const double* efg6, double* nmr, double* U, double* S,
                    int32_t* bad) {
  if (n < 0 || (n > 0 && (!efg6 || !nmr || !U || !S))) {
    g_create_error = "mdffgpu_efg_nmr: NULL array or negative count";
    return MDFFGPU_ERR_ARG;
  }
  if (bad) *bad = 0;
  if (n == 0) return MDFFGPU_OK;
  const size_t b = (size_t)n * sizeof(double);
  double* d = nullptr;
  int* dbad = nullptr;
  int rc = MDFFGPU_OK;
  if (hipMalloc((void**)&d, 16 * b) != hipSuccess || hipMalloc((void**)&dbad, sizeof(int)) != hipSuccess) {
    g_create_error = "mdffgpu_efg_nmr: device allocation failed";
    rc = MDFFGPU_ERR_ALLOC;
  } else if (hipMemcpy(d, efg6, 6 * b, hipMemcpyHostToDevice) != hipSuccess ||
             hipMemset(dbad, 0, sizeof(int)) != hipSuccess) {
    rc = MDFFGPU_ERR_DEVICE;
  } else {
    k_efg_nmr<<<(n + MDFF_BLOCK - 1) / MDFF_BLOCK, MDFF_BLOCK>>>(n, d, d + 6 * n, d + 10 * n,
                                                              d + 15 * n, dbad);
    int hbad = 0;
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(nmr, d + 6 * n, 4 * b, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(U, d + 10 * n, 5 * b, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(S, d + 15 * n, b, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(&hbad, dbad, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
      rc = MDFFGPU_ERR_DEVICE;
    if (bad) *bad = hbad;
  }
  if (rc == MDFFGPU_ERR_DEVICE) g_create_error = "mdffgpu_efg_nmr: HIP runtime failure";
  if (d) (void)hipFree(d);
  if (dbad) (void)hipFree(dbad);
  return rc;
}

int mdffgpu_get_dipole_history(mdffgpu_ctx* c, double* hist) {
  if (!c || !hist) return MDFFGPU_ERR_ARG;
  const int K = c->P.extrapolate_order + 1;
  const size_t cnt = (size_t)K * 3 * c->n;
  try {
    const double* src = c->d_hist;
    if (c->renum) {
      scatter_dev(c, c->d_xbuf, c->d_hist, 3, K);
      src = c->d_xbuf;
    }
    HCK(c, hipMemcpyAsync(hist, src, cnt * sizeof(double), hipMemcpyDeviceToHost, c->st));
    HCK(c, hipStreamSynchronize(c->st));
    return MDFFGPU_OK;
  } catch (Fail& f) {
    return f.code;
  }
}

int mdffgpu_set_dipole_history(mdffgpu_ctx* c, const double* hist) {
  if (!c || !hist) return MDFFGPU_ERR_ARG;
  const int K = c->P.extrapolate_order + 1;
  const size_t cnt = (size_t)K * 3 * c->n;
  try {
    double* dst = c->renum ? c->d_xbuf : c->d_hist;
    HCK(c, hipMemcpyAsync(dst, hist, cnt * sizeof(double), hipMemcpyHostToDevice, c->st));
    if (c->renum) gather_dev(c, c->d_hist, c->d_xbuf, 3, K);
    HCK(c, hipStreamSynchronize(c->st));
    return MDFFGPU_OK;
  } catch (Fail& f) {
    return f.code;
  }
}

int32_t mdffgpu_num_kvectors(const mdffgpu_ctx* c) { return c ? c->nk : 0; }

const char* mdffgpu_last_error(const mdffgpu_ctx* c) {
  if (!c) return g_create_error.c_str();
  return c->err.c_str();
}

int mdffgpu_set_mixed_tuning(mdffgpu_ctx* c, int32_t apply_slices, int32_t split_atoms) {
  if (!c) return MDFFGPU_ERR_ARG;
  if (!c->mx) return MDFFGPU_ERR_UNSUPPORTED;
  if (hipStreamSynchronize(c->st) != hipSuccess) return MDFFGPU_ERR_DEVICE;
  return mixed::set_tuning(c->mx, apply_slices, split_atoms) ? MDFFGPU_OK : MDFFGPU_ERR_ARG;
}

int mdffgpu_profile_enable(mdffgpu_ctx* c, int32_t on) {
  if (!c) return MDFFGPU_ERR_ARG;
  c->prof = on ? 1 : 0;
  c->prof_only = -1;
  for (int k = 0; k < K_COUNT; ++k) {
    c->prof_ms[k] = 0.0;
    c->prof_n[k] = 0;
  }
  return MDFFGPU_OK;
}

int mdffgpu_profile_only(mdffgpu_ctx* c, int32_t kernel) {
  if (!c || kernel < -2 || kernel >= K_COUNT) return MDFFGPU_ERR_ARG;
  c->prof_only = kernel;
  return MDFFGPU_OK;
}

int mdffgpu_profile_read(const mdffgpu_ctx* c, int32_t kernel, double* total_ms,
                         int64_t* launches) {
  if (!c || kernel < 0 || kernel >= K_COUNT || !total_ms || !launches) return MDFFGPU_ERR_ARG;
  *total_ms = c->prof_ms[kernel];
  *launches = c->prof_n[kernel];
  return MDFFGPU_OK;
}

const char* mdffgpu_kernel_name(int32_t kernel) {
  if (kernel < 0 || kernel >= K_COUNT) return "";
  return kKernelNames[kernel];
}

int32_t mdffgpu_kernel_count(void) { return K_COUNT; }

int32_t mdffgpu_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

}  // extern "C"
