/*
 * mdffgpu.h -- C ABI of the MI355X (gfx950) force/field library.
 *
 * This library replaces the body of MDFF's force/field evaluation: the
 * subroutine `engforce_driver` (src/field.f90:1111-1166) and everything it
 * calls -- n-m Lennard-Jones (engforce_nmlj_pbc, src/field.f90:1181-1345),
 * BMHFTD with Tang-Toennies damping (engforce_bmhftd_pbc, :2876-3026), Morse
 * (engforce_morse_pbc, :3039-3185), the polarizable-ion SCF
 * (get_moments :3686, moment_from_pola_scf :3200-3458, induced_moment
 * :1716-1803, extrapolate_dipole_poly/aspc :3807-3992) and the Ewald
 * multipole sum (multipole_ES :1817-2054, multipole_ES_dir :2057-2643,
 * multipole_ES_rec :2646-2873) with per-site electric fields and EFG tensors.
 *
 * The reference has no FFI: engforce_driver takes no arguments and reads and
 * writes Fortran module globals.  Every entry point below takes those globals
 * as plain pointers and sizes instead, with Fortran (column-major) layouts so
 * a Fortran host passes its arrays through ISO_C_BINDING without copies.  No
 * entry point throws; each returns an mdffgpu_status, and the message of the
 * last failure is available from mdffgpu_last_error().
 *
 * Threading: one context per device, calls on a context are synchronous with
 * respect to the host (the *_device entry point is asynchronous on the stream
 * it is given) and are not re-entrant, like the reference's single-threaded
 * MPI ranks.
 */
#ifndef MDFFGPU_H
#define MDFFGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MDFFGPU_NTYPEMAX 16 /* ntypemax, src/config.f90 */
#define MDFFGPU_TT_MAX_ORDER 8 /* maximum_of_TT_expansion, src/tt_damp.f90:8 */
#define MDFFGPU_ABI_VERSION 2

typedef enum mdffgpu_status {
  MDFFGPU_OK = 0,
  MDFFGPU_ERR_ARG = 1,          /* invalid argument or parameter table */
  MDFFGPU_ERR_DEVICE = 2,       /* HIP runtime failure */
  MDFFGPU_ERR_ALLOC = 3,        /* device allocation failure */
  MDFFGPU_ERR_NLIST = 4,        /* neighbour list capacity exceeded
                                   (src/tools.f90:339-342) */
  MDFFGPU_ERR_TT_ORDER = 5,     /* Tang-Toennies order > 8
                                   (src/field.f90:355-358) */
  MDFFGPU_ERR_UNSUPPORTED = 6,  /* an option outside the supported set */
  MDFFGPU_ERR_NO_LIBRARY = 7,   /* reported by host wrappers only */
  MDFFGPU_ERR_WFC = 8           /* an ion without its lwfc Wannier centres
                                   within reach (src/field.f90:3594-3599) */
} mdffgpu_status;

typedef enum mdffgpu_precision {
  MDFFGPU_PREC_FP64 = 0,  /* plain FP64 path: parity with the reference */
  MDFFGPU_PREC_MIXED = 1  /* FP32 pair / erfc / k-space arithmetic, FP64
                             energy, virial and SCF-residual reductions */
} mdffgpu_precision;

/* Bits of mdffgpu_params.mixed_fp64_parts: parts of the mixed path that are
   evaluated with the FP64 kernels instead (accuracy studies). */
typedef enum mdffgpu_mixed_part {
  MDFFGPU_MIXED_FP64_SHORTRANGE = 1, /* BMHFTD / LJ / Morse pair forces */
  MDFFGPU_MIXED_FP64_DIRECT = 2,     /* real-space Ewald (static, final) */
  MDFFGPU_MIXED_FP64_SCF_FIELD = 4,  /* real-space dipole field of the SCF */
  MDFFGPU_MIXED_FP64_KSPACE = 8      /* reciprocal-space sums */
} mdffgpu_mixed_part;

/* SCF solver of the induced dipoles.  JACOBI is the reference's
   moment_from_pola_scf (src/field.f90:3200-3458), the only solver its
   get_moments dispatches (src/field.f90:3716-3718; 'cng' is an accepted
   algo_moment_from_pola value there that computes nothing).  PCG is this
   library's preconditioned conjugate gradient on (alpha^-1 - T) mu = E_stat
   for the same fixed point, with alpha as the preconditioner: fewer Ewald
   calls per solve, results equal to JACOBI's to the SCF tolerance.  The
   batched evaluation (mdffgpu_batch_*) runs its fixed-count Jacobi
   iteration whatever this field says. */
typedef enum mdffgpu_scf_solver {
  MDFFGPU_SCF_JACOBI = 0,
  MDFFGPU_SCF_PCG = 1
} mdffgpu_scf_solver;

typedef enum mdffgpu_ext_algo {
  MDFFGPU_EXT_POLY = 0, /* algo_ext_dipole = 'poly' */
  MDFFGPU_EXT_ASPC = 1  /* algo_ext_dipole = 'aspc' */
} mdffgpu_ext_algo;

/*
 * Force-field parameters as engforce_driver sees them, i.e. AFTER
 * field_check_tag (src/field.f90:282-433: tables symmetrised, isotropic
 * polarizabilities expanded, task_coul set) and ewald_param
 * (src/field.f90:860-908: cutlongrange clamped to half the smallest cell
 * width, alphaES/kES chosen when lautoES).
 *
 * Index conventions (0-based, Fortran column-major, so a Fortran host may
 * pass its module arrays as they are):
 *   pair table      t(it,jt)        -> [it + 16*jt]
 *   dip(it,a)                       -> [it + 16*a]
 *   poldip(it,a,b)                  -> [it + 16*(a + 3*b)]
 *   damping (a,b,c) e.g. ldip_damping(ita,ita,jta)
 *                                   -> [a + 16*(b + 16*c)]
 * Fortran logicals are passed as int32 0/1.
 */
typedef struct mdffgpu_params {
  int32_t natm;  /* number of atoms */
  int32_t ntype; /* number of species, 1..16 */

  /* &controltag (src/control.f90:46-66) */
  int32_t lnmlj, lmorse, lbmhft, lbmhftd, lcoulomb, lsurf;
  /* calc: 0 any other mode, 1 'md', 2 'opt'.  In 'md' and 'opt' the first
     SCF iteration takes the extrapolated dipoles (src/field.f90:3321); the
     ASPC corrector runs in 'md' only (src/field.f90:3370). */
  int32_t calc_mode;
  int32_t trunc; /* ctrunc: 0 notrunc, 1 linear, 2 quadratic */
  double cutshortrange;
  double cutlongrange; /* after the clamp of src/field.f90:879-881 */
  double skindiff;     /* Verlet skin (src/control.f90:225) */
  double coul_unit;    /* 14.3996441494161, or 1 with lreduced */
  double press_unit;   /* 0.00624150964712042, or 1 with lreduced */

  /* n-m Lennard-Jones (src/field.f90:84-95) */
  double plj[256], qlj[256], epslj[256], sigmalj[256];
  /* Morse (src/field.f90:103-107) */
  double epsmor[256], sigmamor[256], rhomor[256];
  /* BMHFTD (src/field.f90:118-122) */
  double Abmhftd[256], Bbmhftd[256], Cbmhftd[256], Dbmhftd[256],
      BDbmhftd[256];

  /* per-type moments (src/field.f90:128-136) */
  double qch[16];
  double dip[48];
  double poldip[144];
  int32_t ldip_polar[16];

  /* Tang-Toennies dipole damping (src/field.f90:144-148) */
  int32_t ldip_damping[4096];
  double dip_pol_damp_b[4096];
  double dip_pol_damp_c[4096];
  int32_t dip_pol_damp_k[4096];

  /* Ewald (src/field.f90:163-167) */
  double alphaES;
  int32_t kES[3];
  /* task_coul(1:3): charge-charge, charge-dipole, dipole-dipole; the
     quadrupole tasks 4..6 are task_coul_quad below. */
  int32_t task_coul[3];

  /* polarizable-ion SCF (src/field.f90:139-143, 150-151) */
  double conv_tol_ind;
  int32_t min_scf_pol_iter, max_scf_pol_iter;
  int32_t extrapolate_order; /* 0..1 for 'poly', 0..5 for 'aspc' */
  int32_t algo_ext_dipole;   /* mdffgpu_ext_algo */
  int32_t doefield, doefg;

  /* this library */
  int32_t precision;        /* mdffgpu_precision */
  int32_t mixed_fp64_parts; /* mdffgpu_mixed_part bits, 0 = all mixed */
  int32_t scf_solver;       /* mdffgpu_scf_solver */
  /* task_coul(4:6) of the reference: charge-quadrupole,
     dipole-quadrupole, quadrupole-quadrupole (src/field.f90:166, set by
     field_check_tag :391-395).  Quadrupoles run on the FP64 path only: a
     mixed-precision context or a batch refuses them with
     MDFFGPU_ERR_UNSUPPORTED. */
  int32_t task_coul_quad[3];
  int32_t reserved[2];

  /* ABI version 2 */
  /* permanent quadrupoles quad(it,a,b) -> [it + 16*(a + 3*b)]
     (src/field.f90:133), symmetric in (a,b) */
  double quad[144];
  /* quadrupole polarizabilities polquad(it,g,a,b)
     -> [it + 16*(g + 3*(a + 3*b))] (src/field.f90:138, polquad_iso already
     expanded as :422-426), symmetric in (a,b); theta_ind(a,b) =
     sum_g polquad(g,a,b) E_g on species with lquad_polar (induced_moment,
     :1716-1803).  The Jacobi SCF only. */
  double polquad[432];
  int32_t lquad_polar[16];
  /* Wannier centres (moment_from_WFc, src/field.f90:3492-3672): lwfc(it) > 0
     centres per ion of species it, < 0 on the species of the centres;
     their dipoles -2 sum (r_c - r_ion) join the total moments when
     ldip_wfc (:3725-3740).  rcut_wfc is compared with the squared distance,
     as the reference does (:3572). */
  int32_t lwfc[16];
  int32_t ldip_wfc;
  int32_t reserved2;
  double rcut_wfc;
} mdffgpu_params;

/* Scalar outputs of one evaluation: the thermodynamic module variables
   written by the hot path (src/thermo.f90:41-76), the stress tensors
   tau_nonb/tau_coul (src/config.f90:51-52, Fortran (3,3) order) and the SCF
   statistics. */
typedef struct mdffgpu_energies {
  double u_lj, vir_lj;
  double u_morse, vir_morse;
  double u_bmhft, vir_bmhft;
  double u_coul;       /* includes u_pol*coul_unit (src/field.f90:1983) */
  double u_pol;        /* internal units, without coul_unit */
  double pvirial_coul; /* tr(tau_coul)/3*press_unit (src/field.f90:2050) */
  double tau_nonb[9];
  double tau_coul[9];
  int32_t n_scf;       /* SCF iterations of the last evaluation */
  int32_t n_es_calls;  /* Ewald evaluations of the last evaluation */
  double rmsd;         /* final SCF residual (src/field.f90:3378-3400) */
  double scf_tol;      /* residual tolerance the SCF applied: conv_tol_ind,
                          or in mixed precision the context's estimate of
                          the FP32 noise floor times 1.5 when that is larger */
} mdffgpu_energies;

typedef struct mdffgpu_ctx mdffgpu_ctx;

/* ---------------------------------------------------------------------
 * Set-up.  Replaces the allocations of initialize_coulomb
 * (src/field.f90:1596-1662: k-vector mesh kpoint_sum_init, ef_t, efg_t,
 * mu_t, theta_t, dipia_ind_t) and the per-atom tables of typeinfo_init
 * (src/read_pos.f90:162-218).  itype is 1-based like config%itype; A is
 * simu_cell%A (column j = lattice vector j).
 *
 * Atom order: every per-atom input and output of a context is in the
 * caller's order (the order of itype).  Inside, a context renumbers its
 * atoms once, at its first evaluation, species by species and by cell,
 * and gathers and scatters across that numbering at each entry point; the
 * reference's pair-order conventions (damping slots, the sign of the
 * quadrupole-quadrupole EFG, the atom a Wannier error names) follow the
 * caller's indices.
 * ------------------------------------------------------------------- */
int mdffgpu_create(const mdffgpu_params* params, const int32_t* itype,
                   const double* A, mdffgpu_ctx** out);

/* Replaces finalize_coulomb (src/field.f90:1668-1706). */
void mdffgpu_destroy(mdffgpu_ctx* ctx);

/* New cell (e.g. NPT): recomputes B, Omega and the k-vector mesh as
   lattice (src/lattice.f90:60-168) and kpoint_sum_init
   (src/kspace.f90:142-207) do. */
int mdffgpu_set_cell(mdffgpu_ctx* ctx, const double* A);

/* ---------------------------------------------------------------------
 * engforce_driver (src/field.f90:1111-1166) on host arrays.
 *   rx,ry,rz  in      Cartesian positions (natm)
 *   itime     in      MD step, used by the dipole extrapolation
 *   fx,fy,fz  in/out  forces; overwritten when LJ or BMHFTD is on, and the
 *                     Coulomb forces are added to them (src/field.f90:1220,
 *                     2904, 1966-1968)
 *   mu_t      out     total dipoles (natm,3)       may be NULL
 *   theta_t   out     total quadrupoles (natm,3,3) may be NULL (zeros)
 *   ef_t      out     electric field (natm,3), without coul_unit
 *                     (src/field.f90:1968-1972)      may be NULL
 *   efg_t     out     EFG (natm,3,3) * coul_unit    may be NULL
 *   en        out     energies, virials, stress, SCF statistics
 * ------------------------------------------------------------------- */
int mdffgpu_engforce(mdffgpu_ctx* ctx, const double* rx, const double* ry,
                     const double* rz, int32_t itime, double* fx, double* fy,
                     double* fz, double* mu_t, double* theta_t, double* ef_t,
                     double* efg_t, mdffgpu_energies* en);

/* The same evaluation on device-resident SoA buffers: d_pos = [x|y|z]
   (3*natm doubles), d_force = [fx|fy|fz] in/out, d_mu = [mx|my|mz],
   d_ef (3*natm), d_efg (9*natm, Fortran (natm,3,3) order); any output may
   be NULL.  stream is a hipStream_t (NULL = default stream).  Returns after
   the scalar outputs are on the host and the device outputs are written. */
int mdffgpu_engforce_device(mdffgpu_ctx* ctx, const double* d_pos,
                            int32_t itime, double* d_force, double* d_mu,
                            double* d_ef, double* d_efg,
                            mdffgpu_energies* en, void* stream);

/* ---------------------------------------------------------------------
 * Device-resident molecular dynamics: the time loop of md_run
 * (src/runmd.f90:233-330) for the integrators below, with positions,
 * velocities and forces kept in HBM between steps; only scalars (energies,
 * kinetic energy, thermostat state) come back to the host each step.
 * Each step is the reference's integrator subroutine with its one
 * engforce_driver call, and the Verlet-list check of the evaluation.
 * ------------------------------------------------------------------- */
typedef enum mdffgpu_integrator {
  MDFFGPU_NVE_VV = 0,   /* 'nve-vv': prop_velocity_verlet
                           (src/integration.f90:131-194) */
  MDFFGPU_NVT_NHC2 = 1, /* 'nvt-nhc2': nhc2 + chain_nhc2 (:205-324) */
  MDFFGPU_NVT_NHCN = 2  /* 'nvt-nhcn': nhcn + chain_nhcn (:476-713), the
                           trotter_order -1 branch nhcn calls */
} mdffgpu_integrator;

#define MDFFGPU_NHC_MAX 16

/* &mdtag after md_check_tag's unit conversions (src/md.f90:316-321). */
typedef struct mdffgpu_md_params {
  int32_t integrator;     /* mdffgpu_integrator */
  int32_t nhc_n;          /* chain length, 2..16 (src/md.f90:302-305; nhc2
                             uses 2 whatever this says) */
  int32_t nhc_yosh_order; /* 1, 3, 5, 7 or 9 (src/integration.f90:596-632) */
  int32_t nhc_mults;      /* >= 1 */
  double dt;              /* dt * time_unit */
  double temp;            /* temp * boltz_unit, i.e. kB T */
  double timesca_thermo;  /* timesca_thermo * time_unit */
  double boltz_unit;      /* for temp_r (src/kinetic.f90:531-556) */
  double mass[MDFFGPU_NTYPEMAX]; /* per species (massia, src/read_pos.f90) */
  int32_t reserved[4];
} mdffgpu_md_params;

/* Scalars of one MD step, after it (the thermodynamic module variables the
   integrators write, src/integration.f90:185-189, 532-541). */
typedef struct mdffgpu_md_step {
  mdffgpu_energies en; /* the step's engforce_driver outputs */
  double e_kin;        /* kinetic energy, 1/2 sum m v^2 */
  double temp_r;       /* 2 e_kin / (3 natm boltz_unit) */
  double e_nvt;        /* thermostat energy (nhc2 :237-243, nhcn :524-541);
                          0 for nve-vv */
} mdffgpu_md_step;

/* nsteps steps from step itime0 (the step index passed to each
 * engforce_driver call, as md_run's itime).
 *   d_pos    in/out  device SoA [x|y|z] positions
 *   d_vel    in/out  device SoA velocities
 *   d_force  in/out  forces at d_pos on entry (md_run evaluates them at
 *                    t = 0, src/runmd.f90:162-168), at the final d_pos on
 *                    return; passed through each evaluation with the
 *                    reference's overwrite/accumulate semantics
 *   xi, vxi  in/out  host thermostat coordinates (nhc_n entries; 2 for
 *                    nhc2), NULL for nve-vv
 *   steps    out     host, nsteps records, may be NULL
 * Returns after the last step's outputs are written. */
int mdffgpu_md_run(mdffgpu_ctx* ctx, const mdffgpu_md_params* mp, int32_t itime0,
                   int32_t nsteps, double* d_pos, double* d_vel, double* d_force,
                   double* xi, double* vxi, mdffgpu_md_step* steps, void* stream);

/* Neighbour-list statistics of the context: list builds so far and
   whether the last one used the cell-grid search (1) or the all-pairs scan
   (0, cells too small for the stencil). */
int mdffgpu_list_info(const mdffgpu_ctx* ctx, int32_t* builds, int32_t* cell_search);

/* ---------------------------------------------------------------------
 * Allocation-free evaluation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code that must not call
 * hipMalloc, e.g. tight serving loops or harnesses that forbid it).  The
 * same evaluation as mdffgpu_create + mdffgpu_engforce_device on a fresh
 * context, with every device buffer taken from a static scratch array of
 * MDFFGPU_NOALLOC_SCRATCH_MIB MiB inside the library (one per device,
 * held until the process ends) instead of hipMalloc; no hipMalloc,
 * hipFree, hipHostMalloc, stream or event creation happens in the call.
 * The dipole history, the only state the context would keep between steps,
 * is the caller's: d_hist (device, dipia_ind_t(extrapolate_order+1, natm,
 * 3) Fortran order, in/out) or NULL for a zero history.
 *   - FP64 precision and the Jacobi SCF only (MDFFGPU_ERR_UNSUPPORTED
 *     otherwise);
 *   - a workload whose buffers (neighbour list, SCF pair table, k mesh and
 *     partial sums) exceed the scratch is refused with MDFFGPU_ERR_ALLOC and
 *     a message (e.g. about 3000 polarizable atoms with a 10 A cutoff fit);
 *   - every call on a device shares that device's scratch: calls that
 *     overlap on one device (two host threads, or two streams) race;
 *   - returns after the outputs are written, like mdffgpu_engforce_device;
 *     the message of a failure is mdffgpu_last_error(NULL).
 * ------------------------------------------------------------------- */
#define MDFFGPU_NOALLOC_SCRATCH_MIB 128
int mdffgpu_engforce_noalloc(const mdffgpu_params* params, const int32_t* itype,
                             const double* A, const double* d_pos, int32_t itime,
                             double* d_hist, double* d_force, double* d_mu,
                             double* d_ef, double* d_efg, mdffgpu_energies* en,
                             void* stream);

/* ---------------------------------------------------------------------
 * Standalone electric field gradients, the per-configuration body of
 * efgcalc (calc = 'efg', src/efg.f90:325-616; ex4 of the reference's
 * examples runs it): positions wrapped into the cell (:424-426), the
 * moments of get_moments (src/field.f90:3686-3765: permanent dipoles plus
 * the induced ones of the SCF, computed by this context as in
 * mdffgpu_engforce with calc_mode 0), then the EFG at every site:
 *   longrange 0 'ewald': multipole_efg_ES (src/efg.f90:1278-1625), every
 *     pair j != i under the minimum image with no cutoff, the reciprocal
 *     sum of the context's k mesh, the self term and the optional charged
 *     cell correction;
 *   longrange 1 'direct': multipole_efg_DS (src/efg.f90:1080-1268), bare
 *     tensors over the image cells -ncelldirect..ncelldirect along each
 *     lattice vector, pairs closer than cutdirect.
 * The context's own Coulomb cutoff, alpha and k mesh serve the SCF and the
 * Ewald reciprocal sum; nothing here changes them.  The direct sum uses none
 * of them (the context still needs valid values, e.g. kES 1 1 1 and a
 * cutlongrange within half the cell).
 * ------------------------------------------------------------------- */
typedef struct mdffgpu_efg_opts {
  int32_t longrange;       /* 0 'ewald', 1 'direct' */
  int32_t ncelldirect;     /* image cells per direction (default 2,
                              src/field.f90:225) */
  double cutdirect;        /* cutlongrange of the direct sum; ewald_param
                              does not clamp it there (src/field.f90:532) */
  int32_t lefg_it_contrib; /* only atoms of species it_efg contribute
                              (src/efg.f90:1163, 1406, 1575, 2922) */
  int32_t it_efg;          /* 1-based species */
  int32_t lmp_correction;  /* + 4 pi Q / (3 V) on the diagonal, Q the
                              total (contributing) charge (:1592-1605) */
  int32_t lefg_vasp_sign;  /* efg_t = -efg_t (src/efg.f90:486-488) */
  int32_t lefg_old;        /* charge-only routines efg_ES / efg_DS:
                              refused with MDFFGPU_ERR_UNSUPPORTED */
  int32_t reserved[3];
} mdffgpu_efg_opts;

/* efg_t out: (natm,3,3) Fortran order, symmetric, times coul_unit and the
   sign option; mu_t out (may be NULL): the moments the EFG was taken with;
   en out (may be NULL): n_scf, rmsd and scf_tol of the SCF, the rest 0.
   rx, ry, rz are host positions. */
int mdffgpu_efg(mdffgpu_ctx* ctx, const mdffgpu_efg_opts* opts,
                const double* rx, const double* ry, const double* rz,
                double* mu_t, double* efg_t, mdffgpu_energies* en);

/* Per-site NMR parameters of efg_stat (src/efg.f90:2233-2269) for n sites:
   efg6 in (n,6) Fortran order with the columns vxx vyy vzz vxy vxz vyz
   (the EFGALL record); nmr out (n,4): principal components ordered by
   nmr_convention (src/efg.f90:2546-2648) as Vxx, Vyy, Vzz and the
   asymmetry eta; U out (n,5) the Czjzek components, S out (n) their norm;
   bad out (may be NULL): sites whose ordering fails the checks of
   src/efg.f90:2618-2629, where the reference stops.  Runs on the current
   device; host arrays. */
int mdffgpu_efg_nmr(int32_t n, const double* efg6, double* nmr, double* U,
                    double* S, int32_t* bad);

/* Induced-dipole history dipia_ind_t(extrapolate_order+1, natm, 3)
   (src/field.f90:182), Fortran order and the caller's atom order, for
   restart and parity runs. */
int mdffgpu_get_dipole_history(mdffgpu_ctx* ctx, double* hist);
int mdffgpu_set_dipole_history(mdffgpu_ctx* ctx, const double* hist);

/* Number of half-space k vectors (src/field.f90:1648). */
int32_t mdffgpu_num_kvectors(const mdffgpu_ctx* ctx);

/* Message of the last failure on ctx (or of the last failed create when
   ctx is NULL).  Never NULL. */
const char* mdffgpu_last_error(const mdffgpu_ctx* ctx);

/* ---------------------------------------------------------------------
 * Kernel instrumentation (no counterpart in the reference, which only has
 * MPI_WTIME accumulators, src/time.f90:35-73).  When enabled, every launch
 * on the context's stream is bracketed by HIP events; the elapsed times are
 * folded in at the next stream synchronisation of an evaluation.
 * ------------------------------------------------------------------- */
int mdffgpu_profile_enable(mdffgpu_ctx* ctx, int32_t on); /* also resets */
/* While enabled, bracket only the launches of one kernel id (-1: all; -2:
 * none, the totals kept), so that a timed run carries the events of the
 * kernel it reports alone, on the steps it samples. */
int mdffgpu_profile_only(mdffgpu_ctx* ctx, int32_t kernel);

/* Launch shape of the mixed path's matrix-core reciprocal sums (picked by
 * the Python autotuner): K slices of the field apply (1, 2 or 4) and atoms
 * per structure-factor split of the sums with charges (a multiple of 128;
 * the dipole-only sums of the SCF and final calls size their splits to fill
 * the device in one round).  MDFFGPU_ERR_UNSUPPORTED on an FP64 context.
 * No counterpart in the reference. */
int mdffgpu_set_mixed_tuning(mdffgpu_ctx* ctx, int32_t apply_slices, int32_t split_atoms);
int mdffgpu_profile_read(const mdffgpu_ctx* ctx, int32_t kernel,
                         double* total_ms, int64_t* launches);
const char* mdffgpu_kernel_name(int32_t kernel);
int32_t mdffgpu_kernel_count(void);

/* ---------------------------------------------------------------------
 * Batched evaluation and force-matching adjoint (no counterpart in the
 * reference, which evaluates one configuration per process; used from
 * the Python tooling for parameter fitting).  Many configurations of
 * uneven atom counts share one model (mdffgpu_params: potential, charges,
 * polarizabilities, damping, precision); each has its own cell, Ewald
 * parameter, k mesh and cutoffs.  Atom arrays are concatenated over the
 * configurations; device arrays are SoA [x|y|z] over all atoms.
 * Supported model: one or two species, BMHFTD/LJ/Morse pair forces,
 * point charges, isotropic induced dipoles with Tang-Toennies damping,
 * no permanent dipoles, quadrupoles or Wannier centres, no surface term.
 * ------------------------------------------------------------------- */
typedef struct mdffgpu_batch mdffgpu_batch;

typedef struct mdffgpu_batch_config {
  int32_t natm;
  int32_t kmax[3];      /* 1..12 */
  double A[9];          /* lattice vectors as columns */
  double alphaES;       /* Ewald alpha */
  double cutlongrange;  /* real-space Ewald cutoff, at most half a width */
  double cutshortrange; /* short-range cutoff, at most half a width */
} mdffgpu_batch_config;

int mdffgpu_batch_create(const mdffgpu_params* model, int32_t nconf,
                         const mdffgpu_batch_config* cfgs,
                         const int32_t* itype, mdffgpu_batch** out);
void mdffgpu_batch_destroy(mdffgpu_batch* b);

/* New parameter values (same species, flags and precision). */
int mdffgpu_batch_set_model(mdffgpu_batch* b, const mdffgpu_params* model);

/* Positions of all atoms (device SoA); builds the neighbour lists, phase
   tables and the SCF pair table.  stream: hipStream_t the inputs were
   produced on (NULL = none). */
int mdffgpu_batch_set_positions(mdffgpu_batch* b, const double* d_pos,
                                void* stream);

/* Forces (device SoA, overwritten) of every atom.  scf_iters > 0: exactly
   that many dipole iterations; 0: until conv_tol_ind (residual over all
   polarizable atoms of the batch). */
int mdffgpu_batch_forces(mdffgpu_batch* b, double* d_force, int32_t scf_iters,
                         int32_t* n_scf, double* rmsd, void* stream);

/* Force-matching loss L = sum |F - F_target|^2 / (3 N) over all atoms and
   its gradient with respect to the parameter vector
     [A, B, C6, C8, BD] per unordered species pair (11, 12, 22),
     [epsmor, sigmamor, rhomor] per unordered pair when lmorse,
     qch per species, polarizability per polarizable species,
     the Tang-Toennies dipole damping b per species t whose own field terms
     are damped (dip_pol_damp_b(t,t,:), one value per species; with
     polarizable species only)
   (mdffgpu_batch_num_params entries).  Pair parameters come from pair
   terms recomputed in the backward pass; charges, polarizabilities and the
   dipole part of b from tangents of the SCF fixed point.  d_force (may be NULL) receives the
   forces.  The SCF iteration count is as in mdffgpu_batch_forces. */
int32_t mdffgpu_batch_num_params(const mdffgpu_batch* b);
int64_t mdffgpu_batch_num_atoms(const mdffgpu_batch* b);
int mdffgpu_batch_loss_grad(mdffgpu_batch* b, const double* d_target,
                            double* d_force, double* loss, double* grad,
                            int32_t scf_iters, int32_t* n_scf, double* rmsd,
                            void* stream);

/* enable != 0: mdffgpu_batch_loss_grad records its launches in a hipGraph
   on the first call (fixed scf_iters > 0 required) and replays it while the
   iteration count, positions and tuning stay the same. */
int mdffgpu_batch_graph(mdffgpu_batch* b, int32_t enable);

/* Launch shape chosen by the autotuner: lanes per atom of the
   reciprocal-space field/force kernel (2, 4, 8 or 16). */
int mdffgpu_batch_set_tuning(mdffgpu_batch* b, int32_t apply_lanes);

/* Launch shape of the FP32 batch's matrix-core reciprocal sums, chosen by
   the autotuner: K slices of the field apply (1, 2 or 4) and atoms per
   structure-factor split (a multiple of 128).  MDFFGPU_ERR_UNSUPPORTED when
   the batch has no such path (FP64, or configurations with different kmax).
   No counterpart in the reference. */
int mdffgpu_batch_set_mixed_tuning(mdffgpu_batch* b, int32_t apply_slices,
                                   int32_t split_atoms);

const char* mdffgpu_batch_last_error(const mdffgpu_batch* b);

/* Number of visible HIP devices (0 without a GPU); never fails. */
int32_t mdffgpu_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* MDFFGPU_H */
