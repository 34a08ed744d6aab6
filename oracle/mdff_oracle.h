/*
 * mdff_oracle.h -- CPU FP64 restatement of MDFF's force/field evaluation.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle exists to check the HIP library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and never as the thing measured or shipped.
 *
 * Every function follows a reference routine (cited file:line, relative to
 * the MDFF source tree) in its formulas, loop structure and summation order,
 * with two deliberate differences: pairs are enumerated as i<j (the
 * reference's lvnlist=.false. branch, the same pair set as its half Verlet
 * list), and positions are not round-tripped through kardir/dirkar in place
 * (src/field.f90:1237,1341), which only moves them by an ULP.
 */
#ifndef MDFF_ORACLE_H
#define MDFF_ORACLE_H

#include "../include/mdffgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_cell {
  double A[9];  /* direct basis, column-major (Fortran simu_cell%A) */
  double B[9];  /* reciprocal basis / Omega, column-major */
  double omega; /* volume */
  double W[3];  /* perpendicular widths WA, WB, WC */
  double anorm[3];
} orc_cell;

/* src/functions.f90:28-40 (Numerical Recipes erfcc) */
double orc_errfc(double x);
/* src/field.f90:3775-3805 + src/tt_damp.f90:13-29 */
void orc_tt_damp(double b, double c, double r, int order, double* f,
                 double* fd);
/* src/lattice.f90:60-168 */
void orc_lattice(const double* A, orc_cell* cell);
/* Number of half-space k vectors, src/field.f90:1648 */
int orc_kmesh_count(const int* kmax);
/* src/kspace.f90:142-207, sorted as reorder_kpt/merge_sort
   (src/kspace.f90:287-345, src/tools.f90:691-734).  Arrays hold
   orc_kmesh_count(kmax) entries; nvec (may be NULL) receives the integer
   triples (nx,ny,nz) of the sorted vectors. */
int orc_kmesh(const orc_cell* cell, const int* kmax, double alpha, double* kx,
              double* ky, double* kz, double* kk, double* Ak, double* kcoe,
              int* nvec);
/* Ewald parameters of ewald_param (src/field.f90:860-908): clamped
   cutlongrange, and the dl_poly-style alpha and kmax. */
void orc_ewald_param(const orc_cell* cell, double epsw, double* cutlongrange,
                     double* alpha, int* kmax);

/* State the reference keeps in modules between calls: the induced-dipole
   history dipia_ind_t(extrapolate_order+1, natm, 3) and the stale u_pol. */
typedef struct orc_state {
  double* hist;
  double u_pol;
} orc_state;

/* Dipole extrapolation of the first SCF iterate in MD
   (src/field.f90:3807-3992), exposed for the tests.  hist is
   dipia_ind_t(extrapolate_order+1, natm, 3) in Fortran order (updated as
   the reference shifts it); mu (natm,3) receives the predictor. */
int orc_aspc_order(const mdffgpu_params* p, int itime);
void orc_aspc_coeffs(int ord, double* b /* [6] */, double* w);
int orc_extrapolate(const mdffgpu_params* p, int itime, double* hist,
                    double* mu);
/* ASPC corrector: mu (in: predictor) <- w alpha E + (1 - w) mu; itype
   1-based; u_pol of alpha E. */
void orc_aspc_correct(const mdffgpu_params* p, const int* itype, int itime,
                      const double* E, double* mu, double* u_pol);

/* The per-configuration body of efgcalc (calc = 'efg', src/efg.f90:
   325-616): positions wrapped (:424-426), get_moments with an SCF from the
   static field when polarizable (src/field.f90:3686-3765), then
   multipole_efg_ES (src/efg.f90:1278-1625) or multipole_efg_DS
   (:1080-1268) as opts->longrange says, symmetrised, the charged-cell
   correction (ewald), times coul_unit and the sign option.  efg_t out
   (natm,3,3), mu_t out (natm,3) or NULL. */
int orc_efgcalc(const mdffgpu_params* p, const mdffgpu_efg_opts* opts,
                const int* itype, const double* A, const double* rx,
                const double* ry, const double* rz, double* mu_t,
                double* efg_t, int* n_scf, double* rmsd);

/* engforce_driver, src/field.f90:1111-1166.  Arguments as
   mdffgpu_engforce; A is the cell.  Returns 0 or an mdffgpu_status. */
int orc_engforce(const mdffgpu_params* p, const int* itype, const double* A,
                 const double* rx, const double* ry, const double* rz,
                 int itime, orc_state* st, double* fx, double* fy, double* fz,
                 double* mu_t, double* theta_t, double* ef_t, double* efg_t,
                 mdffgpu_energies* en);

/* One multipole_ES call (src/field.f90:1817-2054) with explicit moments:
   q (natm), mu (natm,3), theta (natm,3,3) or NULL.  task[6] as
   task_coul(1:6).  ef (natm,3) and efg
   (natm,3,3) are outputs; f? are added to; u_coul, tau_coul (9) and
   pvirial are outputs.  Exposed for the unit tests of the Ewald sum. */
int orc_multipole_es(const mdffgpu_params* p, const int* itype,
                     const double* A, const double* rx, const double* ry,
                     const double* rz, const double* q, const double* mu,
                     const double* theta,
                     const int* task, int damp_ind, int do_efield, int do_efg,
                     int do_forces, int do_stress, double u_pol, double* ef,
                     double* efg, double* fx, double* fy, double* fz,
                     double* u_coul, double* tau_coul, double* pvirial,
                     double* u_parts /* [4]: dir, rec, self, surf (raw) */);

/* The md_run time loop (src/runmd.f90:233-330) for 'nve-vv'
   (prop_velocity_verlet, src/integration.f90:131-194), 'nvt-nhc2' (nhc2 +
   chain_nhc2, :205-324) and 'nvt-nhcn' (nhcn + chain_nhcn, :476-713):
   arguments as mdffgpu_md_run on host arrays; f? are the forces at r? on
   entry.  Returns 0 or an mdffgpu_status. */
int orc_md_run(const mdffgpu_params* p, const int* itype, const double* A,
               const mdffgpu_md_params* mp, int itime0, int nsteps, double* rx,
               double* ry, double* rz, double* vx, double* vy, double* vz,
               double* fx, double* fy, double* fz, double* xi, double* vxi,
               orc_state* st, mdffgpu_md_step* steps);
/* The thermostat pieces, exposed for the tests: Yoshida weights and one
   chain call each (kin in, returns the velocity scale factor). */
int orc_yoshida(int order, double* w);
double orc_chain_nhcn(double kin, double* vxi, double* xi, const double* Q,
                      double L, double temp, double dt, int nhc_n, int yosh,
                      int mults);
double orc_chain_nhc2(double kin, double* vxi, double* xi, const double* Q,
                      double L, double temp, double dt);

#ifdef __cplusplus
}
#endif

#endif
