/*
 * mdff_oracle.c -- CPU FP64 restatement of MDFF's force/field evaluation.
 *
 * TEST INFRASTRUCTURE ONLY (see mdff_oracle.h).  Plain C99, double
 * precision, no vectorisation tricks: each function follows the reference
 * routine cited above it.  Citations are relative to the MDFF source tree.
 *
 * Parts the reference cannot run as written, pinned by textbook closed
 * forms instead (tests/test_oracle_pins.py, dimers):
 *   - Morse: the reference routine is never dispatched and reads
 *     uninitialised erh2/u2 (src/field.f90:3124,3132); this file evaluates
 *     the reference's energy expression, eps ((exp(-rho (r - sigma)) - 1)^2
 *     - 1), which is the Morse potential, and its exact derivative.
 *   - BMHFT without damping (lbmhft): the reference reads uninitialised
 *     fdiff6/fdiff8 (src/field.f90:2900,2975); this file uses 0, which gives
 *     A exp(-B r) - C / r^6 - D / r^8, the limit of the damped form.
 */
#include "mdff_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NT MDFFGPU_NTYPEMAX
#define IX2(i, j) ((i) + NT * (j))
#define IX3(a, b, c) ((a) + NT * ((b) + NT * (c)))

static const double PI = 3.14159265358979323846264338327950;

/* src/functions.f90:28-40 */
double orc_errfc(double x) {
  double z = fabs(x);
  double t = 1.0 / (1.0 + 0.5 * z);
  double r =
      t * exp(-z * z - 1.26551223 +
              t * (1.00002368 +
                   t * (.37409196 +
                        t * (.09678418 +
                             t * (-.18628806 +
                                  t * (.27886807 +
                                       t * (-1.13520398 +
                                            t * (1.48851587 +
                                                 t * (-.82215223 +
                                                      t * .17087277)))))))));
  if (x < 0.0) r = 2.0 - r;
  return r;
}

/* E_TT(k) = 1/k!, src/tt_damp.f90:13-29 */
static double e_tt(int k) {
  double e = 1.0;
  for (int i = 1; i <= k; ++i) e = e / (double)i;
  return e;
}

/* src/field.f90:3775-3805 */
void orc_tt_damp(double b, double c, double r, int order, double* f,
                 double* fd) {
  double br = b * r;
  double expbdr = exp(-br) * c;
  double ff = e_tt(order);
  for (int k = order - 1; k >= 1; --k) ff = ff * br + e_tt(k);
  ff = ff * br + e_tt(0);
  *f = 1.0 - ff * expbdr;
  double brn = 1.0;
  for (int k = 0; k < order; ++k) brn *= br;
  *fd = e_tt(order) * brn * expbdr * b;
}

/* src/lattice.f90:60-168 (EXPRO cross products, volume, widths) */
void orc_lattice(const double* A, orc_cell* c) {
  memcpy(c->A, A, 9 * sizeof(double));
  const double* a1 = A;
  const double* a2 = A + 3;
  const double* a3 = A + 6;
  double* B = c->B;
  /* B(:,1) = a2 x a3, B(:,2) = a3 x a1, B(:,3) = a1 x a2 */
  B[0] = a2[1] * a3[2] - a2[2] * a3[1];
  B[1] = a2[2] * a3[0] - a2[0] * a3[2];
  B[2] = a2[0] * a3[1] - a2[1] * a3[0];
  B[3] = a3[1] * a1[2] - a3[2] * a1[1];
  B[4] = a3[2] * a1[0] - a3[0] * a1[2];
  B[5] = a3[0] * a1[1] - a3[1] * a1[0];
  B[6] = a1[1] * a2[2] - a1[2] * a2[1];
  B[7] = a1[2] * a2[0] - a1[0] * a2[2];
  B[8] = a1[0] * a2[1] - a1[1] * a2[0];
  double omega = B[0] * A[0] + B[1] * A[1] + B[2] * A[2];
  c->omega = omega;
  for (int i = 0; i < 3; ++i) {
    const double* b = B + 3 * i;
    c->W[i] = omega / sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
  }
  for (int i = 0; i < 9; ++i) B[i] = B[i] / omega;
  for (int i = 0; i < 3; ++i) {
    const double* a = A + 3 * i;
    c->anorm[i] = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  }
}

/* src/field.f90:1648 */
int orc_kmesh_count(const int* k) {
  return k[2] + k[1] * (2 * k[2] + 1) + k[0] * (2 * k[1] + 1) * (2 * k[2] + 1);
}

/* stable merge sort of idx by key (src/tools.f90:691-734 keeps equal keys
   in input order) */
static void msort(const double* key, int* idx, int* tmp, int n) {
  if (n < 2) return;
  int na = (n + 1) / 2;
  msort(key, idx, tmp, na);
  msort(key, idx + na, tmp, n - na);
  int i = 0, j = na, k = 0;
  while (i < na && j < n) {
    if (key[idx[i]] <= key[idx[j]])
      tmp[k++] = idx[i++];
    else
      tmp[k++] = idx[j++];
  }
  while (i < na) tmp[k++] = idx[i++];
  while (j < n) tmp[k++] = idx[j++];
  memcpy(idx, tmp, n * sizeof(int));
}

/* src/kspace.f90:142-207, reorder_kpt :287-345 */
int orc_kmesh(const orc_cell* c, const int* kmax, double alpha, double* kxo,
              double* kyo, double* kzo, double* kko, double* Ako,
              double* kcoeo, int* nvec) {
  int nk = orc_kmesh_count(kmax);
  double* buf = (double*)malloc(6 * (size_t)nk * sizeof(double));
  int* nv = (int*)malloc(3 * (size_t)nk * sizeof(int));
  int* idx = (int*)malloc(2 * (size_t)nk * sizeof(int));
  if (!buf || !nv || !idx) {
    free(buf);
    free(nv);
    free(idx);
    return -1;
  }
  double *kx = buf, *ky = buf + nk, *kz = buf + 2 * nk, *kk = buf + 3 * nk,
         *Ak = buf + 4 * nk, *kc = buf + 5 * nk;
  const double tpi = 2.0 * PI;
  const double* B = c->B;
  double alpha2 = alpha * alpha;
  int n = 0;
  for (int nx = 0; nx <= kmax[0]; ++nx) {
    int nymin = (nx == 0) ? 0 : -kmax[1];
    for (int ny = nymin; ny <= kmax[1]; ++ny) {
      int nzmin = (nx == 0 && ny == 0) ? 1 : -kmax[2];
      for (int nz = nzmin; nz <= kmax[2]; ++nz) {
        double x = (double)nx, y = (double)ny, z = (double)nz;
        double vx = tpi * (x * B[0] + y * B[3] + z * B[6]);
        double vy = tpi * (x * B[1] + y * B[4] + z * B[7]);
        double vz = tpi * (x * B[2] + y * B[5] + z * B[8]);
        double k2 = vx * vx + vy * vy + vz * vz;
        kx[n] = vx;
        ky[n] = vy;
        kz[n] = vz;
        kk[n] = k2;
        Ak[n] = exp(-k2 * 0.25 / alpha2) / k2;
        kc[n] = 2.0 * (1.0 / k2 + 1.0 / alpha2 / 4.0);
        nv[3 * n] = nx;
        nv[3 * n + 1] = ny;
        nv[3 * n + 2] = nz;
        ++n;
      }
    }
  }
  for (int i = 0; i < nk; ++i) idx[i] = i;
  msort(kk, idx, idx + nk, nk);
  for (int i = 0; i < nk; ++i) {
    int l = idx[i];
    kxo[i] = kx[l];
    kyo[i] = ky[l];
    kzo[i] = kz[l];
    kko[i] = kk[l];
    Ako[i] = Ak[l];
    kcoeo[i] = kc[l];
    if (nvec) {
      nvec[3 * i] = nv[3 * l];
      nvec[3 * i + 1] = nv[3 * l + 1];
      nvec[3 * i + 2] = nv[3 * l + 2];
    }
  }
  free(buf);
  free(nv);
  free(idx);
  return nk;
}

/* src/field.f90:860-908 (dl_poly-like branch used with lautoES) */
void orc_ewald_param(const orc_cell* c, double epsw, double* cutlongrange,
                     double* alpha, int* kmax) {
  double rcut = 0.5 * fmin(c->W[0], fmin(c->W[1], c->W[2]));
  *cutlongrange = fmin(*cutlongrange, rcut);
  double eps = fmin(fabs(epsw), 0.5);
  double tol = sqrt(fabs(log(eps * (*cutlongrange))));
  double a = sqrt(fabs(log(eps * (*cutlongrange) * tol))) / rcut;
  double tol1 =
      sqrt(-log(eps * (*cutlongrange) * (2.0 * tol * a) * (2.0 * tol * a)));
  for (int i = 0; i < 3; ++i)
    kmax[i] = (int)lround(0.25 + c->anorm[i] * a * tol1 / PI);
  *alpha = a;
}

/* cartesian -> fractional, kardir with BASIS=B (src/lattice.f90:183-213) */
static void kardir(const orc_cell* c, int n, const double* rx,
                   const double* ry, const double* rz, double* s) {
  const double* B = c->B;
  for (int i = 0; i < n; ++i) {
    s[3 * i + 0] = rx[i] * B[0] + ry[i] * B[1] + rz[i] * B[2];
    s[3 * i + 1] = rx[i] * B[3] + ry[i] * B[4] + rz[i] * B[5];
    s[3 * i + 2] = rx[i] * B[6] + ry[i] * B[7] + rz[i] * B[8];
  }
}

/* minimum image in fractional coordinates, then back to Cartesian
   (src/field.f90:1268-1273) */
static double min_image(const orc_cell* c, const double* si, const double* sj,
                        double* r) {
  const double* A = c->A;
  double sx = si[0] - sj[0], sy = si[1] - sj[1], sz = si[2] - sj[2];
  sx = sx - round(sx);
  sy = sy - round(sy);
  sz = sz - round(sz);
  r[0] = sx * A[0] + sy * A[3] + sz * A[6];
  r[1] = sx * A[1] + sy * A[4] + sz * A[7];
  r[2] = sx * A[2] + sy * A[5] + sz * A[8];
  return r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
}

typedef struct sr_out {
  double u, vir, tau[9];
} sr_out;

/* engforce_nmlj_pbc, src/field.f90:1181-1345, tables of
   initialize_param_non_bonded :914-1101 */
static void lj(const mdffgpu_params* p, const int* it, const orc_cell* c,
               const double* s, double* fx, double* fy, double* fz,
               sr_out* o) {
  int n = p->natm;
  double rcutsq = p->cutshortrange * p->cutshortrange;
  double two16 = pow(2.0, 1.0 / 6.0);
  memset(o, 0, sizeof(*o));
  for (int i = 0; i < n; ++i) fx[i] = fy[i] = fz[i] = 0.0;
  for (int ia = 0; ia < n; ++ia) {
    for (int ja = ia + 1; ja < n; ++ja) {
      double r[3];
      double rijsq = min_image(c, s + 3 * ia, s + 3 * ja, r);
      int t = IX2(it[ia], it[ja]);
      if (!(rijsq < rcutsq)) continue;
      double pp = p->plj[t], qq = p->qlj[t];
      double epsp = p->epslj[t] / (qq - pp);
      double sigsq = two16 * two16 * p->sigmalj[t] * p->sigmalj[t];
      double fc = pp * qq * epsp / sigsq;
      double sr2 = sigsq / rijsq;
      double srp = pow(sr2, pp * 0.5);
      double srq = pow(sr2, qq * 0.5);
      double uu = epsp * (pp * srq - qq * srp);
      if (p->trunc == 0) {
        o->u = o->u + uu;
      } else {
        double sr2c = sigsq / rcutsq;
        double src = sqrt(sr2c);
        double srpc = pow(src, pp), srqc = pow(src, qq);
        if (p->trunc == 1) {
          double uc = epsp * (pp * srqc - qq * srpc);
          o->u = o->u + uu - uc;
        } else {
          double uc1 = epsp * (pp * qq) / (2.0 * rcutsq) * (srqc - srpc);
          double uc2 = 0.5 * epsp *
                       ((2.0 * pp + pp * qq) * srqc - (2.0 * qq + pp * qq) * srpc);
          o->u = o->u + uu + uc1 * rijsq - uc2;
        }
      }
      double wij = fc * (srq - srp) * sr2;
      double f[3] = {wij * r[0], wij * r[1], wij * r[2]};
      o->vir = o->vir + wij * rijsq;
      fx[ia] += f[0];
      fy[ia] += f[1];
      fz[ia] += f[2];
      fx[ja] -= f[0];
      fy[ja] -= f[1];
      fz[ja] -= f[2];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          o->tau[a + 3 * b] += (r[a] * f[b] + r[b] * f[a]) * 0.5;
    }
  }
  for (int k = 0; k < 9; ++k) o->tau[k] = o->tau[k] / c->omega / p->press_unit;
  o->vir = o->vir / 3.0;
}

/* engforce_bmhftd_pbc, src/field.f90:2876-3026 */
static void bmhftd(const mdffgpu_params* p, const int* it, const orc_cell* c,
                   const double* s, double* fx, double* fy, double* fz,
                   sr_out* o) {
  int n = p->natm;
  double rcutsq = p->cutshortrange * p->cutshortrange;
  memset(o, 0, sizeof(*o));
  for (int i = 0; i < n; ++i) fx[i] = fy[i] = fz[i] = 0.0;
  for (int ia = 0; ia < n; ++ia) {
    for (int ja = ia + 1; ja < n; ++ja) {
      double r[3];
      double rijsq = min_image(c, s + 3 * ia, s + 3 * ja, r);
      int t = IX2(it[ia], it[ja]);
      if (p->Abmhftd[t] == 0.0) continue;
      if (!(rijsq < rcutsq)) continue;
      double ir2 = 1.0 / rijsq;
      double d = sqrt(rijsq);
      double erh = p->Abmhftd[t] * exp(-p->Bbmhftd[t] * d);
      double ir6 = ir2 * ir2 * ir2;
      double ir8 = ir6 * ir2;
      ir6 = ir6 * p->Cbmhftd[t];
      ir8 = ir8 * p->Dbmhftd[t];
      double f6 = 1.0, f8 = 1.0, fd6 = 0.0, fd8 = 0.0, ir6d = ir6, ir8d = ir8;
      if (p->lbmhftd) {
        orc_tt_damp(p->BDbmhftd[t], 1.0, d, 6, &f6, &fd6);
        orc_tt_damp(p->BDbmhftd[t], 1.0, d, 8, &f8, &fd8);
        ir6d = ir6 * f6;
        ir8d = ir8 * f8;
      }
      double ir7 = 6.0 * ir6d / d;
      double ir9 = 8.0 * ir8d / d;
      o->u = o->u + erh - ir6d - ir8d;
      double wij = p->Bbmhftd[t] * erh - ir7 - ir9 + ir6 * fd6 + ir8 * fd8;
      wij = wij / d;
      double f[3] = {wij * r[0], wij * r[1], wij * r[2]};
      o->vir = o->vir + (r[0] * f[0] + r[1] * f[1] + r[2] * f[2]);
      fx[ia] += f[0];
      fy[ia] += f[1];
      fz[ia] += f[2];
      fx[ja] -= f[0];
      fy[ja] -= f[1];
      fz[ja] -= f[2];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) o->tau[a + 3 * b] += (r[a] * f[b] + r[b] * f[a]);
    }
  }
  o->vir = o->vir / 3.0;
  for (int k = 0; k < 9; ++k)
    o->tau[k] = o->tau[k] / c->omega / p->press_unit * 0.5;
}

/* Morse: energy expression of engforce_morse_pbc (src/field.f90:3120-3122),
   V = eps*((exp(-rho*(r-sigma)) - 1)^2 - 1), with its exact derivative.
   The reference's force line (:3132) reads uninitialised erh2 and omits
   1/r; it is not reproduced.  Forces and stress are added (accumulate). */
static void morse(const mdffgpu_params* p, const int* it, const orc_cell* c,
                  const double* s, double* fx, double* fy, double* fz,
                  sr_out* o) {
  int n = p->natm;
  double rcutsq = p->cutshortrange * p->cutshortrange;
  memset(o, 0, sizeof(*o));
  for (int ia = 0; ia < n; ++ia) {
    for (int ja = ia + 1; ja < n; ++ja) {
      double r[3];
      double rijsq = min_image(c, s + 3 * ia, s + 3 * ja, r);
      int t = IX2(it[ia], it[ja]);
      if (!(rijsq < rcutsq)) continue;
      double rho = p->rhomor[t], eps = p->epsmor[t], sig = p->sigmamor[t];
      double d = sqrt(rijsq);
      double rs = exp(rho * sig);
      double erh = exp(-d * rho);
      double X = erh * rs;
      double expon = (X - 1.0) * (X - 1.0) - 1.0;
      o->u = o->u + eps * expon;
      double wij = 2.0 * eps * rho * X * (X - 1.0) / d;
      double f[3] = {wij * r[0], wij * r[1], wij * r[2]};
      o->vir = o->vir + wij * rijsq;
      fx[ia] += f[0];
      fy[ia] += f[1];
      fz[ia] += f[2];
      fx[ja] -= f[0];
      fy[ja] -= f[1];
      fz[ja] -= f[2];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) o->tau[a + 3 * b] += r[a] * f[b];
    }
  }
  o->vir = o->vir / 3.0;
  for (int k = 0; k < 9; ++k) o->tau[k] = o->tau[k] / c->omega / p->press_unit;
}

/* ---------------------------------------------------------------------- */
/* Ewald multipole sum                                                     */

typedef struct es_ctx {
  const mdffgpu_params* p;
  const int* it;
  orc_cell cell;
  int nk;
  double *kx, *ky, *kz, *kk, *Ak, *kcoe;
  double* s;
} es_ctx;

static int es_init(es_ctx* e, const mdffgpu_params* p, const int* it,
                   const double* A, const double* rx, const double* ry,
                   const double* rz) {
  memset(e, 0, sizeof(*e));
  e->p = p;
  e->it = it;
  orc_lattice(A, &e->cell);
  e->nk = orc_kmesh_count(p->kES);
  e->kx = (double*)malloc(6 * (size_t)e->nk * sizeof(double));
  e->s = (double*)malloc(3 * (size_t)p->natm * sizeof(double));
  if (!e->kx || !e->s) return MDFFGPU_ERR_ALLOC;
  e->ky = e->kx + e->nk;
  e->kz = e->kx + 2 * e->nk;
  e->kk = e->kx + 3 * e->nk;
  e->Ak = e->kx + 4 * e->nk;
  e->kcoe = e->kx + 5 * e->nk;
  orc_kmesh(&e->cell, p->kES, p->alphaES, e->kx, e->ky, e->kz, e->kk, e->Ak,
            e->kcoe, NULL);
  kardir(&e->cell, p->natm, rx, ry, rz, e->s);
  return 0;
}

static void es_free(es_ctx* e) {
  free(e->kx);
  free(e->s);
}

/* multipole_ES_dir, src/field.f90:2057-2643.  mu is (natm,3) and th
   (natm,3,3) Fortran order (th may be NULL: no quadrupoles); task[6] as
   task_coul(1:6).  ef (natm,3), efg (natm,3,3), f (natm) are accumulated;
   tau (9) accumulated before the final scaling.  The quadrupole terms
   (:2542-2601) carry energy, and for quadrupole-quadrupole field and EFG,
   but no force or stress, as in the reference. */
static void es_dir(const es_ctx* e, const double* q, const double* mu,
                   const double* th,
                   const int* task, int damp_ind, int do_efield, int do_efg,
                   int do_forces, int do_stress, double* u_dir, double* ef,
                   double* efg, double* fx, double* fy, double* fz,
                   double* tau) {
  const mdffgpu_params* p = e->p;
  int n = p->natm;
  const int* it = e->it;
  double alpha = p->alphaES;
  double cutsq = p->cutlongrange * p->cutlongrange;
  double alpha2 = alpha * alpha, alpha3 = alpha2 * alpha, alpha5 = alpha3 * alpha2;
  double alpha7 = alpha5 * alpha2;
  double piroot = sqrt(PI);
  double u = 0.0;
  double taud[9] = {0};
  const int lq = th && (task[3] || task[4] || task[5]);
#define EF(i, a) ef[(i) + n * (a)]
#define EFG(i, a, b) efg[(i) + n * ((a) + 3 * (b))]
  for (int ia = 0; ia < n; ++ia) {
    int ita = it[ia];
    double qi = q[ia];
    double mui[3] = {mu[ia], mu[ia + n], mu[ia + 2 * n]};
    int dip_i = (mui[0] != 0.0 || mui[1] != 0.0 || mui[2] != 0.0);
    for (int ja = ia + 1; ja < n; ++ja) {
      int jta = it[ja];
      double qj = q[ja];
      double muj[3] = {mu[ja], mu[ja + n], mu[ja + 2 * n]};
      int dip_j = (muj[0] != 0.0 || muj[1] != 0.0 || muj[2] != 0.0);
      double qij = qi * qj;
      double rij[3];
      double d2 = min_image(&e->cell, e->s + 3 * ia, e->s + 3 * ja, rij);
      if (d2 > cutsq) continue;
      double d = sqrt(d2);
      double d3 = d2 * d;
      double d5 = d3 * d2;
      double dm1 = 1.0 / d;
      double dm3 = dm1 / d2;
      double dm5 = dm3 / d2;
      double dm7 = dm5 / d2;
      int ldamp = p->ldip_damping[IX3(ita, ita, jta)] ||
                  p->ldip_damping[IX3(jta, ita, jta)];
      if (!damp_ind) ldamp = 0;
      double fdamp = 1.0, fdamp2 = 1.0, fdd = 0.0, fdd2 = 0.0;
      if (ldamp) {
        int k1 = IX3(ita, ita, jta), k2 = IX3(jta, ita, jta);
        orc_tt_damp(p->dip_pol_damp_b[k1], p->dip_pol_damp_c[k1], d,
                    p->dip_pol_damp_k[k1], &fdamp, &fdd);
        orc_tt_damp(p->dip_pol_damp_b[k2], p->dip_pol_damp_c[k2], d,
                    p->dip_pol_damp_k[k2], &fdamp2, &fdd2);
      }
      double expon = exp(-alpha2 * d2) / piroot;
      double F0 = orc_errfc(alpha * d);
      double F1 = F0 + 2.0 * alpha * d * expon;
      double F2 = F1 + 4.0 * alpha3 * d3 * expon / 3.0;
      double F3 = F2 + 8.0 * alpha5 * d5 * expon / 15.0;
      double F1d = -fdamp + 1.0;
      double F2d = F1d + (d / 3.0) * fdd;
      double F1d2 = -fdamp2 + 1.0;
      double F2d2 = F1d2 + (d / 3.0) * fdd2;
      double T0 = dm1 * F0;
      double T1[3], T1d[3] = {0}, T1d2[3] = {0};
      for (int a = 0; a < 3; ++a) {
        T1[a] = -rij[a] * dm3;
        if (ldamp) {
          T1d[a] = T1[a] * F1d;
          T1d2[a] = T1[a] * F1d2;
        }
        T1[a] = T1[a] * F1;
      }
      double F2_dm5 = 3.0 * F2 * dm5, F2d_dm5 = 3.0 * F2d * dm5,
             F2d2_dm5 = 3.0 * F2d2 * dm5;
      double F1_dm3 = F1 * dm3, F1d_dm3 = F1d * dm3, F1d2_dm3 = F1d2 * dm3;
      double T2[3][3], T2d[3][3] = {{0}}, T2d2[3][3] = {{0}};
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          T2[a][b] = rij[a] * rij[b] * F2_dm5;
          if (a == b) T2[a][a] = T2[a][a] - F1_dm3;
          if (ldamp) {
            T2d[a][b] = rij[a] * rij[b] * F2d_dm5;
            T2d2[a][b] = rij[a] * rij[b] * F2d2_dm5;
            if (a == b) {
              T2d[a][a] = T2d[a][a] - F1d_dm3;
              T2d2[a][a] = T2d2[a][a] - F1d2_dm3;
            }
          }
        }
      double F3_dm7 = F3 * dm7 * 15.0;
      double T3[3][3][3];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          for (int g = 0; g < 3; ++g) {
            double v = -rij[a] * rij[b] * rij[g] * F3_dm7;
            if (a == b) v = v + rij[g] * F2_dm5;
            if (a == g) v = v + rij[b] * F2_dm5;
            if (b == g) v = v + rij[a] * F2_dm5;
            T3[a][b][g] = v;
          }
      double fij[3] = {0, 0, 0};
      /* charge-charge, :2363-2403 */
      if (task[0]) {
        u = u + qij * T0;
        if (do_efield) {
          for (int a = 0; a < 3; ++a) {
            EF(ia, a) = EF(ia, a) - qj * T1[a];
            EF(ja, a) = EF(ja, a) + qi * T1[a];
          }
          if (ldamp)
            for (int a = 0; a < 3; ++a) {
              EF(ia, a) = EF(ia, a) + qj * T1d[a];
              EF(ja, a) = EF(ja, a) - qi * T1d2[a];
            }
        }
        if (do_efg)
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
              EFG(ia, a, b) = EFG(ia, a, b) - qj * T2[a][b];
              EFG(ja, a, b) = EFG(ja, a, b) - qi * T2[a][b];
            }
        if (do_forces) {
          for (int a = 0; a < 3; ++a) fij[a] = qij * T1[a];
          fx[ia] -= fij[0];
          fy[ia] -= fij[1];
          fz[ia] -= fij[2];
          fx[ja] += fij[0];
          fy[ja] += fij[1];
          fz[ja] += fij[2];
        }
        if (do_stress)
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
              taud[a + 3 * b] = taud[a + 3 * b] - (rij[a] * fij[b] + rij[b] * fij[a]);
      }
      /* dipole-dipole, :2407-2462 */
      if (task[2] && (dip_i || dip_j)) {
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) u = u - mui[a] * T2[a][b] * muj[b];
        if (do_efield)
          for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a) {
              EF(ia, a) = EF(ia, a) + T2[a][b] * muj[b];
              EF(ja, a) = EF(ja, a) + T2[a][b] * mui[b];
            }
        if (do_efg)
          for (int g = 0; g < 3; ++g)
            for (int b = 0; b < 3; ++b)
              for (int a = 0; a < 3; ++a) {
                EFG(ia, a, b) = EFG(ia, a, b) + T3[a][b][g] * muj[g];
                EFG(ja, a, b) = EFG(ja, a, b) - T3[a][b][g] * mui[g];
              }
        if (do_forces) {
          fij[0] = fij[1] = fij[2] = 0.0;
          for (int a = 0; a < 3; ++a)
            for (int g = 0; g < 3; ++g) {
              fij[0] = fij[0] + mui[a] * T3[a][0][g] * muj[g];
              fij[1] = fij[1] + mui[a] * T3[a][1][g] * muj[g];
              fij[2] = fij[2] + mui[a] * T3[a][2][g] * muj[g];
            }
          fx[ia] += fij[0];
          fy[ia] += fij[1];
          fz[ia] += fij[2];
          fx[ja] -= fij[0];
          fy[ja] -= fij[1];
          fz[ja] -= fij[2];
        }
        if (do_stress)
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
              taud[a + 3 * b] = taud[a + 3 * b] + (rij[a] * fij[b] + rij[b] * fij[a]);
      }
      /* charge-dipole, :2468-2538 */
      if (task[1] && (dip_i || dip_j)) {
        for (int g = 0; g < 3; ++g) {
          u = u - qi * T1[g] * muj[g];
          u = u + qj * T1[g] * mui[g];
          if (ldamp) {
            u = u + qi * T1d2[g] * muj[g];
            u = u - qj * T1d[g] * mui[g];
          }
        }
        if (do_forces) {
          fij[0] = fij[1] = fij[2] = 0.0;
          for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a) {
              fij[a] = fij[a] + qi * T2[a][b] * muj[b];
              fij[a] = fij[a] - qj * T2[a][b] * mui[b];
            }
          fx[ia] += fij[0];
          fy[ia] += fij[1];
          fz[ia] += fij[2];
          fx[ja] -= fij[0];
          fy[ja] -= fij[1];
          fz[ja] -= fij[2];
        }
        if (do_stress)
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
              taud[a + 3 * b] = taud[a + 3 * b] + (rij[a] * fij[b] + rij[b] * fij[a]);
        if (ldamp) {
          if (do_forces) {
            fij[0] = fij[1] = fij[2] = 0.0;
            for (int b = 0; b < 3; ++b)
              for (int a = 0; a < 3; ++a) {
                fij[a] = fij[a] + qj * T2d[a][b] * mui[b];
                fij[a] = fij[a] - qi * T2d2[a][b] * muj[b];
              }
            fx[ia] += fij[0];
            fy[ia] += fij[1];
            fz[ia] += fij[2];
            fx[ja] -= fij[0];
            fy[ja] -= fij[1];
            fz[ja] -= fij[2];
          }
          if (do_stress)
            for (int a = 0; a < 3; ++a)
              for (int b = 0; b < 3; ++b)
                taud[a + 3 * b] =
                    taud[a + 3 * b] + (rij[a] * fij[b] + rij[b] * fij[a]);
        }
      }
      if (lq) {
        double quadi[3][3], quadj[3][3];
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) {
            quadi[a][b] = th[ia + n * (a + 3 * b)];
            quadj[a][b] = th[ja + n * (a + 3 * b)];
          }
        /* charge-quadrupole, :2542-2551 */
        if (task[3])
          for (int k = 0; k < 3; ++k)
            for (int j = 0; j < 3; ++j) {
              u = u + qi * T2[k][j] * quadj[k][j] / 3.0;
              u = u + qj * T2[k][j] * quadi[k][j] / 3.0;
            }
        /* dipole-quadrupole, :2552-2565 */
        if (task[4])
          for (int l = 0; l < 3; ++l)
            for (int k = 0; k < 3; ++k)
              for (int j = 0; j < 3; ++j) {
                u = u + mui[l] * T3[l][k][j] * quadj[k][j] / 3.0;
                u = u - muj[l] * T3[l][k][j] * quadi[k][j] / 3.0;
              }
        /* quadrupole-quadrupole, :2566-2601, T4 of :2298-2318 */
        if (task[5]) {
          double d7 = d5 * d2, dm9 = dm7 / d2;
          double F4 = F3 + 16.0 * alpha7 * d7 * expon / 105.0;
          double F4_dm9 = dm9 * F4 * 105.0;
          double T4[3][3][3][3];
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
              for (int g = 0; g < 3; ++g)
                for (int h = 0; h < 3; ++h) {
                  double v = rij[a] * rij[b] * rij[g] * rij[h] * F4_dm9;
                  if (g == h) v = v - rij[a] * rij[b] * F3_dm7;
                  if (b == h) v = v - rij[a] * rij[g] * F3_dm7;
                  if (b == g) v = v - rij[a] * rij[h] * F3_dm7;
                  if (a == h) v = v - rij[b] * rij[g] * F3_dm7;
                  if (a == g) v = v - rij[b] * rij[h] * F3_dm7;
                  if (a == b) v = v - rij[g] * rij[h] * F3_dm7;
                  if (a == b && g == h) v = v + F2_dm5;
                  if (a == g && b == h) v = v + F2_dm5;
                  if (a == h && b == g) v = v + F2_dm5;
                  T4[a][b][g][h] = v;
                }
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
              for (int g = 0; g < 3; ++g)
                for (int h = 0; h < 3; ++h)
                  u = u + quadi[a][b] * T4[a][b][g][h] * quadj[g][h] / 9.0;
          if (do_efield)
            for (int k = 0; k < 3; ++k)
              for (int j = 0; j < 3; ++j)
                for (int a = 0; a < 3; ++a) {
                  EF(ia, a) = EF(ia, a) - T3[a][k][j] * quadj[k][j];
                  EF(ja, a) = EF(ja, a) + T3[a][k][j] * quadi[k][j];
                }
          if (do_efg)
            for (int k = 0; k < 3; ++k)
              for (int j = 0; j < 3; ++j)
                for (int b = 0; b < 3; ++b)
                  for (int a = 0; a < 3; ++a) {
                    EFG(ia, a, b) = EFG(ia, a, b) + T4[a][b][k][j] * quadj[k][j];
                    EFG(ja, a, b) = EFG(ja, a, b) - T4[a][b][k][j] * quadi[k][j];
                  }
        }
      }
    }
  }
  if (do_stress)
    for (int k = 0; k < 9; ++k) tau[k] = taud[k] / e->cell.omega * 0.5;
  *u_dir = u;
#undef EF
#undef EFG
}

/* multipole_ES_rec, src/field.f90:2646-2873 (Cartesian positions) */
static void es_rec(const es_ctx* e, const double* rx, const double* ry,
                   const double* rz, const double* q, const double* mu,
                   const double* th, const int* task, int do_efield, int do_efg, int do_forces,
                   int do_stress, double* u_rec, double* ef, double* efg,
                   double* fx, double* fy, double* fz, double* tau) {
  const mdffgpu_params* p = e->p;
  int n = p->natm;
  double tpi_V = 2.0 * PI / e->cell.omega;
  double fpi_V = tpi_V * 2.0;
  int ldip = task[1] || task[2];
  /* the quadrupoles enter rho only with ldip (:2723-2740 sit behind the
     ldip cycle) */
  int lquad = th && (task[3] || task[4] || task[5]);
  double* ckr = (double*)malloc(2 * (size_t)n * sizeof(double));
  double* skr = ckr + n;
  double u = 0.0;
  double t[9] = {0};
#define EF(i, a) ef[(i) + n * (a)]
#define EFG(i, a, b) efg[(i) + n * ((a) + 3 * (b))]
  for (int ik = 0; ik < e->nk; ++ik) {
    if (e->kk[ik] == 0.0) continue;
    double kx = e->kx[ik], ky = e->ky[ik], kz = e->kz[ik];
    double Ak = e->Ak[ik], kcoe = e->kcoe[ik];
    double rR = 0.0, rI = 0.0;
    for (int ia = 0; ia < n; ++ia) {
      double qi = q[ia];
      double kr = (kx * rx[ia] + ky * ry[ia] + kz * rz[ia]);
      ckr[ia] = cos(kr);
      skr[ia] = sin(kr);
      rR = rR + qi * ckr[ia];
      rI = rI + qi * skr[ia];
      if (!ldip) continue;
      double kmu = (mu[ia] * kx + mu[ia + n] * ky + mu[ia + 2 * n] * kz);
      rR = rR - kmu * skr[ia];
      rI = rI + kmu * ckr[ia];
      if (!lquad) continue;
      double txx = th[ia], tyy = th[ia + 4 * n], tzz = th[ia + 8 * n];
      double txy = th[ia + 3 * n], txz = th[ia + 6 * n], tyz = th[ia + 7 * n];
      double kq = txx * kx * kx + txy * kx * ky + txz * kx * kz;
      kq = kq + txy * kx * ky + tyy * ky * ky + tyz * ky * kz;
      kq = kq + txz * kz * kx + tyz * ky * kz + tzz * kz * kz;
      kq = kq / 3.0;
      rR = rR - kq * ckr[ia];
      rI = rI - kq * skr[ia];
    }
    double str = (rR * rR + rI * rI) * Ak;
    u = u + str;
    for (int ia = 0; ia < n; ++ia) {
      double fxij = 0, fyij = 0, fzij = 0;
      if (do_efield || do_forces) {
        double recarg = Ak * (rI * ckr[ia] - rR * skr[ia]);
        fxij = kx * recarg;
        fyij = ky * recarg;
        fzij = kz * recarg;
      }
      if (do_efield) {
        EF(ia, 0) = EF(ia, 0) - fxij;
        EF(ia, 1) = EF(ia, 1) - fyij;
        EF(ia, 2) = EF(ia, 2) - fzij;
      }
      if (do_efg) {
        double recarg2 = Ak * (rR * ckr[ia] + rI * skr[ia]);
        EFG(ia, 0, 0) = EFG(ia, 0, 0) + kx * kx * recarg2;
        EFG(ia, 1, 1) = EFG(ia, 1, 1) + ky * ky * recarg2;
        EFG(ia, 2, 2) = EFG(ia, 2, 2) + kz * kz * recarg2;
        EFG(ia, 0, 1) = EFG(ia, 0, 1) + kx * ky * recarg2;
        EFG(ia, 0, 2) = EFG(ia, 0, 2) + kx * kz * recarg2;
        EFG(ia, 1, 2) = EFG(ia, 1, 2) + ky * kz * recarg2;
      }
      if (do_forces) {
        double qi = q[ia];
        fx[ia] = fx[ia] - qi * fxij;
        fy[ia] = fy[ia] - qi * fyij;
        fz[ia] = fz[ia] - qi * fzij;
        if (!ldip) continue;
        double recarg2 = Ak * (rR * ckr[ia] + rI * skr[ia]);
        double kmu = (mu[ia] * kx + mu[ia + n] * ky + mu[ia + 2 * n] * kz) * recarg2;
        fx[ia] = fx[ia] + kx * kmu;
        fy[ia] = fy[ia] + ky * kmu;
        fz[ia] = fz[ia] + kz * kmu;
      }
    }
    if (do_stress) {
      double k3[3] = {kx, ky, kz};
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          if (a == b)
            t[a + 3 * b] = t[a + 3 * b] + (1.0 - kcoe * k3[a] * k3[b]) * str;
          else
            t[a + 3 * b] = t[a + 3 * b] - kcoe * k3[a] * k3[b] * str;
        }
    }
  }
  /* half mesh and units, :2820-2867 */
  u = u * 2.0;
  if (do_efield)
    for (int i = 0; i < 3 * n; ++i) ef[i] = ef[i] * 2.0 * fpi_V;
  if (do_efg)
    for (int i = 0; i < 9 * n; ++i) efg[i] = efg[i] * 2.0 * fpi_V;
  if (do_forces)
    for (int i = 0; i < n; ++i) {
      fx[i] = fx[i] * 2.0 * fpi_V;
      fy[i] = fy[i] * 2.0 * fpi_V;
      fz[i] = fz[i] * 2.0 * fpi_V;
    }
  if (do_stress)
    for (int k = 0; k < 9; ++k) tau[k] = t[k] * 2.0 * tpi_V / e->cell.omega;
  *u_rec = u * tpi_V;
  free(ckr);
#undef EF
#undef EFG
}

/* multipole_ES, src/field.f90:1817-2054.  ef/efg overwritten, f added. */
static int es_call(const es_ctx* e, const double* rx, const double* ry,
                   const double* rz, const double* q, const double* mu,
                   const double* th, const int* task, int damp_ind, int do_efield, int do_efg,
                   int do_forces, int do_stress, double u_pol, double* ef,
                   double* efg, double* fx, double* fy, double* fz,
                   double* u_coul, double* tau_coul, double* pvirial,
                   double* parts) {
  const mdffgpu_params* p = e->p;
  int n = p->natm;
  size_t n3 = 3 * (size_t)n, n9 = 9 * (size_t)n;
  double* buf = (double*)calloc(2 * n3 + 2 * n9 + 6 * (size_t)n, sizeof(double));
  if (!buf) return MDFFGPU_ERR_ALLOC;
  double *ef_dir = buf, *ef_rec = ef_dir + n3, *efg_dir = ef_rec + n3,
         *efg_rec = efg_dir + n9, *fdx = efg_rec + n9, *fdy = fdx + n,
         *fdz = fdy + n, *frx = fdz + n, *fry = frx + n, *frz = fry + n;
  double qtot[3] = {0, 0, 0}, mutot[3] = {0, 0, 0}, musq = 0.0, qsq = 0.0;
  for (int ia = 0; ia < n; ++ia) {
    double m0 = mu[ia], m1 = mu[ia + n], m2 = mu[ia + 2 * n];
    mutot[0] += m0;
    mutot[1] += m1;
    mutot[2] += m2;
    musq = musq + (m0 * m0 + m1 * m1 + m2 * m2);
    qtot[0] = qtot[0] + q[ia] * rx[ia];
    qtot[1] = qtot[1] + q[ia] * ry[ia];
    qtot[2] = qtot[2] + q[ia] * rz[ia];
    qsq = qsq + q[ia] * q[ia];
  }
  double qmu[3] = {qtot[0] + mutot[0], qtot[1] + mutot[1], qtot[2] + mutot[2]};
  double tpi_V = 2.0 * PI / e->cell.omega;
  double tpi_3V = tpi_V / 3.0;
  double fpi_3V = tpi_3V * 2.0;
  double alpha2 = p->alphaES * p->alphaES;
  double selfa = p->alphaES / sqrt(PI);
  double selfa2 = 2.0 * selfa * alpha2 / 3.0;
  double u_dir = 0.0, u_rec = 0.0, tau_dir[9] = {0}, tau_rec[9] = {0};
  es_dir(e, q, mu, th, task, damp_ind, do_efield, do_efg, do_forces, do_stress,
         &u_dir, ef_dir, efg_dir, fdx, fdy, fdz, tau_dir);
  es_rec(e, rx, ry, rz, q, mu, th, task, do_efield, do_efg, do_forces, do_stress,
         &u_rec, ef_rec, efg_rec, frx, fry, frz, tau_rec);
  double u_surf = (qtot[0] * qtot[0] + qtot[1] * qtot[1] + qtot[2] * qtot[2]) +
                  2.0 * (qtot[0] * mutot[0] + qtot[1] * mutot[1] + qtot[2] * mutot[2]) +
                  (mutot[0] * mutot[0] + mutot[1] * mutot[1] + mutot[2] * mutot[2]);
  u_surf = u_surf * tpi_3V;
  double u_self = -selfa * qsq - selfa2 * musq;
  double cu = p->coul_unit;
  if (p->lsurf)
    *u_coul = (u_dir + u_rec + u_surf + u_self + u_pol) * cu;
  else
    *u_coul = (u_dir + u_rec + u_self + u_pol) * cu;
  for (int a = 0; a < 3; ++a)
    for (int ia = 0; ia < n; ++ia) {
      size_t k = ia + (size_t)n * a;
      double self = 2.0 * selfa2 * mu[k];
      double v = ef_dir[k] + ef_rec[k];
      if (p->lsurf) v = v + qmu[a];
      ef[k] = v + self;
    }
  for (int b = 0; b < 3; ++b)
    for (int a = 0; a < 3; ++a)
      for (int ia = 0; ia < n; ++ia) {
        size_t k = ia + (size_t)n * (a + 3 * b);
        double self = (a == b) ? -2.0 * selfa2 * q[ia] : 0.0;
        efg[k] = (efg_dir[k] + efg_rec[k] + self) * cu;
      }
  for (int k = 0; k < 9; ++k) tau_coul[k] = (tau_dir[k] + tau_rec[k]) * cu / p->press_unit;
  for (int ia = 0; ia < n; ++ia) {
    if (p->lsurf) {
      fx[ia] = fx[ia] + (frx[ia] + fdx[ia] + (-(q[ia] * qmu[0]) * fpi_3V)) * cu;
      fy[ia] = fy[ia] + (fry[ia] + fdy[ia] + (-(q[ia] * qmu[1]) * fpi_3V)) * cu;
      fz[ia] = fz[ia] + (frz[ia] + fdz[ia] + (-(q[ia] * qmu[2]) * fpi_3V)) * cu;
    } else {
      fx[ia] = fx[ia] + (frx[ia] + fdx[ia]) * cu;
      fy[ia] = fy[ia] + (fry[ia] + fdy[ia]) * cu;
      fz[ia] = fz[ia] + (frz[ia] + fdz[ia]) * cu;
    }
  }
  *pvirial = 1.0 / 3.0 * (tau_coul[0] + tau_coul[4] + tau_coul[8]) * p->press_unit;
  if (parts) {
    parts[0] = u_dir;
    parts[1] = u_rec;
    parts[2] = u_self;
    parts[3] = u_surf;
  }
  free(buf);
  return 0;
}

int orc_multipole_es(const mdffgpu_params* p, const int* itype,
                     const double* A, const double* rx, const double* ry,
                     const double* rz, const double* q, const double* mu,
                     const double* theta,
                     const int* task, int damp_ind, int do_efield, int do_efg,
                     int do_forces, int do_stress, double u_pol, double* ef,
                     double* efg, double* fx, double* fy, double* fz,
                     double* u_coul, double* tau_coul, double* pvirial,
                     double* u_parts) {
  int n = p->natm;
  int* it0 = (int*)malloc((size_t)n * sizeof(int));
  if (!it0) return MDFFGPU_ERR_ALLOC;
  for (int i = 0; i < n; ++i) it0[i] = itype[i] - 1;
  es_ctx e;
  int rc = es_init(&e, p, it0, A, rx, ry, rz);
  if (!rc)
    rc = es_call(&e, rx, ry, rz, q, mu, theta, task, damp_ind, do_efield, do_efg,
                 do_forces, do_stress, u_pol, ef, efg, fx, fy, fz, u_coul,
                 tau_coul, pvirial, u_parts);
  es_free(&e);
  free(it0);
  return rc;
}

/* 3x3 inverse by Gauss-Jordan with partial pivoting (the reference uses
   LAPACK DGETRF/DGETRI, src/read_pos.f90:201-210).  Singular -> zeros. */
static void inv3(const double* m, double* out) {
  double a[3][6];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      a[i][j] = m[i + 3 * j];
      a[i][j + 3] = (i == j) ? 1.0 : 0.0;
    }
  for (int c = 0; c < 3; ++c) {
    int piv = c;
    for (int r = c + 1; r < 3; ++r)
      if (fabs(a[r][c]) > fabs(a[piv][c])) piv = r;
    if (a[piv][c] == 0.0) {
      for (int k = 0; k < 9; ++k) out[k] = 0.0;
      return;
    }
    if (piv != c)
      for (int k = 0; k < 6; ++k) {
        double t = a[c][k];
        a[c][k] = a[piv][k];
        a[piv][k] = t;
      }
    double d = a[c][c];
    for (int k = 0; k < 6; ++k) a[c][k] = a[c][k] / d;
    for (int r = 0; r < 3; ++r)
      if (r != c) {
        double f = a[r][c];
        for (int k = 0; k < 6; ++k) a[r][k] = a[r][k] - f * a[c][k];
      }
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) out[i + 3 * j] = a[i][j + 3];
}

/* induced_moment, src/field.f90:1716-1803.  E and mu are (natm,3); th
   (natm,3,3) receives theta_ind = sum_g polquad(g,a,b) E_g (may be NULL
   when no species has lquad_polar). */
static void induced_moment(const mdffgpu_params* p, const int* it,
                           const double* E, double* mu, double* th, double* u_pol) {
  int n = p->natm;
  double u = 0.0;
  for (int ia = 0; ia < n; ++ia) {
    int t = it[ia];
    for (int a = 0; a < 3; ++a) mu[ia + n * a] = 0.0;
    if (th)
      for (int k = 0; k < 9; ++k) th[ia + n * k] = 0.0;
    if (!p->ldip_polar[t] && !p->lquad_polar[t]) continue;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        mu[ia + n * a] =
            mu[ia + n * a] + p->poldip[t + NT * (a + 3 * b)] * E[ia + n * b];
        if (th)
          for (int g = 0; g < 3; ++g)
            th[ia + n * (a + 3 * b)] = th[ia + n * (a + 3 * b)] +
                                       p->polquad[t + NT * (g + 3 * (a + 3 * b))] * E[ia + n * g];
      }
  }
  for (int ia = 0; ia < n; ++ia) {
    int t = it[ia];
    if (!p->ldip_polar[t]) continue;
    double pol[9], inv[9];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) pol[a + 3 * b] = p->poldip[t + NT * (a + 3 * b)];
    inv3(pol, inv);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        u = u + mu[ia + n * a] * inv[a + 3 * b] * mu[ia + n * b];
  }
  u = u * 0.5;
  double uq = 0.0;
  if (th)
    for (int ia = 0; ia < n; ++ia) {
      int t = it[ia];
      if (!p->lquad_polar[t]) continue;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
          uq = uq + th[ia + n * (a + 3 * b)] * th[ia + n * (a + 3 * b)] / p->polquad[t];
    }
  *u_pol = u + uq / 6.0;
}

static const double B_ASPC[6][6] = {
    {1.0, 0, 0, 0, 0, 0},
    {2.0, -1.0, 0, 0, 0, 0},
    {2.5, -2.0, 0.5, 0, 0, 0},
    {2.8, -2.8, 1.2, -0.2, 0, 0},
    {3.0, -24.0 / 7.0, 27.0 / 14.0, -4.0 / 7.0, 1.0 / 14.0, 0},
    {22.0 / 7.0, -55.0 / 14.0, 55.0 / 21.0, -22.0 / 21.0, 5.0 / 21.0, -1.0 / 42.0}};
static const double W_ASPC[6] = {0.0, 2.0 / 3.0, 0.6, 4.0 / 7.0, 5.0 / 9.0, 6.0 / 11.0};

/* order switch of extrapolate_dipole_aspc (src/field.f90:3918-3927): the
   order ramps up while the history fills */
int orc_aspc_order(const mdffgpu_params* p, int itime) {
  if (itime >= p->extrapolate_order + 1) return p->extrapolate_order;
  if (itime == 1 || itime == 0) return 0;
  return itime - 1;
}

/* B_ASPC(1..ord+1) and w_ASPC of src/field.f90:3929-3967 */
void orc_aspc_coeffs(int ord, double* b, double* w) {
  for (int k = 0; k < 6; ++k) b[k] = (ord >= 0 && ord < 6) ? B_ASPC[ord][k] : 0.0;
  *w = (ord >= 0 && ord < 6) ? W_ASPC[ord] : 0.0;
}

#define H(t, i, a) h[(t) + nh * ((i) + (size_t)n * (a))]
/* The first SCF iterate in MD: extrapolate_dipole_poly (src/field.f90:
   3807-3893, orders 0 and 1) or the ASPC predictor (key 1,
   src/field.f90:3969-3979), with their history shifts.  hist is
   dipia_ind_t(extrapolate_order+1, natm, 3) in Fortran order; mu (natm,3)
   is written. */
int orc_extrapolate(const mdffgpu_params* p, int itime, double* h, double* mu) {
  int n = p->natm;
  int ord = p->extrapolate_order;
  int nh = ord + 1;
  if (p->algo_ext_dipole == MDFFGPU_EXT_POLY) {
    if (ord > 1) return MDFFGPU_ERR_UNSUPPORTED;
    if (ord == 0) {
      for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) mu[i + n * a] = H(0, i, a);
    } else if (itime > ord) {
      for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
          mu[i + n * a] = 2.0 * H(0, i, a) - H(1, i, a);
          H(1, i, a) = H(0, i, a);
        }
    } else {
      for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
          mu[i + n * a] = H(0, i, a);
          H(1, i, a) = H(0, i, a);
        }
    }
    return 0;
  }
  int eo = orc_aspc_order(p, itime);
  for (int i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      double v = 0.0;
      for (int k = 0; k <= eo; ++k) v = v + B_ASPC[eo][k] * H(k, i, a);
      mu[i + n * a] = v;
    }
  for (int k = eo; k >= 1; --k)
    for (int i = 0; i < n; ++i)
      for (int a = 0; a < 3; ++a) H(k, i, a) = H(k - 1, i, a);
  return 0;
}
#undef H

/* ASPC corrector (key 2, src/field.f90:3982-3987): mu = w alpha E +
   (1 - w) mu_p, mu_p the predictor's moments (in mu on entry); u_pol is
   that of alpha E, as induced_moment leaves it.  it holds 0-based types. */
static void aspc_correct(const mdffgpu_params* p, const int* it, int itime,
                         const double* E, double* mu, double* th, double* u_pol) {
  size_t n3 = 3 * (size_t)p->natm;
  double* mup = (double*)malloc(n3 * sizeof(double));
  if (!mup) return;
  int eo = orc_aspc_order(p, itime);
  memcpy(mup, mu, n3 * sizeof(double));
  induced_moment(p, it, E, mu, th, u_pol);
  for (size_t k = 0; k < n3; ++k)
    mu[k] = W_ASPC[eo] * mu[k] + (1.0 - W_ASPC[eo]) * mup[k];
  free(mup);
}

/* the same with 1-based itype (exposed for the tests) */
void orc_aspc_correct(const mdffgpu_params* p, const int* itype, int itime,
                      const double* E, double* mu, double* u_pol) {
  int n = p->natm;
  int* it = (int*)malloc((size_t)n * sizeof(int));
  if (!it) return;
  for (int i = 0; i < n; ++i) it[i] = itype[i] - 1;
  aspc_correct(p, it, itime, E, mu, NULL, u_pol);
  free(it);
}

/* moment_from_pola_scf, src/field.f90:3200-3458 */
static int scf(const es_ctx* e, const double* rx, const double* ry,
               const double* rz, const double* qia, const double* dipia,
               const double* quadia, int itime, orc_state* st, double* mu_ind,
               double* th_ind, double* fx, double* fy, double* fz, int* n_scf,
               double* rmsd_out, int* n_es) {
  const mdffgpu_params* p = e->p;
  int n = p->natm;
  const int* it = e->it;
  size_t n3 = 3 * (size_t)n, n9 = 9 * (size_t)n;
  int linduced = 0, lqpol = 0;
  for (int t = 0; t < p->ntype; ++t) {
    if (p->ldip_polar[t]) linduced = 1;
    if (p->lquad_polar[t]) linduced = lqpol = 1;
  }
  for (size_t k = 0; k < n3; ++k) mu_ind[k] = 0.0;
  for (size_t k = 0; k < n9; ++k) th_ind[k] = 0.0;
  *n_scf = 0;
  *rmsd_out = 0.0;
  if (!linduced) return 0;
  int ldip = 0, lqua = 0;
  for (int k = 0; k < 48; ++k)
    if (p->dip[k] != 0.0) ldip = 1;
  for (int k = 0; k < 144; ++k)
    if (p->quad[k] != 0.0) lqua = 1;
  /* :3255-3268: with static quadrupoles every task of the static call */
  int task_static[6] = {1, ldip || lqua, ldip || lqua, lqua, lqua, lqua};
  int task_ind[6] = {0, 0, 1, 0, 0, 1};
  double* thi = lqpol ? th_ind : NULL;
  double* buf = (double*)calloc(4 * n3 + 9 * (size_t)n + (size_t)n + n3, sizeof(double));
  if (!buf) return MDFFGPU_ERR_ALLOC;
  double *Estat = buf, *Eind = Estat + n3, *E = Eind + n3, *mup = E + n3,
         *efg = mup + n3, *qzero = efg + 9 * (size_t)n, *muzero = qzero + n;
  (void)muzero;
  double u_coul, tau[9], pv;
  for (int i = 0; i < n; ++i) fx[i] = fy[i] = fz[i] = 0.0;
  int rc = es_call(e, rx, ry, rz, qia, dipia, lqua ? quadia : NULL, task_static, 1, 1, 0, 0, 0,
                   st->u_pol, Estat, efg, fx, fy, fz, &u_coul, tau, &pv, NULL);
  if (rc) {
    free(buf);
    return rc;
  }
  ++*n_es;
  for (int i = 0; i < n; ++i) fx[i] = fy[i] = fz[i] = 0.0;
  memcpy(E, Estat, n3 * sizeof(double));
  int ord = p->extrapolate_order;
  int nh = ord + 1;
  double* h = st->hist;
#define H(t, i, a) h[(t) + nh * ((i) + (size_t)n * (a))]
  int iscf = 0;
  double rmsd = 1.79769313486231570e308;
  while ((iscf < p->max_scf_pol_iter && rmsd > p->conv_tol_ind) ||
         iscf < p->min_scf_pol_iter) {
    ++iscf;
    if (iscf != 1 || p->calc_mode == 0) {
      induced_moment(p, it, E, mu_ind, thi, &st->u_pol);
    } else {
      /* first iterate in 'md'/'opt': extrapolated dipoles, :3321-3325;
         theta_ind stays zero (the predictor zeroes it, :3979; the poly
         extrapolation leaves the zeroed array of get_moments, :3711) */
      rc = orc_extrapolate(p, itime, h, mu_ind);
      if (rc) break;
    }
    /* induced field: charges and static moments zeroed, :3290-3303 */
    rc = es_call(e, rx, ry, rz, qzero, mu_ind, thi, task_ind, 0, 1, 0, 0, 0,
                 st->u_pol, Eind, efg, fx, fy, fz, &u_coul, tau, &pv, NULL);
    if (rc) break;
    ++*n_es;
    for (size_t k = 0; k < n3; ++k) E[k] = Estat[k] + Eind[k];
    for (int i = 0; i < n; ++i) fx[i] = fy[i] = fz[i] = 0.0;
    if (iscf == 1 && p->calc_mode == 1 && p->algo_ext_dipole == MDFFGPU_EXT_ASPC)
      aspc_correct(p, it, itime, E, mu_ind, thi, &st->u_pol);
    /* stopping criterion, :3378-3400 */
    rmsd = 0.0;
    int npol = 0;
    for (int ia = 0; ia < n; ++ia) {
      int t = it[ia];
      if (p->ldip_polar[t]) {
        ++npol;
        for (int a = 0; a < 3; ++a) {
          double pa = p->poldip[t + NT * (a + 3 * a)];
          if (pa == 0.0) continue;
          double r = mu_ind[ia + n * a] / pa - E[ia + n * a];
          rmsd = rmsd + r * r;
        }
      }
      if (p->lquad_polar[t]) {
        ++npol;
        for (int a = 0; a < 3; ++a) {
          double pq = p->polquad[t + NT * (a + 3 * (a + 3 * a))];
          if (pq == 0.0) continue;
          double r = th_ind[ia + n * (a + 3 * a)] / pq - E[ia + n * a];
          rmsd = rmsd + r * r;
        }
      }
    }
    rmsd = sqrt(rmsd / (double)npol);
  }
  if (!rc)
    for (int i = 0; i < n; ++i)
      for (int a = 0; a < 3; ++a) H(0, i, a) = mu_ind[i + n * a];
#undef H
  *n_scf = iscf;
  *rmsd_out = rmsd;
  free(buf);
  return rc;
}

/* moment_from_WFc, src/field.f90:3492-3672: the dipole of each ion of a
   species with lwfc > 0 from the Wannier centres (atoms of species with
   lwfc < 0) whose squared minimum-image distance is below rcut_wfc (the
   reference compares the square with rcut_wfc itself, :3572), in index
   order: mu = -2 sum (r_c - r_ion) (:3606-3629).  An ion with a centre
   count other than lwfc is the reference's STOP (:3594-3599):
   MDFFGPU_ERR_WFC.  Ions with lwfc = 0 get no contribution (the
   reference's list pointers of such ions are left unset, :3585-3588). */
static int wfc_dipoles(const es_ctx* e, double* mu) {
  const mdffgpu_params* p = e->p;
  int n = p->natm;
  const int* it = e->it;
  for (size_t k = 0; k < 3 * (size_t)n; ++k) mu[k] = 0.0;
  for (int ia = 0; ia < n; ++ia) {
    int t = it[ia];
    if (p->lwfc[t] <= 0) continue;
    int cnt = 0;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int ja = 0; ja < n; ++ja) {
      if (p->lwfc[it[ja]] >= 0) continue;
      double rij[3];
      double d2 = min_image(&e->cell, e->s + 3 * ia, e->s + 3 * ja, rij);
      if (!(d2 < p->rcut_wfc)) continue;
      ++cnt;
      double rji[3];
      min_image(&e->cell, e->s + 3 * ja, e->s + 3 * ia, rji);
      for (int a = 0; a < 3; ++a) acc[a] = acc[a] + rji[a];
    }
    if (cnt != p->lwfc[t]) return MDFFGPU_ERR_WFC;
    for (int a = 0; a < 3; ++a) mu[ia + (size_t)n * a] = acc[a];
  }
  for (size_t k = 0; k < 3 * (size_t)n; ++k) mu[k] = -2.0 * mu[k];
  return 0;
}

int orc_engforce(const mdffgpu_params* p, const int* itype, const double* A,
                 const double* rx, const double* ry, const double* rz,
                 int itime, orc_state* st, double* fx, double* fy, double* fz,
                 double* mu_t, double* theta_t, double* ef_t, double* efg_t,
                 mdffgpu_energies* en) {
  int n = p->natm;
  if (n <= 0 || p->ntype < 1 || p->ntype > NT) return MDFFGPU_ERR_ARG;
  if (p->algo_ext_dipole == MDFFGPU_EXT_POLY && p->extrapolate_order > 1)
    return MDFFGPU_ERR_UNSUPPORTED;
  if (p->extrapolate_order < 0 || p->extrapolate_order > 5) return MDFFGPU_ERR_ARG;
  memset(en, 0, sizeof(*en));
  int* it = (int*)malloc((size_t)n * sizeof(int));
  if (!it) return MDFFGPU_ERR_ALLOC;
  for (int i = 0; i < n; ++i) it[i] = itype[i] - 1;
  es_ctx e;
  int rc = es_init(&e, p, it, A, rx, ry, rz);
  if (rc) goto done;
  size_t n3 = 3 * (size_t)n, n9 = 9 * (size_t)n;
  sr_out o;
  int any_sr = 0;
  if (p->lnmlj) {
    lj(p, it, &e.cell, e.s, fx, fy, fz, &o);
    en->u_lj = o.u;
    en->vir_lj = o.vir;
    memcpy(en->tau_nonb, o.tau, sizeof(o.tau));
    any_sr = 1;
  }
  if (p->lbmhft || p->lbmhftd) {
    bmhftd(p, it, &e.cell, e.s, fx, fy, fz, &o);
    en->u_bmhft = o.u;
    en->vir_bmhft = o.vir;
    memcpy(en->tau_nonb, o.tau, sizeof(o.tau));
    any_sr = 1;
  }
  if (p->lmorse) {
    if (!any_sr) {
      for (int i = 0; i < n; ++i) fx[i] = fy[i] = fz[i] = 0.0;
      memset(en->tau_nonb, 0, sizeof(en->tau_nonb));
    }
    morse(p, it, &e.cell, e.s, fx, fy, fz, &o);
    en->u_morse = o.u;
    en->vir_morse = o.vir;
    for (int k = 0; k < 9; ++k) en->tau_nonb[k] += o.tau[k];
  }
  if (p->lcoulomb) {
    double* buf = (double*)calloc(6 * n3 + 5 * n9 + (size_t)n, sizeof(double));
    if (!buf) {
      rc = MDFFGPU_ERR_ALLOC;
      goto done;
    }
    double *qia = buf, *dipia = qia + n, *muind = dipia + n3, *mu = muind + n3,
           *ef = mu + n3, *fsave = ef + n3, *efg = fsave + n3, *efgs = efg + n9,
           *quadia = efgs + n9, *thind = quadia + n9, *theta = thind + n9,
           *muwfc = theta + n9;
    (void)efgs;
    for (int i = 0; i < n; ++i) {
      qia[i] = p->qch[it[i]];
      for (int a = 0; a < 3; ++a) dipia[i + n * a] = p->dip[it[i] + NT * a];
      for (int k = 0; k < 9; ++k) quadia[i + n * k] = p->quad[it[i] + NT * k];
    }
    /* get_moments, :3686-3765: forces saved around the SCF */
    memcpy(fsave, fx, n * sizeof(double));
    memcpy(fsave + n, fy, n * sizeof(double));
    memcpy(fsave + 2 * n, fz, n * sizeof(double));
    int n_es = 0;
    rc = scf(&e, rx, ry, rz, qia, dipia, quadia, itime, st, muind, thind, fx, fy, fz,
             &en->n_scf, &en->rmsd, &n_es);
    memcpy(fx, fsave, n * sizeof(double));
    memcpy(fy, fsave + n, n * sizeof(double));
    memcpy(fz, fsave + 2 * n, n * sizeof(double));
    int lwannier = 0;
    for (int t = 0; t < p->ntype; ++t)
      if (p->lwfc[t] > 0) lwannier = 1;
    if (!rc && lwannier) rc = wfc_dipoles(&e, muwfc);
    if (!rc) {
      for (size_t k = 0; k < n3; ++k) mu[k] = dipia[k] + muind[k];
      if (lwannier && p->ldip_wfc)
        for (size_t k = 0; k < n3; ++k) mu[k] = mu[k] + muwfc[k];
      for (size_t k = 0; k < n9; ++k) theta[k] = quadia[k] + thind[k];
      int task[6] = {p->task_coul[0], p->task_coul[1], p->task_coul[2],
                     p->task_coul_quad[0], p->task_coul_quad[1], p->task_coul_quad[2]};
      double pv;
      rc = es_call(&e, rx, ry, rz, qia, mu, theta, task, 1, p->doefield, p->doefg, 1,
                   1, st->u_pol, ef, efg, fx, fy, fz, &en->u_coul,
                   en->tau_coul, &pv, NULL);
      ++n_es;
      en->pvirial_coul = pv;
      en->u_pol = st->u_pol;
      en->n_es_calls = n_es;
      if (mu_t) memcpy(mu_t, mu, n3 * sizeof(double));
      if (theta_t) memcpy(theta_t, theta, n9 * sizeof(double));
      if (ef_t) memcpy(ef_t, ef, n3 * sizeof(double));
      if (efg_t) memcpy(efg_t, efg, n9 * sizeof(double));
    }
    free(buf);
  }
done:
  es_free(&e);
  free(it);
  return rc;
}

/* ---------------------------------------------------------------------- */
/* Standalone EFG (calc = 'efg')                                           */

/* reorder_rpt / merge_sort order of direct_sum_init (src/rspace.f90:64-114):
   image cells n in -nc..nc (x outer, z inner) sorted by |n|^2, stable. */
static int ds_cells(int nc, double* cells) {
  int m = (2 * nc + 1) * (2 * nc + 1) * (2 * nc + 1);
  double* rr = (double*)malloc((size_t)m * sizeof(double));
  double* raw = (double*)malloc(3 * (size_t)m * sizeof(double));
  int* idx = (int*)malloc(2 * (size_t)m * sizeof(int));
  if (!rr || !raw || !idx) {
    free(rr);
    free(raw);
    free(idx);
    return -1;
  }
  int k = 0;
  for (int x = -nc; x <= nc; ++x)
    for (int y = -nc; y <= nc; ++y)
      for (int z = -nc; z <= nc; ++z) {
        raw[3 * k] = x;
        raw[3 * k + 1] = y;
        raw[3 * k + 2] = z;
        rr[k] = (double)(x * x + y * y + z * z);
        idx[k] = k;
        ++k;
      }
  msort(rr, idx, idx + m, m);
  for (int c = 0; c < m; ++c)
    for (int a = 0; a < 3; ++a) cells[3 * c + a] = raw[3 * idx[c] + a];
  free(rr);
  free(raw);
  free(idx);
  return m;
}

/* multipole_efg_DS (src/efg.f90:1080-1268): s wrapped fractional
   coordinates, efg (natm,3,3) accumulated (upper six components). */
static int efg_ds(const mdffgpu_params* p, const mdffgpu_efg_opts* o, const int* it,
                  const orc_cell* c, const double* s, const double* q, const double* mu,
                  double* efg) {
  int n = p->natm;
  int nc = o->ncelldirect;
  int m = (2 * nc + 1) * (2 * nc + 1) * (2 * nc + 1);
  double* cells = (double*)malloc(3 * (size_t)m * sizeof(double));
  if (!cells || ds_cells(nc, cells) != m) {
    free(cells);
    return MDFFGPU_ERR_ALLOC;
  }
  const double* A = c->A;
  double cutsq = o->cutdirect * o->cutdirect;
#define EFG(i, a, b) efg[(i) + n * ((a) + 3 * (b))]
  for (int ia = 0; ia < n; ++ia) {
    const double* si = s + 3 * ia;
    for (int nc_ = 0; nc_ < m; ++nc_) {
      const double* bx = cells + 3 * nc_;
      int central = (bx[0] == 0.0 && bx[1] == 0.0 && bx[2] == 0.0);
      for (int ja = 0; ja < n; ++ja) {
        if (o->lefg_it_contrib && it[ja] != o->it_efg - 1) continue;
        if (central && ja == ia) continue;
        double sx = si[0] - (s[3 * ja] + bx[0]);
        double sy = si[1] - (s[3 * ja + 1] + bx[1]);
        double sz = si[2] - (s[3 * ja + 2] + bx[2]);
        double x = sx * A[0] + sy * A[3] + sz * A[6];
        double y = sx * A[1] + sy * A[4] + sz * A[7];
        double z = sx * A[2] + sy * A[5] + sz * A[8];
        double d2 = x * x + y * y + z * z;
        if (!(d2 < cutsq && d2 != 0.0)) continue;
        double qj = q[ja];
        double mx = mu[ja], my = mu[ja + n], mz = mu[ja + 2 * n];
        double d = sqrt(d2);
        double dm5 = 1.0 / (d2 * d2 * d);
        double dm7 = dm5 / d2 * 3.0;
        double Txx = (3.0 * x * x - d2) * dm5, Tyy = (3.0 * y * y - d2) * dm5,
               Tzz = (3.0 * z * z - d2) * dm5;
        double Txy = (3.0 * x * y) * dm5, Txz = (3.0 * x * z) * dm5, Tyz = (3.0 * y * z) * dm5;
        double Txxx = (5.0 * x * x * x - 3.0 * d2 * x) * dm7;
        double Tyyy = (5.0 * y * y * y - 3.0 * d2 * y) * dm7;
        double Tzzz = (5.0 * z * z * z - 3.0 * d2 * z) * dm7;
        double Txxy = (5.0 * x * x * y - d2 * y) * dm7;
        double Txxz = (5.0 * x * x * z - d2 * z) * dm7;
        double Tyyx = (5.0 * y * y * x - d2 * x) * dm7;
        double Tyyz = (5.0 * y * y * z - d2 * z) * dm7;
        double Tzzx = (5.0 * z * z * x - d2 * x) * dm7;
        double Tzzy = (5.0 * z * z * y - d2 * y) * dm7;
        double Txyz = (5.0 * x * y * z) * dm7;
        EFG(ia, 0, 0) = EFG(ia, 0, 0) - qj * Txx;
        EFG(ia, 1, 1) = EFG(ia, 1, 1) - qj * Tyy;
        EFG(ia, 2, 2) = EFG(ia, 2, 2) - qj * Tzz;
        EFG(ia, 0, 1) = EFG(ia, 0, 1) - qj * Txy;
        EFG(ia, 0, 2) = EFG(ia, 0, 2) - qj * Txz;
        EFG(ia, 1, 2) = EFG(ia, 1, 2) - qj * Tyz;
        EFG(ia, 0, 0) = EFG(ia, 0, 0) - (Txxx * mx + Txxy * my + Txxz * mz);
        EFG(ia, 1, 1) = EFG(ia, 1, 1) - (Tyyx * mx + Tyyy * my + Tyyz * mz);
        EFG(ia, 2, 2) = EFG(ia, 2, 2) - (Tzzx * mx + Tzzy * my + Tzzz * mz);
        EFG(ia, 0, 1) = EFG(ia, 0, 1) - (Txxy * mx + Tyyx * my + Txyz * mz);
        EFG(ia, 0, 2) = EFG(ia, 0, 2) - (Txxz * mx + Txyz * my + Tzzx * mz);
        EFG(ia, 1, 2) = EFG(ia, 1, 2) - (Txyz * mx + Tyyz * my + Tzzy * mz);
      }
    }
  }
#undef EFG
  free(cells);
  return 0;
}

/* multipole_efg_ES (src/efg.f90:1278-1625) without the final symmetrisation,
   correction and units: efg (natm,3,3) upper six = dir + rec + self. */
static int efg_es(const es_ctx* e, const mdffgpu_efg_opts* o, const double* rx,
                  const double* ry, const double* rz, const double* q, const double* mu,
                  double* efg) {
  const mdffgpu_params* p = e->p;
  const int* it = e->it;
  int n = p->natm;
  const int itc = o->lefg_it_contrib ? o->it_efg - 1 : -1;
  double alpha = p->alphaES, alpha2 = alpha * alpha, alpha3 = alpha2 * alpha,
         alpha5 = alpha3 * alpha2;
  double piroot = sqrt(PI);
  double selfa = -4.0 * alpha3 / 3.0 / piroot;
  double fpi_V = 4.0 * PI / e->cell.omega;
  double* rec = (double*)calloc(6 * (size_t)n + 2 * (size_t)e->nk, sizeof(double));
  if (!rec) return MDFFGPU_ERR_ALLOC;
  double* rhoR = rec + 6 * (size_t)n;
  double* rhoI = rhoR + e->nk;
  /* charge_density_k (src/efg.f90:2888-2930) */
  for (int ik = 0; ik < e->nk; ++ik) {
    double kx = e->kx[ik], ky = e->ky[ik], kz = e->kz[ik];
    double sr = 0.0, si = 0.0;
    for (int ia = 0; ia < n; ++ia) {
      double kr = kx * rx[ia] + ky * ry[ia] + kz * rz[ia];
      double kmu = mu[ia] * kx + mu[ia + n] * ky + mu[ia + 2 * n] * kz;
      double c = cos(kr), s = sin(kr);
      if (itc >= 0 && it[ia] != itc) continue;
      /* (q + i k.mu)(c + i s) */
      sr = sr + (q[ia] * c - kmu * s);
      si = si + (q[ia] * s + kmu * c);
    }
    rhoR[ik] = sr;
    rhoI[ik] = si;
  }
  /* direct space, every pair under the minimum image (:1401-1499) */
#define EFG(i, a, b) efg[(i) + n * ((a) + 3 * (b))]
  for (int ia = 0; ia < n; ++ia) {
    for (int ja = 0; ja < n; ++ja) {
      if (itc >= 0 && it[ja] != itc) continue;
      if (ja == ia) continue;
      double r[3];
      double d2 = min_image(&e->cell, e->s + 3 * ia, e->s + 3 * ja, r);
      double x = r[0], y = r[1], z = r[2];
      double qj = q[ja], mx = mu[ja], my = mu[ja + n], mz = mu[ja + 2 * n];
      double d = sqrt(d2), d3 = d2 * d, d5 = d3 * d2;
      double dm5 = 1.0 / d2 / d3;
      double dm7 = dm5 / d2 * 3.0;
      double expon = exp(-alpha2 * d2) / piroot;
      double F1 = orc_errfc(alpha * d) + 2.0 * alpha * d * expon;
      double F2 = F1 + 4.0 * alpha3 * d3 * expon / 3.0;
      double F3 = F2 + 8.0 * alpha5 * d5 * expon / 15.0;
      double Txx = (3.0 * x * x * F2 - d2 * F1) * dm5;
      double Tyy = (3.0 * y * y * F2 - d2 * F1) * dm5;
      double Tzz = (3.0 * z * z * F2 - d2 * F1) * dm5;
      double Txy = (3.0 * x * y * F2) * dm5;
      double Txz = (3.0 * x * z * F2) * dm5;
      double Tyz = (3.0 * y * z * F2) * dm5;
      double Txxx = (-5.0 * x * x * x * F3 + 3.0 * d2 * (x)*F2) * dm7;
      double Tyyy = (-5.0 * y * y * y * F3 + 3.0 * d2 * (y)*F2) * dm7;
      double Tzzz = (-5.0 * z * z * z * F3 + 3.0 * d2 * (z)*F2) * dm7;
      double Txxy = (-5.0 * x * x * y * F3 + d2 * (y)*F2) * dm7;
      double Txxz = (-5.0 * x * x * z * F3 + d2 * (z)*F2) * dm7;
      double Tyyx = (-5.0 * y * y * x * F3 + d2 * (x)*F2) * dm7;
      double Tyyz = (-5.0 * y * y * z * F3 + d2 * (z)*F2) * dm7;
      double Tzzx = (-5.0 * z * z * x * F3 + d2 * (x)*F2) * dm7;
      double Tzzy = (-5.0 * z * z * y * F3 + d2 * (y)*F2) * dm7;
      double Txyz = (-5.0 * x * y * z * F3) * dm7;
      EFG(ia, 0, 0) = EFG(ia, 0, 0) - qj * Txx;
      EFG(ia, 1, 1) = EFG(ia, 1, 1) - qj * Tyy;
      EFG(ia, 2, 2) = EFG(ia, 2, 2) - qj * Tzz;
      EFG(ia, 0, 1) = EFG(ia, 0, 1) - qj * Txy;
      EFG(ia, 0, 2) = EFG(ia, 0, 2) - qj * Txz;
      EFG(ia, 1, 2) = EFG(ia, 1, 2) - qj * Tyz;
      EFG(ia, 0, 0) = EFG(ia, 0, 0) + (Txxx * mx + Txxy * my + Txxz * mz);
      EFG(ia, 1, 1) = EFG(ia, 1, 1) + (Tyyx * mx + Tyyy * my + Tyyz * mz);
      EFG(ia, 2, 2) = EFG(ia, 2, 2) + (Tzzx * mx + Tzzy * my + Tzzz * mz);
      EFG(ia, 0, 1) = EFG(ia, 0, 1) + (Txxy * mx + Tyyx * my + Txyz * mz);
      EFG(ia, 0, 2) = EFG(ia, 0, 2) + (Txxz * mx + Txyz * my + Tzzx * mz);
      EFG(ia, 1, 2) = EFG(ia, 1, 2) + (Txyz * mx + Tyyz * my + Tzzy * mz);
    }
  }
  /* reciprocal space (:1507-1556): recarg = Re(conj(rho) e^{ikr}) A_k */
  for (int ik = 0; ik < e->nk; ++ik) {
    if (e->kk[ik] == 0.0) continue;
    double kx = e->kx[ik], ky = e->ky[ik], kz = e->kz[ik];
    for (int ia = 0; ia < n; ++ia) {
      double kr = rx[ia] * kx + ry[ia] * ky + rz[ia] * kz;
      double c = cos(kr), s = sin(kr);
      double recarg = (rhoR[ik] * c + rhoI[ik] * s) * e->Ak[ik];
      rec[ia] += kx * kx * recarg;
      rec[ia + n] += ky * ky * recarg;
      rec[ia + 2 * n] += kz * kz * recarg;
      rec[ia + 3 * n] += kx * ky * recarg;
      rec[ia + 4 * n] += kx * kz * recarg;
      rec[ia + 5 * n] += ky * kz * recarg;
    }
  }
  for (int ia = 0; ia < n; ++ia) {
    double f = fpi_V * 2.0;
    double self = (itc >= 0 && it[ia] != itc) ? 0.0 : selfa * q[ia];
    EFG(ia, 0, 0) = EFG(ia, 0, 0) + rec[ia] * f + self;
    EFG(ia, 1, 1) = EFG(ia, 1, 1) + rec[ia + n] * f + self;
    EFG(ia, 2, 2) = EFG(ia, 2, 2) + rec[ia + 2 * n] * f + self;
    EFG(ia, 0, 1) = EFG(ia, 0, 1) + rec[ia + 3 * n] * f;
    EFG(ia, 0, 2) = EFG(ia, 0, 2) + rec[ia + 4 * n] * f;
    EFG(ia, 1, 2) = EFG(ia, 1, 2) + rec[ia + 5 * n] * f;
  }
#undef EFG
  free(rec);
  return 0;
}

int orc_efgcalc(const mdffgpu_params* p, const mdffgpu_efg_opts* o, const int* itype,
                const double* A, const double* rx0, const double* ry0, const double* rz0,
                double* mu_t, double* efg_t, int* n_scf, double* rmsd) {
  int n = p->natm;
  if (n <= 0 || p->ntype < 1 || p->ntype > NT) return MDFFGPU_ERR_ARG;
  if (o->lefg_old) return MDFFGPU_ERR_UNSUPPORTED;
  if (o->longrange != 0 && o->longrange != 1) return MDFFGPU_ERR_ARG;
  if (o->longrange == 1 && o->ncelldirect < 0) return MDFFGPU_ERR_ARG;
  int linduced = 0;
  for (int t = 0; t < p->ntype; ++t)
    if (p->ldip_polar[t]) linduced = 1;
  if (linduced && o->longrange == 1) return MDFFGPU_ERR_UNSUPPORTED;
  size_t n3 = 3 * (size_t)n, n9 = 9 * (size_t)n;
  int* it = (int*)malloc((size_t)n * sizeof(int));
  double* buf = (double*)calloc(3 * (size_t)n + 4 * n3 + (size_t)n + n9 + n3, sizeof(double));
  if (!it || !buf) {
    free(it);
    free(buf);
    return MDFFGPU_ERR_ALLOC;
  }
  for (int i = 0; i < n; ++i) it[i] = itype[i] - 1;
  double *rw = buf, *q = rw + n3, *dip = q + n, *muind = dip + n3, *mu = muind + n3,
         *f = mu + n3, *efg = f + n3, *hist = efg + n9;
  /* kardir, periodicbc, dirkar (src/efg.f90:424-426) */
  orc_cell cell;
  orc_lattice(A, &cell);
  for (int i = 0; i < n; ++i) {
    double r3[3] = {rx0[i], ry0[i], rz0[i]}, s[3];
    for (int a = 0; a < 3; ++a)
      s[a] = r3[0] * cell.B[3 * a] + r3[1] * cell.B[3 * a + 1] + r3[2] * cell.B[3 * a + 2];
    for (int a = 0; a < 3; ++a) s[a] = s[a] - round(s[a]);
    for (int a = 0; a < 3; ++a)
      rw[i + n * a] = s[0] * cell.A[a] + s[1] * cell.A[3 + a] + s[2] * cell.A[6 + a];
  }
  const double *rx = rw, *ry = rw + n, *rz = rw + 2 * n;
  for (int i = 0; i < n; ++i) {
    q[i] = p->qch[it[i]];
    for (int a = 0; a < 3; ++a) dip[i + n * a] = p->dip[it[i] + NT * a];
  }
  es_ctx e;
  int rc = es_init(&e, p, it, A, rx, ry, rz);
  *n_scf = 0;
  *rmsd = 0.0;
  if (!rc && linduced) {
    /* get_moments with calc = 'efg': SCF from the static field */
    mdffgpu_params pe = *p;
    pe.calc_mode = 0;
    e.p = &pe;
    orc_state st;
    st.hist = hist;
    st.u_pol = 0.0;
    int n_es = 0;
    /* the standalone EFG takes charges and dipoles only (quadrupoles are
       refused by its caller) */
    double* zq = (double*)calloc(18 * (size_t)n, sizeof(double));
    if (!zq) {
      rc = MDFFGPU_ERR_ALLOC;
    } else {
      rc = scf(&e, rx, ry, rz, q, dip, zq, 0, &st, muind, zq + 9 * (size_t)n, f, f + n,
               f + 2 * n, n_scf, rmsd, &n_es);
      free(zq);
    }
    e.p = p;
  }
  if (!rc) {
    for (size_t k = 0; k < n3; ++k) mu[k] = dip[k] + muind[k];
    if (o->longrange == 0) {
      rc = efg_es(&e, o, rx, ry, rz, q, mu, efg);
    } else {
      double* s = (double*)malloc(n3 * sizeof(double));
      if (!s) {
        rc = MDFFGPU_ERR_ALLOC;
      } else {
        for (int i = 0; i < n; ++i)
          for (int a = 0; a < 3; ++a) s[3 * i + a] = e.s[3 * i + a];
        rc = efg_ds(p, o, it, &cell, s, q, mu, efg);
        free(s);
      }
    }
  }
  if (!rc) {
    double corr = 0.0;
    if (o->lmp_correction && o->longrange == 0) {
      for (int ia = 0; ia < n; ++ia) {
        if (o->lefg_it_contrib && it[ia] != o->it_efg - 1) continue;
        corr = corr + q[ia];
      }
      corr = corr * (4.0 * PI / cell.omega) / 3.0;
    }
    double sign = o->lefg_vasp_sign ? -1.0 : 1.0;
    for (int ia = 0; ia < n; ++ia)
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          int lo = a < b ? a : b, hi = a < b ? b : a;
          double v = efg[ia + n * (lo + 3 * hi)];
          if (a == b) v = v + corr;
          efg_t[ia + n * (a + 3 * b)] = sign * (v * p->coul_unit);
        }
    if (mu_t) memcpy(mu_t, mu, n3 * sizeof(double));
  }
  es_free(&e);
  free(it);
  free(buf);
  return rc;
}

/* ------------------------------------------------------------------------
 * Molecular dynamics time loop: the integrators md_run dispatches
 * (src/runmd.f90:244-251) for 'nve-vv', 'nvt-nhc2' and 'nvt-nhcn'.
 * ---------------------------------------------------------------------- */

/* calc_temp (src/kinetic.f90:531-556): ekin = 1/2 sum m v^2. */
static double orc_kin(int n, const double* m, const double* vx, const double* vy,
                      const double* vz) {
  double e = 0.0;
  for (int i = 0; i < n; ++i) e = e + (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]) * m[i];
  return e * 0.5;
}

/* Yoshida weights of chain_nhcn (src/integration.f90:596-632). */
int orc_yoshida(int order, double* w) {
  switch (order) {
    case 1: w[0] = 1.0; return 0;
    case 3:
      w[0] = 1.0 / (2.0 - pow(2.0, 1.0 / 3.0));
      w[1] = 1.0 - 2.0 * w[0];
      w[2] = w[0];
      return 0;
    case 5:
      w[0] = 1.0 / (4.0 - pow(4.0, 1.0 / 3.0));
      w[1] = w[0]; w[2] = w[0]; w[3] = w[0];
      w[4] = 1.0 - 4.0 * w[0];
      return 0;
    case 7:
      w[0] = .78451361047756; w[1] = .235573213359357; w[2] = -1.17767998417887;
      w[3] = 1.0 - 2.0 * (w[0] + w[1] + w[2]);
      w[4] = w[2]; w[5] = w[1]; w[6] = w[0];
      return 0;
    case 9:
      w[0] = 0.192; w[1] = 0.554910818409783619692725006662999;
      w[2] = 0.124659619941888644216504240951585; w[3] = -0.843182063596933505315033808282941;
      w[4] = 1.0 - 2.0 * (w[0] + w[1] + w[2] + w[3]);
      w[5] = w[3]; w[6] = w[2]; w[7] = w[1]; w[8] = w[0];
      return 0;
    default: return MDFFGPU_ERR_ARG;
  }
}

/* chain_nhcn (src/integration.f90:564-713) with trotter_order = -1 (the
   branch nhcn calls): returns the scale factor; the caller scales the
   velocities and recomputes kin.  G is zeroed per call as the reference's
   allocate + G = 0 (:634-635). */
double orc_chain_nhcn(double kin, double* vxi, double* xi, const double* Q, double L,
                      double temp, double dt, int nhc_n, int yosh, int mults) {
  double w[9], G[MDFFGPU_NHC_MAX];
  orc_yoshida(yosh, w);
  for (int k = 0; k < nhc_n; ++k) G[k] = 0.0;
  double s = 1.0;
  G[0] = 2.0 * kin - L * temp;
  const int N = nhc_n - 1;
  for (int k = 0; k < mults; ++k) {
    for (int j = 0; j < yosh; ++j) {
      double dts = (w[j] * dt) / (double)mults;
      double dts2 = dts * 0.5, dts4 = dts2 * 0.5, dts8 = dts4 * 0.5;
      G[N] = vxi[N - 1] * vxi[N - 1] / Q[N - 1] - temp;
      vxi[N] = vxi[N] + G[N] * dts4;
      for (int h = N - 1; h >= 0; --h) {
        vxi[h] = vxi[h] * exp(-vxi[h + 1] * dts8 / Q[h + 1]);
        vxi[h] = vxi[h] + G[h] * dts4;
        vxi[h] = vxi[h] * exp(-vxi[h + 1] * dts8 / Q[h + 1]);
      }
      s = s * exp(-vxi[0] * dts2 / Q[0]);
      for (int h = 0; h <= N; ++h) xi[h] = xi[h] + vxi[h] * dts2 / Q[h];
      G[0] = 2.0 * kin * s * s - L * temp;
      for (int h = 0; h < N; ++h) {
        vxi[h] = vxi[h] * exp(-vxi[h + 1] * dts8 / Q[h + 1]);
        vxi[h] = vxi[h] + G[h] * dts4;
        vxi[h] = vxi[h] * exp(-vxi[h + 1] * dts8 / Q[h + 1]);
        G[h + 1] = vxi[h] * vxi[h] / Q[h] - temp;
      }
      vxi[N] = vxi[N] + G[N] * dts4;
    }
  }
  return s;
}

/* chain_nhc2 (src/integration.f90:261-324): returns s. */
double orc_chain_nhc2(double kin, double* vxi, double* xi, const double* Q, double L,
                      double temp, double dt) {
  double s = 1.0;
  double dt2 = dt * 0.5, dt4 = dt2 * 0.5, dt8 = dt4 * 0.5;
  double G1 = (2.0 * kin - L * temp) / Q[0];
  double G2 = (Q[0] * vxi[0] * vxi[0] - temp) / Q[1];
  vxi[1] = vxi[1] + G2 * dt4;
  vxi[0] = vxi[0] * exp(-vxi[1] * dt8);
  vxi[0] = vxi[0] + G1 * dt4;
  vxi[0] = vxi[0] * exp(-vxi[1] * dt8);
  s = s * exp(-vxi[0] * dt2);
  kin = s * s * kin;
  xi[0] = xi[0] + vxi[0] * dt2;
  xi[1] = xi[1] + vxi[1] * dt2;
  G1 = (2.0 * kin - L * temp) / Q[0];
  vxi[0] = vxi[0] * exp(-vxi[1] * dt8);
  vxi[0] = vxi[0] + G1 * dt4;
  vxi[0] = vxi[0] * exp(-vxi[1] * dt8);
  G2 = (Q[0] * vxi[0] * vxi[0] - temp) / Q[1];
  vxi[1] = vxi[1] + G2 * dt4;
  return s;
}

int orc_md_run(const mdffgpu_params* p, const int* itype, const double* A,
               const mdffgpu_md_params* mp, int itime0, int nsteps, double* rx, double* ry,
               double* rz, double* vx, double* vy, double* vz, double* fx, double* fy,
               double* fz, double* xi, double* vxi, orc_state* st, mdffgpu_md_step* steps) {
  const int n = p->natm;
  const int integ = mp->integrator;
  if (integ < MDFFGPU_NVE_VV || integ > MDFFGPU_NVT_NHCN) return MDFFGPU_ERR_UNSUPPORTED;
  double w[9];
  if (integ == MDFFGPU_NVT_NHCN &&
      (mp->nhc_n < 2 || mp->nhc_n > MDFFGPU_NHC_MAX || mp->nhc_mults < 1 ||
       orc_yoshida(mp->nhc_yosh_order, w)))
    return MDFFGPU_ERR_ARG;
  double* m = (double*)malloc((size_t)n * sizeof(double));
  double* fs = (double*)malloc(3 * (size_t)n * sizeof(double));
  if (!m || !fs) {
    free(m);
    free(fs);
    return MDFFGPU_ERR_ALLOC;
  }
  for (int i = 0; i < n; ++i) m[i] = mp->mass[itype[i] - 1];
  const double dt = mp->dt, temp = mp->temp;
  /* thermostat masses and degrees of freedom (:218-224, :504-510) */
  const int nhc = (integ == MDFFGPU_NVT_NHC2) ? 2 : mp->nhc_n;
  const double L = (integ == MDFFGPU_NVT_NHC2) ? 3.0 * ((double)n - 1.0) : 3.0 * (double)n;
  double Q[MDFFGPU_NHC_MAX];
  for (int h = 0; h < nhc; ++h) Q[h] = pow(mp->timesca_thermo, 2.0) * temp;
  Q[0] = Q[0] * L;
  int rc = 0;
  for (int k = 0; k < nsteps && rc == 0; ++k) {
    mdffgpu_energies en;
    double kin = 0.0, e_nvt = 0.0;
    if (integ == MDFFGPU_NVE_VV) {
      /* prop_velocity_verlet (src/integration.f90:131-194) */
      double dtsq2 = dt * dt * 0.5, dt2 = dt * 0.5;
      for (int i = 0; i < n; ++i) {
        fs[i] = fx[i];
        fs[i + n] = fy[i];
        fs[i + 2 * n] = fz[i];
      }
      for (int i = 0; i < n; ++i) {
        rx[i] = rx[i] + vx[i] * dt + (fx[i] * dtsq2) / m[i];
        ry[i] = ry[i] + vy[i] * dt + (fy[i] * dtsq2) / m[i];
        rz[i] = rz[i] + vz[i] * dt + (fz[i] * dtsq2) / m[i];
      }
      rc = orc_engforce(p, itype, A, rx, ry, rz, itime0 + k, st, fx, fy, fz, NULL, NULL, NULL,
                        NULL, &en);
      for (int i = 0; i < n; ++i) {
        vx[i] = vx[i] + (fs[i] + fx[i]) * dt2 / m[i];
        vy[i] = vy[i] + (fs[i + n] + fy[i]) * dt2 / m[i];
        vz[i] = vz[i] + (fs[i + 2 * n] + fz[i]) * dt2 / m[i];
      }
      kin = orc_kin(n, m, vx, vy, vz);
    } else {
      /* nhc2 (:205-249) / nhcn (:476-553) */
      double s;
      kin = orc_kin(n, m, vx, vy, vz);
      if (integ == MDFFGPU_NVT_NHC2)
        s = orc_chain_nhc2(kin, vxi, xi, Q, L, temp, dt);
      else
        s = orc_chain_nhcn(kin, vxi, xi, Q, L, temp, dt, nhc, mp->nhc_yosh_order, mp->nhc_mults);
      for (int i = 0; i < n; ++i) {
        vx[i] = s * vx[i];
        vy[i] = s * vy[i];
        vz[i] = s * vz[i];
      }
      /* prop_pos_vel_verlet (:333-385) */
      double dt2 = dt * 0.5;
      for (int i = 0; i < n; ++i) {
        vx[i] = vx[i] + fx[i] * dt2 / m[i];
        vy[i] = vy[i] + fy[i] * dt2 / m[i];
        vz[i] = vz[i] + fz[i] * dt2 / m[i];
        rx[i] = rx[i] + vx[i] * dt;
        ry[i] = ry[i] + vy[i] * dt;
        rz[i] = rz[i] + vz[i] * dt;
      }
      rc = orc_engforce(p, itype, A, rx, ry, rz, itime0 + k, st, fx, fy, fz, NULL, NULL, NULL,
                        NULL, &en);
      for (int i = 0; i < n; ++i) {
        vx[i] = vx[i] + fx[i] * dt2 / m[i];
        vy[i] = vy[i] + fy[i] * dt2 / m[i];
        vz[i] = vz[i] + fz[i] * dt2 / m[i];
      }
      kin = orc_kin(n, m, vx, vy, vz);
      if (integ == MDFFGPU_NVT_NHC2)
        s = orc_chain_nhc2(kin, vxi, xi, Q, L, temp, dt);
      else
        s = orc_chain_nhcn(kin, vxi, xi, Q, L, temp, dt, nhc, mp->nhc_yosh_order, mp->nhc_mults);
      for (int i = 0; i < n; ++i) {
        vx[i] = s * vx[i];
        vy[i] = s * vy[i];
        vz[i] = s * vz[i];
      }
      kin = orc_kin(n, m, vx, vy, vz);
      if (integ == MDFFGPU_NVT_NHC2) {
        e_nvt = 0.0;
        e_nvt = e_nvt + L * temp * xi[0];
        e_nvt = e_nvt + vxi[0] * vxi[0] * 0.5 / Q[0];
        e_nvt = e_nvt + vxi[1] * vxi[1] * 0.5 / Q[1];
        e_nvt = e_nvt + temp * xi[1];
      } else {
        double nvt1 = L * temp * xi[0], nvt2 = vxi[0] * vxi[0] * 0.5 / Q[0], nvt3 = 0.0,
               nvt4 = 0.0;
        for (int h = 1; h < nhc; ++h) {
          nvt3 = nvt3 + vxi[h] * vxi[h] * 0.5 / Q[h];
          nvt4 = nvt4 + temp * xi[h];
        }
        e_nvt = nvt1 + nvt2 + nvt3 + nvt4;
      }
    }
    if (steps) {
      steps[k].en = en;
      steps[k].e_kin = kin;
      steps[k].temp_r = 2.0 * kin / mp->boltz_unit / (3.0 * (double)n);
      steps[k].e_nvt = e_nvt;
    }
  }
  free(m);
  free(fs);
  return rc;
}
