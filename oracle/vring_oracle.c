/* vring_oracle.c -- FP64 CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * A restatement, routine by routine, of the reference Fortran RK3 fractional
 * step on one whole domain (one MPI rank, so every ip_a(*) test takes the
 * "first and last rank" branch).  Each function cites the reference lines it
 * follows.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * may load this code; the product path never does.
 *
 * Expression order follows the Fortran source (left to right) so that the
 * oracle can be compared with the reference modules built from source.
 * Single-precision literals of the reference are reproduced with (float).
 */
#include "vring_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

struct orc {
  int nt, nr, nz, nrc, nzc, NR, NZ;
  double rmax, zmax, Re, dt, slip;
  int wvnex;
  double dtheta, dr, dz, dx1, dx2, dx3, dx1q, dx2q, dx3q, axsym;
  double *rm, *rc, *zm, *zc, *thm, *thc, *qc, *da, *dmr, *dcr;
  /* axial mesh (no reference counterpart: its z mesh is uniform,
   * mesh_tools.F90:219, 321-342): dz/dzeta at centres (dmz) and faces (dcz),
   * k = 0..nz+1; zstr = 0 keeps every z expression of the reference, and
   * the Poisson solve's z direction uses the generalised eigenvectors of
   * the z operator (zvf: V, zvi: V^-1, zlam: eigenvalues) instead of the DCT */
  int zstr;
  double *dmz, *dcz, *zvf, *zvi, *zlam;
  double *b_rm, *b_rc, *b_zm, *b_zc, *b_thm, *b_thc, *b_qc, *b_da, *b_dmr, *b_dcr;
  int *imv, *ipv, *kmv, *kpv, *jmv, *jpv, *jmc, *jpc;
  double gam[4], rom[4], alm[4];
  int irk;
  double ntime, unom0, unom1;
  double *h1d11, *h1d22d, *h1d22n, *h1dare, *h2d11, *h2d22m, *h2d22p, *h2dare;
  double *s1t11, *s1t2n, *s1t2s, *s1t2c, *s2t11, *s2t2p, *s2t2m, *s2t2c, *s3t2n, *s3t2s;
  double *amj1, *acj1, *apj1, *amj2, *acj2, *apj2, *amj3, *acj3, *apj3;
  double *amk1, *ack1, *apk1, *amk2, *ack2, *apk2, *amk3, *ack3, *apk3;
  double *q1, *q2, *q3, *dq1, *dq2, *dq3, *ru1, *ru2, *ru3, *qrhs, *qcap, *pres, *prg;
  double *q1s, *q2s, *q3s, *dq3no;
  double *corr1s, *corr2s, *corr3s, *corr1n, *corr2n, *corr3n, *csort, *tconv;
  double *rhs, *sol, *dph;
  double *aa_base, *bb_base, *cc_base;
  /* tripij scratch ami/aci/api/fi/gam2(0:nrc, 0:nt) (solu_toolsII.F90:183-187) */
  double *ami, *aci, *api, *fi, *gam2;
  /* line scratch */
  double *atri, *btri, *ctri, *rtri, *stri, *beta, *gamm;
};

/* ---- indexing helpers (Fortran column-major, lower bounds as declared) ---- */
#define F3(a, i, j, k) (a)[((i) - 1) + (size_t)o->nt * ((size_t)(j) + (size_t)o->NR * (size_t)(k))]
#define PL(a, i, j) (a)[((i) - 1) + (size_t)o->nt * (size_t)(j)]            /* (nt,0:..) */
#define CR(a, i, j) (a)[((i) - 1) + (size_t)o->nt * (size_t)((j) - 1)]      /* (nt,1:..) */
#define PRG(i, j, k, c) o->prg[((i) - 1) + (size_t)o->nt * (((j) - 1) + (size_t)o->nr * (((k) - 1) + (size_t)o->nz * ((c) - 1)))]
#define RS(a, i, j, k) (a)[((i) - 1) + (size_t)o->nt * (((j) - 1) + (size_t)o->nr * ((k) - 1))]
#define DP(i, j, k) o->dph[((i) - 1) + (size_t)o->nt * ((size_t)(j) + (size_t)(o->nr + 2) * (size_t)(k))]
#define TS(a, j, i) (a)[(j) + (size_t)(o->nrc + 1) * (size_t)(i)]           /* (0:nrc,0:nt) */
/* axial metric: inverse cell width at centre k, inverse centre spacing at
 * face k; second-difference weights at centres (CM: k-1, CP: k+1) and at
 * faces (FM, FP) */
#define ZM3(k) (o->dx3 / o->dmz[k])
#define ZC3(k) (o->dx3 / o->dcz[k])
#define CM(k) (ZM3(k) * ZC3(k))
#define CP(k) (ZM3(k) * ZC3((k) + 1))
#define FM(k) (ZC3(k) * ZM3((k) - 1))
#define FP(k) (ZC3(k) * ZM3(k))

static void coesys_z_mesh(orc_t *o);

static double *dalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

/* ---------------------------------------------------------------------------
 * Mesh and parameters: mesh_tools.F90:217-350, para_tools.F90:103-113.
 * ------------------------------------------------------------------------- */
orc_t *orc_create(int nt, int nr, int nz, double rmax, double zmax, double Re,
                  double dt, double slip, int wvnex) {
  if (nt < 1 || nr < 2 || nz < 2) return NULL;
  orc_t *o = (orc_t *)calloc(1, sizeof(orc_t));
  o->nt = nt; o->nr = nr; o->nz = nz; o->nrc = nr + 1; o->nzc = nz + 1;
  o->NR = nr + 3; o->NZ = nz + 3;
  o->rmax = rmax; o->zmax = zmax; o->Re = Re; o->dt = dt; o->slip = slip;
  o->wvnex = wvnex;
  o->axsym = (nt == 2) ? 0.0 : 1.0;                          /* mesh_tools.F90:98-100 */
  o->dtheta = 2.0 * acos(-1.0) / nt;                          /* :217 */
  o->dr = rmax / nr;                                          /* :218 */
  o->dz = zmax / nz;                                          /* :219 */
  o->dx1 = 1.0 / o->dtheta; o->dx2 = 1.0 / o->dr; o->dx3 = 1.0 / o->dz;   /* :315-320 */
  o->dx1q = o->dx1 * o->dx1; o->dx2q = o->dx2 * o->dx2; o->dx3q = o->dx3 * o->dx3;
  o->b_rm = dalloc(nr + 2); o->rm = o->b_rm;          /* rm(0:nr+1) */
  o->b_rc = dalloc(nr + 2); o->rc = o->b_rc;          /* rc(0:nr+1) */
  o->b_zm = dalloc(nz + 2); o->zm = o->b_zm;          /* zm(0:nz+1) */
  o->b_zc = dalloc(nz + 3); o->zc = o->b_zc + 1;      /* zc(-1:nz+1) */
  o->b_thm = dalloc(nt + 1); o->thm = o->b_thm;       /* 1-based */
  o->b_thc = dalloc(nt + 1); o->thc = o->b_thc;
  o->b_qc = dalloc(nr + 2); o->qc = o->b_qc;          /* qc(0:nr+1) */
  o->b_da = dalloc(nr + 2); o->da = o->b_da;
  o->b_dmr = dalloc(nr + 2); o->dmr = o->b_dmr;
  o->b_dcr = dalloc(nr + 2); o->dcr = o->b_dcr;
  o->dmz = dalloc(nz + 3); o->dcz = dalloc(nz + 3);
  for (int k = 0; k <= nz + 2; k++) { o->dmz[k] = 1.0; o->dcz[k] = 1.0; }
  for (int k = 0; k <= nz + 1; k++) o->zm[k] = 0.0 + (k - 0.5) * o->dz;   /* :321-323 */
  for (int j = 0; j <= nr + 1; j++) { o->dmr[j] = 1.0; o->dcr[j] = 1.0; } /* :325-330 */
  for (int j = 0; j <= nr + 1; j++) o->rm[j] = 0.0 + (j - 0.5) * o->dr;   /* :332-334 */
  for (int i = 1; i <= nt; i++) o->thm[i] = 0.0 + (i - 0.5) * o->dtheta; /* :336-338 */
  for (int k = 0; k <= nz + 1; k++) o->zc[k] = 0.0 + (k - 1.0) * o->dz;   /* :340-342 */
  for (int j = 0; j <= nr + 1; j++) o->rc[j] = 0.0 + (j - 1.0) * o->dr;   /* :344-346 */
  for (int i = 1; i <= nt; i++) o->thc[i] = 0.0 + (i - 1.0) * o->dtheta; /* :348-350 */
  /* index maps, :260-300 (one rank in r and z) */
  o->imv = (int *)calloc(nt + 1, sizeof(int)); o->ipv = (int *)calloc(nt + 1, sizeof(int));
  o->kmv = (int *)calloc(nz + 1, sizeof(int)); o->kpv = (int *)calloc(nz + 1, sizeof(int));
  o->jmv = (int *)calloc(nr + 1, sizeof(int)); o->jpv = (int *)calloc(nr + 1, sizeof(int));
  o->jmc = (int *)calloc(nr + 1, sizeof(int)); o->jpc = (int *)calloc(nr + 1, sizeof(int));
  for (int i = 2; i <= nt; i++) o->imv[i] = i - 1;
  for (int i = 1; i <= nt - 1; i++) o->ipv[i] = i + 1;
  o->imv[1] = nt; o->ipv[nt] = 1;
  for (int k = 1; k <= nz; k++) { o->kmv[k] = k - 1; o->kpv[k] = k + 1; }
  o->kmv[1] = 1; o->kmv[nz] = nz;
  for (int j = 1; j <= nr; j++) { o->jmv[j] = j - 1; o->jpv[j] = j + 1; }
  o->jmv[1] = 1; o->jpv[nr] = nr;
  for (int j = 1; j <= nr; j++) { o->jpc[j] = o->jpv[j] - j; o->jmc[j] = j - o->jmv[j]; }
  /* RK3 coefficients, para_tools.F90:103-113 */
  o->gam[1] = 8.0 / 15.0; o->gam[2] = 5.0 / 12.0; o->gam[3] = 3.0 / 4.0;
  o->rom[1] = 0.0; o->rom[2] = -17.0 / 60.0; o->rom[3] = -5.0 / 12.0;
  for (int i = 1; i <= 3; i++) o->alm[i] = o->gam[i] + o->rom[i];
  o->irk = 1;
  /* ucoeff / coesys vectors */
  o->h1d11 = dalloc(nr + 2); o->h1d22d = dalloc(nr + 2); o->h1d22n = dalloc(nr + 2); o->h1dare = dalloc(nr + 2);
  o->h2d11 = dalloc(nr + 2); o->h2d22m = dalloc(nr + 2); o->h2d22p = dalloc(nr + 2); o->h2dare = dalloc(nr + 2);
  o->s1t11 = dalloc(nr + 2); o->s1t2n = dalloc(nr + 2); o->s1t2s = dalloc(nr + 2); o->s1t2c = dalloc(nr + 2);
  o->s2t11 = dalloc(nr + 2); o->s2t2p = dalloc(nr + 2); o->s2t2m = dalloc(nr + 2); o->s2t2c = dalloc(nr + 2);
  o->s3t2n = dalloc(nr + 2); o->s3t2s = dalloc(nr + 2);
  o->amj1 = dalloc(nr + 3); o->acj1 = dalloc(nr + 3); o->apj1 = dalloc(nr + 3);
  o->amj2 = dalloc(nr + 3); o->acj2 = dalloc(nr + 3); o->apj2 = dalloc(nr + 3);
  o->amj3 = dalloc(nr + 3); o->acj3 = dalloc(nr + 3); o->apj3 = dalloc(nr + 3);
  o->amk1 = dalloc(nz + 3); o->ack1 = dalloc(nz + 3); o->apk1 = dalloc(nz + 3);
  o->amk2 = dalloc(nz + 3); o->ack2 = dalloc(nz + 3); o->apk2 = dalloc(nz + 3);
  o->amk3 = dalloc(nz + 3); o->ack3 = dalloc(nz + 3); o->apk3 = dalloc(nz + 3);
  size_t n3 = (size_t)nt * o->NR * o->NZ;
  o->q1 = dalloc(n3); o->q2 = dalloc(n3); o->q3 = dalloc(n3);
  o->dq1 = dalloc(n3); o->dq2 = dalloc(n3); o->dq3 = dalloc(n3);
  o->ru1 = dalloc(n3); o->ru2 = dalloc(n3); o->ru3 = dalloc(n3);
  o->qrhs = dalloc(n3); o->qcap = dalloc(n3); o->pres = dalloc(n3);
  o->prg = dalloc((size_t)nt * nr * nz * 3);
  o->q1s = dalloc((size_t)nt * (nr + 2)); o->q2s = dalloc((size_t)nt * (nr + 2));
  o->q3s = dalloc((size_t)nt * (nr + 2)); o->dq3no = dalloc((size_t)nt * (nr + 3));
  size_t np = (size_t)nt * (nr + 1);
  o->corr1s = dalloc(np); o->corr2s = dalloc(np); o->corr3s = dalloc(np);
  o->corr1n = dalloc(np); o->corr2n = dalloc(np); o->corr3n = dalloc(np);
  o->csort = dalloc(np); o->tconv = dalloc(np);
  o->rhs = dalloc((size_t)nt * nr * nz); o->sol = dalloc((size_t)nt * nr * nz);
  o->dph = dalloc((size_t)nt * (nr + 2) * (nz + 2));
  /* init_tripar_solvers, poisson.F90:377-386 */
  o->aa_base = dalloc(nr + 1); o->bb_base = dalloc(nr + 1); o->cc_base = dalloc(nr + 1);
  for (int j = 1; j <= nr; j++) {
    o->bb_base[j] = 2.0 / (o->dr * o->dr);
    o->aa_base[j] = -1.0 / (o->dr * o->dr) + 1.0 / (2.0 * o->rm[j] * o->dr);
    o->cc_base[j] = -1.0 / (o->dr * o->dr) - 1.0 / (2.0 * o->rm[j] * o->dr);
  }
  size_t ns = (size_t)(o->nrc + 1) * (nt + 1);
  o->ami = dalloc(ns); o->aci = dalloc(ns); o->api = dalloc(ns); o->fi = dalloc(ns); o->gam2 = dalloc(ns);
  int nl = (nr > nz ? nr : nz) + 4;
  o->atri = dalloc(nl); o->btri = dalloc(nl); o->ctri = dalloc(nl); o->rtri = dalloc(nl);
  o->stri = dalloc(nl); o->beta = dalloc(nl); o->gamm = dalloc(nl);
  return o;
}

void orc_destroy(orc_t *o) {
  if (!o) return;
  double *ptrs[] = {o->b_rm, o->b_rc, o->b_zm, o->b_zc, o->b_thm, o->b_thc, o->b_qc, o->b_da,
                    o->b_dmr, o->b_dcr, o->h1d11, o->h1d22d, o->h1d22n, o->h1dare, o->h2d11,
                    o->h2d22m, o->h2d22p, o->h2dare, o->s1t11, o->s1t2n, o->s1t2s, o->s1t2c,
                    o->s2t11, o->s2t2p, o->s2t2m, o->s2t2c, o->s3t2n, o->s3t2s, o->amj1, o->acj1,
                    o->apj1, o->amj2, o->acj2, o->apj2, o->amj3, o->acj3, o->apj3, o->amk1,
                    o->ack1, o->apk1, o->amk2, o->ack2, o->apk2, o->amk3, o->ack3, o->apk3,
                    o->q1, o->q2, o->q3, o->dq1, o->dq2, o->dq3, o->ru1, o->ru2, o->ru3, o->qrhs,
                    o->qcap, o->pres, o->prg, o->q1s, o->q2s, o->q3s, o->dq3no, o->corr1s,
                    o->corr2s, o->corr3s, o->corr1n, o->corr2n, o->corr3n, o->csort, o->tconv,
                    o->rhs, o->sol, o->dph, o->aa_base, o->bb_base, o->cc_base, o->ami, o->aci,
                    o->api, o->fi, o->gam2, o->atri, o->btri, o->ctri, o->rtri, o->stri, o->beta,
                    o->gamm, o->dmz, o->dcz, o->zvf, o->zvi, o->zlam};
  for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
  free(o->imv); free(o->ipv); free(o->kmv); free(o->kpv); free(o->jmv); free(o->jpv);
  free(o->jmc); free(o->jpc);
  free(o);
}

double *orc_field(orc_t *o, const char *name, long *count) {
  long n3 = (long)o->nt * o->NR * o->NZ, npl = (long)o->nt * (o->nr + 1);
  long n2 = (long)o->nt * (o->nr + 2);
  struct { const char *n; double *p; long c; } t[] = {
      {"q1", o->q1, n3}, {"q2", o->q2, n3}, {"q3", o->q3, n3}, {"dq1", o->dq1, n3},
      {"dq2", o->dq2, n3}, {"dq3", o->dq3, n3}, {"ru1", o->ru1, n3}, {"ru2", o->ru2, n3},
      {"ru3", o->ru3, n3}, {"qrhs", o->qrhs, n3}, {"qcap", o->qcap, n3}, {"pres", o->pres, n3},
      {"prg", o->prg, (long)o->nt * o->nr * o->nz * 3}, {"q1s", o->q1s, n2}, {"q2s", o->q2s, n2},
      {"q3s", o->q3s, n2}, {"dq3no", o->dq3no, (long)o->nt * (o->nr + 3)},
      {"corr1s", o->corr1s, npl}, {"corr2s", o->corr2s, npl}, {"corr3s", o->corr3s, npl},
      {"corr1n", o->corr1n, npl}, {"corr2n", o->corr2n, npl}, {"corr3n", o->corr3n, npl},
      {"csort", o->csort, npl}, {"tconv", o->tconv, npl},
      {"rhs", o->rhs, (long)o->nt * o->nr * o->nz}, {"sol", o->sol, (long)o->nt * o->nr * o->nz},
      {"dph", o->dph, (long)o->nt * (o->nr + 2) * (o->nz + 2)},
      {"rm", o->rm, o->nr + 2}, {"rc", o->rc, o->nr + 2}, {"dmr", o->dmr, o->nr + 2},
      {"dcr", o->dcr, o->nr + 2}, {"zm", o->zm, o->nz + 2},
      {"zc", o->zc, o->nz + 2}, {"amj1", o->amj1, o->nr + 3}, {"acj1", o->acj1, o->nr + 3},
      {"apj1", o->apj1, o->nr + 3}, {"amj2", o->amj2, o->nr + 3}, {"acj2", o->acj2, o->nr + 3},
      {"apj2", o->apj2, o->nr + 3}, {"amj3", o->amj3, o->nr + 3}, {"acj3", o->acj3, o->nr + 3},
      {"apj3", o->apj3, o->nr + 3}, {"amk1", o->amk1, o->nz + 3}, {"ack1", o->ack1, o->nz + 3},
      {"apk1", o->apk1, o->nz + 3}, {"amk2", o->amk2, o->nz + 3}, {"ack2", o->ack2, o->nz + 3},
      {"apk2", o->apk2, o->nz + 3}, {"amk3", o->amk3, o->nz + 3}, {"ack3", o->ack3, o->nz + 3},
      {"apk3", o->apk3, o->nz + 3}, {"s1t11", o->s1t11, o->nr + 2}, {"s2t11", o->s2t11, o->nr + 2},
  };
  for (size_t i = 0; i < sizeof(t) / sizeof(t[0]); i++)
    if (strcmp(t[i].n, name) == 0) { if (count) *count = t[i].c; return t[i].p; }
  if (count) *count = 0;
  return NULL;
}

void orc_set_time(orc_t *o, double ntime) { o->ntime = ntime; }

/* Radial mesh through the reference's metric hooks: rm, rc, dmr, dcr over
 * j = 0..nr+1 replace the uniform mesh of mesh_tools.F90:325-346; every
 * routine here reads them where the reference does (ucoeff, coesys, hdnbc's
 * da, divq, upsol, uppres).  The Poisson radial rows, which
 * poisson.F90:379-381 writes for the uniform mesh only, become minus divq's
 * radial difference (divergence.F90:39) of uppres' radial gradient
 * (divergence.F90:132): the operator that leaves the projected field
 * solenoidal on this mesh, equal to the reference's rows on the uniform
 * mesh.  Pinned by the reference's own modules with the same hooks (every
 * phase but the Poisson solve) and by that solenoidal property. */
void orc_set_radial_mesh(orc_t *o, const double *rm, const double *rc, const double *dmr, const double *dcr) {
  for (int j = 0; j <= o->nr + 1; j++) {
    o->rm[j] = rm[j]; o->rc[j] = rc[j]; o->dmr[j] = dmr[j]; o->dcr[j] = dcr[j];
  }
  for (int j = 1; j <= o->nr; j++) {
    double w = o->dx2q / (o->rm[j] * o->dmr[j]);
    o->aa_base[j] = -w * o->rc[j] / o->dcr[j];
    o->cc_base[j] = -w * o->rc[j + 1] / o->dcr[j + 1];
    o->bb_base[j] = -(o->aa_base[j] + o->cc_base[j]);
  }
}
/* Eigen-decomposition of a dense symmetric n x n matrix A (row-major, destroyed)
 * by cyclic Jacobi rotations: on return A holds the eigenvalues on its
 * diagonal and W (row-major, W[k*n+q] = component k of eigenvector q) the
 * orthonormal eigenvectors. */
static void jacobi_eig(int n, double *A, double *W) {
  for (int i = 0; i < n * n; i++) W[i] = 0.0;
  for (int i = 0; i < n; i++) W[i * n + i] = 1.0;
  for (int sweep = 0; sweep < 60; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        tot += A[i * n + j] * A[i * n + j];
        if (i != j) off += A[i * n + j] * A[i * n + j];
      }
    if (off <= 1e-30 * tot) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        const double apq = A[p * n + q];
        if (apq == 0.0) continue;
        /* rotation angle zeroing A[p][q]: tan(2 phi) = 2 apq / (aqq - app) */
        const double th = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
        const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
        for (int k = 0; k < n; k++) {          /* columns p, q */
          const double akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - sn * akq;
          A[k * n + q] = sn * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {          /* rows p, q */
          const double apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - sn * aqk;
          A[q * n + k] = sn * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          const double wkp = W[k * n + p], wkq = W[k * n + q];
          W[k * n + p] = c * wkp - sn * wkq;
          W[k * n + q] = sn * wkp + c * wkq;
        }
      }
  }
}

/* Non-uniform axial mesh (no reference counterpart; the radial hooks of
 * mesh_tools.F90:325-346 carried over to z): cell centres zm, faces zc
 * (zc(1) = 0, the inlet), and dz/dzeta at centres (dmz) and faces (dcz),
 * k = 0..nz+1, zeta uniform with spacing zmax/nz (derivatives keep
 * dx3 = nz/zmax).  Every z difference then carries the metric of where it
 * sits; the pressure solve's z direction is diagonalised by the
 * generalised eigenvectors of minus its z operator D_z G_z (Neumann ends),
 * K v = lam M v with M the cell widths, which on the uniform mesh are the
 * DCT's cosines with lam = (2/dz^2)(1 - cos(pi k/nz)). */
void orc_set_axial_mesh(orc_t *o, const double *zm, const double *zc, const double *dmz, const double *dcz) {
  const int nz = o->nz;
  for (int k = 0; k <= nz + 1; k++) { o->zm[k] = zm[k]; o->zc[k] = zc[k]; o->dmz[k] = dmz[k]; o->dcz[k] = dcz[k]; }
  o->dmz[nz + 2] = o->dmz[nz + 1];
  o->dcz[nz + 2] = o->dcz[nz + 1];
  o->zstr = 1;
  free(o->zvf); free(o->zvi); free(o->zlam);
  o->zvf = dalloc((size_t)nz * nz); o->zvi = dalloc((size_t)nz * nz); o->zlam = dalloc(nz);
  double *A = dalloc((size_t)nz * nz), *W = dalloc((size_t)nz * nz);
  for (int k = 1; k <= nz; k++) {          /* A = D K D, D = diag(sqrt(ZM3)) */
    double lo = (k > 1) ? ZC3(k) : 0.0, hi = (k < nz) ? ZC3(k + 1) : 0.0;
    A[(k - 1) * nz + (k - 1)] = ZM3(k) * (lo + hi);
    if (k < nz) {
      double e = -sqrt(ZM3(k) * ZM3(k + 1)) * ZC3(k + 1);
      A[(k - 1) * nz + k] = e;
      A[k * nz + (k - 1)] = e;
    }
  }
  jacobi_eig(nz, A, W);
  /* eigenvalues ascending (selection order); the smallest is the constant mode, exactly 0 */
  int *ord = (int *)calloc(nz, sizeof(int));
  for (int q = 0; q < nz; q++) ord[q] = q;
  for (int a = 0; a < nz; a++)
    for (int b = a + 1; b < nz; b++)
      if (A[ord[b] * nz + ord[b]] < A[ord[a] * nz + ord[a]]) { int t = ord[a]; ord[a] = ord[b]; ord[b] = t; }
  for (int q = 0; q < nz; q++) {
    const int c = ord[q];
    o->zlam[q] = (q == 0) ? 0.0 : A[c * nz + c];
    for (int k = 1; k <= nz; k++) {
      o->zvf[(k - 1) * nz + q] = sqrt(ZM3(k)) * W[(k - 1) * nz + c];   /* V = D W */
      o->zvi[q * nz + (k - 1)] = W[(k - 1) * nz + c] / sqrt(ZM3(k));   /* V^-1 = W^T D^-1 */
    }
  }
  free(ord); free(A); free(W);
}

double orc_get_time(const orc_t *o) { return o->ntime; }

/* ---------------------------------------------------------------------------
 * ucoeff, momentum_coef.F90:24-69
 * ------------------------------------------------------------------------- */
static void ucoeff(orc_t *o) {
  const double *rm = o->rm, *rc = o->rc, *dmr = o->dmr, *dcr = o->dcr;
  for (int jc = 1; jc <= o->nr; jc++) {
    o->h1d11[jc] = 0.25 * o->dx1 / (rm[jc] * o->axsym);
    o->h1d22d[jc] = 0.25 * o->dx2 / (rm[jc] * dmr[jc]);
    o->h1d22n[jc] = 0.125 / (rm[jc] * rm[jc]);
    o->h1dare[jc] = o->dx1 / (o->Re * (rm[jc] * rm[jc] * rm[jc]));
    o->s1t11[jc] = o->dx1q / ((rm[jc] * rm[jc]) * o->axsym);
    o->s1t2n[jc] = (rc[jc + 1] * rc[jc + 1] * rc[jc + 1]) / dcr[jc + 1];
    o->s1t2s[jc] = (rc[jc] * rc[jc] * rc[jc]) / dcr[jc];
    o->s1t2c[jc] = o->dx2q / (dmr[jc] * (rm[jc] * rm[jc]));
    o->s3t2n[jc] = o->dx2q * rc[jc + 1] / (rm[jc] * dcr[jc + 1] * dmr[jc]);
    o->s3t2s[jc] = o->dx2q * rc[jc] / (rm[jc] * dcr[jc] * dmr[jc]);
  }
  for (int jc = 2; jc <= o->nr; jc++) {   /* ji=2 on the axis rank */
    o->h2d11[jc] = 0.25 * o->dx1 / rc[jc] * o->axsym;
    o->h2d22m[jc] = 0.25 * o->dx2 / (rm[jc - 1] * dcr[jc]);
    o->h2d22p[jc] = 0.25 * o->dx2 / (rm[jc] * dcr[jc]);
    o->h2dare[jc] = o->dx1 / (o->Re * rc[jc] * o->axsym);
    o->s2t11[jc] = o->axsym * o->dx1q / (rc[jc] * rc[jc]);
    o->s2t2p[jc] = o->dx2q * rc[jc] / rm[jc] / (dcr[jc] * dmr[jc]);
    o->s2t2m[jc] = o->dx2q * rc[jc] / rm[jc - 1] / (dcr[jc] * dmr[jc - 1]);
    o->s2t2c[jc] = o->s2t2p[jc] + o->s2t2m[jc];
  }
}

/* ---------------------------------------------------------------------------
 * coesys, momentum_coef.F90:71-465 (the unused scalar rows are omitted)
 * ------------------------------------------------------------------------- */
static void coesys(orc_t *o) {
  const int nr = o->nr, nrc = o->nrc, nz = o->nz, nzc = o->nzc;
  const double *rm = o->rm, *rc = o->rc, *dcr = o->dcr;
  double betadx = 0.5 * o->alm[o->irk] * o->dt / o->Re;               /* :93 */
  double a22p, a22m;
  /* q1, r direction: axis row (:104-107), wall row (:124-127), interior (:140-146) */
  a22p = betadx * o->s1t2c[1] * o->s1t2n[1];
  o->amj1[1] = 0.0;
  o->acj1[1] = 1.0 + a22p / rm[1];
  o->apj1[1] = -a22p / rm[2];
  a22m = betadx * o->s1t2c[nr] * o->s1t2s[nr];
  o->apj1[nr] = 0.0;
  o->acj1[nr] = 1.0 + a22m / rm[nr] -
                betadx * o->s1t2c[nr] * rc[nrc] * rc[nrc] *
                    (rc[nrc] / rm[nr] * o->slip - rc[nrc] / rm[nrc]) / dcr[nrc];
  o->amj1[nr] = -a22m / rm[nr - 1];
  for (int jc = 2; jc <= nr - 1; jc++) {
    a22p = betadx * o->s1t2c[jc] * o->s1t2n[jc];
    a22m = betadx * o->s1t2c[jc] * o->s1t2s[jc];
    o->apj1[jc] = -a22p / rm[jc + 1];
    o->acj1[jc] = 1.0 + (a22p + a22m) / rm[jc];
    o->amj1[jc] = -a22m / rm[jc - 1];
  }
  /* q2, r direction (:185-208); the loop at :204 overwrites row 1 */
  o->acj2[1] = 1.0; o->amj2[1] = 0.0; o->apj2[1] = 0.0;
  o->acj2[nrc] = 1.0; o->amj2[nrc] = 0.0; o->apj2[nrc] = 0.0;
  for (int jc = 1; jc <= nr; jc++) {
    o->apj2[jc] = -(betadx * o->s2t2p[jc]);
    o->amj2[jc] = -(betadx * o->s2t2m[jc]);
    o->acj2[jc] = 1.0 + betadx * o->s2t2c[jc];
  }
  /* q3, r direction (:251-261) */
  for (int jc = 1; jc <= nr; jc++) {
    o->apj3[jc] = -(betadx * o->s3t2n[jc]);
    o->amj3[jc] = -(betadx * o->s3t2s[jc] * (double)(float)o->jmc[jc]);
    o->acj3[jc] = 1.0 - (o->apj3[jc] + o->amj3[jc]);
  }
  o->apj3[nr] = 0.0;
  o->amj3[nr] = -(betadx * o->s3t2s[nr]);
  o->acj3[nr] = 1.0 - (betadx * o->s3t2n[nr] * (o->slip - 1.0) + o->amj3[nr]);
  /* z direction, :314-465 (if_grad_null = 0) */
  if (o->zstr) { coesys_z_mesh(o); return; }
  betadx = 0.5 * o->dx3q * o->alm[o->irk] * o->dt / o->Re;
  for (int kc = 1; kc <= nz; kc++) {
    o->amk1[kc] = -betadx; o->ack1[kc] = 1.0 + 2.0 * betadx; o->apk1[kc] = -betadx;
  }
  o->amk1[nzc] = 1.0; o->ack1[nzc] = 1.0; o->apk1[nzc] = 0.0;
  o->amk1[1] = -(8.0 / 3.0 * betadx); o->ack1[1] = 1.0 + 4.0 * betadx; o->apk1[1] = -(4.0 / 3.0 * betadx);
  for (int kc = 1; kc <= nz; kc++) {
    o->amk2[kc] = -betadx; o->ack2[kc] = 1.0 + 2.0 * betadx; o->apk2[kc] = -betadx;
  }
  o->amk2[nzc] = 1.0; o->ack2[nzc] = 1.0; o->apk2[nzc] = 0.0;
  o->amk2[1] = -(8.0 / 3.0 * betadx); o->ack2[1] = 1.0 + 4.0 * betadx; o->apk2[1] = -(4.0 / 3.0 * betadx);
  for (int kc = 1; kc <= nz; kc++) {
    o->amk3[kc] = -betadx; o->ack3[kc] = 1.0 + 2.0 * betadx; o->apk3[kc] = -betadx;
  }
  o->amk3[nzc] = 0.0; o->ack3[nzc] = 1.0; o->apk3[nzc] = 0.0;
  o->amk3[1] = 0.0; o->ack3[1] = 1.0; o->apk3[1] = 0.0;
}

/* z sweeps on the axial mesh: centre rows (q1, q2) and face rows (q3) from
 * the second-difference weights; the inlet row of q1, q2 is the one-sided
 * second difference through the wall value at z = 0 at distance h1 (half
 * the first cell) and the next centre at h2,
 * f'' = ws fs - (ws + w2) f1 + w2 f2, ws = 2/(h1 (h1+h2)), w2 = 2/(h2 (h1+h2)),
 * which is momentum_coef.F90's (-8/3, 4, -4/3) beta dx3q row on the
 * uniform mesh; the end rows keep the reference's forms. */
static void coesys_z_mesh(orc_t *o) {
  const int nz = o->nz, nzc = o->nzc;
  const double bz = 0.5 * o->alm[o->irk] * o->dt / o->Re;
  const double h1 = 0.5 / ZM3(1), h2 = 1.0 / ZC3(2);
  const double ws = 2.0 / (h1 * (h1 + h2)), w2 = 2.0 / (h2 * (h1 + h2));
  for (int kc = 1; kc <= nz; kc++) {
    double cm = CM(kc), cp = CP(kc), fm = FM(kc), fp = FP(kc);
    o->amk1[kc] = -(bz * cm); o->ack1[kc] = 1.0 + bz * (cm + cp); o->apk1[kc] = -(bz * cp);
    o->amk2[kc] = o->amk1[kc]; o->ack2[kc] = o->ack1[kc]; o->apk2[kc] = o->apk1[kc];
    o->amk3[kc] = -(bz * fm); o->ack3[kc] = 1.0 + bz * (fm + fp); o->apk3[kc] = -(bz * fp);
  }
  o->amk1[nzc] = 1.0; o->ack1[nzc] = 1.0; o->apk1[nzc] = 0.0;
  o->amk1[1] = -(bz * ws); o->ack1[1] = 1.0 + bz * (ws + w2); o->apk1[1] = -(bz * w2);
  o->amk2[nzc] = 1.0; o->ack2[nzc] = 1.0; o->apk2[nzc] = 0.0;
  o->amk2[1] = o->amk1[1]; o->ack2[1] = o->ack1[1]; o->apk2[1] = o->apk1[1];
  o->amk3[nzc] = 0.0; o->ack3[nzc] = 1.0; o->apk3[nzc] = 0.0;
  o->amk3[1] = 0.0; o->ack3[1] = 1.0; o->apk3[1] = 0.0;
}

/* ---------------------------------------------------------------------------
 * Inlet time profile, gen_tools_swirl.F90:159-175 (0.1 is a default real).
 * profilet/profiler (131-157) always return 0.
 * ------------------------------------------------------------------------- */
static double profile(double t) {
  const double n = 2.0, t1 = (double)0.1f * n;
  double u = 0.0;  /* undefined in the reference for t<=0 and at the breakpoints */
  double tt = t;
  if (tt > 0.0 && tt < t1) { double x = tt / t1; u = 3.0 * (x * x) - 2.0 * (x * x * x); }
  if (tt > t1 && tt < n) u = 1.0;
  if (tt > n && tt < n + t1) { double x = (n + t1 - tt) / t1; u = 3.0 * (x * x) - 2.0 * (x * x * x); }
  if (tt > n + t1) u = 0.0;
  return u;
}

/* ---------------------------------------------------------------------------
 * hdnbc, momentum_terms_swirl.F90:177-296 (single domain: the three
 * allreduces are identities and the outlet sum runs over this rank only)
 * ------------------------------------------------------------------------- */
static void hdnbc(orc_t *o) {
  const int nt = o->nt, nr = o->nr, nrc = o->nrc, nz = o->nz, nzc = o->nzc;
  const int irk = o->irk;
  double dt = o->dt;
  double t0 = o->ntime, t1 = o->ntime + dt;                           /* :192-193 */
  o->unom0 = profile(t0);
  o->unom1 = profile(t1);
  const double lim = (double)0.4999f;
  for (int jc = 1; jc <= nrc; jc++)
    for (int ic = 1; ic <= nt; ic++) CR(o->corr3s, ic, jc) = 0.0;   /* :204 */
  for (int ic = 1; ic <= nt; ic++)
    for (int jc = 1; jc <= nr; jc++)
      if (o->rc[jc] <= lim) {                                         /* :210-212 */
        CR(o->corr1s, ic, jc) = 0.0 * (o->gam[irk] + o->rom[irk]);
        CR(o->corr2s, ic, jc) = 0.0 * (o->gam[irk] + o->rom[irk]);
        CR(o->corr3s, ic, jc) = (o->unom1 - o->unom0) * (o->gam[irk] + o->rom[irk]);
      }
  for (int ic = 1; ic <= nt; ic++)
    for (int jc = 1; jc <= nr; jc++) {                                /* :216-223 */
      PL(o->q1s, ic, jc) = 0.0; PL(o->q2s, ic, jc) = 0.0; PL(o->q3s, ic, jc) = 0.0;
    }
  for (int ic = 1; ic <= nt; ic++)
    for (int jc = 1; jc <= nr; jc++)                                  /* :225-231 */
      if (F3(o->q3, ic, jc, nzc) > 0.0) CR(o->csort, ic, jc) = (double)0.6f * F3(o->q3, ic, jc, nzc);
  for (int ic = 1; ic <= nt; ic++)
    for (int jc = 1; jc <= nrc; jc++) {                               /* :234-241 */
      CR(o->tconv, ic, jc) = -(CR(o->csort, ic, jc) * (o->zstr ? ZM3(nz) : o->dx3)) *
                             (F3(o->q3, ic, jc, nzc) - F3(o->q3, ic, jc, nz));
      CR(o->corr3n, ic, jc) = (o->gam[irk] * CR(o->tconv, ic, jc) + o->rom[irk] * PL(o->dq3no, ic, jc)) * o->dt;
      PL(o->dq3no, ic, jc) = CR(o->tconv, ic, jc);
    }
  for (int jc = 1; jc <= nr; jc++) {                                  /* :247-251 */
    o->da[jc] = o->rm[jc] * o->dmr[jc] / (o->dx1 * o->dx2);
    double x = o->rm[jc] / o->rmax;
    o->qc[jc] = 1.0 - x * x;
  }
  double red = 0.0, qinf = 0.0;
  for (int jc = 1; jc <= nr; jc++)                                    /* :261-267 */
    for (int ic = 1; ic <= nt; ic++) {
      red = red + o->qc[jc] * o->qc[jc] * o->da[jc];
      qinf = qinf + CR(o->corr3s, ic, jc) * o->da[jc];
    }
  double cor = (qinf - 0.0) / (red + (double)1.e-15f);               /* :283 */
  for (int jc = 1; jc <= nrc; jc++)                                   /* :284-291 */
    for (int ic = 1; ic <= nt; ic++) CR(o->corr3n, ic, jc) = CR(o->corr3n, ic, jc) + o->qc[jc] * cor;
}

/* ---------------------------------------------------------------------------
 * hdnl1/hdnl2/hdnl3, momentum_terms_swirl.F90:15-174
 * ------------------------------------------------------------------------- */
static void hdnl1(orc_t *o) {
  double *q1 = o->q1, *q2 = o->q2, *q3 = o->q3;
  for (int ic = 1; ic <= o->nt; ic++) {
    int ip = o->ipv[ic], im = o->imv[ic];
    for (int kc = 1; kc <= o->nz; kc++)
      for (int jc = 1; jc <= o->nr; jc++) {
        double qnn, qss, h11, h12d, h12n, h13, d1q2;
        qnn = F3(q1, ip, jc, kc) + F3(q1, ic, jc, kc);
        qss = F3(q1, im, jc, kc) + F3(q1, ic, jc, kc);
        h11 = (qnn * qnn - qss * qss) * o->h1d11[jc];
        qnn = (F3(q2, ic, jc + 1, kc) + F3(q2, im, jc + 1, kc)) * (F3(q1, ic, jc + 1, kc) + F3(q1, ic, jc, kc));
        qss = (F3(q1, ic, jc, kc) + F3(q1, ic, o->jmv[jc], kc)) * (F3(q2, ic, jc, kc) + F3(q2, im, jc, kc));
        h12d = (qnn - qss) * o->h1d22d[jc];
        qnn = (F3(q1, ip, jc, kc) + F3(q1, ic, jc, kc)) * (F3(q2, ic, jc + 1, kc) + F3(q2, ic, jc, kc));
        qss = (F3(q1, ic, jc, kc) + F3(q1, im, jc, kc)) * (F3(q2, im, jc + 1, kc) + F3(q2, im, jc, kc));
        h12n = (qnn + qss) * o->h1d22n[jc];
        qnn = 0.25 * (F3(q3, ic, jc, kc + 1) + F3(q3, im, jc, kc + 1)) * (F3(q1, ic, jc, kc + 1) + F3(q1, ic, jc, kc));
        if (kc == 1)
          qss = 0.5 * (PL(o->q3s, ic, jc) + PL(o->q3s, im, jc)) * PL(o->q1s, ic, jc);
        else
          qss = 0.25 * (F3(q3, ic, jc, kc) + F3(q3, im, jc, kc)) * (F3(q1, ic, jc, kc) + F3(q1, ic, jc, kc - 1));
        h13 = (qnn - qss) * (o->zstr ? ZM3(kc) : o->dx3);
        qnn = F3(q2, ic, jc + 1, kc) + F3(q2, ic, jc, kc);
        qss = F3(q2, im, jc + 1, kc) + F3(q2, im, jc, kc);
        d1q2 = (qnn - qss) * o->h1dare[jc];
        F3(o->dq1, ic, jc, kc) = d1q2 - (h11 + (h12d + h12n) + h13);
      }
  }
}

static void hdnl2(orc_t *o) {
  double *q1 = o->q1, *q2 = o->q2, *q3 = o->q3;
  for (int ic = 1; ic <= o->nt; ic++) {
    int ip = o->ipv[ic], im = o->imv[ic];
    for (int kc = 1; kc <= o->nz; kc++)
      for (int jc = 2; jc <= o->nr; jc++) {
        double qnn, qss, h21, h22, h23, h11n, d1q1;
        qnn = (F3(q1, ip, jc, kc) + F3(q1, ip, jc - 1, kc)) * (F3(q2, ip, jc, kc) + F3(q2, ic, jc, kc));
        qss = (F3(q1, ic, jc, kc) + F3(q1, ic, jc - 1, kc)) * (F3(q2, ic, jc, kc) + F3(q2, im, jc, kc));
        h21 = (qnn - qss) * o->h2d11[jc];
        qnn = F3(q2, ic, jc + 1, kc) + F3(q2, ic, jc, kc);
        qss = F3(q2, ic, jc, kc) + F3(q2, ic, jc - 1, kc);
        h22 = qnn * qnn * o->h2d22p[jc] - qss * qss * o->h2d22m[jc];
        qnn = 0.5 * (F3(q1, ic, jc, kc) + F3(q1, ip, jc, kc));
        qss = 0.5 * (F3(q1, ic, jc - 1, kc) + F3(q1, ip, jc - 1, kc));
        h11n = 0.25 * (qnn + qss) * (qnn + qss);
        qnn = 0.25 * (F3(q3, ic, jc, kc + 1) + F3(q3, ic, jc - 1, kc + 1)) * (F3(q2, ic, jc, kc + 1) + F3(q2, ic, jc, kc));
        if (kc == 1)
          qss = 0.5 * (PL(o->q3s, ic, jc) + PL(o->q3s, ic, jc - 1)) * PL(o->q2s, ic, jc);
        else
          qss = 0.25 * (F3(q3, ic, jc, kc) + F3(q3, ic, jc - 1, kc)) * (F3(q2, ic, jc, kc) + F3(q2, ic, jc, kc - 1));
        h23 = (qnn - qss) * (o->zstr ? ZM3(kc) : o->dx3);
        qnn = F3(q1, ip, jc, kc) + F3(q1, ip, jc - 1, kc);
        qss = F3(q1, ic, jc, kc) + F3(q1, ic, jc - 1, kc);
        d1q1 = -((qnn - qss) * o->h2dare[jc]);
        F3(o->dq2, ic, jc, kc) = d1q1 - (h21 + h22 + h23 - h11n);
      }
  }
}

static void hdnl3(orc_t *o) {
  double *q1 = o->q1, *q2 = o->q2, *q3 = o->q3;
  for (int ic = 1; ic <= o->nt; ic++) {
    int ip = o->ipv[ic], im = o->imv[ic];
    for (int kc = 2; kc <= o->nz; kc++)
      for (int jc = 1; jc <= o->nr; jc++) {
        double qnn, qss, h31, h32, h33;
        qnn = (F3(q3, ic, jc + 1, kc) + F3(q3, ic, jc, kc)) * (F3(q2, ic, jc + 1, kc) + F3(q2, ic, jc + 1, kc - 1));
        qss = (F3(q3, ic, o->jmv[jc], kc) + F3(q3, ic, jc, kc)) * (F3(q2, ic, jc, kc) + F3(q2, ic, jc, kc - 1));
        h32 = (qnn - qss) * o->h1d22d[jc];
        qnn = (F3(q3, ic, jc, kc) + F3(q3, ip, jc, kc)) * (F3(q1, ip, jc, kc) + F3(q1, ip, jc, kc - 1));
        qss = (F3(q3, im, jc, kc) + F3(q3, ic, jc, kc)) * (F3(q1, ic, jc, kc) + F3(q1, ic, jc, kc - 1));
        h31 = (qnn - qss) * o->h1d11[jc];
        qnn = 0.25 * (F3(q3, ic, jc, kc) + F3(q3, ic, jc, kc + 1)) * (F3(q3, ic, jc, kc) + F3(q3, ic, jc, kc + 1));
        qss = 0.25 * (F3(q3, ic, jc, kc) + F3(q3, ic, jc, kc - 1)) * (F3(q3, ic, jc, kc) + F3(q3, ic, jc, kc - 1));
        h33 = (qnn - qss) * (o->zstr ? ZC3(kc) : o->dx3);
        F3(o->dq3, ic, jc, kc) = -(h31 + h32 + h33);
      }
  }
}

/* ---------------------------------------------------------------------------
 * solve_lap_thom, tripar.F90:120-145.  On one rank solve_lap_par
 * (tripar.F90:11-118) forces aa(1)=0, cc(n)=0 and its final solve is a plain
 * Thomas solve of the unmodified rhs; the three preceding solves only feed
 * the (trivial) reduced system.
 * ------------------------------------------------------------------------- */
static void thomas(orc_t *o, const double *rhs, double *sol, double *aa, const double *bb,
                   double *cc, int n) {
  double *beta = o->beta, *gamm = o->gamm;
  aa[1] = 0.0;
  cc[n] = 0.0;
  beta[1] = bb[1];
  gamm[1] = rhs[1] / bb[1];
  for (int k = 2; k <= n; k++) {
    beta[k] = bb[k] - cc[k - 1] * aa[k] / beta[k - 1];
    gamm[k] = (rhs[k] - aa[k] * gamm[k - 1]) / beta[k];
  }
  sol[n] = gamm[n];
  for (int k = 1; k <= n - 1; k++) sol[n - k] = gamm[n - k] - cc[n - k] * sol[n + 1 - k] / beta[n - k];
}

/* tripij, step.F90:487-534: cyclic tridiagonal along theta on scratch rows. */
static void tripij(orc_t *o, int i1, int i2, int mi, int mf) {
  int ib = i1 + 1, ie = i2 - 1, ibw = i1 + i2 - 1;
  for (int j = mi; j <= mf; j++) {
    TS(o->aci, j, i1) = 1.0 / TS(o->aci, j, i1);
    TS(o->gam2, j, i1) = -TS(o->ami, j, i1) * TS(o->aci, j, i1);
    TS(o->ami, j, i1) = TS(o->fi, j, i1) * TS(o->aci, j, i1);
  }
  for (int j = mi; j <= mf; j++) {
    for (int i = ib; i <= ie; i++) {
      int im = i - 1;
      TS(o->api, j, im) = TS(o->api, j, im) * TS(o->aci, j, im);
      TS(o->aci, j, i) = 1.0 / (TS(o->aci, j, i) - TS(o->ami, j, i) * TS(o->api, j, im));
      TS(o->gam2, j, i) = -TS(o->ami, j, i) * TS(o->gam2, j, im) * TS(o->aci, j, i);
      TS(o->ami, j, i) = (TS(o->fi, j, i) - TS(o->ami, j, i) * TS(o->ami, j, im)) * TS(o->aci, j, i);
    }
    TS(o->gam2, j, ie) = TS(o->gam2, j, ie) - TS(o->api, j, ie) * TS(o->aci, j, ie);
  }
  for (int j = mi; j <= mf; j++) {
    TS(o->fi, j, ie) = TS(o->ami, j, ie);
    TS(o->aci, j, ie) = TS(o->gam2, j, ie);
    for (int i = ib; i <= ie; i++) {
      int ip = ibw - i;
      TS(o->fi, j, ip) = TS(o->ami, j, ip) - TS(o->api, j, ip) * TS(o->fi, j, ip + 1);
      TS(o->aci, j, ip) = TS(o->gam2, j, ip) - TS(o->api, j, ip) * TS(o->aci, j, ip + 1);
    }
  }
  for (int j = mi; j <= mf; j++) {
    double za1 = TS(o->fi, j, i2) - TS(o->api, j, i2) * TS(o->fi, j, i1) - TS(o->ami, j, i2) * TS(o->fi, j, ie);
    double za2 = TS(o->aci, j, i2) + TS(o->api, j, i2) * TS(o->aci, j, i1) + TS(o->ami, j, i2) * TS(o->aci, j, ie);
    TS(o->fi, j, i2) = za1 / za2;
    for (int i = i1; i <= ie; i++) TS(o->fi, j, i) = TS(o->fi, j, i) + TS(o->aci, j, i) * TS(o->fi, j, i2);
  }
}

/* solqi, step.F90:212-263 */
static void solqi(orc_t *o, int ieq) {
  double betadx = 0.5 * o->alm[o->irk] * o->dt / o->Re;
  int ik = 1, ij = 1;
  if (ieq == 2) ij = 2;
  if (ieq == 3) ik = 2;
  for (int k = ik; k <= o->nz; k++) {
    for (int j = ij; j <= o->nr; j++) {
      double acc = (ieq == 2) ? betadx * o->s2t11[j] : betadx * o->s1t11[j];
      for (int i = 1; i <= o->nt; i++) {
        TS(o->ami, j, i) = -acc;
        TS(o->api, j, i) = -acc;
        TS(o->aci, j, i) = 1.0 + 2.0 * acc;
        TS(o->fi, j, i) = F3(o->qrhs, i, j, k);
      }
    }
    tripij(o, 1, o->nt, 1, o->nr);
    for (int j = ij; j <= o->nr; j++)
      for (int i = 1; i <= o->nt; i++) F3(o->qrhs, i, j, k) = TS(o->fi, j, i);
  }
}

/* solqj, step.F90:266-358 */
static void solqj(orc_t *o, int ieq) {
  int fk = (ieq == 2) ? o->nzc : o->nz;
  int nrloc = (ieq == 2) ? o->nrc : o->nr;
  double *am = ieq == 1 ? o->amj1 : ieq == 2 ? o->amj2 : o->amj3;
  double *ac = ieq == 1 ? o->acj1 : ieq == 2 ? o->acj2 : o->acj3;
  double *ap = ieq == 1 ? o->apj1 : ieq == 2 ? o->apj2 : o->apj3;
  for (int ic = 1; ic <= o->nt; ic++)
    for (int kc = 1; kc <= fk; kc++) {
      for (int jc = 1; jc <= o->nrc; jc++) {
        o->rtri[jc] = F3(o->qrhs, ic, jc, kc);
        o->atri[jc] = am[jc]; o->btri[jc] = ac[jc]; o->ctri[jc] = ap[jc];
      }
      if (ieq == 2) {
        o->rtri[1] = 0.0; o->atri[1] = 0.0; o->btri[1] = 1.0; o->ctri[1] = 0.0;
        o->rtri[o->nrc] = 0.0; o->atri[o->nrc] = 0.0; o->btri[o->nrc] = 1.0; o->ctri[o->nrc] = 0.0;
      }
      thomas(o, o->rtri, o->stri, o->atri, o->btri, o->ctri, nrloc);
      if (ieq == 2) {
        for (int jc = 1; jc <= o->nr; jc++) F3(o->dq2, ic, jc, kc) = o->stri[jc] + F3(o->q2, ic, jc, kc);
      } else {
        for (int jc = 1; jc <= o->nr; jc++) F3(o->qrhs, ic, jc, kc) = o->stri[jc];
      }
    }
}

/* solqk, step.F90:360-485 */
static void solqk(orc_t *o, int ieq) {
  const int nzc = o->nzc;
  double *am = ieq == 1 ? o->amk1 : ieq == 2 ? o->amk2 : o->amk3;
  double *ac = ieq == 1 ? o->ack1 : ieq == 2 ? o->ack2 : o->ack3;
  double *ap = ieq == 1 ? o->apk1 : ieq == 2 ? o->apk2 : o->apk3;
  for (int ic = 1; ic <= o->nt; ic++)
    for (int jc = 1; jc <= o->nrc; jc++) {
      for (int kc = 0; kc <= nzc + 1; kc++) o->rtri[kc] = 0.0;
      for (int kc = 1; kc <= nzc; kc++) {
        o->rtri[kc] = F3(o->qrhs, ic, jc, kc);
        o->atri[kc] = am[kc]; o->btri[kc] = ac[kc]; o->ctri[kc] = ap[kc];
      }
      if (ieq == 1) { o->rtri[nzc] = 0.0 * CR(o->corr1n, ic, jc); o->rtri[1] = 1.0 * CR(o->corr1s, ic, jc); }
      if (ieq == 2) { o->rtri[nzc] = 0.0 * CR(o->corr2n, ic, jc); o->rtri[1] = 1.0 * CR(o->corr2s, ic, jc); }
      if (ieq == 3) { o->rtri[nzc] = 1.0 * CR(o->corr3n, ic, jc); o->rtri[1] = 1.0 * CR(o->corr3s, ic, jc); }
      thomas(o, o->rtri, o->stri, o->atri, o->btri, o->ctri, nzc);
      for (int kc = 1; kc <= nzc; kc++) {
        if (ieq == 1) F3(o->dq1, ic, jc, kc) = F3(o->q1, ic, jc, kc) + o->stri[kc];
        if (ieq == 2) F3(o->qrhs, ic, jc, kc) = o->stri[kc];
        if (ieq == 3) F3(o->dq3, ic, jc, kc) = F3(o->q3, ic, jc, kc) + o->stri[kc];
      }
    }
}

/* z second difference of a centre quantity on the axial mesh: the
 * one-sided form through the wall value fs at the inlet cell, the
 * centred weights elsewhere. */
static double zdiff2c(const orc_t *o, double fm1, double f0, double fp1, double fs, int kc) {
  if (kc == 1) {
    const double h1 = 0.5 / ZM3(1), h2 = 1.0 / ZC3(2);
    const double ws = 2.0 / (h1 * (h1 + h2)), w2 = 2.0 / (h2 * (h1 + h2));
    return (fp1 - f0) * w2 - (f0 - fs) * ws;
  }
  return (fp1 - f0) * CP(kc) - (f0 - fm1) * CM(kc);
}

/* step, step.F90:17-210: explicit right-hand sides and the ADI sweeps. */
static void step(orc_t *o) {
  const int irk = o->irk;
  const double g = o->gam[irk], r = o->rom[irk], a = o->alm[irk], dt = o->dt, Re = o->Re;
  /* q1 (:32-63) */
  for (int ic = 1; ic <= o->nt; ic++) {
    int ip = o->ipv[ic], im = o->imv[ic];
    for (int jc = 1; jc <= o->nr; jc++) {
      int jm = o->jmv[jc];
      for (int kc = 1; kc <= o->nz; kc++) {
        double d11, d33, d22, qnn, qss;
        d11 = (F3(o->q1, ip, jc, kc) - 2.0 * F3(o->q1, ic, jc, kc) + F3(o->q1, im, jc, kc)) * o->s1t11[jc];
        if (o->zstr)
          d33 = zdiff2c(o, F3(o->q1, ic, jc, kc - 1), F3(o->q1, ic, jc, kc), F3(o->q1, ic, jc, kc + 1),
                        PL(o->q1s, ic, jc), kc);
        else if (kc == 1)
          d33 = 4.0 / 3.0 * (F3(o->q1, ic, jc, kc + 1) - 3.0 * F3(o->q1, ic, jc, kc) + 2.0 * PL(o->q1s, ic, jc)) * o->dx3q;
        else
          d33 = (F3(o->q1, ic, jc, kc + 1) - 2.0 * F3(o->q1, ic, jc, kc) + F3(o->q1, ic, jc, kc - 1)) * o->dx3q;
        qnn = F3(o->q1, ic, jc + 1, kc) / o->rm[jc + 1] - F3(o->q1, ic, jc, kc) / o->rm[jc];
        qnn = qnn * o->s1t2n[jc];
        qss = F3(o->q1, ic, jc, kc) / o->rm[jc] - F3(o->q1, ic, jm, kc) / o->rm[jm];
        qss = qss * o->s1t2s[jc];
        d22 = (qnn - qss) * o->s1t2c[jc];
        F3(o->qrhs, ic, jc, kc) = (g * F3(o->dq1, ic, jc, kc) + r * F3(o->ru1, ic, jc, kc) -
                                   a * PRG(ic, jc, kc, 1) + a / Re * (d11 + d33 + d22)) * dt;
        F3(o->ru1, ic, jc, kc) = F3(o->dq1, ic, jc, kc);
      }
    }
  }
  solqi(o, 1); solqj(o, 1); solqk(o, 1);
  for (int ic = 1; ic <= o->nt; ic++)
    for (int kc = 1; kc <= o->nzc; kc++) F3(o->dq1, ic, o->nrc, kc) = o->slip * F3(o->dq1, ic, o->nr, kc);
  /* q3 (:91-116) */
  for (int ic = 1; ic <= o->nt; ic++) {
    int ip = o->ipv[ic], im = o->imv[ic];
    for (int kc = 2; kc <= o->nz; kc++)
      for (int jc = 1; jc <= o->nr; jc++) {
        double d11, d33, d22, qnn, qss;
        d11 = (F3(o->q3, ip, jc, kc) - 2.0 * F3(o->q3, ic, jc, kc) + F3(o->q3, im, jc, kc)) * o->s1t11[jc];
        if (o->zstr)
          d33 = (F3(o->q3, ic, jc, kc + 1) - F3(o->q3, ic, jc, kc)) * FP(kc) -
                (F3(o->q3, ic, jc, kc) - F3(o->q3, ic, jc, kc - 1)) * FM(kc);
        else
          d33 = (F3(o->q3, ic, jc, kc + 1) - 2.0 * F3(o->q3, ic, jc, kc) + F3(o->q3, ic, jc, kc - 1)) * o->dx3q;
        qnn = (F3(o->q3, ic, jc + 1, kc) - F3(o->q3, ic, jc, kc)) * o->s3t2n[jc];
        qss = (F3(o->q3, ic, jc, kc) - F3(o->q3, ic, o->jmv[jc], kc)) * o->s3t2s[jc];
        d22 = qnn - qss;
        F3(o->qrhs, ic, jc, kc) = (g * F3(o->dq3, ic, jc, kc) + r * F3(o->ru3, ic, jc, kc) -
                                   a * PRG(ic, jc, kc, 3) + (a / Re) * (d11 + d33 + d22)) * dt;
        F3(o->ru3, ic, jc, kc) = F3(o->dq3, ic, jc, kc);
      }
  }
  solqi(o, 3); solqj(o, 3); solqk(o, 3);
  for (int ic = 1; ic <= o->nt; ic++)
    for (int kc = 1; kc <= o->nzc; kc++) F3(o->dq3, ic, o->nrc, kc) = o->slip * F3(o->dq3, ic, o->nr, kc);
  /* q2 (:142-165) */
  for (int ic = 1; ic <= o->nt; ic++) {
    int ip = o->ipv[ic], im = o->imv[ic];
    for (int kc = 1; kc <= o->nz; kc++)
      for (int jc = 2; jc <= o->nr; jc++) {
        double d11, d33, d22;
        d11 = (F3(o->q2, ip, jc, kc) - 2.0 * F3(o->q2, ic, jc, kc) + F3(o->q2, im, jc, kc)) * o->s2t11[jc];
        if (o->zstr)
          d33 = zdiff2c(o, F3(o->q2, ic, jc, kc - 1), F3(o->q2, ic, jc, kc), F3(o->q2, ic, jc, kc + 1),
                        PL(o->q2s, ic, jc), kc);
        else if (kc == 1)
          d33 = 4.0 / 3.0 * (F3(o->q2, ic, jc, kc + 1) - 3.0 * F3(o->q2, ic, jc, kc) + 2.0 * PL(o->q2s, ic, jc)) * o->dx3q;
        else
          d33 = (F3(o->q2, ic, jc, kc + 1) - 2.0 * F3(o->q2, ic, jc, kc) + F3(o->q2, ic, jc, kc - 1)) * o->dx3q;
        d22 = F3(o->q2, ic, jc + 1, kc) * o->s2t2p[jc] - F3(o->q2, ic, jc, kc) * o->s2t2c[jc] +
              F3(o->q2, ic, jc - 1, kc) * o->s2t2m[jc];
        F3(o->qrhs, ic, jc, kc) = (g * F3(o->dq2, ic, jc, kc) + r * F3(o->ru2, ic, jc, kc) -
                                   a * PRG(ic, jc, kc, 2) + a / Re * (d11 + d33 + d22)) * dt;
        F3(o->ru2, ic, jc, kc) = F3(o->dq2, ic, jc, kc);
      }
  }
  solqi(o, 2); solqk(o, 2); solqj(o, 2);
  for (int ic = 1; ic <= o->nt; ic++)
    for (int kc = 1; kc <= o->nzc; kc++) F3(o->dq2, ic, o->nrc, kc) = 0.0 * o->slip * F3(o->dq2, ic, o->nr, kc);
}

/* divq, divergence.F90:16-67 */
static void divq(orc_t *o) {
  const double den = o->alm[o->irk] * o->dt;
  for (int ic = 1; ic <= o->nt; ic++) {
    int ip = o->ipv[ic];
    for (int jc = 1; jc <= o->nr; jc++)
      for (int kc = 1; kc <= o->nz; kc++) {
        double t1 = (F3(o->dq1, ip, jc, kc) - F3(o->dq1, ic, jc, kc)) * o->dx1 * o->axsym / o->rm[jc];
        double t2 = (F3(o->dq2, ic, jc + 1, kc) - F3(o->dq2, ic, jc, kc)) * o->dx2 / (o->rm[jc] * o->dmr[jc]);
        double t3 = (F3(o->dq3, ic, jc, kc + 1) - F3(o->dq3, ic, jc, kc)) * (o->zstr ? ZM3(kc) : o->dx3);
        RS(o->rhs, ic, jc, kc) = -(t1 + t2 + t3) / den;
      }
  }
}

/* ---------------------------------------------------------------------------
 * Transforms.  The reference calls FFTW (fastfouriertransform.F90:34-37,
 * 79-82, 299-305); these follow FFTW's published definitions:
 *   r2c: F_k = sum_j f_j exp(-2 pi i jk/N); c2r: unnormalised inverse;
 *   REDFT10: Y_k = 2 sum_j x_j cos(pi (j+1/2) k / n);
 *   REDFT01: X_j = Y_0 + 2 sum_{k>=1} Y_k cos(pi k (j+1/2) / n).
 * ------------------------------------------------------------------------- */
static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* in-place radix-2 complex FFT, sign=-1 forward, +1 backward (unnormalised) */
static void cfft(int n, double *re, double *im, int sign) {
  for (int i = 1, j = 0; i < n; i++) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) { double t = re[i]; re[i] = re[j]; re[j] = t; t = im[i]; im[i] = im[j]; im[j] = t; }
  }
  for (int len = 2; len <= n; len <<= 1) {
    double ang = sign * 2.0 * M_PI / len;
    for (int i = 0; i < n; i += len)
      for (int k = 0; k < len / 2; k++) {
        double wr = cos(ang * k), wi = sin(ang * k);
        int a = i + k, b = i + k + len / 2;
        double xr = re[b] * wr - im[b] * wi, xi = re[b] * wi + im[b] * wr;
        re[b] = re[a] - xr; im[b] = im[a] - xi;
        re[a] += xr; im[a] += xi;
      }
  }
}

/* forward real DFT coefficients F_k, k=0..N/2 */
static void rdft(int N, const double *f, double *Fr, double *Fi, double *wr, double *wi) {
  if (is_pow2(N)) {
    for (int j = 0; j < N; j++) { wr[j] = f[j]; wi[j] = 0.0; }
    cfft(N, wr, wi, -1);
    for (int k = 0; k <= N / 2; k++) { Fr[k] = wr[k]; Fi[k] = wi[k]; }
  } else {
    for (int k = 0; k <= N / 2; k++) {
      double sr = 0.0, si = 0.0;
      for (int j = 0; j < N; j++) {
        long p = ((long)j * k) % N;
        double ang = -2.0 * M_PI * (double)p / N;
        sr += f[j] * cos(ang); si += f[j] * sin(ang);
      }
      Fr[k] = sr; Fi[k] = si;
    }
  }
}

/* fft_ptf_multi packing (fastfouriertransform.F90:39-43) on one line.
 * For even N the Nyquist imaginary part is not stored (the reference writes
 * it one element past the line, see its even-N out-of-bounds access).      */
static void ptf_line(int N, double *t, double *Fr, double *Fi, double *wr, double *wi) {
  int nhalf = N / 2 + 1;
  rdft(N, t, Fr, Fi, wr, wi);
  t[0] = Fr[0] / N;
  for (int k = 1; k <= nhalf - 1; k++) {
    t[2 * k - 1] = 2.0 * Fr[k] / N;
    if (2 * k < N) t[2 * k] = -2.0 * Fi[k] / N;
  }
}

/* fft_ftp_multi (fastfouriertransform.F90:74-82) on one line */
static void ftp_line(int N, double *t, double *wr, double *wi) {
  int nhalf = N / 2 + 1;
  if (is_pow2(N)) {
    for (int k = 0; k < N; k++) { wr[k] = 0.0; wi[k] = 0.0; }
    wr[0] = t[0];
    for (int k = 1; k <= nhalf - 1; k++) {
      double a = t[2 * k - 1], b = (2 * k < N) ? t[2 * k] : 0.0;
      wr[k] = a / 2.0; wi[k] = -b / 2.0;
      if (N - k != k) { wr[N - k] = a / 2.0; wi[N - k] = b / 2.0; }
    }
    cfft(N, wr, wi, +1);
    for (int j = 0; j < N; j++) t[j] = wr[j];
  } else {
    for (int j = 0; j < N; j++) {
      double s = t[0];
      for (int k = 1; k <= nhalf - 1; k++) {
        long p = ((long)j * k) % N;
        double ang = 2.0 * M_PI * (double)p / N;
        double a = t[2 * k - 1], b = (2 * k < N) ? t[2 * k] : 0.0;
        if (2 * k == N) s += (a / 2.0) * cos(ang);
        else s += a * cos(ang) + b * sin(ang);
      }
      wr[j] = s;
    }
    for (int j = 0; j < N; j++) t[j] = wr[j];
  }
}

/* REDFT10 of x[0..n-1] into y, times scale */
static void redft10(int n, const double *x, double *y, double scale, double *wr, double *wi) {
  if (is_pow2(n)) {
    int m = 2 * n;
    for (int j = 0; j < n; j++) { wr[j] = x[j]; wr[m - 1 - j] = x[j]; wi[j] = 0.0; wi[m - 1 - j] = 0.0; }
    cfft(m, wr, wi, -1);
    for (int k = 0; k < n; k++) {
      double ang = -M_PI * k / (2.0 * n);
      y[k] = (wr[k] * cos(ang) - wi[k] * sin(ang)) * scale;
    }
  } else {
    for (int k = 0; k < n; k++) {
      double s = 0.0;
      for (int j = 0; j < n; j++) s += x[j] * cos(M_PI * (j + 0.5) * k / n);
      y[k] = 2.0 * s * scale;
    }
  }
}

/* REDFT01 of y[0..n-1] into x */
static void redft01(int n, const double *y, double *x, double *wr, double *wi) {
  if (is_pow2(n)) {
    int m = 2 * n;
    for (int k = 0; k < m; k++) { wr[k] = 0.0; wi[k] = 0.0; }
    for (int k = 0; k < n; k++) {
      double c = (k == 0) ? 1.0 : 2.0, ang = M_PI * k / (2.0 * n);
      wr[k] = c * y[k] * cos(ang); wi[k] = c * y[k] * sin(ang);
    }
    cfft(m, wr, wi, +1);
    for (int j = 0; j < n; j++) x[j] = wr[j];
  } else {
    for (int j = 0; j < n; j++) {
      double s = y[0];
      for (int k = 1; k < n; k++) s += 2.0 * y[k] * cos(M_PI * k * (j + 0.5) / n);
      x[j] = s;
    }
  }
}

/* solve_radial, poisson_solve.F90:270-323 */
static void solve_radial(orc_t *o, int m, int mz, const double *rhs, double *sol, int dn2) {
  const int nr = o->nr;
  double *aa = o->atri, *bb = o->btri, *cc = o->ctri;
  double pi = acos(-1.0), wth, wz;
  if (o->wvnex) {
    wth = (double)((float)m * (float)m);
    wz = 2.0 / (o->dz * o->dz) * (1.0 - cos(mz * pi / o->nz));
  } else {
    wth = 2.0 / (o->dtheta * o->dtheta) * (1.0 - cos(2.0 * pi * m / o->nt));
    wz = 2.0 / (o->dz * o->dz) * (1.0 - cos(mz * pi / o->nz));
  }
  if (o->zstr) wz = o->zlam[mz];
  for (int j = 1; j <= nr; j++) {
    aa[j] = o->aa_base[j];
    bb[j] = o->bb_base[j] + wth / (o->rm[j] * o->rm[j]) + wz;
    cc[j] = o->cc_base[j];
  }
  int bcr = (m + mz != 0) ? 1 : dn2;
  bb[1] = bb[1] + 1 * aa[1];
  bb[nr] = bb[nr] + bcr * cc[nr];
  thomas(o, rhs, sol, aa, bb, cc, nr);
}

/* laplace_solve / solve_lap_tripar, poisson_solve.F90:13-32, 140-179 */
static void poisson_core(orc_t *o, const double *rhs_in, double *sol_out) {
  const int nt = o->nt, nr = o->nr, nz = o->nz;
  int nmax = 2 * (nt > nz ? nt : nz) + 4;
  double *wr = dalloc(nmax), *wi = dalloc(nmax), *Fr = dalloc(nmax), *Fi = dalloc(nmax);
  double *line = dalloc(nmax), *line2 = dalloc(nmax);
  double *act = dalloc((size_t)nt * nr * nz);
  memcpy(act, rhs_in, sizeof(double) * (size_t)nt * nr * nz);
  /* theta transform per (j,k) line: contiguous in memory */
  for (size_t l = 0; l < (size_t)nr * nz; l++) ptf_line(nt, act + l * nt, Fr, Fi, wr, wi);
  /* z DCT-II per (m, j), then /(2 nz) (fastfouriertransform.F90:330-332) */
  for (int m = 1; m <= nt; m++)
    for (int j = 1; j <= nr; j++) {
      for (int k = 1; k <= nz; k++) line[k - 1] = RS(act, m, j, k);
      if (o->zstr) {                 /* generalised eigen-transform: V^-1 line */
        for (int q = 0; q < nz; q++) {
          double acc = 0.0;
          for (int k = 0; k < nz; k++) acc += o->zvi[q * nz + k] * line[k];
          RS(act, m, j, q + 1) = acc;
        }
        continue;
      }
      redft10(nz, line, line2, 1.0, wr, wi);
      for (int k = 1; k <= nz; k++) RS(act, m, j, k) = line2[k - 1] / (2.0 * nz * 1);
    }
  /* radial solves (poisson_solve.F90:162-169) */
  int nhalf = nt / 2;
  double *r1 = dalloc(nr + 2), *s1 = dalloc(nr + 2);
  for (int k = 1; k <= nz; k++) {
    int mz = k - 1;
    for (int idx = 1; idx <= nt; idx++) {
      int m = idx / 2;           /* idx=1 -> m=0; 2m and 2m+1 -> m */
      int dn2 = (idx == 1) ? -1 : 1;
      if (idx > 1 && 2 * nhalf == nt && idx == nt + 1) continue;
      for (int j = 1; j <= nr; j++) r1[j] = RS(act, idx, j, k);
      solve_radial(o, m, mz, r1, s1, dn2);
      for (int j = 1; j <= nr; j++) RS(sol_out, idx, j, k) = s1[j];
    }
  }
  /* inverse z transform (REDFT01, unnormalised) and inverse theta transform */
  for (int m = 1; m <= nt; m++)
    for (int j = 1; j <= nr; j++) {
      for (int k = 1; k <= nz; k++) line[k - 1] = RS(sol_out, m, j, k);
      if (o->zstr) {                 /* V line_hat */
        for (int k = 0; k < nz; k++) {
          double acc = 0.0;
          for (int q = 0; q < nz; q++) acc += o->zvf[k * nz + q] * line[q];
          RS(sol_out, m, j, k + 1) = acc;
        }
        continue;
      }
      redft01(nz, line, line2, wr, wi);
      for (int k = 1; k <= nz; k++) RS(sol_out, m, j, k) = line2[k - 1];
    }
  for (size_t l = 0; l < (size_t)nr * nz; l++) ftp_line(nt, sol_out + l * nt, wr, wi);
  free(act); free(wr); free(wi); free(Fr); free(Fi); free(line); free(line2); free(r1); free(s1);
}

int orc_poisson(orc_t *o, const double *rhs, double *sol) {
  poisson_core(o, rhs, sol);
  return 0;
}

static void laplace_solve(orc_t *o) {
  poisson_core(o, o->rhs, o->sol);
  for (int k = 1; k <= o->nz; k++)                 /* poisson_solve.F90:29 */
    for (int j = 1; j <= o->nr; j++)
      for (int i = 1; i <= o->nt; i++) DP(i, j, k) = RS(o->sol, i, j, k);
}

/* upsol, divergence.F90:69-236 */
static void upsol(orc_t *o) {
  const int nt = o->nt, nr = o->nr, nrc = o->nrc, nz = o->nz, nzc = o->nzc;
  const double adt = o->alm[o->irk] * o->dt;
  for (int ic = 1; ic <= nt; ic++)
    for (int jc = 1; jc <= nr; jc++) DP(ic, jc, nzc) = DP(ic, jc, nz);            /* :93-99 */
  for (int ic = 1; ic <= nt; ic++) {                                              /* :105-112 */
    int im = o->imv[ic];
    for (int jc = 1; jc <= nr; jc++)
      for (int kc = 1; kc <= nzc; kc++)
        F3(o->q1, ic, jc, kc) = F3(o->dq1, ic, jc, kc) - (adt * o->dx1 / o->rm[jc]) * (DP(ic, jc, kc) - DP(im, jc, kc));
  }
  for (int ic = 1; ic <= nt; ic++)
    for (int kc = 1; kc <= nzc; kc++) F3(o->q1, ic, nrc, kc) = o->slip * F3(o->q1, ic, nr, kc);   /* :115-121 */
  for (int ic = 1; ic <= nt; ic++)                                                /* :128-136 */
    for (int jc = 2; jc <= nr; jc++)
      for (int kc = 1; kc <= nzc; kc++)
        F3(o->q2, ic, jc, kc) = F3(o->dq2, ic, jc, kc) -
                                adt * o->rc[jc] * o->dx2 * (DP(ic, jc, kc) - DP(ic, o->jmv[jc], kc)) / o->dcr[jc];
  for (int kc = 1; kc <= nzc; kc++)                                               /* :148-153 */
    for (int ic = 1; ic <= nt; ic++) { F3(o->q2, ic, 1, kc) = 0.0; F3(o->q2, ic, nrc, kc) = 0.0; }
  for (int kc = 2; kc <= nzc; kc++)                                               /* :167-174 */
    for (int jc = 1; jc <= nr; jc++)
      for (int ic = 1; ic <= nt; ic++)
        F3(o->q3, ic, jc, kc) = F3(o->dq3, ic, jc, kc) -
                                adt * (o->zstr ? ZC3(kc) : o->dx3) * (DP(ic, jc, kc) - DP(ic, jc, kc - 1));
  for (int jc = 0; jc <= nrc; jc++)                                               /* :193-200 */
    for (int ic = 1; ic <= nt; ic++) F3(o->q3, ic, jc, 1) = F3(o->dq3, ic, jc, 1);
  for (int ic = 1; ic <= nt; ic++)                                                /* :203-208 */
    for (int kc = 2; kc <= nzc; kc++) F3(o->q3, ic, nrc, kc) = o->slip * F3(o->q3, ic, nr, kc);
}

/* uppres, divergence.F90:239-480.  Returns -2 on the "Flow field crash" NaN. */
static int uppres(orc_t *o) {
  const int nt = o->nt, nr = o->nr, nrc = o->nrc, nz = o->nz, nzc = o->nzc;
  double *qcap = o->qcap;
  for (int ic = 1; ic <= nt; ic++)
    for (int jc = 1; jc <= nr; jc++) DP(ic, jc, nzc) = DP(ic, jc, nz);            /* :254-258 */
  double be = 0.5 * (o->alm[o->irk] * o->dt) / o->Re;                             /* :263 */
  if (nt > 1) {
    for (int ic = 1; ic <= nt; ic++) {                                            /* :272-279 */
      int im = o->imv[ic];
      for (int jc = 1; jc <= nr; jc++)
        for (int kc = 1; kc <= nzc; kc++)
          F3(qcap, ic, jc, kc) = (DP(ic, jc, kc) - DP(im, jc, kc)) * o->dx1 / o->rm[jc];
    }
    for (int kc = 1; kc <= nz; kc++)                                              /* :286-310 */
      for (int jc = 1; jc <= nr; jc++)
        for (int ic = 1; ic <= nt; ic++) {
          int km = o->kmv[kc], kp = kc + 1, jm = o->jmv[jc], jp = o->jpv[jc];
          int ip = o->ipv[ic], im = o->imv[ic];
          double d11 = (F3(qcap, ip, jc, kc) - 2.0 * F3(qcap, ic, jc, kc) + F3(qcap, im, jc, kc)) * o->s1t11[jc];
          double d33 = o->zstr ? (F3(qcap, ic, jc, kp) - F3(qcap, ic, jc, kc)) * CP(kc) -
                                     (F3(qcap, ic, jc, kc) - F3(qcap, ic, jc, km)) * CM(kc)
                               : (F3(qcap, ic, jc, kp) - 2.0 * F3(qcap, ic, jc, kc) + F3(qcap, ic, jc, km)) * o->dx3q;
          double qnn = (F3(qcap, ic, jp, kc) / o->rm[jc + 1] - F3(qcap, ic, jc, kc) / o->rm[jc]);
          double qss = (F3(qcap, ic, jc, kc) / o->rm[jc] - F3(qcap, ic, jm, kc) / o->rm[jm]);
          double d22 = (qnn * o->s1t2n[jc] - qss * o->s1t2s[jc]) * o->s1t2c[jc];
          PRG(ic, jc, kc, 1) = PRG(ic, jc, kc, 1) + F3(qcap, ic, jc, kc) - be * (d11 + d33 + d22);
        }
  }
  for (int ic = 1; ic <= nt; ic++)                                                /* :316-321 */
    for (int kc = 1; kc <= nz; kc++) { F3(qcap, ic, 1, kc) = 0.0; F3(qcap, ic, nrc, kc) = 0.0; }
  for (int kc = 1; kc <= nzc; kc++)                                               /* :324-330 */
    for (int jc = 2; jc <= nr; jc++)
      for (int ic = 1; ic <= nt; ic++)
        F3(qcap, ic, jc, kc) = (DP(ic, jc, kc) - DP(ic, jc - 1, kc)) * o->dx2 * o->rc[jc] / o->dcr[jc];
  for (int kc = 1; kc <= nz; kc++)                                                /* :351-373 */
    for (int jc = 2; jc <= nr; jc++)
      for (int ic = 1; ic <= nt; ic++) {
        int km = o->kmv[kc], kp = kc + 1, ip = o->ipv[ic], im = o->imv[ic];
        double d11 = (F3(qcap, ip, jc, kc) - 2.0 * F3(qcap, ic, jc, kc) + F3(qcap, im, jc, kc)) * o->s2t11[jc];
        double d33 = o->zstr ? (F3(qcap, ic, jc, kp) - F3(qcap, ic, jc, kc)) * CP(kc) -
                                   (F3(qcap, ic, jc, kc) - F3(qcap, ic, jc, km)) * CM(kc)
                             : (F3(qcap, ic, jc, kp) - 2.0 * F3(qcap, ic, jc, kc) + F3(qcap, ic, jc, km)) * o->dx3q;
        double qnn = F3(qcap, ic, jc + 1, kc) - F3(qcap, ic, jc, kc);
        double qss = F3(qcap, ic, jc, kc) - F3(qcap, ic, jc - 1, kc);
        double d22 = qnn * o->s2t2p[jc] - qss * o->s2t2m[jc];
        PRG(ic, jc, kc, 2) = PRG(ic, jc, kc, 2) + F3(qcap, ic, jc, kc) - be * (d11 + d22 + d33);
      }
  for (int jc = 2; jc <= nr; jc++)                                                /* :383-387 */
    for (int ic = 1; ic <= nt; ic++) F3(qcap, ic, jc, 1) = 0.0;
  int crash = 0;
  for (int kc = 2; kc <= nzc; kc++)                                               /* :392-402 */
    for (int jc = 1; jc <= nr; jc++)
      for (int ic = 1; ic <= nt; ic++) {
        F3(qcap, ic, jc, kc) = (DP(ic, jc, kc) - DP(ic, jc, kc - 1)) * (o->zstr ? ZC3(kc) : o->dx3);
        if (isnan(F3(qcap, ic, jc, kc))) crash = 1;
      }
  if (crash) return -2;
  for (int kc = 2; kc <= nz; kc++)                                                /* :410-439 */
    for (int jc = 1; jc <= nr; jc++)
      for (int ic = 1; ic <= nt; ic++) {
        int jm = o->jmv[jc], jp = o->jpv[jc], ip = o->ipv[ic], im = o->imv[ic];
        double d11 = (F3(qcap, ip, jc, kc) - 2.0 * F3(qcap, ic, jc, kc) + F3(qcap, im, jc, kc)) * o->s1t11[jc];
        double d33 = o->zstr ? (F3(qcap, ic, jc, kc + 1) - F3(qcap, ic, jc, kc)) * FP(kc) -
                                   (F3(qcap, ic, jc, kc) - F3(qcap, ic, jc, kc - 1)) * FM(kc)
                             : (F3(qcap, ic, jc, kc + 1) - 2.0 * F3(qcap, ic, jc, kc) + F3(qcap, ic, jc, kc - 1)) * o->dx3q;
        double qnn = (F3(qcap, ic, jp, kc) - F3(qcap, ic, jc, kc));
        double qss = (F3(qcap, ic, jc, kc) - F3(qcap, ic, jm, kc));
        double d22 = qnn * o->s3t2n[jc] - qss * o->s3t2s[jc];
        PRG(ic, jc, kc, 3) = PRG(ic, jc, kc, 3) + F3(qcap, ic, jc, kc) - be * (d11 + d22 + d33);
      }
  for (int kc = 1; kc <= nz; kc++)                                                /* :445-476 */
    for (int jc = 1; jc <= nr; jc++)
      for (int ic = 1; ic <= nt; ic++) {
        int km = o->kmv[kc], kp = kc + 1, jm = o->jmv[jc], jp = o->jpv[jc];
        int ip = o->ipv[ic], im = o->imv[ic];
        double d11 = (DP(ip, jc, kc) + DP(im, jc, kc) - 2.0 * DP(ic, jc, kc)) * o->s1t11[jc];
        double qnn = DP(ic, jp, kc) - DP(ic, jc, kc);
        double qss = DP(ic, jc, kc) - DP(ic, jm, kc);
        double d22 = qnn * o->s3t2n[jc] - qss * o->s3t2s[jc];
        double d33 = o->zstr ? (DP(ic, jc, kp) - DP(ic, jc, kc)) * CP(kc) - (DP(ic, jc, kc) - DP(ic, jc, km)) * CM(kc)
                             : (DP(ic, jc, kp) + DP(ic, jc, km) - 2.0 * DP(ic, jc, kc)) * o->dx3q;
        F3(o->pres, ic, jc, kc) = F3(o->pres, ic, jc, kc) + DP(ic, jc, kc) - be * (d11 + d22 + d33);
      }
  return 0;
}

int orc_phase(orc_t *o, int phase, int irk) {
  if (irk < 1 || irk > 3) return -1;
  o->irk = irk;
  switch (phase) {
    case ORC_UCOEFF: ucoeff(o); return 0;
    case ORC_COESYS: coesys(o); return 0;
    case ORC_HDNBC: hdnbc(o); return 0;
    case ORC_HDNL1: hdnl1(o); return 0;
    case ORC_HDNL2: hdnl2(o); return 0;
    case ORC_HDNL3: hdnl3(o); return 0;
    case ORC_STEP: step(o); return 0;
    case ORC_DIVQ: divq(o); return 0;
    case ORC_LAPLACE: laplace_solve(o); return 0;
    case ORC_UPSOL: upsol(o); return 0;
    case ORC_UPPRES: return uppres(o);
  }
  return -1;
}

/* one RK substep in the order of loop.F90:71-127 */
int orc_substep(orc_t *o, int irk) {
  o->irk = irk;
  ucoeff(o); coesys(o); hdnbc(o); hdnl1(o); hdnl2(o); hdnl3(o); step(o); divq(o);
  laplace_solve(o); upsol(o);
  return uppres(o);
}

int orc_rk3_step(orc_t *o) {
  o->ntime = o->dt + o->ntime;                                    /* loop.F90:53 */
  for (int irk = 1; irk <= 3; irk++) {
    int rc = orc_substep(o, irk);
    if (rc) return rc;
  }
  return 0;
}

void orc_divergence(orc_t *o, const double *q1, const double *q2, const double *q3, double *div) {
  for (int ic = 1; ic <= o->nt; ic++) {
    int ip = o->ipv[ic];
    for (int jc = 1; jc <= o->nr; jc++)
      for (int kc = 1; kc <= o->nz; kc++) {
        double t1 = (F3(q1, ip, jc, kc) - F3(q1, ic, jc, kc)) * o->dx1 * o->axsym / o->rm[jc];
        double t2 = (F3(q2, ic, jc + 1, kc) - F3(q2, ic, jc, kc)) * o->dx2 / (o->rm[jc] * o->dmr[jc]);
        double t3 = (F3(q3, ic, jc, kc + 1) - F3(q3, ic, jc, kc)) * (o->zstr ? ZM3(kc) : o->dx3);
        RS(div, ic, jc, kc) = t1 + t2 + t3;
      }
  }
}

/* ===========================================================================
 * Particles, part_tools.F90:228-371 (constants, fit) and 374-1025 (step).
 * ========================================================================= */

/* 27-point stencil offsets xint (part_tools.F90:296-322): ifst1 (theta) is the
 * fastest index, then ifst2 (r), then ifst3 (z). */
static void part_xint(const orc_t *o, double xint[27][3]) {
  for (int c = -1; c <= 1; c++)
    for (int b = -1; b <= 1; b++)
      for (int a = -1; a <= 1; a++) {
        int iv = (a + 1) + 3 * (b + 1) + 9 * (c + 1);
        xint[iv][0] = a * o->dtheta; xint[iv][1] = b * o->dr; xint[iv][2] = c * o->dz;
      }
}

/* monomials of the fit (part_tools.F90:367-369):
 * x1, x2, x3, x1^2, x1 x2, x2^2, x2 x3, x3^2, x3 x1 */
static void part_basis(const double x[3], double b[9]) {
  b[0] = x[0]; b[1] = x[1]; b[2] = x[2];
  b[3] = x[0] * x[0]; b[4] = x[0] * x[1]; b[5] = x[1] * x[1];
  b[6] = x[1] * x[2]; b[7] = x[2] * x[2]; b[8] = x[2] * x[0];
}

/* mint = (X^T X)^{-1} X^T, the least-squares pseudo-inverse the reference
 * tabulates in closed form at part_tools.F90:324-357. */
void orc_part_mint(orc_t *o, double *mint) {
  double xint[27][3], X[27][9], A[9][18];
  part_xint(o, xint);
  for (int k = 0; k < 27; k++) part_basis(xint[k], X[k]);
  for (int i = 0; i < 9; i++)
    for (int j = 0; j < 18; j++) {
      if (j < 9) { double s = 0; for (int k = 0; k < 27; k++) s += X[k][i] * X[k][j]; A[i][j] = s; }
      else A[i][j] = (j - 9 == i) ? 1.0 : 0.0;
    }
  for (int c = 0; c < 9; c++) {       /* Gauss-Jordan with partial pivoting */
    int p = c;
    for (int r = c + 1; r < 9; r++) if (fabs(A[r][c]) > fabs(A[p][c])) p = r;
    if (p != c) for (int j = 0; j < 18; j++) { double t = A[c][j]; A[c][j] = A[p][j]; A[p][j] = t; }
    double piv = A[c][c];
    for (int j = 0; j < 18; j++) A[c][j] /= piv;
    for (int r = 0; r < 9; r++)
      if (r != c) { double f = A[r][c]; for (int j = 0; j < 18; j++) A[r][j] -= f * A[c][j]; }
  }
  for (int i = 0; i < 9; i++)
    for (int k = 0; k < 27; k++) {
      double s = 0; for (int j = 0; j < 9; j++) s += A[i][9 + j] * X[k][j];
      mint[i * 27 + k] = s;
    }
}

/* theta from (x,y) with the reference's acos branches (part_tools.F90:524-534) */
static double part_theta(double x, double y, double r) {
  double pi = (double)4.0f * atan(1.0);
  if (x >= 0.0 && y >= 0.0) return acos(x / r);
  if (x < 0.0 && y >= 0.0) return acos(x / r);
  return acos(-x / r) + pi;
}

static double sample(orc_t *o, int field_mode, int ivar, int itof, int irof, int izof) {
  if (field_mode == 1) {                     /* dertest/part_tools.F90:580-582 */
    if (ivar == 1) return 0.0;
    if (ivar == 2) return o->rc[irof] * o->rc[irof];
    return (double)1.0f;
  }
  if (ivar == 1) return F3(o->q1, itof, irof, izof);
  if (ivar == 2) return (o->rc[irof] > (double)0.0001f) ? F3(o->q2, itof, irof, izof) / o->rc[irof] : 0.0;
  return F3(o->q3, itof, irof, izof);
}

/* Interpolation and chain rule for one particle: part_tools.F90:519-797.
 * Returns -1 if the stencil leaves the arrays (the reference reads halos). */
int orc_part_interp(orc_t *o, int field_mode, const double xp[3], double vel[3],
                    double dvel[9], double ddvel[27]) {
  static const double zero = 0.0;
  double mint[9 * 27], xint[27][3];
  orc_part_mint(o, mint);
  part_xint(o, xint);
  double x = xp[0], y = xp[1], z = xp[2];
  double r = sqrt(x * x + y * y);
  double theta = part_theta(x, y, r);
  double ff[4], dff[4][4], ddff[4][4][4];
  double sn = 0, cs = 0;
  int it = (int)floor(theta / o->dtheta) + 1;
  if (o->nt == 1) it = 1;
  int ir = (int)floor(r / o->dr) + 1;
  int iz = (int)floor(z / o->dz) + 1;
  if (ir < 0) return -1;
  if (ir + 1 > o->nr + 2 || iz - 1 < 0 || iz + 1 > o->nz + 2 || ir - 1 < 0) return -1;
  for (int ivar = 1; ivar <= 3; ivar++) {
    double yint[27], acoef[9] = {0}, rint[27], mp[9];
    for (int c = -1; c <= 1; c++)
      for (int b = -1; b <= 1; b++)
        for (int a = -1; a <= 1; a++) {
          int iv = (a + 1) + 3 * (b + 1) + 9 * (c + 1);
          int itof = it + a;
          if (itof > o->nt) itof = 1;
          if (itof < 1) itof = o->nt;
          yint[iv] = sample(o, field_mode, ivar, itof, ir + b, iz + c);
        }
    double y0 = yint[13];
    for (int iv = 0; iv < 27; iv++) yint[iv] = yint[iv] - y0;
    for (int iint = 1; iint <= 100; iint++) {           /* nitr=100, :609-624 */
      for (int k = 0; k < 27; k++) {
        const double *xi = xint[k];
        rint[k] = xi[0] * acoef[0] + xi[1] * acoef[1] + xi[2] * acoef[2] + xi[0] * xi[0] * acoef[3] +
                  xi[0] * xi[1] * acoef[4] + xi[1] * xi[1] * acoef[5] + xi[1] * xi[2] * acoef[6] +
                  xi[2] * xi[2] * acoef[7] + xi[2] * xi[0] * acoef[8];
      }
      double nrm = 0.0;
      for (int i = 0; i < 9; i++) {
        mp[i] = 0.0;
        for (int k = 0; k < 27; k++) mp[i] = mp[i] - mint[i * 27 + k] * (rint[k] - yint[k]);
        nrm += mp[i] * mp[i];
      }
      for (int i = 0; i < 9; i++) acoef[i] = acoef[i] + mp[i];
      if (sqrt(nrm) < (double)0.000001f) break;
    }
    double dtp, drp, dzp;
    if (ivar == 1) { dtp = theta - o->thc[it]; drp = r - o->rm[ir]; dzp = z - o->zm[iz]; }
    else if (ivar == 2) { dtp = theta - o->thm[it]; drp = r - o->rc[ir]; dzp = z - o->zm[iz]; }
    else { dtp = theta - o->thm[it]; drp = r - o->rm[ir]; dzp = z - o->zc[iz]; }
    double twopi = 8.0 * atan(1.0);
    if (dtp > o->dtheta) dtp = dtp - twopi;
    if (dtp < -o->dtheta) dtp = dtp + twopi;
    const double *a = acoef;
    double f = dtp * a[0] + drp * a[1] + dzp * a[2] + dtp * dtp * a[3] + dtp * drp * a[4] + drp * drp * a[5] +
               drp * dzp * a[6] + dzp * dzp * a[7] + dzp * dtp * a[8] + y0;
    double FI[12];
    FI[0] = a[0] + 2.0 * dtp * a[3] + drp * a[4] + dzp * a[8];    /* ft  */
    FI[1] = a[1] + dtp * a[4] + 2.0 * drp * a[5] + dzp * a[6];    /* fr  */
    FI[2] = a[2] + drp * a[6] + 2.0 * dzp * a[7] + dtp * a[8];    /* fz  */
    FI[3] = 2.0 * a[3];                                           /* ftt */
    FI[4] = drp * a[4];                                           /* ftr (as :657) */
    FI[5] = dzp * a[8];                                           /* ftz (as :658) */
    FI[6] = a[4];                                                 /* frt */
    FI[7] = 2.0 * a[5];                                           /* frr */
    FI[8] = a[6];                                                 /* frz */
    FI[9] = a[8];                                                 /* fzt */
    FI[10] = a[6];                                                /* fzr */
    FI[11] = 2.0 * a[7];                                          /* fzz */
    sn = sin(theta); cs = cos(theta);
    double sn2 = sin(2.0 * theta), cs2 = cos(2.0 * theta);
    /* chain-rule rows (:671-684): d/dx, d/dy, d/dz, then second derivatives
     * (id1,id2) at row 3+id1+3(id2-1) */
    double M[12][12];
    memset(M, 0, sizeof(M));
    double r2 = r * r;
    M[0][0] = -sn / r; M[0][1] = cs;
    M[1][0] = cs / r; M[1][1] = sn;
    M[2][2] = 1.0;
    double M4[9] = {sn2 / r2, sn * sn / r, 0, sn * sn / r2, -cs * sn / r, 0, -cs * sn / r, cs * cs, 0};
    double M5[9] = {-cs2 / r2, -cs * sn / r, 0, -cs * sn / r2, -sn * sn / r, 0, cs * cs / r, cs * sn, 0};
    double M7[9] = {-cs2 / r2, -cs * sn / r, 0, -cs * sn / r2, cs * cs / r, 0, -sn * sn / r, cs * sn, 0};
    double M8[9] = {-sn2 / r2, cs * cs / r, 0, cs * cs / r2, cs * sn / r, 0, cs * sn / r, sn * sn, 0};
    for (int q = 0; q < 9; q++) { M[3][q] = M4[q]; M[4][q] = M5[q]; M[6][q] = M7[q]; M[7][q] = M8[q]; }
    M[5][5] = -sn / r; M[5][8] = cs;
    M[8][5] = cs / r; M[8][8] = sn;
    M[9][9] = -sn / r; M[9][10] = cs;
    M[10][9] = cs / r; M[10][10] = sn;
    M[11][11] = 1.0;
    ff[ivar] = f;
    for (int id = 1; id <= 3; id++) {
      double s = 0.0;
      for (int im = 0; im < 12; im++) s = s + M[id - 1][im] * FI[im];
      dff[ivar][id] = s;
    }
    for (int id1 = 1; id1 <= 3; id1++)
      for (int id2 = 1; id2 <= 3; id2++) {
        double s = 0.0;
        for (int im = 0; im < 12; im++) s = s + M[3 + id1 + 3 * (id2 - 1) - 1][im] * FI[im];
        ddff[ivar][id1][id2] = s;
      }
  }
  (void)zero;
  double rs = fmax((double)0.000001f, r);
  double NM[4][4] = {{0}}, DN[4][4][4], DDN[4][4][4][4];
  memset(DN, 0, sizeof(DN)); memset(DDN, 0, sizeof(DDN));
  NM[1][1] = -sn; NM[1][2] = cs; NM[1][3] = 0.0;
  NM[2][1] = cs; NM[2][2] = sn; NM[2][3] = 0.0;
  NM[3][3] = 1.0;
  /* DN[j][i][k] = d NM(i,k) / d x_j (part_tools.F90:709-719) */
  DN[1][1][1] = cs * sn / rs; DN[1][1][2] = sn * sn / rs;
  DN[1][2][1] = sn * sn / rs; DN[1][2][2] = -cs * sn / rs;
  DN[2][1][1] = -cs * cs / rs; DN[2][1][2] = -sn * cs / rs;
  DN[2][2][1] = -sn * cs / rs; DN[2][2][2] = cs * cs / rs;
  /* DDN[l][j][i][k] = d2 NM(i,k) / dx_l dx_j (part_tools.F90:725-760) */
  double rs2 = rs * rs;
  double A = sn * (-2.0 * (cs * cs) + sn * sn) / rs2;   /* sn(-2cs^2+sn^2)/rs^2 */
  double B = cs * ((cs * cs) - 2.0 * (sn * sn)) / rs2;  /* cs(cs^2-2sn^2)/rs^2 */
  double C = -3.0 * cs * (sn * sn) / rs2;
  double D = 3.0 * (cs * cs) * sn / rs2;
  DDN[1][1][1][1] = A; DDN[1][1][1][2] = C; DDN[1][1][2][1] = C; DDN[1][1][2][2] = -A;
  DDN[1][2][1][1] = B; DDN[1][2][1][2] = -A; DDN[1][2][2][1] = -A; DDN[1][2][2][2] = -B;
  DDN[2][1][1][1] = B; DDN[2][1][1][2] = -A; DDN[2][1][2][1] = -A; DDN[2][1][2][2] = -B;
  DDN[2][2][1][1] = D; DDN[2][2][1][2] = -B; DDN[2][2][2][1] = -B; DDN[2][2][2][2] = -D;
  for (int id = 1; id <= 3; id++) {                 /* :762-797 */
    double s = 0.0;
    for (int jd = 1; jd <= 3; jd++) s = s + ff[jd] * NM[id][jd];
    vel[id - 1] = s;
  }
  for (int id = 1; id <= 3; id++)
    for (int jd = 1; jd <= 3; jd++) {
      double s = 0.0;
      for (int kd = 1; kd <= 3; kd++) s = s + dff[kd][jd] * NM[id][kd] + ff[kd] * DN[jd][id][kd];
      dvel[(id - 1) + 3 * (jd - 1)] = s;
    }
  for (int id = 1; id <= 3; id++)
    for (int jd = 1; jd <= 3; jd++)
      for (int ld = 1; ld <= 3; ld++) {
        double s = 0.0;
        for (int kd = 1; kd <= 3; kd++)
          s = s + ddff[kd][jd][ld] * NM[id][kd] + dff[kd][jd] * DN[ld][id][kd] + dff[kd][ld] * DN[jd][id][kd] +
              ff[kd] * DDN[ld][jd][id][kd];
        ddvel[(id - 1) + 3 * (jd - 1) + 9 * (ld - 1)] = s;
      }
  return 0;
}

/* ODE right-hand side, part_tools.F90:821-870; u and ddt are 1-based rows */
static void part_rates(double st, const double *u, const double vel[3], const double dvel[9],
                       const double ddvel[27], double *ddt) {
  const int ivel = 3, ijac = 6, ijrt = 15, ihes = 24, ihrt = 51;
  double jac[4][4], hes[4][4][4];
  double inv = (double)1.0f / st;
#define DV(i, j) dvel[((i) - 1) + 3 * ((j) - 1)]
#define DDV(i, j, k) ddvel[((i) - 1) + 3 * ((j) - 1) + 9 * ((k) - 1)]
  for (int i = 1; i <= 3; i++)
    for (int j = 1; j <= 3; j++) {
      jac[i][j] = u[ijac + i + 3 * j - 3];
      for (int k = 1; k <= 3; k++) hes[i][j][k] = u[ihes + i + 3 * j - 3 + 9 * k - 9];
    }
  for (int i = 1; i <= 3; i++) {
    ddt[0 + i] = u[ivel + i];
    ddt[ivel + i] = inv * (vel[i - 1] - u[ivel + i]);
  }
  for (int i = 1; i <= 3; i++)
    for (int j = 1; j <= 3; j++) {
      ddt[ijac + i + 3 * j - 3] = u[ijrt + i + 3 * j - 3];
      double sumj = jac[1][j] * DV(i, 1) + jac[2][j] * DV(i, 2) + jac[3][j] * DV(i, 3);
      ddt[ijrt + i + 3 * j - 3] = inv * (sumj - u[ijrt + i + 3 * j - 3]);
    }
  for (int i = 1; i <= 3; i++)
    for (int j = 1; j <= 3; j++)
      for (int k = 1; k <= 3; k++) {
        ddt[ihes + i + 3 * j - 3 + 9 * k - 9] = u[ihrt + i + 3 * j - 3 + 9 * k - 9];
        double sumh = hes[1][j][k] * DV(i, 1) + hes[2][j][k] * DV(i, 2) + hes[3][j][k] * DV(i, 3) +
                      jac[1][j] * jac[1][k] * DDV(i, 1, 1) + jac[2][j] * jac[1][k] * DDV(i, 2, 1) +
                      jac[3][j] * jac[1][k] * DDV(i, 3, 1) + jac[1][j] * jac[2][k] * DDV(i, 1, 2) +
                      jac[1][j] * jac[3][k] * DDV(i, 1, 3) + jac[2][j] * jac[2][k] * DDV(i, 2, 2) +
                      jac[2][j] * jac[3][k] * DDV(i, 2, 3) + jac[3][j] * jac[2][k] * DDV(i, 3, 2) +
                      jac[3][j] * jac[3][k] * DDV(i, 3, 3);
        ddt[ihrt + i + 3 * j - 3 + 9 * k - 9] = inv * (sumh - u[ihrt + i + 3 * j - 3 + 9 * k - 9]);
      }
#undef DV
#undef DDV
}

int orc_part_advance(orc_t *o, int field_mode, int scheme, int irk, double st, double partdt,
                     int nvar, long np, double *u, double *ddt1, double *ddt2, int *alive) {
  int bad = 0;
  for (long p = 0; p < np; p++) {
    if (!alive[p]) continue;
    double *up = u + 117 * p - 1, *d1 = ddt1 + 117 * p - 1, *d2 = ddt2 + 117 * p - 1;  /* 1-based */
    double vel[3], dvel[9], ddvel[27];
    if (orc_part_interp(o, field_mode, up + 1, vel, dvel, ddvel)) { bad++; continue; }
    part_rates(st, up, vel, dvel, ddvel, d1);
    if (scheme == 0) {                                 /* part_tools.F90:926-945 */
      for (int i = 1; i <= 78; i++) up[i] = up[i] + d1[i] * partdt;
      for (int i = 1; i <= 3; i++) up[i + 78] = vel[i - 1];
      for (int i = 1; i <= 9; i++) up[78 + 3 + i] = dvel[i - 1];
      for (int i = 1; i <= 27; i++) up[78 + 3 + 9 + i] = ddvel[i - 1];
    } else {                                           /* dertest/part_tools.F90:902-906 */
      for (int i = 1; i <= nvar; i++) {
        up[i] = up[i] + o->gam[irk] * d1[i] * partdt + o->rom[irk] * d2[i] * partdt;
        d2[i] = d1[i];
      }
    }
  }
  for (long p = 0; p < np; p++) {                     /* :962-1025, single domain */
    if (!alive[p]) continue;
    double *up = u + 117 * p;
    double x = up[0], y = up[1], z = up[2];
    double r = sqrt(x * x + y * y);
    if (z > o->zmax || z < 0.0 || r > o->rmax || r < 0.0) {
      alive[p] = 0;
      memset(up, 0, sizeof(double) * 117);
      memset(ddt1 + 117 * p, 0, sizeof(double) * 117);
      memset(ddt2 + 117 * p, 0, sizeof(double) * 117);
    }
  }
  return bad ? -1 : 0;
}

/* ---- trilinear option (an alternative interpolation; not in the reference's built code) ---- */
static double tri_comp(orc_t *o, int ivar, double theta, double r, double z) {
  double th0 = (ivar == 1) ? 0.0 : 0.5 * o->dtheta;
  double r0 = (ivar == 2) ? 0.0 : 0.5 * o->dr;
  double z0 = (ivar == 3) ? 0.0 : 0.5 * o->dz;
  double st = (theta - th0) / o->dtheta, sr = (r - r0) / o->dr, sz = (z - z0) / o->dz;
  double ft = floor(st), fr = floor(sr), fz = floor(sz);
  double wt = st - ft, wr = sr - fr, wz = sz - fz;
  int i0 = (int)ft, j0 = (int)fr + 1, k0 = (int)fz + 1;   /* array index of the lower node */
  if (j0 < 0) { j0 = 0; wr = 0.0; }
  if (j0 > o->nr) { j0 = o->nr; wr = 1.0; }
  if (k0 < 0) { k0 = 0; wz = 0.0; }
  if (k0 > o->nz + 1) { k0 = o->nz + 1; wz = 1.0; }
  double s = 0.0;
  for (int c = 0; c <= 1; c++)
    for (int b = 0; b <= 1; b++)
      for (int a = 0; a <= 1; a++) {
        int ii = ((i0 + a) % o->nt + o->nt) % o->nt + 1;
        int jj = j0 + b, kk = k0 + c;
        if (kk > o->nz + 2) kk = o->nz + 2;
        double w = (a ? wt : 1.0 - wt) * (b ? wr : 1.0 - wr) * (c ? wz : 1.0 - wz);
        double v;
        if (ivar == 1) v = F3(o->q1, ii, jj, kk);
        else if (ivar == 2) v = (o->rc[jj] > 1e-12) ? F3(o->q2, ii, jj, kk) / o->rc[jj] : 0.0;
        else v = F3(o->q3, ii, jj, kk);
        s += w * v;
      }
  return s;
}

int orc_part_trilinear(orc_t *o, const double xp[3], double vel[3]) {
  double x = xp[0], y = xp[1], z = xp[2];
  double r = sqrt(x * x + y * y);
  double theta = atan2(y, x);
  if (theta < 0) theta += 2.0 * M_PI;
  double ut = tri_comp(o, 1, theta, r, z), ur = tri_comp(o, 2, theta, r, z), uz = tri_comp(o, 3, theta, r, z);
  double sn = (r > 0) ? y / r : 0.0, cs = (r > 0) ? x / r : 1.0;
  vel[0] = -sn * ut + cs * ur;
  vel[1] = cs * ut + sn * ur;
  vel[2] = uz;
  return 0;
}

int orc_part_rk3_trilinear(orc_t *o, long np, double st, int irk, double dt, double *xv,
                           double *rate2, int *alive) {
  for (long p = 0; p < np; p++) {
    if (!alive[p]) continue;
    double *s = xv + 6 * p, *r2 = rate2 + 6 * p, f[6], u[3];
    orc_part_trilinear(o, s, u);
    for (int i = 0; i < 3; i++) {
      if (st > 0.0) { f[i] = s[3 + i]; f[3 + i] = (u[i] - s[3 + i]) / st; }
      else { f[i] = u[i]; f[3 + i] = 0.0; }
    }
    for (int i = 0; i < 6; i++) {
      s[i] = s[i] + o->gam[irk] * dt * f[i] + o->rom[irk] * dt * r2[i];
      r2[i] = f[i];
    }
    if (st <= 0.0) for (int i = 0; i < 3; i++) s[3 + i] = u[i];
    double r = sqrt(s[0] * s[0] + s[1] * s[1]);
    if (s[2] > o->zmax || s[2] < 0.0 || r > o->rmax) alive[p] = 0;
  }
  return 0;
}
