/* vring_oracle.h -- FP64 CPU oracle for the vring RK3 substep and particle step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a CPU restatement of the reference Fortran
 * algorithm, used as the checker for the HIP product path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product library (libvrh) never links or calls it.
 *
 * Array layouts follow the reference's Fortran column-major declarations:
 *   haloed 3-D fields  a(nt, 0:nr+2, 0:nz+2)   (solu_toolsII.F90:88-111)
 *   prg(nt, nr, nz, 3)                          (solu_toolsII.F90:113)
 *   rhs/sol(nt, nr, nz), dph(nt, 0:nr+1, 0:nz+1) (solp_tools.F90:53-57)
 *   q1s/q2s/q3s(nt, 0:nr+1), dq3no(nt, 0:nr+2)   (solu_toolsII.F90:108,116-118)
 *   corr*, csort, tconv (nt, nr+1)              (solu_toolsII.F90:191-199)
 */
#ifndef VRING_ORACLE_H
#define VRING_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc_t;

enum orc_phase {
  ORC_UCOEFF = 0, ORC_COESYS = 1, ORC_HDNBC = 2, ORC_HDNL1 = 3, ORC_HDNL2 = 4,
  ORC_HDNL3 = 5, ORC_STEP = 6, ORC_DIVQ = 7, ORC_LAPLACE = 8, ORC_UPSOL = 9,
  ORC_UPPRES = 10
};

/* wvnex=1: theta eigenvalue m^2 as poisson.F90:386 ('wvnex');
 * wvnex=0: the discretely consistent (2/dtheta^2)(1-cos(2 pi m/nt)).       */
orc_t *orc_create(int nt, int nr, int nz, double rmax, double zmax, double Re,
                  double dt, double slip, int wvnex);
void orc_destroy(orc_t *o);
/* Pointer to an internal array by name; *count receives its element count. */
double *orc_field(orc_t *o, const char *name, long *count);
void orc_set_time(orc_t *o, double ntime);
/* non-uniform radial mesh (rm, rc, dmr, dcr over j = 0..nr+1) */
void orc_set_radial_mesh(orc_t *o, const double *rm, const double *rc, const double *dmr, const double *dcr);
/* Non-uniform axial mesh: zm, zc, dmz, dcz over k = 0..nz+1 (no reference
 * counterpart; the radial hooks carried over to z). */
void orc_set_axial_mesh(orc_t *o, const double *zm, const double *zc, const double *dmz, const double *dcz);
double orc_get_time(const orc_t *o);
int orc_phase(orc_t *o, int phase, int irk);
int orc_substep(orc_t *o, int irk);
int orc_rk3_step(orc_t *o);          /* loop.F90:53 then 63-131 (no particles) */
/* Standalone Poisson solve of rhs(nt,nr,nz) into sol(nt,nr,nz). */
int orc_poisson(orc_t *o, const double *rhs, double *sol);
/* Discrete divergence of (q1,q2,q3) at cell centres, (nt,nr,nz). */
void orc_divergence(orc_t *o, const double *q1, const double *q2, const double *q3, double *div);

/* ---------------- particles (part_tools.F90) ---------------- */
/* Per-particle state rows u[117] (part_tools.F90:241-252), rate rows ddt[2][117].
 * field_mode 0: sample q1,q2,q3 from the oracle's fields;
 * field_mode 1: the analytic override of dertest/part_tools.F90:580-582
 *               (u_theta = 0, u_r = rc^2, u_z = 1).
 * scheme 0: forward Euler on states 1..78 (part_tools.F90:926-930);
 * scheme 1: low-storage RK3 on states 1..nvar (dertest/part_tools.F90:903).  */
int orc_part_interp(orc_t *o, int field_mode, const double x[3], double vel[3],
                    double dvel[9], double ddvel[27]);
/* One particle sub-step for np particles (rows of u: 117 values each, rates
 * ddt1/ddt2: 117 each).  alive[p]=0 rows are skipped; particles leaving the
 * domain are removed as in part_tools.F90:962-1025 (alive=0, rows zeroed).  */
int orc_part_advance(orc_t *o, int field_mode, int scheme, int irk, double st,
                     double partdt, int nvar, long np, double *u, double *ddt1,
                     double *ddt2, int *alive);
/* Trilinear velocity sample (u_theta,u_r,u_z) -> Cartesian, staggered grids. */
int orc_part_trilinear(orc_t *o, const double x[3], double vel[3]);
/* Inertial particle (st>0) or tracer (st==0) RK3 substep irk with trilinear
 * interpolation; state xv[6] = (x,y,z,vx,vy,vz), previous rate rate2[6].   */
int orc_part_rk3_trilinear(orc_t *o, long np, double st, int irk, double dt, double *xv,
                           double *rate2, int *alive);
/* pseudo-inverse of the 27-point quadratic design matrix (9 x 27, row-major) */
void orc_part_mint(orc_t *o, double *mint);

#ifdef __cplusplus
}
#endif
#endif
