/* vrh.h -- C ABI of libvrh, the MI355X (gfx950) implementation of the vring
 * RK3 fractional-step substep and Lagrangian particle step.
 *
 * The reference has no plugin or FFI layer: its hot path is the sequence of
 * Fortran calls in loop.F90:63-131, each taking `type(run) :: var` with the
 * whole state.  These entry points replace that sequence for a Fortran driver
 * (iso_c_binding) or any C caller.  Conventions:
 *  - every function returns 0 on success, a negative code on error; the
 *    message is available from vrh_last_error;
 *  - host arrays use the reference's Fortran column-major layouts:
 *      haloed fields  q(ntheta, 0:nr+2, 0:nz+2)      solu_toolsII.F90:88-111
 *      prg(ntheta, nr, nz, 3)                         solu_toolsII.F90:113
 *      Poisson rhs/phi(ntheta, nr, nz)                solp_tools.F90:53-55
 *  - the library never keeps host pointers; all device memory is owned by
 *    the handle; one HIP stream per handle; calls that return host data are
 *    synchronous, the others are asynchronous on the handle's stream.
 */
#ifndef VRH_H
#define VRH_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vrh_s vrh_t;

/* Mesh: mesh_tools.F90:86-93, 217-350 (one whole-domain layout). */
typedef struct {
  int ntheta, nr, nz;
  double rmax, zmax;
} vrh_grid;

/* Parameters: para_tools.F90:68-86 (dt, Re, slip) and the theta eigenvalue
 * used by the Poisson solve (poisson.F90:386; poisson_solve.F90:286-292). */
typedef struct {
  double dt, Re, slip;
  int theta_eigen; /* VRH_THETA_EXACT_M2 (reference 'wvnex') or VRH_THETA_CONSISTENT */
} vrh_params;

enum { VRH_THETA_EXACT_M2 = 0, VRH_THETA_CONSISTENT = 1 };
enum { VRH_FP64 = 0, VRH_MIXED = 1 };

/* Substep phases for vrh_phase (loop.F90:71-127 call order). */
enum {
  VRH_PH_HDNBC = 2,    /* momentum_terms_swirl.F90:177 hdnbc                   */
  VRH_PH_EXPLICIT = 3, /* hdnl1-3 (momentum_terms_swirl.F90:15-174) + step's
                          explicit right-hand sides (step.F90:32-165) + the
                          periodic theta sweeps solqi (step.F90:212-263)       */
  VRH_PH_ADI = 6,      /* step's solqj/solqk sweeps (step.F90:266-485)         */
  VRH_PH_POISSON = 8,  /* divq + laplace_solve (divergence.F90:16;
                          poisson_solve.F90:13)                                */
  VRH_PH_PROJECT = 9   /* upsol + uppres (divergence.F90:69, 239)              */
};

/* Lifecycle.  Replaces create_mesh_a/alct_solu/alct_solp/prep_poisson_solver
 * (main.F90:49-68): allocates and zero-initialises every field.  Grids:
 * 3 <= ntheta <= 256 (the explicit kernel holds whole theta rows in LDS),
 * 4 <= nr <= 512 and 4 <= nz <= 1151 (line solves of at most 16 segments of
 * at most 72 rows); power-of-two ntheta and nz take the FFT paths, other
 * sizes direct transforms.  Returns -1 for grids outside these limits. */
int vrh_create(const vrh_grid *grid, const vrh_params *params, int precision, vrh_t **out);
int vrh_destroy(vrh_t *h);
int vrh_last_error(vrh_t *h, char *buf, size_t len);

/* Whole-field transfers (restart_load/restart_save, run.F90:390-539, and
 * the analysis output).  Any pointer may be NULL to skip that field. */
int vrh_upload_fields(vrh_t *h, const double *q1, const double *q2, const double *q3,
                      const double *pres);
int vrh_download_fields(vrh_t *h, double *q1, double *q2, double *q3, double *pres);
/* Any named state array, as doubles (work arrays of the mixed path are
 * converted):
 *   q1 q2 q3 dq1 dq2 dq3 ru1 ru2 ru3 qrhs1 qrhs2 qrhs3 pres dph
 *       haloed (ntheta, 0:nr+2, 0:nz+2)   ntheta*(nr+3)*(nz+3) values
 *   prg                                   (ntheta, nr, nz, 3): 3*ntheta*nr*nz
 *   prg1 prg2 prg3 spec                   (ntheta, nr, nz): ntheta*nr*nz
 *   csort dq3no corr3n                    (ntheta, 0:nr+2): ntheta*(nr+3)
 * (all three take the reference's dq3no shape (nthetac, 0:nrc+1) with
 * nrc = nr + 1, solu_toolsII.F90:108; its csort and corr3n are
 * (nthetac, nrc), solu_toolsII.F90:197-198, i.e. rows 1..nr+1 here).
 * The host buffer must hold vrh_state_size(name) values. */
int vrh_upload_state(vrh_t *h, const char *name, const double *host);
int vrh_download_state(vrh_t *h, const char *name, double *host);
/* Number of doubles vrh_upload_state/vrh_download_state move for `name` on
 * `grid` (host only: no device call).  -1 for an unknown name. */
int vrh_state_size(const vrh_grid *grid, const char *name, long *n);

/* Simulation time (par%ntime).  vrh_step advances it as loop.F90:53 does. */
int vrh_set_time(vrh_t *h, double ntime);
int vrh_get_time(vrh_t *h, double *ntime);

/* Non-uniform radial mesh through the reference's radial metric factors.
 * mesh_tools.F90:325-346 sets rm(j) = (j - 1/2) dr, rc(j) = (j - 1) dr and
 * dmr = dcr = 1; momentum_coef.F90:24-69, divergence.F90:39, 132, 327 and
 * momentum_terms_swirl.F90:248 read them.  Four arrays over j = 0..nr+1
 * (nr + 2 values): cell centres rm, faces rc (rc(1) = 0 on the axis), and
 * dr/dxi of the mapping r(xi) at the centres (dmr) and faces (dcr), xi
 * uniform with spacing rmax/nr (the derivatives keep dx2 = nr/rmax).  The
 * radial Poisson rows become minus divq's radial difference of uppres'
 * radial gradient, which is poisson.F90:379-381 on the uniform mesh.
 * vrh_analysis reads rm from this mesh as run.F90:300-345 reads msh%rm.
 * Particles need the uniform mesh; call before creating particles. */
int vrh_set_radial_mesh(vrh_t *h, const double *rm, const double *rc, const double *dmr, const double *dcr);

/* Non-uniform axial mesh: the z counterpart of vrh_set_radial_mesh (the
 * reference's z mesh is uniform, mesh_tools.F90:321-342, and reads dx3
 * wherever a z difference is taken: momentum_terms_swirl.F90, divergence.F90,
 * momentum_coef.F90:314-465).  Four arrays over k = 0..nz+1 (nz + 2
 * values): cell centres zm, faces zc (zc(1) = 0, the inlet), and dz/dzeta at
 * the centres (dmz) and faces (dcz), zeta uniform with spacing zmax/nz.
 * Every z difference then carries its metric, the z sweeps use the
 * second-difference weights of the mesh, and the pressure solve replaces the
 * z cosine transform by the generalised eigenvectors of its z operator
 * (dense nz x nz transforms, GEMMs on the device), which on the uniform
 * mesh are the same cosines.  No reference counterpart.  Particles need the
 * uniform axial mesh; a handle with an axial mesh steps without a HIP graph
 * and cannot join an ensemble. */
int vrh_set_axial_mesh(vrh_t *h, const double *zm, const double *zc, const double *dmz, const double *dcz);

/* One RK substep irk=1..3 (loop.F90:71-127) of the step that has reached
 * the current time: the reference advances ntime before its substeps
 * (loop.F90:53), so a driver that keeps its own RK loop calls
 * vrh_set_time(t + dt), then for irk = 1, 2, 3 vrh_particles_step(h, irk)
 * (where loop.F90:67 calls iter_part) and vrh_substep(h, irk).  The step's
 * scalars (inlet increment, mass correction) come from the same device
 * computation as vrh_step's, so the result equals one vrh_step from t bit
 * for bit. */
int vrh_substep(vrh_t *h, int irk);
/* nsteps full steps: ntime += dt, then substeps 1,2,3 (loop.F90:53-131). */
int vrh_step(vrh_t *h, int nsteps);
/* vrh_step without the final wait: the steps are queued on the handle's
 * stream and the call returns; vrh_synchronize waits and reports the crash
 * flag.  The host may work meanwhile (e.g. write the files of a snapshot);
 * calls that read or replace device state wait on the stream themselves.
 * No reference counterpart (the reference steps synchronously). */
int vrh_step_async(vrh_t *h, int nsteps);
/* Asynchronous field export (restart_save and the
 * Tecplot writers, run.F90:238-466, read these fields): _begin copies q1,
 * q2, q3, pres on the device (stream-ordered after queued steps) and drains
 * the copy to pinned host memory on a second stream, so stepping may go on
 * at once; _wait blocks until the copy has landed and hands the haloed
 * fields (any pointer may be NULL) and their time out.  One snapshot may be
 * pending per handle. */
int vrh_snapshot_begin(vrh_t *h);
int vrh_snapshot_wait(vrh_t *h, double *q1, double *q2, double *q3, double *pres, double *ntime);
/* vrh_step replays one RK3 step captured as a HIP graph (captured on first
 * use, recaptured after vrh_autotune, vrh_set_tuning, vrh_set_radial_mesh or a
 * particle change; not used while the kernel timer is on or with an axial
 * mesh).  on = 0 launches the kernels one by one instead.  No
 * reference counterpart. */
int vrh_set_graph(vrh_t *h, int on);
/* One phase of substep irk (parity tests and profiling). */
int vrh_phase(vrh_t *h, int phase, int irk);
/* Standalone Poisson solve, host rhs -> host phi, (ntheta,nr,nz) each
 * (solve_lap_tripar, poisson_solve.F90:140-179). */
int vrh_poisson_solve(vrh_t *h, const double *rhs, double *phi);
/* Wait for the handle's stream; returns the deferred "Flow field crash"
 * error (divergence.F90:396-399) if a NaN was met. */
int vrh_synchronize(vrh_t *h);
/* Per-kernel timing with HIP events on the handle's stream: enable (resets
 * the totals), run work, then read kernel `idx` (0..count-1): its name,
 * total milliseconds and launch count.  vrh_ktimer_read returns the number
 * of kernel kinds when idx < 0. */
int vrh_ktimer_enable(vrh_t *h, int on);
int vrh_ktimer_read(vrh_t *h, int idx, char *name, size_t len, double *total_ms, long *count);
/* Autotuner: times every candidate segment length of the six partitioned
 * line sweeps, the FFT/DCT tile sizes and the four forms of the explicit
 * kernel (one z plane per block, or z marching) with HIP events on the live
 * buffers and keeps the fastest.  Call between steps.  choice[9] (optional)
 * receives M for sweeps r1 r3 r2 z1 z3 z2, theta FFT lines/block, z DCT
 * lines/block and the explicit-kernel form (0 plane, 1 marching, 2 marching
 * with one wavefront per r row within 64 KB of LDS per block, 3 the same
 * with the rows per block that keep the most waves resident within the
 * device's LDS per block and per CU). */
int vrh_autotune(vrh_t *h, int reps, int *choice);
/* Apply a tuning returned by vrh_autotune (9 entries; 0 keeps a segment
 * length or tile size, -1 selects the default explicit form). */
int vrh_set_tuning(vrh_t *h, const int *choice);
/* Device address of a named field (for callers that keep data in HBM). */
int vrh_device_ptr(vrh_t *h, const char *name, void **ptr);

/* Vortex-core statistics of analysis (run.F90:238-388) as a device reduction
 * over the current fields (synchronous): out[0] max azimuthal vorticity,
 * out[1] max swirl (the window 0.1 < z/zmax < 0.9, 0.05 < r/rmax < 0.8,
 * j != 1), out[2..9] the sums stat(1:8) = xvort, yvort, vvort, uvort, xswrl,
 * yswrl, vswrl, uswrl of the positive vorticity and swirl. */
int vrh_analysis(vrh_t *h, double *out);
/* Allocation-free variant of vrh_analysis.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same ten numbers, bit for bit, are written to dev_out
 * (device memory, 10 doubles) on the handle's stream, and the call returns
 * without waiting (vrh_synchronize or the stream's order makes them
 * visible).  It calls no allocating routine: the reduction's block partials
 * live in a fixed device array of 1024 x 10 doubles (80 KB) per GPU, held
 * until the process ends.  That size does not depend on the grid, so no
 * grid is refused for size.  Every call on one GPU shares the array:
 * calls on that GPU whose work overlaps (any handles, any streams) race. */
int vrh_analysis_noalloc(vrh_t *h, double *dev_out);

/* Host-side initial condition (no reference counterpart: the reference loads
 * RESTART, run.F90:469-539).  Gaussian-core vortex rings from a discrete
 * Stokes streamfunction; rings[4*r..] = (R, z0, sigma, circulation); the ring
 * radius is R + pert_amp * sum_{m=pert_m0..pert_m1} cos(m theta + phase[m-pert_m0]).
 * Writes haloed q1,q2,q3 (divergence-free to rounding on the staggered grid). */
int vrh_ring_init(const vrh_grid *grid, int nrings, const double *rings, double pert_amp,
                  int pert_m0, int pert_m1, const double *phase, double *q1, double *q2,
                  double *q3);
/* Fixed time step from a CFL number on a haloed host field (with an inlet
 * jet speed floor `ujet`). */
int vrh_cfl_dt(const vrh_grid *grid, const double *q1, const double *q2, const double *q3,
               double cfl, double ujet, double *dt);

/* ---- Ensembles of independent rings ----
 * No reference counterpart (the reference runs one ring per MPI job).  The
 * members (handles of any grid size, Re, precision and particle load) are
 * advanced together: one RK3 step of every member (or K steps, see
 * vrh_ensemble_set_batch) is one HIP graph launch, the members' kernels
 * running concurrently.  Members must not be used
 * elsewhere during vrh_ensemble_step; their times advance as with vrh_step.
 * The ensemble keeps the handle pointers; destroy it before its members. */
typedef struct vrh_ens_s vrh_ens_t;
int vrh_ensemble_create(vrh_t *const *members, int n, vrh_ens_t **out);
int vrh_ensemble_step(vrh_ens_t *e, int nsteps);
/* Steps per graph launch (default 1): with K > 1 vrh_ensemble_step advances
 * every member K steps per launch of one graph (K whole steps of every ring
 * captured in sequence), then single steps for a remainder.  Results are
 * those of stepping each member alone. */
int vrh_ensemble_set_batch(vrh_ens_t *e, int steps_per_launch);
int vrh_ensemble_last_error(vrh_ens_t *e, char *buf, size_t len);
int vrh_ensemble_destroy(vrh_ens_t *e);

/* ---- Lagrangian particles (part_tools.F90) ---- */
enum { VRH_PART_LSQ_EULER = 0, VRH_PART_LSQ_RK3 = 1, VRH_PART_TRILINEAR_RK3 = 2 };
/* Create nbatch particle batches (init_part, part_tools.F90:228-359): batch b
 * holds up to cap[b] particles with Stokes number st[b] and nsub[b] sub-steps
 * per call (part_tools.F90:493-502).  State rows hold 117 values
 * (part_tools.F90:241-252) for the LSQ schemes, 6 for TRILINEAR_RK3.
 * Returns -1 for a bad argument or batch (nothing allocated) and -3 when a
 * device allocation fails (what was allocated is freed); either way the
 * handle keeps no batches and the call may be made again. */
int vrh_particles_create(vrh_t *h, int nbatch, const long *cap, const double *st,
                         const int *nsub, int scheme);
/* Host rows (n x nvar, row-major by particle) and alive flags. */
int vrh_particles_upload(vrh_t *h, int batch, long n, const double *u, const int *alive);
int vrh_particles_download(vrh_t *h, int batch, long n, double *u, int *alive);
/* Advance all batches at RK substep irk (iter_part, part_tools.F90:374-1025;
 * the reference calls it at irk==1 only, loop.F90:67). */
int vrh_particles_step(vrh_t *h, int irk);
/* Pack the live particles of a batch to the front, order preserved (the
 * reference's free list, remv_part/injc_part, part_tools.F90:141-225). */
int vrh_particles_compact(vrh_t *h, int batch);
/* Sort a batch's particles in device memory by their cells (no reference
 * counterpart: part_tools.F90 keeps its slots in injection order), so that
 * the interpolation gathers of neighbouring particles meet the same cache
 * lines.  Upload and compaction sort as well; this call refreshes the order
 * on long runs between them.  The order the API shows (upload, download,
 * compaction, injection) does not change. */
int vrh_particles_sort(vrh_t *h, int batch);
/* Live particles of a batch (synchronous). */
int vrh_particles_count(vrh_t *h, int batch, long *n);
/* Inlet injection (inlt_part, part_tools.F90:95-138), after a compaction:
 * for every inlet cell (rc(ir) < r0, first z plane, theta order per row)
 * inflow(ir) += c a q3 dt inj / nprcl with a = dtheta (rc(ir+1)^2 - rc(ir)^2),
 * c = 1e6 0.02^3, nprcl = 0.001; floor(inflow) particles are appended,
 * uniform in the cell, velocity (0, 0, q3), identity Jacobian.  Positions
 * come from a counter-based stream (splitmix64 of `seed`, the call number
 * and the cell), not Fortran's random_number.  Particles beyond the capacity
 * are dropped.  ninj (optional, synchronous) receives the number injected. */
int vrh_particles_inject(vrh_t *h, int batch, double r0, int inj, unsigned long long seed, long *ninj);

#ifdef __cplusplus
}
#endif
#endif
