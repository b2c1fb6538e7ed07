/* vrh_learn.h -- C ABI of the per-Fourier-mode (r, z) Helmholtz subsystem of
 * libvrh: the pressure operator of one azimuthal mode on a uniform r grid and
 * a tanh-stretched z grid, its direct solve, plain Krylov solvers, and a
 * learned warm start / preconditioner (a 3x3 convolutional network with
 * BF16 MFMA inference and training kernels, FP32 accumulation).
 *
 * No reference counterpart: the reference solves its pressure equation only
 * on a uniform z grid by FFT/DCT and radial tridiagonal solves
 * (poisson_solve.F90:140-179, 270-323).  The operator here is that equation's
 * (r, z) part for one mode m in symmetric finite-volume form,
 *     S_m p = V f,
 * with the radial fluxes of poisson.F90:368-388, the exact theta eigenvalue
 * m^2/r^2 ('wvnex', poisson_solve.F90:286-289) and zero-flux axis, wall and
 * z ends; V is the cell volume r dr dz.  For m = 0 the volume-weighted mean of
 * f is removed and the solution has zero volume-weighted mean.
 *
 * A handle holds `nslots` independent systems (slots), each with its own m.
 * Host arrays are C order (nslots, nz, nr), r fastest.  Every call returns 0
 * or a negative code; vrh_hz_last_error gives the message.
 */
#ifndef VRH_LEARN_H
#define VRH_LEARN_H

#include <stddef.h>

#include "vrh.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vrh_hz_s vrh_hz_t;

typedef struct {
  int nr, nz, nslots;      /* nr a multiple of 16 * pool, nz of pool */
  double rmax, zmax;
  double stretch;          /* z_k = zmax (1 + tanh(s (k/nz - 1)) / tanh s); 0: uniform */
  int pool;                /* the network works on blocks of pool x pool cells (0 or 1: cells) */
} vrh_hz_grid;

enum { VRH_HZ_PCG = 0, VRH_HZ_BICGSTAB = 1, VRH_HZ_LEARNED = 2, VRH_HZ_LEARNED_FCG = 3 };

int vrh_hz_create(const vrh_hz_grid *grid, vrh_hz_t **out);
int vrh_hz_destroy(vrh_hz_t *h);
int vrh_hz_last_error(vrh_hz_t *h, char *buf, size_t len);
/* Azimuthal mode of each slot (nslots values, 0..). */
int vrh_hz_set_modes(vrh_hz_t *h, const int *m);
/* Mesh: z faces (nz+1), z centres (nz), r centres (nr); any may be NULL. */
int vrh_hz_mesh(vrh_hz_t *h, double *zf, double *zm, double *rm);
/* y = S_m x per slot. */
int vrh_hz_apply(vrh_hz_t *h, const double *x, double *y);
/* Direct solve per slot: generalised eigenvectors of the z operator (dense
 * transforms) and one radial tridiagonal solve per (slot, z eigenvalue). */
int vrh_hz_direct(vrh_hz_t *h, const double *f, double *p);
/* Iterative solve of every slot to ||V f - S p|| <= tol ||V f||:
 *   VRH_HZ_PCG       Jacobi-preconditioned CG, zero start;
 *   VRH_HZ_BICGSTAB  Jacobi-preconditioned BiCGSTAB, zero start;
 *   VRH_HZ_LEARNED   plain PCG started from the network's warm start
 *                    M(V f) = beta D^-1 V f + gamma c, c = scale N(V f)
 *                    prolonged bilinearly to the cells, with (beta, gamma)
 *                    the energy-norm optimal (Galerkin) pair -- so the
 *                    start's energy error never exceeds that of the best
 *                    multiple of the Jacobi estimate, nor that of a zero
 *                    start, whatever the network returns;
 *   VRH_HZ_LEARNED_FCG  the same start, then flexible CG (Polak-Ribiere)
 *                    preconditioned by the same Galerkin map M applied to
 *                    each residual; a slot that breaks down or does not
 *                    converge within maxit falls back to plain PCG from where
 *                    it is.
 * The true residual of the returned p is recomputed and reported per slot in
 * relres; iters receives the iterations per slot (fallback included), and
 * ms (optional) the solve time in milliseconds (GPU events). */
int vrh_hz_solve(vrh_hz_t *h, int method, double tol, int maxit, const double *f, double *p, int *iters,
                 double *relres, double *ms);

/* ---- network: 4 layers of 3x3 convolutions, dilations 1, 2, 4, 8, widths
 * 16 (4 input features) -> 32 -> 32 -> 32 -> 16 (channel 0 is the output),
 * ReLU between layers; parameters as fp32, per layer W (9, cout, cin) then
 * b (cout). */
int vrh_hz_net_size(vrh_hz_t *h);
int vrh_hz_net_init(vrh_hz_t *h, unsigned long long seed);
int vrh_hz_net_get(vrh_hz_t *h, float *params);
int vrh_hz_net_set(vrh_hz_t *h, const float *params);
/* Network output channel on the network grid (nslots, nz/pool, nr/pool) for
 * the features of f, and the output scale rms(f) R^2 per slot (R = nr dr):
 * the correction c is scale * out, prolonged bilinearly (in index space,
 * from the block centres; constant beyond the outermost centres). */
int vrh_hz_net_forward(vrh_hz_t *h, const double *f, float *out, double *scale);
/* Gradient of sum(gout * out) with respect to the parameters, for the
 * features of f (gout and out on the network grid; same packing as the
 * parameters). */
int vrh_hz_net_grad(vrh_hz_t *h, const double *f, const float *gout, float *grads);
/* Right-hand sides the trainer draws: family 0 Gaussian ring cores
 * (4 per sample, radius 0.1..0.3 of the r extent scale), family 1 white noise. */
int vrh_hz_sample(vrh_hz_t *h, unsigned long long seed, int family, double *f);
/* Train on nsamples (rhs, p) pairs generated on the GPU (rhs from both
 * families, random m in 0..63 per slot, p from the direct solve), nslots per
 * batch, Adam with learning rate lr, energy-norm loss e^T S e / (p^T V f)
 * of the warm start's error e = M(V f) - p (features taken from the rhs
 * divided by the cell volume, as the solver sees residuals).  loss
 * (optional) receives the mean loss of each batch. */
int vrh_hz_train(vrh_hz_t *h, int nsamples, double lr, unsigned long long seed, float *loss);
/* The same training on the residuals a preconditioner meets: each batch's
 * right-hand sides are replaced by their residuals after k plain PCG
 * iterations from zero, k cycling through 0, 1, 4, 16, ... <= max_cg over the
 * batches; the target is the exact solve of that residual. */
int vrh_hz_train_residuals(vrh_hz_t *h, int nsamples, double lr, unsigned long long seed, int max_cg,
                           float *loss);

/* ---- in the time step.  The pressure solve of a stepper handle h with an
 * axial mesh equal to hz's tanh mesh (vrh_set_axial_mesh), the uniform
 * radial mesh and the exact theta eigenvalue (VRH_THETA_EXACT_M2) solves the
 * theta positions of the modes m >= mmin (pos >= 2 mmin - 1; mmin >= 1, and
 * >= 3 in the mixed path, whose m = 0..2 stay in the FP64 direct chain) by
 * hz's iterative `method` to relative residual tol instead of the direct z
 * transform and radial solve (poisson_solve.F90:140-179, 270-323 on the
 * uniform grid).  hz needs one slot per such position (nt - 2 mmin + 1);
 * this call sets their modes.  A slot that misses tol within maxit makes the
 * step fail (code -4).  hz = NULL switches it off, and so does a later
 * vrh_set_radial_mesh or vrh_set_axial_mesh.  hz must stay alive (and its
 * modes unchanged) while h uses it; vrh_autotune leaves the mode solve out
 * of its timing.  No reference counterpart (the reference has no
 * stretched z grid). */
int vrh_set_mode_solver(vrh_t *h, vrh_hz_t *hz, int method, int mmin, double tol, int maxit);
/* Krylov iterations (summed over slots) and mode solves since the last
 * vrh_set_mode_solver. */
int vrh_mode_solver_stats(vrh_t *h, long long *iterations, long long *solves);

#ifdef __cplusplus
}
#endif
#endif
