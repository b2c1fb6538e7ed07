// vrh_learn.hip -- per-Fourier-mode (r, z) Helmholtz subsystem (include/vrh_learn.h):
// the pressure operator of one azimuthal mode on a uniform r grid and a
// tanh-stretched z grid, its direct solve, Jacobi-preconditioned CG and
// BiCGSTAB, and a convolutional network (BF16 MFMA forward and backward
// kernels, FP32 accumulation) used as warm start and flexible-CG
// preconditioner, trained on the GPU from pairs the direct solver generates.
//
// Device layout: slot-major, r fastest: cell c = (b * nz + k) * nr + j.
// Network tensors are NHWC in bf16: t[c * C + channel].
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "vrh_learn.h"

namespace {

typedef short bf4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16_t;

constexpr int kRedBlocks = 64;   // reduction blocks per slot
constexpr int kNsc = 16;         // scalars per slot
constexpr int kWidth = 32, kCin0 = 16, kCout4 = 16;
constexpr int kDil[4] = {1, 2, 4, 8};
constexpr int kLayerCin[4] = {kCin0, kWidth, kWidth, kWidth};
constexpr int kLayerCout[4] = {kWidth, kWidth, kWidth, kCout4};
constexpr int kDwRows = 16;      // z rows per block of the weight-gradient kernel

// scalar slots
enum { SC_RZ = 0, SC_DONE, SC_IT, SC_BB, SC_RHO, SC_ALPHA, SC_OMEGA, SC_RZP, SC_SCALE, SC_DEN, SC_GB, SC_GA };

__device__ __forceinline__ bf16_t f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u = (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
  return (bf16_t)u;
}
__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float(((unsigned)h) << 16); }

struct Geo {
  int nr, nz, S;
  double dr;
  const double *rm, *rf, *dz, *izm;  // rf: nr+1 face factors rface/dr (0 at both ends); izm: nz+1 (0 at ends)
  const int *m;
};

__device__ __forceinline__ double vol(const Geo &g, int k, int j) { return g.dz[k] * (g.rm[j] * g.dr); }
__device__ __forceinline__ double sdiag(const Geo &g, int b, int k, int j) {
  const double mm = (double)g.m[b] * (double)g.m[b];
  return mm * (g.dz[k] * (g.dr / g.rm[j])) + g.dz[k] * (g.rf[j] + g.rf[j + 1]) +
         (g.rm[j] * g.dr) * (g.izm[k] + g.izm[k + 1]);
}
// (S x)(c) with neighbours read through `at`
template <typename F>
__device__ __forceinline__ double sapply(const Geo &g, int b, int k, int j, F at) {
  const double xc = at(k, j);
  const double mm = (double)g.m[b] * (double)g.m[b];
  double y = mm * (g.dz[k] * (g.dr / g.rm[j])) * xc;
  if (j > 0) y += g.dz[k] * g.rf[j] * (xc - at(k, j - 1));
  if (j < g.nr - 1) y += g.dz[k] * g.rf[j + 1] * (xc - at(k, j + 1));
  const double rv = g.rm[j] * g.dr;
  if (k > 0) y += rv * g.izm[k] * (xc - at(k - 1, j));
  if (k < g.nz - 1) y += rv * g.izm[k + 1] * (xc - at(k + 1, j));
  return y;
}

// Sum of kRedBlocks partials of quantity q for slot b (fixed order: the
// same value in every block that asks).
__device__ __forceinline__ double psum(const double *part, int nq, int b, int q) {
  double s = 0.0;
  for (int i = 0; i < kRedBlocks; i++) s += part[((long)b * kRedBlocks + i) * nq + q];
  return s;
}

// Block reduction of NQ values; thread 0 writes partials.
template <int NQ>
__device__ __forceinline__ void block_partials(double (&v)[NQ], double *part, int nq, int b) {
  __shared__ double red[NQ][256 / 64];
  for (int q = 0; q < NQ; q++) {
    double x = v[q];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int q = 0; q < NQ; q++)
      part[((long)b * kRedBlocks + blockIdx.x) * nq + q] = red[q][0] + red[q][1] + red[q][2] + red[q][3];
}

// grid (kRedBlocks, S), 256 threads: cell loop of slot b
#define CELL_LOOP(g, b, c, k, j)                                                   \
  const long ncs = (long)(g).nz * (g).nr;                                          \
  for (long c_ = (long)blockIdx.x * 256 + threadIdx.x; c_ < ncs; c_ += 256L * kRedBlocks) \
    for (int once_ = 1; once_; once_ = 0)                                          \
      for (const long c = (long)(b) * ncs + c_; once_; once_ = 0)                  \
        for (const int k = (int)(c_ / (g).nr), j = (int)(c_ - (long)k * (g).nr); once_; once_ = 0)

// ---------------------------------------------------------------------------
// elementary kernels
// ---------------------------------------------------------------------------
__global__ void k_apply(Geo g, const double *__restrict__ x, double *__restrict__ y) {
  const int b = blockIdx.y;
  CELL_LOOP(g, b, c, k, j) {
    const double *xb = x + (long)b * ncs;
    y[c] = sapply(g, b, k, j, [&](int kk, int jj) { return xb[(long)kk * g.nr + jj]; });
  }
}

// rhs: b = V f with, for m = 0, f minus its volume-weighted mean (from the
// partials of (V f, V) computed by k_rhs_moments).
__global__ void k_rhs_moments(Geo g, const double *__restrict__ f, double *part) {
  const int b = blockIdx.y;
  double v[2] = {0.0, 0.0};
  CELL_LOOP(g, b, c, k, j) {
    const double w = vol(g, k, j);
    v[0] += w * f[c];
    v[1] += w;
  }
  block_partials<2>(v, part, 2, b);
}
__global__ void k_rhs(Geo g, const double *__restrict__ f, const double *part, double *__restrict__ rhs) {
  const int b = blockIdx.y;
  const double mean = (g.m[b] == 0) ? psum(part, 2, b, 0) / psum(part, 2, b, 1) : 0.0;
  CELL_LOOP(g, b, c, k, j) { rhs[c] = vol(g, k, j) * (f[c] - mean); }
}
// gauge of m = 0 slots: subtract the volume-weighted mean
__global__ void k_gauge_moments(Geo g, const double *__restrict__ p, double *part) {
  const int b = blockIdx.y;
  double v[2] = {0.0, 0.0};
  CELL_LOOP(g, b, c, k, j) {
    const double w = vol(g, k, j);
    v[0] += w * p[c];
    v[1] += w;
  }
  block_partials<2>(v, part, 2, b);
}
__global__ void k_gauge(Geo g, const double *part, double *__restrict__ p) {
  const int b = blockIdx.y;
  if (g.m[b] != 0) return;
  const double mean = psum(part, 2, b, 0) / psum(part, 2, b, 1);
  CELL_LOOP(g, b, c, k, j) { p[c] -= mean; }
}
// true residual norms: ||b - S x||^2, ||b||^2
__global__ void k_resid(Geo g, const double *__restrict__ rhs, const double *__restrict__ x, double *part) {
  const int b = blockIdx.y;
  double v[2] = {0.0, 0.0};
  CELL_LOOP(g, b, c, k, j) {
    const double *xb = x + (long)b * ncs;
    const double r = rhs[c] - sapply(g, b, k, j, [&](int kk, int jj) { return xb[(long)kk * g.nr + jj]; });
    v[0] += r * r;
    v[1] += rhs[c] * rhs[c];
  }
  block_partials<2>(v, part, 2, b);
}

// ---------------------------------------------------------------------------
// Jacobi-preconditioned CG (textbook form), all slots together; per-slot
// scalars live on the device, each kernel sums the partials it needs.
// ---------------------------------------------------------------------------
// r = b - S x, z = D^-1 r, p = z; partials (r.z, r.r, b.b)
__global__ void k_cg_init(Geo g, const double *__restrict__ rhs, const double *__restrict__ x, double *__restrict__ r,
                          double *__restrict__ z, double *__restrict__ p, double *part) {
  const int b = blockIdx.y;
  double v[3] = {0.0, 0.0, 0.0};
  CELL_LOOP(g, b, c, k, j) {
    const double *xb = x + (long)b * ncs;
    const double rr = rhs[c] - sapply(g, b, k, j, [&](int kk, int jj) { return xb[(long)kk * g.nr + jj]; });
    const double zz = rr / sdiag(g, b, k, j);
    r[c] = rr; z[c] = zz; p[c] = zz;
    v[0] += rr * zz; v[1] += rr * rr; v[2] += rhs[c] * rhs[c];
  }
  block_partials<3>(v, part, 3, b);
}
__global__ void k_cg_start(int S, const double *part, double *sc, double tol2) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= S) return;
  sc[b * kNsc + SC_RZ] = psum(part, 3, b, 0);
  sc[b * kNsc + SC_BB] = psum(part, 3, b, 2);
  sc[b * kNsc + SC_DONE] = (psum(part, 3, b, 1) <= tol2 * sc[b * kNsc + SC_BB]) ? 1.0 : 0.0;
  sc[b * kNsc + SC_IT] = 0.0;
}
// after an update: r.z of the new residual, iteration count, done flag
__global__ void k_cg_fold(int S, const double *part_rz, double *sc, double tol2) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= S) return;
  double *s = sc + b * kNsc;
  if (s[SC_DONE] != 0.0) return;
  s[SC_RZ] = psum(part_rz, 3, b, 0);
  s[SC_IT] += 1.0;
  if (psum(part_rz, 3, b, 1) <= tol2 * s[SC_BB]) s[SC_DONE] = 1.0;
}
// q = S p; partial p.q
__global__ void k_cg_q(Geo g, const double *__restrict__ p, double *__restrict__ q, double *part_pq) {
  const int b = blockIdx.y;
  double v[1] = {0.0};
  CELL_LOOP(g, b, c, k, j) {
    const double *pb = p + (long)b * ncs;
    const double y = sapply(g, b, k, j, [&](int kk, int jj) { return pb[(long)kk * g.nr + jj]; });
    q[c] = y;
    v[0] += p[c] * y;
  }
  block_partials<1>(v, part_pq, 1, b);
}
// x += a p, r -= a q, z = D^-1 r; partials (r.z, r.r, 0)
__global__ void k_cg_update(Geo g, double *__restrict__ x, double *__restrict__ r, double *__restrict__ z,
                            const double *__restrict__ p, const double *__restrict__ q, const double *part_pq,
                            double *part_rz, const double *sc) {
  const int b = blockIdx.y;
  const double *s = sc + b * kNsc;
  const bool done = s[SC_DONE] != 0.0;
  const double pq = psum(part_pq, 1, b, 0);
  const double a = (done || pq == 0.0) ? 0.0 : s[SC_RZ] / pq;
  double v[3] = {0.0, 0.0, 0.0};
  CELL_LOOP(g, b, c, k, j) {
    const double rr = r[c] - a * q[c];
    const double zz = rr / sdiag(g, b, k, j);
    if (!done) { x[c] += a * p[c]; r[c] = rr; z[c] = zz; }
    v[0] += rr * zz; v[1] += rr * rr;
  }
  block_partials<3>(v, part_rz, 3, b);
}
// p = z + beta p with beta = (r.z)_new / (r.z)_old (the scalars still hold the old value)
__global__ void k_cg_dir(Geo g, const double *__restrict__ z, double *__restrict__ p, const double *part_rz,
                         const double *sc) {
  const int b = blockIdx.y;
  const double *s = sc + b * kNsc;
  if (s[SC_DONE] != 0.0) return;
  const double beta = psum(part_rz, 3, b, 0) / s[SC_RZ];
  CELL_LOOP(g, b, c, k, j) { p[c] = z[c] + beta * p[c]; }
}

// ---------------------------------------------------------------------------
// Jacobi-preconditioned BiCGSTAB (textbook form, right preconditioning)
// ---------------------------------------------------------------------------
// init: r = b - S x, rh = r, p = v = 0; partials (rh.r, r.r, b.b)
__global__ void k_bi_init(Geo g, const double *__restrict__ rhs, const double *__restrict__ x, double *r, double *rh,
                          double *p, double *v, double *part) {
  const int b = blockIdx.y;
  double w[3] = {0.0, 0.0, 0.0};
  CELL_LOOP(g, b, c, k, j) {
    const double *xb = x + (long)b * ncs;
    const double rr = rhs[c] - sapply(g, b, k, j, [&](int kk, int jj) { return xb[(long)kk * g.nr + jj]; });
    r[c] = rr; rh[c] = rr; p[c] = 0.0; v[c] = 0.0;
    w[0] += rr * rr; w[1] += rr * rr; w[2] += rhs[c] * rhs[c];
  }
  block_partials<3>(w, part, 3, b);
}
__global__ void k_bi_start(int S, const double *part, double *sc) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= S) return;
  double *s = sc + b * kNsc;
  s[SC_BB] = psum(part, 3, b, 2);
  s[SC_RHO] = 1.0; s[SC_ALPHA] = 1.0; s[SC_OMEGA] = 1.0;
  s[SC_DONE] = 0.0; s[SC_IT] = 0.0;
}
// p = r + beta (p - omega v), ph = D^-1 p ; rho2 = rh.r from part3 (q0), r.r (q1)
__global__ void k_bi_p(Geo g, const double *r, double *p, const double *v, double *ph, const double *part3,
                       double *sc, double tol2) {
  const int b = blockIdx.y;
  double *s = sc + b * kNsc;
  const double rho2 = psum(part3, 3, b, 0), rr = psum(part3, 3, b, 1);
  const bool done = s[SC_DONE] != 0.0 || rr <= tol2 * s[SC_BB];
  const double beta = done ? 0.0 : (rho2 / s[SC_RHO]) * (s[SC_ALPHA] / s[SC_OMEGA]);
  const double om = s[SC_OMEGA];
  CELL_LOOP(g, b, c, k, j) {
    if (done) { ph[c] = 0.0; continue; }
    const double pp = r[c] + beta * (p[c] - om * v[c]);
    p[c] = pp;
    ph[c] = pp / sdiag(g, b, k, j);
  }
}
// v = S ph; partial rh.v.  Block 0 records rho2 and the done flag.
__global__ void k_bi_v(Geo g, const double *ph, double *v, const double *rh, double *part1, const double *part3,
                       double *sc, double tol2) {
  const int b = blockIdx.y;
  double *s = sc + b * kNsc;
  double w[1] = {0.0};
  CELL_LOOP(g, b, c, k, j) {
    const double *pb = ph + (long)b * ncs;
    const double y = sapply(g, b, k, j, [&](int kk, int jj) { return pb[(long)kk * g.nr + jj]; });
    v[c] = y;
    w[0] += rh[c] * y;
  }
  block_partials<1>(w, part1, 1, b);
  if (blockIdx.x == 0 && threadIdx.x == 0 && s[SC_DONE] == 0.0) {
    const double rho2 = psum(part3, 3, b, 0), rr = psum(part3, 3, b, 1);
    if (rr <= tol2 * s[SC_BB]) s[SC_DONE] = 1.0;
    else { s[SC_RZ] = rho2; s[SC_IT] += 1.0; }
  }
}
// s = r - alpha v, sh = D^-1 s (alpha = rho2 / rh.v)
__global__ void k_bi_s(Geo g, const double *r, const double *v, double *sv, double *sh, const double *part1,
                       const double *sc) {
  const int b = blockIdx.y;
  const double *s = sc + b * kNsc;
  if (s[SC_DONE] != 0.0) return;
  const double a = s[SC_RZ] / psum(part1, 1, b, 0);
  CELL_LOOP(g, b, c, k, j) {
    const double ss = r[c] - a * v[c];
    sv[c] = ss;
    sh[c] = ss / sdiag(g, b, k, j);
  }
}
// t = S sh; partials (t.s, t.t)
__global__ void k_bi_t(Geo g, const double *sh, double *t, const double *sv, double *part2, const double *sc) {
  const int b = blockIdx.y;
  double w[2] = {0.0, 0.0};
  if (sc[b * kNsc + SC_DONE] == 0.0) {
    CELL_LOOP(g, b, c, k, j) {
      const double *pb = sh + (long)b * ncs;
      const double y = sapply(g, b, k, j, [&](int kk, int jj) { return pb[(long)kk * g.nr + jj]; });
      t[c] = y;
      w[0] += y * sv[c];
      w[1] += y * y;
    }
  }
  block_partials<2>(w, part2, 2, b);
}
// omega = t.s / t.t; x += alpha ph + omega sh; r = s - omega t; partials (rh.r, r.r)
__global__ void k_bi_x(Geo g, double *x, double *r, const double *rh, const double *ph, const double *sh,
                       const double *sv, const double *t, const double *part1, const double *part2, double *part3,
                       const double *sc) {
  const int b = blockIdx.y;
  const double *s = sc + b * kNsc;
  const bool done = s[SC_DONE] != 0.0;
  double w[3] = {0.0, 0.0, 0.0};
  const double a = done ? 0.0 : s[SC_RZ] / psum(part1, 1, b, 0);
  const double tt = psum(part2, 2, b, 1);
  const double om = (done || tt == 0.0) ? 0.0 : psum(part2, 2, b, 0) / tt;
  CELL_LOOP(g, b, c, k, j) {
    double rr = r[c];
    if (!done) {
      x[c] += a * ph[c] + om * sh[c];
      rr = sv[c] - om * t[c];
      r[c] = rr;
    }
    w[0] += rh[c] * rr; w[1] += rr * rr;
  }
  block_partials<3>(w, part3, 3, b);
}
// scalars after the x update: rho <- rho2, alpha, omega (block 0 of the next p kernel reads them)
__global__ void k_bi_scalars(int S, const double *part1, const double *part2, double *sc) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= S) return;
  double *s = sc + b * kNsc;
  if (s[SC_DONE] != 0.0) return;
  const double tt = psum(part2, 2, b, 1);
  s[SC_ALPHA] = s[SC_RZ] / psum(part1, 1, b, 0);
  s[SC_OMEGA] = (tt == 0.0) ? 1.0 : psum(part2, 2, b, 0) / tt;
  s[SC_RHO] = s[SC_RZ];
}

// ---------------------------------------------------------------------------
// Direct solve: radial tridiagonal systems (R_m + lambda_q W_r) p = b per
// (slot, z eigenvalue), one thread each; layout [slot][q][j].  The singular
// (m = 0, lambda = 0) system is pinned at its last row.
// ---------------------------------------------------------------------------
__global__ void k_radial_eig(Geo g, const double *__restrict__ lam, double *__restrict__ x, double *__restrict__ wk) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.S * g.nz) return;
  const int b = (int)(t / g.nz), q = (int)(t - (long)b * g.nz);
  const int nr = g.nr;
  double *xs = x + t * nr, *cp = wk + t * nr;
  const double mm = (double)g.m[b] * (double)g.m[b], l = lam[q];
  const bool pin = (g.m[b] == 0 && q == 0);
  // forward sweep
  double prev_c = 0.0, prev_x = 0.0;
  for (int j = 0; j < nr; j++) {
    const double lo = (j > 0) ? -g.rf[j] : 0.0, up = (j < nr - 1) ? -g.rf[j + 1] : 0.0;
    double d = mm * (g.dr / g.rm[j]) + g.rf[j] + g.rf[j + 1] + l * (g.rm[j] * g.dr);
    double rhs = xs[j];
    double lo2 = lo, up2 = up;
    if (pin && j == nr - 1) { lo2 = 0.0; d = 1.0; rhs = 0.0; }
    const double den = d - lo2 * prev_c;
    const double cj = up2 / den;
    const double xj = (rhs - lo2 * prev_x) / den;
    cp[j] = cj; xs[j] = xj;
    prev_c = cj; prev_x = xj;
  }
  for (int j = nr - 2; j >= 0; j--) xs[j] -= cp[j] * xs[j + 1];
}

// ---------------------------------------------------------------------------
// Network kernels
// ---------------------------------------------------------------------------
// features of f (density) per cell: f / rms(f), m dr/(m dr + rm), dz/dr,
// rm/rmax, zeros; rms from partials of f^2 (part: 1 quantity)
__global__ void k_fsq(Geo g, const double *__restrict__ f, double *part) {
  const int b = blockIdx.y;
  double v[1] = {0.0};
  CELL_LOOP(g, b, c, k, j) { v[0] += f[c] * f[c]; }
  block_partials<1>(v, part, 1, b);
}
// f may be a residual in rhs space: fv = 1 divides by the cell volume first
__global__ void k_rsq(Geo g, const double *__restrict__ r, double *part) {
  const int b = blockIdx.y;
  double v[1] = {0.0};
  CELL_LOOP(g, b, c, k, j) { const double f = r[c] / vol(g, k, j); v[0] += f * f; }
  block_partials<1>(v, part, 1, b);
}
// Features on the network grid, whose cells are blocks of pool x pool
// cells: block mean of f / rms(f), m drn / (m drn + rn), dzn / drn, rn / R
// (drn = pool dr, rn the block's centre radius, dzn the block height).  The
// network output times scale = rms(f) R^2 is the correction, prolonged
// piecewise constant to the cells of each block.
__global__ void k_features(Geo g, int pool, const double *__restrict__ f, int fv, const double *part,
                           bf16_t *__restrict__ X, double *sc) {
  const int b = blockIdx.y;
  const int nrn = g.nr / pool, nzn = g.nz / pool;
  double s = sqrt(psum(part, 1, b, 0) / ((double)g.nz * g.nr));
  if (!(s > 0.0)) s = 1.0;
  const double drn = g.dr * pool, R = g.dr * g.nr;
  // output scale rms(f) R^2: the size of smooth pressure solutions (a scale of
  // drn^2 would leave the last layer ~(R/drn)^2 too small to start)
  if (blockIdx.x == 0 && threadIdx.x == 0) sc[b * kNsc + SC_SCALE] = s * R * R;
  const double mdr = (double)g.m[b] * drn;
  const long ncs = (long)g.nz * g.nr, ncn = (long)nzn * nrn;
  for (long cn = (long)blockIdx.x * 256 + threadIdx.x; cn < ncn; cn += 256L * kRedBlocks) {
    const int kn = (int)(cn / nrn), jn = (int)(cn - (long)kn * nrn);
    double acc = 0.0, dzs = 0.0;
    for (int dk = 0; dk < pool; dk++) {
      const int k = kn * pool + dk;
      dzs += g.dz[k];
      for (int dj = 0; dj < pool; dj++) {
        const int j = jn * pool + dj;
        const long c = (long)b * ncs + (long)k * g.nr + j;
        acc += fv ? f[c] / vol(g, k, j) : f[c];
      }
    }
    const double fc = acc / (double)(pool * pool), rn = (jn + 0.5) * drn;
    bf16_t v[16];
    v[0] = f2bf((float)(fc / s));
    v[1] = f2bf((float)(mdr / (mdr + rn)));
    v[2] = f2bf((float)(dzs / drn));
    v[3] = f2bf((float)(rn / (g.dr * g.nr)));
    for (int q = 4; q < 16; q++) v[q] = 0;
    uint4 *dst = reinterpret_cast<uint4 *>(X + ((long)b * ncn + cn) * 16);
    const unsigned *w = reinterpret_cast<const unsigned *>(v);
    dst[0] = make_uint4(w[0], w[1], w[2], w[3]);
    dst[1] = make_uint4(w[4], w[5], w[6], w[7]);
  }
}
// Bilinear prolongation from the network grid (block centres) to the cells,
// in index space, constant beyond the outermost block centres.  1-D weights
// of cell i: u = (i + 1/2) / pool - 1/2, node floor(u) gets 1 - t and the
// next t (t = u - floor(u)).
__device__ __forceinline__ void pweights(int i, int pool, int nn, int &i0, double &t) {
  const double u = (i + 0.5) / pool - 0.5;
  if (u <= 0.0) { i0 = 0; t = 0.0; return; }
  if (u >= nn - 1) { i0 = nn - 1; t = 0.0; return; }
  i0 = (int)u;
  t = u - i0;
}
__device__ __forceinline__ double pweight(int i, int pool, int nn, int node) {
  int i0;
  double t;
  pweights(i, pool, nn, i0, t);
  return (i0 == node ? 1.0 - t : 0.0) + (i0 + 1 == node ? t : 0.0);
}
__device__ __forceinline__ double prolong(const Geo &g, int pool, const float *__restrict__ out, int b, int k,
                                          int j) {
  const int nrn = g.nr / pool, nzn = g.nz / pool;
  int k0, j0;
  double tk, tj;
  pweights(k, pool, nzn, k0, tk);
  pweights(j, pool, nrn, j0, tj);
  const float *o = out + (long)b * nzn * nrn;
  const int k1 = min(k0 + 1, nzn - 1), j1 = min(j0 + 1, nrn - 1);
  const double a = (1.0 - tj) * (double)o[(long)k0 * nrn + j0] + tj * (double)o[(long)k0 * nrn + j1];
  const double c = (1.0 - tj) * (double)o[(long)k1 * nrn + j0] + tj * (double)o[(long)k1 * nrn + j1];
  return (1.0 - tk) * a + tk * c;
}

// 3x3 dilated convolution as an implicit GEMM on 16x16x16 BF16 MFMA tiles.
// One wave computes 16 consecutive r cells of one (slot, z) row for all
// COUT channels: A = weights (rows cout, k = cin), B = inputs (k = cin,
// columns = cells), K looped over 9 taps x CIN/16.
// EPI 0: + bias, ReLU, bf16 store; 1: + bias, fp32 store of channel 0;
// 2: mask by act > 0 (ReLU backward), bf16 store.
template <int CIN, int COUT, int EPI>
__global__ __launch_bounds__(256) void k_conv(const bf16_t *__restrict__ X, const bf16_t *__restrict__ W,
                                              const float *__restrict__ bias, const bf16_t *__restrict__ act,
                                              void *__restrict__ Y, int nr, int nz, int dil) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int j0 = (blockIdx.x * 4 + wv) * 16;
  if (j0 >= nr) return;
  const int k = blockIdx.y, b = blockIdx.z;
  const int n = lane & 15, kg = lane >> 4;
  f4 acc[COUT / 16];
#pragma unroll
  for (int mt = 0; mt < COUT / 16; mt++) acc[mt] = f4{0.f, 0.f, 0.f, 0.f};
  const long slot0 = (long)b * nz * nr;
#pragma unroll 1
  for (int t = 0; t < 9; t++) {
    const int kk = k + (t / 3 - 1) * dil, jj = j0 + n + (t % 3 - 1) * dil;
    const bool ok = kk >= 0 && kk < nz && jj >= 0 && jj < nr;
    const bf16_t *xp = X + (slot0 + (long)(ok ? kk : 0) * nr + (ok ? jj : 0)) * CIN + kg * 4;
#pragma unroll
    for (int cc = 0; cc < CIN / 16; cc++) {
      bf4 bv = {0, 0, 0, 0};
      if (ok) bv = *reinterpret_cast<const bf4 *>(xp + cc * 16);
#pragma unroll
      for (int mt = 0; mt < COUT / 16; mt++) {
        const bf4 av = *reinterpret_cast<const bf4 *>(W + ((long)(t * COUT + mt * 16 + n) * CIN + cc * 16 + kg * 4));
        acc[mt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[mt], 0, 0, 0);
      }
    }
  }
  const long pix = slot0 + (long)k * nr + j0 + n;
#pragma unroll
  for (int mt = 0; mt < COUT / 16; mt++) {
    const int c0 = mt * 16 + kg * 4;
    if (EPI == 1) {
      if (mt == 0 && kg == 0) reinterpret_cast<float *>(Y)[pix] = acc[0][0] + bias[0];
      continue;
    }
    float y[4];
    if (EPI == 0) {
      for (int v = 0; v < 4; v++) y[v] = fmaxf(acc[mt][v] + bias[c0 + v], 0.f);
    } else {
      const bf4 a = *reinterpret_cast<const bf4 *>(act + pix * COUT + c0);
      for (int v = 0; v < 4; v++) y[v] = (bf2f((bf16_t)a[v]) > 0.f) ? acc[mt][v] : 0.f;
    }
    bf4 o;
    for (int v = 0; v < 4; v++) o[v] = (short)f2bf(y[v]);
    *reinterpret_cast<bf4 *>(reinterpret_cast<bf16_t *>(Y) + pix * COUT + c0) = o;
  }
}

// Weight gradient of a forward layer (CIN -> COUT): per block partial
// dW[t][co][ci] = sum_p G[p][co] A[p + tap_t][ci] over kDwRows z rows of one
// slot, on 16x16x16 MFMA tiles with K = 16 consecutive r cells (A operand
// rows = co, B operand columns = ci); wave w owns taps w, w+4, w+8.  Wave 0
// also sums G for the bias gradient.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_conv_dw(const bf16_t *__restrict__ G, const bf16_t *__restrict__ A,
                                                 float *__restrict__ part, int nr, int nz, int dil) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int n = lane & 15, kg = lane >> 4;
  const int b = blockIdx.y;
  const long slot0 = (long)b * nz * nr;
  constexpr int MT = COUT / 16, NT = CIN / 16;
  f4 acc[3][MT][NT];
  float dbs[MT];
#pragma unroll
  for (int q = 0; q < 3; q++)
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
      for (int nt = 0; nt < NT; nt++) acc[q][mt][nt] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int mt = 0; mt < MT; mt++) dbs[mt] = 0.f;
  const int kbeg = blockIdx.x * kDwRows, kend = min(kbeg + kDwRows, nz);
  for (int k = kbeg; k < kend; k++) {
    for (int j0 = 0; j0 < nr; j0 += 16) {
      bf4 av[MT];
#pragma unroll
      for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const bf16_t gv = G[(slot0 + (long)k * nr + j0 + kg * 4 + v) * COUT + mt * 16 + n];
          av[mt][v] = (short)gv;
          if (wv == 0) dbs[mt] += bf2f(gv);
        }
#pragma unroll
      for (int q = 0; q < 3; q++) {
        const int t = wv + 4 * q;
        if (t >= 9) break;
        const int kk = k + (t / 3 - 1) * dil, dj = (t % 3 - 1) * dil;
        const bool rowok = kk >= 0 && kk < nz;
#pragma unroll
        for (int nt = 0; nt < NT; nt++) {
          bf4 bv;
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const int jj = j0 + kg * 4 + v + dj;
            bv[v] = (rowok && jj >= 0 && jj < nr) ? (short)A[(slot0 + (long)kk * nr + jj) * CIN + nt * 16 + n] : (short)0;
          }
#pragma unroll
          for (int mt = 0; mt < MT; mt++)
            acc[q][mt][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[mt], bv, acc[q][mt][nt], 0, 0, 0);
        }
      }
    }
  }
  const long blk = (long)b * gridDim.x + blockIdx.x;
  float *pp = part + blk * (9 * COUT * CIN + COUT);
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const int t = wv + 4 * q;
    if (t >= 9) break;
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
      for (int nt = 0; nt < NT; nt++)
#pragma unroll
        for (int v = 0; v < 4; v++) pp[(t * COUT + mt * 16 + kg * 4 + v) * CIN + nt * 16 + n] = acc[q][mt][nt][v];
  }
  if (wv == 0) {
#pragma unroll
    for (int mt = 0; mt < MT; mt++) {
      float s = dbs[mt];
      s += __shfl_xor(s, 16);
      s += __shfl_xor(s, 32);
      if (kg == 0) pp[9 * COUT * CIN + mt * 16 + n] = s;
    }
  }
}

// sum of nblk partial vectors of length n (fixed order)
__global__ void k_sum_partials(const float *__restrict__ part, long nblk, int n, float *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (long q = 0; q < nblk; q++) s += part[q * n + i];
  out[i] = (float)s;
}

// fp32 master weights of one layer -> bf16 forward weights W[t][co][ci] and
// backward-data weights Wt[8 - t][ci][co]
__global__ void k_pack_layer(const float *__restrict__ w, int cin, int cout, bf16_t *__restrict__ wf,
                             bf16_t *__restrict__ wb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 9 * cin * cout) return;
  const int t = i / (cin * cout), rem = i - t * cin * cout, co = rem / cin, ci = rem - co * cin;
  const bf16_t v = f2bf(w[i]);
  wf[i] = v;
  wb[((8 - t) * cin + ci) * cout + co] = v;
}

// Adam (bias-corrected) on all parameters
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m1,
                       float *__restrict__ m2, int n, float lr, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = 0.9f * m1[i] + 0.1f * gi, v = 0.999f * m2[i] + 0.001f * gi * gi;
  m1[i] = a; m2[i] = v;
  w[i] -= lr * (a * c1) / (sqrtf(v * c2) + 1e-8f);
}

// Galerkin combination of the Jacobi estimate j = D^-1 r and the network
// correction c = scale * out (prolonged) as an approximation of S^-1 r:
// z = beta j + gamma c with (beta, gamma) minimising the energy error
// ||z - S^-1 r||_S, i.e. solving
//   [j.Sj  j.Sc] [beta ]   [j.r]
//   [j.Sc  c.Sc] [gamma] = [c.r].
// z is never worse in that norm than the best multiple of j alone (gamma = 0
// is admissible), whatever the network returns, and z.r = z.Sz > 0.
// Partials per slot: j.Sj, j.Sc, c.Sc, j.r, c.r.
__global__ void k_gal_dots(Geo g, int pool, const float *__restrict__ out, const double *sc,
                           const double *__restrict__ r, double *part) {
  const int b = blockIdx.y;
  const double s = sc[b * kNsc + SC_SCALE];
  double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  CELL_LOOP(g, b, c, k, j) {
    const long base = (long)b * ncs;
    auto jf = [&](int kk, int jj) { return r[base + (long)kk * g.nr + jj] / sdiag(g, b, kk, jj); };
    auto cf = [&](int kk, int jj) { return s * prolong(g, pool, out, b, kk, jj); };
    const double jc = jf(k, j), cc = cf(k, j);
    const double sj = sapply(g, b, k, j, jf), scv = sapply(g, b, k, j, cf);
    v[0] += jc * sj;
    v[1] += jc * scv;
    v[2] += cc * scv;
    v[3] += jc * r[c];
    v[4] += cc * r[c];
  }
  block_partials<5>(v, part, 5, b);
}
__global__ void k_gal_solve(int S, const double *part, double *sc) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= S) return;
  const double a11 = psum(part, 5, b, 0), a12 = psum(part, 5, b, 1), a22 = psum(part, 5, b, 2);
  const double r1 = psum(part, 5, b, 3), r2 = psum(part, 5, b, 4);
  const double det = a11 * a22 - a12 * a12;
  double beta = a11 > 0.0 ? r1 / a11 : 0.0, gam = 0.0;
  // a nearly dependent pair (or a zero / non-finite correction) keeps the Jacobi term alone
  if (a11 > 0.0 && a22 > 0.0 && det > 1e-10 * a11 * a22 && isfinite(det)) {
    beta = (r1 * a22 - r2 * a12) / det;
    gam = (a11 * r2 - a12 * r1) / det;
  }
  sc[b * kNsc + SC_GB] = beta;
  sc[b * kNsc + SC_GA] = gam;
}
// z = beta D^-1 r + gamma scale out
__global__ void k_gal_combine(Geo g, int pool, const float *__restrict__ out, const double *sc,
                              const double *__restrict__ r, double *__restrict__ z) {
  const int b = blockIdx.y;
  const double s = sc[b * kNsc + SC_SCALE], be = sc[b * kNsc + SC_GB], ga = sc[b * kNsc + SC_GA];
  CELL_LOOP(g, b, c, k, j) { z[c] = be * (r[c] / sdiag(g, b, k, j)) + ga * (s * prolong(g, pool, out, b, k, j)); }
}

// energy-norm loss of the Galerkin estimate of p = S^-1 b:
// e = beta D^-1 b + gamma scale out - p; per slot L = e.Se / den.  The
// coefficients are optimal, so dL/d(out) = 2 gamma scale Se / den (the
// terms through beta and gamma vanish); it goes, divided by the batch size,
// into gf; partial e.Se
__global__ void k_loss(Geo g, int pool, const float *__restrict__ out, const double *__restrict__ p,
                       const double *__restrict__ rhs, const double *sc, double *__restrict__ gf, double *part) {
  const int b = blockIdx.y;
  const double s = sc[b * kNsc + SC_SCALE], den = sc[b * kNsc + SC_DEN];
  const double be = sc[b * kNsc + SC_GB], ga = sc[b * kNsc + SC_GA];
  double v[1] = {0.0};
  CELL_LOOP(g, b, c, k, j) {
    const long base = (long)b * ncs;
    auto e = [&](int kk, int jj) {
      const long cc = base + (long)kk * g.nr + jj;
      return be * (rhs[cc] / sdiag(g, b, kk, jj)) + ga * (s * prolong(g, pool, out, b, kk, jj)) - p[cc];
    };
    const double se = sapply(g, b, k, j, e);
    v[0] += e(k, j) * se;
    gf[c] = 2.0 * ga * s * se / (den * g.S);
  }
  block_partials<1>(v, part, 1, b);
}
// gradient on the network grid: the adjoint of the bilinear prolongation
// applied to the cell gradients (each node gathers the cells within two
// blocks of it with their weights), into channel 0 of G
__global__ void k_pool_grad(Geo g, int pool, const double *__restrict__ gf, bf16_t *__restrict__ G) {
  const int b = blockIdx.y;
  const int nrn = g.nr / pool, nzn = g.nz / pool;
  const long ncs = (long)g.nz * g.nr, ncn = (long)nzn * nrn;
  for (long cn = (long)blockIdx.x * 256 + threadIdx.x; cn < ncn; cn += 256L * kRedBlocks) {
    const int kn = (int)(cn / nrn), jn = (int)(cn - (long)kn * nrn);
    double acc = 0.0;
    const int ka = max(0, (kn - 1) * pool), kb = min(g.nz - 1, (kn + 2) * pool - 1);
    const int ja = max(0, (jn - 1) * pool), jb = min(g.nr - 1, (jn + 2) * pool - 1);
    for (int k = ka; k <= kb; k++) {
      const double wk = pweight(k, pool, nzn, kn);
      if (wk == 0.0) continue;
      double row = 0.0;
      for (int j = ja; j <= jb; j++) {
        const double wj = pweight(j, pool, nrn, jn);
        if (wj != 0.0) row += wj * gf[(long)b * ncs + (long)k * g.nr + j];
      }
      acc += wk * row;
    }
    uint4 *dst = reinterpret_cast<uint4 *>(G + ((long)b * ncn + cn) * 16);
    dst[0] = make_uint4((unsigned)f2bf((float)acc), 0u, 0u, 0u);
    dst[1] = make_uint4(0u, 0u, 0u, 0u);
  }
}
// den = p . (V f) per slot (data of one batch)
__global__ void k_den(Geo g, const double *__restrict__ p, const double *__restrict__ rhs, double *part) {
  const int b = blockIdx.y;
  double v[1] = {0.0};
  CELL_LOOP(g, b, c, k, j) { v[0] += p[c] * rhs[c]; }
  block_partials<1>(v, part, 1, b);
}
__global__ void k_store_den(int S, const double *part, double *sc) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= S) return;
  sc[b * kNsc + SC_DEN] = psum(part, 1, b, 0);
}
__global__ void k_batch_loss(int S, const double *part, const double *sc, float *loss_out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double L = 0.0;
  for (int b = 0; b < S; b++) L += psum(part, 1, b, 0) / sc[b * kNsc + SC_DEN];
  *loss_out = (float)(L / S);
}

// gradient channel 0 from a host array (tests): G[c][0] = bf16(gout[c])
__global__ void k_gout(long n, const float *__restrict__ gout, bf16_t *__restrict__ G) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  uint4 *dst = reinterpret_cast<uint4 *>(G + c * 16);
  dst[0] = make_uint4((unsigned)f2bf(gout[c]), 0u, 0u, 0u);
  dst[1] = make_uint4(0u, 0u, 0u, 0u);
}

// ---------------------------------------------------------------------------
// training data: counter-based random numbers (splitmix64)
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long smix(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__device__ __forceinline__ double u01(unsigned long long seed, unsigned long long i) {
  return ((smix(seed ^ smix(i)) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
// family < 0: even slots ring cores, odd slots white noise
__global__ void k_sample(Geo g, unsigned long long seed, int family, double rmax, double zmax, const double *zm,
                         double *__restrict__ f) {
  const int b = blockIdx.y;
  const int fam = family >= 0 ? family : (b & 1);
  const unsigned long long sb = smix(seed * 1315423911ull + (unsigned long long)b);
  double pr[4][4];
  for (int q = 0; q < 4; q++) {
    pr[q][0] = (0.05 + 0.85 * u01(sb, 4 * q)) * rmax;
    pr[q][1] = (0.05 + 0.9 * u01(sb, 4 * q + 1)) * zmax;
    pr[q][2] = 0.1 + 0.2 * u01(sb, 4 * q + 2);
    pr[q][3] = 2.0 * u01(sb, 4 * q + 3) - 1.0;
  }
  CELL_LOOP(g, b, c, k, j) {
    if (fam == 0) {
      double v = 0.0;
      for (int q = 0; q < 4; q++) {
        const double dr = g.rm[j] - pr[q][0], dz = zm[k] - pr[q][1];
        v += pr[q][3] * exp(-(dr * dr + dz * dz) / (pr[q][2] * pr[q][2]));
      }
      f[c] = v;
    } else {
      const double u1 = u01(sb, 1000 + 2 * (unsigned long long)c), u2 = u01(sb, 1001 + 2 * (unsigned long long)c);
      f[c] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
  }
}
__global__ void k_random_modes(int S, unsigned long long seed, int mmax, int *m) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= S) return;
  m[b] = (int)(u01(smix(seed + 77), (unsigned long long)b) * (mmax + 1));
  if (m[b] > mmax) m[b] = mmax;
}

// ---------------------------------------------------------------------------
// host: symmetric tridiagonal eigenproblem (Sturm bisection + inverse
// iteration), used once per grid for the z transform of the direct solve
// ---------------------------------------------------------------------------
int sturm_count(const std::vector<double> &a, const std::vector<double> &e2, double x) {
  // number of eigenvalues < x (LDL^T of T - x I)
  int cnt = 0;
  double d = 1.0;
  const int n = (int)a.size();
  for (int i = 0; i < n; i++) {
    d = a[i] - x - (i > 0 ? e2[i - 1] / d : 0.0);
    if (d == 0.0) d = -1e-300;
    if (d < 0.0) cnt++;
  }
  return cnt;
}

void tridiag_eig(const std::vector<double> &a, const std::vector<double> &e, std::vector<double> &lam,
                 std::vector<double> &U) {
  const int n = (int)a.size();
  std::vector<double> e2(n > 1 ? n - 1 : 0);
  for (int i = 0; i + 1 < n; i++) e2[i] = e[i] * e[i];
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < n; i++) {
    const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < n ? fabs(e[i]) : 0.0);
    lo = std::min(lo, a[i] - r);
    hi = std::max(hi, a[i] + r);
  }
  lam.assign(n, 0.0);
  for (int q = 0; q < n; q++) {
    double l = lo, h = hi;
    for (int it = 0; it < 200 && h - l > 1e-15 * std::max(1.0, fabs(l) + fabs(h)); it++) {
      const double mid = 0.5 * (l + h);
      if (sturm_count(a, e2, mid) > q) h = mid; else l = mid;
    }
    lam[q] = 0.5 * (l + h);
  }
  U.assign((size_t)n * n, 0.0);
  std::vector<double> v(n), w(n), dd(n), cc(n);
  for (int q = 0; q < n; q++) {
    // shift slightly off the eigenvalue; inverse iteration with partial pivoting-free LU (tridiagonal)
    const double gap = (q > 0 ? lam[q] - lam[q - 1] : 1.0);
    const double sh = lam[q] + 1e-10 * std::max(fabs(lam[q]), 1e-3) * (gap > 0 ? 1.0 : 1.0);
    for (int i = 0; i < n; i++) v[i] = 1.0 + 0.01 * std::sin(1.0 + i * (q + 1.3));
    for (int it = 0; it < 4; it++) {
      // solve (T - sh I) w = v  (Thomas; T - sh I is nearly singular, which is the point)
      for (int i = 0; i < n; i++) {
        const double lo2 = i > 0 ? e[i - 1] : 0.0;
        double den = (a[i] - sh) - (i > 0 ? lo2 * cc[i - 1] : 0.0);
        if (fabs(den) < 1e-300) den = 1e-300;
        cc[i] = (i + 1 < n ? e[i] : 0.0) / den;
        dd[i] = (v[i] - (i > 0 ? lo2 * dd[i - 1] : 0.0)) / den;
      }
      w[n - 1] = dd[n - 1];
      for (int i = n - 2; i >= 0; i--) w[i] = dd[i] - cc[i] * w[i + 1];
      // orthogonalise against the previous vectors (clustered eigenvalues)
      for (int p = std::max(0, q - 8); p < q; p++) {
        double s = 0.0;
        for (int i = 0; i < n; i++) s += U[(size_t)p * n + i] * w[i];
        for (int i = 0; i < n; i++) w[i] -= s * U[(size_t)p * n + i];
      }
      double nrm = 0.0;
      for (int i = 0; i < n; i++) nrm += w[i] * w[i];
      nrm = std::sqrt(nrm);
      for (int i = 0; i < n; i++) v[i] = w[i] / nrm;
    }
    for (int i = 0; i < n; i++) U[(size_t)q * n + i] = v[i];
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// handle
// ---------------------------------------------------------------------------
struct vrh_hz_s {
  int nr = 0, nz = 0, S = 0, pool = 1;
  double rmax = 0, zmax = 0, stretch = 0, dr = 0;
  hipStream_t stream = nullptr;
  rocblas_handle blas = nullptr;
  std::string err;
  std::vector<double> zf, zm, dz, rm;
  std::vector<int> hm;
  // geometry
  double *d_rm = nullptr, *d_rf = nullptr, *d_dz = nullptr, *d_izm = nullptr, *d_zm = nullptr;
  int *d_m = nullptr;
  // direct solve: V (column q = z eigenvector, W-orthonormal), eigenvalues
  double *d_V = nullptr, *d_lam = nullptr;
  // vectors (S * nz * nr doubles each)
  std::vector<double *> vec;
  double *d_part = nullptr, *d_sc = nullptr;
  // network
  float *d_w = nullptr, *d_g = nullptr, *d_m1 = nullptr, *d_m2 = nullptr, *d_dwpart = nullptr;
  bf16_t *d_wf = nullptr, *d_wb = nullptr;
  bf16_t *d_x0 = nullptr, *d_a[3] = {nullptr, nullptr, nullptr}, *d_g4 = nullptr, *d_gt[2] = {nullptr, nullptr};
  float *d_out = nullptr, *d_loss = nullptr;
  long adam_t = 0;
  int nparam = 0;
  int woff[4], boff[4], wfo[4];
  Geo geo() const {
    Geo g;
    g.nr = nr; g.nz = nz; g.S = S; g.dr = dr;
    g.rm = d_rm; g.rf = d_rf; g.dz = d_dz; g.izm = d_izm; g.m = d_m;
    return g;
  }
  size_t ncell() const { return (size_t)S * nz * nr; }
  size_t nnet() const { return (size_t)S * (nz / pool) * (nr / pool); }
};

namespace {
int hfail(vrh_hz_t *h, const std::string &m, int code = -1) {
  if (h) h->err = m;
  return code;
}
#define HZCHK(h, call)                                                                     \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) {                                                                \
      (void)hipGetLastError(); /* reported here, not by the next launch check */           \
      return hfail(h, std::string(#call) + ": " + hipGetErrorString(e_), -3);              \
    }                                                                                      \
  } while (0)
#define BLCHK(h, call)                                                                     \
  do {                                                                                     \
    rocblas_status s_ = (call);                                                            \
    if (s_ != rocblas_status_success) return hfail(h, std::string(#call) + ": rocblas error " + std::to_string((int)s_), -3); \
  } while (0)

template <typename T> int halloc(vrh_hz_t *h, T **p, size_t n) {
  HZCHK(h, hipMalloc((void **)p, n * sizeof(T) + 256));
  HZCHK(h, hipMemsetAsync(*p, 0, n * sizeof(T) + 256, h->stream));
  return 0;
}

// Host-to-device copy on the handle's stream, complete on return: ordered
// after the zero fills halloc queues there (a null-stream hipMemcpy is not).
int hz_h2d(vrh_hz_t *h, void *dst, const void *src, size_t bytes) {
  HZCHK(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream));
  HZCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

dim3 red_grid(const vrh_hz_t *h) { return dim3(kRedBlocks, h->S); }

int upload_host(vrh_hz_t *h, double *d, const double *src) {
  HZCHK(h, hipMemcpyAsync(d, src, h->ncell() * sizeof(double), hipMemcpyHostToDevice, h->stream));
  return 0;
}
int download_host(vrh_hz_t *h, double *dst, const double *d) {
  HZCHK(h, hipMemcpyAsync(dst, d, h->ncell() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HZCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

// rhs = V f (m = 0: mean removed)
void launch_rhs(vrh_hz_t *h, const double *f, double *rhs) {
  Geo g = h->geo();
  hipLaunchKernelGGL(k_rhs_moments, red_grid(h), dim3(256), 0, h->stream, g, f, h->d_part);
  hipLaunchKernelGGL(k_rhs, red_grid(h), dim3(256), 0, h->stream, g, f, h->d_part, rhs);
}
void launch_gauge(vrh_hz_t *h, double *p) {
  Geo g = h->geo();
  hipLaunchKernelGGL(k_gauge_moments, red_grid(h), dim3(256), 0, h->stream, g, p, h->d_part);
  hipLaunchKernelGGL(k_gauge, red_grid(h), dim3(256), 0, h->stream, g, h->d_part, p);
}

// direct solve of S p = rhs (device arrays); uses vec[6], vec[7] as work
int direct_dev(vrh_hz_t *h, const double *rhs, double *p) {
  const int nr = h->nr, nz = h->nz, S = h->S;
  const double one = 1.0, zero = 0.0;
  double *bh = h->vec[6], *wk = h->vec[7];
  BLCHK(h, rocblas_dgemm_strided_batched(h->blas, rocblas_operation_none, rocblas_operation_none, nr, nz, nz, &one,
                                         rhs, nr, (rocblas_stride)nr * nz, h->d_V, nz, 0, &zero, bh, nr,
                                         (rocblas_stride)nr * nz, S));
  Geo g = h->geo();
  const long nsys = (long)S * nz;
  hipLaunchKernelGGL(k_radial_eig, dim3((unsigned)((nsys + 127) / 128)), dim3(128), 0, h->stream, g, h->d_lam, bh, wk);
  BLCHK(h, rocblas_dgemm_strided_batched(h->blas, rocblas_operation_none, rocblas_operation_transpose, nr, nz, nz,
                                         &one, bh, nr, (rocblas_stride)nr * nz, h->d_V, nz, 0, &zero, p, nr,
                                         (rocblas_stride)nr * nz, S));
  launch_gauge(h, p);
  HZCHK(h, hipGetLastError());
  return 0;
}

// network forward on features already in d_x0 (network grid); output fp32 in d_out
void net_forward(vrh_hz_t *h) {
  const int nr = h->nr / h->pool, nz = h->nz / h->pool, S = h->S;
  dim3 grd((unsigned)((nr + 63) / 64), nz, S), blk(256);
  const bf16_t *wf = h->d_wf;
  hipLaunchKernelGGL((k_conv<16, 32, 0>), grd, blk, 0, h->stream, h->d_x0, wf + h->wfo[0], h->d_w + h->boff[0],
                     nullptr, (void *)h->d_a[0], nr, nz, kDil[0]);
  hipLaunchKernelGGL((k_conv<32, 32, 0>), grd, blk, 0, h->stream, h->d_a[0], wf + h->wfo[1], h->d_w + h->boff[1],
                     nullptr, (void *)h->d_a[1], nr, nz, kDil[1]);
  hipLaunchKernelGGL((k_conv<32, 32, 0>), grd, blk, 0, h->stream, h->d_a[1], wf + h->wfo[2], h->d_w + h->boff[2],
                     nullptr, (void *)h->d_a[2], nr, nz, kDil[2]);
  hipLaunchKernelGGL((k_conv<32, 16, 1>), grd, blk, 0, h->stream, h->d_a[2], wf + h->wfo[3], h->d_w + h->boff[3],
                     nullptr, (void *)h->d_out, nr, nz, kDil[3]);
}

// features of a density (fv = 0) or of a residual in rhs space (fv = 1)
void net_features(vrh_hz_t *h, const double *f, int fv) {
  Geo g = h->geo();
  if (fv) hipLaunchKernelGGL(k_rsq, red_grid(h), dim3(256), 0, h->stream, g, f, h->d_part);
  else hipLaunchKernelGGL(k_fsq, red_grid(h), dim3(256), 0, h->stream, g, f, h->d_part);
  hipLaunchKernelGGL(k_features, red_grid(h), dim3(256), 0, h->stream, g, h->pool, f, fv, h->d_part, h->d_x0,
                     h->d_sc);
}

// backward from the 16-channel gradient in d_g4: parameter gradients into d_g
void net_backward(vrh_hz_t *h) {
  const int nr = h->nr / h->pool, nz = h->nz / h->pool, S = h->S;
  dim3 grd((unsigned)((nr + 63) / 64), nz, S), blk(256);
  dim3 gdw((unsigned)((nz + kDwRows - 1) / kDwRows), S);
  const long nblk = (long)gdw.x * S;
  const bf16_t *wb = h->d_wb;
  auto dw = [&](int l, const bf16_t *G, const bf16_t *A) {
    const int cin = kLayerCin[l], cout = kLayerCout[l];
    if (cin == 16) hipLaunchKernelGGL((k_conv_dw<16, 32>), gdw, blk, 0, h->stream, G, A, h->d_dwpart, nr, nz, kDil[l]);
    else if (cout == 16) hipLaunchKernelGGL((k_conv_dw<32, 16>), gdw, blk, 0, h->stream, G, A, h->d_dwpart, nr, nz, kDil[l]);
    else hipLaunchKernelGGL((k_conv_dw<32, 32>), gdw, blk, 0, h->stream, G, A, h->d_dwpart, nr, nz, kDil[l]);
    const int n = 9 * cin * cout + cout;
    // the partial layout is [W (9*cout*cin)][b (cout)], as the parameters
    hipLaunchKernelGGL(k_sum_partials, dim3((n + 255) / 256), dim3(256), 0, h->stream, h->d_dwpart, nblk, n,
                       h->d_g + h->woff[l]);
  };
  // layer 4 (32 -> 16): dW from G4 and a3; gradient to a3: G3 = conv(G4, W4^T) * (a3 > 0)
  dw(3, h->d_g4, h->d_a[2]);
  hipLaunchKernelGGL((k_conv<16, 32, 2>), grd, blk, 0, h->stream, h->d_g4, wb + h->wfo[3], nullptr, h->d_a[2],
                     (void *)h->d_gt[0], nr, nz, kDil[3]);
  dw(2, h->d_gt[0], h->d_a[1]);
  hipLaunchKernelGGL((k_conv<32, 32, 2>), grd, blk, 0, h->stream, h->d_gt[0], wb + h->wfo[2], nullptr, h->d_a[1],
                     (void *)h->d_gt[1], nr, nz, kDil[2]);
  dw(1, h->d_gt[1], h->d_a[0]);
  hipLaunchKernelGGL((k_conv<32, 32, 2>), grd, blk, 0, h->stream, h->d_gt[1], wb + h->wfo[1], nullptr, h->d_a[0],
                     (void *)h->d_gt[0], nr, nz, kDil[1]);
  dw(0, h->d_gt[0], h->d_x0);
}

void pack_weights(vrh_hz_t *h) {
  for (int l = 0; l < 4; l++) {
    const int n = 9 * kLayerCin[l] * kLayerCout[l];
    hipLaunchKernelGGL(k_pack_layer, dim3((n + 255) / 256), dim3(256), 0, h->stream, h->d_w + h->woff[l],
                       kLayerCin[l], kLayerCout[l], h->d_wf + h->wfo[l], h->d_wb + h->wfo[l]);
  }
}

// the network's Galerkin estimate of S^-1 r into z (features, forward, the
// 2x2 energy-optimal coefficients, the combination); z may be null to stop
// after the coefficients
double *gal_part(vrh_hz_t *h) { return h->d_part + (size_t)h->S * kRedBlocks * 16; }
void net_galerkin(vrh_hz_t *h, const double *r, double *z) {
  Geo g = h->geo();
  net_features(h, r, 1);
  net_forward(h);
  hipLaunchKernelGGL(k_gal_dots, red_grid(h), dim3(256), 0, h->stream, g, h->pool, h->d_out, h->d_sc, r, gal_part(h));
  hipLaunchKernelGGL(k_gal_solve, dim3((h->S + 63) / 64), dim3(64), 0, h->stream, h->S, gal_part(h), h->d_sc);
  if (z) hipLaunchKernelGGL(k_gal_combine, red_grid(h), dim3(256), 0, h->stream, g, h->pool, h->d_out, h->d_sc, r, z);
}

int sync_scalars(vrh_hz_t *h, std::vector<double> &sc) {
  sc.resize((size_t)h->S * kNsc);
  HZCHK(h, hipMemcpyAsync(sc.data(), h->d_sc, sc.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HZCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

// PCG from x (device arrays rhs, x); returns iterations per slot in its
int pcg_dev(vrh_hz_t *h, const double *rhs, double *x, double tol, int maxit, std::vector<int> &its) {
  Geo g = h->geo();
  double *r = h->vec[2], *z = h->vec[3], *p = h->vec[4], *q = h->vec[5];
  double *part_rz = h->d_part, *part_pq = h->d_part + (size_t)h->S * kRedBlocks * 4;
  const double tol2 = tol * tol;
  const dim3 sg((h->S + 63) / 64), sb(64);
  hipLaunchKernelGGL(k_cg_init, red_grid(h), dim3(256), 0, h->stream, g, rhs, x, r, z, p, part_rz);
  hipLaunchKernelGGL(k_cg_start, sg, sb, 0, h->stream, h->S, part_rz, h->d_sc, tol2);
  std::vector<double> sc;
  for (int it = 0; it < maxit; it++) {
    hipLaunchKernelGGL(k_cg_q, red_grid(h), dim3(256), 0, h->stream, g, p, q, part_pq);
    hipLaunchKernelGGL(k_cg_update, red_grid(h), dim3(256), 0, h->stream, g, x, r, z, p, q, part_pq, part_rz,
                       h->d_sc);
    hipLaunchKernelGGL(k_cg_dir, red_grid(h), dim3(256), 0, h->stream, g, z, p, part_rz, h->d_sc);
    hipLaunchKernelGGL(k_cg_fold, sg, sb, 0, h->stream, h->S, part_rz, h->d_sc, tol2);
    if ((it & 15) == 15) {
      int rc = sync_scalars(h, sc);
      if (rc) return rc;
      bool all = true;
      for (int b = 0; b < h->S; b++) all = all && sc[(size_t)b * kNsc + SC_DONE] != 0.0;
      if (all) break;
    }
  }
  int rc = sync_scalars(h, sc);
  if (rc) return rc;
  its.resize(h->S);
  for (int b = 0; b < h->S; b++)
    its[b] = sc[(size_t)b * kNsc + SC_DONE] != 0.0 ? (int)sc[(size_t)b * kNsc + SC_IT] : -1;
  HZCHK(h, hipGetLastError());
  return 0;
}

int bicgstab_dev(vrh_hz_t *h, const double *rhs, double *x, double tol, int maxit, std::vector<int> &its) {
  Geo g = h->geo();
  double *r = h->vec[2], *rh = h->vec[3], *p = h->vec[4], *v = h->vec[5], *ph = h->vec[6], *sv = h->vec[7],
         *sh = h->vec[8], *t = h->vec[9];
  double *part1 = h->d_part, *part2 = part1 + (size_t)h->S * kRedBlocks * 4,
         *part3 = part2 + (size_t)h->S * kRedBlocks * 4;
  const double tol2 = tol * tol;
  hipLaunchKernelGGL(k_bi_init, red_grid(h), dim3(256), 0, h->stream, g, rhs, x, r, rh, p, v, part3);
  hipLaunchKernelGGL(k_bi_start, dim3((h->S + 63) / 64), dim3(64), 0, h->stream, h->S, part3, h->d_sc);
  std::vector<double> sc;
  for (int it = 0; it < maxit; it++) {
    hipLaunchKernelGGL(k_bi_p, red_grid(h), dim3(256), 0, h->stream, g, r, p, v, ph, part3, h->d_sc, tol2);
    hipLaunchKernelGGL(k_bi_v, red_grid(h), dim3(256), 0, h->stream, g, ph, v, rh, part1, part3, h->d_sc, tol2);
    hipLaunchKernelGGL(k_bi_s, red_grid(h), dim3(256), 0, h->stream, g, r, v, sv, sh, part1, h->d_sc);
    hipLaunchKernelGGL(k_bi_t, red_grid(h), dim3(256), 0, h->stream, g, sh, t, sv, part2, h->d_sc);
    hipLaunchKernelGGL(k_bi_x, red_grid(h), dim3(256), 0, h->stream, g, x, r, rh, ph, sh, sv, t, part1, part2, part3,
                       h->d_sc);
    hipLaunchKernelGGL(k_bi_scalars, dim3((h->S + 63) / 64), dim3(64), 0, h->stream, h->S, part1, part2, h->d_sc);
    if ((it & 15) == 15) {
      int rc = sync_scalars(h, sc);
      if (rc) return rc;
      bool all = true;
      for (int b = 0; b < h->S; b++) all = all && sc[(size_t)b * kNsc + SC_DONE] != 0.0;
      if (all) break;
    }
  }
  // fold the final residual into the done flags
  hipLaunchKernelGGL(k_bi_p, red_grid(h), dim3(256), 0, h->stream, g, r, p, v, ph, part3, h->d_sc, tol2);
  hipLaunchKernelGGL(k_bi_v, red_grid(h), dim3(256), 0, h->stream, g, ph, v, rh, part1, part3, h->d_sc, tol2);
  int rc = sync_scalars(h, sc);
  if (rc) return rc;
  its.resize(h->S);
  for (int b = 0; b < h->S; b++)
    its[b] = sc[(size_t)b * kNsc + SC_DONE] != 0.0 ? (int)sc[(size_t)b * kNsc + SC_IT] : -1;
  HZCHK(h, hipGetLastError());
  return 0;
}

// flexible CG (Polak-Ribiere beta) preconditioned by the network: z is the
// Galerkin combination of D^-1 r and the network correction for r; partials
// r.z, rold.z
__global__ void k_fcg_z(Geo g, int pool, const float *__restrict__ out, const double *sc, const double *__restrict__ r,
                        const double *__restrict__ rold, double *__restrict__ z, double *part) {
  const int b = blockIdx.y;
  const double s = sc[b * kNsc + SC_SCALE], be = sc[b * kNsc + SC_GB], ga = sc[b * kNsc + SC_GA];
  double v[2] = {0.0, 0.0};
  CELL_LOOP(g, b, c, k, j) {
    const double zz = be * (r[c] / sdiag(g, b, k, j)) + ga * (s * prolong(g, pool, out, b, k, j));
    z[c] = zz;
    v[0] += r[c] * zz;
    v[1] += rold[c] * zz;
  }
  block_partials<2>(v, part, 2, b);
}
__global__ void k_fcg_p(Geo g, const double *__restrict__ z, double *__restrict__ p, const double *part_z,
                        double *sc, int first) {
  const int b = blockIdx.y;
  double *s = sc + b * kNsc;
  if (s[SC_DONE] != 0.0) return;
  const double rz = psum(part_z, 2, b, 0), zro = psum(part_z, 2, b, 1);
  const double beta = first ? 0.0 : (rz - zro) / s[SC_RZP];
  CELL_LOOP(g, b, c, k, j) { p[c] = z[c] + beta * p[c]; }
}
__global__ void k_fcg_q(Geo g, const double *__restrict__ p, double *__restrict__ q, double *part_pq,
                        const double *part_z, double *sc) {
  const int b = blockIdx.y;
  double w[1] = {0.0};
  CELL_LOOP(g, b, c, k, j) {
    const double *pb = p + (long)b * ncs;
    const double y = sapply(g, b, k, j, [&](int kk, int jj) { return pb[(long)kk * g.nr + jj]; });
    q[c] = y;
    w[0] += p[c] * y;
  }
  block_partials<1>(w, part_pq, 1, b);
  (void)part_z; (void)sc;
}
// x += a p, rold = r, r -= a q with a = (r.z) / (p.q); partials r.r
__global__ void k_fcg_update(Geo g, double *__restrict__ x, double *__restrict__ r, double *__restrict__ rold,
                             const double *__restrict__ p, const double *__restrict__ q, const double *part_pq,
                             const double *part_z, const double *sc, double *part_rr) {
  const int b = blockIdx.y;
  const double *s = sc + b * kNsc;
  const bool done = s[SC_DONE] != 0.0;
  const double pq = psum(part_pq, 1, b, 0);
  const double a = (done || pq == 0.0) ? 0.0 : psum(part_z, 2, b, 0) / pq;
  double v[1] = {0.0};
  CELL_LOOP(g, b, c, k, j) {
    const double rr = r[c];
    const double rn = rr - a * q[c];
    if (!done) { x[c] += a * p[c]; rold[c] = rr; r[c] = rn; }
    v[0] += (done ? rr : rn) * (done ? rr : rn);
  }
  block_partials<1>(v, part_rr, 1, b);
}
// bookkeeping after an update: rz_prev <- r.z, iteration count, done flag
// A slot whose preconditioned residual is not a descent direction (r.z <= 0,
// or not finite) is frozen as broken (done = 2) and left to the PCG fallback.
__global__ void k_fcg_scalars(int S, const double *part_z, const double *part_rr, double *sc, double tol2) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= S) return;
  double *s = sc + b * kNsc;
  if (s[SC_DONE] != 0.0) return;
  const double rz = psum(part_z, 2, b, 0), rr = psum(part_rr, 1, b, 0);
  s[SC_RZP] = rz;
  s[SC_IT] += 1.0;
  if (rr <= tol2 * s[SC_BB]) s[SC_DONE] = 1.0;
  else if (!(rz > 0.0) || !isfinite(rr)) s[SC_DONE] = 2.0;
}
// restart a broken slot from zero
__global__ void k_reset_broken(Geo g, const double *sc, double *__restrict__ x) {
  const int b = blockIdx.y;
  if (sc[b * kNsc + SC_DONE] != 2.0) return;
  CELL_LOOP(g, b, c, k, j) { x[c] = 0.0; }
}
__global__ void k_residual_init(Geo g, const double *__restrict__ rhs, const double *__restrict__ x,
                                double *__restrict__ r, double *__restrict__ rold, double *part) {
  const int b = blockIdx.y;
  double v[2] = {0.0, 0.0};
  CELL_LOOP(g, b, c, k, j) {
    const double *xb = x + (long)b * ncs;
    const double rr = rhs[c] - sapply(g, b, k, j, [&](int kk, int jj) { return xb[(long)kk * g.nr + jj]; });
    r[c] = rr; rold[c] = 0.0;
    v[0] += rr * rr; v[1] += rhs[c] * rhs[c];
  }
  block_partials<2>(v, part, 2, b);
}
__global__ void k_fcg_start(int S, const double *part, double *sc, double tol2) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= S) return;
  double *s = sc + b * kNsc;
  s[SC_BB] = psum(part, 2, b, 1);
  s[SC_IT] = 0.0;
  s[SC_DONE] = (psum(part, 2, b, 0) <= tol2 * s[SC_BB]) ? 1.0 : 0.0;
}

int fcg_dev(vrh_hz_t *h, const double *rhs, double *x, double tol, int maxit, std::vector<int> &its) {
  Geo g = h->geo();
  double *r = h->vec[2], *rold = h->vec[3], *p = h->vec[4], *q = h->vec[5], *z = h->vec[8];
  double *pa = h->d_part, *part_z = pa + (size_t)h->S * kRedBlocks * 4, *part_pq = part_z + (size_t)h->S * kRedBlocks * 4,
         *part_rr = part_pq + (size_t)h->S * kRedBlocks * 4;
  const double tol2 = tol * tol;
  hipLaunchKernelGGL(k_residual_init, red_grid(h), dim3(256), 0, h->stream, g, rhs, x, r, rold, pa);
  hipLaunchKernelGGL(k_fcg_start, dim3((h->S + 63) / 64), dim3(64), 0, h->stream, h->S, pa, h->d_sc, tol2);
  std::vector<double> sc;
  for (int it = 0; it < maxit; it++) {
    net_galerkin(h, r, nullptr);  // d_part[0..] and the Galerkin partials as scratch
    hipLaunchKernelGGL(k_fcg_z, red_grid(h), dim3(256), 0, h->stream, g, h->pool, h->d_out, h->d_sc, r, rold, z,
                       part_z);
    hipLaunchKernelGGL(k_fcg_p, red_grid(h), dim3(256), 0, h->stream, g, z, p, part_z, h->d_sc, it == 0 ? 1 : 0);
    hipLaunchKernelGGL(k_fcg_q, red_grid(h), dim3(256), 0, h->stream, g, p, q, part_pq, part_z, h->d_sc);
    hipLaunchKernelGGL(k_fcg_update, red_grid(h), dim3(256), 0, h->stream, g, x, r, rold, p, q, part_pq, part_z,
                       h->d_sc, part_rr);
    hipLaunchKernelGGL(k_fcg_scalars, dim3((h->S + 63) / 64), dim3(64), 0, h->stream, h->S, part_z, part_rr, h->d_sc,
                       tol2);
    if ((it & 7) == 7) {
      int rc = sync_scalars(h, sc);
      if (rc) return rc;
      bool all = true;
      for (int b = 0; b < h->S; b++) all = all && sc[(size_t)b * kNsc + SC_DONE] != 0.0;
      if (all) break;
    }
  }
  int rc = sync_scalars(h, sc);
  if (rc) return rc;
  its.resize(h->S);
  for (int b = 0; b < h->S; b++) its[b] = sc[(size_t)b * kNsc + SC_DONE] == 1.0 ? (int)sc[(size_t)b * kNsc + SC_IT] : -1;
  hipLaunchKernelGGL(k_reset_broken, red_grid(h), dim3(256), 0, h->stream, g, h->d_sc, x);
  HZCHK(h, hipGetLastError());
  return 0;
}

int true_residual(vrh_hz_t *h, const double *rhs, const double *x, std::vector<double> &rel) {
  Geo g = h->geo();
  hipLaunchKernelGGL(k_resid, red_grid(h), dim3(256), 0, h->stream, g, rhs, x, h->d_part);
  std::vector<double> part((size_t)h->S * kRedBlocks * 2);
  HZCHK(h, hipMemcpyAsync(part.data(), h->d_part, part.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HZCHK(h, hipStreamSynchronize(h->stream));
  rel.assign(h->S, 0.0);
  for (int b = 0; b < h->S; b++) {
    double a = 0.0, c = 0.0;
    for (int i = 0; i < kRedBlocks; i++) { a += part[((size_t)b * kRedBlocks + i) * 2]; c += part[((size_t)b * kRedBlocks + i) * 2 + 1]; }
    rel[b] = c > 0.0 ? std::sqrt(a / c) : std::sqrt(a);
  }
  return 0;
}

}  // namespace

namespace {
// Training batches: each draws right-hand sides and random modes, and (with
// max_cg > 0) replaces each right-hand side by the residual after k plain PCG
// iterations from zero, k cycling through 0, 1, 4, 16, ... <= max_cg over the
// batches, so that the network sees the residuals a preconditioner is
// applied to; the target is the exact solve of whatever it sees.
int train_batches(vrh_hz_t *h, int nsamples, double lr, unsigned long long seed, int max_cg, float *loss) {
  const int nb = (nsamples + h->S - 1) / h->S;
  if (nb > (1 << 16)) return hfail(h, "too many batches");
  std::vector<int> saved(h->hm);
  Geo g = h->geo();
  double *f = h->vec[0], *rhs = h->vec[1], *pt = h->vec[11];
  std::vector<int> levels(1, 0);
  for (int k = 1; k <= max_cg; k *= 4) levels.push_back(k);
  for (int it = 0; it < nb; it++) {
    const unsigned long long sd = seed * 1000003ull + (unsigned long long)it;
    hipLaunchKernelGGL(k_random_modes, dim3((h->S + 63) / 64), dim3(64), 0, h->stream, h->S, sd, 63, h->d_m);
    hipLaunchKernelGGL(k_sample, red_grid(h), dim3(256), 0, h->stream, g, sd, -1, h->rmax, h->zmax, h->d_zm, f);
    launch_rhs(h, f, rhs);
    const double *src = rhs;
    const int kcg = levels[(size_t)it % levels.size()];
    if (kcg > 0) {
      std::vector<int> its;
      HZCHK(h, hipMemsetAsync(h->vec[10], 0, h->ncell() * sizeof(double), h->stream));
      int rc = pcg_dev(h, rhs, h->vec[10], 1e-300, kcg, its);
      if (rc) return rc;
      src = h->vec[2];  // the residual after kcg iterations
    }
    int rc = direct_dev(h, src, pt);
    if (rc) return rc;
    hipLaunchKernelGGL(k_den, red_grid(h), dim3(256), 0, h->stream, g, pt, src, h->d_part);
    hipLaunchKernelGGL(k_store_den, dim3((h->S + 63) / 64), dim3(64), 0, h->stream, h->S, h->d_part, h->d_sc);
    net_galerkin(h, src, nullptr);
    hipLaunchKernelGGL(k_loss, red_grid(h), dim3(256), 0, h->stream, g, h->pool, h->d_out, pt, src, h->d_sc,
                       h->vec[9], h->d_part);
    hipLaunchKernelGGL(k_pool_grad, red_grid(h), dim3(256), 0, h->stream, g, h->pool, h->vec[9], h->d_g4);
    hipLaunchKernelGGL(k_batch_loss, dim3(1), dim3(64), 0, h->stream, h->S, h->d_part, h->d_sc, h->d_loss + it);
    net_backward(h);
    h->adam_t++;
    const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)h->adam_t)));
    const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)h->adam_t)));
    hipLaunchKernelGGL(k_adam, dim3((h->nparam + 255) / 256), dim3(256), 0, h->stream, h->d_w, h->d_g, h->d_m1,
                       h->d_m2, h->nparam, (float)lr, c1, c2);
    pack_weights(h);
    HZCHK(h, hipGetLastError());
  }
  if (loss) HZCHK(h, hipMemcpyAsync(loss, h->d_loss, nb * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  HZCHK(h, hipStreamSynchronize(h->stream));
  return vrh_hz_set_modes(h, saved.data());
}
}  // namespace

// ---------------------------------------------------------------------------
// In-step use by the stepper's pressure solve (vrh_set_mode_solver): the
// stepper writes each slot's right-hand side f (the mode's divergence, as the
// direct path receives it) into the input buffer, solves in place ordered
// after its own stream, and reads the solution back.
// ---------------------------------------------------------------------------
int hz_slots(const vrh_hz_t *h) { return h->S; }
double *hz_step_input(vrh_hz_t *h) { return h->vec[0]; }
const double *hz_step_output(vrh_hz_t *h) { return h->vec[10]; }
const char *hz_error(const vrh_hz_t *h) { return h->err.c_str(); }

// The grid must be the stepper's: same nr, nz, rmax, zmax; z faces zc(1..nz+1)
// and centres zm(1..nz) equal to this handle's tanh mesh (1e-12 of zmax).
int hz_step_check(const vrh_hz_t *h, int nr, int nz, double rmax, double zmax, const double *zc, const double *zm,
                  std::string &err) {
  if (h->nr != nr || h->nz != nz) { err = "mode solver: grid size differs from the handle's"; return -1; }
  if (std::fabs(h->rmax - rmax) > 1e-14 * rmax || std::fabs(h->zmax - zmax) > 1e-14 * zmax) {
    err = "mode solver: domain differs from the handle's"; return -1;
  }
  double d = 0.0;
  for (int k = 0; k <= nz; k++) d = std::max(d, std::fabs(h->zf[k] - zc[k + 1]));
  for (int k = 0; k < nz; k++) d = std::max(d, std::fabs(h->zm[k] - zm[k + 1]));
  if (!(d <= 1e-12 * zmax)) { err = "mode solver: the handle's axial mesh is not the Helmholtz handle's tanh mesh"; return -1; }
  return 0;
}

int hz_step_solve(vrh_hz_t *h, hipStream_t caller, int method, double tol, int maxit, long long *iters) {
  hipEvent_t ev;
  HZCHK(h, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  HZCHK(h, hipEventRecord(ev, caller));
  HZCHK(h, hipStreamWaitEvent(h->stream, ev, 0));
  double *rhs = h->vec[1], *x = h->vec[10];
  launch_rhs(h, h->vec[0], rhs);
  std::vector<int> its(h->S, 0), its2;
  int rc = 0;
  if (method == VRH_HZ_LEARNED || method == VRH_HZ_LEARNED_FCG) net_galerkin(h, rhs, x);
  else HZCHK(h, hipMemsetAsync(x, 0, h->ncell() * sizeof(double), h->stream));
  if (method == VRH_HZ_LEARNED_FCG) {
    if ((rc = fcg_dev(h, rhs, x, tol, maxit, its))) return rc;
    bool fallback = false;
    for (int b = 0; b < h->S; b++) fallback = fallback || its[b] < 0;
    if (fallback) {
      if ((rc = pcg_dev(h, rhs, x, tol, maxit, its2))) return rc;
      for (int b = 0; b < h->S; b++)
        if (its[b] < 0) its[b] = its2[b] < 0 ? -1 : maxit + its2[b];
    }
  } else if (method == VRH_HZ_BICGSTAB) {
    if ((rc = bicgstab_dev(h, rhs, x, tol, maxit, its))) return rc;
  } else {
    if ((rc = pcg_dev(h, rhs, x, tol, maxit, its))) return rc;
  }
  launch_gauge(h, x);
  // the caller's stream reads x only after everything queued here
  HZCHK(h, hipEventRecord(ev, h->stream));
  HZCHK(h, hipStreamWaitEvent(caller, ev, 0));
  HZCHK(h, hipEventDestroy(ev));
  long long tot = 0;
  bool ok = true;
  for (int b = 0; b < h->S; b++) {
    ok = ok && its[b] >= 0;
    tot += its[b] > 0 ? its[b] : 0;
  }
  if (iters) *iters += tot;
  if (!ok) return hfail(h, "mode solver: a slot did not converge within maxit", -4);
  return 0;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int vrh_hz_last_error(vrh_hz_t *h, char *buf, size_t len) {
  if (!buf || len == 0) return -1;
  snprintf(buf, len, "%s", h ? h->err.c_str() : "null handle");
  return 0;
}

int vrh_hz_create(const vrh_hz_grid *gr, vrh_hz_t **out) {
  if (!gr || !out) return -1;
  *out = nullptr;
  const int pool = gr->pool > 1 ? gr->pool : 1;
  if (gr->nr < 16 || gr->nr % (16 * pool) != 0 || gr->nz < 4 || gr->nz % pool != 0 || gr->nslots < 1 ||
      gr->rmax <= 0 || gr->zmax <= 0 || gr->stretch < 0)
    return -1;
  vrh_hz_t *h = new (std::nothrow) vrh_hz_t();
  if (!h) return -1;
  *out = h;
  h->nr = gr->nr; h->nz = gr->nz; h->S = gr->nslots; h->pool = pool;
  h->rmax = gr->rmax; h->zmax = gr->zmax; h->stretch = gr->stretch;
  const int nr = h->nr, nz = h->nz;
  h->dr = h->rmax / nr;
  HZCHK(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  if (rocblas_create_handle(&h->blas) != rocblas_status_success) return hfail(h, "rocblas_create_handle failed", -3);
  BLCHK(h, rocblas_set_stream(h->blas, h->stream));
  // mesh (tanh map of the z faces)
  h->zf.resize(nz + 1); h->zm.resize(nz); h->dz.resize(nz); h->rm.resize(nr);
  for (int k = 0; k <= nz; k++)
    h->zf[k] = h->stretch > 0 ? h->zmax * (1.0 + std::tanh(h->stretch * ((double)k / nz - 1.0)) / std::tanh(h->stretch))
                              : h->zmax * k / nz;
  for (int k = 0; k < nz; k++) { h->zm[k] = 0.5 * (h->zf[k] + h->zf[k + 1]); h->dz[k] = h->zf[k + 1] - h->zf[k]; }
  for (int j = 0; j < nr; j++) h->rm[j] = (j + 0.5) * h->dr;
  std::vector<double> rf(nr + 1, 0.0), izm(nz + 1, 0.0);
  for (int j = 1; j < nr; j++) rf[j] = (j * h->dr) / h->dr;
  for (int k = 1; k < nz; k++) izm[k] = 1.0 / (h->zm[k] - h->zm[k - 1]);
  int rc;
  if ((rc = halloc(h, &h->d_rm, nr)) || (rc = halloc(h, &h->d_rf, nr + 1)) || (rc = halloc(h, &h->d_dz, nz)) ||
      (rc = halloc(h, &h->d_izm, nz + 1)) || (rc = halloc(h, &h->d_zm, nz)) || (rc = halloc(h, &h->d_m, h->S)))
    return rc;
  if ((rc = hz_h2d(h, h->d_rm, h->rm.data(), nr * sizeof(double)))) return rc;
  if ((rc = hz_h2d(h, h->d_rf, rf.data(), (nr + 1) * sizeof(double)))) return rc;
  if ((rc = hz_h2d(h, h->d_dz, h->dz.data(), nz * sizeof(double)))) return rc;
  if ((rc = hz_h2d(h, h->d_izm, izm.data(), (nz + 1) * sizeof(double)))) return rc;
  if ((rc = hz_h2d(h, h->d_zm, h->zm.data(), nz * sizeof(double)))) return rc;
  h->hm.assign(h->S, 0);
  // z eigenproblem Z v = lambda W v  ->  T = W^-1/2 Z W^-1/2 (symmetric tridiagonal)
  {
    std::vector<double> a(nz), e(nz > 1 ? nz - 1 : 0), lam, U;
    for (int k = 0; k < nz; k++) a[k] = (izm[k] + izm[k + 1]) / h->dz[k];
    for (int k = 0; k + 1 < nz; k++) e[k] = -izm[k + 1] / std::sqrt(h->dz[k] * h->dz[k + 1]);
    tridiag_eig(a, e, lam, U);
    lam[0] = std::max(lam[0], 0.0);
    std::vector<double> V((size_t)nz * nz);
    for (int q = 0; q < nz; q++)
      for (int k = 0; k < nz; k++) V[(size_t)q * nz + k] = U[(size_t)q * nz + k] / std::sqrt(h->dz[k]);
    if ((rc = halloc(h, &h->d_V, (size_t)nz * nz)) || (rc = halloc(h, &h->d_lam, nz))) return rc;
    if ((rc = hz_h2d(h, h->d_V, V.data(), V.size() * sizeof(double)))) return rc;
    if ((rc = hz_h2d(h, h->d_lam, lam.data(), nz * sizeof(double)))) return rc;
  }
  // work vectors
  h->vec.assign(12, nullptr);
  for (auto &v : h->vec)
    if ((rc = halloc(h, &v, h->ncell()))) return rc;
  if ((rc = halloc(h, &h->d_part, (size_t)h->S * kRedBlocks * 24)) || (rc = halloc(h, &h->d_sc, (size_t)h->S * kNsc)))
    return rc;
  // network
  int off = 0, wo = 0;
  for (int l = 0; l < 4; l++) {
    h->woff[l] = off;
    off += 9 * kLayerCin[l] * kLayerCout[l];
    h->boff[l] = off;
    off += kLayerCout[l];
    h->wfo[l] = wo;
    wo += 9 * kLayerCin[l] * kLayerCout[l];
  }
  h->nparam = off;
  if ((rc = halloc(h, &h->d_w, off)) || (rc = halloc(h, &h->d_g, off)) || (rc = halloc(h, &h->d_m1, off)) ||
      (rc = halloc(h, &h->d_m2, off)) || (rc = halloc(h, &h->d_wf, wo)) || (rc = halloc(h, &h->d_wb, wo)))
    return rc;
  const size_t nc = h->ncell();
  if ((rc = halloc(h, &h->d_x0, nc * 16)) || (rc = halloc(h, &h->d_a[0], nc * 32)) ||
      (rc = halloc(h, &h->d_a[1], nc * 32)) || (rc = halloc(h, &h->d_a[2], nc * 32)) ||
      (rc = halloc(h, &h->d_g4, nc * 16)) || (rc = halloc(h, &h->d_gt[0], nc * 32)) ||
      (rc = halloc(h, &h->d_gt[1], nc * 32)) || (rc = halloc(h, &h->d_out, nc)) || (rc = halloc(h, &h->d_loss, 1 << 16)))
    return rc;
  const long nblk = (long)((nz + kDwRows - 1) / kDwRows) * h->S;
  if ((rc = halloc(h, &h->d_dwpart, (size_t)nblk * (9 * 32 * 32 + 32)))) return rc;
  HZCHK(h, hipStreamSynchronize(h->stream));
  return vrh_hz_net_init(h, 1234);
}

int vrh_hz_destroy(vrh_hz_t *h) {
  if (!h) return -1;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  void *ptrs[] = {h->d_rm, h->d_rf, h->d_dz, h->d_izm, h->d_zm, h->d_m, h->d_V, h->d_lam, h->d_part, h->d_sc,
                  h->d_w, h->d_g, h->d_m1, h->d_m2, h->d_wf, h->d_wb, h->d_x0, h->d_a[0], h->d_a[1], h->d_a[2],
                  h->d_g4, h->d_gt[0], h->d_gt[1], h->d_out, h->d_loss, h->d_dwpart};
  for (void *p : ptrs) if (p) (void)hipFree(p);
  for (double *v : h->vec) if (v) (void)hipFree(v);
  if (h->blas) rocblas_destroy_handle(h->blas);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return 0;
}

int vrh_hz_set_modes(vrh_hz_t *h, const int *m) {
  if (!h || !m) return -1;
  for (int b = 0; b < h->S; b++)
    if (m[b] < 0) return hfail(h, "modes must be >= 0");
  h->hm.assign(m, m + h->S);
  HZCHK(h, hipMemcpyAsync(h->d_m, h->hm.data(), h->S * sizeof(int), hipMemcpyHostToDevice, h->stream));
  HZCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int vrh_hz_mesh(vrh_hz_t *h, double *zf, double *zm, double *rm) {
  if (!h) return -1;
  if (zf) std::copy(h->zf.begin(), h->zf.end(), zf);
  if (zm) std::copy(h->zm.begin(), h->zm.end(), zm);
  if (rm) std::copy(h->rm.begin(), h->rm.end(), rm);
  return 0;
}

int vrh_hz_apply(vrh_hz_t *h, const double *x, double *y) {
  if (!h || !x || !y) return -1;
  int rc = upload_host(h, h->vec[0], x);
  if (rc) return rc;
  hipLaunchKernelGGL(k_apply, red_grid(h), dim3(256), 0, h->stream, h->geo(), h->vec[0], h->vec[1]);
  HZCHK(h, hipGetLastError());
  return download_host(h, y, h->vec[1]);
}

int vrh_hz_direct(vrh_hz_t *h, const double *f, double *p) {
  if (!h || !f || !p) return -1;
  int rc = upload_host(h, h->vec[0], f);
  if (rc) return rc;
  launch_rhs(h, h->vec[0], h->vec[1]);
  if ((rc = direct_dev(h, h->vec[1], h->vec[10]))) return rc;
  return download_host(h, p, h->vec[10]);
}

int vrh_hz_solve(vrh_hz_t *h, int method, double tol, int maxit, const double *f, double *p, int *iters,
                 double *relres, double *ms) {
  if (!h || !f || !p || tol <= 0 || maxit < 1) return -1;
  if (method != VRH_HZ_PCG && method != VRH_HZ_BICGSTAB && method != VRH_HZ_LEARNED && method != VRH_HZ_LEARNED_FCG)
    return hfail(h, "unknown method");
  int rc = upload_host(h, h->vec[0], f);
  if (rc) return rc;
  hipEvent_t e0, e1;
  HZCHK(h, hipEventCreate(&e0));
  HZCHK(h, hipEventCreate(&e1));
  HZCHK(h, hipEventRecord(e0, h->stream));
  double *rhs = h->vec[1], *x = h->vec[10];
  launch_rhs(h, h->vec[0], rhs);
  std::vector<int> its(h->S, 0), its2;
  if (method == VRH_HZ_LEARNED || method == VRH_HZ_LEARNED_FCG) {
    // warm start: the Galerkin combination of the Jacobi estimate and the
    // network correction for V f
    net_galerkin(h, rhs, x);
  }
  if (method == VRH_HZ_LEARNED) {
    if ((rc = pcg_dev(h, rhs, x, tol, maxit, its))) return rc;
  } else if (method == VRH_HZ_LEARNED_FCG) {
    if ((rc = fcg_dev(h, rhs, x, tol, maxit, its))) return rc;
    bool fallback = false;
    for (int b = 0; b < h->S; b++) fallback = fallback || its[b] < 0;
    if (fallback) {
      // plain PCG from the current iterate; slots already within tol stop at once
      if ((rc = pcg_dev(h, rhs, x, tol, maxit, its2))) return rc;
      for (int b = 0; b < h->S; b++)
        if (its[b] < 0) its[b] = its2[b] < 0 ? -1 : maxit + its2[b];
    }
  } else {
    HZCHK(h, hipMemsetAsync(x, 0, h->ncell() * sizeof(double), h->stream));
    rc = method == VRH_HZ_PCG ? pcg_dev(h, rhs, x, tol, maxit, its) : bicgstab_dev(h, rhs, x, tol, maxit, its);
    if (rc) return rc;
  }
  launch_gauge(h, x);
  HZCHK(h, hipEventRecord(e1, h->stream));
  HZCHK(h, hipEventSynchronize(e1));
  float t = 0.f;
  HZCHK(h, hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::vector<double> rel;
  if ((rc = true_residual(h, rhs, x, rel))) return rc;
  if (iters) std::copy(its.begin(), its.end(), iters);
  if (relres) std::copy(rel.begin(), rel.end(), relres);
  if (ms) *ms = t;
  return download_host(h, p, x);
}

int vrh_hz_net_size(vrh_hz_t *h) { return h ? h->nparam : -1; }

int vrh_hz_net_init(vrh_hz_t *h, unsigned long long seed) {
  if (!h) return -1;
  // He-style uniform init on the used input channels; small output layer
  std::vector<float> w(h->nparam, 0.f);
  unsigned long long s = seed * 6364136223846793005ull + 1442695040888963407ull;
  auto uni = [&]() {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return ((s >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  };
  for (int l = 0; l < 4; l++) {
    const int ci = kLayerCin[l], co = kLayerCout[l];
    const int fan = 9 * (l == 0 ? 4 : ci);
    const double lim = std::sqrt(6.0 / fan) * (l == 3 ? 0.01 : 1.0);
    for (int t = 0; t < 9; t++)
      for (int o = 0; o < co; o++)
        for (int i = 0; i < ci; i++) {
          const bool used = (l != 0 || i < 4) && (l != 3 || o == 0);
          w[h->woff[l] + (t * co + o) * ci + i] = used ? (float)((2.0 * uni() - 1.0) * lim) : 0.f;
        }
  }
  h->adam_t = 0;
  HZCHK(h, hipMemsetAsync(h->d_m1, 0, h->nparam * sizeof(float), h->stream));
  HZCHK(h, hipMemsetAsync(h->d_m2, 0, h->nparam * sizeof(float), h->stream));
  return vrh_hz_net_set(h, w.data());
}

int vrh_hz_net_get(vrh_hz_t *h, float *params) {
  if (!h || !params) return -1;
  HZCHK(h, hipMemcpyAsync(params, h->d_w, h->nparam * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  HZCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int vrh_hz_net_set(vrh_hz_t *h, const float *params) {
  if (!h || !params) return -1;
  HZCHK(h, hipMemcpyAsync(h->d_w, params, h->nparam * sizeof(float), hipMemcpyHostToDevice, h->stream));
  pack_weights(h);
  HZCHK(h, hipGetLastError());
  HZCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int vrh_hz_net_forward(vrh_hz_t *h, const double *f, float *out, double *scale) {
  if (!h || !f || !out) return -1;
  int rc = upload_host(h, h->vec[0], f);
  if (rc) return rc;
  net_features(h, h->vec[0], 0);
  net_forward(h);
  HZCHK(h, hipGetLastError());
  HZCHK(h, hipMemcpyAsync(out, h->d_out, h->nnet() * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  std::vector<double> sc;
  if ((rc = sync_scalars(h, sc))) return rc;
  if (scale)
    for (int b = 0; b < h->S; b++) scale[b] = sc[(size_t)b * kNsc + SC_SCALE];
  return 0;
}

int vrh_hz_net_grad(vrh_hz_t *h, const double *f, const float *gout, float *grads) {
  if (!h || !f || !gout || !grads) return -1;
  int rc = upload_host(h, h->vec[0], f);
  if (rc) return rc;
  net_features(h, h->vec[0], 0);
  net_forward(h);
  HZCHK(h, hipMemcpyAsync(h->d_out, gout, h->nnet() * sizeof(float), hipMemcpyHostToDevice, h->stream));
  const long n = (long)h->nnet();
  hipLaunchKernelGGL(k_gout, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, n, h->d_out, h->d_g4);
  net_backward(h);
  HZCHK(h, hipGetLastError());
  HZCHK(h, hipMemcpyAsync(grads, h->d_g, h->nparam * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  HZCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int vrh_hz_sample(vrh_hz_t *h, unsigned long long seed, int family, double *f) {
  if (!h || !f || family < -1 || family > 1) return -1;
  hipLaunchKernelGGL(k_sample, red_grid(h), dim3(256), 0, h->stream, h->geo(), seed, family, h->rmax, h->zmax,
                     h->d_zm, h->vec[0]);
  HZCHK(h, hipGetLastError());
  return download_host(h, f, h->vec[0]);
}

int vrh_hz_train(vrh_hz_t *h, int nsamples, double lr, unsigned long long seed, float *loss) {
  if (!h || nsamples < 1 || lr <= 0) return -1;
  return train_batches(h, nsamples, lr, seed, 0, loss);
}

int vrh_hz_train_residuals(vrh_hz_t *h, int nsamples, double lr, unsigned long long seed, int max_cg, float *loss) {
  if (!h || nsamples < 1 || lr <= 0 || max_cg < 0) return -1;
  return train_batches(h, nsamples, lr, seed, max_cg, loss);
}

}  // extern "C"
