// vrh_particles.hip -- Lagrangian particle step (iter_part, part_tools.F90:374-1025).
//
// One thread per particle.  Particle state lives in HBM as structure of
// arrays, u[v * cap + p] (v = 0..nvar-1), so every lane of a wavefront reads
// consecutive addresses for each state variable.  The 27-point gathers hit
// the L2-resident velocity fields.  T is the arithmetic type of the
// interpolation and right-hand side (double in FP64 mode, float in mixed
// mode); the state update is accumulated in double.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <cmath>
#include <cstring>
#include <vector>

#include "vrh_internal.h"

namespace {

struct PartArgs {
  VGrid g;
  const double *q1, *q2, *q3;
  const double *mint;  // 9 x 27 (device copy; the fit kernels take it by value)
  double *u, *ddt2;
  void *fit;          // LSQ fit rows (39 x cap of the work type)
  int *alive;
  const long *n_dev;  // used slots, read on the device
  long cap, n;
  int nvar, nsub, scheme, irk;
  double st, partdt, gam, rom, dtheta, dr, dz, rmax, zmax;
};

// Interpolation and chain rule of one particle, part_tools.F90:519-797, in
// two kernels per substep: the fit kernel gathers the 27-point stencil of
// each velocity component, fits the quadratic and stores the component's
// value, Cartesian gradient and Hessian (13 numbers, 39 per particle); the
// step kernel rotates them to the Cartesian velocity, gradient and Hessian
// (:700-797) and updates the particle state.  Splitting keeps each kernel's
// live registers small enough for several waves per SIMD.
constexpr int kFitRows = 39;

// the pseudo-inverse by value (a kernel argument); the fit kernel copies it
// to LDS once per block and reads it back in every fit pass
template <typename T> struct MintArg { T m[243]; };

// cell of a particle and its azimuth (part_tools.F90:519-538); false if
// the 27-point stencil would leave the haloed arrays
__device__ inline bool lsq_locate(const PartArgs &a, double xp, double yp, double zp, double &theta, double &r,
                                  int &it, int &ir, int &iz) {
  const VGrid &g = a.g;
  r = sqrt(xp * xp + yp * yp);
  const double pi = (double)4.0f * atan(1.0);
  if (xp >= 0.0 && yp >= 0.0) theta = acos(xp / r);           // part_tools.F90:524-534
  else if (xp < 0.0 && yp >= 0.0) theta = acos(xp / r);
  else theta = acos(-xp / r) + pi;
  it = (int)floor(theta / a.dtheta) + 1;
  if (g.nt == 1) it = 1;
  ir = (int)floor(r / a.dr) + 1;
  iz = (int)floor(zp / a.dz) + 1;
  return (r == r) && ir >= 1 && ir <= g.nr + 1 && iz >= 1 && iz <= g.nz + 1;
}

// sin and cos of the particle's azimuth and of twice it, in the work type
// (shared by the three components' chain rules)
template <typename T> struct Trig { T sn, cs, sn2, cs2; };

// fit of component ivar (1..3) and its value, Cartesian gradient and
// Hessian at the particle (part_tools.F90:540-698)
template <typename T>
__device__ inline void lsq_component(const PartArgs &a, const T *mtl, int ivar, double theta, double r,
                                     double zp, int it, int ir, int iz, const Trig<T> &tg, T &fv, T df[3],
                                     T ddf[3][3]) {
  const VGrid &g = a.g;
  const T sn = tg.sn, cs = tg.cs, sn2 = tg.sn2, cs2 = tg.cs2;
  const T rr = (T)r, r2 = rr * rr;
  const double dx[3] = {a.dtheta, a.dr, a.dz};
  T y[27];
  // the three theta columns of the stencil (periodic) and the component's
  // array; gathers go out one z plane (9 loads) at a time, so only nine
  // addresses are live at once
  int itc[3] = {it - 1, it, it + 1};
#pragma unroll
  for (int q = 0; q < 3; q++) {
    if (itc[q] > g.nt) itc[q] = 1;
    if (itc[q] < 1) itc[q] = g.nt;
  }
  const double *qa = ivar == 1 ? a.q1 : (ivar == 2 ? a.q2 : a.q3);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    double vv[9];
#pragma unroll
    for (int u = 0; u < 9; u++) {
      const int aa = u % 3, b = u / 3;
      vv[u] = qa[hidx(g, itc[aa], ir + b - 1, iz + c - 1)];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < 9; u++) {
      const int b = u / 3, irof = ir + b - 1;
      double v = vv[u];
      if (ivar == 2) {
        const double rc = (irof - 1.0) * a.dr;
        v = (rc > (double)0.0001f) ? v / rc : 0.0;
      }
      y[u + 9 * c] = (T)v;
    }
  }
  const T y0 = y[13];
#pragma unroll
  for (int iv = 0; iv < 27; iv++) y[iv] = y[iv] - y0;
  // least-squares fit, iterated as part_tools.F90:609-624 (converges in
  // one pass since mint is the exact pseudo-inverse; the second pass adds
  // the rounding-level correction)
  T acoef[9];
#pragma unroll
  for (int q = 0; q < 9; q++) acoef[q] = (T)0;
#pragma unroll 1
  for (int iint = 1; iint <= 100; iint++) {
    // the pseudo-inverse entries are read from LDS in every pass: an offset
    // the compiler cannot see through keeps it from hoisting all 243 of them
    // out of the loops (they would not fit the scalar registers)
    int mo = 0;
    __asm__ volatile("" : "+s"(mo));
    T mp[9];
    T nrm = (T)0;
#pragma unroll
    for (int q = 0; q < 9; q++) mp[q] = (T)0;
#pragma unroll
    for (int kk = 0; kk < 27; kk++) {
      const int a1 = kk % 3 - 1, b1 = (kk / 3) % 3 - 1, c1 = kk / 9 - 1;
      const T x0 = (T)(a1 * dx[0]), x1 = (T)(b1 * dx[1]), x2 = (T)(c1 * dx[2]);
      const T rint = x0 * acoef[0] + x1 * acoef[1] + x2 * acoef[2] + x0 * x0 * acoef[3] + x0 * x1 * acoef[4] +
                     x1 * x1 * acoef[5] + x1 * x2 * acoef[6] + x2 * x2 * acoef[7] + x2 * x0 * acoef[8];
      const T res = rint - y[kk];
#pragma unroll
      for (int q = 0; q < 9; q++) mp[q] = mp[q] - mtl[mo + q * 27 + kk] * res;
    }
#pragma unroll
    for (int q = 0; q < 9; q++) { acoef[q] = acoef[q] + mp[q]; nrm += mp[q] * mp[q]; }
    if (sqrt((double)nrm) < (double)0.000001f) break;
    // in FP32 the rounding-level correction of the second pass sits near
    // 1e-7 of the coefficients and can stay above the absolute 1e-6 tolerance
    if (sizeof(T) == sizeof(float) && iint >= 2) break;
  }
  double dtp, drp, dzp;
  if (ivar == 1) { dtp = theta - (it - 1.0) * a.dtheta; drp = r - (ir - 0.5) * a.dr; dzp = zp - (iz - 0.5) * a.dz; }
  else if (ivar == 2) { dtp = theta - (it - 0.5) * a.dtheta; drp = r - (ir - 1.0) * a.dr; dzp = zp - (iz - 0.5) * a.dz; }
  else { dtp = theta - (it - 0.5) * a.dtheta; drp = r - (ir - 0.5) * a.dr; dzp = zp - (iz - 1.0) * a.dz; }
  const double twopi = 8.0 * atan(1.0);
  if (dtp > a.dtheta) dtp -= twopi;
  if (dtp < -a.dtheta) dtp += twopi;
  const T t_ = (T)dtp, r_ = (T)drp, z_ = (T)dzp;
  const T *A = acoef;
  fv = t_ * A[0] + r_ * A[1] + z_ * A[2] + t_ * t_ * A[3] + t_ * r_ * A[4] + r_ * r_ * A[5] + r_ * z_ * A[6] +
       z_ * z_ * A[7] + z_ * t_ * A[8] + y0;
  T FI[12];
  FI[0] = A[0] + (T)2 * t_ * A[3] + r_ * A[4] + z_ * A[8];
  FI[1] = A[1] + t_ * A[4] + (T)2 * r_ * A[5] + z_ * A[6];
  FI[2] = A[2] + r_ * A[6] + (T)2 * z_ * A[7] + t_ * A[8];
  FI[3] = (T)2 * A[3];
  FI[4] = r_ * A[4];   // ftr as part_tools.F90:657
  FI[5] = z_ * A[8];   // ftz as part_tools.F90:658
  FI[6] = A[4]; FI[7] = (T)2 * A[5]; FI[8] = A[6];
  FI[9] = A[8]; FI[10] = A[6]; FI[11] = (T)2 * A[7];
  // chain rule to Cartesian derivatives (:671-698)
  df[0] = -sn / rr * FI[0] + cs * FI[1];
  df[1] = cs / rr * FI[0] + sn * FI[1];
  df[2] = FI[2];
  // second derivatives (id1,id2) = rows 4..12
  ddf[0][0] = sn2 / r2 * FI[0] + sn * sn / rr * FI[1] + sn * sn / r2 * FI[3] - cs * sn / rr * FI[4] -
              cs * sn / rr * FI[6] + cs * cs * FI[7];
  ddf[1][0] = -cs2 / r2 * FI[0] - cs * sn / rr * FI[1] - cs * sn / r2 * FI[3] - sn * sn / rr * FI[4] +
              cs * cs / rr * FI[6] + cs * sn * FI[7];
  ddf[2][0] = -sn / rr * FI[5] + cs * FI[8];
  ddf[0][1] = -cs2 / r2 * FI[0] - cs * sn / rr * FI[1] - cs * sn / r2 * FI[3] + cs * cs / rr * FI[4] -
              sn * sn / rr * FI[6] + cs * sn * FI[7];
  ddf[1][1] = -sn2 / r2 * FI[0] + cs * cs / rr * FI[1] + cs * cs / r2 * FI[3] + cs * sn / rr * FI[4] +
              cs * sn / rr * FI[6] + sn * sn * FI[7];
  ddf[2][1] = cs / rr * FI[5] + sn * FI[8];
  ddf[0][2] = -sn / rr * FI[9] + cs * FI[10];
  ddf[1][2] = cs / rr * FI[9] + sn * FI[10];
  ddf[2][2] = FI[11];
}

// rotation to Cartesian velocity, gradient and Hessian (:700-797) from the
// three components' (f, grad f, Hessian f) in fit rows kd*13 + (0 | 1..3 |
// 4..12), summed over components in order
template <typename T>
__device__ inline void lsq_rotate(double theta, double r, const T *fit, long cap, T vel[3], T dvel[9], T ddvel[27]) {
  const T sn = (T)sin(theta), cs = (T)cos(theta);
  const T rs = (T)fmax((double)0.000001f, r), rs2 = rs * rs;
  const T NM[3][3] = {{-sn, cs, (T)0}, {cs, sn, (T)0}, {(T)0, (T)0, (T)1}};
  T DN[3][3][3];  // DN[j][i][k] = d NM(i,k) / d x_j
#pragma unroll
  for (int q = 0; q < 27; q++) (&DN[0][0][0])[q] = (T)0;
  DN[0][0][0] = cs * sn / rs; DN[0][0][1] = sn * sn / rs;
  DN[0][1][0] = sn * sn / rs; DN[0][1][1] = -cs * sn / rs;
  DN[1][0][0] = -cs * cs / rs; DN[1][0][1] = -sn * cs / rs;
  DN[1][1][0] = -sn * cs / rs; DN[1][1][1] = cs * cs / rs;
  const T Aa = sn * (-(T)2 * (cs * cs) + sn * sn) / rs2;
  const T Bb = cs * ((cs * cs) - (T)2 * (sn * sn)) / rs2;
  const T Cc = -(T)3 * cs * (sn * sn) / rs2;
  const T Dd = (T)3 * (cs * cs) * sn / rs2;
  // DDN[l][j][i][k] = d2 NM(i,k) / dx_l dx_j; only i,k in {0,1} and l,j in {0,1}
  auto DDN = [&](int l, int j, int i, int k) -> T {
    if (l > 1 || j > 1 || i > 1 || k > 1) return (T)0;
    const int key = l * 8 + j * 4 + i * 2 + k;
    switch (key) {
      case 0: return Aa; case 1: return Cc; case 2: return Cc; case 3: return -Aa;
      case 4: return Bb; case 5: return -Aa; case 6: return -Aa; case 7: return -Bb;
      case 8: return Bb; case 9: return -Aa; case 10: return -Aa; case 11: return -Bb;
      case 12: return Dd; case 13: return -Bb; case 14: return -Bb; case 15: return -Dd;
    }
    return (T)0;
  };
#pragma unroll
  for (int q = 0; q < 3; q++) vel[q] = (T)0;
#pragma unroll
  for (int q = 0; q < 9; q++) dvel[q] = (T)0;
#pragma unroll
  for (int q = 0; q < 27; q++) ddvel[q] = (T)0;
#pragma unroll
  for (int kd = 0; kd < 3; kd++) {
    const T *fr = fit + (long)(kd * 13) * cap;
    const T f = fr[0];
    T df[3], ddf[3][3];
#pragma unroll
    for (int q = 0; q < 3; q++) df[q] = fr[(long)(1 + q) * cap];
#pragma unroll
    for (int q = 0; q < 9; q++) ddf[q % 3][q / 3] = fr[(long)(4 + q) * cap];
#pragma unroll
    for (int id = 0; id < 3; id++) vel[id] = vel[id] + f * NM[id][kd];
#pragma unroll
    for (int id = 0; id < 3; id++)
#pragma unroll
      for (int jd = 0; jd < 3; jd++)
        dvel[id + 3 * jd] = dvel[id + 3 * jd] + df[jd] * NM[id][kd] + f * DN[jd][id][kd];
#pragma unroll
    for (int id = 0; id < 3; id++)
#pragma unroll
      for (int jd = 0; jd < 3; jd++)
#pragma unroll
        for (int ld = 0; ld < 3; ld++)
          ddvel[id + 3 * jd + 9 * ld] = ddvel[id + 3 * jd + 9 * ld] + ddf[jd][ld] * NM[id][kd] +
                                        df[jd] * DN[ld][id][kd] + df[ld] * DN[jd][id][kd] + f * DDN(ld, jd, id, kd);
  }
}

// one substep, part 1: the fits of live particle p into 39 rows fit[v fs]
// (a particle whose stencil leaves the arrays is removed, part_tools.F90:
// 519-538, and false returned)
template <typename T>
__device__ inline bool lsq_fit_particle(const PartArgs &a, const T *mtl, long p, T *fit, long fs) {
  const long cap = a.cap;
  double *u = a.u + p;
#define U(v) u[(long)(v) * cap]
  double theta, r;
  int it, ir, iz;
  const double zp = U(2);
  if (!lsq_locate(a, U(0), U(1), zp, theta, r, it, ir, iz)) {
    a.alive[p] = 0;
    for (int v = 0; v < a.nvar; v++) U(v) = 0.0;
    return false;
  }
  const Trig<T> tg{(T)sin(theta), (T)cos(theta), (T)sin(2.0 * theta), (T)cos(2.0 * theta)};
#pragma unroll 1
  for (int kd = 0; kd < 3; kd++) {
    T f, df[3], ddf[3][3];
    lsq_component<T>(a, mtl, kd + 1, theta, r, zp, it, ir, iz, tg, f, df, ddf);
    T *fr = fit + (long)(kd * 13) * fs;
    fr[0] = f;
#pragma unroll
    for (int q = 0; q < 3; q++) fr[(long)(1 + q) * fs] = df[q];
#pragma unroll
    for (int q = 0; q < 9; q++) fr[(long)(4 + q) * fs] = ddf[q % 3][q / 3];
  }
#undef U
  return true;
}

// one substep, part 2: rotation of the fit rows fit[v fs], rates and state
// update of live particle p (part_tools.F90:700-930), then removal of
// particles that left the domain (:962-1025)
template <typename T>
__device__ inline void lsq_step_particle(const PartArgs &a, long p, const T *fit, long fs) {
  const long cap = a.cap;
  double *u = a.u + p;      // u[v*cap]
  double *d2 = a.ddt2 ? a.ddt2 + p : nullptr;
#define U(v) u[(long)(v) * cap]
  {
    T vel[3], dvel[9], ddvel[27];
    {
      const double xp = U(0), yp = U(1);
      const double r = sqrt(xp * xp + yp * yp);
      const double pi = (double)4.0f * atan(1.0);
      double theta;
      if (xp >= 0.0 && yp >= 0.0) theta = acos(xp / r);
      else if (xp < 0.0 && yp >= 0.0) theta = acos(xp / r);
      else theta = acos(-xp / r) + pi;
      lsq_rotate<T>(theta, r, fit, fs, vel, dvel, ddvel);
    }
    // rates (part_tools.F90:836-870) and the update of states 1..78
    // (part_tools.F90:926-930 Euler; dertest RK3 form), state by state: every
    // rate reads only old states, so the Hessian rows go first (they read the
    // old Jacobian, kept in J), then the Jacobian rows, then position and
    // velocity; no array of 78 rates is kept
    const double inv = (double)1.0f / a.st;
    const bool euler = a.scheme == VRH_PART_LSQ_EULER;
    // every state is read once (its old value is passed in), so the stores
    // never have to wait for re-reads of the same rows
    auto upd = [&](int v, double old, double rate) {
      if (euler) {
        U(v) = old + rate * a.partdt;
      } else {
        U(v) = old + a.gam * rate * a.partdt + a.rom * d2[(long)v * cap] * a.partdt;
        d2[(long)v * cap] = rate;
      }
    };
    double J[9];
    for (int q = 0; q < 9; q++) J[q] = U(6 + q);
    for (int k = 0; k < 3; k++)
      for (int j = 0; j < 3; j++) {
        double H[3], ra[3], rb[3];
        for (int l = 0; l < 3; l++) H[l] = U(24 + l + 3 * j + 9 * k);
        for (int i = 0; i < 3; i++) ra[i] = U(51 + i + 3 * j + 9 * k);
        for (int i = 0; i < 3; i++) {
          double sumh = 0.0;
          for (int l = 0; l < 3; l++) sumh += H[l] * (double)dvel[i + 3 * l];
          for (int m = 0; m < 3; m++)
            for (int l = 0; l < 3; l++) sumh += J[l + 3 * j] * J[m + 3 * k] * (double)ddvel[i + 3 * l + 9 * m];
          rb[i] = inv * (sumh - ra[i]);
        }
        for (int i = 0; i < 3; i++) {
          const int ix = i + 3 * j + 9 * k;
          upd(24 + ix, H[i], ra[i]);
          upd(51 + ix, ra[i], rb[i]);
        }
      }
    double x0[3], v0[3];
    for (int i = 0; i < 3; i++) { x0[i] = U(i); v0[i] = U(3 + i); }
    for (int j = 0; j < 3; j++) {
      double ra[3], rb[3];
      for (int i = 0; i < 3; i++) ra[i] = U(15 + i + 3 * j);
      for (int i = 0; i < 3; i++) {
        const double sumj =
            J[0 + 3 * j] * (double)dvel[i + 0] + J[1 + 3 * j] * (double)dvel[i + 3] + J[2 + 3 * j] * (double)dvel[i + 6];
        rb[i] = inv * (sumj - ra[i]);
      }
      for (int i = 0; i < 3; i++) {
        upd(6 + i + 3 * j, J[i + 3 * j], ra[i]);
        upd(15 + i + 3 * j, ra[i], rb[i]);
      }
    }
    for (int i = 0; i < 3; i++) upd(i, x0[i], v0[i]);
    for (int i = 0; i < 3; i++) upd(3 + i, v0[i], inv * ((double)vel[i] - v0[i]));
    if (euler) {
      for (int i = 0; i < 3; i++) U(78 + i) = vel[i];
      for (int i = 0; i < 9; i++) U(81 + i) = dvel[i];
      for (int i = 0; i < 27; i++) U(90 + i) = ddvel[i];
    }
    // removal of particles leaving the domain (part_tools.F90:962-1025)
    const double x = U(0), y = U(1), z = U(2), r = sqrt(x * x + y * y);
    if (z > a.zmax || z < 0.0 || r > a.rmax || r < 0.0 || !(r == r)) {
      a.alive[p] = 0;
      for (int v = 0; v < a.nvar; v++) U(v) = 0.0;
      if (d2) for (int v = 0; v < a.nvar; v++) d2[(long)v * cap] = 0.0;
      return;
    }
    }
#undef U
}

// the pseudo-inverse of a block in LDS, then the fits into global rows
template <typename T>
__global__ __launch_bounds__(128, sizeof(T) == 4 ? 3 : 2) void k_lsq_fit(PartArgs a, MintArg<T> mt) {
  __shared__ T mtl[243];
  for (int q = threadIdx.x; q < 243; q += blockDim.x) mtl[q] = mt.m[q];
  __syncthreads();
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.cap || p >= *a.n_dev || !a.alive[p]) return;
  lsq_fit_particle<T>(a, mtl, p, reinterpret_cast<T *>(a.fit) + p, a.cap);
}

template <typename T>
__global__ __launch_bounds__(128, 2) void k_lsq_step(PartArgs a) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.cap || p >= *a.n_dev || !a.alive[p]) return;
  lsq_step_particle<T>(a, p, reinterpret_cast<const T *>(a.fit) + p, a.cap);
}

// both parts in one kernel: the 39 fit rows of a particle stay in LDS
// (each thread reads back only its own rows), which spares their round trip
// through HBM and lets one wave's fit arithmetic overlap other waves' state
// traffic
template <typename T>
__global__ __launch_bounds__(128, 2) void k_lsq_substep(PartArgs a, MintArg<T> mt) {
  __shared__ T mtl[243];
  __shared__ T fl[kFitRows * 128];
  for (int q = threadIdx.x; q < 243; q += blockDim.x) mtl[q] = mt.m[q];
  __syncthreads();
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.cap || p >= *a.n_dev || !a.alive[p]) return;
  T *fit = fl + threadIdx.x;
  if (!lsq_fit_particle<T>(a, mtl, p, fit, 128)) return;
  lsq_step_particle<T>(a, p, fit, 128);
}

// ---- trilinear option: (u_theta, u_r, u_z) from the staggered grids ----
template <typename T>
__device__ T tri_comp(const PartArgs &a, int ivar, double theta, double r, double z) {
  const VGrid &g = a.g;
  const double th0 = (ivar == 1) ? 0.0 : 0.5 * a.dtheta;
  const double r0 = (ivar == 2) ? 0.0 : 0.5 * a.dr;
  const double z0 = (ivar == 3) ? 0.0 : 0.5 * a.dz;
  const double st = (theta - th0) / a.dtheta, sr = (r - r0) / a.dr, sz = (z - z0) / a.dz;
  const double ft = floor(st), fr = floor(sr), fz = floor(sz);
  double wt = st - ft, wr = sr - fr, wz = sz - fz;
  const int i0 = (int)ft;
  int j0 = (int)fr + 1, k0 = (int)fz + 1;
  if (j0 < 0) { j0 = 0; wr = 0.0; }
  if (j0 > g.nr) { j0 = g.nr; wr = 1.0; }
  if (k0 < 0) { k0 = 0; wz = 0.0; }
  if (k0 > g.nz + 1) { k0 = g.nz + 1; wz = 1.0; }
  T s = (T)0;
  for (int c = 0; c <= 1; c++)
    for (int b = 0; b <= 1; b++)
      for (int aa = 0; aa <= 1; aa++) {
        int ii = (i0 + aa) % g.nt;
        if (ii < 0) ii += g.nt;
        ii += 1;
        const int jj = j0 + b;
        int kk = k0 + c;
        if (kk > g.nz + 2) kk = g.nz + 2;
        const double w = (aa ? wt : 1.0 - wt) * (b ? wr : 1.0 - wr) * (c ? wz : 1.0 - wz);
        const long o = hidx(g, ii, jj, kk);
        double v;
        if (ivar == 1) v = a.q1[o];
        else if (ivar == 2) { const double rc = (jj - 1.0) * a.dr; v = (rc > 1e-12) ? a.q2[o] / rc : 0.0; }
        else v = a.q3[o];
        s += (T)(w * v);
      }
  return s;
}

template <typename T>
__global__ __launch_bounds__(256) void k_part_tri(PartArgs a) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.cap || p >= *a.n_dev || !a.alive[p]) return;
  const long cap = a.cap;
  // previous rates in the work type (float rows in mixed precision)
  double *s = a.u + p;
  T *r2 = reinterpret_cast<T *>(a.ddt2) + p;
#define S(v) s[(long)(v) * cap]
#define R(v) r2[(long)(v) * cap]
  const double x = S(0), y = S(1), z = S(2), r = sqrt(x * x + y * y);
  double theta = atan2(y, x);
  if (theta < 0) theta += 2.0 * M_PI;
  const double ut = (double)tri_comp<T>(a, 1, theta, r, z), ur = (double)tri_comp<T>(a, 2, theta, r, z),
               uz = (double)tri_comp<T>(a, 3, theta, r, z);
  const double sn = (r > 0) ? y / r : 0.0, cs = (r > 0) ? x / r : 1.0;
  const double u[3] = {-sn * ut + cs * ur, cs * ut + sn * ur, uz};
  double f[6];
  for (int i = 0; i < 3; i++) {
    if (a.st > 0.0) { f[i] = S(3 + i); f[3 + i] = (u[i] - S(3 + i)) / a.st; }
    else { f[i] = u[i]; f[3 + i] = 0.0; }
  }
  for (int i = 0; i < 6; i++) {
    S(i) = S(i) + a.gam * a.partdt * f[i] + a.rom * a.partdt * (double)R(i);
    R(i) = (T)f[i];
  }
  if (a.st <= 0.0) for (int i = 0; i < 3; i++) S(3 + i) = u[i];
  const double rn = sqrt(S(0) * S(0) + S(1) * S(1));
  if (S(2) > a.zmax || S(2) < 0.0 || rn > a.rmax) a.alive[p] = 0;
#undef S
#undef R
}

// ---------------------------------------------------------------------------
// Compaction of the live particles and inlet injection (inlt_part, injc_part
// and remv_part, part_tools.F90:95-225).  The reference keeps a free list of
// slots; here the live particles are packed to the front (prefix sums over
// the alive flags, order preserved) and new particles are appended, so the
// kernels only see slots [0, n).
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long psmix(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
// uniform in (0, 1) from a counter-based stream (splitmix64)
__device__ __forceinline__ double pu01(unsigned long long seed, unsigned long long i) {
  return ((psmix(seed ^ psmix(i)) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// nonzero flags among the first n entries, per block of 256
__global__ void k_alive_count(const int *__restrict__ alive, const long *n_dev, long cap, long *__restrict__ bsum) {
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  const long n = *n_dev;
  int x = (p < cap && p < n && alive[p]) ? 1 : 0;
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  __shared__ int red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  if (threadIdx.x == 0) bsum[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// in-place exclusive scan of m values by one block of 1024 threads; the
// total goes to *total
__global__ void k_scan1(long *v, long m, long *total) {
  __shared__ long part[1024];
  const int t = threadIdx.x;
  const long chunk = (m + 1023) / 1024, b0 = t * chunk, b1 = min(m, b0 + chunk);
  long s = 0;
  for (long q = b0; q < b1; q++) s += v[q];
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    long acc = 0;
    for (int q = 0; q < 1024; q++) { const long x = part[q]; part[q] = acc; acc += x; }
    *total = acc;
  }
  __syncthreads();
  long acc = part[t];
  for (long q = b0; q < b1; q++) { const long x = v[q]; v[q] = acc; acc += x; }
}

__global__ void k_set_count(long *n_dev, const long *total) { *n_dev = *total; }

// inlet flux per (r row, theta cell), in the reference's theta order per row:
// inflow(ir) += c a q3 dt inj / nprcl, n = floor(inflow), inflow -= n
__global__ void k_inflow(VGrid g, VCoef c, const double *__restrict__ q3, double *__restrict__ inflow, int jr0,
                         double dth, double dt, int inj, double cc, double nprcl, long *__restrict__ counts) {
  const int ir = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (ir > jr0) return;
  const double a = dth * (c.rc[ir + 1] * c.rc[ir + 1] - c.rc[ir] * c.rc[ir]);
  double acc = inflow[ir];
  for (int it = 1; it <= g.nt; it++) {
    acc = acc + (cc * a * q3[hidx(g, it, ir, 1)] * dt * (double)(float)inj / nprcl);
    const long nn = (long)floor(acc);
    acc = acc - (double)(float)nn;
    counts[(long)(ir - 1) * g.nt + (it - 1)] = nn > 0 ? nn : 0;
  }
  inflow[ir] = acc;
}

// write the new particles of each inlet cell: uniform in the cell, velocity
// (0, 0, q3), identity Jacobian (part_tools.F90:120-134, 164-181)
__global__ void k_inject(VGrid g, VCoef c, const double *__restrict__ q3, const long *__restrict__ counts,
                         const long *__restrict__ offs, const long *n_dev, long cap, int nvar, double *__restrict__ u,
                         double *__restrict__ d2, int d2w, int *__restrict__ alive, int *__restrict__ slot, int jr0, double dth,
                         double dr, double dz, unsigned long long seed, long call) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)jr0 * g.nt) return;
  const int ir = (int)(t / g.nt) + 1, it = (int)(t - (long)(ir - 1) * g.nt) + 1;
  const long cnt = counts[t], base = *n_dev + offs[t];
  const double uz = q3[hidx(g, it, ir, 1)];
  for (long q = 0; q < cnt; q++) {
    const long idx = base + q;
    if (idx >= cap) break;
    const unsigned long long key =
        ((((unsigned long long)call * (unsigned long long)((long)jr0 * g.nt) + (unsigned long long)t) << 20) +
         (unsigned long long)q) * 3ull;
    const double th = (it - 1) * dth + dth * pu01(seed, key);
    const double r = c.rc[ir] + dr * pu01(seed, key + 1);
    const double z = 0.0 + dz * pu01(seed, key + 2);
    for (int v = 0; v < nvar; v++) u[(long)v * cap + idx] = 0.0;
    if (d2 && d2w == 8)
      for (int v = 0; v < nvar; v++) d2[(long)v * cap + idx] = 0.0;
    if (d2 && d2w == 4)
      for (int v = 0; v < nvar; v++) reinterpret_cast<float *>(d2)[(long)v * cap + idx] = 0.0f;
    u[idx] = r * cos(th);
    u[cap + idx] = r * sin(th);
    u[2 * cap + idx] = z;
    u[5 * cap + idx] = uz;
    if (nvar == 117) { u[6 * cap + idx] = 1.0; u[10 * cap + idx] = 1.0; u[14 * cap + idx] = 1.0; }
    alive[idx] = 1;
    slot[idx] = (int)idx;  // appended after a compaction: physical and shown order agree
  }
}

__global__ void k_add_count(long *n_dev, const long *total, long cap) {
  const long n = *n_dev + *total;
  *n_dev = n < cap ? n : cap;
}

// ---------------------------------------------------------------------------
// Spatial order.  The particle kernels run over physical slots; sorting the
// slots by cell (theta fastest, then r, then z) makes the 8- and 27-point
// gathers of one wavefront fall on a few neighbouring rows of the fields.
// The order the API shows (and the reference keeps) travels in slot[]:
// physical slot p holds the particle of shown slot slot[p].
// ---------------------------------------------------------------------------
__global__ void k_iota(int *__restrict__ v, long cap) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < cap) v[p] = (int)p;
}

// cell key of each live particle; dead and unused slots take kmax, so the
// stable sort keeps them behind the live ones in their present order
__global__ void k_sort_keys(const double *__restrict__ u, const int *__restrict__ alive, const long *n_dev, long cap,
                            int nt, int nr, int nz, double dth, double dr, double dz, unsigned kmax,
                            unsigned *__restrict__ key, unsigned *__restrict__ val) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= cap) return;
  unsigned k = kmax;
  if (p < *n_dev && alive[p]) {
    const double x = u[p], y = u[cap + p], z = u[2 * cap + p];
    double th = atan2(y, x);
    if (th < 0.0) th += 2.0 * M_PI;
    // fmin/fmax clamp, and map a NaN coordinate to 0
    const int i = (int)fmin(fmax(th / dth, 0.0), nt - 1.0);
    const int j = (int)fmin(fmax(sqrt(x * x + y * y) / dr, 0.0), nr + 1.0);
    const int kz = (int)fmin(fmax(z / dz, 0.0), nz + 1.0);
    k = ((unsigned)kz * (unsigned)(nr + 2) + (unsigned)j) * (unsigned)nt + (unsigned)i;
  }
  key[p] = k;
  val[p] = (unsigned)p;
}

// physical slot i <- slot order[i]: state rows, previous rates, alive flag
// and shown slot.  Compacting (rank != nullptr): the first nlive slots are
// the live particles, their shown slots renumbered to their ranks; the rest
// are unused, numbered as their position
__global__ void k_permute(const unsigned *__restrict__ order, const double *__restrict__ u,
                          const double *__restrict__ d2, const int *__restrict__ alive, const int *__restrict__ slot,
                          const int *__restrict__ rank, const long *nlive, long cap, int nvar, int d2w,
                          double *__restrict__ ut, double *__restrict__ d2t, int *__restrict__ at,
                          int *__restrict__ st) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cap) return;
  const long src = order[i];
  for (int v = 0; v < nvar; v++) ut[(long)v * cap + i] = u[(long)v * cap + src];
  if (d2 && d2w == 8)
    for (int v = 0; v < nvar; v++) d2t[(long)v * cap + i] = d2[(long)v * cap + src];
  if (d2 && d2w == 4) {
    const float *f2 = reinterpret_cast<const float *>(d2);
    float *f2t = reinterpret_cast<float *>(d2t);
    for (int v = 0; v < nvar; v++) f2t[(long)v * cap + i] = f2[(long)v * cap + src];
  }
  if (rank) {
    const bool live = i < *nlive;
    at[i] = live ? 1 : 0;
    st[i] = live ? rank[slot[src]] : (int)i;
  } else {
    at[i] = alive[src];
    st[i] = slot[src];
  }
}

// flag[s] = 1 for the shown slots s < n whose particle is alive (flag zeroed
// before)
__global__ void k_slot_flags(const int *__restrict__ alive, const int *__restrict__ slot, const long *n_dev, long cap,
                             int *__restrict__ flag) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < cap && p < *n_dev && alive[p]) flag[slot[p]] = 1;
}

// exclusive rank of each flagged shown slot: block offset boff (scan of the
// per-block counts) plus the rank within the block of 256
__global__ void k_slot_rank(const int *__restrict__ flag, const long *n_dev, long cap, const long *__restrict__ boff,
                            int *__restrict__ rank) {
  const long s = (long)blockIdx.x * 256 + threadIdx.x;
  const int f = (s < cap && s < *n_dev && flag[s]) ? 1 : 0;
  __shared__ int cnt[256];
  cnt[threadIdx.x] = f;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {  // inclusive scan by doubling
    const int x = ((int)threadIdx.x >= o) ? cnt[threadIdx.x - o] : 0;
    __syncthreads();
    cnt[threadIdx.x] += x;
    __syncthreads();
  }
  if (s < cap) rank[s] = (int)(boff[blockIdx.x] + cnt[threadIdx.x] - f);
}

}  // namespace

bool launch_particles(vrh_t *h, int irk, hipStream_t st) {
  if (h->nbatch <= 0) return false;
  const bool every = h->pscheme != VRH_PART_LSQ_EULER;
  if (!every && irk != 1) return false;  // loop.F90:67 calls iter_part at irk==1 only
  KTimer kt(h, KT_PARTICLES);
  for (auto &b : h->pb) {
    PartArgs a{};
    a.g = h->g;
    a.q1 = h->f.q1; a.q2 = h->f.q2; a.q3 = h->f.q3;
    a.mint = h->d_mint;
    a.u = b.u; a.ddt2 = b.ddt2; a.alive = b.alive;
    a.cap = b.cap; a.n = b.cap; a.n_dev = b.d_n;
    a.nvar = h->pscheme == VRH_PART_TRILINEAR_RK3 ? 6 : 117;
    a.nsub = b.nsub; a.scheme = h->pscheme; a.irk = irk;
    a.st = b.st;
    a.partdt = h->dt / (double)(float)b.nsub;  // partdt = dt/float(nsubdt), part_tools.F90:505
    a.gam = h->rk[irk].g; a.rom = h->rk[irk].r;
    a.dtheta = 2.0 * acos(-1.0) / h->nt; a.dr = h->rmax / h->nr; a.dz = h->zmax / h->nz;
    a.rmax = h->rmax; a.zmax = h->zmax;
    if (h->pscheme == VRH_PART_TRILINEAR_RK3) {
      a.partdt = h->dt;
      a.nsub = 1;
      dim3 blk(256), grd((unsigned)((b.cap + 255) / 256));
      if (h->prec == VRH_MIXED) hipLaunchKernelGGL(k_part_tri<float>, grd, blk, 0, st, a);
      else hipLaunchKernelGGL(k_part_tri<double>, grd, blk, 0, st, a);
    } else {
      a.fit = b.fit;
      a.nsub = 1;
      dim3 blk(128), grd((unsigned)((b.cap + 127) / 128));
      MintArg<float> mf;
      MintArg<double> md;
      for (int q = 0; q < 243; q++) { mf.m[q] = (float)h->h_mint[q]; md.m[q] = h->h_mint[q]; }
      // mixed: fit and update in one kernel (fit rows in LDS, 2 % faster);
      // FP64: two kernels (the fused one spills at 256 VGPRs, 3 % slower)
      for (int isub = 0; isub < b.nsub; isub++) {
        if (h->prec == VRH_MIXED) {
          hipLaunchKernelGGL(k_lsq_substep<float>, grd, blk, 0, st, a, mf);
        } else {
          hipLaunchKernelGGL(k_lsq_fit<double>, grd, blk, 0, st, a, md);
          hipLaunchKernelGGL(k_lsq_step<double>, grd, blk, 0, st, a);
        }
      }
    }
  }
  return true;
}

void particles_release(vrh_t *h) {
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (auto &b : h->pb) {
    void *bufs[] = {b.u, b.fit, b.ddt2, b.alive, b.d_n, b.inflow, b.u_tmp, b.ddt2_tmp, b.alive_tmp, b.work,
                    b.slot, b.slot_tmp, b.lflag, b.lrank, b.skeys, b.stmp};
    for (void *q : bufs) if (q) (void)hipFree(q);
  }
  h->pb.clear();
  h->nbatch = 0;
}

extern "C" {

static int sort_batch(vrh_t *h, PartBatch &B, bool compact);

int vrh_particles_create(vrh_t *h, int nbatch, const long *cap, const double *st, const int *nsub, int scheme) {
  if (h) graph_invalidate(h);
  if (!h || nbatch < 1 || !cap || !st || !nsub) return -1;
  if (scheme < VRH_PART_LSQ_EULER || scheme > VRH_PART_TRILINEAR_RK3) { h->err = "unknown particle scheme"; return -1; }
  if (!h->pb.empty()) { h->err = "particles already created"; return -1; }
  if (!h->mesh[0].empty()) { h->err = "particles need the uniform radial mesh"; return -1; }
  if (!h->zmesh[0].empty()) { h->err = "particles need the uniform axial mesh"; return -1; }
  const int nvar = scheme == VRH_PART_TRILINEAR_RK3 ? 6 : 117;
  for (int b = 0; b < nbatch; b++)  // every batch is checked before anything is allocated
    if (cap[b] < 1 || cap[b] > 0x7fffffffL || (scheme != VRH_PART_TRILINEAR_RK3 && !(st[b] > 0.0))) {
      h->err = "bad particle batch";
      return -1;
    }
  h->pscheme = scheme;
  h->pb.resize(nbatch);
  // an allocation failure releases what was allocated, so the call can be
  // retried; the failed hipMalloc's error is cleared so that the next launch
  // check does not report it
  auto oom = [&]() {
    (void)hipGetLastError();
    particles_release(h);
    h->err = "particle allocation failed";
    return -3;
  };
  for (int b = 0; b < nbatch; b++) {
    PartBatch &B = h->pb[b];
    B.cap = cap[b]; B.n = 0; B.st = st[b]; B.nsub = nsub[b] < 1 ? 1 : nsub[b];
    if (hipMalloc(&B.u, sizeof(double) * nvar * B.cap) != hipSuccess ||
        hipMalloc(&B.alive, sizeof(int) * B.cap) != hipSuccess) return oom();
    hipMemsetAsync(B.u, 0, sizeof(double) * nvar * B.cap, h->stream);
    hipMemsetAsync(B.alive, 0, sizeof(int) * B.cap, h->stream);
    if (hipMalloc(&B.d_n, sizeof(long)) != hipSuccess ||
        hipMalloc(&B.slot, sizeof(int) * B.cap) != hipSuccess) return oom();
    hipMemsetAsync(B.d_n, 0, sizeof(long), h->stream);
    hipLaunchKernelGGL(k_iota, dim3((unsigned)((B.cap + 255) / 256)), dim3(256), 0, h->stream, B.slot, B.cap);
    B.ddt2 = nullptr;
    B.fit = nullptr;
    // global fit rows: the FP64 LSQ path only (the mixed path keeps them in LDS)
    if (scheme != VRH_PART_TRILINEAR_RK3 && h->prec != VRH_MIXED &&
        hipMalloc(&B.fit, sizeof(double) * kFitRows * B.cap) != hipSuccess) return oom();
    if (scheme != VRH_PART_LSQ_EULER) {
      B.d2w = (scheme == VRH_PART_TRILINEAR_RK3 && h->prec == VRH_MIXED) ? 4 : 8;
      if (hipMalloc(&B.ddt2, (size_t)B.d2w * nvar * B.cap) != hipSuccess) return oom();
      hipMemsetAsync(B.ddt2, 0, (size_t)B.d2w * nvar * B.cap, h->stream);
    }
  }
  h->nbatch = nbatch;
  return hipStreamSynchronize(h->stream) == hipSuccess ? 0 : -3;
}

int vrh_particles_upload(vrh_t *h, int batch, long n, const double *u, const int *alive) {
  if (h) graph_invalidate(h);
  if (!h || batch < 0 || batch >= h->nbatch || !u || n < 0) return -1;
  PartBatch &B = h->pb[batch];
  if (n > B.cap) { h->err = "more particles than the batch capacity"; return -1; }
  const int nvar = h->pscheme == VRH_PART_TRILINEAR_RK3 ? 6 : 117;
  std::vector<double> soa((size_t)nvar * B.cap, 0.0);
  for (long p = 0; p < n; p++)
    for (int v = 0; v < nvar; v++) soa[(size_t)v * B.cap + p] = u[(size_t)p * nvar + v];
  std::vector<int> al(B.cap, 0);
  for (long p = 0; p < n; p++) al[p] = alive ? alive[p] : 1;
  // in order on the handle's (non-blocking) stream, behind any step still
  // running there; the host vectors live until the stream has drained
  bool ok = hipMemcpyAsync(B.u, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice, h->stream) ==
                hipSuccess &&
            hipMemcpyAsync(B.alive, al.data(), al.size() * sizeof(int), hipMemcpyHostToDevice, h->stream) ==
                hipSuccess &&
            hipMemcpyAsync(B.d_n, &n, sizeof(long), hipMemcpyHostToDevice, h->stream) == hipSuccess;
  if (ok && B.ddt2) ok = hipMemsetAsync(B.ddt2, 0, (size_t)B.d2w * nvar * B.cap, h->stream) == hipSuccess;
  if (ok) hipLaunchKernelGGL(k_iota, dim3((unsigned)((B.cap + 255) / 256)), dim3(256), 0, h->stream, B.slot, B.cap);
  if (hipStreamSynchronize(h->stream) != hipSuccess || !ok) {
    h->err = "particle upload failed";
    return -3;
  }
  B.n = n;
  return sort_batch(h, B, false);
}

int vrh_particles_download(vrh_t *h, int batch, long n, double *u, int *alive) {
  if (!h || batch < 0 || batch >= h->nbatch || !u || n < 0) return -1;
  PartBatch &B = h->pb[batch];
  if (n > B.cap) n = B.cap;
  const int nvar = h->pscheme == VRH_PART_TRILINEAR_RK3 ? 6 : 117;
  std::vector<double> soa((size_t)nvar * B.cap);
  std::vector<int> al(B.cap), sl(B.cap);
  if (hipMemcpyAsync(soa.data(), B.u, soa.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
      hipMemcpyAsync(al.data(), B.alive, al.size() * sizeof(int), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
      hipMemcpyAsync(sl.data(), B.slot, sl.size() * sizeof(int), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
      hipStreamSynchronize(h->stream) != hipSuccess) {
    h->err = "particle download failed";
    return -3;
  }
  // physical slot p holds shown slot sl[p] (a permutation of [0, cap))
  for (long p = 0; p < B.cap; p++) {
    const long q = sl[p];
    if (q < 0 || q >= B.cap) { h->err = "particle slot map corrupt"; return -3; }
    if (q >= n) continue;
    for (int v = 0; v < nvar; v++) u[(size_t)q * nvar + v] = soa[(size_t)v * B.cap + p];
    if (alive) alive[q] = al[p];
  }
  return 0;
}

// compaction and sort buffers of a batch (allocated on first use)
static int particle_buffers(vrh_t *h, PartBatch &B) {
  if (B.u_tmp) return 0;
  const int nvar = h->pscheme == VRH_PART_TRILINEAR_RK3 ? 6 : 117;
  const long nblk = (B.cap + 255) / 256;
  // u_tmp is allocated last: it marks the whole set as present.  A failure
  // frees the part that was allocated (and clears the failed hipMalloc's
  // error), so a later call starts again from nothing.
  if (hipMalloc(&B.alive_tmp, sizeof(int) * B.cap) != hipSuccess ||
      hipMalloc(&B.work, sizeof(long) * (nblk + 8 + (long)(h->nr + 1) * h->nt * 2)) != hipSuccess ||
      hipMalloc(&B.inflow, sizeof(double) * (h->nr + 2)) != hipSuccess ||
      hipMalloc(&B.slot_tmp, sizeof(int) * B.cap) != hipSuccess ||
      hipMalloc(&B.lflag, sizeof(int) * B.cap) != hipSuccess ||
      hipMalloc(&B.lrank, sizeof(int) * B.cap) != hipSuccess ||
      hipMalloc(&B.skeys, sizeof(unsigned) * 4 * B.cap) != hipSuccess ||
      (B.ddt2 && hipMalloc(&B.ddt2_tmp, (size_t)B.d2w * nvar * B.cap) != hipSuccess) ||
      hipMalloc(&B.u_tmp, sizeof(double) * nvar * B.cap) != hipSuccess) {
    (void)hipGetLastError();
    void **bufs[] = {(void **)&B.alive_tmp, (void **)&B.work, (void **)&B.inflow, (void **)&B.slot_tmp,
                     (void **)&B.lflag, (void **)&B.lrank, (void **)&B.skeys, (void **)&B.ddt2_tmp,
                     (void **)&B.u_tmp};
    for (void **q : bufs)
      if (*q) { (void)hipFree(*q); *q = nullptr; }
    h->err = "particle allocation failed";
    return -3;
  }
  hipMemsetAsync(B.inflow, 0, sizeof(double) * (h->nr + 2), h->stream);
  return 0;
}

// Sort the physical slots of batch b by cell (stable: particles of one cell,
// and the dead and unused slots, keep their order).  compact: drop the dead
// particles as well, renumbering the shown slots of the live ones to their
// ranks (the reference's order, remv_part/injc_part, part_tools.F90:141-225)
// and setting the count to the live particles.  Device only.
static int sort_batch(vrh_t *h, PartBatch &B, bool compact) {
  int rc = particle_buffers(h, B);
  if (rc) return rc;
  const int nvar = h->pscheme == VRH_PART_TRILINEAR_RK3 ? 6 : 117;
  const long cap = B.cap, nblk = (cap + 255) / 256;
  long *bsum = B.work, *total = B.work + nblk;
  hipStream_t st = h->stream;
  bool ok = true;
  if (compact) {
    ok = hipMemsetAsync(B.lflag, 0, sizeof(int) * cap, st) == hipSuccess;
    hipLaunchKernelGGL(k_slot_flags, dim3((unsigned)nblk), dim3(256), 0, st, B.alive, B.slot, B.d_n, cap, B.lflag);
    hipLaunchKernelGGL(k_alive_count, dim3((unsigned)nblk), dim3(256), 0, st, B.lflag, B.d_n, cap, bsum);
    hipLaunchKernelGGL(k_scan1, dim3(1), dim3(1024), 0, st, bsum, nblk, total);
    hipLaunchKernelGGL(k_slot_rank, dim3((unsigned)nblk), dim3(256), 0, st, B.lflag, B.d_n, cap, bsum, B.lrank);
  }
  // keys: cells (nt x (nr + 2) x (nz + 2)), kmax for dead and unused slots
  const unsigned kmax = (unsigned)((long)h->nt * (h->nr + 2) * (h->nz + 2));
  int bits = 0;
  while (bits < 32 && (kmax >> bits) != 0u) bits++;
  unsigned *kin = B.skeys, *kout = kin + cap, *vin = kout + cap, *vout = vin + cap;
  const double dth = 2.0 * acos(-1.0) / h->nt, dr = h->rmax / h->nr, dz = h->zmax / h->nz;
  hipLaunchKernelGGL(k_sort_keys, dim3((unsigned)nblk), dim3(256), 0, st, B.u, B.alive, B.d_n, cap, h->nt, h->nr,
                     h->nz, dth, dr, dz, kmax, kin, vin);
  if (!B.stmp) {
    size_t bytes = 0;
    if (rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, (unsigned)cap, 0u, (unsigned)bits, st) !=
            hipSuccess ||
        hipMalloc(&B.stmp, bytes ? bytes : 1) != hipSuccess) {
      (void)hipGetLastError();
      B.stmp = nullptr;
      h->err = "particle sort allocation failed";
      return -3;
    }
    B.stmp_bytes = bytes;
  }
  size_t bytes = B.stmp_bytes;
  if (rocprim::radix_sort_pairs(B.stmp, bytes, kin, kout, vin, vout, (unsigned)cap, 0u, (unsigned)bits, st) !=
      hipSuccess) {
    h->err = "particle sort failed";
    return -3;
  }
  hipLaunchKernelGGL(k_permute, dim3((unsigned)nblk), dim3(256), 0, st, vout, B.u, B.ddt2, B.alive, B.slot,
                     compact ? B.lrank : nullptr, total, cap, nvar, B.d2w, B.u_tmp, B.ddt2_tmp, B.alive_tmp, B.slot_tmp);
  ok = ok && hipMemcpyAsync(B.u, B.u_tmp, sizeof(double) * nvar * cap, hipMemcpyDeviceToDevice, st) == hipSuccess;
  if (B.ddt2)
    ok = ok && hipMemcpyAsync(B.ddt2, B.ddt2_tmp, (size_t)B.d2w * nvar * cap, hipMemcpyDeviceToDevice, st) ==
                   hipSuccess;
  ok = ok && hipMemcpyAsync(B.alive, B.alive_tmp, sizeof(int) * cap, hipMemcpyDeviceToDevice, st) == hipSuccess;
  ok = ok && hipMemcpyAsync(B.slot, B.slot_tmp, sizeof(int) * cap, hipMemcpyDeviceToDevice, st) == hipSuccess;
  if (compact) hipLaunchKernelGGL(k_set_count, dim3(1), dim3(1), 0, st, B.d_n, total);
  if (!ok || hipGetLastError() != hipSuccess) { h->err = "particle sort failed"; return -3; }
  return 0;
}

// pack the live particles of batch b to the front (device only)
static int compact_batch(vrh_t *h, PartBatch &B) { return sort_batch(h, B, true); }

int vrh_particles_compact(vrh_t *h, int batch) {
  if (!h || batch < 0 || batch >= h->nbatch) return -1;
  return compact_batch(h, h->pb[batch]);
}

int vrh_particles_sort(vrh_t *h, int batch) {
  if (!h || batch < 0 || batch >= h->nbatch) return -1;
  return sort_batch(h, h->pb[batch], false);
}

int vrh_particles_count(vrh_t *h, int batch, long *n) {
  if (!h || batch < 0 || batch >= h->nbatch || !n) return -1;
  if (hipMemcpyAsync(n, h->pb[batch].d_n, sizeof(long), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
      hipStreamSynchronize(h->stream) != hipSuccess) {
    h->err = "particle count failed";
    return -3;
  }
  return 0;
}

int vrh_particles_inject(vrh_t *h, int batch, double r0, int inj, unsigned long long seed, long *ninj) {
  if (!h || batch < 0 || batch >= h->nbatch || inj < 1) return -1;
  PartBatch &B = h->pb[batch];
  int rc = compact_batch(h, B);
  if (rc) return rc;
  const int nvar = h->pscheme == VRH_PART_TRILINEAR_RK3 ? 6 : 117;
  // inlet rows: rc(ir) < r0 (part_tools.F90:107)
  int jr0 = 0;
  const double dr = h->rmax / h->nr;
  for (int ir = 1; ir <= h->nr; ir++)
    if ((ir - 1) * dr < r0) jr0 = ir;
  if (jr0 == 0) { if (ninj) *ninj = 0; return 0; }
  const long nblk = (B.cap + 255) / 256;
  long *total = B.work + nblk + 1, *counts = B.work + nblk + 8, *offs = counts + (long)jr0 * h->nt;
  const double dth = 2.0 * acos(-1.0) / h->nt, dz = h->zmax / h->nz;
  // c = 10^6 * 0.02^3 and nprcl = 0.001 in single precision (part_tools.F90:275-276)
  const float c02 = 0.02f;
  const double cc = (double)(1000000.0f * ((c02 * c02) * c02)), nprcl = (double)0.001f;
  hipLaunchKernelGGL(k_inflow, dim3((unsigned)((jr0 + 63) / 64)), dim3(64), 0, h->stream, h->g, h->coef, h->f.q3,
                     B.inflow, jr0, dth, h->dt, inj, cc, nprcl, counts);
  hipMemcpyAsync(offs, counts, sizeof(long) * jr0 * h->nt, hipMemcpyDeviceToDevice, h->stream);
  hipLaunchKernelGGL(k_scan1, dim3(1), dim3(1024), 0, h->stream, offs, (long)jr0 * h->nt, total);
  const long ncell = (long)jr0 * h->nt;
  hipLaunchKernelGGL(k_inject, dim3((unsigned)((ncell + 127) / 128)), dim3(128), 0, h->stream, h->g, h->coef,
                     h->f.q3, counts, offs, B.d_n, B.cap, nvar, B.u, B.ddt2, B.d2w, B.alive, B.slot, jr0, dth, dr, dz,
                     seed, B.ncall);
  hipLaunchKernelGGL(k_add_count, dim3(1), dim3(1), 0, h->stream, B.d_n, total, B.cap);
  B.ncall++;
  if (hipGetLastError() != hipSuccess) { h->err = "particle injection failed"; return -3; }
  if (ninj) {
    if (hipMemcpyAsync(ninj, total, sizeof(long), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess) {
      h->err = "particle injection failed";
      return -3;
    }
  }
  return 0;
}

int vrh_particles_step(vrh_t *h, int irk) {
  if (!h || irk < 1 || irk > 3) return -1;
  if (h->nbatch <= 0) { h->err = "no particles"; return -1; }
  launch_particles(h, irk, h->stream);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

}  // extern "C"
