// vrh_kernels.hip -- gfx950 kernels of the vring RK3 substep.
//
// All kernels read and write HBM-resident fields in the reference's Fortran
// layout (theta fastest), so 64-lane wavefronts always span contiguous theta
// rows and every global access is coalesced.  TW is the work type (double in
// the FP64 path, float in the mixed path); the state q, the predictor dq,
// phi and p are always double.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cmath>
#include <mutex>

#include "vrh_internal.h"

namespace {

template <typename T> __device__ __forceinline__ T ldg(const T *p) { return __builtin_nontemporal_load(p); }

// Work-group order across the eight XCDs: the dispatcher hands work group p
// (linear id) to XCD p mod 8, and each XCD has its own L2.  xcd_tile maps p
// to a logical tile so that XCD x runs one contiguous range of tiles, in
// order, one dispatch round apart: neighbouring tiles meet in one L2.  A
// bijection on [0, nblk) for any nblk.
__device__ __forceinline__ long xcd_tile(long b, long nblk) {
  constexpr long kX = 8;
  const long q = nblk / kX, rem = nblk % kX, x = b % kX, r = b / kX;
  return x < rem ? x * (q + 1) + r : rem * (q + 1) + (x - rem) * q + r;
}
// The block's logical (x, y, z) coordinates of a 2-D or 3-D grid in the
// order of xcd_tile, x fastest: tiles adjacent in x (whose halo rows
// overlap) share an L2.
__device__ __forceinline__ void xcd_block(int &bx, int &by, int &bz) {
  const long nx = gridDim.x, ny = gridDim.y, nb = nx * ny * (long)gridDim.z;
  const long l = xcd_tile((long)blockIdx.x + nx * ((long)blockIdx.y + ny * (long)blockIdx.z), nb);
  bx = (int)(l % nx);
  const long r = l / nx;
  by = (int)(r % ny);
  bz = (int)(r / ny);
}

// ---------------------------------------------------------------------------
// hdnbc: outlet convective condition and mass correction, one thread per
// (theta, r) outlet point (momentum_terms_swirl.F90:225-291).  zm3n: the
// outlet cell's z difference dx3/dmz(nz) on an axial mesh, 0 on the uniform one.
// ---------------------------------------------------------------------------
__global__ void k_hdnbc(VGrid g, VCoef c, double zm3n, double gam, double rom, double cor, const double *__restrict__ cor_dev,
                        const double *__restrict__ q3, double *__restrict__ csort, double *__restrict__ dq3no,
                        double *__restrict__ corr3n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y + 1;  // 1..nrc
  if (i > g.nt) return;
  if (cor_dev) cor = *cor_dev;
  long p = (long)(i - 1) + (long)g.nt * j;
  double qn = q3[hidx(g, i, j, g.nzc)], qm = q3[hidx(g, i, j, g.nz)];
  double cs = csort[p];
  if (j <= g.nr && qn > 0.0) cs = (double)0.6f * qn;
  csort[p] = cs;
  double tconv = -(cs * (zm3n > 0.0 ? zm3n : g.dx3)) * (qn - qm);
  double c3n = (gam * tconv + rom * dq3no[p]) * g.dt;
  dq3no[p] = tconv;
  corr3n[p] = c3n + c.qc[j] * cor;
}

// Inlet time profile (gen_tools_swirl.F90:159-175): smooth ramp to 1 over
// 0.1n, hold until n, ramp down (n = 2).
__device__ __host__ inline double profile_fn(double t) {
  const double n = 2.0, t1 = (double)0.1f * n;
  double u = 0.0;
  if (t > 0.0 && t < t1) { double x = t / t1; u = 3.0 * (x * x) - 2.0 * (x * x * x); }
  if (t > t1 && t < n) u = 1.0;
  if (t > n && t < n + t1) { double x = (n + t1 - t) / t1; u = 3.0 * (x * x) - 2.0 * (x * x * x); }
  if (t > n + t1) u = 0.0;
  return u;
}

// Per-step scalars (loop.F90:53; momentum_terms_swirl.F90:192-212, 261-283):
// advance ntime by dt, then for each substep the inlet increment
// c3s = (profile(t+dt) - profile(t)) alpha and the mass-correction factor
// cor = c3s * inlet_area / (sum qc^2 da + 1e-15).
__global__ void k_scalars(double *sc, double dt, double a1, double a2, double a3, double inlet_area,
                          double cor_red, int advance, double t_set) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double t = advance ? dt + sc[0] : t_set;
  sc[0] = t;
  const double u0 = profile_fn(t), u1 = profile_fn(t + dt);
  const double al[3] = {a1, a2, a3};
  for (int q = 0; q < 3; q++) {
    const double c3s = (u1 - u0) * al[q];
    sc[1 + q] = c3s;
    sc[4 + q] = (c3s * inlet_area - 0.0) / (cor_red + (double)1.e-15f);
  }
}

// ---------------------------------------------------------------------------
// Explicit part: hdnl1-3 (momentum_terms_swirl.F90:15-174) fused with the
// explicit right-hand sides of step (step.F90:32-165) and the periodic theta
// solves solqi/tripij (step.F90:212-263, 487-534).
// A block owns whole theta rows (NTP >= nt lanes) of RJ consecutive r rows
// and marches up a chunk of z planes.  The planes k-1, k, k+1 of q1, q2, q3
// (rows j0-1 .. j0+RJ) sit in a 3-slot LDS ring, so every stencil operand
// comes from LDS; the next plane is prefetched into registers while the
// current one is computed.  The theta systems (constant coefficients
// (-acc, 1+2acc, -acc) per r row) are solved by parallel cyclic reduction
// on LDS-resident lines; the reduction ratios of each row are computed once
// per block.  The inlet planes q1s, q2s, q3s are zero
// (momentum_terms_swirl.F90:216-223).
// ---------------------------------------------------------------------------
constexpr int kMaxPcrLevels = 40;
constexpr int kRowCoefs = 24;

// z second differences on the axial mesh (ZMet, vrh_set_axial_mesh): at a
// cell centre (q1, q2, p, the first two pressure-gradient components) with
// the one-sided form through the wall value 0 in the inlet cell, and at a
// face (q3, the third component)
template <typename T> __device__ __forceinline__ T zd2c(const ZMet &z, int k, T fm1, T f0, T fp1) {
  if (k == 1) return (fp1 - f0) * (T)z.w2 - (f0 - (T)0) * (T)z.ws;
  return (fp1 - f0) * (T)z.cp[k] - (f0 - fm1) * (T)z.cm[k];
}
template <typename T> __device__ __forceinline__ T zd2f(const ZMet &z, int k, T fm1, T f0, T fp1) {
  return (fp1 - f0) * (T)z.fp[k] - (f0 - fm1) * (T)z.fm[k];
}

template <typename TW, typename TC, bool ZS>
__global__ __launch_bounds__(512, sizeof(TC) == 4 ? 4 : 3) void k_explicit_march(VGrid g, VCoef c, VRK rk, ZMet zm, double betadx, int nlev, int kchunk,
                                                  const double *__restrict__ q1, const double *__restrict__ q2,
                                                  const double *__restrict__ q3, const double *__restrict__ prg1,
                                                  const double *__restrict__ prg2, const double *__restrict__ prg3,
                                                  TW *__restrict__ ru1, TW *__restrict__ ru2, TW *__restrict__ ru3,
                                                  TW *__restrict__ r1, TW *__restrict__ r2, TW *__restrict__ r3) {
  extern __shared__ unsigned char smem_raw[];
  const int NTP = blockDim.x, RJ = blockDim.y, ti = threadIdx.x, tj = threadIdx.y;
  const int NTH = NTP * RJ, tid = tj * NTP + ti;
  const int nt = g.nt, nr = g.nr, nz = g.nz;
  const long sj = g.sj, sk = g.sk;
  int bx, by, bz;
  xcd_block(bx, by, bz);  // r blocks of one z chunk (sharing halo rows) on one XCD
  const int j0 = 1 + bx * RJ;
  const int kb = 1 + by * kchunk, ke = min(kb + kchunk - 1, nz);
  const int LRW = RJ + 2, PL = LRW * NTP;
  double *pcr = reinterpret_cast<double *>(smem_raw);             // [RJ][kMaxPcrLevels][2] + [RJ][2]
  double *fin = pcr + RJ * kMaxPcrLevels * 2;
  TC *cfs = reinterpret_cast<TC *>(fin + RJ * 2);                 // [RJ][kRowCoefs]
  TC *ring = cfs + RJ * kRowCoefs;                                // [3 slots][3 arrays][LRW][NTP]
  TC *fb = ring + 9 * PL;                                         // [2][3][RJ][NTP]
  const int j = j0 + tj;
  const bool cell = (ti < nt) && (j <= nr);
  const int jc = (j <= nr) ? j : nr;
  // PCR ratios of this row (one lane per row)
  if (ti == 0) {
    double a1 = -(betadx * c.s1t11[jc]), b1 = 1.0 + 2.0 * (betadx * c.s1t11[jc]);
    double a2 = -(betadx * c.s2t11[jc]), b2 = 1.0 + 2.0 * (betadx * c.s2t11[jc]);
    for (int lvl = 0; lvl < nlev; ++lvl) {
      const double al1 = -a1 / b1, al2 = -a2 / b2;
      pcr[(tj * kMaxPcrLevels + lvl) * 2] = al1;
      pcr[(tj * kMaxPcrLevels + lvl) * 2 + 1] = al2;
      b1 = b1 + 2.0 * al1 * a1; a1 = al1 * a1;
      b2 = b2 + 2.0 * al2 * a2; a2 = al2 * a2;
    }
    const bool pow2 = nt >= 2 && (nt & (nt - 1)) == 0;
    fin[tj * 2] = 1.0 / (pow2 ? b1 + 2.0 * a1 : b1);
    fin[tj * 2 + 1] = 1.0 / (pow2 ? b2 + 2.0 * a2 : b2);
    const int jm = (jc == 1) ? 1 : jc - 1;
    TC *cf = cfs + tj * kRowCoefs;
    cf[0] = (TC)c.h1d11[jc]; cf[1] = (TC)c.h1d22d[jc]; cf[2] = (TC)c.h1d22n[jc]; cf[3] = (TC)c.h1dare[jc];
    cf[4] = (TC)c.h2d11[jc]; cf[5] = (TC)c.h2d22p[jc]; cf[6] = (TC)c.h2d22m[jc]; cf[7] = (TC)c.h2dare[jc];
    cf[8] = (TC)c.s1t11[jc]; cf[9] = (TC)c.s1t2n[jc]; cf[10] = (TC)c.s1t2s[jc]; cf[11] = (TC)c.s1t2c[jc];
    cf[12] = (TC)c.s2t11[jc]; cf[13] = (TC)c.s2t2p[jc]; cf[14] = (TC)c.s2t2c[jc]; cf[15] = (TC)c.s2t2m[jc];
    cf[16] = (TC)c.s3t2n[jc]; cf[17] = (TC)c.s3t2s[jc];
    cf[18] = (TC)(1.0 / c.rm[jc + 1]); cf[19] = (TC)(1.0 / c.rm[jc]); cf[20] = (TC)(1.0 / c.rm[jm]);
  }
  // plane fetch: element e = tid + l*NTH of the (LRW x NTP) tile, three arrays
  constexpr int NL = 3;
  long goff[NL];
  int lidx[NL];
  bool lval[NL];
#pragma unroll
  for (int l = 0; l < NL; l++) {
    const int e = tid + l * NTH;
    const int r = e / NTP, cc = e - r * NTP;
    const int jj = j0 - 1 + r;
    lidx[l] = (e < PL) ? e : -1;
    lval[l] = (e < PL) && cc < nt && jj <= nr + 2;
    goff[l] = lval[l] ? (long)cc + (long)jj * sj : 0;
  }
  TC pre[3][NL];
  // unconditional loads (masked after) so that a plane's loads issue together
  auto fetch = [&](int kk) {
#pragma unroll
    for (int l = 0; l < NL; l++) {
      const long o = goff[l] + (long)kk * sk;
      const double a1 = q1[o], a2 = q2[o], a3 = q3[o];
      pre[0][l] = lval[l] ? (TC)a1 : (TC)0;
      pre[1][l] = lval[l] ? (TC)a2 : (TC)0;
      pre[2][l] = lval[l] ? (TC)a3 : (TC)0;
    }
  };
  auto stash = [&](int slot) {
#pragma unroll
    for (int l = 0; l < NL; l++)
      if (lidx[l] >= 0) {
        ring[(slot * 3 + 0) * PL + lidx[l]] = pre[0][l];
        ring[(slot * 3 + 1) * PL + lidx[l]] = pre[1][l];
        ring[(slot * 3 + 2) * PL + lidx[l]] = pre[2][l];
      }
  };
  int sM = (kb - 1) % 3, s0 = kb % 3, sP = (kb + 1) % 3;
  fetch(kb - 1); stash(sM);
  fetch(kb); stash(s0);
  fetch(kb + 1);
  // per-row coefficients
  const TC ga = (TC)rk.g, ro = (TC)rk.r, al = (TC)rk.a, dt = (TC)g.dt, aRe = (TC)(rk.a / g.Re);
  const TC dx3 = (TC)g.dx3, dx3q = (TC)g.dx3q;
  const TC *cf = cfs + tj * kRowCoefs;  // per-row metric coefficients (LDS, broadcast reads)
#define h1d11 cf[0]
#define h1d22d cf[1]
#define h1d22n cf[2]
#define h1dare cf[3]
#define h2d11 cf[4]
#define h2d22p cf[5]
#define h2d22m cf[6]
#define h2dare cf[7]
#define s1t11 cf[8]
#define s1t2n cf[9]
#define s1t2s cf[10]
#define s1t2c cf[11]
#define s2t11 cf[12]
#define s2t2p cf[13]
#define s2t2c cf[14]
#define s2t2m cf[15]
#define s3t2n cf[16]
#define s3t2s cf[17]
#define irmn cf[18]
#define irm cf[19]
#define irmm cf[20]
  const int ci = ti, cip = (ti + 1 >= nt) ? 0 : ti + 1, cim = (ti == 0) ? nt - 1 : ti - 1;
  const int rw = tj + 1, rm1 = (jc == 1) ? rw : rw - 1;   // LDS rows of j and jm
  // ru and prg of the next plane stream in while this one is computed
  // (clamped in-bounds addresses for lanes and planes that do not use them)
  const int tic = min(ti, nt - 1);
  TC n_ru[3], n_pg[3];
  // rho = 0 (the first substep): the previous RHS is not needed
  const bool use_ru = rk.r != 0.0;
  auto load_rp = [&](int kk) {
    const int kc = min(kk, ke);
    const long oo = (long)tic + (long)jc * sj + (long)kc * sk;
    const long pp = (long)tic + (long)nt * ((long)(jc - 1) + (long)nr * (long)(kc - 1));
    if (use_ru) { n_ru[0] = ru1[oo]; n_ru[1] = ru2[oo]; n_ru[2] = ru3[oo]; }
    else { n_ru[0] = (TC)0; n_ru[1] = (TC)0; n_ru[2] = (TC)0; }
    n_pg[0] = prg1[pp]; n_pg[1] = prg2[pp]; n_pg[2] = prg3[pp];
  };
  // FP32: one plane ahead; FP64 (register budget): this plane, issued with
  // the ring fetch
  constexpr bool kAhead = sizeof(TC) == 4;
  if (kAhead) load_rp(kb);
  for (int k = kb; k <= ke; k++) {
    stash(sP);
    TC c_ru1 = n_ru[0], c_ru2 = n_ru[1], c_ru3 = n_ru[2];
    TC c_pg1 = n_pg[0], c_pg2 = n_pg[1], c_pg3 = n_pg[2];
    __syncthreads();
    fetch(k + 2);
    if (kAhead) {
      load_rp(k + 1);
    } else {
      load_rp(k);
      c_ru1 = n_ru[0]; c_ru2 = n_ru[1]; c_ru3 = n_ru[2];
      c_pg1 = n_pg[0]; c_pg2 = n_pg[1]; c_pg3 = n_pg[2];
    }
    TC v1 = (TC)0, v2 = (TC)0, v3 = (TC)0;
    const long o = (long)ti + (long)j * sj + (long)k * sk;
    constexpr bool zs = ZS;  // axial mesh: per-plane metric (a separate instantiation)
    const TC zmk = zs ? (TC)zm.m3[k] : dx3, zck = zs ? (TC)zm.c3[k] : dx3;
    if (cell) {
      const TC *A1m = ring + (sM * 3 + 0) * PL, *A1 = ring + (s0 * 3 + 0) * PL, *A1p = ring + (sP * 3 + 0) * PL;
      const TC *A2m = ring + (sM * 3 + 1) * PL, *A2 = ring + (s0 * 3 + 1) * PL, *A2p = ring + (sP * 3 + 1) * PL;
      const TC *A3m = ring + (sM * 3 + 2) * PL, *A3 = ring + (s0 * 3 + 2) * PL, *A3p = ring + (sP * 3 + 2) * PL;
#define L(A, row, col) A[(row) * NTP + (col)]
      const TC q1c = L(A1, rw, ci), q1ip = L(A1, rw, cip), q1im = L(A1, rw, cim);
      const TC q1n = L(A1, rw + 1, ci), q1jm = L(A1, rm1, ci), q1s = L(A1, rw - 1, ci), q1ips = L(A1, rw - 1, cip);
      const TC q1kp = L(A1p, rw, ci), q1km = L(A1m, rw, ci), q1ipkm = L(A1m, rw, cip);
      const TC q2c = L(A2, rw, ci), q2n = L(A2, rw + 1, ci), q2imn = L(A2, rw + 1, cim), q2im = L(A2, rw, cim);
      const TC q2ip = L(A2, rw, cip), q2s = L(A2, rw - 1, ci), q2kp = L(A2p, rw, ci), q2km = L(A2m, rw, ci);
      const TC q2nkm = L(A2m, rw + 1, ci);
      const TC q3c = L(A3, rw, ci), q3kp = L(A3p, rw, ci), q3imkp = L(A3p, rw, cim), q3im = L(A3, rw, cim);
      const TC q3ip = L(A3, rw, cip), q3s = L(A3, rw - 1, ci), q3skp = L(A3p, rw - 1, ci), q3n = L(A3, rw + 1, ci);
      const TC q3jm = L(A3, rm1, ci), q3km = L(A3m, rw, ci);
#undef L
      // ---- q1 at (theta_c, r_m, z_m) ----
      {
        TC qnn = q1ip + q1c, qss = q1im + q1c;
        TC h11 = (qnn * qnn - qss * qss) * h1d11;
        qnn = (q2n + q2imn) * (q1n + q1c);
        qss = (q1c + q1jm) * (q2c + q2im);
        TC h12d = (qnn - qss) * h1d22d;
        qnn = (q1ip + q1c) * (q2n + q2c);
        qss = (q1c + q1im) * (q2imn + q2im);
        TC h12n = (qnn + qss) * h1d22n;
        qnn = (TC)0.25 * (q3kp + q3imkp) * (q1kp + q1c);
        qss = (k == 1) ? (TC)0 : (TC)0.25 * (q3c + q3im) * (q1c + q1km);
        TC h13 = (qnn - qss) * zmk;
        qnn = q2n + q2c;
        qss = q2imn + q2im;
        TC d1q2 = (qnn - qss) * h1dare;
        TC H1 = d1q2 - (h11 + (h12d + h12n) + h13);
        TC d11 = (q1ip - (TC)2 * q1c + q1im) * s1t11;
        TC d33 = zs ? zd2c<TC>(zm, k, q1km, q1c, q1kp)
                    : (k == 1) ? (TC)(4.0 / 3.0) * (q1kp - (TC)3 * q1c) * dx3q : (q1kp - (TC)2 * q1c + q1km) * dx3q;
        TC qn2 = (q1n * irmn - q1c * irm) * s1t2n;
        TC qs2 = (q1c * irm - q1jm * irmm) * s1t2s;
        TC d22 = (qn2 - qs2) * s1t2c;
        v1 = (ga * H1 + ro * c_ru1 - al * c_pg1 + aRe * (d11 + d33 + d22)) * dt;
        ru1[o] = H1;
      }
      // ---- q2 at (theta_m, r_c, z_m), j >= 2 ----
      if (j >= 2) {
        TC qnn = (q1ip + q1ips) * (q2ip + q2c);
        TC qss = (q1c + q1s) * (q2c + q2im);
        TC h21 = (qnn - qss) * h2d11;
        qnn = q2n + q2c;
        qss = q2c + q2s;
        TC h22 = qnn * qnn * h2d22p - qss * qss * h2d22m;
        qnn = (TC)0.5 * (q1c + q1ip);
        qss = (TC)0.5 * (q1s + q1ips);
        TC h11n = (TC)0.25 * (qnn + qss) * (qnn + qss);
        qnn = (TC)0.25 * (q3kp + q3skp) * (q2kp + q2c);
        qss = (k == 1) ? (TC)0 : (TC)0.25 * (q3c + q3s) * (q2c + q2km);
        TC h23 = (qnn - qss) * zmk;
        qnn = q1ip + q1ips;
        qss = q1c + q1s;
        TC d1q1 = -((qnn - qss) * h2dare);
        TC H2 = d1q1 - (h21 + h22 + h23 - h11n);
        TC d11 = (q2ip - (TC)2 * q2c + q2im) * s2t11;
        TC d33 = zs ? zd2c<TC>(zm, k, q2km, q2c, q2kp)
                    : (k == 1) ? (TC)(4.0 / 3.0) * (q2kp - (TC)3 * q2c) * dx3q : (q2kp - (TC)2 * q2c + q2km) * dx3q;
        TC d22 = q2n * s2t2p - q2c * s2t2c + q2s * s2t2m;
        v2 = (ga * H2 + ro * c_ru2 - al * c_pg2 + aRe * (d11 + d33 + d22)) * dt;
        ru2[o] = H2;
      }
      // ---- q3 at (theta_m, r_m, z_c), k >= 2 ----
      if (k >= 2) {
        TC qnn = (q3n + q3c) * (q2n + q2nkm);
        TC qss = (q3jm + q3c) * (q2c + q2km);
        TC h32 = (qnn - qss) * h1d22d;
        qnn = (q3c + q3ip) * (q1ip + q1ipkm);
        qss = (q3im + q3c) * (q1c + q1km);
        TC h31 = (qnn - qss) * h1d11;
        qnn = (TC)0.25 * (q3c + q3kp) * (q3c + q3kp);
        qss = (TC)0.25 * (q3c + q3km) * (q3c + q3km);
        TC h33 = (qnn - qss) * zck;
        TC H3 = -(h31 + h32 + h33);
        TC d11 = (q3ip - (TC)2 * q3c + q3im) * s1t11;
        TC d33 = zs ? zd2f<TC>(zm, k, q3km, q3c, q3kp) : (q3kp - (TC)2 * q3c + q3km) * dx3q;
        TC qn2 = (q3n - q3c) * s3t2n;
        TC qs2 = (q3c - q3jm) * s3t2s;
        TC d22 = qn2 - qs2;
        v3 = (ga * H3 + ro * c_ru3 - al * c_pg3 + aRe * (d11 + d33 + d22)) * dt;
        ru3[o] = H3;
      }
    }
    // ---- periodic theta solves by PCR (ping-pong buffers: one barrier per level) ----
    const long S3 = 3L * RJ * NTP;
    TC *f1 = fb + tj * NTP, *f2 = f1 + RJ * NTP, *f3 = f2 + RJ * NTP;
    f1[ti] = v1; f2[ti] = v2; f3[ti] = v3;
    __syncthreads();
    int sm = 1 % nt;
    for (int lvl = 0; lvl < nlev; ++lvl) {
      const TC al1 = (TC)pcr[(tj * kMaxPcrLevels + lvl) * 2], al2 = (TC)pcr[(tj * kMaxPcrLevels + lvl) * 2 + 1];
      TC n1 = v1, n2 = v2, n3 = v3;
      if (ti < nt) {
        int imv = ti - sm; if (imv < 0) imv += nt;
        int ipv = ti + sm; if (ipv >= nt) ipv -= nt;
        n1 = v1 + al1 * (f1[imv] + f1[ipv]);
        n2 = v2 + al2 * (f2[imv] + f2[ipv]);
        n3 = v3 + al1 * (f3[imv] + f3[ipv]);
      }
      sm = (2 * sm) % nt;
      const long flip = (lvl & 1) ? -S3 : S3;
      f1 += flip; f2 += flip; f3 += flip;
      f1[ti] = n1; f2[ti] = n2; f3[ti] = n3;
      v1 = n1; v2 = n2; v3 = n3;
      __syncthreads();
    }
    if (cell) {
      const TC d1 = (TC)fin[tj * 2], d2 = (TC)fin[tj * 2 + 1];
      r1[o] = v1 * d1;
      if (j >= 2) r2[o] = v2 * d2;
      if (k >= 2) r3[o] = v3 * d1;
    }
    sM = s0; s0 = sP; sP = (sP == 2) ? 0 : sP + 1;
    __syncthreads();
  }
#undef h1d11
#undef h1d22d
#undef h1d22n
#undef h1dare
#undef h2d11
#undef h2d22p
#undef h2d22m
#undef h2dare
#undef s1t11
#undef s1t2n
#undef s1t2s
#undef s1t2c
#undef s2t11
#undef s2t2p
#undef s2t2c
#undef s2t2m
#undef s3t2n
#undef s3t2s
#undef irmn
#undef irm
#undef irmm
}

// Variant with one z plane per block (no z marching): every stencil operand
// is read through L1/L2.  Fewer registers and LDS than the marching form;
// the autotuner picks whichever is faster for the precision and grid.
// The explicit kernel with one wavefront per r row (nθ <= 64 CPL): each
// lane computes the cells θ = lane + 64 c, and the periodic θ solves run by
// PCR inside the wavefront (LDS exchange ordered by wave-local fences, no
// block barrier per level), which leaves two block barriers per z plane for
// the plane ring.  Same arithmetic as k_explicit_march.
template <typename TW, typename TC, int CPL, int RJT, int WPE, bool ZS>
__global__ __launch_bounds__(64 * RJT, WPE) void k_explicit_wave(VGrid g, VCoef c, VRK rk, ZMet zm, double betadx, int nlev, int kchunk,
                                                  const double *__restrict__ q1, const double *__restrict__ q2,
                                                  const double *__restrict__ q3, const double *__restrict__ prg1,
                                                  const double *__restrict__ prg2, const double *__restrict__ prg3,
                                                  TW *__restrict__ ru1, TW *__restrict__ ru2, TW *__restrict__ ru3,
                                                  TW *__restrict__ r1, TW *__restrict__ r2, TW *__restrict__ r3) {
  extern __shared__ unsigned char smem_raw[];
  // one wave per r row; lane holds theta cells lane + 64 c, c < CPL
  constexpr int NTP = 64 * CPL, RJ = RJT;
  const int lane = threadIdx.x, tj = threadIdx.y;
  const int NTH = 64 * RJ, tid = tj * 64 + lane;
  const int nt = g.nt, nr = g.nr, nz = g.nz;
  const long sj = g.sj, sk = g.sk;
  int bx, by, bz;
  xcd_block(bx, by, bz);  // r blocks of one z chunk (sharing halo rows) on one XCD
  const int j0 = 1 + bx * RJ;
  const int kb = 1 + by * kchunk, ke = min(kb + kchunk - 1, nz);
  const int LRW = RJ + 2, PL = LRW * NTP;
  double *pcr = reinterpret_cast<double *>(smem_raw);             // [RJ][nlev][2] + [RJ][2]
  double *fin = pcr + RJ * nlev * 2;
  TC *cfs = reinterpret_cast<TC *>(fin + RJ * 2);                 // [RJ][kRowCoefs]
  TC *ring = cfs + RJ * kRowCoefs;                                // [3 slots][3 arrays][LRW][NTP]
  TC *fb = ring + 9 * PL;                                         // [2][3][RJ][NTP]
  const int j = j0 + tj;
  const int jc = (j <= nr) ? j : nr;
  // PCR ratios of this row (one lane per row)
  if (lane == 0) {
    double a1 = -(betadx * c.s1t11[jc]), b1 = 1.0 + 2.0 * (betadx * c.s1t11[jc]);
    double a2 = -(betadx * c.s2t11[jc]), b2 = 1.0 + 2.0 * (betadx * c.s2t11[jc]);
    for (int lvl = 0; lvl < nlev; ++lvl) {
      const double al1 = -a1 / b1, al2 = -a2 / b2;
      pcr[(tj * nlev + lvl) * 2] = al1;
      pcr[(tj * nlev + lvl) * 2 + 1] = al2;
      b1 = b1 + 2.0 * al1 * a1; a1 = al1 * a1;
      b2 = b2 + 2.0 * al2 * a2; a2 = al2 * a2;
    }
    const bool pow2 = nt >= 2 && (nt & (nt - 1)) == 0;
    fin[tj * 2] = 1.0 / (pow2 ? b1 + 2.0 * a1 : b1);
    fin[tj * 2 + 1] = 1.0 / (pow2 ? b2 + 2.0 * a2 : b2);
    const int jm = (jc == 1) ? 1 : jc - 1;
    TC *cf = cfs + tj * kRowCoefs;
    cf[0] = (TC)c.h1d11[jc]; cf[1] = (TC)c.h1d22d[jc]; cf[2] = (TC)c.h1d22n[jc]; cf[3] = (TC)c.h1dare[jc];
    cf[4] = (TC)c.h2d11[jc]; cf[5] = (TC)c.h2d22p[jc]; cf[6] = (TC)c.h2d22m[jc]; cf[7] = (TC)c.h2dare[jc];
    cf[8] = (TC)c.s1t11[jc]; cf[9] = (TC)c.s1t2n[jc]; cf[10] = (TC)c.s1t2s[jc]; cf[11] = (TC)c.s1t2c[jc];
    cf[12] = (TC)c.s2t11[jc]; cf[13] = (TC)c.s2t2p[jc]; cf[14] = (TC)c.s2t2c[jc]; cf[15] = (TC)c.s2t2m[jc];
    cf[16] = (TC)c.s3t2n[jc]; cf[17] = (TC)c.s3t2s[jc];
    cf[18] = (TC)(1.0 / c.rm[jc + 1]); cf[19] = (TC)(1.0 / c.rm[jc]); cf[20] = (TC)(1.0 / c.rm[jm]);
  }
  // plane fetch: element e = tid + l*NTH of the (LRW x NTP) tile, three arrays
  constexpr int NL = (CPL * (RJT + 2) + RJT - 1) / RJT;  // ceil((RJ + 2) NTP / (64 RJ))
  long goff[NL];
  int lidx[NL];
  bool lval[NL];
#pragma unroll
  for (int l = 0; l < NL; l++) {
    const int e = tid + l * NTH;
    const int r = e / NTP, cc = e - r * NTP;
    const int jj = j0 - 1 + r;
    lidx[l] = (e < PL) ? e : -1;
    lval[l] = (e < PL) && cc < nt && jj <= nr + 2;
    goff[l] = lval[l] ? (long)cc + (long)jj * sj : 0;
  }
  TC pre[3][NL];
  // unconditional loads (masked after) so that a plane's loads issue together
  auto fetch = [&](int kk) {
#pragma unroll
    for (int l = 0; l < NL; l++) {
      const long o = goff[l] + (long)kk * sk;
      const double a1 = q1[o], a2 = q2[o], a3 = q3[o];
      pre[0][l] = lval[l] ? (TC)a1 : (TC)0;
      pre[1][l] = lval[l] ? (TC)a2 : (TC)0;
      pre[2][l] = lval[l] ? (TC)a3 : (TC)0;
    }
  };
  auto stash = [&](int slot) {
#pragma unroll
    for (int l = 0; l < NL; l++)
      if (lidx[l] >= 0) {
        ring[(slot * 3 + 0) * PL + lidx[l]] = pre[0][l];
        ring[(slot * 3 + 1) * PL + lidx[l]] = pre[1][l];
        ring[(slot * 3 + 2) * PL + lidx[l]] = pre[2][l];
      }
  };
  int sM = (kb - 1) % 3, s0 = kb % 3, sP = (kb + 1) % 3;
  fetch(kb - 1); stash(sM);
  fetch(kb); stash(s0);
  fetch(kb + 1);
  // per-row coefficients
  const TC ga = (TC)rk.g, ro = (TC)rk.r, al = (TC)rk.a, dt = (TC)g.dt, aRe = (TC)(rk.a / g.Re);
  const TC dx3 = (TC)g.dx3, dx3q = (TC)g.dx3q;
  const TC *cf = cfs + tj * kRowCoefs;  // per-row metric coefficients (LDS, broadcast reads)
#define h1d11 cf[0]
#define h1d22d cf[1]
#define h1d22n cf[2]
#define h1dare cf[3]
#define h2d11 cf[4]
#define h2d22p cf[5]
#define h2d22m cf[6]
#define h2dare cf[7]
#define s1t11 cf[8]
#define s1t2n cf[9]
#define s1t2s cf[10]
#define s1t2c cf[11]
#define s2t11 cf[12]
#define s2t2p cf[13]
#define s2t2c cf[14]
#define s2t2m cf[15]
#define s3t2n cf[16]
#define s3t2s cf[17]
#define irmn cf[18]
#define irm cf[19]
#define irmm cf[20]
  const int rw = tj + 1, rm1 = (jc == 1) ? rw : rw - 1;   // LDS rows of j and jm
  // ru and prg of the next plane stream in while this one is computed
  // (clamped in-bounds addresses for lanes and planes that do not use them)
  TC n_ru[CPL][3], n_pg[CPL][3];
  // rho = 0 (the first substep): the previous RHS is not needed
  const bool use_ru = rk.r != 0.0;
  auto load_rp = [&](int kk) {
    const int kc = min(kk, ke);
#pragma unroll
    for (int cc = 0; cc < CPL; cc++) {
      const int tic = min(lane + 64 * cc, nt - 1);
      const long oo = (long)tic + (long)jc * sj + (long)kc * sk;
      const long pp = (long)tic + (long)nt * ((long)(jc - 1) + (long)nr * (long)(kc - 1));
      if (use_ru) { n_ru[cc][0] = ru1[oo]; n_ru[cc][1] = ru2[oo]; n_ru[cc][2] = ru3[oo]; }
      else { n_ru[cc][0] = (TC)0; n_ru[cc][1] = (TC)0; n_ru[cc][2] = (TC)0; }
      n_pg[cc][0] = prg1[pp]; n_pg[cc][1] = prg2[pp]; n_pg[cc][2] = prg3[pp];
    }
  };
  auto wave_sync = [&]() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  };
  load_rp(kb);
  for (int k = kb; k <= ke; k++) {
    stash(sP);
    TC cru[CPL][3], cpg[CPL][3];
#pragma unroll
    for (int cc = 0; cc < CPL; cc++)
#pragma unroll
      for (int q = 0; q < 3; q++) { cru[cc][q] = n_ru[cc][q]; cpg[cc][q] = n_pg[cc][q]; }
    __syncthreads();
    fetch(k + 2);
    load_rp(k + 1);
    TC w1[CPL], w2[CPL], w3[CPL];
#pragma unroll
    for (int cc = 0; cc < CPL; cc++) {
      const int ti = lane + 64 * cc;
      const int ci = ti, cip = (ti + 1 >= nt) ? 0 : ti + 1, cim = (ti == 0) ? nt - 1 : ti - 1;
      const bool cell = (ti < nt) && (j <= nr);
      const TC c_ru1 = cru[cc][0], c_ru2 = cru[cc][1], c_ru3 = cru[cc][2];
      const TC c_pg1 = cpg[cc][0], c_pg2 = cpg[cc][1], c_pg3 = cpg[cc][2];
      TC v1 = (TC)0, v2 = (TC)0, v3 = (TC)0;
      constexpr bool zs = ZS;  // axial mesh: per-plane metric (a separate instantiation)
      const TC zmk = zs ? (TC)zm.m3[k] : dx3, zck = zs ? (TC)zm.c3[k] : dx3;
      const long o = (long)ti + (long)j * sj + (long)k * sk;
      if (cell) {
      const TC *A1m = ring + (sM * 3 + 0) * PL, *A1 = ring + (s0 * 3 + 0) * PL, *A1p = ring + (sP * 3 + 0) * PL;
      const TC *A2m = ring + (sM * 3 + 1) * PL, *A2 = ring + (s0 * 3 + 1) * PL, *A2p = ring + (sP * 3 + 1) * PL;
      const TC *A3m = ring + (sM * 3 + 2) * PL, *A3 = ring + (s0 * 3 + 2) * PL, *A3p = ring + (sP * 3 + 2) * PL;
#define L(A, row, col) A[(row) * NTP + (col)]
      const TC q1c = L(A1, rw, ci), q1ip = L(A1, rw, cip), q1im = L(A1, rw, cim);
      const TC q1n = L(A1, rw + 1, ci), q1jm = L(A1, rm1, ci), q1s = L(A1, rw - 1, ci), q1ips = L(A1, rw - 1, cip);
      const TC q1kp = L(A1p, rw, ci), q1km = L(A1m, rw, ci), q1ipkm = L(A1m, rw, cip);
      const TC q2c = L(A2, rw, ci), q2n = L(A2, rw + 1, ci), q2imn = L(A2, rw + 1, cim), q2im = L(A2, rw, cim);
      const TC q2ip = L(A2, rw, cip), q2s = L(A2, rw - 1, ci), q2kp = L(A2p, rw, ci), q2km = L(A2m, rw, ci);
      const TC q2nkm = L(A2m, rw + 1, ci);
      const TC q3c = L(A3, rw, ci), q3kp = L(A3p, rw, ci), q3imkp = L(A3p, rw, cim), q3im = L(A3, rw, cim);
      const TC q3ip = L(A3, rw, cip), q3s = L(A3, rw - 1, ci), q3skp = L(A3p, rw - 1, ci), q3n = L(A3, rw + 1, ci);
      const TC q3jm = L(A3, rm1, ci), q3km = L(A3m, rw, ci);
#undef L
      // ---- q1 at (theta_c, r_m, z_m) ----
      {
        TC qnn = q1ip + q1c, qss = q1im + q1c;
        TC h11 = (qnn * qnn - qss * qss) * h1d11;
        qnn = (q2n + q2imn) * (q1n + q1c);
        qss = (q1c + q1jm) * (q2c + q2im);
        TC h12d = (qnn - qss) * h1d22d;
        qnn = (q1ip + q1c) * (q2n + q2c);
        qss = (q1c + q1im) * (q2imn + q2im);
        TC h12n = (qnn + qss) * h1d22n;
        qnn = (TC)0.25 * (q3kp + q3imkp) * (q1kp + q1c);
        qss = (k == 1) ? (TC)0 : (TC)0.25 * (q3c + q3im) * (q1c + q1km);
        TC h13 = (qnn - qss) * zmk;
        qnn = q2n + q2c;
        qss = q2imn + q2im;
        TC d1q2 = (qnn - qss) * h1dare;
        TC H1 = d1q2 - (h11 + (h12d + h12n) + h13);
        TC d11 = (q1ip - (TC)2 * q1c + q1im) * s1t11;
        TC d33 = zs ? zd2c<TC>(zm, k, q1km, q1c, q1kp)
                    : (k == 1) ? (TC)(4.0 / 3.0) * (q1kp - (TC)3 * q1c) * dx3q : (q1kp - (TC)2 * q1c + q1km) * dx3q;
        TC qn2 = (q1n * irmn - q1c * irm) * s1t2n;
        TC qs2 = (q1c * irm - q1jm * irmm) * s1t2s;
        TC d22 = (qn2 - qs2) * s1t2c;
        v1 = (ga * H1 + ro * c_ru1 - al * c_pg1 + aRe * (d11 + d33 + d22)) * dt;
        ru1[o] = H1;
      }
      // ---- q2 at (theta_m, r_c, z_m), j >= 2 ----
      if (j >= 2) {
        TC qnn = (q1ip + q1ips) * (q2ip + q2c);
        TC qss = (q1c + q1s) * (q2c + q2im);
        TC h21 = (qnn - qss) * h2d11;
        qnn = q2n + q2c;
        qss = q2c + q2s;
        TC h22 = qnn * qnn * h2d22p - qss * qss * h2d22m;
        qnn = (TC)0.5 * (q1c + q1ip);
        qss = (TC)0.5 * (q1s + q1ips);
        TC h11n = (TC)0.25 * (qnn + qss) * (qnn + qss);
        qnn = (TC)0.25 * (q3kp + q3skp) * (q2kp + q2c);
        qss = (k == 1) ? (TC)0 : (TC)0.25 * (q3c + q3s) * (q2c + q2km);
        TC h23 = (qnn - qss) * zmk;
        qnn = q1ip + q1ips;
        qss = q1c + q1s;
        TC d1q1 = -((qnn - qss) * h2dare);
        TC H2 = d1q1 - (h21 + h22 + h23 - h11n);
        TC d11 = (q2ip - (TC)2 * q2c + q2im) * s2t11;
        TC d33 = zs ? zd2c<TC>(zm, k, q2km, q2c, q2kp)
                    : (k == 1) ? (TC)(4.0 / 3.0) * (q2kp - (TC)3 * q2c) * dx3q : (q2kp - (TC)2 * q2c + q2km) * dx3q;
        TC d22 = q2n * s2t2p - q2c * s2t2c + q2s * s2t2m;
        v2 = (ga * H2 + ro * c_ru2 - al * c_pg2 + aRe * (d11 + d33 + d22)) * dt;
        ru2[o] = H2;
      }
      // ---- q3 at (theta_m, r_m, z_c), k >= 2 ----
      if (k >= 2) {
        TC qnn = (q3n + q3c) * (q2n + q2nkm);
        TC qss = (q3jm + q3c) * (q2c + q2km);
        TC h32 = (qnn - qss) * h1d22d;
        qnn = (q3c + q3ip) * (q1ip + q1ipkm);
        qss = (q3im + q3c) * (q1c + q1km);
        TC h31 = (qnn - qss) * h1d11;
        qnn = (TC)0.25 * (q3c + q3kp) * (q3c + q3kp);
        qss = (TC)0.25 * (q3c + q3km) * (q3c + q3km);
        TC h33 = (qnn - qss) * zck;
        TC H3 = -(h31 + h32 + h33);
        TC d11 = (q3ip - (TC)2 * q3c + q3im) * s1t11;
        TC d33 = zs ? zd2f<TC>(zm, k, q3km, q3c, q3kp) : (q3kp - (TC)2 * q3c + q3km) * dx3q;
        TC qn2 = (q3n - q3c) * s3t2n;
        TC qs2 = (q3c - q3jm) * s3t2s;
        TC d22 = qn2 - qs2;
        v3 = (ga * H3 + ro * c_ru3 - al * c_pg3 + aRe * (d11 + d33 + d22)) * dt;
        ru3[o] = H3;
      }
          }
      w1[cc] = v1; w2[cc] = v2; w3[cc] = v3;
    }
    // ---- periodic theta solves by PCR within the wavefront ----
    const long S3 = 3L * RJ * NTP;
    TC *f1 = fb + tj * NTP, *f2 = f1 + RJ * NTP, *f3 = f2 + RJ * NTP;
#pragma unroll
    for (int cc = 0; cc < CPL; cc++) {
      const int ti = lane + 64 * cc;
      f1[ti] = w1[cc]; f2[ti] = w2[cc]; f3[ti] = w3[cc];
    }
    wave_sync();
    int sm = 1 % nt;
    for (int lvl = 0; lvl < nlev; ++lvl) {
      const TC al1 = (TC)pcr[(tj * nlev + lvl) * 2], al2 = (TC)pcr[(tj * nlev + lvl) * 2 + 1];
      TC n1[CPL], n2[CPL], n3[CPL];
#pragma unroll
      for (int cc = 0; cc < CPL; cc++) {
        const int ti = lane + 64 * cc;
        n1[cc] = w1[cc]; n2[cc] = w2[cc]; n3[cc] = w3[cc];
        if (ti < nt) {
          int imv = ti - sm; if (imv < 0) imv += nt;
          int ipv = ti + sm; if (ipv >= nt) ipv -= nt;
          n1[cc] = w1[cc] + al1 * (f1[imv] + f1[ipv]);
          n2[cc] = w2[cc] + al2 * (f2[imv] + f2[ipv]);
          n3[cc] = w3[cc] + al1 * (f3[imv] + f3[ipv]);
        }
      }
      sm = (2 * sm) % nt;
      const long flip = (lvl & 1) ? -S3 : S3;
      f1 += flip; f2 += flip; f3 += flip;
#pragma unroll
      for (int cc = 0; cc < CPL; cc++) {
        const int ti = lane + 64 * cc;
        f1[ti] = n1[cc]; f2[ti] = n2[cc]; f3[ti] = n3[cc];
        w1[cc] = n1[cc]; w2[cc] = n2[cc]; w3[cc] = n3[cc];
      }
      wave_sync();
    }
    const TC d1 = (TC)fin[tj * 2], d2 = (TC)fin[tj * 2 + 1];
#pragma unroll
    for (int cc = 0; cc < CPL; cc++) {
      const int ti = lane + 64 * cc;
      if (ti < nt && j <= nr) {
        const long o = (long)ti + (long)j * sj + (long)k * sk;
        r1[o] = w1[cc] * d1;
        if (j >= 2) r2[o] = w2[cc] * d2;
        if (k >= 2) r3[o] = w3[cc] * d1;
      }
    }
    sM = s0; s0 = sP; sP = (sP == 2) ? 0 : sP + 1;
    __syncthreads();
  }
#undef h1d11
#undef h1d22d
#undef h1d22n
#undef h1dare
#undef h2d11
#undef h2d22p
#undef h2d22m
#undef h2dare
#undef s1t11
#undef s1t2n
#undef s1t2s
#undef s1t2c
#undef s2t11
#undef s2t2p
#undef s2t2c
#undef s2t2m
#undef s3t2n
#undef s3t2s
#undef irmn
#undef irm
#undef irmm
}

template <typename TW, typename TC, bool ZS>
__global__ __launch_bounds__(1024) void k_explicit_cell(VGrid g, VCoef c, VRK rk, ZMet zm, double betadx, int nlev,
                                                   const double *__restrict__ q1, const double *__restrict__ q2,
                                                   const double *__restrict__ q3, const double *__restrict__ prg1,
                                                   const double *__restrict__ prg2, const double *__restrict__ prg3,
                                                   TW *__restrict__ ru1, TW *__restrict__ ru2, TW *__restrict__ ru3,
                                                   TW *__restrict__ r1, TW *__restrict__ r2, TW *__restrict__ r3) {
  extern __shared__ unsigned char smem_raw[];
  const int NTP = blockDim.x, RJ = blockDim.y, ti = threadIdx.x, tj = threadIdx.y;
  const int i = ti + 1;
  int bx, by, bz;
  xcd_block(bx, by, bz);  // r blocks of one plane (sharing halo rows) on one XCD
  const int j = bx * RJ + tj + 1;
  const int k = by + 1;
  const bool cell = (i <= g.nt) && (j <= g.nr);
  const int jc = cell ? j : 1;  // coefficient row for idle threads
  TC v1 = (TC)0, v2 = (TC)0, v3 = (TC)0;
  if (cell) {
  const int ip = (i == g.nt) ? 1 : i + 1, im = (i == 1) ? g.nt : i - 1;
  const int jm = (j == 1) ? 1 : j - 1;
  const long o = hidx(g, i, j, k);
  const long dip = ip - i, dim = im - i, sj = g.sj, sk = g.sk, djm = (long)(jm - j) * sj;
  const TC ga = (TC)rk.g, ro = (TC)rk.r, al = (TC)rk.a, dt = (TC)g.dt, aRe = (TC)(rk.a / g.Re);
  const bool use_ru = rk.r != 0.0;  // rho = 0 (the first substep): the previous RHS is not needed
  const TC dx3 = (TC)g.dx3, dx3q = (TC)g.dx3q;
  constexpr bool zs = ZS;  // axial mesh: per-plane metric (a separate instantiation)
  const TC zmk = zs ? (TC)zm.m3[k] : dx3, zck = zs ? (TC)zm.c3[k] : dx3;
  const long pc = cidx(g, i, j, k);
#define Q1(d) ((TC)q1[o + (d)])
#define Q2(d) ((TC)q2[o + (d)])
#define Q3(d) ((TC)q3[o + (d)])
  // ---- q1 at (theta_c, r_m, z_m) ----
  {
    TC qnn = Q1(dip) + Q1(0), qss = Q1(dim) + Q1(0);
    TC h11 = (qnn * qnn - qss * qss) * (TC)c.h1d11[j];
    qnn = (Q2(sj) + Q2(dim + sj)) * (Q1(sj) + Q1(0));
    qss = (Q1(0) + Q1(djm)) * (Q2(0) + Q2(dim));
    TC h12d = (qnn - qss) * (TC)c.h1d22d[j];
    qnn = (Q1(dip) + Q1(0)) * (Q2(sj) + Q2(0));
    qss = (Q1(0) + Q1(dim)) * (Q2(dim + sj) + Q2(dim));
    TC h12n = (qnn + qss) * (TC)c.h1d22n[j];
    qnn = (TC)0.25 * (Q3(sk) + Q3(dim + sk)) * (Q1(sk) + Q1(0));
    qss = (k == 1) ? (TC)0 : (TC)0.25 * (Q3(0) + Q3(dim)) * (Q1(0) + Q1(-sk));
    TC h13 = (qnn - qss) * zmk;
    qnn = Q2(sj) + Q2(0);
    qss = Q2(dim + sj) + Q2(dim);
    TC d1q2 = (qnn - qss) * (TC)c.h1dare[j];
    TC H1 = d1q2 - (h11 + (h12d + h12n) + h13);
    TC d11 = (Q1(dip) - (TC)2 * Q1(0) + Q1(dim)) * (TC)c.s1t11[j];
    TC d33 = zs ? zd2c<TC>(zm, k, Q1(-sk), Q1(0), Q1(sk))
                : (k == 1) ? (TC)(4.0 / 3.0) * (Q1(sk) - (TC)3 * Q1(0)) * dx3q
                           : (Q1(sk) - (TC)2 * Q1(0) + Q1(-sk)) * dx3q;
    TC qn2 = (Q1(sj) / (TC)c.rm[j + 1] - Q1(0) / (TC)c.rm[j]) * (TC)c.s1t2n[j];
    TC qs2 = (Q1(0) / (TC)c.rm[j] - Q1(djm) / (TC)c.rm[jm]) * (TC)c.s1t2s[j];
    TC d22 = (qn2 - qs2) * (TC)c.s1t2c[j];
    v1 = (ga * H1 + ro * (use_ru ? ru1[o] : (TC)0) - al * prg1[pc] + aRe * (d11 + d33 + d22)) * dt;
    ru1[o] = H1;
  }
  // ---- q2 at (theta_m, r_c, z_m), j >= 2 ----
  if (j >= 2) {
    TC qnn = (Q1(dip) + Q1(dip - sj)) * (Q2(dip) + Q2(0));
    TC qss = (Q1(0) + Q1(-sj)) * (Q2(0) + Q2(dim));
    TC h21 = (qnn - qss) * (TC)c.h2d11[j];
    qnn = Q2(sj) + Q2(0);
    qss = Q2(0) + Q2(-sj);
    TC h22 = qnn * qnn * (TC)c.h2d22p[j] - qss * qss * (TC)c.h2d22m[j];
    qnn = (TC)0.5 * (Q1(0) + Q1(dip));
    qss = (TC)0.5 * (Q1(-sj) + Q1(dip - sj));
    TC h11n = (TC)0.25 * (qnn + qss) * (qnn + qss);
    qnn = (TC)0.25 * (Q3(sk) + Q3(sk - sj)) * (Q2(sk) + Q2(0));
    qss = (k == 1) ? (TC)0 : (TC)0.25 * (Q3(0) + Q3(-sj)) * (Q2(0) + Q2(-sk));
    TC h23 = (qnn - qss) * zmk;
    qnn = Q1(dip) + Q1(dip - sj);
    qss = Q1(0) + Q1(-sj);
    TC d1q1 = -((qnn - qss) * (TC)c.h2dare[j]);
    TC H2 = d1q1 - (h21 + h22 + h23 - h11n);
    TC d11 = (Q2(dip) - (TC)2 * Q2(0) + Q2(dim)) * (TC)c.s2t11[j];
    TC d33 = zs ? zd2c<TC>(zm, k, Q2(-sk), Q2(0), Q2(sk))
                : (k == 1) ? (TC)(4.0 / 3.0) * (Q2(sk) - (TC)3 * Q2(0)) * dx3q
                           : (Q2(sk) - (TC)2 * Q2(0) + Q2(-sk)) * dx3q;
    TC d22 = Q2(sj) * (TC)c.s2t2p[j] - Q2(0) * (TC)c.s2t2c[j] + Q2(-sj) * (TC)c.s2t2m[j];
    v2 = (ga * H2 + ro * (use_ru ? ru2[o] : (TC)0) - al * prg2[pc] + aRe * (d11 + d33 + d22)) * dt;
    ru2[o] = H2;
  }
  // ---- q3 at (theta_m, r_m, z_c), k >= 2 ----
  if (k >= 2) {
    TC qnn = (Q3(sj) + Q3(0)) * (Q2(sj) + Q2(sj - sk));
    TC qss = (Q3(djm) + Q3(0)) * (Q2(0) + Q2(-sk));
    TC h32 = (qnn - qss) * (TC)c.h1d22d[j];
    qnn = (Q3(0) + Q3(dip)) * (Q1(dip) + Q1(dip - sk));
    qss = (Q3(dim) + Q3(0)) * (Q1(0) + Q1(-sk));
    TC h31 = (qnn - qss) * (TC)c.h1d11[j];
    qnn = (TC)0.25 * (Q3(0) + Q3(sk)) * (Q3(0) + Q3(sk));
    qss = (TC)0.25 * (Q3(0) + Q3(-sk)) * (Q3(0) + Q3(-sk));
    TC h33 = (qnn - qss) * zck;
    TC H3 = -(h31 + h32 + h33);
    TC d11 = (Q3(dip) - (TC)2 * Q3(0) + Q3(dim)) * (TC)c.s1t11[j];
    TC d33 = zs ? zd2f<TC>(zm, k, Q3(-sk), Q3(0), Q3(sk)) : (Q3(sk) - (TC)2 * Q3(0) + Q3(-sk)) * dx3q;
    TC qn2 = (Q3(sj) - Q3(0)) * (TC)c.s3t2n[j];
    TC qs2 = (Q3(0) - Q3(djm)) * (TC)c.s3t2s[j];
    TC d22 = qn2 - qs2;
    v3 = (ga * H3 + ro * (use_ru ? ru3[o] : (TC)0) - al * prg3[pc] + aRe * (d11 + d33 + d22)) * dt;
    ru3[o] = H3;
  }
#undef Q1
#undef Q2
#undef Q3
  }
  // ---- periodic theta solves of the three components (solqi/tripij,
  // step.F90:212-263, 487-534): constant coefficients (-acc, 1+2acc, -acc)
  // per line, parallel cyclic reduction on the LDS-resident lines ----
  TC *f1 = reinterpret_cast<TC *>(smem_raw) + (size_t)tj * NTP;
  TC *f2 = f1 + (size_t)RJ * NTP;
  TC *f3 = f2 + (size_t)RJ * NTP;
  f1[ti] = v1; f2[ti] = v2; f3[ti] = v3;
  double a1 = -(betadx * c.s1t11[jc]), b1 = 1.0 + 2.0 * (betadx * c.s1t11[jc]);
  double a2 = -(betadx * c.s2t11[jc]), b2 = 1.0 + 2.0 * (betadx * c.s2t11[jc]);
  const int N = g.nt;
  const bool pow2 = N >= 2 && (N & (N - 1)) == 0;
  __syncthreads();
  // ping-pong buffers: one barrier per reduction level
  const long S3 = 3L * RJ * NTP;
  int sm = 1 % N;
  for (int lvl = 0; lvl < nlev; ++lvl) {
    const double al1 = -a1 / b1, al2 = -a2 / b2;
    TC n1 = v1, n2 = v2, n3 = v3;
    if (ti < N) {
      int imv = ti - sm; if (imv < 0) imv += N;
      int ipv = ti + sm; if (ipv >= N) ipv -= N;
      n1 = v1 + (TC)al1 * (f1[imv] + f1[ipv]);
      n2 = v2 + (TC)al2 * (f2[imv] + f2[ipv]);
      n3 = v3 + (TC)al1 * (f3[imv] + f3[ipv]);
    }
    b1 = b1 + 2.0 * al1 * a1; a1 = al1 * a1;
    b2 = b2 + 2.0 * al2 * a2; a2 = al2 * a2;
    sm = (2 * sm) % N;
    const long flip = (lvl & 1) ? -S3 : S3;
    f1 += flip; f2 += flip; f3 += flip;
    f1[ti] = n1; f2[ti] = n2; f3[ti] = n3;
    v1 = n1; v2 = n2; v3 = n3;
    __syncthreads();
  }
  if (!cell) return;
  const double d1 = pow2 ? b1 + 2.0 * a1 : b1, d2 = pow2 ? b2 + 2.0 * a2 : b2;
  const long o = hidx(g, i, j, k);
  r1[o] = (TC)(v1 / (TC)d1);
  if (j >= 2) r2[o] = (TC)(v2 / (TC)d2);
  if (k >= 2) r3[o] = (TC)(v3 / (TC)d1);
}

// ---------------------------------------------------------------------------
// Partitioned Thomas solve of the r and z sweeps (solqj/solqk with
// solve_lap_par/solve_lap_thom, step.F90:266-485; tripar.F90:11-145).
// A block owns 64 lines (one per lane, consecutive theta -> coalesced) and P
// waves; wave s owns positions [s*M, s*M+M) of every line and keeps them in
// registers.  The two linear recurrences of the Thomas algorithm are closed
// across segments with precomputed carry products (one LDS exchange each).
// ---------------------------------------------------------------------------
struct LineArgs {
  LineCoef lc;
  VGrid g;
  long nlines;
  int j0;                      // first r index of z-lines
  double slip;
  const double *q;             // state component (z sweeps / q2 r sweep)
  double *dq;                  // predictor output (z sweeps of q1,q3 / r sweep of q2)
  void *r;                     // qrhs of the component (TW)
  const double *corr3n;        // outlet row of the q3 z sweep
  double c3s;                  // inlet row value of the q3 z sweep (uniform in theta)
  const double *c3s_dev;       // same value read from the device (captured steps), or null
  int jlim;                    // inlet rows: j <= jlim
};

template <typename TW, int M, int MODE>
__global__ __attribute__((amdgpu_flat_work_group_size(64, 1024), amdgpu_waves_per_eu(1, 4))) void k_line(LineArgs a) {
  __shared__ TW sh_a[16][64];
  const int lane = threadIdx.x & 63;
  const int s = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long L = (long)blockIdx.x * 64 + lane;
  const bool act = L < a.nlines;
  const long Lc = act ? L : a.nlines - 1;  // addresses stay in bounds; results of !act lanes are dropped
  const VGrid &g = a.g;
  const int n = a.lc.n, P = a.lc.P;
  const int p0 = s * M;
  const int len = min(M, n - p0);
  TW *rr = reinterpret_cast<TW *>(a.r);
  // the coefficients in the work type (the mixed path's float copies)
  auto cw = [](const double *d, const float *f) {
    if constexpr (sizeof(TW) == 4) { (void)d; return f; } else { (void)f; return d; }
  };
  const TW *cA = cw(a.lc.A, a.lc.Af), *cB = cw(a.lc.B, a.lc.Bf), *cC = cw(a.lc.C, a.lc.Cf);
  const TW *cPk = cw(a.lc.Pk, a.lc.Pkf), *cQk = cw(a.lc.Qk, a.lc.Qkf);
  const TW *cPs = cw(a.lc.Pseg, a.lc.Psegf), *cQs = cw(a.lc.Qseg, a.lc.Qsegf);
  // line -> (i, second index)
  const int i = (int)(Lc % g.nt) + 1;
  const int l2 = (int)(Lc / g.nt);
  long base;  // address of position 0 of the line
  long step;  // address stride between positions
  if (MODE <= SW_R2) { base = hidx(g, i, 1, l2 + 1); step = g.sj; }      // r lines at k = l2+1
  else { base = hidx(g, i, l2 + a.j0, 1); step = g.sk; }                  // z lines at j = l2+j0
  const int jz = l2 + a.j0;
  // every load of the segment is issued before the first use (positions
  // beyond the line clamp to its last one and are masked afterwards)
  TW x[M];
#pragma unroll
  for (int t = 0; t < M; ++t) x[t] = rr[base + step * min(p0 + t, n - 1)];
  TW rlast = (TW)0;
  if (MODE == SW_Z3) rlast = (TW)a.corr3n[(long)(i - 1) + (long)g.nt * jz];
  __builtin_amdgcn_sched_barrier(0);
  TW gp = (TW)0;
#pragma unroll
  for (int t = 0; t < M; ++t) {
    const int p = p0 + t;
    TW rv = x[t];
    if (MODE == SW_R2 || MODE == SW_Z1 || MODE == SW_Z2) {
      if (p == 0 || p == n - 1) rv = (TW)0;
    } else if (MODE == SW_Z3) {
      if (p == 0) rv = (jz <= a.jlim) ? (TW)(a.c3s_dev ? *a.c3s_dev : a.c3s) : (TW)0;
      else if (p == n - 1) rv = rlast;
    }
    const TW v = cA[p] * rv + cB[p] * gp;
    const bool in = t < len;
    x[t] = in ? v : (TW)0;
    gp = in ? v : gp;
  }
  sh_a[s][lane] = gp;
  __syncthreads();
  TW G = (TW)0;
  for (int s2 = 0; s2 < s; ++s2) G = sh_a[s2][lane] + cPs[s2] * G;
  __syncthreads();
  TW xp = (TW)0;
#pragma unroll
  for (int t = M - 1; t >= 0; --t) {
    const int p = p0 + t;
    const TW v = x[t] + cPk[p] * G - cC[p] * xp;
    const bool in = t < len;
    x[t] = in ? v : x[t];
    xp = in ? v : xp;
  }
  sh_a[s][lane] = xp;
  __syncthreads();
  TW X = (TW)0;
  for (int s2 = P - 1; s2 > s; --s2) X = sh_a[s2][lane] + cQs[s2] * X;
#pragma unroll
  for (int t = 0; t < M; ++t) x[t] = x[t] + cQk[p0 + t] * X;
  if (MODE == SW_R1 || MODE == SW_R3 || MODE == SW_Z2) {
    if (!act) return;
#pragma unroll
    for (int t = 0; t < M; ++t)
      if (t < len) rr[base + step * (p0 + t)] = x[t];
  } else {
    // dq = q + x, in chunks of 8 positions whose q loads go out together
    // before their stores
    constexpr int CH = 8;
#pragma unroll
    for (int c0 = 0; c0 < M; c0 += CH) {
      double qv[CH];
#pragma unroll
      for (int u = 0; u < CH; ++u)
        qv[u] = (c0 + u < M) ? a.q[base + step * min(p0 + c0 + u, n - 1)] : 0.0;
      __builtin_amdgcn_sched_barrier(0);
      if (act) {
#pragma unroll
        for (int u = 0; u < CH; ++u) {
          const int t = c0 + u;
          if (t < M && t < len) {
            const int p = p0 + t;
            const long ad = base + step * p;
            if (MODE == SW_R2) {
              // dq2 = x + q2 for j = 1..nr; dq2(nrc) = 0 (step.F90:330-333, 183-187)
              a.dq[ad] = (p == n - 1) ? 0.0 : (double)x[t] + qv[u];
            } else {
              // dq = q + x for k = 1..nzc; slip row at nrc (step.F90:84-88, 133-138)
              const double d = qv[u] + (double)x[t];
              a.dq[ad] = d;
              if (jz == g.nr) a.dq[ad + g.sj] = a.slip * d;
            }
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Poisson solve (poisson_solve.F90:140-179) in five passes: divergence +
// theta FFT, z DCT-II, radial Thomas, z DCT-III, inverse theta FFT.
// ---------------------------------------------------------------------------

// complex multiply (x + iy)(c + is)
template <typename T> __device__ __forceinline__ void cmul(T &x, T &y, T c, T s) {
  const T t = x * c - y * s;
  y = x * s + y * c;
  x = t;
}
// in-register DFT of R = 2, 4 or 8 points, v_s = sum_r v_r e^{i sign 2 pi r s / R}
template <typename T> __device__ __forceinline__ void dft2(T &ar, T &ai, T &br, T &bi) {
  const T tr = ar - br, ti = ai - bi;
  ar = ar + br; ai = ai + bi;
  br = tr; bi = ti;
}
template <typename T> __device__ __forceinline__ void dft4(T *vr, T *vi, int st, T sign) {
  // inputs at 0, st, 2 st, 3 st
  T t0r = vr[0], t0i = vi[0], t1r = vr[2 * st], t1i = vi[2 * st];
  dft2(t0r, t0i, t1r, t1i);  // t0 = v0 + v2, t1 = v0 - v2
  T t2r = vr[st], t2i = vi[st], t3r = vr[3 * st], t3i = vi[3 * st];
  dft2(t2r, t2i, t3r, t3i);  // t2 = v1 + v3, t3 = v1 - v3
  // t3 *= (i sign)
  const T u = t3r;
  t3r = -sign * t3i;
  t3i = sign * u;
  vr[0] = t0r + t2r; vi[0] = t0i + t2i;
  vr[2 * st] = t0r - t2r; vi[2 * st] = t0i - t2i;
  vr[st] = t1r + t3r; vi[st] = t1i + t3i;
  vr[3 * st] = t1r - t3r; vi[3 * st] = t1i - t3i;
}
template <typename T> __device__ __forceinline__ void dft8(T *vr, T *vi, T sign) {
  // even points 0,2,4,6 and odd points 1,3,5,7: two 4-point DFTs in place
  dft4(vr, vi, 2, sign);
  dft4(vr + 1, vi + 1, 2, sign);
  const T h = (T)0.70710678118654752440;
  // odd outputs s = 0..3 sit at 1,3,5,7: multiply by w8^s
  cmul(vr[3], vi[3], h, sign * h);
  { const T u = vr[5]; vr[5] = -sign * vi[5]; vi[5] = sign * u; }
  cmul(vr[7], vi[7], -h, sign * h);
  T er[4] = {vr[0], vr[2], vr[4], vr[6]}, ei[4] = {vi[0], vi[2], vi[4], vi[6]};
  T orr[4] = {vr[1], vr[3], vr[5], vr[7]}, oi[4] = {vi[1], vi[3], vi[5], vi[7]};
  for (int q = 0; q < 4; q++) {
    vr[q] = er[q] + orr[q]; vi[q] = ei[q] + oi[q];
    vr[q + 4] = er[q] - orr[q]; vi[q + 4] = ei[q] - oi[q];
  }
}

// radix of pass with `bits` bits still to do: 8 while the rest stays a
// multiple of 3, else 4 (2 for a lone bit)
__host__ __device__ constexpr int fft_pass_bits(int bits) {
  if (bits % 3 == 0) return 3;
  return bits == 1 ? 1 : 2;
}
__host__ inline int fft_min_radix(int logn) {
  int bits = logn, m = 8;
  while (bits > 0) { const int lr = fft_pass_bits(bits); m = m < (1 << lr) ? m : (1 << lr); bits -= lr; }
  return m;
}

// butterflies each thread holds across the barrier of an FFT pass (one: the
// register budget keeps five or more waves per SIMD in both precisions)
template <typename T> constexpr int fft_qt() { return 1; }

// One radix-2^LR pass of the Stockham autosort FFT (ls bits already done):
// every butterfly reads its R inputs at stride n/R, applies the twiddles,
// and after the barrier writes the in-register DFT at stride 2^ls.
// One radix-2^LR pass of the Stockham autosort FFT (ls bits already done):
// every butterfly reads its R inputs at stride n/R, applies the twiddles,
// and after the barrier writes the in-register DFT at stride 2^ls.  The
// length and pass (logn, ls) come either from template arguments (CL > 0:
// logn = CL, ls = CS; the LDS offsets r << lg then fold into immediates) or
// from the run-time arguments (CL = 0).
template <typename T, int LR, int QT, int CL = 0, int CS = 0>
__device__ __forceinline__ void fft_pass(T *re, T *im, int nfft, int logn_rt, int sstr, int ls_rt, const T *twr,
                                         const T *twi, T sign, int tid, int nth) {
  constexpr int R = 1 << LR;
  const int logn = CL > 0 ? CL : logn_rt, ls = CL > 0 ? CS : ls_rt;
  const int nh = 1 << (logn - 1), lg = logn - LR;
  const int tasks = nfft << lg;
  T vr[QT][R], vi[QT][R];
#pragma unroll
  for (int q = 0; q < QT; q++) {
    const int t = tid + q * nth;
    if (t < tasks) {
      const int f = t >> lg, j = t & ((1 << lg) - 1), k = j & ((1 << ls) - 1);
      const T *br = re + f * sstr, *bi = im + f * sstr;
#pragma unroll
      for (int r = 0; r < R; r++) {
        T x = br[j + (r << lg)], y = bi[j + (r << lg)];
        if (r > 0 && k > 0) {
          const int idx = (r * k) << (logn - ls - LR);
          T c, sn;
          if (idx < nh) { c = twr[idx]; sn = twi[idx]; }
          else { c = -twr[idx - nh]; sn = -twi[idx - nh]; }
          cmul(x, y, c, sign * sn);
        }
        vr[q][r] = x; vi[q][r] = y;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < QT; q++) {
    const int t = tid + q * nth;
    if (t < tasks) {
      const int f = t >> lg, j = t & ((1 << lg) - 1), k = j & ((1 << ls) - 1);
      if (LR == 3) dft8(vr[q], vi[q], sign);
      else if (LR == 2) dft4(vr[q], vi[q], 1, sign);
      else dft2(vr[q][0], vi[q][0], vr[q][1], vi[q][1]);
      const int d = ((j >> ls) << (ls + LR)) + k;
      T *br = re + f * sstr, *bi = im + f * sstr;
#pragma unroll
      for (int r = 0; r < R; r++) { br[d + (r << ls)] = vr[q][r]; bi[d + (r << ls)] = vi[q][r]; }
    }
  }
  __syncthreads();
}

// the passes of a transform of compile-time length 2^CL from pass CS on
template <typename T, int QT, int CL, int CS>
__device__ __forceinline__ void lds_fft_c(T *re, T *im, int nfft, int sstr, const T *twr, const T *twi, T sign,
                                          int tid, int nth) {
  if constexpr (CS < CL) {
    constexpr int LR = fft_pass_bits(CL - CS);
    fft_pass<T, LR, QT, CL, CS>(re, im, nfft, CL, sstr, CS, twr, twi, sign, tid, nth);
    lds_fft_c<T, QT, CL, CS + LR>(re, im, nfft, sstr, twr, twi, sign, tid, nth);
  }
}

// In-place Stockham autosort FFT of nfft sequences of length n = 2^logn in
// LDS (re, im; sequence f starts at f * sstr), natural order in and out,
// radix-8/4/2 passes with the DFT in
// registers.  Twiddles w_p = (twr[p], sign twi[p]), p < n/2, from LDS (the
// upper half by symmetry).  Each thread holds at most fft_qt<T>() butterflies
// of a pass across the barrier that separates the reads from the writes; the
// caller keeps nfft n / (min radix) <= fft_qt<T>() nth.
template <typename T>
__device__ __forceinline__ void lds_fft(T *re, T *im, int nfft, int logn, int sstr, const T *twr, const T *twi, T sign,
                                        int tid, int nth) {
  constexpr int QT = fft_qt<T>();
  // the lengths of the workloads' theta and z transforms: passes with
  // compile-time geometry
  switch (logn) {
    case 5: lds_fft_c<T, QT, 5, 0>(re, im, nfft, sstr, twr, twi, sign, tid, nth); return;
    case 6: lds_fft_c<T, QT, 6, 0>(re, im, nfft, sstr, twr, twi, sign, tid, nth); return;
    case 7: lds_fft_c<T, QT, 7, 0>(re, im, nfft, sstr, twr, twi, sign, tid, nth); return;
    case 8: lds_fft_c<T, QT, 8, 0>(re, im, nfft, sstr, twr, twi, sign, tid, nth); return;
    case 9: lds_fft_c<T, QT, 9, 0>(re, im, nfft, sstr, twr, twi, sign, tid, nth); return;
    case 10: lds_fft_c<T, QT, 10, 0>(re, im, nfft, sstr, twr, twi, sign, tid, nth); return;
    default: break;
  }
  int ls = 0;
  while (ls < logn) {
    const int lr = fft_pass_bits(logn - ls);
    if (lr == 3) fft_pass<T, 3, QT>(re, im, nfft, logn, sstr, ls, twr, twi, sign, tid, nth);
    else if (lr == 2) fft_pass<T, 2, QT>(re, im, nfft, logn, sstr, ls, twr, twi, sign, tid, nth);
    else fft_pass<T, 1, QT>(re, im, nfft, logn, sstr, ls, twr, twi, sign, tid, nth);
    ls += lr;
  }
}

// XCD-aware tile index: the dispatcher deals consecutive workgroups round
// robin to the 8 XCDs, each with its own L2.  Renumbering so that tiles t
// and t + 1 run on the same XCD, one dispatch round apart, lets tiles that
// share 128-byte lines (narrow tiles of a strided layout) meet in one L2
// instead of each fetching the whole line from HBM.  A bijection on
// [0, nblk) for any nblk.

// byte offset rounded up to 16
__host__ __device__ constexpr size_t al16(size_t b) { return (b + 15) & ~(size_t)15; }

// divq (divergence.F90:31-41) + fft_ptf_multi packing
// (fastfouriertransform.F90:39-43) for LB theta lines per block.  Power-of-two
// ntheta: two real lines per complex FFT in LDS.  Other ntheta: direct DFT.
// In mixed mode the packed positions 0..4 (m = 0,1,2) are also produced in
// double from the double divergence into S64 (layout (5, nr, nz)).
template <typename TW>
__global__ __launch_bounds__(256) void k_divfft(VGrid g, VCoef c, ZMet zm, double inv_adt, const double *__restrict__ dq1,
                                                const double *__restrict__ dq2, const double *__restrict__ dq3,
                                                const double *__restrict__ ext, TW *__restrict__ S,
                                                double *__restrict__ S64, const double *__restrict__ twr,
                                                const double *__restrict__ twi, int LB, int logn) {
  extern __shared__ unsigned char smem_raw[];
  const int N = g.nt, tid = threadIdx.x, nth = blockDim.x;
  const long nlines = (long)g.nr * g.nz;
  const long l0 = (long)blockIdx.x * LB;
  // LDS: [dv (LB*N double)] | re, im ((LB/2)*N TW each) | [tw (2N double), part (10 LB double)] |
  // twl (N TW) | jk (2 LB int) | 1/rm, dx2/dmr (LB double each).  The double
  // copy dv feeds only the direct DFT (ntheta not a power of two); the FP64
  // low modes of the mixed path are summed in registers.
  const bool pow2 = N >= 2 && (N & (N - 1)) == 0;
  const bool keep_dv = !pow2;
  size_t off = keep_dv ? al16(sizeof(double) * LB * N) : 0;
  double *dv = reinterpret_cast<double *>(smem_raw);
  TW *re = reinterpret_cast<TW *>(smem_raw + off);
  TW *im = re + (size_t)(LB / 2) * N;
  off = al16(off + sizeof(TW) * LB * N);
  double *tw = reinterpret_cast<double *>(smem_raw + off);     // cos (N) then sin (N), S64 only
  double *part = tw + 2 * (size_t)N;                           // (line, coefficient, 2 waves), S64 only
  if (S64) off = al16(off + sizeof(double) * (2 * (size_t)N + 10 * (size_t)LB));
  TW *twl = reinterpret_cast<TW *>(smem_raw + off);            // FFT twiddles: N/2 cos, N/2 sin
  off = al16(off + sizeof(TW) * N);
  int *jk = reinterpret_cast<int *>(smem_raw + off);           // (j, k) of each line
  // 1/rm(j) and dx2/dmr(j) of each line: divq's theta and r differences
  // (divergence.F90:38-39) share the 1/rm factor, so a cell needs no division
  double *lirm = reinterpret_cast<double *>(smem_raw + al16(off + sizeof(int) * 2 * LB));
  double *lcx2 = lirm + LB;
  if (tid < LB) {  // lines past the end are clamped to the last line (never stored)
    const long l = min(l0 + tid, nlines - 1);
    const int j = (int)(l % g.nr) + 1;
    jk[2 * tid] = j;
    jk[2 * tid + 1] = (int)(l / g.nr) + 1;
    lirm[tid] = 1.0 / c.rm[j];
    lcx2[tid] = g.dx2 / c.dmr[j];
  }
  if (pow2)
    for (int q = tid; q < N / 2; q += nth) { twl[q] = (TW)twr[q]; twl[N / 2 + q] = (TW)twi[q]; }
  if (S64)
    for (int q = tid; q < N; q += nth) { tw[q] = twr[q]; tw[N + q] = twi[q]; }
  __syncthreads();
  // TPL threads per line: thread tl of line lpt takes the angles tl,
  // tl + TPL, ... (consecutive lanes read consecutive angles), so its FP64
  // low-mode sums (mixed path) stay within one line
  const int TPL = nth / LB, lpt = tid / TPL, tl = tid - lpt * TPL;
  const int NU = (N + TPL - 1) / TPL;
  const bool lvalid = l0 + lpt < nlines;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
  auto put = [&](int ii, double val) {
    if (keep_dv) dv[lpt * N + ii] = val;
    if (pow2) {
      const int pos = (lpt >> 1) * N + ii;
      if (lpt & 1) im[pos] = (TW)val; else re[pos] = (TW)val;
    }
    if (S64) {  // cos and sin sums of wavenumbers 0, 1, 2
      int q2 = 2 * ii;
      if (q2 >= N) q2 -= N;
      s0 += val;
      s1 += val * tw[ii];
      s2 += val * tw[N + ii];
      s3 += val * tw[q2];
      s4 += val * tw[N + q2];
    }
  };
  const int jl = jk[2 * lpt], kl = jk[2 * lpt + 1];
  if (ext) {
    for (int u = 0; u < NU; ++u) {
      const int ii = tl + TPL * u;
      if (ii < N) put(ii, lvalid ? ext[cidx(g, ii + 1, jl, kl)] : 0.0);
    }
  } else {
    // the loads of U angles are all issued before the first is used;
    // indices are clamped to valid cells, so the loads are unconditional
    constexpr int U = 4;
    const double rmv = lirm[lpt], rdv = lcx2[lpt];
    const double zkv = zm.m3 ? zm.m3[kl] : g.dx3;  // dx3, or the axial mesh's 1/width of plane kl
    const long ob = hidx(g, 1, jl, kl);
    for (int ub = 0; ub < NU; ub += U) {
      double a1[U], b1[U], a2[U], b2[U], a3[U], b3[U];
      #pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = min(tl + TPL * (ub + u), N - 1) + 1;
        const long o = ob + (i - 1);
        a1[u] = dq1[o];
        b1[u] = dq1[o + ((i == N) ? 1 - N : 1)];
        a2[u] = dq2[o];
        b2[u] = dq2[o + g.sj];
        a3[u] = dq3[o];
        b3[u] = dq3[o + g.sk];
      }
      __builtin_amdgcn_sched_barrier(0);
      #pragma unroll
      for (int u = 0; u < U; ++u) {
        const int ii = tl + TPL * (ub + u);
        if (ii < N) {
          const double t12 = ((b1[u] - a1[u]) * (g.dx1 * g.axsym) + (b2[u] - a2[u]) * rdv) * rmv;
          const double t3 = (b3[u] - a3[u]) * zkv;
          put(ii, lvalid ? -(t12 + t3) * inv_adt : 0.0);
        }
      }
    }
  }
  if (S64) {
    // the line's TPL lanes: a fixed butterfly within each wave (deterministic
    // order), then one partial per wave into LDS
    const int grp = TPL < 64 ? TPL : 64;
    for (int o = grp >> 1; o > 0; o >>= 1) {
      s0 += __shfl_xor(s0, o, grp);
      s1 += __shfl_xor(s1, o, grp);
      s2 += __shfl_xor(s2, o, grp);
      s3 += __shfl_xor(s3, o, grp);
      s4 += __shfl_xor(s4, o, grp);
    }
    if ((tl & (grp - 1)) == 0) {
      double *pp = part + (size_t)lpt * 10 + tl / grp;
      pp[0] = s0; pp[2] = s1; pp[4] = s2; pp[6] = s3; pp[8] = s4;
    }
  }
  __syncthreads();
  if (pow2) lds_fft(re, im, LB / 2, logn, N, twl, twl + N / 2, (TW)-1, tid, nth);
  // a power-of-two N scales exactly, so x * (1/N) is x / N bit for bit
  // without a double division per element
  const double inv_n = 1.0 / N;
  for (int e = tid; e < LB * N; e += nth) {
    const int lp = pow2 ? e >> logn : e / N, pos = e - lp * N;
    const long l = l0 + lp;
    if (l >= nlines) continue;
    const int k = (pos + 1) >> 1;  // wavenumber of this packed position
    double Fr, Fi;
    if (pow2) {
      const int q = lp >> 1, kr = (N - k) & (N - 1);
      const double zr = re[q * N + k], zi = im[q * N + k], wr = re[q * N + kr], wi = im[q * N + kr];
      if ((lp & 1) == 0) { Fr = 0.5 * (zr + wr); Fi = 0.5 * (zi - wi); }
      else { Fr = 0.5 * (zi + wi); Fi = -0.5 * (zr - wr); }
    } else {
      Fr = 0.0; Fi = 0.0;
      int ph = 0;
      for (int jj = 0; jj < N; ++jj) {
        const double x = dv[lp * N + jj];
        Fr += x * twr[ph];
        Fi -= x * twi[ph];
        ph += k;
        if (ph >= N) ph -= N;
      }
    }
    const double num = (pos == 0) ? Fr : (pos & 1) ? 2.0 * Fr : -2.0 * Fi;
    const double outv = pow2 ? num * inv_n : num / N;
    S[(size_t)pos + (size_t)N * l] = (TW)outv;
  }
  if (S64) {
    // FP64 modes m = 0, 1, 2 (t(1), t(2k), t(2k+1) packing) from the wave
    // partials, summed in wave order
    const int nw = TPL > 64 ? TPL / 64 : 1;
    for (int e = tid; e < LB * 5; e += nth) {
      const int lp = e / 5, c5 = e - lp * 5;
      const long l = l0 + lp;
      if (l >= nlines || c5 >= N) continue;
      double acc = part[(size_t)e * 2];
      for (int w = 1; w < nw; ++w) acc += part[(size_t)e * 2 + w];
      S64[(size_t)c5 + 5 * l] = (c5 == 0) ? acc / N : 2.0 * acc / N;
    }
  }
}

// z DCT-II (REDFT10, then /(2nz): fastfouriertransform.F90:299-332) of TL
// lines per block; lines are contiguous in memory (line stride 1, position
// stride nl).  Makhoul's reordering turns each pair of real lines into one
// complex FFT of length nz.  dir=+1 forward, -1 inverse (REDFT01).
template <typename TW>
__global__ __launch_bounds__(1024) void k_dctz(TW *__restrict__ S, long nl, int n, int logn, int TL, int logtl, int dir,
                                              const double *__restrict__ twr, const double *__restrict__ twi,
                                              const double *__restrict__ er, const double *__restrict__ ei) {
  extern __shared__ unsigned char smem_raw[];
  TW *re = reinterpret_cast<TW *>(smem_raw);
  // sequences at stride n + 1 in LDS: the line-to-sequence transposes
  // below touch consecutive sequences from consecutive lanes, which would
  // all hit one bank at stride n
  const int sn = n + 1;
  TW *im = re + (size_t)(TL / 2) * sn;
  TW *twl = im + (size_t)(TL / 2) * sn;  // n/2 cos, n/2 sin
  // Makhoul post/pre-twiddles e^{i pi k/2n}: in FP32 an LDS copy (n cos,
  // n sin) spares the per-element global loads; FP64 reads them from global
  // memory (the LDS would cost the tile lines)
  constexpr bool lds_tw = sizeof(TW) == 4;
  TW *etw = twl + n;
  auto ecos = [&](int k) { return lds_tw ? etw[k] : (TW)er[k]; };
  auto esin = [&](int k) { return lds_tw ? etw[n + k] : (TW)ei[k]; };
  const int tid = threadIdx.x, nth = blockDim.x;
  const long l0 = xcd_tile(blockIdx.x, gridDim.x) * TL;
  const bool pow2 = n >= 2 && (n & (n - 1)) == 0;
  if (pow2) {
    for (int q = tid; q < n / 2; q += nth) { twl[q] = (TW)twr[q]; twl[n / 2 + q] = (TW)twi[q]; }
    if (lds_tw)
      for (int q = tid; q < n; q += nth) { etw[q] = (TW)er[q]; etw[n + q] = (TW)ei[q]; }
    // Element u of a thread is line lp = tid mod TL (consecutive lanes,
    // consecutive lines) at position kz0 + u dk, dk = nth / TL (a multiple
    // of 8, so every position of a thread has the parity of kz0).  The
    // launch keeps TL n <= 16 nth; lanes with kz0 >= n (TL n < nth) only
    // help with the LDS phases.  Lines past nl load the last line and are
    // zeroed (forward) or not stored.
    constexpr int PF = 16;
    const int tot = TL * n, lp = tid & (TL - 1), kz0 = tid >> logtl, dk = nth >> logtl;
    const int nu = min(PF, (tot + nth - 1) / nth);
    const bool lin = l0 + lp < nl;
    TW v[PF];
    {
      const int kc = min(kz0, n - 1);
      const TW *src = S + min(l0 + lp, nl - 1) + nl * kc;
      const long step = nl * dk, last = (long)((n - 1 - kc) / dk) * step;
#pragma unroll
      for (int u = 0; u < PF; ++u)
        if (u < nu) v[u] = src[min(u * step, last)];
    }
    __syncthreads();  // twiddle tables
    // Makhoul order of position kz: kz/2 for even kz, n-1-kz/2 for odd kz,
    // for the thread's positions pm0 + u dpm.  Forward: the lines go into
    // the pair's complex sequence in Makhoul order (and come out in natural
    // order); inverse: in natural order (and out in Makhoul order).
    TW *const dpair = ((lp & 1) ? im : re) + (lp >> 1) * sn;
    const int pm0 = (kz0 & 1) ? n - 1 - (kz0 >> 1) : (kz0 >> 1);
    const int dpm = (kz0 & 1) ? -(dk >> 1) : (dk >> 1);
    const int pi0 = dir > 0 ? pm0 : kz0, dpi = dir > 0 ? dpm : dk;
#pragma unroll
    for (int u = 0; u < PF; ++u)
      if (u < nu && kz0 + u * dk < n) dpair[pi0 + u * dpi] = lin ? v[u] : (TW)0;
    __syncthreads();
    if (dir < 0) {
      // in place, one (k, n-k) pair per task (k = 0..n/2, Y[n] = 0):
      // W[k] = e^{i pi k/2n} ((Ya[k] + Yb[n-k]) + i (Yb[k] - Ya[n-k]))
      const int ltq = logtl - 1, TQ = TL >> 1, nh = n >> 1, tq = TQ * (nh + 1);
      for (int e = tid; e < tq; e += nth) {
        const int q = e & (TQ - 1), k = e >> ltq, kn = (n - k) & (n - 1);
        TW *rq = re + q * sn, *iq = im + q * sn;
        const TW a0 = rq[k], b0 = iq[k];
        const TW an = k > 0 ? rq[kn] : (TW)0, bn = k > 0 ? iq[kn] : (TW)0;
        {
          const TW uu = a0 + bn, w = b0 - an, c = ecos(k), sv = esin(k);
          rq[k] = uu * c - w * sv;
          iq[k] = uu * sv + w * c;
        }
        if (k > 0 && k < nh) {
          const TW uu = an + b0, w = bn - a0, c = ecos(kn), sv = esin(kn);
          rq[kn] = uu * c - w * sv;
          iq[kn] = uu * sv + w * c;
        }
      }
      __syncthreads();
    }
    lds_fft(re, im, TL / 2, logn, sn, twl, twl + n / 2, (TW)(dir > 0 ? -1 : 1), tid, nth);
    const TW inv_n = (TW)1 / (TW)n;  // exact: n is a power of two
    if (kz0 < n && lin) {
      TW *dst = S + (l0 + lp) + nl * kz0;
      const long step = nl * dk;
      const TW *zq = re + (lp >> 1) * sn, *zi = im + (lp >> 1) * sn;
#pragma unroll
      for (int u = 0; u < PF; ++u) {
        const int kz = kz0 + u * dk;
        if (u < nu && kz < n) {
          TW outv;
          if (dir > 0) {
            const int kr = (n - kz) & (n - 1);
            const TW zr = zq[kz], zim = zi[kz], wr = zq[kr], wi = zi[kr];
            TW vr, vi;
            if ((lp & 1) == 0) { vr = (TW)0.5 * (zr + wr); vi = (TW)0.5 * (zim - wi); }
            else { vr = (TW)0.5 * (zim + wi); vi = (TW)(-0.5) * (zr - wr); }
            outv = (vr * ecos(kz) + vi * esin(kz)) * inv_n;  // Re(e^{-i pi k/2n} V)/n
          } else {
            outv = dpair[pm0 + u * dpm];
          }
          dst[u * step] = outv;
        }
      }
    }
  } else {
    // direct transform for non-power-of-two nz (small test grids)
    TW *buf = re;  // TL * n
    for (int e = tid; e < TL * n; e += nth) {
      const int lp = e % TL, kz = e / TL;
      const long l = l0 + lp;
      buf[lp * n + kz] = (l < nl) ? S[l + nl * kz] : (TW)0;
    }
    __syncthreads();
    for (int e = tid; e < TL * n; e += nth) {
      const int lp = e % TL, kz = e / TL;
      const long l = l0 + lp;
      double sacc = 0.0;
      if (dir > 0) {
        for (int jj = 0; jj < n; ++jj) sacc += (double)buf[lp * n + jj] * cos(M_PI * (jj + 0.5) * kz / n);
        sacc = 2.0 * sacc / (2.0 * n);
      } else {
        sacc = buf[lp * n];
        for (int kk = 1; kk < n; ++kk) sacc += 2.0 * (double)buf[lp * n + kk] * cos(M_PI * kk * (kz + 0.5) / n);
      }
      if (l < nl) S[l + nl * kz] = (TW)sacc;
    }
  }
}

// Radial tridiagonal solves of every (mode, kz) system (solve_radial,
// poisson_solve.F90:270-323) with precomputed inverse pivots ib; lanes over
// the packed mode index (contiguous), P waves of M rows each.
// FP32: at most 64 VGPRs, so that two 1024-lane blocks share a CU.
template <typename T, int M>
__global__ __launch_bounds__(1024, sizeof(T) == 4 ? 8 : 1) void k_radial(T *__restrict__ S, const T *__restrict__ ib, long nm, long nlines,
                                                 int nr, const T *__restrict__ aab, const T *__restrict__ ccb) {
  __shared__ T sh_v[16][64];
  __shared__ T sh_p[16][64];
  const int lane = threadIdx.x & 63;
  const int s = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int P = blockDim.x >> 6;
  const long L = (long)blockIdx.x * 64 + lane;
  const bool act = L < nlines;
  const long Lc = act ? L : nlines - 1;  // addresses stay in bounds; results of !act lanes are dropped
  const long pos = Lc % nm, kz = Lc / nm;
  const long base = pos + nm * (long)nr * kz;
  const int p0 = s * M, len = min(M, nr - p0);
  // all loads of the segment go out before the first use; positions past
  // the line end clamp to its last row and are masked (uniform selects)
  T x[M], w[M];
#pragma unroll
  for (int t = 0; t < M; ++t) {
    const long ad = base + nm * min(p0 + t, nr - 1);
    w[t] = ib[ad];
    x[t] = S[ad];
  }
  __builtin_amdgcn_sched_barrier(0);
  T gp = (T)0, Pacc = (T)1;
#pragma unroll
  for (int t = 0; t < M; ++t) {
    const int p = p0 + t, pc = min(p, nr - 1);
    const bool in = t < len;
    const T ibv = w[t];
    const T B = -aab[pc + 1] * ibv * (T)(p > 0);
    const T v = ibv * x[t] + B * gp;
    x[t] = in ? v : (T)0;
    gp = in ? v : gp;
    Pacc = in ? Pacc * B : Pacc;
  }
  sh_v[s][lane] = gp;
  sh_p[s][lane] = Pacc;
  __syncthreads();
  T G = (T)0;
  for (int s2 = 0; s2 < s; ++s2) G = sh_v[s2][lane] + sh_p[s2][lane] * G;
  __syncthreads();
  T Pr = (T)1;
#pragma unroll
  for (int t = 0; t < M; ++t) {
    const int p = p0 + t, pc = min(p, nr - 1);
    const bool in = t < len;
    Pr = in ? Pr * (-aab[pc + 1] * w[t] * (T)(p > 0)) : Pr;
    x[t] = in ? x[t] + Pr * G : x[t];
  }
  T xp = (T)0, Qacc = (T)1;
#pragma unroll
  for (int t = M - 1; t >= 0; --t) {
    const int p = p0 + t, pc = min(p, nr - 1);
    const bool in = t < len;
    const T C = ccb[pc + 1] * w[t] * (T)(p < nr - 1);
    const T v = x[t] - C * xp;
    x[t] = in ? v : x[t];
    xp = in ? v : xp;
    Qacc = in ? Qacc * -C : Qacc;
  }
  sh_v[s][lane] = xp;
  sh_p[s][lane] = Qacc;
  __syncthreads();
  T X = (T)0;
  for (int s2 = P - 1; s2 > s; --s2) X = sh_v[s2][lane] + sh_p[s2][lane] * X;
  if (!act) return;
  T Qr = (T)1;
#pragma unroll
  for (int t = M - 1; t >= 0; --t) {
    if (t < len) {
      const int p = p0 + t;
      Qr *= -ccb[p + 1] * w[t] * (T)(p < nr - 1);
      S[base + nm * p] = x[t] + Qr * X;
    }
  }
}

// fft_ftp_multi (fastfouriertransform.F90:74-82) + copy into dph
// (poisson_solve.F90:29) + dph(nzc)=dph(nz) (divergence.F90:93-99).
template <typename TW>
__global__ __launch_bounds__(256) void k_ifft(VGrid g, const TW *__restrict__ S, const double *__restrict__ S64,
                                              double *__restrict__ phi_out, double *__restrict__ dph,
                                              const double *__restrict__ twr, const double *__restrict__ twi, int LB,
                                              int logn) {
  extern __shared__ unsigned char smem_raw[];
  const int N = g.nt, tid = threadIdx.x, nth = blockDim.x;
  const long nlines = (long)g.nr * g.nz;
  const long l0 = (long)blockIdx.x * LB;
  TW *re = reinterpret_cast<TW *>(smem_raw);
  TW *im = re + (size_t)(LB / 2) * N;
  TW *tb = im + (size_t)(LB / 2) * N;  // packed input, LB*N
  TW *twl = tb + (size_t)LB * N;       // N/2 cos, N/2 sin
  const size_t ojk = al16(sizeof(TW) * ((size_t)2 * LB * N + N));
  int *jk = reinterpret_cast<int *>(smem_raw + ojk);
  // mixed path: the FP64 low modes of the block's lines and an FP64 cos/sin
  // table (N each) staged in LDS for the merge
  double *s64l = reinterpret_cast<double *>(smem_raw + al16(ojk + sizeof(int) * 2 * LB));
  double *tw64 = s64l + 5 * (size_t)LB;
  const bool pow2 = N >= 2 && (N & (N - 1)) == 0;
  if (tid < LB) {
    const long l = l0 + tid;
    jk[2 * tid] = (int)(l % g.nr) + 1;
    jk[2 * tid + 1] = (int)(l / g.nr) + 1;
  }
  if (pow2)
    for (int q = tid; q < N / 2; q += nth) { twl[q] = (TW)twr[q]; twl[N / 2 + q] = (TW)twi[q]; }
  if (S64) {
    for (int q = tid; q < N; q += nth) { tw64[q] = twr[q]; tw64[N + q] = twi[q]; }
    for (int e = tid; e < 5 * LB; e += nth) s64l[e] = S64[5 * min(l0 + e / 5, nlines - 1) + e % 5];
  }
  {
    // batches of U elements: all loads (line clamped, then masked) go out
    // before the first is used
    constexpr int U = 8;
    const int tot = LB * N;
    for (int eb = tid; eb < tot; eb += U * nth) {
      TW v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = min(eb + u * nth, tot - 1);
        const int lp = pow2 ? e >> logn : e / N, pos = e - lp * N;
        v[u] = S[(size_t)pos + (size_t)N * min(l0 + lp, nlines - 1)];
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = eb + u * nth;
        if (e < tot) {
          const int lp = pow2 ? e >> logn : e / N, pos = e - lp * N;
          tb[e] = (l0 + lp < nlines && !(S64 && pos < 5)) ? v[u] : (TW)0;
        }
      }
    }
  }
  __syncthreads();
  if (pow2) {
    // Z[k] = Fa[k] + i Fb[k], Fa[k] = (t(2k-1) - i t(2k))/2, Hermitian extension
    for (int e = tid; e < (LB / 2) * N; e += nth) {
      const int q = e >> logn, k = e - q * N;
      double ar, ai, br, bi;
      auto coef = [&](const TW *t, int kk, double &cr, double &ci) {
        if (kk == 0) { cr = t[0]; ci = 0.0; return; }
        int kh = kk <= N / 2 ? kk : N - kk;
        double a = t[2 * kh - 1];
        double b = (2 * kh < N) ? t[2 * kh] : 0.0;
        cr = 0.5 * a;
        ci = (kk <= N / 2) ? -0.5 * b : 0.5 * b;
        if (2 * kh == N) ci = 0.0;
      };
      coef(tb + (2 * q) * N, k, ar, ai);
      coef(tb + (2 * q + 1) * N, k, br, bi);
      const int slot = q * N + k;
      re[slot] = (TW)(ar - bi);
      im[slot] = (TW)(ai + br);
    }
    __syncthreads();
    lds_fft(re, im, LB / 2, logn, N, twl, twl + N / 2, (TW)1, tid, nth);
  }
  if (pow2 && (nth & (N - 1)) == 0) {
    // N divides the block: every element of a thread has the same theta
    // index ii, so the low-mode twiddles of the merge are loaded once
    const int ii = tid & (N - 1), lstep = nth >> logn;
    double cw1 = 0.0, sw1 = 0.0, cw2 = 0.0, sw2 = 0.0;
    if (S64) {
      int ph2 = 2 * ii;
      if (ph2 >= N) ph2 -= N;
      cw1 = tw64[ii]; sw1 = tw64[N + ii];
      if (3 < N) { cw2 = tw64[ph2]; sw2 = tw64[N + ph2]; }
    }
    for (int lp = tid >> logn; lp < LB; lp += lstep) {
      const long l = l0 + lp;
      if (l >= nlines) break;
      const int q = lp >> 1;
      double v = (lp & 1) ? (double)im[q * N + ii] : (double)re[q * N + ii];
      if (S64) {
        // the terms of the generic loop below, in its order
        const double *t = s64l + 5 * lp;
        v += t[0];
        {
          const double a = t[1], b = (2 < N) ? t[2] : 0.0;
          if (2 == N) v += 0.5 * a * cw1;
          else v += a * cw1 + b * sw1;
        }
        if (3 < N) {
          const double a = t[3], b = (4 < N) ? t[4] : 0.0;
          if (4 == N) v += 0.5 * a * cw2;
          else v += a * cw2 + b * sw2;
        }
      }
      const int j = jk[2 * lp], k = jk[2 * lp + 1];
      if (phi_out) phi_out[cidx(g, ii + 1, j, k)] = v;
      if (dph) {
        dph[hidx(g, ii + 1, j, k)] = v;
        if (k == g.nz) dph[hidx(g, ii + 1, j, g.nzc)] = v;
      }
    }
    return;
  }
  for (int e = tid; e < LB * N; e += nth) {
    const int lp = pow2 ? e >> logn : e / N, ii = e - lp * N;
    const long l = l0 + lp;
    if (l >= nlines) continue;
    double v;
    if (pow2) {
      const int q = lp >> 1;
      v = (lp & 1) ? (double)im[q * N + ii] : (double)re[q * N + ii];
    } else {
      const TW *t = tb + lp * N;
      v = t[0];
      for (int kk = 1; 2 * kk - 1 < N; ++kk) {
        const int ph = (int)(((long)ii * kk) % N);
        const double a = t[2 * kk - 1], b = (2 * kk < N) ? t[2 * kk] : 0.0;
        if (2 * kk == N) v += 0.5 * a * twr[ph];
        else v += a * twr[ph] + b * twi[ph];
      }
    }
    if (S64) {
      const double *t = s64l + 5 * lp;
      v += t[0];
      for (int kk = 1; kk <= 2 && 2 * kk - 1 < N; ++kk) {
        int ph = ii * kk;
        if (ph >= N) ph -= N;
        const double a = t[2 * kk - 1], b = (2 * kk < N) ? t[2 * kk] : 0.0;
        if (2 * kk == N) v += 0.5 * a * tw64[ph];
        else v += a * tw64[ph] + b * tw64[N + ph];
      }
    }
    const int j = jk[2 * lp], k = jk[2 * lp + 1];
    if (phi_out) phi_out[cidx(g, ii + 1, j, k)] = v;
    if (dph) {
      dph[hidx(g, ii + 1, j, k)] = v;
      if (k == g.nz) dph[hidx(g, ii + 1, j, g.nzc)] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// upsol (divergence.F90:69-236) fused with uppres (divergence.F90:239-480).
// A block owns a 64 (theta) x R (r) column tile and marches up a chunk of z
// planes.  The planes k-2..k+1 of dph live in a 4-slot LDS ring (with two
// theta columns and two r rows of halo on the low side, one on the high
// side), so each dph value is fetched from HBM about once; the gradient
// components qcap are recomputed from the ring where the reference stores
// them.  Everything that depends on the thread's r row only (metric ratios,
// reciprocals of rm) is hoisted out of the z loop.
// ---------------------------------------------------------------------------
template <typename TW, int R, bool ZS>
__global__ __launch_bounds__(64 * R) void k_project(VGrid g, VCoef c, ZMet zm, double adt, double be, int kchunk,
                                                    const double *__restrict__ dph, const double *__restrict__ dq1,
                                                    const double *__restrict__ dq2, const double *__restrict__ dq3,
                                                    double *__restrict__ q1, double *__restrict__ q2,
                                                    double *__restrict__ q3, double *__restrict__ prg1,
                                                    double *__restrict__ prg2, double *__restrict__ prg3,
                                                    double *__restrict__ pres, int *__restrict__ flag) {
  constexpr int W = 64, LC = W + 3, LR = R + 3, NE = LR * LC, NT = 64 * R;
  constexpr int NL = (NE + NT - 1) / NT;
  __shared__ double ring[4][LR][LC + 1];
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * W + tx;
  const int nt = g.nt, nr = g.nr, nz = g.nz, nrc = g.nrc, nzc = g.nzc;
  const long sj = g.sj, sk = g.sk;
  const int i0 = blockIdx.x * W;  // 0-based theta of the tile's first column
  const int j0 = blockIdx.y * R;  // rows 0..nrc (row 0 only takes the inlet copy of q3)
  const int kb = 1 + blockIdx.z * kchunk;
  const int ke = min(kb + kchunk - 1, nzc);
  // ring fill: element e = tid + l*NT of the (LR x LC) plane tile
  long goff[NL];
  int lrow[NL], lcol[NL];
  bool lval[NL];
#pragma unroll
  for (int l = 0; l < NL; l++) {
    const int e = tid + l * NT;
    const int r = e / LC, cc = e - r * LC;
    const int jj = j0 - 2 + r;
    int ii = i0 - 2 + cc;
    ii %= nt;
    if (ii < 0) ii += nt;
    lrow[l] = r; lcol[l] = cc;
    lval[l] = (e < NE) && jj >= 0 && jj <= nr + 2;
    goff[l] = lval[l] ? (long)ii + (long)jj * sj : 0;
  }
  double pre[NL];
  // unconditional loads from in-bounds (clamped) addresses, masked after:
  // the loads of one plane issue together
  auto fetch = [&](int kk) {
    const long kc = (long)min(max(kk, 0), nz + 2) * sk;
    const bool kin = kk >= 0 && kk <= nz + 2;
#pragma unroll
    for (int l = 0; l < NL; l++) {
      const double v = dph[goff[l] + kc];
      pre[l] = (lval[l] && kin) ? v : 0.0;
    }
  };
  auto stash = [&](int kk) {
#pragma unroll
    for (int l = 0; l < NL; l++)
      if (tid + l * NT < NE) ring[kk & 3][lrow[l]][lcol[l]] = pre[l];
  };
  fetch(kb - 2); stash(kb - 2);
  fetch(kb - 1); stash(kb - 1);
  fetch(kb); stash(kb);
  fetch(kb + 1);
  // the thread's cell column
  const int i = i0 + tx + 1, j = j0 + ty;
  const bool act = (i <= nt) && (j <= nrc);
  const int jr = (j <= nrc) ? j : nrc;           // clamped row for coefficient loads
  const int jq = (jr >= 1 && jr <= nr) ? jr : 1;  // rows of the uppres coefficients
  const int djp = (jq == nr) ? 0 : 1, djm = (jq == 1) ? 0 : -1;
  // radial gradient rc dx2 / dcr (divergence.F90:132, 327)
  const double a1 = adt * g.dx1 / c.rm[jq], a2 = adt * c.rc[jr] * g.dx2 / c.dcr[jr], a3 = adt * g.dx3;
  const double irm = 1.0 / c.rm[jq], irmp = 1.0 / c.rm[jq + djp], irmm = 1.0 / c.rm[jq + djm];
  const double irmn = 1.0 / c.rm[jq + 1];
  const double s1t11 = c.s1t11[jq], s1t2n = c.s1t2n[jq], s1t2s = c.s1t2s[jq], s1t2c = c.s1t2c[jq];
  const double s2t11 = c.s2t11[jq], s2t2p = c.s2t2p[jq], s2t2m = c.s2t2m[jq];
  const double s3t2n = c.s3t2n[jq], s3t2s = c.s3t2s[jq];
  const int jqp = min(jq + 1, nrc), jqm = max(jq - 1, 1);
  const double rcj = c.rc[jq] / c.dcr[jq], rcp = c.rc[jqp] / c.dcr[jqp], rcm = c.rc[jqm] / c.dcr[jqm];
  const double dx1 = g.dx1, dx2 = g.dx2, dx3 = g.dx3, dx3q = g.dx3q, slip = g.slip;
  const int cx = tx + 2, cy = ty + 2;
  // the streamed operands of the next plane are loaded while this one is computed
  const bool rowu = act && j >= 1 && j <= nr;
  double n_dq1 = 0.0, n_dq2 = 0.0, n_dq3 = 0.0, n_pres = 0.0;
  double n_p1 = 0.0, n_p2 = 0.0, n_p3 = 0.0;
  // operands of cell (i, j, kk); lanes, rows and planes where a value is not
  // used read a clamped in-bounds address instead of branching, so that the
  // seven loads issue together
  const int ic = min(i, nt), jc1 = min(max(j, 1), nr);
  auto load_cell = [&](int kk) {
    const int kc = min(kk, ke), kp = min(max(kc, 1), nz);
    const long oo = (long)(ic - 1) + (long)min(j, nrc) * sj + (long)kc * sk;
    const long pp = (long)(ic - 1) + (long)nt * ((long)(jc1 - 1) + (long)nr * (long)(kp - 1));
    n_dq3 = dq3[oo];
    n_dq1 = dq1[oo];
    n_dq2 = dq2[oo];
    n_p1 = prg1[pp];
    n_p2 = prg2[pp];
    n_p3 = prg3[pp];
    n_pres = pres[oo];
  };
  load_cell(kb);
  for (int k = kb; k <= ke; k++) {
    stash(k + 1);
    const double c_dq1 = n_dq1, c_dq2 = n_dq2, c_dq3 = n_dq3, c_pres = n_pres;
    const double c_p1 = n_p1, c_p2 = n_p2, c_p3 = n_p3;
    __syncthreads();
    fetch(k + 2);
    load_cell(k + 1);
    if (act) {
      const double(*P0)[LC + 1] = ring[k & 3];
      const double(*Pm)[LC + 1] = ring[(k - 1) & 3];
      const double(*Pp)[LC + 1] = ring[(k + 1) & 3];
      const double(*Pm2)[LC + 1] = ring[(k - 2) & 3];
      // kmv quirk of uppres: km = 1 at k = 1 and km = nz at k = nz
      const double(*Pkm)[LC + 1] = (k == 1 || k == nz) ? P0 : Pm;
      const long o = (long)(i - 1) + (long)j * sj + (long)k * sk;
      const double p0 = P0[cy][cx];
      // axial mesh: the z gradient at faces k, k+1, k-1 and the z-gradient
      // factor of upsol; the uniform mesh keeps dx3
      constexpr bool zs = ZS;  // axial mesh (a separate instantiation)
      const double zc0 = zs ? zm.c3[k] : dx3, zcp = zs ? zm.c3[k + 1] : dx3, zcm = zs ? zm.c3[k - 1] : dx3;
      const double a3k = zs ? adt * zc0 : a3;
      // ---- upsol ----
      if (j >= 1 && j <= nr) {
        const double v = c_dq1 - a1 * (p0 - P0[cy][cx - 1]);
        q1[o] = v;
        if (j == nr) q1[o + sj] = slip * v;
      }
      if (j >= 2 && j <= nr) q2[o] = c_dq2 - a2 * (p0 - P0[cy - 1][cx]);
      if (j == 1 || j == nrc) q2[o] = 0.0;
      if (k == 1) q3[o] = c_dq3;
      if (k >= 2 && j >= 1 && j <= nr) {
        const double v = c_dq3 - a3k * (p0 - Pm[cy][cx]);
        q3[o] = v;
        if (j == nr) q3[o + sj] = slip * v;
      }
      if (k <= nz && j >= 1 && j <= nr) {
        const long pc = (long)(i - 1) + (long)nt * ((long)(j - 1) + (long)nr * (long)(k - 1));
        // theta component: qcap = (dph(i) - dph(i-1)) dx1 / rm
        {
          const double qc0 = (p0 - P0[cy][cx - 1]) * dx1 * irm;
          const double qip = (P0[cy][cx + 1] - p0) * dx1 * irm;
          const double qim = (P0[cy][cx - 1] - P0[cy][cx - 2]) * dx1 * irm;
          const double qkp = (Pp[cy][cx] - Pp[cy][cx - 1]) * dx1 * irm;
          const double qkm = (Pkm[cy][cx] - Pkm[cy][cx - 1]) * dx1 * irm;
          const double qjp = (P0[cy + djp][cx] - P0[cy + djp][cx - 1]) * dx1 * irmp;
          const double qjm = (P0[cy + djm][cx] - P0[cy + djm][cx - 1]) * dx1 * irmm;
          const double d11 = (qip - 2.0 * qc0 + qim) * s1t11;
          const double d33 = zs ? (qkp - qc0) * zm.cp[k] - (qc0 - qkm) * zm.cm[k] : (qkp - 2.0 * qc0 + qkm) * dx3q;
          const double qnn = qjp * irmn - qc0 * irm;
          const double qss = qc0 * irm - qjm * irmm;
          const double d22 = (qnn * s1t2n - qss * s1t2s) * s1t2c;
          prg1[pc] = c_p1 + qc0 - be * (d11 + d33 + d22);
        }
        // r component: qcap = (dph(j) - dph(j-1)) dx2 rc, zero at j = 1 and nrc
        if (j >= 2) {
          const double qc0 = (p0 - P0[cy - 1][cx]) * dx2 * rcj;
          const double qip = (P0[cy][cx + 1] - P0[cy - 1][cx + 1]) * dx2 * rcj;
          const double qim = (P0[cy][cx - 1] - P0[cy - 1][cx - 1]) * dx2 * rcj;
          const double qkp = (Pp[cy][cx] - Pp[cy - 1][cx]) * dx2 * rcj;
          const double qkm = (Pkm[cy][cx] - Pkm[cy - 1][cx]) * dx2 * rcj;
          const double qjp = (j + 1 < nrc) ? (P0[cy + 1][cx] - p0) * dx2 * rcp : 0.0;
          const double qjm = (j - 1 > 1) ? (P0[cy - 1][cx] - P0[cy - 2][cx]) * dx2 * rcm : 0.0;
          const double d11 = (qip - 2.0 * qc0 + qim) * s2t11;
          const double d33 = zs ? (qkp - qc0) * zm.cp[k] - (qc0 - qkm) * zm.cm[k] : (qkp - 2.0 * qc0 + qkm) * dx3q;
          const double d22 = (qjp - qc0) * s2t2p - (qc0 - qjm) * s2t2m;
          prg2[pc] = c_p2 + qc0 - be * (d11 + d22 + d33);
        }
        // z component: qcap = (dph(k) - dph(k-1)) dx3, zero at k = 1
        {
          const double qkp = (Pp[cy][cx] - p0) * zcp;
          if (qkp != qkp) atomicOr(flag, 1);  // "Flow field crash" (divergence.F90:396-399)
          if (k >= 2) {
            const double qc0 = (p0 - Pm[cy][cx]) * zc0;
            const double qip = (P0[cy][cx + 1] - Pm[cy][cx + 1]) * zc0;
            const double qim = (P0[cy][cx - 1] - Pm[cy][cx - 1]) * zc0;
            const double qkm = (k - 1 > 1) ? (Pm[cy][cx] - Pm2[cy][cx]) * zcm : 0.0;
            const double qjp = (P0[cy + djp][cx] - Pm[cy + djp][cx]) * zc0;
            const double qjm = (P0[cy + djm][cx] - Pm[cy + djm][cx]) * zc0;
            const double d11 = (qip - 2.0 * qc0 + qim) * s1t11;
            const double d33 = zs ? (qkp - qc0) * zm.fp[k] - (qc0 - qkm) * zm.fm[k] : (qkp - 2.0 * qc0 + qkm) * dx3q;
            const double d22 = (qjp - qc0) * s3t2n - (qc0 - qjm) * s3t2s;
            prg3[pc] = c_p3 + qc0 - be * (d11 + d22 + d33);
          }
        }
        {
          const double d11 = (P0[cy][cx + 1] + P0[cy][cx - 1] - 2.0 * p0) * s1t11;
          const double d22 = (P0[cy + djp][cx] - p0) * s3t2n - (p0 - P0[cy + djm][cx]) * s3t2s;
          const double d33 = zs ? (Pp[cy][cx] - p0) * zm.cp[k] - (p0 - Pkm[cy][cx]) * zm.cm[k]
                                : (Pp[cy][cx] + Pkm[cy][cx] - 2.0 * p0) * dx3q;
          pres[o] = c_pres + p0 - be * (d11 + d22 + d33);
        }
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Vortex-core statistics of analysis (run.F90:238-388) as a device
// reduction: per cell the azimuthal vorticity (dvdz - dwdr) and the swirl
// q1m of the cell centre; inside the window 0.1 < z/L < 0.9,
// 0.05 < r/R < 0.8, j != 1: their maxima and the sums of z w^2 r, w^2 r,
// w^2 r, w^2 (w = vorticity where > 0, then swirl where > 0) times dx1 dx2
// dx3.  Per-block partials [block][10], summed in a fixed order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_analysis(VGrid g, double rmax, double zmax, const double *__restrict__ rmv,
                                                  const double *__restrict__ zpos,
                                                  const double *__restrict__ q1, const double *__restrict__ q2,
                                                  const double *__restrict__ q3, double *__restrict__ part) {
  const long ncell = (long)g.nt * g.nr * g.nz;
  const double dz = zmax / g.nz, w = g.dx1 * g.dx2 * g.dx3;
  double v[10] = {-10000.0, -10000.0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (long c = (long)blockIdx.x * 256 + threadIdx.x; c < ncell; c += (long)gridDim.x * 256) {
    const int ith = (int)(c % g.nt) + 1;
    const long rest = c / g.nt;
    const int ir = (int)(rest % g.nr) + 1, iz = (int)(rest / g.nr) + 1;
    // msh%rm, msh%zm: the radial and axial meshes (uniform or set)
    const double rm = rmv[ir], zm = zpos ? zpos[iz] : (iz - 0.5) * dz;
    if (!(zm < 0.90 * zmax && zm > 0.10 * zmax && rm < 0.80 * rmax && rm > rmax * 0.05 && ir != 1)) continue;
    const int ithp1 = ith == g.nt ? 1 : ith + 1;
    const double q1m = g.nt > 1 ? 0.5 * (q1[hidx(g, ith, ir, iz)] + q1[hidx(g, ithp1, ir, iz)]) : q1[hidx(g, ith, ir, iz)];
    const double rp = rmv[ir + 1], rmm = rmv[ir - 1];
    const double dwdr = 0.5 * ((q3[hidx(g, ith, ir + 1, iz)] + q3[hidx(g, ith, ir + 1, iz + 1)]) -
                               (q3[hidx(g, ith, ir - 1, iz)] + q3[hidx(g, ith, ir - 1, iz + 1)])) / (rp - rmm);
    const double zp = zpos ? zpos[iz + 1] : (iz + 0.5) * dz, zmm = zpos ? zpos[iz - 1] : (iz - 1.5) * dz;
    const double dvdz = 0.5 * ((q2[hidx(g, ith, ir, iz + 1)] / rm + q2[hidx(g, ith, ir + 1, iz + 1)] / rp) -
                               (q2[hidx(g, ith, ir + 1, iz - 1)] / rp + q2[hidx(g, ith, ir, iz - 1)] / rm)) / (zp - zmm);
    const double vort = +(dvdz - dwdr), swrl = q1m;
    v[0] = fmax(vort, v[0]);
    v[1] = fmax(swrl, v[1]);
    if (vort > 0.0) {
      v[2] += zm * vort * vort * rm * w;
      v[3] += vort * vort * rm * w;
      v[4] += vort * vort * rm * w;
      v[5] += vort * vort * w;
    }
    if (swrl > 0.0) {
      v[6] += zm * swrl * swrl * rm * w;
      v[7] += swrl * swrl * rm * w;
      v[8] += swrl * swrl * rm * w;
      v[9] += swrl * swrl * w;
    }
  }
  __shared__ double red[10][4];
  for (int q = 0; q < 10; q++) {
    double x = v[q];
    for (int o = 32; o > 0; o >>= 1) {
      const double y = __shfl_xor(x, o);
      x = (q < 2) ? fmax(x, y) : x + y;
    }
    if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  if (threadIdx.x < 10) {
    const int q = threadIdx.x;
    const double a = red[q][0], b = red[q][1], cc = red[q][2], d = red[q][3];
    part[(long)blockIdx.x * 10 + q] = (q < 2) ? fmax(fmax(a, b), fmax(cc, d)) : ((a + b) + (cc + d));
  }
}
__global__ void k_analysis_sum(const double *__restrict__ part, int nblk, double *__restrict__ out) {
  const int q = threadIdx.x;
  if (q >= 10) return;
  double x = (q < 2) ? -10000.0 : 0.0;
  for (int b = 0; b < nblk; b++) x = (q < 2) ? fmax(x, part[(long)b * 10 + q]) : x + part[(long)b * 10 + q];
  out[q] = x;
}

}  // namespace

// ===========================================================================
// launchers
// ===========================================================================
#define CHECK_LAUNCH() (void)hipGetLastError()

static void hdnbc_on(vrh_t *h, int irk, double cor, const double *cor_dev, hipStream_t st) {
  const VGrid &g = h->g;
  dim3 blk(64), grd((g.nt + 63) / 64, g.nrc);
  KTimer kt(h, KT_HDNBC);
  const double zm3n = h->zmesh[2].empty() ? 0.0 : g.dx3 / h->zmesh[2][g.nz];
  hipLaunchKernelGGL(k_hdnbc, grd, blk, 0, st, g, h->coef, zm3n, h->rk[irk].g, h->rk[irk].r, cor, cor_dev, h->f.q3,
                     h->f.csort, h->f.dq3no, h->f.corr3n);
}
template <typename TW> void launch_hdnbc(vrh_t *h, int irk, double cor, const double *cor_dev) {
  hdnbc_on(h, irk, cor, cor_dev, h->stream);
}

void launch_scalars(vrh_t *h) {
  hipLaunchKernelGGL(k_scalars, dim3(1), dim3(64), 0, h->stream, h->d_sc, h->dt, h->rk[1].a, h->rk[2].a,
                     h->rk[3].a, h->inlet_area, h->cor_red, 1, 0.0);
}
// the same scalars for the step that has reached time t (no advance)
void launch_scalars_at(vrh_t *h, double t) {
  hipLaunchKernelGGL(k_scalars, dim3(1), dim3(64), 0, h->stream, h->d_sc, h->dt, h->rk[1].a, h->rk[2].a,
                     h->rk[3].a, h->inlet_area, h->cor_red, 0, t);
}

template <typename TW> void launch_explicit(vrh_t *h, int irk) {
  // compute type of the stencils: the work type (FP32 in the mixed path)
  using TC = TW;
  const VGrid &g = h->g;
  int ntp = 1;
  while (ntp < g.nt) ntp <<= 1;
  if (ntp < 64) ntp = 64;
  const double betadx = 0.5 * h->rk[irk].a * h->dt / h->Re;
  const int nlev = h->theta_levels[irk];
  const int var = h->explicit_variant >= 0 ? h->explicit_variant : (sizeof(TW) == 8 ? 0 : 1);
  KTimer kt(h, KT_EXPLICIT);
  if (var == 0) {
    // one z plane per block
    int rj = 256 / ntp;
    if (rj < 1) rj = 1;
    dim3 blk(ntp, rj), grd((g.nr + rj - 1) / rj, g.nz);
    const size_t sh = sizeof(TC) * 6 * (size_t)rj * ntp;
    auto kern = h->zmet.m3 ? k_explicit_cell<TW, TC, true> : k_explicit_cell<TW, TC, false>;
    hipLaunchKernelGGL(kern, grd, blk, sh, h->stream, g, h->coef, h->rk[irk], h->zmet, betadx, nlev,
                       h->f.q1, h->f.q2, h->f.q3, h->f.prg1, h->f.prg2,
                       h->f.prg3, (TW *)h->f.ru1, (TW *)h->f.ru2, (TW *)h->f.ru3, (TW *)h->f.r1,
                       (TW *)h->f.r2, (TW *)h->f.r3);
    return;
  }
  // z marching: rows per block for 512 threads within 64 KB of LDS
  auto lds = [&](int rj) {
    return sizeof(double) * (size_t)rj * (kMaxPcrLevels * 2 + 2) +
           sizeof(TC) * ((size_t)rj * kRowCoefs + (size_t)9 * (rj + 2) * ntp + (size_t)6 * rj * ntp);
  };
  int rj = 512 / ntp;
  if (rj < 1) rj = 1;
  const int kchunk = 16;
  // the wave form's PCR ratio table holds this substep's levels only
  auto lds_w = [&](int rw) { return lds(rw) - sizeof(double) * (size_t)rw * 2 * (kMaxPcrLevels - nlev); };
  if (var == 2 || var == 3) {
    // one wavefront per r row.  Form 2: the most rows (<= 8) whose LDS fits
    // 64 KB.  Form 3: among the instantiated row counts, the one that keeps
    // the most waves resident on a CU within the device's LDS per block and
    // per CU (ties: more rows, less r-halo re-read).  Shapes without an
    // instantiation fall back to the marching form.  WPE: waves per SIMD the
    // LDS allows (register budget).
    const int cpl = ntp / 64;
    int rw = 8;
    size_t lim = 65536;
    if (var == 2) {
      while (rw > 1 && lds(rw) > lim) rw--;
    } else {
      lim = (size_t)h->lds_block;
      const int cand1[] = {8, 6}, cand2[] = {6, 4, 2}, cand4[] = {2};
      const int *cd = cpl == 1 ? cand1 : cpl == 2 ? cand2 : cand4;
      const int ncd = cpl == 1 ? 2 : cpl == 2 ? 3 : 1;
      int best = 0;
      rw = 0;
      for (int i = 0; i < ncd; i++) {
        const size_t L = lds_w(cd[i]);
        if (L > lim) continue;
        const int waves = cd[i] * (int)((size_t)h->lds_cu / L);
        if (waves > best) { best = waves; rw = cd[i]; }
      }
      if (rw == 0) rw = 1;
    }
    dim3 grd((g.nr + rw - 1) / rw, (g.nz + kchunk - 1) / kchunk);
    bool done = rw > 1 && (var == 2 ? lds(rw) : lds_w(rw)) <= lim;
#define EXPL_WAVE(C, R, W)                                                                                          \
  hipLaunchKernelGGL((h->zmet.m3 ? k_explicit_wave<TW, TC, C, R, W, true> : k_explicit_wave<TW, TC, C, R, W, false>), \
                     grd, dim3(64, R), lds_w(R), h->stream, g, h->coef, h->rk[irk],                                     \
                     h->zmet, betadx, nlev, kchunk, h->f.q1, h->f.q2, h->f.q3, h->f.prg1, h->f.prg2, \
                     h->f.prg3, (TW *)h->f.ru1, (TW *)h->f.ru2, (TW *)h->f.ru3, (TW *)h->f.r1,          \
                     (TW *)h->f.r2, (TW *)h->f.r3)
    if (done) {
      if (cpl == 1 && rw == 8) EXPL_WAVE(1, 8, 4);
      else if (cpl == 1 && rw == 6) EXPL_WAVE(1, 6, 3);
      else if (cpl == 2 && rw == 6) EXPL_WAVE(2, 6, 3);
      else if (cpl == 2 && rw == 4) EXPL_WAVE(2, 4, 2);
      else if (cpl == 2 && rw == 2) EXPL_WAVE(2, 2, 2);
      else if (cpl == 4 && rw == 2) EXPL_WAVE(4, 2, 2);
      else done = false;
    }
#undef EXPL_WAVE
    if (done) return;
  }
  while (rj > 1 && lds(rj) > 65536) rj--;
  dim3 blk(ntp, rj), grd((g.nr + rj - 1) / rj, (g.nz + kchunk - 1) / kchunk);
  auto kern = h->zmet.m3 ? k_explicit_march<TW, TC, true> : k_explicit_march<TW, TC, false>;
  hipLaunchKernelGGL(kern, grd, blk, lds(rj), h->stream, g, h->coef, h->rk[irk], h->zmet, betadx, nlev,
                     kchunk, h->f.q1, h->f.q2, h->f.q3, h->f.prg1, h->f.prg2,
                     h->f.prg3, (TW *)h->f.ru1, (TW *)h->f.ru2, (TW *)h->f.ru3, (TW *)h->f.r1,
                     (TW *)h->f.r2, (TW *)h->f.r3);
}

// Blocks of the explicit kernel's one-wave-per-row form and tuned z-DCT
// tiles may take more than 64 KB of LDS where the device allows it (gfx950:
// 160 KB per CU).  Those instantiations opt in once per handle; if the
// runtime refuses, they keep the 64 KB limit.
void lds_optin(vrh_t *h) {
  int dev = 0, lb = 0, lc = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&lb, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
      hipDeviceGetAttribute(&lc, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess)
    return;
  if (lc < lb) lc = lb;
  bool ok = true;
  if (lb > 65536) {
    const void *fns[] = {
        reinterpret_cast<const void *>(&k_explicit_wave<double, double, 1, 8, 4, false>),
        reinterpret_cast<const void *>(&k_explicit_wave<double, double, 1, 6, 3, false>),
        reinterpret_cast<const void *>(&k_explicit_wave<double, double, 2, 6, 3, false>),
        reinterpret_cast<const void *>(&k_explicit_wave<double, double, 2, 4, 2, false>),
        reinterpret_cast<const void *>(&k_explicit_wave<double, double, 2, 2, 2, false>),
        reinterpret_cast<const void *>(&k_explicit_wave<double, double, 4, 2, 2, false>),
        reinterpret_cast<const void *>(&k_explicit_wave<float, float, 1, 8, 4, false>),
        reinterpret_cast<const void *>(&k_explicit_wave<float, float, 1, 6, 3, false>),
        reinterpret_cast<const void *>(&k_explicit_wave<float, float, 2, 6, 3, false>),
        reinterpret_cast<const void *>(&k_explicit_wave<float, float, 2, 4, 2, false>),
        reinterpret_cast<const void *>(&k_explicit_wave<float, float, 2, 2, 2, false>),
        reinterpret_cast<const void *>(&k_explicit_wave<float, float, 4, 2, 2, false>),
        reinterpret_cast<const void *>(&k_explicit_wave<double, double, 1, 8, 4, true>),
        reinterpret_cast<const void *>(&k_explicit_wave<double, double, 1, 6, 3, true>),
        reinterpret_cast<const void *>(&k_explicit_wave<double, double, 2, 6, 3, true>),
        reinterpret_cast<const void *>(&k_explicit_wave<double, double, 2, 4, 2, true>),
        reinterpret_cast<const void *>(&k_explicit_wave<double, double, 2, 2, 2, true>),
        reinterpret_cast<const void *>(&k_explicit_wave<double, double, 4, 2, 2, true>),
        reinterpret_cast<const void *>(&k_explicit_wave<float, float, 1, 8, 4, true>),
        reinterpret_cast<const void *>(&k_explicit_wave<float, float, 1, 6, 3, true>),
        reinterpret_cast<const void *>(&k_explicit_wave<float, float, 2, 6, 3, true>),
        reinterpret_cast<const void *>(&k_explicit_wave<float, float, 2, 4, 2, true>),
        reinterpret_cast<const void *>(&k_explicit_wave<float, float, 2, 2, 2, true>),
        reinterpret_cast<const void *>(&k_explicit_wave<float, float, 4, 2, 2, true>),
        reinterpret_cast<const void *>(&k_dctz<float>),
        reinterpret_cast<const void *>(&k_dctz<double>)};
    for (const void *f : fns)
      ok = ok && hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lb) == hipSuccess;
  }
  (void)hipGetLastError();
  if (ok) { h->lds_block = lb; h->lds_cu = lc; }
}

template <typename TW, int MODE>
static void line_launch(vrh_t *h, const LineArgs &a, int M, hipStream_t st) {
  const int P = a.lc.P;
  dim3 blk(64 * P), grd((unsigned)((a.nlines + 63) / 64));
  KTimer kt(h, MODE <= SW_R2 ? KT_LINE_R : KT_LINE_Z);
  switch (M) {
    case 16: hipLaunchKernelGGL((k_line<TW, 16, MODE>), grd, blk, 0, st, a); break;
    case 24: hipLaunchKernelGGL((k_line<TW, 24, MODE>), grd, blk, 0, st, a); break;
    case 32: hipLaunchKernelGGL((k_line<TW, 32, MODE>), grd, blk, 0, st, a); break;
    case 40: hipLaunchKernelGGL((k_line<TW, 40, MODE>), grd, blk, 0, st, a); break;
    case 48: hipLaunchKernelGGL((k_line<TW, 48, MODE>), grd, blk, 0, st, a); break;
    case 72: hipLaunchKernelGGL((k_line<TW, 72, MODE>), grd, blk, 0, st, a); break;
    default: hipLaunchKernelGGL((k_line<TW, 64, MODE>), grd, blk, 0, st, a); break;
  }
}

template <typename TW> void launch_sweeps(vrh_t *h, int irk, double c3s, const double *c3s_dev, bool hdnbc,
                                          double cor, const double *cor_dev) {
  const VGrid &g = h->g;
  LineArgs a{};
  a.g = g;
  a.slip = h->slip;
  a.jlim = h->jlim;
  a.c3s = c3s;
  a.c3s_dev = c3s_dev;
  a.corr3n = h->f.corr3n;
  // The three components' chains are independent: q3 and q2 run on the
  // side streams, forked from and joined back into the handle's stream
  // (captured steps record the fork as graph branches).  With the kernel
  // timer on, and inside an ensemble graph (whose members already run side
  // by side), all run in order on the handle's stream.
  const bool par = !h->kt_on && !h->serial_sweeps;
  hipStream_t s1 = par ? h->side[0] : h->stream, s2 = par ? h->side[1] : h->stream;
  const int ie = (irk >= 1 && irk <= 3) ? irk : 0;
  if (par) {
    (void)hipEventRecord(h->fork_ev[ie], h->stream);
    (void)hipStreamWaitEvent(s1, h->fork_ev[ie], 0);
    (void)hipStreamWaitEvent(s2, h->fork_ev[ie], 0);
  }
  // q1: r then z (step.F90:75-78)
  a.lc = h->lc[irk][SW_R1]; a.nlines = (long)g.nt * g.nz; a.r = h->f.r1;
  line_launch<TW, SW_R1>(h, a, h->Msw[SW_R1], h->stream);
  a.lc = h->lc[irk][SW_Z1]; a.nlines = (long)g.nt * g.nr; a.j0 = 1; a.q = h->f.q1; a.dq = h->f.dq1;
  line_launch<TW, SW_Z1>(h, a, h->Msw[SW_Z1], h->stream);
  // q3: r then z (step.F90:124-127).  In a whole step the outlet rows of
  // hdnbc (csort, dq3no, corr3n: read only by the q3 z sweep; computed from
  // q3, which the explicit kernel leaves alone) head this chain, off the
  // critical path.
  if (hdnbc) hdnbc_on(h, irk, cor, cor_dev, s1);
  a.lc = h->lc[irk][SW_R3]; a.nlines = (long)g.nt * g.nz; a.r = h->f.r3;
  line_launch<TW, SW_R3>(h, a, h->Msw[SW_R3], s1);
  a.lc = h->lc[irk][SW_Z3]; a.nlines = (long)g.nt * g.nr; a.j0 = 1; a.q = h->f.q3; a.dq = h->f.dq3;
  line_launch<TW, SW_Z3>(h, a, h->Msw[SW_Z3], s1);
  // q2: z then r (step.F90:171-177)
  a.lc = h->lc[irk][SW_Z2]; a.nlines = (long)g.nt * (g.nr - 1); a.j0 = 2; a.r = h->f.r2;
  line_launch<TW, SW_Z2>(h, a, h->Msw[SW_Z2], s2);
  a.lc = h->lc[irk][SW_R2]; a.nlines = (long)g.nt * g.nzc; a.q = h->f.q2; a.dq = h->f.dq2;
  line_launch<TW, SW_R2>(h, a, h->Msw[SW_R2], s2);
  if (par) {
    (void)hipEventRecord(h->join_ev[ie][0], s1);
    (void)hipEventRecord(h->join_ev[ie][1], s2);
    (void)hipStreamWaitEvent(h->stream, h->join_ev[ie][0], 0);
    (void)hipStreamWaitEvent(h->stream, h->join_ev[ie][1], 0);
  }
}

static int ilog2(int n) { int l = 0; while ((1 << l) < n) ++l; return l; }

template <typename T>
static void radial_launch(vrh_t *h, T *S, const T *ib, long nm, hipStream_t st) {
  const VGrid &g = h->g;
  // rows per wave: 16 keeps the segment's rhs and pivots in registers (a
  // 1024-lane block allows 128 VGPRs) up to nr = 256; 32 beyond (P <= 16)
  const int M = g.nr <= 256 ? 16 : 32;
  const int P = (g.nr + M - 1) / M;
  const long nlines = nm * g.nz;
  dim3 blk(64 * P), grd((unsigned)((nlines + 63) / 64));
  // the radial base coefficients in the solve's type (float copies rounded
  // as the device conversion)
  const T *ra, *rc;
  if constexpr (sizeof(T) == 4) { ra = h->d_aabf; rc = h->d_ccbf; }
  else { ra = h->d_aab; rc = h->d_ccb; }
  KTimer kt(h, KT_RADIAL);
  if (M == 16)
    hipLaunchKernelGGL((k_radial<T, 16>), grd, blk, 0, st, S, ib, nm, nlines, g.nr, ra, rc);
  else
    hipLaunchKernelGGL((k_radial<T, 32>), grd, blk, 0, st, S, ib, nm, nlines, g.nr, ra, rc);
}

// z transform of the pressure solve on an axial mesh: every spectral line
// (the nl lines of one z position are contiguous, so the array is a
// column-major nl x nz matrix) times a dense nz x nz matrix, out of place.
// Forward: out(:, q) = sum_k in(:, k) Vinv(q, k); inverse: out(:, k) =
// sum_q in(:, q) V(k, q).  The row-major matrices read column-major are
// their transposes, the right operand of both products.
// A failed call is reported through the step's error (mode_rc).
static void zgemm_status(vrh_t *h, rocblas_status st) {
  if (st != rocblas_status_success && !h->mode_rc) {
    h->mode_rc = -3;
    h->err = "rocBLAS z transform failed (status " + std::to_string((int)st) + ")";
  }
}
static void zgemm(vrh_t *h, hipStream_t st, const float *in, float *out, long nl, int dir) {
  const int nz = h->g.nz;
  const float one = 1.f, zero = 0.f;
  rocblas_handle b = (rocblas_handle)h->blas;
  zgemm_status(h, rocblas_set_stream(b, st));
  zgemm_status(h, rocblas_sgemm(b, rocblas_operation_none, rocblas_operation_none, (rocblas_int)nl, nz, nz, &one,
                                in, (rocblas_int)nl, h->d_zvf + (dir > 0 ? 0 : (size_t)nz * nz), nz, &zero, out,
                                (rocblas_int)nl));
  zgemm_status(h, rocblas_set_stream(b, h->stream));
}
static void zgemm(vrh_t *h, hipStream_t st, const double *in, double *out, long nl, int dir) {
  const int nz = h->g.nz;
  const double one = 1.0, zero = 0.0;
  rocblas_handle b = (rocblas_handle)h->blas;
  zgemm_status(h, rocblas_set_stream(b, st));
  zgemm_status(h, rocblas_dgemm(b, rocblas_operation_none, rocblas_operation_none, (rocblas_int)nl, nz, nz, &one,
                                in, (rocblas_int)nl, h->d_zv + (dir > 0 ? 0 : (size_t)nz * nz), nz, &zero, out,
                                (rocblas_int)nl));
  zgemm_status(h, rocblas_set_stream(b, h->stream));
}

// Spectral array (position fastest, then r, then z) <-> the Helmholtz
// handle's slots (slot, z, r), for positions pos0..N-1: 64 x 64 tiles of
// (position, r) through LDS, one z plane per blockIdx.z.
template <typename TW>
__global__ __launch_bounds__(256) void k_hz_gather(const TW *__restrict__ S, double *__restrict__ f, int N, int nr,
                                                   int nz, int pos0) {
  __shared__ double t[64][65];
  const int p0 = pos0 + blockIdx.x * 64, j0 = blockIdx.y * 64, k = blockIdx.z;
  const int np = N - pos0;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int pp = e & 63, jj = e >> 6;
    const int pos = p0 + pp, j = j0 + jj;
    t[jj][pp] = (pos < N && j < nr) ? (double)S[(size_t)pos + (size_t)N * ((size_t)j + (size_t)nr * k)] : 0.0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int jj = e & 63, pp = e >> 6;
    const int b = p0 + pp - pos0, j = j0 + jj;
    if (b < np && j < nr) f[((size_t)b * nz + k) * nr + j] = t[jj][pp];
  }
}
template <typename TW>
__global__ __launch_bounds__(256) void k_hz_scatter(const double *__restrict__ x, TW *__restrict__ S, int N, int nr,
                                                    int nz, int pos0) {
  __shared__ double t[64][65];
  const int p0 = pos0 + blockIdx.x * 64, j0 = blockIdx.y * 64, k = blockIdx.z;
  const int np = N - pos0;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int jj = e & 63, pp = e >> 6;
    const int b = p0 + pp - pos0, j = j0 + jj;
    t[pp][jj] = (b < np && j < nr) ? x[((size_t)b * nz + k) * nr + j] : 0.0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int pp = e & 63, jj = e >> 6;
    const int pos = p0 + pp, j = j0 + jj;
    if (pos < N && j < nr) S[(size_t)pos + (size_t)N * ((size_t)j + (size_t)nr * k)] = (TW)t[pp][jj];
  }
}

template <typename TW> void launch_poisson(vrh_t *h, int irk, const double *ext_rhs, double *phi_out) {
  const VGrid &g = h->g;
  const int N = g.nt, logn = ilog2(N);
  const bool pow2 = N >= 2 && (N & (N - 1)) == 0;
  int LB = pow2 ? (h->fft_lines > 0 ? h->fft_lines : 1024 / N) : 2;
  if (LB > 16) LB = 16;
  if (pow2) {  // butterflies per thread of the LDS FFT: (LB/2) N / rmin <= qt * 256
    while (LB > 2 && (LB / 2) * N / fft_min_radix(logn) > 256 * fft_qt<TW>()) LB /= 2;
  }
  if (LB < 2) LB = 2;
  const long nlines = (long)g.nr * g.nz;
  TW *S = (TW *)h->f.S;
  double *S64 = (h->prec == VRH_MIXED) ? h->f.S64 : nullptr;
  const double adt = h->rk[irk].a * h->dt;
  const bool keep_dv = !pow2;   // the kernel's FP64 copy of the lines (direct DFT)
  size_t sh1 = (keep_dv ? al16(sizeof(double) * LB * N) : 0) + al16(sizeof(TW) * LB * N);
  if (S64) sh1 = al16(sh1 + sizeof(double) * (2 * (size_t)N + 10 * (size_t)LB));
  sh1 = al16(al16(sh1 + sizeof(TW) * N) + sizeof(int) * 2 * LB) + sizeof(double) * 2 * LB;
  { KTimer kt(h, KT_DIVFFT);
  hipLaunchKernelGGL(k_divfft<TW>, dim3((unsigned)((nlines + LB - 1) / LB)), dim3(256), sh1, h->stream, g, h->coef,
                     h->zmet, 1.0 / adt, h->f.dq1, h->f.dq2, h->f.dq3, ext_rhs, S, S64, h->d_twt, h->d_twt + N, LB, logn); }
  // z transforms
  const int nz = g.nz, lz = ilog2(nz);
  // lines per block: a power of two, data + twiddles within 64 KB of LDS
  // (a tuned width may take the device's larger per-block LDS: 32 FP32
  // lines read whole 128-byte rows of the strided layout), one FFT
  // butterfly per thread with up to 1024 threads
  auto lines = [&](size_t es, int want) {
    const size_t lim = want > 0 ? (size_t)h->lds_block : 65536;
    const int tmax = want > 0 ? 32 : 16;
    int tl = 2;
    while (tl * 2 <= tmax && (size_t)(tl * 2) * (nz + 1) * es + (size_t)(es == 4 ? 3 : 1) * nz * es <= lim &&
           (tl * 2 / 2) * nz / fft_min_radix(lz) <= 1024)
      tl *= 2;
    if (want > 0) while (tl > 2 && tl > want) tl /= 2;
    return tl;
  };
  auto threads = [&](int tl) {
    const int b = (tl / 2) * nz / fft_min_radix(lz);
    return b > 512 ? 1024 : b > 256 ? 512 : 256;
  };
  const int TL = lines(sizeof(TW), h->dct_lines), TL64 = lines(sizeof(double), 0);
  const int ltl = ilog2(TL), ltl64 = ilog2(TL64);
  const int NTD = threads(TL), NTD64 = threads(TL64);
  const long nl = (long)N * g.nr;
  // tile lines, FFT twiddles, and in FP32 the Makhoul twiddles
  const size_t sh2 = sizeof(TW) * ((size_t)TL * (nz + 1) + (size_t)(sizeof(TW) == 4 ? 3 : 1) * nz),
               sh64 = sizeof(double) * ((size_t)TL64 * (nz + 1) + nz);
  // The FP64 low-mode chain (mixed precision) reads only what k_divfft
  // wrote to S64 and feeds only k_ifft: it runs on a side stream, forked
  // after k_divfft and joined before k_ifft, beside the FP32 transforms.
  // With the kernel timer on, or inside an ensemble graph, it stays in
  // order on the handle's stream.
  const bool par = S64 && !h->kt_on && !h->serial_sweeps;
  hipStream_t s64 = par ? h->side[0] : h->stream;
  const int ie = (irk >= 1 && irk <= 3) ? irk : 0;
  if (par) {
    (void)hipEventRecord(h->poi_ev[ie][0], h->stream);
    (void)hipStreamWaitEvent(s64, h->poi_ev[ie][0], 0);
  }
  if (S64 && h->zmet.m3) {
    const long nl64 = 5L * g.nr;
    zgemm(h, s64, S64, h->S2_64, nl64, 1);
    radial_launch<double>(h, h->S2_64, h->f.ib64, 5, s64);
    zgemm(h, s64, h->S2_64, S64, nl64, -1);
  } else if (S64) {
    const long nl64 = 5L * g.nr;
    hipLaunchKernelGGL(k_dctz<double>, dim3((unsigned)((nl64 + TL64 - 1) / TL64)), dim3(NTD64), sh64, s64, S64,
                       nl64, nz, lz, TL64, ltl64, 1, h->d_twz, h->d_twz + nz, h->d_ez, h->d_ez + nz);
    radial_launch<double>(h, S64, h->f.ib64, 5, s64);
    hipLaunchKernelGGL(k_dctz<double>, dim3((unsigned)((nl64 + TL64 - 1) / TL64)), dim3(NTD64), sh64, s64, S64,
                       nl64, nz, lz, TL64, ltl64, -1, h->d_twz, h->d_twz + nz, h->d_ez, h->d_ez + nz);
  }
  // per-mode iterative solve of the positions >= hz_pos0: stage their
  // right-hand sides before the direct path overwrites them
  const dim3 hzg((unsigned)((N - h->hz_pos0 + 63) / 64), (unsigned)((g.nr + 63) / 64), (unsigned)nz);
  if (h->hz)
    hipLaunchKernelGGL(k_hz_gather<TW>, hzg, dim3(256), 0, h->stream, S, hz_step_input(h->hz), N, g.nr, nz,
                       h->hz_pos0);
  if (h->zmet.m3) {
    TW *S2 = (TW *)h->S2;
    { KTimer kt(h, KT_DCT_FWD); zgemm(h, h->stream, S, S2, nl, 1); }
    radial_launch<TW>(h, S2, (const TW *)h->f.ib, N, h->stream);
    { KTimer kt(h, KT_DCT_INV); zgemm(h, h->stream, S2, S, nl, -1); }
  } else {
  { KTimer kt(h, KT_DCT_FWD);
  hipLaunchKernelGGL(k_dctz<TW>, dim3((unsigned)((nl + TL - 1) / TL)), dim3(NTD), sh2, h->stream, S, nl, nz, lz, TL,
                     ltl, 1, h->d_twz, h->d_twz + nz, h->d_ez, h->d_ez + nz); }
  radial_launch<TW>(h, S, (const TW *)h->f.ib, N, h->stream);
  { KTimer kt(h, KT_DCT_INV);
  hipLaunchKernelGGL(k_dctz<TW>, dim3((unsigned)((nl + TL - 1) / TL)), dim3(NTD), sh2, h->stream, S, nl, nz, lz, TL,
                     ltl, -1, h->d_twz, h->d_twz + nz, h->d_ez, h->d_ez + nz); }
  }
  if (par) {
    (void)hipEventRecord(h->poi_ev[ie][1], s64);
    (void)hipStreamWaitEvent(h->stream, h->poi_ev[ie][1], 0);
  }
  if (h->hz) {
    if (!h->mode_rc) {
      const int rc = hz_step_solve(h->hz, h->stream, h->hz_method, h->hz_tol, h->hz_maxit, &h->hz_iters);
      if (rc) { h->mode_rc = rc; h->err = hz_error(h->hz); }
    }
    h->hz_solves++;
    hipLaunchKernelGGL(k_hz_scatter<TW>, hzg, dim3(256), 0, h->stream, hz_step_output(h->hz), S, N, g.nr, nz,
                       h->hz_pos0);
  }
  size_t sh3 = al16(sizeof(TW) * ((size_t)2 * LB * N + N)) + sizeof(int) * 2 * LB;
  if (S64) sh3 = al16(sh3) + sizeof(double) * (5 * (size_t)LB + 2 * (size_t)N);
  KTimer kti(h, KT_IFFT);
  hipLaunchKernelGGL(k_ifft<TW>, dim3((unsigned)((nlines + LB - 1) / LB)), dim3(256), sh3, h->stream, g, S, S64,
                     phi_out, ext_rhs ? nullptr : h->f.dph, h->d_twt, h->d_twt + N, LB, logn);
}

template <typename TW> void launch_project(vrh_t *h, int irk) {
  const VGrid &g = h->g;
  const double adt = h->rk[irk].a * h->dt;
  const double be = 0.5 * (h->rk[irk].a * h->dt) / h->Re;
  constexpr int R = 4;
  const int kchunk = 32;
  dim3 blk(64, R), grd((g.nt + 63) / 64, (g.nrc + 1 + R - 1) / R, (g.nzc + kchunk - 1) / kchunk);
  KTimer kt(h, KT_PROJECT);
  auto kern = h->zmet.m3 ? k_project<TW, R, true> : k_project<TW, R, false>;
  hipLaunchKernelGGL(kern, grd, blk, 0, h->stream, g, h->coef, h->zmet, adt, be, kchunk, h->f.dph, h->f.dq1,
                     h->f.dq2, h->f.dq3, h->f.q1, h->f.q2, h->f.q3, h->f.prg1, h->f.prg2,
                     h->f.prg3, h->f.pres, h->f.flag);
}

void launch_analysis(vrh_t *h, double *part, int nblk, double *out) {
  hipLaunchKernelGGL(k_analysis, dim3(nblk), dim3(256), 0, h->stream, h->g, h->rmax, h->zmax, h->coef.rm, h->zmet.zpos, h->f.q1,
                     h->f.q2, h->f.q3, part);
  hipLaunchKernelGGL(k_analysis_sum, dim3(1), dim3(64), 0, h->stream, part, nblk, out);
}

// Block partials of the allocation-free analysis (vrh_analysis_noalloc):
// a file-scope device array, one copy per GPU, held until the process
// ends.  Host code never passes the array's name to a kernel (on the host
// it is only a stand-in for the symbol): each device's address is taken
// once, with that device current, by analysis_scratch_init, and used from
// then on.  Every block of k_analysis writes all ten of its partials, so
// the array needs no clearing between calls.
static __device__ __align__(256) double scratch[(size_t)kAnalysisBlocks * 10];
namespace {
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int analysis_scratch_init(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return -1;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev]) return 0;
  void *p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return -1;
  g_scratch_ptr[dev] = static_cast<double *>(p);
  return 0;
}

double *analysis_scratch(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[dev];
}

#define INST(TW)                                                   \
  template void launch_hdnbc<TW>(vrh_t *, int, double, const double *); \
  template void launch_explicit<TW>(vrh_t *, int);                 \
  template void launch_sweeps<TW>(vrh_t *, int, double, const double *, bool, double, const double *); \
  template void launch_poisson<TW>(vrh_t *, int, const double *, double *); \
  template void launch_project<TW>(vrh_t *, int);
INST(double)
INST(float)
