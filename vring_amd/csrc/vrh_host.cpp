// vrh_host.cpp -- host-side setup math of libvrh: mesh vectors, tridiagonal
// coefficients, Poisson radial pivots, inlet profile, and the initial-condition
// and time-step helpers of the C ABI.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "vrh_internal.h"

// Radial metric vectors (momentum_coef.F90:24-69) and inlet/outlet weights
// (momentum_terms_swirl.F90:247-251) on the radial mesh (rm, rc, dmr, dcr):
// the uniform one of mesh_tools.F90:325-346 (dmr = dcr = 1), or the one given
// to vrh_set_radial_mesh, entering through the reference's metric factors.
void host_mesh_vectors(vrh_t *h) {
  const int nr = h->nr;
  const VGrid &g = h->g;
  for (auto &v : h->hv) v.assign(nr + 3, 0.0);
  std::vector<double> &rm = h->hv[0], &rc = h->hv[1], &dmr = h->hv[21], &dcr = h->hv[22];
  for (int j = 0; j <= nr + 1; j++) {
    if (h->mesh[0].empty()) {
      rm[j] = 0.0 + (j - 0.5) * (h->rmax / nr);
      rc[j] = 0.0 + (j - 1.0) * (h->rmax / nr);
      dmr[j] = 1.0;
      dcr[j] = 1.0;
    } else {
      rm[j] = h->mesh[0][j]; rc[j] = h->mesh[1][j]; dmr[j] = h->mesh[2][j]; dcr[j] = h->mesh[3][j];
    }
  }
  for (int j = 1; j <= nr; j++) {
    h->hv[2][j] = 0.25 * g.dx1 / (rm[j] * g.axsym);
    h->hv[3][j] = 0.25 * g.dx2 / (rm[j] * dmr[j]);
    h->hv[4][j] = 0.125 / (rm[j] * rm[j]);
    h->hv[5][j] = g.dx1 / (h->Re * (rm[j] * rm[j] * rm[j]));
    h->hv[10][j] = g.dx1q / ((rm[j] * rm[j]) * g.axsym);
    h->hv[11][j] = (rc[j + 1] * rc[j + 1] * rc[j + 1]) / dcr[j + 1];
    h->hv[12][j] = (rc[j] * rc[j] * rc[j]) / dcr[j];
    h->hv[13][j] = g.dx2q / (dmr[j] * (rm[j] * rm[j]));
    h->hv[18][j] = g.dx2q * rc[j + 1] / (rm[j] * dcr[j + 1] * dmr[j]);
    h->hv[19][j] = g.dx2q * rc[j] / (rm[j] * dcr[j] * dmr[j]);
  }
  for (int j = 2; j <= nr; j++) {
    h->hv[6][j] = 0.25 * g.dx1 / rc[j] * g.axsym;
    h->hv[7][j] = 0.25 * g.dx2 / (rm[j - 1] * dcr[j]);
    h->hv[8][j] = 0.25 * g.dx2 / (rm[j] * dcr[j]);
    h->hv[9][j] = g.dx1 / (h->Re * rc[j] * g.axsym);
    h->hv[14][j] = g.axsym * g.dx1q / (rc[j] * rc[j]);
    h->hv[15][j] = g.dx2q * rc[j] / rm[j] / (dcr[j] * dmr[j]);
    h->hv[16][j] = g.dx2q * rc[j] / rm[j - 1] / (dcr[j] * dmr[j - 1]);
    h->hv[17][j] = h->hv[15][j] + h->hv[16][j];
  }
  h->da.assign(nr + 3, 0.0);
  for (int j = 1; j <= nr; j++) {
    h->da[j] = rm[j] * dmr[j] / (g.dx1 * g.dx2);
    double x = rm[j] / h->rmax;
    h->hv[20][j] = 1.0 - x * x;  // qc(nr+1) stays 0
  }
  double red = 0.0;  // momentum_terms_swirl.F90:261-267 (j outer, theta inner)
  for (int j = 1; j <= nr; j++)
    for (int i = 1; i <= h->nt; i++) red = red + h->hv[20][j] * h->hv[20][j] * h->da[j];
  h->cor_red = red;
  h->jlim = 0;
  for (int j = 1; j <= nr; j++)
    if (rc[j] <= (double)0.4999f) h->jlim = j;
}

// Partitioned-Thomas coefficients from (aa, bb, cc) with aa(1)=0, cc(n)=0
// forced as solve_lap_par does (tripar.F90:30-35); Thomas pivots as
// solve_lap_thom (tripar.F90:128-138).
static void line_coef(const std::vector<double> &aa0, const std::vector<double> &bb,
                      const std::vector<double> &cc0, int M, std::vector<double> out[7]) {
  const int n = (int)bb.size();
  std::vector<double> aa = aa0, cc = cc0, beta(n);
  aa[0] = 0.0;
  cc[n - 1] = 0.0;
  beta[0] = bb[0];
  for (int k = 1; k < n; k++) beta[k] = bb[k] - cc[k - 1] * aa[k] / beta[k - 1];
  const int P = (n + M - 1) / M;
  for (int q = 0; q < 7; q++) out[q].assign(q < 5 ? n : P, 0.0);
  for (int k = 0; k < n; k++) {
    out[0][k] = 1.0 / beta[k];
    out[1][k] = -aa[k] / beta[k];
    out[2][k] = cc[k] / beta[k];
  }
  for (int s = 0; s < P; s++) {
    int p0 = s * M, p1 = std::min(n, (s + 1) * M);
    double acc = 1.0;
    for (int p = p0; p < p1; p++) { acc *= out[1][p]; out[3][p] = acc; }
    out[5][s] = acc;
    acc = 1.0;
    for (int p = p1 - 1; p >= p0; p--) { acc *= -out[2][p]; out[4][p] = acc; }
    out[6][s] = acc;
  }
}

// coesys (momentum_coef.F90:71-465) for substep irk, as six sweeps:
// r-q1, r-q3, r-q2 (identity rows at 1 and nrc, step.F90:315-328),
// z-q1, z-q3, z-q2.
void host_line_coefs(vrh_t *h, int irk, std::vector<double> out[6][7], const int M[6]) {
  const int nr = h->nr, nrc = h->nrc, nz = h->nz, nzc = h->nzc;
  const VGrid &g = h->g;
  const std::vector<double> &rm = h->hv[0], &rc = h->hv[1];
  const std::vector<double> &s1t2n = h->hv[11], &s1t2s = h->hv[12], &s1t2c = h->hv[13];
  const std::vector<double> &s2t2p = h->hv[15], &s2t2m = h->hv[16], &s2t2c = h->hv[17];
  const std::vector<double> &s3t2n = h->hv[18], &s3t2s = h->hv[19];
  const double alm = h->rk[irk].a, slip = h->slip;
  double betadx = 0.5 * alm * h->dt / h->Re;
  std::vector<double> am(nr + 2), ac(nr + 2), ap(nr + 2);
  // q1, r
  double a22p = betadx * s1t2c[1] * s1t2n[1], a22m;
  am[1] = 0.0; ac[1] = 1.0 + a22p / rm[1]; ap[1] = -a22p / rm[2];
  a22m = betadx * s1t2c[nr] * s1t2s[nr];
  ap[nr] = 0.0;
  ac[nr] = 1.0 + a22m / rm[nr] -
           betadx * s1t2c[nr] * rc[nrc] * rc[nrc] * (rc[nrc] / rm[nr] * slip - rc[nrc] / rm[nrc]) / h->hv[22][nrc];
  am[nr] = -a22m / rm[nr - 1];
  for (int j = 2; j <= nr - 1; j++) {
    a22p = betadx * s1t2c[j] * s1t2n[j];
    a22m = betadx * s1t2c[j] * s1t2s[j];
    ap[j] = -a22p / rm[j + 1]; ac[j] = 1.0 + (a22p + a22m) / rm[j]; am[j] = -a22m / rm[j - 1];
  }
  auto slice = [](const std::vector<double> &v, int a, int b) {
    return std::vector<double>(v.begin() + a, v.begin() + b + 1);
  };
  line_coef(slice(am, 1, nr), slice(ac, 1, nr), slice(ap, 1, nr), M[0], out[SW_R1]);
  // q3, r
  for (int j = 1; j <= nr; j++) {
    ap[j] = -(betadx * s3t2n[j]);
    double jmc = (j == 1) ? 0.0 : 1.0;
    am[j] = -(betadx * s3t2s[j] * jmc);
    ac[j] = 1.0 - (ap[j] + am[j]);
  }
  ap[nr] = 0.0;
  am[nr] = -(betadx * s3t2s[nr]);
  ac[nr] = 1.0 - (betadx * s3t2n[nr] * (slip - 1.0) + am[nr]);
  line_coef(slice(am, 1, nr), slice(ac, 1, nr), slice(ap, 1, nr), M[1], out[SW_R3]);
  // q2, r: rows 1 and nrc are identities (step.F90:315-328)
  {
    std::vector<double> a2(nrc), b2(nrc), c2(nrc);
    for (int j = 1; j <= nrc; j++) {
      if (j == 1 || j == nrc) { a2[j - 1] = 0.0; b2[j - 1] = 1.0; c2[j - 1] = 0.0; }
      else { a2[j - 1] = -(betadx * s2t2m[j]); b2[j - 1] = 1.0 + betadx * s2t2c[j]; c2[j - 1] = -(betadx * s2t2p[j]); }
    }
    line_coef(a2, b2, c2, M[2], out[SW_R2]);
  }
  // z sweeps (momentum_coef.F90:314-465, if_grad_null = 0)
  if (!h->zmesh[0].empty()) {
    // axial mesh: centre rows (q1, q2) and face rows (q3) from the second-
    // difference weights; the inlet row of q1, q2 is the one-sided second
    // difference through the wall value, f'' = ws fs - (ws + w2) f1 + w2 f2
    // (momentum_coef.F90's (-8/3, 4, -4/3) row on the uniform mesh); the end
    // rows keep the uniform forms
    const double bz = 0.5 * alm * h->dt / h->Re, dx3 = g.dx3;
    const std::vector<double> &dmz = h->zmesh[2], &dcz = h->zmesh[3];
    auto zm3 = [&](int k) { return dx3 / dmz[k]; };
    auto zc3 = [&](int k) { return dx3 / dcz[k]; };
    const double h1 = 0.5 / zm3(1), h2 = 1.0 / zc3(2);
    const double ws = 2.0 / (h1 * (h1 + h2)), w2 = 2.0 / (h2 * (h1 + h2));
    for (int comp = 0; comp < 3; comp++) {
      std::vector<double> a(nzc), b(nzc), c(nzc);
      for (int k = 1; k <= nz; k++) {
        const double wm = comp == 1 ? zc3(k) * zm3(k - 1) : zm3(k) * zc3(k);
        const double wp = comp == 1 ? zc3(k) * zm3(k) : zm3(k) * zc3(k + 1);
        a[k - 1] = -(bz * wm); b[k - 1] = 1.0 + bz * (wm + wp); c[k - 1] = -(bz * wp);
      }
      if (comp == 1) {
        a[nzc - 1] = 0.0; b[nzc - 1] = 1.0; c[nzc - 1] = 0.0;
        a[0] = 0.0; b[0] = 1.0; c[0] = 0.0;
      } else {
        a[nzc - 1] = 1.0; b[nzc - 1] = 1.0; c[nzc - 1] = 0.0;
        a[0] = -(bz * ws); b[0] = 1.0 + bz * (ws + w2); c[0] = -(bz * w2);
      }
      int sw = comp == 0 ? SW_Z1 : comp == 1 ? SW_Z3 : SW_Z2;
      line_coef(a, b, c, M[sw], out[sw]);
    }
    return;
  }
  betadx = 0.5 * g.dx3q * alm * h->dt / h->Re;
  for (int comp = 0; comp < 3; comp++) {
    std::vector<double> a(nzc), b(nzc), c(nzc);
    for (int k = 1; k <= nz; k++) { a[k - 1] = -betadx; b[k - 1] = 1.0 + 2.0 * betadx; c[k - 1] = -betadx; }
    if (comp == 1) {  // q3: identity rows at both ends
      a[nzc - 1] = 0.0; b[nzc - 1] = 1.0; c[nzc - 1] = 0.0;
      a[0] = 0.0; b[0] = 1.0; c[0] = 0.0;
    } else {          // q1, q2: (1,1) at the outlet, one-sided row at the inlet
      a[nzc - 1] = 1.0; b[nzc - 1] = 1.0; c[nzc - 1] = 0.0;
      a[0] = -(8.0 / 3.0 * betadx); b[0] = 1.0 + 4.0 * betadx; c[0] = -(4.0 / 3.0 * betadx);
    }
    int sw = comp == 0 ? SW_Z1 : comp == 1 ? SW_Z3 : SW_Z2;
    line_coef(a, b, c, M[sw], out[sw]);
  }
}

// Row j (1..nr) of the radial Poisson operator: sub- and super-diagonal and
// the radial part of the diagonal.  Uniform mesh: poisson.F90:379-381.  A
// mesh set by vrh_set_radial_mesh: minus divq's radial difference
// (divergence.F90:39, with dmr) of uppres' radial gradient (divergence.F90:132,
// with dcr), so that the projection leaves a discretely solenoidal field on
// that mesh; on the uniform mesh this is the same operator (rc = rm -+ dr/2).
void host_radial_row(const vrh_t *h, int j, double &aa, double &cc, double &dd) {
  const std::vector<double> &rm = h->hv[0], &rc = h->hv[1], &dmr = h->hv[21], &dcr = h->hv[22];
  if (h->mesh[0].empty()) {
    const double dr = h->rmax / h->nr;
    aa = -1.0 / (dr * dr) + 1.0 / (2.0 * rm[j] * dr);
    cc = -1.0 / (dr * dr) - 1.0 / (2.0 * rm[j] * dr);
    dd = 2.0 / (dr * dr);
  } else {
    const double w = h->g.dx2q / (rm[j] * dmr[j]);
    aa = -w * rc[j] / dcr[j];
    cc = -w * rc[j + 1] / dcr[j + 1];
    dd = -(aa + cc);
  }
}

// Inverse Thomas pivots of every (mode, kz) radial system of the Poisson
// solve (poisson.F90:379-381; poisson_solve.F90:284-319), in the spectral
// layout (position, j, kz).  The operator does not change during a run.
void host_radial_pivots(vrh_t *h, std::vector<double> &ib, std::vector<double> &ib64) {
  const int nt = h->nt, nr = h->nr, nz = h->nz;
  const double dz = h->zmax / nz, dth = 2.0 * acos(-1.0) / nt, pi = acos(-1.0);
  const std::vector<double> &rm = h->hv[0];
  std::vector<double> aa(nr + 1), bb(nr + 1), cc(nr + 1), beta(nr + 1);
  ib.assign((size_t)nt * nr * nz, 0.0);
  ib64.assign((size_t)5 * nr * nz, 0.0);
  for (int kz = 1; kz <= nz; kz++) {
    int mz = kz - 1;
    // axial mesh: the eigenvalue of the z transform's mode mz
    double wz = h->zlam.empty() || h->zmesh[0].empty() ? 2.0 / (dz * dz) * (1.0 - cos(mz * pi / nz)) : h->zlam[mz];
    for (int pos = 0; pos < nt; pos++) {
      int m = (pos + 1) / 2;
      double wth = (h->theta_eigen == VRH_THETA_EXACT_M2)
                       ? (double)((float)m * (float)m)
                       : 2.0 / (dth * dth) * (1.0 - cos(2.0 * pi * m / nt));
      for (int j = 1; j <= nr; j++) {
        double dd;
        host_radial_row(h, j, aa[j], cc[j], dd);
        bb[j] = dd + wth / (rm[j] * rm[j]) + wz;
      }
      int bcr = (m + mz != 0) ? 1 : -1;
      bb[1] = bb[1] + 1 * aa[1];
      bb[nr] = bb[nr] + bcr * cc[nr];
      aa[1] = 0.0;
      cc[nr] = 0.0;
      beta[1] = bb[1];
      for (int j = 2; j <= nr; j++) beta[j] = bb[j] - cc[j - 1] * aa[j] / beta[j - 1];
      for (int j = 1; j <= nr; j++) {
        ib[(size_t)pos + (size_t)nt * ((j - 1) + (size_t)nr * (kz - 1))] = 1.0 / beta[j];
        if (pos < 5) ib64[(size_t)pos + 5 * ((j - 1) + (size_t)nr * (kz - 1))] = 1.0 / beta[j];
      }
    }
  }
}

// Pressure-solve z transforms on an axial mesh.  The z part of minus the
// Laplacian, -(D_z G_z) with Neumann ends (the inlet and outlet faces carry
// no pressure gradient), is M^-1 K with M = diag(1/m3) and K the symmetric
// face-coupling matrix (K(k,k) = c3(k) + c3(k+1) without the end faces,
// K(k,k+1) = -c3(k+1)).  With D = diag(sqrt(m3)) it is similar to the
// symmetric tridiagonal A = D K D: A W = W diag(lam), V = D W, V^-1 = W^T D^-1.
// Eigenvalues by Sturm-sequence bisection (Wilkinson's method); eigenvectors
// by inverse iteration on A - lam I with a partially pivoted tridiagonal LU,
// each orthogonalised against all earlier ones.  The constant mode is exact:
// lam[0] = 0 with w0 ~ 1/sqrt(m3).  On the uniform mesh lam[q] =
// (2/dz^2)(1 - cos(pi q/nz)) and V holds the DCT-II cosines.
int host_axial_transforms(const double *m3, const double *c3, int nz, std::vector<double> &vi,
                          std::vector<double> &vf, std::vector<double> &lam, std::string &err) {
  const int n = nz;
  std::vector<double> a(n), e(n > 1 ? n - 1 : 0), sq(n);
  for (int i = 0; i < n; i++) {
    const int k = i + 1;
    a[i] = m3[k] * ((k > 1 ? c3[k] : 0.0) + (k < nz ? c3[k + 1] : 0.0));
    sq[i] = std::sqrt(m3[k]);
  }
  for (int i = 0; i + 1 < n; i++) e[i] = -std::sqrt(m3[i + 1] * m3[i + 2]) * c3[i + 2];
  double lo = 0.0, hi = 0.0, anorm = 0.0;
  for (int i = 0; i < n; i++) {
    const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(e[i]) : 0.0);
    lo = std::min(lo, a[i] - r);
    hi = std::max(hi, a[i] + r);
    anorm = std::max(anorm, std::fabs(a[i]) + r);
  }
  const double eps = 2.220446049250313e-16, tiny = eps * anorm;
  // number of eigenvalues below x (Sturm count of the LDL^T pivots)
  auto below = [&](double x) {
    int cnt = 0;
    double d = a[0] - x;
    if (d == 0.0) d = -tiny;
    if (d < 0.0) cnt++;
    for (int i = 1; i < n; i++) {
      d = (a[i] - x) - e[i - 1] * e[i - 1] / d;
      if (d == 0.0) d = -tiny;
      if (d < 0.0) cnt++;
    }
    return cnt;
  };
  lam.assign(n, 0.0);
  for (int q = 1; q < n; q++) {
    double l = lo, h = hi;
    for (int it = 0; it < 200 && h - l > 2.0 * eps * std::max(std::fabs(l), std::fabs(h)) + tiny * 1e-3; it++) {
      const double mid = 0.5 * (l + h);
      if (below(mid) > q) h = mid; else l = mid;
    }
    lam[q] = 0.5 * (l + h);
  }
  std::vector<double> W((size_t)n * n), v(n), dl(n), dd(n), du(n), du2(n);
  std::vector<char> piv(n);
  {  // constant mode
    double nrm = 0.0;
    for (int i = 0; i < n; i++) { v[i] = 1.0 / sq[i]; nrm += v[i] * v[i]; }
    nrm = std::sqrt(nrm);
    for (int i = 0; i < n; i++) W[i] = v[i] / nrm;
  }
  auto orth = [&](int q) {  // v minus its projection on W rows 0..q-1 (twice), normalised
    for (int pass = 0; pass < 2; pass++)
      for (int p = 0; p < q; p++) {
        const double *w = &W[(size_t)p * n];
        double s2 = 0.0;
        for (int i = 0; i < n; i++) s2 += w[i] * v[i];
        for (int i = 0; i < n; i++) v[i] -= s2 * w[i];
      }
    double nrm = 0.0;
    for (int i = 0; i < n; i++) nrm += v[i] * v[i];
    nrm = std::sqrt(nrm);
    for (int i = 0; i < n; i++) v[i] /= nrm;
    return nrm;
  };
  for (int q = 1; q < n; q++) {
    // LU with partial pivoting of A - lam I (rows i, i+1 swapped where the
    // subdiagonal dominates; du2 the fill of the second superdiagonal)
    for (int i = 0; i < n; i++) {
      dd[i] = a[i] - lam[q];
      du[i] = i + 1 < n ? e[i] : 0.0;
      dl[i] = i + 1 < n ? e[i] : 0.0;
      du2[i] = 0.0;
      piv[i] = 0;
    }
    for (int i = 0; i + 1 < n; i++) {
      if (std::fabs(dd[i]) >= std::fabs(dl[i])) {
        if (dd[i] == 0.0) dd[i] = tiny;
        const double f = dl[i] / dd[i];
        dl[i] = f;
        dd[i + 1] -= f * du[i];
      } else {
        const double f = dd[i] / dl[i];
        dd[i] = dl[i];
        dl[i] = f;
        const double t = du[i];
        du[i] = dd[i + 1];
        dd[i + 1] = t - f * dd[i + 1];
        if (i + 2 < n) { du2[i] = du[i + 1]; du[i + 1] = -f * du[i + 1]; }
        piv[i] = 1;
      }
    }
    if (dd[n - 1] == 0.0) dd[n - 1] = tiny;
    for (int i = 0; i < n; i++) v[i] = 1.0 + 0.5 * std::sin(0.7 + 1.3 * (double)i * (q + 1));
    orth(q);
    for (int it = 0; it < 3; it++) {
      for (int i = 0; i + 1 < n; i++) {
        if (piv[i]) std::swap(v[i], v[i + 1]);
        v[i + 1] -= dl[i] * v[i];
      }
      v[n - 1] /= dd[n - 1];
      if (n > 1) v[n - 2] = (v[n - 2] - du[n - 2] * v[n - 1]) / dd[n - 2];
      for (int i = n - 3; i >= 0; i--) v[i] = (v[i] - du[i] * v[i + 1] - du2[i] * v[i + 2]) / dd[i];
      orth(q);
    }
    std::memcpy(&W[(size_t)q * n], v.data(), n * sizeof(double));
  }
  // residual max |A w - lam w| and orthogonality of neighbours and of the
  // constant mode, relative to |A|
  double res = 0.0, dot = 0.0;
  for (int q = 0; q < n; q++) {
    const double *w = &W[(size_t)q * n];
    for (int i = 0; i < n; i++) {
      double aw = a[i] * w[i] - lam[q] * w[i];
      if (i > 0) aw += e[i - 1] * w[i - 1];
      if (i + 1 < n) aw += e[i] * w[i + 1];
      res = std::max(res, std::fabs(aw));
    }
    for (int p : {0, q - 1}) {
      if (p < 0 || p == q) continue;
      double s2 = 0.0;
      for (int i = 0; i < n; i++) s2 += w[i] * W[(size_t)p * n + i];
      dot = std::max(dot, std::fabs(s2));
    }
  }
  if (!(res <= 1e-10 * anorm) || !(dot <= 1e-10)) {
    char b[160];
    snprintf(b, sizeof b, "z eigen-decomposition failed its check (residual %.3e of |A| %.3e, overlap %.3e)", res,
             anorm, dot);
    err = b;
    return -1;
  }
  vi.assign((size_t)n * n, 0.0);
  vf.assign((size_t)n * n, 0.0);
  for (int q = 0; q < n; q++)
    for (int i = 0; i < n; i++) {
      const double w = W[(size_t)q * n + i];
      vi[(size_t)q * n + i] = w / sq[i];  // V^-1 = W^T D^-1
      vf[(size_t)i * n + q] = sq[i] * w;  // V = D W
    }
  return 0;
}

// Inlet time profile (gen_tools_swirl.F90:159-175); 0.1 is a default real.
double host_profile(double t) {
  const double n = 2.0, t1 = (double)0.1f * n;
  double u = 0.0, tt = t;
  if (tt > 0.0 && tt < t1) { double x = tt / t1; u = 3.0 * (x * x) - 2.0 * (x * x * x); }
  if (tt > t1 && tt < n) u = 1.0;
  if (tt > n && tt < n + t1) { double x = (n + t1 - tt) / t1; u = 3.0 * (x * x) - 2.0 * (x * x * x); }
  if (tt > n + t1) u = 0.0;
  return u;
}

// Least-squares pseudo-inverse of the 27-point quadratic design matrix that
// part_tools.F90:324-357 tabulates in closed form.
void host_mint(vrh_t *h, double *mint) {
  const double d[3] = {2.0 * acos(-1.0) / h->nt, h->rmax / h->nr, h->zmax / h->nz};
  double X[27][9], A[9][18];
  for (int c = -1; c <= 1; c++)
    for (int b = -1; b <= 1; b++)
      for (int a = -1; a <= 1; a++) {
        int iv = (a + 1) + 3 * (b + 1) + 9 * (c + 1);
        double x0 = a * d[0], x1 = b * d[1], x2 = c * d[2];
        double v[9] = {x0, x1, x2, x0 * x0, x0 * x1, x1 * x1, x1 * x2, x2 * x2, x2 * x0};
        for (int q = 0; q < 9; q++) X[iv][q] = v[q];
      }
  for (int i = 0; i < 9; i++)
    for (int j = 0; j < 18; j++) {
      if (j < 9) { double s = 0; for (int k = 0; k < 27; k++) s += X[k][i] * X[k][j]; A[i][j] = s; }
      else A[i][j] = (j - 9 == i) ? 1.0 : 0.0;
    }
  for (int c = 0; c < 9; c++) {
    int p = c;
    for (int r = c + 1; r < 9; r++) if (fabs(A[r][c]) > fabs(A[p][c])) p = r;
    if (p != c) for (int j = 0; j < 18; j++) std::swap(A[c][j], A[p][j]);
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

// ---- initial condition: discrete Stokes streamfunction of Gaussian rings ----
static void fft_inplace(std::vector<std::complex<double>> &a, int sign) {
  const int n = (int)a.size();
  for (int i = 1, j = 0; i < n; i++) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (int len = 2; len <= n; len <<= 1) {
    double ang = sign * 2.0 * M_PI / len;
    std::complex<double> wl(cos(ang), sin(ang));
    for (int i = 0; i < n; i += len) {
      std::complex<double> w(1.0, 0.0);
      for (int k = 0; k < len / 2; k++) {
        std::complex<double> u = a[i + k], v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
        w *= wl;
      }
    }
  }
}

// DST-I of x[0..m-1] (m = n-1 interior points): X_p = sum_q x_q sin(pi (p+1)(q+1)/n)
static void dst1(std::vector<double> &x, int n) {
  const int m = n - 1;
  if ((n & (n - 1)) == 0) {
    std::vector<std::complex<double>> a(2 * n, 0.0);
    for (int q = 0; q < m; q++) { a[q + 1] = x[q]; a[2 * n - 1 - q] = -x[q]; }
    fft_inplace(a, -1);
    for (int p = 0; p < m; p++) x[p] = -a[p + 1].imag() / 2.0;
  } else {
    std::vector<double> y(m);
    for (int p = 0; p < m; p++) {
      double s = 0;
      for (int q = 0; q < m; q++) s += x[q] * sin(M_PI * (p + 1) * (q + 1) / n);
      y[p] = s;
    }
    x = y;
  }
}

extern "C" int vrh_ring_init(const vrh_grid *grid, int nrings, const double *rings, double pert_amp,
                             int pert_m0, int pert_m1, const double *phase, double *q1, double *q2,
                             double *q3) {
  if (!grid || !q1 || !q2 || !q3 || nrings < 0 || (nrings > 0 && !rings)) return -1;
  const int nt = grid->ntheta, nr = grid->nr, nz = grid->nz;
  if (nt < 1 || nr < 2 || nz < 2) return -1;
  const double dr = grid->rmax / nr, dz = grid->zmax / nz, dth = 2.0 * M_PI / nt;
  const size_t sj = nt, sk = (size_t)nt * (nr + 3), ntot = sk * (nz + 3);
  memset(q1, 0, ntot * sizeof(double));
  memset(q2, 0, ntot * sizeof(double));
  memset(q3, 0, ntot * sizeof(double));
  const bool perturbed = pert_amp != 0.0 && pert_m1 >= pert_m0;
  const int nplanes = perturbed ? nt : 1;
  // interior nodes j=2..nr (rc), k=2..nz (zc); psi = 0 on the boundary nodes
  const int mj = nr - 1, mk = nz - 1;
  std::vector<double> psi((size_t)(nr + 2) * (nz + 2));
  std::vector<double> line(mk), rhs((size_t)mj * mk), cp(mj), dp(mj);
  for (int plane = 0; plane < nplanes; plane++) {
    double th = (plane + 0.5) * dth;  // q2,q3 live at theta_m
    double dR = 0.0;
    if (perturbed)
      for (int m = pert_m0; m <= pert_m1; m++) dR += pert_amp * cos(m * th + (phase ? phase[m - pert_m0] : 0.0));
    // right-hand side -r omega at interior nodes, transformed in z
    for (int j = 2; j <= nr; j++) {
      double r = (j - 1) * dr;
      for (int k = 2; k <= nz; k++) {
        double z = (k - 1) * dz, w = 0.0;
        for (int q = 0; q < nrings; q++) {
          double R = rings[4 * q] + dR, z0 = rings[4 * q + 1], s = rings[4 * q + 2], G = rings[4 * q + 3];
          double d2 = (r - R) * (r - R) + (z - z0) * (z - z0);
          w += G / (M_PI * s * s) * exp(-d2 / (s * s));
        }
        line[k - 2] = -r * w;
      }
      dst1(line, nz);
      for (int p = 0; p < mk; p++) rhs[(size_t)(j - 2) * mk + p] = line[p];
    }
    // per sine mode: tridiagonal in r,
    //   rc_j [(psi_{j+1}-psi_j)/rm_j - (psi_j-psi_{j-1})/rm_{j-1}]/dr^2 - lam_p psi_j = rhs
    for (int p = 0; p < mk; p++) {
      double lam = 2.0 / (dz * dz) * (1.0 - cos(M_PI * (p + 1) / nz));
      for (int jj = 0; jj < mj; jj++) {
        int j = jj + 2;
        double rc = (j - 1) * dr, rmj = (j - 0.5) * dr, rmm = (j - 1.5) * dr;
        double a = rc / (rmm * dr * dr), c = rc / (rmj * dr * dr), b = -(a + c) - lam;
        double f = rhs[(size_t)jj * mk + p];
        if (jj == 0) { cp[0] = c / b; dp[0] = f / b; }
        else {
          double den = b - a * cp[jj - 1];
          cp[jj] = c / den;
          dp[jj] = (f - a * dp[jj - 1]) / den;
        }
      }
      for (int jj = mj - 1; jj >= 0; jj--) {
        double v = dp[jj] - (jj + 1 < mj ? cp[jj] * rhs[(size_t)(jj + 1) * mk + p] : 0.0);
        rhs[(size_t)jj * mk + p] = v;
      }
    }
    std::fill(psi.begin(), psi.end(), 0.0);
    for (int j = 2; j <= nr; j++) {
      for (int p = 0; p < mk; p++) line[p] = rhs[(size_t)(j - 2) * mk + p];
      dst1(line, nz);
      for (int k = 2; k <= nz; k++) psi[(size_t)j * (nz + 2) + k] = line[k - 2] * 2.0 / nz;
    }
    auto PS = [&](int j, int k) { return psi[(size_t)j * (nz + 2) + k]; };
    int i0 = perturbed ? plane + 1 : 1, i1 = perturbed ? plane + 1 : nt;
    for (int i = i0; i <= i1; i++)
      for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= nr; j++) {
          size_t o = (size_t)(i - 1) + sj * j + sk * k;
          q2[o] = -(PS(j, k + 1) - PS(j, k)) / dz;
          q3[o] = (PS(j + 1, k) - PS(j, k)) / (dr * (j - 0.5) * dr);
        }
  }
  return 0;
}

extern "C" int vrh_cfl_dt(const vrh_grid *grid, const double *q1, const double *q2, const double *q3,
                          double cfl, double ujet, double *dt) {
  if (!grid || !q1 || !q2 || !q3 || !dt || cfl <= 0) return -1;
  const int nt = grid->ntheta, nr = grid->nr, nz = grid->nz;
  const double dr = grid->rmax / nr, dz = grid->zmax / nz, dth = 2.0 * M_PI / nt;
  const size_t sj = nt, sk = (size_t)nt * (nr + 3);
  double rate = ujet / dz;
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= nr; j++)
      for (int i = 1; i <= nt; i++) {
        size_t o = (size_t)(i - 1) + sj * j + sk * k;
        double rmj = (j - 0.5) * dr, rcj = (j - 1) * dr;
        double ut = fabs(q1[o]) / (rmj * dth);
        double ur = rcj > 0 ? fabs(q2[o] / rcj) / dr : 0.0;
        double uz = fabs(q3[o]) / dz;
        double s = ut + ur + uz;
        if (s > rate) rate = s;
      }
  *dt = cfl / rate;
  return 0;
}
