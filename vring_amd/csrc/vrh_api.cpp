// vrh_api.cpp -- the C ABI of libvrh (declared in include/vrh.h): handle
// lifecycle, device memory, transfers and the substep driver.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "vrh_internal.h"

namespace {

int fail(vrh_t *h, const char *msg, int code = -1) {
  if (h) h->err = msg;
  return code;
}

#define HIPCHK(h, call)                                                             \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) {                                                         \
      char b_[256];                                                                 \
      snprintf(b_, sizeof b_, "%s: %s", #call, hipGetErrorString(e_));              \
      (void)hipGetLastError(); /* reported here, not by the next launch check */    \
      return fail(h, b_, -3);                                                       \
    }                                                                               \
  } while (0)

size_t haloed(const vrh_t *h) { return (size_t)h->nt * (h->nr + 3) * (h->nz + 3); }
size_t compact(const vrh_t *h) { return (size_t)h->nt * h->nr * h->nz; }
size_t wsize(const vrh_t *h) { return h->prec == VRH_MIXED ? sizeof(float) : sizeof(double); }

template <typename T> int dalloc(vrh_t *h, T **p, size_t n) {
  HIPCHK(h, hipMalloc((void **)p, n * sizeof(T) + 64));
  HIPCHK(h, hipMemsetAsync(*p, 0, n * sizeof(T) + 64, h->stream));
  return 0;
}
int dalloc_bytes(vrh_t *h, void **p, size_t bytes) {
  HIPCHK(h, hipMalloc(p, bytes + 64));
  HIPCHK(h, hipMemsetAsync(*p, 0, bytes + 64, h->stream));
  return 0;
}
// Host-to-device copy on the handle's stream, complete on return.  The
// stream is non-blocking, so a plain hipMemcpy (null stream) would not be
// ordered after the zero fills dalloc queues there, nor after kernels still
// reading the destination.
int h2d(vrh_t *h, void *dst, const void *src, size_t bytes) {
  HIPCHK(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

// host double -> device buffer of the work type
int upload_conv(vrh_t *h, void *dst, const double *src, size_t n, bool work) {
  if (!work || h->prec == VRH_FP64) {
    HIPCHK(h, hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
  }
  std::vector<float> tmp(n);
  for (size_t i = 0; i < n; i++) tmp[i] = (float)src[i];
  HIPCHK(h, hipMemcpyAsync(dst, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}
int download_conv(vrh_t *h, double *dst, const void *src, size_t n, bool work) {
  if (!work || h->prec == VRH_FP64) {
    HIPCHK(h, hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return 0;
  }
  std::vector<float> tmp(n);
  HIPCHK(h, hipMemcpyAsync(tmp.data(), src, n * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (size_t i = 0; i < n; i++) dst[i] = tmp[i];
  return 0;
}

struct Named {
  void *p;
  size_t n;
  bool work;
};
// Element counts of the named state arrays (vrh_state_size): haloed fields
// (nt, 0:nr+2, 0:nz+2), compact arrays (nt, nr, nz), prg (nt, nr, nz, 3)
// and the outlet planes (nt, 0:nr+2).  0 for an unknown name.
enum NameShape { NS_NONE = 0, NS_HALO, NS_COMPACT, NS_PRG, NS_PLANE };
NameShape name_shape(const char *name) {
  static const char *halo[] = {"q1", "q2", "q3", "dq1", "dq2", "dq3", "pres", "dph",
                               "ru1", "ru2", "ru3", "qrhs1", "qrhs2", "qrhs3"};
  static const char *comp[] = {"prg1", "prg2", "prg3", "spec"};
  static const char *plane[] = {"csort", "dq3no", "corr3n"};
  for (const char *n : halo) if (strcmp(n, name) == 0) return NS_HALO;
  for (const char *n : comp) if (strcmp(n, name) == 0) return NS_COMPACT;
  for (const char *n : plane) if (strcmp(n, name) == 0) return NS_PLANE;
  if (strcmp(name, "prg") == 0) return NS_PRG;
  return NS_NONE;
}
size_t shape_count(NameShape s, long nt, long nr, long nz) {
  switch (s) {
    case NS_HALO: return (size_t)(nt * (nr + 3) * (nz + 3));
    case NS_COMPACT: return (size_t)(nt * nr * nz);
    case NS_PRG: return (size_t)(3 * nt * nr * nz);
    case NS_PLANE: return (size_t)(nt * (nr + 3));
    default: return 0;
  }
}

bool lookup(vrh_t *h, const char *name, Named &out) {
  struct { const char *n; void *p; bool w; } t[] = {
      {"q1", h->f.q1, false}, {"q2", h->f.q2, false}, {"q3", h->f.q3, false},
      {"dq1", h->f.dq1, false}, {"dq2", h->f.dq2, false}, {"dq3", h->f.dq3, false},
      {"pres", h->f.pres, false}, {"dph", h->f.dph, false},
      {"ru1", h->f.ru1, true}, {"ru2", h->f.ru2, true}, {"ru3", h->f.ru3, true},
      {"qrhs1", h->f.r1, true}, {"qrhs2", h->f.r2, true}, {"qrhs3", h->f.r3, true},
      {"prg1", h->f.prg1, false}, {"prg2", h->f.prg2, false}, {"prg3", h->f.prg3, false},
      {"spec", h->f.S, true},
      {"csort", h->f.csort, false}, {"dq3no", h->f.dq3no, false}, {"corr3n", h->f.corr3n, false},
  };
  for (auto &e : t)
    if (strcmp(e.n, name) == 0) {
      out = {e.p, shape_count(name_shape(name), h->nt, h->nr, h->nz), e.w};
      return true;
    }
  return false;
}

// One substep whose time-dependent scalars (inlet increment, mass
// correction) are read from d_sc, as written by k_scalars at the start of
// the step: the form that is captured in a graph.  join (optional): work on
// another stream that must end before the projection writes q.
template <typename TW> void substep_dev(vrh_t *h, int irk, hipEvent_t join) {
  launch_explicit<TW>(h, irk);
  launch_sweeps<TW>(h, irk, 0.0, h->d_sc + irk, true, 0.0, h->d_sc + 3 + irk);
  launch_poisson<TW>(h, irk, nullptr, nullptr);
  if (join) (void)hipStreamWaitEvent(h->stream, join, 0);
  launch_project<TW>(h, irk);
}

template <typename TW> int substep_t(vrh_t *h, int irk, int phase) {
  const double t0 = h->ntime, t1 = h->ntime + h->dt;   // momentum_terms_swirl.F90:192-193
  const double unom0 = host_profile(t0), unom1 = host_profile(t1);
  const double alpha = h->rk[irk].g + h->rk[irk].r;
  const double c3s = (unom1 - unom0) * alpha;           // :212
  double qinf = 0.0;                                    // :261-267
  for (int j = 1; j <= h->nr; j++) {
    const double cj = (j <= h->jlim) ? c3s : 0.0;
    for (int i = 1; i <= h->nt; i++) qinf = qinf + cj * h->da[j];
  }
  const double cor = (qinf - 0.0) / (h->cor_red + (double)1.e-15f);  // :283
  if (phase == VRH_PH_HDNBC) launch_hdnbc<TW>(h, irk, cor, nullptr);
  if (phase < 0 || phase == VRH_PH_EXPLICIT) launch_explicit<TW>(h, irk);
  if (phase < 0 || phase == VRH_PH_ADI) launch_sweeps<TW>(h, irk, c3s, nullptr, phase < 0, cor, nullptr);
  if (phase < 0 || phase == VRH_PH_POISSON) launch_poisson<TW>(h, irk, nullptr, nullptr);
  if (phase < 0 || phase == VRH_PH_PROJECT) launch_project<TW>(h, irk);
  HIPCHK(h, hipGetLastError());
  if (h->mode_rc) { const int rc = h->mode_rc; h->mode_rc = 0; return rc; }
  return 0;
}

int substep(vrh_t *h, int irk, int phase) {
  if (irk < 1 || irk > 3) return fail(h, "irk must be 1..3");
  return h->prec == VRH_MIXED ? substep_t<float>(h, irk, phase) : substep_t<double>(h, irk, phase);
}

int check_flag(vrh_t *h) {
  int flag = 0;
  HIPCHK(h, hipMemcpyAsync(&flag, h->f.flag, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (flag) return fail(h, "Flow field crash (NaN in the pressure gradient)", -2);
  return 0;
}

}  // namespace

// A whole RK3 step (loop.F90:53-131): scalars, particles at irk 1, three
// substeps.  Nothing here synchronises or copies, so it can be captured.
int step_body(vrh_t *h) {
  launch_scalars(h);
  // The particle stage reads q as the previous projection left it and
  // writes only particle state; of the substep's kernels only the projection
  // writes q.  So the LSQ stage runs on a side stream beside the explicit,
  // sweep and pressure kernels and is joined before the projection (10^6
  // particles: 2.67 ms per step instead of 2.81).  The short trilinear stage
  // stays in order (beside the fluid kernels its gathers slowed it by 9 %),
  // as does everything with the kernel timer on or inside an ensemble graph.
  const bool ppar = !h->kt_on && !h->serial_sweeps && h->side[2] && h->pscheme != VRH_PART_TRILINEAR_RK3;
  for (int irk = 1; irk <= 3; irk++) {
    // loop.F90:66-70 calls iter_part at irk == 1 (the Euler scheme); the RK3
    // particle schemes take one stage per substep (launch_particles decides)
    hipEvent_t join = nullptr;
    if (h->nbatch > 0) {
      if (ppar) {
        (void)hipEventRecord(h->part_ev[irk][0], h->stream);
        (void)hipStreamWaitEvent(h->side[2], h->part_ev[irk][0], 0);
        if (launch_particles(h, irk, h->side[2])) {
          (void)hipEventRecord(h->part_ev[irk][1], h->side[2]);
          join = h->part_ev[irk][1];
        }
      } else {
        launch_particles(h, irk, h->stream);
      }
    }
    if (h->prec == VRH_MIXED) substep_dev<float>(h, irk, join);
    else substep_dev<double>(h, irk, join);
  }
  if (h->mode_rc) { const int rc = h->mode_rc; h->mode_rc = 0; return rc; }
  return hipGetLastError() == hipSuccess ? 0 : fail(h, "kernel launch failed", -3);
}

void graph_invalidate(vrh_t *h) {
  if (h->gexec) hipGraphExecDestroy(h->gexec);
  if (h->graph) hipGraphDestroy(h->graph);
  h->gexec = nullptr;
  h->graph = nullptr;
  h->graph_ok = false;
  h->epoch++;
}

namespace {
// Everything that depends on the radial mesh, recomputed on the host and
// uploaded into the handle's buffers: the metric vectors, the theta PCR
// depths, the line-sweep coefficients, the radial Poisson rows and pivots,
// the inlet area.
int refresh_mesh(vrh_t *h) {
  host_mesh_vectors(h);
  const int nv = h->nr + 3;
  int rc;
  std::vector<double> vb((size_t)kNumVec * nv);
  for (int q = 0; q < kNumVec; q++) std::memcpy(&vb[(size_t)q * nv], h->hv[q].data(), nv * sizeof(double));
  if ((rc = h2d(h, h->d_vec, vb.data(), vb.size() * sizeof(double)))) return rc;
  // parallel-cyclic-reduction depth of the theta solves: log2(nt) for a
  // power-of-two nt, otherwise until the off-diagonal ratio is below 1e-18
  for (int irk = 1; irk <= 3; irk++) {
    const double betadx = 0.5 * h->rk[irk].a * h->dt / h->Re;
    int lv = 0;
    if ((h->nt & (h->nt - 1)) == 0) {
      while ((1 << lv) < h->nt) lv++;
    } else {
      for (int j = 1; j <= h->nr; j++)
        for (int which = 0; which < 2; which++) {
          double acc = betadx * h->hv[which ? 14 : 10][j];
          double a = -acc, b = 1.0 + 2.0 * acc;
          int l = 0;
          while (fabs(a) >= 1e-18 * fabs(b) && l < 40) { double al = -a / b; b += 2.0 * al * a; a *= al; l++; }
          if (l > lv) lv = l;
        }
    }
    h->theta_levels[irk] = lv;
  }
  if ((rc = build_line_coefs(h))) return rc;
  std::vector<double> aab(h->nr + 1, 0.0), ccb(h->nr + 1, 0.0);
  for (int j = 1; j <= h->nr; j++) { double dd; host_radial_row(h, j, aab[j], ccb[j], dd); }
  if ((rc = h2d(h, h->d_aab, aab.data(), aab.size() * sizeof(double)))) return rc;
  if ((rc = h2d(h, h->d_ccb, ccb.data(), ccb.size() * sizeof(double)))) return rc;
  if (h->d_aabf) {  // round to nearest, as the device conversion
    std::vector<float> af(aab.begin(), aab.end()), cf(ccb.begin(), ccb.end());
    if ((rc = h2d(h, h->d_aabf, af.data(), af.size() * sizeof(float)))) return rc;
    if ((rc = h2d(h, h->d_ccbf, cf.data(), cf.size() * sizeof(float)))) return rc;
  }
  std::vector<double> ib, ib64;
  host_radial_pivots(h, ib, ib64);
  if ((rc = upload_conv(h, h->f.ib, ib.data(), ib.size(), true))) return rc;
  if (h->f.ib64 && (rc = h2d(h, h->f.ib64, ib64.data(), ib64.size() * sizeof(double)))) return rc;
  h->inlet_area = 0.0;  // momentum_terms_swirl.F90:261-267 with c3s factored out
  for (int j = 1; j <= h->jlim; j++)
    for (int i = 1; i <= h->nt; i++) h->inlet_area = h->inlet_area + h->da[j];
  return 0;
}
}  // namespace

namespace {
int build_graph(vrh_t *h) {
  graph_invalidate(h);
  HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  int rc = step_body(h);
  hipGraph_t gr = nullptr;
  hipError_t e = hipStreamEndCapture(h->stream, &gr);
  if (rc) { if (gr) hipGraphDestroy(gr); return rc; }
  HIPCHK(h, e);
  h->graph = gr;
  HIPCHK(h, hipGraphInstantiate(&h->gexec, h->graph, nullptr, nullptr, 0));
  h->graph_ok = true;
  return 0;
}
}  // namespace

extern "C" {

int vrh_last_error(vrh_t *h, char *buf, size_t len) {
  if (!buf || len == 0) return -1;
  const char *m = h ? h->err.c_str() : "null handle";
  snprintf(buf, len, "%s", m);
  return 0;
}

int vrh_create(const vrh_grid *grid, const vrh_params *p, int precision, vrh_t **out) {
  if (!grid || !p || !out) return -1;
  *out = nullptr;
  if (grid->ntheta < 3 || grid->nr < 4 || grid->nz < 4 || grid->nr > 16 * 32 || grid->nz > 16 * 72 - 1 ||
      grid->ntheta > 256 ||
      grid->rmax <= 0 || grid->zmax <= 0 || p->dt <= 0 || p->Re <= 0)
    return -1;
  if (precision != VRH_FP64 && precision != VRH_MIXED) return -1;
  vrh_t *h = new (std::nothrow) vrh_t();
  if (!h) return -1;
  *out = h;
  h->nt = grid->ntheta; h->nr = grid->nr; h->nz = grid->nz; h->nrc = h->nr + 1; h->nzc = h->nz + 1;
  h->rmax = grid->rmax; h->zmax = grid->zmax; h->Re = p->Re; h->dt = p->dt; h->slip = p->slip;
  h->theta_eigen = p->theta_eigen; h->prec = precision; h->ntime = 0.0;
  VGrid &g = h->g;
  g.nt = h->nt; g.nr = h->nr; g.nz = h->nz; g.nrc = h->nrc; g.nzc = h->nzc;
  g.sj = h->nt; g.sk = (long)h->nt * (h->nr + 3); g.ncomp = (long)compact(h);
  const double dth = 2.0 * acos(-1.0) / h->nt;                   // mesh_tools.F90:217-219, 315-320
  g.dx1 = 1.0 / dth; g.dx2 = 1.0 / (h->rmax / h->nr); g.dx3 = 1.0 / (h->zmax / h->nz);
  g.dx1q = g.dx1 * g.dx1; g.dx2q = g.dx2 * g.dx2; g.dx3q = g.dx3 * g.dx3;
  g.axsym = (h->nt == 2) ? 0.0 : 1.0;
  g.dt = h->dt; g.Re = h->Re; g.slip = h->slip;
  const double gam[4] = {0, 8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};   // para_tools.F90:103-113
  const double rom[4] = {0, 0.0, -17.0 / 60.0, -5.0 / 12.0};
  for (int i = 1; i <= 3; i++) h->rk[i] = {gam[i], rom[i], gam[i] + rom[i]};
  HIPCHK(h, hipGetDevice(&h->device));
  if (analysis_scratch_init(h->device) != 0)
    return fail(h, "vrh_create: the analysis scratch array has no device address (hipGetSymbolAddress)", -3);
  HIPCHK(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  lds_optin(h);
  for (int q = 0; q < 3; q++) HIPCHK(h, hipStreamCreateWithFlags(&h->side[q], hipStreamNonBlocking));
  for (int irk = 0; irk < 4; irk++) {
    HIPCHK(h, hipEventCreateWithFlags(&h->fork_ev[irk], hipEventDisableTiming));
    for (int q = 0; q < 2; q++) HIPCHK(h, hipEventCreateWithFlags(&h->part_ev[irk][q], hipEventDisableTiming));
    for (int q = 0; q < 2; q++) HIPCHK(h, hipEventCreateWithFlags(&h->join_ev[irk][q], hipEventDisableTiming));
    for (int q = 0; q < 2; q++) HIPCHK(h, hipEventCreateWithFlags(&h->poi_ev[irk][q], hipEventDisableTiming));
  }
  // radial vectors (values from refresh_mesh below)
  const int nv = h->nr + 3;
  int rc = dalloc(h, &h->d_vec, (size_t)kNumVec * nv);
  if (rc) return rc;
  const double **cv[kNumVec] = {&h->coef.rm, &h->coef.rc, &h->coef.h1d11, &h->coef.h1d22d, &h->coef.h1d22n,
                                &h->coef.h1dare, &h->coef.h2d11, &h->coef.h2d22m, &h->coef.h2d22p, &h->coef.h2dare,
                                &h->coef.s1t11, &h->coef.s1t2n, &h->coef.s1t2s, &h->coef.s1t2c, &h->coef.s2t11,
                                &h->coef.s2t2p, &h->coef.s2t2m, &h->coef.s2t2c, &h->coef.s3t2n, &h->coef.s3t2s,
                                &h->coef.qc, &h->coef.dmr, &h->coef.dcr};
  for (int q = 0; q < kNumVec; q++) *cv[q] = h->d_vec + (size_t)q * nv;
  // line coefficients for the three substeps; default segment lengths
  // minimise the padded work P*M (P <= 16 waves per block)
  {
    const int ns[6] = {h->nr, h->nr, h->nrc, h->nzc, h->nzc, h->nzc};
    for (int sw = 0; sw < 6; sw++) {
      int best = -1, bestw = 1 << 30;
      for (int M : kLineSegs) {
        int P = (ns[sw] + M - 1) / M;
        if (P > 16) continue;
        if (P * M < bestw) { bestw = P * M; best = M; }
      }
      if (best < 0) return fail(h, "grid too large for the line solver (nr+1, nz+1 <= 16 x 72)");
      h->Msw[sw] = best;
    }
  }
  // transform tables
  {
    std::vector<double> tt(2 * (size_t)h->nt), tz(2 * (size_t)h->nz), ez(2 * (size_t)h->nz);
    for (int q = 0; q < h->nt; q++) { tt[q] = cos(2.0 * M_PI * q / h->nt); tt[h->nt + q] = sin(2.0 * M_PI * q / h->nt); }
    for (int q = 0; q < h->nz; q++) {
      tz[q] = cos(2.0 * M_PI * q / h->nz); tz[h->nz + q] = sin(2.0 * M_PI * q / h->nz);
      ez[q] = cos(M_PI * q / (2.0 * h->nz)); ez[h->nz + q] = sin(M_PI * q / (2.0 * h->nz));
    }
    if ((rc = dalloc(h, &h->d_twt, tt.size())) || (rc = dalloc(h, &h->d_twz, tz.size())) ||
        (rc = dalloc(h, &h->d_ez, ez.size())))
      return rc;
    if ((rc = h2d(h, h->d_twt, tt.data(), tt.size() * sizeof(double)))) return rc;
    if ((rc = h2d(h, h->d_twz, tz.data(), tz.size() * sizeof(double)))) return rc;
    if ((rc = h2d(h, h->d_ez, ez.data(), ez.size() * sizeof(double)))) return rc;
    if ((rc = dalloc(h, &h->d_aab, (size_t)h->nr + 1)) || (rc = dalloc(h, &h->d_ccb, (size_t)h->nr + 1))) return rc;
    if (h->prec == VRH_MIXED &&
        ((rc = dalloc_bytes(h, reinterpret_cast<void **>(&h->d_aabf), sizeof(float) * ((size_t)h->nr + 1))) ||
         (rc = dalloc_bytes(h, reinterpret_cast<void **>(&h->d_ccbf), sizeof(float) * ((size_t)h->nr + 1)))))
      return rc;
  }
  // fields
  const size_t H = haloed(h), C = compact(h), W = wsize(h), PL = (size_t)h->nt * (h->nr + 3);
  Fields &f = h->f;
  double **dfs[] = {&f.q1, &f.q2, &f.q3, &f.dq1, &f.dq2, &f.dq3, &f.pres, &f.dph};
  for (auto p2 : dfs) if ((rc = dalloc(h, p2, H))) return rc;
  void **wfs[] = {&f.ru1, &f.ru2, &f.ru3, &f.r1, &f.r2, &f.r3};
  for (auto p2 : wfs) if ((rc = dalloc_bytes(h, p2, H * W))) return rc;
  void **cfs[] = {&f.S, &f.ib};
  for (auto p2 : cfs) if ((rc = dalloc_bytes(h, p2, C * W))) return rc;
  double **pgs[] = {&f.prg1, &f.prg2, &f.prg3};
  for (auto p2 : pgs) if ((rc = dalloc(h, p2, C))) return rc;
  double **pls[] = {&f.csort, &f.dq3no, &f.corr3n};
  for (auto p2 : pls) if ((rc = dalloc(h, p2, PL))) return rc;
  if ((rc = dalloc(h, &f.flag, 1))) return rc;
  if (precision == VRH_MIXED)
    if ((rc = dalloc(h, &f.S64, (size_t)5 * h->nr * h->nz)) || (rc = dalloc(h, &f.ib64, (size_t)5 * h->nr * h->nz)))
      return rc;
  std::vector<double> mint(9 * 27);
  host_mint(h, mint.data());
  if ((rc = dalloc(h, &h->d_mint, mint.size()))) return rc;
  if ((rc = h2d(h, h->d_mint, mint.data(), mint.size() * sizeof(double)))) return rc;
  for (size_t q = 0; q < mint.size() && q < 243; q++) h->h_mint[q] = mint[q];
  if ((rc = dalloc(h, &h->d_sc, 16))) return rc;
  HIPCHK(h, hipHostMalloc((void **)&h->h_pin, 16 * sizeof(double), hipHostMallocDefault));
  if ((rc = refresh_mesh(h))) return rc;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int vrh_destroy(vrh_t *h) {
  if (!h) return -1;
  if (h->stream) hipStreamSynchronize(h->stream);
  Fields &f = h->f;
  void *ptrs[] = {h->d_vec, h->d_line, h->d_linef, h->d_twt, h->d_twz, h->d_ez, h->d_aab, h->d_ccb, h->d_aabf, h->d_ccbf, h->d_mint, f.q1, f.q2,
                  f.q3, f.dq1, f.dq2, f.dq3, f.pres, f.dph, f.ru1, f.ru2, f.ru3, f.r1, f.r2, f.r3, f.prg1, f.prg2,
                  f.prg3, f.S, f.ib, f.S64, f.ib64, f.csort, f.dq3no, f.corr3n, f.flag};
  for (void *p : ptrs) if (p) hipFree(p);
  void *zp[] = {h->d_zmet, h->d_zv, h->d_zvf, h->S2, h->S2_64};
  for (void *p : zp) if (p) hipFree(p);
  if (h->blas) rocblas_destroy_handle((rocblas_handle)h->blas);
  for (auto &e : h->kt_pool) hipEventDestroy(e);
  graph_invalidate(h);
  if (h->d_sc) hipFree(h->d_sc);
  if (h->h_pin) hipHostFree(h->h_pin);
  if (h->copy_stream) hipStreamSynchronize(h->copy_stream);
  if (h->d_snap) hipFree(h->d_snap);
  if (h->d_anl) hipFree(h->d_anl);
  if (h->h_snap) hipHostFree(h->h_snap);
  if (h->snap_ev) hipEventDestroy(h->snap_ev);
  if (h->copy_stream) hipStreamDestroy(h->copy_stream);
  particles_release(h);
  for (int q = 0; q < 3; q++)
    if (h->side[q]) hipStreamDestroy(h->side[q]);
  for (int irk = 0; irk < 4; irk++) {
    if (h->fork_ev[irk]) hipEventDestroy(h->fork_ev[irk]);
    for (int q = 0; q < 2; q++)
      if (h->part_ev[irk][q]) hipEventDestroy(h->part_ev[irk][q]);
    for (int q = 0; q < 2; q++)
      if (h->join_ev[irk][q]) hipEventDestroy(h->join_ev[irk][q]);
    for (int q = 0; q < 2; q++)
      if (h->poi_ev[irk][q]) hipEventDestroy(h->poi_ev[irk][q]);
  }
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
  return 0;
}

int vrh_upload_fields(vrh_t *h, const double *q1, const double *q2, const double *q3, const double *pres) {
  if (!h) return -1;
  const double *src[4] = {q1, q2, q3, pres};
  double *dst[4] = {h->f.q1, h->f.q2, h->f.q3, h->f.pres};
  for (int q = 0; q < 4; q++)
    if (src[q]) { int rc = upload_conv(h, dst[q], src[q], haloed(h), false); if (rc) return rc; }
  return 0;
}

int vrh_download_fields(vrh_t *h, double *q1, double *q2, double *q3, double *pres) {
  if (!h) return -1;
  double *dst[4] = {q1, q2, q3, pres};
  const double *src[4] = {h->f.q1, h->f.q2, h->f.q3, h->f.pres};
  for (int q = 0; q < 4; q++)
    if (dst[q]) { int rc = download_conv(h, dst[q], src[q], haloed(h), false); if (rc) return rc; }
  return 0;
}

int vrh_upload_state(vrh_t *h, const char *name, const double *host) {
  if (!h || !name || !host) return -1;
  if (strcmp(name, "prg") == 0) {
    const size_t C = compact(h);
    void *d[3] = {h->f.prg1, h->f.prg2, h->f.prg3};
    for (int c = 0; c < 3; c++) { int rc = upload_conv(h, d[c], host + c * C, C, false); if (rc) return rc; }
    return 0;
  }
  Named n;
  if (!lookup(h, name, n)) return fail(h, "unknown field name");
  return upload_conv(h, n.p, host, n.n, n.work);
}

int vrh_download_state(vrh_t *h, const char *name, double *host) {
  if (!h || !name || !host) return -1;
  if (strcmp(name, "prg") == 0) {
    const size_t C = compact(h);
    void *d[3] = {h->f.prg1, h->f.prg2, h->f.prg3};
    for (int c = 0; c < 3; c++) { int rc = download_conv(h, host + c * C, d[c], C, false); if (rc) return rc; }
    return 0;
  }
  Named n;
  if (!lookup(h, name, n)) return fail(h, "unknown field name");
  return download_conv(h, host, n.p, n.n, n.work);
}

int vrh_state_size(const vrh_grid *grid, const char *name, long *n) {
  if (!grid || !name || !n) return -1;
  const NameShape s = name_shape(name);
  if (s == NS_NONE || grid->ntheta < 1 || grid->nr < 1 || grid->nz < 1) return -1;
  *n = (long)shape_count(s, grid->ntheta, grid->nr, grid->nz);
  return 0;
}

int vrh_device_ptr(vrh_t *h, const char *name, void **ptr) {
  if (!h || !name || !ptr) return -1;
  Named n;
  if (!lookup(h, name, n)) return fail(h, "unknown field name");
  *ptr = n.p;
  return 0;
}

int vrh_set_time(vrh_t *h, double t) {
  if (!h) return -1;
  h->ntime = t;
  h->time_dirty = true;
  return 0;
}
int vrh_get_time(vrh_t *h, double *t) { if (!h || !t) return -1; *t = h->ntime; return 0; }

int vrh_set_radial_mesh(vrh_t *h, const double *rm, const double *rc, const double *dmr, const double *dcr) {
  if (!h) return -1;
  if (!rm || !rc || !dmr || !dcr) return fail(h, "radial mesh: null array");
  if (!h->pb.empty()) return fail(h, "radial mesh: particles need the uniform mesh (set the mesh first)");
  const int nr = h->nr;
  bool ok = rc[1] == 0.0;
  for (int j = 1; j <= nr; j++)
    ok = ok && rc[j] < rm[j] && rm[j] < rc[j + 1] && dmr[j] > 0.0 && dcr[j] > 0.0;
  ok = ok && dcr[nr + 1] > 0.0 && rm[nr + 1] > rm[nr];
  if (!ok)
    return fail(h, "radial mesh: need rc(1) = 0, rc(j) < rm(j) < rc(j+1), rm(nr+1) > rm(nr), dmr, dcr > 0");
  const double *src[4] = {rm, rc, dmr, dcr};
  for (int q = 0; q < 4; q++) h->mesh[q].assign(src[q], src[q] + nr + 2);
  h->hz = nullptr;  // the per-mode solve needs the uniform radial mesh: switched off
  int rc2 = refresh_mesh(h);
  if (rc2) return rc2;
  return 0;
}

// Axial mesh: metric arrays, line coefficients, and the pressure solve's z
// transforms (generalised eigenvectors of the z operator, host-computed).
int vrh_set_axial_mesh(vrh_t *h, const double *zm, const double *zc, const double *dmz, const double *dcz) {
  if (!h) return -1;
  if (!zm || !zc || !dmz || !dcz) return fail(h, "axial mesh: null array");
  if (!h->pb.empty()) return fail(h, "axial mesh: particles need the uniform mesh");
  const int nz = h->nz;
  bool ok = zc[1] == 0.0 && nz >= 2;
  for (int k = 1; k <= nz; k++)
    ok = ok && zc[k] < zm[k] && zm[k] < zc[k + 1] && dmz[k] > 0.0 && dcz[k] > 0.0;
  ok = ok && dmz[0] > 0.0 && dcz[0] > 0.0 && dmz[nz + 1] > 0.0 && dcz[nz + 1] > 0.0 && zm[nz + 1] > zm[nz] &&
       zm[0] < zm[1];
  for (int k = 0; k <= nz + 1; k++) ok = ok && std::isfinite(zm[k]) && std::isfinite(zc[k]);
  if (!ok)
    return fail(h, "axial mesh: need nz >= 2, zc(1) = 0, zc(k) < zm(k) < zc(k+1), zm(0) < zm(1), zm(nz+1) > zm(nz), "
                   "dmz, dcz > 0 (k = 0..nz+1)");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  // everything is computed and the device buffers allocated before the
  // handle's mesh state changes: a failure leaves the previous mesh in place
  std::vector<double> zmn[4];
  const double *src[4] = {zm, zc, dmz, dcz};
  for (int q = 0; q < 4; q++) {
    zmn[q].assign(src[q], src[q] + nz + 2);
    zmn[q].push_back(zmn[q][nz + 1]);  // k = nz+2 repeats nz+1
  }
  // metric arrays (k = 0..nz+2)
  const int L = nz + 3;
  std::vector<double> mb((size_t)7 * L, 0.0);
  double *m3 = &mb[0], *c3 = &mb[L], *cm = &mb[2 * L], *cp = &mb[3 * L], *fm = &mb[4 * L], *fp = &mb[5 * L],
         *zp = &mb[6 * L];
  const double dx3 = h->g.dx3;
  for (int k = 0; k < L; k++) { m3[k] = dx3 / zmn[2][k]; c3[k] = dx3 / zmn[3][k]; zp[k] = zmn[0][k]; }
  for (int k = 0; k < L; k++) {
    cm[k] = m3[k] * c3[k];
    cp[k] = m3[k] * c3[k + 1 < L ? k + 1 : k];
    fm[k] = c3[k] * m3[k > 0 ? k - 1 : 0];
    fp[k] = c3[k] * m3[k];
  }
  const double h1 = 0.5 / m3[1], h2 = 1.0 / c3[2];
  std::vector<double> vi, vf, lam;
  std::string emsg;
  if (host_axial_transforms(m3, c3, nz, vi, vf, lam, emsg)) return fail(h, ("axial mesh: " + emsg).c_str());
  const size_t nn = (size_t)nz * nz;
  int rc;
  if (!h->d_zmet && (rc = dalloc(h, &h->d_zmet, mb.size()))) return rc;
  if (!h->d_zv && (rc = dalloc(h, &h->d_zv, 2 * nn))) return rc;
  if (h->prec == VRH_MIXED) {
    if (!h->d_zvf && (rc = dalloc(h, &h->d_zvf, 2 * nn))) return rc;
    if (!h->S2_64 && (rc = dalloc(h, &h->S2_64, (size_t)5 * h->nr * nz))) return rc;
  }
  if (!h->S2 && (rc = dalloc_bytes(h, &h->S2, compact(h) * wsize(h)))) return rc;
  if (!h->blas) {
    rocblas_handle b = nullptr;
    if (rocblas_create_handle(&b) != rocblas_status_success) return fail(h, "axial mesh: rocblas_create_handle failed", -3);
    h->blas = b;
    if (rocblas_set_stream(b, h->stream) != rocblas_status_success)
      return fail(h, "axial mesh: rocblas_set_stream failed", -3);
  }
  h->hz = nullptr;  // a per-mode solve set for another mesh: switched off
  for (int q = 0; q < 4; q++) h->zmesh[q] = zmn[q];
  h->zlam = lam;
  if ((rc = h2d(h, h->d_zmet, mb.data(), mb.size() * sizeof(double)))) return rc;
  h->zmet.m3 = h->d_zmet; h->zmet.c3 = h->d_zmet + L; h->zmet.cm = h->d_zmet + 2 * L; h->zmet.cp = h->d_zmet + 3 * L;
  h->zmet.fm = h->d_zmet + 4 * L; h->zmet.fp = h->d_zmet + 5 * L; h->zmet.zpos = h->d_zmet + 6 * L;
  h->zmet.ws = 2.0 / (h1 * (h1 + h2));
  h->zmet.w2 = 2.0 / (h2 * (h1 + h2));
  std::vector<double> zv(2 * nn);
  std::memcpy(zv.data(), vi.data(), nn * sizeof(double));
  std::memcpy(zv.data() + nn, vf.data(), nn * sizeof(double));
  if ((rc = h2d(h, h->d_zv, zv.data(), zv.size() * sizeof(double)))) return rc;
  if (h->prec == VRH_MIXED) {
    std::vector<float> zf(2 * nn);
    for (size_t i = 0; i < 2 * nn; i++) zf[i] = (float)zv[i];
    if ((rc = h2d(h, h->d_zvf, zf.data(), zf.size() * sizeof(float)))) return rc;
  }
  if ((rc = refresh_mesh(h))) return rc;
  graph_invalidate(h);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

// Per-mode iterative pressure solve for the modes m >= mmin through a
// Helmholtz handle (its operator is the pressure equation's (r, z) part of
// one mode on the same tanh axial mesh); hz == NULL switches it off.
int vrh_set_mode_solver(vrh_t *h, vrh_hz_t *hz, int method, int mmin, double tol, int maxit) {
  if (!h) return -1;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->hz = nullptr;
  h->hz_iters = h->hz_solves = 0;
  h->mode_rc = 0;
  if (!hz) return 0;
  if (method != VRH_HZ_PCG && method != VRH_HZ_BICGSTAB && method != VRH_HZ_LEARNED && method != VRH_HZ_LEARNED_FCG)
    return fail(h, "mode solver: unknown method");
  if (!(tol > 0.0) || maxit < 1) return fail(h, "mode solver: need tol > 0 and maxit >= 1");
  if (mmin < 1) return fail(h, "mode solver: mmin must be >= 1 (the (0,0) mode keeps the reference's pinned direct solve)");
  if (h->prec == VRH_MIXED && mmin < 3)
    return fail(h, "mode solver: the mixed path solves m = 0..2 in FP64 directly; mmin must be >= 3");
  if (h->theta_eigen != VRH_THETA_EXACT_M2) return fail(h, "mode solver: needs the exact theta eigenvalue m^2");
  if (!h->mesh[0].empty()) return fail(h, "mode solver: needs the uniform radial mesh");
  if (h->zmesh[0].empty()) return fail(h, "mode solver: needs an axial mesh (vrh_set_axial_mesh)");
  const int N = h->nt, pos0 = 2 * mmin - 1;
  if (pos0 >= N) return fail(h, "mode solver: mmin beyond the grid's modes");
  if (hz_slots(hz) != N - pos0) return fail(h, "mode solver: the Helmholtz handle needs one slot per position pos >= 2 mmin - 1");
  std::string e;
  if (hz_step_check(hz, h->nr, h->nz, h->rmax, h->zmax, h->zmesh[1].data(), h->zmesh[0].data(), e))
    return fail(h, e.c_str());
  std::vector<int> m(N - pos0);
  for (int p = pos0; p < N; p++) m[p - pos0] = (p + 1) / 2;
  if (vrh_hz_set_modes(hz, m.data())) return fail(h, (std::string("mode solver: ") + hz_error(hz)).c_str());
  h->hz = hz;
  h->hz_method = method;
  h->hz_pos0 = pos0;
  h->hz_tol = tol;
  h->hz_maxit = maxit;
  return 0;
}

int vrh_mode_solver_stats(vrh_t *h, long long *iterations, long long *solves) {
  if (!h) return -1;
  if (iterations) *iterations = h->hz_iters;
  if (solves) *solves = h->hz_solves;
  return 0;
}

// One substep of the step that has reached the current time, with the
// step's scalars computed on the device as vrh_step computes them, so that
// three calls after vrh_set_time(t + dt) equal one vrh_step from t bit for
// bit.  The device time follows the host's: a later vrh_step continues
// from it.
int vrh_substep(vrh_t *h, int irk) {
  if (!h) return -1;
  if (irk < 1 || irk > 3) return fail(h, "irk must be 1..3");
  launch_scalars_at(h, h->ntime);
  if (h->prec == VRH_MIXED) substep_dev<float>(h, irk, nullptr);
  else substep_dev<double>(h, irk, nullptr);
  HIPCHK(h, hipGetLastError());
  h->time_dirty = false;
  if (h->mode_rc) { const int rc = h->mode_rc; h->mode_rc = 0; return rc; }
  return 0;
}

int vrh_phase(vrh_t *h, int phase, int irk) {
  if (!h) return -1;
  if (phase != VRH_PH_HDNBC && phase != VRH_PH_EXPLICIT && phase != VRH_PH_ADI && phase != VRH_PH_POISSON &&
      phase != VRH_PH_PROJECT)
    return fail(h, "unknown phase");
  return substep(h, irk, phase);
}

namespace {
// queue nsteps whole steps; the device copy of ntime advances with them and
// is refreshed from the host value only after vrh_set_time (or creation),
// when the handle is idle, so h_pin is never overwritten under a pending copy
int step_enqueue(vrh_t *h, int nsteps) {
  if (h->time_dirty) {
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->h_pin[0] = h->ntime;
    HIPCHK(h, hipMemcpyAsync(h->d_sc, h->h_pin, sizeof(double), hipMemcpyHostToDevice, h->stream));
    h->time_dirty = false;
  }
  // an axial mesh's z transforms are library GEMMs: launched eagerly
  const bool graph = h->use_graph && !h->kt_on && h->zmesh[0].empty();
  if (graph && !h->graph_ok) { int rc = build_graph(h); if (rc) return rc; }
  for (int s = 0; s < nsteps; s++) {
    h->ntime = h->dt + h->ntime;  // loop.F90:53 (k_scalars does the same on the device)
    if (graph) {
      HIPCHK(h, hipGraphLaunch(h->gexec, h->stream));
    } else {
      int rc = step_body(h);
      if (rc) return rc;
    }
  }
  return 0;
}
}  // namespace

int vrh_step(vrh_t *h, int nsteps) {
  if (!h || nsteps < 0) return -1;
  if (nsteps == 0) return 0;
  int rc = step_enqueue(h, nsteps);
  if (rc) return rc;
  return check_flag(h);
}

int vrh_step_async(vrh_t *h, int nsteps) {
  if (!h || nsteps < 0) return -1;
  return nsteps == 0 ? 0 : step_enqueue(h, nsteps);
}

int vrh_snapshot_begin(vrh_t *h) {
  if (!h) return -1;
  if (h->snap_pending) return fail(h, "a snapshot is pending: vrh_snapshot_wait first");
  const size_t n = haloed(h);
  if (!h->snap_ev) {
    // created on first use, the event last: a failure part way leaves what
    // was made for the next call to complete, never a set taken as whole
    if (!h->d_snap) HIPCHK(h, hipMalloc(&h->d_snap, 4 * n * sizeof(double)));
    if (!h->h_snap) HIPCHK(h, hipHostMalloc(&h->h_snap, 4 * n * sizeof(double), hipHostMallocDefault));
    if (!h->copy_stream) HIPCHK(h, hipStreamCreateWithFlags(&h->copy_stream, hipStreamNonBlocking));
    HIPCHK(h, hipEventCreateWithFlags(&h->snap_ev, hipEventDisableTiming));
  }
  const double *src[4] = {h->f.q1, h->f.q2, h->f.q3, h->f.pres};
  for (int q = 0; q < 4; q++)
    HIPCHK(h, hipMemcpyAsync(h->d_snap + q * n, src[q], n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipEventRecord(h->snap_ev, h->stream));
  HIPCHK(h, hipStreamWaitEvent(h->copy_stream, h->snap_ev, 0));
  HIPCHK(h, hipMemcpyAsync(h->h_snap, h->d_snap, 4 * n * sizeof(double), hipMemcpyDeviceToHost, h->copy_stream));
  h->snap_time = h->ntime;
  h->snap_pending = true;
  return 0;
}

int vrh_snapshot_wait(vrh_t *h, double *q1, double *q2, double *q3, double *pres, double *t) {
  if (!h) return -1;
  if (!h->snap_pending) return fail(h, "no snapshot pending");
  HIPCHK(h, hipStreamSynchronize(h->copy_stream));
  h->snap_pending = false;
  const size_t n = haloed(h);
  double *dst[4] = {q1, q2, q3, pres};
  for (int q = 0; q < 4; q++)
    if (dst[q]) std::memcpy(dst[q], h->h_snap + q * n, n * sizeof(double));
  if (t) *t = h->snap_time;
  return 0;
}

int vrh_set_graph(vrh_t *h, int on) {
  if (!h) return -1;
  h->use_graph = on != 0;
  if (!h->use_graph) graph_invalidate(h);
  return 0;
}

int vrh_synchronize(vrh_t *h) {
  if (!h) return -1;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return check_flag(h);
}

int vrh_poisson_solve(vrh_t *h, const double *rhs, double *phi) {
  if (!h || !rhs || !phi) return -1;
  const size_t C = compact(h);
  // freed on every return (hipFree waits for the device)
  struct Buf {
    double *p = nullptr;
    ~Buf() { if (p) (void)hipFree(p); }
  } rb, pb;
  HIPCHK(h, hipMalloc(&rb.p, C * sizeof(double)));
  HIPCHK(h, hipMalloc(&pb.p, C * sizeof(double)));
  double *d_rhs = rb.p, *d_phi = pb.p;
  HIPCHK(h, hipMemcpyAsync(d_rhs, rhs, C * sizeof(double), hipMemcpyHostToDevice, h->stream));
  if (h->prec == VRH_MIXED) launch_poisson<float>(h, 1, d_rhs, d_phi);
  else launch_poisson<double>(h, 1, d_rhs, d_phi);
  HIPCHK(h, hipGetLastError());
  if (h->mode_rc) {
    const int rc = h->mode_rc;
    h->mode_rc = 0;
    (void)hipStreamSynchronize(h->stream);
    return rc;
  }
  HIPCHK(h, hipMemcpyAsync(phi, d_phi, C * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// partitioned-Thomas coefficients for the current segment lengths
// ---------------------------------------------------------------------------
int build_line_coefs(vrh_t *h) {
  graph_invalidate(h);
  std::vector<double> lcv[4][6][7];
  size_t tot = 0;
  for (int irk = 1; irk <= 3; irk++) {
    host_line_coefs(h, irk, lcv[irk], h->Msw);
    for (int s = 0; s < 6; s++)
      for (int q = 0; q < 7; q++) tot += lcv[irk][s][q].size() + (q < 5 ? kLinePad : 0);
  }
  if (h->d_line) { hipStreamSynchronize(h->stream); hipFree(h->d_line); h->d_line = nullptr; }
  if (h->d_linef) { hipFree(h->d_linef); h->d_linef = nullptr; }
  int rc = dalloc(h, &h->d_line, tot);
  if (rc) return rc;
  const bool fcopy = h->prec == VRH_MIXED;
  if (fcopy && (rc = dalloc_bytes(h, reinterpret_cast<void **>(&h->d_linef), tot * sizeof(float)))) return rc;
  std::vector<double> lb(tot, 0.0);
  size_t off = 0;
  for (int irk = 1; irk <= 3; irk++)
    for (int s = 0; s < 6; s++) {
      const double *ptr[7];
      for (int q = 0; q < 7; q++) {
        std::memcpy(&lb[off], lcv[irk][s][q].data(), lcv[irk][s][q].size() * sizeof(double));
        ptr[q] = h->d_line + off;
        // per-position arrays carry kLinePad zeros: a segment reads its M
        // positions without clamping (rows past the line end are inert)
        off += lcv[irk][s][q].size() + (q < 5 ? kLinePad : 0);
      }
      LineCoef &L = h->lc[irk][s];
      L.A = ptr[0]; L.B = ptr[1]; L.C = ptr[2]; L.Pk = ptr[3]; L.Qk = ptr[4]; L.Pseg = ptr[5]; L.Qseg = ptr[6];
      const float *pf[7];
      for (int q = 0; q < 7; q++) pf[q] = fcopy ? h->d_linef + (ptr[q] - h->d_line) : nullptr;
      L.Af = pf[0]; L.Bf = pf[1]; L.Cf = pf[2]; L.Pkf = pf[3]; L.Qkf = pf[4]; L.Psegf = pf[5]; L.Qsegf = pf[6];
      L.n = (int)lcv[irk][s][0].size();
      L.P = (int)lcv[irk][s][5].size();
    }
  if ((rc = h2d(h, h->d_line, lb.data(), tot * sizeof(double)))) return rc;
  if (fcopy) {
    std::vector<float> lf(tot);
    for (size_t i = 0; i < tot; i++) lf[i] = (float)lb[i];  // round to nearest, as the device conversion
    if ((rc = h2d(h, h->d_linef, lf.data(), tot * sizeof(float)))) return rc;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// autotuner: segment length of each line sweep and the transform tile sizes,
// each timed with HIP events on the live buffers (intermediate arrays only;
// the state q, p, prg is not touched except through idempotent writes of the
// predictor dq, which the next substep recomputes).
// ---------------------------------------------------------------------------
template <typename TW> static double time_sweeps(vrh_t *h, int reps) {
  hipEvent_t e0, e1;
  hipEventCreate(&e0); hipEventCreate(&e1);
  launch_sweeps<TW>(h, 1, 0.0, nullptr);
  hipEventRecord(e0, h->stream);
  for (int r = 0; r < reps; r++) launch_sweeps<TW>(h, 1, 0.0, nullptr);
  hipEventRecord(e1, h->stream);
  hipEventSynchronize(e1);
  float ms = 0.f;
  hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0); hipEventDestroy(e1);
  return ms;
}
template <typename TW> static double time_poisson(vrh_t *h, int reps) {
  hipEvent_t e0, e1;
  hipEventCreate(&e0); hipEventCreate(&e1);
  launch_poisson<TW>(h, 1, nullptr, nullptr);
  hipEventRecord(e0, h->stream);
  for (int r = 0; r < reps; r++) launch_poisson<TW>(h, 1, nullptr, nullptr);
  hipEventRecord(e1, h->stream);
  hipEventSynchronize(e1);
  float ms = 0.f;
  hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0); hipEventDestroy(e1);
  return ms;
}

// The explicit kernel at irk = 1 (rho = 0): its outputs r1..r3 and ru are
// rewritten by the next substep before they are read with a nonzero weight.
template <typename TW> static double time_explicit(vrh_t *h, int reps) {
  hipEvent_t e0, e1;
  hipEventCreate(&e0); hipEventCreate(&e1);
  launch_explicit<TW>(h, 1);
  hipEventRecord(e0, h->stream);
  for (int r = 0; r < reps; r++) launch_explicit<TW>(h, 1);
  hipEventRecord(e1, h->stream);
  hipEventSynchronize(e1);
  float ms = 0.f;
  hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0); hipEventDestroy(e1);
  return ms;
}

extern "C" int vrh_autotune(vrh_t *h, int reps, int *choice) {
  if (!h) return -1;
  if (reps < 1) reps = 3;
  // the per-mode iterative pressure solve (if set) is left out of the timed
  // launches: the tuned shapes are those of the direct kernels it sits beside
  struct HzPause {
    vrh_t *h;
    vrh_hz_t *hz;
    explicit HzPause(vrh_t *h_) : h(h_), hz(h_->hz) { h->hz = nullptr; }
    ~HzPause() { h->hz = hz; }
  } pause(h);
  const bool kt = h->kt_on;
  h->kt_on = false;
  const int ns[6] = {h->nr, h->nr, h->nrc, h->nzc, h->nzc, h->nzc};
  for (int sw = 0; sw < 6; sw++) {
    int best = h->Msw[sw];
    double bestt = 1e30;
    for (int M : kLineSegs) {
      if ((ns[sw] + M - 1) / M > 16) continue;
      h->Msw[sw] = M;
      int rc = build_line_coefs(h);
      if (rc) { h->kt_on = kt; return rc; }
      double t = h->prec == VRH_MIXED ? time_sweeps<float>(h, reps) : time_sweeps<double>(h, reps);
      if (t < bestt) { bestt = t; best = M; }
    }
    h->Msw[sw] = best;
  }
  int rc = build_line_coefs(h);
  if (rc) { h->kt_on = kt; return rc; }
  const bool pow2t = (h->nt & (h->nt - 1)) == 0;
  if (pow2t) {
    int best = 0; double bestt = 1e30;
    for (int lb : {2, 4, 8, 16}) {
      if (lb * h->nt > 4096) continue;
      h->fft_lines = lb;
      double t = h->prec == VRH_MIXED ? time_poisson<float>(h, reps) : time_poisson<double>(h, reps);
      if (t < bestt) { bestt = t; best = lb; }
    }
    h->fft_lines = best;
  }
  {
    int best = 0; double bestt = 1e30;
    for (int tl : {2, 4, 8, 16, 32}) {
      h->dct_lines = tl;
      double t = h->prec == VRH_MIXED ? time_poisson<float>(h, reps) : time_poisson<double>(h, reps);
      if (t < bestt) { bestt = t; best = tl; }
    }
    h->dct_lines = best;
  }
  {
    int best = 0; double bestt = 1e30;
    for (int v : {0, 1, 2, 3}) {
      h->explicit_variant = v;
      double t = h->prec == VRH_MIXED ? time_explicit<float>(h, reps) : time_explicit<double>(h, reps);
      if (t < bestt) { bestt = t; best = v; }
    }
    h->explicit_variant = best;
  }
  graph_invalidate(h);
  h->kt_on = kt;
  if (choice) {
    for (int sw = 0; sw < 6; sw++) choice[sw] = h->Msw[sw];
    choice[6] = h->fft_lines;
    choice[7] = h->dct_lines;
    choice[8] = h->explicit_variant;
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

// ---------------------------------------------------------------------------
// per-kernel event timer
// ---------------------------------------------------------------------------
const char *const kKernelNames[KT_COUNT] = {"hdnbc", "explicit", "theta_pcr", "line_r", "line_z", "divfft",
                                            "dct_fwd", "radial", "dct_inv", "ifft", "project", "particles"};

KTimer::KTimer(vrh_s *h_, int id_) : h(h_), id(id_), slot(-1) {
  if (!h || !h->kt_on) return;
  if ((size_t)(2 * h->kt_used + 2) > h->kt_pool.size()) ktimer_harvest(h);
  slot = h->kt_used++;
  h->kt_ids[slot] = id;
  hipEventRecord(h->kt_pool[2 * slot], h->stream);
}
KTimer::~KTimer() {
  if (slot >= 0) hipEventRecord(h->kt_pool[2 * slot + 1], h->stream);
}

void ktimer_harvest(vrh_t *h) {
  if (h->kt_used == 0) return;
  hipStreamSynchronize(h->stream);
  for (int s = 0; s < h->kt_used; s++) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, h->kt_pool[2 * s], h->kt_pool[2 * s + 1]) == hipSuccess) {
      h->kt_ms[h->kt_ids[s]] += ms;
      h->kt_cnt[h->kt_ids[s]] += 1;
    }
  }
  h->kt_used = 0;
}

extern "C" int vrh_ktimer_enable(vrh_t *h, int on) {
  if (!h) return -1;
  if (h->kt_pool.empty()) {
    h->kt_pool.resize(4096);
    h->kt_ids.resize(2048);
    for (auto &e : h->kt_pool) HIPCHK(h, hipEventCreate(&e));
  }
  ktimer_harvest(h);
  for (int k = 0; k < KT_COUNT; k++) { h->kt_ms[k] = 0.0; h->kt_cnt[k] = 0; }
  h->kt_on = on != 0;
  return 0;
}

extern "C" int vrh_ktimer_read(vrh_t *h, int idx, char *name, size_t len, double *total_ms, long *count) {
  if (!h) return -1;
  if (idx < 0) return KT_COUNT;
  if (idx >= KT_COUNT) return -1;
  ktimer_harvest(h);
  if (name && len) snprintf(name, len, "%s", kKernelNames[idx]);
  if (total_ms) *total_ms = h->kt_ms[idx];
  if (count) *count = h->kt_cnt[idx];
  return 0;
}

// ---------------------------------------------------------------------------
// Ensembles: independent rings (each its own handle, grid, Re and stream)
// advanced together.  One RK3 step of every member is captured into a single
// HIP graph, forked from an origin stream into the members' streams and
// joined back, so a step of the whole ensemble is one graph launch and the
// members' kernels run concurrently.  No reference counterpart (the
// reference runs one ring per MPI job).
// ---------------------------------------------------------------------------
struct vrh_ens_s {
  std::vector<vrh_t *> m;
  std::vector<long> epochs;
  hipStream_t origin = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  // graph[0]: one step of every member; graph[1]: `batch` steps of every
  // member (built when batch > 1)
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  int batch = 1;
  std::string err;
};

namespace {
int efail(vrh_ens_t *e, const std::string &msg, int code = -1) {
  if (e) e->err = msg;
  return code;
}
#define ECHK(e, call)                                                               \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) return efail(e, std::string(#call) + ": " + hipGetErrorString(e_), -3); \
  } while (0)

void ens_release_graph(vrh_ens_t *e) {
  for (int g = 0; g < 2; g++) {
    if (e->gexec[g]) hipGraphExecDestroy(e->gexec[g]);
    if (e->graph[g]) hipGraphDestroy(e->graph[g]);
    e->gexec[g] = nullptr;
    e->graph[g] = nullptr;
  }
}

// Capture `nsteps` consecutive steps of every member into graph slot g:
// forked from the origin stream into each member's stream, the member's
// steps in order (each advances its device time), joined back.
int ens_capture(vrh_ens_t *e, int g, int nsteps) {
  ECHK(e, hipStreamBeginCapture(e->origin, hipStreamCaptureModeThreadLocal));
  ECHK(e, hipEventRecord(e->fork, e->origin));
  int rc = 0;
  for (size_t q = 0; q < e->m.size() && rc == 0; q++) {
    vrh_t *h = e->m[q];
    if (hipStreamWaitEvent(h->stream, e->fork, 0) != hipSuccess) { rc = -3; break; }
    h->serial_sweeps = true;
    for (int s = 0; s < nsteps && rc == 0; s++) rc = step_body(h);
    h->serial_sweeps = false;
    if (rc) { e->err = "member " + std::to_string(q) + ": " + h->err; break; }
    if (hipEventRecord(e->join[q], h->stream) != hipSuccess ||
        hipStreamWaitEvent(e->origin, e->join[q], 0) != hipSuccess) rc = -3;
  }
  hipGraph_t gr = nullptr;
  hipError_t ee = hipStreamEndCapture(e->origin, &gr);
  if (rc) { if (gr) hipGraphDestroy(gr); return rc; }
  ECHK(e, ee);
  e->graph[g] = gr;
  ECHK(e, hipGraphInstantiate(&e->gexec[g], e->graph[g], nullptr, nullptr, 0));
  return 0;
}

int ens_build(vrh_ens_t *e) {
  ens_release_graph(e);
  int rc = ens_capture(e, 0, 1);
  if (rc == 0 && e->batch > 1) rc = ens_capture(e, 1, e->batch);
  if (rc) return rc;
  for (size_t q = 0; q < e->m.size(); q++) e->epochs[q] = e->m[q]->epoch;
  return 0;
}
}  // namespace

extern "C" {

int vrh_ensemble_create(vrh_t *const *members, int n, vrh_ens_t **out) {
  if (!members || n < 1 || !out) return -1;
  *out = nullptr;
  vrh_ens_t *e = new (std::nothrow) vrh_ens_t();
  if (!e) return -1;
  *out = e;
  for (int q = 0; q < n; q++) {
    if (!members[q]) return efail(e, "null member handle");
    for (int p = 0; p < q; p++)
      if (members[p] == members[q]) return efail(e, "a handle appears twice in the ensemble");
    e->m.push_back(members[q]);
  }
  e->epochs.assign(n, -1);
  ECHK(e, hipStreamCreateWithFlags(&e->origin, hipStreamNonBlocking));
  ECHK(e, hipEventCreateWithFlags(&e->fork, hipEventDisableTiming));
  e->join.assign(n, nullptr);
  for (int q = 0; q < n; q++) ECHK(e, hipEventCreateWithFlags(&e->join[q], hipEventDisableTiming));
  return 0;
}

int vrh_ensemble_step(vrh_ens_t *e, int nsteps) {
  if (!e || nsteps < 0) return -1;
  if (nsteps == 0) return 0;
  bool stale = !e->gexec[0] || (e->batch > 1 && !e->gexec[1]);
  for (size_t q = 0; q < e->m.size(); q++) {
    vrh_t *h = e->m[q];
    if (h->kt_on) return efail(e, "member " + std::to_string(q) + " has the kernel timer enabled");
    if (!h->zmesh[0].empty())
      return efail(e, "member " + std::to_string(q) + " has an axial mesh (stepped eagerly, not in a graph)");
    ECHK(e, hipStreamSynchronize(h->stream));
    if (h->epoch != e->epochs[q]) stale = true;
  }
  if (stale) { int rc = ens_build(e); if (rc) return rc; }
  for (vrh_t *h : e->m) {
    h->h_pin[0] = h->ntime;
    ECHK(e, hipMemcpyAsync(h->d_sc, h->h_pin, sizeof(double), hipMemcpyHostToDevice, e->origin));
    h->time_dirty = false;
  }
  // whole batches first (one launch advances every member `batch` steps),
  // then single steps; the host times advance step by step as the
  // device's do
  int left = nsteps;
  while (left > 0) {
    const int g = (e->batch > 1 && left >= e->batch) ? 1 : 0;
    const int n = g ? e->batch : 1;
    for (int s = 0; s < n; s++)
      for (vrh_t *h : e->m) h->ntime = h->dt + h->ntime;  // loop.F90:53
    ECHK(e, hipGraphLaunch(e->gexec[g], e->origin));
    left -= n;
  }
  ECHK(e, hipStreamSynchronize(e->origin));
  for (size_t q = 0; q < e->m.size(); q++) {
    int rc = check_flag(e->m[q]);
    if (rc) return efail(e, "member " + std::to_string(q) + ": " + e->m[q]->err, rc);
  }
  return 0;
}

int vrh_ensemble_set_batch(vrh_ens_t *e, int steps_per_launch) {
  if (!e || steps_per_launch < 1) return -1;
  if (steps_per_launch != e->batch) {
    ECHK(e, hipStreamSynchronize(e->origin));
    ens_release_graph(e);
    e->batch = steps_per_launch;
  }
  return 0;
}

int vrh_ensemble_last_error(vrh_ens_t *e, char *buf, size_t len) {
  if (!buf || len == 0) return -1;
  snprintf(buf, len, "%s", e ? e->err.c_str() : "null ensemble");
  return 0;
}

int vrh_ensemble_destroy(vrh_ens_t *e) {
  if (!e) return -1;
  if (e->origin) hipStreamSynchronize(e->origin);
  ens_release_graph(e);
  if (e->fork) hipEventDestroy(e->fork);
  for (auto ev : e->join) if (ev) hipEventDestroy(ev);
  if (e->origin) hipStreamDestroy(e->origin);
  delete e;
  return 0;
}

}  // extern "C"

// Apply a tuning previously returned by vrh_autotune (same 9 entries) without
// timing anything; 0 entries keep the current value.
extern "C" int vrh_set_tuning(vrh_t *h, const int *choice) {
  if (!h || !choice) return -1;
  const int ns[6] = {h->nr, h->nr, h->nrc, h->nzc, h->nzc, h->nzc};
  for (int sw = 0; sw < 6; sw++) {
    const int M = choice[sw];
    if (M == 0) continue;
    bool known = false;
    for (int m : kLineSegs) known = known || m == M;
    if (!known) return fail(h, "segment length must be 16, 24, 32, 40, 48, 64 or 72");
    if ((ns[sw] + M - 1) / M > 16) return fail(h, "segment length too short for the line length");
  }
  if (choice[6] != 0 && choice[6] != 2 && choice[6] != 4 && choice[6] != 8 && choice[6] != 16)
    return fail(h, "fft lines per block must be 2, 4, 8 or 16");
  if (choice[7] != 0 && choice[7] != 2 && choice[7] != 4 && choice[7] != 8 && choice[7] != 16 && choice[7] != 32)
    return fail(h, "dct lines per block must be 2, 4, 8, 16 or 32");
  if (choice[8] < -1 || choice[8] > 3) return fail(h, "explicit form must be -1, 0, 1, 2 or 3");
  for (int sw = 0; sw < 6; sw++) if (choice[sw]) h->Msw[sw] = choice[sw];
  if (choice[6]) h->fft_lines = choice[6];
  if (choice[7]) h->dct_lines = choice[7];
  h->explicit_variant = choice[8];
  int rc = build_line_coefs(h);
  if (rc) return rc;
  graph_invalidate(h);
  return 0;
}

// Vortex-core statistics (run.F90:238-388) reduced on the device; out[10] =
// vortmax, swrlmax, then stat(1:8).
extern "C" int vrh_analysis(vrh_t *h, double *out) {
  if (!h || !out) return -1;
  const int nblk = kAnalysisBlocks;
  // every block writes its ten partials and the sum kernel writes the ten
  // results, so the buffer needs no clearing between calls
  if (!h->d_anl) HIPCHK(h, hipMalloc(&h->d_anl, sizeof(double) * (nblk * 10 + 16)));
  double *part = h->d_anl;
  launch_analysis(h, part, nblk, part + nblk * 10);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipMemcpyAsync(out, part + nblk * 10, 10 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

// The same reduction with the block partials in the per-device scratch
// array (set up at vrh_create) and the results left in device memory:
// nothing is allocated and nothing waits, so the call may be captured.
extern "C" int vrh_analysis_noalloc(vrh_t *h, double *dev_out) {
  if (!h || !dev_out) return -1;
  double *part = analysis_scratch(h->device);
  if (!part) return fail(h, "vrh_analysis_noalloc: no scratch address for the handle's device");
  launch_analysis(h, part, kAnalysisBlocks, dev_out);
  HIPCHK(h, hipGetLastError());
  return 0;
}
