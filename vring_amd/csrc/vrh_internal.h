// vrh_internal.h -- shared declarations of libvrh (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "vrh.h"
#include "vrh_learn.h"

#define VRH_HD __host__ __device__ __forceinline__

// Grid constants and strides.  Haloed 3-D fields use the reference's
// Fortran layout q(nt, 0:nr+2, 0:nz+2) (theta fastest); compact fields
// (prg components, spectral arrays) use (nt, nr, nz).
struct VGrid {
  int nt, nr, nz, nrc, nzc;
  long sj, sk;  // haloed strides: sj = nt, sk = nt*(nr+3)
  long ncomp;   // nt*nr*nz
  double dx1, dx2, dx3, dx1q, dx2q, dx3q, axsym, dt, Re, slip;
};
VRH_HD long hidx(const VGrid &g, int i, int j, int k) { return (long)(i - 1) + g.sj * j + g.sk * k; }
VRH_HD long cidx(const VGrid &g, int i, int j, int k) {
  return (long)(i - 1) + (long)g.nt * ((long)(j - 1) + (long)g.nr * (long)(k - 1));
}

// Radial metric vectors (momentum_coef.F90:24-69), each nr+3 long, indexed by
// the Fortran index j.
struct VCoef {
  const double *rm, *rc, *h1d11, *h1d22d, *h1d22n, *h1dare, *h2d11, *h2d22m, *h2d22p, *h2dare,
      *s1t11, *s1t2n, *s1t2s, *s1t2c, *s2t11, *s2t2p, *s2t2m, *s2t2c, *s3t2n, *s3t2s, *qc,
      *dmr, *dcr;  // radial metric factors at centres and faces (mesh_tools.F90:325-330)
};
constexpr int kNumVec = 23;  // vectors of VCoef, in that order
struct VRK {
  double g, r, a;  // gamma, rho, alpha of the substep (para_tools.F90:103-113)
};
// Axial mesh metric (vrh_set_axial_mesh), each array nz+3 long and indexed by
// the Fortran index k = 0..nz+2: m3 = dx3/dmz and c3 = dx3/dcz (the z
// differences at centres and faces); the second-difference weights of a
// centred variable cm = m3(k) c3(k), cp = m3(k) c3(k+1) and of a face
// variable fm = c3(k) m3(k-1), fp = c3(k) m3(k); zpos the cell centres.
// ws, w2: the one-sided inlet row of a centred variable (wall value at
// distance h1 = dz(1)/2, next centre at h2).  m3 == nullptr: the uniform
// mesh of mesh_tools.F90:325-346.
struct ZMet {
  const double *m3, *c3, *cm, *cp, *fm, *fp, *zpos;
  double ws, w2;
};

// Line-independent coefficients of one tridiagonal sweep for the partitioned
// solver: forward g_p = A_p r_p + B_p g_{p-1}; backward x_p = g_p - C_p x_{p+1};
// Pk/Qk prefix products inside each segment of length M, Pseg/Qseg their
// segment totals.
// A, B, C, Pk, Qk: n positions followed by kLinePad zeros (>= the longest
// segment), so that a segment may read all its positions unclamped.
constexpr int kLinePad = 72;
// segment lengths of the partitioned line solves (at most 16 segments per
// line: one wave each in a block of up to 1024 threads)
constexpr int kLineSegs[] = {16, 24, 32, 40, 48, 64, 72};
struct LineCoef {
  const double *A, *B, *C, *Pk, *Qk, *Pseg, *Qseg;
  // the same arrays rounded to float (mixed path): the FP32 sweeps read
  // half the scalar registers' worth and convert nothing
  const float *Af, *Bf, *Cf, *Pkf, *Qkf, *Psegf, *Qsegf;
  int n, P;
};

enum SweepMode { SW_R1 = 0, SW_R3 = 1, SW_R2 = 2, SW_Z1 = 3, SW_Z3 = 4, SW_Z2 = 5 };

// Work type TW (double in FP64 mode, float in mixed mode); state type TS is
// always double.
struct Fields {
  double *q1, *q2, *q3, *dq1, *dq2, *dq3, *pres, *dph;  // haloed, double
  void *ru1, *ru2, *ru3, *r1, *r2, *r3;                 // haloed, TW
  // compact, double in both paths: the accumulated pressure gradient of
  // uppres is part of the projection correction, which stays FP64
  double *prg1, *prg2, *prg3;
  void *S, *ib;                                         // compact spectral + radial pivots, TW
  double *S64, *ib64;                                   // low modes (5, nr, nz) in mixed mode
  double *csort, *dq3no, *corr3n;                       // planes (nt, nr+2)
  int *flag;                                            // NaN flag
};

struct PartBatch {
  long cap = 0, n = 0;
  double st = 0.0;
  int nsub = 1;
  double *u = nullptr;  // nvar rows of cap (structure of arrays)
  double *ddt2 = nullptr;  // nvar rows of cap, previous rate (RK3 schemes)
  // bytes per element of the ddt2 rows: 8, or 4 for the trilinear scheme in
  // mixed precision (its rates are formed from the FP32 interpolation; the
  // state update stays FP64); ddt2 then holds float rows
  int d2w = 8;
  double *fit = nullptr;  // LSQ schemes: 39 rows of cap, per-component fits of a substep
  int *alive = nullptr;
  long *d_n = nullptr;  // used slots [0, n) on the device (kernels read it)
  // inlet injection and compaction (allocated on first use)
  double *inflow = nullptr;  // accumulated inlet flux per r row (nr + 1)
  double *u_tmp = nullptr, *ddt2_tmp = nullptr;
  int *alive_tmp = nullptr;
  long *work = nullptr;      // scan scratch
  long ncall = 0;            // injection calls so far (random stream index)
  // Physical slots are kept sorted by the particles' cells (upload,
  // compaction, vrh_particles_sort), so that the interpolation gathers of a
  // wavefront meet few cache lines; slot[p] is the slot that physical slot p
  // holds in the order the API shows (upload, download, compaction order).
  int *slot = nullptr;
  int *slot_tmp = nullptr, *lflag = nullptr, *lrank = nullptr;  // compaction renumbering
  unsigned *skeys = nullptr;  // sort keys and values, in and out (4 x cap)
  void *stmp = nullptr;       // radix sort scratch
  size_t stmp_bytes = 0;
};

// Kernel ids for the per-kernel event timer.
enum KernelId {
  KT_HDNBC = 0, KT_EXPLICIT, KT_THETA, KT_LINE_R, KT_LINE_Z, KT_DIVFFT, KT_DCT_FWD, KT_RADIAL, KT_DCT_INV,
  KT_IFFT, KT_PROJECT, KT_PARTICLES, KT_COUNT
};
extern const char *const kKernelNames[KT_COUNT];

struct vrh_s;
// RAII bracket of one kernel launch with HIP events on the handle's stream
// (active only while the timer is enabled).
struct KTimer {
  vrh_s *h;
  int id;
  int slot;
  KTimer(vrh_s *h_, int id_);
  ~KTimer();
};

struct vrh_s {
  int nt, nr, nz, nrc, nzc, prec, theta_eigen;
  int device = 0;  // HIP device current at creation (the handle's GPU)
  double rmax, zmax, Re, dt, slip, ntime;
  VGrid g;
  VRK rk[4];
  hipStream_t stream;
  // side streams for the independent component chains of the line sweeps
  // (forked from and joined back into `stream`; captured with the step)
  // side streams: [0], [1] the sweep chains and the pressure solve's FP64
  // low-mode chain, [2] the particle stage of a substep (beside the fluid
  // kernels up to the projection, which is the only kernel that writes q)
  hipStream_t side[3] = {nullptr, nullptr, nullptr};
  bool serial_sweeps = false;  // set while captured into an ensemble graph
  // one fork and two join events per substep: a captured step records each
  // event once (an event recorded twice in one capture keeps only one of
  // the dependency sets)
  hipEvent_t fork_ev[4] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t join_ev[4][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
  hipEvent_t part_ev[4][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
  // fork and join of the pressure solve's FP64 low-mode chain, per substep
  hipEvent_t poi_ev[4][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
  std::string err;
  // host copies of the radial vectors
  std::vector<double> hv[kNumVec];
  double *d_vec;  // device block holding the kNumVec vectors
  // radial mesh set by vrh_set_radial_mesh (rm, rc, dmr, dcr over j = 0..nr+1);
  // empty: the uniform mesh with dmr = dcr = 1
  std::vector<double> mesh[4];
  // axial mesh set by vrh_set_axial_mesh (zm, zc, dmz, dcz over k = 0..nz+2);
  // empty: the uniform mesh.  zlam: eigenvalues of the pressure solve's z
  // operator (ascending, zlam[0] = 0); d_zmet: the ZMet arrays; d_zv: the z
  // transforms V^-1 then V (nz x nz each, double), d_zvf: their float copies
  // (mixed precision); S2: the spectral work array's twin (out-of-place
  // transforms), S2_64: the same for the FP64 low modes
  std::vector<double> zmesh[4];
  std::vector<double> zlam;
  ZMet zmet{};
  double *d_zmet = nullptr, *d_zv = nullptr;
  float *d_zvf = nullptr;
  void *S2 = nullptr;
  double *S2_64 = nullptr;
  void *blas = nullptr;  // rocBLAS handle of the z transforms (on `stream`)
  // per-mode iterative pressure solve (vrh_set_mode_solver): theta
  // positions pos >= hz_pos0 (modes m >= mmin) are solved by the Helmholtz
  // handle hz instead of the direct z transform + radial solve; hz_f is the
  // double staging copy of their right-hand sides taken before the direct
  // path overwrites the spectral array
  vrh_hz_t *hz = nullptr;
  int hz_method = 0, hz_maxit = 0, hz_pos0 = 0;
  double hz_tol = 0.0;
  long long hz_iters = 0, hz_solves = 0;
  // error of the last in-step mode solve or rocBLAS z transform (err holds
  // the message), returned by the substep or step that met it
  int mode_rc = 0;
  VCoef coef;
  // line coefficients: [irk 1..3][sweep 0..5]
  double *d_line;
  float *d_linef = nullptr;  // float copy of d_line (mixed path)
  LineCoef lc[4][6];
  int Msw[6];        // segment length M of each sweep (partitioned Thomas)
  int fft_lines = 0;  // theta lines per block of the FFT kernels (0: default)
  int dct_lines = 0;  // z lines per block of the DCT kernels (0: default)
  int explicit_variant = -1;  // 0: plane per block, 1: z marching, 2/3: one wave per r row, -1: default
  // LDS a block may take beyond 64 KB (after the kernels' opt-in) and LDS per CU
  int lds_block = 65536, lds_cu = 65536;
  int theta_levels[4];  // PCR levels of the theta solve per substep
  // transform tables (double): theta twiddles (nt/2 complex), z twiddles
  // (nz/2 complex), z post-twiddles (nz complex)
  double *d_twt, *d_twz, *d_ez;
  double *d_aab, *d_ccb;  // Poisson radial base coefficients (nr+1)
  float *d_aabf = nullptr, *d_ccbf = nullptr;  // the same rounded to float (FP32 radial solves)
  Fields f;
  double cor_red;  // sum qc^2 da (hdnbc mass correction denominator)
  int jlim;        // rc(j) <= 0.4999 for j <= jlim (inlet rows)
  std::vector<double> da;
  // particles
  int pscheme, nbatch;
  std::vector<PartBatch> pb;
  double *d_mint;  // 9 x 27 least-squares pseudo-inverse
  double h_mint[243];  // host copy (passed by value to the fit kernels)
  // kernel timer
  bool kt_on = false;
  std::vector<hipEvent_t> kt_pool;
  std::vector<int> kt_ids;  // kernel id per used event pair
  int kt_used = 0;
  double kt_ms[KT_COUNT] = {0};
  long kt_cnt[KT_COUNT] = {0};
  // step scalars on the device: [0] ntime, [1..3] c3s of substeps 1..3,
  // [4..6] mass-correction factor cor of substeps 1..3; inlet_area is the
  // sum of da over the inlet rows (both theta and r)
  double *d_sc = nullptr;
  double *h_pin = nullptr;  // pinned host staging (time)
  bool time_dirty = true;   // host ntime set outside stepping: upload before the next step
  double inlet_area = 0.0;
  // field snapshot for asynchronous export: device copy of q1, q2, q3, pres
  // taken on the compute stream, drained to pinned host memory on a copy
  // stream while stepping continues
  double *d_snap = nullptr, *h_snap = nullptr;
  double *d_anl = nullptr;  // vrh_analysis block partials and result (first use)
  hipStream_t copy_stream = nullptr;
  hipEvent_t snap_ev = nullptr;
  bool snap_pending = false;
  double snap_time = 0.0;
  // one RK3 step captured as a HIP graph (rebuilt when launch shapes,
  // particles or the timer change)
  bool use_graph = true;
  bool graph_ok = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  long epoch = 0;  // bumped whenever captured launch parameters change
};
void graph_invalidate(vrh_t *h);
// queue one whole RK3 step whose scalars come from d_sc (captured or eager)
int step_body(vrh_t *h);
void ktimer_harvest(vrh_t *h);

// host math (vrh_host.cpp)
void host_mesh_vectors(vrh_t *h);
void host_line_coefs(vrh_t *h, int irk, std::vector<double> out[6][7], const int M[6]);
int build_line_coefs(vrh_t *h);
void host_radial_row(const vrh_t *h, int j, double &aa, double &cc, double &dd);
void host_radial_pivots(vrh_t *h, std::vector<double> &ib, std::vector<double> &ib64);
// Pressure-solve z transforms on an axial mesh with z differences m3 (centres)
// and c3 (faces), k = 0..nz+2: vi = V^-1, vf = V (row-major nz x nz) and lam
// (ascending, lam[0] = 0) with (-D_z G_z) V = V diag(lam).  Returns nonzero
// (and err) if the decomposition fails its residual or orthogonality check.
// Helmholtz-handle hooks used by the in-step mode solve (vrh_learn.hip).
int hz_slots(const vrh_hz_t *h);
double *hz_step_input(vrh_hz_t *h);
const double *hz_step_output(vrh_hz_t *h);
const char *hz_error(const vrh_hz_t *h);
int hz_step_check(const vrh_hz_t *h, int nr, int nz, double rmax, double zmax, const double *zc, const double *zm,
                  std::string &err);
int hz_step_solve(vrh_hz_t *h, hipStream_t caller, int method, double tol, int maxit, long long *iters);
int host_axial_transforms(const double *m3, const double *c3, int nz, std::vector<double> &vi,
                          std::vector<double> &vf, std::vector<double> &lam, std::string &err);
double host_profile(double t);
void host_mint(vrh_t *h, double *mint);

// kernel launchers (vrh_kernels.hip); TW = double or float
template <typename TW> void launch_hdnbc(vrh_t *h, int irk, double cor, const double *cor_dev);
void launch_scalars(vrh_t *h);
void launch_scalars_at(vrh_t *h, double t);
void launch_analysis(vrh_t *h, double *part, int nblk, double *out);
// Partial blocks of the analysis reduction (both entry points), and the
// allocation-free variant's per-device scratch (vrh_kernels.hip): set up
// once per device with that device current, then looked up by device id.
constexpr int kAnalysisBlocks = 1024;
constexpr int kMaxDevices = 64;
int analysis_scratch_init(int dev);
double *analysis_scratch(int dev);
template <typename TW> void launch_explicit(vrh_t *h, int irk);
void lds_optin(vrh_t *h);
// hdnbc: also launch the outlet-row kernel of hdnbc at the head of the q3
// sweep chain (whole steps), with the mass correction cor or *cor_dev
template <typename TW> void launch_sweeps(vrh_t *h, int irk, double c3s, const double *c3s_dev, bool hdnbc = false,
                                          double cor = 0.0, const double *cor_dev = nullptr);
template <typename TW> void launch_poisson(vrh_t *h, int irk, const double *ext_rhs, double *phi_out);
template <typename TW> void launch_project(vrh_t *h, int irk);
// the particle stage of substep irk on stream st (nothing launched where
// the scheme has no stage at irk); returns whether anything was launched
bool launch_particles(vrh_t *h, int irk, hipStream_t st);
// frees every particle batch (after the handle's stream drains); no batches remain
void particles_release(vrh_t *h);
