"""ctypes binding of libvrh (include/vrh.h) and a thin Solver class.

Arrays cross the boundary in the reference's Fortran layouts:
haloed fields (ntheta, nr+3, nz+3) with index (i-1, j, k), prg (ntheta, nr,
nz, 3), Poisson arrays (ntheta, nr, nz).  There is no CPU fallback: every
call runs on the GPU through libvrh, and loading fails loudly if the library
is missing."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libvrh.so")

PRECISION = {"fp64": 0, "mixed": 1}
THETA = {"exact_m2": 0, "consistent": 1}
PHASE = {"hdnbc": 2, "explicit": 3, "adi": 6, "poisson": 8, "project": 9}
PART_SCHEME = {"lsq_euler": 0, "lsq_rk3": 1, "trilinear_rk3": 2}


class VrhGrid(ctypes.Structure):
    _fields_ = [("ntheta", ctypes.c_int), ("nr", ctypes.c_int), ("nz", ctypes.c_int),
                ("rmax", ctypes.c_double), ("zmax", ctypes.c_double)]


class VrhParams(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("Re", ctypes.c_double), ("slip", ctypes.c_double),
                ("theta_eigen", ctypes.c_int)]


EXPORTS = ["vrh_create", "vrh_destroy", "vrh_last_error", "vrh_upload_fields", "vrh_download_fields",
           "vrh_upload_state", "vrh_download_state", "vrh_state_size", "vrh_set_time", "vrh_get_time", "vrh_set_radial_mesh",
           "vrh_set_axial_mesh",
           "vrh_substep",
           "vrh_step", "vrh_phase", "vrh_poisson_solve", "vrh_synchronize", "vrh_device_ptr",
           "vrh_ring_init", "vrh_cfl_dt", "vrh_particles_create", "vrh_particles_upload",
           "vrh_particles_download", "vrh_particles_step", "vrh_ktimer_enable", "vrh_ktimer_read", "vrh_autotune",
           "vrh_set_graph", "vrh_set_tuning", "vrh_analysis", "vrh_step_async", "vrh_snapshot_begin",
           "vrh_snapshot_wait", "vrh_particles_compact", "vrh_particles_sort", "vrh_particles_count",
           "vrh_particles_inject", "vrh_ensemble_create", "vrh_ensemble_step", "vrh_ensemble_set_batch",
           "vrh_ensemble_last_error",
           "vrh_ensemble_destroy", "vrh_analysis_noalloc"]

_lib = None


def load_library(path=LIBPATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libvrh not built: {path} missing (run __graft_entry__.build())")
    L = ctypes.CDLL(path)
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    lp = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
    L.vrh_create.argtypes = [ctypes.POINTER(VrhGrid), ctypes.POINTER(VrhParams), I, ctypes.POINTER(P)]
    L.vrh_destroy.argtypes = [P]
    L.vrh_last_error.argtypes = [P, ctypes.c_char_p, ctypes.c_size_t]
    L.vrh_upload_fields.argtypes = [P, P, P, P, P]
    L.vrh_download_fields.argtypes = [P, P, P, P, P]
    L.vrh_upload_state.argtypes = [P, ctypes.c_char_p, dp]
    L.vrh_download_state.argtypes = [P, ctypes.c_char_p, dp]
    L.vrh_state_size.argtypes = [ctypes.POINTER(VrhGrid), ctypes.c_char_p, ctypes.POINTER(ctypes.c_long)]
    L.vrh_set_time.argtypes = [P, D]
    L.vrh_get_time.argtypes = [P, ctypes.POINTER(D)]
    L.vrh_set_radial_mesh.argtypes = [P, dp, dp, dp, dp]
    L.vrh_set_axial_mesh.argtypes = [P, dp, dp, dp, dp]
    L.vrh_substep.argtypes = [P, I]
    L.vrh_step.argtypes = [P, I]
    L.vrh_phase.argtypes = [P, I, I]
    L.vrh_poisson_solve.argtypes = [P, dp, dp]
    L.vrh_synchronize.argtypes = [P]
    L.vrh_device_ptr.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(P)]
    L.vrh_ring_init.argtypes = [ctypes.POINTER(VrhGrid), I, dp, D, I, I, P, dp, dp, dp]
    L.vrh_cfl_dt.argtypes = [ctypes.POINTER(VrhGrid), dp, dp, dp, D, D, ctypes.POINTER(D)]
    L.vrh_particles_create.argtypes = [P, I, lp, dp, ip, I]
    L.vrh_particles_upload.argtypes = [P, I, ctypes.c_long, dp, ip]
    L.vrh_particles_download.argtypes = [P, I, ctypes.c_long, dp, ip]
    L.vrh_particles_step.argtypes = [P, I]
    L.vrh_autotune.argtypes = [P, I, ip]
    L.vrh_ktimer_enable.argtypes = [P, I]
    L.vrh_ktimer_read.argtypes = [P, I, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(D),
                                  ctypes.POINTER(ctypes.c_long)]
    L.vrh_set_graph.argtypes = [P, I]
    L.vrh_analysis.argtypes = [P, dp]
    L.vrh_analysis_noalloc.argtypes = [P, P]
    L.vrh_step_async.argtypes = [P, I]
    L.vrh_snapshot_begin.argtypes = [P]
    L.vrh_snapshot_wait.argtypes = [P, dp, dp, dp, dp, ctypes.POINTER(ctypes.c_double)]
    L.vrh_particles_compact.argtypes = [P, I]
    L.vrh_particles_sort.argtypes = [P, I]
    L.vrh_particles_count.argtypes = [P, I, ctypes.POINTER(ctypes.c_long)]
    L.vrh_particles_inject.argtypes = [P, I, D, I, ctypes.c_ulonglong, ctypes.POINTER(ctypes.c_long)]
    L.vrh_set_tuning.argtypes = [P, ip]
    L.vrh_ensemble_create.argtypes = [ctypes.POINTER(P), I, ctypes.POINTER(P)]
    L.vrh_ensemble_step.argtypes = [P, I]
    L.vrh_ensemble_set_batch.argtypes = [P, I]
    L.vrh_ensemble_last_error.argtypes = [P, ctypes.c_char_p, ctypes.c_size_t]
    L.vrh_ensemble_destroy.argtypes = [P]
    _lib = L
    return L


def state_size(ntheta, nr, nz, name):
    """Number of doubles vrh_upload_state/vrh_download_state move for `name`
    (host query, no device call)."""
    L = load_library()
    n = ctypes.c_long(0)
    if L.vrh_state_size(ctypes.byref(VrhGrid(ntheta, nr, nz, 1.0, 1.0)), name.encode(), ctypes.byref(n)) != 0:
        raise KeyError(name)
    return n.value


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))


def ring_initial_condition(ntheta, nr, nz, rmax, zmax, rings, pert_amp=0.0, pert_modes=(0, -1),
                           phases=None):
    """Haloed (q1, q2, q3) of Gaussian-core rings [(R, z0, sigma, Gamma), ...]."""
    L = load_library()
    g = VrhGrid(ntheta, nr, nz, rmax, zmax)
    shape = (ntheta, nr + 3, nz + 3)
    n = int(np.prod(shape))
    q1, q2, q3 = np.zeros(n), np.zeros(n), np.zeros(n)
    rr = np.ascontiguousarray(np.asarray(rings, dtype=np.float64).ravel())
    ph = None
    if phases is not None:
        ph = np.ascontiguousarray(np.asarray(phases, dtype=np.float64))
    rc = L.vrh_ring_init(ctypes.byref(g), len(rings), rr, pert_amp, int(pert_modes[0]), int(pert_modes[1]),
                         ph.ctypes.data if ph is not None else None, q1, q2, q3)
    if rc:
        raise RuntimeError("vrh_ring_init failed")
    return tuple(a.reshape(shape, order="F") for a in (q1, q2, q3))


def cfl_dt(ntheta, nr, nz, rmax, zmax, q1, q2, q3, cfl=0.3, ujet=1.0):
    L = load_library()
    g = VrhGrid(ntheta, nr, nz, rmax, zmax)
    out = ctypes.c_double(0.0)
    rc = L.vrh_cfl_dt(ctypes.byref(g), _f64(q1), _f64(q2), _f64(q3), cfl, ujet, ctypes.byref(out))
    if rc:
        raise RuntimeError("vrh_cfl_dt failed")
    return out.value


def torus_points(n, R, z0, a, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    out = np.zeros((0, 3))
    while out.shape[0] < n:
        m = 2 * (n - out.shape[0]) + 16
        p = rng.uniform([-(R + a), -(R + a), z0 - a], [R + a, R + a, z0 + a], size=(m, 3))
        rr = np.sqrt(p[:, 0] ** 2 + p[:, 1] ** 2)
        ok = (rr - R) ** 2 + (p[:, 2] - z0) ** 2 <= a * a
        out = np.concatenate([out, p[ok]])
    return out[:n]


def stokes_number(radius, Re):
    """st = (1/18)(998/1.2) r^2 Re with the reference's default-real literals
    (part_tools.F90:273)."""
    f = np.float32
    return float(np.float64(f(f(1.0) / f(18.0)) * f(f(998.0) / f(1.2))) * radius ** 2 * Re)


class Solver:
    """One whole-domain solver state on the GPU (the reference's `type(run)`)."""

    def __init__(self, ntheta, nr, nz, rmax, zmax, Re, dt, slip=0.0, theta_eigen="exact_m2",
                 precision="fp64"):
        self.L = load_library()
        self.nt, self.nr, self.nz = ntheta, nr, nz
        self.rmax, self.zmax, self.Re, self.dt = rmax, zmax, Re, dt
        self.precision = precision
        self.h = ctypes.c_void_p()
        g = VrhGrid(ntheta, nr, nz, rmax, zmax)
        p = VrhParams(dt, Re, slip, THETA[theta_eigen])
        rc = self.L.vrh_create(ctypes.byref(g), ctypes.byref(p), PRECISION[precision], ctypes.byref(self.h))
        self._check(rc, "vrh_create")
        self.halo_shape = (ntheta, nr + 3, nz + 3)
        self.comp_shape = (ntheta, nr, nz)
        self.plane_shape = (ntheta, nr + 3)

    def _check(self, rc, what):
        if rc != 0:
            buf = ctypes.create_string_buffer(512)
            if self.h:
                self.L.vrh_last_error(self.h, buf, 512)
            raise RuntimeError(f"{what} failed ({rc}): {buf.value.decode(errors='replace')}")

    def close(self):
        if self.h:
            self.L.vrh_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape_of(self, name):
        if name in ("prg",):
            return self.comp_shape + (3,)
        if name in ("prg1", "prg2", "prg3", "spec"):
            return self.comp_shape
        if name in ("csort", "dq3no", "corr3n"):
            return self.plane_shape
        return self.halo_shape

    def upload(self, name, arr):
        a = _f64(arr)
        if a.size != state_size(self.nt, self.nr, self.nz, name):
            raise ValueError(f"{name}: {a.size} values, the library moves {state_size(self.nt, self.nr, self.nz, name)}")
        self._check(self.L.vrh_upload_state(self.h, name.encode(), a), f"upload {name}")

    def download(self, name):
        shp = self.shape_of(name)
        out = np.zeros(state_size(self.nt, self.nr, self.nz, name))
        self._check(self.L.vrh_download_state(self.h, name.encode(), out), f"download {name}")
        return out.reshape(shp, order="F")

    def set_time(self, t):
        self._check(self.L.vrh_set_time(self.h, float(t)), "set_time")

    def set_radial_mesh(self, rm, rc, dmr, dcr):
        """Non-uniform radial mesh: rm, rc, dmr, dcr over j = 0..nr+1 (cell
        centres, faces with rc[1] = 0, and dr/dxi at centres and faces)."""
        arrs = [_f64(a) for a in (rm, rc, dmr, dcr)]
        assert all(a.size == self.nr + 2 for a in arrs)
        self._check(self.L.vrh_set_radial_mesh(self.h, *arrs), "set_radial_mesh")

    def set_mode_solver(self, hz, method="pcg", mmin=1, tol=1e-8, maxit=20000):
        """Solve the pressure modes m >= mmin with the Helmholtz handle `hz`
        (one slot per theta position 2 mmin - 1 .. ntheta - 1, on this
        solver's tanh axial mesh) by `method` instead of the direct z
        transform; hz None switches it off."""
        L = _hz_bind(self.L)
        hh = hz.h if hz is not None else None
        self._check(L.vrh_set_mode_solver(self.h, hh, HZ_METHOD[method], int(mmin), float(tol), int(maxit)),
                    "set_mode_solver")

    def mode_solver_stats(self):
        """(Krylov iterations summed over slots, mode solves) since set_mode_solver."""
        it, ns = ctypes.c_longlong(0), ctypes.c_longlong(0)
        self._check(_hz_bind(self.L).vrh_mode_solver_stats(self.h, ctypes.byref(it), ctypes.byref(ns)),
                    "mode_solver_stats")
        return int(it.value), int(ns.value)

    def set_axial_mesh(self, zm, zc, dmz, dcz):
        """Non-uniform axial mesh: zm, zc, dmz, dcz over k = 0..nz+1 (cell
        centres, faces with zc[1] = 0, and dz/dzeta at centres and faces)."""
        arrs = [_f64(a) for a in (zm, zc, dmz, dcz)]
        assert all(a.size == self.nz + 2 for a in arrs)
        self._check(self.L.vrh_set_axial_mesh(self.h, *arrs), "set_axial_mesh")

    def time(self):
        t = ctypes.c_double(0.0)
        self._check(self.L.vrh_get_time(self.h, ctypes.byref(t)), "get_time")
        return t.value

    def substep(self, irk):
        self._check(self.L.vrh_substep(self.h, irk), "substep")

    def phase(self, name, irk):
        self._check(self.L.vrh_phase(self.h, PHASE[name], irk), f"phase {name}")

    def step(self, n=1):
        self._check(self.L.vrh_step(self.h, int(n)), "step")

    def synchronize(self):
        self._check(self.L.vrh_synchronize(self.h), "synchronize")

    def step_async(self, n=1):
        """Queue n steps and return; synchronize() waits and checks."""
        self._check(self.L.vrh_step_async(self.h, int(n)), "step_async")

    def synchronize(self):
        self._check(self.L.vrh_synchronize(self.h), "synchronize")

    def snapshot_begin(self):
        self._check(self.L.vrh_snapshot_begin(self.h), "snapshot_begin")

    def snapshot_wait(self):
        """(q1, q2, q3, pres, time) of the pending snapshot, haloed."""
        out = [np.zeros(int(np.prod(self.halo_shape))) for _ in range(4)]
        t = ctypes.c_double()
        self._check(self.L.vrh_snapshot_wait(self.h, *out, ctypes.byref(t)), "snapshot_wait")
        q = [a.reshape(self.halo_shape, order="F") for a in out]
        return q[0], q[1], q[2], q[3], t.value

    def analysis(self):
        """Device-reduced vortex-core statistics: (vortmax, swrlmax, stat[0:8])."""
        out = np.zeros(10)
        self._check(self.L.vrh_analysis(self.h, out), "analysis")
        return out

    def set_graph(self, on=True):
        """vrh_step replays a captured HIP graph of one RK3 step (default) or
        launches the kernels one by one."""
        self._check(self.L.vrh_set_graph(self.h, 1 if on else 0), "set_graph")

    def poisson(self, rhs):
        r = _f64(rhs)
        out = np.zeros_like(r)
        self._check(self.L.vrh_poisson_solve(self.h, r, out), "poisson")
        return out.reshape(self.comp_shape, order="F")

    def kernel_times(self):
        """{kernel: (total_ms, launches)} accumulated since enable_ktimer(True)."""
        out = {}
        n = self.L.vrh_ktimer_read(self.h, -1, None, 0, None, None)
        for k in range(n):
            name = ctypes.create_string_buffer(64)
            ms, cnt = ctypes.c_double(0), ctypes.c_long(0)
            self._check(self.L.vrh_ktimer_read(self.h, k, name, 64, ctypes.byref(ms), ctypes.byref(cnt)),
                        "ktimer_read")
            if cnt.value:
                out[name.value.decode()] = (ms.value, cnt.value)
        return out

    def enable_ktimer(self, on=True):
        self._check(self.L.vrh_ktimer_enable(self.h, 1 if on else 0), "ktimer_enable")

    def autotune(self, reps=3):
        """Times the candidate line-solver segment lengths and FFT/DCT tile
        sizes on the GPU and keeps the fastest; returns the choice."""
        ch = np.zeros(9, np.int32)
        self._check(self.L.vrh_autotune(self.h, int(reps), ch), "autotune")
        keys = ["M_r1", "M_r3", "M_r2", "M_z1", "M_z3", "M_z2", "fft_lines", "dct_lines", "explicit_form"]
        return dict(zip(keys, ch.tolist()))

    def set_tuning(self, M=(0, 0, 0, 0, 0, 0), fft_lines=0, dct_lines=0, explicit_form=-1):
        """Apply launch shapes without timing (the values vrh_autotune returns)."""
        ch = np.array(list(M) + [fft_lines, dct_lines, explicit_form], np.int32)
        self._check(self.L.vrh_set_tuning(self.h, ch), "set_tuning")

    # ---- particles (part_tools.F90) ----
    def particles_create(self, caps, st, nsub, scheme="lsq_euler"):
        self.pscheme = scheme
        self.nvar = 6 if scheme == "trilinear_rk3" else 117
        rc = self.L.vrh_particles_create(self.h, len(caps), np.ascontiguousarray(caps, dtype=np.int64),
                                         np.ascontiguousarray(st, dtype=np.float64),
                                         np.ascontiguousarray(nsub, dtype=np.int32), PART_SCHEME[scheme])
        self._check(rc, "particles_create")

    def particles_upload(self, batch, u, alive=None):
        u = np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(-1, self.nvar))
        n = u.shape[0]
        al = np.ones(n, np.int32) if alive is None else np.ascontiguousarray(alive, dtype=np.int32)
        self._check(self.L.vrh_particles_upload(self.h, batch, n, u.ravel(), al), "particles_upload")

    def particles_download(self, batch, n):
        u = np.zeros(n * self.nvar)
        al = np.zeros(n, np.int32)
        self._check(self.L.vrh_particles_download(self.h, batch, n, u, al), "particles_download")
        return u.reshape(n, self.nvar), al

    def particles_step(self, irk):
        self._check(self.L.vrh_particles_step(self.h, irk), "particles_step")

    def seed_particles_torus(self, n, R=1.0, z0=3.0, a=0.6, seed=0, st=None, nsub=1, scheme="lsq_euler"):
        """n particles uniform in the torus (sqrt(x^2+y^2)-R)^2+(z-z0)^2 <= a^2
        (rejection sampling, numpy PCG64(seed)), at rest, identity Jacobian."""
        xyz = torus_points(n, R, z0, a, seed)
        if st is None:
            st = stokes_number(5.0e-6 / 0.02, self.Re)
        self.particles_create([n], [st], [nsub], scheme)
        u = np.zeros((n, self.nvar))
        u[:, :3] = xyz
        if self.nvar == 117:
            u[:, [6, 10, 14]] = 1.0
        self.particles_upload(0, u)
        return xyz

    def particles_compact(self, batch=0):
        self._check(self.L.vrh_particles_compact(self.h, batch), "particles_compact")

    def particles_sort(self, batch=0):
        """Sort the batch's particles in device memory by cell (the order
        downloads show is unchanged)."""
        self._check(self.L.vrh_particles_sort(self.h, batch), "particles_sort")

    def particles_count(self, batch=0):
        n = ctypes.c_long(0)
        self._check(self.L.vrh_particles_count(self.h, batch, ctypes.byref(n)), "particles_count")
        return n.value

    def particles_inject(self, batch=0, r0=0.5, inj=10, seed=0):
        """Inlet injection (inlt_part); returns the number of new particles."""
        n = ctypes.c_long(0)
        self._check(self.L.vrh_particles_inject(self.h, batch, float(r0), int(inj), int(seed), ctypes.byref(n)),
                    "particles_inject")
        return n.value

    def device_ptr(self, name):
        p = ctypes.c_void_p()
        self._check(self.L.vrh_device_ptr(self.h, name.encode(), ctypes.byref(p)), "device_ptr")
        return p.value


class Ensemble:
    """Independent rings (Solver objects of any size, Re and precision)
    advanced together: one HIP graph launch per RK3 step for all members."""

    def __init__(self, solvers):
        self.L = load_library()
        self.members = list(solvers)
        arr = (ctypes.c_void_p * len(self.members))(*[m.h.value for m in self.members])
        self.e = ctypes.c_void_p()
        rc = self.L.vrh_ensemble_create(arr, len(self.members), ctypes.byref(self.e))
        self._check(rc, "vrh_ensemble_create")

    def _check(self, rc, what):
        if rc != 0:
            buf = ctypes.create_string_buffer(512)
            if self.e:
                self.L.vrh_ensemble_last_error(self.e, buf, 512)
            raise RuntimeError(f"{what} failed ({rc}): {buf.value.decode(errors='replace')}")

    def step(self, n=1):
        self._check(self.L.vrh_ensemble_step(self.e, int(n)), "ensemble_step")

    def set_batch(self, steps_per_launch):
        """Steps of every member per graph launch."""
        self._check(self.L.vrh_ensemble_set_batch(self.e, int(steps_per_launch)), "ensemble_set_batch")

    def close(self):
        if self.e:
            self.L.vrh_ensemble_destroy(self.e)
            self.e = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# per-mode (r, z) Helmholtz subsystem (include/vrh_learn.h)
# ---------------------------------------------------------------------------
class VrhHzGrid(ctypes.Structure):
    _fields_ = [("nr", ctypes.c_int), ("nz", ctypes.c_int), ("nslots", ctypes.c_int),
                ("rmax", ctypes.c_double), ("zmax", ctypes.c_double), ("stretch", ctypes.c_double),
                ("pool", ctypes.c_int)]


HZ_EXPORTS = ["vrh_hz_create", "vrh_hz_destroy", "vrh_hz_last_error", "vrh_hz_set_modes", "vrh_hz_mesh",
              "vrh_hz_apply", "vrh_hz_direct", "vrh_hz_solve", "vrh_hz_net_size", "vrh_hz_net_init",
              "vrh_hz_net_get", "vrh_hz_net_set", "vrh_hz_net_forward", "vrh_hz_net_grad", "vrh_hz_sample",
              "vrh_hz_train", "vrh_hz_train_residuals", "vrh_set_mode_solver", "vrh_mode_solver_stats"]
HZ_METHOD = {"pcg": 0, "bicgstab": 1, "learned": 2, "learned_fcg": 3}


def _hz_bind(L):
    if getattr(L, "_hz_bound", False):
        return L
    P, D, I = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
    U = ctypes.c_ulonglong
    dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    L.vrh_hz_create.argtypes = [ctypes.POINTER(VrhHzGrid), ctypes.POINTER(P)]
    L.vrh_hz_destroy.argtypes = [P]
    L.vrh_hz_last_error.argtypes = [P, ctypes.c_char_p, ctypes.c_size_t]
    L.vrh_hz_set_modes.argtypes = [P, ip]
    L.vrh_hz_mesh.argtypes = [P, dp, dp, dp]
    L.vrh_hz_apply.argtypes = [P, dp, dp]
    L.vrh_hz_direct.argtypes = [P, dp, dp]
    L.vrh_hz_solve.argtypes = [P, I, D, I, dp, dp, ip, dp, ctypes.POINTER(D)]
    L.vrh_hz_net_size.argtypes = [P]
    L.vrh_hz_net_init.argtypes = [P, U]
    L.vrh_hz_net_get.argtypes = [P, fp]
    L.vrh_hz_net_set.argtypes = [P, fp]
    L.vrh_hz_net_forward.argtypes = [P, dp, fp, dp]
    L.vrh_hz_net_grad.argtypes = [P, dp, fp, fp]
    L.vrh_hz_sample.argtypes = [P, U, I, dp]
    L.vrh_hz_train.argtypes = [P, I, D, U, fp]
    L.vrh_hz_train_residuals.argtypes = [P, I, D, U, I, fp]
    L.vrh_set_mode_solver.argtypes = [P, P, I, I, D, I]
    L.vrh_mode_solver_stats.argtypes = [P, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]
    L._hz_bound = True
    return L


class Helmholtz:
    """Per-mode (r, z) pressure systems S_m p = V f on a tanh-stretched z grid
    (nslots independent slots), with direct, Krylov and learned solvers."""

    def __init__(self, nr, nz, nslots, rmax=4.0, zmax=12.0, stretch=2.0, pool=1):
        self.L = _hz_bind(load_library())
        self.nr, self.nz, self.S, self.pool = nr, nz, nslots, pool
        self.h = ctypes.c_void_p()
        g = VrhHzGrid(nr, nz, nslots, rmax, zmax, stretch, pool)
        self._check(self.L.vrh_hz_create(ctypes.byref(g), ctypes.byref(self.h)), "vrh_hz_create")
        self.shape = (nslots, nz, nr)
        self.net_shape = (nslots, nz // pool, nr // pool)

    def _check(self, rc, what):
        if rc != 0:
            buf = ctypes.create_string_buffer(512)
            if self.h:
                self.L.vrh_hz_last_error(self.h, buf, 512)
            raise RuntimeError(f"{what} failed ({rc}): {buf.value.decode(errors='replace')}")

    def close(self):
        if self.h:
            self.L.vrh_hz_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _arr(self, a):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float64), self.shape))
        return a

    def set_modes(self, m):
        self._check(self.L.vrh_hz_set_modes(self.h, np.ascontiguousarray(m, np.int32)), "set_modes")

    def mesh(self):
        zf, zm, rm = np.zeros(self.nz + 1), np.zeros(self.nz), np.zeros(self.nr)
        self._check(self.L.vrh_hz_mesh(self.h, zf, zm, rm), "mesh")
        return zf, zm, rm

    def apply(self, x):
        y = np.zeros(self.shape)
        self._check(self.L.vrh_hz_apply(self.h, self._arr(x), y), "apply")
        return y

    def direct(self, f):
        p = np.zeros(self.shape)
        self._check(self.L.vrh_hz_direct(self.h, self._arr(f), p), "direct")
        return p

    def solve(self, f, method="pcg", tol=1e-8, maxit=5000):
        p = np.zeros(self.shape)
        it = np.zeros(self.S, np.int32)
        rel = np.zeros(self.S)
        ms = ctypes.c_double(0.0)
        self._check(self.L.vrh_hz_solve(self.h, HZ_METHOD[method], float(tol), int(maxit), self._arr(f), p, it, rel,
                                        ctypes.byref(ms)), f"solve {method}")
        return p, it, rel, ms.value

    def net_size(self):
        return self.L.vrh_hz_net_size(self.h)

    def net_init(self, seed):
        self._check(self.L.vrh_hz_net_init(self.h, int(seed)), "net_init")

    def net_get(self):
        w = np.zeros(self.net_size(), np.float32)
        self._check(self.L.vrh_hz_net_get(self.h, w), "net_get")
        return w

    def net_set(self, w):
        self._check(self.L.vrh_hz_net_set(self.h, np.ascontiguousarray(w, np.float32)), "net_set")

    def net_forward(self, f):
        out = np.zeros(self.net_shape, np.float32)
        sc = np.zeros(self.S)
        self._check(self.L.vrh_hz_net_forward(self.h, self._arr(f), out, sc), "net_forward")
        return out, sc

    def net_grad(self, f, gout):
        g = np.zeros(self.net_size(), np.float32)
        self._check(self.L.vrh_hz_net_grad(self.h, self._arr(f), np.ascontiguousarray(gout, np.float32), g),
                    "net_grad")
        return g

    def sample(self, seed, family=-1):
        f = np.zeros(self.shape)
        self._check(self.L.vrh_hz_sample(self.h, int(seed), int(family), f), "sample")
        return f

    def train(self, nsamples, lr=1e-3, seed=0):
        nb = (nsamples + self.S - 1) // self.S
        loss = np.zeros(nb, np.float32)
        self._check(self.L.vrh_hz_train(self.h, int(nsamples), float(lr), int(seed), loss), "train")
        return loss

    def train_residuals(self, nsamples, lr=1e-3, seed=0, max_cg=256):
        """Training on the residuals of 0, 1, 4, 16, ... <= max_cg plain PCG
        iterations (the inputs a preconditioner is applied to)."""
        nb = (nsamples + self.S - 1) // self.S
        loss = np.zeros(nb, np.float32)
        self._check(self.L.vrh_hz_train_residuals(self.h, int(nsamples), float(lr), int(seed), int(max_cg), loss),
                    "train_residuals")
        return loss
