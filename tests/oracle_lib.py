"""ctypes binding to the FP64 CPU oracle (test infrastructure) and the
reference-module harness built by oracle/build_ref.sh."""
import ctypes
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")
REF_HARNESS = os.path.join(ORACLE_DIR, "_ref", "ref_harness")

PHASES = dict(ucoeff=0, coesys=1, hdnbc=2, hdnl1=3, hdnl2=4, hdnl3=5, step=6, divq=7,
              laplace=8, upsol=9, uppres=10)

_lib = None


def build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build_oracle()
        L = ctypes.CDLL(ORACLE_SO)
        P = ctypes.c_void_p
        D = ctypes.c_double
        I = ctypes.c_int
        L.orc_create.restype = P
        L.orc_create.argtypes = [I, I, I, D, D, D, D, D, I]
        L.orc_destroy.argtypes = [P]
        L.orc_field.restype = ctypes.POINTER(ctypes.c_double)
        L.orc_field.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_long)]
        L.orc_set_time.argtypes = [P, D]
        L.orc_get_time.argtypes = [P]
        L.orc_get_time.restype = D
        L.orc_phase.argtypes = [P, I, I]
        L.orc_substep.argtypes = [P, I]
        L.orc_rk3_step.argtypes = [P]
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        L.orc_poisson.argtypes = [P, dp, dp]
        L.orc_divergence.argtypes = [P, dp, dp, dp, dp]
        L.orc_set_radial_mesh.argtypes = [P, dp, dp, dp, dp]
        L.orc_set_axial_mesh.argtypes = [P, dp, dp, dp, dp]
        L.orc_part_interp.argtypes = [P, I, dp, dp, dp, dp]
        L.orc_part_advance.argtypes = [P, I, I, I, D, D, I, ctypes.c_long, dp, dp, dp, ip]
        L.orc_part_trilinear.argtypes = [P, dp, dp]
        L.orc_part_rk3_trilinear.argtypes = [P, ctypes.c_long, D, I, D, dp, dp, ip]
        L.orc_part_mint.argtypes = [P, dp]
        _lib = L
    return _lib


# Fortran shapes (column-major) of the state arrays, in harness file order.
def shapes(nt, nr, nz):
    h3 = (nt, nr + 3, nz + 3)
    pl = (nt, nr + 1)
    return [
        ("q1", h3), ("q2", h3), ("q3", h3), ("dq1", h3), ("dq2", h3), ("dq3", h3),
        ("ru1", h3), ("ru2", h3), ("ru3", h3), ("qrhs", h3), ("qcap", h3), ("pres", h3),
        ("prg", (nt, nr, nz, 3)), ("rhs", (nt, nr, nz)), ("dph", (nt, nr + 2, nz + 2)),
        ("q1s", (nt, nr + 2)), ("q2s", (nt, nr + 2)), ("q3s", (nt, nr + 2)), ("dq3no", (nt, nr + 3)),
        ("corr1s", pl), ("corr2s", pl), ("corr3s", pl), ("corr1n", pl), ("corr2n", pl),
        ("corr3n", pl), ("csort", pl), ("tconv", pl),
    ]


class Oracle:
    """Owns one oracle state; fields are numpy views (Fortran order) of the C arrays."""

    def __init__(self, nt, nr, nz, rmax, zmax, Re, dt, slip=0.0, wvnex=1):
        self.L = lib()
        self.nt, self.nr, self.nz = nt, nr, nz
        self.h = self.L.orc_create(nt, nr, nz, rmax, zmax, Re, dt, slip, wvnex)
        assert self.h, "orc_create failed"
        self.shapes = dict(shapes(nt, nr, nz))

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def field(self, name):
        cnt = ctypes.c_long(0)
        p = self.L.orc_field(self.h, name.encode(), ctypes.byref(cnt))
        assert p, name
        arr = np.ctypeslib.as_array(p, shape=(cnt.value,))
        if name in self.shapes:
            return arr.reshape(self.shapes[name], order="F")
        return arr

    def set_time(self, t):
        self.L.orc_set_time(self.h, t)

    def set_radial_mesh(self, rm, rc, dmr, dcr):
        a = [np.ascontiguousarray(np.asarray(x, dtype=np.float64)) for x in (rm, rc, dmr, dcr)]
        assert all(x.size == self.nr + 2 for x in a)
        self.L.orc_set_radial_mesh(self.h, *a)

    def set_axial_mesh(self, zm, zc, dmz, dcz):
        a = [np.ascontiguousarray(np.asarray(x, dtype=np.float64)) for x in (zm, zc, dmz, dcz)]
        assert all(x.size == self.nz + 2 for x in a)
        self.L.orc_set_axial_mesh(self.h, *a)

    def time(self):
        return self.L.orc_get_time(self.h)

    def phase(self, name, irk):
        return self.L.orc_phase(self.h, PHASES[name], irk)

    def substep(self, irk):
        return self.L.orc_substep(self.h, irk)

    def rk3_step(self):
        return self.L.orc_rk3_step(self.h)

    def poisson(self, rhs):
        rhs = np.ascontiguousarray(np.asarray(rhs, dtype=np.float64).ravel(order="F"))
        sol = np.zeros_like(rhs)
        self.L.orc_poisson(self.h, rhs, sol)
        return sol.reshape((self.nt, self.nr, self.nz), order="F")


def tanh_axial_mesh(nz, zmax, s):
    """Axial mesh with faces z(zeta) = zmax (1 + tanh(s (zeta/zmax - 1)) / tanh s)
    (cells cluster at the inlet z = 0), zeta uniform with spacing zmax/nz:
    over k = 0..nz+1 the centres zm (midpoints of the faces), faces zc
    (zc(1) = 0), and dmz, dcz = the cell widths and centre spacings over
    zmax/nz, with mirrored ghost cells at both ends."""
    k = np.arange(nz + 1)
    zf = zmax * (1.0 + np.tanh(s * (k / nz - 1.0)) / np.tanh(s)) if s > 0 else zmax * k / nz
    zf[0], zf[-1] = 0.0, zmax
    zc = np.zeros(nz + 2)
    zc[1:] = zf
    zc[0] = -zc[2]
    zm = np.zeros(nz + 2)
    zm[1 : nz + 1] = 0.5 * (zf[1:] + zf[:-1])
    zm[0] = -zm[1]
    zm[nz + 1] = 2.0 * zmax - zm[nz]
    dz = zmax / nz
    dmz = np.zeros(nz + 2)
    dmz[1 : nz + 1] = (zf[1:] - zf[:-1]) / dz
    dmz[0], dmz[nz + 1] = dmz[1], dmz[nz]
    dcz = np.zeros(nz + 2)
    dcz[1:] = (zm[1:] - zm[:-1]) / dz
    dcz[0] = dcz[1]
    return zm, zc, dmz, dcz


def sinh_radial_mesh(nr, rmax, s):
    """Radial mesh r(xi) = rmax sinh(s xi / rmax) / sinh(s), xi uniform on
    [0, rmax] with spacing rmax / nr (cells cluster at the axis): rm, rc at
    xi = (j - 1/2) dxi and (j - 1) dxi, dmr, dcr = dr/dxi there, j = 0..nr+1."""
    dxi = rmax / nr
    j = np.arange(nr + 2)
    xm, xc = (j - 0.5) * dxi, (j - 1.0) * dxi
    f = lambda x: rmax * np.sinh(s * x / rmax) / np.sinh(s)  # noqa: E731
    fp = lambda x: s * np.cosh(s * x / rmax) / np.sinh(s)  # noqa: E731
    rc = f(xc)
    rc[1] = 0.0
    return f(xm), rc, fp(xm), fp(xc)


def run_ref_harness(nt, nr, nz, rmax, zmax, Re, dt, slip, ntime, phase_list, state, mesh=None):
    """Run the reference modules on `state` (dict of Fortran-ordered arrays);
    `mesh` = (rm, rc, dmr, dcr) over j = 0..nr+1 fills the mesh module's
    radial arrays after they are built."""
    assert os.path.exists(REF_HARNESS), "oracle/_ref/ref_harness not built"
    sh = shapes(nt, nr, nz)
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            np.array([nt, nr, nz, len(phase_list), 0 if mesh is None else 1], np.int32).tofile(f)
            np.array([rmax, zmax, Re, dt, slip, ntime], np.float64).tofile(f)
            np.array([x for p in phase_list for x in (PHASES[p[0]], p[1])], np.int32).tofile(f)
            if mesh is not None:
                for a in mesh:
                    np.asarray(a, np.float64).ravel().tofile(f)
            for name, s in sh:
                np.asarray(state[name], np.float64).ravel(order="F").tofile(f)
        subprocess.run([REF_HARNESS, fin, fout], cwd=d, check=True, stdout=subprocess.DEVNULL)
        raw = np.fromfile(fout, np.float64)
    out, off = {}, 0
    for name, s in sh:
        n = int(np.prod(s))
        out[name] = raw[off:off + n].reshape(s, order="F")
        off += n
    out["ntime"], out["unom0"], out["unom1"] = raw[off:off + 3]
    off += 3
    for name in ["amj1", "acj1", "apj1", "amj2", "acj2", "apj2", "amj3", "acj3", "apj3"]:
        out[name] = raw[off:off + nr + 2]
        off += nr + 2
    for name in ["amk1", "ack1", "apk1", "amk2", "ack2", "apk2", "amk3", "ack3", "apk3"]:
        out[name] = raw[off:off + nz + 2]
        off += nz + 2
    assert off == raw.size
    return out
