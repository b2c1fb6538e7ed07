"""The C-ABI library loads and exports every function declared in include/vrh.h
(no GPU needed: nothing is called that touches a device)."""
import ctypes
import os
import re

import numpy as np

import vring_amd
from vring_amd.lib import EXPORTS, HZ_EXPORTS, LIBPATH

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header="vrh.h"):
    with open(os.path.join(ROOT, "include", header)) as f:
        src = f.read()
    return sorted(set(re.findall(r"^\s*int\s+(vrh_\w+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIBPATH)
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(EXPORTS), set(names) ^ set(EXPORTS)
    hz = declared_functions("vrh_learn.h")
    for n in hz:
        assert hasattr(lib, n), n
    assert set(hz) == set(HZ_EXPORTS), set(hz) ^ set(HZ_EXPORTS)
    headers = sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h"))
    assert headers == ["vrh.h", "vrh_learn.h"], headers


def test_ring_initial_condition_is_divergence_free_on_host():
    """vrh_ring_init is host code: the discrete divergence of the staggered
    field vanishes to rounding and the peak vorticity is Gamma/(pi sigma^2)."""
    nt, nr, nz, rmax, zmax = 8, 48, 96, 4.0, 12.0
    q1, q2, q3 = vring_amd.ring_initial_condition(nt, nr, nz, rmax, zmax, [(1.0, 3.0, 0.25, 1.0)],
                                                  pert_amp=0.02, pert_modes=(2, 3), phases=np.array([0.1, 0.7]))
    dr, dz = rmax / nr, zmax / nz
    rm = (np.arange(1, nr + 1) - 0.5) * dr
    div = (q2[:, 2 : nr + 2, 1 : nz + 1] - q2[:, 1 : nr + 1, 1 : nz + 1]) / dr / rm[None, :, None]
    div += (q3[:, 1 : nr + 1, 2 : nz + 2] - q3[:, 1 : nr + 1, 1 : nz + 1]) / dz
    div[:, :, nz - 1] = 0.0  # outlet ghost q3(nz+1) is not part of the field
    assert np.abs(div).max() < 1e-10 * np.abs(q3).max() / dz
    assert np.all(q1 == 0.0)


def test_create_rejects_grids_outside_the_limits_before_any_device_call():
    """vrh_create checks the grid limits (3 <= ntheta <= 256, 4 <= nr <= 512,
    4 <= nz <= 1151) and the parameters before it touches a device, so the
    rejections run here without a GPU."""
    from vring_amd.lib import VrhGrid, VrhParams
    lib = ctypes.CDLL(LIBPATH)
    lib.vrh_create.argtypes = [ctypes.POINTER(VrhGrid), ctypes.POINTER(VrhParams), ctypes.c_int,
                               ctypes.POINTER(ctypes.c_void_p)]
    ok_p = VrhParams(0.01, 1000.0, 0.0, 0)
    for nt, nr, nz in [(2, 32, 64), (257, 32, 64), (16, 3, 64), (16, 513, 64), (16, 32, 3), (16, 32, 1152)]:
        h = ctypes.c_void_p(1)
        assert lib.vrh_create(ctypes.byref(VrhGrid(nt, nr, nz, 4.0, 12.0)), ctypes.byref(ok_p), 0,
                              ctypes.byref(h)) == -1, (nt, nr, nz)
        assert not h.value
    for p in [VrhParams(0.0, 1000.0, 0.0, 0), VrhParams(0.01, -1.0, 0.0, 0)]:
        h = ctypes.c_void_p(1)
        assert lib.vrh_create(ctypes.byref(VrhGrid(16, 32, 64, 4.0, 12.0)), ctypes.byref(p), 0, ctypes.byref(h)) == -1
    h = ctypes.c_void_p(1)
    assert lib.vrh_create(ctypes.byref(VrhGrid(16, 32, 64, 4.0, 12.0)), ctypes.byref(ok_p), 7, ctypes.byref(h)) == -1


def test_state_sizes_match_the_header():
    """vrh_state_size (host only) gives the element counts include/vrh.h
    documents for every named array, and the Python shapes agree."""
    nt, nr, nz = 5, 7, 11
    H, C, PL = nt * (nr + 3) * (nz + 3), nt * nr * nz, nt * (nr + 3)
    want = {n: H for n in ("q1", "q2", "q3", "dq1", "dq2", "dq3", "ru1", "ru2", "ru3",
                           "qrhs1", "qrhs2", "qrhs3", "pres", "dph")}
    want.update({n: C for n in ("prg1", "prg2", "prg3", "spec")})
    want.update({"prg": 3 * C, "csort": PL, "dq3no": PL, "corr3n": PL})
    for name, n in want.items():
        assert vring_amd.lib.state_size(nt, nr, nz, name) == n, name
    try:
        vring_amd.lib.state_size(nt, nr, nz, "nope")
        raise AssertionError("unknown name accepted")
    except KeyError:
        pass
    s = object.__new__(vring_amd.Solver)  # shape_of only: no device handle
    s.nt, s.nr, s.nz = nt, nr, nz
    s.halo_shape, s.comp_shape, s.plane_shape = (nt, nr + 3, nz + 3), (nt, nr, nz), (nt, nr + 3)
    for name, n in want.items():
        assert int(np.prod(s.shape_of(name))) == n, name
