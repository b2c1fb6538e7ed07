"""vrh_analysis_noalloc (the allocation-free variant of vrh_analysis, its
block partials in a per-device static array) against vrh_analysis through
the C ABI: the same ten numbers bit for bit, on two calls in a row, the
second on a smaller grid (so the scratch holds the first call's partials
when the second starts)."""
import ctypes

import numpy as np
import pytest

import vring_amd
from vring_amd.lib import load_library

pytestmark = pytest.mark.gpu


def _hip():
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


def test_noalloc_analysis_equals_vrh_analysis():
    L = load_library()
    hip = _hip()
    dev = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dev), 10 * 8) == 0
    try:
        for nt, nr, nz in [(33, 96, 200), (16, 40, 60)]:
            print(f"case noalloc analysis {nt}x{nr}x{nz}", flush=True)
            q = vring_amd.ring_initial_condition(nt, nr, nz, 4.0, 12.0, [(1.0, 3.0, 0.3, 1.0)], pert_amp=0.05,
                                                 pert_modes=(1, 3), phases=np.arange(3.0))
            s = vring_amd.Solver(nt, nr, nz, 4.0, 12.0, 1000.0, 0.01)
            for n, a in zip(("q1", "q2", "q3"), q):
                s.upload(n, a)
            s.step(2)
            ref = s.analysis()
            # poison the output so that a call that writes nothing fails
            nan = np.full(10, np.nan)
            assert hip.hipMemcpy(dev, nan.ctypes.data, 80, 1) == 0  # host to device
            assert hip.hipDeviceSynchronize() == 0
            rc = L.vrh_analysis_noalloc(s.h, dev)
            assert rc == 0, rc
            s.synchronize()
            got = np.zeros(10)
            assert hip.hipMemcpy(got.ctypes.data, dev, 80, 2) == 0  # device to host
            s.close()
            assert np.isfinite(ref).all()
            assert np.array_equal(got, ref), (got, ref)
    finally:
        hip.hipFree(dev)


def test_noalloc_analysis_rejects_null_arguments():
    L = load_library()
    s = vring_amd.Solver(8, 16, 32, 4.0, 12.0, 1000.0, 0.01)
    assert L.vrh_analysis_noalloc(s.h, None) == -1
    assert L.vrh_analysis_noalloc(None, None) == -1
    s.close()
