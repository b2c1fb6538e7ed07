"""Device-reduced vortex-core statistics (vrh_analysis, run.F90:300-360)
against the CPU oracle (oracle/analysis_oracle.py) on random haloed fields and
on a ring after a few steps."""
import os
import sys

import numpy as np
import pytest

import vring_amd

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import analysis_oracle as ao  # noqa: E402

pytestmark = pytest.mark.gpu


def check(s, q1, q2, q3, nr, nz, R, L, rm=None):
    got = s.analysis()
    vmax, smax, st = ao.analysis(q1, q2, q3, nr, nz, R, L, rm=rm)
    assert abs(got[0] - vmax) <= 1e-13 * abs(vmax) and abs(got[1] - smax) <= 1e-13 * abs(smax), (got[:2], vmax, smax)
    err = np.abs(got[2:] - st) / np.abs(st)
    assert np.all(err < 1e-12), (got[2:], st)


@pytest.mark.parametrize("nt,nr,nz", [(16, 40, 60), (3, 32, 64), (33, 96, 200)])
def test_analysis_matches_oracle_on_random_fields(nt, nr, nz):
    print(f"case analysis random {nt}x{nr}x{nz}", flush=True)
    R, L = 4.0, 12.0
    s = vring_amd.Solver(nt, nr, nz, R, L, 1000.0, 0.01)
    rng = np.random.default_rng(nt + nr)
    q = [rng.standard_normal((nt, nr + 3, nz + 3)) for _ in range(3)]
    for n, a in zip(("q1", "q2", "q3"), q):
        s.upload(n, a)
    check(s, *q, nr, nz, R, L)
    s.close()


def test_analysis_on_a_stretched_radial_mesh():
    """The statistics read rm from the radial mesh set on the handle."""
    from oracle_lib import sinh_radial_mesh
    print("case analysis stretched mesh 16x48x60", flush=True)
    nt, nr, nz, R, L = 16, 48, 60, 4.0, 12.0
    mesh = sinh_radial_mesh(nr, R, 1.5)
    s = vring_amd.Solver(nt, nr, nz, R, L, 1000.0, 0.01)
    s.set_radial_mesh(*mesh)
    rng = np.random.default_rng(5)
    q = [rng.standard_normal((nt, nr + 3, nz + 3)) for _ in range(3)]
    for n, a in zip(("q1", "q2", "q3"), q):
        s.upload(n, a)
    check(s, *q, nr, nz, R, L, rm=mesh[0])
    s.close()


def test_analysis_of_a_ring_after_steps():
    print("case analysis ring 32x64x128", flush=True)
    nt, nr, nz, R, L = 32, 64, 128, 4.0, 12.0
    q1, q2, q3 = vring_amd.ring_initial_condition(nt, nr, nz, R, L, [(1.0, 3.0, 0.2, 1.0)], pert_amp=0.02,
                                                  pert_modes=(2, 4))
    s = vring_amd.Solver(nt, nr, nz, R, L, 1000.0, 0.01)
    for n, a in (("q1", q1), ("q2", q2), ("q3", q3)):
        s.upload(n, a)
    s.step(4)
    check(s, s.download("q1"), s.download("q2"), s.download("q3"), nr, nz, R, L)
    s.close()


def test_snapshot_is_the_state_at_its_time_while_stepping_goes_on():
    """vrh_snapshot_begin/_wait hand out exactly the fields of the moment the
    snapshot was taken, although steps queued after it (vrh_step_async)
    change them; asynchronous and synchronous stepping give identical
    states."""
    print("case snapshot ring 32x48x96", flush=True)
    nt, nr, nz, R, L = 32, 48, 96, 4.0, 12.0
    q = vring_amd.ring_initial_condition(nt, nr, nz, R, L, [(1.0, 3.0, 0.2, 1.0)], pert_amp=0.02,
                                         pert_modes=(2, 4))
    a = vring_amd.Solver(nt, nr, nz, R, L, 1000.0, 0.01)
    b = vring_amd.Solver(nt, nr, nz, R, L, 1000.0, 0.01)
    for s in (a, b):
        for n, arr in zip(("q1", "q2", "q3"), q):
            s.upload(n, arr)
    a.step(2)
    ref = [a.download(n) for n in ("q1", "q2", "q3", "pres")]
    t_ref = a.time()
    a.snapshot_begin()
    a.step_async(3)
    snap = a.snapshot_wait()
    a.synchronize()
    for x, y in zip(snap[:4], ref):
        assert np.array_equal(x, y)
    assert snap[4] == t_ref
    assert not np.array_equal(a.download("q3"), ref[2])
    b.step_async(2)
    b.step_async(3)
    b.synchronize()
    assert a.time() == b.time()
    for n in ("q1", "q2", "q3", "pres"):
        assert np.array_equal(a.download(n), b.download(n)), n
    a.close()
    b.close()
