"""Non-uniform radial mesh on the GPU (vrh_set_radial_mesh: the reference's
rm, rc, dmr, dcr hooks) against the oracle with the same mesh, phase by phase
(1e-11 per substep, every explicit form), over whole captured steps (1e-10),
the mixed path against FP64, the identity mesh against the default path, and
the solenoidal property of the projected field on the device."""
import numpy as np
import pytest

from gpu_common import compare, make_pair, seed_random
from oracle_lib import Oracle, sinh_radial_mesh

import vring_amd

pytestmark = pytest.mark.gpu
R, L = 4.0, 12.0


@pytest.mark.parametrize("form,slip", [(0, 0.0), (1, 0.0), (2, 0.0), (3, 0.0), (-1, 0.5)])
def test_substep_phases_on_stretched_mesh_match_oracle(form, slip):
    """form -1: the default explicit form; slip 0.5: the wall-slip rows,
    whose q1 r-sweep coefficient carries dcr(nrc)."""
    nt, nr, nz = 16, 24, 40
    print(f"case mesh phases form={form} slip={slip}", flush=True)
    mesh = sinh_radial_mesh(nr, R, 1.5)
    o, s = make_pair(nt, nr, nz, slip=slip)
    o.set_radial_mesh(*mesh)
    s.set_radial_mesh(*mesh)
    s.set_tuning(explicit_form=form)
    seed_random(o, s, seed=91)
    o.set_time(0.13)
    s.set_time(0.13)
    for irk in (1, 2, 3):
        for p in ["ucoeff", "coesys", "hdnbc", "hdnl1", "hdnl2", "hdnl3", "step"]:
            assert o.phase(p, irk) == 0
        for p in ["hdnbc", "explicit", "adi"]:
            s.phase(p, irk)
        compare(o, s, ["dq1", "dq2", "dq3", "ru1", "ru2", "ru3"], 1e-11)
        for p in ["divq", "laplace", "upsol", "uppres"]:
            assert o.phase(p, irk) == 0
        for p in ["poisson", "project"]:
            s.phase(p, irk)
        compare(o, s, ["dph", "q1", "q2", "q3", "pres", "prg"], 1e-11)


def ring(nt, nr, nz):
    return vring_amd.ring_initial_condition(nt, nr, nz, R, L, [(1.0, 3.0, 0.3, 1.0)], pert_amp=1e-2,
                                            pert_modes=(1, 4), phases=np.arange(4.0))


def test_captured_steps_on_stretched_mesh_match_oracle_and_stay_solenoidal():
    nt, nr, nz, dt = 32, 40, 64, 0.01
    print("case mesh steps 32x40x64", flush=True)
    mesh = sinh_radial_mesh(nr, R, 2.0)
    q = ring(nt, nr, nz)
    o = Oracle(nt, nr, nz, R, L, 1000.0, dt, 0.0, 0)
    o.set_radial_mesh(*mesh)
    s = vring_amd.Solver(nt, nr, nz, R, L, 1000.0, dt, theta_eigen="consistent")
    s.set_radial_mesh(*mesh)
    for n, a in zip(("q1", "q2", "q3"), q):
        o.field(n)[...] = a
        s.upload(n, a)
    # past the inlet ramp (profile constant for 0.2 < t < 2): while the inflow
    # ramps, the outlet rows carry the reference's mass correction and are
    # not solenoidal on any mesh
    o.set_time(0.5)
    s.set_time(0.5)
    for _ in range(10):
        assert o.rk3_step() == 0
    s.step(10)
    vscale = max(np.abs(o.field(n)).max() for n in ("q1", "q2", "q3"))
    errs = compare(o, s, ["q1", "q2", "q3"], 1e-10, scale=vscale)
    errs.update(compare(o, s, ["pres"], 1e-10))
    print("stretched mesh 10-step errors", errs, flush=True)
    # the device's projected field, run through the oracle's divergence with
    # the same metrics
    div = np.zeros(nt * nr * nz)
    f = [np.ascontiguousarray(s.download(n).ravel(order="F")) for n in ("q1", "q2", "q3")]
    o.L.orc_divergence(o.h, f[0], f[1], f[2], div)
    assert np.abs(div).max() <= 1e-11 * vscale * nr


def test_mixed_path_on_stretched_mesh_vs_fp64():
    nt, nr, nz, dt = 32, 40, 64, 0.01
    print("case mesh mixed 32x40x64", flush=True)
    mesh = sinh_radial_mesh(nr, R, 2.0)
    q = ring(nt, nr, nz)
    out = {}
    for prec in ("fp64", "mixed"):
        s = vring_amd.Solver(nt, nr, nz, R, L, 1000.0, dt, precision=prec)
        s.set_radial_mesh(*mesh)
        for n, a in zip(("q1", "q2", "q3"), q):
            s.upload(n, a)
        s.step(20)
        out[prec] = {n: s.download(n) for n in ("q1", "q2", "q3")}
        s.close()
    vs = max(np.abs(out["fp64"][n]).max() for n in ("q1", "q2", "q3"))
    errs = {n: float(np.abs(out["mixed"][n] - out["fp64"][n]).max() / vs) for n in ("q1", "q2", "q3")}
    print("stretched mesh mixed vs fp64 after 20 steps", errs, flush=True)
    assert max(errs.values()) < 1e-6, errs


def test_identity_mesh_equals_default_path_and_guards():
    nt, nr, nz, dt = 16, 24, 40, 0.01
    print("case mesh identity", flush=True)
    dr = R / nr
    j = np.arange(nr + 2)
    ident = ((j - 0.5) * dr, (j - 1.0) * dr, np.ones(nr + 2), np.ones(nr + 2))
    q = ring(nt, nr, nz)
    out = []
    for mesh in (None, ident):
        s = vring_amd.Solver(nt, nr, nz, R, L, 1000.0, dt)
        if mesh is not None:
            s.set_radial_mesh(*mesh)
        for n, a in zip(("q1", "q2", "q3"), q):
            s.upload(n, a)
        s.step(3)
        out.append({n: s.download(n) for n in ("q1", "q2", "q3", "pres")})
        if mesh is not None:
            # the particles need the uniform mesh
            with pytest.raises(RuntimeError):
                s.particles_create([16], [1.0], [1], "lsq_euler")
            bad = (ident[0], ident[1] + 0.5, ident[2], ident[3])
            with pytest.raises(RuntimeError):
                s.set_radial_mesh(*bad)
        s.close()
    vs = max(np.abs(out[0][n]).max() for n in ("q1", "q2", "q3"))
    for n in ("q1", "q2", "q3"):
        assert np.abs(out[0][n] - out[1][n]).max() <= 1e-13 * vs, n
    p0, p1 = out[0]["pres"], out[1]["pres"]
    assert np.abs(p0 - p1).max() <= 1e-12 * np.abs(p0).max()
