"""Non-uniform axial mesh on the GPU (vrh_set_axial_mesh: every z difference
with its metric, the z sweeps with the mesh's second-difference weights, and
the pressure solve's z transform as dense GEMMs with the generalised
eigenvectors of its z operator) against the oracle with the same mesh,
phase by phase (1e-11 per substep, every explicit form), over whole steps
(1e-10), the mixed path against FP64, the uniform mesh through the axial
hook against the default (DCT) path, the solenoidal property of the
projected field, and the guards."""
import numpy as np
import pytest

from gpu_common import compare, make_pair, seed_random
from oracle_lib import Oracle, tanh_axial_mesh, sinh_radial_mesh

import vring_amd

pytestmark = pytest.mark.gpu
R, L = 4.0, 12.0


@pytest.mark.parametrize("form,theta,nt", [(0, "exact_m2", 16), (1, "exact_m2", 16), (2, "consistent", 16),
                                           (3, "exact_m2", 16), (-1, "consistent", 16), (-1, "exact_m2", 12)])
def test_substep_phases_on_axial_mesh_match_oracle(form, theta, nt):
    """nt = 12: the direct theta transforms of a non-power-of-two grid."""
    nr, nz = 24, 40
    print(f"case axial phases form={form} theta={theta} nt={nt}", flush=True)
    mesh = tanh_axial_mesh(nz, L, 2.0)
    o, s = make_pair(nt, nr, nz, theta=theta)
    o.set_axial_mesh(*mesh)
    s.set_axial_mesh(*mesh)
    s.set_tuning(explicit_form=form)
    seed_random(o, s, seed=93)
    o.set_time(0.13)
    s.set_time(0.13)
    for irk in (1, 2, 3):
        for p in ["ucoeff", "coesys", "hdnbc", "hdnl1", "hdnl2", "hdnl3", "step"]:
            assert o.phase(p, irk) == 0
        for p in ["hdnbc", "explicit", "adi"]:
            s.phase(p, irk)
        compare(o, s, ["dq1", "dq2", "dq3", "ru1", "ru2", "ru3"], 1e-11)
        # outlet convective term (hdnbc's tconv with the outlet cell's metric)
        g3, o3 = s.download("dq3no")[:, 1 : nr + 2], o.field("dq3no")[:, 1 : nr + 2]
        assert np.abs(g3 - o3).max() <= 1e-11 * np.abs(o3).max()
        for p in ["divq", "laplace", "upsol", "uppres"]:
            assert o.phase(p, irk) == 0
        for p in ["poisson", "project"]:
            s.phase(p, irk)
        compare(o, s, ["dph", "q1", "q2", "q3", "pres", "prg"], 1e-11)


def ring(nt, nr, nz):
    return vring_amd.ring_initial_condition(nt, nr, nz, R, L, [(1.0, 3.0, 0.3, 1.0)], pert_amp=1e-2,
                                            pert_modes=(1, 4), phases=np.arange(4.0))


@pytest.mark.parametrize("radial", [False, True])
def test_steps_on_axial_mesh_match_oracle_and_stay_solenoidal(radial):
    """radial: both meshes stretched (sinh in r, tanh in z)."""
    nt, nr, nz, dt = 32, 40, 64, 0.01
    print(f"case axial steps 32x40x64 radial={radial}", flush=True)
    zmesh = tanh_axial_mesh(nz, L, 2.0)
    q = ring(nt, nr, nz)
    o = Oracle(nt, nr, nz, R, L, 1000.0, dt, 0.0, 0)
    s = vring_amd.Solver(nt, nr, nz, R, L, 1000.0, dt, theta_eigen="consistent")
    if radial:
        rmesh = sinh_radial_mesh(nr, R, 1.5)
        o.set_radial_mesh(*rmesh)
        s.set_radial_mesh(*rmesh)
    o.set_axial_mesh(*zmesh)
    s.set_axial_mesh(*zmesh)
    for n, a in zip(("q1", "q2", "q3"), q):
        o.field(n)[...] = a
        s.upload(n, a)
    if radial:
        s.autotune(reps=1)  # launch shapes only; the direct Poisson kernels it times sit beside the GEMMs
    o.set_time(0.5)
    s.set_time(0.5)
    for _ in range(10):
        assert o.rk3_step() == 0
    s.step(10)
    vscale = max(np.abs(o.field(n)).max() for n in ("q1", "q2", "q3"))
    errs = compare(o, s, ["q1", "q2", "q3"], 1e-10, scale=vscale)
    errs.update(compare(o, s, ["pres"], 1e-10))
    print("axial mesh 10-step errors", errs, flush=True)
    div = np.zeros(nt * nr * nz)
    f = [np.ascontiguousarray(s.download(n).ravel(order="F")) for n in ("q1", "q2", "q3")]
    o.L.orc_divergence(o.h, f[0], f[1], f[2], div)
    assert np.abs(div).max() <= 1e-11 * vscale * nz


def test_mixed_path_on_axial_mesh_vs_fp64():
    nt, nr, nz, dt = 32, 40, 64, 0.01
    print("case axial mixed 32x40x64", flush=True)
    mesh = tanh_axial_mesh(nz, L, 2.0)
    q = ring(nt, nr, nz)
    out = {}
    for prec in ("fp64", "mixed"):
        s = vring_amd.Solver(nt, nr, nz, R, L, 1000.0, dt, precision=prec)
        s.set_axial_mesh(*mesh)
        for n, a in zip(("q1", "q2", "q3"), q):
            s.upload(n, a)
        s.step(20)
        out[prec] = {n: s.download(n) for n in ("q1", "q2", "q3")}
        s.close()
    vs = max(np.abs(out["fp64"][n]).max() for n in ("q1", "q2", "q3"))
    errs = {n: float(np.abs(out["mixed"][n] - out["fp64"][n]).max() / vs) for n in ("q1", "q2", "q3")}
    print("axial mesh mixed vs fp64 after 20 steps", errs, flush=True)
    assert max(errs.values()) < 1e-6, errs


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_uniform_axial_mesh_equals_default_path_and_guards(prec):
    """The uniform mesh through the axial hook: GEMMs with the host
    eigenvectors in place of the LDS cosine transform."""
    nt, nr, nz, dt = 16, 24, 40, 0.01
    print(f"case axial identity {prec}", flush=True)
    ident = tanh_axial_mesh(nz, L, 0.0)
    q = ring(nt, nr, nz)
    out = []
    for mesh in (None, ident):
        s = vring_amd.Solver(nt, nr, nz, R, L, 1000.0, dt, precision=prec)
        if mesh is not None:
            s.set_axial_mesh(*mesh)
        for n, a in zip(("q1", "q2", "q3"), q):
            s.upload(n, a)
        s.step(3)
        out.append({n: s.download(n) for n in ("q1", "q2", "q3", "pres")})
        if mesh is not None:
            with pytest.raises(RuntimeError):
                s.particles_create([16], [1.0], [1], "lsq_euler")
            bad = (ident[0], ident[1] + 0.5, ident[2], ident[3])
            with pytest.raises(RuntimeError):
                s.set_axial_mesh(*bad)
            with pytest.raises(RuntimeError):
                vring_amd.Ensemble([s]).step(1)
        s.close()
    vs = max(np.abs(out[0][n]).max() for n in ("q1", "q2", "q3"))
    tol = 1e-13 if prec == "fp64" else 1e-5
    for n in ("q1", "q2", "q3"):
        assert np.abs(out[0][n] - out[1][n]).max() <= tol * vs, n
    p0, p1 = out[0]["pres"], out[1]["pres"]
    assert np.abs(p0 - p1).max() <= (1e-12 if prec == "fp64" else 1e-4) * np.abs(p0).max()
