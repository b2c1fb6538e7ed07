"""The Fortran driver end to end on the GPU: inp.txt with a Gaussian ring,
11 RK3 steps (istep = 0..10), restart files in the reference's format (run.F90:390-466)
and the analysis.dat records (run.F90:380-384; device-reduced statistics except
on field-export steps), compared with the FP64 oracle run from the same
initial condition."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest
from scipy.io import FortranFile

from oracle_lib import Oracle

import vring_amd

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import analysis_oracle as ao  # noqa: E402

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "driver", "vring_driver")

DECK = """===Comm===
dim                  3
npth                 1     1    1

===Mesh===
nx                  {nt}    {nr}   {nz}
maxr                 4.0
maxz                12.0

===Flag===
poisson              fftp

===Para===
dt                   {dt}
totstp               {totstp}
expstp               5
anlstp               {anlstp}
restrt               {restrt}
verbose              0
u                    0.0
v                    0.0
w                    0.0
w                    0.0   0.0  0.0 16.0
rw.r0.rh             1.0   0.5  5.0
p                    1.0
Re                1000.0
Pe                   1.0
nrkstep              3
slip                 0.0

===Ring===
nring 1
ring 1.0 3.0 0.2 1.0
pert 0.0 0 -1
cfl 0.3 1.0
"""


def read_restart(path, nt, nr, nz):
    f = FortranFile(path, "r")
    hdr = f.read_record(np.dtype([("n", "<i4"), ("t", "<f8"), ("a", "<i4"), ("b", "<i4"), ("c", "<i4")]))
    q = np.zeros((3, nt, nr, nz))
    for iz in range(nz):
        for ir in range(nr):
            for ith in range(nt):
                q[:, ith, ir, iz] = f.read_reals("<f8")
    f.close()
    return int(hdr["n"][0]), float(hdr["t"][0]), q


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")
def test_driver_ring_run_matches_oracle(tmp_path):
    nt, nr, nz = 33, 32, 64
    print("case driver ring 33x32x64", flush=True)
    (tmp_path / "inp.txt").write_text(DECK.format(nt=nt, nr=nr, nz=nz, totstp=10, restrt=0, dt=0.01, anlstp=1))
    r = subprocess.run([DRIVER], cwd=tmp_path, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    # exports after nstep = 0, 5, 10 (mod(nstep, expstp) == 0, loop.F90:136)
    for name in ["out.txt", "analysis.dat", "restart00005RK03.dat", "restart00010RK03.dat",
                 "2Dfld00005RK03.dat", "3Dfld00005RK03PR00.dat"]:
        assert (tmp_path / name).exists(), name
    nstep, ntime, qd = read_restart(tmp_path / "restart00010RK03.dat", nt, nr, nz)
    q1, q2, q3 = vring_amd.ring_initial_condition(nt, nr, nz, 4.0, 12.0, [(1.0, 3.0, 0.2, 1.0)])
    dt = vring_amd.cfl_dt(nt, nr, nz, 4.0, 12.0, q1, q2, q3, cfl=0.3, ujet=1.0)
    assert nstep == 10 and abs(ntime - 11 * dt) < 1e-12
    o = Oracle(nt, nr, nz, 4.0, 12.0, 1000.0, dt)
    for n, a in (("q1", q1), ("q2", q2), ("q3", q3)):
        o.field(n)[...] = a
    # nstep = 1 starts analysis.dat afresh (run.F90:378-379): records 1..10
    rows = np.loadtxt(tmp_path / "analysis.dat", ndmin=2)
    assert rows.shape[0] == 10 and np.all(rows[:, 0] == np.arange(1, 11))
    for n in range(11):
        assert o.rk3_step() == 0
        if n == 0:
            continue
        vmax, smax, st = ao.analysis(o.field("q1"), o.field("q2"), o.field("q3"), nr, nz, 4.0, 12.0)
        # vorticity columns: vortmax, xvort/vvort, sqrt(yvort/uvort), vvort, uvort
        ref = [vmax, st[0] / st[2], np.sqrt(st[1] / st[3]), st[2], st[3]]
        got = rows[n - 1, [2, 4, 5, 6, 7]]
        # E16.8 records: 8 significant digits
        assert np.allclose(got, ref, rtol=1e-7, atol=0), (n, got, ref)
        # the ring has no swirl: its columns hold round-off only
        assert abs(rows[n - 1, 3]) < 1e-12 and abs(rows[n - 1, 11]) < 1e-12, rows[n - 1]
    ref = np.stack([o.field(n)[:, 1 : nr + 1, 1 : nz + 1] for n in ("q1", "q2", "q3")])
    assert np.abs(qd - ref).max() <= 1e-10 * np.abs(ref).max()
    # restart: load the step-5 file as RESTART and take one step with the
    # first run's dt (a restart keeps the deck's dt; CFL applies to a fresh ring)
    shutil.copy(tmp_path / "restart00005RK03.dat", tmp_path / "RESTART")
    (tmp_path / "inp.txt").write_text(DECK.format(nt=nt, nr=nr, nz=nz, totstp=0, restrt=1, dt=repr(dt), anlstp=5))
    r = subprocess.run([DRIVER], cwd=tmp_path, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    n5, t5, q5 = read_restart(tmp_path / "RESTART", nt, nr, nz)
    n6, t6, q6 = read_restart(tmp_path / "restart00005RK03.dat", nt, nr, nz)
    o2 = Oracle(nt, nr, nz, 4.0, 12.0, 1000.0, dt)
    for c, n in enumerate(("q1", "q2", "q3")):
        o2.field(n)[:, 1 : nr + 1, 1 : nz + 1] = q5[c]
    o2.set_time(t5)
    assert o2.rk3_step() == 0
    ref2 = np.stack([o2.field(n)[:, 1 : nr + 1, 1 : nz + 1] for n in ("q1", "q2", "q3")])
    assert abs(t6 - (t5 + dt)) < 1e-12
    assert np.abs(q6 - ref2).max() <= 1e-10 * np.abs(ref2).max()


PART_DECK = """
===GPU===
precision            0
theta                0
autotune             0

===Part===
npart              300
ptorus             1.0 3.0 0.3
pradius            0.0001
nsub                 5
pscheme              0
pseed                7
"""


def read_part(path):
    """part*.dat (part_export): two header lines, then id, time, c, u(1:nvar),
    d(1:10) per live particle."""
    with open(path) as f:
        head = [f.readline(), f.readline()]
        rows = np.loadtxt(f, ndmin=2)
    return head, rows


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")
def test_driver_particle_run_exports_particles(tmp_path):
    """Driver with LSQ particles seeded in the ring's torus (part_tools.F90's
    117-value state): export files at nstep = 0, 2, 4 with one row per live
    particle, positions inside the domain, the particles carried by the
    flow, and the deformation gradient's determinant (export column d1)
    near 1 early on.  Radius 1e-4 gives a Stokes number near 1
    (part_tools.F90:273), so the particles pick up the flow within the run."""
    nt, nr, nz = 32, 32, 64
    print("case driver particles 32x32x64", flush=True)
    deck = DECK.format(nt=nt, nr=nr, nz=nz, totstp=4, restrt=0, dt=0.01, anlstp=0).replace(
        "expstp               5", "expstp               2") + PART_DECK
    (tmp_path / "inp.txt").write_text(deck)
    r = subprocess.run([DRIVER], cwd=tmp_path, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    out = {}
    for n in (0, 2, 4):
        head, rows = read_part(tmp_path / f"part{n:05d}size01.dat")
        assert head[0].strip().startswith("title") and head[1].strip().startswith("variables")
        assert rows.shape[1] == 3 + 117 + 10
        out[n] = rows
    p0, p4 = out[0], out[4]
    assert p0.shape[0] == 300 and p4.shape[0] >= 290
    for rows in out.values():
        x, y, z = rows[:, 3], rows[:, 4], rows[:, 5]
        assert np.all(np.hypot(x, y) <= 4.0) and np.all((z >= 0.0) & (z <= 12.0))
    ids = np.intersect1d(p0[:, 2], p4[:, 2])
    a = p0[np.isin(p0[:, 2], ids)][:, 3:6]
    b = p4[np.isin(p4[:, 2], ids)][:, 3:6]
    assert np.abs(b - a).max() > 1e-4          # carried by the flow
    det0 = p0[:, 3 + 117]
    assert np.all(np.abs(det0 - 1.0) < 0.05), (det0.min(), det0.max())


INLET_DECK = """
===GPU===
precision            0
theta                0
autotune             0

===Part===
npart                0
ptorus             1.0 3.0 0.3
pradius            0.0001
nsub                 5
pscheme              0
pseed                3
inlet                8   20000
"""


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")
def test_driver_inlet_batches_inject_every_ten_steps(tmp_path):
    """The driver with the reference's particle set-up (loop.F90:39-59): the
    eight inlet batches of init_part (radii 0.5..100 um, part_tools.F90:
    262-269), inlet injection every 10 steps from the jet (r0 = 0.5 from the
    deck's rw.r0.rh line), LSQ-Euler with each batch's sub-steps.  Export
    files partNNNNNsizeBB.dat for every batch at nstep = 10 and 20: the
    injections at nstep = 10 and 20 (after the inlet ramp, t > 0.2) put
    particles in the inlet cells, and the count grows between the exports.
    Re = 1e5 with dt = 0.021 keeps the smallest droplets' Euler relaxation
    stable (partdt / st = 0.36; the reference's deck has Re = 1e4,
    dt = 0.001)."""
    nt, nr, nz = 32, 32, 64
    print("case driver inlet batches 32x32x64", flush=True)
    deck = DECK.format(nt=nt, nr=nr, nz=nz, totstp=20, restrt=0, dt=0.021, anlstp=0).replace(
        "expstp               5", "expstp              10").replace("cfl 0.3 1.0", "cfl 0.0 1.0").replace(
        "Re                1000.0", "Re              100000.0") + INLET_DECK
    (tmp_path / "inp.txt").write_text(deck)
    r = subprocess.run([DRIVER], cwd=tmp_path, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "inlet particles:  8 batches" in r.stdout, r.stdout[:2000]
    counts = {}
    for n in (10, 20):
        for b in range(1, 9):
            head, rows = read_part(tmp_path / f"part{n:05d}size{b:02d}.dat")
            counts[(n, b)] = rows.shape[0] if rows.size else 0
            if rows.size:
                assert rows.shape[1] == 3 + 117 + 10
                x, y, z = rows[:, 3], rows[:, 4], rows[:, 5]
                assert np.all(np.hypot(x, y) < 0.5 + 4.0 / nr + 1e-12), (n, b)
                assert np.all((z >= 0.0) & (z < 12.0)), (n, b)
    print("particles per (step, batch)", counts, flush=True)
    for b in range(1, 9):
        assert counts[(10, b)] > 0 and counts[(20, b)] > counts[(10, b)], (b, counts)


SUBSTEP_CHECK = os.path.join(ROOT, "driver", "substep_check")


@pytest.mark.skipif(not os.path.exists(SUBSTEP_CHECK), reason="substep check not built")
def test_reference_style_rk_loop_equals_vrh_step():
    """driver/substep_check (flang): the reference's RK loop through the C
    ABI (vrh_set_time(t + dt); vrh_particles_step and vrh_substep for irk =
    1..3, loop.F90:53-131) against vrh_step: q1, q2, q3, pres, prg and two
    LSQ-Euler particle batches bit for bit, FP64 and mixed."""
    print("case substep loop vs vrh_step (flang)", flush=True)
    r = subprocess.run([SUBSTEP_CHECK], capture_output=True, text=True, timeout=240)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "substep check: bit-identical" in r.stdout


MESH_DECK = """
===GPU===
precision            0
theta                0
autotune             0
mesh                 1.5  2.0
"""


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")
def test_driver_stretched_meshes_match_oracle(tmp_path):
    """`mesh 1.5 2.0` in ===GPU===: the driver builds a sinh radial mesh and a
    tanh axial mesh and hands them to the library; six steps of the ring
    against the oracle on the same meshes (restart fields), and the
    analysis pass at the export step reads the stretched centres."""
    from oracle_lib import sinh_radial_mesh, tanh_axial_mesh
    nt, nr, nz = 32, 32, 64
    print("case driver stretched meshes 32x32x64", flush=True)
    (tmp_path / "inp.txt").write_text(DECK.format(nt=nt, nr=nr, nz=nz, totstp=5, restrt=0, dt=0.01, anlstp=5)
                                      + MESH_DECK)
    r = subprocess.run([DRIVER], cwd=tmp_path, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "radial mesh: sinh stretch" in r.stdout and "axial mesh: tanh stretch" in r.stdout
    for name in ["restart00005RK03.dat", "analysis.dat", "3Dfld00005RK03PR00.dat"]:
        assert (tmp_path / name).exists(), name
    nstep, ntime, qd = read_restart(tmp_path / "restart00005RK03.dat", nt, nr, nz)
    q1, q2, q3 = vring_amd.ring_initial_condition(nt, nr, nz, 4.0, 12.0, [(1.0, 3.0, 0.2, 1.0)])
    dt = vring_amd.cfl_dt(nt, nr, nz, 4.0, 12.0, q1, q2, q3, cfl=0.3, ujet=1.0)
    o = Oracle(nt, nr, nz, 4.0, 12.0, 1000.0, dt)
    o.set_radial_mesh(*sinh_radial_mesh(nr, 4.0, 1.5))
    o.set_axial_mesh(*tanh_axial_mesh(nz, 12.0, 2.0))
    for n, a in (("q1", q1), ("q2", q2), ("q3", q3)):
        o.field(n)[...] = a
    for _ in range(6):
        assert o.rk3_step() == 0
    ref = np.stack([o.field(n)[:, 1 : nr + 1, 1 : nz + 1] for n in ("q1", "q2", "q3")])
    err = np.abs(qd - ref).max() / np.abs(ref).max()
    print("driver stretched meshes vs oracle after 6 steps", err, flush=True)
    assert nstep == 5 and err <= 1e-10
    # the 3-D export's coordinates are the stretched centres
    rows = np.loadtxt(tmp_path / "3Dfld00005RK03PR00.dat", skiprows=3, max_rows=nt * nr * 2)
    zm = tanh_axial_mesh(nz, 12.0, 2.0)[0]
    zs = np.unique(rows[:, 2])
    assert zs.size == 2 and np.allclose(zs, zm[1:3], rtol=1e-7, atol=0), (zs, zm[1:3])
