"""Graph-captured steps and ensembles against eager launches of the same
kernels (bit-identical), and against the FP64 oracle."""
import numpy as np
import pytest

from oracle_lib import Oracle, sinh_radial_mesh

import vring_amd

pytestmark = pytest.mark.gpu

FIELDS = ("q1", "q2", "q3", "pres")


def ring_solver(nt, nr, nz, Re, precision="fp64", graph=True, amp=0.0, stretch=0.0):
    phases = np.linspace(0.3, 2.0, 3)
    q = vring_amd.ring_initial_condition(nt, nr, nz, 4.0, 12.0, [(1.0, 3.0, 0.3, 1.0)], pert_amp=amp,
                                         pert_modes=(2, 4), phases=phases)
    dt = vring_amd.cfl_dt(nt, nr, nz, 4.0, 12.0, *q, cfl=0.3, ujet=1.0)
    s = vring_amd.Solver(nt, nr, nz, 4.0, 12.0, Re, dt, precision=precision)
    if stretch > 0.0:
        s.set_radial_mesh(*sinh_radial_mesh(nr, 4.0, stretch))
    for n, a in zip(("q1", "q2", "q3"), q):
        s.upload(n, a)
    s.set_graph(graph)
    return s, q, dt


def test_graph_replays_match_eager_and_oracle():
    nt, nr, nz, Re = 32, 32, 64, 1000.0
    print(f"case graph nt={nt} nr={nr} nz={nz}", flush=True)
    sg, q, dt = ring_solver(nt, nr, nz, Re, graph=True, amp=1e-2)
    se, _, _ = ring_solver(nt, nr, nz, Re, graph=False, amp=1e-2)
    o = Oracle(nt, nr, nz, 4.0, 12.0, Re, dt)
    for n, a in zip(("q1", "q2", "q3"), q):
        o.field(n)[...] = a
    for rep in range(4):
        sg.step(1 + rep % 2)
        se.step(1 + rep % 2)
        for _ in range(1 + rep % 2):
            assert o.rk3_step() == 0
        assert sg.time() == se.time()
        assert abs(sg.time() - o.time()) < 1e-14
        scale = max(np.abs(o.field(n)).max() for n in ("q1", "q2", "q3"))
        for n in FIELDS:
            g, e = sg.download(n), se.download(n)
            assert np.array_equal(g, e), (rep, n)
            sc = scale if n != "pres" else np.abs(o.field("pres")).max()
            assert np.abs(g - o.field(n)).max() < 1e-11 * sc, (rep, n)
    sg.close()
    se.close()


@pytest.mark.parametrize("stretch,batch", [(0.0, 1), (1.5, 1), (0.0, 3)])
def test_ensemble_matches_members_stepped_alone(stretch, batch):
    """stretch > 0: every other member on a sinh-stretched radial mesh.
    batch 3: three steps of every member per graph launch (each call of 2,
    then 7 steps: single-step launches, then batches and a remainder)."""
    cases = [(16, 24, 48, 500.0, "fp64"), (33, 32, 64, 2000.0, "fp64"), (32, 48, 96, 1000.0, "mixed"),
             (64, 32, 128, 4000.0, "mixed")]
    print("case ensemble", cases, "stretch", stretch, flush=True)
    ens_members, alone = [], []
    for q, (nt, nr, nz, Re, prec) in enumerate(cases):
        st = stretch if q % 2 else 0.0
        ens_members.append(ring_solver(nt, nr, nz, Re, prec, graph=True, amp=1e-2, stretch=st)[0])
        alone.append(ring_solver(nt, nr, nz, Re, prec, graph=False, amp=1e-2, stretch=st)[0])
    ens = vring_amd.Ensemble(ens_members)
    ens.set_batch(batch)
    for rep, n in enumerate((2, 7, 2)):
        ens.step(n)
        for s in alone:
            s.step(n)
        for m, s in zip(ens_members, alone):
            assert m.time() == s.time()
            for n in FIELDS:
                assert np.array_equal(m.download(n), s.download(n)), (rep, m.nt, n)
    ens.close()
    for s in ens_members + alone:
        s.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_captured_step_after_set_time_equals_eager(precision):
    """A handle's first step after vrh_set_time, captured into a graph, equals
    the eager step bit for bit (both fork the independent sweep chains onto
    the side streams; the captured step records the fork as graph branches)."""
    nt, nr, nz = 33, 32, 64
    print(f"case captured step after set_time {precision}", flush=True)
    q, dt = ring_solver(nt, nr, nz, 1000.0, precision, graph=False, amp=1e-2)[1:]
    out = []
    for graph in (True, False):
        s = vring_amd.Solver(nt, nr, nz, 4.0, 12.0, 1000.0, dt, precision=precision)
        for n, a in zip(("q1", "q2", "q3"), q):
            s.upload(n, a)
        s.set_time(0.37)
        s.set_graph(graph)
        s.step(1)
        out.append([s.download(n) for n in FIELDS])
        s.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)


@pytest.mark.timeout(600)
def test_batched_ensemble_workload_matches_members_stepped_alone():
    """The batched-ensemble workload's draw (bench.ensemble_specs: shapes
    from {64..192} x {32..128} x {128..512}, Re ~ U[500, 5000], sigma ~
    U[0.1, 0.3], PCG64(0)) for 8 rings, mixed precision, 50 000 trilinear-RK3
    particles each: one graph launch per step for all members equals every
    member stepped alone, fields and particles bit for bit."""
    from bench import ensemble_member, ensemble_specs
    specs = ensemble_specs(8, 0)
    print("case ensemble workload, 8 rings:", [s[:3] for s in specs], flush=True)
    npart = 50000
    ens_members = [ensemble_member(sp, npart, i) for i, sp in enumerate(specs)]
    alone = [ensemble_member(sp, npart, i) for i, sp in enumerate(specs)]
    ens = vring_amd.Ensemble(ens_members)
    for rep in range(2):
        ens.step(2)
        for s in alone:
            s.step(2)
        for m, s in zip(ens_members, alone):
            assert m.time() == s.time()
            for n in FIELDS:
                assert np.array_equal(m.download(n), s.download(n)), (rep, m.nt, m.nr, m.nz, n)
            um, am = m.particles_download(0, npart)
            us, as_ = s.particles_download(0, npart)
            assert np.array_equal(um, us) and np.array_equal(am, as_), (rep, m.nt, m.nr, m.nz)
    ens.close()
    for s in ens_members + alone:
        s.close()
