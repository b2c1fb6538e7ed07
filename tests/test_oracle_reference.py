"""The oracle against the reference's own Fortran modules.

Every substep phase except the FFTW-backed Poisson solve (ucoeff, coesys,
hdnbc, hdnl1-3, step with its ADI sweeps, divq, upsol, uppres) is compared
with the reference modules compiled from /root/reference (oracle/_ref), or,
where that build is absent, with the committed outputs of that build
(tests/golden/ref_substep.npz, written by tests/golden/make_ref_golden.py)."""
import os

import numpy as np
import pytest

from oracle_lib import REF_HARNESS, Oracle, run_ref_harness, shapes, sinh_radial_mesh

CASE = dict(nt=7, nr=10, nz=14, rmax=4.0, zmax=12.0, Re=1000.0, dt=0.01, slip=0.0, ntime=0.15)
PHASE_SEQ = [(p, irk) for irk in (1, 2, 3) for p in
             ["ucoeff", "coesys", "hdnbc", "hdnl1", "hdnl2", "hdnl3", "step", "divq", "upsol", "uppres"]]
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_substep.npz")
GOLDEN_MESH = os.path.join(os.path.dirname(__file__), "golden", "ref_substep_mesh.npz")
MESH_STRETCH = 1.5   # sinh-stretched radial mesh for the metric-hook cases


def make_state(seed=0):
    c = CASE
    rng = np.random.default_rng(seed)
    st = {n: rng.standard_normal(s) for n, s in shapes(c["nt"], c["nr"], c["nz"])}
    for n in ["corr1s", "corr2s", "corr3s", "corr1n", "corr2n", "corr3n", "csort", "tconv", "qcap", "qrhs"]:
        st[n][:] = 0.0
    return st


def run_oracle(st, seq, slip=0.0, mesh=None):
    c = CASE
    o = Oracle(c["nt"], c["nr"], c["nz"], c["rmax"], c["zmax"], c["Re"], c["dt"], slip)
    if mesh is not None:
        o.set_radial_mesh(*mesh)
    for n, _ in shapes(c["nt"], c["nr"], c["nz"]):
        o.field(n)[...] = st[n]
    o.set_time(c["ntime"])
    for p, irk in seq:
        assert o.phase(p, irk) == 0
    return o


def _compare(o, ref):
    c = CASE
    for n, _ in shapes(c["nt"], c["nr"], c["nz"]):
        a, b = o.field(n), ref[n]
        scale = max(np.abs(b).max(), 1e-300)
        assert np.abs(a - b).max() <= 1e-14 * scale, n
    for n in ["amj1", "acj1", "apj1", "amj2", "acj2", "apj2", "amj3", "acj3", "apj3"]:
        np.testing.assert_array_equal(o.field(n)[: c["nr"] + 2], ref[n])
    for n in ["amk1", "ack1", "apk1", "amk2", "ack2", "apk2", "amk3", "ack3", "apk3"]:
        np.testing.assert_array_equal(o.field(n)[: c["nz"] + 2], ref[n])


def test_oracle_matches_committed_reference_outputs():
    ref = dict(np.load(GOLDEN))
    o = run_oracle(make_state(), PHASE_SEQ)
    _compare(o, ref)


@pytest.mark.skipif(not os.path.exists(REF_HARNESS), reason="oracle/_ref not built")
@pytest.mark.parametrize("seed,slip", [(1, 0.0), (2, 1.0)])
def test_oracle_matches_reference_modules_live(seed, slip):
    c = CASE
    st = make_state(seed)
    ref = run_ref_harness(c["nt"], c["nr"], c["nz"], c["rmax"], c["zmax"], c["Re"], c["dt"], slip,
                          c["ntime"], PHASE_SEQ, st)
    o = run_oracle(st, PHASE_SEQ, slip)
    _compare(o, ref)


@pytest.mark.skipif(not os.path.exists(REF_HARNESS), reason="oracle/_ref not built")
def test_planted_mistake_is_caught():
    """A dropped term (the rho_2 ru1 history term, by zeroing ru1 in the
    oracle's copy of the input) must fail the comparison."""
    c = CASE
    st = make_state(3)
    ref = run_ref_harness(c["nt"], c["nr"], c["nz"], c["rmax"], c["zmax"], c["Re"], c["dt"], 0.0,
                          c["ntime"], PHASE_SEQ[10:17], st)
    st2 = {k: v.copy() for k, v in st.items()}
    st2["ru1"][:] = 0.0
    o = run_oracle(st2, PHASE_SEQ[10:17])
    with pytest.raises(AssertionError):
        np.testing.assert_allclose(o.field("dq1"), ref["dq1"], rtol=1e-12, atol=0)


def test_oracle_on_stretched_mesh_matches_committed_reference_outputs():
    """The metric hooks (rm, rc, dmr, dcr) filled with a sinh-stretched
    radial mesh: every phase but the Poisson solve against the reference's
    own modules run with the same mesh."""
    c = CASE
    mesh = sinh_radial_mesh(c["nr"], c["rmax"], MESH_STRETCH)
    ref = dict(np.load(GOLDEN_MESH))
    o = run_oracle(make_state(5), PHASE_SEQ, 0.0, mesh)
    _compare(o, ref)


@pytest.mark.skipif(not os.path.exists(REF_HARNESS), reason="oracle/_ref not built")
@pytest.mark.parametrize("seed,slip", [(6, 0.0), (7, 1.0)])
def test_oracle_on_stretched_mesh_matches_reference_modules_live(seed, slip):
    c = CASE
    mesh = sinh_radial_mesh(c["nr"], c["rmax"], MESH_STRETCH)
    st = make_state(seed)
    ref = run_ref_harness(c["nt"], c["nr"], c["nz"], c["rmax"], c["zmax"], c["Re"], c["dt"], slip,
                          c["ntime"], PHASE_SEQ, st, mesh=mesh)
    o = run_oracle(st, PHASE_SEQ, slip, mesh)
    _compare(o, ref)
