"""GPU (libvrh through its C ABI) against the FP64 oracle, phase by phase and
over whole steps.  FP64 path: 1e-11 per substep; 1e-10 after 100 steps."""
import numpy as np
import pytest

from gpu_common import compare, make_pair, seed_fields, seed_random

import vring_amd

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form", [0, 1, 2, 3])
@pytest.mark.parametrize("nt,nr,nz", [(9, 16, 32), (32, 32, 64), (33, 24, 40), (16, 64, 128), (128, 20, 40),
                                      (64, 96, 128)])
def test_substep_phases_match_oracle(nt, nr, nz, form):
    """form: explicit kernel with one z plane per block (0), z marching (1),
    or z marching with one wavefront per r row within 64 KB of LDS (2) or
    within the device's LDS limits (3)."""
    print(f"case phases nt={nt} nr={nr} nz={nz} form={form}", flush=True)
    o, s = make_pair(nt, nr, nz, slip=0.0)
    s.set_tuning(explicit_form=form)
    seed_random(o, s, seed=nt + nr + nz)
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


@pytest.mark.parametrize("nt,nr,nz", [(3, 4, 4), (8, 512, 8), (4, 8, 1023), (4, 8, 1024), (3, 4, 1151)])
def test_substep_phases_at_size_limits_match_oracle(nt, nr, nz):
    """The smallest grid vrh_create accepts (ntheta = 3: direct theta DFT),
    the largest nr (512: radial solves of 16 segments of 32 rows, r sweeps
    of 513 rows), z lines of 1024 rows (nz = 1023: 16 segments of 64, direct
    z DCT for a length that is not a power of two), of 1025 rows (nz = 1024,
    the large workload's length: 15 segments of 72, FFT-based DCT) and the
    longest (nz = 1151: 16 segments of 72)."""
    print(f"case phases at size limits nt={nt} nr={nr} nz={nz}", flush=True)
    o, s = make_pair(nt, nr, nz, slip=0.0)
    seed_random(o, s, seed=nt + nr + nz)
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


@pytest.mark.parametrize("slip", [1.0, 0.5])
def test_substep_phases_with_wall_slip_match_oracle(slip):
    """Wall slip (the q1, q3 copies into the ghost row at nrc and the slip
    rows of the sweeps) against the oracle, every phase of three substeps."""
    nt, nr, nz = 16, 24, 40
    print(f"case phases slip={slip}", flush=True)
    o, s = make_pair(nt, nr, nz, slip=slip)
    seed_random(o, s, seed=77)
    o.set_time(0.21)
    s.set_time(0.21)
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


@pytest.mark.parametrize("nt,nr,nz,theta", [(9, 16, 32, "exact_m2"), (16, 32, 64, "consistent"),
                                            (33, 20, 24, "exact_m2"), (128, 64, 128, "exact_m2"),
                                            (64, 96, 256, "consistent"), (32, 192, 512, "exact_m2")])
def test_poisson_matches_oracle(nt, nr, nz, theta):
    print(f"case poisson nt={nt} nr={nr} nz={nz} {theta}", flush=True)
    o, s = make_pair(nt, nr, nz, theta=theta)
    rhs = np.random.default_rng(5).standard_normal((nt, nr, nz))
    ref = o.poisson(rhs)
    got = s.poisson(rhs)
    assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()


@pytest.mark.parametrize("precision,lines", [("fp64", 8), ("mixed", 16), ("mixed", 2), ("mixed", 32), ("fp64", 32)])
def test_poisson_long_z_lines_match_oracle(precision, lines):
    """nz = 512 z transforms at the tile widths the tuner may pick (16 mixed
    lines per block run with 512 threads, 32 lines with 1024 threads and
    more than 64 KB of LDS) against the oracle; FP64 to 1e-11, mixed to the
    FP32 transforms' rounding."""
    nt, nr, nz = 8, 12, 512
    print(f"case poisson long z {precision} lines={lines}", flush=True)
    o, s = make_pair(nt, nr, nz, precision=precision)
    s.set_tuning(dct_lines=lines)
    rhs = np.random.default_rng(9).standard_normal((nt, nr, nz))
    ref = o.poisson(rhs)
    got = s.poisson(rhs)
    tol = 1e-11 if precision == "fp64" else 2e-5
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()


def test_ring_100_steps_fp64_matches_oracle():
    """Config-1 shape at odd ntheta (the reference's Poisson packing needs odd
    ntheta): nr=64, ntheta=33, nz=128, r in [0,4], z in [0,12], Re=1000, one
    Gaussian ring (R=1, sigma=0.2, Gamma=1) at z=3, dt at CFL 0.3, 100 steps."""
    nt, nr, nz, rmax, zmax = 33, 64, 128, 4.0, 12.0
    q1, q2, q3 = vring_amd.ring_initial_condition(nt, nr, nz, rmax, zmax, [(1.0, 3.0, 0.2, 1.0)])
    dt = vring_amd.cfl_dt(nt, nr, nz, rmax, zmax, q1, q2, q3, cfl=0.3, ujet=1.0)
    o, s = make_pair(nt, nr, nz, rmax, zmax, Re=1000.0, dt=dt)
    seed_fields(o, s, q1, q2, q3)
    for step in range(100):
        assert o.rk3_step() == 0
    s.step(100)
    vscale = max(np.abs(o.field("q2")).max(), np.abs(o.field("q3")).max())
    errs = compare(o, s, ["q1", "q2", "q3"], 1e-10, scale=vscale)
    errs.update(compare(o, s, ["pres"], 1e-10))
    print("100-step relative errors", errs, flush=True)


def test_mixed_precision_ring_error_vs_fp64():
    """Mixed path (FP32 stencils/sweeps/transforms, FP64 low modes, predictor,
    projection and state) against the FP64 path after 50 steps of the ring:
    relative field error below 1e-6."""
    nt, nr, nz, rmax, zmax = 32, 64, 128, 4.0, 12.0
    q1, q2, q3 = vring_amd.ring_initial_condition(nt, nr, nz, rmax, zmax, [(1.0, 3.0, 0.2, 1.0)],
                                                  pert_amp=1e-3, pert_modes=(2, 8), phases=np.arange(7.0))
    dt = vring_amd.cfl_dt(nt, nr, nz, rmax, zmax, q1, q2, q3, cfl=0.3, ujet=1.0)
    out = {}
    for prec in ("fp64", "mixed"):
        s = vring_amd.Solver(nt, nr, nz, rmax, zmax, 1000.0, dt, precision=prec)
        for n, a in (("q1", q1), ("q2", q2), ("q3", q3)):
            s.upload(n, a)
        s.step(50)
        out[prec] = {n: s.download(n) for n in ("q1", "q2", "q3", "pres")}
    vs = max(np.abs(out["fp64"][n]).max() for n in ("q1", "q2", "q3"))
    errs = {n: np.abs(out["mixed"][n] - out["fp64"][n]).max() / vs for n in ("q1", "q2", "q3")}
    errs["pres"] = np.abs(out["mixed"]["pres"] - out["fp64"]["pres"]).max() / np.abs(out["fp64"]["pres"]).max()
    print("mixed vs fp64 after 50 steps", errs, flush=True)
    assert max(errs.values()) < 1e-6, errs


@pytest.mark.parametrize("nt,nr,nz,slip", [(33, 32, 64, 0.0), (9, 16, 32, 0.0), (48, 24, 40, 0.0), (32, 24, 40, 0.5)])
def test_mixed_precision_at_other_ntheta_vs_fp64(nt, nr, nz, slip):
    """The mixed path where theta or z is not a power of two (direct
    transforms; the FP64 low modes of a line from two waves' partials), and
    with wall slip: relative field error against FP64 after 20 ring steps
    below 1e-6."""
    print(f"case mixed vs fp64 nt={nt} nr={nr} nz={nz} slip={slip}", flush=True)
    rmax, zmax = 4.0, 12.0
    q1, q2, q3 = vring_amd.ring_initial_condition(nt, nr, nz, rmax, zmax, [(1.0, 3.0, 0.3, 1.0)],
                                                  pert_amp=1e-2, pert_modes=(1, 3), phases=np.arange(3.0))
    dt = vring_amd.cfl_dt(nt, nr, nz, rmax, zmax, q1, q2, q3, cfl=0.3, ujet=1.0)
    out = {}
    for prec in ("fp64", "mixed"):
        s = vring_amd.Solver(nt, nr, nz, rmax, zmax, 1000.0, dt, slip, precision=prec)
        for n, a in (("q1", q1), ("q2", q2), ("q3", q3)):
            s.upload(n, a)
        s.step(20)
        out[prec] = {n: s.download(n) for n in ("q1", "q2", "q3", "pres")}
        s.close()
    vs = max(np.abs(out["fp64"][n]).max() for n in ("q1", "q2", "q3"))
    errs = {n: float(np.abs(out["mixed"][n] - out["fp64"][n]).max() / vs) for n in ("q1", "q2", "q3")}
    errs["pres"] = float(np.abs(out["mixed"]["pres"] - out["fp64"]["pres"]).max() / np.abs(out["fp64"]["pres"]).max())
    print("mixed vs fp64 after 20 steps", errs, flush=True)
    assert max(errs.values()) < 1e-6, errs


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("nt,nr,nz", [(128, 40, 64), (128, 22, 40), (48, 36, 64), (256, 16, 32)])
def test_explicit_forms_agree_bitwise(precision, nt, nr, nz):
    """The marching forms of the explicit kernel (one block row per theta
    line, or one wavefront per r row with either row count per block) do the
    same arithmetic per cell in the same order: whole ring steps agree bit
    for bit in both precisions; the one-plane form agrees to rounding."""
    print(f"case explicit forms {precision} {nt}x{nr}x{nz}", flush=True)
    q = vring_amd.ring_initial_condition(nt, nr, nz, 4.0, 12.0, [(1.0, 3.0, 0.3, 1.0)], pert_amp=1e-2,
                                         pert_modes=(2, 4))
    out = []
    for form in (1, 2, 0, 3):
        s = vring_amd.Solver(nt, nr, nz, 4.0, 12.0, 1000.0, 0.01, precision=precision)
        s.set_tuning(explicit_form=form)
        for n, a in zip(("q1", "q2", "q3"), q):
            s.upload(n, a)
        s.step(3)
        out.append([s.download(n) for n in ("q1", "q2", "q3", "pres")])
        s.close()
    for a, b, c in zip(out[0], out[1], out[3]):
        assert np.array_equal(a, b)
        assert np.array_equal(a, c)
    tol = 1e-12 if precision == "fp64" else 1e-5
    for a, b in zip(out[0], out[2]):
        assert np.abs(a - b).max() <= tol * np.abs(a).max()
