"""The per-mode Helmholtz solvers inside the time step (vrh_set_mode_solver):
on a tanh axial mesh the pressure modes m >= mmin are solved by the
Helmholtz handle's Krylov or learned solvers instead of the direct z
eigen-transform.  Its operator is the stepper's pressure operator of one
mode, so a tight tolerance must reproduce the direct path; a loose one stays
within the tolerance's reach; the guards reject every mismatch."""
import numpy as np
import pytest

from oracle_lib import tanh_axial_mesh

import vring_amd

pytestmark = pytest.mark.gpu
R, L = 4.0, 12.0
NT, NR, NZ = 16, 32, 64


def ring():
    return vring_amd.ring_initial_condition(NT, NR, NZ, R, L, [(1.0, 3.0, 0.3, 1.0)], pert_amp=5e-2,
                                            pert_modes=(1, 2, 3, 4), phases=np.arange(4.0))


def run(prec, hz=None, method="pcg", mmin=1, tol=1e-12, steps=3):
    s = vring_amd.Solver(NT, NR, NZ, R, L, 1000.0, 0.01, precision=prec)
    s.set_axial_mesh(*tanh_axial_mesh(NZ, L, 2.0))
    if hz is not None:
        s.set_mode_solver(hz, method, mmin, tol, 20000)
    for n, a in zip(("q1", "q2", "q3"), ring()):
        s.upload(n, a)
    s.set_time(0.5)
    s.step(steps)
    out = {n: s.download(n) for n in ("q1", "q2", "q3", "pres")}
    stats = s.mode_solver_stats()
    s.close()
    return out, stats


def err(a, b):
    vs = max(np.abs(b[n]).max() for n in ("q1", "q2", "q3"))
    e = {n: float(np.abs(a[n] - b[n]).max() / vs) for n in ("q1", "q2", "q3")}
    e["pres"] = float(np.abs(a["pres"] - b["pres"]).max() / np.abs(b["pres"]).max())
    return e


@pytest.mark.parametrize("prec,mmin,tol,bar", [("fp64", 1, 1e-13, 1e-9), ("mixed", 3, 1e-10, 2e-6)])
def test_krylov_mode_solve_in_step_matches_direct(prec, mmin, tol, bar):
    print(f"case mode solver pcg {prec} mmin={mmin}", flush=True)
    ref, _ = run(prec)
    hz = vring_amd.Helmholtz(NR, NZ, NT - (2 * mmin - 1), R, L, 2.0)
    got, (its, ns) = run(prec, hz, "pcg", mmin, tol)
    hz.close()
    e = err(got, ref)
    print("pcg in step vs direct", e, "iterations", its, "solves", ns, flush=True)
    assert ns == 9 and its > 0
    assert max(e.values()) < bar, e


def test_learned_mode_solve_in_step():
    """A briefly trained network: PCG from its warm start to 1e-10."""
    print("case mode solver learned fp64", flush=True)
    ref, _ = run("fp64")
    hz = vring_amd.Helmholtz(NR, NZ, NT - 1, R, L, 2.0, pool=2)
    hz.net_init(3)
    hz.train(256, lr=1e-3, seed=3)
    got, (its, ns) = run("fp64", hz, "learned", 1, 1e-10)
    got2, (its2, _) = run("fp64", hz, "learned_fcg", 1, 1e-10)
    hz.close()
    e, e2 = err(got, ref), err(got2, ref)
    print("learned in step vs direct", e, its, "fcg", e2, its2, flush=True)
    assert ns == 9 and its > 0 and its2 > 0
    assert max(e.values()) < 1e-7 and max(e2.values()) < 1e-7


def test_mode_solver_guards():
    print("case mode solver guards", flush=True)
    s = vring_amd.Solver(NT, NR, NZ, R, L, 1000.0, 0.01)
    hz = vring_amd.Helmholtz(NR, NZ, NT - 1, R, L, 2.0)
    with pytest.raises(RuntimeError):   # no axial mesh yet
        s.set_mode_solver(hz, "pcg", 1)
    s.set_axial_mesh(*tanh_axial_mesh(NZ, L, 1.0))
    with pytest.raises(RuntimeError):   # a different stretch
        s.set_mode_solver(hz, "pcg", 1)
    s.set_axial_mesh(*tanh_axial_mesh(NZ, L, 2.0))
    with pytest.raises(RuntimeError):   # slot count of mmin = 2
        s.set_mode_solver(hz, "pcg", 2)
    with pytest.raises(RuntimeError):   # the pinned (0,0) mode stays direct
        s.set_mode_solver(hz, "pcg", 0)
    s.set_mode_solver(hz, "pcg", 1, 1e-14, 3)
    for n, a in zip(("q1", "q2", "q3"), ring()):
        s.upload(n, a)
    with pytest.raises(RuntimeError):   # 3 iterations cannot reach 1e-14
        s.step(1)
    s.set_mode_solver(None)
    s.step(1)
    # a new mesh switches the mode solve off (no iterations counted)
    s.set_mode_solver(hz, "pcg", 1, 1e-10, 20000)
    s.set_axial_mesh(*tanh_axial_mesh(NZ, L, 2.0))
    s.step(1)
    assert s.mode_solver_stats() == (0, 0)
    s.close()
    m = vring_amd.Solver(NT, NR, NZ, R, L, 1000.0, 0.01, precision="mixed")
    m.set_axial_mesh(*tanh_axial_mesh(NZ, L, 2.0))
    with pytest.raises(RuntimeError):   # mixed keeps m = 0..2 in the FP64 chain
        m.set_mode_solver(hz, "pcg", 1)
    m.close()
    c = vring_amd.Solver(NT, NR, NZ, R, L, 1000.0, 0.01, theta_eigen="consistent")
    c.set_axial_mesh(*tanh_axial_mesh(NZ, L, 2.0))
    with pytest.raises(RuntimeError):   # the Helmholtz operator has the exact m^2
        c.set_mode_solver(hz, "pcg", 1)
    c.close()
    hz.close()
