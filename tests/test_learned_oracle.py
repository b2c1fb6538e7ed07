"""Pins the CPU oracle of the learned Helmholtz subsystem (oracle/learned_oracle.py):
operator symmetry and null space, second-order convergence on manufactured
solutions on the stretched grid, Krylov solvers against the direct solve, and
the network's backward pass against finite differences."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import learned_oracle as lo  # noqa: E402


def test_operator_symmetric_positive_and_neumann_null_space():
    g = lo.mesh(12, 20, 4.0, 12.0, 2.0)
    for m in (0, 1, 5):
        A = lo.matrix(g, m).toarray()
        assert np.abs(A - A.T).max() < 1e-12 * np.abs(A).max()
        x = np.random.default_rng(m).standard_normal((20, 12))
        assert np.allclose(A @ x.ravel(), lo.apply(g, m, x).ravel(), rtol=1e-13, atol=1e-13)
        ev = np.linalg.eigvalsh(A)
        if m == 0:
            assert abs(ev[0]) < 1e-10 * ev[-1] and ev[1] > 0
            assert np.abs(A @ np.ones(A.shape[0])).max() < 1e-12
        else:
            assert ev[0] > 0


def test_stretched_mesh_is_tanh_map():
    g = lo.mesh(8, 64, 4.0, 12.0, 2.0)
    zf = g["zf"]
    assert zf[0] == 0.0 and abs(zf[-1] - 12.0) < 1e-14
    assert np.all(np.diff(zf) > 0)
    k = 16
    assert abs(zf[k] - 12.0 * (1 + np.tanh(2.0 * (k / 64 - 1)) / np.tanh(2.0))) < 1e-14
    # clustered towards z = 0: first cell much smaller than the last
    assert g["dz"][0] < 0.3 * g["dz"][-1]


def _manufactured(nr, nz, m):
    """p = (r/R)^m (1 - (r/R)^2)^2 cos(pi z / L) on the stretched grid;
    f = -[(1/r)(r p_r)_r - m^2 p / r^2 + p_zz]; returns max error of the
    discrete solution relative to max |p|."""
    R, L = 4.0, 12.0
    g = lo.mesh(nr, nz, R, L, 2.0)
    r = g["rm"][None, :]
    z = g["zm"][:, None]
    x = r / R
    a = x ** m * (1 - x * x) ** 2
    # radial part: a(x); d/dr = a'(x)/R
    da = m * x ** (m - 1) * (1 - x * x) ** 2 - 4 * x ** (m + 1) * (1 - x * x) if m > 0 else -4 * x * (1 - x * x)
    d2a = (m * (m - 1) * x ** (m - 2) * (1 - x * x) ** 2 if m > 1 else 0.0) \
        - 4 * m * x ** m * (1 - x * x) * (1 if m > 0 else 0) \
        - 4 * (m + 1) * x ** m * (1 - x * x) + 8 * x ** (m + 2)
    cz = np.cos(np.pi * z / L)
    lap_r = (d2a + da / x) / R ** 2 - m * m * a / r ** 2
    p = a * cz
    f = -(lap_r * cz - a * (np.pi / L) ** 2 * cz)
    ph = lo.direct(g, m, f)
    pe = lo.gauge(g, m, p)
    return np.abs(ph - pe).max() / np.abs(pe).max()


@pytest.mark.parametrize("m", [0, 2])
def test_manufactured_solution_second_order(m):
    e1 = _manufactured(16, 32, m)
    e2 = _manufactured(32, 64, m)
    e3 = _manufactured(64, 128, m)
    assert e3 < 2e-3
    assert np.log2(e1 / e2) > 1.7 and np.log2(e2 / e3) > 1.7, (e1, e2, e3)


@pytest.mark.parametrize("m", [0, 3])
def test_cg_and_bicgstab_reach_direct_solution(m):
    g = lo.mesh(16, 32, 4.0, 12.0, 2.0)
    f = np.random.default_rng(7 + m).standard_normal((32, 16))
    pd = lo.direct(g, m, f)
    b = lo.rhs(g, m, f)
    assert np.abs(lo.apply(g, m, pd) - b).max() < 1e-9 * np.abs(b).max()
    for solver in (lo.cg, lo.bicgstab):
        p, it = solver(g, m, f, 1e-10, 4000)
        assert it < 4000
        assert np.abs(p - pd).max() < 1e-6 * np.abs(pd).max(), solver.__name__


def test_network_backward_matches_finite_differences():
    """Float64, no rounding: d/dtheta of sum(gout * out) by central differences."""
    g = lo.mesh(6, 10, 4.0, 12.0, 2.0)
    rng = np.random.default_rng(1)
    x, _ = lo.features(g, 2, rng.standard_normal((10, 6)))
    x = x.astype(np.float64)
    Ws, bs = lo.init_params(3)
    Ws = [w.astype(np.float64) for w in Ws]
    bs = [b.astype(np.float64) + 0.05 * rng.standard_normal(b.shape) for b in bs]
    Ws[-1][:, 0, :] += 0.1 * rng.standard_normal(Ws[-1][:, 0, :].shape)
    gout = rng.standard_normal((10, 6))
    out, acts = lo.forward((Ws, bs), x, rounding=False)
    gW, gb = lo.backward((Ws, bs), acts, gout, rounding=False)

    def L(Wv, bv):
        return float((lo.forward((Wv, bv), x, rounding=False)[0] * gout).sum())

    h = 1e-6
    for l in range(4):
        for (t, co, ci) in [(0, 0, 0), (4, 3 % Ws[l].shape[1], 2), (8, 0, 1)]:
            Wp = [w.copy() for w in Ws]; Wm = [w.copy() for w in Ws]
            Wp[l][t, co, ci] += h; Wm[l][t, co, ci] -= h
            fd = (L(Wp, bs) - L(Wm, bs)) / (2 * h)
            assert abs(fd - gW[l][t, co, ci]) < 1e-6 * max(1.0, abs(fd)), (l, t, co, ci, fd, gW[l][t, co, ci])
        bp = [b.copy() for b in bs]; bm = [b.copy() for b in bs]
        bp[l][0] += h; bm[l][0] -= h
        fd = (L(Ws, bp) - L(Ws, bm)) / (2 * h)
        assert abs(fd - gb[l][0]) < 1e-6 * max(1.0, abs(fd))


def test_energy_loss_gradient_matches_finite_differences():
    g = lo.mesh(6, 10, 4.0, 12.0, 2.0)
    rng = np.random.default_rng(2)
    f = rng.standard_normal((10, 6))
    p = lo.direct(g, 1, f)
    out = rng.standard_normal((10, 6))
    scale = 0.3
    L0, gr = lo.energy_loss_grad(g, 1, out, p, f, scale)
    assert L0 > 0
    h = 1e-6
    for idx in [(0, 0), (4, 3), (9, 5)]:
        op = out.copy(); om = out.copy()
        op[idx] += h; om[idx] -= h
        fd = (lo.energy_loss_grad(g, 1, op, p, f, scale)[0] - lo.energy_loss_grad(g, 1, om, p, f, scale)[0]) / (2 * h)
        assert abs(fd - gr[idx]) < 1e-6 * max(1.0, abs(fd))
    # zero when the correction completes the Jacobi estimate to the exact
    # solution (beta = gamma = 1), whatever multiple of it the network returns
    corr = (p - lo.rhs(g, 1, f) / lo.diag(g, 1)) / scale
    for k in (1.0, -3.0):
        L1, g1 = lo.energy_loss_grad(g, 1, k * corr, p, f, scale)
        assert L1 < 1e-20 and np.abs(g1).max() < 1e-9
    # never above the best multiple of the Jacobi estimate alone, which in
    # turn is not above a zero start (L = 1)
    b = lo.rhs(g, 1, f)
    j = b / lo.diag(g, 1)
    ej = (j * b).sum() / (j * lo.apply(g, 1, j)).sum() * j - p
    LJ = (ej * lo.apply(g, 1, ej)).sum() / (p * b).sum()
    assert L0 <= LJ + 1e-14 and LJ < 1.0
    z, beta, gam = lo.galerkin(g, 1, b, np.zeros_like(b))
    assert gam == 0.0 and abs(beta - (j * b).sum() / (j * lo.apply(g, 1, j)).sum()) < 1e-14 * abs(beta)


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 3.1415927], np.float32)
    y = lo.bf16(x)
    assert y[0] == 1.0 and y[1] == 1.0            # tie -> even
    assert y[2] == 1.0 + 2 ** -7
    assert y[3] == 1.0 + 2 ** -6                  # tie -> even (upwards)
    assert y[4] == -2.5
    assert abs(y[5] - 3.140625) < 1e-7


def test_prolongation_is_bilinear_and_partition_of_unity():
    P = lo.prolong_matrix(16, 4)
    assert np.allclose(P.sum(1), 1.0) and P.min() >= 0.0
    # reproduces a linear function of the block centres between the outer centres
    nodes = (np.arange(4) + 0.5) * 4 - 0.5
    cells = np.arange(16.0)
    inside = (cells >= nodes[0]) & (cells <= nodes[-1])
    assert np.allclose((P @ (3.0 * nodes + 1.0))[inside], 3.0 * cells[inside] + 1.0)
    assert np.allclose(lo.prolong_matrix(8, 1), np.eye(8))
    out = np.random.default_rng(0).standard_normal((3, 5))
    y = lo.prolong(out, 2)
    assert y.shape == (6, 10) and abs(y[0, 0] - out[0, 0]) < 1e-15
