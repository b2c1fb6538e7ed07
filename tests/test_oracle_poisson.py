"""The oracle's Poisson path (poisson_solve.F90:140-179, 270-323;
fastfouriertransform.F90:10-87, 250-339).  The reference's path calls FFTW,
which is absent here, so these checks pin it by definitions and mathematics:
FFTW's r2c/REDFT10/REDFT01 definitions via numpy/scipy, the discrete
divergence-of-gradient stencil it must invert, and a manufactured solution."""
import numpy as np
import pytest
import scipy.fft

from oracle_lib import Oracle


def grad_div(phi, o):
    """D(G phi) with the stencils of upsol (divergence.F90:105-183) and divq
    (divergence.F90:31-60): phi (nt,nr,nz) at cell centres."""
    nt, nr, nz = phi.shape
    dth, dr, dz = 2 * np.pi / nt, o["rmax"] / nr, o["zmax"] / nz
    rm = (np.arange(1, nr + 1) - 0.5) * dr
    rc = (np.arange(1, nr + 2) - 1.0) * dr
    g1 = (phi - np.roll(phi, 1, axis=0)) / dth / rm[None, :, None]
    g2 = np.zeros((nt, nr + 1, nz))
    g2[:, 1:nr, :] = rc[None, 1:nr, None] * (phi[:, 1:, :] - phi[:, :-1, :]) / dr
    g3 = np.zeros((nt, nr, nz + 1))
    g3[:, :, 1:nz] = (phi[:, :, 1:] - phi[:, :, :-1]) / dz
    div = (np.roll(g1, -1, axis=0) - g1) / dth / rm[None, :, None]
    div += (g2[:, 1:, :] - g2[:, :-1, :]) / dr / rm[None, :, None]
    div += (g3[:, :, 1:] - g3[:, :, :-1]) / dz
    return div


def test_fftw_definitions_match_numpy_scipy():
    """Pins the transform conventions the oracle implements: for a random line,
    REDFT10 = scipy dct type 2 (norm=None) and REDFT01 = dct type 3."""
    rng = np.random.default_rng(0)
    x = rng.standard_normal(64)
    n = x.size
    y = np.array([2 * np.sum(x * np.cos(np.pi * (np.arange(n) + 0.5) * k / n)) for k in range(n)])
    np.testing.assert_allclose(scipy.fft.dct(x, type=2), y, rtol=1e-12, atol=1e-12)
    xx = np.array([y[0] + 2 * np.sum(y[1:] * np.cos(np.pi * np.arange(1, n) * (j + 0.5) / n)) for j in range(n)])
    np.testing.assert_allclose(scipy.fft.dct(y, type=3), xx, rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(xx, 2 * n * x, rtol=1e-11, atol=1e-9)


@pytest.mark.parametrize("nt,nr,nz", [(9, 12, 16), (8, 10, 32), (33, 16, 24)])
def test_poisson_inverts_discrete_div_grad_consistent_mode(nt, nr, nz):
    """Consistent mode (modified theta wavenumber): D G solve(-D G phi) = D G phi
    for arbitrary 3-D phi, odd and even nt, power-of-two and other nz."""
    p = dict(rmax=4.0, zmax=12.0)
    rng = np.random.default_rng(1)
    phi = rng.standard_normal((nt, nr, nz))
    f = grad_div(phi, p)
    o = Oracle(nt, nr, nz, p["rmax"], p["zmax"], 1000.0, 0.01, 0.0, wvnex=0)
    sol = o.poisson(-f)
    np.testing.assert_allclose(grad_div(sol, p), f, rtol=0, atol=1e-10 * np.abs(f).max())
    # the (m=0, kz=0) mode is pinned at the wall (poisson_solve.F90:304-317):
    # the solution equals phi up to a constant
    d = sol - phi
    assert np.abs(d - d.mean()).max() < 1e-9 * np.abs(phi).max()


@pytest.mark.parametrize("nt", [9, 33])
def test_poisson_parity_mode_exact_for_axisymmetric_content(nt):
    """Reference ('wvnex') mode: exact inverse for m=0 content."""
    p = dict(rmax=4.0, zmax=12.0)
    rng = np.random.default_rng(2)
    nr, nz = 16, 32
    phi = np.broadcast_to(rng.standard_normal((1, nr, nz)), (nt, nr, nz)).copy()
    f = grad_div(phi, p)
    o = Oracle(nt, nr, nz, p["rmax"], p["zmax"], 1000.0, 0.01, 0.0, wvnex=1)
    sol = o.poisson(-f)
    np.testing.assert_allclose(grad_div(sol, p), f, rtol=0, atol=1e-10 * np.abs(f).max())


def test_poisson_parity_mode_uses_exact_m2():
    """Reference mode solves mode m with eigenvalue m^2/r^2 (poisson_solve.F90:286-289),
    the consistent mode with (2/dth^2)(1-cos(m dth))/r^2: identical for m=0
    content, different for a cos(3 theta) right-hand side."""
    nt, nr, nz = 9, 12, 16
    th = (np.arange(nt) + 0.5) * 2 * np.pi / nt
    rng = np.random.default_rng(3)
    prof = rng.standard_normal((nr, nz))
    rhs3 = np.cos(3 * th)[:, None, None] * prof[None]
    rhs0 = np.ones(nt)[:, None, None] * prof[None]
    a = Oracle(nt, nr, nz, 4.0, 12.0, 1000.0, 0.01, 0.0, wvnex=1)
    b = Oracle(nt, nr, nz, 4.0, 12.0, 1000.0, 0.01, 0.0, wvnex=0)
    np.testing.assert_allclose(a.poisson(rhs0), b.poisson(rhs0), rtol=1e-12, atol=1e-12)
    assert np.abs(a.poisson(rhs3) - b.poisson(rhs3)).max() > 1e-6


def _manufactured(n):
    """solp_tools.F90:95-126: sol = cos(2 pi z)/(4 pi^2) (r^3 - 1.5 r^2)(1 - cos(5 th)/2)
    on r in [0,1], z in [0,1] with rhs = -lap(sol)."""
    nt, nr, nz = 11, n, n
    dr, dz = 1.0 / nr, 1.0 / nz
    th = (np.arange(nt) + 0.5) * 2 * np.pi / nt
    r = (np.arange(nr) + 0.5) * dr
    z = (np.arange(nz) + 0.5) * dz
    T, R, Z = np.meshgrid(th, r, z, indexing="ij")
    pi = np.pi
    sol = np.cos(2 * pi * Z) / (4 * pi ** 2) * (R ** 3 - 1.5 * R ** 2) * (1 - 0.5 * np.cos(5 * T))
    rhs = (np.cos(2 * pi * Z) * (R ** 3 - 1.5 * R ** 2) * (1 - 0.5 * np.cos(5 * T))
           + np.cos(2 * pi * Z) * (6 - 9 * R) / (4 * pi ** 2) * (1 - 0.5 * np.cos(5 * T))
           + np.cos(2 * pi * Z) / (4 * pi ** 2) * (R ** 3 - 1.5 * R ** 2) * (-0.5 * 25 * np.cos(5 * T)) / R ** 2)
    o = Oracle(nt, nr, nz, 1.0, 1.0, 1000.0, 0.01, 0.0, wvnex=1)
    num = o.poisson(rhs)
    return np.abs(num - sol).max()


def test_manufactured_solution_second_order():
    e1, e2 = _manufactured(16), _manufactured(32)
    order = np.log2(e1 / e2)
    assert 1.8 < order < 2.3, (e1, e2, order)
