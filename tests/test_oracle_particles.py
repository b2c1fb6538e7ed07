"""The oracle's particle step (part_tools.F90:228-371, 374-1025) against the
reference's own particle dumps (dertest/VR2D/fort.3000-3300, subset in
tests/golden/dertest_particle.npz) and closed forms."""
import os

import numpy as np

from oracle_lib import Oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden", "dertest_particle.npz")
GAM = [None, 8 / 15, 5 / 12, 3 / 4]


def dertest_oracle():
    # dertest/VR2D/inp.txt: nx 1 256 512, maxr 1.5, maxz 3.0, dt 1e-3, Re 1000
    return Oracle(1, 256, 512, 1.5, 3.0, 1000.0, 1e-3)


def stokes_dertest():
    """prt%r = 5.000*(10.0**(-6.0))/0.02 and st=(1.0/18.0)*(998.0/1.2)*r**2*Re, with
    default-real literals as in part_tools.F90:262-273."""
    f = np.float32
    r = np.float64(f(f(5.0) * f(10.0) ** f(-6.0) / f(0.02)))
    return np.float64(f(f(1.0) / f(18.0)) * f(f(998.0) / f(1.2))) * r ** 2 * 1000.0


def _rel(a, b, floor):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))


def test_interpolation_and_chain_rule_match_reference_dumps():
    g = np.load(GOLD)
    o = dertest_oracle()
    for row in range(g["f3000"].shape[0]):
        x = np.ascontiguousarray(g["f3000"][row, :3])
        vel, dv, ddv = np.zeros(3), np.zeros(9), np.zeros(27)
        assert o.L.orc_part_interp(o.h, 1, x, vel, dv, ddv) == 0
        D = dv.reshape(3, 3, order="F")
        DD = ddv.reshape(3, 3, 3, order="F")
        # E20.8 dumps: 8 significant digits in both the outputs and the input
        # positions (which are read back from the dump), so 1e-7 relative;
        # entries below 1e-3 are rounding noise of exact zeros
        assert _rel(vel, g["f3000"][row, 3:6], 1e-3) < 1e-7
        assert _rel(D.reshape(-1), g["f3100"][row, 3:12], 1e-3) < 1e-7
        for i, key in enumerate(["f3210", "f3220", "f3230"]):
            assert _rel(DD[i].reshape(-1), g[key][row, 3:12], 1e-3) < 1e-7


def test_rk3_update_matches_reference_dump():
    """fort.3300: low-storage RK3 (dertest/part_tools.F90:903) with iter_part
    called every substep, nsubdt=1, batch radius 5 um."""
    g = np.load(GOLD)["f3300"]
    o = dertest_oracle()
    st = stokes_dertest()
    u = np.zeros(117)
    u[:3] = [0.2, 0.2, 0.1]
    u[3:6] = [0.2, 0.2, 0.1]
    u[[6, 10, 14]] = 1.0
    d1, d2, alive = np.zeros(117), np.zeros(117), np.ones(1, np.int32)
    for row in range(g.shape[0]):
        irk = row % 3 + 1
        pre = u[:3].copy()
        assert o.L.orc_part_advance(o.h, 1, 1, irk, st, 1e-3, 78, 1, u, d1, d2, alive) == 0
        got = np.concatenate([pre, u[:7], [d1[0], d2[0], d1[3], d2[3], d1[6], d2[6]]])
        assert _rel(got, g[row], 1e-3) < 6e-8, row


def test_mint_is_least_squares_pseudo_inverse_and_closed_form_rows():
    o = Oracle(7, 8, 8, 1.0, 1.0, 100.0, 1e-3)
    M = np.zeros(9 * 27)
    o.L.orc_part_mint(o.h, M)
    M = M.reshape(9, 27)
    dx, dy, dz = 2 * np.pi / 7, 1.0 / 8, 1.0 / 8
    off = np.array([[a * dx, b * dy, c * dz] for c in (-1, 0, 1) for b in (-1, 0, 1) for a in (-1, 0, 1)])
    X = np.stack([off[:, 0], off[:, 1], off[:, 2], off[:, 0] ** 2, off[:, 0] * off[:, 1], off[:, 1] ** 2,
                  off[:, 1] * off[:, 2], off[:, 2] ** 2, off[:, 2] * off[:, 0]], axis=1)
    np.testing.assert_allclose(M @ X, np.eye(9), atol=1e-10)
    np.testing.assert_allclose(M, np.linalg.pinv(X), rtol=1e-9, atol=1e-9 * np.abs(M).max())
    # part_tools.F90:324-326: first row is (-1,0,1)/(18 dx) repeated
    np.testing.assert_allclose(M[0], np.tile([-1, 0, 1], 9) / (18 * dx), rtol=1e-12, atol=1e-14)
    # part_tools.F90:336-338: fourth row starts 1/(42 dx^2), -2/(21 dx^2), 1/(42 dx^2)
    np.testing.assert_allclose(M[3, :3], np.array([1 / 42, -2 / 21, 1 / 42]) / dx ** 2, rtol=1e-12)


def test_euler_substep_closed_form_relaxation():
    """Uniform axial flow (u_z = 1 from the dertest override), particle at rest:
    forward Euler on v' = (u - v)/st gives v_n = 1 - (1 - dt/st)^n exactly."""
    o = dertest_oracle()
    st, dt = 0.05, 1e-3
    u = np.zeros(117)
    u[:3] = [1e-9, 0.0, 1.0]          # on the axis line: u_r = r^2 ~ 0
    u[[6, 10, 14]] = 1.0
    d1, d2, alive = np.zeros(117), np.zeros(117), np.ones(1, np.int32)
    for n in range(1, 21):
        o.L.orc_part_advance(o.h, 1, 0, 1, st, dt, 117, 1, u, d1, d2, alive)
        assert abs(u[5] - (1 - (1 - dt / st) ** n)) < 1e-12
    # states 79..81 hold the sampled velocity after the update (part_tools.F90:931-933)
    assert abs(u[80] - 1.0) < 1e-12


def test_particles_leaving_domain_are_removed():
    o = dertest_oracle()
    u = np.zeros(117)
    u[:3] = [0.2, 0.2, 2.9999]
    u[3:6] = [0.0, 0.0, 5.0]
    u[[6, 10, 14]] = 1.0
    d1, d2, alive = np.zeros(117), np.zeros(117), np.ones(1, np.int32)
    o.L.orc_part_advance(o.h, 1, 0, 1, 1.0, 1e-3, 117, 1, u, d1, d2, alive)
    assert alive[0] == 0 and np.all(u == 0)


def _linear_fields(o, nt, nr, nz, rmax, zmax, cf):
    """q1 = u_theta, q2 = rc * u_r, q3 = u_z with each velocity component
    linear in (theta, r, z) at its own staggered nodes (mesh_tools.F90:217-350:
    q1 at (thc, rm, zm), q2 at (thm, rc, zm), q3 at (thm, rm, zc))."""
    dth, dr, dz = 2 * np.pi / nt, rmax / nr, zmax / nz
    i = np.arange(1, nt + 1)[:, None, None]
    j = np.arange(nr + 3)[None, :, None]
    k = np.arange(nz + 3)[None, None, :]
    thc, thm = (i - 1) * dth, (i - 0.5) * dth
    rc, rm = (j - 1) * dr, (j - 0.5) * dr
    zc, zm = (k - 1) * dz, (k - 0.5) * dz
    lin = lambda c, th, r, z: c[0] + c[1] * th + c[2] * r + c[3] * z
    o.field("q1")[...] = lin(cf[0], thc, rm, zm)
    o.field("q2")[...] = rc * lin(cf[1], thm, rc, zm)
    o.field("q3")[...] = lin(cf[2], thm, rm, zc)
    return lin


def test_trilinear_exact_for_linear_fields():
    """Trilinear interpolation reproduces fields linear in (theta, r, z) at
    interior points (away from the theta wrap and the axis row where the
    oracle sets u_r = q2/rc to 0), and rotates (u_theta, u_r) to Cartesian."""
    nt, nr, nz, rmax, zmax = 16, 20, 24, 2.0, 3.0
    o = Oracle(nt, nr, nz, rmax, zmax, 1000.0, 1e-3)
    cf = [(0.3, 0.05, -0.4, 0.7), (-0.2, 0.11, 0.9, -0.3), (1.1, -0.07, 0.25, 0.5)]
    lin = _linear_fields(o, nt, nr, nz, rmax, zmax, cf)
    rng = np.random.Generator(np.random.PCG64(3))
    dth, dr, dz = 2 * np.pi / nt, rmax / nr, zmax / nz
    for _ in range(200):
        th = rng.uniform(dth, 2 * np.pi - 2 * dth)
        r = rng.uniform(1.5 * dr, rmax - 1.5 * dr)
        z = rng.uniform(1.5 * dz, zmax - 1.5 * dz)
        x = np.array([r * np.cos(th), r * np.sin(th), z])
        vel = np.zeros(3)
        o.L.orc_part_trilinear(o.h, x, vel)
        ut, ur, uz = (lin(c, th, r, z) for c in cf)
        want = [-np.sin(th) * ut + np.cos(th) * ur, np.cos(th) * ut + np.sin(th) * ur, uz]
        assert np.max(np.abs(vel - want)) < 1e-12, (th, r, z, vel, want)


def test_trilinear_rk3_is_cubic_taylor_map_of_affine_ode():
    """Uniform flow u: x' = v, v' = (u - v)/st is affine, so one RK3 step
    (three substeps with gamma, rho of para_tools.F90:103-113) equals the
    cubic Taylor polynomial of the exact flow:  w = v - u, lam = -1/st,
    w1 = w0 (1 + h lam + (h lam)^2/2 + (h lam)^3/6),
    x1 = x0 + h u + w0 (h + lam h^2/2 + lam^2 h^3/6)."""
    nt, nr, nz, rmax, zmax = 16, 20, 24, 2.0, 3.0
    o = Oracle(nt, nr, nz, rmax, zmax, 1000.0, 1e-3)
    u_z = 0.8
    _linear_fields(o, nt, nr, nz, rmax, zmax, [(0, 0, 0, 0), (0, 0, 0, 0), (u_z, 0, 0, 0)])
    st, h = 0.03, 0.01
    x0 = np.array([0.4, -0.3, 1.2])
    v0 = np.array([0.2, 0.1, -0.5])
    s = np.concatenate([x0, v0])
    r2, alive = np.zeros(6), np.ones(1, np.int32)
    for irk in (1, 2, 3):
        o.L.orc_part_rk3_trilinear(o.h, 1, st, irk, h, s, r2, alive)
    u = np.array([0.0, 0.0, u_z])
    lam, w0 = -1.0 / st, v0 - u
    z = h * lam
    w1 = w0 * (1 + z + z * z / 2 + z ** 3 / 6)
    x1 = x0 + h * u + w0 * (h + lam * h * h / 2 + lam * lam * h ** 3 / 6)
    assert alive[0] == 1
    assert np.max(np.abs(s[3:] - (u + w1))) < 1e-14
    assert np.max(np.abs(s[:3] - x1)) < 1e-14
