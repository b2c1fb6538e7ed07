"""Pins the vortex-core statistics oracle (oracle/analysis_oracle.py) by
closed forms: a constant swirl, the vorticity of a z-linear radial flux and of
an r-quadratic axial velocity, and the symmetry of the window."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import analysis_oracle as ao  # noqa: E402

NT, NR, NZ, R, L = 4, 40, 60, 4.0, 12.0


def fields():
    z = np.zeros((NT, NR + 3, NZ + 3))
    return z.copy(), z.copy(), z.copy()


def window_cells():
    """Independent count of window cells: centres (j - 1/2) dr strictly inside
    (0.05 R, 0.8 R) with j >= 2, (k - 1/2) dz strictly inside (0.1 L, 0.9 L)."""
    dr, dz = R / NR, L / NZ
    js = [j for j in range(2, NR + 1) if 0.05 * R < (j - 0.5) * dr < 0.8 * R]
    ks = [k for k in range(1, NZ + 1) if 0.1 * L < (k - 0.5) * dz < 0.9 * L]
    return js, ks


def test_constant_swirl():
    q1, q2, q3 = fields()
    q1[...] = 0.7
    vmax, smax, st = ao.analysis(q1, q2, q3, NR, NZ, R, L)
    js, ks = window_cells()
    dr, dz = R / NR, L / NZ
    w = NT / (2 * np.pi) / dr / dz
    assert vmax == 0.0 and smax == 0.7
    assert np.all(st[:4] == 0.0)       # zero vorticity is not > 0
    n = NT * len(js) * len(ks)
    assert abs(st[7] - 0.49 * w * n) < 1e-12 * st[7]
    # window symmetric in z: x/v moment is L/2; y/u is the mean window radius
    assert abs(st[4] / st[6] - L / 2) < 1e-12
    assert abs(st[5] / st[7] - np.mean([(j - 0.5) * dr for j in js])) < 1e-12
    assert st[5] == st[6]


def test_vorticity_of_linear_flux_and_quadratic_axial_velocity():
    """q2 = a z at the r faces (z at cell centres) gives dvdz = a (1/r_j +
    1/r_j+1) / 2; q3 = b r^2 (r at cell centres) gives dwdr = 2 b r_j; the
    vorticity is their difference."""
    a, b = 0.3, 0.02
    q1, q2, q3 = fields()
    dr, dz = R / NR, L / NZ
    zc = (np.arange(NZ + 3) - 0.5) * dz
    rc = (np.arange(NR + 3) - 0.5) * dr
    q2[...] = a * zc[None, None, :]
    q3[...] = b * rc[None, :, None] ** 2
    vmax, smax, st = ao.analysis(q1, q2, q3, NR, NZ, R, L)
    js, ks = window_cells()
    v = np.array([0.5 * a * (1 / ((j - 0.5) * dr) + 1 / ((j + 0.5) * dr)) - 2 * b * (j - 0.5) * dr for j in js])
    assert abs(vmax - v.max()) < 1e-12 * abs(v.max())
    assert smax == 0.0 and np.all(st[4:] == 0.0)
    w = NT / (2 * np.pi) / dr / dz
    pos = v > 0
    assert pos.any() and not pos.all()   # the sign test cuts through the window
    u = NT * len(ks) * w * (v[pos] ** 2).sum()
    assert abs(st[3] - u) < 1e-12 * u
    assert abs(st[0] / st[2] - L / 2) < 1e-12


def test_negative_vorticity_contributes_nothing():
    q1, q2, q3 = fields()
    dz = L / NZ
    zc = (np.arange(NZ + 3) - 0.5) * dz
    q2[...] = -0.3 * zc[None, None, :]
    q1[...] = -1.0
    vmax, smax, st = ao.analysis(q1, q2, q3, NR, NZ, R, L)
    assert vmax < 0 and smax == -1.0 and np.all(st == 0.0)


def test_vorticity_on_a_stretched_radial_mesh():
    """With the radial mesh rm(j) of a sinh stretching: q3 = b rm^2 gives
    dwdr = b (rm(j+1) + rm(j-1)) exactly, and a z-linear q2 gives
    dvdz = a (1/rm(j) + 1/rm(j+1)) / 2; the window follows rm."""
    a, b, s = 0.3, 0.02, 1.5
    dxi = R / NR
    j = np.arange(NR + 2)
    rmesh = R * np.sinh(s * (j - 0.5) * dxi / R) / np.sinh(s)
    q1, q2, q3 = fields()
    dz = L / NZ
    zc = (np.arange(NZ + 3) - 0.5) * dz
    q2[...] = a * zc[None, None, :]
    rext = np.append(rmesh, 2 * rmesh[-1] - rmesh[-2])   # j = nr+2 (outside the window)
    q3[...] = b * rext[None, :, None] ** 2
    vmax, smax, st = ao.analysis(q1, q2, q3, NR, NZ, R, L, rm=rmesh)
    js = [jj for jj in range(2, NR + 1) if 0.05 * R < rmesh[jj] < 0.8 * R]
    v = np.array([0.5 * a * (1 / rmesh[jj] + 1 / rmesh[jj + 1]) - b * (rmesh[jj + 1] + rmesh[jj - 1]) for jj in js])
    assert abs(vmax - v.max()) < 1e-12 * abs(v.max())
    ks = [k for k in range(1, NZ + 1) if 0.1 * L < (k - 0.5) * dz < 0.9 * L]
    w = NT / (2 * np.pi) / dxi / dz
    pos = v > 0
    u = NT * len(ks) * w * (v[pos] ** 2).sum()
    assert abs(st[3] - u) < 1e-12 * u
    # the stretched window holds other rows than the uniform one
    assert len(js) != len([jj for jj in range(2, NR + 1) if 0.05 * R < (jj - 0.5) * dxi < 0.8 * R])
