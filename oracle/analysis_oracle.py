"""CPU oracle (test infrastructure only) of the vortex-core statistics of
analysis, run.F90:300-360: per cell the azimuthal vorticity
dvdz - dwdr (run.F90:317-320) and the swirl q1m (run.F90:314-315), their
maxima inside the window 0.1 < z/zmax < 0.9, 0.05 < r/rmax < 0.8, j != 1
(run.F90:324), and the moment sums stat(1:8) of the positive parts
(run.F90:327-341).  The run.F90:307-309 copy of j = 1 into the j = 0 halo does
not enter: the window excludes j = 1 and j = 2 reads j = 1.

Fields are haloed, Fortran order, as numpy arrays of shape
(nt, nr + 3, nz + 3) indexed [ith - 1, j, k].

Pinned by closed forms in tests/test_analysis_oracle.py (a swirl constant,
dvdz of a z-linear radial flux, dwdr of an r-quadratic axial velocity, the
window mean of z and r).
"""
import numpy as np


def window(nr, nz, rmax, zmax, rm=None):
    """Cell centres and the mask of cells inside the statistics window; rm
    (j = 0..nr+1) is the radial mesh, uniform when not given."""
    dr, dz = rmax / nr, zmax / nz
    if rm is None:
        rm = (np.arange(0, nr + 2) - 0.5) * dr   # index j = 0..nr+1
    zm = (np.arange(0, nz + 2) - 0.5) * dz   # index k = 0..nz+1
    j = np.arange(1, nr + 1)
    k = np.arange(1, nz + 1)
    inr = (rm[j] < 0.80 * rmax) & (rm[j] > rmax * 0.05) & (j != 1)
    inz = (zm[k] < 0.90 * zmax) & (zm[k] > 0.10 * zmax)
    return rm, zm, inr, inz


def analysis(q1, q2, q3, nr, nz, rmax, zmax, rm=None):
    """Returns (vortmax, swrlmax, stat[0:8]) with stat = xvort, yvort, vvort,
    uvort, xswrl, yswrl, vswrl, uswrl.  rm (j = 0..nr+1): a non-uniform
    radial mesh, read where run.F90 reads msh%rm (the weights keep dx2)."""
    nt = q1.shape[0]
    dth, dr, dz = 2.0 * np.pi / nt, rmax / nr, zmax / nz
    w = (1.0 / dth) * (1.0 / dr) * (1.0 / dz)
    rm, zm, inr, inz = window(nr, nz, rmax, zmax, None if rm is None else np.asarray(rm, dtype=np.float64))
    J = np.arange(1, nr + 1)[:, None]
    K = np.arange(1, nz + 1)[None, :]
    if nt > 1:
        q1m = 0.5 * (q1[:, 1:nr + 1, 1:nz + 1] + np.roll(q1, -1, axis=0)[:, 1:nr + 1, 1:nz + 1])
    else:
        q1m = q1[:, 1:nr + 1, 1:nz + 1].copy()
    dwdr = 0.5 * ((q3[:, J + 1, K] + q3[:, J + 1, K + 1]) - (q3[:, J - 1, K] + q3[:, J - 1, K + 1])) \
        / (rm[J + 1] - rm[J - 1])
    dvdz = 0.5 * ((q2[:, J, K + 1] / rm[J] + q2[:, J + 1, K + 1] / rm[J + 1])
                  - (q2[:, J + 1, K - 1] / rm[J + 1] + q2[:, J, K - 1] / rm[J])) / (zm[K + 1] - zm[K - 1])
    vort = dvdz - dwdr
    swrl = q1m
    mask = np.broadcast_to(inr[:, None] & inz[None, :], vort.shape)
    r = np.broadcast_to(rm[J], vort.shape)
    z = np.broadcast_to(zm[K], vort.shape)
    if not mask.any():
        return -10000.0, -10000.0, np.zeros(8)
    out = [max(vort[mask].max(), -10000.0), max(swrl[mask].max(), -10000.0)]
    stat = []
    for a in (vort, swrl):
        sel = mask & (a > 0.0)
        a2 = a[sel] ** 2
        stat += [(z[sel] * a2 * r[sel] * w).sum(), (a2 * r[sel] * w).sum(), (a2 * r[sel] * w).sum(),
                 (a2 * w).sum()]
    return out[0], out[1], np.array(stat)
