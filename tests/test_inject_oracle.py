"""Pins the injection restatement (oracle/inject_oracle.py) by closed forms:
with a uniform inlet velocity the particles injected per r row over K calls
equal floor(K nt x) (x the per-cell flux), every particle lies in its inlet
cell with velocity (0, 0, q3), and the single-precision constants are those
of part_tools.F90:275-277."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import inject_oracle as io  # noqa: E402


def test_constants_are_single_precision_literals():
    c, nprcl = io.constants()
    # (10.0**6)*(0.02**3) rounds to exactly 8 in single precision; 0.001 does not
    assert c == 8.0
    assert nprcl == float(np.float32(0.001)) and nprcl != 0.001


def test_uniform_inlet_counts_and_cells():
    nt, nr, nz, rmax, zmax, dt, r0, inj = 8, 16, 32, 4.0, 12.0, 1e-3, 0.9, 10
    U = 1.3
    q3 = np.full((nt, nr + 3), U)
    inflow = np.zeros(nr + 2)
    c, nprcl = io.constants()
    dth, dr, dz = 2 * np.pi / nt, rmax / nr, zmax / nz
    K = 7
    tot = np.zeros(nr + 1, np.int64)
    for call in range(K):
        rows, counts = io.inject(q3, inflow, nt, nr, rmax, zmax, nz, dt, r0, inj, 11, call)
        tot[1:counts.shape[0] + 1] += counts.sum(1)
        r = np.hypot(rows[:, 0], rows[:, 1])
        th = np.mod(np.arctan2(rows[:, 1], rows[:, 0]), 2 * np.pi)
        assert np.all(rows[:, 2] >= 0) and np.all(rows[:, 2] < dz)
        assert np.all(rows[:, 5] == U) and np.all(rows[:, 3:5] == 0)
        assert np.all(rows[:, [6, 10, 14]] == 1.0)
        # each particle in the theta cell and r row that produced it
        k = 0
        for ir in range(1, counts.shape[0] + 1):
            for it in range(1, nt + 1):
                for _ in range(counts[ir - 1, it - 1]):
                    assert (ir - 1) * dr <= r[k] < ir * dr + 1e-12
                    assert (it - 1) * dth - 1e-12 <= th[k] < it * dth + 1e-12
                    k += 1
    jr0 = int(np.ceil(r0 / dr))
    assert jr0 == 4
    for ir in range(1, jr0 + 1):
        a = dth * ((ir * dr) ** 2 - ((ir - 1) * dr) ** 2)
        x = c * a * U * dt * 10.0 / nprcl
        assert abs(tot[ir] - np.floor(K * nt * x)) <= 1, (ir, tot[ir], K * nt * x)
    assert tot[jr0 + 1:].sum() == 0
