"""CPU restatement of the inlet particle injection and the compaction of
live particles (TEST INFRASTRUCTURE ONLY; only tests/ import it).

Follows inlt_part / injc_part (part_tools.F90:95-181) for one batch on one
domain: for every r row ir with rc(ir) < r0 and every theta cell it (in that
order), inflow(ir) += c a q3(it, ir, 1) dt float(inj) / nprcl with
a = dtheta (rc(ir+1)^2 - rc(ir)^2), c = 1e6 * 0.02^3 and nprcl = 0.001 as
single-precision literals (part_tools.F90:275-277); n = floor(inflow) new
particles, inflow -= n.  New particles sit uniformly in the cell
(theta in [thc(it), thc(it) + dtheta), r in [rc(ir), rc(ir) + dr),
z in [0, dz)) with velocity (0, 0, q3) and an identity Jacobian.  The
positions use the library's counter-based stream (splitmix64), since
Fortran's random_number cannot be reproduced; live particles are packed to
the front in order before new ones are appended (the reference's free list
reuses holes instead; remv_part, part_tools.F90:196-225).
"""
import numpy as np

M64 = (1 << 64) - 1


def smix(x):
    x = (x + 0x9E3779B97F4A7C15) & M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


def u01(seed, i):
    return ((smix(seed ^ smix(i)) >> 11) + 0.5) * (1.0 / 9007199254740992.0)


def constants():
    f = np.float32
    c02 = f(0.02)
    c = float(f(f(1000000.0) * f(f(c02 * c02) * c02)))
    return c, float(f(0.001))


def inject(q3_inlet, inflow, nt, nr, rmax, zmax, nz, dt, r0, inj, seed, call, nvar=117):
    """q3_inlet: (nt, nr+3) values q3(it, ir, 1) (ir index as the haloed r
    index).  inflow: (nr+2,) float64 state, updated in place.  Returns the
    rows (n, nvar) of the new particles in injection order."""
    c, nprcl = constants()
    dth = 2.0 * np.pi / nt
    dr, dz = rmax / nr, zmax / nz
    rc = lambda j: (j - 1) * dr
    jr0 = max([ir for ir in range(1, nr + 1) if rc(ir) < r0], default=0)
    counts = np.zeros((jr0, nt), np.int64)
    for ir in range(1, jr0 + 1):
        a = dth * (rc(ir + 1) * rc(ir + 1) - rc(ir) * rc(ir))
        acc = inflow[ir]
        for it in range(1, nt + 1):
            acc = acc + (c * a * q3_inlet[it - 1, ir] * dt * float(np.float32(inj)) / nprcl)
            n = int(np.floor(acc))
            acc = acc - float(np.float32(n))
            counts[ir - 1, it - 1] = max(n, 0)
        inflow[ir] = acc
    rows = []
    ncell = jr0 * nt
    for t in range(ncell):
        ir, it = t // nt + 1, t % nt + 1
        for q in range(counts[ir - 1, it - 1]):
            key = ((((call * ncell + t) & M64) << 20) + q) * 3 & M64
            th = (it - 1) * dth + dth * u01(seed, key)
            r = rc(ir) + dr * u01(seed, key + 1)
            z = 0.0 + dz * u01(seed, key + 2)
            u = np.zeros(nvar)
            u[0], u[1], u[2] = r * np.cos(th), r * np.sin(th), z
            u[5] = q3_inlet[it - 1, ir]
            if nvar == 117:
                u[[6, 10, 14]] = 1.0
            rows.append(u)
    return np.array(rows).reshape(-1, nvar), counts
