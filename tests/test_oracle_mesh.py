"""The oracle's non-uniform radial mesh (the reference's rm, rc, dmr, dcr
hooks filled by orc_set_radial_mesh) where the reference's own modules cannot
check it: the Poisson rows.  The reference writes them for the uniform mesh
only (poisson.F90:379-381); the oracle's rows are minus divq's radial
difference of uppres' radial gradient, so (a) a substep must leave a
discretely solenoidal field on the stretched mesh, which an inconsistent row
(the uniform rows kept on the stretched mesh, a dropped dmr or dcr) does not,
and (b) on the uniform mesh they are the reference's rows up to rounding."""
import numpy as np
import pytest

from oracle_lib import Oracle, sinh_radial_mesh

import vring_amd

NT, NR, NZ, R, L = 16, 24, 40, 4.0, 12.0


def ring():
    return vring_amd.ring_initial_condition(NT, NR, NZ, R, L, [(1.0, 3.0, 0.4, 1.0)], pert_amp=0.05,
                                            pert_modes=(1, 3))


def divergence(o):
    out = np.zeros(NT * NR * NZ)
    q = [np.ascontiguousarray(o.field(n).ravel(order="F")) for n in ("q1", "q2", "q3")]
    o.L.orc_divergence(o.h, q[0], q[1], q[2], out)
    return out.reshape((NT, NR, NZ), order="F")


def make(mesh, wvnex=0):
    o = Oracle(NT, NR, NZ, R, L, 1000.0, 0.01, 0.0, wvnex)
    if mesh is not None:
        o.set_radial_mesh(*mesh)
    for n, a in zip(("q1", "q2", "q3"), ring()):
        o.field(n)[...] = a
    o.set_time(0.5)
    return o


@pytest.mark.parametrize("s", [0.8, 1.5, 2.5])
def test_projection_is_solenoidal_on_stretched_mesh(s):
    """Discrete eigenvalues in theta (wvnex = 0) and z make the Poisson
    operator exactly divq of the uppres gradient; the ring, built for the
    uniform mesh, is far from solenoidal on the stretched one."""
    o = make(sinh_radial_mesh(NR, R, s))
    d0 = np.abs(divergence(o)).max()
    assert d0 > 1e-2
    for irk in (1, 2, 3):
        assert o.substep(irk) == 0
        scale = max(np.abs(o.field(n)).max() for n in ("q1", "q2", "q3"))
        assert np.abs(divergence(o)).max() <= 1e-12 * scale * NR, irk


def test_uniform_rows_on_stretched_mesh_are_caught():
    """The mesh set but the Poisson rows left as the uniform ones (a plausible
    slip): the divergence after a substep stays far above round-off."""
    mesh = sinh_radial_mesh(NR, R, 1.5)
    o = make(None)
    # fill only the metric hooks, as the reference's own mesh module would,
    # keeping the uniform-mesh Poisson rows
    for name, a in zip(("rm", "rc", "dmr", "dcr"), mesh):
        o.field(name)[: NR + 2] = a
    assert o.substep(1) == 0
    scale = max(np.abs(o.field(n)).max() for n in ("q1", "q2", "q3"))
    assert np.abs(divergence(o)).max() > 1e-6 * scale


def test_identity_mesh_equals_default_path():
    dr = R / NR
    j = np.arange(NR + 2)
    mesh = ((j - 0.5) * dr, (j - 1.0) * dr, np.ones(NR + 2), np.ones(NR + 2))
    a, b = make(None, wvnex=1), make(mesh, wvnex=1)
    for irk in (1, 2, 3):
        assert a.substep(irk) == 0 and b.substep(irk) == 0
    for n in ("q1", "q2", "q3", "pres"):
        x, y = a.field(n), b.field(n)
        assert np.abs(x - y).max() <= 1e-13 * np.abs(x).max(), n
