"""The non-uniform axial mesh (orc_set_axial_mesh in the oracle,
vrh_set_axial_mesh in the library) where the reference has no counterpart
(its z mesh is uniform, mesh_tools.F90:321-342): (a) a substep must leave a
discretely solenoidal field on a stretched axial mesh, which any z metric
left out of divq, uppres or the pressure solve's z transform breaks; (b) on
the uniform mesh every z expression is the reference's up to rounding; (c)
the library's host eigen-decomposition of the pressure solve's z operator
against the closed form on the uniform mesh (DCT-II cosines, eigenvalues
(2/dz^2)(1 - cos(pi q/nz))) and against LAPACK (numpy.linalg.eigh) on a
stretched one."""
import os
import subprocess

import numpy as np
import pytest

from oracle_lib import Oracle, tanh_axial_mesh

import vring_amd

NT, NR, NZ, R, L = 16, 24, 40, 4.0, 12.0
HERE = os.path.dirname(os.path.abspath(__file__))


def ring():
    return vring_amd.ring_initial_condition(NT, NR, NZ, R, L, [(1.0, 3.0, 0.4, 1.0)], pert_amp=0.05,
                                            pert_modes=(1, 3))


def divergence(o):
    out = np.zeros(NT * NR * NZ)
    q = [np.ascontiguousarray(o.field(n).ravel(order="F")) for n in ("q1", "q2", "q3")]
    o.L.orc_divergence(o.h, q[0], q[1], q[2], out)
    return out.reshape((NT, NR, NZ), order="F")


def make(mesh, t=0.5):
    o = Oracle(NT, NR, NZ, R, L, 1000.0, 0.01, 0.0, 0)
    if mesh is not None:
        o.set_axial_mesh(*mesh)
    for n, a in zip(("q1", "q2", "q3"), ring()):
        o.field(n)[...] = a
    o.set_time(t)
    return o


@pytest.mark.parametrize("s", [1.0, 2.0, 3.0])
def test_projection_is_solenoidal_on_stretched_axial_mesh(s):
    """Discrete eigenvalues in theta (wvnex = 0) and the generalised
    eigen-transform in z make the Poisson operator exactly divq of the uppres
    gradient on the tanh mesh; past the inlet ramp (constant inflow, t = 0.5)
    the outlet mass correction balances the inflow, so the whole domain,
    wall row included, is solenoidal."""
    o = make(tanh_axial_mesh(NZ, L, s))
    d0 = np.abs(divergence(o)).max()
    assert d0 > 1e-2
    for irk in (1, 2, 3):
        assert o.substep(irk) == 0
        scale = max(np.abs(o.field(n)).max() for n in ("q1", "q2", "q3"))
        assert np.abs(divergence(o)).max() <= 1e-12 * scale * NZ, irk


def test_uniform_axial_mesh_equals_default_path():
    """tanh stretch 0 is the uniform mesh: every z expression and the
    eigen-transform (against the DCT) reduce to the reference's forms; a
    dropped or misplaced metric factor (m3 where c3 belongs, k where k+1
    belongs) would show on the stretched meshes above, a wrong uniform
    reduction here."""
    ref, ax = make(None, 0.13), make(tanh_axial_mesh(NZ, L, 0.0), 0.13)
    for irk in (1, 2, 3):
        assert ref.substep(irk) == 0 and ax.substep(irk) == 0
    vs = max(np.abs(ref.field(n)).max() for n in ("q1", "q2", "q3"))
    for n in ("q1", "q2", "q3"):
        assert np.abs(ref.field(n) - ax.field(n)).max() <= 1e-13 * vs, n
    p0, p1 = ref.field("pres"), ax.field("pres")
    assert np.abs(p0 - p1).max() <= 1e-12 * np.abs(p0).max()


def test_stretched_axial_mesh_changes_the_step():
    """The metrics are live: a stretched mesh gives a different state (a
    hook read but never used would pass the reduction test above)."""
    ref, ax = make(None), make(tanh_axial_mesh(NZ, L, 2.0))
    assert ref.substep(1) == 0 and ax.substep(1) == 0
    assert np.abs(ref.field("q3") - ax.field("q3")).max() > 1e-3


def metric(nz, zmax, s):
    zm, zc, dmz, dcz = tanh_axial_mesh(nz, zmax, s)
    dx3 = nz / zmax
    m3 = np.append(dx3 / dmz, dx3 / dmz[-1])
    c3 = np.append(dx3 / dcz, dx3 / dcz[-1])
    return m3, c3


@pytest.fixture(scope="module")
def eig_harness(tmp_path_factory):
    lib = os.path.dirname(os.path.abspath(vring_amd.lib.LIBPATH))
    if not os.path.exists(os.path.join(lib, "libvrh.so")):
        pytest.skip("libvrh.so is not built")
    exe = str(tmp_path_factory.mktemp("axeig") / "axial_eig_check")
    subprocess.run(["g++", "-O2", "-o", exe, os.path.join(HERE, "axial_eig_check.cpp"), "-L" + lib, "-lvrh",
                    "-Wl,-rpath," + lib], check=True)
    return exe


def run_eig(exe, m3, c3):
    nz = m3.size - 3
    inp = f"{nz}\n" + "\n".join(f"{x:.17e}" for x in np.concatenate([m3, c3])) + "\n"
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, timeout=120, check=True).stdout
    v = np.array(out.split(), dtype=float)
    lam, vi, vf = v[:nz], v[nz:nz + nz * nz].reshape(nz, nz), v[nz + nz * nz:].reshape(nz, nz)
    return lam, vi, vf


def test_host_eigen_transform_uniform_is_the_dct(eig_harness):
    nz = 48
    m3, c3 = metric(nz, L, 0.0)
    lam, vi, vf = run_eig(eig_harness, m3, c3)
    dz = L / nz
    q = np.arange(nz)
    assert np.abs(lam - 2.0 / dz**2 * (1.0 - np.cos(np.pi * q / nz))).max() <= 1e-12 * lam.max()
    assert lam[0] == 0.0
    # columns of V are the DCT-II cosines cos(pi q (k - 1/2)/nz) up to scale and sign
    k = np.arange(1, nz + 1)
    for qq in range(nz):
        c = np.cos(np.pi * qq * (k - 0.5) / nz)
        col = vf[:, qq]
        a = col @ c / (c @ c)
        assert np.abs(col - a * c).max() <= 1e-10 * np.abs(col).max(), qq
    assert np.abs(vi @ vf - np.eye(nz)).max() <= 1e-12


@pytest.mark.parametrize("nz,s", [(64, 2.0), (200, 3.0), (1151, 3.0)])
def test_host_eigen_transform_stretched_against_lapack(eig_harness, nz, s):
    m3, c3 = metric(nz, L, s)
    lam, vi, vf = run_eig(eig_harness, m3, c3)
    # the operator -(D_z G_z) with Neumann ends: M^-1 K
    K = np.zeros((nz, nz))
    for k in range(1, nz + 1):
        lo = c3[k] if k > 1 else 0.0
        hi = c3[k + 1] if k < nz else 0.0
        K[k - 1, k - 1] = lo + hi
        if k < nz:
            K[k - 1, k] = K[k, k - 1] = -c3[k + 1]
    A = np.diag(m3[1:nz + 1]) @ K
    ref = np.linalg.eigvalsh(np.sqrt(m3[1:nz + 1])[:, None] * K * np.sqrt(m3[1:nz + 1])[None, :])
    assert np.abs(lam - np.sort(ref)).max() <= 1e-11 * ref.max()
    assert lam[0] == 0.0
    assert np.abs(A @ vf - vf * lam[None, :]).max() <= 1e-10 * lam.max() * np.abs(vf).max()
    assert np.abs(vi @ vf - np.eye(nz)).max() <= 1e-11
    # the constant mode maps to the constant
    assert np.ptp(vf[:, 0]) <= 1e-13 * abs(vf[0, 0])
