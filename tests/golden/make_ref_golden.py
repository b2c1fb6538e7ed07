"""Writes tests/golden/ref_substep.npz and ref_substep_mesh.npz: the
reference's own Fortran modules (oracle/_ref/ref_harness, built by
oracle/build_ref.sh from /root/reference) run over one RK3 step's worth of
substep phases (every phase except the FFTW Poisson solve) on a seeded random
state, on the uniform radial mesh and on a sinh-stretched one written into
the mesh module's rm, rc, dmr, dcr.  Only the outputs are stored; the inputs
are regenerated from the seeds by tests/test_oracle_reference.py."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle_lib import run_ref_harness, sinh_radial_mesh  # noqa: E402
from test_oracle_reference import CASE, MESH_STRETCH, PHASE_SEQ, make_state  # noqa: E402

if __name__ == "__main__":
    c = CASE
    st = make_state()
    out = run_ref_harness(c["nt"], c["nr"], c["nz"], c["rmax"], c["zmax"], c["Re"], c["dt"],
                          c["slip"], c["ntime"], PHASE_SEQ, st)
    np.savez_compressed(os.path.join(HERE, "ref_substep.npz"), **out)
    mesh = sinh_radial_mesh(c["nr"], c["rmax"], MESH_STRETCH)
    out = run_ref_harness(c["nt"], c["nr"], c["nz"], c["rmax"], c["zmax"], c["Re"], c["dt"],
                          c["slip"], c["ntime"], PHASE_SEQ, make_state(5), mesh=mesh)
    np.savez_compressed(os.path.join(HERE, "ref_substep_mesh.npz"), **out)
    print("wrote ref_substep.npz, ref_substep_mesh.npz")
