"""Writes tests/golden/ref_substep.npz: the reference's own Fortran modules
(oracle/_ref/ref_harness, built by oracle/build_ref.sh from /root/reference)
run over one RK3 step's worth of substep phases (every phase except the FFTW
Poisson solve) on a seeded random state.  Only the outputs are stored; the
input is regenerated from the seed by tests/test_oracle_reference.py."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle_lib import run_ref_harness  # noqa: E402
from test_oracle_reference import CASE, PHASE_SEQ, make_state  # noqa: E402

if __name__ == "__main__":
    c = CASE
    st = make_state()
    out = run_ref_harness(c["nt"], c["nr"], c["nz"], c["rmax"], c["zmax"], c["Re"], c["dt"],
                          c["slip"], c["ntime"], PHASE_SEQ, st)
    np.savez_compressed(os.path.join(HERE, "ref_substep.npz"), **out)
    print("wrote ref_substep.npz")
