"""Extracts a subset of the reference's particle dumps under
/root/reference/dertest/VR2D (data files, not source) into
tests/golden/dertest_particle.npz:
  fort.3000 rows: x, y, z, vel(1:3)
  fort.3100 rows: x, y, z, dvel(1,1..3), dvel(2,1..3), dvel(3,1..3)
  fort.3210/3220/3230 rows: x, y, z, ddvel(i,1,1..3), ddvel(i,2,1..3), ddvel(i,3,1..3)
  fort.3300 rows 1-60: x, y, z (before), u(1:7) after, ddt(1,1), ddt(2,1),
                       ddt(1,4), ddt(2,4), ddt(1,7), ddt(2,7)
Field override u_theta=0, u_r=r^2, u_z=1 (dertest/part_tools.F90:580-582);
mesh nx 1 256 512, maxr 1.5, maxz 3.0 (dertest/VR2D/inp.txt)."""
import os

import numpy as np

SRC = "/root/reference/dertest/VR2D"
HERE = os.path.dirname(os.path.abspath(__file__))

if __name__ == "__main__":
    rows = np.arange(0, 4208, 29)
    out = {}
    for n in ["3000", "3100", "3210", "3220", "3230"]:
        out["f" + n] = np.loadtxt(os.path.join(SRC, "fort." + n))[rows]
    out["f3300"] = np.loadtxt(os.path.join(SRC, "fort.3300"))[:60]
    np.savez_compressed(os.path.join(HERE, "dertest_particle.npz"), **out)
    print({k: v.shape for k, v in out.items()})
