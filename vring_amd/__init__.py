"""vring_amd: MI355X (gfx950) implementation of the vring RK3 substep and
Lagrangian particle step behind a C ABI (include/vrh.h, libvrh.so).

`vring_amd.lib` binds the C ABI with ctypes; `vring_amd.Solver` mirrors the
reference driver's view of the state (Fortran-ordered numpy arrays)."""
from .lib import Ensemble, Helmholtz, Solver, load_library, ring_initial_condition, cfl_dt  # noqa: F401
