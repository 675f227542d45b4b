"""MI355X-native SWE Euler step on Sierpinski SFC meshes (sam(oa)^2 drop-in).

The compute path is the HIP library samoa_amd/lib/libswe_mi355x.so behind the
C ABI of include/swe_mi355x.h.
"""
from .swe import (BC_OUTFLOW, BC_WALL, G, SOLVER_AUG_RIEMANN, SOLVER_FWAVE,  # noqa: F401
                  Batch, EulerTimestep, Grid, Mesh, StepResult)
