"""MI355X-native fused esoteric-pull stream+collide step (D2Q9 / D3Q27).

The compute path is libm3lb_hip.so (hand-written gfx950 HIP kernels behind the
C-ABI declared in include/m3lb.h); this package is its host-side mirror of the
reference's FluidSolver API.  Importing it loads nothing; the first solver
created loads the library and raises if it has not been built.
"""
from . import abi  # noqa: F401
