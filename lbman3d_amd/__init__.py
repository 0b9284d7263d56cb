"""lbman3d_amd — MI355X-native lbman3d solver step (D3Q15 LB + Beris-Edwards Q-tensor).

The compute path is the HIP library lbman3d_amd/liblbman3d.so, reached through
its C ABI (include/lbman3d.h). This package holds the Python mirror of that ABI
(abi.py), the ctypes loader and host-side solver wrappers (solver.py) and the
build script (build.py).
"""
from .abi import *  # noqa: F401,F403
