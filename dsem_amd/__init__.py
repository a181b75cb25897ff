"""MI355X-native DSEM hot path: the staggered-grid spectral-element RHS and
the low-storage RK4 step, as HIP kernels behind a C ABI (include/dsem_hip.h).
"""
from .mesh import Mesh, periodic_box, read_mesh, build_topology  # noqa: F401

__all__ = ["Mesh", "periodic_box", "read_mesh", "build_topology", "DSEM"]


def __getattr__(name):
    if name == "DSEM":
        from .solver import DSEM
        return DSEM
    raise AttributeError(name)
