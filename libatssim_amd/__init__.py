"""MI355X-native Vertex Block Descent cloth backend (HIP, gfx950).

The product is libatssim_amd/lib/libatssim_hip.so (C ABI:
include/libatssim_hip.h); this package is its Python front-end.
"""
from . import _lib  # noqa: F401
from .vbd import Cloth, GpuVBD, Params, SolverInterface, Topology, make_grid, make_wind  # noqa: F401
from .warm_start import WarmStartMLP  # noqa: F401
