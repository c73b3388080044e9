"""wtomo3d on MI355X: elastic frequency-domain forward solve (HIP, gfx950).

Thin Python host layer over ``libwtomo3d_hip.so`` (C ABI in
``include/wtomo3d.h``).  The library is required: importing the compute
entry points without it raises, there is no CPU fallback.
"""
from ._lib import lib, LibraryMissing, grid_params, pml_params  # noqa: F401
from .assembly import graph, graph_nnz, assemble, lame_from_velocity  # noqa: F401
