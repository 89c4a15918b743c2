"""MI355X-native screened Poisson surface reconstruction (drop-in for mojo_poissonrecon).

Same public surface as the reference package (python/mojo_poissonrecon/__init__.py:17-28);
the kernels run as hand-written HIP on gfx950 through libmpr_hip.so.
"""

from ._lib import build
from .core import (
    FlatOctree,
    GridSolution,
    Job,
    Mesh,
    build_octree,
    extract_isosurface,
    oriented_points_from_mesh,
    reconstruct,
    reconstruct_mesh,
    release_device_memory,
    solve_screened_poisson,
)

__version__ = "0.1.0"
__all__ = [
    "FlatOctree",
    "GridSolution",
    "Mesh",
    "build",
    "build_octree",
    "extract_isosurface",
    "oriented_points_from_mesh",
    "reconstruct",
    "reconstruct_mesh",
    "solve_screened_poisson",
]
