"""MI355X (gfx950) voxel hot path for MesoEngine: seeded voxelization, brick-map
DDA raymarching and culled/greedy chunk meshing behind the C-ABI of
include/meso_hip.h.  This package is the Python harness over that library."""
from . import _lib, runtime, workloads  # noqa: F401
from ._lib import MesoError  # noqa: F401
from .runtime import (Context, Server, camera, chunk_source, mesh_chunk, mesh_chunks,  # noqa: F401
                      raymarch, shading)
