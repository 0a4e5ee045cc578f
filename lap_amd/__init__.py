"""lap_amd -- PCG for sparse SPD graph-Laplacian systems on AMD Instinct MI355X.

The compute path is liblapgfx.so (hand-written HIP for gfx950, C ABI in
include/lap_gfx950.h).  ``lap_amd.solver`` is the torch-facing host layer;
``lap_amd.problems`` builds the synthetic inputs.
"""
from .lib import LapError, load  # noqa: F401

__all__ = ["LapError", "load"]
