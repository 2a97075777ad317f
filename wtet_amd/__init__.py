"""wtet_amd: MI355X (gfx950) 3D Delaunay triangulation, a drop-in for the
reference's ``wtet`` Python package (python/wtet/__init__.py:47-69)."""
from ._core import (  # noqa: F401
    INVALID,
    BatchResult,
    Config,
    Context,
    DelaunayResult,
    __version__,
    check_quality,
    decode_adjacency,
    delaunay,
    delaunay_batch,
    encode_adjacency,
    gpu_info,
    initialize_gpu,
)

__all__ = [
    "delaunay",
    "Config",
    "DelaunayResult",
    "decode_adjacency",
    "encode_adjacency",
    "initialize_gpu",
    "gpu_info",
    "check_quality",
    "INVALID",
    "__version__",
    "Context",
    "delaunay_batch",
    "BatchResult",
]
