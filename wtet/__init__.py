"""wtet: GPU 3D Delaunay triangulation, the reference package's import name.

The reference ships ``wtet`` with a compiled ``wtet._core`` module
(python/wtet/__init__.py:47-69, src/python.rs:377-402).  Here the same names
come from the MI355X library (``wtet_amd``), so ``import wtet`` code runs
unchanged:

    >>> import numpy as np, wtet
    >>> result = wtet.delaunay(np.random.rand(1000, 3).astype(np.float32))
    >>> result.num_tets, result.num_failed
"""
from ._core import (  # noqa: F401
    INVALID,
    Config,
    DelaunayResult,
    __version__,
    decode_adjacency,
    delaunay,
    encode_adjacency,
    gpu_info,
    initialize_gpu,
)

# the reference re-exports these nine names; check_quality stays in _core
__all__ = [
    "delaunay",
    "Config",
    "DelaunayResult",
    "decode_adjacency",
    "encode_adjacency",
    "initialize_gpu",
    "gpu_info",
    "INVALID",
    "__version__",
]
