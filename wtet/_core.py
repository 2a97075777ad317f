"""wtet._core: the names of the reference's extension module
(src/python.rs:377-402 registers delaunay, check_quality, initialize_gpu,
gpu_info, decode_adjacency, encode_adjacency, Config, DelaunayResult,
INVALID, __version__), served by the HIP library through wtet_amd._core."""
from wtet_amd._core import (  # noqa: F401
    INVALID,
    Config,
    DelaunayResult,
    __version__,
    check_quality,
    decode_adjacency,
    delaunay,
    encode_adjacency,
    gpu_info,
    initialize_gpu,
)

__all__ = [
    "delaunay",
    "check_quality",
    "initialize_gpu",
    "gpu_info",
    "decode_adjacency",
    "encode_adjacency",
    "Config",
    "DelaunayResult",
    "INVALID",
    "__version__",
]
