"""meshes_amd: MI355X-native batched propagation through meshes of 2x2 photonic crossings.

The compute path is libmeshprop.so (hand-written HIP for gfx950, C ABI in
include/meshprop.h); this package is the host side that mirrors the reference's
``meshes.gpu`` interface (``MeshNetworkGPU``), plus the autograd wrapper (``autograd``),
the grouped launch of many meshes (``batched``, ``autotune``), the request server
(``serving``) and the bridge from the reference's CPU mesh objects (``bridge``).
"""
from ._lib import MeshpropError, MeshpropUnavailable  # noqa: F401
from .mesh import MeshNetworkGPU  # noqa: F401
from . import geometry  # noqa: F401

__all__ = ["MeshNetworkGPU", "MeshpropError", "MeshpropUnavailable", "geometry"]
