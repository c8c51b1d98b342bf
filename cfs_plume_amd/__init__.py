"""MI355X-native compressed-subspace step (decode + project + unproject).

The compute path is libcfs_plume.so: hand-written gfx950 HIP kernels behind
the C ABI in include/cfs_plume.h.  This package is its Python host side.
"""
from ._lib import CfsError, lib  # noqa: F401
from .basis import BasisBatch, CompressedBasis, EncodedBasis, encode_block, encode_matrix, step  # noqa: F401
from .hip import DeviceArray, Event, Graph, Stream, gpu_available  # noqa: F401
from .closure import Closure, ReducedPlume, advect_field, advect_scalar, reduced_advect  # noqa: F401

__all__ = ["CfsError", "BasisBatch", "CompressedBasis", "EncodedBasis", "encode_matrix", "step", "DeviceArray", "Event", "Graph",
           "Stream", "gpu_available", "lib", "Closure", "ReducedPlume", "advect_field", "advect_scalar", "reduced_advect"]
