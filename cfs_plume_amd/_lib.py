"""Loader for libcfs_plume.so (the C ABI declared in include/cfs_plume.h).

The library is built in-tree by cfs_plume_amd/build.py.  There is no fallback:
if the library is missing or cannot be loaded, every call raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcfs_plume.so")

CFS_OK = 0
STATUS_NAMES = {
    -1: "CFS_ERR_ARG", -2: "CFS_ERR_IO", -3: "CFS_ERR_FORMAT", -4: "CFS_ERR_HIP",
    -5: "CFS_ERR_STREAM", -6: "CFS_ERR_NOMEM",
}

# Every symbol declared in include/cfs_plume.h, with (restype, argtypes).
_vp, _i, _sz, _cp = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, ctypes.c_char_p
SIGNATURES = {
    "cfs_last_error": (ctypes.c_char_p, []),
    "cfs_abi_version": (_i, []),
    "cfs_basis_open_files": (_i, [_cp, _cp, _cp, ctypes.POINTER(_vp)]),
    "cfs_basis_open_memory": (_i, [_vp, _sz, _vp, _sz, _vp, _sz, ctypes.POINTER(_vp)]),
    "cfs_basis_close": (_i, [_vp]),
    "cfs_basis_get_info": (_i, [_vp, _vp]),
    "cfs_project": (_i, [_vp, _vp, _vp, _vp]),
    "cfs_unproject": (_i, [_vp, _vp, _vp, _vp]),
    "cfs_step": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "cfs_decode_columns": (_i, [_vp, _i, _i, _vp, _vp]),
    "cfs_unproject_spatial": (_i, [_vp, _vp, _vp, _vp, _i, _vp]),
    "cfs_untransform_svd": (_i, [_vp, _i, ctypes.c_int64, _vp, _vp]),
    "cfs_decode_block_ints": (_i, [_vp, _i, _i, _i, _i, _vp, _vp]),
    "cfs_decode_block_coeffs": (_i, [_vp, _i, _i, _i, _i, _vp, _vp]),
    "cfs_get_submatrix": (_i, [_vp, _i, _vp, _vp]),
    "cfs_check_stream_errors": (_i, [_vp, ctypes.POINTER(ctypes.c_uint32)]),
    "cfs_batch_create": (_i, [ctypes.POINTER(_vp), _i, ctypes.POINTER(_vp)]),
    "cfs_batch_bind": (_i, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "cfs_batch_step": (_i, [_vp, _vp]),
    "cfs_batch_tune": (_i, [_vp, _i, _vp]),
    "cfs_batch_get_config": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(ctypes.c_float)]),
    "cfs_batch_set_config": (_i, [_vp, _i, ctypes.c_float]),
    "cfs_batch_get_shape": (_i, [_vp, ctypes.POINTER(_i)]),
    "cfs_batch_set_shape": (_i, [_vp, _i]),
    "cfs_batch_graph_capture": (_i, [_vp, _vp]),
    "cfs_batch_graph_launch": (_i, [_vp, _vp]),
    "cfs_batch_destroy": (_i, [_vp]),
    "cfs_reduced_advect": (_i, [_vp, _vp, _vp, _vp, _vp, ctypes.c_float, ctypes.c_double, _vp, _vp]),
    "cfs_reduced_advect_noalloc": (_i, [_vp, _vp, _vp, ctypes.c_float, ctypes.c_double, _vp, _vp]),
    "cfs_cubature_create": (_i, [_i, _vp, _vp, _vp, ctypes.POINTER(_vp)]),
    "cfs_cubature_create_rows": (_i, [_i, _vp, _vp, _i, _i, _i, _i, _vp, ctypes.POINTER(_vp)]),
    "cfs_cubature_open_file": (_i, [_cp, _vp, ctypes.POINTER(_vp)]),
    "cfs_cubature_write_file": (_i, [_cp, _i, _vp, _vp]),
    "cfs_cubature_get_info": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "cfs_cubature_get_cells": (_i, [_vp, _vp, _vp]),
    "cfs_cubature_advect": (_i, [_vp, _vp, _vp, _vp, _vp, ctypes.c_float, ctypes.c_double, _vp, _vp]),
    "cfs_cubature_advect_field": (_i, [_vp, _vp, _vp, ctypes.c_float, ctypes.c_double, _vp, _vp]),
    "cfs_cubature_destroy": (_i, [_vp]),
    "cfs_nnls": (_i, [_vp, _i, _i, _vp, _i, _vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_i)]),
    "cfs_advect_field": (_i, [_vp, _vp, _i, _i, _i, ctypes.c_float, _vp]),
    "cfs_advect_scalar": (_i, [_vp, _vp, _vp, _i, _i, _i, ctypes.c_float, _vp]),
    "cfs_closure_create": (_i, [_i, _i, _i, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    "cfs_closure_destroy": (_i, [_vp]),
    "cfs_closure_get_dims": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i),
                                  ctypes.POINTER(ctypes.c_int64)]),
    "cfs_closure_get_params": (_i, [_vp, _vp]),
    "cfs_closure_set_params": (_i, [_vp, _vp]),
    "cfs_closure_set_scales": (_i, [_vp, _vp, _vp]),
    "cfs_closure_get_scales": (_i, [_vp, _vp, _vp]),
    "cfs_closure_set_linear": (_i, [_vp, _vp]),
    "cfs_closure_forward": (_i, [_vp, _vp, _vp, _i, _vp]),
    "cfs_closure_gradients": (_i, [_vp, _vp, _vp, _i, _vp, _vp, _vp]),
    "cfs_closure_train_step": (_i, [_vp, _vp, _vp, _vp, _i, ctypes.c_float, _vp, _vp]),
    "cfs_closure_step": (_i, [_vp, _vp, _vp, _i, ctypes.c_double, _vp, _vp]),
    "cfs_plume_default_params": (None, [_i, _i, _i, _vp]),
    "cfs_plume_create": (_i, [_vp, _vp, _i, _i, _vp, _vp, _vp, _vp, ctypes.POINTER(_vp)]),
    "cfs_plume_create_dense": (_i, [_vp, _i, _vp, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, ctypes.POINTER(_vp)]),
    "cfs_plume_set_exact_after": (_i, [_vp, _vp]),
    "cfs_plume_set_advection": (_i, [_vp, _i, _vp, _vp]),
    "cfs_plume_add_smoke": (_i, [_vp, _vp, _vp, _vp]),
    "cfs_plume_step": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "cfs_plume_set_params": (_i, [_vp, _vp]),
    "cfs_plume_advection_state": (_i, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "cfs_plume_destroy": (_i, [_vp]),
    "cfs_fluid_default_params": (None, [_i, _i, _i, _vp]),
    "cfs_fluid_create": (_i, [_i, _i, _i, _vp, ctypes.POINTER(_vp)]),
    "cfs_fluid_fields": (_i, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                              ctypes.POINTER(_vp)]),
    "cfs_fluid_add_smoke": (_i, [_vp, _vp]),
    "cfs_fluid_step": (_i, [_vp, _vp, _vp]),
    "cfs_fluid_last_solve": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(ctypes.c_double)]),
    "cfs_fluid_destroy": (_i, [_vp]),
    "cfs_box_init": (_i, [_vp, _vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, _vp]),
    "cfs_box_advance": (_i, [_vp]),
    "cfs_fluid_set_initial_velocity": (_i, [_vp, _vp, _vp]),
    "cfs_fluid_step_obstacle": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "cfs_reduced_iop_scratch_bytes": (ctypes.c_int64, [_i, _i, _i]),
    "cfs_reduced_iop": (_i, [_vp, _i, _vp, _i, _vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp]),
    "cfs_plume_set_obstacle": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp]),
    "cfs_plume_step_obstacle": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "cfs_gemm_tn": (_i, [_vp, ctypes.c_int64, _i, ctypes.c_int64, _vp, _i, ctypes.c_int64, _vp, _vp]),
    "cfs_gemm_nn": (_i, [_vp, ctypes.c_int64, _i, ctypes.c_int64, _vp, _i, _vp, ctypes.c_int64, _vp]),
    "cfs_apply_operator": (_i, [_i, _i, _i, _i, _vp, ctypes.c_double, ctypes.c_double, _vp, ctypes.c_int64, _i, _vp,
                                ctypes.c_int64, _vp]),
    "cfs_encode_matrix": (_i, [_vp, _i, _i, _i, _i, _i, ctypes.c_double, _i, _i, ctypes.POINTER(_vp)]),
    "cfs_encode_block": (_i, [_vp, _i, ctypes.c_double, _i, ctypes.POINTER(ctypes.c_double),
                              ctypes.POINTER(ctypes.c_double), _vp, ctypes.POINTER(_i)]),
    "cfs_encoded_buffer": (_i, [_vp, _i, ctypes.POINTER(_vp), ctypes.POINTER(_sz)]),
    "cfs_encoded_write": (_i, [_vp, _cp]),
    "cfs_encoded_free": (None, [_vp]),
}


class PlumeParams(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double),
        ("dx", ctypes.c_double),
        ("buoyancy", ctypes.c_double),
        ("vorticity_eps", ctypes.c_double),
        ("open_x", ctypes.c_int),
        ("open_y", ctypes.c_int),
        ("open_z", ctypes.c_int),
        ("smoke_offset_x", ctypes.c_double),
        ("smoke_offset_z", ctypes.c_double),
    ]


class FluidParams(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double),
        ("dx", ctypes.c_double),
        ("buoyancy", ctypes.c_double),
        ("vorticity_eps", ctypes.c_double),
        ("damping_w", ctypes.c_double),
        ("solver_eps", ctypes.c_double),
        ("iterations", ctypes.c_int),
        ("boundaries", ctypes.c_int * 6),
    ]


class Box(ctypes.Structure):
    _fields_ = [
        ("center", ctypes.c_double * 3),
        ("original_center", ctypes.c_double * 3),
        ("half_lengths", ctypes.c_double * 3),
        ("period", ctypes.c_double),
        ("translation_period", ctypes.c_double),
        ("theta", ctypes.c_double),
        ("phi", ctypes.c_double),
        ("time", ctypes.c_double),
        ("dt", ctypes.c_double),
    ]


class FluidSnapshots(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in
                ("preadvect", "prediffuse", "preproject", "final_velocity", "pressure", "divergence")]


class CfsError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class BasisInfo(ctypes.Structure):
    _fields_ = [
        ("n_bits", ctypes.c_int32),
        ("dims", ctypes.c_int32 * 3),
        ("padded", ctypes.c_int32 * 3),
        ("num_cols", ctypes.c_int32),
        ("num_blocks", ctypes.c_int32),
        ("hazards", ctypes.c_int32),
        ("tokens", ctypes.c_int64 * 3),
        ("device_bytes", ctypes.c_int64),
    ]


_lib = None


def lib() -> ctypes.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m cfs_plume_amd.build` "
                "(the HIP path has no CPU fallback)")
        h = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def check(status: int, what: str = ""):
    if status != CFS_OK:
        msg = lib().cfs_last_error().decode(errors="replace")
        raise CfsError(status, f"{what}: {msg}" if what else msg)
