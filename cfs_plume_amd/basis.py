"""Python host interface over the C ABI, mirroring the reference's call surface.

Reference (aadjones/CFS_PLUME, src/compression/COMPRESSION.h) -> here:

  ReadBinaryFileToMemory x3 + MATRIX_COMPRESSION_DATA   CompressedBasis.from_files / from_buffers
  PeeledCompressedProjectTransformNoSVD                 CompressedBasis.project
  PeeledCompressedUnprojectTransform                    CompressedBasis.unproject
  DecodeVectorField                                     CompressedBasis.decode_columns
  RunLengthDecodeBinary + ZigzagUnflatten               CompressedBasis.decode_block_ints
  DecodeBlockWithCompressionData                        CompressedBasis.decode_block_coeffs
  GetSubmatrixNoSVDSparse                               CompressedBasis.get_submatrix
  stepPlume's project/unproject pair                    step
  CompressAndWriteMatrixComponents                      encode_matrix

Arrays passed to the per-step operators are hip.DeviceArray objects (FP32
fields, FP64 q).  Errors raise CfsError where the reference asserted,
returned NULL or called exit().
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from . import _lib
from .hip import DeviceArray, Stream

STREAM_BITS = {
    0x2: "a stream did not decode to 512 values per segment (device memory corrupted?)",
}


def _sptr(stream: Optional[Stream]):
    return stream.handle if stream is not None else None


class EncodedBasis:
    """Three in-memory component files produced by the encoder."""

    def __init__(self, handle):
        self._h = handle

    def buffer(self, comp: int) -> bytes:
        data, size = ctypes.c_void_p(), ctypes.c_size_t()
        _lib.check(_lib.lib().cfs_encoded_buffer(self._h, comp, ctypes.byref(data), ctypes.byref(size)), "encoded buffer")
        return ctypes.string_at(data.value, size.value)

    def buffers(self):
        return [self.buffer(k) for k in range(3)]

    def _views(self):
        out = []
        for k in range(3):
            data, size = ctypes.c_void_p(), ctypes.c_size_t()
            _lib.check(_lib.lib().cfs_encoded_buffer(self._h, k, ctypes.byref(data), ctypes.byref(size)))
            out.append((data.value, size.value))
        return out

    def write(self, prefix: str):
        """Writes <prefix>0, <prefix>1, <prefix>2."""
        _lib.check(_lib.lib().cfs_encoded_write(self._h, prefix.encode()), "write component files")
        return [prefix + str(k) for k in range(3)]

    def open(self) -> "CompressedBasis":
        return CompressedBasis._open_views(self._views())

    def close(self):
        if self._h:
            _lib.lib().cfs_encoded_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def encode_block(F: np.ndarray, n_bits: int = 32, percent: float = 0.9995, max_iterations: int = 50):
    """EncodeBlock (with PreprocessBlock and TuneGammaQuantized) of one
    forward-DCT block F (512 FP64, u+8v+64w order): returns (s, gamma,
    RLE tokens of the zigzag-ordered ints)."""
    B = np.ascontiguousarray(F, dtype=np.float64).reshape(-1)
    if B.size != 512:
        raise ValueError("a block holds 512 coefficients")
    s, g, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
    tok = np.zeros(768, np.int32)
    _lib.check(_lib.lib().cfs_encode_block(B.ctypes.data, n_bits, float(percent), max_iterations, ctypes.byref(s),
                                           ctypes.byref(g), tok.ctypes.data, ctypes.byref(n)), "EncodeBlock")
    return s.value, g.value, tok[:n.value].copy()


def encode_matrix(U: np.ndarray, dims, n_bits: int = 32, percent: float = 0.9995, max_iterations: int = 50,
                  threads: int = 0) -> EncodedBasis:
    """CompressAndWriteMatrixComponents on U given as (r, 3*X*Y*Z) FP64 rows
    (row c = column c of the reference's column-major MatrixXd)."""
    X, Y, Z = (int(d) for d in dims)
    U = np.ascontiguousarray(U, dtype=np.float64)
    if U.ndim != 2 or U.shape[1] != 3 * X * Y * Z:
        raise ValueError(f"U must be (r, {3 * X * Y * Z}), got {U.shape}")
    h = ctypes.c_void_p()
    _lib.check(_lib.lib().cfs_encode_matrix(U.ctypes.data, X, Y, Z, U.shape[0], n_bits, percent, max_iterations,
                                            threads, ctypes.byref(h)), "encode")
    return EncodedBasis(h.value)


class CompressedBasis:
    """A compressed basis resident on the GPU (MATRIX_COMPRESSION_DATA analogue)."""

    def __init__(self, handle):
        self._h = handle
        info = _lib.BasisInfo()
        _lib.check(_lib.lib().cfs_basis_get_info(self._h, ctypes.byref(info)), "basis info")
        self.n_bits = info.n_bits
        self.dims = tuple(info.dims)
        self.padded = tuple(info.padded)
        self.r = info.num_cols
        self.nB = info.num_blocks
        self.tokens = tuple(info.tokens)
        self.device_bytes = info.device_bytes
        self.hazards = info.hazards
        X, Y, Z = self.dims
        self.full_cells = (X + 2) * (Y + 2) * (Z + 2)
        self.peeled_cells = X * Y * Z

    @classmethod
    def from_files(cls, paths: Sequence[str]) -> "CompressedBasis":
        h = ctypes.c_void_p()
        p = [str(x).encode() for x in paths]
        _lib.check(_lib.lib().cfs_basis_open_files(p[0], p[1], p[2], ctypes.byref(h)), "ReadBinaryFileToMemory")
        return cls(h.value)

    @classmethod
    def from_buffers(cls, bufs: Sequence[bytes]) -> "CompressedBasis":
        keep = [ctypes.create_string_buffer(bytes(b), len(b)) for b in bufs]
        return cls._open_views([(ctypes.addressof(k), len(b)) for k, b in zip(keep, bufs)])

    @classmethod
    def _open_views(cls, views) -> "CompressedBasis":
        h = ctypes.c_void_p()
        (p0, n0), (p1, n1), (p2, n2) = views
        _lib.check(_lib.lib().cfs_basis_open_memory(p0, n0, p1, n1, p2, n2, ctypes.byref(h)), "open basis")
        return cls(h.value)

    # ---- algorithmic bytes of one fused step over this basis ------------------
    def step_bytes(self) -> int:
        """Reference-format bytes read for one decode (tokens + 24 B metadata per
        segment) plus the FP32 field in and out and q in and out."""
        segs = 3 * self.r * self.nB
        return 4 * sum(self.tokens) + 24 * segs + 2 * 4 * 3 * self.peeled_cells + 16 * self.r

    # ---- operators -------------------------------------------------------------
    def project(self, v: DeviceArray, q: Optional[DeviceArray] = None, stream: Optional[Stream] = None) -> DeviceArray:
        assert v.dtype == np.float32 and v.nbytes == 12 * self.full_cells
        if q is None:
            q = DeviceArray((self.r,), np.float64)
        _lib.check(_lib.lib().cfs_project(self._h, v.ptr, q.ptr, _sptr(stream)), "project")
        return q

    def unproject(self, q: DeviceArray, field: DeviceArray, stream: Optional[Stream] = None) -> DeviceArray:
        assert q.dtype == np.float64 and q.nbytes == 8 * self.r
        assert field.dtype == np.float32 and field.nbytes == 12 * self.full_cells
        _lib.check(_lib.lib().cfs_unproject(self._h, q.ptr, field.ptr, _sptr(stream)), "unproject")
        return field

    def unproject_spatial(self, q: DeviceArray, field: DeviceArray, group: int = 8,
                          scratch: Optional[DeviceArray] = None, stream: Optional[Stream] = None) -> DeviceArray:
        """PeeledCompressedUnproject (COMPRESSION.cpp:2677-2711): the columns
        decoded spatially `group` at a time and summed with weights q into the
        interior of the full-grid field (boundary cells untouched)."""
        assert q.dtype == np.float64 and q.nbytes == 8 * self.r
        assert field.dtype == np.float32 and field.nbytes == 12 * self.full_cells
        g = min(group, self.r)
        if scratch is not None:
            assert scratch.dtype == np.float32 and scratch.nbytes >= 12 * g * self.peeled_cells
        _lib.check(_lib.lib().cfs_unproject_spatial(self._h, q.ptr, field.ptr, scratch.ptr if scratch else None,
                                                    group, _sptr(stream)), "PeeledCompressedUnproject")
        return field

    def decode_columns(self, col0: int = 0, ncols: Optional[int] = None, out: Optional[DeviceArray] = None,
                       stream: Optional[Stream] = None) -> DeviceArray:
        ncols = self.r - col0 if ncols is None else ncols
        if out is None:
            out = DeviceArray((ncols, 3 * self.peeled_cells), np.float32)
        _lib.check(_lib.lib().cfs_decode_columns(self._h, col0, ncols, out.ptr, _sptr(stream)), "DecodeVectorField")
        return out

    def untransform_svd(self, columns: DeviceArray, V: DeviceArray, stream: Optional[Stream] = None) -> DeviceArray:
        """UntransformVectorFieldSVD on decoded columns (ncols x 3*cells FP32,
        in place) with one row-major 3 x 3 FP64 matrix per column in V."""
        ncols = columns.shape[0] if len(columns.shape) > 1 else 1
        cells = columns.nbytes // columns.dtype.itemsize // (3 * ncols)
        if V.nbytes != 72 * ncols or V.dtype != np.float64:
            raise ValueError("V must hold one FP64 3 x 3 matrix per column")
        _lib.check(_lib.lib().cfs_untransform_svd(columns.ptr, ncols, cells, V.ptr, _sptr(stream)),
                   "UntransformVectorFieldSVD")
        return columns

    def decode_block_ints(self, comp: int, col: int, block0: int = 0, nblocks: Optional[int] = None,
                          stream: Optional[Stream] = None) -> np.ndarray:
        nblocks = self.nB - block0 if nblocks is None else nblocks
        out = DeviceArray((nblocks, 512), np.int32)
        _lib.check(_lib.lib().cfs_decode_block_ints(self._h, comp, col, block0, nblocks, out.ptr, _sptr(stream)),
                   "decode ints")
        return out.numpy()

    def decode_block_coeffs(self, comp: int, col: int, block0: int = 0, nblocks: Optional[int] = None,
                            stream: Optional[Stream] = None) -> np.ndarray:
        nblocks = self.nB - block0 if nblocks is None else nblocks
        out = DeviceArray((nblocks, 512), np.float32)
        _lib.check(_lib.lib().cfs_decode_block_coeffs(self._h, comp, col, block0, nblocks, out.ptr, _sptr(stream)),
                   "decode coeffs")
        return out.numpy()

    def get_submatrix(self, start_row: int, stream: Optional[Stream] = None) -> np.ndarray:
        out = DeviceArray((self.r, 3), np.float64)
        _lib.check(_lib.lib().cfs_get_submatrix(self._h, start_row, out.ptr, _sptr(stream)), "GetSubmatrixNoSVDSparse")
        return out.numpy().T.copy()  # 3 x r

    def check_stream_errors(self) -> int:
        bits = ctypes.c_uint32(0)
        _lib.check(_lib.lib().cfs_check_stream_errors(self._h, ctypes.byref(bits)), "check stream errors")
        return bits.value

    def raise_on_stream_errors(self):
        bits = self.check_stream_errors()
        if bits:
            names = [v for k, v in STREAM_BITS.items() if bits & k]
            raise _lib.CfsError(-5, "malformed token stream: " + "; ".join(names))

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().cfs_basis_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def step(pre: CompressedBasis, fin: CompressedBasis, v_in: DeviceArray, q_out: DeviceArray, q_in: DeviceArray,
         field: DeviceArray, stream: Optional[Stream] = None):
    """stepPlume's pair: q_out = U_pre^T v_in, field interior = U_fin q_in."""
    _lib.check(_lib.lib().cfs_step(pre._h, fin._h, v_in.ptr, q_out.ptr, q_in.ptr, field.ptr, _sptr(stream)), "step")


class BasisBatch:
    """Several bases stepped in one persistent launch (cfs_batch_*).

    bind() gives every basis its v_in, q_out, q_in and field; step() runs the
    fused step of all of them; tune() runs the autotuner; capture() records
    one step into a HIP graph that replay() launches."""

    def __init__(self, bases: Sequence[CompressedBasis]):
        self.bases = list(bases)
        arr = (ctypes.c_void_p * len(self.bases))(*[b._h for b in self.bases])
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().cfs_batch_create(arr, len(self.bases), ctypes.byref(h)), "cfs_batch_create")
        self._h = h.value
        self._bound = None

    def bind(self, v_in, q_out, q_in, field):
        n = len(self.bases)
        for b, v, qo, qi, f in zip(self.bases, v_in, q_out, q_in, field):
            assert v.dtype == np.float32 and v.nbytes == 12 * b.full_cells
            assert f.dtype == np.float32 and f.nbytes == 12 * b.full_cells
            assert qo.dtype == np.float64 and qo.nbytes == 8 * b.r and qi.dtype == np.float64 and qi.nbytes == 8 * b.r
        ptrs = [(ctypes.c_void_p * n)(*[x.ptr for x in arrs]) for arrs in (v_in, q_out, q_in, field)]
        _lib.check(_lib.lib().cfs_batch_bind(self._h, *ptrs), "cfs_batch_bind")
        self._bound = (list(v_in), list(q_out), list(q_in), list(field))  # keep the arrays alive

    def step(self, stream: Optional[Stream] = None):
        _lib.check(_lib.lib().cfs_batch_step(self._h, _sptr(stream)), "cfs_batch_step")

    def tune(self, reps: int = 3, stream: Optional[Stream] = None):
        _lib.check(_lib.lib().cfs_batch_tune(self._h, reps, _sptr(stream)), "cfs_batch_tune")
        return self.config()

    def config(self):
        g, w, iw, wpw = ctypes.c_int(), ctypes.c_int(), ctypes.c_float(), ctypes.c_int()
        _lib.check(_lib.lib().cfs_batch_get_config(self._h, ctypes.byref(g), ctypes.byref(w), ctypes.byref(iw)))
        _lib.check(_lib.lib().cfs_batch_get_shape(self._h, ctypes.byref(wpw)))
        return {"grid": g.value, "wg_per_cu": w.value, "item_weight": iw.value, "waves_per_wg": wpw.value}

    def set_config(self, wg_per_cu: int, item_weight: float):
        _lib.check(_lib.lib().cfs_batch_set_config(self._h, wg_per_cu, item_weight), "cfs_batch_set_config")

    def set_shape(self, waves_per_wg: int):
        _lib.check(_lib.lib().cfs_batch_set_shape(self._h, waves_per_wg), "cfs_batch_set_shape")

    def capture(self, stream: Stream):
        _lib.check(_lib.lib().cfs_batch_graph_capture(self._h, _sptr(stream)), "cfs_batch_graph_capture")

    def replay(self, stream: Stream):
        _lib.check(_lib.lib().cfs_batch_graph_launch(self._h, _sptr(stream)), "cfs_batch_graph_launch")

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().cfs_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
