"""ctypes binding of the FP64 CPU oracle (oracle/cfs_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  Nothing in cfs_plume_amd/ imports
it.  Each wrapper names the reference function it restates; the C++ source
carries the file:line citations.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libcfs_oracle.so")
BC = 512

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build() -> str:
    """Compile the oracle with its Makefile (g++), returning the library path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    c_int, c_double, c_void_p, c_char_p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_char_p
    sig = {
        "orc_zigzag_tables": (None, [_i32p, _i32p]),
        "orc_damping_list": (c_int, [c_int, _f64p]),
        "orc_modified_cumsum": (None, [_i32p, c_int, _i32p]),
        "orc_get_paddings": (None, [_i32p, _i32p]),
        "orc_compute_block_number": (c_int, [c_int, _i32p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
        "orc_rle_encode": (c_int, [_i32p, _i32p]),
        "orc_rle_decode": (c_int, [_i32p, c_int, _i32p, ctypes.POINTER(c_int)]),
        "orc_zigzag_flatten": (None, [_i32p, _i32p]),
        "orc_zigzag_unflatten": (None, [_i32p, _i32p]),
        "orc_dct_matrix": (None, [_f64p]),
        "orc_dct3": (None, [_f64p, c_int, c_int]),
        "orc_get_blocks": (None, [_f64p, c_int, c_int, c_int, c_int, _f64p]),
        "orc_assimilate_unpadded": (None, [_f64p, c_int, c_int, c_int, _f64p]),
        "orc_encode_block": (c_int, [_f64p, c_int, c_double, c_int, ctypes.POINTER(c_double), ctypes.POINTER(c_double), _i32p]),
        "orc_compress_matrix": (c_int, [_f64p, c_int, c_int, c_int, c_int, c_int, c_double, c_int, c_char_p]),
        "orc_basis_open": (c_void_p, [c_char_p, c_char_p, c_char_p, ctypes.POINTER(c_int)]),
        "orc_basis_close": (None, [c_void_p]),
        "orc_basis_open_mem": (c_void_p, [c_char_p, ctypes.c_size_t, c_char_p, ctypes.c_size_t, c_char_p,
                                          ctypes.c_size_t, ctypes.POINTER(c_int)]),
        "orc_basis_info": (None, [c_void_p, _i64p]),
        "orc_decode_block_ints": (c_int, [c_void_p, c_int, c_int, c_int, _i32p, ctypes.POINTER(c_int)]),
        "orc_decode_block_coeffs": (c_int, [c_void_p, c_int, c_int, c_int, c_int, _f64p]),
        "orc_decode_vector_field": (c_int, [c_void_p, c_int, _f64p]),
        "orc_decode_vector_fields_mt": (c_int, [c_void_p, c_int, c_int, c_int, _f64p]),
        "orc_project": (c_int, [c_void_p, _f64p, _f64p, c_int, c_int]),
        "orc_project_mt": (c_int, [c_void_p, _f64p, _f64p, c_int, c_int, c_int]),
        "orc_unproject_mt": (c_int, [c_void_p, _f64p, _f64p, c_int, c_int, c_int]),
        "orc_unproject": (c_int, [c_void_p, _f64p, _f64p, c_int, c_int]),
        "orc_get_submatrix": (c_int, [c_void_p, c_int, _f64p]),
        "orc_untransform_svd": (None, [_f64p, _f64p, ctypes.c_int64]),
        "orc_transform_svd": (None, [_f64p, _f64p, ctypes.c_int64]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


class OracleError(RuntimeError):
    pass


def _check(st: int, what: str):
    if st != 0:
        raise OracleError(f"{what} failed with status {st}")


# --- tables ----------------------------------------------------------------
def zigzag_tables():
    """set_zigzagArray: (zigzag[x+8y+64z] = j, reverse[j] = x+8y+64z)."""
    z = np.zeros(BC, np.int32)
    r = np.zeros(BC, np.int32)
    lib().orc_zigzag_tables(z, r)
    return z, r


def damping_list(n_bits: int) -> np.ndarray:
    """set_dampingArrayList: row k = (1+u+v+w)^(k/4), k = 0..4*nBits."""
    out = np.zeros((4 * n_bits + 1) * BC, np.float64)
    lib().orc_damping_list(n_bits, out)
    return out.reshape(4 * n_bits + 1, BC)


def modified_cumsum(v) -> np.ndarray:
    v = np.ascontiguousarray(v, np.int32)
    out = np.zeros_like(v)
    lib().orc_modified_cumsum(v, len(v), out)
    return out


def get_paddings(dims) -> tuple:
    d = np.ascontiguousarray(dims, np.int32)
    p = np.zeros(3, np.int32)
    lib().orc_get_paddings(d, p)
    return tuple(int(x) for x in p)


def compute_block_number(row: int, dims) -> tuple:
    d = np.ascontiguousarray(dims, np.int32)
    bn, bi = ctypes.c_int32(), ctypes.c_int32()
    _check(lib().orc_compute_block_number(row, d, ctypes.byref(bn), ctypes.byref(bi)), "ComputeBlockNumber")
    return bn.value, bi.value


# --- RLE / zigzag ---------------------------------------------------------------
def rle_encode(zz) -> np.ndarray:
    zz = np.ascontiguousarray(zz, np.int32)
    assert zz.shape == (BC,)
    out = np.zeros(3 * BC, np.int32)
    n = lib().orc_rle_encode(zz, out)
    return out[:n].copy()


def rle_decode(tokens):
    """RunLengthDecodeBinary: returns (ints[512] in zigzag order, underfull)."""
    t = np.ascontiguousarray(tokens, np.int32)
    pad = np.zeros(len(t) + 4, np.int32)
    pad[: len(t)] = t
    out = np.zeros(BC, np.int32)
    uf = ctypes.c_int(0)
    _check(lib().orc_rle_decode(pad, len(t), out, ctypes.byref(uf)), "RunLengthDecodeBinary")
    return out, bool(uf.value)


def rle_decode_status(tokens) -> int:
    t = np.ascontiguousarray(tokens, np.int32)
    pad = np.zeros(len(t) + 4, np.int32)
    pad[: len(t)] = t
    out = np.zeros(BC, np.int32)
    uf = ctypes.c_int(0)
    return lib().orc_rle_decode(pad, len(t), out, ctypes.byref(uf))


def zigzag_flatten(field) -> np.ndarray:
    f = np.ascontiguousarray(field, np.int32)
    out = np.zeros(BC, np.int32)
    lib().orc_zigzag_flatten(f, out)
    return out


def zigzag_unflatten(zz) -> np.ndarray:
    z = np.ascontiguousarray(zz, np.int32)
    out = np.zeros(BC, np.int32)
    lib().orc_zigzag_unflatten(z, out)
    return out


# --- DCT / geometry -------------------------------------------------------------
def dct_matrix() -> np.ndarray:
    c = np.zeros(64, np.float64)
    lib().orc_dct_matrix(c)
    return c.reshape(8, 8)


def dct3(blocks, direction: int) -> np.ndarray:
    """UnitaryBlockDCTEigen(direction) on (n, 512) x-fastest blocks."""
    b = np.array(blocks, np.float64, copy=True).reshape(-1)
    lib().orc_dct3(b, len(b) // BC, direction)
    return b.reshape(-1, BC)


def get_blocks(field, dims, zero_pad: bool) -> np.ndarray:
    X, Y, Z = dims
    f = np.ascontiguousarray(field, np.float64).reshape(-1)
    assert f.size == X * Y * Z
    nB = 1
    for n in dims:
        nB *= (n + (8 - n % 8) % 8) // 8
    out = np.zeros(nB * BC, np.float64)
    lib().orc_get_blocks(f, X, Y, Z, int(zero_pad), out)
    return out.reshape(nB, BC)


def assimilate_unpadded(blocks, dims) -> np.ndarray:
    X, Y, Z = dims
    b = np.ascontiguousarray(blocks, np.float64).reshape(-1)
    out = np.zeros(X * Y * Z, np.float64)
    lib().orc_assimilate_unpadded(b, X, Y, Z, out)
    return out


# --- encoder ----------------------------------------------------------------------
def encode_block(F, n_bits=32, percent=0.9995, max_iterations=50):
    """PreprocessBlock + TuneGammaQuantized + EncodeBlock + zigzag + RLE."""
    f = np.ascontiguousarray(F, np.float64)
    s, g = ctypes.c_double(), ctypes.c_double()
    tok = np.zeros(3 * BC, np.int32)
    n = lib().orc_encode_block(f, n_bits, percent, max_iterations, ctypes.byref(s), ctypes.byref(g), tok)
    return s.value, g.value, tok[:n].copy()


def compress_matrix(U_colmajor, dims, n_bits, percent, max_iterations, prefix: str):
    """CompressAndWriteMatrixComponents: U is (r, 3*X*Y*Z) C-order == column-major."""
    X, Y, Z = dims
    U = np.ascontiguousarray(U_colmajor, np.float64)
    r = U.shape[0]
    assert U.shape[1] == 3 * X * Y * Z
    _check(lib().orc_compress_matrix(U.reshape(-1), X, Y, Z, r, n_bits, percent, max_iterations,
                                     prefix.encode()), "CompressAndWriteMatrixComponents")
    return [prefix + str(k) for k in range(3)]


# --- decoder / basis ------------------------------------------------------------------
class Basis:
    """Three component files read as by ReadBinaryFileToMemory x3."""

    def __init__(self, paths=None, buffers=None):
        st = ctypes.c_int(0)
        if buffers is not None:
            b = [bytes(x) for x in buffers]
            h = lib().orc_basis_open_mem(b[0], len(b[0]), b[1], len(b[1]), b[2], len(b[2]), ctypes.byref(st))
        else:
            h = lib().orc_basis_open(*(p.encode() for p in paths), ctypes.byref(st))
        if not h:
            raise OracleError(f"ReadBinaryFileToMemory failed with status {st.value}")
        self._h = h
        info = np.zeros(12, np.int64)
        lib().orc_basis_info(h, info)
        self.n_bits, X, Y, Z, Px, Py, Pz, self.r, self.nB = (int(x) for x in info[:9])
        self.dims = (X, Y, Z)
        self.padded = (Px, Py, Pz)
        self.tokens = tuple(int(x) for x in info[9:12])

    def close(self):
        if self._h:
            lib().orc_basis_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def decode_ints(self, comp, block, col):
        out = np.zeros(BC, np.int32)
        uf = ctypes.c_int(0)
        _check(lib().orc_decode_block_ints(self._h, comp, block, col, out, ctypes.byref(uf)), "decode ints")
        return out

    def decode_coeffs(self, comp, block, col, table=False):
        out = np.zeros(BC, np.float64)
        _check(lib().orc_decode_block_coeffs(self._h, comp, block, col, int(table), out), "decode coeffs")
        return out

    def decode_vector_field(self, col):
        X, Y, Z = self.dims
        out = np.zeros(3 * X * Y * Z, np.float64)
        _check(lib().orc_decode_vector_field(self._h, col, out), "DecodeVectorField")
        return out

    def decode_vector_fields_mt(self, col0, col1, threads):
        """decode_vector_field for columns [col0, col1) on `threads` threads
        (identical results), as an array [col1 - col0][3*X*Y*Z]."""
        X, Y, Z = self.dims
        out = np.zeros((col1 - col0, 3 * X * Y * Z), np.float64)
        _check(lib().orc_decode_vector_fields_mt(self._h, col0, col1, threads, out), "DecodeVectorField (threads)")
        return out

    def project(self, V_full, col0=0, col1=None):
        col1 = self.r if col1 is None else col1
        V = np.ascontiguousarray(V_full, np.float64).reshape(-1)
        X, Y, Z = self.dims
        assert V.size == 3 * (X + 2) * (Y + 2) * (Z + 2)
        q = np.zeros(col1 - col0, np.float64)
        _check(lib().orc_project(self._h, V, q, col0, col1), "PeeledCompressedProjectTransformNoSVD")
        return q

    def unproject(self, q, V_full, col0=0, col1=None):
        col1 = self.r if col1 is None else col1
        qq = np.ascontiguousarray(q, np.float64)
        assert qq.size == col1 - col0
        V = np.array(V_full, np.float64, copy=True).reshape(-1)
        _check(lib().orc_unproject(self._h, qq, V, col0, col1), "PeeledCompressedUnprojectTransform")
        return V

    def project_mt(self, V_full, threads, col0=0, col1=None):
        """project() split over `threads` threads (identical results)."""
        col1 = self.r if col1 is None else col1
        V = np.ascontiguousarray(V_full, np.float64).reshape(-1)
        q = np.zeros(col1 - col0, np.float64)
        _check(lib().orc_project_mt(self._h, V, q, col0, col1, threads), "project (threads)")
        return q

    def unproject_mt(self, q, V_full, threads, col0=0, col1=None):
        """unproject() split over `threads` threads (identical results)."""
        col1 = self.r if col1 is None else col1
        qq = np.ascontiguousarray(q, np.float64)
        V = np.array(V_full, np.float64, copy=True).reshape(-1)
        _check(lib().orc_unproject_mt(self._h, qq, V, col0, col1, threads), "unproject (threads)")
        return V

    def get_submatrix(self, start_row):
        out = np.zeros(3 * self.r, np.float64)
        _check(lib().orc_get_submatrix(self._h, start_row, out), "GetSubmatrixNoSVDSparse")
        return out.reshape(self.r, 3).T.copy()  # 3 x r


def untransform_svd(Vm, field):
    f = np.array(field, np.float64, copy=True).reshape(-1)
    lib().orc_untransform_svd(np.ascontiguousarray(Vm, np.float64).reshape(-1), f, f.size // 3)
    return f


def transform_svd(Vm, field):
    f = np.array(field, np.float64, copy=True).reshape(-1)
    lib().orc_transform_svd(np.ascontiguousarray(Vm, np.float64).reshape(-1), f, f.size // 3)
    return f
