"""Host-side checks of the product library (no GPU): the C ABI exports every
symbol include/cfs_plume.h declares, and the host encoder writes the same
bytes as the oracle's restatement of CompressAndWriteMatrixComponents."""
import os
import re

import numpy as np
import pytest

from cfs_plume_amd import _lib, encode_matrix, synth
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "cfs_plume.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cfs_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("cfs_basis_open_files", "cfs_basis_open_memory", "cfs_project", "cfs_unproject", "cfs_step",
                 "cfs_decode_columns", "cfs_decode_block_ints", "cfs_get_submatrix", "cfs_encode_matrix"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    names = declared_functions()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes table covers exactly the declared set
    assert sorted(_lib.SIGNATURES) == names


def test_abi_version():
    src = open(os.path.join(ROOT, "include", "cfs_plume.h")).read()
    v = int(re.search(r"#define CFS_ABI_VERSION (\d+)", src).group(1))
    assert _lib.lib().cfs_abi_version() == v


def test_open_missing_file_reports_io_error(tmp_path):
    from cfs_plume_amd import CompressedBasis, CfsError
    with pytest.raises(CfsError) as e:
        CompressedBasis.from_files([str(tmp_path / "none0"), str(tmp_path / "none1"), str(tmp_path / "none2")])
    assert "CFS_ERR_IO" in str(e.value)


@pytest.mark.parametrize("dims,r,percent", [((14, 10, 9), 4, 0.9995), ((8, 8, 8), 2, 0.99), ((17, 9, 12), 3, 0.999999)])
def test_encoder_bytes_match_the_oracle(tmp_path, dims, r, percent):
    U = synth.synthetic_basis(dims, r, seed=sum(dims) + r)
    got = encode_matrix(U, dims, percent=percent, threads=3).buffers()
    paths = O.compress_matrix(U, dims, 32, percent, 50, str(tmp_path / "U.oracle.component"))
    for k in range(3):
        want = open(paths[k], "rb").read()
        assert got[k] == want, f"component {k}"


def test_encoder_writes_files(tmp_path):
    dims = (8, 8, 8)
    U = synth.synthetic_basis(dims, 2, seed=1)
    enc = encode_matrix(U, dims)
    paths = enc.write(str(tmp_path / "U.x.component"))
    for k, p in enumerate(paths):
        assert open(p, "rb").read() == enc.buffer(k)
    ob = O.Basis(paths)
    assert ob.dims == dims and ob.r == 2 and ob.nB == 1
    ob.close()


def test_encoder_rejects_bad_shapes():
    with pytest.raises(ValueError):
        encode_matrix(np.zeros((2, 10)), (8, 8, 8))


def test_synthetic_basis_is_orthonormal_and_matches_sampled_qr():
    """The coefficient-space QR equals a QR of the sampled modes."""
    dims = (12, 9, 10)
    A = synth.synthetic_basis(dims, 6, seed=2)
    B = synth.synthetic_basis(dims, 6, seed=2, direct_qr=True)
    assert np.max(np.abs(A @ A.T - np.eye(6))) < 1e-12
    assert np.max(np.abs(A - B)) < 1e-12


def test_closure_and_advection_entry_points_reject_bad_arguments():
    """Argument checks of the closure / reduced-advection ABI happen before any
    device call, so they are testable without a GPU."""
    import ctypes
    lib = _lib.lib()
    h = ctypes.c_void_p()
    for r, hidden, batch in ((0, 256, 16), (300, 256, 16), (52, 128, 16), (52, 256, 0)):
        assert lib.cfs_closure_create(r, hidden, batch, 1, ctypes.byref(h)) == -1, (r, hidden, batch)
        assert not h.value
    assert lib.cfs_closure_create(52, 256, 16, 1, None) == -1
    assert lib.cfs_advect_field(None, None, 8, 8, 8, 1.0, None) == -1
    buf = (ctypes.c_float * 3)()
    assert lib.cfs_advect_field(buf, buf, 3, 8, 8, 1.0, None) == -1   # dims < 4
    buf2 = (ctypes.c_float * 3)()
    assert lib.cfs_advect_scalar(None, buf, buf2, 8, 8, 8, 1.0, None) == -1
    assert lib.cfs_advect_scalar(buf, buf2, buf2, 8, 8, 8, 1.0, None) == -1   # in place
    assert "in place" in lib.cfs_last_error().decode()
    assert lib.cfs_advect_scalar(buf, buf, buf2, 8, 3, 8, 1.0, None) == -1   # dims < 4
    assert lib.cfs_reduced_advect(None, None, None, None, None, 1.0, 1.0, None, None) == -1
    assert lib.cfs_closure_step(None, None, None, 1, 1.0, None, None) == -1
    assert "null" in lib.cfs_last_error().decode()
