"""Loader validation on the host (no GPU needed): malformed token streams are
rejected with CFS_ERR_FORMAT before anything is uploaded."""
import numpy as np
import pytest

from cfs_plume_amd import CompressedBasis
from oracle import oracle as O
from tests import helpers as H
from tests.test_gpu_parity import adversarial_zz


def _bad_basis(tmp_path, mutate):
    dims = (9, 8, 8)  # 2 blocks
    zz = [[O.rle_encode(adversarial_zz(0, b, c)) for b in range(2)] for c in range(2)]
    mutate(zz)
    s = np.full((2, 2), 1e9)
    g = np.full((2, 2), 2.0)
    img = H.component_bytes(dims, 32, s, g, zz)
    good = H.basis_from_ints(dims, 2, adversarial_zz)
    return H.write_files([img, good[1], good[2]], str(tmp_path / "bad."))


@pytest.mark.parametrize("name,mutate,reason", [
    ("run_length_too_long", lambda zz: zz[0].__setitem__(0, np.array([0, 0, 600], np.int32)), "run length"),
    ("run_length_one", lambda zz: zz[0].__setitem__(0, np.concatenate([[4, 4, 1], np.arange(1, 512)]).astype(np.int32)),
     "run length"),
    ("short_segment", lambda zz: zz[1].__setitem__(1, np.array([0, 0, 511], np.int32)), "fewer than 512"),
    ("long_segment", lambda zz: zz[1].__setitem__(0, np.array([3, 0, 0, 512], np.int32)), "more than 512"),
    ("truncated_run", lambda zz: zz[0].__setitem__(1, np.concatenate([np.arange(1, 511), [9, 9]]).astype(np.int32)),
     "inside a run"),
])
def test_malformed_streams_are_rejected(tmp_path, name, mutate, reason):
    """Where the reference asserts (COMPRESSION.cpp:1206-1207, 1267) or fills a
    short segment with zeros, the loader refuses the file."""
    from cfs_plume_amd import CfsError
    paths = _bad_basis(tmp_path, mutate)
    with pytest.raises(CfsError) as e:
        CompressedBasis.from_files(paths)
    assert "CFS_ERR_FORMAT" in str(e.value) and reason in str(e.value), str(e.value)
    # the oracle decoder agrees that the segment is not a valid 512-value block
    bad = 0
    for b in range(2):
        for c in range(2):
            toks = ob_segment(paths[0], b, c)
            if O.rle_decode_status(toks) != 0 or O.rle_decode(toks)[1]:
                bad += 1
    assert bad == 1


def ob_segment(path, b, c):
    import struct
    raw = open(path, "rb").read()
    h = struct.unpack("<10i", raw[:40])
    r, nB = h[8], h[9]
    m = r * nB
    lens = np.frombuffer(raw, np.int32, m, 40 + 16 * m)
    idx = np.frombuffer(raw, np.int32, m, 40 + 20 * m)
    tok = np.frombuffer(raw, np.int32, offset=40 + 24 * m)
    e = b + nB * c
    return tok[idx[e]: idx[e] + lens[e]]


def test_more_columns_than_supported_are_rejected(tmp_path):
    """The step keeps four column slots per lane (256 columns); a basis with
    more columns is refused at load, before anything is uploaded."""
    from cfs_plume_amd import CfsError
    dims = (8, 8, 8)
    bufs = H.basis_from_ints(dims, 257, lambda k, b, c: np.zeros(512, np.int32))
    with pytest.raises(CfsError) as e:
        CompressedBasis.from_files(H.write_files(bufs, str(tmp_path / "wide.")))
    assert "CFS_ERR_FORMAT" in str(e.value) and "exceeds the supported maximum 256" in str(e.value), str(e.value)


def test_header_sizes_that_would_wrap_are_rejected():
    """A header whose numCols * numBlocks * 24 wraps a 64-bit size must not
    pass the size check (it used to make a 72-byte component look whole and
    send the segment loop to wild addresses); numCols is checked before any
    size arithmetic, and numBlocks against the padded dims in 64 bits."""
    import ctypes
    import struct
    from cfs_plume_amd import _lib
    lib = _lib.lib()

    def try_open(header, payload=b"\0" * 32):
        img = struct.pack("<10i", *header) + payload
        buf = ctypes.create_string_buffer(img, len(img))
        h = ctypes.c_void_p()
        st = lib.cfs_basis_open_memory(buf, len(img), buf, len(img), buf, len(img), ctypes.byref(h))
        return st, lib.cfs_last_error().decode(), h.value

    wrap = [1, 32, 8, 16, 1684887088, 8, 16, 1684887088, 1824726041, 421221772]
    st, msg, h = try_open(wrap)
    assert st == -3 and not h, (st, msg)
    assert "numCols 1824726041 exceeds the supported maximum" in msg
    # the same header with a plausible column count still fails on size, not on memory
    st, msg, h = try_open(wrap[:8] + [2, 421221772])
    assert st == -3 and "truncated" in msg and not h, (st, msg)
    # padded dims whose block count overflows 64 bits unless computed in steps
    big = 2147483640
    st, msg, h = try_open([1, 32, big, big, big, big, big, big, 1, 8])
    assert st == -3 and "numBlocks" in msg and not h, (st, msg)
    # numBlocks <= 0
    st, msg, h = try_open([1, 32, 8, 8, 8, 8, 8, 8, 1, 0])
    assert st == -3 and "numBlocks" in msg and not h, (st, msg)
