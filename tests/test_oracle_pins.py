"""Pins the FP64 CPU oracle before it is trusted as the GPU path's checker.

Each test states what fixes the answer independently of the oracle's own code:
the reference's pure unit tests (tests/test_compression_data.cpp, cited by
line), scipy's DCT golden vectors (tests/golden/dct_scipy.npz, made by
tests/golden/make_dct_golden.py), closed forms, and brute force on small
inputs.  CPU only.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dct_scipy.npz")


# ---- block geometry: reference test_compression_data.cpp:1087-1205 --------------------------

def test_modified_cumsum_known_answers():
    # :1087-1123
    assert list(O.modified_cumsum([10, 20, 30, 40])) == [0, 10, 30, 60]
    assert list(O.modified_cumsum([42])) == [0]
    assert list(O.modified_cumsum(np.ones(8, np.int32))) == list(range(8))


def test_get_paddings_known_answers():
    # :1179-1205
    assert O.get_paddings((16, 24, 8)) == (0, 0, 0)
    assert O.get_paddings((46, 62, 46)) == (2, 2, 2)
    assert O.get_paddings((1, 1, 1)) == (7, 7, 7)


def test_compute_block_number_known_answers():
    # :1125-1177 on the 46x62x46 peeled grid
    dims = (46, 62, 46)
    assert O.compute_block_number(0, dims) == (0, 0)
    assert O.compute_block_number(3, dims) == (0, 1)
    assert O.compute_block_number(24, dims) == (1, 0)
    assert O.compute_block_number(1104, dims) == (6, 0)
    bn, bi = O.compute_block_number(3 * 46 * 62 * 46 - 1, dims)
    assert 0 <= bn < 6 * 8 * 6 and 0 <= bi < 512


def test_compute_block_number_agrees_with_block_gather():
    """Every cell's (block, index) from ComputeBlockNumber addresses the same
    value in GetBlocksEigen's zero-padded blocks (two independent index maps)."""
    dims = (13, 9, 10)
    X, Y, Z = dims
    field = np.arange(X * Y * Z, dtype=np.float64) + 1.0
    blocks = O.get_blocks(field, dims, zero_pad=True)
    for cell in range(X * Y * Z):
        for comp in range(3):
            bn, bi = O.compute_block_number(3 * cell + comp, dims)
            assert blocks[bn, bi] == field[cell]


def test_block_decomposition_roundtrip():
    # :1465-1537: 46x62x46, sin(i*0.01)*100, zero padding, reassembly
    dims = (46, 62, 46)
    X, Y, Z = dims
    f = np.sin(np.arange(X * Y * Z) * 0.01) * 100.0
    blocks = O.get_blocks(f, dims, zero_pad=True)
    assert blocks.shape == (6 * 8 * 6, 512)
    back = O.assimilate_unpadded(blocks, dims)
    assert np.max(np.abs(back - f)) < 1e-10
    # padded region is zero: rebuild the padded grid from the blocks
    cube = blocks.reshape(6, 8, 6, 8, 8, 8)  # (bz, by, bx, z, y, x)
    full = cube.transpose(0, 3, 1, 4, 2, 5).reshape(48, 64, 48)
    assert np.max(np.abs(full[Z:])) == 0 and np.max(np.abs(full[:, Y:])) == 0 and np.max(np.abs(full[:, :, X:])) == 0
    assert np.array_equal(full[:Z, :Y, :X].reshape(-1), f)


def test_edge_padding_replicates_the_boundary():
    """GetBlocks pads by replicating the last plane (FIELD_3D::pad_xyz)."""
    dims = (11, 8, 9)
    X, Y, Z = dims
    rng = np.random.default_rng(2)
    f = rng.standard_normal(X * Y * Z)
    blocks = O.get_blocks(f, dims, zero_pad=False)
    cube = blocks.reshape(2, 1, 2, 8, 8, 8).transpose(0, 3, 1, 4, 2, 5).reshape(16, 8, 16)
    g = f.reshape(Z, Y, X)
    want = g[np.minimum(np.arange(16), Z - 1)][:, :, np.minimum(np.arange(16), X - 1)]
    assert np.array_equal(cube, want)


# ---- zigzag: :1210-1288 and set_zigzagArray (COMPRESSION_DATA.cpp:95-117) ---------------

def test_zigzag_is_the_diagonal_order():
    zig, rev = O.zigzag_tables()
    assert sorted(zig) == list(range(512)) and sorted(rev) == list(range(512))
    assert np.array_equal(zig[rev], np.arange(512))
    x, y, z = rev % 8, (rev // 8) % 8, rev // 64
    d = x + y + z
    # non-decreasing diagonal, then z, y, x ascending within a diagonal
    key = d * 512 + z * 64 + y * 8 + x
    assert np.all(np.diff(key) > 0)
    sizes = np.bincount(d, minlength=22)
    assert list(sizes) == [1, 3, 6, 10, 15, 21, 28, 36, 42, 46, 48, 48, 46, 42, 36, 28, 21, 15, 10, 6, 3, 1]


def test_zigzag_roundtrip():
    # :1210-1252 with the reference's pattern
    field = ((np.arange(512) * 7 + 13) % 200 - 100).astype(np.int32)
    zz = O.zigzag_flatten(field)
    zig, _ = O.zigzag_tables()
    assert np.array_equal(zz[zig], field)
    assert np.array_equal(O.zigzag_unflatten(zz), field)


# ---- RLE: :1293-1460, :1743-1814, grammar COMPRESSION.cpp:1139-1169, 1241-1282 ----------

def test_rle_known_tokens():
    # :1293-1355 -- DC 2147483646, a run of two 42s, -7, a run of four 3s, zeros
    v = np.zeros(512, np.int32)
    v[0] = 2147483646
    v[1] = v[2] = 42
    v[3] = -7
    v[10:14] = 3
    tok = O.rle_encode(v)
    assert list(tok) == [2147483646, 42, 42, 2, -7, 0, 0, 6, 3, 3, 4, 0, 0, 498]
    out, underfull = O.rle_decode(tok)
    assert np.array_equal(out, v) and not underfull


def test_rle_all_zeros_is_three_tokens():
    # :1360-1403
    tok = O.rle_encode(np.zeros(512, np.int32))
    assert list(tok) == [0, 0, 512]
    assert np.array_equal(O.rle_decode(tok)[0], np.zeros(512, np.int32))


def test_rle_no_runs_is_512_tokens():
    # :1408-1460
    v = np.arange(1, 513, dtype=np.int32)
    tok = O.rle_encode(v)
    assert np.array_equal(tok, v)
    assert np.array_equal(O.rle_decode(tok)[0], v)


def test_rle_pairs_give_the_768_token_maximum():
    v = np.repeat(np.arange(1, 257, dtype=np.int32), 2)
    tok = O.rle_encode(v)
    assert len(tok) == 768
    assert np.array_equal(O.rle_decode(tok)[0], v)


def test_rle_roles_need_the_state_machine():
    """[a, a, 5, 5, 5, k]: run(a, 5) then run(5, k); a test of token equality
    alone would read index 2 as the start of a run."""
    a, k = 9, 507
    tok = np.array([a, a, 5, 5, 5, k], np.int32)
    out, underfull = O.rle_decode(tok)
    want = np.concatenate([np.full(5, a), np.full(k, 5)]).astype(np.int32)
    assert np.array_equal(out, want) and not underfull


def _brute_rle_decode(tok):
    """Sequential decoder written from the grammar: a token equal to its
    successor opens a run whose length is the token after that."""
    out, i = [], 0
    while i < len(tok):
        if i + 1 < len(tok) and tok[i] == tok[i + 1]:
            out += [tok[i]] * tok[i + 2]
            i += 3
        else:
            out.append(tok[i])
            i += 1
    return np.array(out, np.int32)


@pytest.mark.parametrize("seed", range(6))
def test_rle_random_roundtrip_against_brute_force(seed):
    rng = np.random.default_rng(seed)
    # runs of random lengths over a small alphabet: many equal neighbours across runs' ends
    vals, n = [], 0
    while n < 512:
        L = int(min(512 - n, rng.choice([1, 1, 1, 2, 3, 5, 40])))
        vals += [int(rng.integers(-3, 4))] * L
        n += L
    v = np.array(vals, np.int32)
    tok = O.rle_encode(v)
    # maximal runs: no two consecutive runs share a value
    assert np.array_equal(_brute_rle_decode(tok), v)
    assert np.array_equal(O.rle_decode(tok)[0], v)


def test_rle_decode_rejects_malformed():
    assert O.rle_decode_status([0, 0, 600]) != 0                   # run longer than a block
    out, underfull = O.rle_decode([0, 0, 511])                      # a short segment
    assert underfull


def test_rle_multiblock_multicolumn_table(tmp_path):
    # :1743-1814: 16x16x8 grid (4 blocks), 3 columns, per-(col, block) patterns
    dims = (16, 16, 8)
    nB, r = 4, 3

    def zz(k, b, c):
        seed = c * nB + b
        v = np.zeros(512, np.int32)
        v[0] = 1000 + seed
        v[1] = v[2] = 50 + seed
        v[seed % 512] = -seed
        return v

    bufs = H.basis_from_ints(dims, r, zz)
    ob = O.Basis(buffers=bufs)
    assert (ob.nB, ob.r) == (nB, r)
    for k in range(3):
        for c in range(r):
            for b in range(nB):
                assert np.array_equal(ob.decode_ints(k, b, c), O.zigzag_unflatten(zz(k, b, c)))
    ob.close()


# ---- DCT: closed form, scipy golden vectors, :527-578, :1542-1582, :1820-1870 --------------

def test_dct_matrix_closed_form():
    C = O.dct_matrix()
    n = np.arange(8)
    want = np.array([np.sqrt((1 if k == 0 else 2) / 8.0) * np.cos(np.pi * (2 * n + 1) * k / 16) for k in range(8)])
    assert np.max(np.abs(C - want)) < 4e-15  # a few ulp: cos of differently rounded arguments
    assert np.max(np.abs(C @ C.T - np.eye(8))) < 1e-14


def test_dct_matches_scipy_golden_vectors():
    g = np.load(GOLDEN)
    fwd = O.dct3(g["x"], 1)
    inv = O.dct3(g["x"], -1)
    assert np.max(np.abs(fwd - g["forward"])) < 1e-12
    assert np.max(np.abs(inv - g["inverse"])) < 1e-12


def test_dct_roundtrip_and_unnormalised_pair():
    g = np.load(GOLDEN)
    x = g["x"]
    # :1542-1582, UnitaryBlockDCTEigen(1) then (-1)
    back = O.dct3(O.dct3(x[:4], 1), -1)
    assert np.max(np.abs(back - x[:4])) < 1e-10
    assert np.sum(np.linalg.norm(O.dct3(x[:4], 1) - x[:4], axis=1)) > 1.0
    # :527-578: FFTW's REDFT10 then REDFT01 scale by 8*N = 4096
    assert np.max(np.abs(g["unnormalised_pair"] / (8.0 * 512) - x)) < 1e-10


def test_dct_energy_compaction():
    # :1820-1870
    z, y, x = np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")
    blk = (np.sin(x * 0.5) + np.cos(y * 0.3) + np.sin(z * 0.2)).reshape(1, -1)
    F = O.dct3(blk, 1)[0]
    e = np.sort(F ** 2)[::-1]
    assert e[:51].sum() / (blk ** 2).sum() > 0.90
    assert abs(F[0]) == np.max(np.abs(F))


# ---- dequantisation tables: set_dampingArrayList (COMPRESSION_DATA.cpp:25-43) --------------

def test_damping_list_closed_form():
    D = O.damping_list(32)
    assert D.shape == (129, 512)
    i = np.arange(512)
    d = 1 + i % 8 + (i // 8) % 8 + i // 64
    assert np.all(D[0] == 1.0)
    assert np.array_equal(D[4], d.astype(np.float64))                  # gamma = 1
    assert np.array_equal(D[8], (d * d).astype(np.float64))            # gamma = 2
    assert np.max(np.abs(D[2] / np.sqrt(d) - 1)) < 1e-15              # gamma = 1/2


def test_table_and_pow_dequant_agree():
    """DecodeBlockWithCompressionData (pow) and ...SparseQuantized (table)
    agree to 1e-10, as the reference's :834-906 requires."""
    rng = np.random.default_rng(4)
    dims = (9, 9, 8)

    def zz(k, b, c):
        return rng.integers(-1000, 1000, size=512).astype(np.int32)

    ob = O.Basis(buffers=H.basis_from_ints(dims, 2, zz, seed=5))
    for k in range(3):
        for b in range(ob.nB):
            for c in range(ob.r):
                a = ob.decode_coeffs(k, b, c)
                t = ob.decode_coeffs(k, b, c, table=True)
                assert np.max(np.abs(a - t)) <= 1e-10 * np.max(np.abs(a))
    ob.close()


# ---- project / unproject: brute force against dense U built from spatial decode ------------

@pytest.fixture(scope="module")
def small_basis():
    from cfs_plume_amd import encode_matrix, synth
    dims = (14, 10, 9)
    U = synth.synthetic_basis(dims, 5, seed=7)
    bufs = encode_matrix(U, dims, threads=4).buffers()
    ob = O.Basis(buffers=bufs)
    yield ob, U
    ob.close()


def _full(dims, interior):
    X, Y, Z = dims
    f = np.zeros((Z + 2, Y + 2, X + 2, 3))
    f[1:-1, 1:-1, 1:-1] = interior.reshape(Z, Y, X, 3)
    return f.reshape(-1)


def test_spatial_decode_is_close_to_the_encoded_matrix(small_basis):
    ob, U = small_basis
    for c in range(ob.r):
        col = ob.decode_vector_field(c)
        assert np.linalg.norm(col - U[c]) / np.linalg.norm(U[c]) < 1e-2


def test_project_equals_dense_transpose_product(small_basis):
    """Parseval: the DCT-domain project equals U^T v with U decoded spatially."""
    ob, _ = small_basis
    X, Y, Z = ob.dims
    Ud = np.stack([ob.decode_vector_field(c) for c in range(ob.r)])
    v = np.random.default_rng(1).standard_normal(3 * X * Y * Z)
    q = ob.project(_full(ob.dims, v))
    assert np.max(np.abs(q - Ud @ v)) <= 1e-12 * np.abs(Ud).sum(axis=1).max() * np.abs(v).max()


def test_unproject_equals_dense_product_and_keeps_boundary(small_basis):
    ob, _ = small_basis
    X, Y, Z = ob.dims
    Ud = np.stack([ob.decode_vector_field(c) for c in range(ob.r)])
    q = np.random.default_rng(2).standard_normal(ob.r)
    init = np.full(3 * (X + 2) * (Y + 2) * (Z + 2), 3.5)
    out = ob.unproject(q, init).reshape(Z + 2, Y + 2, X + 2, 3)
    assert np.max(np.abs(out[1:-1, 1:-1, 1:-1].reshape(-1) - q @ Ud)) < 1e-12 * np.abs(q).sum()
    inner = np.zeros(out.shape, bool)
    inner[1:-1, 1:-1, 1:-1] = True
    assert np.all(out[~inner] == 3.5)


def test_unit_q_unprojects_to_the_decoded_column(small_basis):
    ob, _ = small_basis
    X, Y, Z = ob.dims
    for c in (0, ob.r - 1):
        e = np.zeros(ob.r)
        e[c] = 1.0
        out = ob.unproject(e, np.zeros(3 * (X + 2) * (Y + 2) * (Z + 2))).reshape(Z + 2, Y + 2, X + 2, 3)
        assert np.max(np.abs(out[1:-1, 1:-1, 1:-1].reshape(-1) - ob.decode_vector_field(c))) < 1e-13


def test_project_after_unproject_is_the_gram_matrix(small_basis):
    """U^T (U q) = q for the (near-)orthonormal decoded basis; the reference's
    round-trip test (:1005-1081) only asks rel < 0.5."""
    ob, _ = small_basis
    X, Y, Z = ob.dims
    q = np.random.default_rng(3).standard_normal(ob.r)
    f = ob.unproject(q, np.zeros(3 * (X + 2) * (Y + 2) * (Z + 2)))
    q2 = ob.project(f)
    assert np.linalg.norm(q2 - q) / np.linalg.norm(q) < 2e-2


def test_submatrix_is_the_rows_of_the_decoded_columns(small_basis):
    ob, _ = small_basis
    X, Y, Z = ob.dims
    Ud = np.stack([ob.decode_vector_field(c) for c in range(ob.r)])
    for cell in (0, 5, X * Y * Z - 1):
        S = ob.get_submatrix(3 * cell)
        assert np.max(np.abs(S - Ud[:, 3 * cell:3 * cell + 3].T)) < 1e-13


# ---- encoder invariants: TuneGammaQuantized (COMPRESSION.cpp:841-916) ------------------------

@pytest.mark.parametrize("seed", range(4))
def test_encoder_invariants(seed):
    rng = np.random.default_rng(seed)
    F = O.dct3(rng.standard_normal((1, 512)) * np.exp(-np.arange(512) / 60.0), 1)[0]
    s, gamma, tok = O.encode_block(F, 32, 0.9995, 50)
    assert gamma * 4 == int(gamma * 4) and 0 <= gamma <= 32
    assert s == pytest.approx((2 ** 31 - 1) / np.max(np.abs(F)), rel=1e-15)
    ints = O.zigzag_unflatten(O.rle_decode(tok)[0])
    d = O.damping_list(32)[int(4 * gamma)]
    # rint quantisation of the scaled, damped coefficients
    assert np.array_equal(ints, np.rint(F * s / d).astype(np.int64).astype(np.int32))


def test_threaded_project_unproject_are_identical():
    """The all-cores CPU baseline splits the oracle's work without changing any
    summation order: results equal the single-threaded ones bit for bit."""
    from cfs_plume_amd import encode_matrix, synth
    dims = (14, 10, 9)
    bufs = encode_matrix(synth.synthetic_basis(dims, 5, seed=2), dims).buffers()
    ob = O.Basis(buffers=bufs)
    rng = np.random.default_rng(4)
    X, Y, Z = dims
    v = rng.normal(size=3 * (X + 2) * (Y + 2) * (Z + 2))
    q = rng.normal(size=5)
    for threads in (1, 3, 8):
        assert np.array_equal(ob.project_mt(v, threads), ob.project(v))
        assert np.array_equal(ob.unproject_mt(q, np.zeros_like(v), threads), ob.unproject(q, np.zeros_like(v)))
