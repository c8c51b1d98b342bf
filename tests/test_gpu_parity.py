"""GPU parity: the HIP path (through the C ABI) against the FP64 CPU oracle.

Integer stages (RLE tokens -> zigzag -> ints) must match bit for bit.  FP32
results are compared with relative L2 error against the oracle's FP64:
  decoded coefficients / columns / submatrix  <= 1e-5
  projected q, unprojected field              <= 1e-5   (the step's accuracy bound)
"""
import numpy as np
import pytest

from cfs_plume_amd import CompressedBasis, DeviceArray, encode_matrix, step, synth
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _pair(bufs, tmp_path_factory, name):
    d = tmp_path_factory.mktemp(name)
    paths = H.write_files(bufs, str(d / "U.test.component"))
    return O.Basis(paths), CompressedBasis.from_files(paths)


@pytest.fixture(scope="module")
def smooth(tmp_path_factory, gpu):
    dims = (30, 22, 14)
    U = synth.synthetic_basis(dims, 8, seed=3)
    print(f"[case] smooth basis dims={dims} r=8", flush=True)
    return _pair(encode_matrix(U, dims).buffers(), tmp_path_factory, "smooth")


@pytest.fixture(scope="module")
def rough(tmp_path_factory, gpu):
    # high wavenumbers on a tiny grid: segments of ~500-770 tokens (several chunks)
    dims = (14, 14, 6)
    U = synth.synthetic_basis(dims, 3, seed=5)
    print(f"[case] rough basis dims={dims} r=3", flush=True)
    return _pair(encode_matrix(U, dims).buffers(), tmp_path_factory, "rough")


def adversarial_zz(k, b, c):
    """Zigzag int vectors exercising the RLE grammar's corner cases."""
    rng = np.random.default_rng(1000 * k + 10 * b + c)
    kind = (b + c + k) % 7
    z = np.zeros(512, np.int32)
    if kind == 0:        # all zeros -> [0, 0, 512]
        pass
    elif kind == 1:      # no runs: 512 singletons
        z[:] = np.arange(1, 513) * (1 if c % 2 else -1)
    elif kind == 2:      # pairs: a, a, b, b, ... -> 768 tokens (the maximum)
        z[:] = np.repeat(np.arange(1, 257), 2)
    elif kind == 3:      # run lengths equal to neighbouring values: [5,5,5,k] style
        z[:3] = 2147483647
        z[3:8] = 5
        z[8] = 5 + 1
        z[9:11] = 2
        z[11:13] = 2 + 1  # value 3 after a run of length 2 ending in a length token 2
        z[13:15] = 3
        z[15:515 - 100] = 0
        z[412:] = rng.integers(-3, 4, size=100)
    elif kind == 4:      # long nonzero run then zeros
        z[0] = -7
        z[1:300] = 1
    elif kind == 5:      # random small values with many short zero runs
        z[:] = rng.integers(-2, 3, size=512)
    else:                # dense DCT-like decay
        z[:] = np.round(rng.standard_normal(512) * 2.0e9 / (1 + np.arange(512)) ** 2).astype(np.int64).clip(-2**31 + 1, 2**31 - 1)
    return z


@pytest.fixture(scope="module")
def adversarial(tmp_path_factory, gpu):
    dims = (17, 9, 8)  # 3 x 2 x 1 blocks
    print(f"[case] adversarial token streams dims={dims} r=5", flush=True)
    bufs = H.basis_from_ints(dims, 5, adversarial_zz, seed=11)
    return _pair(bufs, tmp_path_factory, "adv")


def hazard_zz(k, b, c):
    """Every segment starts and ends with the singleton 5 (z[510] != 5), so each
    boundary inside a block stream reads [.., 5][5, ..]: without a reset the
    parser would take the pair for a run."""
    z = (np.arange(512, dtype=np.int32) % 97) + 10 + 3 * k
    z[0] = 5
    z[511] = 5
    if c % 2:
        z[200:300] = 0  # a zero run in the middle of every other column
    return z


@pytest.fixture(scope="module")
def hazard(tmp_path_factory, gpu):
    dims = (16, 8, 8)
    print(f"[case] hazard boundaries dims={dims} r=6", flush=True)
    return _pair(H.basis_from_ints(dims, 6, hazard_zz, seed=12), tmp_path_factory, "hazard")


def runs_zz(k, b, c):
    """Nonzero runs of every length 1..7 at random offsets, so run tokens
    (value, duplicate, length) straddle lane and chunk boundaries."""
    rng = np.random.default_rng(7000 + 100 * k + 10 * b + c)
    out, n = [], 0
    while n < 512:
        L = int(min(512 - n, rng.integers(1, 8)))
        v = int(rng.choice([-3, -2, -1, 1, 2, 3, 0]))
        if out and out[-1] == v:
            v = v + 5  # keep runs maximal
        out += [v] * L
        n += L
    return np.array(out, np.int32)


@pytest.fixture(scope="module")
def runs(tmp_path_factory, gpu):
    dims = (16, 16, 8)
    print(f"[case] nonzero runs of length 1..7 dims={dims} r=6", flush=True)
    return _pair(H.basis_from_ints(dims, 6, runs_zz, seed=13), tmp_path_factory, "runs")


def longruns_zz(k, b, c):
    """Nonzero runs of length 4..9 only, about 75 per segment, so a block
    stream of 6 columns holds several hundred runs whose values past the
    third are added from the loader's long-run list: the fused step reads
    that list 64 runs at a time and refills it several times per stream."""
    rng = np.random.default_rng(9000 + 100 * k + 10 * b + c)
    out, n = [], 0
    while n < 512:
        L = int(min(512 - n, rng.integers(4, 10)))
        v = int(rng.choice([-9, -4, -1, 1, 4, 9]))
        if out and out[-1] == v:
            v = -v  # keep runs maximal
        out += [v] * L
        n += L
    return np.array(out, np.int32)


@pytest.fixture(scope="module")
def longruns(tmp_path_factory, gpu):
    dims = (16, 8, 8)
    print(f"[case] long nonzero runs dims={dims} r=6", flush=True)
    return _pair(H.basis_from_ints(dims, 6, longruns_zz, seed=15), tmp_path_factory, "longruns")


def wide_zz(k, b, c):
    """256 columns (the supported maximum): stretches of all-zero segments
    (3 tokens each, so more than 64 segments close inside one 512-token
    chunk), dense segments and runs, so every lane's four column slots and the
    repeated segment-close pass are used."""
    kind = (c // 80 + k + b) % 4
    if kind == 0 or (kind == 1 and c % 3):
        return np.zeros(512, np.int32)
    if kind == 1:
        return adversarial_zz(k, b, c % 7)
    if kind == 2:
        return runs_zz(k, b, c)
    return hazard_zz(k, b, c)


@pytest.fixture(scope="module")
def wide(tmp_path_factory, gpu):
    dims = (16, 8, 8)
    print(f"[case] wide basis dims={dims} r=256", flush=True)
    return _pair(H.basis_from_ints(dims, 256, wide_zz, seed=14), tmp_path_factory, "wide")


@pytest.fixture(scope="module")
def ragged(tmp_path_factory, gpu):
    # peeled dims that are not multiples of 8 on any axis (partial blocks on
    # every face) and 130 columns (three column slots per lane, the last
    # partly filled)
    dims = (23, 13, 9)
    U = synth.synthetic_basis(dims, 130, seed=8)
    print(f"[case] ragged basis dims={dims} r=130", flush=True)
    return _pair(encode_matrix(U, dims).buffers(), tmp_path_factory, "ragged")


@pytest.fixture(scope="module")
def tiny(tmp_path_factory, gpu):
    # one partial block (5 x 3 x 2 cells of an 8^3 block) and one column:
    # the smallest basis the format holds
    dims = (5, 3, 2)
    U = synth.synthetic_basis(dims, 1, seed=9)
    print(f"[case] tiny basis dims={dims} r=1", flush=True)
    return _pair(encode_matrix(U, dims).buffers(), tmp_path_factory, "tiny")


CASES = ["smooth", "rough", "adversarial", "hazard", "runs", "longruns", "wide", "ragged", "tiny"]


def test_hazards_are_listed(hazard):
    ob, gb = hazard
    assert gb.hazards == 3 * gb.nB * (gb.r - 1)


@pytest.mark.parametrize("case", CASES)
def test_decode_ints_bit_exact(case, request):
    ob, gb = request.getfixturevalue(case)
    for k in range(3):
        for c in range(ob.r):
            got = gb.decode_block_ints(k, c)
            want = np.stack([ob.decode_ints(k, b, c) for b in range(ob.nB)])
            assert np.array_equal(got, want), f"comp {k} col {c}"
    assert gb.check_stream_errors() == 0


@pytest.mark.parametrize("case", CASES)
def test_decode_coeffs(case, request):
    ob, gb = request.getfixturevalue(case)
    for k in range(3):
        for c in range(ob.r):
            got = gb.decode_block_coeffs(k, c)
            want = np.stack([ob.decode_coeffs(k, b, c) for b in range(ob.nB)])
            assert rel(got, want) <= TOL, (k, c, rel(got, want))
            # the table path (DecodeBlockWithCompressionDataSparseQuantized) agrees with pow
            want2 = np.stack([ob.decode_coeffs(k, b, c, table=True) for b in range(ob.nB)])
            assert rel(want2, want) <= 1e-14


# The spatial decode runs its inverse DCT on MFMA with BF16 hi + lo split
# products (FP32 accumulation): per-column relative L2 error against the FP64
# oracle is bounded by SPATIAL_TOL.
SPATIAL_TOL = 3e-5


@pytest.mark.parametrize("case", CASES)
def test_decode_columns(case, request):
    ob, gb = request.getfixturevalue(case)
    cols = gb.decode_columns().numpy()
    errs = []
    for c in range(ob.r):
        want = ob.decode_vector_field(c)
        errs.append(rel(cols[c], want))
        assert errs[-1] <= SPATIAL_TOL, (c, errs[-1])
    print(f"[{case}] decode_columns per-column rel-L2: max {max(errs):.2e} mean {np.mean(errs):.2e}")
    # a column range that starts inside a group of columns
    if ob.r > 2:
        part = gb.decode_columns(1, ob.r - 2).numpy()
        for c in range(1, ob.r - 1):
            assert np.array_equal(part[c - 1], cols[c]), c


def _fields(ob, seed):
    v = synth.gaussian_field(ob.dims, seed)
    q = synth.gaussian_q(ob.r, seed + 1)
    return v, q


@pytest.mark.parametrize("case", CASES)
def test_project(case, request):
    ob, gb = request.getfixturevalue(case)
    v, _ = _fields(ob, 42)
    want = ob.project(v)
    dv = DeviceArray.from_numpy(v.astype(np.float32))
    got = gb.project(dv).numpy()
    # the oracle sees the same FP32-rounded field
    want32 = ob.project(v.astype(np.float32).astype(np.float64))
    assert rel(got, want32) <= TOL, rel(got, want32)
    assert rel(got, want) <= 1e-4
    assert gb.check_stream_errors() == 0


@pytest.mark.parametrize("case", CASES)
def test_unproject_keeps_boundary(case, request):
    ob, gb = request.getfixturevalue(case)
    _, q = _fields(ob, 43)
    X, Y, Z = ob.dims
    init = np.full((Z + 2) * (Y + 2) * (X + 2) * 3, 7.25)
    want = ob.unproject(q, init)
    dq = DeviceArray.from_numpy(q)
    df = DeviceArray.from_numpy(init.astype(np.float32))
    got = gb.unproject(dq, df).numpy().astype(np.float64)
    g4 = got.reshape(Z + 2, Y + 2, X + 2, 3)
    w4 = want.reshape(Z + 2, Y + 2, X + 2, 3)
    assert rel(g4[1:-1, 1:-1, 1:-1], w4[1:-1, 1:-1, 1:-1]) <= TOL
    inner = np.zeros_like(g4, bool)
    inner[1:-1, 1:-1, 1:-1] = True
    assert np.all(g4[~inner] == 7.25), "unproject wrote boundary cells"
    assert gb.check_stream_errors() == 0


@pytest.mark.parametrize("case", CASES)
def test_unproject_spatial(case, request):
    """PeeledCompressedUnproject: spatial decode of every column, then the
    weighted sum; equal to the oracle's unproject and to the fused one."""
    ob, gb = request.getfixturevalue(case)
    _, q = _fields(ob, 45)
    X, Y, Z = ob.dims
    init = np.full((Z + 2) * (Y + 2) * (X + 2) * 3, -3.5)
    want = ob.unproject(q, init).reshape(Z + 2, Y + 2, X + 2, 3)
    dq = DeviceArray.from_numpy(q)
    fused = gb.unproject(dq, DeviceArray.from_numpy(init.astype(np.float32))).numpy().reshape(want.shape)
    inner = np.zeros(want.shape, bool)
    inner[1:-1, 1:-1, 1:-1] = True
    scratch = DeviceArray((3, 3 * gb.peeled_cells), np.float32)
    for group, buf in ((1, None), (3, scratch), (ob.r, None)):
        df = DeviceArray.from_numpy(init.astype(np.float32))
        got = gb.unproject_spatial(dq, df, group=group, scratch=buf).numpy().astype(np.float64).reshape(want.shape)
        e = rel(got[inner], want[inner])
        print(f"[{case}] unproject_spatial group {group}: rel {e:.2e} vs oracle, "
              f"{rel(got[inner], fused[inner]):.2e} vs fused", flush=True)
        assert e <= SPATIAL_TOL, (group, e)
        assert rel(got[inner], fused[inner]) <= SPATIAL_TOL
        assert np.all(got[~inner] == -3.5), "spatial unproject wrote boundary cells"
    scratch.free()
    assert gb.check_stream_errors() == 0


@pytest.mark.parametrize("case", CASES)
def test_fused_step_matches_separate(case, request):
    ob, gb = request.getfixturevalue(case)
    v, q = _fields(ob, 44)
    dv = DeviceArray.from_numpy(v.astype(np.float32))
    dq = DeviceArray.from_numpy(q)
    qo = DeviceArray.zeros((ob.r,), np.float64)
    fo = DeviceArray.zeros((len(v),), np.float32)
    results = []
    for _ in range(3):  # repeated launches: the persistent work counter must reset
        step(gb, gb, dv, qo, dq, fo)
        results.append((qo.numpy(), fo.numpy()))
    q_sep = gb.project(dv).numpy()
    f_sep = gb.unproject(dq, DeviceArray.zeros((len(v),), np.float32)).numpy()
    for qq, ff in results:
        assert rel(qq, q_sep) <= 1e-6
        assert rel(ff, f_sep) <= 1e-6
    want_q = ob.project(v.astype(np.float32).astype(np.float64))
    want_f = ob.unproject(q, np.zeros(len(v)))
    assert rel(results[-1][0], want_q) <= TOL
    assert rel(results[-1][1], want_f) <= TOL
    assert gb.check_stream_errors() == 0


@pytest.mark.parametrize("case", ["smooth", "adversarial"])
def test_submatrix(case, request):
    ob, gb = request.getfixturevalue(case)
    X, Y, Z = ob.dims
    rng = np.random.default_rng(9)
    cells = [0, 1, X - 1, X * Y * Z - 1] + list(rng.integers(0, X * Y * Z, size=6))
    for cell in cells:
        got = gb.get_submatrix(3 * int(cell))
        want = ob.get_submatrix(3 * int(cell))
        assert rel(got, want) <= TOL, (cell, rel(got, want))




def test_fused_step_replays_in_a_hip_graph(smooth):
    """cfs_step issues kernel launches only: captured in a HIP graph, every
    replay (with eager steps and synchronisations between the replays) equals
    the eager step bit for bit."""
    from cfs_plume_amd.hip import Graph, Stream
    ob, gb = smooth
    v, q = _fields(ob, 45)
    dv = DeviceArray.from_numpy(v.astype(np.float32))
    dq = DeviceArray.from_numpy(q)
    q_e, q_g = DeviceArray.zeros((ob.r,), np.float64), DeviceArray.zeros((ob.r,), np.float64)
    f_e, f_g = DeviceArray.zeros((len(v),), np.float32), DeviceArray.zeros((len(v),), np.float32)
    s = Stream()
    step(gb, gb, dv, q_e, dq, f_e, s)
    s.synchronize()
    eager = (q_e.numpy(), f_e.numpy())
    graph = Graph(s, lambda: step(gb, gb, dv, q_g, dq, f_g, s))
    for rep in range(3):
        q_g.fill_zero()
        f_g.fill_zero()
        graph.launch()
        s.synchronize()
        assert np.array_equal(q_g.numpy(), eager[0]), f"replay {rep}: q"
        assert np.array_equal(f_g.numpy(), eager[1]), f"replay {rep}: field"
        step(gb, gb, dv, q_e, dq, f_e, s)
        s.synchronize()
        assert np.array_equal(q_e.numpy(), eager[0])
    assert gb.check_stream_errors() == 0


def test_step_with_two_bases(smooth, tmp_path_factory):
    """stepPlume projects with U_preadvect and unprojects with U_final
    (SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:477,490): cfs_step(pre, fin)
    gives q_out = pre^T v and field = fin q_in, each against its own basis."""
    ob_pre, gb_pre = smooth
    dims = ob_pre.dims
    U2 = synth.synthetic_basis(dims, 5, seed=31)
    print(f"[case] second basis dims={dims} r=5 (the unproject basis)", flush=True)
    ob_fin, gb_fin = _pair(encode_matrix(U2, dims).buffers(), tmp_path_factory, "final")
    v = synth.gaussian_field(dims, 46)
    q = synth.gaussian_q(gb_fin.r, 47)
    dv = DeviceArray.from_numpy(v.astype(np.float32))
    dq = DeviceArray.from_numpy(q)
    qo = DeviceArray.zeros((gb_pre.r,), np.float64)
    fo = DeviceArray.zeros((len(v),), np.float32)
    step(gb_pre, gb_fin, dv, qo, dq, fo)
    want_q = ob_pre.project(v.astype(np.float32).astype(np.float64))
    want_f = ob_fin.unproject(q, np.zeros(len(v)))
    assert rel(qo.numpy(), want_q) <= TOL
    assert rel(fo.numpy(), want_f) <= TOL
    # the two halves used different bases: swapping them changes both results
    assert rel(ob_fin.project(v.astype(np.float32).astype(np.float64)), want_q[:gb_fin.r]) > 0.1
    assert gb_pre.check_stream_errors() == 0 and gb_fin.check_stream_errors() == 0


def test_decode_columns_empty_range(smooth):
    """An empty column range decodes nothing and succeeds, as the block
    decoders do for an empty block range."""
    import ctypes
    from cfs_plume_amd import _lib
    ob, gb = smooth
    out = DeviceArray.from_numpy(np.full(16, 3.5, np.float32))
    assert _lib.lib().cfs_decode_columns(gb._h, 0, 0, ctypes.c_void_p(out.ptr), None) == 0
    assert _lib.lib().cfs_decode_columns(gb._h, ob.r, 0, ctypes.c_void_p(out.ptr), None) == 0
    assert np.all(out.numpy() == 3.5)


def test_untransform_svd_matches_oracle(ragged):
    """UntransformVectorFieldSVD (COMPRESSION.cpp:713-727) on decoded columns:
    each column's cells multiplied by its own 3 x 3 V, against the oracle's
    restatement applied to the oracle's FP64 decode."""
    ob, gb = ragged
    ncols = 7
    rng = np.random.default_rng(5)
    Vs = np.stack([np.linalg.qr(rng.standard_normal((3, 3)))[0] for _ in range(ncols)])
    cols = gb.decode_columns(3, ncols)
    gb.untransform_svd(cols, DeviceArray.from_numpy(Vs.reshape(-1)))
    got = cols.numpy()
    for i in range(ncols):
        want = O.untransform_svd(Vs[i], ob.decode_vector_field(3 + i))
        assert rel(got[i], want) <= SPATIAL_TOL, (i, rel(got[i], want))
