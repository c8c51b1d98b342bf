"""Cubature reduced advection on the GPU (reducedAdvectCompressionFriendly,
SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:1781-1854) against the FP64 oracle
(oracle/cubature_oracle.py), through the C ABI.  The device interpolates an
FP32 unprojected field where the oracle multiplies FP64 decoded rows by q, so
each key cell's term After_i a_i carries the field's relative error (about
1e-6).  Bounds: |q' - q'_oracle| <= 4e-6 * sum_i |After_i a_i| always (the
sum of the terms' sizes: terms of a large, weakly correlated backtrace
cancel), and rel-L2 <= 1e-5 at displacements of about a cell per step."""
import numpy as np
import pytest

from cfs_plume_amd import CompressedBasis, DeviceArray, Stream, encode_matrix, reduced_advect, synth
from cfs_plume_amd import cubature as CB
from cfs_plume_amd.hip import Graph
from oracle import cubature_oracle as QO
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def setup(gpu):
    dims = (30, 22, 14)
    print(f"[case] cubature on the basis dims={dims} r=8", flush=True)
    bufs = encode_matrix(synth.synthetic_basis(dims, 8, seed=3), dims).buffers()
    ob = O.Basis(buffers=bufs)
    gb = CompressedBasis.from_buffers(bufs)
    rows = QO.dense_rows(ob, 8)
    X, Y, Z = dims
    field = DeviceArray.zeros((3 * (X + 2) * (Y + 2) * (Z + 2),), np.float32)
    yield dims, ob, gb, rows, field
    ob.close()


def _keys(dims, K, seed):
    rng = np.random.default_rng(seed)
    n = dims[0] * dims[1] * dims[2]
    return np.sort(rng.choice(n, size=K, replace=False)).astype(np.int32), rng.uniform(0.5, 2.0, size=K)


@pytest.mark.parametrize("K,scale", [(300, 40.0), (1000, 40.0), (1000, 150.0)])
def test_cubature_matches_the_oracle(setup, K, scale):
    dims, ob, gb, rows, field = setup
    cells, w = _keys(dims, K, K)
    cub = CB.Cubature.from_cells(cells, w, gb)
    assert cub.K == K and cub.r == gb.r
    c2, w2 = cub.cells_and_weights()
    assert np.array_equal(c2, cells) and np.array_equal(w2, w)
    after = QO.after_matrices(rows, cells, w)
    for seed in range(3):
        q = scale * synth.gaussian_q(gb.r, 70 + seed)
        f = synth.gaussian_q(gb.r, 80 + seed)
        dq, dout, df = DeviceArray.from_numpy(q), DeviceArray.zeros((gb.r,), np.float64), DeviceArray.from_numpy(f)
        cub.advect(gb, dq, dout, field, 1.0, 0.97, df)
        want, adv = QO.cubature_advect(rows, after, cells, dims, q, 1.0, return_advected=True)
        got = (dout.numpy() - f) / 0.97
        e = rel(got, want)
        terms = float(sum(np.linalg.norm(after[i] @ adv[i]) for i in range(K)))
        bound = float(np.linalg.norm(got - want)) / terms
        print(f"[cubature] K={K} displacement scale {scale}: seed {seed}: rel-L2 {e:.2e}, "
              f"error / sum of term sizes {bound:.2e} (cancellation {terms / np.linalg.norm(want):.1f})", flush=True)
        assert bound <= 4e-6
        if scale <= 40.0:
            assert e <= TOL
    assert gb.check_stream_errors() == 0
    cub.close()


def test_every_cell_is_the_exact_reduced_advection(setup):
    """Every peeled cell a key cell of weight 1: the device cubature equals the
    device's exact reduced advection U^T A(U q)."""
    dims, ob, gb, rows, field = setup
    n = dims[0] * dims[1] * dims[2]
    cub = CB.Cubature.from_cells(np.arange(n, dtype=np.int32), np.ones(n), gb)
    q = 60.0 * synth.gaussian_q(gb.r, 5)
    dq = DeviceArray.from_numpy(q)
    a, b = DeviceArray.zeros((gb.r,), np.float64), DeviceArray.zeros((gb.r,), np.float64)
    fb = DeviceArray.zeros(field.shape, np.float32)
    cub.advect(gb, dq, a, field, 1.0)
    reduced_advect(gb, dq, b, field, fb, 1.0)
    e = rel(a.numpy(), b.numpy())
    print(f"[cubature] all {n} cells, weight 1, vs exact reduced advection: rel-L2 {e:.2e}", flush=True)
    assert e <= TOL
    cub.close()


def test_cubature_file_round_trip_and_graph_replay(setup, tmp_path):
    """A cubature written by writeCubature's layout and read back advects bit
    for bit like the one built from the arrays; the three launches captured in
    a HIP graph replay bit for bit between eager calls."""
    dims, ob, gb, rows, field = setup
    cells, w = _keys(dims, 500, 9)
    path = str(tmp_path / "cubature")
    CB.write_cubature(path, cells, w)
    a = CB.Cubature.from_cells(cells, w, gb)
    b = CB.Cubature.from_file(path, gb)
    q = DeviceArray.from_numpy(40.0 * synth.gaussian_q(gb.r, 11))
    qa, qb = DeviceArray.zeros((gb.r,), np.float64), DeviceArray.zeros((gb.r,), np.float64)
    s = Stream()
    a.advect(gb, q, qa, field, 1.0, stream=s)
    b.advect(gb, q, qb, field, 1.0, stream=s)
    s.synchronize()
    eager = qa.numpy()
    assert np.array_equal(eager, qb.numpy())
    graph = Graph(s, lambda: b.advect(gb, q, qb, field, 1.0, stream=s))
    for rep in range(3):
        qb.fill_zero()
        graph.launch()
        s.synchronize()
        assert np.array_equal(qb.numpy(), eager), f"replay {rep}"
        a.advect(gb, q, qa, field, 1.0, stream=s)
        s.synchronize()
    a.close()
    b.close()


def test_cubature_refuses_a_basis_of_another_grid(setup):
    from cfs_plume_amd import CfsError
    dims, ob, gb, rows, field = setup
    other = CompressedBasis.from_buffers(encode_matrix(synth.synthetic_basis((14, 14, 6), 3, seed=5),
                                                       (14, 14, 6)).buffers())
    cub = CB.Cubature.from_cells(np.array([0, 5], np.int32), np.array([1.0, 1.0]), gb)
    q = DeviceArray.zeros((gb.r,), np.float64)
    with pytest.raises(CfsError):
        cub.advect(other, q, q, field, 1.0)
    with pytest.raises(CfsError):
        CB.Cubature.from_cells(np.array([dims[0] * dims[1] * dims[2]], np.int32), np.array([1.0]), gb)
    cub.close()
