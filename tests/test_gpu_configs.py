"""GPU parity at the benchmark configurations beside the headline one.

  48x64x48, r = 52     the plume-basis shape, here with synthetic modes (a
                       stand-in: the reference's basis comes from its SVD of
                       plume snapshots); fused step and all 52 columns
  96x128x96, r = 150   fused step against the FP64 oracle and all 150 columns
  16 uneven bases      one persistent batched launch replayed from a HIP
                       graph, bit-equal to each basis's own fused step, and
                       the largest and smallest bases against the FP64 oracle

Synthetic modes: 16 seeded Fourier terms per mode and component, wavenumbers
1..6, orthonormalised in FP64, coded at nBits 32 with the stated percent and
maxIterations 50.  The oracle (FP64, reference loop order) runs on 16 host
threads with results identical to its single-thread form.

Tolerances: projected q and unprojected field <= 1e-5 rel-L2 (the step's
accuracy bound); each spatially decoded column <= 3e-5 rel-L2 (BF16 hi+lo
MFMA inverse DCT with FP32 accumulation).
"""
import numpy as np
import pytest

from cfs_plume_amd import BasisBatch, CompressedBasis, DeviceArray, Stream, encode_matrix, step, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5
SPATIAL_TOL = 3e-5
THREADS = 16


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def encoded(grid, r, percent, seed):
    dims = tuple(g - 2 for g in grid)
    U = synth.synthetic_basis(dims, r, seed=seed)
    bufs = encode_matrix(U, dims, 32, percent, 50, threads=THREADS).buffers()
    return dims, bufs


def check_step(ob, gb, label, seed=42):
    """Fused step against the threaded FP64 oracle; returns (q rel-L2,
    largest per-entry q error relative to max |q|, field rel-L2)."""
    r = gb.r
    v = synth.gaussian_field(gb.dims, seed)
    q = synth.gaussian_q(r, seed + 1)
    v32 = v.astype(np.float32)
    dv, dq = DeviceArray.from_numpy(v32), DeviceArray.from_numpy(q)
    qo = DeviceArray.zeros((r,), np.float64)
    fo = DeviceArray.zeros((len(v),), np.float32)
    step(gb, gb, dv, qo, dq, fo)
    q_gpu, f_gpu = qo.numpy(), fo.numpy()
    assert gb.check_stream_errors() == 0
    want_q = ob.project_mt(v32.astype(np.float64), THREADS)
    want_f = ob.unproject_mt(q, np.zeros(len(v)), THREADS)
    eq = rel(q_gpu, want_q)
    eq_entry = float(np.max(np.abs(q_gpu - want_q)) / np.max(np.abs(want_q)))
    ef = rel(f_gpu, want_f)
    print(f"[{label}] fused step vs FP64 oracle: q rel-L2 {eq:.2e} (max per-entry {eq_entry:.2e}), "
          f"field rel-L2 {ef:.2e}", flush=True)
    assert eq <= TOL and ef <= TOL
    return eq, eq_entry, ef


def check_all_columns(ob, gb, label, chunk=16):
    """Every column's spatial decode against the FP64 oracle's DecodeVectorField."""
    errs = []
    for c0 in range(0, gb.r, chunk):
        n = min(chunk, gb.r - c0)
        got = gb.decode_columns(c0, n).numpy()
        want = ob.decode_vector_fields_mt(c0, c0 + n, THREADS)
        for i in range(n):
            errs.append(rel(got[i], want[i]))
    assert gb.check_stream_errors() == 0
    errs = np.array(errs)
    worst = int(np.argmax(errs))
    print(f"[{label}] all {gb.r} decoded columns vs FP64 oracle: rel-L2 max {errs.max():.2e} (column {worst}), "
          f"mean {errs.mean():.2e}", flush=True)
    assert np.all(errs <= SPATIAL_TOL), (worst, errs.max())
    return errs


@pytest.fixture(scope="module")
def c1(gpu):
    print("[case] 48x64x48 r=52 (plume-basis shape, synthetic modes): encoding", flush=True)
    dims, bufs = encoded((48, 64, 48), 52, 0.9995, 7)
    return O.Basis(buffers=bufs), CompressedBasis.from_buffers(bufs)


@pytest.fixture(scope="module")
def c2(gpu):
    print("[case] 96x128x96 r=150: encoding", flush=True)
    dims, bufs = encoded((96, 128, 96), 150, 0.9995, 7)
    return O.Basis(buffers=bufs), CompressedBasis.from_buffers(bufs)


def test_c1_shape_step(c1):
    ob, gb = c1
    assert gb.dims == (46, 62, 46) and gb.r == 52 and gb.nB == 288
    check_step(ob, gb, "48x64x48 r52")


def test_c1_shape_all_columns(c1):
    check_all_columns(*c1, "48x64x48 r52")


def test_c2_step(c2):
    ob, gb = c2
    assert gb.dims == (94, 126, 94) and gb.r == 150 and gb.nB == 2304
    check_step(ob, gb, "96x128x96 r150")
    # repeated launches of the persistent kernel agree
    check_step(ob, gb, "96x128x96 r150 (second seed)", seed=52)


def test_c2_all_columns(c2):
    check_all_columns(*c2, "96x128x96 r150")


# --- 16 uneven bases in one persistent launch ---------------------------------------------

def batch_specs(n=16, seed=2024):
    """The batched workload's bases: grids in [32, 128] x [32, 192] x [32, 128]
    (multiples of 8), r in [20, 200], percent in [0.99, 0.9995], seeded."""
    import bench
    return bench.batch_specs(n, seed)


@pytest.fixture(scope="module")
def batch16(gpu):
    specs = batch_specs()
    out = []
    for i, (grid, r, percent, seed) in enumerate(specs):
        print(f"[case] batch basis {i}: grid {grid} r={r} percent={percent:.5f}", flush=True)
        dims, bufs = encoded(grid, r, percent, seed)
        out.append((grid, r, bufs, CompressedBasis.from_buffers(bufs)))
    return out


def test_c4_batch_graph_replay(batch16):
    """The whole batched step captured in a HIP graph: every replay equals the
    eager batched step bit for bit (with eager steps and synchronisations
    between replays); the batched fields equal each basis's own fused step bit
    for bit, and its q agrees to the FP64 summation order (the batch
    gives a basis fewer workgroups than its own launch, so its items are
    summed in other groupings); the largest and the smallest bases agree with
    the FP64 oracle."""
    bases = [b[3] for b in batch16]
    arrs, own = [], []
    for i, gb in enumerate(bases):
        v = synth.gaussian_field(gb.dims, 42 + i)
        q = synth.gaussian_q(gb.r, 43 + i)
        a = (DeviceArray.from_numpy(v.astype(np.float32)), DeviceArray.zeros((gb.r,), np.float64),
             DeviceArray.from_numpy(q), DeviceArray.zeros((len(v),), np.float32))
        arrs.append((v, q) + a)
        qo = DeviceArray.zeros((gb.r,), np.float64)
        fo = DeviceArray.zeros((len(v),), np.float32)
        step(gb, gb, a[0], qo, a[2], fo)
        own.append((qo.numpy(), fo.numpy()))
    bt = BasisBatch(bases)
    bt.bind([a[2] for a in arrs], [a[3] for a in arrs], [a[4] for a in arrs], [a[5] for a in arrs])
    s = Stream()
    cfg = bt.tune(reps=2, stream=s)
    print(f"[batch16] autotuned launch: {cfg}", flush=True)
    bt.step(s)
    s.synchronize()
    eager = [(a[3].numpy(), a[5].numpy()) for a in arrs]
    worst = 0.0
    for i, ((qb, fb), (q1, f1)) in enumerate(zip(eager, own)):
        assert np.array_equal(fb, f1), f"basis {i}: batched field differs from its own step"
        worst = max(worst, rel(qb, q1))
    print(f"[batch16] batched q vs each basis's own step: max rel-L2 {worst:.2e}", flush=True)
    assert worst <= 1e-6
    bt.capture(s)
    for rep in range(3):
        for a in arrs:
            a[3].fill_zero()
            a[5].fill_zero()
        bt.replay(s)
        s.synchronize()
        for i, (a, (qb, fb)) in enumerate(zip(arrs, eager)):
            assert np.array_equal(a[5].numpy(), fb), f"replay {rep}, basis {i}: field"
            assert np.array_equal(a[3].numpy(), qb), f"replay {rep}, basis {i}: q"
        bt.step(s)  # eager work between replays
        s.synchronize()
    for gb in bases:
        assert gb.check_stream_errors() == 0
    sizes = [gb.r * gb.nB for gb in bases]
    for i in (int(np.argmax(sizes)), int(np.argmin(sizes))):
        grid, r, bufs, gb = batch16[i]
        v, q = arrs[i][0], arrs[i][1]
        ob = O.Basis(buffers=bufs)
        want_q = ob.project_mt(v.astype(np.float32).astype(np.float64), THREADS)
        want_f = ob.unproject_mt(q, np.zeros(len(v)), THREADS)
        ob.close()
        eq, ef = rel(arrs[i][3].numpy(), want_q), rel(arrs[i][5].numpy(), want_f)
        print(f"[batch16] basis {i} (grid {grid}, r={r}) vs FP64 oracle: q {eq:.2e}, field {ef:.2e}", flush=True)
        assert eq <= TOL and ef <= TOL
    bt.close()
