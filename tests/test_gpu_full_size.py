"""GPU parity at the benchmark's full size: grid 128x192x128 (126x190x126
peeled cells), r = 200, percent 0.9995, nBits 32 -- 239 M tokens.

The fused step (PeeledCompressedProjectTransformNoSVD + PeeledCompressedUnprojectTransform,
COMPRESSION.cpp:2978-3016, 2739-2900) is compared with the FP64 oracle run
on 16 host threads, and with the dense twin of the same decoded basis
(VECTOR3_FIELD_3D::peeledProject / peeledUnproject, VECTOR3_FIELD_3D.cpp:3434-3606):
q = D^T v and the field's interior = D q in FP64 over the columns D that the
spatial decoder (DecodeVectorField, COMPRESSION.cpp:2640-2670) returns.
Tolerances: q and field <= 1e-5 rel-L2 against the oracle (the step's
accuracy bound); <= 3e-5 against the dense twin, whose FP32 columns carry
the spatial decoder's own error (<= 3e-5 per column).
"""
import numpy as np
import pytest

from cfs_plume_amd import CompressedBasis, DeviceArray, encode_matrix, step, synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5
TWIN_TOL = 3e-5
THREADS = 16
DIMS = (126, 190, 126)
R = 200


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def peel(full, dims):
    X, Y, Z = dims
    return full.reshape(Z + 2, Y + 2, X + 2, 3)[1:-1, 1:-1, 1:-1].reshape(-1)


@pytest.fixture(scope="module")
def full(gpu):
    print(f"[case] full-size basis dims={DIMS} r={R}: encoding", flush=True)
    U = synth.synthetic_basis(DIMS, R, seed=7)
    bufs = encode_matrix(U, DIMS, 32, 0.9995, 50, threads=THREADS).buffers()
    del U
    ob = O.Basis(buffers=bufs)
    gb = CompressedBasis.from_buffers(bufs)
    del bufs
    print(f"[case] full-size basis opened: tokens={gb.tokens}", flush=True)
    v = synth.gaussian_field(DIMS, 42)
    q = synth.gaussian_q(R, 43)
    v32 = v.astype(np.float32)
    dv = DeviceArray.from_numpy(v32)
    dq = DeviceArray.from_numpy(q)
    qo = DeviceArray.zeros((R,), np.float64)
    fo = DeviceArray.zeros((len(v),), np.float32)
    results = []
    for _ in range(3):  # repeated launches of the persistent kernel
        step(gb, gb, dv, qo, dq, fo)
        results.append((qo.numpy(), fo.numpy()))
    assert gb.check_stream_errors() == 0
    yield ob, gb, v32.astype(np.float64), q, dv, dq, results
    ob.close()


def test_step_matches_the_oracle(full):
    ob, gb, v, q, dv, dq, results = full
    want_q = ob.project_mt(v, THREADS)
    want_f = ob.unproject_mt(q, np.zeros(len(v)), THREADS)
    eq, ef = rel(results[-1][0], want_q), rel(results[-1][1], want_f)
    print(f"[full] fused step vs FP64 oracle: q rel-L2 {eq:.2e}, field rel-L2 {ef:.2e}", flush=True)
    assert eq <= TOL and ef <= TOL
    for qq, ff in results[:-1]:
        assert rel(qq, results[-1][0]) <= 1e-12
        assert np.array_equal(ff, results[-1][1])


def test_step_matches_separate_launches(full):
    ob, gb, v, q, dv, dq, results = full
    q_sep = gb.project(dv).numpy()
    init = np.full(len(v), 7.25, np.float32)
    f_sep = gb.unproject(dq, DeviceArray.from_numpy(init)).numpy()
    assert rel(results[-1][0], q_sep) <= 1e-6
    X, Y, Z = DIMS
    g4 = f_sep.reshape(Z + 2, Y + 2, X + 2, 3)
    inner = np.zeros(g4.shape, bool)
    inner[1:-1, 1:-1, 1:-1] = True
    assert np.all(g4[~inner] == 7.25), "unproject wrote boundary cells"
    assert rel(peel(f_sep, DIMS), peel(results[-1][1], DIMS)) <= 1e-6
    assert gb.check_stream_errors() == 0


def test_step_matches_the_dense_twin(full):
    ob, gb, v, q, dv, dq, results = full
    cols = gb.decode_columns().numpy()
    assert cols.shape == (R, 3 * DIMS[0] * DIMS[1] * DIMS[2])
    # every column against the FP64 oracle's DecodeVectorField
    errs = []
    for c0 in range(0, R, 16):
        want = ob.decode_vector_fields_mt(c0, min(R, c0 + 16), THREADS)
        errs += [rel(cols[c0 + i], want[i]) for i in range(len(want))]
        del want
    errs = np.array(errs)
    print(f"[full] all {R} decoded columns vs FP64 oracle: rel-L2 max {errs.max():.2e} "
          f"(column {int(np.argmax(errs))}), mean {errs.mean():.2e}", flush=True)
    assert np.all(errs <= TWIN_TOL)
    vp = peel(v, DIMS)
    q_twin = np.zeros(R)
    f_twin = np.zeros(cols.shape[1])
    for c0 in range(0, R, 10):
        D = cols[c0:c0 + 10].astype(np.float64)
        q_twin[c0:c0 + 10] = D @ vp
        f_twin += q[c0:c0 + 10] @ D
    eq = rel(results[-1][0], q_twin)
    ef = rel(peel(results[-1][1], DIMS), f_twin)
    print(f"[full] fused step vs dense twin: q rel-L2 {eq:.2e}, field rel-L2 {ef:.2e}", flush=True)
    assert eq <= TWIN_TOL and ef <= TWIN_TOL
    assert gb.check_stream_errors() == 0
