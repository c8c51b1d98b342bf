"""Batched step: several bases of different grids and column counts in one
persistent launch, the autotuner, and HIP-graph capture of the whole step.
Every basis's result must equal its own fused step (field bit for bit, q to
FP64 re-association) and the FP64 oracle within the step's bound."""
import numpy as np
import pytest

from cfs_plume_amd import BasisBatch, CompressedBasis, DeviceArray, Stream, encode_matrix, step, synth
from oracle import oracle as O
from tests import helpers as H
from tests.test_gpu_parity import adversarial_zz, runs_zz

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def bases(gpu):
    print("[case] batch of 5 bases: (30,22,14) r8, (14,14,6) r3, (17,9,8) r5 adversarial, (16,16,8) r6 runs, "
          "(5,3,2) r1", flush=True)
    bufs = [
        encode_matrix(synth.synthetic_basis((30, 22, 14), 8, seed=3), (30, 22, 14)).buffers(),
        encode_matrix(synth.synthetic_basis((14, 14, 6), 3, seed=5), (14, 14, 6)).buffers(),
        H.basis_from_ints((17, 9, 8), 5, adversarial_zz, seed=11),
        H.basis_from_ints((16, 16, 8), 6, runs_zz, seed=13),
        encode_matrix(synth.synthetic_basis((5, 3, 2), 1, seed=9), (5, 3, 2)).buffers(),
    ]
    return [(O.Basis(buffers=b), CompressedBasis.from_buffers(b)) for b in bufs]


def _arrays(gb, seed):
    X, Y, Z = gb.dims
    v = synth.gaussian_field(gb.dims, seed)
    q = synth.gaussian_q(gb.r, seed + 1)
    return (v, q, DeviceArray.from_numpy(v.astype(np.float32)), DeviceArray.from_numpy(q),
            DeviceArray.zeros((gb.r,), np.float64), DeviceArray.zeros((len(v),), np.float32))


def _reference(bases):
    """Each basis through its own fused step, and through the oracle."""
    out = []
    for i, (ob, gb) in enumerate(bases):
        v, q, dv, dq, qo, fo = _arrays(gb, 100 + 7 * i)
        step(gb, gb, dv, qo, dq, fo)
        out.append((qo.numpy(), fo.numpy(), ob.project(v.astype(np.float32).astype(np.float64)),
                    ob.unproject(q, np.zeros(len(v)))))
    return out


def _batch(bases):
    arrs = [_arrays(gb, 100 + 7 * i) for i, (_, gb) in enumerate(bases)]
    bt = BasisBatch([gb for _, gb in bases])
    bt.bind([a[2] for a in arrs], [a[4] for a in arrs], [a[3] for a in arrs], [a[5] for a in arrs])
    return bt, arrs


def _check(bases, arrs, ref):
    for (ob, gb), a, (q1, f1, qo_want, fo_want) in zip(bases, arrs, ref):
        q, f = a[4].numpy(), a[5].numpy()
        assert np.array_equal(f, f1), "batched field differs from the basis's own step"
        assert rel(q, q1) <= 1e-12
        assert rel(q, qo_want) <= TOL and rel(f, fo_want) <= TOL
        assert gb.check_stream_errors() == 0


def test_batch_matches_individual_steps(bases):
    ref = _reference(bases)
    bt, arrs = _batch(bases)
    cfg = bt.config()
    assert cfg["grid"] >= len(bases)
    bt.step()
    _check(bases, arrs, ref)
    # every fixed configuration gives the same results
    for w in (1, 2):
        for iw in (0.0, 1024.0):
            bt.set_config(w, iw)
            for a in arrs:
                a[5].fill_zero()
            bt.step()
            _check(bases, arrs, ref)
    bt.close()


def test_batch_autotuner(bases):
    ref = _reference(bases)
    bt, arrs = _batch(bases)
    cfg = bt.tune(reps=2)
    assert cfg["wg_per_cu"] >= 1 and cfg["grid"] >= len(bases) and cfg["waves_per_wg"] in (2, 4)
    print(f"[batch] tuned {cfg}", flush=True)
    bt.set_shape(4)
    bt.step()
    _check(bases, arrs, ref)
    bt.close()


def test_batch_tile_shapes(bases):
    """Two-wave workgroups: the same field bit for bit (unproject does not
    depend on how items are grouped), q within 1e-6 of the basis's own step
    (the FP32 per-wave sums are grouped differently) and within the oracle
    tolerance."""
    ref = _reference(bases)
    bt, arrs = _batch(bases)
    for w in (1, 3):
        bt.set_shape(2)
        bt.set_config(w, 256.0)
        assert bt.config()["waves_per_wg"] == 2
        for a in arrs:
            a[5].fill_zero()
        bt.step()
        for (ob, gb), a, (q1, f1, qo_want, fo_want) in zip(bases, arrs, ref):
            q, f = a[4].numpy(), a[5].numpy()
            assert np.array_equal(f, f1)
            assert rel(q, q1) <= 1e-6
            assert rel(q, qo_want) <= TOL and rel(f, fo_want) <= TOL
    bt.close()


def test_batch_graph_replays(bases):
    """The captured step replays with eager steps and synchronisations in
    between; every replay equals the eager result."""
    ref = _reference(bases)
    bt, arrs = _batch(bases)
    s = Stream()
    bt.step(s)
    s.synchronize()
    eager = [(a[4].numpy(), a[5].numpy()) for a in arrs]
    bt.capture(s)
    for rep in range(3):
        for a in arrs:
            a[4].fill_zero()
            a[5].fill_zero()
        bt.replay(s)
        s.synchronize()
        for (qe, fe), a in zip(eager, arrs):
            assert np.array_equal(a[4].numpy(), qe), f"replay {rep}: q"
            assert np.array_equal(a[5].numpy(), fe), f"replay {rep}: field"
        bt.step(s)  # eager work between replays
        s.synchronize()
    _check(bases, arrs, ref)
    bt.close()
