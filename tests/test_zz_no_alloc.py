"""cfs_reduced_advect_noalloc (scratch fields in a fixed device array) against
cfs_reduced_advect (scratch fields passed by the caller): two calls in a row,
the second on a smaller grid, each bit-equal to the existing entry point and
within the step's bound of the FP64 oracle; a grid beyond the fixed scratch is
refused before any device work."""
import ctypes

import numpy as np
import pytest

from cfs_plume_amd import CompressedBasis, DeviceArray, _lib, encode_matrix, reduced_advect, synth
from oracle import closure_oracle as CO
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def _noalloc(gb, q_in, q_out, dt0, alpha, forcing):
    return _lib.lib().cfs_reduced_advect_noalloc(gb._h, ctypes.c_void_p(q_in.ptr), ctypes.c_void_p(q_out.ptr),
                                                 dt0, alpha, ctypes.c_void_p(forcing.ptr) if forcing else None,
                                                 None)


def test_noalloc_matches_the_workspace_entry_point(gpu):
    cases = [((46, 62, 46), 52, 7, True), ((28, 20, 12), 8, 3, False)]  # the second call is smaller
    for dims, r, seed, with_forcing in cases:
        print(f"[case] no-alloc reduced advection dims={dims} r={r}", flush=True)
        bufs = encode_matrix(synth.synthetic_basis(dims, r, seed=seed), dims, threads=16).buffers()
        gb = CompressedBasis.from_buffers(bufs)
        X, Y, Z = dims
        n = 3 * (X + 2) * (Y + 2) * (Z + 2)
        q = 40.0 * synth.gaussian_q(r, seed + 1)
        f = synth.gaussian_q(r, seed + 2) if with_forcing else None
        dq = DeviceArray.from_numpy(q)
        df = DeviceArray.from_numpy(f) if with_forcing else None
        q_ws = DeviceArray.zeros((r,), np.float64)
        q_na = DeviceArray.zeros((r,), np.float64)
        fa, fb = DeviceArray.zeros((n,), np.float32), DeviceArray.zeros((n,), np.float32)
        reduced_advect(gb, dq, q_ws, fa, fb, 1.0, 0.99, df)
        st = _noalloc(gb, dq, q_na, 1.0, 0.99, df)
        assert st == 0, _lib.lib().cfs_last_error().decode()
        got, want = q_na.numpy(), q_ws.numpy()
        assert np.array_equal(got, want), rel(got, want)
        ob = O.Basis(buffers=bufs)
        ref = CO.reduced_advect(ob, q, 1.0, 0.99, f)
        ob.close()
        e = rel(got, ref)
        print(f"[no-alloc] dims={dims} r={r}: equal to cfs_reduced_advect, FP64 oracle rel-L2 {e:.2e}", flush=True)
        assert e <= TOL
        assert gb.check_stream_errors() == 0


def test_noalloc_refuses_grids_beyond_its_scratch(gpu):
    dims = (110, 110, 110)  # full grid 112^3: 3 * 1,404,928 floats > 4,194,304
    bufs = encode_matrix(synth.synthetic_basis(dims, 1, seed=2), dims, threads=16).buffers()
    gb = CompressedBasis.from_buffers(bufs)
    q = DeviceArray.zeros((1,), np.float64)
    st = _noalloc(gb, q, q, 1.0, 1.0, None)
    assert st == -1
    assert "exceeds the no-allocation scratch" in _lib.lib().cfs_last_error().decode()
