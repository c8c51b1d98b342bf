"""The box of the C ABI (cfs_box_init / cfs_box_advance, host code in
csrc/obstacle.hip) against the oracle's BOX over a long schedule, and the
ABI's argument checks (CPU only: these entry points do not touch the GPU)."""
import ctypes

import numpy as np

from cfs_plume_amd import _lib
from cfs_plume_amd import obstacle as OBX
from oracle import obstacle_oracle as OB


def test_box_schedule_matches_the_oracle_exactly():
    lib_box = OBX.Box(dt=0.1)
    ref = OB.Box(dt=0.1)
    for step in range(130):  # past a full spin (40 steps) and a full slide (100 steps)
        assert np.array_equal(lib_box.center, ref.center), step
        assert lib_box.theta == ref.theta and lib_box.phi == ref.phi, step
        lib_box.advance()
        ref.advance()
    assert ref.theta < 2 * np.pi and ref.phi < 2 * np.pi


def test_box_copy_is_independent():
    a = OBX.Box()
    b = a.copy()
    a.advance()
    assert b.theta == 0.0 and a.theta > 0.0


def test_box_arguments_are_checked():
    lib = _lib.lib()
    st = _lib.Box()
    c = (ctypes.c_double * 3)(0.5, 0.5, 0.5)
    ln = (ctypes.c_double * 3)(0.4, 0.05, 0.15)
    assert lib.cfs_box_init(c, ln, 0.0, 10.0, 0.1, ctypes.byref(st)) == -1
    assert lib.cfs_box_init(c, ln, 4.0, 10.0, 0.1, None) == -1
    bad = (ctypes.c_double * 3)(0.4, -1.0, 0.15)
    assert lib.cfs_box_init(c, bad, 4.0, 10.0, 0.1, ctypes.byref(st)) == -1
    assert lib.cfs_box_advance(None) == -1
    assert lib.cfs_reduced_iop_scratch_bytes(0, 4, 4) == -1
    assert lib.cfs_reduced_iop_scratch_bytes(4, 4, 4) >= 3 * 64 * 8 + 64
    assert lib.cfs_reduced_iop(None, 1, None, 1, None, 4, 4, 4, None, None, None, None, None) == -1
