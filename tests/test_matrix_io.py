"""The reference's `.matrix` layout (EIGEN::read / EIGEN::write,
src/linearalgebra/EIGEN.cpp:24-121, 245-290) and the per-step comparison of
compareSubspace (SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:271-289)."""
import struct

import numpy as np
import pytest

from cfs_plume_amd.accuracy import relative_l2_per_step
from cfs_plume_amd.matrix_io import TrajectoryWriter, read_matrix, write_matrix


def test_bytes_follow_the_reference_layout(tmp_path):
    p = tmp_path / "m.matrix"
    write_matrix(p, np.array([[1.5, -2.0, 0.25], [3.0, 4.0, -8.5]]))
    # int32 rows, int32 cols, then FP64 row by row
    want = struct.pack("<ii", 2, 3) + struct.pack("<6d", 1.5, -2.0, 0.25, 3.0, 4.0, -8.5)
    assert p.read_bytes() == want


def test_round_trip(tmp_path):
    M = np.random.default_rng(3).standard_normal((17, 5))
    p = tmp_path / "r.matrix"
    write_matrix(p, M)
    assert np.array_equal(read_matrix(p), M)


def test_trajectory_writer_matches_a_whole_matrix(tmp_path):
    Q = np.random.default_rng(4).standard_normal((6, 4))
    w = TrajectoryWriter(tmp_path / "qDot.matrix", 6, 4)
    for row in Q:
        w.append(row)
    write_matrix(tmp_path / "whole.matrix", Q)
    assert (tmp_path / "qDot.matrix").read_bytes() == (tmp_path / "whole.matrix").read_bytes()
    with pytest.raises(ValueError):
        w.append(np.zeros(3))


def test_bad_files_are_refused(tmp_path):
    with pytest.raises(ValueError):
        write_matrix(tmp_path / "x.matrix.gz", np.ones((2, 2)))
    short = tmp_path / "short.matrix"
    short.write_bytes(struct.pack("<i", 3))
    with pytest.raises(ValueError):
        read_matrix(short)
    zero = tmp_path / "zero.matrix"
    zero.write_bytes(struct.pack("<ii", 0, 4))
    with pytest.raises(ValueError):
        read_matrix(zero)
    trunc = tmp_path / "trunc.matrix"
    trunc.write_bytes(struct.pack("<ii", 2, 2) + struct.pack("<3d", 1, 2, 3))
    with pytest.raises(ValueError):
        read_matrix(trunc)


def test_relative_l2_per_step_closed_form():
    q_true = np.array([[3.0, 4.0], [0.0, 2.0], [0.0, 0.0]])
    q = np.array([[3.0, 4.5], [0.0, 2.0], [0.0, 0.0]])
    e = relative_l2_per_step(q_true, q)
    assert e[0] == pytest.approx(0.5 / 5.0, rel=1e-15)
    assert e[1] == 0.0 and e[2] == 0.0
    assert np.isinf(relative_l2_per_step(np.zeros((1, 2)), np.ones((1, 2))))[0]
    with pytest.raises(ValueError):
        relative_l2_per_step(np.zeros((2, 3)), np.zeros((3, 2)))
