"""Pins of the moving-obstacle oracle (oracle/obstacle_oracle.py), CPU only:
the box geometry and schedule against closed forms, the reduced IOP pair of
buildObstacleMatrices against the full-space stomp of
setPeeledSparseMovingIOP, and the full-rank obstacle step's stomp, solve and
rest state."""
import math

import numpy as np

from oracle import fluid_oracle as FO
from oracle import obstacle_oracle as OB

FULL = (14, 16, 12)


def test_rotation_is_a_proper_rotation_about_z():
    for theta in (0.3, 1.0, 2.5):
        R = OB.rotation((0, 0, 1), -theta)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-15)
        assert math.isclose(np.linalg.det(R), 1.0, rel_tol=1e-14)
        assert np.allclose(R @ [1.0, 0.0, 0.0], [math.cos(theta), -math.sin(theta), 0.0], atol=1e-15)
        assert np.allclose(R @ [0.0, 0.0, 1.0], [0.0, 0.0, 1.0])


def test_quarter_turn_swaps_the_box_extents():
    box = OB.Box()
    c = box.center
    along_x = c + [0.15, 0.0, 0.0]
    along_y = c + [0.0, 0.15, 0.0]
    assert box.inside([along_x])[0] and not box.inside([along_y])[0]
    box.theta = math.pi / 2
    box.R = OB.rotation((0, 0, 1), -box.theta)
    assert not box.inside([along_x])[0] and box.inside([along_y])[0]
    # the z extent does not turn
    assert box.inside([c + [0.0, 0.0, 0.07]])[0] and not box.inside([c + [0.0, 0.0, 0.08]])[0]


def test_box_velocity_is_a_rigid_rotation_plus_the_slide():
    box = OB.Box()
    for _ in range(7):
        box.advance()
    rng = np.random.default_rng(0)
    pts = box.center + rng.uniform(-0.2, 0.2, size=(50, 3))
    v = box.velocity_at(pts) - box.slide_velocity()
    r = pts - box.center
    r[:, 2] = 0.0
    omega = 2 * math.pi / box.period
    assert np.allclose(np.sum(v * r, axis=1), 0.0, atol=1e-15)  # tangential
    assert np.allclose(np.linalg.norm(v, axis=1), omega * np.linalg.norm(r, axis=1), rtol=1e-14)
    assert np.allclose(v[:, 2], 0.0)
    # counter-clockwise about +z: at r = +x the velocity points along +y
    vx = box.velocity_at([box.center + [0.1, 0.0, 0.0]])[0] - box.slide_velocity()
    assert vx[1] > 0 and abs(vx[0]) < 1e-15


def test_schedule_follows_the_closed_forms():
    box = OB.Box(dt=0.1)
    for k in range(1, 40):
        box.advance()
        t = 0.1 * k
        assert math.isclose(box.center[0], 0.5 + 0.25 * math.sin(2 * math.pi * t / 10.0), rel_tol=1e-12, abs_tol=1e-14)
        assert math.isclose(box.theta, 2 * math.pi * t / 4.0, rel_tol=1e-12)
        want = 0.25 * 2 * math.pi / 10.0 * math.cos(2 * math.pi * t / 10.0)
        assert math.isclose(box.slide_velocity()[0], want, rel_tol=1e-12, abs_tol=1e-14)
    for _ in range(10):
        box.advance()
        assert 0.0 <= box.theta < 2 * math.pi  # wraps to zero at a full turn


def test_stomp_sets_the_box_cells_only():
    box = OB.Box()
    for _ in range(3):
        box.advance()
    rng = np.random.default_rng(1)
    vp = rng.standard_normal(3 * (FULL[0] - 2) * (FULL[1] - 2) * (FULL[2] - 2))
    inside, nvec = OB.complement(FULL, box)
    assert 0 < inside.sum() < len(inside)
    out = OB.stomp_peeled(FULL, box, vp).reshape(-1, 3)
    pts = OB.cell_centers(FULL)
    assert np.array_equal(out[inside], box.velocity_at(pts[inside]))
    assert np.array_equal(out[~inside], vp.reshape(-1, 3)[~inside])
    assert np.array_equal(nvec.reshape(-1, 3)[~inside], np.zeros(((~inside).sum(), 3)))


def test_reduced_pair_equals_the_projected_stomp():
    """buildObstacleMatrices' R q + n equals U_iop^T stomp(U_proj q)."""
    box = OB.Box()
    for _ in range(5):
        box.advance()
    rng = np.random.default_rng(2)
    rows = 3 * (FULL[0] - 2) * (FULL[1] - 2) * (FULL[2] - 2)
    U_iop = np.linalg.qr(rng.standard_normal((rows, 9)))[0]
    U_proj = np.linalg.qr(rng.standard_normal((rows, 7)))[0]
    R, n = OB.reduced_iop(FULL, box, U_iop, U_proj)
    for _ in range(3):
        q = rng.standard_normal(7)
        want = U_iop.T @ OB.stomp_peeled(FULL, box, U_proj @ q)
        assert np.allclose(R @ q + n, want, rtol=1e-13, atol=1e-14)
    # with no cell inside, R is U_iop^T U_proj and n vanishes
    box.center[:] = (5.0, 5.0, 5.0)
    R0, n0 = OB.reduced_iop(FULL, box, U_iop, U_proj)
    assert np.allclose(R0, U_iop.T @ U_proj) and not np.any(n0)


def test_obstacle_step_stomps_then_projects():
    solver = FO.Solver(FULL)
    X, Y, Z = FULL
    n = X * Y * Z
    rng = np.random.default_rng(3)
    vel = 0.1 * rng.standard_normal(3 * n)
    dens = np.zeros(n)
    heat = np.zeros(n)
    box = OB.Box(dt=solver.prm["dt"])
    box.advance()
    v, d, h, snaps, (its, mx) = OB.step_moving_obstacle(solver, vel, dens, heat, box)
    assert set(snaps) == {"preadvect", "prediffuse", "preproject", "iop", "final_velocity", "pressure", "divergence"}
    inside, nvec = OB.complement(FULL, box)
    iop = snaps["iop"].reshape(-1, 3)
    assert np.array_equal(iop[inside], nvec.reshape(-1, 3)[inside])
    assert np.array_equal(iop[~inside], snaps["preproject"].reshape(-1, 3)[~inside])
    # the pressure solve of the stomped field converged
    assert its > 0 and mx <= 1e-10
    assert np.any(snaps["divergence"])
    # no buoyancy: a fluid at rest with the box out of the domain stays at rest
    far = OB.Box(center=(5.0, 5.0, 5.0), dt=solver.prm["dt"])
    v2, _, _, _, _ = OB.step_moving_obstacle(solver, np.zeros(3 * n), dens, np.ones(n), far)
    assert not np.any(v2)
