"""GPU: the reduced advection and the learned closure against the FP64 oracle.

* advect_field (VECTOR3_FIELD_3D::advect) vs closure_oracle.advect: rel-L2
  <= 1e-6, and the linear closed form of the oracle's pin on the device.
* reduced_advect (U^T A(U q) on the full grid) vs the oracle: rel-L2 <= 1e-5.
* MLP forward on MFMA (BF16 weights, BF16 hi + lo activations, FP32
  accumulation) vs the FP64 forward with the same BF16 weights: <= 1e-4.
* Fused closure step (BF16 weights, FP32 arithmetic) vs the oracle: <= 1e-5.
* Gradients of the device backward pass vs the oracle's (themselves pinned
  against finite differences on the CPU): per-layer rel-L2 <= 1e-3.
* Adam training on the device lowers the loss of a fixed regression.
* A closure trained on exact reduced-plume trajectories tracks held-out runs
  far better than the step without advection, and the checked rollout (exact
  step as fallback) stays within its tolerance.
"""
import numpy as np
import pytest

from cfs_plume_amd import Closure, CompressedBasis, DeviceArray, ReducedPlume, advect_field, encode_matrix, \
    reduced_advect, synth
from cfs_plume_amd import closure as CL
from oracle import closure_oracle as C
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def small_basis(gpu):
    dims = (14, 12, 10)
    print(f"[case] closure basis dims={dims} r=6", flush=True)
    bufs = encode_matrix(synth.synthetic_basis(dims, 6, seed=21), dims).buffers()
    return O.Basis(buffers=bufs), CompressedBasis.from_buffers(bufs)


def test_advect_matches_oracle(gpu):
    full = (20, 18, 16)
    rng = np.random.default_rng(5)
    v = rng.normal(0, 1.5, 3 * np.prod(full)).astype(np.float32)
    dv = DeviceArray.from_numpy(v)
    out = DeviceArray((v.size,), np.float32)
    print("[case] advect random field 20x18x16 dt0=1.3", flush=True)
    advect_field(dv, out, full, 1.3)
    assert rel(out.numpy(), C.advect(v.astype(np.float64), full, 1.3)) <= 1e-6


def test_advect_linear_closed_form_on_device(gpu):
    full = (14, 12, 10)
    X, Y, Z = full
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    c, eps, dt0 = 0.75, 0.0625, 1.25
    v = np.zeros((Z, Y, X, 3), np.float32)
    v[..., 1] = c
    v[..., 2] = eps * y
    out = DeviceArray((v.size,), np.float32)
    advect_field(DeviceArray.from_numpy(v.reshape(-1)), out, full, dt0)
    o = out.numpy().reshape(Z, Y, X, 3)
    ty = y - dt0 * c
    tz = z - dt0 * eps * y
    ok = (ty >= 1.5) & (ty <= Y - 2.5) & (tz >= 1.5) & (tz <= Z - 2.5) & (x >= 2) & (x <= X - 3)
    np.testing.assert_allclose(o[..., 1][ok], c, atol=1e-6)
    np.testing.assert_allclose(o[..., 2][ok], (eps * (y - dt0 * c))[ok], atol=1e-6)


def test_reduced_advect_matches_oracle(small_basis):
    ob, gb = small_basis
    X, Y, Z = gb.dims
    n = 3 * (X + 2) * (Y + 2) * (Z + 2)
    rng = np.random.default_rng(7)
    q = rng.normal(0, 40.0, gb.r)
    f = rng.normal(0, 1.0, gb.r)
    dq = DeviceArray.from_numpy(q)
    out = DeviceArray((gb.r,), np.float64)
    fa, fb = DeviceArray((n,), np.float32), DeviceArray((n,), np.float32)
    print("[case] reduced advect r=6 dt0=0.8 alpha=0.97", flush=True)
    reduced_advect(gb, dq, out, fa, fb, 0.8, 0.97, DeviceArray.from_numpy(f))
    want = C.reduced_advect(ob, q, 0.8, 0.97, f)
    assert rel(out.numpy(), want) <= 1e-5


def _net(r, seed):
    cl = Closure(r, max_batch=256, seed=seed)
    p = cl.get_params()
    rng = np.random.default_rng(seed)
    # non-zero biases so that every term of the gradient is exercised
    o = 0
    for l, (K, M) in enumerate(cl.layer_shapes()):
        mout = r if l == 3 else M
        p[o + K * M:o + K * M + mout] = rng.normal(0, 0.1, mout)
        o += K * M + M
    cl.set_params(p)
    return cl, p


def test_forward_matches_oracle(gpu):
    r = 52
    cl, p = _net(r, 3)
    rng = np.random.default_rng(4)
    x = cl.pad(rng.normal(size=(37, r)))
    print("[case] MLP forward n=37 r=52 (MFMA)", flush=True)
    y = cl.forward(DeviceArray.from_numpy(x), 37).numpy()
    want, _ = C.mlp_forward(p, x.astype(np.float64), r)
    assert rel(y[:, :r], want[:, :r]) <= 1e-4
    assert np.all(y[:, r:] == 0.0)


def test_closure_step_matches_oracle(gpu):
    r = 20
    cl, p = _net(r, 5)
    rng = np.random.default_rng(6)
    s_in = rng.uniform(0.5, 2.0, r).astype(np.float32)
    s_out = rng.uniform(0.5, 2.0, r).astype(np.float32)
    cl.set_scales(s_in, s_out)
    q = rng.normal(size=(7, r))
    f = rng.normal(size=(7, r))
    out = DeviceArray((7, r), np.float64)
    print("[case] closure step n=7 r=20", flush=True)
    cl.step(DeviceArray.from_numpy(q), out, 7, 0.95, DeviceArray.from_numpy(f))
    want = C.closure_step(p, s_in, s_out, q, r, 0.95, f)
    assert rel(out.numpy(), want) <= 1e-5
    # one sample addressed as a row of a larger array
    out1 = DeviceArray((r,), np.float64)
    cl.step((DeviceArray.from_numpy(q), 3), out1, 1, 0.95, (DeviceArray.from_numpy(f), 3))
    assert rel(out1.numpy(), want[3]) <= 1e-5
    # with the FP32 linear skip
    L = rng.normal(0, 0.3, (r, r)).astype(np.float32)
    cl.set_linear(L)
    cl.step(DeviceArray.from_numpy(q), out, 7, 0.95, DeviceArray.from_numpy(f))
    assert rel(out.numpy(), C.closure_step(p, s_in, s_out, q, r, 0.95, f, L)) <= 1e-5


def test_gradients_match_oracle(gpu):
    r = 52
    cl, p = _net(r, 8)
    rng = np.random.default_rng(9)
    n = 45
    x = cl.pad(rng.normal(size=(n, r)))
    t = cl.pad(rng.normal(size=(n, r)))
    print("[case] MLP gradients n=45 r=52", flush=True)
    g, sse = cl.gradients(DeviceArray.from_numpy(x), DeviceArray.from_numpy(t), n)
    want, sse_want = C.mlp_gradients(p, x.astype(np.float64), t.astype(np.float64), r)
    assert abs(sse - sse_want) <= 1e-4 * sse_want
    o = 0
    for l, (K, M) in enumerate(cl.layer_shapes()):
        for a, b_ in ((o, o + K * M), (o + K * M, o + K * M + M)):
            assert rel(g[a:b_], want[a:b_]) <= 1e-3, (l, a, b_)
        o += K * M + M


def test_training_lowers_loss(gpu):
    r = 16
    rng = np.random.default_rng(11)
    A = rng.normal(0, 0.5, (r, r))
    B = rng.normal(0, 0.5, (r, r))
    x = rng.normal(size=(2048, r))
    t = np.tanh(x @ A) @ B
    cl = Closure(r, max_batch=256, seed=12)
    print("[case] training 2048 samples r=16", flush=True)
    losses = cl.train(x, t, epochs=25, batch=128, lr=2e-3, lr_final=2e-4, seed=13)
    assert losses[-1] < 0.1 * losses[0], losses


@pytest.fixture(scope="module")
def plume(gpu):
    dims = (22, 30, 22)
    print(f"[case] reduced plume dims={dims} r=20", flush=True)
    bufs = encode_matrix(synth.synthetic_basis(dims, 20, seed=31), dims).buffers()
    U = CompressedBasis.from_buffers(bufs)
    return ReducedPlume(U, dt0=1.0, alpha=0.98, force=25.0)


def test_closure_tracks_held_out_runs(plume):
    runs = ReducedPlume.seeded_runs(52, seed=41)
    train, held = runs[:48], runs[48:]
    steps = 30
    b_train = plume.forcing(train)
    traj = plume.rollout_exact(b_train, len(train), steps)
    q, d = CL.increments(traj, b_train.numpy(), plume.alpha)
    # exact increments also at states perturbed by 10 % around the trajectories
    rng = np.random.default_rng(23)
    qp = q + 0.1 * rng.normal(size=q.shape) * np.linalg.norm(q, axis=1, keepdims=True) / np.sqrt(q.shape[1])
    q, d = np.concatenate([q, qp]), np.concatenate([d, plume.exact_increments(qp)])
    cl = Closure(plume.U.r, max_batch=256, seed=17)
    print("[case] plume closure training 2880 samples", flush=True)
    losses = cl.fit(q, d, epochs=200, batch=128, lr=2e-3, lr_final=1e-5, seed=19)
    assert losses[-1] < 0.2 * losses[0]
    b_held = plume.forcing(held)
    bh = b_held.numpy()
    exact = plume.rollout_exact(b_held, len(held), steps)
    learned = plume.rollout_closure(cl, b_held, len(held), steps)
    s_in, s_out = cl.get_scales()
    # the method without the network: its least-squares linear part alone,
    # and no advection at all (q_{t+1} = alpha q_t + b)
    linear = np.zeros_like(exact)
    plain = np.zeros_like(exact)
    for t in range(steps):
        x = linear[:, t] * s_in
        linear[:, t + 1] = plume.alpha * (linear[:, t] + (x @ cl.L.astype(np.float64)) * s_out) + bh
        plain[:, t + 1] = plume.alpha * plain[:, t] + bh
    err_cl = rel(learned[:, -1], exact[:, -1])
    err_lin = rel(linear[:, -1], exact[:, -1])
    err_plain = rel(plain[:, -1], exact[:, -1])
    print(f"held-out rel error at step {steps}: closure {err_cl:.3e}, linear part alone {err_lin:.3e}, "
          f"no advection {err_plain:.3e}")
    assert err_cl < err_lin
    assert err_cl < 0.1 * err_plain
    checked, falls = plume.rollout_closure_checked(cl, b_held, 0, steps, check_every=1, tol=0.01)
    # with a check every step the fallback keeps each step within 1 % of the exact step
    assert rel(checked[-1], exact[0, -1]) <= 0.05
    assert 0 <= falls <= steps


def test_per_step_paths_replay_in_a_hip_graph(small_basis):
    """The exact reduced advection and the closure's gradient pass (both
    clear device memory with a fill kernel, not a memset node) captured in a
    HIP graph: every replay, with eager work and synchronisations between the
    replays, equals the eager result."""
    from cfs_plume_amd import _lib
    from cfs_plume_amd.hip import Graph, Stream
    ob, gb = small_basis
    X, Y, Z = gb.dims
    n = 3 * (X + 2) * (Y + 2) * (Z + 2)
    rng = np.random.default_rng(17)
    dq = DeviceArray.from_numpy(rng.normal(0, 40.0, gb.r))
    df = DeviceArray.from_numpy(rng.normal(0, 1.0, gb.r))
    fa, fb = DeviceArray((n,), np.float32), DeviceArray((n,), np.float32)
    q_e, q_g = DeviceArray((gb.r,), np.float64), DeviceArray((gb.r,), np.float64)
    r, nb = 52, 40
    cl, _ = _net(r, 23)
    x = DeviceArray.from_numpy(cl.pad(rng.normal(size=(nb, r))))
    t = DeviceArray.from_numpy(cl.pad(rng.normal(size=(nb, r))))
    g_e, g_g = DeviceArray((cl.n_params,), np.float32), DeviceArray((cl.n_params,), np.float32)
    l_e, l_g = DeviceArray((1,), np.float64), DeviceArray((1,), np.float64)
    s = Stream()
    lib = _lib.lib()

    def work(q_out, grads, loss):
        reduced_advect(gb, dq, q_out, fa, fb, 0.8, 0.97, df, s)
        _lib.check(lib.cfs_closure_gradients(cl.handle, x.ptr, t.ptr, nb, grads.ptr, loss.ptr, s.handle),
                   "cfs_closure_gradients")

    print("[case] graph replay of reduced advect r=6 and closure gradients n=40", flush=True)
    work(q_e, g_e, l_e)
    s.synchronize()
    eager = (q_e.numpy(), g_e.numpy(), l_e.numpy())
    graph = Graph(s, lambda: work(q_g, g_g, l_g))
    for rep in range(3):
        for a in (q_g, g_g, l_g):
            a.fill_zero()
        graph.launch()
        s.synchronize()
        assert np.array_equal(q_g.numpy(), eager[0]), f"replay {rep}: q"
        assert rel(g_g.numpy(), eager[1]) <= 1e-6, f"replay {rep}: gradients"
        assert rel(l_g.numpy(), eager[2]) <= 1e-12, f"replay {rep}: loss"
        work(q_e, g_e, l_e)  # eager work between the replays
        s.synchronize()
        assert np.array_equal(q_e.numpy(), eager[0])
    cl.close()


def test_advect_scalar_matches_oracle(gpu):
    """cfs_advect_scalar against the FP64 oracle (VECTOR3_FIELD_3D.cpp:2082-2160)."""
    from cfs_plume_amd import advect_scalar
    full = (20, 16, 12)
    X, Y, Z = full
    rng = np.random.default_rng(8)
    vel = (rng.standard_normal(3 * X * Y * Z) * 1.5).astype(np.float32)
    dens = rng.random(X * Y * Z).astype(np.float32)
    out = DeviceArray.zeros((X * Y * Z,), np.float32)
    advect_scalar(DeviceArray.from_numpy(vel), DeviceArray.from_numpy(dens), out, full, 1.25)
    want = C.advect_scalar(vel.astype(np.float64), dens.astype(np.float64), full, 1.25)
    got = out.numpy().astype(np.float64)
    assert np.max(np.abs(got - want)) <= 1e-5 * max(1.0, np.max(np.abs(want)))
