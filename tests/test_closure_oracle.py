"""Pins of the closure oracle (oracle/closure_oracle.py) on the CPU.

* The MLP backward pass against central finite differences of the FP64 loss.
* GELU' against finite differences.
* advect() (VECTOR3_FIELD_3D::advect, reference VECTOR3_FIELD_3D.cpp:2001-2077)
  against closed forms: a field linear in one coordinate under a velocity
  that is uniform along that axis is advected exactly by trilinear
  interpolation (value at x - dt0 * v), which fixes the sign of the
  backtrace, the axis of each component and the interpolation; a constant
  field stays constant; dt0 = 0 is the identity away from the clamp.
"""
import numpy as np
import pytest

from oracle import closure_oracle as C


def test_gelu_grad_matches_finite_differences():
    z = np.linspace(-6, 6, 241)
    h = 1e-6
    fd = (C.gelu(z + h) - C.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - C.gelu_grad(z))) < 1e-8


def _random_params(r, seed, scale=0.1):
    rng = np.random.default_rng(seed)
    layers = []
    for l, (K, M) in enumerate(C.layer_shapes(r)):
        W = np.zeros((K, M))
        kin = r if l == 0 else K
        mout = r if l == 3 else M
        W[:kin, :mout] = rng.normal(0, scale, (kin, mout))
        b = np.zeros(M)
        b[:mout] = rng.normal(0, scale, mout)
        layers.append((W, b))
    return C.pack(layers)


def test_mlp_gradients_match_finite_differences():
    r = 5
    rp = C.padded(r)
    rng = np.random.default_rng(1)
    p = _random_params(r, 2, scale=0.2).astype(np.float64)
    x = np.zeros((4, rp))
    x[:, :r] = rng.normal(size=(4, r))
    t = np.zeros((4, rp))
    t[:, :r] = rng.normal(size=(4, r))

    def loss(pp):
        y, _ = C.mlp_forward(pp, x, r, bf16=False)
        return float(((y[:, :r] - t[:, :r]) ** 2).mean())

    # FP64 parameters for the finite differences (the FP32 cast inside is avoided)
    def fwd64(pp):
        layers = []
        o = 0
        for K, M in C.layer_shapes(r):
            layers.append((pp[o:o + K * M].reshape(K, M), pp[o + K * M:o + K * M + M]))
            o += K * M + M
        h = x
        for l, (W, b) in enumerate(layers):
            z = h @ W + b
            h = C.gelu(z) if l < 3 else z
        return float(((h[:, :r] - t[:, :r]) ** 2).mean())

    g, sse = C.mlp_gradients(p, x, t, r, bf16=False)
    assert abs(sse / (4 * r) - loss(p)) < 1e-6 * max(1.0, loss(p))
    # every layer's weights and biases, live entries only
    offs, o = [], 0
    for l, (K, M) in enumerate(C.layer_shapes(r)):
        kin = r if l == 0 else K
        mout = r if l == 3 else M
        for _ in range(6):
            k, m = rng.integers(kin), rng.integers(mout)
            offs.append(o + k * M + m)
        for _ in range(3):
            offs.append(o + K * M + rng.integers(mout))
        o += K * M + M
    h = 1e-6
    for i in offs:
        e = np.zeros_like(p)
        e[i] = h
        fd = (fwd64(p + e) - fwd64(p - e)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_round_is_nearest_even():
    # BF16 keeps 7 fraction bits: ulp(1) = 2^-7; ties go to the even neighbour
    a = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5, 3.14159265],
                 np.float32)
    got = C.bf16_round(a)
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0, 1.0 + 2 ** -7, -2.5, 3.140625])
    assert np.array_equal(got, want)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_advect_linear_field_closed_form(axis):
    full = (14, 12, 10)
    X, Y, Z = full
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    coords = [x, y, z]
    c, eps, dt0 = 0.7, 0.05, 1.3
    other = (axis + 1) % 3
    v = np.zeros((Z, Y, X, 3))
    v[..., axis] = c                       # uniform along `axis`
    v[..., other] = eps * coords[axis]     # linear in the `axis` coordinate
    out = C.advect(v.reshape(-1), full, dt0).reshape(Z, Y, X, 3)
    res = [X, Y, Z]
    # cells whose backtrace is not clamped
    tr = [coords[d] - dt0 * v[..., d] for d in range(3)]
    ok = np.ones_like(x, bool)
    for d in range(3):
        ok &= (tr[d] >= 1.5) & (tr[d] <= res[d] - 2.5)
    assert ok.sum() > 100
    np.testing.assert_allclose(out[..., axis][ok], c, rtol=0, atol=1e-12)
    np.testing.assert_allclose(out[..., other][ok], (eps * (coords[axis] - dt0 * c))[ok], rtol=0, atol=1e-12)


def test_advect_constant_and_zero_step():
    full = (9, 8, 7)
    X, Y, Z = full
    rng = np.random.default_rng(3)
    v = np.tile(np.array([0.3, -0.2, 0.5]), X * Y * Z)
    np.testing.assert_allclose(C.advect(v, full, 2.0), v, atol=1e-12)
    f = rng.normal(size=3 * X * Y * Z)
    out = C.advect(f, full, 0.0).reshape(Z, Y, X, 3)
    f = f.reshape(Z, Y, X, 3)
    np.testing.assert_array_equal(out[2:Z - 2, 2:Y - 2, 2:X - 2], f[2:Z - 2, 2:Y - 2, 2:X - 2])


def test_advect_scalar_agrees_with_the_vector_advection():
    """advect_scalar (VECTOR3_FIELD_3D.cpp:2082-2160) of one component of a
    field along that same field equals that component of the self-advection
    (VECTOR3_FIELD_3D.cpp:2001-2077), whose closed forms are pinned above."""
    full = (9, 8, 7)
    X, Y, Z = full
    v = np.random.default_rng(5).standard_normal(3 * X * Y * Z) * 0.8
    vec = C.advect(v, full, 1.3).reshape(-1, 3)
    for c in range(3):
        s = C.advect_scalar(v, v.reshape(-1, 3)[:, c], full, 1.3)
        np.testing.assert_allclose(s, vec[:, c], rtol=0, atol=1e-14)


def test_advect_scalar_linear_density_closed_form():
    """A density linear in x moved by a uniform x velocity u: inside the clamp
    window the result is a (x - dt0 u) + b exactly."""
    full = (12, 7, 6)
    X, Y, Z = full
    a, b, u, dt0 = 0.75, -2.0, 0.6, 1.5
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    dens = (a * x + b).astype(np.float64)
    vel = np.zeros((Z, Y, X, 3))
    vel[..., 0] = u
    out = C.advect_scalar(vel.reshape(-1), dens.reshape(-1), full, dt0).reshape(Z, Y, X)
    xt = x - dt0 * u
    inside = (xt >= 1.5) & (xt <= X - 2.5) & (y >= 2) & (y <= Y - 3) & (z >= 2) & (z <= Z - 3)
    np.testing.assert_allclose(out[inside], (a * xt + b)[inside], rtol=0, atol=1e-12)
