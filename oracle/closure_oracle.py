"""FP64 CPU oracle of the reduced advection and the learned closure.

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py as the checker.
Nothing in cfs_plume_amd/ imports it.

advect():          VECTOR3_FIELD_3D::advect (reference VECTOR3_FIELD_3D.cpp:2001-2077),
                   first-order semi-Lagrangian self-advection of a full-grid
                   interleaved field: backtrace by dt0 * v, clamp to
                   [1.5, res - 2.5], trilinear interpolation in the reference's
                   evaluation order.
advect_scalar():   the same for a scalar field along a velocity field
                   (VECTOR3_FIELD_3D.cpp:2082-2160).
reduced_advect():  the quantity the cubature step reducedAdvectCompressionFriendly
                   (SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:1781-1854) approximates:
                   U^T A(U q) with the FP64 decoder (oracle.Basis).
mlp_*():           the closure network r -> 256 -> 256 -> 256 -> r with GELU
                   (erf form) and its mean-squared-error gradients, evaluated in
                   FP64 with the BF16-rounded weights the device uses.

Pins (tests/test_closure_oracle.py): the backward pass against central finite
differences, GELU' against finite differences, and advect() against closed
forms (a linear field under a uniform velocity, a constant field, dt0 = 0);
advect_scalar() against advect() component by component and against a
linear density under a uniform velocity.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

HIDDEN = 256


# ---------------------------------------------------------------------------
# advection
# ---------------------------------------------------------------------------
def advect(vel: np.ndarray, full_dims, dt0: float) -> np.ndarray:
    X, Y, Z = full_dims
    v = np.asarray(vel, np.float64).reshape(Z, Y, X, 3)
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    xt = np.clip(x - dt0 * v[..., 0], 1.5, X - 2.5)
    yt = np.clip(y - dt0 * v[..., 1], 1.5, Y - 2.5)
    zt = np.clip(z - dt0 * v[..., 2], 1.5, Z - 2.5)
    x0, y0, z0 = xt.astype(np.int64), yt.astype(np.int64), zt.astype(np.int64)
    s1, t1, u1 = xt - x0, yt - y0, zt - z0
    s0, t0, u0 = 1.0 - s1, 1.0 - t1, 1.0 - u1
    out = np.empty_like(v)
    for c in range(3):
        f = v[..., c]

        def at(dx, dy, dz):
            return f[z0 + dz, y0 + dy, x0 + dx]

        out[..., c] = (u0 * (s0 * (t0 * at(0, 0, 0) + t1 * at(0, 1, 0)) + s1 * (t0 * at(1, 0, 0) + t1 * at(1, 1, 0))) +
                       u1 * (s0 * (t0 * at(0, 0, 1) + t1 * at(0, 1, 1)) + s1 * (t0 * at(1, 0, 1) + t1 * at(1, 1, 1))))
    return out.reshape(-1)


def advect_scalar(vel: np.ndarray, old: np.ndarray, full_dims, dt0: float) -> np.ndarray:
    """VECTOR3_FIELD_3D::advect for a FIELD_3D (reference VECTOR3_FIELD_3D.cpp:2082-2160):
    the scalar field `old` sampled trilinearly at each cell's clamped backtrace along vel."""
    X, Y, Z = full_dims
    v = np.asarray(vel, np.float64).reshape(Z, Y, X, 3)
    f = np.asarray(old, np.float64).reshape(Z, Y, X)
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    xt = np.clip(x - dt0 * v[..., 0], 1.5, X - 2.5)
    yt = np.clip(y - dt0 * v[..., 1], 1.5, Y - 2.5)
    zt = np.clip(z - dt0 * v[..., 2], 1.5, Z - 2.5)
    x0, y0, z0 = xt.astype(np.int64), yt.astype(np.int64), zt.astype(np.int64)
    s1, t1, u1 = xt - x0, yt - y0, zt - z0
    s0, t0, u0 = 1.0 - s1, 1.0 - t1, 1.0 - u1

    def at(dx, dy, dz):
        return f[z0 + dz, y0 + dy, x0 + dx]

    out = (u0 * (s0 * (t0 * at(0, 0, 0) + t1 * at(0, 1, 0)) + s1 * (t0 * at(1, 0, 0) + t1 * at(1, 1, 0))) +
           u1 * (s0 * (t0 * at(0, 0, 1) + t1 * at(0, 1, 1)) + s1 * (t0 * at(1, 0, 1) + t1 * at(1, 1, 1))))
    return out.reshape(-1)


def reduced_advect(basis, q: np.ndarray, dt0: float, alpha: float = 1.0, forcing=None) -> np.ndarray:
    """alpha * U^T A(U q) + forcing with the FP64 oracle basis."""
    X, Y, Z = basis.dims
    full = (X + 2, Y + 2, Z + 2)
    v = basis.unproject(q, np.zeros(3 * full[0] * full[1] * full[2]))
    a = advect(v, full, dt0)
    out = alpha * basis.project(a)
    if forcing is not None:
        out = out + forcing
    return out


# ---------------------------------------------------------------------------
# the closure network
# ---------------------------------------------------------------------------
def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round FP32 values to BF16 (nearest even), returned as FP64."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def padded(r: int) -> int:
    return (r + 63) // 64 * 64


def layer_shapes(r: int):
    rp = padded(r)
    return [(rp, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, rp)]


def unpack(params: np.ndarray, r: int):
    """Flat FP32 parameters -> [(W_l, b_l)] (W input-major, K x M)."""
    out, o = [], 0
    for K, M in layer_shapes(r):
        W = params[o:o + K * M].reshape(K, M)
        o += K * M
        b = params[o:o + M]
        o += M
        out.append((W, b))
    assert o == params.size
    return out


def pack(layers) -> np.ndarray:
    return np.concatenate([np.concatenate([W.reshape(-1), b]) for W, b in layers]).astype(np.float32)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def mlp_forward(params: np.ndarray, x: np.ndarray, r: int, bf16: bool = True):
    """x: [n][r_padded].  Returns (output [n][r_padded], cache)."""
    layers = unpack(np.asarray(params, np.float32), r)
    h = np.asarray(x, np.float64)
    cache = []
    for l, (W, b) in enumerate(layers):
        Wd = bf16_round(W) if bf16 else W.astype(np.float64)
        z = h @ Wd + b.astype(np.float64)
        cache.append((h, z, Wd))
        h = gelu(z) if l < 3 else z
    return h, cache


def mlp_gradients(params: np.ndarray, x: np.ndarray, t: np.ndarray, r: int, bf16: bool = True):
    """Gradients of mean((y - t)^2) over the first r outputs and all rows,
    laid out like the parameters, and the sum of squared errors."""
    y, cache = mlp_forward(params, x, r, bf16)
    n = x.shape[0]
    d = np.zeros_like(y)
    d[:, :r] = y[:, :r] - np.asarray(t, np.float64)[:, :r]
    sse = float((d * d).sum())
    dz = 2.0 * d / (n * r)
    grads = []
    for l in range(3, -1, -1):
        h, z, Wd = cache[l]
        gW = h.T @ dz
        gb = dz.sum(axis=0)
        grads.append((gW, gb))
        if l > 0:
            dz = (dz @ Wd.T) * gelu_grad(cache[l - 1][1])
    grads.reverse()
    return np.concatenate([np.concatenate([gW.reshape(-1), gb]) for gW, gb in grads]), sse


def closure_step(params, s_in, s_out, q: np.ndarray, r: int, alpha: float = 1.0, forcing=None, L=None):
    """alpha * (q + s_out * (MLP(x) + x L)) + forcing, x = s_in * q, for rows
    of q ([n][r]); L the optional linear skip (r x r, input-major)."""
    q = np.atleast_2d(np.asarray(q, np.float64))
    x = np.zeros((q.shape[0], padded(r)))
    x[:, :r] = q * np.asarray(s_in, np.float64)
    y, _ = mlp_forward(params, x, r)
    y = y[:, :r]
    if L is not None:
        y = y + x[:, :r] @ np.asarray(L, np.float64)
    out = alpha * (q + y * np.asarray(s_out, np.float64))
    if forcing is not None:
        out = out + forcing
    return out
