"""Synthetic plume-like bases: seeded sums of random Fourier modes.

Mode m, component k = sum over 16 terms of a cos(2 pi (kx x/X + ky y/Y + kz z/Z) + phi)
on the peeled grid (x in [0, X)), with kx, ky, kz uniform integers in [1, 6],
a ~ N(0, 1), phi ~ U[0, 2 pi).  The term parameters come from a counter-based
generator keyed by (seed, m, k, term) so any subset of modes is reproducible on
its own.  Columns are stacked interleaved (row 3*(x + X*(y + Y*z)) + k) and
orthonormalised by a thin FP64 QR.

The field evaluation groups the terms of one (m, k) on the 6x6x6 wavenumber
lattice and contracts it axis by axis (separable complex exponentials), which
costs O(6 * cells) per (m, k) instead of O(16 * cells * trig).  The QR runs in
the coefficient space of these orthogonal functions (see synthetic_basis).
"""
from __future__ import annotations

import numpy as np

TERMS = 16
KMAX = 6


def _mix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser on uint64 counters."""
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    x = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    x = ((x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    return x ^ (x >> np.uint64(31))


def _uniform(seed: int, m: int, k: int, j: int, slot: int) -> float:
    key = np.uint64(((seed & 0xFFFF) << 48) | ((m & 0xFFFFF) << 28) | ((k & 0x3) << 26) | ((j & 0xFFFF) << 8) | slot)
    with np.errstate(over="ignore"):
        u = _mix64(np.array([key], dtype=np.uint64))[0]
    return (int(u >> np.uint64(11)) + 0.5) / float(1 << 53)


def mode_terms(seed: int, m: int, k: int):
    """(kx, ky, kz, amplitude, phase) of the 16 terms of mode m, component k."""
    out = []
    for j in range(TERMS):
        kx = 1 + int(_uniform(seed, m, k, j, 0) * KMAX)
        ky = 1 + int(_uniform(seed, m, k, j, 1) * KMAX)
        kz = 1 + int(_uniform(seed, m, k, j, 2) * KMAX)
        u1, u2 = _uniform(seed, m, k, j, 3), _uniform(seed, m, k, j, 4)
        amp = float(np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2))  # Box-Muller
        phase = 2.0 * np.pi * _uniform(seed, m, k, j, 5)
        out.append((kx, ky, kz, amp, phase))
    return out


def _lattice_index(k: int, kx: int, ky: int, kz: int) -> int:
    return ((k * KMAX + (kz - 1)) * KMAX + (ky - 1)) * KMAX + (kx - 1)


NFUN = 3 * KMAX ** 3  # (component, wavevector) pairs; each has a cos and a sin function


def mode_coefficients(seed: int, r: int) -> np.ndarray:
    """(2*NFUN, r) coefficients of the modes on the functions
    cos(theta_kappa) and sin(theta_kappa), theta_kappa = 2 pi (kx x/X + ky y/Y + kz z/Z):
    a cos(theta + phi) = (a cos phi) cos(theta) + (-a sin phi) sin(theta)."""
    M = np.zeros((2 * NFUN, r), np.float64)
    for m in range(r):
        for k in range(3):
            for kx, ky, kz, amp, ph in mode_terms(seed, m, k):
                f = _lattice_index(k, kx, ky, kz)
                M[f, m] += amp * np.cos(ph)
                M[NFUN + f, m] -= amp * np.sin(ph)
    return M


def _evaluate(C: np.ndarray, dims) -> np.ndarray:
    """Columns sum_f C[f, m] phi_f on the peeled grid, interleaved: (r, 3*X*Y*Z)."""
    X, Y, Z = dims
    r = C.shape[1]
    ks = np.arange(1, KMAX + 1)
    ex = np.exp(2j * np.pi * np.outer(ks, np.arange(X)) / X)  # (6, X)
    ey = np.exp(2j * np.pi * np.outer(ks, np.arange(Y)) / Y)
    ez = np.exp(2j * np.pi * np.outer(ks, np.arange(Z)) / Z)
    out = np.empty((r, Z, Y, X, 3), np.float64)
    for m in range(r):
        for k in range(3):
            sl = slice(k * KMAX ** 3, (k + 1) * KMAX ** 3)
            # Re(A e^{i theta}) with A = c_cos - i c_sin
            A = (C[sl, m] - 1j * C[NFUN:][sl, m]).reshape(KMAX, KMAX, KMAX)  # [kz, ky, kx]
            t = (A.reshape(KMAX * KMAX, KMAX) @ ex).reshape(KMAX, KMAX, X)  # (kz, ky, X)
            t = np.einsum("abx,by->axy", t, ey, optimize=True).reshape(KMAX, X * Y)  # (kz, X*Y)
            t = (ez.T @ t).reshape(Z, X, Y)  # (Z, X, Y)
            out[m, :, :, :, k] = t.real.transpose(0, 2, 1)
    return out.reshape(r, -1)


def mode_field(seed: int, m: int, dims) -> np.ndarray:
    """Mode m as an interleaved column of 3*X*Y*Z FP64 values (before QR)."""
    M = mode_coefficients(seed, m + 1)[:, m:m + 1]
    return _evaluate(M, dims)[0]


def synthetic_basis(dims, r: int, seed: int = 7, direct_qr: bool = False) -> np.ndarray:
    """Orthonormal (r, 3*X*Y*Z) FP64 basis (row c = column c of U).

    The functions cos(theta_kappa), sin(theta_kappa) with kappa in {1..6}^3
    are mutually orthogonal on the periodic peeled grid with squared norm
    X*Y*Z/2 (dims >= 13 keep 2*kappa off the Nyquist aliases), so the thin QR
    of the sampled modes equals the QR of their scaled coefficient matrix;
    signs are fixed by diag(R) > 0.  direct_qr=True runs numpy's QR on the
    sampled matrix instead (for checking this equivalence at small sizes)."""
    X, Y, Z = dims
    if direct_qr or min(dims) < 13:
        A = _evaluate(mode_coefficients(seed, r), dims).T
        Q, R = np.linalg.qr(A, mode="reduced")
        s = np.sign(np.diag(R))
        s[s == 0] = 1.0
        return np.ascontiguousarray((Q * s[None, :]).T)
    scale = np.sqrt(X * Y * Z / 2.0)
    M = mode_coefficients(seed, r) * scale
    Qc, R = np.linalg.qr(M, mode="reduced")
    s = np.sign(np.diag(R))
    s[s == 0] = 1.0
    Qc = Qc * s[None, :]
    return _evaluate(Qc / scale, dims)


def gaussian_field(dims, seed: int) -> np.ndarray:
    """Full-grid (X+2)(Y+2)(Z+2) x 3 FP64 velocity: N(0,1) inside, 0 on the boundary."""
    X, Y, Z = dims
    rng = np.random.default_rng(seed)
    v = np.zeros((Z + 2, Y + 2, X + 2, 3), np.float64)
    v[1:-1, 1:-1, 1:-1, :] = rng.standard_normal((Z, Y, X, 3))
    return v.reshape(-1)


def gaussian_q(r: int, seed: int) -> np.ndarray:
    return np.random.default_rng(seed).standard_normal(r)
