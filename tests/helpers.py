"""Test helpers: build reference-format component files from chosen integers."""
from __future__ import annotations

import struct

import numpy as np

from oracle import oracle as O


def padded(n: int) -> int:
    return n + (8 - n % 8) % 8


def num_blocks(dims) -> int:
    X, Y, Z = dims
    return (padded(X) // 8) * (padded(Y) // 8) * (padded(Z) // 8)


def component_bytes(dims, n_bits, s, gamma, segments) -> bytes:
    """File image (format v1) from per-segment data.

    s, gamma: (nB, r) arrays; segments[c][b]: int32 token array of (b, c).
    Tokens are stored column-major as the reference writes them.
    """
    X, Y, Z = dims
    nB, r = s.shape
    lens = np.array([[len(segments[c][b]) for c in range(r)] for b in range(nB)], np.int32)  # (nB, r)
    flat = lens.T.reshape(-1)  # column-major
    idx = O.modified_cumsum(flat)
    head = struct.pack("<10i", 1, n_bits, X, Y, Z, padded(X), padded(Y), padded(Z), r, nB)
    toks = np.concatenate([np.asarray(segments[c][b], np.int32) for c in range(r) for b in range(nB)])
    return (head + np.asarray(s, np.float64).T.tobytes() + np.asarray(gamma, np.float64).T.tobytes()
            + flat.astype(np.int32).tobytes() + idx.astype(np.int32).tobytes() + toks.tobytes())


def basis_from_ints(dims, r, zz_of, n_bits=32, seed=0, max_gamma=8.0):
    """Three component images whose segment (k, b, c) encodes zz_of(k, b, c)
    (a zigzag-ordered int32[512]); s and gamma are drawn from `seed`.
    gamma <= max_gamma keeps int * 22^gamma / s inside the FP32 range, as it
    is for every stream the encoder writes (nonzero ints need damp^gamma < 2^31)."""
    rng = np.random.default_rng(seed)
    nB = num_blocks(dims)
    out = []
    for k in range(3):
        s = rng.uniform(1e8, 1e10, size=(nB, r))
        gamma = rng.integers(0, int(4 * max_gamma) + 1, size=(nB, r)) / 4.0
        segs = [[O.rle_encode(zz_of(k, b, c)) for b in range(nB)] for c in range(r)]
        out.append(component_bytes(dims, n_bits, s, gamma, segs))
    return out


def write_files(bufs, prefix):
    paths = []
    for k, b in enumerate(bufs):
        p = f"{prefix}{k}"
        with open(p, "wb") as f:
            f.write(b)
        paths.append(p)
    return paths
