"""Writes dct_scipy.npz: scipy.fft's orthonormal 3-D DCT-II / DCT-III of a few
8x8x8 blocks, the golden vectors for the oracle's block DCT.

scipy's dctn(type=2, norm='ortho') is the orthonormal DCT-II that the
reference's UnitaryBlockDCT computes with FFTW REDFT10 plus its normalisation
(COMPRESSION.cpp:173-201), and idctn(type=2, norm='ortho') the inverse
(REDFT01 with UndoNormalize, :207-272).  The unnormalised pair (norm=None, FFTW's
convention) is stored too, for the 8*N round-trip factor the reference's
fftw_dct_roundtrip test checks.

Blocks are flattened x fastest (index x + 8y + 64z), so a block reshaped
C-order is indexed [z, y, x]; dctn over all axes does not care.

Run from the repository root:  python tests/golden/make_dct_golden.py
"""
import os

import numpy as np
import scipy
import scipy.fft

HERE = os.path.dirname(os.path.abspath(__file__))


def inputs():
    i = np.arange(512, dtype=np.float64)
    blocks = [np.sin((b + 1) * i * 0.03) * (b + 1) for b in range(4)]  # the reference's dct_roundtrip blocks
    blocks.append(np.sin(0.1 * i) + 0.5 * np.cos(0.03 * i))            # its fftw_dct_roundtrip pattern
    z, y, x = np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")
    blocks.append((np.sin(x * 0.5) + np.cos(y * 0.3) + np.sin(z * 0.2)).reshape(-1))  # energy compaction block
    rng = np.random.default_rng(20240611)
    blocks.append(rng.standard_normal(512))
    e = np.zeros(512)
    e[1 + 8 * 2 + 64 * 3] = 1.0  # a single basis function
    blocks.append(e)
    return np.stack(blocks)


def main():
    x = inputs()
    cube = x.reshape(-1, 8, 8, 8)
    fwd = scipy.fft.dctn(cube, type=2, norm="ortho", axes=(1, 2, 3)).reshape(x.shape)
    inv = scipy.fft.idctn(cube, type=2, norm="ortho", axes=(1, 2, 3)).reshape(x.shape)
    raw_fwd = scipy.fft.dctn(cube, type=2, norm=None, axes=(1, 2, 3))
    raw_pair = scipy.fft.dctn(raw_fwd, type=3, norm=None, axes=(1, 2, 3)).reshape(x.shape)
    out = os.path.join(HERE, "dct_scipy.npz")
    np.savez_compressed(out, x=x, forward=fwd, inverse=inv, unnormalised_pair=raw_pair,
                        scipy_version=np.array(scipy.__version__))
    print(out)


if __name__ == "__main__":
    main()
