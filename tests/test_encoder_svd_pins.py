"""Known answers for the encoder's gamma search and for the SVD coordinate
transform, computed by hand (CPU, no GPU).

Encoder (TuneGammaQuantized, reference COMPRESSION.cpp:841-916, with
PreprocessBlock :733-759 and EncodeBlock :924-949): the block F has DC 1.0
and 0.52 at u=1 (damping base 1+u+v+w = 2), every other coefficient 0;
nBits = 4, so s = (2^3 - 1) / 1.0 = 7 and the scaled block is (7, 3.64).
The search starts at gamma = 0 and, while |percent - pe| > 1e-6 and fewer
than maxIterations steps were taken, moves `upper` (pe < percent) or `lower`
to gamma and bisects.  pe = 1 - |E - sum (d^g trunc(F / d^g))^2| / E with
E = 49 + 3.64^2 = 62.2496.  By hand, at percent 0.9:

  g = 0     trunc(3.64) = 3          pe = 1 - 4.2496/62.2496 = 0.9317  -> lower = 0,   g = 2
  g = 2     trunc(3.64/4) = 0        pe = 0.7872                       -> upper = 2,   g = 1
  g = 1     trunc(1.82) = 1          pe = 0.8514                       -> upper = 1,   g = 0.5
  g = 0.5   trunc(2.574) = 2         pe = 0.9157                       -> lower = 0.5, g = 0.75
  g = 0.75  trunc(2.164) = 2         pe = 0.9689                       -> lower = 0.75, g = 0.875
  g = 0.875 trunc(1.985) = 1         pe = 0.8412                       -> upper,       g = 0.8125
  ... converging to log2(3.64/2) = 0.864 from below: rint(4 g)/4 = 0.75.

The final ints are rint(F / 2^gamma): 7 and rint(3.64 / 2^0.75) = rint(2.164) = 2.
With rint in place of trunc inside the search the bisection converges to
2.142 instead (gamma 2.25, ints 7 and 1), so a trunc -> rint slip fails.
Cut off after 0..6 steps, gamma is 0, 2, 1, 0.5, 0.75, rint(3.5)/4 = 1 (ties
to even), rint(3.25)/4 = 0.75.  With percent equal to pe(0) the loop never
runs (gamma 0, ints 7 and rint(3.64) = 4); with percent 2e-6 below it the
search runs and converges to log2(3.64/3) = 0.279, gamma 0.25, ints 7 and
rint(3.64 / 2^0.25) = rint(3.061) = 3.

The zigzag order puts u=1 at position 1, so the tokens are [7, q1, 0, 0, 510].
"""
import numpy as np
import pytest

from cfs_plume_amd import encode_block as product_encode_block
from oracle import oracle as O


def oracle_encode_block(F, n_bits, percent, max_iterations):
    s, g, tok = O.encode_block(F, n_bits, percent, max_iterations)
    return s, g, np.asarray(tok)


ENCODERS = {"oracle": oracle_encode_block, "product": product_encode_block}


def block(*vals):
    F = np.zeros(512)
    for i, v in enumerate(vals):
        F[i] = v
    return F


@pytest.mark.parametrize("which", sorted(ENCODERS))
@pytest.mark.parametrize("max_iterations,gamma,q1", [
    (0, 0.0, 4), (1, 2.0, 1), (2, 1.0, 2), (3, 0.5, 3), (4, 0.75, 2), (5, 1.0, 2), (6, 0.75, 2), (50, 0.75, 2),
])
def test_gamma_search_known_answers(which, max_iterations, gamma, q1):
    s, g, tok = ENCODERS[which](block(1.0, 0.52), 4, 0.9, max_iterations)
    assert s == 7.0
    assert g == gamma
    assert list(tok) == [7, q1, 0, 0, 510]


@pytest.mark.parametrize("which", sorted(ENCODERS))
def test_gamma_search_stopping_tolerance(which):
    """|percent - pe| <= 1e-6 stops the search before its first step."""
    pe0 = 1.0 - (3.64 ** 2 - 9.0) / (49.0 + 3.64 ** 2)
    _, g, tok = ENCODERS[which](block(1.0, 0.52), 4, pe0, 50)
    assert g == 0.0 and list(tok) == [7, 4, 0, 0, 510]
    _, g, tok = ENCODERS[which](block(1.0, 0.52), 4, pe0 - 2e-6, 50)
    assert g == 0.25 and list(tok) == [7, 3, 0, 0, 510]


@pytest.mark.parametrize("which", sorted(ENCODERS))
def test_final_quantisation_rounds_ties_to_even(which):
    """RoundFieldToInt (:22-30) under FE_TONEAREST: with nBits 5 and absMax
    15/16 the scale is exactly 16; at gamma 0 (no search step) the scaled
    values 2.5, 3.5, -2.5, 0.5 round to 2, 4, -2, 0.  u = 1, 2, 3, 4 lie at
    zigzag positions 1, 4, 10 and 20."""
    s, g, tok = ENCODERS[which](block(0.9375, 0.15625, 0.21875, -0.15625, 0.03125), 5, 0.9, 0)
    assert s == 16.0 and g == 0.0
    assert list(tok) == [15, 2, 0, 0, 2, 4, 0, 0, 5, -2, 0, 0, 501]


def test_encoder_block_entry_is_the_matrix_encoder():
    """cfs_encode_block is the per-block step of cfs_encode_matrix: on each
    block's forward DCT it gives the ints the encoded basis holds."""
    from cfs_plume_amd import encode_matrix, synth
    dims = (8, 8, 8)
    U = synth.synthetic_basis(dims, 2, seed=4)
    bufs = encode_matrix(U, dims).buffers()
    ob = O.Basis(buffers=bufs)
    for col in range(2):
        for k in range(3):
            f = U[col].reshape(-1, 3)[:, k]                  # scalar field x-fastest, one 8^3 block
            F = O.dct3(O.get_blocks(f, dims, False), 1)[0]   # UnitaryBlockDCT(1), u + 8v + 64w order
            s, g, tok = product_encode_block(F, 32, 0.9995, 50)
            assert np.array_equal(O.rle_decode(tok)[0], O.zigzag_flatten(ob.decode_ints(k, 0, col)))
    ob.close()


# --- SVD coordinate transform (COMPRESSION.cpp:566-727) ---------------------------------

def test_svd_transform_known_rotation():
    """V = rotation by +90 degrees about z (columns: e_y, -e_x, e_z).
    TransformVectorFieldSVDCached multiplies each cell by V^T
    (SparseBlockDiagonal(v->transpose()) :623-643): V^T (1, 0, 0) = (0, -1, 0)
    and V^T (0, 1, 0) = (1, 0, 0).  UntransformVectorFieldSVD multiplies by V
    (:713-727): V (0, -1, 0) = (1, 0, 0)."""
    V = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    u = np.array([1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 2.0])
    assert np.array_equal(O.transform_svd(V, u), [0.0, -1.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 2.0])
    assert np.array_equal(O.untransform_svd(V, [0.0, -1.0, 0.0]), [1.0, 0.0, 0.0])
    assert np.array_equal(O.untransform_svd(V, O.transform_svd(V, u)), u)


def test_svd_transform_decorrelates_the_components():
    """TransformVectorFieldSVD (:623-643) with V from the SVD of the N x 3
    matrix of the field (BuildXYZMatrix :604-617): the transformed field's
    component columns are X V = U S, orthogonal with norms equal to the
    singular values in decreasing order; the round trip is < 1e-10
    (the reference's test, tests/test_compression_data.cpp:1662-1706)."""
    n = 8 ** 3
    i = np.arange(n)
    field = np.stack([np.sin(i * 0.1), np.cos(i * 0.2), np.sin(i * 0.3 + 1.0)], axis=1)
    field = field @ np.array([[1.0, 0.3, -0.2], [0.0, 1.0, 0.5], [0.0, 0.0, 1.0]])  # correlate them
    _, S, Vt = np.linalg.svd(field, full_matrices=False)
    V = Vt.T
    t = O.transform_svd(V, field.reshape(-1)).reshape(n, 3)
    G = t.T @ t
    assert np.allclose(G, np.diag(S ** 2), rtol=0, atol=1e-9 * S[0] ** 2)
    assert np.all(np.diff(np.diag(G)) < 0)
    # the wrong direction (V instead of V^T) does not decorrelate a correlated field
    w = O.untransform_svd(V, field.reshape(-1)).reshape(n, 3)
    assert np.max(np.abs((w.T @ w) - np.diag(np.diag(w.T @ w)))) > 1e-3 * S[0] ** 2
    back = O.untransform_svd(V, t.reshape(-1))
    assert np.max(np.abs(back - field.reshape(-1))) < 1e-10
