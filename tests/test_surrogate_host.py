"""Host side of the surrogate: the network's input features (CPU only)."""
import math

import numpy as np

from mosaic_amd import surrogate as SG


def test_features_ranges_and_unit_vectors():
    rng = np.random.default_rng(0)
    P = SG.sample_params(500, rng)
    X = SG.features(P)
    assert X.shape == (500, SG.N_FEAT) and X.dtype == np.float32
    assert np.all(np.abs(X) <= 1.0 + 1e-6)
    for c in (2, 5):
        assert np.allclose(np.linalg.norm(X[:, c:c + 3], axis=1), 1.0, atol=1e-6)
    # the Burgers vector of the defect block is along the same unit vector
    for p, x in zip(P[:5], X[:5]):
        b = SG._unit(p[2], p[3])
        assert np.allclose(b, x[2:5], atol=1e-6)
    # the angle, strain and fault shift are the normalised parameters
    assert np.allclose(X[:, 8:11], SG.normalise(P)[:, 6:9], atol=1e-6)


def test_features_continuous_across_azimuth_wrap():
    p = SG.sample_params(1, np.random.default_rng(1))[0]
    a, b = p.copy(), p.copy()
    a[3], b[3] = 1e-9, 2 * math.pi - 1e-9
    a[5], b[5] = 2 * math.pi - 1e-9, 1e-9
    xa, xb = SG.features(a), SG.features(b)
    assert np.max(np.abs(xa - xb)) < 1e-6


def test_anchor_weights_interpolate_linearly():
    angles = np.linspace(-1.0, 1.0, 9)
    W = SG.anchor_weights(angles, angles[[0, 8]])
    assert np.allclose(W.sum(axis=1), 1.0)
    assert np.allclose(W @ np.array([2.0, 4.0]), 3.0 + angles)
    W3 = SG.anchor_weights(angles, angles[[0, 4, 8]])
    vals = np.array([1.0, -1.0, 5.0])
    got = W3 @ vals
    assert np.allclose(got[[0, 4, 8]], vals)
    assert np.isclose(got[2], 0.0) and np.isclose(got[6], 2.0)
    # beyond the outer anchors the end value holds
    Wout = SG.anchor_weights([-2.0, 2.0], [-1.0, 1.0])
    assert np.allclose(Wout, [[1, 0], [0, 1]])
