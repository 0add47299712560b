"""Host-side assembly of the predictor's training data."""
import numpy as np

from nedyson_amd import predictor


def test_dataset_rows_and_scaling():
    rng = np.random.default_rng(0)
    R = [rng.normal(size=(9, 128)).astype(np.float32) * (t + 1) for t in range(10)]
    R[6][:, 64:] *= 1e-6  # the groups keep separate scales
    X, Y, s = predictor.dataset(R)
    assert X.shape == (9 * 4, 768) and Y.shape == (9 * 4, 128) and s.shape == (9 * 4, 128)
    # newest step first; every (component, zone) group of 32 at unit RMS over the six steps
    assert np.allclose(X[0, :128] * s[0], R[5][0], rtol=1e-6)
    assert np.allclose(X[0, 128:256] * s[0], R[4][0], rtol=1e-6)
    assert np.allclose(Y[0] * s[0], R[6][0], rtol=1e-6)
    g = X.reshape(36, 6, 4, 32).astype(np.float64)
    assert np.allclose(np.sqrt(np.mean(g ** 2, axis=(1, 3))), 1.0, rtol=1e-5)


def test_dataset_too_short_is_empty():
    X, Y, s = predictor.dataset([np.zeros((4, 128), np.float32)] * 6)
    assert X.shape == (0, 768) and Y.shape == (0, 128)
