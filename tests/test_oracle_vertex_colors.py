"""Pins of the oracle's vertex colours (mesh/DrawMesh.cu:179-211): the 4
nearest points weighted by squared distance over the four's sum."""
import numpy as np


def _pts():
    P = np.array([[1, 0, 0], [-1, 0, 0], [0, 2, 0], [0, -2, 0], [0, 0, 50]], np.float32)
    C = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1], [9, 9, 9]], np.float32)
    return P, C


def test_distance_proportional_weights(orc):
    P, C = _pts()
    out, d4, d5 = orc.vertex_colors(np.zeros((1, 3), np.float32), P, C)
    # squared distances 1, 1, 4, 4 over 10: weights 0.1, 0.1, 0.4, 0.4
    want = 0.1 * C[0] + 0.1 * C[1] + 0.4 * C[2] + 0.4 * C[3]
    assert np.allclose(out[0], want, atol=1e-6) and d4[0] == 2.0 and d5[0] == 50.0


def test_equal_distances_average(orc):
    P = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 7]], np.float32)
    _, C = _pts()
    out, *_ = orc.vertex_colors(np.zeros((1, 3), np.float32), P, C)
    assert np.allclose(out[0], C[:4].mean(axis=0), atol=1e-6)


def test_coincident_points_equal_weights(orc):
    P = np.zeros((4, 3), np.float32)
    C = np.eye(4, 3, dtype=np.float32)
    out, *_ = orc.vertex_colors(np.zeros((1, 3), np.float32), P, C)
    assert np.allclose(out[0], C.mean(axis=0), atol=1e-7)
