"""Pins of the oracle's point-normal estimation (the published algorithm of
PCL's NormalEstimation, which the reference calls at
mesh/PoissonReconstruction.cpp:137-186): closed forms on planes and spheres,
brute-force neighbours, and the viewpoint rule."""
import numpy as np


def test_neighbours_match_brute_force(orc):
    rng = np.random.default_rng(5)
    P = rng.random((400, 3))
    _, _, idx, dk, dk1, _ = orc.estimate_normals(P, k=10)
    D = np.linalg.norm(P[:, None] - P[None], axis=2)
    want = np.argsort(D, axis=1, kind="stable")[:, :10]
    for i in range(len(P)):
        assert set(idx[i]) == set(want[i])
        assert idx[i][0] == i  # the point itself, at distance 0
    assert np.all(dk <= dk1)


def test_plane_normal_closed_form(orc):
    rng = np.random.default_rng(6)
    xy = rng.random((2000, 2))
    P = np.c_[xy, 0.3 * xy[:, 0] + 0.1 * xy[:, 1] + 2.0]
    n, curv, *_ = orc.estimate_normals(P, k=10)
    true = np.array([-0.3, -0.1, 1.0]) / np.linalg.norm([-0.3, -0.1, 1.0])
    # the viewpoint (origin) lies below the plane: normals face down
    assert np.allclose(n, -true, atol=1e-9)
    assert np.all(np.abs(curv) < 1e-12)


def test_sphere_normals_radial_and_viewpoint(orc):
    rng = np.random.default_rng(7)
    v = rng.normal(size=(20000, 3))
    P = v / np.linalg.norm(v, axis=1, keepdims=True) + np.array([3.0, 0, 0])
    n, *_ = orc.estimate_normals(P, k=10, viewpoint=(3.0, 0, 0))
    radial = P - np.array([3.0, 0, 0])
    cos = np.sum(n * radial, axis=1)
    assert np.median(cos) < -0.999  # facing the centre, where the viewpoint is
    assert np.all(np.sum((np.array([3.0, 0, 0]) - P) * n, axis=1) >= 0)
