"""Pins of the oracle's per-vertex normals (mesh/DrawMesh.cu:80-177) by
closed forms and by a plain-loop restatement on small meshes."""
import numpy as np


def _octahedron():
    v = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float32)
    t = []
    for sx in (0, 1):
        for sy in (2, 3):
            for sz in (4, 5):
                tri = [sx, sy, sz]
                a, b, c = v[tri]
                if np.dot(np.cross(b - a, c - a), a + b + c) < 0:
                    tri = [sx, sz, sy]
                t.append(tri)
    return v, np.array(t, np.int32)


def test_single_triangle(orc):
    n = orc.mesh_normals(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([[0, 1, 2]], np.int32))
    assert np.array_equal(n, np.tile(np.array([0, 0, 1], np.float32), (3, 1)))


def test_octahedron_vertex_normals_are_the_axes(orc):
    v, t = _octahedron()
    n = orc.mesh_normals(v, t)
    assert np.allclose(n, v, atol=1e-7)


def test_isolated_vertex_degenerate_and_invalid_triangles(orc):
    v = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0], [0, 1, 0], [5, 5, 5]], np.float32)
    t = np.array([[0, 1, 2], [0, 1, 3], [0, 1, 7], [-1, 0, 1]], np.int32)
    n = orc.mesh_normals(v, t)
    # vertex 0 and 1: one zero (collinear) face and one +z face, averaged then normalised
    assert np.array_equal(n[0], np.array([0, 0, 1], np.float32)) and np.array_equal(n[1], n[0])
    assert np.array_equal(n[2], np.zeros(3, np.float32))  # only the collinear face
    assert np.array_equal(n[3], np.array([0, 0, 1], np.float32))
    assert np.array_equal(n[4], np.zeros(3, np.float32))  # touched by no triangle


def test_matches_plain_loops(orc):
    rng = np.random.default_rng(3)
    v = rng.random((40, 3), dtype=np.float32)
    t = rng.integers(0, 40, size=(90, 3), dtype=np.int32)
    f32 = np.float32
    fn = []
    for a, b, c in t:
        o, p, q = v[a], v[b], v[c]
        u = [f32(p[k] - o[k]) for k in range(3)]
        w = [f32(q[k] - o[k]) for k in range(3)]
        cr = [f32(u[1] * w[2] - u[2] * w[1]), f32(u[2] * w[0] - u[0] * w[2]), f32(u[0] * w[1] - u[1] * w[0])]
        ln = np.sqrt(f32(f32(cr[0] * cr[0] + cr[1] * cr[1]) + cr[2] * cr[2]))
        fn.append([f32(x / ln) for x in cr] if ln != 0 else [f32(0)] * 3)
    want = np.zeros((40, 3), np.float32)
    for i in range(40):
        s, cnt = [f32(0)] * 3, 0
        for ti, tri in enumerate(t):
            for c in tri:
                if c == i:
                    s = [f32(s[k] + fn[ti][k]) for k in range(3)]
                    cnt += 1
        if cnt:
            a = [f32(s[k] / f32(cnt)) for k in range(3)]
            ln = np.sqrt(f32(f32(a[0] * a[0] + a[1] * a[1]) + a[2] * a[2]))
            if ln != 0:
                want[i] = [f32(x / ln) for x in a]
    assert np.array_equal(orc.mesh_normals(v, t), want)
