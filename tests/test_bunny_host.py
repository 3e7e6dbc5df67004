"""The reference's own input cloud (PointCloudData/data_ply_normal/bunny/
bunny.ply, 35,947 points with normals) on the CPU: the tests-side ASCII PLY
reader, the committed fixture, and the oracle's reconstruction of it at
depth 7 checked by invariants."""
import os

import numpy as np
import pytest
from scipy.spatial import cKDTree

from ply_ascii import read_ascii_ply

HERE = os.path.dirname(os.path.abspath(__file__))
REF_BUNNY = "/root/reference/PointCloudData/data_ply_normal/bunny/bunny.ply"


def test_reader_on_a_pcl_style_file(tmp_path):
    p = tmp_path / "t.ply"
    p.write_text("ply\nformat ascii 1.0\ncomment PCL generated\nelement vertex 2\nproperty float x\nproperty float y\n"
                 "property float z\nproperty float nx\nproperty float ny\nproperty float nz\nproperty float curvature\n"
                 "element face 0\nproperty list uchar int vertex_indices\nelement camera 1\nproperty float view_px\n"
                 "end_header\n0.5 -1 2.25 0 0 1 0.01\n1e-3 3 4 1 0 0 0\n0 0 0\n")
    d = read_ascii_ply(p)
    np.testing.assert_array_equal(d["x"], np.float32([0.5, 1e-3]))
    np.testing.assert_array_equal(d["z"], np.float32([2.25, 4]))
    np.testing.assert_array_equal(d["nz"], np.float32([1, 0]))
    assert set(d) == {"x", "y", "z", "nx", "ny", "nz", "curvature"}


def test_reader_rejects_binary(tmp_path):
    p = tmp_path / "b.ply"
    p.write_bytes(b"ply\nformat binary_little_endian 1.0\nend_header\n")
    with pytest.raises(ValueError):
        read_ascii_ply(p)


@pytest.mark.skipif(not os.path.exists(REF_BUNNY), reason="the reference tree is not present")
def test_fixture_equals_the_reference_file():
    d = read_ascii_ply(REF_BUNNY)
    fx = np.load(os.path.join(HERE, "golden", "bunny_points.npz"))
    np.testing.assert_array_equal(fx["xyz"], np.stack([d["x"], d["y"], d["z"]], 1))
    np.testing.assert_array_equal(fx["normals"], np.stack([d["nx"], d["ny"], d["nz"]], 1))


def bunny():
    fx = np.load(os.path.join(HERE, "golden", "bunny_points.npz"))
    return fx["xyz"], fx["normals"]


def mesh_invariants(v, t, corner, edge, xyz, depth):
    """Checks shared with the GPU test: a non-empty mesh with valid indices,
    in the normalised frame; half its vertices within one depth-D cell of an
    input point (the scan is open underneath and the implicit surface closes
    it away from the points), and 99% of the input points within two cells of
    the mesh."""
    assert len(t) > 1000 and t.min() >= 0 and t.max() < len(v)
    assert np.all(v >= 0.0) and np.all(v <= 1.0)
    world = corner + v.astype(np.float64) * edge
    cell = edge / (1 << depth)
    d_v = cKDTree(xyz).query(world)[0]
    d_p = cKDTree(world).query(xyz)[0]
    assert np.median(d_v) <= cell, np.median(d_v) / cell
    assert np.percentile(d_p, 99) <= 2 * cell, np.percentile(d_p, 99) / cell
    return d_v, d_p


def test_oracle_reconstructs_the_bunny(orc):
    xyz, nrm = bunny()
    r = orc.Result(xyz, nrm, 7, precision=0)
    v = r.array("mesh_v").reshape(-1, 3)
    t = r.array("mesh_t").reshape(-1, 3)
    corner = np.array([r.scalar("corner%d" % a) for a in range(3)])
    mesh_invariants(v, t, corner, r.scalar("edge"), xyz, 7)
    assert all(k < 10000 for k in r.array("cg_iters"))
