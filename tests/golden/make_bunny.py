"""Writes tests/golden/bunny_points.npz from the reference's own input cloud
(PointCloudData/data_ply_normal/bunny/bunny.ply: 35,947 points with normals,
ASCII PLY written by PCL): float32 xyz and normals, as data only.  Run from
the repository root where the reference tree is present:
    python tests/golden/make_bunny.py /root/reference"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), ".."))
from ply_ascii import read_ascii_ply  # noqa: E402

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
p = read_ascii_ply(os.path.join(ref, "PointCloudData", "data_ply_normal", "bunny", "bunny.ply"))
xyz = np.stack([p["x"], p["y"], p["z"]], 1)
nrm = np.stack([p["nx"], p["ny"], p["nz"]], 1)
np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "bunny_points.npz"), xyz=xyz, normals=nrm)
print(xyz.shape, xyz.min(0), xyz.max(0))
