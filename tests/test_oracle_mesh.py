"""Pins the oracle's mesh stages: MC table self-consistency
(mesh/ComputeTriangleIndices.cu:41-315), vertex / edge / face ownership counts
against brute-force distinct positions (BuildMeshGeometry.cu:42-400), vertex
values at depth D against a global sum, and the final mesh against the
analytic surface.
"""
import json
import os

import numpy as np
import pytest

from conftest import bspline_np, func_cw

HERE = os.path.dirname(os.path.abspath(__file__))
# Bourke corner order: corner id -> (x, y, z) offset bits
BOURKE = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
EDGE_VERTEX = [(0, 1), (2, 3), (4, 5), (6, 7), (0, 3), (1, 2), (4, 7), (5, 6), (0, 4), (1, 5), (3, 7), (2, 6)]


def test_mc_table_consistency(orc):
    tab = json.load(open(os.path.join(HERE, "golden", "reference_tables.json")))
    counts = tab["trianglesCount"]
    for cube in range(256):
        w = orc.mc_word(cube)
        edges = []
        for k in range(15):
            e = (w >> (4 * k)) & 0xF
            if e == 0xF:
                break
            edges.append(e)
        assert len(edges) % 3 == 0 and len(edges) // 3 == counts[cube]
        inside = [(cube >> v) & 1 for v in range(8)]
        crossing = {e for e, (a, b) in enumerate(EDGE_VERTEX) if inside[a] != inside[b]}
        assert set(edges) == crossing, cube


def _int_pos(p, depth):
    return np.round(np.asarray(p, dtype=np.float64) * (1 << (depth + 1))).astype(np.int64)


def test_ownership_counts(orc, pkg):
    xyz, nrm = pkg.synth.sphere(n=3000)
    r = orc.Result(xyz, nrm, 5, precision=1)
    D = 5
    key, depth = r.array("node_key"), r.array("depth")
    center = r.array("center").reshape(-1, 3)
    verts, edges, faces = r.array("verts"), r.array("edges"), r.array("faces")
    corners = set()
    for o in range(len(key)):
        d = depth[o]
        h = 1.0 / (1 << (d + 1))
        for bx, by, bz in BOURKE:
            p = center[o] + np.array([2 * bx - 1, 2 * by - 1, 2 * bz - 1]) * h
            corners.add((d,) + tuple(_int_pos(p, D)))
    assert len(verts) == len(corners)
    vpos = verts[:, :3].copy().view(np.float32)
    got = {(int(vd),) + tuple(_int_pos(vp, D)) for vp, vd in zip(vpos, verts[:, 5])}
    assert got == corners
    # depth-D edges: distinct centres
    base = r.base
    ec = set()
    h = 1.0 / (1 << (D + 1))
    for o in range(base[D], base[D + 1]):
        for kind in range(12):
            orient, o0, o1 = kind >> 2, kind & 1, (kind >> 1) & 1
            m, dim = [0, 0, 0], 0
            for a in range(3):
                if a != orient:
                    m[a] = 2 * (o0 if dim == 0 else o1) - 1
                    dim += 1
            ec.add(tuple(_int_pos(center[o] + np.array(m) * h, D)))
    assert len(edges) == len(ec)
    fc = set()
    for o in range(len(key)):
        d = depth[o]
        hh = 1.0 / (1 << (d + 1))
        for kind in range(6):
            m = np.zeros(3)
            m[kind >> 1] = 2 * (kind & 1) - 1
            fc.add((d,) + tuple(_int_pos(center[o] + m * hh, D)))
    assert len(faces) == len(fc)


def test_depth_d_vertex_values_brute_force(small_sphere_ref):
    r = small_sphere_ref
    D = r.depth
    verts = r.array("verts")
    vv = r.array("vvalue")
    dx = r.array("dx")
    enc = r.array("encidx")
    s1 = 1 << (D + 1)
    f = np.stack([enc % s1, (enc // s1) % s1, enc // (s1 * s1)], axis=1)
    cw = [func_cw(f[:, a]) for a in range(3)]
    iso = r.scalar("iso")
    pos = verts[:, :3].copy().view(np.float32).astype(np.float64)
    deep = np.nonzero((verts[:, 5] == D) & (verts[:, 3] > 0))[0]
    for v in deep[::53]:
        w = np.ones(len(dx))
        for a in range(3):
            w *= bspline_np((pos[v, a] - cw[a][0]) / cw[a][1])
        assert abs(vv[v] - (w @ dx - iso)) <= 1e-9 * np.abs(vv).max()


@pytest.mark.parametrize("shape", ["sphere", "torus"])
def test_mesh_near_analytic_surface(orc, pkg, shape):
    if shape == "sphere":
        xyz, nrm = pkg.synth.sphere(n=20000)
        D = 6
    else:
        xyz, nrm = pkg.synth.torus(n=60000)
        D = 6
    r = orc.Result(xyz, nrm, D, precision=1)
    v = r.array("mesh_v").reshape(-1, 3).astype(np.float64)
    t = r.array("mesh_t").reshape(-1, 3)
    assert len(t) > 1000 and t.min() >= 0 and t.max() < len(v)
    corner = np.array([r.scalar("corner%d" % a) for a in range(3)])
    p = v * r.scalar("edge") + corner  # back to the input frame
    c = p - 0.5
    if shape == "sphere":
        dist = np.abs(np.linalg.norm(c, axis=1) - 0.4)
    else:
        rho = np.hypot(c[:, 0], c[:, 1])
        dist = np.abs(np.hypot(rho - 0.3, c[:, 2]) - 0.1)
    voxel = r.scalar("edge") / (1 << D)
    assert dist.mean() < 0.25 * voxel and dist.max() < 1.0 * voxel
