"""Independent pins of the oracle's mesh stages (CPU).

Each test recomputes a result of the oracle by a different route and
requires agreement:
  * the marching-cubes triangulation and the quirk tables (childrenVertexKind,
    edgeVertex, parentFaceKind) against the values the reference declares
    (tests/golden/reference_tables.json), both in the oracle and in the HIP
    sources;
  * vertex / edge / face ownership (mesh/BuildMeshGeometry.cu:42-400) by brute
    force over cell coordinates: the owner of a corner, edge or face is the
    existing node of the same depth with the smallest key among those sharing
    it;
  * vertex implicit values at every depth, including the coarse depths'
    descent through key-ordered children with the Bourke map (quirk Q8,
    mesh/ComputeTriangleIndices.cu:346-399), as global sums over dense
    per-depth grids of d_x with the descent chain built by key arithmetic;
  * the adaptive subdivision's sets (ComputeTriangleIndices.cu:1475-2194):
    vertex and edge owners as the smallest Morton key among the virtual
    depth-D cells of the set sharing the corner / edge, and vertex values as
    the same dense-grid sums over the real nodes only (virtual nodes carry
    d_x = 0, ComputeTriangleIndices.cu:978, 1021).
"""
import json
import os
import re

import numpy as np
import pytest

from conftest import bspline_np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "poisson-surface-reconstruction-gpu_amd", "csrc")


@pytest.fixture(scope="module")
def ref_tables():
    with open(os.path.join(HERE, "golden", "reference_tables.json")) as f:
        return json.load(f)


def _decode_words(words):
    rows = []
    for w in words:
        edges = []
        for k in range(15):  # at most 5 triangles: 15 edge ids, 0xF ends a shorter list
            e = (w >> (4 * k)) & 0xF
            if e == 0xF:
                break
            edges.append(int(e))
        rows.append(edges)
    return rows


def _reference_rows(ref):
    tri = np.array(ref["triangles"]).reshape(256, 16)
    return [[int(e) for e in row if e != -1] for row in tri]


def test_reference_triangle_table_is_consistent(ref_tables):
    """The fixture itself: every row ends at the first -1 and holds
    3 x trianglesCount edges."""
    tri = np.array(ref_tables["triangles"]).reshape(256, 16)
    for c in range(256):
        row = tri[c]
        n = int((row != -1).sum())
        assert np.all(row[:n] != -1) and np.all(row[n:] == -1)
        assert n == 3 * ref_tables["trianglesCount"][c]


def test_oracle_mc_table_equals_reference(orc, ref_tables):
    got = _decode_words([orc.mc_word(c) for c in range(256)])
    assert got == _reference_rows(ref_tables)


def test_hip_mc_table_equals_reference(ref_tables):
    src = open(os.path.join(CSRC, "psr_mc_table_body.inc")).read()
    words = [int(w, 16) for w in re.findall(r"0x([0-9a-fA-F]+)ULL", src)]
    assert len(words) == 256
    assert _decode_words(words) == _reference_rows(ref_tables)


def _c_array(src, name, count):
    m = re.search(name + r"[^=]*=\s*\{(.*?)\};", src, re.S)
    nums = [int(x) for x in re.findall(r"-?\d+", m.group(1))]
    assert len(nums) == count, (name, len(nums))
    return nums


def test_quirk_tables_equal_reference(orc, ref_tables):
    L = orc.lib()
    L.orc_small_table.restype = __import__("ctypes").c_int
    L.orc_small_table.argtypes = [__import__("ctypes").c_int, __import__("ctypes").c_int]
    oracle = {"childrenVertexKind": [L.orc_small_table(0, i) for i in range(8)],
              "edgeVertex": [L.orc_small_table(1, i) for i in range(24)],
              "parentFaceKind": [L.orc_small_table(2, i) for i in range(48)]}
    src = open(os.path.join(CSRC, "psr_mesh_common.h")).read()
    hip = {"childrenVertexKind": _c_array(src, r"kChildrenVertexKind\[8\]", 8),
           "edgeVertex": _c_array(src, r"kEdgeVertex\[12\]\[2\]", 24),
           "parentFaceKind": _c_array(src, r"kParentFaceKind\[8\]\[6\]", 48)}
    for k in oracle:
        assert oracle[k] == ref_tables[k], k
        assert hip[k] == ref_tables[k], k
    # faceEdges is derived, not declared as data: face i (axis i >> 1, side
    # i & 1) holds the edges of the other two axes lying in that plane
    face_edges = [L.orc_small_table(3, i) for i in range(24)]
    for f in range(6):
        axis, side = f >> 1, f & 1
        want = []
        for e in range(12):
            ea = e >> 2
            if ea == axis:
                continue
            others = [a for a in range(3) if a != ea]
            off = [e & 1, (e >> 1) & 1]
            if off[others.index(axis)] == side:
                want.append(e)
        assert sorted(face_edges[4 * f:4 * f + 4]) == want, f


# ---------------------------------------------------------------------------
# helpers: keys, cells and dense grids
# ---------------------------------------------------------------------------

def key_xyz(keys, D):
    """Depth-D cell coordinates of the lower corner of each key (level l at
    bit 3(D - l): x + 2, y + 1, z + 0)."""
    k = np.asarray(keys, np.int64)
    out = np.zeros((k.size, 3), np.int64)
    for lvl in range(1, D + 1):
        b = 3 * (D - lvl)
        for a in range(3):
            out[:, a] |= ((k >> (b + 2 - a)) & 1) << (D - lvl)
    return out


def xyz_key(c, D):
    c = np.asarray(c, np.int64)
    k = np.zeros(len(c), np.int64)
    for lvl in range(1, D + 1):
        b = 3 * (D - lvl)
        for a in range(3):
            k |= ((c[:, a] >> (D - lvl)) & 1) << (b + 2 - a)
    return k


class Tree:
    """Node cells per depth from the oracle's exported arrays."""

    def __init__(self, ref, D):
        self.D = D
        self.base = ref.base
        self.key = ref.array("node_key").astype(np.int64)
        self.pnum = ref.array("node_pnum")
        self.M = len(self.key)
        self.depth = np.zeros(self.M, np.int64)
        for d in range(D + 1):
            self.depth[self.base[d]:self.base[d + 1]] = d
        full = key_xyz(self.key, D)
        self.cell = full >> (D - self.depth)[:, None]
        # dense index grid per depth: node at each cell, -1 where none
        self.grid = []
        for d in range(D + 1):
            g = np.full((1 << d,) * 3, -1, np.int64)
            sl = slice(self.base[d], self.base[d + 1])
            c = self.cell[sl]
            g[c[:, 0], c[:, 1], c[:, 2]] = np.arange(self.base[d], self.base[d + 1])
            self.grid.append(g)

    def lookup(self, d, c):
        """Node index at depth-d cells c (n, 3), -1 outside or where none."""
        n = 1 << d
        ok = np.all((c >= 0) & (c < n), axis=1)
        out = np.full(len(c), -1, np.int64)
        cc = c[ok]
        out[ok] = self.grid[d][cc[:, 0], cc[:, 1], cc[:, 2]]
        return out

    def dx_grids(self, dx):
        out = []
        for d in range(self.D + 1):
            g = np.zeros((1 << d,) * 3, np.float64)
            sl = slice(self.base[d], self.base[d + 1])
            c = self.cell[sl]
            g[c[:, 0], c[:, 1], c[:, 2]] = dx[sl]
            out.append(g)
        return out


CORNERS = np.array([[i & 1, (i >> 1) & 1, (i >> 2) & 1] for i in range(8)], np.int64)
OFFS27 = np.array([[a, b, c] for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)], np.int64)


def chain_sum(grids, d, cells, pos):
    """Sum over the 27 cells around `cells` (depth d) of d_x times the node
    function at pos (FP64)."""
    n = 1 << d
    w = 1.0 / n
    acc = np.zeros(len(cells))
    for off in OFFS27:
        c = cells + off
        ok = np.all((c >= 0) & (c < n), axis=1)
        if not ok.any():
            continue
        v = np.zeros(len(cells))
        cc = c[ok]
        v[ok] = grids[d][cc[:, 0], cc[:, 1], cc[:, 2]]
        f = np.ones(len(cells))
        for a in range(3):
            f *= bspline_np((pos[:, a] - (c[:, a] + 0.5) * w) / w)
        acc += v * f
    return acc


@pytest.fixture(scope="module", params=["cloud3_d6", "cloud11_d7"])
def case(request, pkg, orc):
    """Clouds with coarse and finer subdivision sets: batch cloud 3 at depth 6
    (a coarse root at depth 1) and cloud 11 at depth 7 (coarse roots at depth 2)."""
    idx, D = {"cloud3_d6": (3, 6), "cloud11_d7": (11, 7)}[request.param]
    xyz, nrm = pkg.synth.batch64()[idx]
    ref = orc.Result(xyz, nrm, D, precision=1)
    assert ref.scalar("q1") == 0
    sets = ref.array("sub_sets").reshape(-1, 4)
    assert (sets[:, 1] == 1).any() and (sets[:, 1] == 0).any()
    return ref, Tree(ref, D), D


def _rows(a):
    a = np.asarray(a, np.int64)
    return a[np.lexsort(a.T[::-1])]


def test_vertex_ownership_brute_force(case):
    ref, T, D = case
    verts = ref.array("verts")
    pos = verts[:, :3].copy().view(np.float32)
    gp = np.rint(pos.astype(np.float64) * (1 << verts[:, 5])[:, None]).astype(np.int64)
    got = _rows(np.column_stack([verts[:, 5], gp, verts[:, 3], verts[:, 4]]))
    want = []
    for d in range(D + 1):
        sl = slice(T.base[d], T.base[d + 1])
        g = np.unique((T.cell[sl][:, None, :] + CORNERS[None]).reshape(-1, 3), axis=0)
        best = np.full(len(g), -1, np.int64)
        bestKey = np.full(len(g), np.iinfo(np.int64).max)
        for dl in CORNERS:
            nb = T.lookup(d, g - dl)
            k = np.where(nb >= 0, T.key[np.maximum(nb, 0)], np.iinfo(np.int64).max)
            better = k < bestKey
            best[better], bestKey[better] = nb[better], k[better]
        kind = g - T.cell[best]
        kind = kind[:, 0] | (kind[:, 1] << 1) | (kind[:, 2] << 2)
        want.append(np.column_stack([np.full(len(g), d), g, best, kind]))
    np.testing.assert_array_equal(got, _rows(np.concatenate(want)))


def test_edge_and_face_ownership_brute_force(case, ref_tables):
    ref, T, D = case
    edges = ref.array("edges")  # kind, owner
    sl = slice(T.base[D], T.base[D + 1])
    cells = T.cell[sl]
    # every depth-D edge (axis, lower end) of every depth-D node, its owner the
    # smallest key among the (up to 4) nodes sharing it
    want = []
    for axis in range(3):
        others = [a for a in range(3) if a != axis]
        ends = []
        for o0 in (0, 1):
            for o1 in (0, 1):
                e = cells.copy()
                e[:, others[0]] += o0
                e[:, others[1]] += o1
                ends.append(e)
        ends = np.unique(np.concatenate(ends), axis=0)
        best = np.full(len(ends), -1, np.int64)
        bestKey = np.full(len(ends), np.iinfo(np.int64).max)
        for o0 in (0, 1):
            for o1 in (0, 1):
                c = ends.copy()
                c[:, others[0]] -= o0
                c[:, others[1]] -= o1
                nb = T.lookup(D, c)
                k = np.where(nb >= 0, T.key[np.maximum(nb, 0)], np.iinfo(np.int64).max)
                better = k < bestKey
                best[better], bestKey[better] = nb[better], k[better]
        off = ends - T.cell[best]
        kind = (axis << 2) | off[:, others[0]] | (off[:, others[1]] << 1)
        want.append(np.column_stack([kind, best]))
    np.testing.assert_array_equal(_rows(edges), _rows(np.concatenate(want)))
    # faces at every depth: owner the smaller key of the two nodes sharing it
    faces = ref.array("faces")  # kind, owner, hasParentFace
    pfk = np.array(ref_tables["parentFaceKind"]).reshape(8, 6)
    want = []
    for d in range(D + 1):
        sl = slice(T.base[d], T.base[d + 1])
        idx = np.arange(T.base[d], T.base[d + 1])
        for f in range(6):
            axis, side = f >> 1, f & 1
            other = T.cell[sl].copy()
            other[:, axis] += 2 * side - 1
            nb = T.lookup(d, other)
            mine = np.where(nb >= 0, T.key[np.maximum(nb, 0)], np.iinfo(np.int64).max) > T.key[idx]
            own = idx[mine]
            son = (T.key[own] >> (3 * (D - d))) & 7
            hp = np.where(d == 0, -1, np.where(pfk[son, f] != -1, 1, -1))
            want.append(np.column_stack([np.full(len(own), f), own, hp]))
    np.testing.assert_array_equal(_rows(faces), _rows(np.concatenate(want)))


def test_vertex_values_dense_grid_with_descent(case, ref_tables):
    """Vertex values at every depth (Q8): the owner's ancestors by key
    truncation, then the descent into key-ordered child childrenVertexKind[kind]
    while the node has children, each chain node adding its 27 neighbours'
    d_x times their functions; owner 0 gives 0; minus the iso-value."""
    ref, T, D = case
    verts = ref.array("verts")
    pos = verts[:, :3].copy().view(np.float32).astype(np.float64)
    owner, kind, vd = verts[:, 3].astype(np.int64), verts[:, 4], verts[:, 5].astype(np.int64)
    grids = T.dx_grids(ref.array("dx"))
    cvk = np.array(ref_tables["childrenVertexKind"])
    val = np.zeros(len(verts))
    oc = T.cell[owner]
    for d in range(D + 1):
        up = vd >= d
        val[up] += chain_sum(grids, d, oc[up] >> (vd[up] - d)[:, None], pos[up])
    ex = cvk[kind]
    step = np.stack([(ex >> 2) & 1, (ex >> 1) & 1, ex & 1], 1)
    cur, curCell, alive = owner.copy(), oc.copy(), owner > 0
    for d in range(1, D + 1):
        go = alive & (vd < d) & (T.pnum[cur] > 0)
        alive &= ~((vd < d) & ~go)
        if not go.any():
            continue
        curCell[go] = 2 * curCell[go] + step[go]
        nxt = T.lookup(d, curCell[go])
        assert np.all(nxt >= 0)
        cur[go] = nxt
        val[go] += chain_sum(grids, d, curCell[go], pos[go])
    val[owner == 0] = 0.0
    want = val - ref.scalar("iso")
    got = ref.array("vvalue")
    scale = np.abs(got).max()
    assert np.abs(got - want).max() <= 1e-10 * scale
    # the coarse depths are covered, and the descent does occur
    assert (vd < D - 2).sum() > 100


def _virtual_cells(T, D, roots, rd):
    base = key_xyz(T.key[roots], D)
    s = 1 << (D - rd)
    off = np.stack(np.meshgrid(np.arange(s), np.arange(s), np.arange(s), indexing="ij"), -1).reshape(-1, 3)
    return (base[:, None, :] + off[None]).reshape(-1, 3)


def test_subdivision_owners_brute_force(case):
    """Owner of a virtual corner (edge): the virtual depth-D cell of the set
    with the smallest Morton key among those sharing it; every corner (edge)
    of a virtual cell is listed once, with the owner's key and its kind."""
    ref, T, D = case
    sets = ref.array("sub_sets").reshape(-1, 4)
    roots = ref.array("sub_roots")
    sv = ref.array("sub_v").reshape(-1, 7)
    se = ref.array("sub_e").reshape(-1, 3)
    n = 1 << D
    for s, (rd, coarse, first, nr) in enumerate(sets):
        # depth 7: the coarse sets and the finer sets of up to 2^17 cells (the
        # depth-6 case checks every set)
        if D >= 7 and not coarse and nr << (3 * (D - rd)) > (1 << 17):
            continue
        vc = _virtual_cells(T, D, roots[first:first + nr], rd)
        vkeys = np.sort(xyz_key(vc, D))

        def owner_of(points, deltas):
            bestKey = np.full(len(points), np.iinfo(np.int64).max)
            for dl in deltas:
                c = points - dl
                ok = np.all((c >= 0) & (c < n), axis=1)
                k = np.full(len(points), -1, np.int64)
                k[ok] = xyz_key(c[ok], D)
                j = np.searchsorted(vkeys, k)
                virt = ok & (j < len(vkeys)) & (vkeys[np.minimum(j, len(vkeys) - 1)] == k)
                better = virt & (k < bestKey)
                bestKey[better] = k[better]
            assert np.all(bestKey < np.iinfo(np.int64).max)
            return bestKey

        g = np.unique((vc[:, None, :] + CORNERS[None]).reshape(-1, 3), axis=0)
        ok_ = owner_of(g, CORNERS)
        kd = g - key_xyz(ok_, D)
        kind = kd[:, 0] | (kd[:, 1] << 1) | (kd[:, 2] << 2)
        want = _rows(np.column_stack([g, ok_, kind]))
        mine = sv[sv[:, 0] == s]
        p = mine[:, 3:6].copy().view(np.float32).astype(np.float64)
        gp = np.rint(p * n).astype(np.int64)
        np.testing.assert_array_equal(_rows(np.column_stack([gp, mine[:, 1], mine[:, 2]])), want)
        # edges: per axis, the cells sharing an edge differ in the other two axes
        wantE = []
        for axis in range(3):
            others = [a for a in range(3) if a != axis]
            deltas = []
            for o0 in (0, 1):
                for o1 in (0, 1):
                    dl = np.zeros(3, np.int64)
                    dl[others[0]], dl[others[1]] = o0, o1
                    deltas.append(dl)
            ends = np.unique(np.concatenate([vc + dl for dl in deltas]), axis=0)
            ok_ = owner_of(ends, deltas)
            off = ends - key_xyz(ok_, D)
            kindE = (axis << 2) | off[:, others[0]] | (off[:, others[1]] << 1)
            wantE.append(np.column_stack([ok_, kindE]))
        mineE = se[se[:, 0] == s][:, 1:]
        np.testing.assert_array_equal(_rows(mineE), _rows(np.concatenate(wantE)))


def test_subdivision_values_dense_grid(case):
    """Virtual vertex values: over every depth, the 27 cells around the
    owner's ancestor, real nodes only (virtual cells and everything below a
    root hold no node), minus the iso-value; stored as float."""
    ref, T, D = case
    sv = ref.array("sub_v").reshape(-1, 7)
    val = sv[:, 6].copy().view(np.float32)
    have = np.isfinite(val)
    assert have.sum() > 1000
    sv, val = sv[have], val[have]
    if len(sv) > 300000:  # a seeded sample at depth 7
        pick = np.random.default_rng(5).choice(len(sv), 300000, replace=False)
        sv, val = sv[pick], val[pick]
    pos = sv[:, 3:6].copy().view(np.float32).astype(np.float64)
    oc = key_xyz(sv[:, 1], D)
    grids = T.dx_grids(ref.array("dx"))
    acc = np.zeros(len(sv))
    for d in range(D + 1):
        acc += chain_sum(grids, d, oc >> (D - d), pos)
    want = (acc - ref.scalar("iso")).astype(np.float32)
    scale = float(np.abs(val).max())
    assert np.abs(val.astype(np.float64) - want).max() <= 2e-6 * scale
