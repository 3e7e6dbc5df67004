"""Pins the oracle's octree (normalisation, Morton codes, sort, node arrays,
neighbours, packed function index) with constructions independent of its
code: ceil-based cell indices, stable argsort, key-lookup neighbours and
subtree counting.  Reference: mesh/BuildOctree.cu, mesh/BuildOctree.cpp:75-194.
"""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def normalise_np(xyz, depth):
    # float32 arithmetic of BuildOctree.cpp:87-107 / BuildOctree.cu:144-165
    x = xyz.astype(np.float32)
    mx, mn = x.max(0), x.min(0)
    ext = (mx - mn).astype(np.float32)
    edge = np.float32(ext[0])
    for a in (1, 2):
        if edge < ext[a]:
            edge = ext[a]
    center = ((mx + mn).astype(np.float32) / np.float32(2)).astype(np.float32)
    edge = np.float32(edge * np.float32(1.25))
    corner = (center - edge / np.float32(2)).astype(np.float32)
    return ((x - corner).astype(np.float32) / edge).astype(np.float32), corner, edge


def cell_index(p, depth):
    # the level bits select the upper half iff p > centre; for a dyadic grid that
    # is ceil(p * 2^D) - 1 clamped to the grid
    k = np.ceil(p.astype(np.float64) * (1 << depth)).astype(np.int64) - 1
    return np.clip(k, 0, (1 << depth) - 1)


def interleave(ix, iy, iz, depth):
    code = np.zeros_like(ix)
    for b in range(depth):
        code |= ((ix >> b) & 1) << (3 * b + 2)
        code |= ((iy >> b) & 1) << (3 * b + 1)
        code |= ((iz >> b) & 1) << (3 * b)
    return code


@pytest.mark.parametrize("n,depth", [(3000, 5), (20000, 6)])
def test_codes_and_sort_match_cell_indices(pkg, orc, n, depth):
    xyz, nrm = pkg.synth.sphere(n=n)
    r = orc.Result(xyz, nrm, depth, stages=orc.STAGE_OCTREE)
    p, corner, edge = normalise_np(xyz, depth)
    assert np.float32(r.scalar("edge")) == edge
    assert all(np.float32(r.scalar("corner%d" % a)) == corner[a] for a in range(3))
    ci = [cell_index(p[:, a], depth) for a in range(3)]
    code = interleave(*ci, depth)
    order = np.argsort(code, kind="stable")
    np.testing.assert_array_equal(r.array("order"), order)
    np.testing.assert_array_equal(r.array("codes"), code[order])
    np.testing.assert_array_equal(r.array("spos").reshape(-1, 3), p[order])


def test_node_array_invariants(small_sphere_ref):
    r = small_sphere_ref
    D = r.depth
    base, counts = r.base, r.counts
    key = r.array("node_key")
    pidx, pnum = r.array("node_pidx"), r.array("node_pnum")
    dnum, didx = r.array("node_dnum"), r.array("node_didx")
    parent, children = r.array("node_parent"), r.array("node_children")
    codes = r.array("codes")
    N = len(codes)
    assert counts[0] == 1 and all(c % 8 == 0 for c in counts[1:])
    for d in range(D + 1):
        sl = slice(base[d], base[d + 1])
        assert pnum[sl].sum() == N                       # every point once per level
        sh = 3 * (D - d)
        ks = key[sl]
        assert np.all(ks & ((1 << sh) - 1) == 0)         # level-aligned keys
        assert np.all(np.diff(ks) > 0)                   # Morton order, unique
        occ = np.nonzero(pnum[sl])[0] + base[d]
        for o in occ[:: max(1, len(occ) // 200)]:
            seg = codes[pidx[o]:pidx[o] + pnum[o]]
            assert np.all((seg >> sh) << sh == key[o])
        # points of the level partition [0, N) in key order
        assert np.all(pidx[sl][1:] == (pidx[sl] + pnum[sl])[:-1])
    # depth-D slots below each node: dnum counts them, didx is the first one
    nD = counts[D]
    for d in range(D):
        sh = 3 * (D - d)
        for o in range(base[d], base[d + 1]):
            if pnum[o] == 0:
                assert dnum[o] == 0 and np.all(children[o] == -1)
                continue
            under = np.nonzero((key[base[D]:base[D] + nD] >> sh) << sh == key[o])[0]
            assert dnum[o] == len(under) and didx[o] == under[0]
            if d < D:
                np.testing.assert_array_equal(children[o], children[o][0] + np.arange(8))
                assert np.all(parent[children[o]] == o) or r.scalar("q1") > 0
    assert np.all(dnum[base[D]:] == 1)
    np.testing.assert_array_equal(didx[base[D]:], np.arange(nD))
    assert r.scalar("q1") == 0 and r.scalar("q1b") == 0


def test_neighbors_match_key_lookup(small_sphere_ref):
    r = small_sphere_ref
    D = r.depth
    base = r.base
    key = r.array("node_key")
    neighs = r.array("node_neighs")
    for d in range(1, D + 1):
        sh = 3 * (D - d)
        lut = {}
        ks = key[base[d]:base[d + 1]] >> sh
        for i, k in enumerate(ks):
            lut[int(k)] = base[d] + i
        size = 1 << d

        def decode(k):
            x = y = z = 0
            for b in range(d):
                x |= ((k >> (3 * b + 2)) & 1) << b
                y |= ((k >> (3 * b + 1)) & 1) << b
                z |= ((k >> (3 * b)) & 1) << b
            return x, y, z

        def encode(x, y, z):
            k = 0
            for b in range(d):
                k |= ((x >> b) & 1) << (3 * b + 2) | ((y >> b) & 1) << (3 * b + 1) | ((z >> b) & 1) << (3 * b)
            return k

        for i, k in enumerate(ks):
            x, y, z = decode(int(k))
            want = []
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dz in (-1, 0, 1):
                        X, Y, Z = x + dx, y + dy, z + dz
                        if 0 <= X < size and 0 <= Y < size and 0 <= Z < size:
                            want.append(lut.get(encode(X, Y, Z), -1))
                        else:
                            want.append(-1)
            np.testing.assert_array_equal(neighs[base[d] + i], want)


def test_encoded_function_index(small_sphere_ref):
    r = small_sphere_ref
    D = r.depth
    key, enc, depth = r.array("node_key"), r.array("encidx"), r.array("depth")
    s1 = 1 << (D + 1)
    fx, fy, fz = enc % s1, (enc // s1) % s1, enc // (s1 * s1)
    for o in range(len(key)):
        d = depth[o]
        cell = [0, 0, 0]
        for b in range(d):
            lvl_bit = 3 * (D - d + b)
            cell[0] |= ((key[o] >> (lvl_bit + 2)) & 1) << b
            cell[1] |= ((key[o] >> (lvl_bit + 1)) & 1) << b
            cell[2] |= ((key[o] >> lvl_bit) & 1) << b
        off = (1 << d) - 1
        assert (fx[o], fy[o], fz[o]) == (off + cell[0], off + cell[1], off + cell[2])


def test_lut_derivation_matches_reference_tables(orc):
    tab = json.load(open(os.path.join(HERE, "golden", "reference_tables.json")))
    lp = np.array(tab["LUTparent"]).reshape(8, 27)
    lc = np.array(tab["LUTchild"]).reshape(8, 27)
    for s in range(8):
        for i in range(27):
            assert orc.lut(0, s, i) == lp[s, i]
            assert orc.lut(1, s, i) == lc[s, i]


def test_point_to_node_map(small_sphere_ref):
    r = small_sphere_ref
    D = r.depth
    base = r.base
    p2n = r.array("point2node")
    key, pidx, pnum = r.array("node_key"), r.array("node_pidx"), r.array("node_pnum")
    codes = r.array("codes")
    np.testing.assert_array_equal(key[base[D] + p2n], codes)
    i = np.arange(len(codes))
    assert np.all((pidx[base[D] + p2n] <= i) & (i < pidx[base[D] + p2n] + pnum[base[D] + p2n]))


def test_level_local_child0_quirk_detected(pkg, orc):
    """Cloud 36 of the 64-cloud batch (a rotated torus, depth 7) leaves octant 0
    empty while block 0 of depth 2 holds an occupied child 0: the reference's
    parent rule (BuildOctree.cu:384-389, 412) gives that block an empty parent
    (quirk Q1).  The oracle reports it and stays inside its arrays."""
    x, n = pkg.synth.batch64()[36]
    r = orc.Result(x, n, 7, stages=orc.STAGE_OCTREE, precision=0)
    assert r.scalar("q1") == 1
    base = r.base
    par = r.array("node_parent")
    assert all(par[base[2] + k] == base[1] for k in range(8))
