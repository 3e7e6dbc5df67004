"""GPU checks of mesh welding and per-vertex normals through the C ABI.

Every mesh vertex carries the key of its depth-D edge; the check is
geometric: the vertex lies on that edge (its two off-axis coordinates are
exactly the edge's grid coordinates, the on-axis one lies within the edge).
The welded mesh is compared bit for bit with a numpy restatement of the
weld (np.unique over the keys: first vertex of each key, ascending keys,
triangle corners through the inverse map).  Its edges are then counted: the
welded torus is closed (every edge shared by two triangles).
Per-vertex normals (mesh/DrawMesh.cu:80-177) are compared within 4e-6 with
the oracle's float32 restatement, on the raw and the welded mesh, and the
welded ones against the input's oriented normals at the nearest point.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = {"sphere": 6, "torus": 7}


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a GPU"
    return t


@pytest.fixture(scope="module")
def welded(pkg, torch):
    out = {}
    for name, depth in CASES.items():
        xyz, nrm, d = pkg.synth.workload(name)
        print("case", name, flush=True)
        rec = pkg.PoissonReconstruction(depth=d)
        rec.SolvePoissionReconstructionMesh(torch.from_numpy(xyz).cuda(), torch.from_numpy(nrm).cuda())
        torch.cuda.synchronize()
        v, t, k = rec.export("MESH_VERTICES"), rec.export("MESH_TRIANGLES"), rec.export("MESH_VERTEX_KEYS")
        n_raw = rec.MeshNormals()
        wm = rec.WeldMesh()
        wv, wt = rec.mesh_to_host(wm)
        assert np.array_equal(wv, rec.export("WELDED_VERTICES")) and np.array_equal(wt, rec.export("WELDED_TRIANGLES"))
        n_w = rec.MeshNormals(wm)
        norm = rec.export("NORM_PARAMS")
        out[name] = (d, v, t, k, wv, wt, n_raw, n_w, norm, xyz, nrm)
        rec.close()
    return out


def _decode(k, D):
    S = (1 << D) + 1
    axis = k % 3
    r = k // 3
    return np.stack([r // (S * S), (r // S) % S, r % S], 1).astype(np.int64), axis.astype(np.int64)


@pytest.mark.parametrize("case", list(CASES))
def test_vertex_lies_on_its_key_edge(welded, case):
    D, v, _, k = welded[case][:4]
    assert k.shape[0] == v.shape[0] > 0
    lo, axis = _decode(k.astype(np.int64), D)
    assert lo.max() <= (1 << D) and np.all(axis < 3)
    g = v.astype(np.float64) * (1 << D)
    for a in range(3):
        off = axis != a
        assert np.array_equal(g[off, a], lo[off, a].astype(np.float64)), "off-axis coordinate not on the grid"
        on = ~off
        assert np.all(g[on, a] >= lo[on, a]) and np.all(g[on, a] <= lo[on, a] + 1)


@pytest.mark.parametrize("case", list(CASES))
def test_weld_matches_numpy_restatement(welded, case):
    _, v, t, k, wv, wt = welded[case][:6]
    _, first, inv = np.unique(k, return_index=True, return_inverse=True)
    want_v = v[first]
    tt = t.reshape(-1)
    ok = (tt >= 0) & (tt < v.shape[0])
    want_t = np.where(ok, inv[np.clip(tt, 0, v.shape[0] - 1)], -1).reshape(-1, 3)
    assert wv.tobytes() == want_v.tobytes()
    assert np.array_equal(wt, want_t)
    assert wv.shape[0] < v.shape[0]  # the subdivision batches share edges with the depth-D batch


@pytest.mark.parametrize("case", list(CASES))
def test_welded_edges(welded, case):
    """The torus comes out a closed 2-manifold (every edge shared by two
    triangles, V - E + F = 0).  The depth-6 sphere keeps a few edges used once
    (11 of 36835 when measured): openings the reference's subdivision leaves;
    no edge is ever shared by more than two triangles."""
    wv, wt = welded[case][4:6]
    assert np.all(wt >= 0)
    e = np.sort(np.concatenate([wt[:, [0, 1]], wt[:, [1, 2]], wt[:, [2, 0]]]), axis=1)
    _, cnt = np.unique(e, axis=0, return_counts=True)
    print(case, "edge use histogram", np.bincount(cnt).tolist(), flush=True)
    assert cnt.max() <= 2
    if case == "torus":
        assert np.all(cnt == 2)
        assert wv.shape[0] - cnt.shape[0] + wt.shape[0] == 0
    else:
        assert np.count_nonzero(cnt == 1) <= cnt.shape[0] // 1000


@pytest.mark.parametrize("case", list(CASES))
def test_normals_match_oracle(welded, orc, case):
    _, v, t, _, wv, wt, n_raw, n_w = welded[case][:8]
    # within 4e-6 (unit vectors, about 32 ulp): the GPU's square root and
    # division round differently from numpy's in the last bits of every face
    # normal; the summation order is the same (a wrong grouping errs by 0.1-1)
    for got, want in ((n_raw, orc.mesh_normals(v, t)), (n_w, orc.mesh_normals(wv, wt))):
        d = np.abs(got.astype(np.float64) - want)
        print(case, "normals max diff", float(d.max()), "elements differing", int(np.count_nonzero(d)), flush=True)
        assert d.max() <= 4e-6
    assert np.allclose(np.linalg.norm(n_w, axis=1), 1, atol=1e-6)


@pytest.mark.parametrize("case", list(CASES))
def test_normals_follow_the_input_orientation(welded, case):
    from scipy.spatial import cKDTree
    wv, _, _, n_w, norm, xyz, nrm = welded[case][4:11]
    world = wv.astype(np.float64) * norm[3] + norm[:3]
    _, j = cKDTree(xyz).query(world)
    ref = nrm[j] / np.linalg.norm(nrm[j], axis=1, keepdims=True)
    cos = np.sum(n_w * ref, axis=1)
    sign = np.sign(np.median(cos))  # the indicator's sign convention may face either way
    print(case, "median |cos|", float(np.median(np.abs(cos))), "orientation", sign, flush=True)
    assert np.mean(np.abs(cos) > 0.9) > 0.98
    assert np.mean(np.sign(cos) == sign) > 0.99  # one consistent orientation


def test_mesh_tensors_alias_device_buffers(pkg):
    import torch
    xyz, nrm, d = pkg.synth.workload("sphere")
    rec = pkg.PoissonReconstruction(depth=d)
    mesh = rec.SolvePoissionReconstructionMesh(torch.from_numpy(xyz).cuda(), torch.from_numpy(nrm).cuda())
    v, t = rec.mesh_tensors(mesh)
    assert v.data_ptr() == mesh.vertices and t.data_ptr() == mesh.triangles  # no copy
    assert np.array_equal(v.cpu().numpy(), rec.GetRebuildMeshVertices())
    assert np.array_equal(t.cpu().numpy(), rec.GetRebuildMeshTriangleIndices())
    back = torch.utils.dlpack.from_dlpack(torch.utils.dlpack.to_dlpack(v))
    assert back.data_ptr() == mesh.vertices
    wv, wt = rec.mesh_tensors(rec.WeldMesh())
    assert wv.shape[1] == 3 and wt.shape[1] == 3 and int(wt.max()) < wv.shape[0]
    rec.close()
