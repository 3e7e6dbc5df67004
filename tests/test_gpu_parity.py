"""GPU parity of the HIP pipeline against the oracle, through the C ABI.

Bars: integer/index arrays bit-exact; sorted points bit-exact; FP64 tables
within 1e-12 relative; vector field and divergence within 1e-5 of max|.|;
per-depth solution within 1e-5 (relative L2) of the FP64 CG; implicit values
within 1e-4 of max|.|; marching-cubes cube classes identical wherever no
corner value is within 1e-4 of zero (relative); mesh vertices within a
symmetric Hausdorff distance of 1e-4 of the bounding-box diagonal.
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

pytestmark = pytest.mark.gpu

CASES = {"sphere": 6, "torus": 7}


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a GPU"
    return t


def _upload(torch, xyz, nrm):
    return torch.from_numpy(xyz).cuda(), torch.from_numpy(nrm).cuda()


@pytest.fixture(scope="module")
def runs(pkg, orc, torch):
    out = {}
    for name, depth in CASES.items():
        xyz, nrm, d = pkg.synth.workload(name)
        assert d == depth
        print("case", name, flush=True)
        tx, tn = _upload(torch, xyz, nrm)
        rec = pkg.PoissonReconstruction(depth=depth)
        rec.SolvePoissionReconstructionMesh(tx, tn)
        torch.cuda.synchronize()
        ref = orc.Result(xyz, nrm, depth, precision=1)
        out[name] = (rec, ref, xyz, nrm)
    yield out
    for rec, *_ in out.values():
        rec.close()


EXACT = [("CODES", "codes"), ("ORDER", "order"), ("POINT2NODE", "point2node"), ("NODE_KEY", "node_key"),
         ("NODE_PIDX", "node_pidx"), ("NODE_PNUM", "node_pnum"), ("NODE_PARENT", "node_parent"),
         ("NODE_CHILDREN", "node_children"), ("NODE_NEIGHS", "node_neighs"), ("NODE_DIDX", "node_didx"),
         ("NODE_DNUM", "node_dnum"), ("NODE_ENCIDX", "encidx"), ("VERTS", "verts"), ("EDGES", "edges"),
         ("FACES", "faces"), ("NODE_VERTICES", "node_vertices"), ("NODE_EDGES", "node_edges"),
         ("NODE_FACES", "node_faces")]


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("gname,oname", EXACT)
def test_integer_arrays_bit_exact(runs, case, gname, oname):
    rec, ref, *_ = runs[case]
    got, want = rec.export(gname), ref.array(oname)
    assert got.shape[0] == want.shape[0]
    np.testing.assert_array_equal(got.reshape(want.shape), want)


@pytest.mark.parametrize("case", list(CASES))
def test_levels_and_sorted_points(runs, case):
    rec, ref, *_ = runs[case]
    np.testing.assert_array_equal(rec.export("LEVEL_BASE"), ref.base)
    np.testing.assert_array_equal(rec.export("LEVEL_COUNT"), ref.counts)
    np.testing.assert_array_equal(rec.export("SORTED_POS")[:, :3], ref.array("spos").reshape(-1, 3))
    # normals feed only the floating-point field; the length normalisation
    # (sqrt, divide) may round differently by one ulp
    np.testing.assert_allclose(rec.export("SORTED_NRM")[:, :3], ref.array("snrm").reshape(-1, 3), rtol=2.5e-7, atol=0)
    prm = rec.export("NORM_PARAMS")
    assert prm[3] == np.float32(ref.scalar("edge"))
    assert all(prm[a] == np.float32(ref.scalar("corner%d" % a)) for a in range(3))


def _relmax(got, want):
    got = np.asarray(got, np.float64)
    want = np.asarray(want, np.float64)
    return np.abs(got - want).max() / max(np.abs(want).max(), 1e-300)


@pytest.mark.parametrize("case", list(CASES))
def test_tables(runs, case):
    rec, ref, *_ = runs[case]
    for g, o in (("TABLE_DOT", "dot"), ("TABLE_DDOT", "ddot"), ("TABLE_D2DOT", "d2dot")):
        assert _relmax(rec.export(g), ref.array(o)) <= 1e-12


@pytest.mark.parametrize("case", list(CASES))
def test_field_and_divergence(runs, case):
    rec, ref, *_ = runs[case]
    assert _relmax(rec.export("VFIELD"), ref.array("vfield").reshape(-1, 3)) <= 1e-5
    assert _relmax(rec.export("DIVERGENCE"), ref.array("div")) <= 1e-5


@pytest.mark.parametrize("case", list(CASES))
def test_solution_per_depth(runs, case):
    rec, ref, *_ = runs[case]
    dx, dxo = rec.export("DX").astype(np.float64), ref.array("dx")
    base = ref.base
    for d in range(len(base) - 1):
        a, b = base[d], base[d + 1]
        rel = np.linalg.norm(dx[a:b] - dxo[a:b]) / max(np.linalg.norm(dxo[a:b]), 1e-300)
        assert rel <= 1e-5, (d, rel)
    st = rec.stats()
    assert all(1 <= k < 10000 for k in st["cg_iterations"][1:])
    assert all(r <= 1e-5 for r in st["cg_updated_residual"])


@pytest.mark.parametrize("case", list(CASES))
def test_implicit_values(runs, case):
    rec, ref, *_ = runs[case]
    assert _relmax(rec.export("CHI_POINTS"), ref.array("chi_pts")) <= 1e-4
    assert _relmax(rec.export("VVALUE"), ref.array("vvalue")) <= 1e-4
    iso = ref.scalar("iso")
    assert abs(rec.stats()["iso_value"] - iso) <= 1e-5 * abs(iso)


@pytest.mark.parametrize("case", list(CASES))
def test_cube_topology(runs, case):
    rec, ref, *_ = runs[case]
    cube, want = rec.export("CUBE"), ref.array("cube")
    vv = ref.array("vvalue")
    nv = ref.array("node_vertices")
    base = ref.base
    corner = np.abs(vv[nv[base[-2]:base[-1]] - 1]).min(axis=1)
    safe = corner > 1e-4 * np.abs(vv).max()
    np.testing.assert_array_equal(cube[safe], want[safe])


def triangle_overlap(v, t, ov, ot, tol=1e-6):
    """Fractions of the triangles of each mesh found in the other, order
    independent: vertices of both meshes closer than tol are one point
    (connected components of the pairs within tol), a triangle is the set of
    its three points."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    allv = np.concatenate([v, ov]).astype(np.float64)
    pairs = cKDTree(allv).query_pairs(tol, output_type="ndarray")
    n = len(allv)
    g = coo_matrix((np.ones(len(pairs)), (pairs[:, 0], pairs[:, 1])), shape=(n, n))
    _, lab = connected_components(g, directed=False)
    a = np.sort(lab[t], axis=1)
    b = np.sort(lab[ot + len(v)], axis=1)
    sa = set(map(tuple, a))
    sb = set(map(tuple, b))
    return len(sa & sb) / max(len(sb), 1), len(sa & sb) / max(len(sa), 1)


@pytest.mark.parametrize("case", list(CASES))
def test_mesh(runs, case):
    rec, ref, *_ = runs[case]
    v, t = rec.GetRebuildMeshVertices(), rec.GetRebuildMeshTriangleIndices()
    ov, ot = ref.array("mesh_v").reshape(-1, 3), ref.array("mesh_t").reshape(-1, 3)
    assert t.min() >= 0 and t.max() < len(v)
    lo, hi = ref.array("spos").reshape(-1, 3).min(0), ref.array("spos").reshape(-1, 3).max(0)
    diag = float(np.linalg.norm(hi - lo))
    h = max(cKDTree(ov).query(v)[0].max(), cKDTree(v).query(ov)[0].max())
    assert h <= 1e-4 * diag, h
    # every triangle, order independent: the oracle's triangles (as vertex
    # sets) are the GPU's, except around corners whose value is near zero
    found, extra = triangle_overlap(v, t, ov, ot)
    print("%s: oracle triangles found %.6f, GPU triangles found %.6f" % (case, found, extra))
    assert found >= 0.998 and extra >= 0.998, (found, extra)
    # on these inputs no corner value lies near zero: the cube classes, the
    # counts and the triangle list (index for index) equal the oracle's
    np.testing.assert_array_equal(rec.export("CUBE"), ref.array("cube"))
    assert len(v) == len(ov) and len(t) == len(ot)
    np.testing.assert_array_equal(t, ot)
    assert np.abs(v - ov).max() <= 1e-4 * diag


@pytest.mark.parametrize("case", list(CASES))
def test_operator_entries_per_depth(runs, case):
    """Stored operator entries per depth (the |L| <= 1e-6 drop,
    LaplacianSolver.cu:62) equal the oracle's CSR counts."""
    rec, ref, *_ = runs[case]
    nnz = ref.array("nnz")
    st = rec.stats()
    np.testing.assert_array_equal(np.array(st["cg_nnz"]), nnz[:len(st["cg_nnz"])])


def test_deterministic_and_reusable(pkg, torch, runs):
    rec, ref, xyz, nrm = runs["sphere"]
    tx, tn = _upload(torch, xyz, nrm)
    first = {k: rec.export(k) for k in ("NODE_NEIGHS", "DX", "VVALUE", "MESH_VERTICES", "MESH_TRIANGLES")}
    # a smaller cloud in between exercises buffer reuse in the same context
    sx, sn = pkg.synth.sphere(n=5000)
    sx_t, sn_t = _upload(torch, sx, sn)
    rec.SolvePoissionReconstructionMesh(sx_t, sn_t)
    fresh = pkg.PoissonReconstruction(depth=6)
    fresh.SolvePoissionReconstructionMesh(sx_t, sn_t)
    for k in ("NODE_NEIGHS", "DX", "MESH_VERTICES", "MESH_TRIANGLES"):
        np.testing.assert_array_equal(rec.export(k), fresh.export(k))
    fresh.close()
    rec.SolvePoissionReconstructionMesh(tx, tn)
    for k, v in first.items():
        np.testing.assert_array_equal(rec.export(k), v)


def test_depthsurfel_records(pkg, torch, runs):
    rec, ref, xyz, nrm = runs["sphere"]
    n = len(xyz)
    recs = np.zeros((n, 16), np.float32)  # DepthSurfel: pixel coord + pad, xyz@4, normal@8, colour@12
    recs[:, 4:7] = xyz
    recs[:, 7] = 1.0
    recs[:, 8:11] = nrm
    other = pkg.PoissonReconstruction(depth=6)
    other.SolvePoissionReconstructionMesh(torch.from_numpy(recs).cuda(), None)
    np.testing.assert_array_equal(other.export("MESH_TRIANGLES"), ref.array("mesh_t").reshape(-1, 3))
    np.testing.assert_array_equal(other.export("NODE_NEIGHS"), ref.array("node_neighs"))
    other.close()


def test_error_paths(pkg, torch):
    b = pkg.binding
    rec = pkg.PoissonReconstruction(depth=5, max_points=1000)
    xyz, nrm = pkg.synth.sphere(n=2000)
    tx, tn = _upload(torch, xyz, nrm)
    with pytest.raises(b.PsrError) as e:
        rec.SolvePoissionReconstructionMesh(tx, tn)
    assert e.value.code == b.PSR_ERR_CAPACITY
    with pytest.raises(b.PsrError) as e:
        rec.SolvePoissionReconstructionMesh(tx[:5].contiguous(), tn[:5].contiguous())
    assert e.value.code == b.PSR_ERR_INVALID
    same = torch.full((100, 3), 0.25, device="cuda")
    with pytest.raises(b.PsrError) as e:
        rec.SolvePoissionReconstructionMesh(same, tn[:100].contiguous())
    assert e.value.code == b.PSR_ERR_INVALID
    # the context stays usable after errors
    rec.SolvePoissionReconstructionMesh(tx[:1000].contiguous(), tn[:1000].contiguous())
    assert rec.stats()["mesh_triangles"] > 0
    rec.close()


def test_cg_iteration_cap_reported_not_raised(pkg, torch):
    xyz, nrm = pkg.synth.sphere(n=3000)
    tx, tn = _upload(torch, xyz, nrm)
    rec = pkg.PoissonReconstruction(depth=5, cg_max_iter=1)
    rec.SolvePoissionReconstructionMesh(tx, tn)
    st = rec.stats()
    assert max(st["cg_iterations"]) <= 1
    assert max(st["cg_updated_residual"]) > 1e-5
    rec.close()


def test_subdivision_chunking_does_not_change_the_mesh(pkg, torch, runs):
    """The subdivision sets run in working sets of bounded size; a tiny bound
    (many chunks, each set alone) gives the same vertices and triangles."""
    rec, ref, xyz, nrm = runs["torus"]
    st = rec.stats()
    assert st["roots"] > 0
    tx, tn = _upload(torch, xyz, nrm)
    small = pkg.PoissonReconstruction(depth=7, subdiv_chunk_nodes=1000)
    small.SolvePoissionReconstructionMesh(tx, tn)
    np.testing.assert_array_equal(small.GetRebuildMeshVertices(), rec.GetRebuildMeshVertices())
    np.testing.assert_array_equal(small.GetRebuildMeshTriangleIndices(), rec.GetRebuildMeshTriangleIndices())
    small.close()


def _true_relres(rec):
    """||b - A x|| / ||b|| per depth in FP64, with A rebuilt on the host from
    exported arrays by the reference's entry rule (LaplacianSolver.cu:27-74)."""
    from residual_host import true_relres
    return np.array(true_relres(rec, rec.depth))


@pytest.mark.parametrize("case", list(CASES))
def test_plain_fp32_true_residual(runs, case):
    rec, *_ = runs[case]
    rel = _true_relres(rec)
    assert rel.max() <= 1e-5, rel


@pytest.fixture(scope="module")
def mixed_runs(pkg, runs, torch):
    out = {}
    for name, (rec, ref, xyz, nrm) in runs.items():
        tx, tn = _upload(torch, xyz, nrm)
        m = pkg.PoissonReconstruction(depth=rec.depth, precision=pkg.binding.PRECISION_MIXED)
        m.SolvePoissionReconstructionMesh(tx, tn)
        torch.cuda.synchronize()
        out[name] = m
    yield out
    for m in out.values():
        m.close()


@pytest.mark.parametrize("case", list(CASES))
def test_mixed_precision_meets_residual_and_mesh_targets(runs, mixed_runs, case):
    """Mixed form (FP16 field storage, compensated FP32 CG, FP64 refinement):
    true relative residual <= 1e-5 at every depth, checked independently in
    FP64 on the host; solution close to the FP64 oracle's; mesh within 1e-4
    of the bounding-box diagonal of the FP64 oracle's mesh."""
    rec, ref, *_ = runs[case]
    m = mixed_runs[case]
    st = m.stats()
    rel = _true_relres(m)
    assert rel.max() <= 1e-5, rel
    reported = np.array(st["cg_true_relres"])
    np.testing.assert_allclose(reported, rel, rtol=0.05, atol=1e-9)
    assert all(k < 10000 for k in st["cg_iterations"])
    # FP16 field storage: the divergence carries FP16 rounding (2^-11 relative per field entry)
    assert _relmax(m.export("DIVERGENCE"), ref.array("div")) <= 2e-3
    dx, dxo = m.export("DX").astype(np.float64), ref.array("dx")
    assert np.linalg.norm(dx - dxo) / np.linalg.norm(dxo) <= 2e-3
    v = m.GetRebuildMeshVertices()
    ov = ref.array("mesh_v").reshape(-1, 3)
    lo, hi = ref.array("spos").reshape(-1, 3).min(0), ref.array("spos").reshape(-1, 3).max(0)
    diag = float(np.linalg.norm(hi - lo))
    h = max(cKDTree(ov).query(v)[0].max(), cKDTree(v).query(ov)[0].max())
    assert h <= 1e-4 * diag, h


def test_batch_matches_single_reconstructions(pkg, torch):
    """psr_reconstruct_batch with the single-cloud CG plan: every cloud's mesh
    equals its own single-cloud reconstruction bit for bit (same kernel
    bodies, same reduction order), whatever runs beside it; three clouds a
    pass, so the clouds go through several passes."""
    s = pkg.synth
    clouds = [s.sphere(n=6000), s.torus(n=30000)] + [tuple(c) for c in s.batch64(count=6)[:6]]
    dev = [(torch.from_numpy(x).cuda(), torch.from_numpy(n).cuda()) for x, n in clouds]
    batch = pkg.PoissonReconstruction(depth=6, batch_workers=3, batch_cg_plan=0)
    meshes = batch.ReconstructBatch(dev)
    assert len(meshes) == len(clouds)
    single = pkg.PoissonReconstruction(depth=6)
    for i, (x, n) in enumerate(dev):
        single.SolvePoissionReconstructionMesh(x, n)
        v, t = batch.mesh_to_host(meshes[i])
        np.testing.assert_array_equal(v, single.GetRebuildMeshVertices())
        np.testing.assert_array_equal(t, single.GetRebuildMeshTriangleIndices())
        st = batch.batch_stats(i)
        assert st["n_points"] == x.shape[0] and st["mesh_triangles"] == t.shape[0]
    # a second, smaller batch reuses the buffers
    meshes = batch.ReconstructBatch(dev[:2])
    single.SolvePoissionReconstructionMesh(*dev[1])
    v, t = batch.mesh_to_host(meshes[1])
    np.testing.assert_array_equal(t, single.GetRebuildMeshTriangleIndices())
    single.close()
    batch.close()


def test_batch_with_a_failing_cloud(pkg, torch):
    """A cloud that fails (fewer than 8 points) stops no other cloud: the call
    reports that cloud, and every other cloud's mesh equals its single-cloud
    reconstruction."""
    s = pkg.synth
    good = [s.sphere(n=5000), s.torus(n=8000), s.sphere(n=7000), s.torus(n=6000)]
    bad = (np.ascontiguousarray(good[0][0][:4]), np.ascontiguousarray(good[0][1][:4]))
    clouds = [good[0], bad, good[1], good[2], good[3]]
    dev = [(torch.from_numpy(x).cuda(), torch.from_numpy(n).cuda()) for x, n in clouds]
    batch = pkg.PoissonReconstruction(depth=6, batch_workers=2, batch_cg_plan=0)
    with pytest.raises(pkg.binding.PsrError, match="cloud 1"):
        batch.ReconstructBatch(dev)
    single = pkg.PoissonReconstruction(depth=6)
    for i in (0, 2, 3, 4):
        single.SolvePoissionReconstructionMesh(*dev[i])
        v, t = batch.mesh_to_host(batch.batch[i])
        np.testing.assert_array_equal(v, single.GetRebuildMeshVertices())
        np.testing.assert_array_equal(t, single.GetRebuildMeshTriangleIndices())
    # the context stays usable
    meshes = batch.ReconstructBatch([dev[0], dev[2], dev[3]])
    single.SolvePoissionReconstructionMesh(*dev[3])
    np.testing.assert_array_equal(batch.mesh_to_host(meshes[2])[1], single.GetRebuildMeshTriangleIndices())
    single.close()
    batch.close()


def test_level_local_child0_quirk_parity(pkg, orc, torch):
    """Quirk Q1 (BuildOctree.cu:384-389, 412) on a cloud that fires it: the node
    arrays and neighbour tables reproduce it bit-exactly, and the mesh stages,
    where the reference would read unassigned slots, follow the same defined
    rule as the oracle (absent slots are skipped)."""
    x, n = pkg.synth.batch64()[36]
    tx, tn = _upload(torch, x, n)
    rec = pkg.PoissonReconstruction(depth=7)
    rec.SolvePoissionReconstructionMesh(tx, tn)
    st = rec.stats()
    # FP32 restatement: with the broken tables the reference's FP32 CG rule
    # decides the solution (and with it the roots and the mesh)
    ref = orc.Result(x, n, 7, precision=0)
    assert st["q1_fired"] == 1 and ref.scalar("q1") == 1
    for g, o in (("NODE_PARENT", "node_parent"), ("NODE_NEIGHS", "node_neighs"), ("NODE_VERTICES", "node_vertices"),
                 ("NODE_FACES", "node_faces")):
        np.testing.assert_array_equal(rec.export(g).reshape(-1), ref.array(o).reshape(-1))
    assert st["roots"] == int(ref.scalar("n_roots"))
    assert st["mesh_triangles"] == ref.count("mesh_t") // 3
    rec.close()
