"""The reference's own input cloud (bunny.ply, 35,947 points with normals,
the committed fixture tests/golden/bunny_points.npz) on the GPU at depth 7:
octree and neighbour arrays bit-exact against the oracle, the FP32 mesh
within 1e-4 of the bounding-box diagonal of the FP64 oracle's mesh, the mesh
invariants of the CPU test, and the mixed form's true residual <= 1e-5 at
every depth."""
import numpy as np
import pytest
from scipy.spatial import cKDTree

from test_bunny_host import bunny, mesh_invariants

pytestmark = pytest.mark.gpu

INT_ARRAYS = [("CODES", "codes"), ("ORDER", "order"), ("NODE_KEY", "node_key"), ("NODE_PARENT", "node_parent"),
              ("NODE_NEIGHS", "node_neighs"), ("NODE_DIDX", "node_didx"), ("NODE_DNUM", "node_dnum"),
              ("VERTS", "verts"), ("EDGES", "edges"), ("FACES", "faces")]


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a GPU"
    return t


@pytest.mark.parametrize("prec", [0, 1])
def test_bunny_depth7(pkg, orc, torch, prec):
    from residual_host import true_relres
    xyz, nrm = bunny()
    print("bunny depth 7, precision", prec, flush=True)
    rec = pkg.PoissonReconstruction(depth=7, precision=prec)
    rec.SolvePoissionReconstructionMesh(torch.from_numpy(xyz).cuda(), torch.from_numpy(nrm).cuda())
    ref = orc.Result(xyz, nrm, 7, precision=1)
    for g, o in INT_ARRAYS:
        np.testing.assert_array_equal(rec.export(g).reshape(-1), ref.array(o).reshape(-1))
    v, t = rec.GetRebuildMeshVertices(), rec.GetRebuildMeshTriangleIndices()
    prm = rec.export("NORM_PARAMS").astype(np.float64)
    mesh_invariants(v, t, prm[:3], prm[3], xyz, 7)
    ov = ref.array("mesh_v").reshape(-1, 3)
    diag = float(np.linalg.norm(np.ptp(ref.array("spos").reshape(-1, 3), axis=0)))
    h = max(cKDTree(ov).query(v)[0].max(), cKDTree(v).query(ov)[0].max())
    print("Hausdorff / diag %.2e" % (h / diag), flush=True)
    assert h <= 1e-4 * diag, h / diag
    st = rec.stats()
    assert all(k < 10000 for k in st["cg_iterations"])
    rel = true_relres(rec, 7)
    assert max(rel) <= (1e-5 if prec == 1 else 1e-3), rel
    rec.close()
