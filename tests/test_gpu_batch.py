"""The batched form at the 64-cloud configuration (5k-150k points, random
ellipsoids and tori, depth 7), through the C ABI.

With the single-cloud CG plan (batch_cg_plan = 0) every cloud's mesh equals its
psr_reconstruct mesh bit for bit, and the batch contexts' node arrays and
neighbour tables equal the oracle's for a sample of clouds (including cloud
36, where the level-local child 0 rule fires).  With the tuner's plans
(every depth's rows spread evenly over workgroups so that all solves fit a
few waves; the dot products are then summed in another order) the integer
stages stay bit-exact and the solution and mesh stay within
the parity tolerances of the single call: per-depth solution within 2e-5
relative L2, symmetric Hausdorff within 1e-4 of the bounding-box diagonal.
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a GPU"
    return t


@pytest.fixture(scope="module")
def clouds64(pkg, torch):
    host = pkg.synth.batch64()
    dev = [(torch.from_numpy(x).cuda(), torch.from_numpy(n).cuda()) for x, n in host]
    return host, dev


@pytest.fixture(scope="module")
def singles(pkg, torch, clouds64):
    """Every cloud reconstructed alone: mesh, solution and level layout."""
    host, dev = clouds64
    rec = pkg.PoissonReconstruction(depth=7)
    out = []
    for x, n in dev:
        rec.SolvePoissionReconstructionMesh(x, n)
        out.append({k: rec.export(k) for k in ("MESH_VERTICES", "MESH_TRIANGLES", "DX", "LEVEL_BASE", "NODE_KEY",
                                                 "NODE_NEIGHS")})
    rec.close()
    return out


INT_ARRAYS = [("CODES", "codes"), ("ORDER", "order"), ("NODE_KEY", "node_key"), ("NODE_PARENT", "node_parent"),
              ("NODE_NEIGHS", "node_neighs"), ("NODE_DIDX", "node_didx"), ("NODE_DNUM", "node_dnum"),
              ("NODE_PIDX", "node_pidx"), ("NODE_PNUM", "node_pnum")]


def test_batch64_single_plan_is_bit_exact(pkg, orc, torch, clouds64, singles):
    host, dev = clouds64
    batch = pkg.PoissonReconstruction(depth=7, batch_cg_plan=0)
    for call in range(2):  # the second call replays the captured graphs
        print("batch call", call, flush=True)
        meshes = batch.ReconstructBatch(dev)
        assert batch.batch_cg_plan() == 0
        for i, m in enumerate(meshes):
            v, t = batch.mesh_to_host(m)
            np.testing.assert_array_equal(v, singles[i]["MESH_VERTICES"])
            np.testing.assert_array_equal(t, singles[i]["MESH_TRIANGLES"])
            st = batch.batch_stats(i)
            assert st["n_points"] == host[i][0].shape[0] and st["mesh_triangles"] == t.shape[0]
    for i in (0, 36, 63):
        view = batch.batch_cloud(i)
        ref = orc.Result(host[i][0], host[i][1], 7, stages=orc.STAGE_OCTREE, precision=1)
        np.testing.assert_array_equal(view.export("LEVEL_COUNT"), ref.counts)
        for g, o in INT_ARRAYS:
            np.testing.assert_array_equal(view.export(g).reshape(-1), ref.array(o).reshape(-1))
        assert view.stats()["q1_fired"] == (1 if i == 36 else 0)
    batch.close()


def test_batch64_tuned_plans_within_tolerance(pkg, torch, clouds64, singles):
    host, dev = clouds64
    batch = pkg.PoissonReconstruction(depth=7)
    plans = []
    # the shape's first call is untimed (plan 0), then the tuner tries its 5
    # plans and keeps the fastest
    for call in range(8):
        meshes = batch.ReconstructBatch(dev)
        plans.append(batch.batch_cg_plan())
        print("batch call", call, "plan", plans[-1], flush=True)
        check = range(64) if call in (2, 7) else range(0, 64, 9)
        for i in check:
            view = batch.batch_cloud(i)
            np.testing.assert_array_equal(view.export("NODE_NEIGHS"), singles[i]["NODE_NEIGHS"])
            dx, dxs = view.export("DX").astype(np.float64), singles[i]["DX"].astype(np.float64)
            base = singles[i]["LEVEL_BASE"]
            if i != 36:  # cloud 36: the reference's CG fails at depth 2 (Q1), no solution to compare
                for d in range(len(base) - 1):
                    a, b = base[d], base[d + 1]
                    e = np.linalg.norm(dx[a:b] - dxs[a:b]) / max(np.linalg.norm(dxs[a:b]), 1e-30)
                    assert e <= 2e-5, (i, d, e)
                v, t = batch.mesh_to_host(meshes[i])
                sv = singles[i]["MESH_VERTICES"]
                assert t.size == 0 or (t.min() >= 0 and t.max() < len(v))
                diag = float(np.linalg.norm(np.ptp(sv, axis=0)))
                h = max(cKDTree(sv).query(v)[0].max(), cKDTree(v).query(sv)[0].max())
                assert h <= 1e-4 * diag, (i, h)
    assert plans[:6] == [0, 0, 1, 2, 4, 8] and plans[6] == plans[7]
    batch.close()


def test_batch_mixed_precision_single_plan(pkg, torch, clouds64):
    host, dev = clouds64
    sub = dev[:10]
    batch = pkg.PoissonReconstruction(depth=7, precision=pkg.binding.PRECISION_MIXED, batch_cg_plan=0)
    meshes = batch.ReconstructBatch(sub)
    single = pkg.PoissonReconstruction(depth=7, precision=pkg.binding.PRECISION_MIXED)
    for i, (x, n) in enumerate(sub):
        single.SolvePoissionReconstructionMesh(x, n)
        v, t = batch.mesh_to_host(meshes[i])
        np.testing.assert_array_equal(v, single.GetRebuildMeshVertices())
        np.testing.assert_array_equal(t, single.GetRebuildMeshTriangleIndices())
        assert max(batch.batch_stats(i)["cg_true_relres"]) <= 1e-5
    single.close()
    batch.close()
