"""The serving loop on the GPU: a short seeded Poisson-arrival stream
(sizes 10k-200k, depths 6-8) served with HBM-budget admission into batched
calls; every served mesh equals the mesh of the same cloud reconstructed
alone (psr_reconstruct, mesh/PoissonReconstruction.cpp:195-245), bit for bit
(the batches run the single-cloud CG plan)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a GPU"
    return t


@pytest.mark.parametrize("budget", [None, 400 << 20])
def test_served_meshes_equal_single_calls(pkg, torch, budget):
    reqs = pkg.synth.request_stream(count=16, rate_hz=40.0, hi=200000)
    dev = [(a, torch.from_numpy(x).cuda(), torch.from_numpy(n).cuda(), d) for a, x, n, d in reqs]
    print("serving", len(dev), "requests, budget", budget, flush=True)
    srv = pkg.serve.Server(depths=(6, 7, 8), precision=pkg.binding.PRECISION_FP32, budget_bytes=budget)
    got = {}
    recs = srv.serve(dev, realtime=True,
                     on_mesh=lambda i, m: got.__setitem__(i, pkg.PoissonReconstruction.mesh_to_host(m)))
    learned = dict(srv.bytes_per_point)
    srv.close()
    assert sorted(got) == list(range(len(dev)))
    assert all(r["latency"] > 0 and r["done"] >= r["arrival"] for r in recs)
    if budget is not None:
        by_batch = {}
        for r in recs:
            by_batch.setdefault(r["batch"], []).append(r)
        for b in by_batch.values():
            # a batch of several requests stays within the budget by the
            # estimates in force at admission
            if len(b) > 1:
                assert sum(r["estimate"] for r in b) <= budget
    assert all(learned[d] > 0 for d in (6, 7, 8) if any(r["depth"] == d for r in recs))
    single = {}
    for i, (a, x, n, d) in enumerate(dev):
        rec = single.setdefault(d, pkg.PoissonReconstruction(depth=d))
        rec.SolvePoissionReconstructionMesh(x, n)
        v, t = rec.GetRebuildMeshVertices(), rec.GetRebuildMeshTriangleIndices()
        gv, gt = got[i]
        np.testing.assert_array_equal(gv, v)
        np.testing.assert_array_equal(gt, t)
    for rec in single.values():
        rec.close()
