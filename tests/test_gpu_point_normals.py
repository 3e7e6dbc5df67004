"""GPU point-normal estimation through the C ABI, against the oracle's
restatement of PCL's NormalEstimation (k nearest neighbours from a kd-tree,
smallest-eigenvalue eigenvector of their covariance, viewpoint flip;
mesh/PoissonReconstruction.cpp:137-186).

Bars: neighbour sets identical wherever the k-th and (k+1)-th distances are
not within 1e-5 relative (float32 vs float64 distances may order near-ties
differently); normals within 1e-4 (1 - |cos| <= 5e-9) and curvature within
1e-5 wherever the eigenvalue gap is at least 1e-3 of the trace (below it the
eigenvector is ill-conditioned); every GPU normal faces the viewpoint.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a GPU"
    return t


@pytest.mark.parametrize("case,n,vp", [("sphere", 30000, (0.0, 0.0, 0.0)), ("torus", 50000, (0.0, 0.0, 0.05)),
                                       ("torus", 200000, (0.0, 0.0, 0.0))])
def test_point_normals_match_oracle(pkg, orc, torch, case, n, vp):
    xyz, true_n = getattr(pkg.synth, case)(n=n)
    print("case", case, n, flush=True)
    rec = pkg.PoissonReconstruction(depth=7)
    nrm, cur, knn = rec.EstimateNormals(torch.from_numpy(xyz).cuda(), k=10, viewpoint=vp, with_knn=True)
    nrm, cur, knn = nrm.cpu().numpy(), cur.cpu().numpy(), knn.cpu().numpy()
    rec.close()
    wn, wc, widx, dk, dk1, gap = orc.estimate_normals(xyz, k=10, viewpoint=vp)
    tie = (dk1 - dk) <= 1e-5 * np.maximum(dk1, 1e-30)
    same = np.array([set(a) == set(b) for a, b in zip(knn, widx)])
    print("neighbour sets equal", float(same.mean()), "near ties", int(tie.sum()), flush=True)
    assert np.all(same | tie)
    assert np.all(knn[:, 0] == np.arange(len(xyz))) or np.mean(knn[:, 0] == np.arange(len(xyz))) > 0.999
    ok = same & (gap >= 1e-3)
    cos = np.sum(nrm.astype(np.float64) * wn, axis=1)
    print("well-conditioned", float(ok.mean()), "min cos", float(cos[ok].min()), flush=True)
    assert ok.mean() > 0.95
    assert np.all(cos[ok] >= 1 - 5e-9 - 1e-6)
    assert np.all(np.abs(cur[ok] - wc[ok]) <= 1e-5)
    assert np.all(np.sum((np.asarray(vp) - xyz.astype(np.float64)) * nrm, axis=1) >= -1e-6)
    # the synthetic clouds' own normals (rotated <= 2 degrees): the same lines
    t = true_n / np.linalg.norm(true_n, axis=1, keepdims=True)
    assert np.median(np.abs(np.sum(nrm * t, axis=1))) > 0.99


def test_rejects_large_k(pkg, torch):
    xyz, _ = pkg.synth.sphere(n=1000)
    rec = pkg.PoissonReconstruction(depth=5)
    with pytest.raises(pkg.binding.PsrError):
        rec.EstimateNormals(torch.from_numpy(xyz).cuda(), k=17)
    nrm, _, _ = rec.EstimateNormals(torch.from_numpy(xyz).cuda(), k=16)  # the context stays usable
    assert nrm.shape == (1000, 3)
    rec.close()
