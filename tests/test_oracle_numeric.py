"""Pins the oracle's floating-point stages with brute-force global sums:

* vector field V(o) = sum over ALL points of F_o(p) n_p (the 27-neighbour
  restriction of ComputeVectorField.cu:61-100 only drops zero terms);
* divergence: sum over the depth-D nodes lying inside o's 3x3x3 neighbourhood
  (ComputeNodesDivergence.cu:25-136), found geometrically;
* per-depth solution against a dense solve of the geometrically assembled
  operator (LaplacianSolver.cu:27-74);
* implicit function at points = sum over ALL nodes of dx_o F_o(p)
  (LaplacianSolver.cu:85-118 restricted to the non-zero terms).
"""
import numpy as np

from conftest import bspline_np, func_cw


def _node_funcs(r):
    D = r.depth
    enc = r.array("encidx")
    s1 = 1 << (D + 1)
    return np.stack([enc % s1, (enc // s1) % s1, enc // (s1 * s1)], axis=1)


def test_vector_field_brute_force(small_sphere_ref):
    r = small_sphere_ref
    D = r.depth
    base = r.base
    f = _node_funcs(r)[base[D]:]
    p = r.array("spos").reshape(-1, 3).astype(np.float64)
    n = r.array("snrm").reshape(-1, 3).astype(np.float64)
    V = r.array("vfield").reshape(-1, 3)
    sel = np.arange(0, len(f), 7)
    for o in sel:
        w = np.ones(len(p))
        for a in range(3):
            c, wd = func_cw(f[o, a])
            w *= bspline_np((p[:, a] - c) / wd)
        np.testing.assert_allclose(V[o], w @ n, rtol=1e-10, atol=1e-9 * np.abs(V).max())


def test_divergence_brute_force(small_sphere_ref):
    r = small_sphere_ref
    D = r.depth
    base = r.base
    res = (1 << (D + 1)) - 1
    ddot = r.array("ddot").reshape(res, res)  # ddot[b, a] = <F_a', F_b>
    f = _node_funcs(r)
    V = r.array("vfield").reshape(-1, 3)
    div = r.array("div")
    fD = f[base[D]:]
    cD = np.stack([func_cw(fD[:, a])[0] for a in range(3)], axis=1)
    for o in range(0, r.count("nodes"), 13):
        c = np.array([func_cw(f[o, a])[0] for a in range(3)])
        w = func_cw(f[o, 0])[1]
        inside = np.all(np.abs(cD - c) < 1.5 * w, axis=1)
        terms = sum(V[inside, a] * ddot[fD[inside, a], f[o, a]] for a in range(3))
        want = terms.sum()
        assert abs(div[o] - want) <= 1e-10 * max(abs(div).max(), 1e-30), o


def test_cg_against_dense_solve(small_sphere_ref):
    r = small_sphere_ref
    D = r.depth
    base = r.base
    res = (1 << (D + 1)) - 1
    dot = r.array("dot").reshape(res, res)
    d2 = r.array("d2dot").reshape(res, res)
    f = _node_funcs(r)
    div, dx = r.array("div"), r.array("dx")
    for d in range(2, D):
        sl = slice(base[d], base[d + 1])
        fd = f[sl]
        n = len(fd)
        c = np.stack([func_cw(fd[:, a])[0] for a in range(3)], axis=1)
        w = func_cw(fd[0, 0])[1]
        A = np.zeros((n, n))
        for i in range(n):
            near = np.nonzero(np.all(np.abs(c - c[i]) < 1.5 * w, axis=1))[0]
            for j in near:
                v = (dot[fd[i, 0], fd[j, 0]] * dot[fd[i, 1], fd[j, 1]] * dot[fd[i, 2], fd[j, 2]] *
                     (d2[fd[i, 0], fd[j, 0]] + d2[fd[i, 1], fd[j, 1]] + d2[fd[i, 2], fd[j, 2]]))
                if abs(v) > float(np.float32(1e-6)):
                    A[i, j] = v
        x = np.linalg.lstsq(A, div[sl], rcond=None)[0]
        rel = np.linalg.norm(A @ dx[sl] - div[sl]) / max(np.linalg.norm(div[sl]), 1e-300)
        assert rel < 1e-10, (d, rel)
        np.testing.assert_allclose(dx[sl], x, rtol=1e-6, atol=1e-8 * np.abs(x).max())
    rr = r.array("cg_relres")
    assert np.all(rr <= 1e-10)


def test_point_chi_brute_force(small_sphere_ref):
    r = small_sphere_ref
    f = _node_funcs(r)
    dx = r.array("dx")
    p = r.array("spos").reshape(-1, 3).astype(np.float64)
    chi = r.array("chi_pts")
    cw = [func_cw(f[:, a]) for a in range(3)]
    for i in range(0, len(p), 97):
        w = np.ones(len(dx))
        for a in range(3):
            w *= bspline_np((p[i, a] - cw[a][0]) / cw[a][1])
        want = w @ dx
        assert abs(chi[i] - want) <= 1e-9 * np.abs(chi).max()
    assert r.scalar("iso") == np.float64(chi.mean()) or abs(r.scalar("iso") - chi.mean()) < 1e-9 * abs(chi.mean())


def test_fp32_restatement_close_to_fp64(orc, small_sphere):
    xyz, nrm, d = small_sphere
    a = orc.Result(xyz, nrm, d, stages=orc.STAGE_OCTREE | orc.STAGE_FIELD | orc.STAGE_SOLVE, precision=0)
    b = orc.Result(xyz, nrm, d, stages=orc.STAGE_OCTREE | orc.STAGE_FIELD | orc.STAGE_SOLVE, precision=1)
    it = a.array("cg_iters")
    assert np.all(it >= 1) and np.all(it < 10000)
    assert np.all(a.array("cg_relres") < 1e-5)
    np.testing.assert_allclose(a.array("chi_pts"), b.array("chi_pts"), rtol=0, atol=1e-4 * np.abs(b.array("chi_pts")).max())
