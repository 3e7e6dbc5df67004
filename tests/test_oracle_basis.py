"""Pins the oracle's B-spline and inner-product tables
(mesh/FunctionData.h:83-312, math/PPolynomial.h:568-582) against closed forms:
the cardinal quadratic B-spline B (F = B / B(0) = 4B/3) has autocorrelations
<B, B(.-k)> = M6(k) (quintic B-spline: 11/20, 13/60, 1/120) and
<B', B'(.-k)> = -M6''(k) (1, -1/3, -1/6); cross-depth entries are checked with
adaptive quadrature of the closed form.
"""
import numpy as np
import pytest
from scipy import integrate

from conftest import bspline_np, func_cw

M6 = {0: 11 / 20, 1: 13 / 60, 2: 1 / 120}
M6dd = {0: 1.0, 1: -1 / 3, 2: -1 / 6}


def test_bspline_values(orc):
    for x, want in [(0.0, 1.0), (0.5, 2 / 3), (-0.5, 2 / 3), (1.0, 1 / 6), (1.5, 0.0), (2.0, 0.0)]:
        assert orc.basis(0, 0.5 + x) == pytest.approx(want, abs=1e-15)
    xs = np.linspace(-3, 3, 1201)
    # integral 4/3 and partition of unity sum_k F(x - k) = 4/3
    assert integrate.quad(lambda u: float(bspline_np(u)), -1.5, 1.5, points=[-0.5, 0.5])[0] == pytest.approx(4 / 3)
    s = sum(bspline_np(xs - k) for k in range(-5, 6))
    np.testing.assert_allclose(s, 4 / 3, rtol=1e-13)


@pytest.mark.parametrize("depth", [0, 3, 7])
@pytest.mark.parametrize("k", [0, 1, 2])
def test_same_depth_closed_forms(orc, depth, k):
    a = (1 << depth) - 1 + 0
    b = a + k
    if k >= (1 << depth):
        pytest.skip("offset beyond the level")
    w = 1.0 / (1 << depth)
    assert orc.inner(a, b, 0, 0) == pytest.approx(w * 16 / 9 * M6[k], rel=1e-13)
    assert orc.inner(a, b, 1, 1) == pytest.approx(16 / 9 * M6dd[k] / w, rel=1e-13)
    assert orc.inner(a, b, 1, 0) == pytest.approx(-orc.inner(b, a, 1, 0), abs=1e-13)


def _quad(a, b, da, db):
    ca, wa = func_cw(a)
    cb, wb = func_cw(b)
    lo, hi = max(ca - 1.5 * wa, cb - 1.5 * wb), min(ca + 1.5 * wa, cb + 1.5 * wb)
    if lo >= hi:
        return 0.0

    def f(x, c, w, d):
        u = (x - c) / w
        if d == 0:
            return float(bspline_np(u))
        h = 1e-7
        return float((bspline_np(u + h) - bspline_np(u - h)) / (2 * h) / w)

    pts = sorted({ca + k * wa for k in (-0.5, 0.5)} | {cb + k * wb for k in (-0.5, 0.5)})
    pts = [p for p in pts if lo < p < hi]
    return integrate.quad(lambda x: f(x, ca, wa, da) * f(x, cb, wb, db), lo, hi, points=pts or None,
                          epsabs=1e-14, epsrel=1e-12, limit=200)[0]


def test_cross_depth_entries_against_quadrature(orc):
    rng = np.random.default_rng(7)
    D = 6
    res = (1 << (D + 1)) - 1
    checked = 0
    while checked < 40:
        a, b = rng.integers(0, res, 2)
        ca, wa = func_cw(a)
        cb, wb = func_cw(b)
        if abs(ca - cb) > 1.5 * (wa + wb):
            continue
        for da, db in ((0, 0), (1, 0), (1, 1)):
            want = _quad(int(a), int(b), da, db)
            got = orc.inner(int(a), int(b), da, db)
            scale = max(abs(want), 1e-3 * (1 / max(wa, wb) if da + db else min(wa, wb)))
            assert abs(got - want) <= 2e-5 * scale, (a, b, da, db, got, want)
        checked += 1


def test_tables_symmetry(small_sphere_ref):
    r = small_sphere_ref
    res = (1 << (r.depth + 1)) - 1
    dot = r.array("dot").reshape(res, res)
    ddot = r.array("ddot").reshape(res, res)
    d2 = r.array("d2dot").reshape(res, res)
    np.testing.assert_array_equal(dot, dot.T)
    np.testing.assert_array_equal(d2, d2.T)
    off = ~np.eye(res, dtype=bool)
    np.testing.assert_array_equal(ddot[off], -ddot.T[off])
    assert np.abs(np.diag(ddot)).max() < 1e-13 * np.abs(ddot).max()  # <F', F> = 0 analytically

