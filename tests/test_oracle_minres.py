"""Pins of the oracle's solver choice for indefinite depths (CPU).

The reference drops every Laplacian entry with |L| <= 1e-6
(solver/LaplacianSolver.cu:62).  The entries depend only on the neighbour
offset, so each depth's operator is a constant 27-point stencil; at depth 9
only the diagonal and the six face entries survive the drop and the stencil's
symbol diag + 2 face (cos a + cos b + cos c) turns negative at (pi, pi, pi):
the operator is indefinite and conjugate gradients has no convergence
guarantee there.  The oracle's FP64 reference solves such depths with MINRES.

  * the stencil values are recomputed here from the oracle's 1-D inner
    products (FunctionData.h:134-190 tables) and the drop applied in numpy;
    the symbol's minimum is taken at the checkerboard mode, where a stencil
    with non-negative off-diagonal entries attains it;
  * MINRES is checked against a dense LU solve on random symmetric
    indefinite and definite sparse matrices, so a wrong rotation or a dropped
    Lanczos term fails.
"""
import ctypes

import numpy as np
import pytest


@pytest.fixture(scope="module")
def lib(orc):
    L = orc.lib()
    L.orc_inner.restype = ctypes.c_double
    L.orc_inner.argtypes = [ctypes.c_int] * 4
    L.orc_depth_indefinite.restype = ctypes.c_int
    L.orc_depth_indefinite.argtypes = [ctypes.c_int, ctypes.c_int]
    L.orc_minres.restype = ctypes.c_int
    return L


def _stencil(lib, depth):
    f = (1 << depth) - 1 + (1 << (depth - 1))
    dot = [lib.orc_inner(f, f + k, 0, 0) for k in (0, 1)]
    d2 = [lib.orc_inner(f, f + k, 1, 1) for k in (0, 1)]
    st = {}
    for a in (-1, 0, 1):
        for b in (-1, 0, 1):
            for c in (-1, 0, 1):
                o = (abs(a), abs(b), abs(c))
                v = dot[o[0]] * dot[o[1]] * dot[o[2]] * (d2[o[0]] + d2[o[1]] + d2[o[2]])
                if abs(v) > float(np.float32(1e-6)):
                    st[(a, b, c)] = v
    return st


def test_depth9_keeps_diagonal_and_faces_only(lib):
    st = _stencil(lib, 9)
    assert sorted(st) == sorted([(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)])
    diag, face = st[(0, 0, 0)], st[(1, 0, 0)]
    assert face > 0 and diag - 6 * face < 0  # symbol at (pi, pi, pi)


@pytest.mark.parametrize("D", [6, 7, 8, 9])
def test_indefinite_depths(lib, D):
    flags = [lib.orc_depth_indefinite(D, d) for d in range(D + 1)]
    expect = []
    for d in range(D + 1):
        if d < 2:
            expect.append(0)
            continue
        st = _stencil(lib, d)
        # every kept off-diagonal entry of these stencils is positive or every
        # one is negative; the symbol's extremes are then at theta = 0 and at
        # the checkerboard mode, where cos(pi o) = (-1)^|o|_1
        check = sum(v * (-1) ** (abs(a) + abs(b) + abs(c)) for (a, b, c), v in st.items())
        at0 = sum(st.values())
        expect.append(int(min(check, at0) < 0))
    assert flags == expect
    assert flags[-1] == (1 if D == 9 else 0)
    assert sum(flags[:-1]) == 0


def _random_sym(n, nnz_row, shift, seed):
    rng = np.random.default_rng(seed)
    A = np.zeros((n, n))
    for i in range(n):
        for j in rng.choice(n, nnz_row, replace=False):
            v = rng.standard_normal()
            A[i, j] += v
            A[j, i] += v
    A += np.diag(rng.uniform(-1, 1, n) + shift)
    return A


def _csr(A):
    rp, col, val = [0], [], []
    for i in range(A.shape[0]):
        nz = np.nonzero(A[i])[0]
        col.extend(nz.tolist())
        val.extend(A[i, nz].tolist())
        rp.append(len(col))
    return (np.array(rp, np.int32), np.array(col, np.int32), np.array(val, np.float64))


@pytest.mark.parametrize("shift,seed", [(0.0, 1), (0.5, 2), (12.0, 3)])
def test_minres_matches_dense_solve(lib, shift, seed):
    n = 120
    A = _random_sym(n, 3, shift, seed)
    ev = np.linalg.eigvalsh(A)
    if shift < 1:
        assert ev.min() < 0 < ev.max()  # indefinite cases really are
    else:
        assert ev.min() > 0
    rp, col, val = _csr(A)
    b = np.random.default_rng(seed + 10).standard_normal(n)
    x = np.zeros(n)
    P = lambda a, t: a.ctypes.data_as(ctypes.POINTER(t))
    it = lib.orc_minres(ctypes.c_int(n), P(rp, ctypes.c_int), P(col, ctypes.c_int), P(val, ctypes.c_double),
                        P(b, ctypes.c_double), P(x, ctypes.c_double), ctypes.c_double(1e-12), ctypes.c_int(10 * n))
    assert 0 < it <= 10 * n
    ref = np.linalg.solve(A, b)
    assert np.linalg.norm(A @ x - b) <= 1e-9 * np.linalg.norm(b)
    assert np.linalg.norm(x - ref) <= 1e-6 * np.linalg.norm(ref) * np.linalg.cond(A) / 1e3 + 1e-9 * np.linalg.norm(ref)
