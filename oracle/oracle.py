"""ctypes loader for the oracle library -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and the benchmark's CPU-baseline leg use this
module, as the checker or as the timed CPU port.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

STAGE_OCTREE, STAGE_FIELD, STAGE_SOLVE, STAGE_MESH = 1, 2, 4, 8
STAGE_ALL = 15
STAGE_GEOMETRY = 16  # with 1|2|4 and without 8: ownership and vertex values, no marching cubes


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def _declare(L):
    L.orc_run.restype = ctypes.c_void_p
    L.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                          ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.orc_free.argtypes = [ctypes.c_void_p]
    L.orc_scalar.restype = ctypes.c_double
    L.orc_scalar.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    return L


def native_port(out_dir):
    """The FP32 port compiled for the host it runs on (-march=native) into
    out_dir: the timed CPU baseline, built where it is timed."""
    so = os.path.join(out_dir, "liboracle_native.so")
    subprocess.check_call(["make", "-s", "-B", "-C", _HERE, "native", "NATIVE_SO=" + so])
    return _declare(ctypes.CDLL(so))


def time_port(L, xyz, nrm, depth, threads, precision=0):
    """Milliseconds of one whole reconstruction by library L, summed over its
    stage timers (normalisation and octree through marching cubes and
    subdivision; the one-time basis tables are excluded)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float32)
    nrm = np.ascontiguousarray(nrm, dtype=np.float32)
    h = L.orc_run(xyz.ctypes.data, nrm.ctypes.data, xyz.shape[0], depth, STAGE_ALL, precision, threads)
    try:
        return sum(L.orc_scalar(h, k.encode()) for k in ("t_octree", "t_field", "t_solve", "t_mesh"))
    finally:
        L.orc_free(h)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = _declare(ctypes.CDLL(_LIB_PATH))
        L.orc_count.restype = ctypes.c_longlong
        L.orc_count.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        L.orc_copy.restype = ctypes.c_int
        L.orc_copy.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_longlong]
        L.orc_inner.restype = ctypes.c_double
        L.orc_inner.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_basis.restype = ctypes.c_double
        L.orc_basis.argtypes = [ctypes.c_int, ctypes.c_double]
        L.orc_lut.restype = ctypes.c_int
        L.orc_lut.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_mc_word.restype = ctypes.c_longlong
        L.orc_mc_word.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


_DTYPES = {4: {"f": np.float32, "i": np.int32}, 8: {"f": np.float64, "i": np.int64}}
_FLOAT_FIELDS = {"spos", "snrm", "center", "vfield", "div", "dx", "chi_pts", "cg_relres",
                 "cg_recres", "vvalue", "mesh_v", "dot", "ddot", "d2dot"}
_NODE_WIDTH = {"children": 8, "vertices": 8, "neighs": 27, "edges": 12, "faces": 6}


class Result:
    """One oracle run; arrays are fetched lazily by name."""

    def __init__(self, xyz, nrm, depth, stages=STAGE_ALL, precision=1, threads=0):
        xyz = np.ascontiguousarray(xyz, dtype=np.float32)
        nrm = np.ascontiguousarray(nrm, dtype=np.float32)
        self.n = xyz.shape[0]
        self.depth = depth
        self._h = lib().orc_run(xyz.ctypes.data, nrm.ctypes.data, self.n, depth, stages, precision, threads)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_free(self._h)
            self._h = None

    def scalar(self, name):
        return lib().orc_scalar(self._h, name.encode())

    def count(self, name):
        return int(lib().orc_count(self._h, name.encode()))

    def array(self, name):
        if name.startswith("node_"):
            f = name[5:]
            m = self.count("nodes")
            w = _NODE_WIDTH.get(f, 1)
            out = np.empty((m, w) if w > 1 else m, dtype=np.int32)
        elif name in ("verts", "edges", "faces"):
            k = self.count(name)
            width = {"verts": 6, "edges": 2, "faces": 3}[name]
            out = np.empty((k, width), dtype=np.int32)
        else:
            k = self.count(name)
            if k < 0:
                raise KeyError(name)
            esz = 8 if name in ("vfield", "div", "dx", "chi_pts", "cg_relres", "cg_recres", "vvalue",
                                "dot", "ddot", "d2dot") else 4
            kind = "f" if name in _FLOAT_FIELDS else "i"
            out = np.empty(k, dtype=_DTYPES[esz][kind])
        rc = lib().orc_copy(self._h, name.encode(), out.ctypes.data, out.nbytes)
        if rc != 0:
            raise RuntimeError("orc_copy(%s) = %d" % (name, rc))
        return out

    @property
    def base(self):
        return [int(self.scalar("base%d" % d)) for d in range(self.depth + 2)]

    @property
    def counts(self):
        return [int(self.scalar("count%d" % d)) for d in range(self.depth + 1)]


def inner(a, b, da, db):
    return lib().orc_inner(a, b, da, db)


def basis(idx, x):
    return lib().orc_basis(idx, x)


def lut(which, s, i):
    return lib().orc_lut(which, s, i)


def mc_word(cube):
    return lib().orc_mc_word(cube) & ((1 << 64) - 1)


def mesh_normals(vertices, triangles):
    """Per-vertex normals in float32 (mesh/DrawMesh.cu:80-177): each
    triangle's unit normal cross(A - O, B - O) (CalculateMeshNormalsKernel
    :102-124, CrossProduct :93-100, VectorNormalize :80-91: zero stays zero),
    summed per vertex in triangle order (the reference's float atomics add in
    no fixed order), divided by the vertex's triangle count and normalised
    (:135-177).  A vertex no triangle touches gets zero (the reference divides
    0 by 0); a triangle with a corner outside the mesh is skipped."""
    import numpy as np
    V = np.asarray(vertices, np.float32).reshape(-1, 3)
    T = np.asarray(triangles, np.int64).reshape(-1, 3)
    nv = V.shape[0]
    T = T[np.all((T >= 0) & (T < nv), axis=1)]

    def unit(x):
        ln = np.sqrt(x[:, 0] * x[:, 0] + x[:, 1] * x[:, 1] + x[:, 2] * x[:, 2])
        out = np.zeros_like(x)
        nz = ln != 0
        out[nz] = x[nz] / ln[nz, None]
        return out

    O, A, B = V[T[:, 0]], V[T[:, 1]], V[T[:, 2]]
    a, b = A - O, B - O
    cr = np.stack([a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1], a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
                   a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]], 1).astype(np.float32)
    fn = unit(cr)
    S = np.zeros((nv, 3), np.float32)
    np.add.at(S, T.reshape(-1), np.repeat(fn, 3, axis=0))
    cnt = np.bincount(T.reshape(-1), minlength=nv).astype(np.float32)
    avg = np.zeros_like(S)
    has = cnt > 0
    avg[has] = S[has] / cnt[has, None]
    return unit(avg)


def estimate_normals(xyz, k=10, viewpoint=(0.0, 0.0, 0.0)):
    """Point normals by the published algorithm of PCL's NormalEstimation (the
    reference calls NormalEstimationOMP with k = 10 from a kd-tree and the
    default viewpoint, mesh/PoissonReconstruction.cpp:137-186; PCL itself is
    absent): the k nearest points (the point itself included), their mean and
    covariance (divided by k), the eigenvector of the smallest eigenvalue,
    flipped so that (viewpoint - p) . n >= 0, and curvature
    lambda_0 / (lambda_0 + lambda_1 + lambda_2).  FP64 throughout.
    Returns normals, curvature, neighbour indices (nearest first) and the
    distances to the k-th and (k+1)-th nearest points (for ties)."""
    import numpy as np
    from scipy.spatial import cKDTree
    P = np.asarray(xyz, np.float64).reshape(-1, 3)
    kk = min(k, P.shape[0])
    kq = min(kk + 1, P.shape[0])
    dist, idx = cKDTree(P).query(P, k=kq)
    dist, idx = dist.reshape(P.shape[0], -1), idx.reshape(P.shape[0], -1)
    nb = P[idx[:, :kk]]
    d = nb - nb.mean(axis=1, keepdims=True)
    C = np.einsum("nki,nkj->nij", d, d) / kk
    w, V = np.linalg.eigh(C)
    n = V[:, :, 0]
    flip = np.sum((np.asarray(viewpoint, np.float64) - P) * n, axis=1) < 0
    n[flip] *= -1
    tr = w.sum(axis=1)
    curv = np.where(tr != 0, w[:, 0] / np.where(tr != 0, tr, 1), 0.0)
    gap = (w[:, 1] - w[:, 0]) / np.where(tr != 0, tr, 1)
    dk1 = dist[:, kk] if kq > kk else np.full(P.shape[0], np.inf)
    return n, curv, idx[:, :kk], dist[:, kk - 1], dk1, gap


def vertex_colors(vertices, xyz, colors):
    """Vertex colours as mesh/DrawMesh.cu:179-211 forms them (float32): the 4
    nearest points, weights = squared distance / sum of the four (the farther
    point weighs more), colour = sum of weight x colour in neighbour order;
    equal weights where the four distances are all zero.  Returns the colours
    and the 4th / 5th nearest distances (for ties)."""
    import numpy as np
    from scipy.spatial import cKDTree
    V = np.asarray(vertices, np.float32).reshape(-1, 3)
    P = np.asarray(xyz, np.float32).reshape(-1, 3)
    C = np.asarray(colors, np.float32).reshape(-1, 3)
    kq = min(5, len(P))
    dist, idx = cKDTree(P.astype(np.float64)).query(V.astype(np.float64), k=kq)
    dist, idx = dist.reshape(len(V), -1), idx.reshape(len(V), -1)
    kk = min(4, len(P))
    t = V[:, None, :] - P[idx[:, :kk]]
    d = t[:, :, 0] * t[:, :, 0] + t[:, :, 1] * t[:, :, 1] + t[:, :, 2] * t[:, :, 2]
    s = np.zeros(len(V), np.float32)
    for j in range(kk):
        s = s + d[:, j]
    pos = s > 0
    inv = np.zeros_like(s)
    inv[pos] = np.float32(1) / s[pos]
    w = np.where(pos[:, None], d * inv[:, None], np.float32(1) / np.float32(kk)).astype(np.float32)
    out = np.zeros((len(V), 3), np.float32)
    for j in range(kk):
        out = out + C[idx[:, j]] * w[:, j:j + 1]
    d5 = dist[:, kk] if kq > kk else np.full(len(V), np.inf)
    return out, dist[:, kk - 1], d5
