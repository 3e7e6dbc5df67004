"""ctypes binding of libpsr.so with the reference's class interface.

``PoissonReconstruction.SolvePoissionReconstructionMesh`` mirrors
mesh/PoissonReconstruction.h:124 (name kept, misspelling included): it takes
an oriented point array resident on the GPU and leaves the mesh in device
buffers owned by the object, read back with ``GetRebuildMeshVertices`` /
``GetRebuildMeshTriangleIndices`` (mesh/ComputeTriangleIndices.h:331-342).
There is no CPU fallback: a missing library raises.
"""
import atexit
import ctypes
import os
import sys
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpsr.so")

PSR_OK, PSR_ERR_INVALID, PSR_ERR_CAPACITY, PSR_ERR_HIP, PSR_ERR_INTERNAL = range(5)
PRECISION_FP32, PRECISION_MIXED = 0, 1

ARRAYS = ["SORTED_POS", "SORTED_NRM", "CODES", "ORDER", "POINT2NODE", "NODE_KEY", "NODE_PIDX", "NODE_PNUM",
          "NODE_PARENT", "NODE_CHILDREN", "NODE_NEIGHS", "NODE_DIDX", "NODE_DNUM", "NODE_ENCIDX", "LEVEL_BASE",
          "LEVEL_COUNT", "NORM_PARAMS", "VFIELD", "DIVERGENCE", "DX", "CHI_POINTS", "VERTS", "EDGES", "FACES",
          "NODE_VERTICES", "NODE_EDGES", "NODE_FACES", "VVALUE", "MESH_VERTICES", "MESH_TRIANGLES", "TABLE_DOT",
          "TABLE_DDOT", "TABLE_D2DOT", "CUBE", "MESH_VERTEX_KEYS", "WELDED_VERTICES", "WELDED_TRIANGLES"]
ARRAY_ID = {n: i for i, n in enumerate(ARRAYS)}
_ARRAY_LAYOUT = {
    "SORTED_POS": (np.float32, 4), "SORTED_NRM": (np.float32, 4), "NODE_CHILDREN": (np.int32, 8),
    "NODE_NEIGHS": (np.int32, 27), "NORM_PARAMS": (np.float32, 1), "VFIELD": (np.float32, 3),
    "DIVERGENCE": (np.float32, 1), "DX": (np.float32, 1), "CHI_POINTS": (np.float32, 1), "VERTS": (np.int32, 6),
    "EDGES": (np.int32, 2), "FACES": (np.int32, 3), "NODE_VERTICES": (np.int32, 8), "NODE_EDGES": (np.int32, 12),
    "NODE_FACES": (np.int32, 6), "VVALUE": (np.float32, 1), "MESH_VERTICES": (np.float32, 3),
    "MESH_TRIANGLES": (np.int32, 3), "TABLE_DOT": (np.float64, 1), "TABLE_DDOT": (np.float64, 1),
    "TABLE_D2DOT": (np.float64, 1), "MESH_VERTEX_KEYS": (np.uint32, 1), "WELDED_VERTICES": (np.float32, 3),
    "WELDED_TRIANGLES": (np.int32, 3)}

EXPORTED_SYMBOLS = ["psr_default_config", "psr_create", "psr_reconstruct", "psr_reconstruct_batch",
                    "psr_get_batch_stats", "psr_batch_cloud", "psr_batch_cg_plan", "psr_sync", "psr_get_stats", "psr_export", "psr_destroy",
                    "psr_status_string", "psr_last_error", "psr_weld_mesh", "psr_write_ply", "psr_mesh_normals", "psr_estimate_normals",
                    "psr_vertex_colors", "psr_mesh_normals_static", "psr_mesh_normals_static_max_triangles"]


class Config(ctypes.Structure):
    _fields_ = [("depth", ctypes.c_int), ("max_points", ctypes.c_uint32), ("precision", ctypes.c_int),
                ("cg_tol", ctypes.c_float), ("cg_max_iter", ctypes.c_int), ("refine_rel_tol", ctypes.c_float),
                ("refine_max_outer", ctypes.c_int), ("subdiv_chunk_nodes", ctypes.c_uint32),
                ("batch_workers", ctypes.c_uint32), ("batch_cg_plan", ctypes.c_int32)]


class Points(ctypes.Structure):
    _fields_ = [("xyz", ctypes.c_void_p), ("normal", ctypes.c_void_p), ("count", ctypes.c_uint32),
                ("stride_bytes", ctypes.c_uint32)]


class Mesh(ctypes.Structure):
    _fields_ = [("vertices", ctypes.c_void_p), ("triangles", ctypes.c_void_p), ("n_vertices", ctypes.c_uint32),
                ("n_triangles", ctypes.c_uint32)]


_MAXD = 11


class Stats(ctypes.Structure):
    _fields_ = [("depth", ctypes.c_int), ("n_points", ctypes.c_uint32), ("n_nodes", ctypes.c_uint32),
                ("nodes_per_depth", ctypes.c_uint32 * _MAXD), ("cg_iterations", ctypes.c_int * _MAXD),
                ("cg_updated_residual", ctypes.c_double * _MAXD), ("cg_true_relres", ctypes.c_double * _MAXD),
                ("iso_value", ctypes.c_float), ("n_vertex_nodes", ctypes.c_uint32),
                ("n_edge_nodes", ctypes.c_uint32), ("n_face_nodes", ctypes.c_uint32),
                ("n_subdivision_roots", ctypes.c_uint32), ("n_coarse_roots", ctypes.c_uint32),
                ("n_mesh_vertices", ctypes.c_uint32), ("n_mesh_triangles", ctypes.c_uint32),
                ("q1_fired", ctypes.c_int), ("ms_octree", ctypes.c_float), ("ms_field", ctypes.c_float),
                ("ms_solve", ctypes.c_float), ("ms_mesh", ctypes.c_float), ("ms_total", ctypes.c_float),
                ("ms_cg_kernel", ctypes.c_float), ("cg_workgroups", ctypes.c_uint32),
                ("cg_nnz", ctypes.c_int * _MAXD), ("cg_outer", ctypes.c_int * _MAXD),
                ("cg_row_form", ctypes.c_int), ("ms_cg_kernel_sum", ctypes.c_double), ("n_runs", ctypes.c_uint32),
                ("device_bytes", ctypes.c_uint64)]

    def as_dict(self):
        d = self.depth
        return {
            "depth": d, "n_points": self.n_points, "n_nodes": self.n_nodes,
            "nodes_per_depth": list(self.nodes_per_depth)[:d + 1],
            "cg_iterations": list(self.cg_iterations)[:d + 1],
            "cg_updated_residual": list(self.cg_updated_residual)[:d + 1],
            "cg_true_relres": list(self.cg_true_relres)[:d + 1],
            "iso_value": self.iso_value, "V": self.n_vertex_nodes, "E": self.n_edge_nodes, "F": self.n_face_nodes,
            "roots": self.n_subdivision_roots, "coarse_roots": self.n_coarse_roots,
            "mesh_vertices": self.n_mesh_vertices, "mesh_triangles": self.n_mesh_triangles,
            "ms_octree": self.ms_octree, "ms_field": self.ms_field, "ms_solve": self.ms_solve,
            "ms_mesh": self.ms_mesh, "ms_total": self.ms_total, "ms_cg_kernel": self.ms_cg_kernel,
            "cg_workgroups": self.cg_workgroups, "cg_nnz": list(self.cg_nnz)[:d + 1],
            "cg_outer": list(self.cg_outer)[:d + 1], "cg_row_form": self.cg_row_form,
            "ms_cg_kernel_sum": self.ms_cg_kernel_sum, "n_runs": self.n_runs, "q1_fired": self.q1_fired,
            "device_bytes": self.device_bytes,
        }


_lib = None
_hip = None


def _memcpy_d2h(dst, src_ptr):
    """hipMemcpy of a device pointer into a numpy array (device-to-host), through
    the HIP runtime libpsr.so is linked against (resolved via its handle)."""
    global _hip
    if _hip is None:
        fn = lib().hipMemcpy
        fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        fn.restype = ctypes.c_int
        _hip = fn
    rc = _hip(dst.ctypes.data, ctypes.c_void_p(src_ptr), dst.nbytes, 2)  # hipMemcpyDeviceToHost
    if rc != 0:
        raise RuntimeError("hipMemcpy failed: %d" % rc)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libpsr.so is not built (%s); run build()" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        L.psr_default_config.argtypes = [ctypes.POINTER(Config)]
        L.psr_create.argtypes = [ctypes.POINTER(Config), ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
        L.psr_reconstruct.argtypes = [ctypes.c_void_p, ctypes.POINTER(Points), ctypes.POINTER(Mesh)]
        L.psr_reconstruct_batch.argtypes = [ctypes.c_void_p, ctypes.POINTER(Points), ctypes.c_uint32,
                                            ctypes.POINTER(Mesh)]
        L.psr_get_batch_stats.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(Stats)]
        L.psr_batch_cloud.restype = ctypes.c_void_p
        L.psr_batch_cloud.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
        L.psr_batch_cg_plan.argtypes = [ctypes.c_void_p]
        L.psr_sync.argtypes = [ctypes.c_void_p]
        L.psr_get_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(Stats)]
        L.psr_export.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64,
                                 ctypes.POINTER(ctypes.c_uint64)]
        L.psr_destroy.argtypes = [ctypes.c_void_p]
        L.psr_weld_mesh.argtypes = [ctypes.c_void_p, ctypes.POINTER(Mesh)]
        L.psr_mesh_normals.argtypes = [ctypes.c_void_p, ctypes.POINTER(Mesh), ctypes.c_void_p]
        L.psr_mesh_normals_static.argtypes = [ctypes.POINTER(Mesh), ctypes.c_void_p, ctypes.c_void_p]
        L.psr_mesh_normals_static_max_triangles.restype = ctypes.c_uint32
        L.psr_mesh_normals_static_max_triangles.argtypes = []
        L.psr_estimate_normals.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                           ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_void_p]
        L.psr_vertex_colors.argtypes = [ctypes.c_void_p, ctypes.POINTER(Mesh), ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]
        L.psr_write_ply.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                    ctypes.c_uint32]
        L.psr_status_string.restype = ctypes.c_char_p
        L.psr_status_string.argtypes = [ctypes.c_int]
        L.psr_last_error.restype = ctypes.c_char_p
        L.psr_last_error.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


# Contexts still open; closed at interpreter exit, before the HIP runtime and
# torch tear down, so that no HIP call runs during finalisation.
_live = weakref.WeakSet()


@atexit.register
def _close_all():
    for rec in list(_live):
        try:
            rec.close()
        except Exception:
            pass


def _check_points(t, width, name):
    """A contiguous float32 (N, width) tensor resident on the GPU."""
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda):
        raise TypeError("%s must be a CUDA tensor" % name)
    if t.dtype != torch.float32 or t.dim() != 2 or t.shape[1] != width:
        raise ValueError("%s must be float32 of shape (N, %d), got %s %s" % (name, width, t.dtype, tuple(t.shape)))
    if not t.is_contiguous():
        raise ValueError("%s must be contiguous" % name)


class PsrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (lib().psr_status_string(code).decode(), code, msg))
        self.code = code


class PoissonReconstruction:
    """Host-side mirror of SparseSurfelFusion::PoissonReconstruction."""

    def __init__(self, depth=7, precision=PRECISION_FP32, max_points=0, cg_tol=1e-5, cg_max_iter=10000,
                 stream=None, subdiv_chunk_nodes=0, batch_workers=0, batch_cg_plan=-1):
        cfg = Config()
        lib().psr_default_config(ctypes.byref(cfg))
        cfg.depth = depth
        cfg.precision = precision
        cfg.max_points = max_points
        cfg.cg_tol = cg_tol
        cfg.cg_max_iter = cg_max_iter
        cfg.subdiv_chunk_nodes = subdiv_chunk_nodes
        cfg.batch_workers = batch_workers
        cfg.batch_cg_plan = batch_cg_plan
        self._ctx = ctypes.c_void_p()
        # The context works on a stream of its own, which waits for the
        # caller's current stream before a call while that stream has work in
        # flight (inputs written by torch kernels still running are read only
        # once they are complete).
        import torch
        self._stream = torch.cuda.Stream() if stream is None else None
        # one event, recorded on the caller's stream when the context must wait
        # for it (an event created per call costs more than the rest of the facade)
        self._ready = torch.cuda.Event() if stream is None else None
        handle = self._stream.cuda_stream if stream is None else stream
        rc = lib().psr_create(ctypes.byref(cfg), ctypes.c_void_p(handle), ctypes.byref(self._ctx))
        if rc != PSR_OK:
            raise PsrError(rc, "psr_create")
        self.depth = depth
        self.mesh = Mesh()
        _live.add(self)

    def close(self):
        if self._ctx:
            lib().psr_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self, _finalizing=sys.is_finalizing):
        # during interpreter finalisation the atexit hook has closed the
        # context already; no HIP call may run then (module globals may be
        # gone, so the check is bound at definition)
        if _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def _order_after_caller(self):
        # The context's stream waits for the caller's current stream only when
        # that stream still has work in flight: once it is idle, everything
        # queued on it before this call (the kernels writing the inputs) is
        # complete, and a wait would only delay the call's first kernel.
        if self._stream is not None:
            import torch
            cur = torch.cuda.current_stream()
            if not cur.query():
                self._ready.record(cur)
                self._stream.wait_event(self._ready)

    def solve_ptr(self, xyz_ptr, nrm_ptr, count, stride_bytes=12):
        self._order_after_caller()
        pts = Points(xyz_ptr, nrm_ptr, count, stride_bytes)
        rc = lib().psr_reconstruct(self._ctx, ctypes.byref(pts), ctypes.byref(self.mesh))
        if rc != PSR_OK:
            raise PsrError(rc, lib().psr_last_error(self._ctx).decode())
        return self.mesh

    def SolvePoissionReconstructionMesh(self, xyz, normals):
        """xyz, normals: contiguous float32 GPU tensors of shape (N, 3), or one
        (N, 16) float32 tensor of DepthSurfel records (xyz at 4..6, normal at 8..10)."""
        if normals is None:
            _check_points(xyz, 16, "DepthSurfel records")
            base = xyz.data_ptr()
            return self.solve_ptr(base + 16, base + 32, xyz.shape[0], 64)
        _check_points(xyz, 3, "xyz")
        _check_points(normals, 3, "normals")
        if normals.shape[0] != xyz.shape[0]:
            raise ValueError("xyz and normals differ in length")
        return self.solve_ptr(xyz.data_ptr(), normals.data_ptr(), xyz.shape[0], 12)

    def ReconstructBatch(self, clouds):
        """clouds: list of (xyz, normals) contiguous float32 GPU tensor pairs.
        Returns one psr_mesh per cloud (device buffers valid until the next batch)."""
        n = len(clouds)
        for x, nn in clouds:
            _check_points(x, 3, "xyz")
            _check_points(nn, 3, "normals")
            if nn.shape[0] != x.shape[0]:
                raise ValueError("xyz and normals differ in length")
        self._order_after_caller()
        pts = (Points * n)(*[Points(x.data_ptr(), nn.data_ptr(), x.shape[0], 12) for x, nn in clouds])
        self.batch = (Mesh * n)()
        rc = lib().psr_reconstruct_batch(self._ctx, pts, n, self.batch)
        if rc != PSR_OK:
            raise PsrError(rc, lib().psr_last_error(self._ctx).decode())
        return list(self.batch)

    def batch_cloud(self, i):
        """A view of the context that reconstructed cloud i of the last batch
        (export / stats), valid until the next batch call or close."""
        h = lib().psr_batch_cloud(self._ctx, i)
        if not h:
            raise PsrError(PSR_ERR_INVALID, "no batch cloud %d" % i)
        view = PoissonReconstruction.__new__(PoissonReconstruction)
        view._ctx = ctypes.c_void_p(h)
        view._stream = None
        view.depth = self.depth
        view.mesh = Mesh()
        view.close = lambda: None  # owned by the batch context
        return view

    def batch_cg_plan(self):
        """The CG plan the last batch call ran with (0: the single-cloud plan;
        w: every solve in w waves of the resident budget)."""
        return lib().psr_batch_cg_plan(self._ctx)

    def batch_stats(self, i):
        st = Stats()
        rc = lib().psr_get_batch_stats(self._ctx, i, ctypes.byref(st))
        if rc != PSR_OK:
            raise PsrError(rc, "batch stats %d" % i)
        return st.as_dict()

    @staticmethod
    def mesh_to_host(mesh):
        """Copies a psr_mesh (device buffers) to numpy (vertices (V, 3), triangles (T, 3))."""
        import torch
        v = np.empty((mesh.n_vertices, 3), np.float32)
        t = np.empty((mesh.n_triangles, 3), np.int32)
        torch.cuda.synchronize()
        if mesh.n_vertices:
            _memcpy_d2h(v, mesh.vertices)
        if mesh.n_triangles:
            _memcpy_d2h(t, mesh.triangles)
        return v, t

    def stats(self):
        st = Stats()
        lib().psr_get_stats(self._ctx, ctypes.byref(st))
        return st.as_dict()

    def export(self, name):
        aid = ARRAY_ID[name]
        need = ctypes.c_uint64(0)
        rc = lib().psr_export(self._ctx, aid, None, 0, ctypes.byref(need))
        if rc != PSR_OK:
            raise PsrError(rc, "export %s" % name)
        dtype, width = _ARRAY_LAYOUT.get(name, (np.int32, 1))
        n = need.value // np.dtype(dtype).itemsize
        out = np.empty(n, dtype=dtype)
        rc = lib().psr_export(self._ctx, aid, out.ctypes.data, out.nbytes, ctypes.byref(need))
        if rc != PSR_OK:
            raise PsrError(rc, "export %s" % name)
        return out.reshape(-1, width) if width > 1 else out

    def WeldMesh(self):
        """Welds the last mesh (one vertex per depth-D edge) on the device and
        returns its psr_mesh (device buffers valid until the next reconstruction)."""
        welded = Mesh()
        rc = lib().psr_weld_mesh(self._ctx, ctypes.byref(welded))
        if rc != PSR_OK:
            raise PsrError(rc, lib().psr_last_error(self._ctx).decode())
        return welded

    def MeshNormals(self, mesh=None):
        """Per-vertex normals (numpy (V, 3) float32) of a psr_mesh in device
        memory, by default the last reconstruction's."""
        import torch
        m = self.mesh if mesh is None else mesh
        out = torch.empty((max(m.n_vertices, 1), 3), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        rc = lib().psr_mesh_normals(self._ctx, ctypes.byref(m), out.data_ptr())
        if rc != PSR_OK:
            raise PsrError(rc, lib().psr_last_error(self._ctx).decode())
        return out[:m.n_vertices].cpu().numpy()

    def EstimateNormals(self, xyz, k=10, viewpoint=(0.0, 0.0, 0.0), with_knn=False):
        """Normals of an unoriented cloud (contiguous float32 (N, 3) GPU tensor),
        the step before SolvePoissionReconstructionMesh (the reference's PCL
        NormalEstimationOMP, mesh/PoissonReconstruction.cpp:137-186).  Returns
        (normals (N, 3), curvature (N,), knn (N, k) int32 or None), GPU tensors."""
        import torch
        assert xyz.is_cuda and xyz.is_contiguous() and xyz.dtype == torch.float32
        n = xyz.shape[0]
        nrm = torch.empty((max(n, 1), 3), dtype=torch.float32, device=xyz.device)
        cur = torch.empty((max(n, 1),), dtype=torch.float32, device=xyz.device)
        knn = torch.empty((max(n, 1), k), dtype=torch.int32, device=xyz.device) if with_knn else None
        vp = (ctypes.c_float * 3)(*viewpoint)
        torch.cuda.synchronize()
        rc = lib().psr_estimate_normals(self._ctx, xyz.data_ptr(), n, 12, k, vp, nrm.data_ptr(), cur.data_ptr(),
                                        knn.data_ptr() if with_knn else None)
        if rc != PSR_OK:
            raise PsrError(rc, lib().psr_last_error(self._ctx).decode())
        return nrm[:n], cur[:n], (knn[:n] if with_knn else None)

    def VertexColors(self, vertices, xyz, colors):
        """Colours (numpy (V, 3)) of mesh vertices (float32 (V, 3) GPU tensor, in
        the points' frame) from coloured points (float32 (N, 3) GPU tensors
        xyz and colors), as DrawMesh.cu:179-211 weights its 4 nearest points."""
        import torch
        v = vertices.contiguous()
        m = Mesh(v.data_ptr(), None, v.shape[0], 0)
        out = torch.empty((max(v.shape[0], 1), 3), dtype=torch.float32, device=v.device)
        torch.cuda.synchronize()
        rc = lib().psr_vertex_colors(self._ctx, ctypes.byref(m), xyz.data_ptr(), colors.data_ptr(), xyz.shape[0], 12,
                                     out.data_ptr())
        if rc != PSR_OK:
            raise PsrError(rc, lib().psr_last_error(self._ctx).decode())
        return out[:v.shape[0]].cpu().numpy()

    @staticmethod
    def mesh_tensors(mesh):
        """Zero-copy torch views (vertices (V, 3) float32, triangles (T, 3)
        int32) of a psr_mesh's device buffers, through
        __cuda_array_interface__; torch.utils.dlpack.to_dlpack hands them to
        other consumers.  They alias memory owned by the context: valid until
        its next reconstruction (or weld, for a welded mesh) or its close.
        Replaces the reference's GL buffer mapping (mesh/DrawMesh.cpp:146-209)."""
        import torch

        class _View:
            def __init__(self, ptr, shape, typestr):
                self.__cuda_array_interface__ = {"shape": shape, "typestr": typestr, "data": (ptr, False),
                                                 "version": 2, "strides": None}

        if mesh.n_vertices:
            v = torch.as_tensor(_View(mesh.vertices, (mesh.n_vertices, 3), "<f4"), device="cuda")
        else:
            v = torch.empty((0, 3), dtype=torch.float32, device="cuda")
        if mesh.n_triangles:
            t = torch.as_tensor(_View(mesh.triangles, (mesh.n_triangles, 3), "<i4"), device="cuda")
        else:
            t = torch.empty((0, 3), dtype=torch.int32, device="cuda")
        return v, t

    def GetRebuildMeshVertices(self):
        return self.export("MESH_VERTICES")

    def GetRebuildMeshTriangleIndices(self):
        return self.export("MESH_TRIANGLES")


def write_ply(path, vertices, triangles):
    """Writes host arrays (float32 (V, 3), int32 (T, 3)) as a binary PLY."""
    v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, dtype=np.int32).reshape(-1, 3)
    rc = lib().psr_write_ply(os.fsencode(path), v.ctypes.data, v.shape[0], t.ctypes.data, t.shape[0])
    if rc != PSR_OK:
        raise PsrError(rc, "write_ply %s" % path)


def read_ply(path):
    """Reads a binary little-endian PLY with float x y z vertices and
    uchar-counted int32 triangle lists (the layout write_ply produces)."""
    with open(path, "rb") as f:
        data = f.read()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    header = data[:end].decode("ascii").splitlines()
    nv = nt = 0
    for line in header:
        w = line.split()
        if w[:2] == ["element", "vertex"]:
            nv = int(w[2])
        elif w[:2] == ["element", "face"]:
            nt = int(w[2])
    v = np.frombuffer(data, np.float32, 3 * nv, end).reshape(nv, 3)
    rec = np.frombuffer(data, np.uint8, 13 * nt, end + 12 * nv).reshape(nt, 13)
    if nt and not np.all(rec[:, 0] == 3):
        raise ValueError("non-triangle face")
    t = rec[:, 1:].copy().view(np.int32).reshape(nt, 3)
    return v, t
