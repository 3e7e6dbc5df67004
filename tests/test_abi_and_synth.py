"""CPU checks of the C ABI library (loads, exports every declared symbol,
struct layouts) and of the seeded synthetic workloads."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol(pkg):
    header = open(os.path.join(ROOT, "include", "psr.h")).read()
    declared = set(re.findall(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b(psr_\w+)\s*\(", header, re.M))
    assert declared == set(pkg.binding.EXPORTED_SYMBOLS)
    lib = ctypes.CDLL(pkg.binding.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name


def test_struct_layouts_match_header(pkg):
    b = pkg.binding
    assert ctypes.sizeof(b.Points) == 24
    assert ctypes.sizeof(b.Mesh) == 24
    assert ctypes.sizeof(b.Config) == 40
    cfg = b.Config()
    b.lib().psr_default_config(ctypes.byref(cfg))
    assert cfg.depth == 7 and abs(cfg.cg_tol - 1e-5) < 1e-12 and cfg.cg_max_iter == 10000


def test_status_strings(pkg):
    L = pkg.binding.lib()
    assert L.psr_status_string(0) == b"ok"
    assert L.psr_status_string(2) == b"capacity exceeded"


def test_synthetic_workloads_deterministic(pkg):
    s = pkg.synth
    a = s.torus(n=5000)
    b = s.torus(n=5000)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert a[0].flags["C_CONTIGUOUS"] and a[0].dtype == np.float32
    c = a[0].astype(np.float64) - 0.5
    rho = np.hypot(c[:, 0], c[:, 1])
    assert np.abs(np.hypot(rho - 0.3, c[:, 2]) - 0.1).max() < 1e-6
    n = a[1].astype(np.float64)
    assert np.abs(np.linalg.norm(n, axis=1) - 1).max() < 1e-6
    p, q = s.sphere(n=4000)
    r = np.linalg.norm(p.astype(np.float64) - 0.5, axis=1)
    assert abs(r.mean() - 0.4) < 1e-3 and r.std() < 0.003


def test_batch_sizes(pkg):
    clouds = pkg.synth.batch64()
    sizes = [c[0].shape[0] for c in clouds]
    assert len(clouds) == 64 and min(sizes) >= 5000 and max(sizes) <= 150000


def test_cpp_facade_compiles_links_and_layouts_agree(pkg, tmp_path):
    """The C++ facade builds against psr.h and links libpsr.so; the C struct
    sizes and field offsets match the ctypes mirror."""
    import subprocess
    b = pkg.binding
    src = tmp_path / "facade.cpp"
    src.write_text(r'''
#include <cstddef>
#include <cstdio>
#include "psr_facade.hpp"
int main() {
    using R = SparseSurfelFusion::PoissonReconstruction;  // the extra operations compile against psr.h
    auto m1 = &R::WeldMesh; auto m2 = &R::CalculateMeshNormals; auto m3 = &R::CalculatePointCloudNormal;
    auto m4 = &R::WriteMeshPly; auto m5 = &R::GetWeldedMeshVertices;
    (void)m1; (void)m2; (void)m3; (void)m4; (void)m5;
    psr_config c; psr_default_config(&c);
    std::printf("%zu %zu %zu %zu %zu %zu %d %s\n", sizeof(psr_config), sizeof(psr_points), sizeof(psr_mesh),
                sizeof(psr_stats), offsetof(psr_stats, ms_cg_kernel), offsetof(psr_stats, cg_nnz), c.depth,
                psr_status_string(PSR_ERR_INVALID));
    return 0;
}
''')
    exe = tmp_path / "facade"
    libdir = os.path.dirname(b.LIB_PATH)
    subprocess.check_call(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                           "-L", libdir, "-lpsr", "-Wl,-rpath," + libdir, "-Wl,-rpath,/opt/rocm/lib"])
    out = subprocess.check_output([str(exe)]).decode().split()
    assert [int(x) for x in out[:4]] == [ctypes.sizeof(b.Config), ctypes.sizeof(b.Points), ctypes.sizeof(b.Mesh),
                                         ctypes.sizeof(b.Stats)]
    assert int(out[4]) == b.Stats.ms_cg_kernel.offset and int(out[5]) == b.Stats.cg_nnz.offset
    assert out[6] == "7" and " ".join(out[7:]) == b.lib().psr_status_string(1).decode()


def test_post_processing_entry_points_reject_bad_arguments(pkg):
    """Argument checks of the welding, normal, colour and PLY entry points
    (they return before touching the device)."""
    b = pkg.binding
    L = b.lib()
    mesh = b.Mesh(None, None, 5, 0)  # vertices announced, no buffer
    assert L.psr_weld_mesh(None, ctypes.byref(b.Mesh())) == b.PSR_ERR_INVALID
    assert L.psr_mesh_normals(None, ctypes.byref(mesh), None) == b.PSR_ERR_INVALID
    vp = (ctypes.c_float * 3)(0, 0, 0)
    assert L.psr_estimate_normals(None, None, 10, 12, 10, vp, None, None, None) == b.PSR_ERR_INVALID
    assert L.psr_vertex_colors(None, ctypes.byref(mesh), None, None, 10, 12, None) == b.PSR_ERR_INVALID
    assert L.psr_write_ply(None, None, 0, None, 0) == b.PSR_ERR_INVALID
