"""psr_mesh_normals_static (no allocation, scratch in a file-scope device
array) against psr_mesh_normals on the same meshes, through the C ABI: two
calls in a row, the second on a smaller mesh, each bit-identical to the
allocating entry point (same kernels, same order); a mesh beyond the static
capacity is refused before any launch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available(), "GPU tests need a GPU"
    return t


def _mesh_of(pkg, torch, name):
    xyz, nrm, d = pkg.synth.workload(name)
    rec = pkg.PoissonReconstruction(depth=d)
    m = rec.SolvePoissionReconstructionMesh(torch.from_numpy(xyz).cuda(), torch.from_numpy(nrm).cuda())
    return rec, m


def test_static_normals_equal_allocating_entry_point(pkg, torch):
    b = pkg.binding
    L = b.lib()
    cap = L.psr_mesh_normals_static_max_triangles()
    for name in ("torus", "sphere"):  # the second call on a smaller mesh
        rec, m = _mesh_of(pkg, torch, name)
        print(name, "vertices", m.n_vertices, "triangles", m.n_triangles, "capacity", cap, flush=True)
        assert m.n_triangles <= cap
        want = rec.MeshNormals()
        out = torch.full((m.n_vertices, 3), float("nan"), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        rc = L.psr_mesh_normals_static(ctypes.byref(m), out.data_ptr(), None)
        assert rc == b.PSR_OK, rc
        np.testing.assert_array_equal(out.cpu().numpy(), want)
        rec.close()


def test_static_normals_refuse_meshes_beyond_capacity(pkg, torch):
    b = pkg.binding
    L = b.lib()
    cap = L.psr_mesh_normals_static_max_triangles()
    v = torch.zeros((3, 3), dtype=torch.float32, device="cuda")
    t = torch.zeros((1, 3), dtype=torch.int32, device="cuda")
    big = b.Mesh(v.data_ptr(), t.data_ptr(), 3, cap + 1)  # announced, refused before any launch
    out = torch.zeros((3, 3), dtype=torch.float32, device="cuda")
    assert L.psr_mesh_normals_static(ctypes.byref(big), out.data_ptr(), None) == b.PSR_ERR_CAPACITY
