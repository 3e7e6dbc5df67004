// Mesh welding on the device, and the binary PLY writer.
//
// The reference emits one vertex per marching-cubes batch and edge, so a
// surface crossing shared by two batches appears twice
// (ComputeTriangleIndices.cu:1247-1282 indexes each batch's vertices on their
// own).  Every vertex of this build's mesh carries the global key of its
// depth-D edge (segment_key); welding keeps one vertex per key:
//   * (key, vertex) pairs sorted by key (stable: equal keys keep vertex order);
//   * head flags where the key changes, their exclusive scan = welded index;
//   * the first vertex of each run (the lowest-numbered vertex on that edge)
//     is kept, every vertex maps to its run's welded index, and triangle
//     corners are remapped through that map.
// Welded vertices come out in ascending key order.
#include "psr_internal.h"
#include "psr_mesh_common.h"
#include "psr_prims.h"

#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace psr {
namespace {

constexpr int kThreads = 256;
constexpr int kMailWeld = 44;  // mailbox slot of the welded vertex count
static_assert(kMailWeld > kMailGeo + 2 && kMailWeld < kMailWords, "mailbox slot of the weld");

__global__ void __launch_bounds__(kThreads) weld_init(const unsigned* __restrict__ key, int n,
                                                      unsigned* __restrict__ k, unsigned* __restrict__ v) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    k[i] = key[i];
    v[i] = unsigned(i);
}

__global__ void __launch_bounds__(kThreads) weld_heads(const unsigned* __restrict__ k, int n, int* __restrict__ head) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    head[i] = (i == 0 || k[i] != k[i - 1]) ? 1 : 0;
}

// Sorted position i: vertex v[i] maps to run grp[i] + head[i] - 1; a run's
// first vertex is copied to it.
__global__ void __launch_bounds__(kThreads) weld_vertices(const unsigned* __restrict__ v, const int* __restrict__ head,
                                                          const int* __restrict__ grp, int n,
                                                          const float* __restrict__ inV, int* __restrict__ remap,
                                                          float* __restrict__ outV) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const int h = head[i], g = grp[i] + h - 1, src = int(v[i]);
    remap[src] = g;
    if (h) {
        outV[3 * size_t(g)] = inV[3 * size_t(src)];
        outV[3 * size_t(g) + 1] = inV[3 * size_t(src) + 1];
        outV[3 * size_t(g) + 2] = inV[3 * size_t(src) + 2];
    }
}

// Triangle corners through the map; a corner outside [0, nv) (the reference
// does not re-index triangles around dropped vertices) becomes -1.
__global__ void __launch_bounds__(kThreads) weld_triangles(const int* __restrict__ inT, size_t n3, int nv,
                                                           const int* __restrict__ remap, int* __restrict__ outT) {
    const size_t i = size_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n3) return;
    const int a = inT[i];
    outT[i] = a >= 0 && a < nv ? remap[a] : -1;
}

}  // namespace

void weld_mesh(Ctx& c) {
    Mesh& m = c.mesh;
    hipStream_t s = c.stream;
    const int nv = m.nv, nt = m.nt;
    m.weldedVertices = c.buf<float>("weld_vertices", size_t(3) * std::max(nv, 1));
    m.weldedTriangles = c.buf<int>("weld_triangles", size_t(3) * std::max(nt, 1));
    m.weldedNv = 0;
    m.weldedNt = nt;
    if (nv > 0) {
        unsigned* k = c.buf<unsigned>("weld_k", nv);
        unsigned* v = c.buf<unsigned>("weld_v", nv);
        unsigned* kt = c.buf<unsigned>("weld_kt", nv);
        unsigned* vt = c.buf<unsigned>("weld_vt", nv);
        int* head = c.buf<int>("weld_head", nv);
        int* grp = c.buf<int>("weld_grp", nv);
        int* remap = c.buf<int>("weld_remap", nv);
        weld_init<<<div_up(nv, kThreads), kThreads, 0, s>>>(m.keys, nv, k, v);
        int* scratch = c.scratch_for(sort_scratch_ints(size_t(nv)));
        radix_sort_pairs(k, v, kt, vt, nv, 32, scratch, false, c.scan_for(sort_scan_elems(size_t(nv))), SortGather{},
                         s);
        weld_heads<<<div_up(nv, kThreads), kThreads, 0, s>>>(k, nv, head);
        c.hostPinned[kMailWeld] = 0;
        scan_exclusive(head, grp, nv, c.mail + kMailWeld, c.scan_for(size_t(nv)), s);
        weld_vertices<<<div_up(nv, kThreads), kThreads, 0, s>>>(v, head, grp, nv, m.vertices, remap,
                                                                m.weldedVertices);
        if (nt > 0)
            weld_triangles<<<div_up(size_t(3) * nt, kThreads), kThreads, 0, s>>>(m.triangles, size_t(3) * nt, nv,
                                                                                 remap, m.weldedTriangles);
        c.wait();
        m.weldedNv = c.hostPinned[kMailWeld];  // the scan's total: the run count
    }
    PSR_HIP_CHECK(hipGetLastError());
}

}  // namespace psr

// Binary little-endian PLY: float x y z per vertex, a uchar-counted int list
// per triangle.
int psr_write_ply(const char* path, const float* vertices, uint32_t n_vertices, const int32_t* triangles,
                  uint32_t n_triangles) {
    if (!path || (n_vertices && !vertices) || (n_triangles && !triangles)) return PSR_ERR_INVALID;
    FILE* f = std::fopen(path, "wb");
    if (!f) return PSR_ERR_INVALID;
    std::fprintf(f,
                 "ply\nformat binary_little_endian 1.0\ncomment libpsr mesh\nelement vertex %u\n"
                 "property float x\nproperty float y\nproperty float z\nelement face %u\n"
                 "property list uchar int vertex_indices\nend_header\n",
                 n_vertices, n_triangles);
    bool ok = std::fwrite(vertices, sizeof(float) * 3, n_vertices, f) == n_vertices;
    std::vector<unsigned char> rec(13 * size_t(1024));
    for (uint32_t t0 = 0; ok && t0 < n_triangles; t0 += 1024) {
        const uint32_t cnt = std::min<uint32_t>(1024, n_triangles - t0);
        for (uint32_t t = 0; t < cnt; ++t) {
            rec[13 * t] = 3;
            std::memcpy(&rec[13 * t + 1], triangles + 3 * size_t(t0 + t), 12);
        }
        ok = std::fwrite(rec.data(), 13, cnt, f) == cnt;
    }
    ok = (std::fclose(f) == 0) && ok;
    return ok ? PSR_OK : PSR_ERR_INTERNAL;
}
