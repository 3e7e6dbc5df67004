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
#include <mutex>
#include <stdexcept>
#include <string>
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

// --- per-vertex normals (DrawMesh.cu:80-177) ---
// Unit normal of triangle t, cross(A - O, B - O) with (O, A, B) its corners,
// or nothing for a triangle with a corner outside the mesh; the (vertex,
// corner) pairs for the grouping sort (key nv: no vertex).  Also clears the
// output, which the vertex kernel writes only where a vertex has triangles.
__device__ __forceinline__ float3 unit_or_zero(float x, float y, float z) {
    // VectorNormalize (DrawMesh.cu:80-91): zero stays zero
    const float len = __fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(x, x), __fmul_rn(y, y)), __fmul_rn(z, z)));
    if (len == 0.f) return make_float3(0.f, 0.f, 0.f);
    return make_float3(__fdiv_rn(x, len), __fdiv_rn(y, len), __fdiv_rn(z, len));
}

__global__ void __launch_bounds__(kThreads) normals_faces(const float* __restrict__ V, int nv,
                                                          const int* __restrict__ T, int nt,
                                                          float3* __restrict__ fn, unsigned* __restrict__ key,
                                                          unsigned* __restrict__ val, float* __restrict__ out) {
    const size_t i = size_t(blockIdx.x) * kThreads + threadIdx.x;
    for (size_t z = i; z < 3 * size_t(nv); z += size_t(gridDim.x) * kThreads) out[z] = 0.f;
    if (i >= size_t(nt)) return;
    const int c0 = T[3 * i], c1 = T[3 * i + 1], c2 = T[3 * i + 2];
    const bool ok = c0 >= 0 && c0 < nv && c1 >= 0 && c1 < nv && c2 >= 0 && c2 < nv;
    if (ok) {
        const float ox = V[3 * size_t(c0)], oy = V[3 * size_t(c0) + 1], oz = V[3 * size_t(c0) + 2];
        const float ax = __fsub_rn(V[3 * size_t(c1)], ox), ay = __fsub_rn(V[3 * size_t(c1) + 1], oy),
                    az = __fsub_rn(V[3 * size_t(c1) + 2], oz);
        const float bx = __fsub_rn(V[3 * size_t(c2)], ox), by = __fsub_rn(V[3 * size_t(c2) + 1], oy),
                    bz = __fsub_rn(V[3 * size_t(c2) + 2], oz);
        // CrossProduct (DrawMesh.cu:93-100)
        fn[i] = unit_or_zero(__fsub_rn(__fmul_rn(ay, bz), __fmul_rn(az, by)),
                             __fsub_rn(__fmul_rn(az, bx), __fmul_rn(ax, bz)),
                             __fsub_rn(__fmul_rn(ax, by), __fmul_rn(ay, bx)));
    }
    key[3 * i] = ok ? unsigned(c0) : unsigned(nv);
    key[3 * i + 1] = ok ? unsigned(c1) : unsigned(nv);
    key[3 * i + 2] = ok ? unsigned(c2) : unsigned(nv);
    val[3 * i] = val[3 * i + 1] = val[3 * i + 2] = unsigned(i);
}

// The first pair of each vertex's run (pairs sorted stably by vertex, so in
// triangle order) sums the run's face normals, divides by the count and
// normalises (VerticesNormalsSumKernel, CalculateVerticesAverageNormals).
__global__ void __launch_bounds__(kThreads) normals_vertices(const unsigned* __restrict__ key,
                                                             const unsigned* __restrict__ val, size_t n, int nv,
                                                             const float3* __restrict__ fn, float* __restrict__ out) {
    const size_t i = size_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n) return;
    const unsigned k = key[i];
    if (k >= unsigned(nv) || (i > 0 && key[i - 1] == k)) return;
    float sx = 0.f, sy = 0.f, sz = 0.f;
    int cnt = 0;
    for (size_t j = i; j < n && key[j] == k; ++j, ++cnt) {
        const float3 f = fn[val[j]];
        sx = __fadd_rn(sx, f.x);
        sy = __fadd_rn(sy, f.y);
        sz = __fadd_rn(sz, f.z);
    }
    const float fc = float(cnt);
    const float3 u = unit_or_zero(__fdiv_rn(sx, fc), __fdiv_rn(sy, fc), __fdiv_rn(sz, fc));
    out[3 * size_t(k)] = u.x;
    out[3 * size_t(k) + 1] = u.y;
    out[3 * size_t(k) + 2] = u.z;
}

}  // namespace

void mesh_normals(Ctx& c, const float* V, int nv, const int* T, int nt, float* out) {
    hipStream_t s = c.stream;
    if (nv <= 0) return;
    const size_t n = size_t(3) * std::max(nt, 0);
    float3* fn = c.buf<float3>("nrm_faces", std::max(nt, 1));
    unsigned* k = c.buf<unsigned>("nrm_k", std::max<size_t>(n, 1));
    unsigned* v = c.buf<unsigned>("nrm_v", std::max<size_t>(n, 1));
    unsigned* kt = c.buf<unsigned>("nrm_kt", std::max<size_t>(n, 1));
    unsigned* vt = c.buf<unsigned>("nrm_vt", std::max<size_t>(n, 1));
    const unsigned grid = std::max(div_up(size_t(std::max(nt, 1)), kThreads), div_up(3 * size_t(nv), kThreads * 8));
    normals_faces<<<grid, kThreads, 0, s>>>(V, nv, T, nt, fn, k, v, out);
    if (n > 0) {
        int bits = 1;
        while (bits < 32 && (unsigned(nv) >> bits) != 0) ++bits;  // keys 0..nv
        int* scratch = c.scratch_for(sort_scratch_ints(n));
        radix_sort_pairs(k, v, kt, vt, int(n), bits, scratch, false, c.scan_for(sort_scan_elems(n)), SortGather{}, s);
        normals_vertices<<<div_up(n, kThreads), kThreads, 0, s>>>(k, v, n, nv, fn, out);
    }
    c.wait();
    PSR_HIP_CHECK(hipGetLastError());
}

// ---- per-vertex normals with every scratch buffer in one file-scope device
// array (psr_mesh_normals_static): for callers that cannot allocate.  The
// array holds the face normals, the four sort buffers, the sort's histograms,
// the scan state and an error word for meshes of up to kStaticTriangles
// triangles; the call clears what must start at zero.  Every call on a device
// shares that device's array: calls that overlap on one device race.
namespace {

constexpr size_t kStaticTriangles = size_t(1) << 18;
constexpr size_t kStaticVertices = size_t(1) << 24;
constexpr size_t kStaticCorners = 3 * kStaticTriangles;
constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kStaticSortTiles = (kStaticCorners + kSortTileKeys - 1) / kSortTileKeys;
constexpr size_t kStaticHistInts = 256 * kStaticSortTiles + 64;      // sort_scratch_ints(kStaticCorners)
constexpr size_t kStaticScanTiles = (256 * kStaticSortTiles + 2047) / 2048;  // tiles of the histogram scan
constexpr size_t kOffFaces = 0;
constexpr size_t kOffKeys = kOffFaces + align256(sizeof(float3) * kStaticTriangles);
constexpr size_t kOffVals = kOffKeys + align256(4 * kStaticCorners);
constexpr size_t kOffKeysTmp = kOffVals + align256(4 * kStaticCorners);
constexpr size_t kOffValsTmp = kOffKeysTmp + align256(4 * kStaticCorners);
constexpr size_t kOffHist = kOffValsTmp + align256(4 * kStaticCorners);
constexpr size_t kOffScanStatus = kOffHist + align256(4 * kStaticHistInts);
constexpr size_t kOffScanCounters = kOffScanStatus + align256(8 * kStaticScanTiles);
constexpr size_t kOffError = kOffScanCounters + align256(4 * kScanRing);
constexpr size_t kStaticBytes = kOffError + 256;

static __device__ __align__(256) unsigned char scratch[kStaticBytes];

// The device address of `scratch` on each device, looked up once per device
// with that device current (the array's name is only a host-side stand-in).
constexpr int kMaxDevices = 64;
std::mutex& static_mutex() {
    static std::mutex m;
    return m;
}
void* g_static_scratch[kMaxDevices] = {};

void* static_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) throw std::invalid_argument("device index out of range");
    std::lock_guard<std::mutex> lk(static_mutex());
    if (!g_static_scratch[device]) {
        void* p = nullptr;
        PSR_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_static_scratch[device] = p;
    }
    return g_static_scratch[device];
}

}  // namespace

size_t mesh_normals_static_max_triangles() { return kStaticTriangles; }

int mesh_normals_static(const float* V, int nv, const int* T, int nt, float* out, hipStream_t s, std::string& err) {
    if (size_t(nt) > kStaticTriangles) return err = "more triangles than the static scratch holds", PSR_ERR_CAPACITY;
    if (size_t(nv) >= kStaticVertices) return err = "more vertices than the static scratch holds", PSR_ERR_CAPACITY;
    if (nv <= 0) return PSR_OK;
    int device = 0;
    PSR_HIP_CHECK(hipGetDevice(&device));
    char* base = static_cast<char*>(static_scratch(device));
    const size_t n = size_t(3) * std::max(nt, 0);
    if (n > 0 && (sort_scratch_ints(n) > kStaticHistInts || scan_tiles(sort_scan_elems(n)) > kStaticScanTiles))
        return err = "static scratch layout does not fit the sort", PSR_ERR_INTERNAL;
    float3* fn = reinterpret_cast<float3*>(base + kOffFaces);
    unsigned* k = reinterpret_cast<unsigned*>(base + kOffKeys);
    unsigned* v = reinterpret_cast<unsigned*>(base + kOffVals);
    unsigned* kt = reinterpret_cast<unsigned*>(base + kOffKeysTmp);
    unsigned* vt = reinterpret_cast<unsigned*>(base + kOffValsTmp);
    ScanState st;
    st.status = reinterpret_cast<unsigned long long*>(base + kOffScanStatus);
    st.counters = reinterpret_cast<unsigned*>(base + kOffScanCounters);
    st.capacity = kStaticScanTiles;
    st.epoch = 0;
    int* errWord = reinterpret_cast<int*>(base + kOffError);
    st.err = errWord;
    // the array keeps the last call's contents: the scan's status words, tile
    // counters and the error word start at zero (a fill kernel, no memset node)
    zero_spans({{st.status, 2 * kStaticScanTiles}, {st.counters, size_t(kScanRing)}, {errWord, 1}}, s);
    const unsigned grid = std::max(div_up(size_t(std::max(nt, 1)), kThreads), div_up(3 * size_t(nv), kThreads * 8));
    normals_faces<<<grid, kThreads, 0, s>>>(V, nv, T, nt, fn, k, v, out);
    if (n > 0) {
        int bits = 1;
        while (bits < 32 && (unsigned(nv) >> bits) != 0) ++bits;  // keys 0..nv
        radix_sort_pairs(k, v, kt, vt, int(n), bits, reinterpret_cast<int*>(base + kOffHist), false, st, SortGather{},
                         s);
        normals_vertices<<<div_up(n, kThreads), kThreads, 0, s>>>(k, v, n, nv, fn, out);
    }
    PSR_HIP_CHECK(hipGetLastError());
    int failed = 0;
    PSR_HIP_CHECK(hipMemcpyAsync(&failed, errWord, sizeof(int), hipMemcpyDeviceToHost, s));
    PSR_HIP_CHECK(hipStreamSynchronize(s));
    if (failed) return err = "scan look-back timed out", PSR_ERR_INTERNAL;
    return PSR_OK;
}

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
