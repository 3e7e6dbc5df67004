// Normal estimation for an unoriented point cloud on the device: the step
// before the reconstruction, which the reference runs on the host with PCL
// (mesh/PoissonReconstruction.cpp:137-186: NormalEstimationOMP, k = 10
// nearest neighbours from a kd-tree, default viewpoint (0, 0, 0)).  PCL is a
// third-party library absent here; its published algorithm is restated:
//   * the k nearest neighbours of each point, the point itself included
//     (distance 0);
//   * their mean and covariance, the eigenvector of the smallest eigenvalue
//     as the normal and lambda_0 / (lambda_0 + lambda_1 + lambda_2) as the
//     curvature (computePointNormal, solvePlaneParameters);
//   * the normal flipped to face the viewpoint: n = -n when (vp - p) . n < 0
//     (flipNormalTowardsViewpoint).
// Nearest neighbours are exact: points are bucketed in a uniform grid, a point
// searches the (2r+1)^3 cells around its own, and r grows until the k-th
// distance lies within the searched box.  Ties go to the lower point index.
#include "psr_internal.h"
#include "psr_prims.h"

#include <cmath>
#include <stdexcept>

namespace psr {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxK = 16;

struct GridSpec {
    float3 lo;
    float h, invH;
    float eps;  // bound of the rounding in a coordinate's cell assignment
    int G;      // cells per axis
};

__device__ __forceinline__ int3 cell_of(float3 p, const GridSpec& g) {
    return make_int3(min(g.G - 1, max(0, int((p.x - g.lo.x) * g.invH))), min(g.G - 1, max(0, int((p.y - g.lo.y) * g.invH))),
                     min(g.G - 1, max(0, int((p.z - g.lo.z) * g.invH))));
}

__device__ __forceinline__ float3 load_point(const char* xyz, unsigned stride, int i) {
    const float* p = reinterpret_cast<const float*>(xyz + size_t(i) * stride);
    return make_float3(p[0], p[1], p[2]);
}

// Per-block bounds of the points.
__global__ void __launch_bounds__(kThreads) ne_bounds(const char* __restrict__ xyz, unsigned stride, int n,
                                                      float* __restrict__ part) {
    __shared__ float s[6][kThreads];
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
        const float3 p = load_point(xyz, stride, i);
        mn[0] = fminf(mn[0], p.x), mn[1] = fminf(mn[1], p.y), mn[2] = fminf(mn[2], p.z);
        mx[0] = fmaxf(mx[0], p.x), mx[1] = fmaxf(mx[1], p.y), mx[2] = fmaxf(mx[2], p.z);
    }
    for (int a = 0; a < 3; ++a) s[a][threadIdx.x] = mn[a], s[3 + a][threadIdx.x] = mx[a];
    __syncthreads();
    for (int st = kThreads / 2; st > 0; st >>= 1) {
        if (int(threadIdx.x) < st)
            for (int a = 0; a < 3; ++a) {
                s[a][threadIdx.x] = fminf(s[a][threadIdx.x], s[a][threadIdx.x + st]);
                s[3 + a][threadIdx.x] = fmaxf(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + st]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

// Grid cell (x-major linear index) of every point, and the identity values.
__global__ void __launch_bounds__(kThreads) ne_cells(const char* __restrict__ xyz, unsigned stride, int n, GridSpec g,
                                                     unsigned* __restrict__ key, unsigned* __restrict__ val) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const int3 c = cell_of(load_point(xyz, stride, i), g);
    key[i] = (unsigned(c.x) * unsigned(g.G) + unsigned(c.y)) * unsigned(g.G) + unsigned(c.z);
    val[i] = unsigned(i);
}

// start[c] = first sorted position whose cell is >= c, for c in [0, G^3].
__global__ void __launch_bounds__(kThreads) ne_cell_starts(const unsigned* __restrict__ key, int n, unsigned cells,
                                                           int* __restrict__ start) {
    const size_t c = size_t(blockIdx.x) * kThreads + threadIdx.x;
    if (c > cells) return;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (key[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    start[c] = lo;
}

__device__ __forceinline__ bool closer(float d, int i, float e, int j) { return d < e || (d == e && i < j); }

// Cyclic Jacobi on a symmetric 3 x 3 matrix: eigenvalues on the diagonal,
// eigenvectors in the columns of V.
__device__ void jacobi3(double (&A)[3][3], double (&V)[3][3]) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) V[i][j] = i == j ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 32; ++sweep) {
        const double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
        const double diag = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
        if (off <= 1e-30 * diag || off == 0.0) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (A[p][q] == 0.0) continue;
                const double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
                for (int k = 0; k < 3; ++k) {  // A <- A J
                    const double akp = A[k][p], akq = A[k][q];
                    A[k][p] = cs * akp - sn * akq;
                    A[k][q] = sn * akp + cs * akq;
                }
                for (int k = 0; k < 3; ++k) {  // A <- J^T A
                    const double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = cs * apk - sn * aqk;
                    A[q][k] = sn * apk + cs * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    const double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = cs * vkp - sn * vkq;
                    V[k][q] = sn * vkp + cs * vkq;
                }
            }
    }
}

// Exact k nearest points of p among the n grid points (top kMaxK kept sorted
// in registers, ties to the lower index): the (2r+1)^3 cells around p's cell,
// r growing until the k-th distance lies inside the searched box.  p may lie
// anywhere (its cell is clamped into the grid).
__device__ __forceinline__ void knn_search(float3 p, const char* __restrict__ xyz, unsigned stride, int kk,
                                           const GridSpec& g, const unsigned* __restrict__ order,
                                           const int* __restrict__ start, float (&dist)[kMaxK], int (&idx)[kMaxK]) {
    const int3 c = cell_of(p, g);
    for (int r = 1;; ++r) {
#pragma unroll
        for (int j = 0; j < kMaxK; ++j) dist[j] = INFINITY, idx[j] = 0x7fffffff;
        const int x0 = max(0, c.x - r), x1 = min(g.G - 1, c.x + r), y0 = max(0, c.y - r), y1 = min(g.G - 1, c.y + r),
                  z0 = max(0, c.z - r), z1 = min(g.G - 1, c.z + r);
        for (int x = x0; x <= x1; ++x)
            for (int y = y0; y <= y1; ++y) {
                const unsigned row = (unsigned(x) * unsigned(g.G) + unsigned(y)) * unsigned(g.G);
                const int b = start[row + unsigned(z0)], e = start[row + unsigned(z1) + 1];
                for (int s = b; s < e; ++s) {
                    const int q = int(order[s]);
                    const float3 o = load_point(xyz, stride, q);
                    const float dx = o.x - p.x, dy = o.y - p.y, dz = o.z - p.z;
                    const float d = dx * dx + dy * dy + dz * dz;
                    if (!closer(d, q, dist[kMaxK - 1], idx[kMaxK - 1])) continue;
                    bool placed = false;
#pragma unroll
                    for (int j = kMaxK - 1; j >= 0; --j) {
                        if (placed) continue;
                        if (j > 0 && closer(d, q, dist[j - 1], idx[j - 1])) {
                            dist[j] = dist[j - 1];
                            idx[j] = idx[j - 1];
                        } else {
                            dist[j] = d;
                            idx[j] = q;
                            placed = true;
                        }
                    }
                }
            }
        // every point outside the box lies at least cover - eps from p (eps:
        // the rounding of the cell assignment), so the k found are the nearest
        const bool all = x0 == 0 && y0 == 0 && z0 == 0 && x1 == g.G - 1 && y1 == g.G - 1 && z1 == g.G - 1;
        float cover = INFINITY;
        if (x0 > 0) cover = fminf(cover, p.x - (g.lo.x + x0 * g.h));
        if (y0 > 0) cover = fminf(cover, p.y - (g.lo.y + y0 * g.h));
        if (z0 > 0) cover = fminf(cover, p.z - (g.lo.z + z0 * g.h));
        if (x1 < g.G - 1) cover = fminf(cover, g.lo.x + (x1 + 1) * g.h - p.x);
        if (y1 < g.G - 1) cover = fminf(cover, g.lo.y + (y1 + 1) * g.h - p.y);
        if (z1 < g.G - 1) cover = fminf(cover, g.lo.z + (z1 + 1) * g.h - p.z);
        float dk = INFINITY;
#pragma unroll
        for (int j = 0; j < kMaxK; ++j)
            if (j == kk - 1) dk = dist[j];
        if (all || (dk < INFINITY && sqrtf(dk) * (1.0f + 1e-6f) + g.eps < cover)) return;
    }
}

// One thread per point: its k nearest neighbours, then the plane fit.
__global__ void __launch_bounds__(kThreads) ne_estimate(const char* __restrict__ xyz, unsigned stride, int n, int k,
                                                        GridSpec g, const unsigned* __restrict__ order,
                                                        const int* __restrict__ start, float3 vp,
                                                        float* __restrict__ normals, float* __restrict__ curvature,
                                                        int* __restrict__ knn) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const float3 p = load_point(xyz, stride, i);
    float dist[kMaxK];
    int idx[kMaxK];
    const int kk = min(k, n);
    knn_search(p, xyz, stride, kk, g, order, start, dist, idx);
    // mean and covariance of the neighbours in FP64
    double m[3] = {0, 0, 0};
    float3 nb[kMaxK];
#pragma unroll
    for (int j = 0; j < kMaxK; ++j) {
        if (j >= kk) continue;
        nb[j] = load_point(xyz, stride, idx[j]);
        m[0] += nb[j].x, m[1] += nb[j].y, m[2] += nb[j].z;
        if (knn) knn[size_t(i) * k + j] = idx[j];
    }
    for (int a = 0; a < 3; ++a) m[a] /= kk;
    double A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
    for (int j = 0; j < kMaxK; ++j) {
        if (j >= kk) continue;
        const double d[3] = {nb[j].x - m[0], nb[j].y - m[1], nb[j].z - m[2]};
        for (int a = 0; a < 3; ++a)
            for (int b = a; b < 3; ++b) A[a][b] += d[a] * d[b];
    }
    for (int a = 0; a < 3; ++a)
        for (int b = a; b < 3; ++b) A[a][b] /= kk, A[b][a] = A[a][b];
    double V[3][3];
    jacobi3(A, V);
    int s = 0;
    if (A[1][1] < A[s][s]) s = 1;
    if (A[2][2] < A[s][s]) s = 2;
    double nx = V[0][s], ny = V[1][s], nz = V[2][s];
    const double len = sqrt(nx * nx + ny * ny + nz * nz);
    nx /= len, ny /= len, nz /= len;
    if ((vp.x - p.x) * nx + (vp.y - p.y) * ny + (vp.z - p.z) * nz < 0) nx = -nx, ny = -ny, nz = -nz;
    normals[3 * size_t(i)] = float(nx);
    normals[3 * size_t(i) + 1] = float(ny);
    normals[3 * size_t(i) + 2] = float(nz);
    if (curvature) {
        const double tr = A[0][0] + A[1][1] + A[2][2];
        curvature[i] = tr != 0.0 ? float(A[s][s] / tr) : 0.f;
    }
}

// --- vertex colours (DrawMesh.cu:35-79, 179-211) ---
// The colour of a mesh vertex from its 4 nearest points, weighted as the
// reference weights them: each squared distance over their sum (the farther
// point weighs more), colour = sum of weight x point colour.  Where all four
// distances are zero the weights are equal (the reference divides by zero).
__global__ void __launch_bounds__(kThreads) vc_colors(const float* __restrict__ V, int nv,
                                                      const char* __restrict__ xyz, unsigned stride, int n,
                                                      const float* __restrict__ pc, GridSpec g,
                                                      const unsigned* __restrict__ order,
                                                      const int* __restrict__ start, float* __restrict__ out) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= nv) return;
    const float3 p = make_float3(V[3 * size_t(i)], V[3 * size_t(i) + 1], V[3 * size_t(i) + 2]);
    float dist[kMaxK];
    int idx[kMaxK];
    const int kk = min(4, n);
    knn_search(p, xyz, stride, kk, g, order, start, dist, idx);
    float d[4] = {0.f, 0.f, 0.f, 0.f}, sum = 0.f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (j >= kk) continue;
        const float3 q = load_point(xyz, stride, idx[j]);
        const float tx = __fsub_rn(p.x, q.x), ty = __fsub_rn(p.y, q.y), tz = __fsub_rn(p.z, q.z);
        d[j] = __fadd_rn(__fadd_rn(__fmul_rn(tx, tx), __fmul_rn(ty, ty)), __fmul_rn(tz, tz));
        sum = __fadd_rn(sum, d[j]);
    }
    float w[4];
    const float inv = sum > 0.f ? __fdiv_rn(1.0f, sum) : 0.f;
#pragma unroll
    for (int j = 0; j < 4; ++j) w[j] = j >= kk ? 0.f : (sum > 0.f ? __fmul_rn(d[j], inv) : __fdiv_rn(1.0f, float(kk)));
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float acc = 0.f;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (j < kk) acc = __fadd_rn(acc, __fmul_rn(pc[3 * size_t(idx[j]) + a], w[j]));
        out[3 * size_t(i) + a] = acc;
    }
}

// Grid of the n points for exact nearest-neighbour searches: cell size from
// about k points per occupied surface cell; points sorted by cell (order) and
// each cell's first sorted position (start, G^3 + 1 entries).
void build_knn_grid(Ctx& c, const char* base, unsigned stride, int n, int k, GridSpec& g, unsigned*& order,
                    int*& start) {
    hipStream_t s = c.stream;
    const int nb = std::min<int>(128, int(div_up(n, kThreads)));
    float* part = c.buf<float>("ne_bounds", size_t(6) * nb);
    ne_bounds<<<nb, kThreads, 0, s>>>(base, stride, n, part);
    std::vector<float> hp(size_t(6) * nb);
    PSR_HIP_CHECK(hipMemcpyAsync(hp.data(), part, sizeof(float) * hp.size(), hipMemcpyDeviceToHost, s));
    c.wait();
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int b = 0; b < nb; ++b)
        for (int a = 0; a < 3; ++a) lo[a] = std::min(lo[a], hp[6 * b + a]), hi[a] = std::max(hi[a], hp[6 * b + 3 + a]);
    for (int a = 0; a < 3; ++a)
        if (!std::isfinite(lo[a]) || !std::isfinite(hi[a])) throw std::invalid_argument("non-finite point coordinates");
    const float ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-30f});
    // about k points per occupied cell on a surface: 4^L cells of a level meet it
    int L = int(std::lround(std::log(std::max(1.0, double(n) / double(k))) / std::log(4.0)));
    L = std::min(8, std::max(1, L));
    g.G = 1 << L;
    g.h = ext * 1.0001f / float(g.G);
    g.invH = 1.0f / g.h;
    g.lo = make_float3(lo[0], lo[1], lo[2]);
    const float mag = std::max({std::fabs(lo[0]), std::fabs(lo[1]), std::fabs(lo[2]), std::fabs(hi[0]),
                                std::fabs(hi[1]), std::fabs(hi[2])});
    g.eps = 8e-7f * (mag + ext);
    const unsigned cells = unsigned(g.G) * unsigned(g.G) * unsigned(g.G);
    unsigned* key = c.buf<unsigned>("ne_key", n);
    unsigned* val = c.buf<unsigned>("ne_val", n);
    unsigned* kt = c.buf<unsigned>("ne_kt", n);
    unsigned* vt = c.buf<unsigned>("ne_vt", n);
    start = c.buf<int>("ne_start", size_t(cells) + 1);
    ne_cells<<<div_up(n, kThreads), kThreads, 0, s>>>(base, stride, n, g, key, val);
    int* scratch = c.scratch_for(sort_scratch_ints(size_t(n)));
    radix_sort_pairs(key, val, kt, vt, n, 3 * L, scratch, false, c.scan_for(sort_scan_elems(size_t(n))), SortGather{},
                     s);
    ne_cell_starts<<<div_up(size_t(cells) + 1, kThreads), kThreads, 0, s>>>(key, n, cells, start);
    order = val;
}

}  // namespace

void estimate_normals(Ctx& c, const float* xyz, int n, unsigned stride, int k, float3 vp, float* normals,
                      float* curvature, int* knn) {
    if (k < 1 || k > kMaxK) throw std::invalid_argument("k must lie in 1..16");
    if (n <= 0) return;
    const char* base = reinterpret_cast<const char*>(xyz);
    GridSpec g;
    unsigned* order = nullptr;
    int* start = nullptr;
    build_knn_grid(c, base, stride, n, k, g, order, start);
    ne_estimate<<<div_up(n, kThreads), kThreads, 0, c.stream>>>(base, stride, n, k, g, order, start, vp, normals,
                                                                 curvature, knn);
    c.wait();
    PSR_HIP_CHECK(hipGetLastError());
}

void vertex_colors(Ctx& c, const float* V, int nv, const float* xyz, unsigned stride, const float* pc, int n,
                   float* out) {
    if (nv <= 0) return;
    if (n <= 0) throw std::invalid_argument("no points to colour from");
    const char* base = reinterpret_cast<const char*>(xyz);
    GridSpec g;
    unsigned* order = nullptr;
    int* start = nullptr;
    build_knn_grid(c, base, stride, n, 10, g, order, start);  // cells sized as for 10 (the search is exact for 4)
    vc_colors<<<div_up(nv, kThreads), kThreads, 0, c.stream>>>(V, nv, base, stride, n, pc, g, order, start, out);
    c.wait();
    PSR_HIP_CHECK(hipGetLastError());
}

}  // namespace psr
