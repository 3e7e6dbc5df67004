// Mesh geometry (vertex / edge / face ownership), vertex implicit values,
// depth-D marching cubes, face-intersection propagation and subdivision roots.
#include "psr_mesh_common.h"
#include "psr_prims.h"
#include "psr_tile_scan.h"

namespace psr {

namespace {

constexpr int kThreads = 256;  // a tile of the self-scanning kernels is one workgroup
// vertex values: one wave per workgroup, so that the chains of dependent loads
// of a hundred thousand vertices spread over every CU
constexpr int kValueThreads = PSR_WAVE;

// --- ownership (BuildMeshGeometry.cu:42-400) -------------------------------
// A corner / edge / face of node o is owned by the smallest-key neighbour whose
// centre lies within one node width of it; the owner emits it once.

struct NodeView {
    int M, D;
    const int* key;
    const int* depth;
    const int* neigh;
    const float4* basis;  // per node: centre (x, y, z), 2^depth
};

// Node centre (BuildOctree.cu:568-586): the exact dyadic (cell + 1/2) / 2^depth,
// which the basis parameters hold bit for bit.
__device__ __forceinline__ float3 node_center(const NodeView& v, int o) {
    const float4 b = v.basis[o];
    return make_float3(b.x, b.y, b.z);
}

// Centres and keys of the 27 neighbours of o (absent: key INT_MAX), loaded together.
__device__ __forceinline__ void load_neighbourhood(const NodeView& v, int o, int (&nb)[27], float3 (&c)[27],
                                                   int (&k)[27]) {
#pragma unroll
    for (int i = 0; i < 27; ++i) nb[i] = v.neigh[i * v.M + o];
#pragma unroll
    for (int i = 0; i < 27; ++i) {
        const bool ok = nb[i] != -1;
        const float4 b = ok ? v.basis[nb[i]] : make_float4(0.f, 0.f, 0.f, 0.f);
        c[i] = make_float3(b.x, b.y, b.z);
        k[i] = ok ? v.key[nb[i]] : 0x7fffffff;
    }
}

// Per node, the corners (all depths), edges (depth D) and faces (all depths)
// it owns, each kind compacted in (node, slot) order by its own scan (the
// three scans share the launch's tile order), and the node-side tables
// nodeVert / nodeEdge / nodeFace written by the owner for every node sharing
// the element (BuildMeshGeometry.cu:42-400: ownership, then maintenance).
struct GeoOut {
    int* vOwner;
    int* vKind;
    float4* vPos;
    int* nodeVert;
    int* eOwner;
    int* eKind;
    int* nodeEdge;
    int* fOwner;
    int* fKind;
    int* fHasParent;
    int* nodeFace;
    int2* edgeOf;
};

// Ownership on the regular neighbour grid (an octree in which the
// level-local child 0 rule changed no parent, and o its own neighbour 13):
// neighbour i lies at offset (i / 9 - 1, (i / 3) % 3 - 1, i % 3 - 1) widths, so
// the neighbours within one width of a corner (edge, face) centre are exactly
// the 8 (4, 2) sharing it -- every other one is at least sqrt(5) / 2 widths
// away -- and the smallest key among them owns it.  kk: neighbour keys
// (INT_MAX where absent); the result is that of the distance test.
__device__ __forceinline__ void regular_owned(const int (&kk)[27], bool depthD, unsigned& vm, unsigned& em,
                                              unsigned& fm) {
    constexpr int kStride[3] = {9, 3, 1};
    const int self = kk[13];
#pragma unroll
    for (int i = 0; i < 8; ++i) {  // corner side per axis: bit a of i
        int mn = 0x7fffffff;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            mn = min(mn, kk[9 * ((i & 1) + (q & 1)) + 3 * (((i >> 1) & 1) + ((q >> 1) & 1)) +
                            (((i >> 2) & 1) + ((q >> 2) & 1))]);
        vm |= (mn == self ? 1u : 0u) << i;
    }
#pragma unroll
    for (int f = 0; f < 6; ++f)
        fm |= (kk[13 + (2 * (f & 1) - 1) * kStride[f >> 1]] > self ? 1u : 0u) << f;
    if (!depthD) return;
#pragma unroll
    for (int e = 0; e < 12; ++e) {  // axis e >> 2; sides of the other two axes in bits 0, 1
        const int a = e >> 2, u0 = a == 0 ? 1 : 0, u1 = a == 2 ? 1 : 2;
        int mn = 0x7fffffff;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            mn = min(mn, kk[13 + ((e & 1) - 1 + (q & 1)) * kStride[u0] + (((e >> 1) & 1) - 1 + (q >> 1)) * kStride[u1]]);
        em |= (mn == self ? 1u : 0u) << e;
    }
}

// The iso value: the mean implicit value over the points (LaplacianSolver.cu:269-305),
// from the per-block sums of points_chi_block, in one fixed order.
__device__ __forceinline__ void iso_final_body(int psr_bid, int psr_nbid, const double* __restrict__ part, int nb, int N,
                                                   double* __restrict__ isoSum, float* __restrict__ iso) {
    __shared__ double sh[1024 / PSR_WAVE];
    double s = 0;
    for (int k = threadIdx.x; k < nb; k += 1024) s += part[k];
    for (int off = PSR_WAVE / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, PSR_WAVE);
    if (__lane_id() == 0) sh[threadIdx.x / PSR_WAVE] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int w = 0; w < 1024 / PSR_WAVE; ++w) t += sh[w];
        *isoSum = t;
        *iso = float(t / double(N));
    }
}
__global__ void __launch_bounds__(1024) iso_final(const double* __restrict__ part, int nb, int N,
                                                   double* __restrict__ isoSum, float* __restrict__ iso) {
    iso_final_body(blockIdx.x, gridDim.x, part, nb, N, isoSum, iso);
}
PSR_BATCHABLE(iso_final, 1024)


// dense: the dense d_x grids (or null), where a regular octree's nodes put dx.
// Blocks from geoBlocks on compute the points' implicit values instead
// (points_chi_block): independent work that fills the GPU beside this
// kernel's few workgroups.  Compiled for 4 waves per SIMD (128 registers): the
// general ownership path, which only octrees with a changed parent take,
// would hold 229 and keep every block, the points' included, at 2; it spills
// instead.
__device__ __forceinline__ void geometry_emit_body(int psr_bid, int psr_nbid, const NodeView& v, int offD, int nD,
                                                          const int* __restrict__ parent,
                                                          const int* __restrict__ q1Fired, const ScanTicket& tv,
                                                          const ScanTicket& te, const ScanTicket& tf, const GeoOut& out,
                                                          const float* __restrict__ dx, float* __restrict__ dense,
                                                          int geoBlocks, const PointsChiArgs& pc) {
    static_assert(kThreads == kPointsChiThreads, "one block size for both roles");
    if (int(psr_bid) >= geoBlocks) {
        points_chi_block(pc, int(psr_bid) - geoBlocks);
        return;
    }
    const int tile = ticket_tile(tv);
    const int o = tile * kThreads + threadIdx.x;
    unsigned vm = 0, em = 0, fm = 0;
    int depth = 0;
    bool done = false;
    if (o < v.M && *q1Fired == 0) {
        depth = v.depth[o];
        if (dense) dense[dense_grid_index(v.key[o], depth, v.D)] = dx[o];
        int kk[27];
#pragma unroll
        for (int i = 0; i < 27; ++i) {
            const int nb = v.neigh[i * v.M + o];
            kk[i] = nb != -1 ? v.key[nb] : 0x7fffffff;
        }
        if (v.neigh[13 * v.M + o] == o) {
            regular_owned(kk, o >= offD, vm, em, fm);
            done = true;
        }
    }
    if (o < v.M && !done) {
        depth = v.depth[o];
        const float half = 1.0f / float(1 << (depth + 1)), W = 1.0f / float(1 << depth), W2 = W * W;
        int nb[27], kk[27];
        float3 cc[27];
        load_neighbourhood(v, o, nb, cc, kk);
        const float3 nc = cc[13];  // (0, 0, 0) when the node has no self-neighbour
        {
            int ownKey[8], own[8];
            for (int i = 0; i < 8; ++i) {
                ownKey[i] = 0x7fffffff;
                own[i] = -1;
            }
#pragma unroll
            for (int j = 0; j < 27; ++j) {
                if (nb[j] == -1) continue;
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    float3 p = corner_pos(nc, i, half);
                    if (sqdist(p, cc[j]) < double(W2) && ownKey[i] > kk[j]) {
                        ownKey[i] = kk[j];
                        own[i] = nb[j];
                    }
                }
            }
            for (int i = 0; i < 8; ++i) vm |= (own[i] == o ? 1u : 0u) << i;
        }
        {
            int ownKey[6], own[6];
            for (int f = 0; f < 6; ++f) {
                ownKey[f] = 0x7fffffff;
                own[f] = -1;
            }
#pragma unroll
            for (int j = 0; j < 27; ++j) {
                if (nb[j] == -1) continue;
#pragma unroll
                for (int f = 0; f < 6; ++f) {
                    float3 p = face_center(nc, f, half);
                    if (sqdist(p, cc[j]) < double(W2) && ownKey[f] > kk[j]) {
                        ownKey[f] = kk[j];
                        own[f] = nb[j];
                    }
                }
            }
            for (int f = 0; f < 6; ++f) fm |= (own[f] == o ? 1u : 0u) << f;
        }
        if (o >= offD) {  // depth D (the half width and width are those of depth D)
            int ownKey[12], own[12];
            for (int e = 0; e < 12; ++e) {
                ownKey[e] = 0x7fffffff;
                own[e] = -1;
            }
#pragma unroll
            for (int j = 0; j < 27; ++j) {
                if (nb[j] == -1) continue;
#pragma unroll
                for (int e = 0; e < 12; ++e) {
                    float3 p = edge_center(nc, e, half);
                    if (sqdist(p, cc[j]) < double(W2) && ownKey[e] > kk[j]) {
                        ownKey[e] = kk[j];
                        own[e] = nb[j];
                    }
                }
            }
            for (int e = 0; e < 12; ++e) em |= (own[e] == o ? 1u : 0u) << e;
        }
    }
    int av, ae, af;
    ticket_prefix3<kThreads>(tv, te, tf, tile, __popc(vm), __popc(em), __popc(fm), av, ae, af);
    if (o >= offD && o < v.M) out.edgeOf[o - offD] = make_int2(ae, int(em));
    if (!(vm | em | fm)) return;
    const float half = 1.0f / float(1 << (depth + 1)), W = 1.0f / float(1 << depth), W2 = W * W;
    if (done) {
        // regular grid: the sharing neighbours by index, the element's slot in
        // each of them by its side (what the distance and sign tests give)
        constexpr int kStride[3] = {9, 3, 1};
        int nb[27];
#pragma unroll
        for (int i = 0; i < 27; ++i) nb[i] = v.neigh[i * v.M + o];
        const float4 b = v.basis[o];
        const float3 nc = make_float3(b.x, b.y, b.z);
        for (int i = 0; i < 8; ++i) {
            if (!((vm >> i) & 1u)) continue;
            const float3 p = corner_pos(nc, i, half);
            out.vOwner[av] = o;
            out.vKind[av] = i;
            out.vPos[av] = make_float4(p.x, p.y, p.z, __int_as_float(depth));
            ++av;
            const int sx = i & 1, sy = (i >> 1) & 1, sz = (i >> 2) & 1;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int qx = q & 1, qy = (q >> 1) & 1, qz = q >> 2;
                const int j = 9 * (sx + qx) + 3 * (sy + qy) + (sz + qz);
                if (nb[j] == -1) continue;
                // the corner lies on the high side of neighbour j in axis a iff q_a = 0 (Bourke slot)
                int slot = (1 - qx) | ((1 - qz) << 2);
                if (!qy) slot += (slot & 1) ? 1 : 3;
                out.nodeVert[slot * v.M + nb[j]] = av;  // index + 1
            }
        }
        for (int e = 0; e < 12; ++e) {
            if (!((em >> e) & 1u)) continue;
            out.eOwner[ae] = o;
            out.eKind[ae] = e;
            ++ae;
            const int a = e >> 2, u0 = a == 0 ? 1 : 0, u1 = a == 2 ? 1 : 2;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int j = 13 + ((e & 1) - 1 + (q & 1)) * kStride[u0] + (((e >> 1) & 1) - 1 + (q >> 1)) * kStride[u1];
                if (nb[j] == -1) continue;
                out.nodeEdge[((a << 2) | (1 - (q & 1)) | ((1 - (q >> 1)) << 1)) * nD + (nb[j] - offD)] = ae;
            }
        }
        if (fm) {
            const int son = (v.key[o] >> (3 * (v.D - depth))) & 7;
            const bool hasParent = parent[o] != -1;
            for (int f = 0; f < 6; ++f) {
                if (!((fm >> f) & 1u)) continue;
                out.fOwner[af] = o;
                out.fKind[af] = f;
                out.fHasParent[af] = hasParent && kParentFaceKind[son][f] != -1 ? 1 : -1;
                ++af;
                out.nodeFace[f * v.M + o] = af;
                const int j = 13 + (2 * (f & 1) - 1) * kStride[f >> 1];
                if (nb[j] != -1) out.nodeFace[(f ^ 1) * v.M + nb[j]] = af;
            }
        }
        return;
    }
    // emission and maintenance; the neighbourhood is read again (cached)
    int nb[27], kk[27];
    float3 cc[27];
    load_neighbourhood(v, o, nb, cc, kk);
    const float3 nc = cc[13];
    for (int i = 0; i < 8; ++i) {
        if (!((vm >> i) & 1u)) continue;
        const float3 p = corner_pos(nc, i, half);
        out.vOwner[av] = o;
        out.vKind[av] = i;
        out.vPos[av] = make_float4(p.x, p.y, p.z, __int_as_float(depth));
        ++av;
#pragma unroll
        for (int j = 0; j < 27; ++j) {
            if (nb[j] == -1 || !(sqdist(p, cc[j]) < double(W2))) continue;
            out.nodeVert[corner_slot(cc[j], p) * v.M + nb[j]] = av;  // index + 1
        }
    }
    for (int e = 0; e < 12; ++e) {
        if (!((em >> e) & 1u)) continue;
        const float3 p = edge_center(nc, e, half);
        out.eOwner[ae] = o;
        out.eKind[ae] = e;
        ++ae;
#pragma unroll
        for (int j = 0; j < 27; ++j) {
            if (nb[j] == -1 || !(sqdist(p, cc[j]) < double(W2))) continue;
            out.nodeEdge[edge_slot(cc[j], p, e) * nD + (nb[j] - offD)] = ae;
        }
    }
    if (fm) {
        const int son = (v.key[o] >> (3 * (v.D - depth))) & 7;
        const bool hasParent = parent[o] != -1;
        for (int f = 0; f < 6; ++f) {
            if (!((fm >> f) & 1u)) continue;
            const float3 p = face_center(nc, f, half);
            const int orient = f >> 1;
            out.fOwner[af] = o;
            out.fKind[af] = f;
            out.fHasParent[af] = hasParent && kParentFaceKind[son][f] != -1 ? 1 : -1;
            ++af;
#pragma unroll
            for (int j = 0; j < 27; ++j) {
                if (nb[j] == -1 || !(sqdist(p, cc[j]) < double(W2))) continue;
                const float cv = orient == 0 ? cc[j].x : (orient == 1 ? cc[j].y : cc[j].z);
                const float pv = orient == 0 ? p.x : (orient == 1 ? p.y : p.z);
                out.nodeFace[((orient << 1) + (cv - pv < 0 ? 1 : 0)) * v.M + nb[j]] = af;
            }
        }
    }
}
constexpr int kGeometryWaves = 4;
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kGeometryWaves))) geometry_emit(NodeView v, int offD, int nD,
                                                          const int* __restrict__ parent,
                                                          const int* __restrict__ q1Fired, ScanTicket tv,
                                                          ScanTicket te, ScanTicket tf, GeoOut out,
                                                          const float* __restrict__ dx, float* __restrict__ dense,
                                                          int geoBlocks, PointsChiArgs pc) {
    geometry_emit_body(blockIdx.x, gridDim.x, v, offD, nD, parent, q1Fired, tv, te, tf, out, dx, dense, geoBlocks, pc);
}
PSR_BATCHABLE_WAVES(geometry_emit, kThreads, kGeometryWaves)


// --- vertex implicit values (ComputeTriangleIndices.cu:346-399), quirk Q8 --
__device__ __forceinline__ void vertex_values_body(int psr_bid, int psr_nbid, const NodeView& v, const int* __restrict__ count,
                                                          const int* __restrict__ vOwner,
                                                          const int* __restrict__ vKind,
                                                          const float4* __restrict__ vPos,
                                                          const int* __restrict__ parent,
                                                          const int* __restrict__ childBase,
                                                          const float4* __restrict__ nbrCol,
                                                          const float* __restrict__ dCol,
                                                          const int* __restrict__ q1Fired,
                                                          const float* __restrict__ iso, float* __restrict__ vv,
                                                          int* __restrict__ mail, const int* __restrict__ results,
                                                          int* __restrict__ resultsHost, int resultWords) {
    int t = psr_bid * kValueThreads + threadIdx.x;
    if (t < 3) mail[kMailGeo + t] = count[t];  // vertex, edge and face counts for the host
    // the run's device results (CG statistics, iso value) are final here
    if (t < resultWords) resultsHost[t] = results[t];
    if (t >= *count) return;
    int owner = vOwner[t];
    float4 q = vPos[t];
    float3 p = make_float3(q.x, q.y, q.z);
    int depth = __float_as_int(q.w);
    float val = 0.f;
    const bool regular = *q1Fired == 0;
    if (owner > 0) {
        int now = owner;
        for (int d = depth; now != -1; --d) {  // a parent is one level up
            val += regular ? chi_node_sum_regular(now, v.basis[now], p, v.M, dCol)
                           : chi_node_sum(now, float(1 << d), p, v.M, nbrCol);
            now = parent[now];
        }
        // descent into the key-ordered child picked by the Bourke corner id
        int ex = kChildrenVertexKind[vKind[t]];
        now = owner;
        while (depth < v.D) {
            ++depth;
            int cb = childBase[now];
            if (cb < 0) break;
            now = cb + ex;
            val += regular ? chi_node_sum_regular(now, v.basis[now], p, v.M, dCol)
                           : chi_node_sum(now, float(1 << depth), p, v.M, nbrCol);
        }
    }
    vv[t] = val - *iso;
}
__global__ void __launch_bounds__(kValueThreads) vertex_values(NodeView v, const int* __restrict__ count,
                                                          const int* __restrict__ vOwner,
                                                          const int* __restrict__ vKind,
                                                          const float4* __restrict__ vPos,
                                                          const int* __restrict__ parent,
                                                          const int* __restrict__ childBase,
                                                          const float4* __restrict__ nbrCol,
                                                          const float* __restrict__ dCol,
                                                          const int* __restrict__ q1Fired,
                                                          const float* __restrict__ iso, float* __restrict__ vv,
                                                          int* __restrict__ mail, const int* __restrict__ results,
                                                          int* __restrict__ resultsHost, int resultWords) {
    vertex_values_body(blockIdx.x, gridDim.x, v, count, vOwner, vKind, vPos, parent, childBase, nbrCol, dCol, q1Fired, iso, vv, mail, results, resultsHost, resultWords);
}
PSR_BATCHABLE(vertex_values, kValueThreads)


// --- depth-D marching cubes (ComputeTriangleIndices.cu:401-526) ------------
// The owned edges of each depth-D node that cross the surface (their corner
// values differ in sign), compacted in edge order -- (node, edge kind), the
// order of the edge records -- with their crossing points: per edge record a
// crossing flag and its crossing index.
__device__ __forceinline__ void edge_crossings_body(int psr_bid, int psr_nbid, int offD, int nD, int M, const int2* __restrict__ edgeOf,
                                                           const int* __restrict__ nodeVert,
                                                           const float* __restrict__ vv,
                                                           const float4* __restrict__ vPos, const ScanTicket& t,
                                                           int* __restrict__ flag, int* __restrict__ addr,
                                                           float* __restrict__ out, int D, unsigned* __restrict__ key) {
    const int tile = ticket_tile(t);
    const int n = tile * kThreads + threadIdx.x, o = offD + n;
    const int2 eo = n < nD ? edgeOf[n] : make_int2(0, 0);
    // the node's 8 corner vertices and their values, read once for all its
    // edges (two memory round trips)
    int cv[8];
    float cval[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) cv[i] = eo.y ? nodeVert[i * M + o] - 1 : -1;
#pragma unroll
    for (int i = 0; i < 8; ++i) cval[i] = cv[i] >= 0 ? vv[cv[i]] : 0.f;
    unsigned cross = 0;
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        if (!((eo.y >> k) & 1)) continue;
        const int c1 = kEdgeVertex[k][0], c2 = kEdgeVertex[k][1];
        // a corner slot left unassigned (the reference's level-local child 0
        // parent rule breaks neighbour tables) makes the edge absent, not a crossing
        if (cv[c1] >= 0 && cv[c2] >= 0 && cval[c1] * cval[c2] <= 0.f) cross |= 1u << k;
    }
    int c = ticket_prefix<kThreads>(t, tile, __popc(cross));
    int a = eo.x;
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        if (!((eo.y >> k) & 1)) continue;
        const bool x = (cross >> k) & 1u;
        flag[a] = x ? 1 : 0;
        addr[a] = c;
        ++a;
        if (!x) continue;
        const int c1 = kEdgeVertex[k][0], c2 = kEdgeVertex[k][1];
        const float4 p1 = vPos[cv[c1]], p2 = vPos[cv[c2]];
        const float3 p = interpolate_point(make_float3(p1.x, p1.y, p1.z), make_float3(p2.x, p2.y, p2.z), k >> 2,
                                           cval[c1], cval[c2]);
        out[3 * c] = p.x;
        out[3 * c + 1] = p.y;
        out[3 * c + 2] = p.z;
        // corner positions are exact multiples of 2^-D
        const float sc = float(1 << D);
        key[c] = segment_key(make_int3(int(p1.x * sc), int(p1.y * sc), int(p1.z * sc)),
                             make_int3(int(p2.x * sc), int(p2.y * sc), int(p2.z * sc)), D);
        ++c;
    }
}
__global__ void __launch_bounds__(kThreads) edge_crossings(int offD, int nD, int M, const int2* __restrict__ edgeOf,
                                                           const int* __restrict__ nodeVert,
                                                           const float* __restrict__ vv,
                                                           const float4* __restrict__ vPos, ScanTicket t,
                                                           int* __restrict__ flag, int* __restrict__ addr,
                                                           float* __restrict__ out, int D, unsigned* __restrict__ key) {
    edge_crossings_body(blockIdx.x, gridDim.x, offD, nD, M, edgeOf, nodeVert, vv, vPos, t, flag, addr, out, D, key);
}
PSR_BATCHABLE(edge_crossings, kThreads)


// A cube emits triangles only when all its corners and edges are present and
// every edge its triangles use crosses; otherwise (only where the reference's
// level-local child 0 parent rule has broken the tables and the reference
// reads out of bounds) it emits none.  The triangles of each node are written
// in node order, compacted in the same launch, and the faces bounding an
// emitted triangle's edges are marked up the parent chain.
__device__ __forceinline__ void cube_triangles_body(int psr_bid, int psr_nbid, 
    int offD, int nD, int M, const int* __restrict__ nodeVert, const int* __restrict__ nodeEdge,
    const int* __restrict__ crossFlag, const float* __restrict__ vv, const int* __restrict__ vexAddr,
    const int* __restrict__ nodeFace, const int* __restrict__ parent, const int* __restrict__ faceHasParent,
    const int* __restrict__ q1Fired, const ScanTicket& t, int* __restrict__ cube, int* __restrict__ tri,
    int* __restrict__ hasInter, int* __restrict__ mail) {
    const int tile = ticket_tile(t);
    const int n = tile * kThreads + threadIdx.x, o = offD + n;
    if (n == 0) mail[kMailQ1] = *q1Fired;
    // The reads come in three rounds, each issued together: the node's corner
    // and edge slots; the corner values; the 12 edges' crossing flags and
    // crossing indices (before the look-back, which they then overlap).
    int c = 0;
    int va[12];  // crossing index of each edge, read when every slot is assigned
#pragma unroll
    for (int e = 0; e < 12; ++e) va[e] = 0;
    if (n < nD) {
        int vtx[8], ne[12];
#pragma unroll
        for (int i = 0; i < 8; ++i) vtx[i] = nodeVert[i * M + o];
#pragma unroll
        for (int e = 0; e < 12; ++e) ne[e] = nodeEdge[e * nD + n];
        bool ok = true;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if (vtx[i] == 0) ok = false;
            else if (vv[vtx[i] - 1] < 0.f) c |= 1 << i;
        }
#pragma unroll
        for (int e = 0; e < 12; ++e)
            if (ne[e] == 0) ok = false;
        if (ok) {
            unsigned crossed = 0;
#pragma unroll
            for (int e = 0; e < 12; ++e) {
                crossed |= (crossFlag[ne[e] - 1] ? 1u : 0u) << e;
                va[e] = vexAddr[ne[e] - 1];
            }
            const uint64_t w = mc_word(c);
            for (int k = 0, m = 3 * mc_tri_count(c); k < m; ++k)
                if (!((crossed >> int((w >> (4 * k)) & 0xF)) & 1u)) ok = false;
        }
        if (!ok) c = 0;
        cube[n] = c;
    }
    const int nt = mc_tri_count(c);
    const int base = ticket_prefix<kThreads>(t, tile, nt);
    unsigned faces = 0;  // faces holding an edge of an emitted triangle
    if (nt > 0) {
        const uint64_t w = mc_word(c);
        int edgeHas = 0;
        for (int k = 0; k < 3 * nt; ++k) {
            const int e = int((w >> (4 * k)) & 0xF);
            edgeHas |= 1 << e;
            int a = 0;  // va[e], by selects rather than an indexed register read
#pragma unroll
            for (int j = 0; j < 12; ++j) a = e == j ? va[j] : a;
            tri[3 * base + k] = a;
        }
        for (int f = 0; f < 6; ++f) {
            bool mark = false;
            for (int j = 0; j < 4; ++j) mark |= (edgeHas >> kFaceEdges[f][j]) & 1;
            faces |= (mark ? 1u : 0u) << f;
        }
    }
    // each marked face walks up the parent chain, as an item of its own shared
    // out over the block (the walks are chains of dependent loads; marks are
    // idempotent, so their order does not matter)
    __shared__ unsigned short lWalk[kThreads * 6];
    int first;
    const int nWalks = block_excl_scan<kThreads>(__popc(faces), first);
    for (int f = 0; f < 6; ++f)
        if ((faces >> f) & 1u) lWalk[first++] = (threadIdx.x << 3) | f;
    __syncthreads();
    for (int it = threadIdx.x; it < nWalks; it += kThreads) {
        const int f = lWalk[it] & 7, o = offD + tile * kThreads + (lWalk[it] >> 3);
        int pn = parent[o];
        int cf = nodeFace[f * M + o] - 1;
        if (cf < 0) continue;  // face slot left unassigned (broken tables): nothing to mark
        hasInter[cf] = 1;
        // the walk stops where the reference would leave the tree or read an
        // unassigned slot; a step's three reads depend only on the step before,
        // so they are issued together (one memory round trip a level)
        while (pn >= 0) {
            const int up = faceHasParent[cf], cfUp = nodeFace[f * M + pn] - 1, pnUp = parent[pn];
            if (up == -1 || cfUp < 0) break;
            cf = cfUp;
            pn = pnUp;
            hasInter[cf] = 1;
        }
    }
}
__global__ void __launch_bounds__(kThreads) cube_triangles(
    int offD, int nD, int M, const int* __restrict__ nodeVert, const int* __restrict__ nodeEdge,
    const int* __restrict__ crossFlag, const float* __restrict__ vv, const int* __restrict__ vexAddr,
    const int* __restrict__ nodeFace, const int* __restrict__ parent, const int* __restrict__ faceHasParent,
    const int* __restrict__ q1Fired, ScanTicket t, int* __restrict__ cube, int* __restrict__ tri,
    int* __restrict__ hasInter, int* __restrict__ mail) {
    cube_triangles_body(blockIdx.x, gridDim.x, offD, nD, M, nodeVert, nodeEdge, crossFlag, vv, vexAddr, nodeFace, parent, faceHasParent, q1Fired, t, cube, tri, hasInter, mail);
}
PSR_BATCHABLE(cube_triangles, kThreads)


// --- subdivision roots (ProcessLeafNodesAtOtherDepth :555-587, copy_if) -----
// Roots in ascending node order, compacted in the same launch.  The first node
// of each level records how many roots precede it (levelStart), so the roots
// of a depth are a difference of two starts: precomputeSubdivideDepth reads a
// root's depth through neighs[13], and a root is its own neighbour here, so
// that is the root's level.
__device__ __forceinline__ void root_select_body(int psr_bid, int psr_nbid, int count, int M, const Levels& L, const int* __restrict__ neigh,
                                                        const int* __restrict__ nodeVert,
                                                        const int* __restrict__ nodeFace,
                                                        const int* __restrict__ childBase,
                                                        const float* __restrict__ vv,
                                                        const int* __restrict__ hasInter, const ScanTicket& t,
                                                        int* __restrict__ roots, int* __restrict__ rootPos,
                                                        int* __restrict__ levelStart) {
    const int tile = ticket_tile(t);
    const int o = tile * kThreads + threadIdx.x;
    int flag = 0;
    if (o < count && childBase[o] < 0 && neigh[13 * M + o] == o) {
        // unassigned corner and face slots are skipped; the sign reference is
        // the first present corner (corner 0 in a consistent tree).  All slots
        // are read at once, then all values they name: two memory round trips
        // instead of one or two per corner and face.
        int vtx[8], fs[6];
#pragma unroll
        for (int i = 0; i < 8; ++i) vtx[i] = nodeVert[i * M + o];
#pragma unroll
        for (int f = 0; f < 6; ++f) fs[f] = nodeFace[f * M + o];
        float val[8];
        int hit = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) val[i] = vtx[i] != 0 ? vv[vtx[i] - 1] : 0.f;
#pragma unroll
        for (int f = 0; f < 6; ++f) hit |= fs[f] != 0 ? hasInter[fs[f] - 1] : 0;
        bool seen = false;
        float sign = 1.f;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if (vtx[i] == 0) continue;
            if (!seen) {
                seen = true;
                sign = val[i] < 0.f ? -1.f : 1.f;
            } else if (sign * val[i] < 0.f) {
                flag = 1;
            }
        }
        if (hit) flag = 1;
    }
    const int a = ticket_prefix<kThreads>(t, tile, flag);
    if (o >= count) return;
    for (int d = 0; d < L.D; ++d)
        if (L.base[d] == o) levelStart[d] = a;
    if (flag) roots[a] = o;
    rootPos[o] = flag ? a : -1;  // position in roots, -1 for other nodes
}
__global__ void __launch_bounds__(kThreads) root_select(int count, int M, Levels L, const int* __restrict__ neigh,
                                                        const int* __restrict__ nodeVert,
                                                        const int* __restrict__ nodeFace,
                                                        const int* __restrict__ childBase,
                                                        const float* __restrict__ vv,
                                                        const int* __restrict__ hasInter, ScanTicket t,
                                                        int* __restrict__ roots, int* __restrict__ rootPos,
                                                        int* __restrict__ levelStart) {
    root_select_body(blockIdx.x, gridDim.x, count, M, L, neigh, nodeVert, nodeFace, childBase, vv, hasInter, t, roots, rootPos, levelStart);
}
PSR_BATCHABLE(root_select, kThreads)


}  // namespace

void build_geometry(Ctx& c) {
    Octree& o = c.oct;
    Geometry& g = c.geo;
    const Levels& L = o.L;
    const int M = o.M, D = L.D, nD = L.count[D], offD = L.base[D];
    hipStream_t s = c.stream;
    NodeView v{M, D, o.key, o.depth, o.neigh, o.basis};
    g.counts = c.buf<int>("geo_counts", 4);
    g.vtxOwner = c.buf<int>("vtx_owner", size_t(8) * M);
    g.vtxKind = c.buf<int>("vtx_kind", size_t(8) * M);
    g.vtxPos = c.buf<float4>("vtx_pos", size_t(8) * M);
    // nodeVert, nodeFace and nodeEdge were cleared with the neighbour columns
    // vertices, edges, faces: one launch, three scans over the same tiles
    g.edgeOwner = c.buf<int>("edge_owner", size_t(12) * nD);
    g.edgeKind = c.buf<int>("edge_kind", size_t(12) * nD);
    g.faceOwner = c.buf<int>("face_owner", size_t(6) * M);
    g.faceKind = c.buf<int>("face_kind", size_t(6) * M);
    g.faceHasParent = c.buf<int>("face_hasparent", size_t(6) * M);
    const int tilesM = div_up(M, kThreads);
    ScanState& scan = c.scan_for_tiles(size_t(3) * tilesM);
    const ScanTicket tv = scan_ticket(scan, tilesM, g.counts + 0);
    ScanTicket te = tv, tf = tv;  // one epoch; status words of their own
    te.status += tilesM;
    te.total = g.counts + 1;
    tf.status += 2 * size_t(tilesM);
    tf.total = g.counts + 2;
    g.edgeOf = c.buf<int2>("edge_of", nD);
    const GeoOut out{g.vtxOwner, g.vtxKind,  g.vtxPos,   g.nodeVert,      g.edgeOwner, g.edgeKind,
                     g.nodeEdge, g.faceOwner, g.faceKind, g.faceHasParent, g.nodeFace,  g.edgeOf};
    PSR_LAUNCH(geometry_emit, tilesM + c.fld.chiBlocks, 0, s, v, offD, nD, o.parent, o.q1Fired, tv, te, tf, out,
                                                                c.fld.dx, c.geo.dxGrid, tilesM, c.fld.chiArgs);
    PSR_LAUNCH(iso_final, 1, 0, s, c.fld.chiArgs.part, c.fld.chiBlocks, o.N, c.fld.isoSum, &c.results()->iso);
    // vertex implicit values
    float* isoDev = &c.results()->iso;
    g.vvalue = c.buf<float>("vvalue", size_t(8) * M);
    PSR_LAUNCH(vertex_values, div_up(size_t(8) * M, kValueThreads), 0, s, v, g.counts + 0, g.vtxOwner, g.vtxKind,
                                                                       g.vtxPos, o.parent, o.childBase,
                                                                       c.fld.nbrCol, c.fld.dCol, o.q1Fired, isoDev, g.vvalue, c.mail,
                                                                       reinterpret_cast<const int*>(c.results()),
                                                                       reinterpret_cast<int*>(c.mailResults),
                                                                       int(sizeof(DevResults) / sizeof(int)));
    // no wait: the stages that follow size their buffers by the bounds 8M, 12 n_D
    // and 6M, and the counts reach the host with the marching cubes' wait
    PSR_HIP_CHECK(hipGetLastError());
}

void depth_mc_launch(Ctx& c, MainMc& mm, float* rawV, unsigned* rawKey, int* rawT) {
    Octree& o = c.oct;
    Geometry& g = c.geo;
    const Levels& L = o.L;
    const int M = o.M, D = L.D, nD = L.count[D], offD = L.base[D];
    hipStream_t s = c.stream;
    const size_t Emax = size_t(12) * nD;  // bound of the edge count
    mm.hasInter = g.faceMarks;            // cleared with the neighbour columns
    // crossing, triangle and root totals and the level starts go to the mailbox;
    // launches that may be skipped leave zeros
    for (int k = 0; k < kMailLevelStart + PSR_MAX_DEPTH + 1; ++k) c.hostPinned[k] = 0;
    int* cflag = c.buf<int>("mc_cross_flag", Emax);
    mm.vexAddr = c.buf<int>("mc_vex_addr", Emax);
    mm.cube = c.buf<int>("mc_cube", nD);
    const int tilesD = div_up(nD, kThreads), tilesR = div_up(offD, kThreads);
    ScanState& scan = c.scan_for_tiles(size_t(std::max(tilesD, tilesR)));
    PSR_LAUNCH(edge_crossings, tilesD, 0, s, offD, nD, M, g.edgeOf, g.nodeVert, g.vvalue, g.vtxPos,
                                               scan_ticket(scan, tilesD, c.mail + 0), cflag, mm.vexAddr, rawV, L.D,
                                               rawKey);
    PSR_LAUNCH(cube_triangles, tilesD, 0, s, offD, nD, M, g.nodeVert, g.nodeEdge, cflag, g.vvalue, mm.vexAddr,
                                               g.nodeFace, o.parent, g.faceHasParent, o.q1Fired,
                                               scan_ticket(scan, tilesD, c.mail + 1), mm.cube, rawT, mm.hasInter,
                                               c.mail);
    // subdivision roots: leaves above depth D with a sign change or a marked face
    mm.roots = c.buf<int>("roots", offD);
    mm.rootPos = c.buf<int>("root_pos", offD);
    if (offD > 0)
        PSR_LAUNCH(root_select, tilesR, 0, s, offD, M, L, o.neigh, g.nodeVert, g.nodeFace, o.childBase, g.vvalue,
                                                mm.hasInter, scan_ticket(scan, tilesR, c.mail + 2), mm.roots,
                                                mm.rootPos, c.mail + kMailLevelStart);
    PSR_HIP_CHECK(hipGetLastError());
}

// After the wait: the counts the depth-D marching cubes left in the mailbox.
void depth_mc_read(Ctx& c, MainMc& mm) {
    Geometry& g = c.geo;
    const Levels& L = c.oct.L;
    const int D = L.D, offD = L.base[D];
    mm.q1Fired = c.hostPinned[kMailQ1] != 0;
    g.V = c.hostPinned[kMailGeo];
    g.E = c.hostPinned[kMailGeo + 1];
    g.F = c.hostPinned[kMailGeo + 2];
    mm.nCross = c.hostPinned[0];
    mm.nTri = c.hostPinned[1];
    mm.nRoots = c.hostPinned[2];
    // roots of depth d: the difference of the level starts (the total past the last level)
    auto start = [&](int d) { return L.base[d] < offD ? c.hostPinned[kMailLevelStart + d] : mm.nRoots; };
    for (int d = 0; d < D; ++d) mm.rootsPerDepth[d] = start(d + 1) - start(d);
    mm.rootsPerDepth[D] = 0;
    PSR_HIP_CHECK(hipGetLastError());
}

}  // namespace psr
