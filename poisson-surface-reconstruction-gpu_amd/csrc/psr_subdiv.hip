// Adaptive subdivision and final mesh assembly.
//
// The reference rebuilds every coarse subdivision root alone
// (CoarserSubdivideNodeAndRebuildMesh, ComputeTriangleIndices.cu:1475-1799)
// and the finer roots once per root depth (FinerSubdivideNodeAndRebuildMesh,
// :1801-2194), each as a full virtual subtree down to depth D whose ids start
// at NodeArraySize.  Here every such rebuild is a "set" with a disjoint virtual
// id range, and all sets run in the same launches: one launch per absolute
// depth for the neighbour tables, one per stage for the depth-D work.  A child
// look-up through a real parent is redirected only to the roots of the same
// set, which is what the reference's temporary child-pointer rewrite gives
// (:1525 and :1784 for the coarse path; :1137 for the finer path).
#include "psr_mesh_common.h"
#include "psr_prims.h"

#include <algorithm>

namespace psr {

namespace {

constexpr int kThreads = 256;

struct SetDesc {
    int rd;           // root depth
    int nr;           // roots in the set
    int rootOffset;   // first root in the root list
    int coarse;       // 1: coarse path semantics
    int vbase;        // first virtual index of the set
    int dEnum;        // first depth-D enumeration index of the set
    int levelStart[PSR_MAX_DEPTH + 2];  // relative to vbase
};

struct Span {
    int set, start, count, prefix;
};

struct SubView {
    int NAS, D, M, nSets;
    const SetDesc* sets;
    const int* roots;      // real node ids
    const int* rootPos;    // real node -> position in roots, or -1 (nodes above depth D)
    int rootPosSize;
    const int* key;        // real arrays
    const int* depth;
    const int* parent;
    const int* childBase;
    const int* neigh;
    const int* enc;
    const float* dx;
    int* vkey;             // virtual arrays
    int* vparent;
    int* vneigh;           // [27][VN]
    int VN;
};

__device__ __forceinline__ int find_set(const SubView& s, int v) {
    int lo = 0, hi = s.nSets - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (s.sets[mid].vbase <= v) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int find_span(const Span* sp, int n, int t) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (sp[mid].prefix <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Per-root count of a set's level l and the node's (root, offset) coordinates.
__device__ __forceinline__ void level_coords(const SetDesc& S, int l, int rel, int& q, int& j) {
    int per = 1 << (3 * (l - S.rd));
    int local = rel - S.levelStart[l];
    q = local / per;
    j = local - q * per;
}

__device__ __forceinline__ int level_of(const SetDesc& S, int rel, int D) {
    int l = S.rd;
    while (l < D && rel >= S.levelStart[l + 1]) ++l;
    return l;
}

// key / parent of every virtual node (singleRebuildArray :599-635, wholeRebuildArray :1110-1168)
__global__ void __launch_bounds__(kThreads) vnode_init(SubView s) {
    int v = blockIdx.x * kThreads + threadIdx.x;
    if (v >= s.VN) return;
    int si = find_set(s, v);
    const SetDesc& S = s.sets[si];
    int rel = v - S.vbase;
    int l = level_of(S, rel, s.D);
    int q, j;
    level_coords(S, l, rel, q, j);
    int root = s.roots[S.rootOffset + q];
    s.vkey[v] = s.key[root] | (j << (3 * (s.D - l)));
    if (l == S.rd) {
        s.vparent[v] = s.parent[root];
    } else {
        int perUp = 1 << (3 * (l - 1 - S.rd));
        s.vparent[v] = s.NAS + S.vbase + S.levelStart[l - 1] + q * perUp + (j >> 3);
    }
}

// neighbours of the virtual nodes of absolute depth l (computeRebuildNeighbor :669-696)
__global__ void __launch_bounds__(kThreads) vnode_neighbors(SubView s, const Span* __restrict__ spans, int nSpans,
                                                            int total, int l, NeighborLut lut) {
    int t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= total) return;
    const Span& sp = spans[find_span(spans, nSpans, t)];
    const SetDesc& S = s.sets[sp.set];
    int v = sp.start + (t - sp.prefix);
    int son = (s.vkey[v] >> (3 * (s.D - l))) & 7;
    int p = s.vparent[v];
    for (int i = 0; i < 27; ++i) {
        int pi = lut.parent[son][i], ci = lut.child[son][i];
        int np = (p < s.NAS) ? s.neigh[pi * s.M + p] : s.vneigh[size_t(pi) * s.VN + (p - s.NAS)];
        int r = -1;
        if (np != -1) {
            if (np < s.NAS) {
                int cb = s.childBase[np];
                if (cb >= 0) {
                    r = cb + ci;
                    int rp = (r < s.rootPosSize) ? s.rootPos[r] : -1;
                    if (rp >= S.rootOffset && rp < S.rootOffset + S.nr)
                        r = s.NAS + S.vbase + S.levelStart[S.rd] + (rp - S.rootOffset);
                }
            } else {
                int q, j;
                level_coords(S, l - 1, np - s.NAS - S.vbase, q, j);
                r = s.NAS + S.vbase + S.levelStart[l] + q * (1 << (3 * (l - S.rd))) + 8 * j + ci;
            }
        }
        s.vneigh[size_t(i) * s.VN + v] = r;
    }
}

__device__ __forceinline__ float3 any_center(const SubView& s, int id, int depth) {
    return id < s.NAS ? key_center(s.key[id], s.depth[id], s.D) : key_center(s.vkey[id - s.NAS], depth, s.D);
}

// depth-D enumeration e -> virtual index
__device__ __forceinline__ int venum(const Span* sp, int n, int e, int& set) {
    const Span& x = sp[find_span(sp, n, e)];
    set = x.set;
    return x.start + (e - x.prefix);
}

__device__ __forceinline__ int enum_of(const SubView& s, const SetDesc& S, int vid) {
    return S.dEnum + (vid - s.NAS - S.vbase - S.levelStart[s.D]);
}

// vertex ownership among virtual neighbours only (initSubdivideVertexOwner :698-763)
__global__ void __launch_bounds__(kThreads) vvertex_owner(SubView s, const Span* __restrict__ sp, int nsp, int nVD,
                                                          int* __restrict__ flag) {
    int e = blockIdx.x * kThreads + threadIdx.x;
    if (e >= nVD) return;
    int si;
    int v = venum(sp, nsp, e, si);
    const int D = s.D;
    float half = 1.0f / float(1 << (D + 1)), W = 1.0f / float(1 << D), W2 = W * W;
    int nb[27];
    for (int i = 0; i < 27; ++i) nb[i] = s.vneigh[size_t(i) * s.VN + v];
    float3 nc = (nb[13] >= s.NAS) ? key_center(s.vkey[nb[13] - s.NAS], D, D) : make_float3(0.f, 0.f, 0.f);
    int ownKey[8], own[8];
    for (int i = 0; i < 8; ++i) {
        ownKey[i] = 0x7fffffff;
        own[i] = -1;
    }
    for (int j = 0; j < 27; ++j) {
        if (nb[j] < s.NAS) continue;  // real neighbours never own (:735)
        float3 c = key_center(s.vkey[nb[j] - s.NAS], D, D);
        int k = s.vkey[nb[j] - s.NAS];
        for (int i = 0; i < 8; ++i) {
            if (sqdist(corner_pos(nc, i, half), c) < double(W2) && ownKey[i] > k) {
                ownKey[i] = k;
                own[i] = nb[j];
            }
        }
    }
    for (int i = 0; i < 8; ++i) flag[8 * e + i] = (own[i] == s.NAS + v) ? 1 : 0;
}

__global__ void __launch_bounds__(kThreads) vvertex_emit(SubView s, const Span* __restrict__ sp, int nsp, int nVD,
                                                         const int* __restrict__ flag, const int* __restrict__ addr,
                                                         int* __restrict__ vOwner, int* __restrict__ vKind,
                                                         float4* __restrict__ vPos) {
    int t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= 8 * nVD || !flag[t]) return;
    int e = t >> 3, i = t & 7, si;
    int v = venum(sp, nsp, e, si);
    const int D = s.D;
    float half = 1.0f / float(1 << (D + 1));
    int self = s.vneigh[size_t(13) * s.VN + v];
    float3 nc = (self >= s.NAS) ? key_center(s.vkey[self - s.NAS], D, D) : make_float3(0.f, 0.f, 0.f);
    float3 p = corner_pos(nc, i, half);
    int a = addr[t];
    vOwner[a] = s.NAS + v;
    vKind[a] = i;
    vPos[a] = make_float4(p.x, p.y, p.z, __int_as_float(si));
}

// maintainSubdivideVertexNodePointer (:765-810): vertex slots of virtual nodes only
__global__ void __launch_bounds__(kThreads) vvertex_maintain(SubView s, int nV, const int* __restrict__ vOwner,
                                                             const float4* __restrict__ vPos, int nVD,
                                                             int* __restrict__ vvert) {
    int t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= nV) return;
    int owner = vOwner[t];
    float4 q = vPos[t];
    const SetDesc& S = s.sets[__float_as_int(q.w)];
    const int D = s.D;
    float W = 1.0f / float(1 << D), W2 = W * W;
    float3 p = make_float3(q.x, q.y, q.z);
    for (int i = 0; i < 27; ++i) {
        int nb = s.vneigh[size_t(i) * s.VN + (owner - s.NAS)];
        if (nb == -1) continue;
        float3 c = any_center(s, nb, D);
        if (!(sqdist(p, c) < double(W2))) continue;
        if (nb >= s.NAS) vvert[size_t(corner_slot(c, p)) * nVD + enum_of(s, S, nb)] = t + 1;
    }
}

__global__ void __launch_bounds__(kThreads) vedge_owner(SubView s, const Span* __restrict__ sp, int nsp, int nVD,
                                                        int* __restrict__ flag) {
    int e = blockIdx.x * kThreads + threadIdx.x;
    if (e >= nVD) return;
    int si;
    int v = venum(sp, nsp, e, si);
    const int D = s.D;
    float half = 1.0f / float(1 << (D + 1)), W = 1.0f / float(1 << D), W2 = W * W;
    int nb[27];
    for (int i = 0; i < 27; ++i) nb[i] = s.vneigh[size_t(i) * s.VN + v];
    float3 nc = (nb[13] >= s.NAS) ? key_center(s.vkey[nb[13] - s.NAS], D, D) : make_float3(0.f, 0.f, 0.f);
    int ownKey[12], own[12];
    for (int k = 0; k < 12; ++k) {
        ownKey[k] = 0x7fffffff;
        own[k] = -1;
    }
    for (int j = 0; j < 27; ++j) {
        if (nb[j] < s.NAS) continue;
        float3 c = key_center(s.vkey[nb[j] - s.NAS], D, D);
        int kk = s.vkey[nb[j] - s.NAS];
        for (int k = 0; k < 12; ++k) {
            if (sqdist(edge_center(nc, k, half), c) < double(W2) && ownKey[k] > kk) {
                ownKey[k] = kk;
                own[k] = nb[j];
            }
        }
    }
    for (int k = 0; k < 12; ++k) flag[12 * e + k] = (own[k] == s.NAS + v) ? 1 : 0;
}

__global__ void __launch_bounds__(kThreads) vedge_emit(SubView s, const Span* __restrict__ sp, int nsp, int nVD,
                                                       const int* __restrict__ flag, const int* __restrict__ addr,
                                                       int* __restrict__ eOwner, int* __restrict__ eKind) {
    int t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= 12 * nVD || !flag[t]) return;
    int e = t / 12, si;
    int v = venum(sp, nsp, e, si);
    int a = addr[t];
    eOwner[a] = s.NAS + v;
    eKind[a] = (t % 12) | (si << 4);
}

__global__ void __launch_bounds__(kThreads) vedge_maintain(SubView s, int nE, const int* __restrict__ eOwner,
                                                           const int* __restrict__ eKind, int nVD,
                                                           int* __restrict__ vedge) {
    int t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= nE) return;
    int owner = eOwner[t], kind = eKind[t] & 15;
    const SetDesc& S = s.sets[eKind[t] >> 4];
    const int D = s.D;
    float half = 1.0f / float(1 << (D + 1)), W = 1.0f / float(1 << D), W2 = W * W;
    int self = s.vneigh[size_t(13) * s.VN + (owner - s.NAS)];
    float3 nc = (self != -1) ? any_center(s, self, D) : make_float3(0.f, 0.f, 0.f);
    float3 p = edge_center(nc, kind, half);
    for (int i = 0; i < 27; ++i) {
        int nb = s.vneigh[size_t(i) * s.VN + (owner - s.NAS)];
        if (nb == -1) continue;
        float3 c = any_center(s, nb, D);
        if (!(sqdist(p, c) < double(W2))) continue;
        if (nb >= s.NAS) vedge[size_t(edge_slot(c, p, kind)) * nVD + enum_of(s, S, nb)] = t + 1;
    }
}

// computeSubdivideVertexImplicitFunctionValue (:956-1038)
__global__ void __launch_bounds__(kThreads) vvertex_values(SubView s, int nV, const int* __restrict__ vOwner,
                                                           const float4* __restrict__ vPos,
                                                           const float* __restrict__ iso, float* __restrict__ vv) {
    int t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= nV) return;
    float4 q = vPos[t];
    const SetDesc& S = s.sets[__float_as_int(q.w)];
    float3 p = make_float3(q.x, q.y, q.z);
    const int rootLo = s.NAS + S.vbase + S.levelStart[S.rd];
    float val = 0.f;
    int now = vOwner[t];
    while (now != -1) {
        for (int i = 0; i < 27; ++i) {
            int nb = (now < s.NAS) ? s.neigh[i * s.M + now] : s.vneigh[size_t(i) * s.VN + (now - s.NAS)];
            if (nb == -1) continue;
            if (nb >= s.NAS) {
                if (nb >= rootLo && nb < rootLo + S.nr) {
                    // the root copy stands for the real root node (coarse: rootId, finer: ReplacedNodeId)
                    int root = s.roots[S.rootOffset + (nb - rootLo)];
                    nb = s.neigh[13 * s.M + root];
                    if (nb < 0) continue;
                } else {
                    continue;  // d_x = 0 in the subdivision space
                }
            }
            int3 f = decode_func_index(s.enc[nb], s.D);
            val += s.dx[nb] * basis_value(f.x, p.x) * basis_value(f.y, p.y) * basis_value(f.z, p.z);
        }
        now = (now < s.NAS) ? s.parent[now] : s.vparent[now - s.NAS];
    }
    vv[t] = val - *iso;
}

__global__ void __launch_bounds__(kThreads) vedge_crossings(SubView s, int nE, const int* __restrict__ eOwner,
                                                            const int* __restrict__ eKind, int nVD,
                                                            const int* __restrict__ vvert,
                                                            const float* __restrict__ vv, int* __restrict__ flag) {
    int t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= nE) return;
    int kind = eKind[t] & 15;
    const SetDesc& S = s.sets[eKind[t] >> 4];
    int e = enum_of(s, S, eOwner[t]);
    int v1 = vvert[size_t(kEdgeVertex[kind][0]) * nVD + e] - 1, v2 = vvert[size_t(kEdgeVertex[kind][1]) * nVD + e] - 1;
    flag[t] = (vv[v1] * vv[v2] <= 0.f) ? 1 : 0;
}

__global__ void __launch_bounds__(kThreads) vcrossing_points(SubView s, int nE, const int* __restrict__ eOwner,
                                                             const int* __restrict__ eKind, int nVD,
                                                             const int* __restrict__ vvert,
                                                             const float* __restrict__ vv,
                                                             const float4* __restrict__ vPos,
                                                             const int* __restrict__ flag,
                                                             const int* __restrict__ addr, float* __restrict__ out) {
    int t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= nE || !flag[t]) return;
    int kind = eKind[t] & 15;
    const SetDesc& S = s.sets[eKind[t] >> 4];
    int e = enum_of(s, S, eOwner[t]);
    int v1 = vvert[size_t(kEdgeVertex[kind][0]) * nVD + e] - 1, v2 = vvert[size_t(kEdgeVertex[kind][1]) * nVD + e] - 1;
    float4 a = vPos[v1], b = vPos[v2];
    float3 p = interpolate_point(make_float3(a.x, a.y, a.z), make_float3(b.x, b.y, b.z), kind >> 2, vv[v1], vv[v2]);
    int k = addr[t];
    out[3 * k] = p.x;
    out[3 * k + 1] = p.y;
    out[3 * k + 2] = p.z;
}

__global__ void __launch_bounds__(kThreads) vcube_classes(int nVD, const int* __restrict__ vvert,
                                                          const float* __restrict__ vv, int* __restrict__ cube,
                                                          int* __restrict__ triNum) {
    int e = blockIdx.x * kThreads + threadIdx.x;
    if (e >= nVD) return;
    int c = 0;
    for (int i = 0; i < 8; ++i)
        if (vv[vvert[size_t(i) * nVD + e] - 1] < 0.f) c |= 1 << i;
    cube[e] = c;
    triNum[e] = mc_tri_count(c);
}

// First crossing of each set of a chunk: the crossing scan value at the set's first edge.
__global__ void set_cross_start(int nSets, const SetDesc* __restrict__ sets, const int* __restrict__ edgeAddr,
                                int nEflag, const int* __restrict__ vexAddr, int nE, int nCross,
                                int* __restrict__ crossStart) {
    int si = blockIdx.x * blockDim.x + threadIdx.x;
    if (si >= nSets) return;
    int f = 12 * sets[si].dEnum;
    int firstEdge = f < nEflag ? edgeAddr[f] : nE;
    crossStart[si] = firstEdge < nE ? vexAddr[firstEdge] : nCross;
}

// Triangles of the subdivision batches (insertTriangle :1247-1282): an index is
// the number of vertices kept before the batch plus the raw offset inside the
// batch; indices are not remapped around dropped vertices.
__global__ void __launch_bounds__(kThreads) vemit_triangles(const Span* __restrict__ sp, int nsp, int nVD,
                                                            const int* __restrict__ cube,
                                                            const int* __restrict__ triAddr,
                                                            const int* __restrict__ vedge,
                                                            const int* __restrict__ vexAddr,
                                                            const int* __restrict__ crossStart,
                                                            const int* __restrict__ localKept,
                                                            const int* __restrict__ keptBase, int triBase,
                                                            int* __restrict__ tri) {
    int e = blockIdx.x * kThreads + threadIdx.x;
    if (e >= nVD) return;
    int c = cube[e];
    int nt = mc_tri_count(c);
    if (!nt) return;
    int si = sp[find_span(sp, nsp, e)].set;
    int bs = crossStart[si];  // first raw vertex of this batch within the chunk
    int keptBefore = *keptBase + localKept[bs];
    uint64_t w = mc_word(c);
    size_t base = size_t(triBase) + triAddr[e];
    for (int k = 0; k < 3 * nt; ++k) {
        int ed = int((w >> (4 * k)) & 0xF);
        int raw = vexAddr[vedge[size_t(ed) * nVD + e] - 1];
        tri[3 * base + k] = keptBefore + (raw - bs);
    }
}

// markValidMeshVertexIndex (:1170-1181): a vertex with |x| < 1e-6 is dropped;
// triangle indices are not remapped (insertTriangle :1247-1282).
__global__ void __launch_bounds__(kThreads) keep_flags(const float* __restrict__ v, int n, int* __restrict__ keep) {
    int i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) keep[i] = fabsf(v[3 * size_t(i)]) < 1e-6f ? 0 : 1;
    if (i == n) keep[i] = 0;
}

__global__ void __launch_bounds__(kThreads) compact_vertices(const float* __restrict__ v, int n,
                                                             const int* __restrict__ keep,
                                                             const int* __restrict__ addr,
                                                             const int* __restrict__ keptBase,
                                                             float* __restrict__ out) {
    int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n || !keep[i]) return;
    size_t a = size_t(*keptBase) + addr[i];
    out[3 * a] = v[3 * size_t(i)];
    out[3 * a + 1] = v[3 * size_t(i) + 1];
    out[3 * a + 2] = v[3 * size_t(i) + 2];
}

__global__ void add_count(int* __restrict__ acc, const int* __restrict__ add) { *acc += *add; }

__global__ void fill_root_pos(const int* __restrict__ roots, int n, int* __restrict__ rootPos) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) rootPos[roots[i]] = i;
}

__global__ void __launch_bounds__(kThreads) fill_int(int* __restrict__ a, size_t n, int v) {
    size_t i = size_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) a[i] = v;
}

// Default virtual nodes per chunk of subdivision sets; a set larger than the
// chunk size forms a chunk of its own.
constexpr size_t kChunkNodes = size_t(1) << 24;

}  // namespace

void extract_mesh(Ctx& c) {
    Octree& o = c.oct;
    Geometry& g = c.geo;
    const Levels& L = o.L;
    const int M = o.M, D = L.D, nD = L.count[D], offD = L.base[D];
    hipStream_t st = c.stream;
    // ---- depth-D marching cubes: the first output batch
    float* mainV = c.buf<float>("raw_main_v", size_t(3) * std::max(g.E, 1));
    int* mainT = c.buf<int>("raw_main_t", size_t(15) * nD);
    MainMc mm;
    depth_mc(c, mm, mainV, mainT);
    c.stats.n_subdivision_roots = mm.nRoots;
    int* acc = c.buf<int>("mesh_counts", 4);  // [0] vertices kept so far, [1] batch kept count
    fill_int<<<1, kThreads, 0, st>>>(acc, 4, 0);
    size_t rawUpper = 0, triTotal = 0;
    float* outV = nullptr;
    int* outT = nullptr;
    // Appends one batch of raw vertices: keep flags, local scan, compaction at
    // the running kept count.  localKept holds n + 1 entries afterwards.
    auto append_vertices = [&](const float* rawV, int n, int* keep, int* localKept, int* scratch) {
        outV = c.buf_keep<float>("mesh_vertices", 3 * (rawUpper + n), 3 * rawUpper);
        keep_flags<<<div_up(size_t(n) + 1, kThreads), kThreads, 0, st>>>(rawV, n, keep);
        scan_exclusive(keep, localKept, n + 1, acc + 1, scratch, st);
        compact_vertices<<<div_up(std::max(n, 1), kThreads), kThreads, 0, st>>>(rawV, n, keep, localKept, acc, outV);
        rawUpper += n;
    };
    {
        int* keep = c.buf<int>("keep_flag", size_t(mm.nCross) + 1);
        int* kept = c.buf<int>("keep_scan", size_t(mm.nCross) + 1);
        int* scratch = c.scratch_for(scan_scratch_ints(size_t(mm.nCross) + 1) + 64);
        append_vertices(mainV, mm.nCross, keep, kept, scratch);
        add_count<<<1, 1, 0, st>>>(acc, acc + 1);
        // depth-D batch: indices are raw crossing ids (batch start 0, nothing kept before it)
        outT = c.buf_keep<int>("mesh_triangles", size_t(3) * std::max(mm.nTri, 1), 0);
        PSR_HIP_CHECK(hipMemcpyAsync(outT, mainT, sizeof(int) * 3 * mm.nTri, hipMemcpyDeviceToDevice, st));
        triTotal = mm.nTri;
    }
    // ---- subdivision sets: coarse roots one per set, finer roots one set per depth
    const int finerDepth = D - 4;
    std::vector<SetDesc> sets;
    std::vector<size_t> setNodes;
    int rootCursor = 0, nCoarse = 0;
    for (int d = 0; d < D; ++d) {
        int cnt = mm.rootsPerDepth[d];
        if (cnt == 0) continue;
        auto make = [&](int nr, bool coarse) {
            SetDesc S{};
            S.rd = d;
            S.nr = nr;
            S.rootOffset = rootCursor;
            S.coarse = coarse ? 1 : 0;
            size_t a = 0;
            for (int l = d; l <= D; ++l) {
                S.levelStart[l] = int(a);
                a += size_t(nr) << (3 * (l - d));
            }
            S.levelStart[D + 1] = int(a);
            if (a + size_t(M) > size_t(INT32_MAX)) throw std::runtime_error("subdivision set exceeds 32-bit node ids");
            rootCursor += nr;
            sets.push_back(S);
            setNodes.push_back(a);
        };
        if (d < finerDepth) {
            for (int k = 0; k < cnt; ++k) make(1, true);
            nCoarse += cnt;
        } else {
            make(cnt, false);
        }
    }
    c.stats.n_coarse_roots = nCoarse;
    const int nSets = int(sets.size());
    if (nSets > 0) {
        int* rootPos = c.buf<int>("root_pos", offD);
        fill_int<<<div_up(offD, kThreads), kThreads, 0, st>>>(rootPos, size_t(offD), -1);
        fill_root_pos<<<div_up(mm.nRoots, kThreads), kThreads, 0, st>>>(mm.roots, mm.nRoots, rootPos);
        NeighborLut lut;
        make_neighbor_lut(lut);
        const size_t chunkNodes = c.cfg.subdiv_chunk_nodes ? c.cfg.subdiv_chunk_nodes : kChunkNodes;
        std::vector<SetDesc> cs;
        std::vector<Span> spans;
        for (int s0 = 0; s0 < nSets;) {
            // ---- one chunk of sets [s0, s1), virtual ids local to the chunk
            size_t vnTotal = 0, vdTotal = 0;
            int s1 = s0;
            while (s1 < nSets && (s1 == s0 || vnTotal + setNodes[s1] <= chunkNodes)) {
                vnTotal += setNodes[s1];
                vdTotal += size_t(sets[s1].nr) << (3 * (D - sets[s1].rd));
                ++s1;
            }
            const int nCs = s1 - s0, VN = int(vnTotal), nVD = int(vdTotal);
            cs.assign(sets.begin() + s0, sets.begin() + s1);
            {
                int vb = 0, de = 0;
                for (int k = 0; k < nCs; ++k) {
                    cs[k].vbase = vb;
                    cs[k].dEnum = de;
                    vb += int(setNodes[s0 + k]);
                    de += cs[k].nr << (3 * (D - cs[k].rd));
                }
            }
            SetDesc* dsets = c.buf<SetDesc>("sub_sets", nCs);
            PSR_HIP_CHECK(hipMemcpyAsync(dsets, cs.data(), sizeof(SetDesc) * nCs, hipMemcpyHostToDevice, st));
            // spans per absolute depth
            spans.clear();
            std::vector<int> spanStart(D + 2, 0), spanTotal(D + 1, 0);
            for (int l = 0; l <= D; ++l) {
                spanStart[l] = int(spans.size());
                int prefix = 0;
                for (int si = 0; si < nCs; ++si) {
                    const SetDesc& S = cs[si];
                    if (l < S.rd) continue;
                    int cnt = S.nr << (3 * (l - S.rd));
                    spans.push_back({si, S.vbase + S.levelStart[l], cnt, prefix});
                    prefix += cnt;
                }
                spanTotal[l] = prefix;
            }
            spanStart[D + 1] = int(spans.size());
            Span* dspans = c.buf<Span>("sub_spans", spans.size());
            PSR_HIP_CHECK(hipMemcpyAsync(dspans, spans.data(), sizeof(Span) * spans.size(), hipMemcpyHostToDevice,
                                         st));
            SubView sv{};
            sv.NAS = M;
            sv.D = D;
            sv.M = M;
            sv.nSets = nCs;
            sv.sets = dsets;
            sv.roots = mm.roots;
            sv.rootPos = rootPos;
            sv.rootPosSize = offD;
            sv.key = o.key;
            sv.depth = o.depth;
            sv.parent = o.parent;
            sv.childBase = o.childBase;
            sv.neigh = o.neigh;
            sv.enc = o.encIdx;
            sv.dx = c.fld.dx;
            sv.VN = VN;
            sv.vkey = c.buf<int>("sub_vkey", VN);
            sv.vparent = c.buf<int>("sub_vparent", VN);
            sv.vneigh = c.buf<int>("sub_vneigh", size_t(27) * VN);
            vnode_init<<<div_up(VN, kThreads), kThreads, 0, st>>>(sv);
            for (int l = 0; l <= D; ++l) {
                int n = spanTotal[l];
                if (n == 0) continue;
                vnode_neighbors<<<div_up(n, kThreads), kThreads, 0, st>>>(sv, dspans + spanStart[l],
                                                                         spanStart[l + 1] - spanStart[l], n, l, lut);
            }
            Span* dSpans = dspans + spanStart[D];
            const int nDspans = spanStart[D + 1] - spanStart[D];
            // vertices of the virtual depth-D nodes
            const size_t nflag = size_t(12) * nVD;
            int* flag = c.buf<int>("sub_flag", nflag + 1);
            int* addr = c.buf<int>("sub_addr", nflag + 1);
            int* cnts = c.buf<int>("sub_counts", 8);
            int* scratch = c.scratch_for(scan_scratch_ints(nflag + 1) + 64);
            int* vOwner = c.buf<int>("sub_vowner", size_t(8) * nVD);
            int* vKind = c.buf<int>("sub_vkind", size_t(8) * nVD);
            float4* vPos = c.buf<float4>("sub_vpos", size_t(8) * nVD);
            int* vvert = c.buf<int>("sub_vvert", size_t(8) * nVD);
            int* vedge = c.buf<int>("sub_vedge", size_t(12) * nVD);
            fill_int<<<div_up(size_t(8) * nVD, kThreads), kThreads, 0, st>>>(vvert, size_t(8) * nVD, 0);
            fill_int<<<div_up(size_t(12) * nVD, kThreads), kThreads, 0, st>>>(vedge, size_t(12) * nVD, 0);
            vvertex_owner<<<div_up(nVD, kThreads), kThreads, 0, st>>>(sv, dSpans, nDspans, nVD, flag);
            scan_exclusive(flag, addr, 8 * nVD, cnts + 0, scratch, st);
            vvertex_emit<<<div_up(size_t(8) * nVD, kThreads), kThreads, 0, st>>>(sv, dSpans, nDspans, nVD, flag,
                                                                                addr, vOwner, vKind, vPos);
            // edges
            int* eflag = c.buf<int>("sub_eflag", nflag);
            int* eflagAddr = c.buf<int>("sub_eaddr", nflag);
            int* eOwner = c.buf<int>("sub_eowner", nflag);
            int* eKind = c.buf<int>("sub_ekind", nflag);
            vedge_owner<<<div_up(nVD, kThreads), kThreads, 0, st>>>(sv, dSpans, nDspans, nVD, eflag);
            scan_exclusive(eflag, eflagAddr, 12 * nVD, cnts + 1, scratch, st);
            vedge_emit<<<div_up(nflag, kThreads), kThreads, 0, st>>>(sv, dSpans, nDspans, nVD, eflag, eflagAddr,
                                                                     eOwner, eKind);
            PSR_HIP_CHECK(hipMemcpyAsync(c.hostPinned, cnts, sizeof(int) * 2, hipMemcpyDeviceToHost, st));
            PSR_HIP_CHECK(hipStreamSynchronize(st));
            const int nVsub = c.hostPinned[0], nEsub = c.hostPinned[1];
            vvertex_maintain<<<div_up(nVsub, kThreads), kThreads, 0, st>>>(sv, nVsub, vOwner, vPos, nVD, vvert);
            vedge_maintain<<<div_up(nEsub, kThreads), kThreads, 0, st>>>(sv, nEsub, eOwner, eKind, nVD, vedge);
            float* vv = c.buf<float>("sub_vvalue", nVsub);
            float* isoDev = c.buf<float>("iso_value", 1);
            vvertex_values<<<div_up(nVsub, kThreads), kThreads, 0, st>>>(sv, nVsub, vOwner, vPos, isoDev, vv);
            int* cflag = c.buf<int>("sub_cflag", nEsub);
            int* vexAddr = c.buf<int>("sub_vexaddr", nEsub);
            vedge_crossings<<<div_up(nEsub, kThreads), kThreads, 0, st>>>(sv, nEsub, eOwner, eKind, nVD, vvert, vv,
                                                                         cflag);
            scan_exclusive(cflag, vexAddr, nEsub, cnts + 2, scratch, st);
            int* cube = c.buf<int>("sub_cube", nVD);
            int* triNum = c.buf<int>("sub_trinum", nVD);
            int* triAddr = c.buf<int>("sub_triaddr", nVD);
            vcube_classes<<<div_up(nVD, kThreads), kThreads, 0, st>>>(nVD, vvert, vv, cube, triNum);
            scan_exclusive(triNum, triAddr, nVD, cnts + 3, scratch, st);
            PSR_HIP_CHECK(hipMemcpyAsync(c.hostPinned, cnts + 2, sizeof(int) * 2, hipMemcpyDeviceToHost, st));
            PSR_HIP_CHECK(hipStreamSynchronize(st));
            const int nCross = c.hostPinned[0], nTri = c.hostPinned[1];
            float* subV = c.buf<float>("raw_sub_v", size_t(3) * std::max(nCross, 1));
            vcrossing_points<<<div_up(nEsub, kThreads), kThreads, 0, st>>>(sv, nEsub, eOwner, eKind, nVD, vvert, vv,
                                                                          vPos, cflag, vexAddr, subV);
            int* crossStart = c.buf<int>("sub_cross_start", nCs);
            set_cross_start<<<div_up(nCs, 64), 64, 0, st>>>(nCs, dsets, eflagAddr, int(nflag), vexAddr, nEsub, nCross,
                                                            crossStart);
            // this chunk's batches: vertices, then triangles against the kept count before the chunk
            append_vertices(subV, nCross, flag, addr, scratch);
            if (nTri > 0) {
                outT = c.buf_keep<int>("mesh_triangles", 3 * (triTotal + nTri), 3 * triTotal);
                vemit_triangles<<<div_up(nVD, kThreads), kThreads, 0, st>>>(dSpans, nDspans, nVD, cube, triAddr,
                                                                            vedge, vexAddr, crossStart, addr, acc,
                                                                            int(triTotal), outT);
                triTotal += nTri;
            }
            add_count<<<1, 1, 0, st>>>(acc, acc + 1);
            s0 = s1;
        }
    }
    PSR_HIP_CHECK(hipMemcpyAsync(c.hostPinned, acc, sizeof(int), hipMemcpyDeviceToHost, st));
    PSR_HIP_CHECK(hipStreamSynchronize(st));
    if (triTotal > size_t(INT32_MAX) / 3) throw std::runtime_error("mesh exceeds 32-bit triangle indexing");
    c.mesh.vertices = outV;
    c.mesh.triangles = outT;
    c.mesh.nv = c.hostPinned[0];
    c.mesh.nt = int(triTotal);
    c.mesh.counts = acc;
    c.stats.n_mesh_vertices = c.mesh.nv;
    c.stats.n_mesh_triangles = c.mesh.nt;
    PSR_HIP_CHECK(hipGetLastError());
}

}  // namespace psr
