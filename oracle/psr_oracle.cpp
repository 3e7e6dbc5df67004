// Oracle for the Poisson reconstruction hot path -- TEST INFRASTRUCTURE ONLY.
//
// A sequential CPU restatement of the reference pipeline
// (mesh/PoissonReconstruction.cpp:195-245): normalisation, Morton keys, sort,
// octree node arrays, neighbour tables, basis inner-product tables, vector
// field, divergence, per-depth Laplacian + CG, implicit function at points,
// vertex/edge/face ownership, vertex implicit values, marching cubes and the
// adaptive subdivision.  Only tests/, the smoke check and the benchmark's CPU
// baseline load it, as the checker or as the timed CPU port; the product never
// does.
//
// Each function cites the reference lines it restates.  Integer stages follow
// the reference step by step (including its quirks Q1, Q8, the key-ordered
// parentFaceKind table); floating-point stages are templated on the
// accumulation type so one build provides
//   * Real = double : the FP64 accuracy oracle (CG to ||r||/||b|| <= 1e-12), and
//   * Real = float  : the FP32 restatement with the reference's CG stopping rule
//                     (absolute 1e-5 on the recursively updated residual,
//                     <= 10000 iterations, double dot accumulation), which is
//                     timed as the CPU baseline and supplies the iteration
//                     counts k_d.
// B-spline values are evaluated in the node's local coordinate (x-c)/w; the
// reference evaluates shifted float polynomials in the global coordinate
// (mesh/ConfirmedPPolynomial.h:69-82), which is the same function with more
// rounding.  Inner-product tables are exact FP64 integrals (3-point
// Gauss-Legendre per polynomial piece, exact for degree <= 5) of the products
// the reference integrates in float (mesh/FunctionData.h:134-312).
//
// Pinning (see tests/): Morton/sort/unique against brute force, node arrays
// against invariants and a key-arithmetic neighbour search, tables against the
// closed-form B-spline autocorrelations, CG against a dense solve, MC table
// against its own crossing-edge property.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle_mc_table.h"

namespace orc {

static const int kMaxD = 10;

// ---------------------------------------------------------------------------
// Small tables.  LUTparent/LUTchild are derived from the 3x3x3 neighbourhood
// rule (mesh/BuildOctree.cu:20-42 lists the same values).
// ---------------------------------------------------------------------------
struct Luts {
    int parent[8][27];
    int child[8][27];
    Luts() {
        for (int s = 0; s < 8; ++s) {
            int sb[3] = {(s >> 2) & 1, (s >> 1) & 1, s & 1};
            for (int i = 0; i < 27; ++i) {
                int d[3] = {i / 9 - 1, (i / 3) % 3 - 1, i % 3 - 1};
                int pi = 0, ci = 0;
                for (int a = 0; a < 3; ++a) {
                    int c = sb[a] + d[a];               // -1..2 in child units
                    int p = (c < 0) ? -1 : (c > 1 ? 1 : 0);
                    int cm = c - 2 * p;                  // 0..1
                    pi = pi * 3 + (p + 1);
                    ci = ci * 2 + cm;
                }
                parent[s][i] = pi;
                child[s][i] = ci;
            }
        }
    }
};
static const Luts kLut;

// Bourke corner order used for OctNode.vertices (mesh/BuildMeshGeometry.cu:133-144).
static const int kChildrenVertexKind[8] = {0, 1, 3, 2, 4, 5, 7, 6};  // ComputeTriangleIndices.cu:31 (Q8)
static const int kEdgeVertex[12][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 3}, {1, 2},
                                       {4, 7}, {5, 6}, {0, 4}, {1, 5}, {3, 7}, {2, 6}};
static const int kFaceEdges[6][4] = {{4, 6, 8, 10}, {5, 7, 9, 11}, {0, 2, 8, 9},
                                     {1, 3, 10, 11}, {0, 1, 4, 5}, {2, 3, 6, 7}};
// BuildMeshGeometry.cu:19-26, indexed by the key-ordered sonKey exactly as the
// reference does (the rows are in x-fastest order; replicated, not corrected).
static const int kParentFaceKind[8][6] = {{0, -1, 2, -1, 4, -1}, {-1, 1, 2, -1, 4, -1},
                                          {0, -1, -1, 3, 4, -1}, {-1, 1, -1, 3, 4, -1},
                                          {0, -1, 2, -1, -1, 5}, {-1, 1, 2, -1, -1, 5},
                                          {0, -1, -1, 3, -1, 5}, {0, 1, -1, 3, -1, 5}};

static int mc_count(int cube) {
    uint64_t w = kOracleMcTable[cube];
    int n = 0;
    while (n < 15 && ((w >> (4 * n)) & 0xF) != 0xF) ++n;
    return n / 3;
}
static int mc_edge(int cube, int k) { return int((kOracleMcTable[cube] >> (4 * k)) & 0xF); }

// ---------------------------------------------------------------------------
// Types (mesh/OctNode.cuh:15-106).
// ---------------------------------------------------------------------------
struct OctNode {
    int key, pidx, pnum, parent, children[8], neighs[27], didx, dnum;
    int vertices[8], edges[12], faces[6], hasTriangle, hasIntersection;
};
struct Vtx { float pos[3]; int owner, kind, depth; };
struct Edge { int kind, owner; };
struct Face { int kind, owner, hasParentFace; };
// Virtual node of the subdivision (EasyOctNode, OctNode.cuh:44-77).
struct VNode { int key, parent, children[8], neighs[27], vertices[8], edges[12]; };

struct Result {
    int N = 0, D = 0;
    float corner[3] = {0, 0, 0}, edge = 0;
    std::vector<float> spos, snrm;      // sorted, normalised positions / scaled normals (N*3)
    std::vector<int> order;             // sorted -> input index
    std::vector<int> codes;             // sorted Morton codes (3D bits)
    std::vector<OctNode> nodes;
    std::vector<int> childrenBuilt;     // children after the octree build (before subdivision)
    int base[kMaxD + 2] = {0}, count[kMaxD + 2] = {0};
    int M = 0;
    std::vector<int> point2node, encIdx, depthOf;
    std::vector<float> center;          // M*3
    int q1_count = 0, q1b_count = 0, q1_levels = 0, q16_flag = 0;
    // numerics
    std::vector<double> vfield, div, dx, chiPts;
    double iso = 0;
    std::vector<int> cgIters;
    std::vector<double> cgRelRes, cgRecRes;
    std::vector<int> nnz;
    // geometry
    std::vector<Vtx> verts;
    std::vector<Edge> edges;
    std::vector<Face> faces;
    std::vector<double> vvalue;
    std::vector<int> cube;              // depth-D cube categories
    // mesh
    std::vector<float> meshV;
    std::vector<int> meshT;
    std::vector<int> batchV, batchT;    // vertex/triangle count per appended batch
    int nRoots = 0, nCoarseRoots = 0;
    std::vector<int> rootNodes;
    // subdivision record (for the tests' independent owner and value checks):
    //   subSets:  per set rd, coarse (0/1), first entry in subRoots, root count
    //   subRoots: the real node index of every root, set by set
    //   subV:     per virtual vertex set, owner key, corner kind, position (3 float bits),
    //             value (float bits; NaN when the set emits no edge)
    //   subE:     per virtual edge set, owner key, edge kind
    std::vector<int> subSets, subRoots, subV, subE;
    // timings (ms)
    double tOctree = 0, tField = 0, tSolve = 0, tMesh = 0;
};

static inline double now_ms() {
    return std::chrono::duration<double, std::milli>(
        std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------------------
// Octree (mesh/BuildOctree.cu / BuildOctree.cpp:75-194).
// ---------------------------------------------------------------------------
static int depth_of(const Result& r, int idx) {
    // BuildOctree.cu:398-402: the search loop leaves depth = D for the last level.
    int d;
    for (d = 0; d < r.D; ++d)
        if (r.base[d] <= idx && idx < r.base[d + 1]) break;
    return d;
}

static void node_center(int key, int depth, int D, float c[3]) {
    // BuildOctree.cu:568-586
    c[0] = c[1] = c[2] = 0.5f;
    float w = 0.25f;
    for (int i = D - 1; i >= D - depth; --i) {
        c[0] += ((key >> (3 * i + 2)) & 1) ? w : -w;
        c[1] += ((key >> (3 * i + 1)) & 1) ? w : -w;
        c[2] += ((key >> (3 * i)) & 1) ? w : -w;
        w /= 2;
    }
}

static int encoded_index(int key, int depth, int D) {
    // BuildOctree.cu:522-550
    int idx = ((1 << depth) - 1) * (1 + (1 << (D + 1)) + (1 << (2 * (D + 1))));
    for (int l = depth; l >= 1; --l) {
        int sx = (key >> (3 * (D - l) + 2)) & 1, sy = (key >> (3 * (D - l) + 1)) & 1,
            sz = (key >> (3 * (D - l))) & 1;
        idx += sx * (1 << (depth - l)) + sy * (1 << (depth - l)) * (1 << (D + 1)) +
               sz * (1 << (depth - l)) * (1 << (2 * (D + 1)));
    }
    return idx;
}

static void build_octree(Result& r, const float* xyz, const float* nrm, int N, int D) {
    r.N = N;
    r.D = D;
    // bbox (BuildOctree.cu:51-133) and host adjustment (:619-631)
    float mx[3] = {-1e6f, -1e6f, -1e6f}, mn[3] = {1e6f, 1e6f, 1e6f};
    for (int i = 0; i < N; ++i)
        for (int a = 0; a < 3; ++a) {
            float v = xyz[3 * i + a];
            mx[a] = mx[a] < v ? v : mx[a];
            mn[a] = mn[a] > v ? v : mn[a];
        }
    float maxEdge = 1.0f, center[3];
    for (int a = 0; a < 3; ++a) {
        if (a == 0 || maxEdge < (mx[a] - mn[a])) maxEdge = float(mx[a] - mn[a]);
        center[a] = float(mx[a] + mn[a]) / 2.0f;
    }
    maxEdge *= 1.25f * 1.0f;
    for (int a = 0; a < 3; ++a) center[a] -= maxEdge / 2.0f;
    for (int a = 0; a < 3; ++a) r.corner[a] = center[a];
    r.edge = maxEdge;
    // normalise (BuildOctree.cu:144-165)
    std::vector<float> pos(3 * N), nn(3 * N);
    const double eps = double(float(1e-6));
    for (int i = 0; i < N; ++i) {
        for (int a = 0; a < 3; ++a) pos[3 * i + a] = (xyz[3 * i + a] - center[a]) * 1.0f / maxEdge;
        float x = nrm[3 * i], y = nrm[3 * i + 1], z = nrm[3 * i + 2];
        float len = sqrtf(x * x + y * y + z * z);
        if (len > eps) len = 1.0f / len;
        len *= float(2 << D);
        nn[3 * i] = x * len;
        nn[3 * i + 1] = y * len;
        nn[3 * i + 2] = z * len;
    }
    // Morton code (BuildOctree.cu:167-208)
    std::vector<int> code(N);
    for (int i = 0; i < N; ++i) {
        long long key = 0;
        float c[3] = {0.5f, 0.5f, 0.5f}, w = 0.25f;
        for (int l = D - 1; l >= 0; --l) {
            for (int a = 0; a < 3; ++a) {
                if (pos[3 * i + a] > c[a]) {
                    key |= 1ll << (3 * l + 2 - a);
                    c[a] += w;
                } else {
                    c[a] -= w;
                }
            }
            w /= 2.0f;
        }
        code[i] = int(key);
    }
    // sort on (code, index) (core/AlgorithmTypes.cuh:32-50 via BuildOctree.cu:657)
    r.order.resize(N);
    std::iota(r.order.begin(), r.order.end(), 0);
    std::stable_sort(r.order.begin(), r.order.end(), [&](int a, int b) { return code[a] < code[b]; });
    r.codes.resize(N);
    r.spos.resize(3 * N);
    r.snrm.resize(3 * N);
    for (int i = 0; i < N; ++i) {
        int o = r.order[i];
        r.codes[i] = code[o];
        for (int a = 0; a < 3; ++a) {
            r.spos[3 * i + a] = pos[3 * o + a];
            r.snrm[3 * i + a] = nn[3 * o + a];
        }
    }
    // unique leaves (BuildOctree.cu:210-265)
    std::vector<int> uKey, uPidx, uPnum;
    for (int i = 0; i < N; ++i) {
        if (i == 0 || r.codes[i] != r.codes[i - 1]) {
            uKey.push_back(r.codes[i]);
            uPidx.push_back(i);
        }
    }
    int U = int(uKey.size());
    uPnum.resize(U);
    for (int u = 0; u < U; ++u) uPnum[u] = (u + 1 < U ? uPidx[u + 1] : N) - uPidx[u];
    // D-level node array (BuildOctree.cu:267-313, 715-761)
    std::vector<unsigned> nodeAddr(U);
    unsigned acc = 0;
    for (int u = 0; u < U; ++u) {
        if (u > 0 && (uKey[u - 1] >> 3) != (uKey[u] >> 3)) acc += 8;
        nodeAddr[u] = acc;
    }
    std::vector<std::vector<OctNode>> levels(D + 1);
    OctNode zero;
    memset(&zero, 0, sizeof(zero));
    unsigned total = nodeAddr[U - 1] + 8;
    levels[D].assign(total, zero);
    std::vector<unsigned> addrFull(total, 0);
    r.point2node.assign(N, -1);
    for (int u = 0; u < U; ++u) {
        int idx = int(nodeAddr[u]) + (uKey[u] & 7);
        OctNode& nd = levels[D][idx];
        nd = zero;  // fields other than key/pidx/pnum of uniqueNodeD taken as zero
        nd.key = uKey[u];
        nd.pidx = uPidx[u];
        nd.pnum = uPnum[u];
        addrFull[idx] = nodeAddr[u];
        r.point2node[uPidx[u]] = idx;
    }
    for (unsigned i = 0; i < total; ++i) {
        levels[D][i].dnum = 1;
        levels[D][i].didx = int(i);
    }
    for (int i = 0; i < N; ++i) {
        int k = i;
        while (r.point2node[k] == -1) --k;
        r.point2node[i] = r.point2node[k];
    }
    // upper levels (BuildOctree.cu:315-390, 763-841)
    int cnt[kMaxD + 2] = {0};
    unsigned allD = total;
    for (int depth = D; depth >= 1; --depth) {
        cnt[depth] = int(allD);
        int UP = int(allD / 8);
        std::vector<OctNode> up(UP, zero);
        for (auto& n : up) n.pidx = n.didx = 0x7fffffff;
        std::vector<OctNode>& cur = levels[depth];
        for (unsigned idx = 0; idx < allD; ++idx) {
            if (cur[idx].pnum == 0) {
                unsigned first = idx - idx % 8, valid = first;
                for (int i = 0; i < 8; ++i) {
                    valid = first + i;
                    if (cur[valid].pnum != 0) break;
                }
                unsigned f = addrFull[valid] / 8;
                if (cur[idx].dnum != 0) {
                    up[f].dnum += cur[idx].dnum;
                    up[f].didx = std::min(up[f].didx, cur[idx].didx);
                }
            } else {
                int sh = 3 * (D - depth);
                int fk = cur[idx].key & (~(7 << sh));
                int f = int(addrFull[idx] / 8);
                int son = (cur[idx].key >> sh) & 7;
                up[f].key = fk;
                up[f].pnum += cur[idx].pnum;
                up[f].pidx = std::min(up[f].pidx, cur[idx].pidx);
                up[f].dnum += cur[idx].dnum;
                up[f].didx = std::min(up[f].didx, cur[idx].didx);
                up[f].children[son] = int(idx);
            }
        }
        std::vector<unsigned> upAddr(UP);
        unsigned a2 = 0;
        int sh2 = 3 * (D - depth + 2);
        for (int u = 0; u < UP; ++u) {
            if (u > 0 && (up[u - 1].key >> sh2) != (up[u].key >> sh2)) a2 += 8;
            upAddr[u] = a2;
        }
        if (depth > 1) {
            unsigned tp = upAddr[UP - 1] + 8;
            levels[depth - 1].assign(tp, zero);
            addrFull.assign(tp, 0);
            int sh1 = 3 * (D - depth + 1);
            for (int u = 0; u < UP; ++u) {
                int index = int(upAddr[u]) + ((up[u].key >> sh1) & 7);
                levels[depth - 1][index] = up[u];
                addrFull[index] = upAddr[u];
                for (int i = 0; i < 8; ++i) {
                    int c = up[u].children[i];
                    if (c != 0) cur[c].parent = index;  // Q1: local index 0 reads as "no child"
                }
            }
            allD = tp;
        } else {
            levels[0].assign(1, up[0]);
            allD = 1;
        }
    }
    cnt[0] = 1;
    r.base[0] = 0;
    for (int i = 1; i <= D; ++i) r.base[i] = r.base[i - 1] + cnt[i - 1];
    r.base[D + 1] = r.base[D] + cnt[D];
    for (int i = 0; i <= D; ++i) r.count[i] = cnt[i];
    r.M = r.base[D] + cnt[D];
    r.nodes.resize(r.M);
    for (int d = 0; d <= D; ++d)
        std::copy(levels[d].begin(), levels[d].end(), r.nodes.begin() + r.base[d]);
    std::vector<OctNode>& NA = r.nodes;
    // Q1 bookkeeping: a valid level-local slot 0 whose parent is not level-local 0
    r.q1_count = 0;
    r.q1_levels = 0;
    // updateNodeArrayParentAndChildrenKernel (BuildOctree.cu:392-424)
    for (int idx = 0; idx < r.M; ++idx) {
        if (NA[idx].pnum == 0) continue;
        int depth = depth_of(r, idx);
        if (idx == 0) {
            NA[idx].parent = -1;
            for (int c = 0; c < 8; ++c) NA[idx].children[c] = c + 1;
        } else {
            NA[idx].parent += r.base[depth - 1];
            if (depth < D)
                for (int c = 0; c < 8; ++c)
                    if (NA[idx].children[c] != 0) NA[idx].children[c] += r.base[depth + 1];
        }
    }
    // updateEmptyNodeInfo (BuildOctree.cu:426-488)
    for (int idx = 1; idx < r.M; ++idx) {
        if ((idx - 1) % 8 != 0) continue;
        int v;
        for (v = 0; v < 8; ++v)
            if (NA[idx + v].pnum != 0) break;
        int nowP = NA[idx + v].pidx, nowD = NA[idx + v].didx, commonParent = NA[idx + v].parent;
        int depth = depth_of(r, idx);
        int sh = 3 * (D - depth);
        int baseKey = NA[idx + v].key - (NA[idx + v].key & (7 << sh));
        if (idx == r.base[depth] && v == 0 && depth >= 2) {
            int trueParent = r.base[depth - 1] + ((NA[idx].key >> (sh + 3)) & 7);
            // the first block's parent is the first occupied node of the level above,
            // whose level-local index is its own son key
            if (commonParent != trueParent) {
                r.q1_count += 1;
                r.q1_levels |= 1 << depth;
            }
        }
        for (int j = 0; j < 8; ++j) {
            int index = idx + j;
            if (NA[index].pnum == 0) {
                for (int k = 0; k < 8; ++k) NA[index].children[k] = -1;
            } else {
                int basePos = -1;
                for (int k = 0; k < 8; ++k)
                    if (NA[index].children[k] > 0) {
                        basePos = NA[index].children[k] - k;
                        break;
                    }
                if (basePos < 0 && depth != D) {
                    // Q1b: the reference reads an uninitialised basePos here; the
                    // defined result is the first block of the level below.
                    r.q1b_count += 1;
                    basePos = r.base[depth + 1];
                }
                for (int k = 0; k < 8; ++k) NA[index].children[k] = (depth != D) ? basePos + k : -1;
            }
            NA[index].key = baseKey + (j << sh);
            NA[index].pidx = nowP;
            nowP += NA[index].pnum;
            if (depth != D) {
                NA[index].didx = nowD;
                nowD += NA[index].dnum;
            }
            NA[index].parent = commonParent;
        }
    }
    // neighbours (BuildOctree.cu:490-512, 870-892)
    for (int i = 0; i < 27; ++i) NA[0].neighs[i] = (i == 13) ? 0 : -1;
    for (int depth = 1; depth <= D; ++depth) {
        int sh = 3 * (D - depth);
        for (int o = r.base[depth]; o < r.base[depth] + cnt[depth]; ++o) {
            int s = (NA[o].key >> sh) & 7;
            int p = NA[o].parent;
            for (int i = 0; i < 27; ++i) {
                int np = NA[p].neighs[kLut.parent[s][i]];
                NA[o].neighs[i] = (np != -1) ? NA[np].children[kLut.child[s][i]] : -1;
            }
        }
    }
    r.childrenBuilt.resize(size_t(8) * r.M);
    for (int i = 0; i < r.M; ++i) memcpy(&r.childrenBuilt[8 * size_t(i)], NA[i].children, 32);
    // depth, centre, packed function index (BuildOctree.cu:514-586)
    r.depthOf.resize(r.M);
    r.center.resize(3 * r.M);
    r.encIdx.resize(r.M);
    for (int i = 0; i < r.M; ++i) {
        int d = depth_of(r, i);
        r.depthOf[i] = d;
        node_center(NA[i].key, d, D, &r.center[3 * i]);
        r.encIdx[i] = encoded_index(NA[i].key, d, D);
    }
}

// ---------------------------------------------------------------------------
// Basis (mesh/FunctionData.h, mesh/BinaryNode.h).
// ---------------------------------------------------------------------------
static inline void func_cw(int idx, double& c, double& w) {
    int i = idx + 1, depth = -1;
    while (i) {
        i >>= 1;
        ++depth;
    }
    int off = idx + 1 - (1 << depth);
    w = 1.0 / double(1 << depth);
    c = (0.5 + off) * w;
}

template <class T>
static inline T bspline(T u) {
    // PPolynomial<2>::GaussianApproximation(0.5) normalised to F(0)=1
    // (math/PPolynomial.h:568-582, ComputeVectorField.cpp:81-95)
    u = u < T(0) ? -u : u;
    if (u < T(0.5)) return T(1) - T(4.0 / 3.0) * u * u;
    if (u < T(1.5)) {
        T t = T(1.5) - u;
        return T(2.0 / 3.0) * t * t;
    }
    return T(0);
}
static inline double bspline_d(double u) {
    if (u <= -1.5 || u >= 1.5) return 0.0;
    if (u < -0.5) return (4.0 / 3.0) * (1.5 + u);
    if (u < 0.5) return -(8.0 / 3.0) * u;
    return -(4.0 / 3.0) * (1.5 - u);
}

// <F_a^(da), F_b^(db)> over the real line; exact for the piecewise quartic product.
static double inner(int a, int b, int da, int db) {
    double ca, wa, cb, wb;
    func_cw(a, ca, wa);
    func_cw(b, cb, wb);
    double lo = std::max(ca - 1.5 * wa, cb - 1.5 * wb), hi = std::min(ca + 1.5 * wa, cb + 1.5 * wb);
    if (lo >= hi) return 0.0;
    double br[10];
    int nb = 0;
    br[nb++] = lo;
    br[nb++] = hi;
    for (int k = -1; k <= 1; k += 2) {
        double x1 = ca + 0.5 * k * wa, x2 = cb + 0.5 * k * wb;
        if (x1 > lo && x1 < hi) br[nb++] = x1;
        if (x2 > lo && x2 < hi) br[nb++] = x2;
    }
    std::sort(br, br + nb);
    static const double gx[3] = {-0.7745966692414834, 0.0, 0.7745966692414834};
    static const double gw[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
    double s = 0;
    for (int k = 0; k + 1 < nb; ++k) {
        double x0 = br[k], x1 = br[k + 1];
        if (x1 <= x0) continue;
        double hm = 0.5 * (x1 - x0), xm = 0.5 * (x1 + x0);
        for (int g = 0; g < 3; ++g) {
            double x = xm + hm * gx[g];
            double ua = (x - ca) / wa, ub = (x - cb) / wb;
            double fa = da ? bspline_d(ua) / wa : bspline<double>(ua);
            double fb = db ? bspline_d(ub) / wb : bspline<double>(ub);
            s += gw[g] * hm * fa * fb;
        }
    }
    return s;
}

struct Tables {
    int res = 0;
    std::vector<double> dot, ddot, d2dot;  // [a + res*b]
};
static void build_tables(Tables& t, int D) {
    // FunctionData::setDotTables (mesh/FunctionData.h:134-190), useDotRatios = 0
    int res = (1 << (D + 1)) - 1;
    t.res = res;
    t.dot.assign(size_t(res) * res, 0.0);
    t.ddot.assign(size_t(res) * res, 0.0);
    t.d2dot.assign(size_t(res) * res, 0.0);
#pragma omp parallel for schedule(dynamic, 8)
    for (int i = 0; i < res; ++i) {
        for (int j = 0; j <= i; ++j) {
            double dt = inner(i, j, 0, 0);
            if (std::fabs(dt) < 1e-15) continue;
            size_t i1 = size_t(i) + size_t(res) * j, i2 = size_t(j) + size_t(res) * i;
            t.dot[i1] = t.dot[i2] = dt;
            double dd = inner(i, j, 1, 0);
            t.ddot[i1] = dd;
            t.ddot[i2] = -dd;
            t.d2dot[i1] = t.d2dot[i2] = inner(i, j, 1, 1);
        }
    }
}

static inline void decode3(int e, int D, int o[3]) {
    int s1 = 1 << (D + 1), s2 = 1 << (2 * (D + 1));
    o[0] = e % s1;
    o[1] = (e / s1) % s1;
    o[2] = e / s2;
}

// F_idx evaluated at x, in the function's local coordinate.
template <class T>
static inline T basis_at(int idx, T x) {
    double c, w;
    func_cw(idx, c, w);
    return bspline<T>((x - T(c)) * T(1.0 / w));
}

// ---------------------------------------------------------------------------
// Vector field (mesh/ComputeVectorField.cu:61-100), divergence
// (mesh/ComputeNodesDivergence.cu:25-136).
// ---------------------------------------------------------------------------
template <class Real>
static void vector_field(Result& r) {
    int D = r.D, nD = r.count[D], off = r.base[D];
    r.vfield.assign(size_t(3) * nD, 0.0);
#pragma omp parallel for schedule(dynamic, 256)
    for (int o = 0; o < nD; ++o) {
        const OctNode& nd = r.nodes[off + o];
        int f[3];
        decode3(r.encIdx[off + o], D, f);
        Real v[3] = {0, 0, 0};
        for (int i = 0; i < 27; ++i) {
            int nb = nd.neighs[i];
            if (nb == -1) continue;
            for (int j = 0; j < r.nodes[nb].pnum; ++j) {
                int p = r.nodes[nb].pidx + j;
                Real wgt = basis_at<Real>(f[0], Real(r.spos[3 * p])) * basis_at<Real>(f[1], Real(r.spos[3 * p + 1])) *
                           basis_at<Real>(f[2], Real(r.spos[3 * p + 2]));
                for (int a = 0; a < 3; ++a) v[a] += wgt * Real(r.snrm[3 * p + a]);
            }
        }
        for (int a = 0; a < 3; ++a) r.vfield[3 * size_t(o) + a] = double(v[a]);
    }
}

template <class Real>
static void divergence(Result& r, const Tables& t) {
    int D = r.D, offD = r.base[D];
    r.div.assign(r.M, 0.0);
#pragma omp parallel for schedule(dynamic, 64)
    for (int o = 0; o < r.M; ++o) {
        int f1[3];
        decode3(r.encIdx[o], D, f1);
        Real acc = 0;
        for (int i = 0; i < 27; ++i) {
            int nb = r.nodes[o].neighs[i];
            if (nb == -1) continue;
            for (int j = 0; j < r.nodes[nb].dnum; ++j) {
                int dj = r.nodes[nb].didx + j;
                int f2[3];
                decode3(r.encIdx[offD + dj], D, f2);
                Real s = 0;
                for (int a = 0; a < 3; ++a)
                    s += Real(r.vfield[3 * size_t(dj) + a]) * Real(t.ddot[size_t(f1[a]) + size_t(t.res) * f2[a]]);
                acc += s;
            }
        }
        r.div[o] = double(acc);
    }
}

// ---------------------------------------------------------------------------
// Per-depth Laplacian (mesh/solver/LaplacianSolver.cu:27-74, 120-244) and CG
// (mesh/solver/CGAlgorithm.cu:10-147, 296-408).
// ---------------------------------------------------------------------------
struct Csr {
    int n = 0;
    std::vector<int> rp, col;
    std::vector<double> val;
};
static void assemble(const Result& r, const Tables& t, int depth, Csr& A) {
    int begin = r.base[depth], n = r.count[depth], D = r.D;
    A.n = n;
    A.rp.assign(n + 1, 0);
    A.col.clear();
    A.val.clear();
    for (int row = 0; row < n; ++row) {
        int f1[3];
        decode3(r.encIdx[begin + row], D, f1);
        for (int i = 0; i < 27; ++i) {
            int nb = r.nodes[begin + row].neighs[i];
            if (nb == -1) continue;
            int f2[3];
            decode3(r.encIdx[nb], D, f2);
            size_t s[3];
            for (int a = 0; a < 3; ++a) s[a] = size_t(f1[a]) * t.res + f2[a];
            double L = t.dot[s[0]] * t.dot[s[1]] * t.dot[s[2]] * (t.d2dot[s[0]] + t.d2dot[s[1]] + t.d2dot[s[2]]);
            if (std::fabs(L) > double(float(1e-6))) {
                A.col.push_back(nb - begin);
                A.val.push_back(L);
            }
        }
        A.rp[row + 1] = int(A.col.size());
    }
}

template <class Real>
static void spmv(const Csr& A, const std::vector<Real>& vals, const std::vector<Real>& x, std::vector<Real>& y) {
#pragma omp parallel for schedule(static) if (A.n > 4096)
    for (int i = 0; i < A.n; ++i) {
        Real s = 0;
        for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) s += vals[k] * x[A.col[k]];
        y[i] = s;
    }
}
template <class Real>
static double dotp(const std::vector<Real>& a, const std::vector<Real>& b) {
    double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static) if (a.size() > 4096)
    for (long i = 0; i < long(a.size()); ++i) s += double(a[i] * b[i]);
    return s;
}

// CG from x0 = 0.  Real=float: the reference loop (abs tol on the recursively
// updated residual); Real=double: relative tolerance rel_tol on the same.
template <class Real>
static int cg(const Csr& A, const std::vector<double>& b, std::vector<double>& xout, double abs_tol, double rel_tol,
              int max_iter, double& recres) {
    int n = A.n;
    std::vector<Real> vals(A.val.size());
    for (size_t k = 0; k < vals.size(); ++k) vals[k] = Real(A.val[k]);
    std::vector<Real> x(n, 0), r(n), p(n, 0), Ap(n, 0);
    for (int i = 0; i < n; ++i) r[i] = Real(b[i]);
    double bnorm2 = dotp(r, r);
    Real r1 = Real(bnorm2), r0 = 0;
    // reference: while (r1 > tol * tol) with float tol = 1e-5f (CGAlgorithm.cu:36, 298)
    double thr = abs_tol > 0 ? double(Real(abs_tol) * Real(abs_tol)) : rel_tol * rel_tol * bnorm2;
    int k = 1;
    while (double(r1) > thr && k <= max_iter) {
        if (k > 1) {
            Real beta = r1 / r0;
            for (int i = 0; i < n; ++i) p[i] = r[i] + beta * p[i];
        } else {
            p = r;
        }
        spmv(A, vals, p, Ap);
        double pAp = dotp(p, Ap);
        Real alpha = Real(double(r1) / pAp);
        for (int i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * Ap[i];
        }
        r0 = r1;
        r1 = Real(dotp(r, r));
        ++k;
    }
    recres = std::sqrt(double(r1));
    xout.resize(n);
    for (int i = 0; i < n; ++i) xout[i] = double(x[i]);
    return k - 1;
}

static double true_relres(const Csr& A, const std::vector<double>& b, const std::vector<double>& x) {
    double num = 0, den = 0;
    for (int i = 0; i < A.n; ++i) {
        double s = 0;
        for (int k = A.rp[i]; k < A.rp[i + 1]; ++k) s += A.val[k] * x[A.col[k]];
        num += (b[i] - s) * (b[i] - s);
        den += b[i] * b[i];
    }
    return den > 0 ? std::sqrt(num / den) : std::sqrt(num);
}

// Whether the depth's operator, as the 1e-6 drop leaves it
// (LaplacianSolver.cu:62), is indefinite: its entries depend only on the
// neighbour offset (the inner products are shift-invariant), so it is a
// constant 27-point stencil whose symbol sum_o L_o cos(theta . o) is sampled
// over [0, pi]^3; a negative value means eigenvalues of both signs (at depth
// 9 only the diagonal and the six face entries survive the drop, and
// 4.99 + 2.18 (cos + cos + cos) changes sign).
static bool stencil_indefinite(const Tables& t, int depth) {
    if (depth < 2) return false;
    const int f = (1 << depth) - 1 + (1 << (depth - 1));  // an interior 1-D function index
    double L[3][3][3];
    for (int a = -1; a <= 1; ++a)
        for (int b = -1; b <= 1; ++b)
            for (int c = -1; c <= 1; ++c) {
                const int o[3] = {a, b, c};
                size_t s[3];
                for (int k = 0; k < 3; ++k) s[k] = size_t(f) * t.res + size_t(f + o[k]);
                double v = t.dot[s[0]] * t.dot[s[1]] * t.dot[s[2]] * (t.d2dot[s[0]] + t.d2dot[s[1]] + t.d2dot[s[2]]);
                L[a + 1][b + 1][c + 1] = std::fabs(v) > double(float(1e-6)) ? v : 0.0;
            }
    const int S = 32;
    double lo = 1e300;
    for (int i = 0; i <= S; ++i)
        for (int j = 0; j <= S; ++j)
            for (int k = 0; k <= S; ++k) {
                const double th[3] = {M_PI * i / S, M_PI * j / S, M_PI * k / S};
                double sym = 0;
                for (int a = -1; a <= 1; ++a)
                    for (int b = -1; b <= 1; ++b)
                        for (int c = -1; c <= 1; ++c)
                            sym += L[a + 1][b + 1][c + 1] * std::cos(th[0] * a) * std::cos(th[1] * b) *
                                   std::cos(th[2] * c);
                lo = std::min(lo, sym);
            }
    return lo < 0;
}

// MINRES (Paige and Saunders), FP64, x0 = 0, for the symmetric indefinite
// depths: Lanczos vectors, Givens rotations on the tridiagonal, stopping when
// the estimated residual norm falls to rel_tol ||b||.
static int minres(const Csr& A, const std::vector<double>& b, std::vector<double>& x, double rel_tol, int max_iter,
                  double& recres) {
    const int n = A.n;
    std::vector<double> r1(b), r2(b), y(b), v(n), w(n, 0), w1(n, 0), w2(n, 0);
    x.assign(n, 0.0);
    const double beta1 = std::sqrt(dotp(b, b));
    double beta = beta1, oldb = 0, dbar = 0, epsln = 0, phibar = beta1, cs = -1, sn = 0;
    int itn = 0;
    while (phibar > rel_tol * beta1 && itn < max_iter && beta > 0) {
        ++itn;
        const double s = 1.0 / beta;
        for (int i = 0; i < n; ++i) v[i] = s * y[i];
        spmv(A, A.val, v, y);
        if (itn >= 2)
            for (int i = 0; i < n; ++i) y[i] -= (beta / oldb) * r1[i];
        const double alfa = dotp(v, y);
        for (int i = 0; i < n; ++i) y[i] -= (alfa / beta) * r2[i];
        r1.swap(r2);
        r2 = y;
        oldb = beta;
        beta = std::sqrt(dotp(y, y));
        const double oldeps = epsln;
        const double delta = cs * dbar + sn * alfa;
        const double gbar = sn * dbar - cs * alfa;
        epsln = sn * beta;
        dbar = -cs * beta;
        const double gamma = std::max(std::sqrt(gbar * gbar + beta * beta), 1e-300);
        cs = gbar / gamma;
        sn = beta / gamma;
        const double phi = cs * phibar;
        phibar = sn * phibar;
        for (int i = 0; i < n; ++i) {
            w1[i] = w2[i];
            w2[i] = w[i];
            w[i] = (v[i] - oldeps * w1[i] - delta * w2[i]) / gamma;
            x[i] += phi * w[i];
        }
    }
    recres = phibar;
    return itn;
}

template <class Real>
static void solve_all(Result& r, const Tables& t) {
    int D = r.D;
    r.dx.assign(r.M, 0.0);
    r.cgIters.assign(D + 1, 0);
    r.cgRelRes.assign(D + 1, 0.0);
    r.cgRecRes.assign(D + 1, 0.0);
    r.nnz.assign(D + 1, 0);
    for (int d = 0; d <= D; ++d) {
        Csr A;
        assemble(r, t, d, A);
        r.nnz[d] = int(A.col.size());
        std::vector<double> b(r.div.begin() + r.base[d], r.div.begin() + r.base[d] + r.count[d]), x;
        double rec = 0;
        int it;
        if (sizeof(Real) == 4)
            it = cg<float>(A, b, x, 1e-5, 0, 10000, rec);  // the reference's rule, also where it cannot converge
        else if (stencil_indefinite(t, d))
            it = minres(A, b, x, 1e-12, 200000, rec);
        else
            it = cg<double>(A, b, x, 0, 1e-12, 100000, rec);
        r.cgIters[d] = it;
        r.cgRecRes[d] = rec;
        r.cgRelRes[d] = true_relres(A, b, x);
        std::copy(x.begin(), x.end(), r.dx.begin() + r.base[d]);
    }
}

// ---------------------------------------------------------------------------
// Implicit function at points and iso-value (LaplacianSolver.cu:85-118, 269-305).
// ---------------------------------------------------------------------------
template <class Real>
static Real chi_at(const Result& r, int startNode, const float* pos) {
    Real val = 0;
    int now = startNode;
    while (now != -1) {
        for (int i = 0; i < 27; ++i) {
            int nb = r.nodes[now].neighs[i];
            if (nb == -1) continue;
            int f[3];
            decode3(r.encIdx[nb], r.D, f);
            val += Real(r.dx[nb]) * basis_at<Real>(f[0], Real(pos[0])) * basis_at<Real>(f[1], Real(pos[1])) *
                   basis_at<Real>(f[2], Real(pos[2]));
        }
        now = r.nodes[now].parent;
    }
    return val;
}

template <class Real>
static void points_chi(Result& r) {
    r.chiPts.assign(r.N, 0.0);
    double sum = 0;
#pragma omp parallel for reduction(+ : sum) schedule(dynamic, 1024)
    for (int i = 0; i < r.N; ++i) {
        Real v = chi_at<Real>(r, r.base[r.D] + r.point2node[i], &r.spos[3 * i]);
        r.chiPts[i] = double(v);
        sum += double(v);
    }
    r.iso = sum / r.N;
}

// ---------------------------------------------------------------------------
// Vertex / edge / face ownership (mesh/BuildMeshGeometry.cu:42-400).
// ---------------------------------------------------------------------------
static inline double sqdist(const float* a, const float* b) {
    float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return double(dx * dx + dy * dy + dz * dz);
}

// The largest value stored to *p.  The pointer loops below run their index
// upward, so in a sequential run the last store is the largest: taking the
// maximum gives that result under any thread schedule.
static inline void store_max(int* p, int v) {
    int cur = __atomic_load_n(p, __ATOMIC_RELAXED);
    while (cur < v && !__atomic_compare_exchange_n(p, &cur, v, true, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
    }
}

// Ownership runs per node in parallel and records a bit per corner / edge /
// face; the owned elements are then appended in (node, slot) order, the order
// of the reference's stream compaction, so the arrays do not depend on the
// thread count.
static void geometry(Result& r) {
    int M = r.M, D = r.D;
    std::vector<OctNode>& NA = r.nodes;
    auto ctr = [&](int i) { return &r.center[3 * size_t(i)]; };
    static const float kZero[3] = {0, 0, 0};
    auto centre_of = [&](int idx) -> const float* { return NA[idx].neighs[13] != -1 ? ctr(NA[idx].neighs[13]) : kZero; };
    auto corner = [&](int idx, int i, float vp[3]) {
        float half = 1.0f / float(1 << (r.depthOf[idx] + 1));
        const float* nc = centre_of(idx);
        vp[0] = nc[0] + (2 * (i & 1) - 1) * half;
        vp[1] = nc[1] + (2 * ((i & 2) >> 1) - 1) * half;
        vp[2] = nc[2] + (2 * ((i & 4) >> 2) - 1) * half;
    };
    // the node among idx's 27 neighbours closer than W to p with the smallest key
    auto owner_near = [&](int idx, const float* p, float W2) {
        int ownKey = 0x7fffffff, own = -1;
        for (int j = 0; j < 27; ++j) {
            int nb = NA[idx].neighs[j];
            if (nb != -1 && sqdist(p, ctr(nb)) < double(W2) && ownKey > NA[nb].key) {
                ownKey = NA[nb].key;
                own = nb;
            }
        }
        return own;
    };
    std::vector<uint16_t> mask(size_t(M), 0);
    // vertices
#pragma omp parallel for schedule(dynamic, 1024)
    for (int idx = 0; idx < M; ++idx) {
        float W = 1.0f / float(1 << r.depthOf[idx]);
        uint16_t m = 0;
        for (int i = 0; i < 8; ++i) {
            float vp[3];
            corner(idx, i, vp);
            if (owner_near(idx, vp, W * W) == idx) m |= uint16_t(1u << i);
        }
        mask[idx] = m;
    }
    r.verts.clear();
    for (int idx = 0; idx < M; ++idx)
        for (int i = 0; i < 8; ++i)
            if (mask[idx] >> i & 1) {
                float vp[3];
                corner(idx, i, vp);
                r.verts.push_back({{vp[0], vp[1], vp[2]}, idx, i, r.depthOf[idx]});
            }
    const int nv = int(r.verts.size());
#pragma omp parallel for schedule(dynamic, 1024)
    for (int v = 0; v < nv; ++v) {
        const Vtx& vx = r.verts[v];
        int owner = vx.owner, depth = r.depthOf[owner];
        float W = 1.0f / float(1 << depth), W2 = W * W;
        for (int i = 0; i < 27; ++i) {
            int nb = NA[owner].neighs[i];
            if (nb == -1 || !(sqdist(vx.pos, ctr(nb)) < double(W2))) continue;
            const float* c = ctr(nb);
            int index = 0;
            if (c[0] - vx.pos[0] < 0) index |= 1;
            if (c[2] - vx.pos[2] < 0) index |= 4;
            if (c[1] - vx.pos[1] < 0) index += (index & 1) ? 1 : 3;
            store_max(&NA[nb].vertices[index], v + 1);
        }
    }
    // edges (depth D only)
    float halfD = 1.0f / float(1 << (D + 1)), WD = 1.0f / float(1 << D), WD2 = WD * WD;
    auto edge_center = [&](const float* nc, int kind, float half, float out[3]) {
        int orient = kind >> 2, off[2] = {kind & 1, (kind & 2) >> 1}, dim = 0;
        for (int a = 0; a < 3; ++a) {
            int m = 0;
            if (orient != a) m = 2 * off[dim++] - 1;
            out[a] = nc[a] + m * half;
        }
    };
    const int oD = r.base[D], nD = r.count[D];
#pragma omp parallel for schedule(dynamic, 1024)
    for (int o = oD; o < oD + nD; ++o) {
        uint16_t m = 0;
        for (int e = 0; e < 12; ++e) {
            float ec[3];
            edge_center(centre_of(o), e, halfD, ec);
            if (owner_near(o, ec, WD2) == o) m |= uint16_t(1u << e);
        }
        mask[o] = m;
    }
    r.edges.clear();
    for (int o = oD; o < oD + nD; ++o)
        for (int e = 0; e < 12; ++e)
            if (mask[o] >> e & 1) r.edges.push_back({e, o});
    const int ne = int(r.edges.size());
#pragma omp parallel for schedule(dynamic, 1024)
    for (int k = 0; k < ne; ++k) {
        int owner = r.edges[k].owner, kind = r.edges[k].kind, orient = kind >> 2;
        float ec[3];
        edge_center(centre_of(owner), kind, halfD, ec);
        for (int i = 0; i < 27; ++i) {
            int nb = NA[owner].neighs[i];
            if (nb == -1 || !(sqdist(ec, ctr(nb)) < double(WD2))) continue;
            int index = orient << 2, dim = 0;
            for (int a = 0; a < 3; ++a)
                if (orient != a) {
                    if (ctr(nb)[a] - ec[a] < 0) index |= (1 << dim);
                    ++dim;
                }
            store_max(&NA[nb].edges[index], k + 1);
        }
    }
    // faces
    auto face_center = [&](int idx, int i, float fc[3]) {
        float half = 1.0f / float(1 << (r.depthOf[idx] + 1));
        const float* nc = centre_of(idx);
        fc[0] = nc[0];
        fc[1] = nc[1];
        fc[2] = nc[2];
        fc[i >> 1] += (2 * (i & 1) - 1) * half;
    };
#pragma omp parallel for schedule(dynamic, 1024)
    for (int idx = 0; idx < M; ++idx) {
        float W = 1.0f / float(1 << r.depthOf[idx]);
        uint16_t m = 0;
        for (int i = 0; i < 6; ++i) {
            float fc[3];
            face_center(idx, i, fc);
            if (owner_near(idx, fc, W * W) == idx) m |= uint16_t(1u << i);
        }
        mask[idx] = m;
    }
    r.faces.clear();
    for (int idx = 0; idx < M; ++idx) {
        if (!mask[idx]) continue;
        int parent = NA[idx].parent, son = (NA[idx].key >> (3 * (D - r.depthOf[idx]))) & 7;
        for (int i = 0; i < 6; ++i)
            if (mask[idx] >> i & 1) {
                int hp = (parent == -1) ? -1 : (kParentFaceKind[son][i] != -1 ? 1 : -1);
                r.faces.push_back({i, idx, hp});
            }
    }
    const int nf = int(r.faces.size());
#pragma omp parallel for schedule(dynamic, 1024)
    for (int k = 0; k < nf; ++k) {
        int owner = r.faces[k].owner, kind = r.faces[k].kind, orient = kind >> 1;
        float W = 1.0f / float(1 << r.depthOf[owner]), W2 = W * W;
        float fc[3];
        face_center(owner, kind, fc);
        for (int i = 0; i < 27; ++i) {
            int nb = NA[owner].neighs[i];
            if (nb == -1 || !(sqdist(fc, ctr(nb)) < double(W2))) continue;
            int index = orient << 1;
            if (ctr(nb)[orient] - fc[orient] < 0) ++index;
            store_max(&NA[nb].faces[index], k + 1);
        }
    }
}

// ---------------------------------------------------------------------------
// Vertex implicit values, MC and subdivision (mesh/ComputeTriangleIndices.cu).
// ---------------------------------------------------------------------------
template <class Real>
static Real chi_walk_sum(const Result& r, int now, const float* pos) {
    Real val = 0;
    for (int i = 0; i < 27; ++i) {
        int nb = r.nodes[now].neighs[i];
        if (nb == -1) continue;
        int f[3];
        decode3(r.encIdx[nb], r.D, f);
        val += Real(r.dx[nb]) * basis_at<Real>(f[0], Real(pos[0])) * basis_at<Real>(f[1], Real(pos[1])) *
               basis_at<Real>(f[2], Real(pos[2]));
    }
    return val;
}

template <class Real>
static void vertex_values(Result& r) {
    // ComputeVertexImplicitFunctionValueKernel (ComputeTriangleIndices.cu:346-399), Q8
    int nv = int(r.verts.size());
    r.vvalue.assign(nv, 0.0);
#pragma omp parallel for schedule(dynamic, 256)
    for (int v = 0; v < nv; ++v) {
        const Vtx& vx = r.verts[v];
        Real val = 0;
        int depth = vx.depth, ex = kChildrenVertexKind[vx.kind], now = vx.owner;
        if (now > 0) {
            while (now != -1) {
                val += chi_walk_sum<Real>(r, now, vx.pos);
                now = r.nodes[now].parent;
            }
            now = vx.owner;
            while (depth < r.D) {
                ++depth;
                now = r.nodes[now].children[ex];
                if (now == -1) break;
                val += chi_walk_sum<Real>(r, now, vx.pos);
            }
        }
        r.vvalue[v] = double(val - Real(r.iso));
    }
}

static inline void interp(const float* p1, const float* p2, int dim, float v1, float v2, float* out) {
    // ComputeTriangleIndices.cu:464-474
    for (int i = 0; i < 3; ++i)
        if (i != dim) out[i] = p1[i];
    float pivot = v1 / (v1 - v2), another = 1 - pivot;
    out[dim] = p2[dim] * pivot + p1[dim] * another;
}

static void append_batch(Result& r, const std::vector<float>& V, const std::vector<int>& T) {
    // insertTriangle (ComputeTriangleIndices.cu:1170-1204, 1247-1282)
    int prev = int(r.meshV.size() / 3), nv = int(V.size() / 3), nt = int(T.size() / 3);
    int keptV = 0, keptT = 0;
    for (int i = 0; i < nv; ++i) {
        if (std::fabs(V[3 * i]) < float(1e-6)) continue;
        r.meshV.insert(r.meshV.end(), V.begin() + 3 * i, V.begin() + 3 * i + 3);
        ++keptV;
    }
    for (int i = 0; i < nt; ++i) {
        bool ok = true;
        int t[3];
        for (int k = 0; k < 3; ++k) {
            t[k] = T[3 * i + k] + prev;
            if (t[k] < 0 || t[k] >= nv + prev) ok = false;
        }
        if (!ok) continue;
        r.meshT.insert(r.meshT.end(), t, t + 3);
        ++keptT;
    }
    r.batchV.push_back(keptV);
    r.batchT.push_back(keptT);
}

// A subdivision set: roots rebuilt together into one virtual array whose ids
// start at NodeArraySize (ComputeTriangleIndices.cu:599-635, 1110-1168).
struct SubSet {
    std::vector<VNode> nodes;
    std::vector<int> depth;
    std::vector<float> center;   // 3 per node
    std::vector<int> replaced;   // finer path: real id of the copied root
    std::vector<int> isRoot;
    int levelStart[kMaxD + 2] = {0}, levelCount[kMaxD + 2] = {0};
    int coarseRootReal = -1;     // coarse path: real id that virtual id NAS maps to
};

template <class Real>
static void mc_on_set(Result& r, SubSet& S, bool coarse) {
    int NAS = r.M, D = r.D;
    std::vector<OctNode>& NA = r.nodes;
    auto vctr = [&](int gid) -> const float* { return &S.center[3 * size_t(gid - NAS)]; };
    // neighbours (computeRebuildNeighbor, :669-696)
    int firstDepth = coarse ? S.depth[0] : 0;
    for (int depth = firstDepth; depth <= D; ++depth) {
        int sh = 3 * (D - depth);
        // a level reads only the level above it
#pragma omp parallel for schedule(static)
        for (int o = S.levelStart[depth]; o < S.levelStart[depth] + S.levelCount[depth]; ++o) {
            VNode& n = S.nodes[o];
            int s = (n.key >> sh) & 7, p = n.parent;
            for (int i = 0; i < 27; ++i) {
                int np = (p < NAS) ? NA[p].neighs[kLut.parent[s][i]] : S.nodes[p - NAS].neighs[kLut.parent[s][i]];
                if (np != -1)
                    n.neighs[i] = (np < NAS) ? NA[np].children[kLut.child[s][i]]
                                             : S.nodes[np - NAS].children[kLut.child[s][i]];
                else
                    n.neighs[i] = -1;
            }
        }
    }
    float half = 1.0f / float(1 << (D + 1)), W = 1.0f / float(1 << D), W2 = W * W;
    int dStart = S.levelStart[D], dCount = S.levelCount[D];
    static const float kZero[3] = {0, 0, 0};
    auto vcenter_or_zero = [&](int nb) -> const float* { return (nb != -1 && nb >= NAS) ? vctr(nb) : kZero; };
    auto any_center = [&](int nb) -> const float* { return nb < NAS ? &r.center[3 * size_t(nb)] : vctr(nb); };
    // the virtual node among o's 27 neighbours closer than W to p with the smallest key
    auto owner_near = [&](int o, const float* p) {
        int ownKey = 0x7fffffff, own = -1;
        for (int j = 0; j < 27; ++j) {
            int nb = S.nodes[o].neighs[j];
            if (nb != -1 && sqdist(p, vcenter_or_zero(nb)) < double(W2)) {
                if (nb < NAS) continue;
                int k = S.nodes[nb - NAS].key;
                if (ownKey > k) {
                    ownKey = k;
                    own = nb;
                }
            }
        }
        return own;
    };
    auto corner = [&](int o, int i, float vp[3]) {
        const float* nc = vcenter_or_zero(S.nodes[o].neighs[13]);
        vp[0] = nc[0] + (2 * (i & 1) - 1) * half;
        vp[1] = nc[1] + (2 * ((i & 2) >> 1) - 1) * half;
        vp[2] = nc[2] + (2 * ((i & 4) >> 2) - 1) * half;
    };
    // ownership per node in parallel, appended in (node, slot) order as in geometry()
    std::vector<uint16_t> mask(size_t(dCount), 0);
    // vertex owners (initSubdivideVertexOwner, :698-763)
#pragma omp parallel for schedule(dynamic, 1024)
    for (int o = dStart; o < dStart + dCount; ++o) {
        uint16_t m = 0;
        for (int i = 0; i < 8; ++i) {
            float vp[3];
            corner(o, i, vp);
            if (owner_near(o, vp) == NAS + o) m |= uint16_t(1u << i);
        }
        mask[o - dStart] = m;
    }
    std::vector<Vtx> V;
    for (int o = dStart; o < dStart + dCount; ++o)
        for (int i = 0; i < 8; ++i)
            if (mask[o - dStart] >> i & 1) {
                float vp[3];
                corner(o, i, vp);
                V.push_back({{vp[0], vp[1], vp[2]}, NAS + o, i, D});
            }
    const int setId = int(r.subSets.size() / 4) - 1;
    const size_t subV0 = r.subV.size();
    for (const Vtx& vx : V) {
        int rec[7] = {setId, S.nodes[vx.owner - NAS].key, vx.kind, 0, 0, 0, 0x7fc00000};
        memcpy(&rec[3], vx.pos, 12);
        r.subV.insert(r.subV.end(), rec, rec + 7);
    }
    if (V.empty()) return;
    // pointers land on depth-D virtual nodes only, whose slots start at 0
    const int nv = int(V.size());
#pragma omp parallel for schedule(dynamic, 1024)
    for (int v = 0; v < nv; ++v) {
        VNode& ow = S.nodes[V[v].owner - NAS];
        for (int i = 0; i < 27; ++i) {
            int nb = ow.neighs[i];
            if (nb == -1) continue;
            const float* c = any_center(nb);
            if (!(sqdist(V[v].pos, c) < double(W2))) continue;
            int index = 0;
            if (c[0] - V[v].pos[0] < 0) index |= 1;
            if (c[2] - V[v].pos[2] < 0) index |= 4;
            if (c[1] - V[v].pos[1] < 0) index += (index & 1) ? 1 : 3;
            if (nb >= NAS) store_max(&S.nodes[nb - NAS].vertices[index], v + 1);
        }
    }
    // edges (initSubdivideEdgeArray / maintainSubdivideEdgeNodePointer, :812-949)
    auto edge_center = [&](const float* nc, int kind, float out[3]) {
        int orient = kind >> 2, off[2] = {kind & 1, (kind & 2) >> 1}, dim = 0;
        for (int a = 0; a < 3; ++a) {
            int m = 0;
            if (orient != a) m = 2 * off[dim++] - 1;
            out[a] = nc[a] + m * half;
        }
    };
#pragma omp parallel for schedule(dynamic, 1024)
    for (int o = dStart; o < dStart + dCount; ++o) {
        const float* nc = vcenter_or_zero(S.nodes[o].neighs[13]);
        uint16_t m = 0;
        for (int e = 0; e < 12; ++e) {
            float ec[3];
            edge_center(nc, e, ec);
            if (owner_near(o, ec) == NAS + o) m |= uint16_t(1u << e);
        }
        mask[o - dStart] = m;
    }
    std::vector<Edge> E;
    for (int o = dStart; o < dStart + dCount; ++o)
        for (int e = 0; e < 12; ++e)
            if (mask[o - dStart] >> e & 1) E.push_back({e, NAS + o});
    for (const Edge& ed : E) {
        const int rec[3] = {setId, S.nodes[ed.owner - NAS].key, ed.kind};
        r.subE.insert(r.subE.end(), rec, rec + 3);
    }
    if (E.empty()) return;
    const int ne = int(E.size());
#pragma omp parallel for schedule(dynamic, 1024)
    for (int k = 0; k < ne; ++k) {
        VNode& ow = S.nodes[E[k].owner - NAS];
        int orient = E[k].kind >> 2;
        const float* nc = any_center(ow.neighs[13]);
        float ec[3];
        edge_center(nc, E[k].kind, ec);
        for (int i = 0; i < 27; ++i) {
            int nb = ow.neighs[i];
            if (nb == -1) continue;
            const float* c = any_center(nb);
            if (!(sqdist(ec, c) < double(W2))) continue;
            int index = orient << 2, dim = 0;
            for (int a = 0; a < 3; ++a)
                if (orient != a) {
                    if (c[a] - ec[a] < 0) index |= (1 << dim);
                    ++dim;
                }
            if (nb >= NAS) store_max(&S.nodes[nb - NAS].edges[index], k + 1);
        }
    }
    // vertex values (computeSubdivideVertexImplicitFunctionValue, :956-1038)
    std::vector<float> vv(V.size());
#pragma omp parallel for schedule(dynamic, 256)
    for (int v = 0; v < nv; ++v) {
        Real val = 0;
        int now = V[v].owner;
        if (now > 0) {
            while (now != -1) {
                for (int k = 0; k < 27; ++k) {
                    int nb = (now < NAS) ? NA[now].neighs[k] : S.nodes[now - NAS].neighs[k];
                    if (nb == -1) continue;
                    if (coarse) {
                        if (nb == NAS) nb = S.coarseRootReal;
                    } else if (nb >= NAS && S.isRoot[nb - NAS]) {
                        nb = S.replaced[nb - NAS];
                    }
                    if (nb >= NAS) continue;
                    int f[3];
                    decode3(r.encIdx[nb], D, f);
                    val += Real(r.dx[nb]) * basis_at<Real>(f[0], Real(V[v].pos[0])) *
                           basis_at<Real>(f[1], Real(V[v].pos[1])) * basis_at<Real>(f[2], Real(V[v].pos[2]));
                }
                now = (now < NAS) ? NA[now].parent : S.nodes[now - NAS].parent;
            }
        }
        vv[v] = float(double(val) - r.iso);
        memcpy(&r.subV[subV0 + 7 * v + 6], &vv[v], 4);
    }
    // crossings and triangles (:1040-1096, 528-553)
    std::vector<int> vexAddr(E.size());
    std::vector<float> VB;
    int acc = 0;
    for (size_t k = 0; k < E.size(); ++k) {
        VNode& ow = S.nodes[E[k].owner - NAS];
        int v1 = ow.vertices[kEdgeVertex[E[k].kind][0]] - 1, v2 = ow.vertices[kEdgeVertex[E[k].kind][1]] - 1;
        vexAddr[k] = acc;
        if (vv[v1] * vv[v2] <= 0) {
            float p[3];
            interp(V[v1].pos, V[v2].pos, E[k].kind >> 2, vv[v1], vv[v2], p);
            VB.insert(VB.end(), p, p + 3);
            ++acc;
        }
    }
    if (coarse && acc == 0) return;
    std::vector<int> TB;
    for (int o = dStart; o < dStart + dCount; ++o) {
        VNode& n = S.nodes[o];
        int cube = 0;
        for (int i = 0; i < 8; ++i)
            if (vv[n.vertices[i] - 1] < 0) cube |= 1 << i;
        int nt = mc_count(cube);
        for (int t = 0; t < nt; ++t)
            for (int k = 0; k < 3; ++k) TB.push_back(vexAddr[n.edges[mc_edge(cube, 3 * t + k)] - 1]);
    }
    append_batch(r, VB, TB);
}

template <class Real>
static void marching_cubes(Result& r) {
    int D = r.D, offD = r.base[D], nD = r.count[D];
    std::vector<OctNode>& NA = r.nodes;
    vertex_values<Real>(r);
    r.subSets.clear();
    r.subRoots.clear();
    r.subV.clear();
    r.subE.clear();
    r.meshV.clear();
    r.meshT.clear();
    r.batchV.clear();
    r.batchT.clear();
    // depth-D crossings (:401-422, 440-474)
    int nE = int(r.edges.size());
    std::vector<int> vexAddr(nE), crossing(nE, 0);
    std::vector<float> VB;
    int acc = 0;
    for (int k = 0; k < nE; ++k) {
        const OctNode& ow = NA[r.edges[k].owner];
        int kind = r.edges[k].kind;
        int v1 = ow.vertices[kEdgeVertex[kind][0]] - 1, v2 = ow.vertices[kEdgeVertex[kind][1]] - 1;
        vexAddr[k] = acc;
        // Where the level-local child 0 parent rule (quirk Q1) has broken the
        // neighbour tables, slots stay unassigned and the reference reads out of
        // bounds; here an unassigned slot makes the edge, cube, face or corner
        // absent (the same rule as the GPU kernels).
        if (v1 < 0 || v2 < 0) continue;
        float f1 = float(r.vvalue[v1]), f2 = float(r.vvalue[v2]);
        if (f1 * f2 <= 0) {
            crossing[k] = 1;
            float p[3];
            interp(r.verts[v1].pos, r.verts[v2].pos, kind >> 2, f1, f2, p);
            VB.insert(VB.end(), p, p + 3);
            ++acc;
        }
    }
    // triangles and face-intersection propagation (:424-438, 476-526)
    std::vector<int> TB;
    std::vector<int> hasInter(r.faces.size(), 0);
    r.cube.assign(nD, 0);
    for (int o = 0; o < nD; ++o) {
        const OctNode& n = NA[offD + o];
        int cube = 0;
        bool ok = true;
        for (int i = 0; i < 8; ++i) {
            if (n.vertices[i] == 0) ok = false;
            else if (float(r.vvalue[n.vertices[i] - 1]) < 0) cube |= 1 << i;
        }
        for (int e = 0; e < 12; ++e)
            if (n.edges[e] == 0) ok = false;
        if (ok)
            for (int k = 0; k < 3 * mc_count(cube); ++k)
                if (!crossing[n.edges[mc_edge(cube, k)] - 1]) ok = false;
        if (!ok) cube = 0;
        r.cube[o] = cube;
        int nt = mc_count(cube), edgeHas[12] = {0};
        for (int t = 0; t < nt; ++t)
            for (int k = 0; k < 3; ++k) {
                int e = mc_edge(cube, 3 * t + k);
                edgeHas[e] = 1;
                TB.push_back(vexAddr[n.edges[e] - 1]);
            }
        for (int i = 0; i < 6; ++i) {
            int mark = 0;
            for (int j = 0; j < 4; ++j) mark |= edgeHas[kFaceEdges[i][j]];
            if (!mark) continue;
            int pn = n.parent, cf = n.faces[i] - 1;
            if (cf < 0) continue;
            hasInter[cf] = 1;
            while (r.faces[cf].hasParentFace != -1 && pn >= 0) {
                cf = NA[pn].faces[i] - 1;
                if (cf < 0) break;
                pn = NA[pn].parent;
                hasInter[cf] = 1;
            }
        }
    }
    append_batch(r, VB, TB);
    // subdivision roots (ProcessLeafNodesAtOtherDepth :555-587, copy_if :1284-1346)
    std::vector<int> roots;
    for (int idx = 0; idx < offD; ++idx) {
        OctNode& n = NA[idx];
        int i0 = 0, hasTri = 0;
        while (i0 < 8 && n.vertices[i0] == 0) ++i0;
        if (i0 < 8) {
            int sign = float(r.vvalue[n.vertices[i0] - 1]) < 0 ? -1 : 1;
            for (int i = i0 + 1; i < 8; ++i)
                if (n.vertices[i] != 0 && sign * float(r.vvalue[n.vertices[i] - 1]) < 0) {
                    hasTri = 1;
                    break;
                }
        }
        n.hasTriangle = hasTri;
        int hi = 0;
        for (int i = 0; i < 6; ++i)
            if (n.faces[i] != 0 && hasInter[n.faces[i] - 1]) {
                hi = 1;
                break;
            }
        n.hasIntersection = hi;
        bool leaf = true;
        for (int c = 0; c < 8; ++c)
            if (n.children[c] != -1) leaf = false;
        // a root must be its own neighbour: the virtual subtree is built from it
        if (leaf && (hasTri || hi) && n.neighs[13] == idx) roots.push_back(idx);
    }
    r.rootNodes = roots;
    r.nRoots = int(roots.size());
    std::vector<int> rootDepth(roots.size());
    for (size_t i = 0; i < roots.size(); ++i) {
        int self = NA[roots[i]].neighs[13];
        // precomputeSubdivideDepth (:589-597) reads the depth through neighs[13]
        rootDepth[i] = r.depthOf[self >= 0 ? self : roots[i]];
        if (self != roots[i]) r.q1_count += 0;
    }
    int finerDepth = D - 4;
    int NAS = r.M;
    // coarse roots, one at a time (:1475-1799)
    bool sawDepth1 = false;
    r.nCoarseRoots = 0;
    for (size_t i = 0; i < roots.size(); ++i) {
        int rd = rootDepth[i];
        if (rd >= finerDepth) break;
        r.nCoarseRoots++;
        if (rd == 1) sawDepth1 = true;
        if (rd >= 2 && sawDepth1) r.q16_flag = 1;
        const OctNode& rootNode = NA[roots[i]];
        int rootIndex = rootNode.neighs[13], rootParent = rootNode.parent, rootKey = rootNode.key;
        int rootSon = (rootKey >> (3 * (D - rd))) & 7;
        SubSet S;
        int total = 0, num = 1;
        for (int d = rd; d <= D; ++d) {
            S.levelStart[d] = total;
            S.levelCount[d] = num;
            total += num;
            num <<= 3;
        }
        VNode vz;
        memset(&vz, 0, sizeof(vz));
        S.nodes.assign(total, vz);
        S.depth.assign(total, 0);
        S.center.assign(3 * size_t(total), 0.f);
        S.coarseRootReal = rootIndex;
        for (int d = rd; d <= D; ++d) {
            for (int off = 0; off < S.levelCount[d]; ++off) {
                int idx = S.levelStart[d] + off;
                VNode& n = S.nodes[idx];
                for (int k = 0; k < 8; ++k)
                    n.children[k] = (d < D) ? NAS + S.levelStart[d + 1] + (off << 3) + k : -1;
                if (idx != 0) n.parent = NAS + S.levelStart[d - 1] + (off >> 3);
                n.key = rootKey | (off << (3 * (D - d)));
                S.depth[idx] = d;
                node_center(n.key, d, D, &S.center[3 * size_t(idx)]);
            }
        }
        S.nodes[0].parent = rootParent;
        {
            const int desc[4] = {rd, 1, int(r.subRoots.size()), 1};
            r.subSets.insert(r.subSets.end(), desc, desc + 4);
            r.subRoots.push_back(roots[i]);
        }
        NA[rootParent].children[rootSon] = NAS;
        mc_on_set<Real>(r, S, true);
        NA[rootParent].children[rootSon] = rootIndex;
    }
    // finer roots, batched per root depth (:1801-2194)
    for (int rd = std::max(finerDepth, 1); rd < D; ++rd) {
        std::vector<int> sel;
        for (size_t i = 0; i < roots.size(); ++i)
            if (rootDepth[i] == rd) sel.push_back(int(i));
        if (sel.empty()) continue;
        int nr = int(sel.size());
        SubSet S;
        int total = 0;
        for (int d = 0; d <= D; ++d) {
            int c = (d >= rd) ? nr * (1 << (3 * (d - rd))) : 0;
            S.levelStart[d] = total;
            S.levelCount[d] = c;
            total += c;
        }
        VNode vz;
        memset(&vz, 0, sizeof(vz));
        S.nodes.assign(total, vz);
        S.depth.assign(total, 0);
        S.center.assign(3 * size_t(total), 0.f);
        S.replaced.assign(total, 0);
        S.isRoot.assign(total, 0);
        for (int q = 0; q < nr; ++q) {
            const OctNode& rn = NA[roots[sel[q]]];
            int nowIdx = S.levelStart[rd] + q;
            VNode& v = S.nodes[nowIdx];
            v.key = rn.key;
            v.parent = rn.parent;
            memcpy(v.children, rn.children, sizeof(v.children));
            memcpy(v.neighs, rn.neighs, sizeof(v.neighs));
            memcpy(v.vertices, rn.vertices, sizeof(v.vertices));
            memcpy(v.edges, rn.edges, sizeof(v.edges));
            int replacedId = rn.neighs[13];
            v.neighs[13] = NAS + nowIdx;
            S.replaced[nowIdx] = replacedId;
            S.depth[nowIdx] = rd;
            S.isRoot[nowIdx] = 1;
            node_center(rn.key, rd, D, &S.center[3 * size_t(nowIdx)]);
            int son = (rn.key >> (3 * (D - rd))) & 7;
            NA[rn.parent].children[son] = NAS + nowIdx;  // Q13: not restored by the reference
            int per = 1;
            for (int d = rd + 1; d <= D; ++d) {
                int now = S.levelStart[d] + q * (per * 8);
                for (int j = 0; j < per * 8; j += 8) {
                    int pIdx = S.levelStart[d - 1] + q * per + j / 8;
                    int pGid = S.nodes[pIdx].neighs[13], pKey = S.nodes[pIdx].key;
                    for (int k = 0; k < 8; ++k) {
                        int t = now + j + k;
                        int key = pKey | (k << (3 * (D - d)));
                        S.nodes[t].parent = pGid;
                        S.nodes[t].key = key;
                        S.nodes[t].neighs[13] = NAS + t;
                        S.replaced[t] = replacedId;
                        S.depth[t] = d;
                        node_center(key, d, D, &S.center[3 * size_t(t)]);
                        S.nodes[pIdx].children[k] = NAS + t;
                    }
                }
                per *= 8;
            }
        }
        {
            const int desc[4] = {rd, 0, int(r.subRoots.size()), nr};
            r.subSets.insert(r.subSets.end(), desc, desc + 4);
            for (int q = 0; q < nr; ++q) r.subRoots.push_back(roots[sel[q]]);
        }
        mc_on_set<Real>(r, S, false);
    }
}

// ---------------------------------------------------------------------------
// Driver.
// ---------------------------------------------------------------------------
template <class Real>
static void run_numeric(Result& r, const Tables& t, int stages) {
    double t0 = now_ms();
    if (stages & 2) {
        vector_field<Real>(r);
        divergence<Real>(r, t);
    }
    double t1 = now_ms();
    if (stages & 4) {
        solve_all<Real>(r, t);
        points_chi<Real>(r);
    }
    double t2 = now_ms();
    if (stages & 8) {
        geometry(r);
        marching_cubes<Real>(r);
    } else if (stages & 16) {  // ownership and vertex values only, no marching cubes
        geometry(r);
        vertex_values<Real>(r);
    }
    double t3 = now_ms();
    r.tField = t1 - t0;
    r.tSolve = t2 - t1;
    r.tMesh = t3 - t2;
}

}  // namespace orc

using namespace orc;

struct Handle {
    Result r;
    Tables t;
    int tablesDepth = -1;
};

extern "C" {

// stages: 1 octree, 2 field+divergence, 4 solve+point chi, 8 geometry+MC+subdivision,
// 16 (without 8) geometry and vertex values only.
// precision: 0 = FP32 restatement, 1 = FP64 oracle.
void* orc_run(const float* xyz, const float* nrm, int n, int depth, int stages, int precision, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    Handle* h = new Handle();
    double t0 = now_ms();
    build_octree(h->r, xyz, nrm, n, depth);
    h->r.tOctree = now_ms() - t0;
    if (stages & 30) {
        build_tables(h->t, depth);
        h->tablesDepth = depth;
        if (precision)
            run_numeric<double>(h->r, h->t, stages);
        else
            run_numeric<float>(h->r, h->t, stages);
    }
    return h;
}

void orc_free(void* p) { delete static_cast<Handle*>(p); }

static bool field(Handle* h, const std::string& nm, const void*& ptr, long long& count, int& elem) {
    Result& r = h->r;
#define F(name, vec, esz)             \
    if (nm == name) {                 \
        ptr = (vec).data();           \
        count = (long long)(vec).size(); \
        elem = esz;                   \
        return true;                  \
    }
    F("spos", r.spos, 4) F("snrm", r.snrm, 4) F("order", r.order, 4) F("codes", r.codes, 4)
    F("point2node", r.point2node, 4) F("encidx", r.encIdx, 4) F("depth", r.depthOf, 4) F("center", r.center, 4)
    F("vfield", r.vfield, 8) F("div", r.div, 8) F("dx", r.dx, 8) F("chi_pts", r.chiPts, 8)
    F("cg_iters", r.cgIters, 4) F("cg_relres", r.cgRelRes, 8) F("cg_recres", r.cgRecRes, 8) F("nnz", r.nnz, 4)
    F("vvalue", r.vvalue, 8) F("cube", r.cube, 4) F("mesh_v", r.meshV, 4) F("mesh_t", r.meshT, 4)
    F("batch_v", r.batchV, 4) F("batch_t", r.batchT, 4) F("roots", r.rootNodes, 4)
    F("sub_sets", r.subSets, 4) F("sub_roots", r.subRoots, 4) F("sub_v", r.subV, 4) F("sub_e", r.subE, 4)
    F("dot", h->t.dot, 8) F("ddot", h->t.ddot, 8) F("d2dot", h->t.d2dot, 8)
#undef F
    return false;
}

long long orc_count(void* p, const char* name) {
    const void* ptr;
    long long c;
    int e;
    Handle* h = static_cast<Handle*>(p);
    std::string nm(name);
    Result& r = h->r;
    if (nm == "nodes") return (long long)r.nodes.size();
    if (nm == "verts") return (long long)r.verts.size();
    if (nm == "edges") return (long long)r.edges.size();
    if (nm == "faces") return (long long)r.faces.size();
    return field(h, nm, ptr, c, e) ? c : -1;
}

// Copies a named array.  Node arrays are exported column-wise:
// "node_<field>" with field in key,pidx,pnum,parent,didx,dnum,children(8),
// neighs(27),vertices(8),edges(12),faces(6).
int orc_copy(void* p, const char* name, void* dst, long long bytes) {
    Handle* h = static_cast<Handle*>(p);
    Result& r = h->r;
    std::string nm(name);
    if (nm.rfind("node_", 0) == 0) {
        std::string f = nm.substr(5);
        int* out = static_cast<int*>(dst);
        long long M = (long long)r.nodes.size();
        int w = (f == "children" || f == "vertices") ? 8 : f == "neighs" ? 27 : f == "edges" ? 12 : f == "faces" ? 6 : 1;
        if (bytes < M * w * 4) return -2;
        for (long long i = 0; i < M; ++i) {
            const OctNode& n = r.nodes[i];
            const int* src;
            if (f == "key") src = &n.key;
            else if (f == "pidx") src = &n.pidx;
            else if (f == "pnum") src = &n.pnum;
            else if (f == "parent") src = &n.parent;
            else if (f == "didx") src = &n.didx;
            else if (f == "dnum") src = &n.dnum;
            else if (f == "children") src = &r.childrenBuilt[8 * size_t(i)];
            else if (f == "neighs") src = n.neighs;
            else if (f == "vertices") src = n.vertices;
            else if (f == "edges") src = n.edges;
            else if (f == "faces") src = n.faces;
            else return -1;
            memcpy(out + i * w, src, size_t(w) * 4);
        }
        return 0;
    }
    if (nm == "verts") {  // pos xyz as float bits, owner, kind, depth: 6 ints per vertex
        if (bytes < (long long)r.verts.size() * 24) return -2;
        memcpy(dst, r.verts.data(), r.verts.size() * 24);
        return 0;
    }
    if (nm == "edges") {
        if (bytes < (long long)r.edges.size() * 8) return -2;
        memcpy(dst, r.edges.data(), r.edges.size() * 8);
        return 0;
    }
    if (nm == "faces") {
        if (bytes < (long long)r.faces.size() * 12) return -2;
        memcpy(dst, r.faces.data(), r.faces.size() * 12);
        return 0;
    }
    const void* ptr;
    long long c;
    int e;
    if (!field(h, nm, ptr, c, e)) return -1;
    if (bytes < c * e) return -2;
    memcpy(dst, ptr, size_t(c) * e);
    return 0;
}

double orc_scalar(void* p, const char* name) {
    Handle* h = static_cast<Handle*>(p);
    Result& r = h->r;
    std::string nm(name);
    if (nm == "M") return r.M;
    if (nm == "iso") return r.iso;
    if (nm == "edge") return r.edge;
    if (nm.size() == 7 && nm.rfind("corner", 0) == 0) return r.corner[nm[6] - '0'];
    if (nm == "q1") return r.q1_count;
    if (nm == "q1b") return r.q1b_count;
    if (nm == "q1_levels") return r.q1_levels;
    if (nm == "q16") return r.q16_flag;
    if (nm == "n_roots") return r.nRoots;
    if (nm == "n_coarse_roots") return r.nCoarseRoots;
    if (nm == "t_octree") return r.tOctree;
    if (nm == "t_field") return r.tField;
    if (nm == "t_solve") return r.tSolve;
    if (nm == "t_mesh") return r.tMesh;
    if (nm.rfind("base", 0) == 0) return r.base[atoi(nm.c_str() + 4)];
    if (nm.rfind("count", 0) == 0) return r.count[atoi(nm.c_str() + 5)];
    return NAN;
}

// Exposed for the basis tests: <F_a^(da), F_b^(db)> and F_idx(x).
double orc_inner(int a, int b, int da, int db) { return inner(a, b, da, db); }
double orc_basis(int idx, double x) { return basis_at<double>(idx, x); }
int orc_lut(int which, int s, int i) { return which ? kLut.child[s][i] : kLut.parent[s][i]; }
long long orc_mc_word(int cube) { return (long long)kOracleMcTable[cube]; }
// Whether depth `depth` of a depth-D octree gets MINRES (its dropped stencil is indefinite).
int orc_depth_indefinite(int D, int depth) {
    Tables t;
    build_tables(t, D);
    return stencil_indefinite(t, depth) ? 1 : 0;
}
// MINRES on a CSR matrix (the tests' check against a dense solve); returns the iterations.
int orc_minres(int n, const int* rp, const int* col, const double* val, const double* b, double* x, double rel_tol,
               int max_iter) {
    Csr A;
    A.n = n;
    A.rp.assign(rp, rp + n + 1);
    A.col.assign(col, col + rp[n]);
    A.val.assign(val, val + rp[n]);
    std::vector<double> bb(b, b + n), xx;
    double rec = 0;
    const int it = minres(A, bb, xx, rel_tol, max_iter, rec);
    std::copy(xx.begin(), xx.end(), x);
    return it;
}
// Entry i of a small quirk-defining table: 0 childrenVertexKind[8], 1 edgeVertex[12][2]
// (flattened), 2 parentFaceKind[8][6] (flattened), 3 faceEdges[6][4] (flattened).
int orc_small_table(int which, int i) {
    switch (which) {
        case 0: return (i >= 0 && i < 8) ? kChildrenVertexKind[i] : -99;
        case 1: return (i >= 0 && i < 24) ? kEdgeVertex[i / 2][i % 2] : -99;
        case 2: return (i >= 0 && i < 48) ? kParentFaceKind[i / 6][i % 6] : -99;
        case 3: return (i >= 0 && i < 24) ? kFaceEdges[i / 4][i % 4] : -99;
        default: return -99;
    }
}
}
