// Octree construction: normalisation, Morton codes, sort, node arrays,
// neighbour tables, packed function index.
//
// Same node arrays as the reference's level-by-level build
// (mesh/BuildOctree.cu:135-592, mesh/BuildOctree.cpp:75-194), computed without
// a per-level host loop: every sorted point knows the first level at which
// its code differs from its predecessor, one multi-level counting scan gives
// the rank of each occupied node at every level, and each node's final record
// is then written directly.
#include "psr_internal.h"
#include "psr_prims.h"

#include <cmath>
#include <cstring>
#include <stdexcept>

namespace psr {

namespace {

constexpr int kThreads = 256;

// --- bounding box (BuildOctree.cu:51-133) and host adjustment (:619-631) ---
// Reduces the nb block partials and returns the normalisation (x0, y0, z0,
// edge) to every thread of the block; block 0 also stores it to params and mail.
__device__ __forceinline__ float4 bbox_reduce(const float* __restrict__ part, int nb, float* __restrict__ params,
                                              float* __restrict__ mail, int bid) {
#pragma clang fp contract(off)
    __shared__ float s[6][kThreads];
    __shared__ float4 sOut;
    float mx[3] = {-1e6f, -1e6f, -1e6f}, mn[3] = {1e6f, 1e6f, 1e6f};
    for (int b = threadIdx.x; b < nb; b += kThreads)
        for (int a = 0; a < 3; ++a) {
            float v = part[b * 6 + a];
            mx[a] = mx[a] < v ? v : mx[a];
            v = part[b * 6 + 3 + a];
            mn[a] = mn[a] > v ? v : mn[a];
        }
    for (int a = 0; a < 3; ++a) {
        s[a][threadIdx.x] = mx[a];
        s[3 + a][threadIdx.x] = mn[a];
    }
    __syncthreads();
    for (int st = kThreads / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            for (int a = 0; a < 3; ++a) {
                float l = s[a][threadIdx.x], r = s[a][threadIdx.x + st];
                s[a][threadIdx.x] = l < r ? r : l;
                l = s[3 + a][threadIdx.x];
                r = s[3 + a][threadIdx.x + st];
                s[3 + a][threadIdx.x] = l > r ? r : l;
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        float maxEdge = 1.0f, center[3], lo[3];
        for (int a = 0; a < 3; ++a) {
            float ext = __fsub_rn(s[a][0], s[3 + a][0]);
            if (a == 0 || maxEdge < ext) maxEdge = ext;
            center[a] = __fdiv_rn(__fadd_rn(s[a][0], s[3 + a][0]), 2.0f);
        }
        maxEdge = __fmul_rn(maxEdge, 1.25f);
        for (int a = 0; a < 3; ++a) lo[a] = __fsub_rn(center[a], __fdiv_rn(maxEdge, 2.0f));
        sOut = make_float4(lo[0], lo[1], lo[2], maxEdge);
        if (bid == 0) {
            for (int a = 0; a < 3; ++a) params[a] = mail[a] = lo[a];
            params[3] = mail[3] = maxEdge;
        }
    }
    __syncthreads();
    return sOut;
}

// Block partial bounds (reduced by every block of the encoding, which needs
// no launch of its own).
__device__ __forceinline__ void bbox_partial_body(int psr_bid, int psr_nbid, const char* __restrict__ xyz, unsigned stride, int n,
                                                         float* __restrict__ part) {
    float mn[3] = {1e6f, 1e6f, 1e6f}, mx[3] = {-1e6f, -1e6f, -1e6f};
    for (int i = psr_bid * kThreads + threadIdx.x; i < n; i += psr_nbid * kThreads) {
        const float* p = reinterpret_cast<const float*>(xyz + size_t(i) * stride);
        for (int a = 0; a < 3; ++a) {
            float v = p[a];
            mx[a] = mx[a] < v ? v : mx[a];
            mn[a] = mn[a] > v ? v : mn[a];
        }
    }
    __shared__ float s[6][kThreads];
    for (int a = 0; a < 3; ++a) {
        s[a][threadIdx.x] = mx[a];
        s[3 + a][threadIdx.x] = mn[a];
    }
    __syncthreads();
    for (int st = kThreads / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st)
            for (int a = 0; a < 3; ++a) {
                float l = s[a][threadIdx.x], r = s[a][threadIdx.x + st];
                s[a][threadIdx.x] = l < r ? r : l;
                l = s[3 + a][threadIdx.x];
                r = s[3 + a][threadIdx.x + st];
                s[3 + a][threadIdx.x] = l > r ? r : l;
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[psr_bid * 6 + threadIdx.x] = s[threadIdx.x][0];
}
__global__ void __launch_bounds__(kThreads) bbox_partial(const char* __restrict__ xyz, unsigned stride, int n,
                                                         float* __restrict__ part) {
    bbox_partial_body(blockIdx.x, gridDim.x, xyz, stride, n, part);
}
PSR_BATCHABLE(bbox_partial, kThreads)


// --- normalise + Morton code (BuildOctree.cu:144-208) ---
__device__ __forceinline__ unsigned encode_point(const char* __restrict__ xyz, unsigned xs,
                                                 const char* __restrict__ nrm, unsigned ns, int i, int D,
                                                 float4 prm, float4* __restrict__ posOut,
                                                 float4* __restrict__ nrmOut) {
#pragma clang fp contract(off)
    const float* p = reinterpret_cast<const float*>(xyz + size_t(i) * xs);
    const float* q = reinterpret_cast<const float*>(nrm + size_t(i) * ns);
    float edge = prm.w;
    float x = __fdiv_rn(__fsub_rn(p[0], prm.x), edge);
    float y = __fdiv_rn(__fsub_rn(p[1], prm.y), edge);
    float z = __fdiv_rn(__fsub_rn(p[2], prm.z), edge);
    float nx = q[0], ny = q[1], nz = q[2];
    float len = __fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(nx, nx), __fmul_rn(ny, ny)), __fmul_rn(nz, nz)));
    if (double(len) > double(1e-6f)) len = __fdiv_rn(1.0f, len);
    len = __fmul_rn(len, float(2 << D));
    posOut[i] = make_float4(x, y, z, 0.f);
    nrmOut[i] = make_float4(__fmul_rn(nx, len), __fmul_rn(ny, len), __fmul_rn(nz, len), 0.f);
    unsigned code = 0;
    float cx = 0.5f, cy = 0.5f, cz = 0.5f, w = 0.25f;
    for (int l = D - 1; l >= 0; --l) {
        if (x > cx) { code |= 4u << (3 * l); cx += w; } else { cx -= w; }
        if (y > cy) { code |= 2u << (3 * l); cy += w; } else { cy -= w; }
        if (z > cz) { code |= 1u << (3 * l); cz += w; } else { cz -= w; }
        w *= 0.5f;
    }
    return code;
}

// The normalisation comes from the nb bounding-box partials (block 0 stores
// it to params and mail).  A block encodes one tile of the sort (two points
// a thread) and takes the tile's histogram of the sort's first digit
// (radix_sort_pairs), tiles of them at hist.
__device__ __forceinline__ void normalize_encode_body(int psr_bid, int psr_nbid, const char* __restrict__ xyz, unsigned xs,
                                                             const char* __restrict__ nrm, unsigned ns, int n,
                                                             int D, const float* __restrict__ part, int nb,
                                                             float* __restrict__ params, float* __restrict__ mail,
                                                             float4* __restrict__ posOut,
                                                             float4* __restrict__ nrmOut,
                                                             unsigned* __restrict__ keys,
                                                             unsigned* __restrict__ vals, int tiles,
                                                             int* __restrict__ hist, unsigned digitMask) {
    __shared__ int lh[256];
    static_assert(kThreads == 256 && kSortTileKeys == 2 * kThreads, "one thread per digit, two keys each");
    lh[threadIdx.x] = 0;
    const float4 prm = bbox_reduce(part, nb, params, mail, psr_bid);  // its barriers publish lh
    for (int k = 0; k < 2; ++k) {
        const int i = psr_bid * kSortTileKeys + k * kThreads + threadIdx.x;
        if (i >= n) continue;
        const unsigned code = encode_point(xyz, xs, nrm, ns, i, D, prm, posOut, nrmOut);
        keys[i] = code;
        vals[i] = unsigned(i);
        atomicAdd(&lh[code & digitMask], 1);
    }
    __syncthreads();
    if (threadIdx.x <= digitMask) hist[threadIdx.x * tiles + psr_bid] = lh[threadIdx.x];
}
__global__ void __launch_bounds__(kThreads) normalize_encode(const char* __restrict__ xyz, unsigned xs,
                                                             const char* __restrict__ nrm, unsigned ns, int n,
                                                             int D, const float* __restrict__ part, int nb,
                                                             float* __restrict__ params, float* __restrict__ mail,
                                                             float4* __restrict__ posOut,
                                                             float4* __restrict__ nrmOut,
                                                             unsigned* __restrict__ keys,
                                                             unsigned* __restrict__ vals, int tiles,
                                                             int* __restrict__ hist, unsigned digitMask) {
    normalize_encode_body(blockIdx.x, gridDim.x, xyz, xs, nrm, ns, n, D, part, nb, params, mail, posOut, nrmOut, keys, vals, tiles, hist, digitMask);
}
PSR_BATCHABLE(normalize_encode, kThreads)


// --- level ranks -----------------------------------------------------------
// lvl(i): first level 1..D at which the code of sorted point i differs from
// point i-1; D+1 when both lie in the same leaf; 1 for i = 0.
__device__ __forceinline__ int first_level(const unsigned* codes, int i, int D) {
    if (i == 0) return 1;
    unsigned x = codes[i] ^ codes[i - 1];
    if (x == 0) return D + 1;
    int hb = 31 - __clz(int(x));  // highest differing bit
    return D - hb / 3;           // bit 3(D-l)+k belongs to level l
}

constexpr int kRankItems = 8;
constexpr int kRankTile = kThreads * kRankItems;

__device__ __forceinline__ void rank_tile_hist_body(int psr_bid, int psr_nbid, const unsigned* __restrict__ codes, int n, int D,
                                                           int tiles, int* __restrict__ hist) {
    __shared__ int h[PSR_MAX_DEPTH + 2];
    if (threadIdx.x < PSR_MAX_DEPTH + 2) h[threadIdx.x] = 0;
    __syncthreads();
    int base = psr_bid * kRankTile;
    for (int k = 0; k < kRankItems; ++k) {
        int i = base + k * kThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[first_level(codes, i, D)], 1);
    }
    __syncthreads();
    if (threadIdx.x <= D) hist[threadIdx.x * tiles + psr_bid] = h[threadIdx.x];
}
__global__ void __launch_bounds__(kThreads) rank_tile_hist(const unsigned* __restrict__ codes, int n, int D,
                                                           int tiles, int* __restrict__ hist) {
    rank_tile_hist_body(blockIdx.x, gridDim.x, codes, n, D, tiles, hist);
}
PSR_BATCHABLE(rank_tile_hist, kThreads)


// One workgroup per level d: exclusive offsets of the tiles' new-node counts.
__device__ __forceinline__ void rank_tile_scan_body(int psr_bid, int psr_nbid, const int* __restrict__ hist, int tiles, int D,
                                                       int* __restrict__ tileOff, int* __restrict__ U,
                                                       int* __restrict__ mail) {
    const int d = psr_bid + 1;
    const int lane = __lane_id(), w = threadIdx.x / PSR_WAVE;
    __shared__ int wsum[1024 / PSR_WAVE];
    int run = 0;  // the same in every thread
    for (int t0 = 0; t0 < tiles; t0 += 1024) {
        const int t = t0 + threadIdx.x;
        int c = 0;
        if (t < tiles) {
            // every level's count read at once (the histogram holds PSR_MAX_DEPTH + 2
            // rows, those past D unwritten), the levels past d dropped after
            int v[PSR_MAX_DEPTH];
#pragma unroll
            for (int l = 1; l <= PSR_MAX_DEPTH; ++l) v[l - 1] = hist[l * tiles + t];
#pragma unroll
            for (int l = 1; l <= PSR_MAX_DEPTH; ++l) c += l <= d ? v[l - 1] : 0;
        }
        // block exclusive scan: wave scans, then the waves' totals (one barrier)
        int inc = c;
#pragma unroll
        for (int off = 1; off < PSR_WAVE; off <<= 1) {
            const int y = __shfl_up(inc, off, PSR_WAVE);
            if (lane >= off) inc += y;
        }
        if (lane == PSR_WAVE - 1) wsum[w] = inc;
        __syncthreads();
        int pre = 0, tot = 0;
#pragma unroll
        for (int k = 0; k < 1024 / PSR_WAVE; ++k) {
            const int sk = wsum[k];
            pre += k < w ? sk : 0;
            tot += sk;
        }
        if (t < tiles) tileOff[d * tiles + t] = run + pre + inc - c;
        run += tot;
        __syncthreads();  // wsum is written again by the next chunk
    }
    if (threadIdx.x == 0) U[d] = mail[d] = run;
}
__global__ void __launch_bounds__(1024) rank_tile_scan(const int* __restrict__ hist, int tiles, int D,
                                                       int* __restrict__ tileOff, int* __restrict__ U,
                                                       int* __restrict__ mail) {
    rank_tile_scan_body(blockIdx.x, gridDim.x, hist, tiles, D, tileOff, U, mail);
}
PSR_BATCHABLE(rank_tile_scan, 1024)


struct LevelOffsets {
    int lo[PSR_MAX_DEPTH + 2];  // start of level d in the per-level node lists
};

// Per point: rank of its node at every level.  New nodes record their first
// point and their parent's rank; every point records its level-(D-1) rank and
// its level-local depth-D slot.
__device__ __forceinline__ void rank_points_body(int psr_bid, int psr_nbid, const unsigned* __restrict__ codes, int n, int D, int tiles,
                                                        const int* __restrict__ tileOff, const LevelOffsets& lo,
                                                        int* __restrict__ firstPoint, int* __restrict__ parentRank,
                                                        int* __restrict__ rankDm1, int* __restrict__ point2node,
                                                        int* __restrict__ pnum, int M, int* __restrict__ q1Fired) {
    // clears the point counts (only occupied nodes get one) and the Q1 flag
    for (int i = psr_bid * kThreads + threadIdx.x; i < M; i += psr_nbid * kThreads) pnum[i] = 0;
    if (psr_bid == 0 && threadIdx.x == 0) *q1Fired = 0;
    int base = psr_bid * kRankTile + threadIdx.x * kRankItems;
    int lv[kRankItems];
    int cnt[PSR_MAX_DEPTH + 1];
    for (int d = 1; d <= D; ++d) cnt[d] = 0;
    for (int k = 0; k < kRankItems; ++k) {
        int i = base + k;
        lv[k] = (i < n) ? first_level(codes, i, D) : PSR_MAX_DEPTH + 2;
        for (int d = 1; d <= D; ++d) cnt[d] += (lv[k] <= d);
    }
    // block exclusive scan per level
    __shared__ int wsum[kThreads / PSR_WAVE];
    int excl[PSR_MAX_DEPTH + 1];
    int lane = __lane_id(), wid = threadIdx.x / PSR_WAVE;
    for (int d = 1; d <= D; ++d) {
        int v = cnt[d], inc = v;
        for (int off = 1; off < PSR_WAVE; off <<= 1) {
            int t = __shfl_up(inc, off, PSR_WAVE);
            if (lane >= off) inc += t;
        }
        if (lane == PSR_WAVE - 1) wsum[wid] = inc;
        __syncthreads();
        if (threadIdx.x == 0) {
            int r = 0;
            for (int w = 0; w < kThreads / PSR_WAVE; ++w) {
                int sw = wsum[w];
                wsum[w] = r;
                r += sw;
            }
        }
        __syncthreads();
        excl[d] = tileOff[d * tiles + psr_bid] + wsum[wid] + inc - v;
        __syncthreads();
    }
    for (int k = 0; k < kRankItems; ++k) {
        int i = base + k;
        if (i >= n) break;
        unsigned code = codes[i];
        int rank[PSR_MAX_DEPTH + 1];
        rank[0] = 0;
        for (int d = 1; d <= D; ++d) {
            if (lv[k] <= d) {
                int q = excl[d];
                firstPoint[lo.lo[d] + q] = i;
                parentRank[lo.lo[d] + q] = rank[d - 1];
                excl[d] = q + 1;
            }
            rank[d] = excl[d] - 1;
        }
        rankDm1[i] = rank[D - 1];
        point2node[i] = 8 * rank[D - 1] + int(code & 7u);
    }
}
__global__ void __launch_bounds__(kThreads) rank_points(const unsigned* __restrict__ codes, int n, int D, int tiles,
                                                        const int* __restrict__ tileOff, LevelOffsets lo,
                                                        int* __restrict__ firstPoint, int* __restrict__ parentRank,
                                                        int* __restrict__ rankDm1, int* __restrict__ point2node,
                                                        int* __restrict__ pnum, int M, int* __restrict__ q1Fired) {
    rank_points_body(blockIdx.x, gridDim.x, codes, n, D, tiles, tileOff, lo, firstPoint, parentRank, rankDm1, point2node, pnum, M, q1Fired);
}
PSR_BATCHABLE(rank_points, kThreads)


// Writes the record of every occupied node (levels 1..D).
__device__ __forceinline__ void fill_occupied_body(int psr_bid, int psr_nbid, const unsigned* __restrict__ codes, int n, const Levels& L,
                                                          const LevelOffsets& lo, const int* __restrict__ U,
                                                          const int* __restrict__ firstPoint,
                                                          const int* __restrict__ parentRank,
                                                          const int* __restrict__ rankDm1, int total,
                                                          int* __restrict__ key, int* __restrict__ pidx,
                                                          int* __restrict__ pnum, int* __restrict__ parent,
                                                          int* __restrict__ childBase, int* __restrict__ didx,
                                                          int* __restrict__ dnum, int* __restrict__ q1Fired) {
    int t = psr_bid * kThreads + threadIdx.x;
    if (t >= total) return;
    const int D = L.D;
    int d = 1;
    while (d < D && t >= lo.lo[d + 1]) ++d;
    int q = t - lo.lo[d];
    int first = firstPoint[t];
    int next = (q + 1 < U[d]) ? firstPoint[t + 1] : n;
    unsigned code = codes[first];
    int sh = 3 * (D - d);
    int son = int((code >> sh) & 7u);
    int pr = parentRank[t];
    int local = 8 * pr + son;
    int slot = L.base[d] + local;
    key[slot] = int((code >> sh) << sh);
    pidx[slot] = first;
    pnum[slot] = next - first;
    int par;
    if (d == 1) {
        par = 0;
    } else {
        int sonUp = int((code >> (sh + 3)) & 7u);
        int ppr = (d >= 3) ? parentRank[lo.lo[d - 1] + pr] : 0;
        par = L.base[d - 1] + 8 * ppr + sonUp;
        if (local == 0) {
            // Q1: level-local child 0 reads as "no child" (BuildOctree.cu:384-389, 412)
            if (par != L.base[d - 1]) atomicOr(q1Fired, 1);
            par = L.base[d - 1];
        }
    }
    parent[slot] = par;
    childBase[slot] = (d < D) ? L.base[d + 1] + 8 * q : -1;
    if (d == D) {
        didx[slot] = local;
        dnum[slot] = 1;
    } else {
        didx[slot] = 8 * rankDm1[first];
        dnum[slot] = 8 * (rankDm1[next - 1] - rankDm1[first] + 1);
    }
}
__global__ void __launch_bounds__(kThreads) fill_occupied(const unsigned* __restrict__ codes, int n, Levels L,
                                                          LevelOffsets lo, const int* __restrict__ U,
                                                          const int* __restrict__ firstPoint,
                                                          const int* __restrict__ parentRank,
                                                          const int* __restrict__ rankDm1, int total,
                                                          int* __restrict__ key, int* __restrict__ pidx,
                                                          int* __restrict__ pnum, int* __restrict__ parent,
                                                          int* __restrict__ childBase, int* __restrict__ didx,
                                                          int* __restrict__ dnum, int* __restrict__ q1Fired) {
    fill_occupied_body(blockIdx.x, gridDim.x, codes, n, L, lo, U, firstPoint, parentRank, rankDm1, total, key, pidx, pnum, parent, childBase, didx, dnum, q1Fired);
}
PSR_BATCHABLE(fill_occupied, kThreads)


// Completes each block of 8 siblings: empty slots and the shared parent
// (updateEmptyNodeInfo, BuildOctree.cu:426-488).  Block b of level d holds the
// children of the b-th occupied node of level d-1.
__device__ __forceinline__ void fill_blocks_body(int psr_bid, int psr_nbid, const Levels& L, int totalBlocks, int N, int nDm1,
                                                        int* __restrict__ key, int* __restrict__ pidx,
                                                        int* __restrict__ pnum, int* __restrict__ parent,
                                                        int* __restrict__ childBase, int* __restrict__ didx,
                                                        int* __restrict__ dnum) {
    int t = psr_bid * kThreads + threadIdx.x;
    if (t > totalBlocks) return;
    const int D = L.D;
    if (t == totalBlocks) {  // root
        key[0] = 0;
        pidx[0] = 0;
        pnum[0] = N;
        parent[0] = -1;
        childBase[0] = 1;
        didx[0] = 0;
        dnum[0] = 8 * nDm1;
        return;
    }
    int idx = 1 + 8 * t;
    int d = 1;
    while (d < D && idx >= L.base[d + 1]) ++d;
    int v = 0;
    while (v < 7 && pnum[idx + v] == 0) ++v;
    int sh = 3 * (D - d);
    int vk = key[idx + v];
    int baseKey = vk - (vk & (7 << sh));
    int nowP = pidx[idx + v], nowD = didx[idx + v], common = parent[idx + v];
    for (int j = 0; j < 8; ++j) {
        int o = idx + j;
        int pn = pnum[o];
        if (pn == 0) {
            childBase[o] = -1;
            if (d == D) {
                didx[o] = o - L.base[D];
                dnum[o] = 1;
            } else {
                dnum[o] = 0;
            }
        }
        key[o] = baseKey + (j << sh);
        pidx[o] = nowP;
        nowP += pn;
        if (d != D) {
            didx[o] = nowD;
            nowD += dnum[o];
        }
        parent[o] = common;
    }
}
__global__ void __launch_bounds__(kThreads) fill_blocks(Levels L, int totalBlocks, int N, int nDm1,
                                                        int* __restrict__ key, int* __restrict__ pidx,
                                                        int* __restrict__ pnum, int* __restrict__ parent,
                                                        int* __restrict__ childBase, int* __restrict__ didx,
                                                        int* __restrict__ dnum) {
    fill_blocks_body(blockIdx.x, gridDim.x, L, totalBlocks, N, nDm1, key, pidx, pnum, parent, childBase, didx, dnum);
}
PSR_BATCHABLE(fill_blocks, kThreads)


__device__ __forceinline__ void zero_int_body(int psr_bid, int psr_nbid, int* __restrict__ a, int n) {
    int i = psr_bid * kThreads + threadIdx.x;
    if (i < n) a[i] = 0;
}
__global__ void __launch_bounds__(kThreads) zero_int(int* __restrict__ a, int n) {
    zero_int_body(blockIdx.x, gridDim.x, a, n);
}
PSR_BATCHABLE(zero_int, kThreads)


// Same-depth neighbours from the parent's neighbours (BuildOctree.cu:490-512).
// Neighbours of the depth-`depth` nodes from their parents' (the LUT recurrence
// of BuildOctree.cu:20-42, 413-440): neighbour i of a child with son key s is
// child LUTchild[s][i] of the parent's neighbour LUTparent[s][i].  Both are
// derived per axis: the child-level offset c = s_a + d_a lies in parent cell
// floor(c / 2) at child bit c - 2 floor(c / 2); only the 8 parent neighbours
// with per-axis offsets in {s_a - 1, s_a} are touched, each read once.
// Neighbour i of a node with son key s, from the child bases of the 8 parent
// neighbours u with per-axis offsets {s_a - 1, s_a} (cb[u], -1 if none).
__device__ __forceinline__ int child_neighbour(const int (&cb)[8], int s, int i) {
    const int sx = (s >> 2) & 1, sy = (s >> 1) & 1, sz = s & 1;
    const int cx = sx + i / 9 - 1, cy = sy + (i / 3) % 3 - 1, cz = sz + i % 3 - 1;
    const int fx = cx >> 1, fy = cy >> 1, fz = cz >> 1;  // floor(c / 2) for c in [-1, 2]
    const int u = ((fx - sx + 1) << 2) | ((fy - sy + 1) << 1) | (fz - sz + 1);
    const int b = cb[u];
    return b < 0 ? -1 : b + (((cx - 2 * fx) << 2) | ((cy - 2 * fy) << 1) | (cz - 2 * fz));
}
// Index of parent neighbour u of a node with son key s.
__device__ __forceinline__ int parent_slot(int s, int u) {
    const int px = ((s >> 2) & 1) - 1 + ((u >> 2) & 1), py = ((s >> 1) & 1) - 1 + ((u >> 1) & 1),
              pz = (s & 1) - 1 + (u & 1);
    return 9 * (px + 1) + 3 * (py + 1) + (pz + 1);
}

// Packed 1-D function index (BuildOctree.cu:514-550), depth and basis
// parameters (centres, inverse width) of node o at depth d.
__device__ __forceinline__ void encode_node(int o, int d, int D, const int* __restrict__ key, int* __restrict__ enc,
                                            int* __restrict__ depth, float4* __restrict__ basis) {
    depth[o] = d;
    const int3 f = key_func_index(key[o], d, D);
    enc[o] = f.x | (f.y << (D + 1)) | (f.z << (2 * (D + 1)));
    basis[o] = node_basis_params(f);
}

// Every node in one launch: a node at depth d walks its ancestors from the
// root down, keeping at each level only the 8 neighbours of the ancestor that
// its next descendant's 27 entries touch, and ends with its own 27.  The
// recurrence is the one above, applied d times from the root's row (itself at
// 13, nothing else) instead of reading the parent's finished row, so no level
// waits for the one before it; each level costs one round of 8 independent
// childBase reads.  A row depends on the chain of parent links up to the root
// and on the son key of each node on it.  Unless the level-local child 0 rule
// changed a parent (q1Fired), that chain is the node's true ancestry and its
// son keys are the bits of the node's own key; otherwise they are read along
// the parent links, as the reference's level-by-level build uses them.  Every
// thread also encodes its node, and thread 0 writes the root's row.
__device__ __forceinline__ void neighbors_all_body(int psr_bid, int psr_nbid, const Levels& L, int M,
                                                   const int* __restrict__ key, const int* __restrict__ parent,
                                                   const int* __restrict__ childBase,
                                                   const int* __restrict__ q1Fired, int* __restrict__ neigh,
                                                   int* __restrict__ enc, int* __restrict__ depthOf,
                                                   float4* __restrict__ basis) {
    const int o = psr_bid * kThreads + threadIdx.x;
    if (o >= M) return;
    const int D = L.D;
    if (o == 0) {
        encode_node(0, 0, D, key, enc, depthOf, basis);
        for (int i = 0; i < 27; ++i) neigh[i * M] = i == 13 ? 0 : -1;
        return;
    }
    const int d = depth_of(L, o);
    encode_node(o, d, D, key, enc, depthOf, basis);
    const int k = key[o];
    static_assert(3 * PSR_MAX_DEPTH <= 32, "son keys of a chain fit one word");
    unsigned sons = 0;  // son key of the chain's depth-l node in bits 3(l - 1) .. 3l - 1
    if (*q1Fired == 0) {
        for (int l = 1; l <= d; ++l) sons |= unsigned((k >> (3 * (D - l))) & 7) << (3 * (l - 1));
    } else {
        sons = unsigned((k >> (3 * (D - d))) & 7) << (3 * (d - 1));
        for (int l = d - 1, p = parent[o]; l >= 1; --l, p = parent[p])
            sons |= unsigned((key[p] >> (3 * (D - l))) & 7) << (3 * (l - 1));
    }
    auto son = [&](int l) { return int(sons >> (3 * (l - 1))) & 7; };
    int s = son(1);
    int np[8];  // neighbours of the depth l - 1 ancestor at slots parent_slot(s, u)
#pragma unroll
    for (int u = 0; u < 8; ++u) np[u] = parent_slot(s, u) == 13 ? 0 : -1;
    int cb[8];
    for (int l = 1; l < d; ++l) {
#pragma unroll
        for (int u = 0; u < 8; ++u) cb[u] = np[u] == -1 ? -1 : childBase[np[u]];
        const int sn = son(l + 1);
#pragma unroll
        for (int u = 0; u < 8; ++u) np[u] = child_neighbour(cb, s, parent_slot(sn, u));
        s = sn;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) cb[u] = np[u] == -1 ? -1 : childBase[np[u]];
#pragma unroll
    for (int i = 0; i < 27; ++i) neigh[i * M + o] = child_neighbour(cb, s, i);
}
__global__ void __launch_bounds__(kThreads) neighbors_all(Levels L, int M, const int* __restrict__ key,
                                                          const int* __restrict__ parent,
                                                          const int* __restrict__ childBase,
                                                          const int* __restrict__ q1Fired,
                                                          int* __restrict__ neigh, int* __restrict__ enc,
                                                          int* __restrict__ depthOf, float4* __restrict__ basis) {
    neighbors_all_body(blockIdx.x, gridDim.x, L, M, key, parent, childBase, q1Fired, neigh, enc, depthOf, basis);
}
PSR_BATCHABLE(neighbors_all, kThreads)


}  // namespace

void build_octree_a(Ctx& c, const psr_points& pts) {
    Octree& o = c.oct;
    const int N = int(pts.count), D = c.cfg.depth;
    hipStream_t s = c.stream;
    o.N = N;
    const unsigned xs = pts.stride_bytes ? pts.stride_bytes : 12, ns = xs;
    // bbox + normalisation parameters (block partials: every encoding block reduces them)
    const int nbb = std::min<int>(128, int(div_up(N, kThreads)));
    float* part = c.buf<float>("bbox_part", size_t(nbb) * 6);
    o.normParams = c.buf<float>("norm_params", 4);
    PSR_LAUNCH(bbox_partial, nbb, 0, s, reinterpret_cast<const char*>(pts.xyz), xs, N, part);
    // normalise + encode
    float4* posU = c.buf<float4>("pos_unsorted", N);
    float4* nrmU = c.buf<float4>("nrm_unsorted", N);
    unsigned* keys = c.buf<unsigned>("sort_keys", N);
    unsigned* vals = c.buf<unsigned>("sort_vals", N);
    unsigned* ktmp = c.buf<unsigned>("sort_ktmp", N);
    unsigned* vtmp = c.buf<unsigned>("sort_vtmp", N);
    // the encoding also takes the sort's first histogram
    int* sortHist = c.scratch_for(sort_scratch_ints(N));
    const int sortTiles = int(div_up(N, kSortTileKeys));
    PSR_LAUNCH(normalize_encode, sortTiles, 0, s,
        reinterpret_cast<const char*>(pts.xyz), xs, reinterpret_cast<const char*>(pts.normal), ns, N, D, part, nbb,
        o.normParams, reinterpret_cast<float*>(c.mail + kMailNorm), posU, nrmU, keys, vals, sortTiles, sortHist,
        (1u << sort_digit_bits(3 * D)) - 1u);
    // sorted by code; the last pass gathers the points into code order
    o.spos = c.buf<float4>("spos", N);
    o.snrm = c.buf<float4>("snrm", N);
    radix_sort_pairs(keys, vals, ktmp, vtmp, N, 3 * D, sortHist, true, c.scan_for(sort_scan_elems(N)),
                     SortGather{posU, nrmU, o.spos, o.snrm}, s);
    o.codes = reinterpret_cast<int*>(keys);
    o.order = reinterpret_cast<int*>(vals);
    // per-level occupied-node counts
    int tiles = int(div_up(N, kRankTile));
    int* hist = c.buf<int>("rank_hist", size_t(PSR_MAX_DEPTH + 2) * tiles);
    int* tileOff = c.buf<int>("rank_tileoff", size_t(PSR_MAX_DEPTH + 2) * tiles);
    int* Ud = c.buf<int>("level_U", PSR_MAX_DEPTH + 2);
    PSR_LAUNCH(rank_tile_hist, tiles, 0, s, keys, N, D, tiles, hist);
    PSR_LAUNCH(rank_tile_scan, D, 0, s, hist, tiles, D, tileOff, Ud, c.mail);
}

// After the wait for the level counts: the node arrays and neighbour tables.
void build_octree_b(Ctx& c) {
    Octree& o = c.oct;
    const int N = o.N, D = c.cfg.depth;
    hipStream_t s = c.stream;
    const unsigned* keys = reinterpret_cast<const unsigned*>(o.codes);
    const int tiles = int(div_up(N, kRankTile));
    int* tileOff = c.buf<int>("rank_tileoff", size_t(PSR_MAX_DEPTH + 2) * tiles);
    int* Ud = c.buf<int>("level_U", PSR_MAX_DEPTH + 2);
    float edge;
    std::memcpy(&edge, c.hostPinned + kMailNorm + 3, sizeof(float));
    if (!(edge > 0.f) || !std::isfinite(edge))  // the reference divides by it (BuildOctree.cu:149)
        throw std::invalid_argument("degenerate input: zero or non-finite bounding box");
    o.U[0] = 1;
    for (int d = 1; d <= D; ++d) o.U[d] = c.hostPinned[d];
    // level layout
    Levels& L = o.L;
    L.D = D;
    L.count[0] = 1;
    for (int d = 1; d <= D; ++d) L.count[d] = 8 * o.U[d - 1];
    L.base[0] = 0;
    for (int d = 1; d <= D + 1; ++d) L.base[d] = L.base[d - 1] + L.count[d - 1];
    o.M = L.base[D + 1];
    const int M = o.M;
    LevelOffsets lo{};
    lo.lo[1] = 0;
    for (int d = 1; d <= D; ++d) lo.lo[d + 1] = lo.lo[d] + o.U[d];
    int totalNodes = lo.lo[D + 1];
    int* firstPoint = c.buf<int>("first_point", totalNodes);
    int* parentRank = c.buf<int>("parent_rank", totalNodes);
    int* rankDm1 = c.buf<int>("rank_dm1", N);
    o.point2node = c.buf<int>("point2node", N);
    // node arrays
    o.key = c.buf<int>("node_key", M);
    o.pidx = c.buf<int>("node_pidx", M);
    o.pnum = c.buf<int>("node_pnum", M);
    o.parent = c.buf<int>("node_parent", M);
    o.childBase = c.buf<int>("node_childbase", M);
    o.didx = c.buf<int>("node_didx", M);
    o.dnum = c.buf<int>("node_dnum", M);
    o.neigh = c.buf<int>("node_neigh", size_t(27) * M);
    o.encIdx = c.buf<int>("node_enc", M);
    o.q1Fired = &c.results()->q1Fired;
    PSR_LAUNCH(rank_points, tiles, 0, s, keys, N, D, tiles, tileOff, lo, firstPoint, parentRank, rankDm1,
                                           o.point2node, o.pnum, M, o.q1Fired);
    PSR_LAUNCH(fill_occupied, div_up(totalNodes, kThreads), 0, s, keys, N, L, lo, Ud, firstPoint, parentRank,
                                                                     rankDm1, totalNodes, o.key, o.pidx, o.pnum,
                                                                     o.parent, o.childBase, o.didx, o.dnum,
                                                                     o.q1Fired);
    int totalBlocks = (M - 1) / 8;
    PSR_LAUNCH(fill_blocks, div_up(totalBlocks + 1, kThreads), 0, s, L, totalBlocks, N, o.U[D - 1], o.key, o.pidx,
                                                                        o.pnum, o.parent, o.childBase, o.didx,
                                                                        o.dnum);
    // neighbours, packed index, depth and basis of every node
    o.depth = c.buf<int>("node_depth", M);
    o.basis = c.buf<float4>("node_basis", M);
    PSR_LAUNCH(neighbors_all, div_up(M, kThreads), 0, s, L, M, o.key, o.parent, o.childBase, o.q1Fired, o.neigh,
               o.encIdx, o.depth, o.basis);
    PSR_HIP_CHECK(hipGetLastError());
}

void build_octree(Ctx& c, const psr_points& pts) {
    build_octree_a(c, pts);
    c.wait();
    build_octree_b(c);
}

}  // namespace psr
