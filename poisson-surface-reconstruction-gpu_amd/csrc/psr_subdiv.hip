// Adaptive subdivision and final mesh assembly.
//
// The reference rebuilds every coarse subdivision root alone
// (CoarserSubdivideNodeAndRebuildMesh, ComputeTriangleIndices.cu:1475-1799)
// and the finer roots once per root depth (FinerSubdivideNodeAndRebuildMesh,
// :1801-2194), each as a full virtual subtree down to depth D whose ids start
// at NodeArraySize, with neighbour tables (computeRebuildNeighbor :669-696),
// vertex and edge ownership (:698-954), vertex values (:956-1038) and marching
// cubes on the virtual depth-D cells.  Here every such rebuild is a "set", and
// the virtual subtrees are never materialised: they are full octrees, so a
// virtual cell is an integer coordinate, its key its Morton code, and all
// structure follows by arithmetic:
//   * a depth-D cell n is virtual in set S iff its depth-rd ancestor cell is
//     the cell of a root of S (the same root, or a neighbour of it in the
//     real table that is a root of S: child look-ups through a real parent are
//     redirected only to roots of the same set, :1137 and :1525/:1784);
//   * the owner of a corner (edge) is the virtual cell of smallest key among
//     the 8 (4) cells sharing it; real cells never own (:735);
//   * the value at a corner walks the owner's chain: at virtual depths only
//     real neighbours contribute (virtual d_x = 0, :978, :1021); they lie in
//     the root's real neighbours and are found by child descent; at the root
//     depth the root copy's neighbours are the root's real neighbours with
//     root copies mapped to neighs[13] of their root; above it the real chain.
// The order of every floating-point sum and of every emitted vertex and
// triangle is that of the materialised formulation.
#include "psr_mesh_common.h"
#include "psr_prims.h"
#include "psr_tile_scan.h"

#include <algorithm>
#include <cstring>

namespace psr {

namespace {

constexpr int kThreads = 256;

struct SetDesc {
    int rd;          // root depth
    int nr;          // roots in the set
    int rootOffset;  // first root in the root list
    int dEnum;       // first depth-D cell enumeration index of the set (chunk-local)
    int firstTile;   // tile holding the set's first cell (tiles are in enumeration order)
};

struct Grid {
    int D, M, nSets;
    const SetDesc* sets;
    const int* roots;      // real node ids, ascending
    const int* rootPos;    // real node -> position in roots, or -1 (nodes above depth D)
    const int* key;
    const int* parent;
    const int* childBase;
    const int* neigh;
    const float4* basis;
    const float* dx;
    const float4* nbrCol;    // [27][M] centre (x, y, z) and d_x of neighbour i of each node (0 where absent),
                             // written only for octrees in which the level-local child 0 rule changed a parent
    const int* rootNbr;      // [root][27] position in roots of the root's real neighbour, or -1
    const int* rootChain;    // [root][kChain] the root, then its ancestors (-1 past the tree root)
    const int4* rootCell;    // [root] cell at the root depth (x, y, z) and the root node (w)
    const unsigned* rootSet; // [root] bit i: neighbour i of the root belongs to the root's set
    const float* dxAt;       // per depth l, at (8^l - 1) / 7: d_x of the node at each cell, x-major; 0 where none
};

// Neighbour i of node as a column (centre, d_x; zeros where absent), read
// through the neighbour table: the regular octree's tiles (the dense-halo
// kernels) have no float4 columns.
__device__ __forceinline__ float4 nbr_col(const Grid& g, int i, int node) {
    const int nb = g.neigh[size_t(i) * g.M + node];
    if (nb < 0) return make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 b = g.basis[nb];
    return make_float4(b.x, b.y, b.z, g.dx[nb]);
}

constexpr int kChain = PSR_MAX_DEPTH + 1;

// Bourke corner order: slot -> (x, y, z) offset (BuildMeshGeometry.cu:133-144).
__device__ __forceinline__ int3 bourke_offset(int b) {
    int x = (b == 1 || b == 2 || b == 5 || b == 6) ? 1 : 0;
    int y = (b == 2 || b == 3 || b == 6 || b == 7) ? 1 : 0;
    return make_int3(x, y, b >> 2);
}
__device__ __forceinline__ int bourke_slot(int ox, int oy, int oz) {
    int i = ox | (oz << 2);
    if (oy) i += ox ? 1 : 3;
    return i;
}

// A virtual depth-D cell and its root.
struct Leaf {
    SetDesc S;
    int si, q, root, sh;  // sh = D - rd
    int3 Rc;              // root cell at depth rd
    int3 c;               // cell at depth D
};

// Work is issued in tiles of at most 256 consecutive cells of one set; a set
// whose roots hold at least 512 cells is tiled along its roots, so such a tile
// lies in one root.
struct Tile {
    int set, start, count;
    int root;  // position in roots of the tile's first root
};
constexpr int kTile = 256;

__device__ __forceinline__ Leaf make_leaf_in(const Grid& g, int si, const SetDesc& S, int e) {
    Leaf L;
    L.si = si;
    L.S = S;
    L.sh = g.D - L.S.rd;
    int rel = e - L.S.dEnum;
    L.q = rel >> (3 * L.sh);
    int j = rel - (L.q << (3 * L.sh));
    L.root = g.roots[L.S.rootOffset + L.q];
    L.Rc = demorton(g.key[L.root] >> (3 * L.sh), L.S.rd);
    int3 lc = demorton(j, L.sh);
    L.c = make_int3((L.Rc.x << L.sh) + lc.x, (L.Rc.y << L.sh) + lc.y, (L.Rc.z << L.sh) + lc.z);
    return L;
}

__device__ __forceinline__ int nbr_index(int dx, int dy, int dz) { return 9 * (dx + 1) + 3 * (dy + 1) + (dz + 1); }

__device__ __forceinline__ bool same_set_root(const Grid& g, const SetDesc& S, int node, int& q) {
    int rp = g.rootPos[node];
    q = rp - S.rootOffset;
    return rp >= S.rootOffset && rp < S.rootOffset + S.nr;
}

// Enumeration index of depth-D cell n if it is virtual in L's set, else -1.
__device__ __forceinline__ int virtual_enum(const Grid& g, const Leaf& L, int3 n) {
    const int lim = 1 << g.D;
    if (n.x < 0 || n.y < 0 || n.z < 0 || n.x >= lim || n.y >= lim || n.z >= lim) return -1;
    int3 A = make_int3(n.x >> L.sh, n.y >> L.sh, n.z >> L.sh);
    int q = L.q;
    if (A.x != L.Rc.x || A.y != L.Rc.y || A.z != L.Rc.z) {
        int N = g.neigh[nbr_index(A.x - L.Rc.x, A.y - L.Rc.y, A.z - L.Rc.z) * g.M + L.root];
        if (N < 0 || !same_set_root(g, L.S, N, q)) return -1;
    }
    int3 lc = make_int3(n.x - (A.x << L.sh), n.y - (A.y << L.sh), n.z - (A.z << L.sh));
    return L.S.dEnum + (q << (3 * L.sh)) + morton(lc, L.sh);
}

__device__ __forceinline__ bool inside_root(const Leaf& L, int lo, int hi, int axisRoot) {
    return (lo >> L.sh) == axisRoot && (hi >> L.sh) == axisRoot;
}

// Key order of the cells sharing a grid point p: the cells p - d (d in {0,1}
// per axis) differ in axis a only in bits 0..ctz(p_a), so two of them compare
// on their differing axis of highest ctz (x before y before z at the same
// level), where the one with d_a = 1 has the smaller key.  The cells in
// increasing key order are therefore d in decreasing lexicographic order over
// the axes ranked by axis_rank; p_a = 0 (no cell below) ranks last.
__device__ __forceinline__ int axis_rank(int pa, int prio) { return pa > 0 ? ((__builtin_ctz(pa) << 2) | prio) : -1; }

// Axes of p by decreasing rank, as bit positions (0 = x, 1 = y, 2 = z).
__device__ __forceinline__ void rank_axes(int3 p, int& a0, int& a1, int& a2) {
    int r0 = axis_rank(p.x, 2), r1 = axis_rank(p.y, 1), r2 = axis_rank(p.z, 0);
    a0 = 0;
    a1 = 1;
    a2 = 2;
    if (r1 > r0) {
        int t = r0; r0 = r1; r1 = t;
        t = a0; a0 = a1; a1 = t;
    }
    if (r2 > r1) {
        int t = r1; r1 = r2; r2 = t;
        t = a1; a1 = a2; a2 = t;
    }
    if (r1 > r0) {
        int t = a0; a0 = a1; a1 = t;
    }
}

// Whether depth-D cell n is virtual (a look-up may answer this more cheaply
// than the enumeration).
template <class VE>
__device__ __forceinline__ bool ve_valid(const VE& ve, int3 n) {
    return ve(n) >= 0;
}

// The virtual cell of smallest key among those sharing grid point p (and its
// offset d as a bit mask, bit a = d_a), or -1.
template <class VE>
__device__ __forceinline__ int point_owner(int3 p, int& dmask, const VE& ve) {
    int a0, a1, a2;
    rank_axes(p, a0, a1, a2);
    for (int j = 7; j >= 0; --j) {
        const int m = (((j >> 2) & 1) << a0) | (((j >> 1) & 1) << a1) | ((j & 1) << a2);
        const int3 n = make_int3(p.x - (m & 1), p.y - ((m >> 1) & 1), p.z - ((m >> 2) & 1));
        if (ve_valid(ve, n)) {
            dmask = m;
            return ve(n);
        }
    }
    dmask = 0;
    return -1;
}

// Owner (enumeration, Bourke slot within the owner) of the grid point c + off(b);
// ve(n) is the enumeration of depth-D cell n if it is virtual in L's set, else -1.
template <class VE>
__device__ __forceinline__ int owner_of_point(const Leaf& L, int3 p, int& slot, VE ve);
template <class VE>
__device__ __forceinline__ int corner_owner(const Leaf& L, int D, int b, int& slot, VE ve) {
    const int3 o = bourke_offset(b);
    return owner_of_point(L, make_int3(L.c.x + o.x, L.c.y + o.y, L.c.z + o.z), slot, ve);
}

// Owner of grid point p, a corner of a cell of L's root (only the root, its
// cell and L's set are read from L).
template <class VE>
__device__ __forceinline__ int owner_of_point(const Leaf& L, int3 p, int& slot, VE ve) {
    if (inside_root(L, p.x - 1, p.x, L.Rc.x) && inside_root(L, p.y - 1, p.y, L.Rc.y) &&
        inside_root(L, p.z - 1, p.z, L.Rc.z)) {
        // all 8 sharing cells are in the root's cube; the smallest key is the lowest cell
        slot = bourke_slot(1, 1, 1);
        int3 lc = make_int3(p.x - 1 - (L.Rc.x << L.sh), p.y - 1 - (L.Rc.y << L.sh), p.z - 1 - (L.Rc.z << L.sh));
        return L.S.dEnum + (L.q << (3 * L.sh)) + morton(lc, L.sh);
    }
    if (L.S.nr == 1) {
        // a set of one root: the virtual cells sharing p are those in the root's
        // cube, and the lowest of them takes d_a = 1 unless p is on the cube's
        // low face in axis a
        const int3 lp = make_int3(p.x - (L.Rc.x << L.sh), p.y - (L.Rc.y << L.sh), p.z - (L.Rc.z << L.sh));
        const int3 d = make_int3(lp.x > 0 ? 1 : 0, lp.y > 0 ? 1 : 0, lp.z > 0 ? 1 : 0);
        slot = bourke_slot(d.x, d.y, d.z);
        return L.S.dEnum + (L.q << (3 * L.sh)) + morton(make_int3(lp.x - d.x, lp.y - d.y, lp.z - d.z), L.sh);
    }
    int m;
    const int best = point_owner(p, m, ve);
    slot = best < 0 ? 0 : bourke_slot(m & 1, (m >> 1) & 1, (m >> 2) & 1);
    return best;
}

// Whether cell p - off(b) owns grid point p (no virtual cell sharing p comes
// before it in key order).
template <class VE>
__device__ __forceinline__ bool owns_corner(const Leaf& L, int b, const VE& ve) {
    const int3 o = bourke_offset(b);
    const int3 lc = make_int3(L.c.x - (L.Rc.x << L.sh), L.c.y - (L.Rc.y << L.sh), L.c.z - (L.Rc.z << L.sh));
    if (L.S.nr == 1) return (o.x || lc.x == 0) && (o.y || lc.y == 0) && (o.z || lc.z == 0);
    const int3 p = make_int3(L.c.x + o.x, L.c.y + o.y, L.c.z + o.z);
    const int om = o.x | (o.y << 1) | (o.z << 2);
    int a0, a1, a2;
    rank_axes(p, a0, a1, a2);
    for (int j = 7; j >= 0; --j) {
        const int m = (((j >> 2) & 1) << a0) | (((j >> 1) & 1) << a1) | ((j & 1) << a2);
        if (m == om) return true;
        if (ve_valid(ve, make_int3(p.x - (m & 1), p.y - ((m >> 1) & 1), p.z - ((m >> 2) & 1)))) return false;
    }
    return false;
}

// Owner (enumeration, edge kind within the owner) of edge slot k of L.
// Edge kind = 4 * axis + offsets of the other two axes (bit set: positive side).
template <class VE>
__device__ __forceinline__ int edge_owner(const Leaf& L, int D, int k, int& kind, VE ve) {
    const int a = k >> 2;
    const int u0 = a == 0 ? 1 : 0, u1 = a == 2 ? 1 : 2;
    const int o0 = k & 1, o1 = (k >> 1) & 1;
    int cc[3] = {L.c.x, L.c.y, L.c.z}, rc[3] = {L.Rc.x, L.Rc.y, L.Rc.z};
    int lo0 = cc[u0] + o0 - 1, lo1 = cc[u1] + o1 - 1;
    if (inside_root(L, lo0, lo0 + 1, rc[u0]) && inside_root(L, lo1, lo1 + 1, rc[u1])) {
        kind = (a << 2) | 3;
        int n[3];
        n[a] = cc[a];
        n[u0] = lo0;
        n[u1] = lo1;
        int3 lc = make_int3(n[0] - (L.Rc.x << L.sh), n[1] - (L.Rc.y << L.sh), n[2] - (L.Rc.z << L.sh));
        return L.S.dEnum + (L.q << (3 * L.sh)) + morton(lc, L.sh);
    }
    if (L.S.nr == 1) {
        // a set of one root: the lowest cell of the root's cube sharing the edge
        const int b0 = lo0 + 1 > (rc[u0] << L.sh) ? 1 : 0, b1 = lo1 + 1 > (rc[u1] << L.sh) ? 1 : 0;
        kind = (a << 2) | b0 | (b1 << 1);
        int n[3];
        n[a] = cc[a];
        n[u0] = lo0 + 1 - b0;
        n[u1] = lo1 + 1 - b1;
        const int3 lc = make_int3(n[0] - (L.Rc.x << L.sh), n[1] - (L.Rc.y << L.sh), n[2] - (L.Rc.z << L.sh));
        return L.S.dEnum + (L.q << (3 * L.sh)) + morton(lc, L.sh);
    }
    // the 4 cells in increasing key order (axis_rank of the two perpendicular
    // axes at the point lo + 1); b = 1: the cell below the edge in that axis
    const bool firstU1 = axis_rank(lo1 + 1, 2 - u1) > axis_rank(lo0 + 1, 2 - u0);
    kind = 0;
    for (int j = 3; j >= 0; --j) {
        const int hi = (j >> 1) & 1, lo = j & 1;
        const int b0 = firstU1 ? lo : hi, b1 = firstU1 ? hi : lo;
        int n[3];
        n[a] = cc[a];
        n[u0] = lo0 + 1 - b0;
        n[u1] = lo1 + 1 - b1;
        const int3 nn = make_int3(n[0], n[1], n[2]);
        if (ve_valid(ve, nn)) {
            kind = (a << 2) | b0 | (b1 << 1);
            return ve(nn);
        }
    }
    return -1;
}

__device__ __forceinline__ float grid_term(const Grid& g, int nb, float3 p) { return node_term(g.basis[nb], g.dx[nb], p); }

// Contributor k of a root's fixed list, given the chain node k / 27 (the
// root, then its ancestors): k / 27 = 0 the root copy's neighbours (root
// copies standing for neighs[13] of their root), above it the ancestors'
// neighbours; -1 where there is none (computeSubdivideVertexImplicitFunctionValue
// :956-1038, the part of the chain from the root upwards, which all cells of a
// root share).
__device__ __forceinline__ int root_list_node(const Grid& g, const SetDesc& S, int chainNode, int k) {
    if (chainNode < 0) return -1;
    int nb = g.neigh[(k % 27) * g.M + chainNode];
    int qq;
    if (k < 27 && nb >= 0 && same_set_root(g, S, nb, qq)) nb = g.neigh[13 * g.M + nb];
    return nb < 0 ? -1 : nb;
}

// Real node for the depth-l cell n next to root (cell Rc at depth rd): -1 when
// the cell is virtual (inside a root of the set) or absent; otherwise the real
// child reached from the root's real neighbour (virtual d_x = 0, :978, :1021).
__device__ __forceinline__ int halo_node(const Grid& g, const SetDesc& S, int root, int3 Rc, int l, int3 n) {
    const int lim = 1 << l, s = l - S.rd;
    if (n.x < 0 || n.y < 0 || n.z < 0 || n.x >= lim || n.y >= lim || n.z >= lim) return -1;
    int3 A = make_int3(n.x >> s, n.y >> s, n.z >> s);
    if (A.x == Rc.x && A.y == Rc.y && A.z == Rc.z) return -1;
    int node = g.neigh[nbr_index(A.x - Rc.x, A.y - Rc.y, A.z - Rc.z) * g.M + root];
    int qq;
    if (node < 0 || same_set_root(g, S, node, qq)) return -1;
    for (int k = 1; k <= s && node >= 0; ++k) {  // real children down to depth l
        int cb = g.childBase[node];
        int sh = s - k;
        node = cb < 0 ? -1 : cb + ((((n.x >> sh) & 1) << 2) | (((n.y >> sh) & 1) << 1) | ((n.z >> sh) & 1));
    }
    return node;
}

// Tiles.  A tile holds whole roots (small roots: 8 or 64 cells) or 256 cells
// of one root.  What its cells share is resolved once into LDS by all threads:
//   * vbox: per root, the virtual enumeration (or -1) of every depth-D cell of
//     the box around the tile's cells of that root (all owner candidates);
//   * per root, the fixed contributor list (27 (rd + 1) nodes);
//   * halo: per root and virtual depth l, the real node (or -1) of every cell
//     of the box around the tile's cells at depth l (the chain's neighbours);
//   * for a tile in one root, d_x of the list and the 1-D B-spline factors of
//     every list node at the tile's corner coordinates (256 Morton-consecutive
//     cells span 4 x 8 x 8, their corners 5 x 9 x 9).
// The terms and their order are those of the chain walk.
constexpr int kTx = 5, kTy = 9, kTz = 9, kTw = kTx + kTy + kTz;
constexpr int kTilePoints = kTx * kTy * kTz;  // corner points of a one-root tile
constexpr int kMaxRootsPerTile = 32;
constexpr int kClassifyPoints = 864;  // corner points of a tile: 32 roots of 27, 4 of 125 or one box of 5 x 9 x 9
constexpr int kLdsWordsOne = 7168;     // one-root tiles: list, halo, vbox, d_x, factors (<= 6954 words)
constexpr int kLdsWordsSmall = 5376;   // small-root tiles of 64 cells: halo, vbox, staged blocks (<= 5344 words)
constexpr int kStagedBlocks = 32;  // contributor blocks of 27 staged in LDS per small-root tile
constexpr int kSmallThreads = 128;  // threads of a small-root tile (4 x 27 or 125 corner points)
// threads of a small-root tile in sub_corner_values_small_grid: the points take
// at most 125; the others share out the tile's tables (the launch has only a
// few blocks per CU, so the tables' set-up is latency-bound: 128 threads 31.1,
// 256 25.4, 512 23.6 us on the torus)
constexpr int kSmallGridThreads = 512;
constexpr int kVboxWords = 2048;

__device__ __forceinline__ int3 box_ext(int3 extD, int D, int l) {
    int sh = D - l;
    return make_int3(max(1, extD.x >> sh), max(1, extD.y >> sh), max(1, extD.z >> sh));
}

struct TileGeom {
    int nRoots, q0, K;
    bool oneRoot;
    int3 ext;         // cells per root covered by the tile at depth D (per axis)
    int haloPerRoot;  // sum over virtual depths of the box sizes
    int vboxPerRoot;  // depth-D box size
};

__device__ __forceinline__ TileGeom tile_geom(const Tile& T, const SetDesc& S, int D) {
    TileGeom tg;
    const int sh = D - S.rd;
    tg.K = 27 * (S.rd + 1);
    tg.q0 = (T.start - S.dEnum) >> (3 * sh);
    tg.oneRoot = sh >= 3;
    tg.nRoots = tg.oneRoot ? 1 : T.count >> (3 * sh);
    tg.ext = tg.oneRoot ? make_int3(4, 8, 8) : make_int3(1 << sh, 1 << sh, 1 << sh);
    tg.haloPerRoot = 0;
    for (int l = D; l > S.rd; --l) {
        int3 x = box_ext(tg.ext, D, l);
        tg.haloPerRoot += (x.x + 2) * (x.y + 2) * (x.z + 2);
    }
    tg.vboxPerRoot = (tg.ext.x + 2) * (tg.ext.y + 2) * (tg.ext.z + 2);
    return tg;
}

// Lowest depth-D cell of root rr's cells in the tile.
__device__ __forceinline__ Leaf tile_root_first(const Grid& g, const Tile& T, const SetDesc& S, const TileGeom& tg,
                                                int rr) {
    return make_leaf_in(g, T.set, S, tg.oneRoot ? T.start : S.dEnum + ((tg.q0 + rr) << (3 * (g.D - S.rd))));
}

// Per-root facts of a tile, kept in LDS: root node, root cell, lowest tile cell.
struct TileRoot {
    int root, q;
    int3 Rc, first;
};

__device__ __forceinline__ void fill_roots(const Grid& g, const Tile& T, const SetDesc& S, const TileGeom& tg,
                                           TileRoot* tr) {
    for (int rr = threadIdx.x; rr < tg.nRoots; rr += blockDim.x) {
        Leaf F = tile_root_first(g, T, S, tg, rr);
        tr[rr] = TileRoot{F.root, F.q, F.Rc, F.c};
    }
}

// fill_roots from the root records.
__device__ __forceinline__ void fill_roots_rec(const Grid& g, const Tile& T, const SetDesc& S, const TileGeom& tg,
                                               TileRoot* tr) {
    const int sh = g.D - S.rd;
    for (int rr = threadIdx.x; rr < tg.nRoots; rr += blockDim.x) {
        const int q = tg.q0 + rr;
        const int4 rc = g.rootCell[T.root + rr];
        const int3 Rc = make_int3(rc.x, rc.y, rc.z);
        int3 first = make_int3(Rc.x << sh, Rc.y << sh, Rc.z << sh);
        if (tg.oneRoot) {
            const int3 lc = demorton(T.start - S.dEnum - (q << (3 * sh)), sh);
            first = make_int3(first.x + lc.x, first.y + lc.y, first.z + lc.z);
        }
        tr[rr] = TileRoot{rc.w, q, Rc, first};
    }
}

// Leaf of cell e, which lies in tile root R.
__device__ __forceinline__ Leaf leaf_of(const Grid& g, int si, const SetDesc& S, const TileRoot& R, int e) {
    Leaf L;
    L.si = si;
    L.S = S;
    L.sh = g.D - S.rd;
    L.q = R.q;
    L.root = R.root;
    L.Rc = R.Rc;
    const int3 lc = demorton(e - S.dEnum - (R.q << (3 * L.sh)), L.sh);
    L.c = make_int3((R.Rc.x << L.sh) + lc.x, (R.Rc.y << L.sh) + lc.y, (R.Rc.z << L.sh) + lc.z);
    return L;
}

// Depth-D cell of enumeration e, which lies in tile root R.
__device__ __forceinline__ int3 cell_of(const Grid& g, const SetDesc& S, const TileRoot& R, int e) {
    const int sh = g.D - S.rd;
    const int3 lc = demorton(e - S.dEnum - (R.q << (3 * sh)), sh);
    return make_int3((R.Rc.x << sh) + lc.x, (R.Rc.y << sh) + lc.y, (R.Rc.z << sh) + lc.z);
}

__device__ __forceinline__ Leaf root_leaf(const Grid& g, const SetDesc& S, int si, const TileRoot& R) {
    Leaf F;
    F.S = S;
    F.si = si;
    F.q = R.q;
    F.root = R.root;
    F.sh = g.D - S.rd;
    F.Rc = R.Rc;
    F.c = R.first;
    return F;
}

// Virtual enumerations of the depth-D box cells of every root of the tile
// (needs the roots in LDS).
__device__ __forceinline__ void fill_vbox(const Grid& g, const Tile& T, const SetDesc& S, const TileGeom& tg,
                                          const TileRoot* tr, int* vbox) {
    const int by = tg.ext.y + 2, bz = tg.ext.z + 2;
#pragma unroll 4
    for (int t = threadIdx.x; t < tg.nRoots * tg.vboxPerRoot; t += blockDim.x) {
        int rr = t / tg.vboxPerRoot, u = t - rr * tg.vboxPerRoot;
        const Leaf F = root_leaf(g, S, T.set, tr[rr]);
        int dz = u % bz, dy = (u / bz) % by, dx = u / (bz * by);
        vbox[t] = virtual_enum(g, F, make_int3(F.c.x - 1 + dx, F.c.y - 1 + dy, F.c.z - 1 + dz));
    }
}

// Virtual-cell lookup from the root's neighbour positions (nbr: its 27
// entries, in LDS): the same test as virtual_enum, without global loads.
struct RecVE {
    const int* nbr;
    SetDesc S;
    int3 Rc;
    int q, sh, lim;
    unsigned inSet;  // bit i: neighbour cube i of the root belongs to the set (bit 13: the root)
    __device__ __forceinline__ bool valid(int3 n) const {
        if (n.x < 0 || n.y < 0 || n.z < 0 || n.x >= lim || n.y >= lim || n.z >= lim) return false;
        return (inSet >> nbr_index((n.x >> sh) - Rc.x, (n.y >> sh) - Rc.y, (n.z >> sh) - Rc.z)) & 1u;
    }
    __device__ __forceinline__ int operator()(int3 n) const {
        if (n.x < 0 || n.y < 0 || n.z < 0 || n.x >= lim || n.y >= lim || n.z >= lim) return -1;
        const int3 A = make_int3(n.x >> sh, n.y >> sh, n.z >> sh);
        int qq = q;
        if (A.x != Rc.x || A.y != Rc.y || A.z != Rc.z) {
            const int rp = nbr[nbr_index(A.x - Rc.x, A.y - Rc.y, A.z - Rc.z)];
            if (!(rp >= S.rootOffset && rp < S.rootOffset + S.nr)) return -1;
            qq = rp - S.rootOffset;
        }
        const int3 lc = make_int3(n.x - (A.x << sh), n.y - (A.y << sh), n.z - (A.z << sh));
        return S.dEnum + (qq << (3 * sh)) + morton(lc, sh);
    }
};
__device__ __forceinline__ RecVE rec_ve(const int* nbr, unsigned inSet, const Leaf& L, int D) {
    return RecVE{nbr, L.S, L.Rc, L.q, L.sh, 1 << D, inSet};
}
__device__ __forceinline__ bool ve_valid(const RecVE& ve, int3 n) { return ve.valid(n); }

// Virtual-cell lookup in the tile's depth-D box of this cell's root.
struct BoxVE {
    const int* vbox;
    int3 lo;  // box origin (one below the lowest tile cell of the root)
    int by, bz;
    __device__ __forceinline__ int operator()(int3 n) const {
        return vbox[((n.x - lo.x) * by + (n.y - lo.y)) * bz + (n.z - lo.z)];
    }
};

__device__ __forceinline__ BoxVE box_ve(const int* vboxAll, const TileGeom& tg, int r, int3 rlo) {
    BoxVE v;
    v.vbox = vboxAll + r * tg.vboxPerRoot;
    v.lo = make_int3(rlo.x - 1, rlo.y - 1, rlo.z - 1);
    v.by = tg.ext.y + 2;
    v.bz = tg.ext.z + 2;
    return v;
}

// val += node terms of ids[0..n) in order, with the node loads of a batch in flight together.
template <int B>
__device__ __forceinline__ float accumulate_nodes(const Grid& g, const int* ids, int n, float3 p, float val) {
    for (int k0 = 0; k0 < n; k0 += B) {
        float4 bb[B];
        float dd[B];
        bool ok[B];
#pragma unroll
        for (int j = 0; j < B; ++j) {
            int nb = (k0 + j < n) ? ids[k0 + j] : -1;
            ok[j] = nb >= 0;
            bb[j] = ok[j] ? g.basis[nb] : make_float4(0.f, 0.f, 0.f, 0.f);
            dd[j] = ok[j] ? g.dx[nb] : 0.f;
        }
#pragma unroll
        for (int j = 0; j < B; ++j)
            if (ok[j]) val += node_term(bb[j], dd[j], p);
    }
    return val;
}

// Values of the corners each cell owns: cval[slot][e].  ONE: tiles lying in
// one root (roots of at least 512 cells); otherwise tiles of whole small roots.
template <bool ONE, int NT>
__device__ __forceinline__ void sub_corner_values_body(int psr_bid, int psr_nbid, const Grid& g, const Tile* __restrict__ tiles, int nVD,
                                                        const float* __restrict__ iso, float* __restrict__ cval) {
    constexpr int kWords = ONE ? kLdsWordsOne : kLdsWordsSmall;
    __shared__ float4 ldsStore[kWords / 4];  // 16-byte aligned; every region below starts at a multiple of 4 words
    int* lds = reinterpret_cast<int*>(ldsStore);
    __shared__ int lChain[(ONE ? 1 : kMaxRootsPerTile) * (PSR_MAX_DEPTH + 1)];  // [nRoots][rd + 1]: root, ancestors
    __shared__ TileRoot lRoots[ONE ? 1 : kMaxRootsPerTile];
    const Tile T = tiles[psr_bid];
    const SetDesc S = g.sets[T.set];
    const int D = g.D, sh = D - S.rd;
    TileGeom tg = tile_geom(T, S, D);
    tg.oneRoot = ONE;
    const int K = tg.K;
    // one root:   list [K] | halo | vbox | dx [K] | B-spline factors [K][kTw]
    // small roots: halo | vbox | staged block bases [kStagedBlocks * 27] (float4) | staged d_x
    int* lList = lds;
    int* lHalo = lList + (tg.oneRoot ? K : 0);
    int* lVbox = lHalo + tg.nRoots * tg.haloPerRoot;
    float* lDx = reinterpret_cast<float*>(lVbox + tg.nRoots * tg.vboxPerRoot);
    float* lTab = lDx + K;
    float4* sBasis = reinterpret_cast<float4*>(lDx);  // small roots: halo and vbox sizes are multiples of 4 words
    float* sDx = reinterpret_cast<float*>(sBasis + kStagedBlocks * 27);
    __shared__ int lBlk[ONE ? 1 : kMaxRootsPerTile * (PSR_MAX_DEPTH + 1)];  // [nRoots][rd + 1] -> staged block
    __shared__ int lBlkNode[ONE ? 1 : kMaxRootsPerTile * (PSR_MAX_DEPTH + 1)];
    __shared__ int lNumBlk;
    const bool active = int(threadIdx.x) < T.count;
    const int e = T.start + (active ? int(threadIdx.x) : 0);
    Leaf L = make_leaf_in(g, T.set, S, e);
    const int r = L.q - tg.q0;  // this cell's root within the tile
    const float inv = 1.0f / float(1 << D);

    // roots and their chains
    fill_roots(g, T, S, tg, lRoots);
    for (int rr = threadIdx.x; rr < tg.nRoots; rr += NT) {
        int node = g.roots[S.rootOffset + tg.q0 + rr];
        for (int u = 0; u <= S.rd; ++u) {
            lChain[rr * (S.rd + 1) + u] = node;
            node = node < 0 ? -1 : g.parent[node];
        }
    }
    __syncthreads();
    fill_vbox(g, T, S, tg, lRoots, lVbox);
    // halo boxes, depth D first
#pragma unroll 4
    for (int t = threadIdx.x; t < tg.nRoots * tg.haloPerRoot; t += NT) {
        int rr = t / tg.haloPerRoot, u = t - rr * tg.haloPerRoot;
        const Leaf F = root_leaf(g, S, T.set, lRoots[rr]);
        int l = D;
        int3 x = box_ext(tg.ext, D, l);
        while (u >= (x.x + 2) * (x.y + 2) * (x.z + 2)) {
            u -= (x.x + 2) * (x.y + 2) * (x.z + 2);
            --l;
            x = box_ext(tg.ext, D, l);
        }
        int dz = u % (x.z + 2), dy = (u / (x.z + 2)) % (x.y + 2), dx = u / ((x.z + 2) * (x.y + 2));
        int3 n = make_int3((F.c.x >> (D - l)) - 1 + dx, (F.c.y >> (D - l)) - 1 + dy, (F.c.z >> (D - l)) - 1 + dz);
        lHalo[t] = halo_node(g, S, F.root, F.Rc, l, n);
    }
    const int C = S.rd + 1;  // chain length: the root, then its ancestors
    if (tg.oneRoot) {
        for (int k = threadIdx.x; k < K; k += NT) {
            int nb = root_list_node(g, S, lChain[k / 27], k);
            lList[k] = nb;
            lDx[k] = nb < 0 ? 0.f : g.dx[nb];
        }
    } else {
        // contributor blocks: the root level of each root, and each distinct
        // ancestor (roots are in key order, so shared ancestors are adjacent)
        if (threadIdx.x == 0) {
            int nblk = 0;
            for (int rr = 0; rr < tg.nRoots; ++rr)
                for (int u = 0; u < C; ++u) {
                    int id;
                    if (u > 0 && rr > 0 && lChain[rr * C + u] == lChain[(rr - 1) * C + u]) {
                        id = lBlk[(rr - 1) * C + u];
                    } else {
                        id = nblk++;
                        lBlkNode[id] = u == 0 ? rr : -1 - (rr * C + u);  // root level: root; else chain slot
                    }
                    lBlk[rr * C + u] = id;
                }
            lNumBlk = nblk;
        }
        __syncthreads();
        const int nStage = min(lNumBlk, kStagedBlocks);
#pragma unroll 4
        for (int t = threadIdx.x; t < nStage * 27; t += NT) {
            int j = t / 27, i = t - j * 27;
            int tag = lBlkNode[j];
            if (tag >= 0) {  // root level
                int nb = root_list_node(g, S, lChain[tag * C], i);
                sBasis[t] = nb < 0 ? make_float4(0.f, 0.f, 0.f, 0.f) : g.basis[nb];
                sDx[t] = nb < 0 ? 0.f : g.dx[nb];
            } else {         // ancestor: its neighbour column, at the ancestor's depth
                const int slotChain = -1 - tag;
                const int node = lChain[slotChain];
                const float4 col = node < 0 ? make_float4(0.f, 0.f, 0.f, 0.f) : g.nbrCol[size_t(i) * g.M + node];
                sBasis[t] = make_float4(col.x, col.y, col.z, node < 0 ? 0.f : float(1 << (S.rd - slotChain % C)));
                sDx[t] = col.w;
            }
        }
    }
    int3 lo = make_int3(0, 0, 0);
    if (tg.oneRoot) {
        lo = lRoots[0].first;
        __syncthreads();
        // B-spline factors of the list nodes at the tile's corner coordinates
        for (int t = threadIdx.x; t < K * kTw; t += NT) {
            int k = t / kTw, c = t - k * kTw;
            int nb = lList[k];
            float v = 0.f;
            if (nb >= 0) {
                float4 b = g.basis[nb];
                if (c < kTx)
                    v = bspline((float(lo.x + c) * inv - b.x) * b.w);
                else if (c < kTx + kTy)
                    v = bspline((float(lo.y + c - kTx) * inv - b.y) * b.w);
                else
                    v = bspline((float(lo.z + c - kTx - kTy) * inv - b.z) * b.w);
            }
            lTab[t] = v;
        }
    }
    __syncthreads();
    const float isoV = *iso;
    // Value at grid point pt through the chain of owner cell Lo (in tile root rr,
    // whose lowest tile cell is rlo).
    auto value_at = [&](const Leaf& Lo, int rr, int3 rlo, int3 pt) -> float {
        float3 p = make_float3(float(pt.x) * inv, float(pt.y) * inv, float(pt.z) * inv);
        const int* halo = lHalo + rr * tg.haloPerRoot;
        float val = 0.f;
        // virtual depths D .. rd+1: real neighbours of the chain node
        int hb = 0;
        for (int l = D; l > S.rd; --l) {
            const int s = l - S.rd;
            const int3 x = box_ext(tg.ext, D, l);
            const int by = x.y + 2, bz = x.z + 2;
            const int3 cl = make_int3(Lo.c.x >> (D - l), Lo.c.y >> (D - l), Lo.c.z >> (D - l));
            const bool interior = ((cl.x - 1) >> s) == Lo.Rc.x && ((cl.x + 1) >> s) == Lo.Rc.x &&
                                  ((cl.y - 1) >> s) == Lo.Rc.y && ((cl.y + 1) >> s) == Lo.Rc.y &&
                                  ((cl.z - 1) >> s) == Lo.Rc.z && ((cl.z + 1) >> s) == Lo.Rc.z;
            if (!interior) {
                const int3 bl = make_int3((rlo.x >> (D - l)) - 1, (rlo.y >> (D - l)) - 1, (rlo.z >> (D - l)) - 1);
                int ids[27];
#pragma unroll
                for (int i = 0; i < 27; ++i) {
                    int3 n = make_int3(cl.x + i / 9 - 1 - bl.x, cl.y + (i / 3) % 3 - 1 - bl.y,
                                       cl.z + i % 3 - 1 - bl.z);
                    ids[i] = halo[hb + (n.x * by + n.y) * bz + n.z];
                }
                val = accumulate_nodes<9>(g, ids, 27, p, val);
            }
            hb += (x.x + 2) * by * bz;
        }
        // root depth and real ancestors; absent contributors are zero terms,
        // which leave the sum unchanged (it never holds -0)
        if (tg.oneRoot) {
            const int ix = pt.x - lo.x, iy = kTx + pt.y - lo.y, iz = kTx + kTy + pt.z - lo.z;
            for (int k = 0; k < K; ++k) {
                const float* t = lTab + k * kTw;
                val += lDx[k] * t[ix] * t[iy] * t[iz];
            }
        } else {
            for (int u = 0; u < C; ++u) {
                const int j = lBlk[rr * C + u];
                if (j < kStagedBlocks) {
                    const float4* bb = sBasis + j * 27;
                    const float* dd = sDx + j * 27;
                    for (int i = 0; i < 27; ++i) val += node_term(bb[i], dd[i], p);
                } else if (u > 0) {
                    const int node = lChain[rr * C + u];
                    if (node >= 0) {
                        const float invw = float(1 << (S.rd - u));
                        for (int i = 0; i < 27; ++i) val += col_term(g.nbrCol[size_t(i) * g.M + node], invw, p);
                    }
                } else {
                    int ids[27];
                    const int node = lChain[rr * C];
                    for (int i = 0; i < 27; ++i) ids[i] = root_list_node(g, S, node, i);
                    val = accumulate_nodes<9>(g, ids, 27, p, val);
                }
            }
        }
        return val;
    };
    if (ONE) {
        // one thread per cell: the corners it owns (one, away from the root's faces)
        if (!active) return;
        const BoxVE ve = box_ve(lVbox, tg, r, lo);
        for (int b = 0; b < 8; ++b) {
            int slot;
            if (corner_owner(L, D, b, slot, ve) != e) continue;
            int3 o = bourke_offset(b);
            cval[size_t(b) * nVD + e] = value_at(L, 0, lo, make_int3(L.c.x + o.x, L.c.y + o.y, L.c.z + o.z)) - isoV;
        }
    } else {
        // one thread per corner point of each small root; the point is evaluated
        // for the root holding its owner
        const int n = (1 << sh) + 1, P = n * n * n;
        const int rr = threadIdx.x / P, pi = threadIdx.x - rr * P;
        if (rr >= tg.nRoots) return;
        const TileRoot R = lRoots[rr];
        const int3 pt = make_int3(R.first.x + pi / (n * n), R.first.y + (pi / n) % n, R.first.z + pi % n);
        const BoxVE ve = box_ve(lVbox, tg, rr, R.first);
        int m;
        const int best = point_owner(pt, m, ve);
        const int slot = bourke_slot(m & 1, (m >> 1) & 1, (m >> 2) & 1);
        const int3 bestCell = make_int3(pt.x - (m & 1), pt.y - ((m >> 1) & 1), pt.z - ((m >> 2) & 1));
        if (best < 0 || ((best - S.dEnum) >> (3 * sh)) != tg.q0 + rr) return;
        Leaf Lo = root_leaf(g, S, T.set, R);
        Lo.c = bestCell;
        cval[size_t(slot) * nVD + best] = value_at(Lo, rr, R.first, pt) - isoV;
    }
}
template <bool ONE, int NT>
__global__ void __launch_bounds__(NT) sub_corner_values(Grid g, const Tile* __restrict__ tiles, int nVD,
                                                        const float* __restrict__ iso, float* __restrict__ cval) {
    sub_corner_values_body<ONE, NT>(blockIdx.x, gridDim.x, g, tiles, nVD, iso, cval);
}
template <bool ONE, int NT>
struct sub_corner_values_k {
    static constexpr int kBlock = NT;
    template <class... A>
    __device__ __forceinline__ static void run(int bid, int nbid, const A&... a) { sub_corner_values_body<ONE, NT>(bid, nbid, a...); }
};
// the two instantiations launched: one-root tiles and small roots
using sub_corner_values_one_k = sub_corner_values_k<true, kThreads>;
using sub_corner_values_small_k = sub_corner_values_k<false, kSmallThreads>;
constexpr auto sub_corner_values_one = sub_corner_values<true, kThreads>;
constexpr auto sub_corner_values_small = sub_corner_values<false, kSmallThreads>;



// Quotient by a small divisor through a multiply: exact for u * d < 2^32.
struct SmallDiv {
    unsigned m;
    int d;
    __device__ __forceinline__ explicit SmallDiv(int dd) : m(0xFFFFFFFFu / unsigned(dd) + 1u), d(dd) {}
    __device__ __forceinline__ int div(int u) const { return int(__umulhi(unsigned(u), m)); }
};

// One-root tiles when every parent is the true one (the level-local child 0
// rule changed none).  The octree is then geometric: the real node at depth l
// and cell n, if any, is the one the halo descent from the root's neighbour
// reaches, and its basis centre is (n + 0.5) / 2^l with inverse width 2^l.  So
// the halo is read as d_x from the dense per-depth grid, and the B-spline
// factors of every halo node and every list node at the tile's corner
// coordinates come from tables in LDS.  The list tables are stored per corner
// coordinate with the list index fastest, d_x folded into the x factor (the
// product d_x F(x) is the first one the term forms), so four terms are read
// with one 16-byte load per axis.  Terms, their order and their association
// are those of sub_corner_values.
constexpr int kLdsWordsGrid = 7168;  // upper bound of grid_tile_words for depths up to 9
constexpr int kLdsWordsSmallGrid = 12288;  // upper bound of small_tile_words for depths up to 10
constexpr int kHaloPerThread = 4;     // halo entries per thread (<= 1024 for depths up to 9)
// The single-cloud kernel at 7 waves per SIMD (72 registers, 3 spilled): 8%
// faster than at 5 (82 registers).  The batched form, which would spill 25
// there, keeps the compiler's choice (the same batch time).
constexpr int kCornerGridWaves = 7;
__device__ __forceinline__ void sub_corner_values_grid_body(int psr_bid, int psr_nbid, const Grid& g, const Tile* __restrict__ tiles, int nVD,
                                                                   const float* __restrict__ iso,
                                                                   float* __restrict__ cval,
                                                                   unsigned* __restrict__ ptOwner) {
    extern __shared__ float4 ldsDyn[];  // grid_tile_words(D, rd) words
    __shared__ unsigned short lItem[kThreads * 8];  // owned (cell << 3 | corner), cell order
    __shared__ int lNbr[27];
    __shared__ int lChain[kChain];
    __shared__ TileRoot lRoot[1];
    __shared__ unsigned lSet;
    __shared__ unsigned lNz;  // bit l: some halo d_x of depth l is nonzero
    float* lds = reinterpret_cast<float*>(ldsDyn);
    const Tile T = tiles[psr_bid];
    if (threadIdx.x < 27) lNbr[threadIdx.x] = g.rootNbr[27 * size_t(T.root) + threadIdx.x];
    if (threadIdx.x == 32) lSet = g.rootSet[T.root];
    if (threadIdx.x == 33) lNz = 0u;
    if (threadIdx.x < kChain) lChain[threadIdx.x] = g.rootChain[kChain * size_t(T.root) + threadIdx.x];
    const SetDesc S = g.sets[T.set];
    const int D = g.D;
    const TileGeom tg = tile_geom(T, S, D);
    const int K = tg.K, Kp = (K + 3) & ~3;
    // list x (with d_x) [kTx][Kp] | list y [kTy][Kp] | list z [kTz][Kp] | halo d_x | halo factors
    float* lLx = lds;
    float* lLy = lLx + kTx * Kp;
    float* lLz = lLy + kTy * Kp;
    float* lHdx = lLz + kTz * Kp;
    float* lFac = lHdx + tg.haloPerRoot;
    int facWords = 0;
    for (int l = D; l > S.rd; --l) {
        const int3 x = box_ext(tg.ext, D, l);
        facWords += (x.x + 2) * kTx + (x.y + 2) * kTy + (x.z + 2) * kTz;
    }
    fill_roots_rec(g, T, S, tg, lRoot);
    __syncthreads();
    const TileRoot R = lRoot[0];
    const int3 lo = R.first;
    const float inv = 1.0f / float(1 << D);
    // the root list: the 27 neighbours of the root and of each ancestor; thread k
    // loads list node k and writes its factors at the tile's corner coordinates
    if (int(threadIdx.x) < Kp) {
        const int k = threadIdx.x;
        const int node = k < K ? lChain[k / 27] : -1;
        float4 b = make_float4(0.f, 0.f, 0.f, 0.f);
        float d = 0.f;
        if (node >= 0) {
            const float4 col = nbr_col(g, k % 27, node);
            b = make_float4(col.x, col.y, col.z, float(1 << (S.rd - k / 27)));
            d = col.w;
        }
        for (int cc = 0; cc < kTx; ++cc)
            lLx[cc * Kp + k] = node < 0 ? 0.f : d * bspline((float(lo.x + cc) * inv - b.x) * b.w);
        for (int cc = 0; cc < kTy; ++cc)
            lLy[cc * Kp + k] = node < 0 ? 0.f : bspline((float(lo.y + cc) * inv - b.y) * b.w);
        for (int cc = 0; cc < kTz; ++cc)
            lLz[cc * Kp + k] = node < 0 ? 0.f : bspline((float(lo.z + cc) * inv - b.z) * b.w);
    }
    // halo d_x, depth D first: 0 inside the roots of the set and where no node is;
    // a thread's loads are issued together
    unsigned nz = 0;
#pragma unroll
    for (int j = 0; j < kHaloPerThread; ++j) {
        const int t = threadIdx.x + j * kThreads;
        if (t >= tg.haloPerRoot) break;
        int u = t, l = D;
        int3 x = box_ext(tg.ext, D, l);
        while (u >= (x.x + 2) * (x.y + 2) * (x.z + 2)) {
            u -= (x.x + 2) * (x.y + 2) * (x.z + 2);
            --l;
            x = box_ext(tg.ext, D, l);
        }
        const SmallDiv byz((x.y + 2) * (x.z + 2)), bz(x.z + 2);
        const int dx = byz.div(u), r = u - dx * byz.d, dy = bz.div(r), dz = r - dy * bz.d;
        const int3 n = make_int3((lo.x >> (D - l)) - 1 + dx, (lo.y >> (D - l)) - 1 + dy, (lo.z >> (D - l)) - 1 + dz);
        const int lim = 1 << l, s = l - S.rd;
        float v = 0.f;
        if (n.x >= 0 && n.y >= 0 && n.z >= 0 && n.x < lim && n.y < lim && n.z < lim) {
            const int3 A = make_int3(n.x >> s, n.y >> s, n.z >> s);
            if (A.x != R.Rc.x || A.y != R.Rc.y || A.z != R.Rc.z) {
                const int rp = lNbr[nbr_index(A.x - R.Rc.x, A.y - R.Rc.y, A.z - R.Rc.z)];
                if (!(rp >= S.rootOffset && rp < S.rootOffset + S.nr)) {
                    const size_t cell = (size_t(n.x) << (2 * l)) | (size_t(n.y) << l) | size_t(n.z);
                    v = g.dxAt[((size_t(1) << (3 * l)) - 1) / 7 + cell];
                }
            }
        }
        lHdx[t] = v;
        if (v != 0.f) nz |= 1u << l;
    }
    if (nz) atomicOr(&lNz, nz);
    // halo factor tables: per depth, node cell (box-relative) x corner coordinate
    for (int t = threadIdx.x; t < facWords; t += kThreads) {
        int u = t, l = D;
        int3 x = box_ext(tg.ext, D, l);
        int w = (x.x + 2) * kTx + (x.y + 2) * kTy + (x.z + 2) * kTz;
        while (u >= w) {
            u -= w;
            --l;
            x = box_ext(tg.ext, D, l);
            w = (x.x + 2) * kTx + (x.y + 2) * kTy + (x.z + 2) * kTz;
        }
        const float invw = float(1 << l);
        int cell, corner;
        if (u < (x.x + 2) * kTx) {
            cell = (lo.x >> (D - l)) - 1 + u / kTx;
            corner = lo.x + u % kTx;
        } else if ((u -= (x.x + 2) * kTx) < (x.y + 2) * kTy) {
            cell = (lo.y >> (D - l)) - 1 + u / kTy;
            corner = lo.y + u % kTy;
        } else {
            u -= (x.y + 2) * kTy;
            cell = (lo.z >> (D - l)) - 1 + u / kTz;
            corner = lo.z + u % kTz;
        }
        const float c = (0.5f + float(cell)) / invw;
        lFac[t] = bspline((float(corner) * inv - c) * invw);
    }
    // The (cell, corner) items of the tile, from its corner points: a point is
    // owned by the virtual cell of smallest key among the 8 sharing it
    // (owner_of_point; owns_corner asks the same of one cell), and kept when
    // that cell is one of the tile's.  One owner search per point rather than
    // per cell and corner: a cell on a face of the root's cube may own several
    // corners, and the search for the points on those faces diverges.  The
    // items are compacted and shared out over the block.
    const Leaf F = root_leaf(g, S, T.set, R);
    const RecVE ve = rec_ve(lNbr, lSet, F, D);
    constexpr int kPoints = kTilePoints, kPointsPerThread = (kPoints + kThreads - 1) / kThreads;
    int item[kPointsPerThread];
    int nOwn = 0;
#pragma unroll
    for (int j = 0; j < kPointsPerThread; ++j) {
        const int pt = int(threadIdx.x) + j * kThreads;
        item[j] = -1;
        if (pt < kPoints) {
            const int3 p = make_int3(lo.x + pt / (kTy * kTz), lo.y + (pt / kTz) % kTy, lo.z + pt % kTz);
            int slot;
            const int e = owner_of_point(F, p, slot, ve);
            // every point's value index, for sub_classify (when it has the map)
            if (ptOwner)
                ptOwner[size_t(psr_bid) * kPoints + pt] = e >= 0 ? unsigned(slot) * unsigned(nVD) + unsigned(e) : ~0u;
            if (e >= T.start && e < T.start + T.count) {
                item[j] = ((e - T.start) << 3) | slot;
                ++nOwn;
            }
        }
    }
    int first;
    const int nItems = block_excl_scan<kThreads>(nOwn, first);
#pragma unroll
    for (int j = 0; j < kPointsPerThread; ++j)
        if (item[j] >= 0) lItem[first++] = (unsigned short)item[j];
    __syncthreads();
    const float isoV = *iso;
    for (int it = threadIdx.x; it < nItems; it += kThreads) {
        const int b = lItem[it] & 7, e = T.start + (lItem[it] >> 3);
        const Leaf L = leaf_of(g, T.set, S, R, e);
        const int3 o = bourke_offset(b);
        const int3 pt = make_int3(L.c.x + o.x, L.c.y + o.y, L.c.z + o.z);
        const int px = pt.x - lo.x, py = pt.y - lo.y, pz = pt.z - lo.z;
        float val = 0.f;
        int hb = 0, fb = 0;
        for (int l = D; l > S.rd; --l) {
            const int s = l - S.rd;
            const int3 x = box_ext(tg.ext, D, l);
            const int bx = x.x + 2, by = x.y + 2, bz = x.z + 2;
            const int3 cl = make_int3(L.c.x >> (D - l), L.c.y >> (D - l), L.c.z >> (D - l));
            const bool interior = ((cl.x - 1) >> s) == L.Rc.x && ((cl.x + 1) >> s) == L.Rc.x &&
                                  ((cl.y - 1) >> s) == L.Rc.y && ((cl.y + 1) >> s) == L.Rc.y &&
                                  ((cl.z - 1) >> s) == L.Rc.z && ((cl.z + 1) >> s) == L.Rc.z;
            // a depth whose halo d_x are all zero adds only zero terms, which
            // leave the sum as it is (it never holds -0): skipped
            if (!interior && ((lNz >> l) & 1u)) {
                const int3 bl = make_int3((lo.x >> (D - l)) - 1, (lo.y >> (D - l)) - 1, (lo.z >> (D - l)) - 1);
                const float* fx = lFac + fb;
                const float* fy = fx + bx * kTx;
                const float* fz = fy + by * kTy;
                const float* hd = lHdx + hb;
#pragma unroll
                for (int i = 0; i < 27; ++i) {
                    const int nx = cl.x + i / 9 - 1 - bl.x, ny = cl.y + (i / 3) % 3 - 1 - bl.y,
                              nz = cl.z + i % 3 - 1 - bl.z;
                    val += hd[(nx * by + ny) * bz + nz] * fx[nx * kTx + px] * fy[ny * kTy + py] * fz[nz * kTz + pz];
                }
            }
            hb += bx * by * bz;
            fb += bx * kTx + by * kTy + bz * kTz;
        }
        const float4* X4 = reinterpret_cast<const float4*>(lLx + px * Kp);
        const float4* Y4 = reinterpret_cast<const float4*>(lLy + py * Kp);
        const float4* Z4 = reinterpret_cast<const float4*>(lLz + pz * Kp);
        for (int k4 = 0; k4 < Kp / 4; ++k4) {
            const float4 a = X4[k4], yb = Y4[k4], zc = Z4[k4];
            val += a.x * yb.x * zc.x;
            val += a.y * yb.y * zc.y;
            val += a.z * yb.z * zc.z;
            val += a.w * yb.w * zc.w;
        }
        cval[size_t(b) * nVD + e] = val - isoV;
    }
}
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kCornerGridWaves))) sub_corner_values_grid(Grid g, const Tile* __restrict__ tiles, int nVD,
                                                                   const float* __restrict__ iso,
                                                                   float* __restrict__ cval,
                                                                   unsigned* __restrict__ ptOwner) {
    sub_corner_values_grid_body(blockIdx.x, gridDim.x, g, tiles, nVD, iso, cval, ptOwner);
}
PSR_BATCHABLE(sub_corner_values_grid, kThreads)


// LDS words of a small-root tile whose roots (nRoots of them) are at depth rd
// with D - rd <= 2: per root the list tables [3][n][Kp] (n = 2^(D-rd) + 1
// corner coordinates per axis), the halo d_x boxes and their factor tables,
// rounded to 4 words so that every root's tables are 16-byte aligned.
__host__ __device__ inline int small_tile_words(int D, int rd, int nRoots) {
    const int sh = D - rd, n = (1 << sh) + 1, Kp = (27 * (rd + 1) + 3) & ~3;
    int halo = 0, fac = 0;
    for (int l = D; l > rd; --l) {
        const int b = max(1, (1 << sh) >> (D - l)) + 2;
        halo += b * b * b;
        fac += 3 * b * n;
    }
    return nRoots * ((3 * n * Kp + halo + fac + 3) & ~3);
}

// Small roots (8 or 64 cells) when the octree is geometric: the tables of
// sub_corner_values_grid for each root of the tile (the root list at the
// root's corner coordinates, with d_x folded into x; the halo d_x from the
// dense per-depth grids and the halo factors), one thread per corner point of
// a root, evaluated for the root holding the point's owner.  Terms, their
// order and their association are those of sub_corner_values.
__device__ __forceinline__ void sub_corner_values_small_grid_body(int psr_bid, int psr_nbid, const Grid& g, const Tile* __restrict__ tiles,
                                                                              int nVD, const float* __restrict__ iso,
                                                                              float* __restrict__ cval) {
    extern __shared__ float4 ldsDyn[];  // small_tile_words(D, rd, roots of the tile) words
    constexpr int kR = 4;               // roots per tile at most (32 cells of 8)
    __shared__ TileRoot lRoot[kR];
    __shared__ int lNbr[kR * 27];
    __shared__ int lChain[kR * kChain];
    __shared__ unsigned lSet[kR];
    float* lds = reinterpret_cast<float*>(ldsDyn);
    const Tile T = tiles[psr_bid];
    const SetDesc S = g.sets[T.set];
    const int D = g.D, sh = D - S.rd, n = (1 << sh) + 1, P = n * n * n;
    const TileGeom tg = tile_geom(T, S, D);
    const int nR = tg.nRoots, K = tg.K, Kp = (K + 3) & ~3;
    for (int t = threadIdx.x; t < nR * 27; t += kSmallGridThreads) lNbr[t] = g.rootNbr[27 * size_t(T.root) + t];
    for (int t = threadIdx.x; t < nR * kChain; t += kSmallGridThreads)
        lChain[t] = g.rootChain[kChain * size_t(T.root) + t];
    if (int(threadIdx.x) < nR) lSet[threadIdx.x] = g.rootSet[T.root + threadIdx.x];
    fill_roots_rec(g, T, S, tg, lRoot);
    __syncthreads();
    // per root: list x (with d_x) [n][Kp] | list y | list z | halo d_x | halo factors
    const int listWords = 3 * n * Kp, perRoot = small_tile_words(D, S.rd, 1);
    const float inv = 1.0f / float(1 << D);
    for (int t = threadIdx.x; t < nR * Kp; t += kSmallGridThreads) {
        const int r = t / Kp, k = t - r * Kp;
        const int node = k < K ? lChain[r * kChain + k / 27] : -1;
        float4 b = make_float4(0.f, 0.f, 0.f, 0.f);
        float d = 0.f;
        if (node >= 0) {
            const float4 col = nbr_col(g, k % 27, node);
            b = make_float4(col.x, col.y, col.z, float(1 << (S.rd - k / 27)));
            d = col.w;
        }
        const int3 lo = lRoot[r].first;
        float* L = lds + r * perRoot;
        for (int cc = 0; cc < n; ++cc) {
            L[cc * Kp + k] = node < 0 ? 0.f : d * bspline((float(lo.x + cc) * inv - b.x) * b.w);
            L[(n + cc) * Kp + k] = node < 0 ? 0.f : bspline((float(lo.y + cc) * inv - b.y) * b.w);
            L[(2 * n + cc) * Kp + k] = node < 0 ? 0.f : bspline((float(lo.z + cc) * inv - b.z) * b.w);
        }
    }
    // halo d_x, depth D first: 0 inside the roots of the set and where no node is
    for (int t = threadIdx.x; t < nR * tg.haloPerRoot; t += kSmallGridThreads) {
        const int r = t / tg.haloPerRoot;
        int u = t - r * tg.haloPerRoot, l = D;
        int3 x = box_ext(tg.ext, D, l);
        while (u >= (x.x + 2) * (x.y + 2) * (x.z + 2)) {
            u -= (x.x + 2) * (x.y + 2) * (x.z + 2);
            --l;
            x = box_ext(tg.ext, D, l);
        }
        const int bz = x.z + 2, by = x.y + 2;
        const int dx = u / (by * bz), rem = u - dx * by * bz, dy = rem / bz, dz = rem - dy * bz;
        const TileRoot R = lRoot[r];
        const int3 c = make_int3((R.first.x >> (D - l)) - 1 + dx, (R.first.y >> (D - l)) - 1 + dy,
                                 (R.first.z >> (D - l)) - 1 + dz);
        const int lim = 1 << l, s = l - S.rd;
        float v = 0.f;
        if (c.x >= 0 && c.y >= 0 && c.z >= 0 && c.x < lim && c.y < lim && c.z < lim) {
            const int3 A = make_int3(c.x >> s, c.y >> s, c.z >> s);
            if (A.x != R.Rc.x || A.y != R.Rc.y || A.z != R.Rc.z) {
                const int rp = lNbr[r * 27 + nbr_index(A.x - R.Rc.x, A.y - R.Rc.y, A.z - R.Rc.z)];
                if (!(rp >= S.rootOffset && rp < S.rootOffset + S.nr)) {
                    const size_t cell = (size_t(c.x) << (2 * l)) | (size_t(c.y) << l) | size_t(c.z);
                    v = g.dxAt[((size_t(1) << (3 * l)) - 1) / 7 + cell];
                }
            }
        }
        lds[r * perRoot + listWords + (t - r * tg.haloPerRoot)] = v;
    }
    // halo factors: per depth, per axis, node cell (box-relative) x corner coordinate
    int facPerRoot = 0;
    for (int l = D; l > S.rd; --l) facPerRoot += 3 * (box_ext(tg.ext, D, l).x + 2) * n;
    for (int t = threadIdx.x; t < nR * facPerRoot; t += kSmallGridThreads) {
        const int r = t / facPerRoot;
        int u = t - r * facPerRoot, l = D;
        int3 x = box_ext(tg.ext, D, l);
        while (u >= 3 * (x.x + 2) * n) {
            u -= 3 * (x.x + 2) * n;
            --l;
            x = box_ext(tg.ext, D, l);
        }
        const int b = x.x + 2, axis = u / (b * n), w = u - axis * b * n, ci = w / n, cc = w - ci * n;
        const int3 f = lRoot[r].first;
        const int fa = axis == 0 ? f.x : (axis == 1 ? f.y : f.z);
        const float invw = float(1 << l);
        const int cell = (fa >> (D - l)) - 1 + ci;
        const float c = (0.5f + float(cell)) / invw;
        lds[r * perRoot + listWords + tg.haloPerRoot + (t - r * facPerRoot)] =
            bspline((float(fa + cc) * inv - c) * invw);
    }
    __syncthreads();
    const int rr = threadIdx.x / P, pi = threadIdx.x - rr * P;
    if (rr >= nR) return;
    const TileRoot R = lRoot[rr];
    const int3 pt = make_int3(R.first.x + pi / (n * n), R.first.y + (pi / n) % n, R.first.z + pi % n);
    const Leaf Lr = root_leaf(g, S, T.set, R);
    const RecVE ve = rec_ve(lNbr + 27 * rr, lSet[rr], Lr, D);
    int m;
    const int best = point_owner(pt, m, ve);
    if (best < 0 || ((best - S.dEnum) >> (3 * sh)) != tg.q0 + rr) return;
    const int slot = bourke_slot(m & 1, (m >> 1) & 1, (m >> 2) & 1);
    const int3 oc = make_int3(pt.x - (m & 1), pt.y - ((m >> 1) & 1), pt.z - ((m >> 2) & 1));
    const int px = pt.x - R.first.x, py = pt.y - R.first.y, pz = pt.z - R.first.z;
    const float* base = lds + rr * perRoot;
    const float* hd = base + listWords;
    const float* fac = hd + tg.haloPerRoot;
    float val = 0.f;
    for (int l = D; l > S.rd; --l) {
        const int s = l - S.rd;
        const int3 x = box_ext(tg.ext, D, l);
        const int b = x.x + 2;
        const int3 cl = make_int3(oc.x >> (D - l), oc.y >> (D - l), oc.z >> (D - l));
        const bool interior = ((cl.x - 1) >> s) == R.Rc.x && ((cl.x + 1) >> s) == R.Rc.x &&
                              ((cl.y - 1) >> s) == R.Rc.y && ((cl.y + 1) >> s) == R.Rc.y &&
                              ((cl.z - 1) >> s) == R.Rc.z && ((cl.z + 1) >> s) == R.Rc.z;
        if (!interior) {
            const int3 bl = make_int3((R.first.x >> (D - l)) - 1, (R.first.y >> (D - l)) - 1,
                                      (R.first.z >> (D - l)) - 1);
            const float* fx = fac;
            const float* fy = fx + b * n;
            const float* fz = fy + b * n;
#pragma unroll
            for (int i = 0; i < 27; ++i) {
                const int nx = cl.x + i / 9 - 1 - bl.x, ny = cl.y + (i / 3) % 3 - 1 - bl.y, nz = cl.z + i % 3 - 1 - bl.z;
                val += hd[(nx * b + ny) * b + nz] * fx[nx * n + px] * fy[ny * n + py] * fz[nz * n + pz];
            }
        }
        hd += b * b * b;
        fac += 3 * b * n;
    }
    const float4* X4 = reinterpret_cast<const float4*>(base + px * Kp);
    const float4* Y4 = reinterpret_cast<const float4*>(base + (n + py) * Kp);
    const float4* Z4 = reinterpret_cast<const float4*>(base + (2 * n + pz) * Kp);
    for (int k4 = 0; k4 < Kp / 4; ++k4) {
        const float4 a = X4[k4], yb = Y4[k4], zc = Z4[k4];
        val += a.x * yb.x * zc.x;
        val += a.y * yb.y * zc.y;
        val += a.z * yb.z * zc.z;
        val += a.w * yb.w * zc.w;
    }
    cval[size_t(slot) * nVD + best] = val - *iso;
}
__global__ void __launch_bounds__(kSmallGridThreads) sub_corner_values_small_grid(Grid g, const Tile* __restrict__ tiles,
                                                                              int nVD, const float* __restrict__ iso,
                                                                              float* __restrict__ cval) {
    sub_corner_values_small_grid_body(blockIdx.x, gridDim.x, g, tiles, nVD, iso, cval);
}
PSR_BATCHABLE(sub_corner_values_small_grid, kSmallGridThreads)


// Exclusive prefix of v over the block's threads and the block total (every
// thread of the block must call it).
template <int NT>
__device__ __forceinline__ int block_exclusive(int v, int* sh, int& total) {
    constexpr int kW = NT / PSR_WAVE;
    const int lane = __lane_id(), w = threadIdx.x / PSR_WAVE;
    int x = v;
#pragma unroll
    for (int o = 1; o < PSR_WAVE; o <<= 1) {
        int y = __shfl_up(x, o, PSR_WAVE);
        if (lane >= o) x += y;
    }
    if (lane == PSR_WAVE - 1) sh[w] = x;
    __syncthreads();
    int before = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < kW; ++k) {
        int t = sh[k];
        before += k < w ? t : 0;
        tot += t;
    }
    __syncthreads();
    total = tot;
    return before + x - v;
}

// Crossing point of edge k of the depth-D cell c with corner values v
// (generateIntersectionPoint :440-474).
__device__ __forceinline__ float3 cell_crossing(int3 c, int k, const float* v, float inv) {
    const int b1 = kEdgeVertex[k][0], b2 = kEdgeVertex[k][1];
    const int3 o1 = bourke_offset(b1), o2 = bourke_offset(b2);
    const float3 p1 = make_float3(float(c.x + o1.x) * inv, float(c.y + o1.y) * inv, float(c.z + o1.z) * inv);
    const float3 p2 = make_float3(float(c.x + o2.x) * inv, float(c.y + o2.y) * inv, float(c.z + o2.z) * inv);
    return interpolate_point(p1, p2, k >> 2, v[b1], v[b2]);
}

// Cell word: bits 0-11 the owned edges that cross the surface, 12-23 those
// whose vertex is kept (markValidMeshVertexIndex :1170-1181 drops |x| < 1e-6),
// 24-31 the cube class.
constexpr unsigned kEdgeBits = 0xFFFu;

// Per cell: corner values, cube class, owned crossing edges and kept vertices;
// per tile: crossing, kept-vertex and triangle counts, and each cell's first
// crossing within the tile.  Corner values of crossing cells are kept
// for the emission pass.
__device__ __forceinline__ void sub_classify_body(int psr_bid, int psr_nbid, const Grid& g, const Tile* __restrict__ tiles, int nVD,
                                                         const float* __restrict__ cval, float* __restrict__ vals,
                                                         unsigned* __restrict__ cellWord,
                                                         int* __restrict__ crossPrefix, int4* __restrict__ tileCnt,
                                                         const unsigned* __restrict__ ptOwner, int nMapped) {
    __shared__ TileRoot lRoots[kMaxRootsPerTile];
    __shared__ int lNbr[kMaxRootsPerTile * 27];
    __shared__ unsigned lSet[kMaxRootsPerTile];
    __shared__ int lSum[kThreads / PSR_WAVE];
    __shared__ float lVal[kClassifyPoints];
    __shared__ unsigned short lItem[kThreads * 12];  // crossing edges as (cell << 4 | edge), cell order
    __shared__ unsigned lOwn[kThreads];              // per cell: owned crossing edges, kept ones << 12
    const Tile T = tiles[psr_bid];
    const SetDesc S = g.sets[T.set];
    const int D = g.D;
    const TileGeom tg = tile_geom(T, S, D);
    fill_roots_rec(g, T, S, tg, lRoots);
    for (int t = threadIdx.x; t < tg.nRoots * 27; t += kThreads) lNbr[t] = g.rootNbr[27 * size_t(T.root) + t];
    if (int(threadIdx.x) < tg.nRoots) lSet[threadIdx.x] = g.rootSet[T.root + threadIdx.x];
    __syncthreads();
    // the values at the corner points of each root's part of the tile, read
    // once per point (its owner is the same whichever cell asks); the first
    // nMapped tiles (one-root tiles) take each point's value index from the
    // owners that sub_corner_values_grid found for the same points
    const int ex = tg.ext.x + 1, ey = tg.ext.y + 1, ez = tg.ext.z + 1, PP = ex * ey * ez;
    if (psr_bid < nMapped) {
        for (int t = threadIdx.x; t < kTilePoints; t += kThreads) {
            const unsigned oi = ptOwner[size_t(psr_bid) * kTilePoints + t];
            lVal[t] = oi != ~0u ? cval[oi] : 0.f;
        }
    } else
    for (int t = threadIdx.x; t < tg.nRoots * PP; t += kThreads) {
        const int rr = t / PP, u = t - rr * PP;
        const int ux = u / (ey * ez), uy = (u / ez) % ey, uz = u % ez;
        const TileRoot R = lRoots[rr];
        const Leaf Lr = root_leaf(g, S, T.set, R);
        int slot;
        const int owner = owner_of_point(Lr, make_int3(R.first.x + ux, R.first.y + uy, R.first.z + uz), slot,
                                         rec_ve(lNbr + 27 * rr, lSet[rr], Lr, D));
        lVal[t] = owner >= 0 ? cval[size_t(slot) * nVD + owner] : 0.f;
    }
    __syncthreads();
    const bool active = int(threadIdx.x) < T.count;
    const int e = T.start + (active ? int(threadIdx.x) : 0);
    const int rr = tg.oneRoot ? 0 : int(threadIdx.x) >> (3 * (D - S.rd));
    const float inv = 1.0f / float(1 << D);
    // corner values, cube class and the edges whose ends differ in sign
    float v[8];
    int cls = 0;
    unsigned cross = 0;
    if (active) {
        const int3 c = cell_of(g, S, lRoots[rr], e);
        const int3 f = lRoots[rr].first;
        const float* pv = lVal + rr * PP;
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const int3 o = bourke_offset(b);
            v[b] = pv[((c.x - f.x + o.x) * ey + (c.y - f.y + o.y)) * ez + (c.z - f.z + o.z)];
            if (v[b] < 0.f) cls |= 1 << b;
        }
#pragma unroll
        for (int k = 0; k < 12; ++k)
            if (v[kEdgeVertex[k][0]] * v[kEdgeVertex[k][1]] <= 0.f) cross |= 1u << k;
    }
    // ownership of those edges: (cell, edge) items shared out over the block,
    // owned and kept edges gathered per cell in LDS
    lOwn[threadIdx.x] = 0u;
    int first;
    const int nItems = block_excl_scan<kThreads>(__popc(cross), first);
    for (int k = 0; k < 12; ++k)
        if ((cross >> k) & 1u) lItem[first++] = (threadIdx.x << 4) | k;
    __syncthreads();
    for (int it = threadIdx.x; it < nItems; it += kThreads) {
        const int cell = lItem[it] >> 4, k = lItem[it] & 15, r = tg.oneRoot ? 0 : cell >> (3 * (D - S.rd));
        const int ec = T.start + cell;
        const Leaf L = leaf_of(g, T.set, S, lRoots[r], ec);
        int kind;
        if (edge_owner(L, D, k, kind, rec_ve(lNbr + 27 * r, lSet[r], L, D)) != ec || kind != k) continue;
        const int3 f = lRoots[r].first;
        const float* pv = lVal + r * PP;
        float w[8];
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const int3 o = bourke_offset(b);
            w[b] = pv[((L.c.x - f.x + o.x) * ey + (L.c.y - f.y + o.y)) * ez + (L.c.z - f.z + o.z)];
        }
        unsigned bits = 1u << k;
        if (!(fabsf(cell_crossing(L.c, k, w, inv).x) < 1e-6f)) bits |= 1u << (12 + k);
        atomicOr(&lOwn[cell], bits);
    }
    __syncthreads();
    const unsigned mask = lOwn[threadIdx.x] & kEdgeBits, keep = (lOwn[threadIdx.x] >> 12) & kEdgeBits;
    int nt = 0;
    if (active) {
        cellWord[e] = mask | (keep << 12) | (unsigned(cls) << 24);
        if (mask)
#pragma unroll
            for (int b = 0; b < 8; ++b) vals[size_t(b) * nVD + e] = v[b];
        nt = mc_tri_count(cls);
    }
    // crossings (prefix and total) and triangles in one scan, kept vertices in another
    int tct, tk;
    const int pre = block_exclusive<kThreads>((__popc(mask) << 16) | nt, lSum, tct);
    if (active) crossPrefix[e] = pre >> 16;
    (void)block_exclusive<kThreads>(__popc(keep), lSum, tk);
    if (threadIdx.x == 0) tileCnt[psr_bid] = make_int4(tct >> 16, tk, tct & 0xFFFF, 0);
}
__global__ void __launch_bounds__(kThreads) sub_classify(Grid g, const Tile* __restrict__ tiles, int nVD,
                                                         const float* __restrict__ cval, float* __restrict__ vals,
                                                         unsigned* __restrict__ cellWord,
                                                         int* __restrict__ crossPrefix, int4* __restrict__ tileCnt,
                                                         const unsigned* __restrict__ ptOwner, int nMapped) {
    sub_classify_body(blockIdx.x, gridDim.x, g, tiles, nVD, cval, vals, cellWord, crossPrefix, tileCnt, ptOwner, nMapped);
}
PSR_BATCHABLE(sub_classify, kThreads)


// Exclusive scan of the per-tile counts in one workgroup; totals[0..2] are the
// chunk's crossings, kept vertices and triangles; totals[3] = *keptMain, the
// vertices kept by the depth-D batch.
constexpr int kScanThreads = 1024;
__device__ __forceinline__ void sub_tile_scan_body(int psr_bid, int psr_nbid, const int4* __restrict__ cnt, int n,
                                                              int4* __restrict__ base, const int* __restrict__ keptMain,
                                                              int* __restrict__ totals) {
    constexpr int kW = kScanThreads / PSR_WAVE;
    __shared__ int3 lSum3[kW];
    const int per = (n + kScanThreads - 1) / kScanThreads;
    const int i0 = min(n, int(threadIdx.x) * per), i1 = min(n, i0 + per);
    // up to kPerMax counts a thread are held in registers, read at once
    constexpr int kPerMax = 8;
    const bool held = per <= kPerMax;
    int4 v[kPerMax];
    int a = 0, b = 0, c = 0;
    if (held) {
#pragma unroll
        for (int k = 0; k < kPerMax; ++k) v[k] = i0 + k < i1 ? cnt[i0 + k] : make_int4(0, 0, 0, 0);
#pragma unroll
        for (int k = 0; k < kPerMax; ++k) {
            a += v[k].x;
            b += v[k].y;
            c += v[k].z;
        }
    } else {
        for (int i = i0; i < i1; ++i) {
            const int4 t = cnt[i];
            a += t.x;
            b += t.y;
            c += t.z;
        }
    }
    // the three exclusive scans in one pass over the block (one barrier pair)
    const int lane = __lane_id(), w = threadIdx.x / PSR_WAVE;
    int xa = a, xb = b, xc = c;
#pragma unroll
    for (int o = 1; o < PSR_WAVE; o <<= 1) {
        const int ya = __shfl_up(xa, o, PSR_WAVE), yb = __shfl_up(xb, o, PSR_WAVE), yc = __shfl_up(xc, o, PSR_WAVE);
        if (lane >= o) xa += ya, xb += yb, xc += yc;
    }
    if (lane == PSR_WAVE - 1) lSum3[w] = make_int3(xa, xb, xc);
    __syncthreads();
    int ta = 0, tb = 0, tc = 0, pa = 0, pb = 0, pc = 0;
#pragma unroll
    for (int k = 0; k < kW; ++k) {
        const int3 t = lSum3[k];
        if (k < w) pa += t.x, pb += t.y, pc += t.z;
        ta += t.x, tb += t.y, tc += t.z;
    }
    a = pa + xa - a;
    b = pb + xb - b;
    c = pc + xc - c;
    if (held) {
#pragma unroll
        for (int k = 0; k < kPerMax; ++k) {
            if (i0 + k >= i1) break;
            base[i0 + k] = make_int4(a, b, c, 0);
            a += v[k].x;
            b += v[k].y;
            c += v[k].z;
        }
    } else {
        for (int i = i0; i < i1; ++i) {
            const int4 t = cnt[i];
            base[i] = make_int4(a, b, c, 0);
            a += t.x;
            b += t.y;
            c += t.z;
        }
    }
    if (threadIdx.x == 0) {
        totals[0] = ta;
        totals[1] = tb;
        totals[2] = tc;
        totals[3] = *keptMain;
    }
}
__global__ void __launch_bounds__(kScanThreads) sub_tile_scan(const int4* __restrict__ cnt, int n,
                                                              int4* __restrict__ base, const int* __restrict__ keptMain,
                                                              int* __restrict__ totals) {
    sub_tile_scan_body(blockIdx.x, gridDim.x, cnt, n, base, keptMain, totals);
}
PSR_BATCHABLE(sub_tile_scan, kScanThreads)


// Kept crossing points and the triangles of the subdivision batches.  A
// triangle index is the number of vertices kept before the batch plus the raw
// offset inside the batch; indices are not remapped around dropped vertices
// (insertTriangle :1247-1282).  *keptMain + keptBefore / triBefore: vertices
// kept and triangles written before this chunk (*keptMain: those of the
// depth-D batch).
__device__ __forceinline__ void sub_emit_body(int psr_bid, int psr_nbid, const Grid& g, const Tile* __restrict__ tiles, int nVD,
                                                     const float* __restrict__ vals,
                                                     const unsigned* __restrict__ cellWord,
                                                     const int* __restrict__ crossPrefix,
                                                     const int4* __restrict__ tileCnt,
                                                     const int4* __restrict__ tileBase,
                                                     const int* __restrict__ keptMain, int keptBefore, int triBefore,
                                                     float* __restrict__ outV, int* __restrict__ outT,
                                                     unsigned* __restrict__ outKey) {
    __shared__ int lSum[kThreads / PSR_WAVE];
    __shared__ TileRoot lRoots[kMaxRootsPerTile];
    __shared__ int lNbr[kMaxRootsPerTile * 27];
    __shared__ unsigned lSet[kMaxRootsPerTile];
    const int4 cnt = tileCnt[psr_bid];
    if (cnt.y == 0 && cnt.z == 0) return;  // the same for the whole block
    const Tile T = tiles[psr_bid];
    const SetDesc S = g.sets[T.set];
    const TileGeom tg = tile_geom(T, S, g.D);
    // the tile's roots (made visible by the block scans' barriers)
    fill_roots_rec(g, T, S, tg, lRoots);
    for (int t = threadIdx.x; t < tg.nRoots * 27; t += kThreads) lNbr[t] = g.rootNbr[27 * size_t(T.root) + t];
    if (int(threadIdx.x) < tg.nRoots) lSet[threadIdx.x] = g.rootSet[T.root + threadIdx.x];
    const bool active = int(threadIdx.x) < T.count;
    const int e = T.start + (active ? int(threadIdx.x) : 0);
    const unsigned w = active ? cellWord[e] : 0u;
    const unsigned keep = (w >> 12) & kEdgeBits;
    const int nt = mc_tri_count(int(w >> 24));
    int tk, tt;
    const int kp = block_exclusive<kThreads>(__popc(keep), lSum, tk);
    const int tp = block_exclusive<kThreads>(nt, lSum, tt);
    if (!active || (keep == 0 && nt == 0)) return;
    const int rr = tg.oneRoot ? 0 : int(threadIdx.x) >> (3 * (g.D - S.rd));
    const Leaf L = leaf_of(g, T.set, S, lRoots[rr], e);
    const RecVE ve = rec_ve(lNbr + 27 * rr, lSet[rr], L, g.D);
    const int4 tb = tileBase[psr_bid];
    keptBefore += *keptMain;
    if (keep) {
        float v[8];
#pragma unroll
        for (int b = 0; b < 8; ++b) v[b] = vals[size_t(b) * nVD + e];
        const float inv = 1.0f / float(1 << g.D);
        size_t a = size_t(keptBefore) + tb.y + kp;
        for (int k = 0; k < 12; ++k) {
            if (!((keep >> k) & 1u)) continue;
            const float3 p = cell_crossing(L.c, k, v, inv);
            outV[3 * a] = p.x;
            outV[3 * a + 1] = p.y;
            outV[3 * a + 2] = p.z;
            const int3 o1 = bourke_offset(kEdgeVertex[k][0]), o2 = bourke_offset(kEdgeVertex[k][1]);
            outKey[a] = segment_key(make_int3(L.c.x + o1.x, L.c.y + o1.y, L.c.z + o1.z),
                                    make_int3(L.c.x + o2.x, L.c.y + o2.y, L.c.z + o2.z), g.D);
            ++a;
        }
    }
    if (nt) {
        const int4 sb = tileBase[S.firstTile];  // the batch's first raw vertex and kept count before it
        const int batchKept = keptBefore + sb.y;
        const uint64_t mw = mc_word(int(w >> 24));
        const size_t base = size_t(triBefore) + tb.z + tp;
        for (int k = 0; k < 3 * nt; ++k) {
            const int ed = int((mw >> (4 * k)) & 0xF);
            int kind;
            const int owner = edge_owner(L, g.D, ed, kind, ve);
            const unsigned ow = cellWord[owner];
            const int raw = tileBase[S.firstTile + (owner - S.dEnum) / kTile].x + crossPrefix[owner] +
                            __popc(ow & kEdgeBits & ((1u << kind) - 1u));
            outT[3 * base + k] = batchKept + (raw - sb.x);
        }
    }
}
__global__ void __launch_bounds__(kThreads) sub_emit(Grid g, const Tile* __restrict__ tiles, int nVD,
                                                     const float* __restrict__ vals,
                                                     const unsigned* __restrict__ cellWord,
                                                     const int* __restrict__ crossPrefix,
                                                     const int4* __restrict__ tileCnt,
                                                     const int4* __restrict__ tileBase,
                                                     const int* __restrict__ keptMain, int keptBefore, int triBefore,
                                                     float* __restrict__ outV, int* __restrict__ outT,
                                                     unsigned* __restrict__ outKey) {
    sub_emit_body(blockIdx.x, gridDim.x, g, tiles, nVD, vals, cellWord, crossPrefix, tileCnt, tileBase, keptMain, keptBefore, triBefore, outV, outT, outKey);
}
PSR_BATCHABLE(sub_emit, kThreads)


// Per root: the position in roots of each real neighbour (or -1), and the
// chain of the root and its ancestors.
// Its set: coarse roots (depth < finerDepth) form sets of one; the finer roots
// of a depth form one set, so a neighbour belongs to a finer root's set iff it
// is a root (it has the root's depth).
struct RootRecArgs {
    int nRoots, M;
    Levels Lv;
    int finerDepth;
    const int* roots;
    const int* rootPos;
    const int* neigh;
    const int* parent;
    const int* key;
    int* rootNbr;
    int* rootChain;
    int4* rootCell;
    unsigned* rootSet;
};
__device__ __forceinline__ void root_records(const RootRecArgs& a, int block) {
    const int nRoots = a.nRoots, M = a.M, finerDepth = a.finerDepth;
    const Levels& Lv = a.Lv;
    const int* __restrict__ roots = a.roots;
    const int* __restrict__ rootPos = a.rootPos;
    const int* __restrict__ neigh = a.neigh;
    const int* __restrict__ parent = a.parent;
    const int* __restrict__ key = a.key;
    const size_t t = size_t(block) * kThreads + threadIdx.x;
    if (t >= size_t(27) * nRoots) return;
    const int q = int(t / 27), i = int(t - size_t(q) * 27);
    const int root = roots[q];
    const int nb = neigh[i * M + root];
    a.rootNbr[t] = nb < 0 ? -1 : rootPos[nb];
    if (i == 1) {
        int rd = 0;
        while (rd < Lv.D && root >= Lv.base[rd + 1]) ++rd;
        const int3 c = demorton(key[root] >> (3 * (Lv.D - rd)), rd);
        a.rootCell[q] = make_int4(c.x, c.y, c.z, root);
        unsigned m = 1u << 13;
        if (rd >= finerDepth) {
            int nbs[27];
#pragma unroll
            for (int k = 0; k < 27; ++k) nbs[k] = neigh[k * M + root];
#pragma unroll
            for (int k = 0; k < 27; ++k)
                if (nbs[k] >= 0 && rootPos[nbs[k]] >= 0) m |= 1u << k;
        }
        a.rootSet[q] = m;
    }
    if (i == 0) {
        int node = root;
        for (int u = 0; u < kChain; ++u) {
            a.rootChain[kChain * size_t(q) + u] = node;
            node = node < 0 ? -1 : parent[node];
        }
    }
}

// d_x of every node at its cell of the dense grid of its depth.
__device__ __forceinline__ void scatter_dx_grid_body(int psr_bid, int psr_nbid, int M, const Levels& L, const int* __restrict__ key,
                                                            const float* __restrict__ dx, float* __restrict__ grid) {
    const int i = psr_bid * kThreads + threadIdx.x;
    if (i >= M) return;
    int l = 0;
    while (l < L.D && i >= L.base[l + 1]) ++l;
    grid[dense_grid_index(key[i], l, L.D)] = dx[i];
}
__global__ void __launch_bounds__(kThreads) scatter_dx_grid(int M, Levels L, const int* __restrict__ key,
                                                            const float* __restrict__ dx, float* __restrict__ grid) {
    scatter_dx_grid_body(blockIdx.x, gridDim.x, M, L, key, dx, grid);
}
PSR_BATCHABLE(scatter_dx_grid, kThreads)


__device__ __forceinline__ void zero_floats_body(int psr_bid, int psr_nbid, float* __restrict__ a, size_t n) {
    const size_t i = size_t(psr_bid) * kThreads + threadIdx.x;
    if (i < n) a[i] = 0.f;
}
__global__ void __launch_bounds__(kThreads) zero_floats(float* __restrict__ a, size_t n) {
    zero_floats_body(blockIdx.x, gridDim.x, a, n);
}
PSR_BATCHABLE(zero_floats, kThreads)


// markValidMeshVertexIndex (:1170-1181): a vertex with |x| < 1e-6 is dropped;
// triangle indices are not remapped (insertTriangle :1247-1282).  The kept
// vertices are compacted in order in the same launch; their count goes to t.total.
struct KeepArgs {
    const float* v;
    const unsigned* key;
    int n;
    ScanTicket t;
    float* out;
    unsigned* outKey;
};
__device__ __forceinline__ void keep_compact(const KeepArgs& ka) {
    const float* __restrict__ v = ka.v;
    float* __restrict__ out = ka.out;
    const int n = ka.n;
    const ScanTicket& t = ka.t;
    const int tile = ticket_tile(t);
    const int i = tile * kThreads + threadIdx.x;
    float x = 0.f, y = 0.f, z = 0.f;
    int keep = 0;
    if (i < n) {
        x = v[3 * size_t(i)];
        keep = fabsf(x) < 1e-6f ? 0 : 1;
        if (keep) {
            y = v[3 * size_t(i) + 1];
            z = v[3 * size_t(i) + 2];
        }
    }
    const int a = ticket_prefix<kThreads>(t, tile, keep);
    if (!keep) return;
    out[3 * size_t(a)] = x;
    out[3 * size_t(a) + 1] = y;
    out[3 * size_t(a) + 2] = z;
    ka.outKey[a] = ka.key[i];
}

// The chunk's sets, read from mapped host memory on the compute queue (a copy
// engine would add a hand-over between queues), and their tiles: set k
// (blocks (k, *)) has cells = nr << 3(D - rd) cells cut into tiles of kTile
// cells from firstTile on and, for roots less than 3 levels above depth D,
// into small tiles of 32 (one level) or 64 (two levels) cells from
// firstSmall[k] on.  A tile's root is the root holding its first cell.
struct TileArgs {
    const SetDesc* hostSets;
    const int* firstSmall;
    int D, parts;  // parts: blocks per set
    SetDesc* sets;
    Tile* tiles;
    Tile* small;
};
__device__ __forceinline__ void make_tiles(const TileArgs& a, int block) {
    const int k = block / a.parts, part = block - k * a.parts, D = a.D;
    const SetDesc S = a.hostSets[k];
    Tile* __restrict__ tiles = a.tiles;
    Tile* __restrict__ small = a.small;
    if (part == 0 && threadIdx.x == 0) a.sets[k] = S;
    const int sh = 3 * (D - S.rd), cells = S.nr << sh;
    const int j0 = part * kThreads + int(threadIdx.x), stride = a.parts * kThreads;
    for (int j = j0; j < (cells + kTile - 1) / kTile; j += stride) {
        const int t = j * kTile;
        tiles[S.firstTile + j] = Tile{k, S.dEnum + t, min(kTile, cells - t), S.rootOffset + (t >> sh)};
    }
    if (sh >= 9) return;
    const int step = sh == 3 ? 32 : 64, f = a.firstSmall[k];
    for (int j = j0; j < (cells + step - 1) / step; j += stride) {
        const int t = j * step;
        small[f + j] = Tile{k, S.dEnum + t, min(step, cells - t), S.rootOffset + (t >> sh)};
    }
}

// The subdivision's set-up in one launch: blocks [0, keepBlocks) compact the
// vertices kept by the depth-D batch, the next rootBlocks blocks write the
// root records, the remaining blocks expand the tiles (parts blocks per set).
__device__ __forceinline__ void sub_setup_body(int psr_bid, int psr_nbid, const KeepArgs& ka, int keepBlocks, const RootRecArgs& ra, int rootBlocks,
                                                      const TileArgs& ta) {
    const int b = psr_bid;
    if (b < keepBlocks) {
        keep_compact(ka);
    } else if (b < keepBlocks + rootBlocks) {
        root_records(ra, b - keepBlocks);
    } else {
        make_tiles(ta, b - keepBlocks - rootBlocks);
    }
}
__global__ void __launch_bounds__(kThreads) sub_setup(KeepArgs ka, int keepBlocks, RootRecArgs ra, int rootBlocks,
                                                      TileArgs ta) {
    sub_setup_body(blockIdx.x, gridDim.x, ka, keepBlocks, ra, rootBlocks, ta);
}
PSR_BATCHABLE(sub_setup, kThreads)


// LDS words of a sub_corner_values_grid tile whose root is at depth rd, and
// its halo entries.
int grid_tile_words(int D, int rd, int* haloOut = nullptr) {
    int halo = 0, fac = 0;
    for (int l = D; l > rd; --l) {
        const int x = std::max(1, 4 >> (D - l)) + 2, y = std::max(1, 8 >> (D - l)) + 2, z = y;
        halo += x * y * z;
        fac += x * kTx + y * kTy + z * kTz;
    }
    if (haloOut) *haloOut = halo;
    return halo + fac + ((27 * (rd + 1) + 3) & ~3) * kTw;
}

// Default depth-D virtual cells per chunk of subdivision sets; a set larger
// than the chunk size forms a chunk of its own.
constexpr size_t kChunkCells = size_t(1) << 24;

}  // namespace

// ---- depth-D marching cubes: the first output batch; its triangles are
// written in place (raw crossing ids: batch start 0, nothing kept before it)
void extract_mesh_a(Ctx& c) {
    const Levels& L = c.oct.L;
    const int nD = L.count[L.D];
    float* mainV = c.buf<float>("raw_main_v", size_t(36) * nD);  // at most one crossing per depth-D edge
    unsigned* mainKey = c.buf<unsigned>("raw_main_key", size_t(12) * nD);
    int* outT = c.buf_keep<int>("mesh_triangles", size_t(15) * nD, 0);
    c.mc = MainMc{};
    c.mesh.deferred = false;
    depth_mc_launch(c, c.mc, mainV, mainKey, outT);
}

// After the wait for the depth-D counts: the subdivision and the output.
void extract_mesh_b(Ctx& c) {
    Octree& o = c.oct;
    const Levels& L = o.L;
    const int M = o.M, D = L.D, nD = L.count[D], offD = L.base[D];
    hipStream_t st = c.stream;
    MainMc& mm = c.mc;
    depth_mc_read(c, mm);
    float* mainV = c.buf<float>("raw_main_v", size_t(36) * nD);
    unsigned* mainKey = c.buf<unsigned>("raw_main_key", size_t(12) * nD);
    int* outT = c.buf_keep<int>("mesh_triangles", size_t(15) * nD, 0);
    c.stats.n_subdivision_roots = mm.nRoots;
    // ---- subdivision sets: coarse roots one per set, finer roots one set per depth
    // (ComputeTriangleIndices.h:371: roots above depth D-4 are coarse)
    const int finerDepth = D - 4;
    std::vector<SetDesc> sets;
    std::vector<size_t> setCells;
    int rootCursor = 0, nCoarse = 0;
    for (int d = 0; d < D; ++d) {
        int cnt = mm.rootsPerDepth[d];
        if (cnt == 0) continue;
        auto make = [&](int nr) {
            SetDesc S{};
            S.rd = d;
            S.nr = nr;
            S.rootOffset = rootCursor;
            size_t cells = size_t(nr) << (3 * (D - d));
            if (cells * 12 + 1 > size_t(INT32_MAX)) throw std::runtime_error("subdivision set exceeds 32-bit indexing");
            rootCursor += nr;
            sets.push_back(S);
            setCells.push_back(cells);
        };
        if (d < finerDepth) {
            for (int k = 0; k < cnt; ++k) make(1);
            nCoarse += cnt;
        } else {
            make(cnt);
        }
    }
    c.stats.n_coarse_roots = nCoarse;
    const int nSets = int(sets.size());
    const size_t chunkCells = std::min<size_t>(c.cfg.subdiv_chunk_nodes ? c.cfg.subdiv_chunk_nodes : kChunkCells,
                                               size_t(INT32_MAX - 1) / 12);
    // One chunk: nothing waits before the emission.  Vertices and triangles go
    // to buffers sized by bound (a root of s^3 cells has 3s(s+1)^2 edges, a cell
    // at most 5 triangles) and the counts stay in the mailbox until the call's
    // final wait: slot 3 the vertices kept by the depth-D batch, slots 1 and 2
    // the subdivision's kept vertices and triangles.
    size_t cellsAll = 0, vBound = 0;
    for (int k = 0; k < nSets; ++k) {
        const size_t sk = size_t(1) << (D - sets[k].rd);
        cellsAll += setCells[k];
        vBound += size_t(sets[k].nr) * 3 * sk * (sk + 1) * (sk + 1);
    }
    const size_t vCap = size_t(mm.nCross) + vBound, tCap = size_t(mm.nTri) + 5 * cellsAll;
    const bool deferred =
        (nSets <= 1 || cellsAll <= chunkCells) && vCap <= size_t(INT32_MAX) && tCap <= size_t(INT32_MAX) / 3;
    if (!deferred && current_recorder()) throw NotRecordable{};  // it waits between working sets
    constexpr int kMailKept = 3;
    c.hostPinned[1] = c.hostPinned[2] = c.hostPinned[kMailKept] = 0;
    // vertices kept by the depth-D batch: a device int read by the subdivision
    // kernels, or the mailbox slot when there is no subdivision
    int* keptMain = nSets > 0 ? c.buf<int>("sub_kept_main", 1) : c.mail + kMailKept;
    size_t triTotal = mm.nTri, keptSub = 0, keptMainHost = 0;
    float* outV = c.buf_keep<float>("mesh_vertices", 3 * (deferred ? vCap : size_t(std::max(mm.nCross, 1))), 0);
    unsigned* outKey = c.buf_keep<unsigned>("mesh_vertex_keys", deferred ? vCap : size_t(std::max(mm.nCross, 1)), 0);
    // the depth-D batch's kept vertices: compacted by the first set-up launch
    const int keepBlocks = std::max(1, int(div_up(mm.nCross, kThreads)));
    const KeepArgs ka{mainV, mainKey, mm.nCross,
                      scan_ticket(c.scan_for_tiles(size_t(keepBlocks)), keepBlocks, keptMain), outV, outKey};
    if (nSets == 0) PSR_LAUNCH(sub_setup, keepBlocks, 0, st, ka, keepBlocks, RootRecArgs{}, 0, TileArgs{});
    if (deferred && nSets > 0) outT = c.buf_keep<int>("mesh_triangles", 3 * tCap, 3 * size_t(mm.nTri));
    if (nSets > 0) {
        const int* rootPos = mm.rootPos;
        float* isoDev = &c.results()->iso;
        const float4* nbrCol = c.fld.nbrCol;
        int* rootNbr = c.buf<int>("sub_root_nbr", size_t(27) * mm.nRoots);
        int* rootChain = c.buf<int>("sub_root_chain", size_t(kChain) * mm.nRoots);
        int4* rootCell = c.buf<int4>("sub_root_cell", mm.nRoots);
        unsigned* rootSet = c.buf<unsigned>("sub_root_set", mm.nRoots);
        const RootRecArgs ra{mm.nRoots, M,        L,       finerDepth, mm.roots, rootPos,  o.neigh,
                             o.parent,  o.key,    rootNbr, rootChain,  rootCell, rootSet};
        const int rootBlocks = int(div_up(size_t(27) * mm.nRoots, kThreads));
        // tiles read d_x from dense per-depth grids when the octree is geometric
        const bool denseHalo = !mm.q1Fired;
        for (const SetDesc& S : sets) {
            int halo = 0;
            if (denseHalo && D - S.rd >= 3 &&
                (grid_tile_words(D, S.rd, &halo) > kLdsWordsGrid || halo > kHaloPerThread * kThreads))
                throw std::runtime_error("subdivision tile exceeds its LDS budget");
            if (denseHalo && D - S.rd < 3 && small_tile_words(D, S.rd, D - S.rd == 1 ? 4 : 1) > kLdsWordsSmallGrid)
                throw std::runtime_error("subdivision tile exceeds its LDS budget");
        }
        float* dxAt = denseHalo ? c.geo.dxGrid : nullptr;  // built before the wait up to kDenseEarlyDepth
        if (denseHalo && !dxAt) {
            const size_t cellsAll = dense_grid_cells(D);
            dxAt = c.buf<float>("sub_dx_grid", cellsAll);
            PSR_LAUNCH(zero_floats, div_up(cellsAll, kThreads), 0, st, dxAt, cellsAll);
            PSR_LAUNCH(scatter_dx_grid, div_up(M, kThreads), 0, st, M, L, o.key, c.fld.dx, dxAt);
        }
        std::vector<SetDesc> cs;
        std::vector<int> firstSmall;
        for (int s0 = 0; s0 < nSets;) {
            // ---- one chunk of sets [s0, s1), cell enumeration local to the chunk
            size_t cells = 0;
            int s1 = s0;
            while (s1 < nSets && (s1 == s0 || cells + setCells[s1] <= chunkCells)) cells += setCells[s1++];
            const int nCs = s1 - s0, nVD = int(cells);
            cs.assign(sets.begin() + s0, sets.begin() + s1);
            for (int k = 0, de = 0; k < nCs; ++k) {
                cs[k].dEnum = de;
                de += int(setCells[s0 + k]);
            }
            // tiles: those lying in one root first, then those of small roots
            // (a tile per 4 roots of 8 cells or per root of 64 cells, for their
            // corner values); sets of one-root tiles come first, so the tiles are
            // in enumeration order.  Counted here, written by sub_setup.
            int nTiles = 0, nOne = 0, nSmall = 0, perSet = 1;
            firstSmall.assign(nCs, 0);
            for (int pass = 0; pass < 2; ++pass)
                for (int k = 0; k < nCs; ++k) {
                    const int shk = D - cs[k].rd;
                    if ((shk >= 3) != (pass == 0)) continue;
                    const int n = int(div_up(setCells[s0 + k], size_t(kTile)));
                    cs[k].firstTile = nTiles;
                    nTiles += n;
                    perSet = std::max(perSet, n);
                    if (pass == 0) {
                        nOne += n;
                    } else {
                        const int ns = int(div_up(setCells[s0 + k], size_t(shk == 1 ? 32 : 64)));
                        firstSmall[k] = nSmall;
                        nSmall += ns;
                        perSet = std::max(perSet, ns);
                    }
                }
            for (int k = 1; k < nCs; ++k)
                if (cs[k].firstTile < cs[k - 1].firstTile) throw std::runtime_error("subdivision tiles out of order");
            // sets and first small tiles go up through pinned staging (reused
            // only after a synchronisation)
            const size_t bSets = (sizeof(SetDesc) * nCs + 255) & ~size_t(255);
            const size_t bTiles = (sizeof(Tile) * nTiles + 255) & ~size_t(255);
            char* hs = static_cast<char*>(c.stage(bSets + sizeof(int) * nCs));
            std::memcpy(hs, cs.data(), sizeof(SetDesc) * nCs);
            std::memcpy(hs + bSets, firstSmall.data(), sizeof(int) * nCs);
            char* dAll = c.buf<char>("sub_upload", bSets + bTiles + sizeof(Tile) * std::max(nSmall, 1));
            SetDesc* dsets = reinterpret_cast<SetDesc*>(dAll);
            {
                // the first chunk's launch also compacts the depth-D batch and
                // writes the root records
                const char* hsDev = static_cast<const char*>(c.stageDev);
                const int parts = std::min(16, int(div_up(perSet, kThreads)));
                const TileArgs ta{reinterpret_cast<const SetDesc*>(hsDev), reinterpret_cast<const int*>(hsDev + bSets),
                                  D, parts, dsets, reinterpret_cast<Tile*>(dAll + bSets),
                                  reinterpret_cast<Tile*>(dAll + bSets + bTiles)};
                const int kb = s0 == 0 ? keepBlocks : 0, rb = s0 == 0 ? rootBlocks : 0;
                PSR_LAUNCH(sub_setup, unsigned(kb + rb + nCs * parts), 0, st, ka, kb, ra, rb, ta);
            }
            Tile* dtiles = reinterpret_cast<Tile*>(dAll + bSets);
            Tile* dsmall = reinterpret_cast<Tile*>(dAll + bSets + bTiles);
            Grid gr{D,       M,       nCs,      dsets,    mm.roots, rootPos,   o.key,    o.parent, o.childBase,
                    o.neigh, o.basis, c.fld.dx, nbrCol,   rootNbr,  rootChain, rootCell, rootSet,   dxAt};
            float* cval = c.buf<float>("sub_cval", size_t(8) * nVD);
            float* vals = c.buf<float>("sub_vals", size_t(8) * nVD);
            unsigned* cellWord = c.buf<unsigned>("sub_cell_word", nVD);
            int* crossPre = c.buf<int>("sub_cross_prefix", nVD);
            int4* tileCnt = c.buf<int4>("sub_tile_cnt", nTiles);
            int4* tileBase = c.buf<int4>("sub_tile_base", nTiles);
            // the one-root tiles' point owners, found by sub_corner_values_grid
            // and reused by sub_classify (value indices slot * nVD + cell fit 32 bits)
            unsigned* ptOwner = nullptr;
            int nMapped = 0;
            if (nOne > 0 && denseHalo) {
                int words = 0;
                for (int k = 0; k < nCs; ++k)
                    if (D - cs[k].rd >= 3) words = std::max(words, grid_tile_words(D, cs[k].rd));
                if (size_t(8) * size_t(nVD) <= size_t(UINT32_MAX)) {
                    ptOwner = c.buf<unsigned>("sub_point_owner", size_t(nOne) * kTilePoints);
                    nMapped = nOne;
                }
                PSR_LAUNCH(sub_corner_values_grid, nOne, sizeof(float) * words, st, gr, dtiles, nVD, isoDev, cval,
                           ptOwner);
            }
            else if (nOne > 0)
                PSR_LAUNCH_K(sub_corner_values_one, sub_corner_values_one_k, nOne, 0, st, gr, dtiles, nVD, isoDev, cval);
            if (nSmall > 0 && denseHalo) {
                int words = 0;
                for (int k = 0; k < nCs; ++k)
                    if (D - cs[k].rd < 3)
                        words = std::max(words, small_tile_words(D, cs[k].rd, D - cs[k].rd == 1 ? 4 : 1));
                PSR_LAUNCH(sub_corner_values_small_grid, nSmall, sizeof(float) * words, st, gr, dsmall, nVD, isoDev, cval);
            } else if (nSmall > 0) {
                PSR_LAUNCH_K(sub_corner_values_small, sub_corner_values_small_k, nSmall, 0, st, gr, dsmall, nVD, isoDev, cval);
            }
            PSR_LAUNCH(sub_classify, nTiles, 0, st, gr, dtiles, nVD, cval, vals, cellWord, crossPre, tileCnt,
                       static_cast<const unsigned*>(ptOwner), nMapped);
            // totals: slots 0-3
            PSR_LAUNCH(sub_tile_scan, 1, 0, st, tileCnt, nTiles, tileBase, keptMain, c.mail);
            if (deferred) {
                PSR_LAUNCH(sub_emit, nTiles, 0, st, gr, dtiles, nVD, vals, cellWord, crossPre, tileCnt, tileBase,
                                                      keptMain, 0, mm.nTri, outV, outT, outKey);
                s0 = s1;
                continue;
            }
            c.wait();
            const int nKept = c.hostPinned[1], nTri = c.hostPinned[2];
            if (s0 == 0) keptMainHost = size_t(c.hostPinned[kMailKept]);
            const size_t keptTotal = keptMainHost + keptSub;
            if (keptTotal + nKept > size_t(INT32_MAX) || triTotal + nTri > size_t(INT32_MAX) / 3)
                throw std::runtime_error("mesh exceeds 32-bit indexing");
            outV = c.buf_keep<float>("mesh_vertices", 3 * (keptTotal + nKept), 3 * keptTotal);
            outKey = c.buf_keep<unsigned>("mesh_vertex_keys", keptTotal + nKept, keptTotal);
            outT = c.buf_keep<int>("mesh_triangles", 3 * (triTotal + nTri), 3 * triTotal);
            if (nKept > 0 || nTri > 0)
                PSR_LAUNCH(sub_emit, nTiles, 0, st, gr, dtiles, nVD, vals, cellWord, crossPre, tileCnt, tileBase,
                                                      keptMain, int(keptSub), int(triTotal), outV, outT, outKey);
            keptSub += nKept;
            triTotal += nTri;
            s0 = s1;
        }
    }
    c.mesh.vertices = outV;
    c.mesh.triangles = outT;
    c.mesh.keys = outKey;
    PSR_HIP_CHECK(hipGetLastError());
    if (deferred) {
        c.mesh.deferred = true;
        c.mesh.triBefore = mm.nTri;
        c.mesh.vCap = vCap;
        c.mesh.tCap = nSets > 0 ? tCap : size_t(mm.nTri);
        return;
    }
    c.mesh.nv = int(keptMainHost + keptSub);
    c.mesh.nt = int(triTotal);
    c.stats.n_mesh_vertices = c.mesh.nv;
    c.stats.n_mesh_triangles = c.mesh.nt;
}

void extract_mesh(Ctx& c) {
    extract_mesh_a(c);
    c.wait();
    extract_mesh_b(c);
}

void finish_mesh(Ctx& c) {
    Mesh& m = c.mesh;
    if (!m.deferred) return;
    m.deferred = false;
    const size_t kept = size_t(c.hostPinned[3]) + size_t(c.hostPinned[1]);
    const size_t tri = size_t(m.triBefore) + size_t(c.hostPinned[2]);
    if (kept > m.vCap || tri > m.tCap) throw std::runtime_error("mesh counts exceed their bounds");
    m.nv = int(kept);
    m.nt = int(tri);
    c.stats.n_mesh_vertices = m.nv;
    c.stats.n_mesh_triangles = m.nt;
}

}  // namespace psr
