// Geometry helpers and tables shared by the mesh and subdivision kernels.
#pragma once
#include "psr_internal.h"

namespace psr {

// Bourke corner order of OctNode.vertices (BuildMeshGeometry.cu:133-144).
static __constant__ int kEdgeVertex[12][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 3}, {1, 2},
                                              {4, 7}, {5, 6}, {0, 4}, {1, 5}, {3, 7}, {2, 6}};
static __constant__ int kFaceEdges[6][4] = {{4, 6, 8, 10}, {5, 7, 9, 11}, {0, 2, 8, 9},
                                            {1, 3, 10, 11}, {0, 1, 4, 5}, {2, 3, 6, 7}};
// Child picked by the vertex-value descent for a Bourke corner id
// (ComputeTriangleIndices.cu:31; applied to key-ordered children, quirk Q8).
static __constant__ int kChildrenVertexKind[8] = {0, 1, 3, 2, 4, 5, 7, 6};
// BuildMeshGeometry.cu:19-26, indexed by the key-ordered son key as the
// reference does.
static __constant__ int kParentFaceKind[8][6] = {{0, -1, 2, -1, 4, -1}, {-1, 1, 2, -1, 4, -1},
                                                 {0, -1, -1, 3, 4, -1}, {-1, 1, -1, 3, 4, -1},
                                                 {0, -1, 2, -1, -1, 5}, {-1, 1, 2, -1, -1, 5},
                                                 {0, -1, -1, 3, -1, 5}, {0, 1, -1, 3, -1, 5}};
static __constant__ uint64_t kMcWords[256] = {
#include "psr_mc_table_body.inc"
};

__device__ __forceinline__ uint64_t mc_word(int cube) { return kMcWords[cube]; }
__device__ __forceinline__ int mc_tri_count(int cube) {
    uint64_t w = kMcWords[cube];
    int n = 0;
    while (n < 15 && ((w >> (4 * n)) & 0xF) != 0xF) ++n;
    return n / 3;
}

__device__ __forceinline__ double sqdist(float3 a, float3 b) {
    float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    return double(dx * dx + dy * dy + dz * dz);
}

// Corner i of the node centred at c: bit0 -> x, bit1 -> y, bit2 -> z.
__device__ __forceinline__ float3 corner_pos(float3 c, int i, float half) {
    return make_float3(c.x + float(2 * (i & 1) - 1) * half, c.y + float(2 * ((i >> 1) & 1) - 1) * half,
                       c.z + float(2 * ((i >> 2) & 1) - 1) * half);
}

// Centre of edge `kind` = 4 * axis + 2-bit offset of the other two axes.
__device__ __forceinline__ float3 edge_center(float3 c, int kind, float half) {
    int orient = kind >> 2, o0 = kind & 1, o1 = (kind >> 1) & 1;
    float m[3];
    int dim = 0;
    for (int a = 0; a < 3; ++a) {
        if (a == orient) {
            m[a] = 0.f;
        } else {
            m[a] = float(2 * (dim == 0 ? o0 : o1) - 1);
            ++dim;
        }
    }
    return make_float3(c.x + m[0] * half, c.y + m[1] * half, c.z + m[2] * half);
}

__device__ __forceinline__ float3 face_center(float3 c, int kind, float half) {
    float d = float(2 * (kind & 1) - 1) * half;
    int o = kind >> 1;
    return make_float3(c.x + (o == 0 ? d : 0.f), c.y + (o == 1 ? d : 0.f), c.z + (o == 2 ? d : 0.f));
}

// Bourke slot of the vertex at p within the node centred at c
// (maintainVertexNodePointer, BuildMeshGeometry.cu:133-144).
__device__ __forceinline__ int corner_slot(float3 c, float3 p) {
    int index = 0;
    if (c.x - p.x < 0) index |= 1;
    if (c.z - p.z < 0) index |= 4;
    if (c.y - p.y < 0) index += (index & 1) ? 1 : 3;
    return index;
}

// Edge slot within the node centred at c (maintainEdgeNodePointer, :264-277).
__device__ __forceinline__ int edge_slot(float3 c, float3 p, int kind) {
    int orient = kind >> 2, index = orient << 2, dim = 0;
    float cc[3] = {c.x, c.y, c.z}, pp[3] = {p.x, p.y, p.z};
    for (int a = 0; a < 3; ++a)
        if (a != orient) {
            if (cc[a] - pp[a] < 0) index |= 1 << dim;
            ++dim;
        }
    return index;
}

// Linear crossing along axis dim (ComputeTriangleIndices.cu:464-474).
__device__ __forceinline__ float3 interpolate_point(float3 p1, float3 p2, int dim, float v1, float v2) {
    float pivot = v1 / (v1 - v2), another = 1.f - pivot;
    float a[3] = {p1.x, p1.y, p1.z}, b[3] = {p2.x, p2.y, p2.z};
    a[dim] = b[dim] * pivot + a[dim] * another;
    return make_float3(a[0], a[1], a[2]);
}

// Per-call state of the depth-D marching cubes.


// Morton code of a cell at a depth with `bits` levels (x high in each triplet):
// the node key format of the reference.
__device__ __forceinline__ unsigned spread3(unsigned v) {  // bit i -> bit 3i, i < 10
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__device__ __forceinline__ unsigned compact3(unsigned v) {  // bit 3i -> bit i, i < 10
    v &= 0x09249249u;
    v = (v | (v >> 2)) & 0x030C30C3u;
    v = (v | (v >> 4)) & 0x0300F00Fu;
    v = (v | (v >> 8)) & 0x030000FFu;
    v = (v | (v >> 16)) & 0x3FFu;
    return v;
}
// The low `bits` bits of each coordinate, interleaved.
__device__ __forceinline__ int morton(int3 c, int bits) {
    const unsigned m = (1u << bits) - 1u;
    return int((spread3(unsigned(c.x) & m) << 2) | (spread3(unsigned(c.y) & m) << 1) | spread3(unsigned(c.z) & m));
}
__device__ __forceinline__ int3 demorton(int k, int bits) {
    const unsigned u = unsigned(k) & ((1u << (3 * bits)) - 1u);
    return make_int3(int(compact3(u >> 2)), int(compact3(u >> 1)), int(compact3(u)));
}

// Position of node (key, depth l) in the dense per-depth grids of d_x: the
// depth-l grid starts at (8^l - 1) / 7 and holds cell (x, y, z) at x 4^l + y 2^l + z.
__device__ __forceinline__ size_t dense_grid_index(int key, int l, int D) {
    const int3 c = demorton(key >> (3 * (D - l)), l);
    return ((size_t(1) << (3 * l)) - 1) / 7 + ((size_t(c.x) << (2 * l)) | (size_t(c.y) << l) | size_t(c.z));
}

// Global key of a depth-D grid edge, carried by every mesh vertex for
// welding: lower endpoint (x, y, z) in [0, 2^D]^3 and axis a (0 x, 1 y, 2 z)
// give ((x (2^D + 1) + y) (2^D + 1) + z) 3 + a, below 3 (2^D + 1)^3 < 2^32
// for D <= 10.
__device__ __forceinline__ unsigned segment_key(int3 p1, int3 p2, int D) {
    const unsigned S = (1u << D) + 1u;
    const int axis = p1.x != p2.x ? 0 : (p1.y != p2.y ? 1 : 2);
    return ((unsigned(min(p1.x, p2.x)) * S + unsigned(min(p1.y, p2.y))) * S + unsigned(min(p1.z, p2.z))) * 3u +
           unsigned(axis);
}

// Depth-D marching cubes into the raw mesh buffers (vertices at [0, nCross),
// triangles at [0, nTri)), then the subdivision roots.  Ends with one wait
// for the counts.
void depth_mc_launch(Ctx& c, MainMc& mm, float* rawV, unsigned* rawKey, int* rawT);
void depth_mc_read(Ctx& c, MainMc& mm);

}  // namespace psr
