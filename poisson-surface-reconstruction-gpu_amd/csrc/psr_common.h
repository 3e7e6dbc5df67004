// Shared device helpers for the MI355X (gfx950) Poisson reconstruction kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <stdexcept>
#include <string>

#define PSR_MAX_DEPTH 10
#define PSR_WAVE 64

#define PSR_HIP_CHECK(expr)                                                                        \
    do {                                                                                           \
        hipError_t e__ = (expr);                                                                   \
        if (e__ != hipSuccess)                                                                     \
            throw psr::HipError(std::string(#expr) + ": " + hipGetErrorString(e__), e__);          \
    } while (0)

namespace psr {

struct HipError : std::runtime_error {
    hipError_t code;
    HipError(const std::string& m, hipError_t c) : std::runtime_error(m), code(c) {}
};

inline unsigned div_up(size_t a, unsigned b) { return unsigned((a + b - 1) / b); }

// Per-depth layout of the node array: level d occupies [base[d], base[d] + count[d]).
struct Levels {
    int D;
    int base[PSR_MAX_DEPTH + 2];
    int count[PSR_MAX_DEPTH + 1];
};

// The depth search of the reference (BuildOctree.cu:398-402): the last level is
// reported as D because the loop runs d < D.
__device__ __forceinline__ int depth_of(const Levels& L, int idx) {
    int d = 0;
    for (; d < L.D; ++d)
        if (L.base[d] <= idx && idx < L.base[d + 1]) break;
    return d;
}

// Node centre from its level-aligned key (BuildOctree.cu:568-586); exact dyadic.
__device__ __forceinline__ float3 key_center(int key, int depth, int D) {
    float3 c = make_float3(0.5f, 0.5f, 0.5f);
    float w = 0.25f;
    for (int i = D - 1; i >= D - depth; --i) {
        c.x += ((key >> (3 * i + 2)) & 1) ? w : -w;
        c.y += ((key >> (3 * i + 1)) & 1) ? w : -w;
        c.z += ((key >> (3 * i)) & 1) ? w : -w;
        w *= 0.5f;
    }
    return c;
}

// Per-axis 1-D basis index (BinaryNode order: (2^d - 1) + cell coordinate).
__device__ __forceinline__ int3 key_func_index(int key, int depth, int D) {
    int base = (1 << depth) - 1;
    int3 f = make_int3(base, base, base);
    for (int l = depth; l >= 1; --l) {
        int sh = 3 * (D - l);
        int wgt = 1 << (depth - l);
        f.x += ((key >> (sh + 2)) & 1) * wgt;
        f.y += ((key >> (sh + 1)) & 1) * wgt;
        f.z += ((key >> sh) & 1) * wgt;
    }
    return f;
}

__device__ __forceinline__ int3 decode_func_index(int e, int D) {
    int s1 = 1 << (D + 1);
    return make_int3(e & (s1 - 1), (e >> (D + 1)) & (s1 - 1), e >> (2 * (D + 1)));
}

// Centre and inverse width of a 1-D basis function index.
__device__ __forceinline__ void func_center_invw(int idx, float& c, float& invw) {
    int depth = 31 - __clz(idx + 1);
    int off = idx + 1 - (1 << depth);
    invw = float(1 << depth);
    c = (0.5f + float(off)) / invw;
}

// Quadratic B-spline (box * box * box) normalised to F(0) = 1, evaluated in the
// local coordinate u = (x - c) / w.  Same function as the reference's
// PPolynomial<2>::GaussianApproximation (math/PPolynomial.h:568-582) divided by
// F(0) (mesh/ComputeVectorField.cpp:81-95).
__device__ __forceinline__ float bspline(float u) {
    u = fabsf(u);
    float t = 1.5f - u;
    float outer = (2.0f / 3.0f) * t * t;
    float inner = 1.0f - (4.0f / 3.0f) * u * u;
    return u < 0.5f ? inner : (u < 1.5f ? outer : 0.0f);
}

__device__ __forceinline__ float basis_value(int fidx, float x) {
    float c, invw;
    func_center_invw(fidx, c, invw);
    return bspline((x - c) * invw);
}

// Basis parameters of a node: per-axis centres of its 1-D functions and the
// shared inverse width (func_center_invw of each axis index), so that
// node_term(...) equals dx * basis_value(fx, x) * basis_value(fy, y) * basis_value(fz, z).
__device__ __forceinline__ float4 node_basis_params(int3 f) {
    float4 b;
    func_center_invw(f.x, b.x, b.w);
    func_center_invw(f.y, b.y, b.w);
    func_center_invw(f.z, b.z, b.w);
    return b;
}
__device__ __forceinline__ float node_term(float4 b, float dx, float3 p) {
    return dx * bspline((p.x - b.x) * b.w) * bspline((p.y - b.y) * b.w) * bspline((p.z - b.z) * b.w);
}

// sum over the 27 neighbours of `node` of dx * F(x) F(y) F(z)
__device__ __forceinline__ float chi_node_sum(int node, float3 p, int M, const int* __restrict__ neigh,
                                              const float4* __restrict__ basis, const float* __restrict__ dx) {
    float v = 0.f;
    for (int i = 0; i < 27; ++i) {
        int nb = neigh[i * M + node];
        if (nb == -1) continue;
        v += node_term(basis[nb], dx[nb], p);
    }
    return v;
}

// 3x3x3 neighbourhood look-up tables (derived; BuildOctree.cu:20-42 lists the
// same values): neighbour i of a child with son key s is child LUTchild of the
// parent's neighbour LUTparent.
struct NeighborLut {
    unsigned char parent[8][27];
    unsigned char child[8][27];
};

__host__ __device__ inline void make_neighbor_lut(NeighborLut& t) {
    for (int s = 0; s < 8; ++s) {
        int sb[3] = {(s >> 2) & 1, (s >> 1) & 1, s & 1};
        for (int i = 0; i < 27; ++i) {
            int dd[3] = {i / 9 - 1, (i / 3) % 3 - 1, i % 3 - 1};
            int pi = 0, ci = 0;
            for (int a = 0; a < 3; ++a) {
                int c = sb[a] + dd[a];
                int p = (c < 0) ? -1 : (c > 1 ? 1 : 0);
                pi = pi * 3 + (p + 1);
                ci = ci * 2 + (c - 2 * p);
            }
            t.parent[s][i] = (unsigned char)pi;
            t.child[s][i] = (unsigned char)ci;
        }
    }
}

}  // namespace psr
