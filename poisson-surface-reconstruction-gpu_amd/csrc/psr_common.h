// Shared device helpers for the MI355X (gfx950) Poisson reconstruction kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "psr_launch.h"

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

// --- wave reductions through DPP ---------------------------------------------
// Four DPP butterfly steps (swap neighbours, swap pairs, mirror within 8,
// mirror within 16) leave every lane of a row of 16 with its row's sum; the
// four row sums (lanes 0, 16, 32, 48) are then read into scalar registers and
// added in a fixed order, so every lane gets the same bits.  DPP moves run in
// the vector ALU; a shuffle (ds_bpermute) takes a round trip through the LDS
// unit, twice for a double.  The whole wave must be active.
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, int(b), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, int(b >> 32), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double readlane_f64(double v, int lane) {
    const long long b = __double_as_longlong(v);
    return __hiloint2double(__builtin_amdgcn_readlane(int(b >> 32), lane), __builtin_amdgcn_readlane(int(b), lane));
}
__device__ __forceinline__ float readlane_f32(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
constexpr int kDppSwap1 = 0xB1, kDppSwap2 = 0x4E, kDppHalfMirror = 0x141, kDppMirror = 0x140;
// Sum over the wave, the same in every lane.
// Workgroup b of an n-workgroup launch runs on XCD b mod 8 (the dispatcher
// deals workgroups out round-robin).  The single-cloud vector-field splat uses
// block xcd_block(b, n) instead of b, so that each XCD takes one contiguous
// range of blocks: nodes close in Morton order, whose 27-neighbourhoods share
// points, then share that XCD's L2 (8% faster on the torus).  The divergence,
// ownership, vertex-value and neighbour-column launches measured slower with
// it (kernels whose tiles are ordered by an arrival counter, or whose work is
// uneven along the Morton order) and keep b.  A bijection on [0, n).
__device__ __forceinline__ int xcd_block(int b, int n) {
    constexpr int kXcds = 8;
    const int x = b % kXcds, idx = b / kXcds, q = n / kXcds, r = n % kXcds;
    return x * q + min(x, r) + idx;
}

__device__ __forceinline__ double wave_sum_f64(double v) {
    v += dpp_f64<kDppSwap1>(v);
    v += dpp_f64<kDppSwap2>(v);
    v += dpp_f64<kDppHalfMirror>(v);
    v += dpp_f64<kDppMirror>(v);
    return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}
// Sum over the half-wave h (lanes 32 h .. 32 h + 31), the same in every lane
// of that half.
__device__ __forceinline__ float half_sum_f32(float v, int h) {
    v += dpp_f32<kDppSwap1>(v);
    v += dpp_f32<kDppSwap2>(v);
    v += dpp_f32<kDppHalfMirror>(v);
    v += dpp_f32<kDppMirror>(v);
    const float lo = readlane_f32(v, 0) + readlane_f32(v, 16), hi = readlane_f32(v, 32) + readlane_f32(v, 48);
    return h ? hi : lo;
}
__device__ __forceinline__ float wave_sum_f32(float v) {
    v += dpp_f32<kDppSwap1>(v);
    v += dpp_f32<kDppSwap2>(v);
    v += dpp_f32<kDppHalfMirror>(v);
    v += dpp_f32<kDppMirror>(v);
    return (readlane_f32(v, 0) + readlane_f32(v, 16)) + (readlane_f32(v, 32) + readlane_f32(v, 48));
}

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

// node_term of a neighbour-column entry {centre x, y, z, d_x}; invw is the
// inverse width of the node's depth, which all its neighbours share.
__device__ __forceinline__ float col_term(float4 c, float invw, float3 p) {
    return c.w * bspline((p.x - c.x) * invw) * bspline((p.y - c.y) * invw) * bspline((p.z - c.z) * invw);
}

// sum over the 27 neighbours of `node` (at the depth of inverse width invw) of
// dx * F(x) F(y) F(z), from the neighbour columns ([27][M] {centre, d_x} of
// neighbour i, zero where absent).  An absent neighbour adds +0 or -0, which
// leaves the sum unchanged: it starts at +0 and can never become -0 by adding.
__device__ __forceinline__ float chi_node_sum(int node, float invw, float3 p, int M,
                                              const float4* __restrict__ nbrCol) {
    float4 c[27];
#pragma unroll
    for (int i = 0; i < 27; ++i) c[i] = nbrCol[size_t(i) * M + node];
    float v = 0.f;
#pragma unroll
    for (int i = 0; i < 27; ++i) v += col_term(c[i], invw, p);
    return v;
}

// chi_node_sum when the node's neighbours lie on the regular grid around it
// (an octree in which the level-local child 0 rule changed no parent).  Node
// basis b = {centre x, y, z, inverse width}; neighbour i's centre is the
// node's plus (i / 9 - 1, (i / 3) % 3 - 1, i % 3 - 1) widths, exactly (the
// centres are dyadic), so the 27 terms take 3 B-spline values per axis.  Each
// term is formed as col_term forms it, so the sum equals chi_node_sum's.
// The d_x come from the dense [27][M] copy of the columns' d_x (a wave's 27
// loads then touch a quarter of the cache lines).
__device__ __forceinline__ float chi_node_sum_regular(int node, float4 b, float3 p, int M,
                                                      const float* __restrict__ dCol) {
    float d[27];
#pragma unroll
    for (int i = 0; i < 27; ++i) d[i] = dCol[size_t(i) * M + node];
    const float w = 1.0f / b.w;  // exact: b.w is a power of two
    float fx[3], fy[3], fz[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        fx[j] = bspline((p.x - (b.x + float(j - 1) * w)) * b.w);
        fy[j] = bspline((p.y - (b.y + float(j - 1) * w)) * b.w);
        fz[j] = bspline((p.z - (b.z + float(j - 1) * w)) * b.w);
    }
    float v = 0.f;
#pragma unroll
    for (int i = 0; i < 27; ++i) v += d[i] * fx[i / 9] * fy[(i / 3) % 3] * fz[i % 3];
    return v;
}

// Implicit function at the sorted points (LaplacianSolver.cu:269-305): one
// thread per point sums its leaf's and every ancestor's 27 neighbour terms;
// each block of kPointsChiThreads points leaves the FP64 sum of its values in
// part[block] for the iso value.
constexpr int kPointsChiThreads = 256;
struct PointsChiArgs {
    int N, offD, M, D;
    const float4* pos;
    const int* p2n;
    const int* parent;
    const float4* nbrCol;
    const float* dCol;
    const float4* basis;
    const int* q1Fired;
    float* chi;
    double* part;
};
__device__ __forceinline__ void points_chi_block(const PointsChiArgs& a, int block) {
    __shared__ double sh[kPointsChiThreads / PSR_WAVE];
    const int i = block * kPointsChiThreads + int(threadIdx.x);
    double mine = 0;
    const bool regular = *a.q1Fired == 0;
    if (i < a.N) {
        const float4 q = a.pos[i];
        const float3 p = make_float3(q.x, q.y, q.z);
        int now = a.offD + a.p2n[i];
        float v = 0.f;
        for (int d = a.D; now != -1; --d) {  // a parent is one level up
            v += regular ? chi_node_sum_regular(now, a.basis[now], p, a.M, a.dCol)
                         : chi_node_sum(now, float(1 << d), p, a.M, a.nbrCol);
            now = a.parent[now];
        }
        a.chi[i] = v;
        mine = double(v);
    }
    mine = wave_sum_f64(mine);
    if (__lane_id() == 0) sh[threadIdx.x / PSR_WAVE] = mine;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int w = 0; w < kPointsChiThreads / PSR_WAVE; ++w) s += sh[w];
        a.part[block] = s;
    }
}

// --- operator (LaplacianSolver.cu:27-74) -----------------------------------
// L(o, o') = <F,F>_x <F,F>_y <F,F>_z (<F',F'>_x + <F',F'>_y + <F',F'>_z) for the
// same-depth neighbours o' of o; |L| <= 1e-6 is dropped (stored as 0); values
// kept as float.  ELL layout [27][M] over all nodes; the column of slot i is
// neigh[i][o] - base[depth].
// Also counts the stored (nonzero) entries per depth into nnz (zeroed before).
// Blocks of kAssembleThreads threads, launched beside the divergence's last
// kernel (they need only the octree and the tables).
constexpr int kAssembleThreads = 256;
struct AssembleArgs {
    Levels L;
    int M;
    const int* neigh;
    const int* enc;
    const double* dot;
    const double* d2dot;
    int res;
    float* val;
    int* nnz;        // per depth; the slot masks follow the CG error flag after it
    unsigned* slots;
};
__device__ __forceinline__ void assemble_rows(const AssembleArgs& a, int block) {
    const Levels& L = a.L;
    const int M = a.M, res = a.res;
    const int* __restrict__ neigh = a.neigh;
    const int* __restrict__ enc = a.enc;
    const double* __restrict__ dot = a.dot;
    const double* __restrict__ d2dot = a.d2dot;
    float* __restrict__ val = a.val;
    // three threads per row, each the 9 neighbours of one x offset (thread
    // t = j * M + o: neighbours 9 j .. 9 j + 8 of row o), so that three times as
    // many waves hide the loads; a thread's table rows stay in its cache
    __shared__ int cnt[PSR_MAX_DEPTH + 1];
    __shared__ unsigned msk[PSR_MAX_DEPTH + 1];
    if (threadIdx.x <= PSR_MAX_DEPTH) cnt[threadIdx.x] = 0, msk[threadIdx.x] = 0u;
    __syncthreads();
    const int t = block * kAssembleThreads + int(threadIdx.x);
    const bool valid = t < 3 * M;
    int d = -1, c = 0;
    unsigned bits = 0u;
    if (valid) {
        const int j = t / M, o = t - j * M;
        const int3 f1 = decode_func_index(enc[o], L.D);
#pragma unroll 3
        for (int i = 9 * j; i < 9 * j + 9; ++i) {
            const int nb = neigh[i * M + o];
            float v = 0.f;
            if (nb != -1) {
                const int3 f2 = decode_func_index(enc[nb], L.D);
                size_t sx = size_t(f1.x) * res + f2.x, sy = size_t(f1.y) * res + f2.y, sz = size_t(f1.z) * res + f2.z;
                double e = dot[sx] * dot[sy] * dot[sz] * (d2dot[sx] + d2dot[sy] + d2dot[sz]);
                if (fabs(e) > double(1e-6f)) v = float(e);
            }
            val[i * M + o] = v;
            c += v != 0.f;
            bits |= v != 0.f ? 1u << i : 0u;
        }
        d = depth_of(L, o);
    }
    // stored entries per depth: one LDS add per depth present in the wave
    unsigned long long active = __ballot(valid);
    while (active) {
        const int leader = __builtin_ctzll(active);
        const int dl = __shfl(d, leader, PSR_WAVE);
        const unsigned long long same = __ballot(d == dl) & active;
        int s = d == dl ? c : 0;
        unsigned b = d == dl ? bits : 0u;
#pragma unroll
        for (int off = PSR_WAVE / 2; off > 0; off >>= 1) {
            s += __shfl_xor(s, off, PSR_WAVE);
            b |= unsigned(__shfl_xor(int(b), off, PSR_WAVE));
        }
        if (__lane_id() == leader) {
            atomicAdd(&cnt[dl], s);
            atomicOr(&msk[dl], b);
        }
        active &= ~same;
    }
    __syncthreads();
    if (threadIdx.x <= PSR_MAX_DEPTH && cnt[threadIdx.x]) atomicAdd(&a.nnz[threadIdx.x], cnt[threadIdx.x]);
    if (threadIdx.x <= PSR_MAX_DEPTH && msk[threadIdx.x]) atomicOr(&a.slots[threadIdx.x], msk[threadIdx.x]);
}

}  // namespace psr
