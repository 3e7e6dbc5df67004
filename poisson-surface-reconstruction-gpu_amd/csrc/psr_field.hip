// Basis inner-product tables, vector-field splat and divergence.
#include "psr_internal.h"

#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>

namespace psr {

namespace {

// --- 1-D tables (FunctionData::setDotTables, mesh/FunctionData.h:134-190) ---
// Exact FP64 integrals over the real line of products of the scaled/shifted
// quadratic B-spline and its derivative, one thread per (a, b) entry:
//   dot[a + res b]   = <F_a,  F_b>
//   ddot[a + res b]  = <F_a', F_b>
//   d2dot[a + res b] = <F_a', F_b'>
// Each product is a polynomial of degree <= 4 between breakpoints; 3-point
// Gauss-Legendre per piece integrates it exactly.
__device__ double bsp_d(double u) {
    u = fabs(u);
    if (u < 0.5) return 1.0 - (4.0 / 3.0) * u * u;
    if (u < 1.5) return (2.0 / 3.0) * (1.5 - u) * (1.5 - u);
    return 0.0;
}
__device__ double bsp_deriv(double u) {
    if (u <= -1.5 || u >= 1.5) return 0.0;
    if (u < -0.5) return (4.0 / 3.0) * (1.5 + u);
    if (u < 0.5) return -(8.0 / 3.0) * u;
    return -(4.0 / 3.0) * (1.5 - u);
}
__device__ void cw_of(int idx, double& c, double& w) {
    int depth = 31 - __clz(idx + 1);
    int off = idx + 1 - (1 << depth);
    w = 1.0 / double(1 << depth);
    c = (0.5 + off) * w;
}

__global__ void __launch_bounds__(256) build_tables_kernel(int res, double* __restrict__ dot,
                                                           double* __restrict__ ddot, double* __restrict__ d2dot) {
    long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long long)res * res) return;
    int a = int(t % res), b = int(t / res);
    double ca, wa, cb, wb;
    cw_of(a, ca, wa);
    cw_of(b, cb, wb);
    double lo = fmax(ca - 1.5 * wa, cb - 1.5 * wb), hi = fmin(ca + 1.5 * wa, cb + 1.5 * wb);
    double r0 = 0, r1 = 0, r2 = 0;
    if (lo < hi) {
        double br[6];
        int nb = 0;
        br[nb++] = lo;
        double cand[4] = {ca - 0.5 * wa, ca + 0.5 * wa, cb - 0.5 * wb, cb + 0.5 * wb};
        for (int k = 0; k < 4; ++k)
            if (cand[k] > lo && cand[k] < hi) br[nb++] = cand[k];
        br[nb++] = hi;
        for (int i = 1; i < nb; ++i)  // insertion sort, <= 6 entries
            for (int j = i; j > 0 && br[j - 1] > br[j]; --j) {
                double x = br[j];
                br[j] = br[j - 1];
                br[j - 1] = x;
            }
        const double gx[3] = {-0.77459666924148337704, 0.0, 0.77459666924148337704};
        const double gw[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
        for (int k = 0; k + 1 < nb; ++k) {
            double x0 = br[k], x1 = br[k + 1];
            if (!(x1 > x0)) continue;
            double h = 0.5 * (x1 - x0), m = 0.5 * (x1 + x0);
            for (int g = 0; g < 3; ++g) {
                double x = m + h * gx[g];
                double ua = (x - ca) / wa, ub = (x - cb) / wb;
                double fa = bsp_d(ua), fb = bsp_d(ub);
                double da = bsp_deriv(ua) / wa, db = bsp_deriv(ub) / wb;
                double wg = gw[g] * h;
                r0 += wg * fa * fb;
                r1 += wg * da * fb;
                r2 += wg * da * db;
            }
        }
        if (fabs(r0) < 1e-15) r0 = r1 = r2 = 0;  // FunctionData.h:172
    }
    dot[t] = r0;
    ddot[t] = r1;
    d2dot[t] = r2;
}

// Concatenated ranges of the 27 neighbours of a node, held in LDS per wave.
struct NbRanges {
    int start[28];
    int first[27];
};

// Fills r for node o with range (first, len) per neighbour; returns total.
template <class LenFn>
__device__ __forceinline__ int wave_ranges(NbRanges& r, const int* __restrict__ neigh, int M, int o, int lane,
                                           LenFn fn) {
    int nb = -1, len = 0, first = 0;
    if (lane < 27) {
        nb = neigh[lane * M + o];
        if (nb != -1) fn(nb, first, len);
    }
    int inc = len;
    for (int off = 1; off < PSR_WAVE; off <<= 1) {
        int t = __shfl_up(inc, off, PSR_WAVE);
        if (lane >= off) inc += t;
    }
    if (lane < 27) {
        r.start[lane] = inc - len;
        r.first[lane] = first;
    }
    int total = __shfl(inc, 26, PSR_WAVE);
    if (lane == 0) r.start[27] = total;
    return total;
}

// wave_ranges for two nodes, one per half-wave (hl: the lane within the half;
// an invalid node has no range).
template <class LenFn>
__device__ __forceinline__ int half_ranges(NbRanges& r, const int* __restrict__ neigh, int M, int o, int hl,
                                           bool valid, LenFn fn) {
    constexpr int kHalf = PSR_WAVE / 2;
    int len = 0, first = 0;
    if (valid && hl < 27) {
        const int nb = neigh[hl * M + o];
        if (nb != -1) fn(nb, first, len);
    }
    int inc = len;
    for (int off = 1; off < kHalf; off <<= 1) {
        const int t = __shfl_up(inc, off, kHalf);
        if (hl >= off) inc += t;
    }
    if (hl < 27) {
        r.start[hl] = inc - len;
        r.first[hl] = first;
    }
    const int total = __shfl(inc, 26, kHalf);
    if (hl == 0) r.start[27] = total;
    return total;
}

__device__ __forceinline__ int range_of(const NbRanges& r, int t) {
    int i = 0;
    for (int step = 16; step > 0; step >>= 1)
        if (i + step < 28 && r.start[i + step] <= t) i += step;
    return i;
}

// --- vector field (ComputeVectorField.cu:61-100) ---------------------------
// V(o) = sum over the 27 neighbours' points p of F_o(x)F_o(y)F_o(z) n_p,
// two depth-D nodes per wave, one per half: a node's chain of dependent loads
// (neighbours, their point ranges, the points) is latency-bound, and the two
// chains of a wave overlap.
// Field storage: FP32, or FP16 holding V * 2^-(D+1) (the normals' scale, an
// exact power of two, so the stored value is the FP16 rounding of the sum of
// weighted unit normals).
__device__ __forceinline__ void store_field(float* p, float v, float) { *p = v; }
__device__ __forceinline__ void store_field(__half* p, float v, float invScale) { *p = __float2half_rn(v * invScale); }
__device__ __forceinline__ float load_field(const float* p, float) { return *p; }
__device__ __forceinline__ float load_field(const __half* p, float scale) { return __half2float(*p) * scale; }

constexpr int kVfWaves = 4;
template <class FieldT>
__device__ __forceinline__ void vector_field_kernel_body(int psr_bid, int psr_nbid, 
    int nD, int offD, int M, int D, const int* __restrict__ neigh, const int* __restrict__ pidx,
    const int* __restrict__ pnum, const int* __restrict__ key, const float4* __restrict__ pos,
    const float4* __restrict__ nrm, FieldT* __restrict__ vf, int* __restrict__ clear) {
    // the operator's per-depth entry counts and slot masks, which the
    // divergence launch's assembly adds to, and the CG error flag between
    // them in the results block
    if (psr_bid == 0 && threadIdx.x < kResultsClearWords) clear[threadIdx.x] = 0;
    __shared__ NbRanges R[kVfWaves][2];
    constexpr int kHalf = PSR_WAVE / 2;
    const int wid = threadIdx.x / PSR_WAVE, lane = __lane_id(), half = lane / kHalf, hl = lane % kHalf;
    const int node0 = (psr_bid * kVfWaves + wid) * 2;
    if (node0 >= nD) return;
    const int node = node0 + half;
    const bool valid = node < nD;
    const int o = offD + (valid ? node : node0);
    NbRanges& r = R[wid][half];
    const int total = half_ranges(r, neigh, M, o, hl, valid, [&](int nb, int& first, int& len) {
        first = pidx[nb];
        len = pnum[nb];
    });
    __builtin_amdgcn_wave_barrier();
    int3 f = key_func_index(key[o], D, D);
    float cx, cy, cz, invw;
    func_center_invw(f.x, cx, invw);
    func_center_invw(f.y, cy, invw);
    func_center_invw(f.z, cz, invw);
    float ax = 0.f, ay = 0.f, az = 0.f;
    for (int t = hl; t < total; t += kHalf) {
        int i = range_of(r, t);
        int p = r.first[i] + (t - r.start[i]);
        float4 q = pos[p];
        float4 n = nrm[p];
        float w = bspline((q.x - cx) * invw) * bspline((q.y - cy) * invw) * bspline((q.z - cz) * invw);
        ax += w * n.x;
        ay += w * n.y;
        az += w * n.z;
    }
    ax = half_sum_f32(ax, half);
    ay = half_sum_f32(ay, half);
    az = half_sum_f32(az, half);
    if (hl == 0 && valid) {
        const float inv = 1.0f / float(2 << D);
        store_field(&vf[node], ax, inv);
        store_field(&vf[nD + node], ay, inv);
        store_field(&vf[2 * nD + node], az, inv);
    }
}
template <class FieldT>
__global__ void __launch_bounds__(kVfWaves * PSR_WAVE) vector_field_kernel(
    int nD, int offD, int M, int D, const int* __restrict__ neigh, const int* __restrict__ pidx,
    const int* __restrict__ pnum, const int* __restrict__ key, const float4* __restrict__ pos,
    const float4* __restrict__ nrm, FieldT* __restrict__ vf, int* __restrict__ clear) {
    vector_field_kernel_body<FieldT>(xcd_block(blockIdx.x, gridDim.x), gridDim.x, nD, offD, M, D, neigh, pidx, pnum, key, pos, nrm, vf, clear);
}
template <class FieldT>
struct vector_field_kernel_k {
    static constexpr int kBlock = kVfWaves * PSR_WAVE;
    template <class... A>
    __device__ __forceinline__ static void run(int bid, int nbid, const A&... a) { vector_field_kernel_body<FieldT>(bid, nbid, a...); }
};


// --- divergence (ComputeNodesDivergence.cu:25-136) -------------------------
// b(o) = sum over the 27 neighbours' depth-D descendants j of
//        V_j . (<F'_ox, F_jx>, <F'_oy, F_jy>, <F'_oz, F_jz>)     (reference Q3)
template <class FieldT>
__device__ __forceinline__ double div_term(int j, int offD, int nD, int res, int D, int3 fo,
                                           const int* __restrict__ enc, const FieldT* __restrict__ vf,
                                           const double* __restrict__ ddot) {
    const float sc = float(2 << D);
    int3 fj = decode_func_index(enc[offD + j], D);
    double s = double(load_field(&vf[j], sc)) * ddot[fo.x + res * fj.x];
    s += double(load_field(&vf[nD + j], sc)) * ddot[fo.y + res * fj.y];
    s += double(load_field(&vf[2 * nD + j], sc)) * ddot[fo.z + res * fj.z];
    return s;
}

constexpr int kDivWaves = 4;
template <class FieldT>
__device__ __forceinline__ void divergence_fine_body(int psr_bid, int psr_nbid, 
    int begin, int count, int M, int offD, int nD, int res, int D, const int* __restrict__ neigh,
    const int* __restrict__ didx, const int* __restrict__ dnum, const int* __restrict__ enc,
    const FieldT* __restrict__ vf, const double* __restrict__ ddot, float* __restrict__ div) {
    // (two nodes per wave, as the vector field splats them, measured 64%
    // slower here: a node at depth 4 or 5 has hundreds to thousands of terms)
    __shared__ NbRanges R[kDivWaves];
    int wid = threadIdx.x / PSR_WAVE, lane = __lane_id();
    int node = psr_bid * kDivWaves + wid;
    if (node >= count) return;
    int o = begin + node;
    NbRanges& r = R[wid];
    int total = wave_ranges(r, neigh, M, o, lane, [&](int nb, int& first, int& len) {
        first = didx[nb];
        len = dnum[nb];
    });
    __builtin_amdgcn_wave_barrier();
    int3 fo = decode_func_index(enc[o], D);
    double acc = 0;
    for (int t = lane; t < total; t += PSR_WAVE) {
        int i = range_of(r, t);
        acc += div_term(r.first[i] + (t - r.start[i]), offD, nD, res, D, fo, enc, vf, ddot);
    }
    acc = wave_sum_f64(acc);
    if (lane == 0) div[o] = float(acc);
}


// Coarse nodes (few, each covering a large share of depth D): `parts`
// workgroups each, every one summing a strided share of the node's terms; the
// shares are folded in a fixed order by divergence_fold.
constexpr int kDivCoarseThreads = 256;
constexpr int kMaxCoarse = 4;  // coarse depths 0..3
struct CoarsePlan {
    int depths;                      // coarse depths 0..depths-1
    int blockStart[kMaxCoarse + 1];  // first workgroup (and partial) of each depth
    int parts[kMaxCoarse];           // workgroups per node
    int base[kMaxCoarse + 1];        // node range of each depth
};
template <class FieldT>
__device__ __forceinline__ void divergence_coarse_body(int psr_bid, int psr_nbid, 
    const CoarsePlan& P, int M, int offD, int nD, int res, int D, const int* __restrict__ neigh,
    const int* __restrict__ didx, const int* __restrict__ dnum, const int* __restrict__ enc,
    const FieldT* __restrict__ vf, const double* __restrict__ ddot, double* __restrict__ partial) {
    int d = 0;
    while (d + 1 < P.depths && int(psr_bid) >= P.blockStart[d + 1]) ++d;
    const int parts = P.parts[d];
    const int rel = psr_bid - P.blockStart[d];
    const int node = rel / parts, part = rel - node * parts;
    const int o = P.base[d] + node;
    __shared__ int start[28], first[27];
    __shared__ double red[kDivCoarseThreads / PSR_WAVE];
    if (threadIdx.x == 0) {
        int s = 0;
        for (int i = 0; i < 27; ++i) {
            int nb = neigh[i * M + o];
            start[i] = s;
            first[i] = nb != -1 ? didx[nb] : 0;
            s += nb != -1 ? dnum[nb] : 0;
        }
        start[27] = s;
    }
    __syncthreads();
    const int total = start[27];
    const int3 fo = decode_func_index(enc[o], D);
    double acc = 0;
    int i = 0;
    for (int t = part * kDivCoarseThreads + threadIdx.x; t < total; t += parts * kDivCoarseThreads) {
        while (start[i + 1] <= t) ++i;
        acc += div_term(first[i] + (t - start[i]), offD, nD, res, D, fo, enc, vf, ddot);
    }
    acc = wave_sum_f64(acc);
    if (__lane_id() == 0) red[threadIdx.x / PSR_WAVE] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int w = 0; w < kDivCoarseThreads / PSR_WAVE; ++w) s += red[w];
        partial[psr_bid] = s;
    }
}
// Depth-D nodes: their neighbours are depth-D nodes, each its own only
// descendant, so a thread sums the 27 terms of its node.  The first threads
// also fold the coarse nodes' shares (thread t: coarse node t), in order.
// Blocks from leafBlocks on assemble the operator instead (assemble_rows):
// independent work that fills the GPU beside this kernel's few workgroups.
// Depth-D nodes: their neighbours are depth-D nodes, each its own only
// descendant, so a thread sums the 27 terms of its node.
template <class FieldT>
__device__ __forceinline__ void divergence_leaf_nodes(int block, int offD, int nD, int M, int res, int D,
                                                      const int* __restrict__ neigh, const int* __restrict__ enc,
                                                      const FieldT* __restrict__ vf,
                                                      const double* __restrict__ ddot, float* __restrict__ div) {
    const int t = block * 256 + threadIdx.x;
    if (t >= nD) return;
    const int o = offD + t;
    const int3 fo = decode_func_index(enc[o], D);
    int nb[27];
#pragma unroll
    for (int i = 0; i < 27; ++i) nb[i] = neigh[i * M + o];
    double acc = 0;
#pragma unroll
    for (int i = 0; i < 27; ++i)
        if (nb[i] != -1) acc += div_term(nb[i] - offD, offD, nD, res, D, fo, enc, vf, ddot);
    div[o] = float(acc);
}

// Every depth in one launch, each role a range of 256-thread workgroups: the
// coarse nodes' shares, a middle-depth node per wave, a depth-D node per
// thread, and the operator's assembly (assemble_rows).  They read only the
// vector field, the octree and the tables and write disjoint outputs, so the
// roles fill the GPU side by side instead of one launch after another; the
// coarse shares are folded by divergence_fold.
struct DivRoles {
    int coarseBlocks, midBlocks, leafBlocks;
};
static_assert(kDivCoarseThreads == kDivWaves * PSR_WAVE && kAssembleThreads == kDivCoarseThreads,
              "one block size for every role");
template <class FieldT>
__device__ __forceinline__ void divergence_all_body(int psr_bid, int psr_nbid, const CoarsePlan& P, const DivRoles& R,
                                                    int begin, int count, int M, int offD, int nD, int res, int D,
                                                    const int* __restrict__ neigh, const int* __restrict__ didx,
                                                    const int* __restrict__ dnum, const int* __restrict__ enc,
                                                    const FieldT* __restrict__ vf, const double* __restrict__ ddot,
                                                    double* __restrict__ partial, float* __restrict__ div,
                                                    const AssembleArgs& aa) {
    int b = psr_bid;
    if (b < R.coarseBlocks) {
        divergence_coarse_body<FieldT>(b, R.coarseBlocks, P, M, offD, nD, res, D, neigh, didx, dnum, enc, vf, ddot,
                                       partial);
        return;
    }
    b -= R.coarseBlocks;
    if (b < R.midBlocks) {
        divergence_fine_body<FieldT>(b, R.midBlocks, begin, count, M, offD, nD, res, D, neigh, didx, dnum, enc, vf,
                                     ddot, div);
        return;
    }
    b -= R.midBlocks;
    if (b < R.leafBlocks) {
        divergence_leaf_nodes<FieldT>(b, offD, nD, M, res, D, neigh, enc, vf, ddot, div);
        return;
    }
    assemble_rows(aa, b - R.leafBlocks);
}
template <class FieldT>
__global__ void __launch_bounds__(kDivCoarseThreads) divergence_all(
    CoarsePlan P, DivRoles R, int begin, int count, int M, int offD, int nD, int res, int D,
    const int* __restrict__ neigh, const int* __restrict__ didx, const int* __restrict__ dnum,
    const int* __restrict__ enc, const FieldT* __restrict__ vf, const double* __restrict__ ddot,
    double* __restrict__ partial, float* __restrict__ div, AssembleArgs aa) {
    divergence_all_body<FieldT>(blockIdx.x, gridDim.x, P, R, begin, count, M, offD, nD, res, D, neigh, didx, dnum, enc,
                                vf, ddot, partial, div, aa);
}
template <class FieldT>
struct divergence_all_k {
    static constexpr int kBlock = kDivCoarseThreads;
    template <class... A>
    __device__ __forceinline__ static void run(int bid, int nbid, const A&... a) { divergence_all_body<FieldT>(bid, nbid, a...); }
};

// The coarse nodes' divergence: thread t folds node t's shares, in order.
__device__ __forceinline__ void divergence_fold_body(int psr_bid, int psr_nbid, const CoarsePlan& P,
                                                     const double* __restrict__ partial, float* __restrict__ div) {
    const int t = psr_bid * 256 + threadIdx.x;
    if (t >= P.base[P.depths]) return;
    int d = 0;
    while (t >= P.base[d + 1]) ++d;
    const int parts = P.parts[d];
    const double* pp = partial + P.blockStart[d] + size_t(t - P.base[d]) * parts;
    double s = 0;
    for (int k = 0; k < parts; ++k) s += pp[k];
    div[t] = float(s);
}
__global__ void __launch_bounds__(256) divergence_fold(CoarsePlan P, const double* __restrict__ partial,
                                                       float* __restrict__ div) {
    divergence_fold_body(blockIdx.x, gridDim.x, P, partial, div);
}
PSR_BATCHABLE(divergence_fold, 256)


}  // namespace

void build_tables(Ctx& c) {
    const int D = c.cfg.depth;
    const int res = (1 << (D + 1)) - 1;
    Tables& t = c.tables;
    t.res = res;
    size_t n = size_t(res) * res;
    t.dot = c.buf<double>("table_dot", n);
    t.ddot = c.buf<double>("table_ddot", n);
    t.d2dot = c.buf<double>("table_d2dot", n);
    build_tables_kernel<<<div_up(n, 256), 256, 0, c.stream>>>(res, t.dot, t.ddot, t.d2dot);
    PSR_HIP_CHECK(hipGetLastError());
    // Which depths' operators are indefinite once entries with |L| <= 1e-6 are
    // dropped (LaplacianSolver.cu:62).  An entry depends only on the neighbour
    // offset (the inner products are shift-invariant inside the domain), so a
    // depth's operator is a constant 27-point stencil; its symbol
    // sum_o L_o cos(theta . o), sampled over [0, pi]^3, takes both signs
    // exactly when the operator has eigenvalues of both signs on large
    // interiors.  At depth 9 only the diagonal and the six face entries
    // survive and the symbol at (pi, pi, pi) is negative.
    for (int d = 2; d <= D; ++d) {
        const int f = (1 << d) - 1 + (1 << (d - 1));  // an interior function of depth d
        double dt[2], d2[2];
        for (int k = 0; k < 2; ++k) {
            const size_t at = size_t(f) * res + size_t(f + k);
            PSR_HIP_CHECK(hipMemcpyAsync(&dt[k], t.dot + at, sizeof(double), hipMemcpyDeviceToHost, c.stream));
            PSR_HIP_CHECK(hipMemcpyAsync(&d2[k], t.d2dot + at, sizeof(double), hipMemcpyDeviceToHost, c.stream));
        }
        PSR_HIP_CHECK(hipStreamSynchronize(c.stream));
        double L[2][2][2];  // by |offset| per axis
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int e = 0; e < 2; ++e) {
                    const double v = dt[a] * dt[b] * dt[e] * (d2[a] + d2[b] + d2[e]);
                    L[a][b][e] = std::fabs(v) > double(1e-6f) ? v : 0.0;
                }
        constexpr int S = 32;
        double lo = 1e300;
        for (int i = 0; i <= S; ++i)
            for (int j = 0; j <= S; ++j)
                for (int k = 0; k <= S; ++k) {
                    // cos(theta o) summed over o = -1, 0, 1 per axis: 1 for |o| = 0, 2 cos theta for |o| = 1
                    const double cx[2] = {1.0, 2.0 * std::cos(M_PI * i / S)};
                    const double cy[2] = {1.0, 2.0 * std::cos(M_PI * j / S)};
                    const double cz[2] = {1.0, 2.0 * std::cos(M_PI * k / S)};
                    double sym = 0;
                    for (int a = 0; a < 2; ++a)
                        for (int b = 0; b < 2; ++b)
                            for (int e = 0; e < 2; ++e) sym += L[a][b][e] * cx[a] * cy[b] * cz[e];
                    lo = std::min(lo, sym);
                }
        c.indefinite[d] = lo < 0;
    }
}

namespace {

template <class FieldT>
void field_stages(Ctx& c, FieldT* vf) {
    Octree& o = c.oct;
    const Levels& L = o.L;
    const int D = L.D, nD = L.count[D], offD = L.base[D], M = o.M;
    hipStream_t s = c.stream;
    PSR_LAUNCH_K(vector_field_kernel<FieldT>, vector_field_kernel_k<FieldT>,
        std::max<size_t>(1, div_up(nD, 2 * kVfWaves)), 0, s,
        nD, offD, M, D, o.neigh, o.pidx, o.pnum, o.key, o.spos, o.snrm, vf, c.results()->cgNnz);
    // depths 0..kCoarse: several workgroups per node, per depth as many as the
    // node's expected share of depth D asks for; depth D: a thread per node;
    // between: a wave per node
    const int kCoarse = D >= 5 ? kMaxCoarse - 1 : 1;
    const int coarseCount = L.base[kCoarse + 1];
    CoarsePlan P{};
    double* partial = nullptr;
    {
        P.depths = kCoarse + 1;
        int blocks = 0;
        for (int d = 0; d <= kCoarse; ++d) {
            const double share = std::min(double(nD), 27.0 * nD / std::max(1, L.count[d]));
            P.parts[d] = std::max(1, std::min(64, int(share / 2048.0)));
            P.blockStart[d] = blocks;
            P.base[d] = L.base[d];
            blocks += P.parts[d] * L.count[d];
        }
        P.blockStart[kCoarse + 1] = blocks;
        P.base[kCoarse + 1] = L.base[kCoarse + 1];
        partial = c.buf<double>("div_partial", blocks);
    }
    // every depth's divergence and the operator's assembly (it needs only the
    // octree and the tables) in one launch, then the coarse shares' fold
    const int midCount = std::max(offD - coarseCount, 0);
    const DivRoles R{P.blockStart[kCoarse + 1], int(div_up(midCount, kDivWaves)), int(div_up(nD, 256))};
    const AssembleArgs aa{L,     M,     o.neigh, o.encIdx, c.tables.dot, c.tables.d2dot, c.tables.res,
                          c.buf<float>("lap_val", size_t(27) * M), c.results()->cgNnz, c.results()->cgSlotMask};
    const int assembleBlocks = int(div_up(size_t(3) * M, kAssembleThreads));
    PSR_LAUNCH_K(divergence_all<FieldT>, divergence_all_k<FieldT>,
                 R.coarseBlocks + R.midBlocks + R.leafBlocks + assembleBlocks, 0, s, P, R, coarseCount, midCount, M,
                 offD, nD, c.tables.res, D, o.neigh, o.didx, o.dnum, o.encIdx, vf, c.tables.ddot, partial, c.fld.div,
                 aa);
    PSR_LAUNCH(divergence_fold, div_up(coarseCount, 256), 0, s, P, partial, c.fld.div);
    PSR_HIP_CHECK(hipGetLastError());
}

}  // namespace

void compute_field(Ctx& c) {
    const Levels& L = c.oct.L;
    const int nD = L.count[L.D];
    c.fld.div = c.buf<float>("divergence", c.oct.M);
    if (c.cfg.precision == PSR_PRECISION_MIXED) {
        c.fld.vfield = nullptr;
        c.fld.vhalf = c.buf<__half>("vfield_half", size_t(3) * nD);
        field_stages(c, c.fld.vhalf);
    } else {
        c.fld.vhalf = nullptr;
        c.fld.vfield = c.buf<float>("vfield", size_t(3) * nD);
        field_stages(c, c.fld.vfield);
    }
}

}  // namespace psr
