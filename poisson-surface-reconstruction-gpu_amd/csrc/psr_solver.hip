// Per-depth Laplacian, concurrent CG over all depths, implicit function at the
// input points and the iso-value.
#include "psr_internal.h"

#include <algorithm>
#include <mutex>
#include <chrono>
#include <stdexcept>
#include <thread>
#include <vector>

namespace psr {

namespace {

constexpr int kThreads = 256;

// --- concurrent CG (CGAlgorithm.cu:10-147, 296-408) -------------------------
// Work item = a range of rows of one depth.  All work items of a depth form a
// group that iterates in lock step through a bounded-spin group barrier; a
// group of one workgroup uses __syncthreads only.
struct CgGroup {
    int depth;
    int rowBegin;   // global node index of the depth's first row
    int rows;
    int firstWg;    // first workgroup of the group
    int wgs;        // workgroups in the group
    int rowsPerWg;
    int indef;      // mixed form: the depth's operator is indefinite, solved with MINRES
};

struct CgParams {
    int ngroups;
    CgGroup g[PSR_MAX_DEPTH + 1];
    int M;
    float tolSq;     // reference: r1 > tol * tol with float tol (CGAlgorithm.cu:36)
    int maxIter;
    unsigned spinLimit;
    double relTol;   // mixed form: target ||b - A x|| / ||b|| per depth
    int maxOuter;    // mixed form: refinement steps
};

constexpr int kCgThreads = 256;
constexpr int kCgWaves = kCgThreads / PSR_WAVE;
// The FP32 generic form is compiled for 4 waves per SIMD (its persistent
// launch's resident budget follows from that occupancy; 21% faster sweeps at
// depth 9 than at 2), the mixed generic CG for 3.
constexpr int kCgFp32GenericWaves = 4, kCgMixedGenericWaves = 3;
// The MINRES kernel (indefinite depths of the mixed form) carries only its
// own path and is compiled for 4 waves per SIMD.
constexpr int kCgMinresWaves = 4;
// The mixed form's inner CG ends when its residual has not reached a new
// minimum for this many iterations, and its solution goes back to the
// iterate of the smallest residual.  A definite depth converges long before
// (20-70 iterations here); an operator that the reference's neighbour rules
// leave nonsymmetric or singular at a coarse depth (seen on ellipsoids of the
// 64-cloud batch) lets CG wander without converging, and the FP64 oracle's
// own CG diverges there.
constexpr int kCgStall = 128;

// Vectors read by other workgroups (r, p) and the reduction partials are
// accessed with agent-scope relaxed atomics, which are coherent across the
// XCDs without flushing or invalidating the L2; the barrier then only has to
// order them (s_waitcnt) before the arrival.
__device__ __forceinline__ float ld_shared(const float* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_shared(float* p, float v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Orders a wave's LDS and memory writes before its later reads of other
// lanes' values: the steps of a group that fits one wave need no workgroup
// barrier.
__device__ __forceinline__ void wave_sync_memory() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ double block_sum_double(double v, double* sh) {
    v = wave_sum_f64(v);
    int w = threadIdx.x / PSR_WAVE;
    if (__lane_id() == 0) sh[w] = v;
    __syncthreads();
    double s = 0;
    for (int k = 0; k < kCgWaves; ++k) s += sh[k];
    __syncthreads();
    return s;
}

// Per-group control block; each group's counters sit on their own cache lines.
struct GroupCtl {
    unsigned cnt;         // arrivals, monotonic: phase p is complete at p * workgroups
    unsigned pad0[31];
    unsigned gen;         // last phase whose total has been published
    unsigned pad1[31];
    unsigned res[2][32];  // published totals by phase parity
};
static_assert(sizeof(GroupCtl) == 512, "GroupCtl layout");

template <class T>
__device__ __forceinline__ void st_agent(T* p, const T& v) {
    static_assert(sizeof(T) % 4 == 0, "word-sized values");
    const unsigned* s = reinterpret_cast<const unsigned*>(&v);
    unsigned* d = reinterpret_cast<unsigned*>(p);
#pragma unroll
    for (int k = 0; k < int(sizeof(T) / 4); ++k) __hip_atomic_store(&d[k], s[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ T ld_agent(const T* p) {
    T v;
    unsigned* d = reinterpret_cast<unsigned*>(&v);
    const unsigned* s = reinterpret_cast<const unsigned*>(p);
#pragma unroll
    for (int k = 0; k < int(sizeof(T) / 4); ++k) d[k] = __hip_atomic_load(&s[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return v;
}

// Group all-reduce and barrier in one.  Every workgroup of the group stores its
// block total `mine` and arrives; the last to arrive folds the partials in a
// fixed order (fold(parts, n), called by the whole block, returns the same
// value in every thread), publishes the total and the phase; the others wait
// for the phase and read the total.  Every workgroup therefore sees the same
// bits.  The wait is bounded by spinLimit; on expiry (or when another
// workgroup has failed) the error flag is set and false returned.
template <class T, class Fold>
__device__ __forceinline__ bool group_allreduce(const T& mine, T& out, T* partial, GroupCtl* ctl, unsigned& phase,
                                                int G, int local, unsigned spinLimit, int* errFlag, int* shInt,
                                                T* shT, Fold fold) {
    ++phase;
    T* half = partial + (phase & 1u) * unsigned(G);
    if (threadIdx.x == 0) st_agent(&half[local], mine);
    __builtin_amdgcn_s_waitcnt(0);  // this wave's agent-scope stores have completed
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned old = __hip_atomic_fetch_add(&ctl->cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        *shInt = (old == phase * unsigned(G) - 1u) ? 1 : 0;
    }
    __syncthreads();
    T* res = reinterpret_cast<T*>(ctl->res[phase & 1u]);
    if (*shInt) {
        T t = fold(half, G);
        if (threadIdx.x == 0) {
            st_agent(res, t);
            __builtin_amdgcn_s_waitcnt(0);
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            __hip_atomic_store(&ctl->gen, phase, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        out = t;
        __syncthreads();
        return true;
    }
    if (threadIdx.x == 0) {
        unsigned spins = 0;
        int ok = 1;
        while (__hip_atomic_load(&ctl->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < phase) {
            if (++spins > spinLimit || __hip_atomic_load(errFlag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                ok = 0;
                __hip_atomic_store(errFlag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (ok) *shT = ld_agent(res);
        *shInt = ok;
    }
    __syncthreads();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    const bool ok = *shInt != 0;
    out = *shT;
    __syncthreads();
    return ok;
}

__device__ __forceinline__ double fold_double(const double* parts, int n, double* sh) {
    double t = 0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) t += ld_agent(&parts[k]);
    return block_sum_double(t, sh);
}

__device__ __forceinline__ float p_next(float r, float beta, float p) { return r + beta * p; }

// The stored entries of row i in slot order, acc = term(acc, v, j) for each:
// the N slots of `mask` (the slots some row of the depth stores, assembled
// with the operator).  Every weight and column load of the row is issued
// before the gathers that term makes; an entry this row does not store
// (v = 0: dropped or missing) gathers row i itself with weight zero, which
// adds exactly nothing.
template <int N, class T, class F>
__device__ __forceinline__ T row_product_n(unsigned mask, int i, int M, const float* __restrict__ val,
                                           const int* __restrict__ neigh, T acc, F term) {
    // in chunks of at most 9 slots, so that the loads in flight stay within
    // the register budget of the occupancy the CG launch is sized for
    constexpr int C = N < 9 ? N : 9;
    static_assert(N % C == 0, "whole chunks");
    unsigned m = mask;
#pragma unroll
    for (int c0 = 0; c0 < N; c0 += C) {
        float vv[C];
        int jj[C];
#pragma unroll
        for (int k = 0; k < C; ++k) {
            const int q = N == 27 ? c0 + k : __builtin_ctz(m);
            m &= m - 1;
            vv[k] = val[q * M + i];
            jj[k] = neigh[q * M + i];
        }
#pragma unroll
        for (int k = 0; k < C; ++k) acc = term(acc, vv[k], vv[k] != 0.f ? jj[k] : i);
    }
    return acc;
}
// Depths whose operator keeps only the diagonal and the six face entries
// (seven slots: depth 9 and finer after the 1e-6 drop) read those slots only;
// every other depth reads all 27.
template <class T, class F>
__device__ __forceinline__ T row_product(unsigned mask, int i, int M, const float* __restrict__ val,
                                         const int* __restrict__ neigh, T acc, F term) {
    if (__popc(mask) == 7) return row_product_n<7>(mask, i, M, val, neigh, acc, term);
    return row_product_n<27>(0x7ffffffu, i, M, val, neigh, acc, term);
}

// Vectors that other workgroups read are exchanged as packed 8- or 16-byte
// records through buffer loads and stores with the cache policy of the
// agent-scope atomics (sc1), so one load brings all of a neighbour's values.
constexpr int kCoherent = 16;  // cache-policy bits: sc1
__device__ __forceinline__ __amdgpu_buffer_rsrc_t vec_rsrc(const void* base, size_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), short(0), int(bytes), 0x00020000);
}
__device__ __forceinline__ float2 ld_vec2(__amdgpu_buffer_rsrc_t r, int idx) {
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, idx * 8, 0, kCoherent));
}
__device__ __forceinline__ void st_vec2(__amdgpu_buffer_rsrc_t r, int idx, float2 v) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, v), r,
                                          idx * 8, 0, kCoherent);
}
__device__ __forceinline__ float3 ld_vec3(__amdgpu_buffer_rsrc_t r, int idx) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b96(r, idx * 12, 0, kCoherent);
    return make_float3(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]));
}
__device__ __forceinline__ void st_vec3(__amdgpu_buffer_rsrc_t r, int idx, float3 v) {
    __attribute__((ext_vector_type(3))) unsigned u;
    u[0] = __float_as_uint(v.x);
    u[1] = __float_as_uint(v.y);
    u[2] = __float_as_uint(v.z);
    __builtin_amdgcn_raw_buffer_store_b96(u, r, idx * 12, 0, kCoherent);
}
__device__ __forceinline__ float4 ld_vec4(__amdgpu_buffer_rsrc_t r, int idx) {
    return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, idx * 16, 0, kCoherent));
}
__device__ __forceinline__ void st_vec4(__amdgpu_buffer_rsrc_t r, int idx, float4 v) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned, v), r,
                                           idx * 16, 0, kCoherent);
}

// Atomic-free group all-reduce of the mixed one-row kernel (measured faster
// there than the arrival counter of group_allreduce; the FP32 kernel, with two
// reductions and little work between them, is faster with the counter).  Each
// workgroup stores its block total in its own 16-byte slot(s) of the group's
// region, stamped with the phase (one store per slot, so a reader sees a slot
// whole); every
// workgroup polls the group's slots until all carry the phase and then folds
// them in the same fixed order, so every workgroup holds the same bits.  Slots
// are double-buffered by phase parity (a workgroup runs at most one phase
// ahead of the slowest) and zeroed before the launch (phases start at 1).
// S: slots per workgroup (two payload words each); thread t polls the slots
// of workgroups t and t + kCgThreads.
constexpr int kPollPerThread = 2;  // groups of up to kPollPerThread * kCgThreads workgroups
// Every workgroup owns kSlotStride slots per parity whatever S an exchange
// uses: the layout of a phase never overlaps that of the other parity, so a
// workgroup that has passed phase p and writes phase p + 1 cannot overwrite
// a phase-p slot that a slower workgroup has yet to read.  (With a stride of
// S, an S = 2 exchange right after an S = 1 one wrote into the S = 1 region
// of the other parity, and a slow reader could wait for a stamp that was
// already gone.)
constexpr int kSlotStride = 2;
template <int S>
__device__ __forceinline__ bool slot_exchange(float4* region, int G, int local, unsigned& phase,
                                              const unsigned (&mine)[2 * S], float4 (&got)[kPollPerThread][S],
                                              unsigned spinLimit, int* errFlag) {
    static_assert(S <= kSlotStride, "slots per workgroup");
    ++phase;
    const __amdgpu_buffer_rsrc_t rs =
        vec_rsrc(region + size_t(phase & 1u) * G * kSlotStride, sizeof(float4) * size_t(G) * kSlotStride);
    const float stamp = __uint_as_float(phase);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int h = 0; h < S; ++h)
            st_vec4(rs, local * kSlotStride + h,
                    make_float4(__uint_as_float(mine[2 * h]), __uint_as_float(mine[2 * h + 1]), stamp, 0.f));
    }
    for (unsigned spins = 0;; ++spins) {
        int ok = 1;
#pragma unroll
        for (int j = 0; j < kPollPerThread; ++j) {
            const int k = int(threadIdx.x) + j * kCgThreads;
            if (k < G) {
#pragma unroll
                for (int h = 0; h < S; ++h) {
                    got[j][h] = ld_vec4(rs, k * kSlotStride + h);
                    ok &= __float_as_uint(got[j][h].z) == phase ? 1 : 0;
                }
            }
        }
        if (__syncthreads_and(ok)) return true;
        if (spins >= spinLimit) {
            if (threadIdx.x == 0) __hip_atomic_store(errFlag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        if ((spins & 63u) == 63u) {
            const int e =
                threadIdx.x == 0 ? __hip_atomic_load(errFlag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
            if (__syncthreads_or(e)) return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// Sum of the group's block totals (v: this workgroup's, the same in every
// thread), folded as fold_double folds.
__device__ __forceinline__ bool slot_sum_double(double v, double& out, float4* region, int G, int local,
                                                unsigned& phase, unsigned spinLimit, int* errFlag, double* sh) {
    const unsigned long long u = static_cast<unsigned long long>(__double_as_longlong(v));
    const unsigned mine[2] = {unsigned(u), unsigned(u >> 32)};
    float4 got[kPollPerThread][1];
    if (!slot_exchange<1>(region, G, local, phase, mine, got, spinLimit, errFlag)) return false;
    double t = 0;
#pragma unroll
    for (int j = 0; j < kPollPerThread; ++j)
        if (int(threadIdx.x) + j * kCgThreads < G)
            t += __hiloint2double(__float_as_int(got[j][0].y), __float_as_int(got[j][0].x));
    out = block_sum_double(t, sh);
    return true;
}

// A row's 27 stencil entries, held in registers for the whole solve; a
// dropped or missing entry reads the row itself with weight 0.
struct Stencil {
    float v[27];
    int j[27];
    __device__ __forceinline__ void load(const float* __restrict__ val, const int* __restrict__ neigh, int M, int i) {
#pragma unroll
        for (int q = 0; q < 27; ++q) {
            v[q] = val[q * M + i];
            const int c = neigh[q * M + i];
            j[q] = (v[q] != 0.f) ? c : i;
        }
    }
};

// The reference's CG (CGAlgorithm.cu:10-80), one launch for all depths, one
// row per thread.  Per iteration k the record of row i in buffer k & 1 holds
// {r_k, p_{k-1}}; a row forms its neighbours' p_k = r_k + beta p_{k-1} on the
// fly, so the update of p and the product A p need no barrier in between
// (CGAlgorithm.cu:37-53).  The row's stencil, x, r, p and A p stay in
// registers; one 8-byte record per neighbour is read and one written per
// iteration.
__device__ __forceinline__ void cg_fp32_row_body(int psr_bid, int psr_nbid, const CgParams& P, const int* __restrict__ neigh,
                                                          const float* __restrict__ val, const float* __restrict__ b,
                                                          float* __restrict__ x, float2* state, double* partial,
                                                          GroupCtl* ctlAll, int* __restrict__ iters,
                                                          double* __restrict__ resid, int* errFlag) {
    __shared__ double sh[kCgWaves];
    __shared__ int shFlag;
    __shared__ double shRes;
    int gi = 0;
    while (gi + 1 < P.ngroups && int(psr_bid) >= P.g[gi + 1].firstWg) ++gi;
    const CgGroup G = P.g[gi];
    const int local = psr_bid - G.firstWg;
    const int i = G.rowBegin + local * G.rowsPerWg + int(threadIdx.x);
    const bool own = int(threadIdx.x) < G.rowsPerWg && i < G.rowBegin + G.rows;
    const int M = P.M;
    double* part = partial + 2 * G.firstWg;
    GroupCtl* ctl = ctlAll + gi;
    unsigned phase = 0;
    // a group of one workgroup exchanges its records through LDS (its rows
    // are the depth's, all in this workgroup): no memory round trip per sweep
    __shared__ float2 lrec2[2][kCgThreads];
    const bool multi = G.wgs > 1;
    // a group of at most one wave of rows (the coarsest depths) iterates in
    // wave 0 alone, with wave reductions and no workgroup barrier per step
    // (a wave's sum equals the block's, whose other waves add exact zeros)
    const bool oneWave = !multi && G.rows <= PSR_WAVE;
    if (oneWave && int(threadIdx.x) >= PSR_WAVE) return;
    const __amdgpu_buffer_rsrc_t st = vec_rsrc(state, sizeof(float2) * 2 * size_t(M));
    auto group_sum = [&](double v, double& out) -> bool {
        if (oneWave) {
            out = wave_sum_f64(v);
            wave_sync_memory();
            return true;
        }
        double s = block_sum_double(v, sh);
        if (!multi) {
            out = s;
            return true;
        }
        return group_allreduce(s, out, part, ctl, phase, G.wgs, local, P.spinLimit, errFlag, &shFlag, &shRes,
                               [&](const double* parts, int n) { return fold_double(parts, n, sh); });
    };
    Stencil sr;
    if (own) sr.load(val, neigh, M, i);
    // x = 0, r = b, rr = r.r  (CGAlgorithm.cu:21-33 with x0 = 0)
    float xi = 0.f, ri = 0.f, pi = 0.f, api = 0.f;
    double loc = 0;
    if (own) {
        ri = b[i];
        if (multi) st_vec2(st, i, make_float2(ri, 0.f));
        else lrec2[0][threadIdx.x] = make_float2(ri, 0.f);
        loc = double(ri * ri);
    }
    double rrD;
    if (!group_sum(loc, rrD)) return;
    float rr = float(rrD), rrOld = 1.f;
    int k = 1;
    while (rr > P.tolSq && k <= P.maxIter) {
        const float beta = (k > 1) ? rr / rrOld : 0.f;
        const int cur = ((k - 1) & 1) * M, nxt = M - cur, curL = (k - 1) & 1;
        // p = r + beta p (CGAlgorithm.cu:37-43), Ap = A p (:47), p.Ap (:53)
        double lpAp = 0;
        if (own) {
            float s = 0.f;
            if (oneWave) {
                // a straight-line LDS loop for the one-wave groups (the choice
                // per slot below costs them two branches a slot)
#pragma unroll
                for (int q = 0; q < 27; ++q) {
                    const float2 nb = lrec2[curL][sr.j[q] - G.rowBegin];
                    s += sr.v[q] * p_next(nb.x, beta, nb.y);
                }
            } else {
#pragma unroll
                for (int q = 0; q < 27; ++q) {
                    const float2 nb = multi ? ld_vec2(st, cur + sr.j[q]) : lrec2[curL][sr.j[q] - G.rowBegin];
                    s += sr.v[q] * p_next(nb.x, beta, nb.y);
                }
            }
            pi = p_next(ri, beta, pi);
            api = s;
            lpAp = double(pi * s);
        }
        double pAp;
        if (!group_sum(lpAp, pAp)) return;
        const float alpha = float(double(rr) / pAp);
        double lrr = 0;
        if (own) {
            xi = xi + alpha * pi;
            ri = ri + (-alpha) * api;
            if (multi) st_vec2(st, nxt + i, make_float2(ri, pi));
            else lrec2[curL ^ 1][threadIdx.x] = make_float2(ri, pi);
            lrr = double(ri * ri);
        }
        rrOld = rr;
        double rrNew;
        if (!group_sum(lrr, rrNew)) return;
        rr = float(rrNew);
        ++k;
    }
    if (own) x[i] = xi;
    if (local == 0 && threadIdx.x == 0) {
        iters[G.depth] = k - 1;
        iters[PSR_MAX_DEPTH + 1 + G.depth] = 1;
        resid[G.depth] = sqrt(double(rr));
    }
}
__global__ void __launch_bounds__(kCgThreads) cg_fp32_row(CgParams P, const int* __restrict__ neigh,
                                                          const float* __restrict__ val, const float* __restrict__ b,
                                                          float* __restrict__ x, float2* state, double* partial,
                                                          GroupCtl* ctlAll, int* __restrict__ iters,
                                                          double* __restrict__ resid, int* errFlag) {
    cg_fp32_row_body(blockIdx.x, gridDim.x, P, neigh, val, b, x, state, partial, ctlAll, iters, resid, errFlag);
}
PSR_BATCHABLE(cg_fp32_row, kCgThreads)


// A group of one workgroup with at most one row a thread (the coarse depths,
// and a batch's small clouds), solved by that workgroup alone with its
// vectors in LDS: cg_all_depths' arithmetic without a memory round trip per
// sweep.
// Makes p opaque to the compiler: loads through it are repeated after this
// point instead of being kept in registers from an earlier sweep.
template <class T>
__device__ __forceinline__ void opaque(const T*& p) {
    asm volatile("" : "+v"(p));
}

__device__ __forceinline__ void cg_small_fp32(const CgParams& P, const CgGroup& G, const int* __restrict__ neigh,
                                              const float* __restrict__ val, const float* __restrict__ b,
                                              float* __restrict__ x, int* __restrict__ iters,
                                              double* __restrict__ resid, double* sh) {
    // the small groups' stencil is re-read every sweep (it stays in cache) rather than
    // held in registers: this path shares its kernel with the large-depth sweeps
    const float* vs = val;
    const int* ns = neigh;
    __shared__ float lr[kCgThreads], lp[2][kCgThreads];
    const int M = P.M, t = int(threadIdx.x), i = G.rowBegin + t, rb = G.rowBegin;
    const bool own = t < G.rows;
    float xi = 0.f, ri = 0.f, pi = 0.f, api = 0.f;
    double loc = 0;
    if (own) {
        ri = b[i];
        lr[t] = ri;
        lp[0][t] = 0.f;
        loc = double(ri * ri);
    }
    float rr = float(block_sum_double(loc, sh)), rrOld = 1.f;
    int k = 1;
    while (rr > P.tolSq && k <= P.maxIter) {
        const float beta = (k > 1) ? rr / rrOld : 0.f;
        const int cur = (k - 1) & 1;
        double lpAp = 0;
        if (own) {
            float s = 0.f;
            opaque(vs);
            opaque(ns);
            s = row_product_n<27>(0x7ffffffu, i, M, vs, ns, s, [&](float a, float v, int j) {
                return a + v * p_next(lr[j - rb], beta, lp[cur][j - rb]);
            });
            pi = p_next(ri, beta, lp[cur][t]);
            lp[cur ^ 1][t] = pi;
            api = s;
            lpAp = double(pi * s);
        }
        const double pAp = block_sum_double(lpAp, sh);
        const float alpha = float(double(rr) / pAp);
        double lrr = 0;
        if (own) {
            xi = xi + alpha * pi;
            ri = ri + (-alpha) * api;
            lr[t] = ri;
            lrr = double(ri * ri);
        }
        rrOld = rr;
        rr = float(block_sum_double(lrr, sh));
        ++k;
    }
    if (own) x[i] = xi;
    if (threadIdx.x == 0) {
        iters[G.depth] = k - 1;
        iters[PSR_MAX_DEPTH + 1 + G.depth] = 1;
        resid[G.depth] = sqrt(double(rr));
    }
}

// The same iteration with several rows per thread (large depths): vectors in
// memory, the stencil read every sweep.
__device__ __forceinline__ void cg_all_depths_body(int psr_bid, int psr_nbid, const CgParams& P, const int* __restrict__ neigh,
                                                            const float* __restrict__ val,
                                                            const float* __restrict__ b, float* __restrict__ x,
                                                            float* r, float* pA, float* pB,
                                                            float* __restrict__ Ap, double* partial,
                                                            GroupCtl* ctlAll, int* __restrict__ iters,
                                                            double* __restrict__ resid, int* errFlag,
                                                            const unsigned* __restrict__ slotMasks) {
    __shared__ double sh[kCgWaves];
    __shared__ int shFlag;
    __shared__ double shRes;
    int gi = 0;
    while (gi + 1 < P.ngroups && int(psr_bid) >= P.g[gi + 1].firstWg) ++gi;
    const CgGroup G = P.g[gi];
    const int local = psr_bid - G.firstWg;
    const int r0 = G.rowBegin + local * G.rowsPerWg;
    const int r1 = min(G.rowBegin + G.rows, r0 + G.rowsPerWg);
    const int M = P.M;
    double* part = partial + 2 * G.firstWg;
    GroupCtl* ctl = ctlAll + gi;
    unsigned phase = 0;
    const bool multi = G.wgs > 1;
    if (!multi && G.rows <= kCgThreads) {
        cg_small_fp32(P, G, neigh, val, b, x, iters, resid, sh);
        return;
    }
    const unsigned slotMask = slotMasks[G.depth];

    // Sum over the group (partials double-buffered by phase parity).
    auto group_sum = [&](double v, double& out) -> bool {
        double s = block_sum_double(v, sh);
        if (!multi) {
            out = s;
            return true;
        }
        return group_allreduce(s, out, part, ctl, phase, G.wgs, local, P.spinLimit, errFlag, &shFlag, &shRes,
                               [&](const double* parts, int n) { return fold_double(parts, n, sh); });
    };

    // x = 0, r = b, rr = r.r  (CGAlgorithm.cu:21-33 with x0 = 0)
    double loc = 0;
    for (int i = r0 + threadIdx.x; i < r1; i += kCgThreads) {
        float bi = b[i];
        x[i] = 0.f;
        st_shared(&r[i], bi);
        st_shared(&pA[i], 0.f);
        loc += double(bi * bi);
    }
    double rrD;
    if (!group_sum(loc, rrD)) return;
    float rr = float(rrD), rrOld = 1.f;
    int k = 1;
    float* pOld = pA;
    float* pNew = pB;
    while (rr > P.tolSq && k <= P.maxIter) {
        const float beta = (k > 1) ? rr / rrOld : 0.f;
        // p = r + beta p (CGAlgorithm.cu:37-43), Ap = A p (:47), p.Ap (:53).  The
        // neighbours' p entries are formed from the previous r and p, so no
        // barrier is needed between the update and the product.
        double lpAp = 0;
        for (int i = r0 + threadIdx.x; i < r1; i += kCgThreads) {
            float s = 0.f;
            s = row_product(slotMask, i, M, val, neigh, s, [&](float a, float v, int j) {
                return a + v * p_next(ld_shared(&r[j]), beta, ld_shared(&pOld[j]));
            });
            float pi = p_next(ld_shared(&r[i]), beta, ld_shared(&pOld[i]));
            st_shared(&pNew[i], pi);
            Ap[i] = s;
            lpAp += double(pi * s);
        }
        double pAp;
        if (!group_sum(lpAp, pAp)) return;
        const float alpha = float(double(rr) / pAp);
        double lrr = 0;
        for (int i = r0 + threadIdx.x; i < r1; i += kCgThreads) {
            float pi = ld_shared(&pNew[i]);
            x[i] = x[i] + alpha * pi;
            float ri = ld_shared(&r[i]) + (-alpha) * Ap[i];
            st_shared(&r[i], ri);
            lrr += double(ri * ri);
        }
        rrOld = rr;
        double rrNew;
        if (!group_sum(lrr, rrNew)) return;
        rr = float(rrNew);
        float* t = pOld;
        pOld = pNew;
        pNew = t;
        ++k;
    }
    if (local == 0 && threadIdx.x == 0) {
        iters[G.depth] = k - 1;
        iters[PSR_MAX_DEPTH + 1 + G.depth] = 1;
        resid[G.depth] = sqrt(double(rr));
    }
}
__global__ void __launch_bounds__(kCgThreads) __attribute__((amdgpu_waves_per_eu(kCgFp32GenericWaves))) cg_all_depths(CgParams P, const int* __restrict__ neigh,
                                                            const float* __restrict__ val,
                                                            const float* __restrict__ b, float* __restrict__ x,
                                                            float* r, float* pA, float* pB,
                                                            float* __restrict__ Ap, double* partial,
                                                            GroupCtl* ctlAll, int* __restrict__ iters,
                                                            double* __restrict__ resid, int* errFlag,
                                                            const unsigned* __restrict__ slotMasks) {
    cg_all_depths_body(blockIdx.x, gridDim.x, P, neigh, val, b, x, r, pA, pB, Ap, partial, ctlAll, iters, resid, errFlag,
                       slotMasks);
}
PSR_BATCHABLE_WAVES(cg_all_depths, kCgThreads, kCgFp32GenericWaves)


// --- mixed-precision CG ------------------------------------------------------
// Same system per depth, solved to a true relative residual:
//   outer (FP64): r = b - A x accumulated in double, ||r|| / ||b|| in double;
//   inner (FP32): Chronopoulos-Gear CG on A e = r, one group reduction per
//                 iteration, dot products as compensated FP32 sums (Dot2:
//                 TwoProduct by fma, TwoSum accumulation).
// One barrier per inner iteration is possible because every value a row
// reads from its neighbours was written before the previous barrier: the
// neighbour's next residual is formed on the fly as
//   r_{k+1} = r_k - alpha_k (w_k + beta_k s_{k-1}),  w = A r,  s = A p,
// with r, w and s double-buffered by iteration parity.

__device__ __forceinline__ void two_sum_acc(float& s, float& c, float v) {
#pragma clang fp contract(off)
    float t = s + v;
    float bv = t - s;
    float e = (s - (t - bv)) + (v - bv);
    s = t;
    c += e;
}
__device__ __forceinline__ void dot2_acc(float& s, float& c, float a, float b) {
#pragma clang fp contract(off)
    float p = a * b;
    float e = __fmaf_rn(a, b, -p);
    two_sum_acc(s, c, p);
    c += e;
}
__device__ __forceinline__ void comp_merge(float& s, float& c, float s2, float c2) {
#pragma clang fp contract(off)
    two_sum_acc(s, c, s2);
    c += c2;
}
// s_k = w_k + beta s_{k-1};  r_{k+1} = r_k - alpha s_k (owner and neighbours use the same operations)
__device__ __forceinline__ float s_next(float w, float beta, float sPrev) { return __fmaf_rn(beta, sPrev, w); }
__device__ __forceinline__ float r_next(float r, float alpha, float s) { return __fmaf_rn(-alpha, s, r); }

struct Comp2 {
    float gs, gc, ds, dc;  // gamma = gs + gc, delta = ds + dc
};

__device__ __forceinline__ void comp2_merge(Comp2& a, const Comp2& b) {
    comp_merge(a.gs, a.gc, b.gs, b.gc);
    comp_merge(a.ds, a.dc, b.ds, b.dc);
}

// Block-wide merge in a fixed order; every thread returns the same value.
template <int CTRL>
__device__ __forceinline__ void comp2_dpp_step(Comp2& v) {
    const Comp2 o{dpp_f32<CTRL>(v.gs), dpp_f32<CTRL>(v.gc), dpp_f32<CTRL>(v.ds), dpp_f32<CTRL>(v.dc)};
    comp2_merge(v, o);
}
__device__ __forceinline__ Comp2 readlane_comp2(const Comp2& v, int lane) {
    return Comp2{readlane_f32(v.gs, lane), readlane_f32(v.gc, lane), readlane_f32(v.ds, lane), readlane_f32(v.dc, lane)};
}
// The wave's merge through DPP (psr_common.h: rows of 16, then the four rows
// in a fixed order); every lane returns the same value.
__device__ __forceinline__ Comp2 wave_comp2(Comp2 v) {
    comp2_dpp_step<kDppSwap1>(v);
    comp2_dpp_step<kDppSwap2>(v);
    comp2_dpp_step<kDppHalfMirror>(v);
    comp2_dpp_step<kDppMirror>(v);
    Comp2 r = readlane_comp2(v, 0);
    comp2_merge(r, readlane_comp2(v, 16));
    comp2_merge(r, readlane_comp2(v, 32));
    comp2_merge(r, readlane_comp2(v, 48));
    return r;
}
__device__ __forceinline__ Comp2 block_comp2(Comp2 v, Comp2* sh) {
    v = wave_comp2(v);
    int w = threadIdx.x / PSR_WAVE;
    if (__lane_id() == 0) sh[w] = v;
    __syncthreads();
    Comp2 t = sh[0];
    for (int k = 1; k < kCgWaves; ++k) comp2_merge(t, sh[k]);
    __syncthreads();
    return t;
}

// The row form's two dot products, gamma = r.r and delta = w.r, summed in
// FP64: a lane's product of two floats is exact in FP64, and the wave, block
// and group sums take the order of wave_sum_f64, block_sum_double and
// slot_sum_double, so gamma of a right-hand side b is ||b||^2 to the bit.
struct Dots {
    double g, d;
};
__device__ __forceinline__ Dots wave_sum_dots(Dots v) {
    v.g += dpp_f64<kDppSwap1>(v.g);
    v.d += dpp_f64<kDppSwap1>(v.d);
    v.g += dpp_f64<kDppSwap2>(v.g);
    v.d += dpp_f64<kDppSwap2>(v.d);
    v.g += dpp_f64<kDppHalfMirror>(v.g);
    v.d += dpp_f64<kDppHalfMirror>(v.d);
    v.g += dpp_f64<kDppMirror>(v.g);
    v.d += dpp_f64<kDppMirror>(v.d);
    return Dots{(readlane_f64(v.g, 0) + readlane_f64(v.g, 16)) + (readlane_f64(v.g, 32) + readlane_f64(v.g, 48)),
                (readlane_f64(v.d, 0) + readlane_f64(v.d, 16)) + (readlane_f64(v.d, 32) + readlane_f64(v.d, 48))};
}
// fenceStores: the block's earlier stores complete before its barrier (a
// workgroup barrier alone does not wait for them); the wave's sum is formed
// while they are in flight.
__device__ __forceinline__ Dots block_sum_dots(Dots v, double2* sh, bool fenceStores = false) {
    v = wave_sum_dots(v);
    if (fenceStores) __builtin_amdgcn_s_waitcnt(0);
    const int w = threadIdx.x / PSR_WAVE;
    if (__lane_id() == 0) sh[w] = make_double2(v.g, v.d);
    __syncthreads();
    Dots s{0.0, 0.0};
    for (int k = 0; k < kCgWaves; ++k) {
        s.g += sh[k].x;
        s.d += sh[k].y;
    }
    __syncthreads();
    return s;
}

// Mixed form, one launch for all depths, one row per thread.  Per inner
// iteration k the 12-byte record of row i in buffer k & 1 holds {r_k, w_k, s_{k-1}};
// the row writes {r_{k+1}, w_{k+1}, s_k} to the other buffer, keeps its r, w,
// s, p and x in registers (x is also stored for the outer residual) and its
// stencil in LDS.
__device__ __forceinline__ void cg_mixed_row_body(int psr_bid, int psr_nbid, const CgParams& P, const int* __restrict__ neigh,
                                                           const float* __restrict__ val,
                                                           const float* __restrict__ b, float* x, float* rOut,
                                                           float3* state, float4* slots, int* __restrict__ iters,
                                                           double* __restrict__ resid, double* __restrict__ trueRel,
                                                           int* errFlag) {
    __shared__ double shd[kCgWaves];
    __shared__ double2 shp[kCgWaves];
    __shared__ float lV[27][kCgThreads];  // the row's stencil, read from LDS every sweep
    __shared__ int lJ[27][kCgThreads];
    int gi = 0;
    while (gi + 1 < P.ngroups && int(psr_bid) >= P.g[gi + 1].firstWg) ++gi;
    const CgGroup G = P.g[gi];
    const int local = psr_bid - G.firstWg;
    const int i = G.rowBegin + local * G.rowsPerWg + int(threadIdx.x);
    const bool own = int(threadIdx.x) < G.rowsPerWg && i < G.rowBegin + G.rows;
    const int M = P.M;
    float4* region = slots + size_t(4) * G.firstWg;  // [2][wgs][2]
    unsigned phase = 0;
    // a group of one workgroup exchanges its records through LDS
    __shared__ float3 lrec3[2][kCgThreads];
    const bool multi = G.wgs > 1;
    // a group of at most one wave of rows iterates in wave 0 alone (as in the
    // FP32 row form)
    const bool oneWave = !multi && G.rows <= PSR_WAVE;
    if (oneWave && int(threadIdx.x) >= PSR_WAVE) return;
    const __amdgpu_buffer_rsrc_t st = vec_rsrc(state, sizeof(float3) * 2 * size_t(M));
    auto sum_double = [&](double v, double& out) -> bool {
        if (oneWave) {
            out = wave_sum_f64(v);
            wave_sync_memory();
            return true;
        }
        // the stores of x and of the outer residual complete before the slot
        // (as the records before the dots' slot)
        if (multi) __builtin_amdgcn_s_waitcnt(0);
        double s = block_sum_double(v, shd);
        if (!multi) {
            out = s;
            return true;
        }
        return slot_sum_double(s, out, region, G.wgs, local, phase, P.spinLimit, errFlag, shd);
    };
    auto sum_dots = [&](Dots v, Dots& out) -> bool {
        if (oneWave) {
            out = wave_sum_dots(v);
            wave_sync_memory();
            return true;
        }
        // Every wave's record stores complete before the barrier, so the slot
        // stored after it is never seen before them: a workgroup that has seen
        // every slot of a phase reads the records of that phase.
        const Dots s = block_sum_dots(v, shp, multi);
        if (!multi) {
            out = s;
            return true;
        }
        const unsigned long long ug = static_cast<unsigned long long>(__double_as_longlong(s.g));
        const unsigned long long ud = static_cast<unsigned long long>(__double_as_longlong(s.d));
        const unsigned mine[4] = {unsigned(ug), unsigned(ug >> 32), unsigned(ud), unsigned(ud >> 32)};
        float4 got[kPollPerThread][2];
        if (!slot_exchange<2>(region, G.wgs, local, phase, mine, got, P.spinLimit, errFlag)) return false;
        Dots t{0.0, 0.0};
#pragma unroll
        for (int j = 0; j < kPollPerThread; ++j)
            if (int(threadIdx.x) + j * kCgThreads < G.wgs) {
                t.g += __hiloint2double(__float_as_int(got[j][0].y), __float_as_int(got[j][0].x));
                t.d += __hiloint2double(__float_as_int(got[j][1].y), __float_as_int(got[j][1].x));
            }
        out = block_sum_dots(t, shp);
        return true;
    };
    if (own) {
        Stencil sr;
        sr.load(val, neigh, M, i);
#pragma unroll
        for (int q = 0; q < 27; ++q) {
            lV[q][threadIdx.x] = sr.v[q];
            lJ[q][threadIdx.x] = sr.j[q];
        }
    }
    float xi = 0.f, ro = 0.f, wo = 0.f, so = 0.f, po = 0.f;
    // x = 0.  ||b||^2 in FP64 (the outer residual of x = 0 is b itself) is
    // gamma of the first outer step, whose sum is bit for bit that of b.b;
    // a solve that takes no step sums it alone.
    if (own) st_shared(&x[i], 0.f);
    double bb = 0;
    const bool stepping = P.maxOuter > 0 && P.maxIter > 0;
    if (!stepping) {
        double loc = 0;
        if (own) {
            const double bi = b[i];
            loc = bi * bi;
        }
        if (!sum_double(loc, bb)) return;
    }
    double rr = bb, rel = (bb > 0 || stepping) ? 1.0 : 0.0;
    const float* rhs = b;
    int total = 0, outer = 0;
    float gamma = 0.f;
    while (rel > P.relTol && outer < P.maxOuter && total < P.maxIter) {
        // r_0 = rhs, w_0 = A r_0, s_{-1} = 0, p_{-1} = 0
        Dots d0{0.0, 0.0};
        if (own) {
            float w = 0.f;
#pragma unroll
            for (int q = 0; q < 27; ++q) w = __fmaf_rn(lV[q][threadIdx.x], ld_shared(&rhs[lJ[q][threadIdx.x]]), w);
            ro = ld_shared(&rhs[i]);
            wo = w;
            so = 0.f;
            po = 0.f;
            if (multi) st_vec3(st, i, make_float3(ro, w, 0.f));
            else lrec3[0][threadIdx.x] = make_float3(ro, w, 0.f);
            d0.g = double(ro) * double(ro);
            d0.d = double(w) * double(ro);
        }
        Dots g;
        if (!sum_dots(d0, g)) return;
        if (outer == 0) {
            bb = rr = g.g;
            if (!(bb > 0)) {
                rel = 0.0;
                break;
            }
        }
        // inner target on the updated residual: half the tolerance left, no tighter than FP32 can follow
        const double tgt = fmax(0.5 * P.relTol * sqrt(bb), 1e-6 * sqrt(rr));
        const float tgt2 = float(tgt * tgt);
        gamma = float(g.g);
        float delta = float(g.d);
        float alpha = gamma / delta, beta = 0.f;
        int k = 0, kBest = 0;
        float gBest = gamma, xBest = xi;
        while (gamma > tgt2 && total + k < P.maxIter && k - kBest < kCgStall) {
            const int cur = (k & 1) * M, nxt = M - cur;
            Dots dl{0.0, 0.0};
            if (own) {
                float w = 0.f;
                if (oneWave) {
                    // a straight-line LDS loop for the one-wave groups
#pragma unroll
                    for (int q = 0; q < 27; ++q) {
                        const float3 nb = lrec3[k & 1][lJ[q][threadIdx.x] - G.rowBegin];
                        w = __fmaf_rn(lV[q][threadIdx.x], r_next(nb.x, alpha, s_next(nb.y, beta, nb.z)), w);
                    }
                } else {
#pragma unroll
                    for (int q = 0; q < 27; ++q) {
                        const int jq = lJ[q][threadIdx.x];
                        const float3 nb = multi ? ld_vec3(st, cur + jq) : lrec3[k & 1][jq - G.rowBegin];
                        w = __fmaf_rn(lV[q][threadIdx.x], r_next(nb.x, alpha, s_next(nb.y, beta, nb.z)), w);
                    }
                }
                const float si = s_next(wo, beta, so);
                const float rni = r_next(ro, alpha, si);
                // the record first: the exchange waits for it
                if (multi) st_vec3(st, nxt + i, make_float3(rni, w, si));
                else lrec3[(k & 1) ^ 1][threadIdx.x] = make_float3(rni, w, si);
                po = __fmaf_rn(beta, po, ro);
                xi = __fmaf_rn(alpha, po, xi);
                st_shared(&x[i], xi);
                ro = rni;
                wo = w;
                so = si;
                dl.g = double(rni) * double(rni);
                dl.d = double(w) * double(rni);
            }
            if (!sum_dots(dl, g)) return;
            const float gNew = float(g.g), dNew = float(g.d);
            const float bNew = gNew / gamma;
            alpha = gNew / (dNew - bNew * gNew / alpha);
            beta = bNew;
            gamma = gNew;
            ++k;
            if (gamma < gBest) {  // the same in every workgroup of the group
                gBest = gamma;
                kBest = k;
                xBest = xi;
            }
        }
        total += k;
        ++outer;
        if (kBest < k) {
            // back to the iterate of the smallest residual; one more exchange
            // orders the restored x before the outer residual reads it
            gamma = gBest;
            xi = xBest;
            if (own) st_shared(&x[i], xi);
            double fence;
            if (!sum_double(0.0, fence)) return;
        }
        // outer residual in FP64
        double loc = 0;
        if (own) {
            double acc = b[i];
#pragma unroll
            for (int q = 0; q < 27; ++q)
                acc -= double(lV[q][threadIdx.x]) * double(ld_shared(&x[lJ[q][threadIdx.x]]));
            st_shared(&rOut[i], float(acc));
            loc = acc * acc;
        }
        if (!sum_double(loc, rr)) return;
        rel = bb > 0 ? sqrt(rr / bb) : 0.0;
        rhs = rOut;
        // an inner solve that found no smaller residual leaves x as it was:
        // another outer step would repeat it
        if (kBest == 0) break;
    }
    if (local == 0 && threadIdx.x == 0) {
        iters[G.depth] = total;
        iters[PSR_MAX_DEPTH + 1 + G.depth] = outer;
        resid[G.depth] = sqrt(double(gamma));
        trueRel[G.depth] = rel;
    }
}
__global__ void __launch_bounds__(kCgThreads) cg_mixed_row(CgParams P, const int* __restrict__ neigh,
                                                           const float* __restrict__ val,
                                                           const float* __restrict__ b, float* x, float* rOut,
                                                           float3* state, float4* slots, int* __restrict__ iters,
                                                           double* __restrict__ resid, double* __restrict__ trueRel,
                                                           int* errFlag) {
    cg_mixed_row_body(blockIdx.x, gridDim.x, P, neigh, val, b, x, rOut, state, slots, iters, resid, trueRel, errFlag);
}
PSR_BATCHABLE(cg_mixed_row, kCgThreads)


// The mixed form's group of one workgroup with at most one row a thread,
// vectors in LDS: cg_mixed_all_depths' arithmetic without a memory round trip
// per sweep.
__device__ __forceinline__ void cg_small_mixed(const CgParams& P, const CgGroup& G, const int* __restrict__ neigh,
                                               const float* __restrict__ val, const float* __restrict__ b,
                                               float* __restrict__ x, int* __restrict__ iters,
                                               double* __restrict__ resid, double* __restrict__ trueRel,
                                               double* shd, Comp2* shc) {
    // the small groups' stencil is re-read every sweep (it stays in cache) rather than
    // held in registers: this path shares its kernel with the large-depth sweeps
    const float* vs = val;
    const int* ns = neigh;
    __shared__ float lx[kCgThreads], lro[kCgThreads], lr[2][kCgThreads], lw[2][kCgThreads], ls[2][kCgThreads];
    const int M = P.M, t = int(threadIdx.x), i = G.rowBegin + t, rb = G.rowBegin;
    const bool own = t < G.rows;
    // x = 0; ||b||^2 in FP64 (the outer residual of x = 0 is b itself)
    double loc = 0;
    if (own) {
        lx[t] = 0.f;
        const double bi = b[i];
        loc = bi * bi;
    }
    const double bb = block_sum_double(loc, shd);
    double rr = bb, rel = bb > 0 ? 1.0 : 0.0;
    bool fromB = true;  // the first outer step solves for b, the later ones for the residual in lro
    int total = 0, outer = 0;
    float gamma = 0.f, pv = 0.f;
    while (rel > P.relTol && outer < P.maxOuter && total < P.maxIter) {
        const double tgt = fmax(0.5 * P.relTol * sqrt(bb), 1e-6 * sqrt(rr));
        const float tgt2 = float(tgt * tgt);
        // r_0 = rhs, w_0 = A r_0, s_{-1} = 0, p_{-1} = 0
        Comp2 d0{0.f, 0.f, 0.f, 0.f};
        if (own) {
            float w = 0.f;
            opaque(vs);
            opaque(ns);
            w = row_product_n<27>(0x7ffffffu, i, M, vs, ns, w, [&](float a, float v, int j) {
                return __fmaf_rn(v, fromB ? b[j] : lro[j - rb], a);
            });
            const float ri = fromB ? b[i] : lro[t];
            lr[0][t] = ri;
            lw[0][t] = w;
            ls[1][t] = 0.f;
            pv = 0.f;
            dot2_acc(d0.gs, d0.gc, ri, ri);
            dot2_acc(d0.ds, d0.dc, w, ri);
        }
        Comp2 g = block_comp2(d0, shc);
        gamma = g.gs + g.gc;
        const float delta = g.ds + g.dc;
        float alpha = gamma / delta, beta = 0.f;
        int k = 0, kBest = 0;
        float gBest = gamma, xBest = own ? lx[t] : 0.f;
        while (gamma > tgt2 && total + k < P.maxIter && k - kBest < kCgStall) {
            const int c = k & 1, n = c ^ 1;  // r_k, w_k in c; s_{k-1} in n
            Comp2 dl{0.f, 0.f, 0.f, 0.f};
            if (own) {
                float w = 0.f;
                opaque(vs);
                opaque(ns);
                w = row_product_n<27>(0x7ffffffu, i, M, vs, ns, w, [&](float a, float v, int j) {
                    j -= rb;
                    return __fmaf_rn(v, r_next(lr[c][j], alpha, s_next(lw[c][j], beta, ls[n][j])), a);
                });
                const float ri = lr[c][t];
                const float si = s_next(lw[c][t], beta, ls[n][t]);
                const float rni = r_next(ri, alpha, si);
                pv = __fmaf_rn(beta, pv, ri);
                lx[t] = __fmaf_rn(alpha, pv, lx[t]);
                ls[c][t] = si;
                lr[n][t] = rni;
                lw[n][t] = w;
                dot2_acc(dl.gs, dl.gc, rni, rni);
                dot2_acc(dl.ds, dl.dc, w, rni);
            }
            g = block_comp2(dl, shc);
            const float gNew = g.gs + g.gc, dNew = g.ds + g.dc;
            const float bNew = gNew / gamma;
            alpha = gNew / (dNew - bNew * gNew / alpha);
            beta = bNew;
            gamma = gNew;
            ++k;
            if (gamma < gBest) {
                gBest = gamma;
                kBest = k;
                if (own) xBest = lx[t];
            }
        }
        total += k;
        ++outer;
        if (kBest < k) {  // back to the iterate of the smallest residual
            gamma = gBest;
            if (own) lx[t] = xBest;
            __syncthreads();
        }
        // outer residual in FP64
        loc = 0;
        if (own) {
            double acc = b[i];
            opaque(vs);
            opaque(ns);
            acc = row_product_n<27>(0x7ffffffu, i, M, vs, ns, acc, [&](double a, float v, int j) {
                return a - double(v) * double(lx[j - rb]);
            });
            lro[t] = float(acc);
            loc = acc * acc;
        }
        rr = block_sum_double(loc, shd);
        rel = bb > 0 ? sqrt(rr / bb) : 0.0;
        fromB = false;
        if (kBest == 0) break;  // no smaller residual: another outer step would repeat this one
    }
    if (own) x[i] = lx[t];
    if (threadIdx.x == 0) {
        iters[G.depth] = total;
        iters[PSR_MAX_DEPTH + 1 + G.depth] = outer;
        resid[G.depth] = sqrt(double(gamma));
        trueRel[G.depth] = rel;
    }
}

// The mixed form with several rows per thread (large depths): vectors in
// memory, the stencil read every sweep.  Instantiated twice: the CG kernel
// for definite depths (kMinres false) and the MINRES kernel for the depths
// whose operator is indefinite (kMinres true), each compiled without the
// other's path so that each keeps its own register budget.
template <bool kMinres>
__device__ __forceinline__ void cg_mixed_generic_body(int psr_bid, int psr_nbid, 
    const CgParams& P, const int* __restrict__ neigh, const float* __restrict__ val, const float* __restrict__ b,
    float* x, float* rOut, float* rbuf, float* wbuf, float* sbuf, float* __restrict__ pvec, double* partialD,
    Comp2* partialC, GroupCtl* ctlAll, int* __restrict__ iters, double* __restrict__ resid,
    double* __restrict__ trueRel, int* errFlag, const unsigned* __restrict__ slotMasks) {
    __shared__ double shd[kCgWaves];
    __shared__ Comp2 shc[kCgWaves];
    __shared__ int shFlag;
    __shared__ double shResD;
    __shared__ Comp2 shResC;
    int gi = 0;
    while (gi + 1 < P.ngroups && int(psr_bid) >= P.g[gi + 1].firstWg) ++gi;
    const CgGroup G = P.g[gi];
    const int local = psr_bid - G.firstWg;
    const int r0 = G.rowBegin + local * G.rowsPerWg;
    const int r1 = min(G.rowBegin + G.rows, r0 + G.rowsPerWg);
    const int M = P.M;
    double* partD = partialD + 2 * G.firstWg;
    Comp2* partC = partialC + 2 * G.firstWg;
    GroupCtl* ctl = ctlAll + gi;
    unsigned phase = 0;
    const bool multi = G.wgs > 1;
    if constexpr (!kMinres) {
        if (!multi && G.rows <= kCgThreads) {
            cg_small_mixed(P, G, neigh, val, b, x, iters, resid, trueRel, shd, shc);
            return;
        }
    }
    const unsigned slotMask = slotMasks[G.depth];

    auto sum_double = [&](double v, double& out) -> bool {
        double s = block_sum_double(v, shd);
        if (!multi) {
            out = s;
            return true;
        }
        return group_allreduce(s, out, partD, ctl, phase, G.wgs, local, P.spinLimit, errFlag, &shFlag, &shResD,
                               [&](const double* parts, int n) { return fold_double(parts, n, shd); });
    };
    auto sum_comp = [&](Comp2 v, Comp2& out) -> bool {
        Comp2 s = block_comp2(v, shc);
        if (!multi) {
            out = s;
            return true;
        }
        return group_allreduce(s, out, partC, ctl, phase, G.wgs, local, P.spinLimit, errFlag, &shFlag, &shResC,
                               [&](const Comp2* parts, int n) {
                                   Comp2 t{0.f, 0.f, 0.f, 0.f};
                                   for (int k = threadIdx.x; k < n; k += kCgThreads) comp2_merge(t, ld_agent(&parts[k]));
                                   return block_comp2(t, shc);
                               });
    };

    // x = 0; ||b||^2 in FP64 (the outer residual of x = 0 is b itself)
    double loc = 0;
    for (int i = r0 + threadIdx.x; i < r1; i += kCgThreads) {
        st_shared(&x[i], 0.f);
        double bi = b[i];
        loc += bi * bi;
    }
    double bb;
    if (!sum_double(loc, bb)) return;
    double rr = bb, rel = bb > 0 ? 1.0 : 0.0;
    const float* rhs = b;
    int total = 0, outer = 0;
    float gamma = 0.f;
    bool stalled = false;
    while (rel > P.relTol && outer < P.maxOuter && total < P.maxIter) {
        // inner target on the updated residual: half the tolerance left, no tighter than FP32 can follow
        const double tgt = fmax(0.5 * P.relTol * sqrt(bb), 1e-6 * sqrt(rr));
        const float tgt2 = float(tgt * tgt);
        if constexpr (kMinres) {
            // MINRES (Paige and Saunders) on A e = rhs, e_0 = 0, FP32 with
            // compensated sums.  Sweep k forms y_{k-1} = y'_{k-1} - c_{k-1} y_{k-2}
            // on the fly (c = alpha / beta), stores it, and computes
            // y'_k = A y_{k-1} / beta_k - (beta_k / beta_{k-1}) y_{k-2} with the
            // sums v_k . y'_k (= alpha_k) and ||y'_k||^2; v_k is a unit vector,
            // so beta_{k+1}^2 = ||y'_k||^2 - alpha_k^2 and one group reduction
            // per iteration suffices.  Iteration k's rotation is known only
            // after its reduction: its direction w_k = (v_k - eps w_{k-2} -
            // delta w_{k-1}) / gamma and the update e += phi w_k are applied in
            // the next sweep, and after the last one.
            float* Pb0 = rbuf;      // y'_k, k even
            float* Pb1 = rbuf + M;  // y'_k, k odd
            float* Qb0 = wbuf;      // y_k by k mod 3
            float* Qb1 = wbuf + M;
            float* Qb2 = sbuf;
            float* W1 = sbuf + M;   // w_{k-2}, own rows only
            float* W2 = pvec;       // w_{k-1}, own rows only
            auto Qsel = [&](int k) -> float* { const int m = k % 3; return m == 0 ? Qb0 : (m == 1 ? Qb1 : Qb2); };
            const float beta1 = float(sqrt(rr));
            const float tgtf = float(tgt);
            float beta = beta1, oldb = 1.f, cPrev = 0.f;
            float cs = -1.f, sn = 0.f, dbar = 0.f, epsln = 0.f, phibar = beta1;
            float lagEps = 0.f, lagDelta = 0.f, lagGamma = 1.f, lagPhi = 0.f;
            int it = 1, kdone = 0;
            while (beta1 > tgtf && total < P.maxIter) {
                const float* Pprev = (it & 1) ? Pb0 : Pb1;  // y'_{it-1}
                float* Pnew = (it & 1) ? Pb1 : Pb0;         // y'_it
                const float* Qold = Qsel(it + 1);           // y_{it-2}
                float* Qcur = Qsel(it + 2);                 // y_{it-1}
                const float invB = 1.f / beta;
                const float r1c = it >= 2 ? beta / oldb : 0.f;
                const float invOld = 1.f / oldb;
                Comp2 dl{0.f, 0.f, 0.f, 0.f};
                for (int i = r0 + threadIdx.x; i < r1; i += kCgThreads) {
                    float av = 0.f;
                    if (it == 1)
                        av = row_product(slotMask, i, M, val, neigh, av, [&](float a, float v, int j) {
                            return __fmaf_rn(v, ld_shared(&rhs[j]), a);
                        });
                    else
                        av = row_product(slotMask, i, M, val, neigh, av, [&](float a, float v, int j) {
                            return __fmaf_rn(v, __fmaf_rn(-cPrev, ld_shared(&Qold[j]), ld_shared(&Pprev[j])), a);
                        });
                    const float r1i = it >= 2 ? ld_shared(&Qold[i]) : 0.f;  // y_{it-2}
                    const float yi = it == 1 ? ld_shared(&rhs[i]) : __fmaf_rn(-cPrev, r1i, ld_shared(&Pprev[i]));
                    st_shared(&Qcur[i], yi);
                    const float ypi = __fmaf_rn(-r1c, r1i, av * invB);
                    st_shared(&Pnew[i], ypi);
                    dot2_acc(dl.gs, dl.gc, yi * invB, ypi);
                    dot2_acc(dl.ds, dl.dc, ypi, ypi);
                    float w1 = 0.f, w2 = 0.f;
                    if (it >= 2) {
                        w1 = W1[i];
                        w2 = W2[i];
                        const float wk = (r1i * invOld - lagEps * w1 - lagDelta * w2) / lagGamma;
                        st_shared(&x[i], __fmaf_rn(lagPhi, wk, ld_shared(&x[i])));
                        w1 = w2;
                        w2 = wk;
                    }
                    W1[i] = w1;
                    W2[i] = w2;
                }
                Comp2 g;
                if (!sum_comp(dl, g)) return;
                const float alfa = g.gs + g.gc, yy = g.ds + g.dc;
                const float betaN = sqrtf(fmaxf(yy - alfa * alfa, 0.f));
                const float oldeps = epsln;
                const float delta = cs * dbar + sn * alfa;
                const float gbar = sn * dbar - cs * alfa;
                epsln = sn * betaN;
                dbar = -cs * betaN;
                const float gam = fmaxf(sqrtf(gbar * gbar + betaN * betaN), 1e-30f);
                cs = gbar / gam;
                sn = betaN / gam;
                const float phi = cs * phibar;
                phibar = sn * phibar;
                lagEps = oldeps;
                lagDelta = delta;
                lagGamma = gam;
                lagPhi = phi;
                cPrev = alfa / beta;
                oldb = beta;
                beta = betaN;
                ++kdone;
                gamma = phibar * phibar;
                if (!(phibar > tgtf) || total + kdone >= P.maxIter || !(betaN > 0.f)) break;
                ++it;
            }
            if (kdone > 0) {
                // the last iteration's update: v = y_{it-1} / beta_it
                const float* Qcur = Qsel(it + 2);
                const float invOld = 1.f / oldb;
                for (int i = r0 + threadIdx.x; i < r1; i += kCgThreads) {
                    const float wk = (ld_shared(&Qcur[i]) * invOld - lagEps * W1[i] - lagDelta * W2[i]) / lagGamma;
                    st_shared(&x[i], __fmaf_rn(lagPhi, wk, ld_shared(&x[i])));
                }
            }
            total += kdone;
            ++outer;
            if (kdone == 0) break;
            // the last update wrote x after the last group reduction: one more
            // reduction orders it before the outer residual reads neighbours' x
            double fence;
            if (!sum_double(0.0, fence)) return;
        } else {
            // r_0 = rhs, w_0 = A r_0, s_{-1} = 0, p_{-1} = 0
            Comp2 d0{0.f, 0.f, 0.f, 0.f};
            for (int i = r0 + threadIdx.x; i < r1; i += kCgThreads) {
                float w = 0.f;
                w = row_product(slotMask, i, M, val, neigh, w, [&](float a, float v, int j) {
                    return __fmaf_rn(v, ld_shared(&rhs[j]), a);
                });
                float ri = ld_shared(&rhs[i]);
                st_shared(&rbuf[i], ri);
                st_shared(&wbuf[i], w);
                st_shared(&sbuf[M + i], 0.f);
                pvec[i] = 0.f;
                dot2_acc(d0.gs, d0.gc, ri, ri);
                dot2_acc(d0.ds, d0.dc, w, ri);
            }
            Comp2 g;
            if (!sum_comp(d0, g)) return;
            gamma = g.gs + g.gc;
            float delta = g.ds + g.dc;
            float alpha = gamma / delta, beta = 0.f;
            // the iterate of the smallest residual goes to rOut (free until
            // the outer residual) in the sweep after its reduction found it
            int k = 0, kBest = 0;
            float gBest = gamma;
            bool curBest = true;
            while (gamma > tgt2 && total + k < P.maxIter && k - kBest < kCgStall) {
                const float* rk = rbuf + (k & 1) * M;
                const float* wk = wbuf + (k & 1) * M;
                const float* sp = sbuf + ((k + 1) & 1) * M;  // s_{k-1}
                float* rn = rbuf + ((k + 1) & 1) * M;
                float* wn = wbuf + ((k + 1) & 1) * M;
                float* sk = sbuf + (k & 1) * M;
                Comp2 dl{0.f, 0.f, 0.f, 0.f};
                for (int i = r0 + threadIdx.x; i < r1; i += kCgThreads) {
                    float w = 0.f;
                    w = row_product(slotMask, i, M, val, neigh, w, [&](float a, float v, int j) {
                        const float rj = r_next(ld_shared(&rk[j]), alpha, s_next(ld_shared(&wk[j]), beta, ld_shared(&sp[j])));
                        return __fmaf_rn(v, rj, a);
                    });
                    float ri = ld_shared(&rk[i]);
                    float si = s_next(ld_shared(&wk[i]), beta, ld_shared(&sp[i]));
                    float rni = r_next(ri, alpha, si);
                    float pi = __fmaf_rn(beta, pvec[i], ri);
                    pvec[i] = pi;
                    const float xo = ld_shared(&x[i]);
                    if (curBest) st_shared(&rOut[i], xo);
                    st_shared(&x[i], __fmaf_rn(alpha, pi, xo));
                    st_shared(&sk[i], si);
                    st_shared(&rn[i], rni);
                    st_shared(&wn[i], w);
                    dot2_acc(dl.gs, dl.gc, rni, rni);
                    dot2_acc(dl.ds, dl.dc, w, rni);
                }
                if (!sum_comp(dl, g)) return;
                const float gNew = g.gs + g.gc, dNew = g.ds + g.dc;
                const float bNew = gNew / gamma;
                alpha = gNew / (dNew - bNew * gNew / alpha);
                beta = bNew;
                gamma = gNew;
                ++k;
                curBest = gamma < gBest;
                if (curBest) {
                    gBest = gamma;
                    kBest = k;
                }
            }
            total += k;
            ++outer;
            if (!curBest) {
                // back to the iterate of the smallest residual (kept in rOut);
                // one more reduction orders it before the outer residual
                gamma = gBest;
                for (int i = r0 + threadIdx.x; i < r1; i += kCgThreads) st_shared(&x[i], ld_shared(&rOut[i]));
                double fence;
                if (!sum_double(0.0, fence)) return;
            }
            stalled = kBest == 0;
        }
        // outer residual in FP64
        loc = 0;
        for (int i = r0 + threadIdx.x; i < r1; i += kCgThreads) {
            double acc = b[i];
            acc = row_product(slotMask, i, M, val, neigh, acc, [&](double a, float v, int j) {
                return a - double(v) * double(ld_shared(&x[j]));
            });
            st_shared(&rOut[i], float(acc));
            loc += acc * acc;
        }
        if (!sum_double(loc, rr)) return;
        rel = bb > 0 ? sqrt(rr / bb) : 0.0;
        rhs = rOut;
        if (stalled) break;  // no smaller residual: another outer step would repeat this one
    }
    if (local == 0 && threadIdx.x == 0) {
        iters[G.depth] = total;
        iters[PSR_MAX_DEPTH + 1 + G.depth] = outer;
        resid[G.depth] = sqrt(double(gamma));
        trueRel[G.depth] = rel;
    }
}
template <class... A>
__device__ __forceinline__ void cg_mixed_all_depths_body(int bid, int nbid, const A&... a) {
    cg_mixed_generic_body<false>(bid, nbid, a...);
}
template <class... A>
__device__ __forceinline__ void cg_mixed_minres_body(int bid, int nbid, const A&... a) {
    cg_mixed_generic_body<true>(bid, nbid, a...);
}
__global__ void __launch_bounds__(kCgThreads) __attribute__((amdgpu_waves_per_eu(kCgMixedGenericWaves))) cg_mixed_all_depths(
    CgParams P, const int* __restrict__ neigh, const float* __restrict__ val, const float* __restrict__ b,
    float* x, float* rOut, float* rbuf, float* wbuf, float* sbuf, float* __restrict__ pvec, double* partialD,
    Comp2* partialC, GroupCtl* ctlAll, int* __restrict__ iters, double* __restrict__ resid,
    double* __restrict__ trueRel, int* errFlag, const unsigned* __restrict__ slotMasks) {
    cg_mixed_all_depths_body(blockIdx.x, gridDim.x, P, neigh, val, b, x, rOut, rbuf, wbuf, sbuf, pvec, partialD, partialC, ctlAll, iters, resid, trueRel, errFlag, slotMasks);
}
PSR_BATCHABLE_WAVES(cg_mixed_all_depths, kCgThreads, kCgMixedGenericWaves)
__global__ void __launch_bounds__(kCgThreads) __attribute__((amdgpu_waves_per_eu(kCgMinresWaves))) cg_mixed_minres(
    CgParams P, const int* __restrict__ neigh, const float* __restrict__ val, const float* __restrict__ b,
    float* x, float* rOut, float* rbuf, float* wbuf, float* sbuf, float* __restrict__ pvec, double* partialD,
    Comp2* partialC, GroupCtl* ctlAll, int* __restrict__ iters, double* __restrict__ resid,
    double* __restrict__ trueRel, int* errFlag, const unsigned* __restrict__ slotMasks) {
    cg_mixed_minres_body(blockIdx.x, gridDim.x, P, neigh, val, b, x, rOut, rbuf, wbuf, sbuf, pvec, partialD, partialC, ctlAll, iters, resid, trueRel, errFlag, slotMasks);
}
PSR_BATCHABLE_WAVES(cg_mixed_minres, kCgThreads, kCgMinresWaves)


// --- implicit function at points (LaplacianSolver.cu:85-118, 269-305) ------
// Neighbour columns: d_x of neighbour i of every node, and, when the
// level-local child 0 rule changed a parent, its basis parameters as well, so
// the 27 terms of a node need no dependent index loads.  (The kernels of a
// regular octree take the centres from the node grid; the few that need an
// ancestor's neighbour centres read them through the neighbour table.)
// Also clears the nZero words at zero (the mesh stage's node tables and face
// marks) and, when the octree is regular, the nDense dense d_x grid cells,
// which saves the mesh stage launches of its own.
__device__ __forceinline__ void neighbour_columns_body(int psr_bid, int psr_nbid, int M, const int* __restrict__ neigh,
                                                              const float4* __restrict__ basis,
                                                              const float* __restrict__ dx,
                                                              float4* __restrict__ nbrCol, float* __restrict__ dCol,
                                                              unsigned* __restrict__ zero, size_t nZero,
                                                              float* __restrict__ dense, size_t nDense,
                                                              const int* __restrict__ q1Fired,
                                                              unsigned* __restrict__ ctl, size_t nCtl,
                                                              unsigned* __restrict__ slots, size_t nSlots) {
    const size_t t = size_t(psr_bid) * kThreads + threadIdx.x;
    for (size_t i = t; i < nZero; i += size_t(psr_nbid) * kThreads) zero[i] = 0u;
    // the CG's group controls and slots, used: zero again for the next solve
    if (t < nCtl) ctl[t] = 0u;
    if (t < nSlots) slots[t] = 0u;
    if (nDense && *q1Fired == 0)
        for (size_t i = t; i < nDense; i += size_t(psr_nbid) * kThreads) dense[i] = 0.f;
    if (t >= size_t(27) * M) return;
    const bool cols = *q1Fired != 0;
    const int nb = neigh[t];
    if (nb < 0) {
        if (cols) nbrCol[t] = make_float4(0.f, 0.f, 0.f, 0.f);
        dCol[t] = 0.f;
    } else {
        const float d = dx[nb];
        if (cols) {
            const float4 b = basis[nb];
            nbrCol[t] = make_float4(b.x, b.y, b.z, d);
        }
        dCol[t] = d;
    }
}
__global__ void __launch_bounds__(kThreads) neighbour_columns(int M, const int* __restrict__ neigh,
                                                              const float4* __restrict__ basis,
                                                              const float* __restrict__ dx,
                                                              float4* __restrict__ nbrCol, float* __restrict__ dCol,
                                                              unsigned* __restrict__ zero, size_t nZero,
                                                              float* __restrict__ dense, size_t nDense,
                                                              const int* __restrict__ q1Fired,
                                                              unsigned* __restrict__ ctl, size_t nCtl,
                                                              unsigned* __restrict__ slots, size_t nSlots) {
    neighbour_columns_body(blockIdx.x, gridDim.x, M, neigh, basis, dx, nbrCol, dCol, zero, nZero, dense, nDense, q1Fired, ctl, nCtl, slots, nSlots);
}
PSR_BATCHABLE(neighbour_columns, kThreads)



// stored (non-dropped) operator entries per depth
__device__ __forceinline__ void zero_u32_body(int psr_bid, int psr_nbid, unsigned* a, int n) {
    int i = psr_bid * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0u;
}
__global__ void __launch_bounds__(256) zero_u32(unsigned* a, int n) {
    zero_u32_body(blockIdx.x, gridDim.x, a, n);
}
PSR_BATCHABLE(zero_u32, 256)


}  // namespace

std::mutex& cg_order_lock() {
    static std::mutex m;
    return m;
}

// Resident-workgroup accounting of the CG launches in flight on the device
// (one GPU per process; guarded by cg_order_lock): a launch holds its share of
// the device's resident workgroups, in units of kCgUnits for the whole device,
// until its completion event (its context's CG end event) has fired.
struct CgInflight {
    hipEvent_t done;
    long long units;
};
constexpr long long kCgUnits = 1ll << 20;
std::vector<CgInflight>& cg_inflight() {
    static std::vector<CgInflight> v;
    return v;
}

int cg_batch_budget(Ctx& c) {
    const void* fn = c.cfg.precision == PSR_PRECISION_MIXED
                         ? multi_queue_kernel_of<cg_mixed_all_depths_k>(cg_mixed_all_depths)
                         : multi_queue_kernel_of<cg_all_depths_k>(cg_all_depths);
    int per = 0, cus = 0;
    PSR_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, kCgThreads, 0));
    PSR_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device));
    return std::max(1, per * cus);
}

void cg_admit_device(hipEvent_t self) {
    std::unique_lock<std::mutex> order(cg_order_lock());
    for (;;) {
        auto& v = cg_inflight();
        v.erase(std::remove_if(v.begin(), v.end(),
                               [&](const CgInflight& f) { return f.done == self || hipEventQuery(f.done) == hipSuccess; }),
                v.end());
        if (v.empty()) return;
        order.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(20));
        order.lock();
    }
}

void cg_hold_device(hipEvent_t done) {
    std::lock_guard<std::mutex> order(cg_order_lock());
    cg_inflight().push_back(CgInflight{done, kCgUnits});
}

void cg_forget_event(hipEvent_t e) {
    std::lock_guard<std::mutex> order(cg_order_lock());
    auto& v = cg_inflight();
    v.erase(std::remove_if(v.begin(), v.end(), [&](const CgInflight& f) { return f.done == e; }), v.end());
}

void solve_depths(Ctx& c) {
    Octree& o = c.oct;
    const Levels& L = o.L;
    const int D = L.D, M = o.M;
    hipStream_t s = c.stream;
    float* val = c.buf<float>("lap_val", size_t(27) * M);
    DevResults* res = c.results();
    int* nnz = res->cgNnz;
    c.fld.dx = c.buf<float>("dx", M);
    if (size_t(M) * 2 * sizeof(float4) > size_t(INT32_MAX)) throw std::runtime_error("CG state exceeds 2 GiB");
    CgParams P{};
    P.M = M;
    P.tolSq = c.cfg.cg_tol * c.cfg.cg_tol;
    P.maxIter = c.cfg.cg_max_iter;
    P.spinLimit = 1u << 26;  // row form: polls of about a microsecond each, set below
    const bool mixed = c.cfg.precision == PSR_PRECISION_MIXED;
    const void* rowKernel = mixed ? reinterpret_cast<const void*>(cg_mixed_row) : reinterpret_cast<const void*>(cg_fp32_row);
    const void* genKernel =
        mixed ? reinterpret_cast<const void*>(cg_mixed_all_depths) : reinterpret_cast<const void*>(cg_all_depths);
    // workgroup budgets: every workgroup of a CG launch must be resident at once
    // (the groups meet in barriers).  The launch is an ordinary one: the budget,
    // with the accounting of concurrent launches below, is the residency
    // guarantee, and the barriers' spin bound turns a broken one into an error.
    // A batch merges the clouds' launches into the kernels' batched forms,
    // which the compiler may give other register counts: a budget is the
    // smaller occupancy of the two forms, so that a plan made for one cloud
    // also fits the merged launch.
    if (c.cgMaxWgs <= 0) {
        auto occupancy = [](const void* single, const void* batched) {
            int a = 0, b = 0;
            PSR_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, single, kCgThreads, 0));
            PSR_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, batched, kCgThreads, 0));
            return std::min(a, b);
        };
        const void* rowQueue = mixed ? multi_queue_kernel_of<cg_mixed_row_k>(cg_mixed_row)
                                     : multi_queue_kernel_of<cg_fp32_row_k>(cg_fp32_row);
        const void* genQueue = mixed ? multi_queue_kernel_of<cg_mixed_all_depths_k>(cg_mixed_all_depths)
                                     : multi_queue_kernel_of<cg_all_depths_k>(cg_all_depths);
        int cus = 0;
        const int perRow = occupancy(rowKernel, rowQueue), perGen = occupancy(genKernel, genQueue);
        const int perMinres = occupancy(reinterpret_cast<const void*>(cg_mixed_minres),
                                        multi_queue_kernel_of<cg_mixed_minres_k>(cg_mixed_minres));
        PSR_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device));
        c.cgMaxWgsRow = std::max(0, perRow * cus);
        c.cgMaxWgs = std::max(1, perGen * cus);
        c.cgMaxWgsMinres = std::max(1, perMinres * cus);
    }
    // the mixed form solves indefinite depths with MINRES, in a launch of its
    // own (the MINRES kernel) after the CG launch of the other depths
    bool isMinres[PSR_MAX_DEPTH + 1];
    bool anyIndef = false;
    for (int d = 0; d <= D; ++d) anyIndef |= (isMinres[d] = mixed && c.indefinite[d]);
    auto need_of = [&](int rpw, bool minresSet) {
        long long w = 0;
        for (int d = 0; d <= D; ++d)
            if (isMinres[d] == minresSet) w += std::max<long long>(1, (L.count[d] + rpw - 1) / rpw);
        return w;
    };
    auto need = [&](int rpw) { return need_of(rpw, false) + need_of(rpw, true); };
    // one row per thread (stencil in registers) when every group fits at
    // kCgThreads rows per workgroup; otherwise widen the row blocks
    int rowsPerWg = c.cgRowsPerWg > 0 ? c.cgRowsPerWg : kCgThreads;
    bool rowForm = rowsPerWg <= kCgThreads && need(rowsPerWg) <= c.cgMaxWgsRow;
    for (int d = 0; d <= D; ++d)
        if (div_up(L.count[d], kCgThreads) > unsigned(kPollPerThread * kCgThreads)) rowForm = false;
    // MINRES runs in the several-rows-per-thread form only
    if (anyIndef) rowForm = false;
    int rowsPerWgMinres = rowsPerWg;
    if (!rowForm) {
        while (need_of(rowsPerWg, false) > c.cgMaxWgs) rowsPerWg *= 2;
        while (need_of(rowsPerWgMinres, true) > c.cgMaxWgsMinres) rowsPerWgMinres *= 2;
    }
    // A batch sets the rows a workgroup takes instead (cgRowTarget, from the
    // batch's total rows and the resident budget): a depth of n rows gets
    // ceil(n / target) workgroups of equal shares; one row per thread when
    // every share fits 256 rows.
    int capRows[PSR_MAX_DEPTH + 1];
    if (c.cgRowTarget > 0) {
        rowForm = true;
        for (int d = 0; d <= D; ++d) {
            const int w = std::max(1, int(div_up(L.count[d], unsigned(c.cgRowTarget))));
            capRows[d] = std::max(1, int(div_up(L.count[d], unsigned(w))));
            if (capRows[d] > kCgThreads) rowForm = false;
        }
        if (anyIndef) rowForm = false;
    }
    // largest depths first so their groups start on separate CUs early; the
    // MINRES depths form a second launch (PM) with its own workgroup numbering
    CgParams PM = P;
    int wg = 0, wgM = 0;
    for (int d = D; d >= 0; --d) {
        CgParams& Q = isMinres[d] ? PM : P;
        int& w = isMinres[d] ? wgM : wg;
        CgGroup& g = Q.g[Q.ngroups++];
        g.depth = d;
        g.rowBegin = L.base[d];
        g.rows = L.count[d];
        g.firstWg = w;
        g.rowsPerWg = c.cgRowTarget > 0 ? capRows[d] : (isMinres[d] ? rowsPerWgMinres : rowsPerWg);
        g.wgs = std::max(1, int(div_up(g.rows, g.rowsPerWg)));
        g.indef = isMinres[d] ? 1 : 0;
        w += g.wgs;
    }
    if (wg > (rowForm ? c.cgMaxWgsRow : c.cgMaxWgs) || wgM > c.cgMaxWgsMinres)
        throw std::runtime_error("CG launch exceeds the resident workgroup budget");
    c.stats.cg_workgroups = uint32_t(wg + wgM);
    double* partial = c.buf<double>("cg_partial", size_t(2) * (wg + wgM));
    GroupCtl* ctl = c.buf<GroupCtl>("cg_ctl", PSR_MAX_DEPTH + 1);
    const int ctlWords = int(sizeof(GroupCtl) / 4) * (PSR_MAX_DEPTH + 1);
    int* iters = res->cgIters;  // [inner iterations | outer steps] per depth
    double* resid = res->cgResid;
    double* trueRel = res->cgTrueRel;
    int* err = &res->cgErr;
    float4* slots = rowForm && mixed ? c.buf<float4>("cg_slots", size_t(4) * wg) : nullptr;
    // The group controls and slots start at zero: the neighbour-column kernel
    // clears them after every solve, so they need a launch of their own only
    // when the last call on this context did not complete or the slots moved.
    // The CG writes every depth's results itself; its error flag is cleared
    // with the operator counts (by the vector-field kernel).
    const size_t slotWords = slots ? size_t(16) * wg : 0;
    if (!c.cgClean || slots != c.cgSlots || slotWords != c.cgSlotWords)
        zero_spans({{ctl, size_t(ctlWords)}, {slots, slotWords}}, s);
    c.cgClean = false;
    c.cgCtl = ctl;
    c.cgCtlWords = size_t(ctlWords);
    c.cgSlots = slots;
    c.cgSlotWords = slotWords;
    // CG kernels of different contexts (streams) never compete for residency
    // while spinning in their group barriers: a launch is queued only once the
    // launches still in flight leave room for all its workgroups beside them
    // (the host waits otherwise), so concurrent contexts -- the batch workers --
    // run their solves side by side whenever they fit.
    // the two launches of a split solve run one after the other on the stream:
    // the account holds the larger
    const long long cap = rowForm ? c.cgMaxWgsRow : c.cgMaxWgs;
    const long long units = std::max((wg * kCgUnits + cap - 1) / cap,
                                     (wgM * kCgUnits + c.cgMaxWgsMinres - 1) / c.cgMaxWgsMinres);
    // While a batch records, the batch merges every cloud's solve into launches
    // that fit the resident budget and accounts for them itself.
    const bool recording = current_recorder() != nullptr;
    if (!recording) {
        std::unique_lock<std::mutex> order(cg_order_lock());
        for (;;) {
            auto& v = cg_inflight();
            // this context's earlier launch is ahead of this one in its stream
            v.erase(std::remove_if(v.begin(), v.end(),
                                   [&](const CgInflight& f) {
                                       return f.done == c.ev[9] || hipEventQuery(f.done) == hipSuccess;
                                   }),
                    v.end());
            long long used = 0;
            for (const CgInflight& f : v) used += f.units;
            if (used + units <= kCgUnits) break;
            order.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(20));
            order.lock();
        }
        PSR_HIP_CHECK(hipEventRecord(c.ev[8], s));
    }
    if (rowForm && mixed) P.spinLimit = 1u << 22;
    if (c.cgSpinLimit >= 0) P.spinLimit = PM.spinLimit = unsigned(c.cgSpinLimit);
    static const char siteMixedRow = 0, siteMixedGen = 0, siteMinres = 0, siteRow = 0, siteGen = 0;
    if (mixed) {
        P.relTol = PM.relTol = c.cfg.refine_rel_tol;
        P.maxOuter = PM.maxOuter = std::max(1, c.cfg.refine_max_outer);
        float* rOut = c.buf<float>("cg_rout", M);
        Comp2* partialC = c.buf<Comp2>("cg_partial_c", size_t(2) * (wg + wgM));
        if (rowForm) {
            float3* state = c.buf<float3>("cg_state3", size_t(2) * M);
            launch_kernel<cg_mixed_row_k>(&siteMixedRow, cg_mixed_row, wg, 0, s, true, P, o.neigh, val, c.fld.div,
                                          c.fld.dx, rOut, state, slots, iters, resid, trueRel, err);
        } else {
            float* rbuf = c.buf<float>("cg_rbuf", size_t(2) * M);
            float* wbuf = c.buf<float>("cg_wbuf", size_t(2) * M);
            float* sbuf = c.buf<float>("cg_sbuf", size_t(2) * M);
            float* pvec = c.buf<float>("cg_pa", M);
            if (P.ngroups > 0)
                launch_kernel<cg_mixed_all_depths_k>(&siteMixedGen, cg_mixed_all_depths, wg, 0, s, true, P, o.neigh,
                                                     val, c.fld.div, c.fld.dx, rOut, rbuf, wbuf, sbuf, pvec, partial,
                                                     partialC, ctl, iters, resid, trueRel, err, res->cgSlotMask);
            // the MINRES launch uses the group controls and partials after the CG launch's
            if (PM.ngroups > 0)
                launch_kernel<cg_mixed_minres_k>(&siteMinres, cg_mixed_minres, wgM, 0, s, true, PM, o.neigh, val,
                                                 c.fld.div, c.fld.dx, rOut, rbuf, wbuf, sbuf, pvec,
                                                 partial + size_t(2) * wg, partialC + size_t(2) * wg,
                                                 ctl + P.ngroups, iters, resid, trueRel, err, res->cgSlotMask);
        }
    } else if (rowForm) {
        float2* state = c.buf<float2>("cg_state2", size_t(2) * M);
        launch_kernel<cg_fp32_row_k>(&siteRow, cg_fp32_row, wg, 0, s, true, P, o.neigh, val, c.fld.div, c.fld.dx,
                                     state, partial, ctl, iters, resid, err);
    } else {
        float* r = c.buf<float>("cg_r", M);
        float* pA = c.buf<float>("cg_pa", M);
        float* pB = c.buf<float>("cg_pb", M);
        float* Ap = c.buf<float>("cg_ap", M);
        launch_kernel<cg_all_depths_k>(&siteGen, cg_all_depths, wg, 0, s, true, P, o.neigh, val, c.fld.div, c.fld.dx, r,
                                       pA, pB, Ap, partial, ctl, iters, resid, err, res->cgSlotMask);
    }
    c.stats.cg_row_form = rowForm ? 1 : 0;
    if (!recording) {
        PSR_HIP_CHECK(hipEventRecord(c.ev[9], s));
        std::lock_guard<std::mutex> order(cg_order_lock());
        cg_inflight().push_back(CgInflight{c.ev[9], units});
    }
    PSR_HIP_CHECK(hipGetLastError());
}

void points_implicit(Ctx& c) {
    Octree& o = c.oct;
    const Levels& L = o.L;
    hipStream_t s = c.stream;
    int N = o.N, nb = int(div_up(N, kPointsChiThreads));
    c.fld.chiPts = c.buf<float>("chi_points", N);
    double* part = c.buf<double>("iso_part", nb);
    c.fld.isoSum = c.buf<double>("iso_sum", 1);
    c.fld.nbrCol = c.buf<float4>("nbr_col", size_t(27) * o.M);
    c.fld.dCol = c.buf<float>("nbr_dcol", size_t(27) * o.M);
    // the mesh stage's node tables and face marks, cleared with the columns:
    // nodeVert [8][M], nodeFace [6][M], nodeEdge [12][n_D], face marks (6M, the face bound)
    const size_t M = size_t(o.M), nD = size_t(L.count[L.D]), nZero = 8 * M + 6 * M + 12 * nD + 6 * M;
    int* zeroed = c.buf<int>("mesh_zeroed", nZero);
    c.geo.nodeVert = zeroed;
    c.geo.nodeFace = zeroed + 8 * M;
    c.geo.nodeEdge = zeroed + 14 * M;
    c.geo.faceMarks = zeroed + 14 * M + 12 * nD;
    // dense d_x grids for the subdivision of a regular octree (filled by the
    // mesh stage's first kernel)
    const size_t nDense = L.D <= kDenseEarlyDepth ? dense_grid_cells(L.D) : 0;
    c.geo.dxGrid = nDense ? c.buf<float>("sub_dx_grid", nDense) : nullptr;
    PSR_LAUNCH(neighbour_columns, div_up(size_t(27) * o.M, kThreads), 0, s,
        o.M, o.neigh, o.basis, c.fld.dx, c.fld.nbrCol, c.fld.dCol, reinterpret_cast<unsigned*>(zeroed), nZero, c.geo.dxGrid,
        nDense, o.q1Fired, reinterpret_cast<unsigned*>(c.cgCtl), c.cgCtlWords,
        reinterpret_cast<unsigned*>(c.cgSlots), c.cgSlotWords);
    // the implicit function at the points runs beside the mesh stage's first
    // kernel (build_geometry), then the iso value
    c.fld.chiArgs = PointsChiArgs{N,        L.base[L.D], o.M,          L.D,      o.spos,        o.point2node,
                                  o.parent, c.fld.nbrCol, c.fld.dCol, o.basis, o.q1Fired, c.fld.chiPts, part};
    c.fld.chiBlocks = nb;
    PSR_HIP_CHECK(hipGetLastError());
}

}  // namespace psr
