// Per-depth Laplacian, concurrent CG over all depths, implicit function at the
// input points and the iso-value.
#include "psr_internal.h"

#include <algorithm>
#include <stdexcept>

namespace psr {

namespace {

constexpr int kThreads = 256;

// --- operator (LaplacianSolver.cu:27-74) -----------------------------------
// L(o, o') = <F,F>_x <F,F>_y <F,F>_z (<F',F'>_x + <F',F'>_y + <F',F'>_z) for the
// same-depth neighbours o' of o; |L| <= 1e-6 is dropped (stored as 0); values
// kept as float.  ELL layout [27][M] over all nodes; the column of slot i is
// neigh[i][o] - base[depth].
__global__ void __launch_bounds__(kThreads) assemble_operator(Levels L, int M, const int* __restrict__ neigh,
                                                              const int* __restrict__ enc,
                                                              const double* __restrict__ dot,
                                                              const double* __restrict__ d2dot, int res,
                                                              float* __restrict__ val) {
    int o = blockIdx.x * kThreads + threadIdx.x;
    if (o >= M) return;
    const int D = L.D;
    int3 f1 = decode_func_index(enc[o], D);
    for (int i = 0; i < 27; ++i) {
        int nb = neigh[i * M + o];
        float v = 0.f;
        if (nb != -1) {
            int3 f2 = decode_func_index(enc[nb], D);
            size_t sx = size_t(f1.x) * res + f2.x, sy = size_t(f1.y) * res + f2.y, sz = size_t(f1.z) * res + f2.z;
            double e = dot[sx] * dot[sy] * dot[sz] * (d2dot[sx] + d2dot[sy] + d2dot[sz]);
            if (fabs(e) > double(1e-6f)) v = float(e);
        }
        val[i * M + o] = v;
    }
}

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
};

struct CgParams {
    int ngroups;
    CgGroup g[PSR_MAX_DEPTH + 1];
    int M;
    float tolSq;     // reference: r1 > tol * tol with float tol (CGAlgorithm.cu:36)
    int maxIter;
    unsigned spinLimit;
};

constexpr int kCgThreads = 256;
constexpr int kCgWaves = kCgThreads / PSR_WAVE;

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

__device__ __forceinline__ double block_sum_double(double v, double* sh) {
    for (int off = PSR_WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, PSR_WAVE);
    int w = threadIdx.x / PSR_WAVE;
    if (__lane_id() == 0) sh[w] = v;
    __syncthreads();
    double s = 0;
    for (int k = 0; k < kCgWaves; ++k) s += sh[k];
    __syncthreads();
    return s;
}

// Group barrier: monotonic arrival counter per group.  Returns false if the
// spin limit was hit (then the error flag is set and the caller stops).
__device__ __forceinline__ bool group_barrier(unsigned* counter, unsigned target, unsigned spinLimit,
                                              int* errFlag, int* shFlag) {
    __builtin_amdgcn_s_waitcnt(0);  // this wave's agent-scope stores have completed
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        unsigned spins = 0;
        int ok = 1;
        while (__hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (++spins > spinLimit || __hip_atomic_load(errFlag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                ok = 0;
                __hip_atomic_store(errFlag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        *shFlag = ok;
    }
    __syncthreads();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    return *shFlag != 0;
}

__device__ __forceinline__ float p_next(float r, float beta, float p) { return r + beta * p; }

__global__ void __launch_bounds__(kCgThreads) cg_all_depths(CgParams P, const int* __restrict__ neigh,
                                                            const float* __restrict__ val,
                                                            const float* __restrict__ b, float* __restrict__ x,
                                                            float* r, float* pA, float* pB,
                                                            float* __restrict__ Ap, double* partial,
                                                            unsigned* bar, int* __restrict__ iters,
                                                            double* __restrict__ resid, int* errFlag) {
    __shared__ double sh[kCgWaves];
    __shared__ int shFlag;
    int gi = 0;
    while (gi + 1 < P.ngroups && int(blockIdx.x) >= P.g[gi + 1].firstWg) ++gi;
    const CgGroup G = P.g[gi];
    const int local = blockIdx.x - G.firstWg;
    const int r0 = G.rowBegin + local * G.rowsPerWg;
    const int r1 = min(G.rowBegin + G.rows, r0 + G.rowsPerWg);
    const int M = P.M;
    double* part = partial + 2 * G.firstWg;
    unsigned* cnt = bar + gi;
    unsigned phase = 0;
    const bool multi = G.wgs > 1;

    // Sum over the group.  Partials are double-buffered by phase parity: a
    // workgroup that runs ahead writes the other half while slower ones still
    // read this one.
    auto group_sum = [&](double v, double& out) -> bool {
        double s = block_sum_double(v, sh);
        if (!multi) {
            out = s;
            return true;
        }
        double* half = part + (phase & 1u) * unsigned(G.wgs);
        if (threadIdx.x == 0) __hip_atomic_store(&half[local], s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ++phase;
        if (!group_barrier(cnt, phase * unsigned(G.wgs), P.spinLimit, errFlag, &shFlag)) return false;
        double t = 0;
        for (int k = threadIdx.x; k < G.wgs; k += kCgThreads)
            t += __hip_atomic_load(&half[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        out = block_sum_double(t, sh);
        return true;
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
#pragma unroll
            for (int q = 0; q < 27; ++q) {
                float v = val[q * M + i];
                int j = neigh[q * M + i];
                j = (v != 0.f) ? j : i;  // dropped or missing entry: safe index, zero weight
                float pj = p_next(ld_shared(&r[j]), beta, ld_shared(&pOld[j]));
                s += v * pj;
            }
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
        resid[G.depth] = sqrt(double(rr));
    }
}

// --- implicit function at points (LaplacianSolver.cu:85-118, 269-305) ------
__device__ __forceinline__ float chi_sum_node(int node, float3 p, int M, int D, const int* __restrict__ neigh,
                                              const int* __restrict__ enc, const float* __restrict__ dx) {
    float v = 0.f;
    for (int i = 0; i < 27; ++i) {
        int nb = neigh[i * M + node];
        if (nb == -1) continue;
        int3 f = decode_func_index(enc[nb], D);
        v += dx[nb] * basis_value(f.x, p.x) * basis_value(f.y, p.y) * basis_value(f.z, p.z);
    }
    return v;
}

__global__ void __launch_bounds__(kThreads) points_chi_kernel(int N, int offD, int M, int D,
                                                              const float4* __restrict__ pos,
                                                              const int* __restrict__ p2n,
                                                              const int* __restrict__ parent,
                                                              const int* __restrict__ neigh,
                                                              const int* __restrict__ enc,
                                                              const float* __restrict__ dx,
                                                              float* __restrict__ chi, double* __restrict__ part) {
    __shared__ double sh[kThreads / PSR_WAVE];
    int i = blockIdx.x * kThreads + threadIdx.x;
    double mine = 0;
    if (i < N) {
        float4 q = pos[i];
        float3 p = make_float3(q.x, q.y, q.z);
        int now = offD + p2n[i];
        float v = 0.f;
        while (now != -1) {
            v += chi_sum_node(now, p, M, D, neigh, enc, dx);
            now = parent[now];
        }
        chi[i] = v;
        mine = double(v);
    }
    for (int off = PSR_WAVE / 2; off > 0; off >>= 1) mine += __shfl_xor(mine, off, PSR_WAVE);
    if (__lane_id() == 0) sh[threadIdx.x / PSR_WAVE] = mine;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int w = 0; w < kThreads / PSR_WAVE; ++w) s += sh[w];
        part[blockIdx.x] = s;
    }
}

__global__ void __launch_bounds__(1024) iso_final(const double* __restrict__ part, int nb, int N,
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

// stored (non-dropped) operator entries per depth
__global__ void __launch_bounds__(kThreads) count_nonzeros(Levels L, int M, const float* __restrict__ val,
                                                           int* __restrict__ nnz) {
    int o = blockIdx.x * kThreads + threadIdx.x;
    int c = 0;
    if (o < M)
        for (int q = 0; q < 27; ++q) c += val[q * M + o] != 0.f;
    int d = o < M ? depth_of(L, o) : 0;
    // rows of one wave usually share a depth: reduce per depth value in the wave
    for (int dd = 0; dd <= L.D; ++dd) {
        int v = (o < M && d == dd) ? c : 0;
        for (int off = PSR_WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, PSR_WAVE);
        if (__lane_id() == 0 && v) atomicAdd(&nnz[dd], v);
    }
}

__global__ void zero_u32(unsigned* a, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0u;
}

}  // namespace

void solve_depths(Ctx& c) {
    Octree& o = c.oct;
    const Levels& L = o.L;
    const int D = L.D, M = o.M;
    hipStream_t s = c.stream;
    float* val = c.buf<float>("lap_val", size_t(27) * M);
    assemble_operator<<<div_up(M, kThreads), kThreads, 0, s>>>(L, M, o.neigh, o.encIdx, c.tables.dot, c.tables.d2dot,
                                                              c.tables.res, val);
    int* nnz = c.buf<int>("cg_nnz", PSR_MAX_DEPTH + 1);
    zero_u32<<<1, 64, 0, s>>>(reinterpret_cast<unsigned*>(nnz), PSR_MAX_DEPTH + 1);
    count_nonzeros<<<div_up(M, kThreads), kThreads, 0, s>>>(L, M, val, nnz);
    c.fld.dx = c.buf<float>("dx", M);
    float* r = c.buf<float>("cg_r", M);
    float* pA = c.buf<float>("cg_pa", M);
    float* pB = c.buf<float>("cg_pb", M);
    float* Ap = c.buf<float>("cg_ap", M);
    CgParams P{};
    P.M = M;
    P.tolSq = c.cfg.cg_tol * c.cfg.cg_tol;
    P.maxIter = c.cfg.cg_max_iter;
    P.spinLimit = 1u << 26;
    // workgroup budget: every workgroup of the cooperative launch must be resident
    if (c.cgMaxWgs <= 0) {
        int perCu = 0, cus = 0;
        PSR_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, cg_all_depths, kCgThreads, 0));
        PSR_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device));
        c.cgMaxWgs = std::max(1, perCu * cus);
    }
    int rowsPerWg = c.cgRowsPerWg > 0 ? c.cgRowsPerWg : 512;
    {
        // widen the row blocks until every group fits the resident budget
        auto need = [&](int rpw) {
            long long w = 0;
            for (int d = 0; d <= D; ++d) w += std::max<long long>(1, (L.count[d] + rpw - 1) / rpw);
            return w;
        };
        while (need(rowsPerWg) > c.cgMaxWgs) rowsPerWg *= 2;
    }
    int wg = 0;
    // largest depths first so their groups start on separate CUs early
    for (int d = D; d >= 0; --d) {
        CgGroup& g = P.g[P.ngroups++];
        g.depth = d;
        g.rowBegin = L.base[d];
        g.rows = L.count[d];
        g.firstWg = wg;
        g.rowsPerWg = rowsPerWg;
        g.wgs = std::max(1, int(div_up(g.rows, rowsPerWg)));
        wg += g.wgs;
    }
    if (wg > c.cgMaxWgs) throw std::runtime_error("CG launch exceeds the resident workgroup budget");
    c.stats.cg_workgroups = uint32_t(wg);
    double* partial = c.buf<double>("cg_partial", size_t(2) * wg);
    unsigned* bar = c.buf<unsigned>("cg_bar", PSR_MAX_DEPTH + 1);
    int* iters = c.buf<int>("cg_iters", PSR_MAX_DEPTH + 1);
    double* resid = c.buf<double>("cg_resid", PSR_MAX_DEPTH + 1);
    int* err = c.buf<int>("cg_err", 1);
    zero_u32<<<1, 64, 0, s>>>(bar, PSR_MAX_DEPTH + 1);
    zero_u32<<<1, 64, 0, s>>>(reinterpret_cast<unsigned*>(err), 1);
    void* args[] = {&P, &o.neigh, &val, &c.fld.div, &c.fld.dx, &r, &pA, &pB, &Ap, &partial, &bar, &iters, &resid, &err};
    PSR_HIP_CHECK(hipEventRecord(c.ev[8], s));
    PSR_HIP_CHECK(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(cg_all_depths), dim3(wg), dim3(kCgThreads),
                                             args, 0, s));
    PSR_HIP_CHECK(hipEventRecord(c.ev[9], s));
    PSR_HIP_CHECK(hipGetLastError());
}

void points_implicit(Ctx& c) {
    Octree& o = c.oct;
    const Levels& L = o.L;
    hipStream_t s = c.stream;
    int N = o.N, nb = int(div_up(N, kThreads));
    c.fld.chiPts = c.buf<float>("chi_points", N);
    double* part = c.buf<double>("iso_part", nb);
    c.fld.isoSum = c.buf<double>("iso_sum", 1);
    float* isoDev = c.buf<float>("iso_value", 1);
    points_chi_kernel<<<nb, kThreads, 0, s>>>(N, L.base[L.D], o.M, L.D, o.spos, o.point2node, o.parent, o.neigh,
                                              o.encIdx, c.fld.dx, c.fld.chiPts, part);
    iso_final<<<1, 1024, 0, s>>>(part, nb, N, c.fld.isoSum, isoDev);
    PSR_HIP_CHECK(hipGetLastError());
}

}  // namespace psr
