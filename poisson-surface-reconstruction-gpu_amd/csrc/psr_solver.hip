// Per-depth Laplacian, concurrent CG over all depths, implicit function at the
// input points and the iso-value.
#include "psr_internal.h"

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

constexpr int kCgThreads = 512;
constexpr int kCgWaves = kCgThreads / PSR_WAVE;

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
    __threadfence_block();
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();  // release the workgroup's stores at device scope
        atomicAdd(counter, 1u);
        unsigned spins = 0;
        int ok = 1;
        while (__hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (++spins > spinLimit || __hip_atomic_load(errFlag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                ok = 0;
                atomicExch(errFlag, 1);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __threadfence();  // acquire: later loads see the other workgroups' stores
        *shFlag = ok;
    }
    __syncthreads();
    return *shFlag != 0;
}

__device__ __forceinline__ float p_next(float r, float beta, float p) { return r + beta * p; }

__global__ void __launch_bounds__(kCgThreads) cg_all_depths(CgParams P, const int* __restrict__ neigh,
                                                            const float* __restrict__ val,
                                                            const float* __restrict__ b, float* __restrict__ x,
                                                            float* __restrict__ r, float* __restrict__ pA,
                                                            float* __restrict__ pB, float* __restrict__ Ap,
                                                            double* __restrict__ partial, unsigned* __restrict__ bar,
                                                            int* __restrict__ iters, double* __restrict__ resid,
                                                            int* __restrict__ errFlag) {
    __shared__ double sh[kCgWaves];
    __shared__ int shFlag;
    // locate the group of this workgroup
    int gi = 0;
    while (gi + 1 < P.ngroups && int(blockIdx.x) >= P.g[gi + 1].firstWg) ++gi;
    const CgGroup G = P.g[gi];
    const int local = blockIdx.x - G.firstWg;
    const int r0 = G.rowBegin + local * G.rowsPerWg;
    const int r1 = min(G.rowBegin + G.rows, r0 + G.rowsPerWg);
    const int M = P.M, base = G.rowBegin;
    double* part = partial + 2 * G.firstWg;
    unsigned* cnt = bar + gi;
    unsigned phase = 0;
    const bool multi = G.wgs > 1;

    auto group_sum = [&](double v, double& out) -> bool {
        double s = block_sum_double(v, sh);
        if (!multi) {
            out = s;
            return true;
        }
        // partials double-buffered by phase parity: a workgroup that runs ahead
        // writes the other half while slower ones still read this one
        double* half = part + (phase & 1u) * unsigned(G.wgs);
        if (threadIdx.x == 0) half[local] = s;
        ++phase;
        if (!group_barrier(cnt, phase * unsigned(G.wgs), P.spinLimit, errFlag, &shFlag)) return false;
        double t = 0;
        for (int k = 0; k < G.wgs; ++k) t += __hip_atomic_load(&half[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        out = t;
        return true;
    };

    // x = 0, r = b, rr = r.r
    double loc = 0;
    for (int i = r0 + threadIdx.x; i < r1; i += kCgThreads) {
        float bi = b[i];
        x[i] = 0.f;
        r[i] = bi;
        pA[i] = 0.f;
        loc += double(bi * bi);
    }
    double rrD;
    if (!group_sum(loc, rrD)) return;
    float rr = float(rrD), rrOld = 1.f;
    int k = 1;
    float* pOld = pA;
    float* pNew = pB;
    while (rr > P.tolSq && k <= P.maxIter) {
        float beta = (k > 1) ? rr / rrOld : 0.f;
        // p = r + beta p; Ap = A p; pAp
        double lpAp = 0;
        for (int i = r0 + threadIdx.x; i < r1; i += kCgThreads) {
            float s = 0.f;
            for (int q = 0; q < 27; ++q) {
                float v = val[q * M + i];
                if (v != 0.f) {
                    int j = neigh[q * M + i];
                    float pj = (k > 1) ? p_next(r[j], beta, pOld[j]) : r[j];
                    s += v * pj;
                }
            }
            float pi = (k > 1) ? p_next(r[i], beta, pOld[i]) : r[i];
            pNew[i] = pi;
            Ap[i] = s;
            lpAp += double(pi * s);
        }
        double pAp;
        if (!group_sum(lpAp, pAp)) return;
        float alpha = float(double(rr) / pAp);
        double lrr = 0;
        for (int i = r0 + threadIdx.x; i < r1; i += kCgThreads) {
            float pi = pNew[i];
            x[i] = x[i] + alpha * pi;
            float ri = r[i] + (-alpha) * Ap[i];
            r[i] = ri;
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
    (void)base;
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
    const int kSingle = 4096, kRowsPerWg = 1024;
    int wg = 0;
    // largest depths first so their groups start on separate CUs early
    for (int d = D; d >= 0; --d) {
        CgGroup& g = P.g[P.ngroups++];
        g.depth = d;
        g.rowBegin = L.base[d];
        g.rows = L.count[d];
        g.firstWg = wg;
        if (g.rows <= kSingle) {
            g.wgs = 1;
            g.rowsPerWg = g.rows;
        } else {
            g.wgs = int(div_up(g.rows, kRowsPerWg));
            g.rowsPerWg = kRowsPerWg;
        }
        wg += g.wgs;
    }
    double* partial = c.buf<double>("cg_partial", size_t(2) * wg);
    unsigned* bar = c.buf<unsigned>("cg_bar", PSR_MAX_DEPTH + 1);
    int* iters = c.buf<int>("cg_iters", PSR_MAX_DEPTH + 1);
    double* resid = c.buf<double>("cg_resid", PSR_MAX_DEPTH + 1);
    int* err = c.buf<int>("cg_err", 1);
    zero_u32<<<1, 64, 0, s>>>(bar, PSR_MAX_DEPTH + 1);
    zero_u32<<<1, 64, 0, s>>>(reinterpret_cast<unsigned*>(err), 1);
    void* args[] = {&P, &o.neigh, &val, &c.fld.div, &c.fld.dx, &r, &pA, &pB, &Ap, &partial, &bar, &iters, &resid, &err};
    PSR_HIP_CHECK(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(cg_all_depths), dim3(wg), dim3(kCgThreads),
                                             args, 0, s));
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
