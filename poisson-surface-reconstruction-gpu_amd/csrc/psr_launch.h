// Launches that run at once for one cloud, or are recorded so that a batch
// of clouds runs each stage as one launch.
//
// Every kernel of the pipeline is a device body run(bid, nbid, args...) that
// uses bid / nbid in place of blockIdx.x / gridDim.x.  For one cloud the
// ordinary __global__ wrapper launches it on the context's stream.  While a
// batch records (a Recorder is installed for the calling thread), the launch
// is appended to the recording instead; the batch then merges the launches of
// all its clouds that come from the same launch site (and the same occurrence
// of it) into one launch of multi_kernel, whose workgroups are the clouds'
// grids laid end to end: workgroup b finds its cloud c by a binary search over
// the grid starts and runs the body with that cloud's arguments and
// bid = b - start[c].  A cloud's arguments are packed (Pack) into a device
// blob uploaded once per batch phase.
#pragma once
#include <hip/hip_runtime.h>

#include <cstring>
#include <type_traits>
#include <utility>
#include <vector>

namespace psr {

// Arguments of one cloud's launch, laid out identically on host and device.
template <class... T>
struct Pack;
template <>
struct Pack<> {};
template <class H, class... T>
struct Pack<H, T...> {
    H h;
    Pack<T...> t;
};

template <class... T>
struct PackMaker;
template <>
struct PackMaker<> {
    static Pack<> make() { return {}; }
};
template <class H, class... T>
struct PackMaker<H, T...> {
    template <class A, class... R>
    static Pack<H, T...> make(A&& a, R&&... r) {
        Pack<H, T...> p;
        p.h = H(std::forward<A>(a));
        p.t = PackMaker<T...>::make(std::forward<R>(r)...);
        return p;
    }
};

template <class F, class... A>
__device__ __forceinline__ void apply_pack(F&& f, const Pack<>&, const A&... a) {
    f(a...);
}
template <class F, class H, class... T, class... A>
__device__ __forceinline__ void apply_pack(F&& f, const Pack<H, T...>& p, const A&... a) {
    apply_pack(f, p.t, a..., p.h);
}

// The waves per SIMD a body's kernels are compiled for (K::kMinWaves where the
// descriptor names it, else no bound): the register budget of that occupancy.
template <class K, class = void>
struct min_waves {
    static constexpr int value = 1;
};
template <class K>
struct min_waves<K, std::void_t<decltype(K::kMinWaves)>> {
    static constexpr int value = K::kMinWaves;
};

// The grid starts of the clouds of one batched launch, passed by value (the
// search reads them from the kernel arguments, not from memory).
constexpr int kMaxMultiClouds = 64;
struct MultiStarts {
    unsigned s[kMaxMultiClouds + 1];  // s[c]: first workgroup of cloud c; s[n]: the grid
    int n;
};

// The batched form of a kernel body K::run (K::kBlock threads a block).
template <class K, class... T>
__global__ void __launch_bounds__(K::kBlock) __attribute__((amdgpu_waves_per_eu(min_waves<K>::value)))
    multi_kernel(const Pack<T...>* __restrict__ packs, MultiStarts st) {
    const unsigned b = blockIdx.x;
    int lo = 0;  // the last c with s[c] <= b
#pragma unroll
    for (int step = kMaxMultiClouds / 2; step > 0; step >>= 1)
        if (lo + step < st.n && st.s[lo + step] <= b) lo += step;
    const int bid = int(b - st.s[lo]), nbid = int(st.s[lo + 1] - st.s[lo]);
    apply_pack([&](const auto&... a) { K::run(bid, nbid, a...); }, packs[lo]);
}

// The batched form of a body whose workgroups meet in group barriers (the
// CG): a persistent launch of at most the resident budget of workgroups,
// each taking the merged grid's workgroup indices from a counter in
// increasing order until none is left.  Every group's workgroups are taken
// one after another, so while a group waits for its last members, every
// workgroup ahead of them belongs to a group that is complete and finishes;
// a group therefore never waits for a workgroup that cannot start, provided
// no group is larger than the launch.  A workgroup that finishes early (a
// small group) takes the next one instead of idling until the slowest group
// of a wave is done.
template <class K, class... T>
__global__ void __launch_bounds__(K::kBlock) __attribute__((amdgpu_waves_per_eu(min_waves<K>::value)))
    multi_queue_kernel(const Pack<T...>* __restrict__ packs, MultiStarts st, unsigned* __restrict__ counter) {
    __shared__ unsigned sTask;
    const unsigned total = st.s[st.n];
    for (;;) {
        if (threadIdx.x == 0) sTask = atomicAdd(counter, 1u);
        __syncthreads();
        const unsigned b = sTask;
        __syncthreads();  // read by every thread before thread 0 takes the next
        if (b >= total) return;
        int lo = 0;
#pragma unroll
        for (int step = kMaxMultiClouds / 2; step > 0; step >>= 1)
            if (lo + step < st.n && st.s[lo + step] <= b) lo += step;
        const int bid = int(b - st.s[lo]), nbid = int(st.s[lo + 1] - st.s[lo]);
        apply_pack([&](const auto&... a) { K::run(bid, nbid, a...); }, packs[lo]);
        __syncthreads();  // the body's shared memory is free before the next task
    }
}

// The batched form of kernel fn (for occupancy queries).
template <class K, class... T>
inline const void* multi_queue_kernel_of(void (*)(T...)) {
    return reinterpret_cast<const void*>(&multi_queue_kernel<K, std::decay_t<T>...>);
}

// One recorded launch of one cloud.
struct LaunchRec {
    const void* site;      // identity of the launch site
    int occurrence;        // how many launches from this site the cloud recorded before
    const void* multiFn;   // multi_kernel<K, T...>, or multi_queue_kernel<K, T...> when coResident
    unsigned grid, block, shmem;
    bool coResident;       // every workgroup of the merged launch must be resident at once
    std::vector<char> pack;
};

// The launches one cloud recorded during a batch phase.
struct Recorder {
    std::vector<std::vector<LaunchRec>> clouds;
    int cloud = 0;  // the cloud being recorded
    void add(const void* site, const void* fn, unsigned grid, unsigned block, unsigned shmem, bool coResident,
             const void* pack, size_t bytes) {
        auto& v = clouds[size_t(cloud)];
        int occ = 0;
        for (const LaunchRec& r : v)
            if (r.site == site) ++occ;
        LaunchRec r{site, occ, fn, grid, block, shmem, coResident, {}};
        r.pack.assign(static_cast<const char*>(pack), static_cast<const char*>(pack) + bytes);
        v.push_back(std::move(r));
    }
};

// The recorder of the calling thread, or null (launch at once).
Recorder*& current_recorder();

template <class K, class... T, class... A>
inline void launch_kernel(const void* site, void (*fn)(T...), unsigned grid, unsigned shmem, hipStream_t s,
                          bool coResident, A&&... a) {
    static_assert(sizeof...(T) == sizeof...(A), "argument count");
    if (Recorder* r = current_recorder()) {
        using P = Pack<std::decay_t<T>...>;
        static_assert(std::is_trivially_copyable<P>::value, "packed arguments are copied as bytes");
        const P p = PackMaker<std::decay_t<T>...>::make(std::forward<A>(a)...);
        const void* fn = coResident ? reinterpret_cast<const void*>(&multi_queue_kernel<K, std::decay_t<T>...>)
                                    : reinterpret_cast<const void*>(&multi_kernel<K, std::decay_t<T>...>);
        r->add(site, fn, grid, K::kBlock, shmem, coResident, &p, sizeof(P));
        return;
    }
    if (grid == 0) return;
    hipLaunchKernelGGL(fn, dim3(grid), dim3(K::kBlock), shmem, s, std::forward<A>(a)...);
}

}  // namespace psr

// Declares the batch descriptor NAME##_k of the body NAME##_body (NT threads).
#define PSR_BATCHABLE(NAME, NT)                                                         \
    struct NAME##_k {                                                                   \
        static constexpr int kBlock = NT;                                                \
        template <class... A>                                                           \
        __device__ __forceinline__ static void run(int bid, int nbid, const A&... a) { \
            NAME##_body(bid, nbid, a...);                                               \
        }                                                                               \
    };

// As PSR_BATCHABLE, for a body whose kernels are compiled for WAVES waves per SIMD.
#define PSR_BATCHABLE_WAVES(NAME, NT, WAVES)                                            \
    struct NAME##_k {                                                                   \
        static constexpr int kBlock = NT;                                                \
        static constexpr int kMinWaves = WAVES;                                          \
        template <class... A>                                                           \
        __device__ __forceinline__ static void run(int bid, int nbid, const A&... a) { \
            NAME##_body(bid, nbid, a...);                                               \
        }                                                                               \
    };

// Launches kernel NAME (descriptor KDESC) at once, or records it for a batch.
#define PSR_LAUNCH_K(NAME, KDESC, GRID, SHMEM, STREAM, ...)                                                      \
    do {                                                                                                         \
        static const char psr_site_ = 0;                                                                         \
        ::psr::launch_kernel<KDESC>(&psr_site_, NAME, unsigned(GRID), unsigned(SHMEM), STREAM, false, __VA_ARGS__); \
    } while (0)
#define PSR_LAUNCH(NAME, GRID, SHMEM, STREAM, ...) PSR_LAUNCH_K(NAME, NAME##_k, GRID, SHMEM, STREAM, __VA_ARGS__)
