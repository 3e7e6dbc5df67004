// Wave64 scan and radix sort.
#include "psr_common.h"
#include "psr_prims.h"

#include <algorithm>
#include <stdexcept>

namespace psr {

// ---------------------------------------------------------------------------
// Exclusive scan.  A tile is 256 threads x 8 items = 2048 elements.
// ---------------------------------------------------------------------------
static constexpr int kScanThreads = 256;
static constexpr int kScanItems = 8;
static constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ int wave_incl_scan(int v) {
    int lane = __lane_id();
#pragma unroll
    for (int off = 1; off < PSR_WAVE; off <<= 1) {
        int t = __shfl_up(v, off, PSR_WAVE);
        if (lane >= off) v += t;
    }
    return v;
}

// Block-wide exclusive scan of one value per thread; returns the block total.
template <int NT>
__device__ __forceinline__ int block_excl_scan(int v, int& excl) {
    __shared__ int wsum[NT / PSR_WAVE];
    __shared__ int wtotal;
    int lane = __lane_id(), wid = threadIdx.x / PSR_WAVE;
    int inc = wave_incl_scan(v);
    if (lane == PSR_WAVE - 1) wsum[wid] = inc;
    __syncthreads();
    if (wid == 0) {
        int w = lane < NT / PSR_WAVE ? wsum[lane] : 0;
        int wi = wave_incl_scan(w);
        if (lane < NT / PSR_WAVE) wsum[lane] = wi - w;
        if (lane == NT / PSR_WAVE - 1) wtotal = wi;
    }
    __syncthreads();
    excl = inc - v + wsum[wid];
    int total = wtotal;
    __syncthreads();
    return total;
}

// Single-pass scan with decoupled look-back.  Tiles take their index from an
// arrival counter, so every tile a tile waits on has already started; a tile
// publishes its own sum first (flag 1), then walks back over its predecessors
// until one has published an inclusive prefix (flag 2), and publishes its own
// inclusive prefix.  A status word is one 64-bit atomic: epoch (bits 34-63),
// flag (32-33), value (0-31), so words of earlier calls read as not ready.
__global__ void __launch_bounds__(kScanThreads) scan_single_pass(const int* in, int* out, int n, int tiles,
                                                                 int* __restrict__ total,
                                                                 unsigned long long* status,
                                                                 unsigned* counters, unsigned epoch) {
    __shared__ int sTile, sPrefix;
    if (threadIdx.x == 0) {
        const int t = int(atomicAdd(&counters[epoch % kScanRing], 1u));
        if (t == 0) counters[(epoch + 1) % kScanRing] = 0u;  // the next call's counter
        sTile = t;
    }
    __syncthreads();
    const int tile = sTile;
    // each thread owns kScanItems consecutive elements of the tile
    const int base = tile * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int i = base + k;
        v[k] = (i < n) ? in[i] : 0;
        sum += v[k];
    }
    int excl;
    const int agg = block_excl_scan<kScanThreads>(sum, excl);
    if (threadIdx.x < PSR_WAVE) {
        // the first wave looks back over windows of 64 predecessors at once
        const int lane = threadIdx.x;
        const unsigned long long tag = static_cast<unsigned long long>(epoch) << 34;
        int prefix = 0;
        if (tile > 0) {
            if (lane == 0)
                __hip_atomic_store(&status[tile], tag | (1ull << 32) | unsigned(agg), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            // every predecessor has started and publishes without waiting, so each
            // wait ends; the spin bound only guards against a broken state
            for (int j0 = tile - 1;; j0 -= PSR_WAVE) {
                const int j = j0 - lane;
                unsigned flag = 2u;  // before the first tile: an inclusive zero
                int val = 0;
                if (j >= 0) {
                    for (unsigned spins = 0;; ++spins) {
                        const unsigned long long w =
                            __hip_atomic_load(&status[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        flag = unsigned(w >> 32) & 3u;
                        if (unsigned(w >> 34) == epoch && flag != 0u) {
                            val = int(unsigned(w));
                            break;
                        }
                        if (spins > (1u << 24)) {
                            flag = 2u;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                }
                const unsigned long long incl = __ballot(flag == 2u);
                const int stop = incl ? __builtin_ctzll(incl) : PSR_WAVE;  // closest inclusive prefix
                int c = lane <= stop ? val : 0;
#pragma unroll
                for (int off = PSR_WAVE / 2; off > 0; off >>= 1) c += __shfl_xor(c, off, PSR_WAVE);
                prefix += c;
                if (incl) break;
            }
        }
        if (lane == 0) {
            __hip_atomic_store(&status[tile], tag | (2ull << 32) | unsigned(prefix + agg), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            if (total && tile == tiles - 1) *total = prefix + agg;
            sPrefix = prefix;
        }
    }
    __syncthreads();
    int run = excl + sPrefix;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int i = base + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
}

__global__ void set_int(int* p, int v) { *p = v; }

struct ZeroList {
    unsigned* p[kZeroSpans];
    size_t end[kZeroSpans];  // cumulative word counts
    int n;
};

__global__ void __launch_bounds__(256) zero_list(ZeroList z) {
    const size_t total = z.end[z.n - 1];
    for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < total; i += size_t(gridDim.x) * 256) {
        int k = 0;
        while (i >= z.end[k]) ++k;
        z.p[k][i - (k ? z.end[k - 1] : 0)] = 0u;
    }
}

void zero_spans(std::initializer_list<ZeroSpan> spans, hipStream_t s) {
    ZeroList z{};
    size_t acc = 0;
    for (const ZeroSpan& sp : spans) {
        if (sp.words == 0) continue;
        if (z.n == kZeroSpans) throw std::runtime_error("too many zero spans");
        acc += sp.words;
        z.p[z.n] = static_cast<unsigned*>(sp.p);
        z.end[z.n] = acc;
        ++z.n;
    }
    if (z.n == 0) return;
    const unsigned blocks = unsigned(std::min<size_t>(div_up(acc, 256), 2048));
    zero_list<<<blocks, 256, 0, s>>>(z);
    PSR_HIP_CHECK(hipGetLastError());
}

size_t scan_tiles(size_t n) { return div_up(n, kScanTile); }

void scan_exclusive(const int* in, int* out, int n, int* total, ScanState& st, hipStream_t s) {
    if (n <= 0) {
        if (total) set_int<<<1, 1, 0, s>>>(total, 0);
        return;
    }
    const int tiles = int(scan_tiles(size_t(n)));
    if (size_t(tiles) > st.capacity || st.epoch + 1 >= kScanEpochMax)
        throw std::runtime_error("scan state not reserved for this size");
    ++st.epoch;
    scan_single_pass<<<tiles, kScanThreads, 0, s>>>(in, out, n, tiles, total, st.status, st.counters, st.epoch);
    PSR_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort, 8-bit digits.
//   hist:    per tile 256-bin histogram, stored digit-major (digit * tiles + tile)
//   scan:    exclusive scan of the digit-major histogram -> global offsets
//   scatter: per tile, stable rank by wave ballots (multisplit) + offsets
// ---------------------------------------------------------------------------
static constexpr int kSortThreads = 256;
static constexpr int kSortIters = 2;  // small tiles: enough workgroups to fill the GPU at 200k keys
static constexpr int kSortTile = kSortThreads * kSortIters;  // 512 keys
static constexpr int kSortWaves = kSortThreads / PSR_WAVE;

__global__ void __launch_bounds__(kSortThreads) radix_hist(const unsigned* __restrict__ keys, int n, int shift,
                                                           int tiles, int* __restrict__ hist) {
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    int base = blockIdx.x * kSortTile;
#pragma unroll 4
    for (int k = 0; k < kSortIters; ++k) {
        int i = base + k * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255], 1);
    }
    __syncthreads();
    hist[threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(kSortThreads) radix_scatter(const unsigned* __restrict__ kin,
                                                              const unsigned* __restrict__ vin,
                                                              unsigned* __restrict__ kout,
                                                              unsigned* __restrict__ vout, int n, int shift,
                                                              int tiles, const int* __restrict__ offs) {
    __shared__ int running[256];      // digit count of earlier iterations + global base
    __shared__ int waveCnt[kSortWaves][256];
    const int lane = __lane_id(), wid = threadIdx.x / PSR_WAVE;
    running[threadIdx.x] = offs[threadIdx.x * tiles + blockIdx.x];
    const unsigned long long ltmask = (1ull << lane) - 1ull;
    int base = blockIdx.x * kSortTile;
    for (int it = 0; it < kSortIters; ++it) {
        for (int w = 0; w < kSortWaves; ++w) waveCnt[w][threadIdx.x] = 0;
        __syncthreads();
        int i = base + it * kSortThreads + threadIdx.x;
        bool valid = i < n;
        unsigned key = valid ? kin[i] : 0u;
        unsigned val = valid ? vin[i] : 0u;
        int d = int((key >> shift) & 255u);
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            unsigned long long bb = __ballot((d >> b) & 1);
            peers &= ((d >> b) & 1) ? bb : ~bb;
        }
        int lower = __popcll(peers & ltmask);
        if (valid && lower == 0) waveCnt[wid][d] = __popcll(peers);
        __syncthreads();
        int before = 0;
        for (int w = 0; w < wid; ++w) before += waveCnt[w][d];
        if (valid) {
            int pos = running[d] + before + lower;
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        int add = 0;
        for (int w = 0; w < kSortWaves; ++w) add += waveCnt[w][threadIdx.x];
        running[threadIdx.x] += add;
        __syncthreads();
    }
}

size_t sort_scan_elems(size_t n) { return 256 * size_t(div_up(n, kSortTile)); }

size_t sort_scratch_ints(size_t n) {
    size_t tiles = div_up(n, kSortTile);
    size_t h = 256 * tiles;
    return h + 64;
}

void radix_sort_pairs(unsigned*& keys, unsigned*& vals, unsigned*& ktmp, unsigned*& vtmp, int n, int bits,
                      int* scratch, ScanState& st, hipStream_t s) {
    if (n <= 1) return;
    int tiles = int(div_up(n, kSortTile));
    int* hist = scratch;
    int passes = (bits + 7) / 8;
    unsigned *ka = keys, *va = vals, *kb = ktmp, *vb = vtmp;
    for (int p = 0; p < passes; ++p) {
        int shift = 8 * p;
        radix_hist<<<tiles, kSortThreads, 0, s>>>(ka, n, shift, tiles, hist);
        scan_exclusive(hist, hist, 256 * tiles, nullptr, st, s);
        radix_scatter<<<tiles, kSortThreads, 0, s>>>(ka, va, kb, vb, n, shift, tiles, hist);
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (ka != keys) {
        std::swap(keys, ktmp);
        std::swap(vals, vtmp);
    }
    PSR_HIP_CHECK(hipGetLastError());
}

}  // namespace psr
