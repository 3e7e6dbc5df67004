// Wave64 scan and radix sort.
#include "psr_common.h"
#include "psr_prims.h"
#include "psr_tile_scan.h"

#include <algorithm>
#include <stdexcept>

namespace psr {

// ---------------------------------------------------------------------------
// Exclusive scan.  A tile is 256 threads x 8 items = 2048 elements.
// ---------------------------------------------------------------------------
static constexpr int kScanThreads = 256;
static constexpr int kScanItems = 8;
static constexpr int kScanTile = kScanThreads * kScanItems;

// Single-pass scan with decoupled look-back (psr_tile_scan.h); each thread
// owns kScanItems consecutive elements of its tile.
__device__ __forceinline__ void scan_single_pass_body(int psr_bid, int psr_nbid, const int* in, int* out, int n, const ScanTicket& t) {
    const int tile = ticket_tile(t);
    const int base = tile * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int i = base + k;
        v[k] = (i < n) ? in[i] : 0;
        sum += v[k];
    }
    int run = ticket_prefix<kScanThreads>(t, tile, sum);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int i = base + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
}
__global__ void __launch_bounds__(kScanThreads) scan_single_pass(const int* in, int* out, int n, ScanTicket t) {
    scan_single_pass_body(blockIdx.x, gridDim.x, in, out, n, t);
}
PSR_BATCHABLE(scan_single_pass, kScanThreads)


__device__ __forceinline__ void set_int_body(int psr_bid, int psr_nbid, int* p, int v) { *p = v; }
__global__ void __launch_bounds__(1) set_int(int* p, int v) {
    set_int_body(blockIdx.x, gridDim.x, p, v);
}
PSR_BATCHABLE(set_int, 1)


struct ZeroList {
    unsigned* p[kZeroSpans];
    size_t end[kZeroSpans];  // cumulative word counts
    int n;
};

__device__ __forceinline__ void zero_list_body(int psr_bid, int psr_nbid, const ZeroList& z) {
    const size_t total = z.end[z.n - 1];
    for (size_t i = size_t(psr_bid) * 256 + threadIdx.x; i < total; i += size_t(psr_nbid) * 256) {
        int k = 0;
        while (i >= z.end[k]) ++k;
        z.p[k][i - (k ? z.end[k - 1] : 0)] = 0u;
    }
}
__global__ void __launch_bounds__(256) zero_list(ZeroList z) {
    zero_list_body(blockIdx.x, gridDim.x, z);
}
PSR_BATCHABLE(zero_list, 256)


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
    PSR_LAUNCH(zero_list, blocks, 0, s, z);
    PSR_HIP_CHECK(hipGetLastError());
}

size_t scan_tiles(size_t n) { return div_up(n, kScanTile); }

ScanTicket scan_ticket(ScanState& st, int tiles, int* total) {
    if (size_t(tiles) > st.capacity || st.epoch + 1 >= kScanEpochMax)
        throw std::runtime_error("scan state not reserved for this size");
    ++st.epoch;
    return ScanTicket{st.status, st.counters, st.epoch, tiles, total, st.err};
}

void scan_exclusive(const int* in, int* out, int n, int* total, ScanState& st, hipStream_t s) {
    if (n <= 0) {
        if (total) PSR_LAUNCH(set_int, 1, 0, s, total, 0);
        return;
    }
    const int tiles = int(scan_tiles(size_t(n)));
    PSR_LAUNCH(scan_single_pass, tiles, 0, s, in, out, n, scan_ticket(st, tiles, total));
    PSR_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort, digits of at most 8 bits: the fewest passes of 8
// bits, with the key's bits spread evenly over them (21 bits: 7 + 7 + 7), so
// that each pass scans as few histogram entries as it can.
//   hist:    per tile 2^b-bin histogram, stored digit-major (digit * tiles + tile)
//   scan:    exclusive scan of the digit-major histogram -> global offsets
//   scatter: per tile, stable rank by wave ballots (multisplit) + offsets;
//            the last pass can also gather the points into the sorted order
// ---------------------------------------------------------------------------
static constexpr int kSortThreads = 256;
static constexpr int kSortIters = 2;  // small tiles: enough workgroups to fill the GPU at 200k keys
static constexpr int kSortTile = kSortThreads * kSortIters;  // 512 keys
static_assert(kSortTile == kSortTileKeys, "tile of the sort's interface");
static constexpr int kSortWaves = kSortThreads / PSR_WAVE;

__device__ __forceinline__ void radix_hist_body(int psr_bid, int psr_nbid, const unsigned* __restrict__ keys, int n, int shift,
                                                           unsigned mask, int tiles, int* __restrict__ hist) {
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    int base = psr_bid * kSortTile;
#pragma unroll 4
    for (int k = 0; k < kSortIters; ++k) {
        int i = base + k * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & mask], 1);
    }
    __syncthreads();
    if (threadIdx.x <= mask) hist[threadIdx.x * tiles + psr_bid] = h[threadIdx.x];
}
__global__ void __launch_bounds__(kSortThreads) radix_hist(const unsigned* __restrict__ keys, int n, int shift,
                                                           unsigned mask, int tiles, int* __restrict__ hist) {
    radix_hist_body(blockIdx.x, gridDim.x, keys, n, shift, mask, tiles, hist);
}
PSR_BATCHABLE(radix_hist, kSortThreads)


__device__ __forceinline__ void radix_scatter_body(int psr_bid, int psr_nbid, const unsigned* __restrict__ kin,
                                                              const unsigned* __restrict__ vin,
                                                              unsigned* __restrict__ kout,
                                                              unsigned* __restrict__ vout, int n, int shift,
                                                              unsigned mask, int tiles, const int* __restrict__ offs,
                                                              const SortGather& g) {
    __shared__ int running[256];      // digit count of earlier iterations + global base
    __shared__ int waveCnt[kSortWaves][256];
    const int lane = __lane_id(), wid = threadIdx.x / PSR_WAVE;
    running[threadIdx.x] = threadIdx.x <= mask ? offs[threadIdx.x * tiles + psr_bid] : 0;
    const unsigned long long ltmask = (1ull << lane) - 1ull;
    int base = psr_bid * kSortTile;
    for (int it = 0; it < kSortIters; ++it) {
        for (int w = 0; w < kSortWaves; ++w) waveCnt[w][threadIdx.x] = 0;
        __syncthreads();
        int i = base + it * kSortThreads + threadIdx.x;
        bool valid = i < n;
        unsigned key = valid ? kin[i] : 0u;
        unsigned val = valid ? vin[i] : 0u;
        int d = int((key >> shift) & mask);
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            if (!((mask >> b) & 1u)) break;  // the digit's bits end here
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
            if (g.posOut) {
                g.posOut[pos] = g.posIn[val];
                g.nrmOut[pos] = g.nrmIn[val];
            }
        }
        __syncthreads();
        int add = 0;
        for (int w = 0; w < kSortWaves; ++w) add += waveCnt[w][threadIdx.x];
        running[threadIdx.x] += add;
        __syncthreads();
    }
}
__global__ void __launch_bounds__(kSortThreads) radix_scatter(const unsigned* __restrict__ kin,
                                                              const unsigned* __restrict__ vin,
                                                              unsigned* __restrict__ kout,
                                                              unsigned* __restrict__ vout, int n, int shift,
                                                              unsigned mask, int tiles, const int* __restrict__ offs,
                                                              SortGather g) {
    radix_scatter_body(blockIdx.x, gridDim.x, kin, vin, kout, vout, n, shift, mask, tiles, offs, g);
}
PSR_BATCHABLE(radix_scatter, kSortThreads)


int sort_digit_bits(int bits) {
    const int passes = (bits + 7) / 8;
    return passes > 0 ? (bits + passes - 1) / passes : 0;
}

size_t sort_scan_elems(size_t n) { return 256 * size_t(div_up(n, kSortTile)); }

size_t sort_scratch_ints(size_t n) {
    size_t tiles = div_up(n, kSortTile);
    size_t h = 256 * tiles;
    return h + 64;
}

void radix_sort_pairs(unsigned*& keys, unsigned*& vals, unsigned*& ktmp, unsigned*& vtmp, int n, int bits,
                      int* scratch, bool firstHist, ScanState& st, const SortGather& g, hipStream_t s) {
    if (n <= 0) return;
    if (bits <= 0) return;
    int tiles = int(div_up(n, kSortTile));
    int* hist = scratch;
    const int passes = (bits + 7) / 8, width = sort_digit_bits(bits);
    unsigned *ka = keys, *va = vals, *kb = ktmp, *vb = vtmp;
    for (int p = 0; p < passes; ++p) {
        const int shift = width * p, b = std::min(width, bits - shift);
        const unsigned mask = (1u << b) - 1u;
        if (p > 0 || !firstHist) PSR_LAUNCH(radix_hist, tiles, 0, s, ka, n, shift, mask, tiles, hist);
        scan_exclusive(hist, hist, int(mask + 1) * tiles, nullptr, st, s);
        const SortGather none{};
        PSR_LAUNCH(radix_scatter, tiles, 0, s, ka, va, kb, vb, n, shift, mask, tiles, hist,
                   p == passes - 1 ? g : none);
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
