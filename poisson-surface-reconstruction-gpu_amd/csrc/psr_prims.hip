// Wave64 scan and radix sort.
#include "psr_common.h"
#include "psr_prims.h"

namespace psr {

// ---------------------------------------------------------------------------
// Exclusive scan: reduce tiles -> scan tile sums (one workgroup) -> downsweep.
// A tile is 256 threads x 8 items = 2048 elements.
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

__global__ void __launch_bounds__(kScanThreads) scan_reduce_tiles(const int* __restrict__ in, int n,
                                                                  int* __restrict__ tileSums) {
    int base = blockIdx.x * kScanTile;
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int i = base + k * kScanThreads + threadIdx.x;
        s += (i < n) ? in[i] : 0;
    }
    int excl;
    int tot = block_excl_scan<kScanThreads>(s, excl);
    if (threadIdx.x == 0) tileSums[blockIdx.x] = tot;
}

// Single workgroup scans up to 1024 * 64 tile sums in place; writes the grand total.
__global__ void __launch_bounds__(1024) scan_tile_sums(int* __restrict__ sums, int m, int* __restrict__ total) {
    constexpr int NT = 1024;
    int per = (m + NT - 1) / NT;
    int start = threadIdx.x * per;
    int s = 0;
    for (int k = 0; k < per; ++k) {
        int i = start + k;
        if (i < m) s += sums[i];
    }
    int excl;
    int tot = block_excl_scan<NT>(s, excl);
    int run = excl;
    for (int k = 0; k < per; ++k) {
        int i = start + k;
        if (i < m) {
            int v = sums[i];
            sums[i] = run;
            run += v;
        }
    }
    if (threadIdx.x == 0 && total) *total = tot;
}

// in and out may be the same array: each thread reads its items before it
// writes them, and the tiles are disjoint.
__global__ void __launch_bounds__(kScanThreads) scan_downsweep(const int* in, int* out, int n,
                                                               const int* __restrict__ tileOffsets) {
    // Each thread owns kScanItems consecutive elements of the tile.
    int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int i = base + k;
        v[k] = (i < n) ? in[i] : 0;
        s += v[k];
    }
    int excl;
    block_excl_scan<kScanThreads>(s, excl);
    int run = excl + tileOffsets[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        int i = base + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
}

// Up to this many tile sums are scanned by one workgroup; more go through a
// second level of the same three-pass scan.
static constexpr int kScanSingleLevel = 1024 * 64;

size_t scan_scratch_ints(size_t n) {
    size_t tiles = div_up(n, kScanTile);
    if (tiles <= size_t(kScanSingleLevel)) return tiles + 64;
    return 2 * tiles + 64 + scan_scratch_ints(tiles);
}

void scan_exclusive(const int* in, int* out, int n, int* total, int* scratch, hipStream_t s) {
    if (n <= 0) {
        if (total) PSR_HIP_CHECK(hipMemsetAsync(total, 0, sizeof(int), s));
        return;
    }
    int tiles = int(div_up(n, kScanTile));
    scan_reduce_tiles<<<tiles, kScanThreads, 0, s>>>(in, n, scratch);
    if (tiles <= kScanSingleLevel) {
        scan_tile_sums<<<1, 1024, 0, s>>>(scratch, tiles, total);
        scan_downsweep<<<tiles, kScanThreads, 0, s>>>(in, out, n, scratch);
    } else {
        // layout: tile sums [tiles] | their exclusive scan [tiles] | inner scratch
        int* offs = scratch + tiles;
        scan_exclusive(scratch, offs, tiles, total, offs + tiles + 64, s);
        scan_downsweep<<<tiles, kScanThreads, 0, s>>>(in, out, n, offs);
    }
    PSR_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort, 8-bit digits.
//   hist:    per tile 256-bin histogram, stored digit-major (digit * tiles + tile)
//   scan:    exclusive scan of the digit-major histogram -> global offsets
//   scatter: per tile, stable rank by wave ballots (multisplit) + offsets
// ---------------------------------------------------------------------------
static constexpr int kSortThreads = 256;
static constexpr int kSortIters = 16;
static constexpr int kSortTile = kSortThreads * kSortIters;  // 4096 keys
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

size_t sort_scratch_ints(size_t n) {
    size_t tiles = div_up(n, kSortTile);
    size_t h = 256 * tiles;
    return h + scan_scratch_ints(h) + 64;
}

void radix_sort_pairs(unsigned*& keys, unsigned*& vals, unsigned*& ktmp, unsigned*& vtmp, int n, int bits,
                      int* scratch, hipStream_t s) {
    if (n <= 1) return;
    int tiles = int(div_up(n, kSortTile));
    int* hist = scratch;
    int* scanScratch = scratch + 256 * size_t(tiles);
    int passes = (bits + 7) / 8;
    unsigned *ka = keys, *va = vals, *kb = ktmp, *vb = vtmp;
    for (int p = 0; p < passes; ++p) {
        int shift = 8 * p;
        radix_hist<<<tiles, kSortThreads, 0, s>>>(ka, n, shift, tiles, hist);
        scan_exclusive(hist, hist, 256 * tiles, nullptr, scanScratch, s);
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
