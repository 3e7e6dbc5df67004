// Device-wide primitives written for wave64 CDNA: exclusive scan, stable LSD
// radix sort of (u32 key, u32 value) pairs.  They replace the CUB calls of the
// reference (core/AlgorithmTypes.cuh:32-50, core/AlgorithmTypes.cu:189-201).
#pragma once
#include <hip/hip_runtime.h>

#include <initializer_list>

namespace psr {

// State of the single-pass scan of one stream: per-tile status words
// (epoch | flag | value) and a ring of tile counters.  A call takes the next
// epoch, so status words of earlier calls never need clearing; the state is
// zeroed when it is (re)allocated.
struct ScanState {
    unsigned long long* status = nullptr;
    unsigned* counters = nullptr;
    size_t capacity = 0;  // tiles
    unsigned epoch = 0;
    int* err = nullptr;   // set to 1 by a look-back whose wait reached its spin bound (may be null)
};
constexpr int kScanRing = 64;             // tile counters, one per epoch modulo the ring
constexpr unsigned kScanEpochMax = 1u << 29;
size_t scan_tiles(size_t n);
// One call's share of the scan state, for kernels that scan their own counts
// (psr_tile_scan.h): tiles of the launch, and where the grand total goes.
struct ScanTicket {
    unsigned long long* status;
    unsigned* counters;
    unsigned epoch;
    int tiles;
    int* total;
    int* err;  // set to 1 when a look-back wait reaches its spin bound (may be null)
};
// Takes the next epoch for a launch of `tiles` workgroups (the state must hold them).
ScanTicket scan_ticket(ScanState& st, int tiles, int* total);
// out[i] = sum(in[0..i)); *total = sum(in[0..n)) if total != nullptr.  in may equal out.
void scan_exclusive(const int* in, int* out, int n, int* total, ScanState& st, hipStream_t s);

// Sets up to kZeroSpans word ranges to zero with one launch.
struct ZeroSpan {
    void* p;
    size_t words;  // 4-byte words
};
constexpr int kZeroSpans = 8;
void zero_spans(std::initializer_list<ZeroSpan> spans, hipStream_t s);

// Scratch needed by radix_sort_pairs for n elements, in int32 units, and the
// length of the histogram it scans (for the scan state).
size_t sort_scratch_ints(size_t n);
size_t sort_scan_elems(size_t n);
// Tiles of the sort: the first pass's histogram, when its producer takes it
// (firstHist), holds at scratch[d * tiles + t] the keys of tile t (keys
// [512 t, 512 t + 512)) whose low sort_digit_bits(bits) bits are d.
constexpr int kSortTileKeys = 512;
// Width of each pass's digit for keys of `bits` bits: the fewest passes of at
// most 8 bits, the bits spread evenly over them (the last may be narrower).
int sort_digit_bits(int bits);
// Optional gather fused into the last pass: posOut[i] = posIn[v], nrmOut[i] =
// nrmIn[v] for the value v that lands at i.
struct SortGather {
    const float4* posIn;
    const float4* nrmIn;
    float4* posOut;
    float4* nrmOut;
};
// Stable ascending sort of keys (only the low `bits` bits are examined) carrying
// values.  The sorted data ends in (keys, vals): the four buffers are swapped
// in place of a final copy, so keys/vals may come back as the former ktmp/vtmp.
void radix_sort_pairs(unsigned*& keys, unsigned*& vals, unsigned*& ktmp, unsigned*& vtmp, int n, int bits,
                      int* scratch, bool firstHist, ScanState& st, const SortGather& g, hipStream_t s);

}  // namespace psr
