// Device side of the single-pass scan: block scans and the decoupled
// look-back, shared by the scan kernel and by kernels that produce and consume
// their own counts in one launch.
#pragma once
#include "psr_common.h"
#include "psr_prims.h"

namespace psr {

__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = __lane_id();
#pragma unroll
    for (int off = 1; off < PSR_WAVE; off <<= 1) {
        const int t = __shfl_up(v, off, PSR_WAVE);
        if (lane >= off) v += t;
    }
    return v;
}

// Block-wide exclusive scan of one value per thread; returns the block total
// (every thread of the block must call it).
template <int NT>
__device__ __forceinline__ int block_excl_scan(int v, int& excl) {
    __shared__ int wsum[NT / PSR_WAVE];
    __shared__ int wtotal;
    const int lane = __lane_id(), wid = threadIdx.x / PSR_WAVE;
    const int inc = wave_incl_scan(v);
    if (lane == PSR_WAVE - 1) wsum[wid] = inc;
    __syncthreads();
    if (wid == 0) {
        const int w = lane < NT / PSR_WAVE ? wsum[lane] : 0;
        const int wi = wave_incl_scan(w);
        if (lane < NT / PSR_WAVE) wsum[lane] = wi - w;
        if (lane == NT / PSR_WAVE - 1) wtotal = wi;
    }
    __syncthreads();
    excl = inc - v + wsum[wid];
    const int total = wtotal;
    __syncthreads();
    return total;
}

// The tile index of this workgroup, in arrival order (every thread of the
// block must call it).  Every tile a tile waits on has therefore started.
__device__ __forceinline__ int ticket_tile(const ScanTicket& t) {
    __shared__ int sTile;
    if (threadIdx.x == 0) {
        const int v = int(atomicAdd(&t.counters[t.epoch % kScanRing], 1u));
        if (v == 0) t.counters[(t.epoch + 1) % kScanRing] = 0u;  // the next call's counter
        sTile = v;
    }
    __syncthreads();
    return sTile;
}

// The look-back of one tile, run by one whole wave: the tile publishes its
// sum agg first (flag 1), then walks back over 64 predecessors at a time until
// one has published an inclusive prefix (flag 2), and publishes its own.  A
// status word is one 64-bit atomic: epoch (bits 34-63), flag (32-33), value
// (0-31), so words of earlier calls read as not ready.  The last tile writes
// the grand total.  Returns the tile's exclusive prefix (every lane).
__device__ __forceinline__ int wave_lookback(const ScanTicket& t, int tile, int agg) {
    const int lane = __lane_id();
    const unsigned long long tag = static_cast<unsigned long long>(t.epoch) << 34;
    int prefix = 0;
    if (tile > 0) {
        if (lane == 0)
            __hip_atomic_store(&t.status[tile], tag | (1ull << 32) | unsigned(agg), __ATOMIC_RELAXED,
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
                        __hip_atomic_load(&t.status[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    flag = unsigned(w >> 32) & 3u;
                    if (unsigned(w >> 34) == t.epoch && flag != 0u) {
                        val = int(unsigned(w));
                        break;
                    }
                    if (spins > (1u << 24)) {
                        // a broken state: the prefix is wrong, so the call must fail
                        if (t.err) __hip_atomic_store(t.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
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
        __hip_atomic_store(&t.status[tile], tag | (2ull << 32) | unsigned(prefix + agg), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        if (t.total && tile == t.tiles - 1) *t.total = prefix + agg;
    }
    return prefix;
}
// Exclusive prefix of v over all threads of all earlier tiles and the earlier
// threads of this one (every thread of the block must call it); the first
// wave does the look-back.
template <int NT>
__device__ __forceinline__ int ticket_prefix(const ScanTicket& t, int tile, int v) {
    __shared__ int sPrefix;
    int excl;
    const int agg = block_excl_scan<NT>(v, excl);
    if (threadIdx.x < PSR_WAVE) {
        const int prefix = wave_lookback(t, tile, agg);
        if (threadIdx.x == 0) sPrefix = prefix;
    }
    __syncthreads();
    return sPrefix + excl;
}
// Three such prefixes of one tile (scans of their own on one epoch), their
// look-backs run side by side by the first three waves.
template <int NT>
__device__ __forceinline__ void ticket_prefix3(const ScanTicket& t0, const ScanTicket& t1, const ScanTicket& t2,
                                               int tile, int v0, int v1, int v2, int& r0, int& r1, int& r2) {
    static_assert(NT >= 3 * PSR_WAVE, "a wave per look-back");
    __shared__ int sPrefix3[3];
    int e0, e1, e2;
    const int g0 = block_excl_scan<NT>(v0, e0);
    const int g1 = block_excl_scan<NT>(v1, e1);
    const int g2 = block_excl_scan<NT>(v2, e2);
    const int w = threadIdx.x / PSR_WAVE;
    if (w < 3) {
        const int prefix = w == 0 ? wave_lookback(t0, tile, g0)
                                  : (w == 1 ? wave_lookback(t1, tile, g1) : wave_lookback(t2, tile, g2));
        if (__lane_id() == 0) sPrefix3[w] = prefix;
    }
    __syncthreads();
    r0 = sPrefix3[0] + e0;
    r1 = sPrefix3[1] + e1;
    r2 = sPrefix3[2] + e2;
}
}  // namespace psr
