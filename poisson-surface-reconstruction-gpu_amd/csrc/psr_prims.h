// Device-wide primitives written for wave64 CDNA: exclusive scan, stable LSD
// radix sort of (u32 key, u32 value) pairs.  They replace the CUB calls of the
// reference (core/AlgorithmTypes.cuh:32-50, core/AlgorithmTypes.cu:189-201).
#pragma once
#include <hip/hip_runtime.h>

namespace psr {

// Scratch needed by scan_exclusive for n elements, in int32 units.
size_t scan_scratch_ints(size_t n);
// out[i] = sum(in[0..i)); *total = sum(in[0..n)) if total != nullptr.  in may equal out.
void scan_exclusive(const int* in, int* out, int n, int* total, int* scratch, hipStream_t s);

// Scratch needed by radix_sort_pairs for n elements, in int32 units.
size_t sort_scratch_ints(size_t n);
// Stable ascending sort of keys (only the low `bits` bits are examined) carrying
// values.  The sorted data ends in (keys, vals): the four buffers are swapped
// in place of a final copy, so keys/vals may come back as the former ktmp/vtmp.
void radix_sort_pairs(unsigned*& keys, unsigned*& vals, unsigned*& ktmp, unsigned*& vtmp, int n, int bits,
                      int* scratch, hipStream_t s);

}  // namespace psr
