// Device primitives shared by the connectivities, Leiden and spatial stages.
//
// All are deterministic and use no global atomics: scans are three-phase (block reduce,
// scan of block totals, block scan with offsets); the radix sort is a stable LSD sort whose
// in-block ranks come from wave ballots, so every output position is a pure function of
// the input order.
#pragma once

#include <cstdint>

#include "common.h"

namespace msc {
namespace prims {

// Exclusive prefix sum out[i] = sum_{j<i} in[i]; out may alias in.  Returns the workspace
// bytes when ws == nullptr.  T in {int32_t, int64_t}.
template <typename T>
size_t scan_workspace(int64_t n);
template <typename T>
int exclusive_scan(const T* in, T* out, int64_t n, void* ws, size_t ws_bytes, hipStream_t st);

// Stable LSD radix sort of (key, value) u32 pairs on the low `key_bits` bits of the key.
// Inputs are left untouched; the sorted result lands in (keys_out, vals_out).
size_t radix_workspace(int64_t n);
int radix_sort_pairs(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                     uint32_t* vals_out, int64_t n, int key_bits, void* ws, size_t ws_bytes,
                     hipStream_t st);

// seg_start[j] = first position p in the ascending `keys` (length m) with keys[p] >= j, for
// j = 0..nseg (keys >= nseg count as nseg).  seg_start has nseg + 1 entries.
int segment_starts(const uint32_t* keys, int64_t m, int64_t nseg, int64_t* seg_start,
                   hipStream_t st);

int bits_for(int64_t max_value);

}  // namespace prims
}  // namespace msc
