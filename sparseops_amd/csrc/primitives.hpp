// Device-wide primitives used by the LeSpGEMM pipeline (all hand-written for gfx950):
//   exclusive_scan_i64   3-phase tile scan (tile sums -> scan of sums -> add back)
//   radix_sort_u64       stable LSD radix sort of 64-bit keys with optional 32/64-bit
//                        payload, 8-bit digits, wave-ballot multisplit ranking.
#pragma once

#include "common.hpp"

namespace spops {

// out[i] = sum_{j<i} in[i]; returns the total (synchronises the stream to read it when
// `total` is not null).  in and out may alias.  Temporary storage is allocated here.
void exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, hipStream_t s, int64_t *total);

// Stable sort of keys (and payload, if non-null) by bits [begin_bit, end_bit) of the key.
// Ping-pongs between the given buffers and the internal temporaries; the result ends in
// keys / vals.  Payload element size is 4 or 8 bytes.
void radix_sort_u64(uint64_t *keys, void *vals, int val_bytes, int64_t n, int begin_bit, int end_bit, hipStream_t s);

// Number of bits needed to represent v (0 for v == 0).
inline int bit_width_u64(uint64_t v) {
  int b = 0;
  while (v) { ++b; v >>= 1; }
  return b;
}

}  // namespace spops
