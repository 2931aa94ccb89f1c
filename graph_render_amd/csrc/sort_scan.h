// Host entry points of the hand-written scan and radix sort (sort_scan.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace gr {

// Exclusive scan of in[0, count) into out; count = *d_count (clamped to max_n) when d_count
// is non-null, else max_n. The grand total goes to *d_total when non-null.
size_t scan_tmp_bytes(uint32_t max_n);
void scan_exclusive(const uint32_t *in, uint32_t *out, uint32_t max_n, const uint32_t *d_count,
                    uint32_t *d_total, void *tmp, hipStream_t stream);

// Stable LSD sort of n (key, value) pairs by key bits [0, end_bit), radix_digits(end_bit)
// passes; the result is in (keys, vals) after an even number of passes and in
// (keys_alt, vals_alt) after an odd number.
size_t radix_tmp_bytes(uint32_t n);
int radix_digits(int end_bit);
int radix_bits();  // key bits per sort pass
// items: keys per thread of a tile (4, 8 or 16; a tile is 256 * items keys)
void radix_sort_pairs(uint64_t *keys, uint32_t *vals, uint64_t *keys_alt, uint32_t *vals_alt,
                      uint32_t n, int end_bit, void *tmp, hipStream_t stream, int items = 8);

}  // namespace gr
