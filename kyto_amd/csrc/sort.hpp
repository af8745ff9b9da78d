#pragma once

#include "common.hpp"

namespace ky {

// Stable sort of (key, value) pairs by the low `end_bit` bits of the key.
void sort_pairs(hipStream_t st, const uint32_t *keys_in, uint32_t *keys_out, const uint32_t *vals_in,
                uint32_t *vals_out, size_t n, unsigned end_bit, DevBuf<unsigned char> &tmp);

// offsets[i] = counts[0] + ... + counts[i - 1] for i <= n (counts holds n + 1 entries;
// the last is cleared).
void scan_counts(hipStream_t st, uint32_t *counts, size_t n, uint64_t *offsets, DevBuf<unsigned char> &tmp);

}  // namespace ky
