#pragma once

#include "common.hpp"

namespace ky {

// Temporary bytes sort_pairs needs for n pairs.
size_t sort_pairs_bytes(size_t n);
// Stable sort of (key, value) pairs by the low `end_bit` bits of the key (the keys'
// other bits ignored), with sort_pairs_bytes(n) bytes of device memory at tmp.
void sort_pairs(hipStream_t st, const uint32_t *keys_in, uint32_t *keys_out, const uint32_t *vals_in,
                uint32_t *vals_out, size_t n, unsigned end_bit, void *tmp);
void sort_pairs(hipStream_t st, const uint32_t *keys_in, uint32_t *keys_out, const uint32_t *vals_in,
                uint32_t *vals_out, size_t n, unsigned end_bit, DevBuf<unsigned char> &tmp);

// Temporary bytes scan_counts needs for n counts.
size_t scan_counts_bytes(size_t n);
// offsets[i] = counts[0] + ... + counts[i - 1] for i <= n (u64).
void scan_counts(hipStream_t st, const uint32_t *counts, size_t n, uint64_t *offsets, void *tmp);
void scan_counts(hipStream_t st, uint32_t *counts, size_t n, uint64_t *offsets, DevBuf<unsigned char> &tmp);

}  // namespace ky
