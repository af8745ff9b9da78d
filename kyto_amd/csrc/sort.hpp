#pragma once

#include "common.hpp"

namespace ky {

// Stable sort of (key, value) pairs by the low `end_bit` bits of the key.
void sort_pairs(hipStream_t st, const uint32_t *keys_in, uint32_t *keys_out, const uint32_t *vals_in,
                uint32_t *vals_out, size_t n, unsigned end_bit, DevBuf<unsigned char> &tmp);

}  // namespace ky
