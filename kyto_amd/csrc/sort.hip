// Stable device radix sort of (u32 key, u32 value) pairs (rocPRIM): groups rows by
// k-means list with rows ascending within a list, the same order on every run.
#include <hip/hip_runtime.h>

#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>

#include "common.hpp"
#include "sort.hpp"

namespace ky {

void sort_pairs(hipStream_t st, const uint32_t *keys_in, uint32_t *keys_out, const uint32_t *vals_in,
                uint32_t *vals_out, size_t n, unsigned end_bit, DevBuf<unsigned char> &tmp) {
    size_t bytes = 0;
    KY_HIP(rocprim::radix_sort_pairs(nullptr, bytes, keys_in, keys_out, vals_in, vals_out, n, 0u, end_bit, st));
    tmp.reserve(std::max<size_t>(bytes, 1));
    KY_HIP(rocprim::radix_sort_pairs(tmp.ptr, bytes, keys_in, keys_out, vals_in, vals_out, n, 0u, end_bit, st));
}

}  // namespace ky
