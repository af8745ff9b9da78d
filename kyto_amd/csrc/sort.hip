// Device primitives from rocPRIM: a stable radix sort of (u32 key, u32 value) pairs
// (groups rows by k-means list with rows ascending within a list, the same order on
// every run) and an exclusive scan of counts into u64 offsets.
#include <hip/hip_runtime.h>

#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

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

void scan_counts(hipStream_t st, uint32_t *counts, size_t n, uint64_t *offsets, DevBuf<unsigned char> &tmp) {
    // counts[n] is read but does not reach the output: offsets[n] sums counts[0..n).
    KY_HIP(hipMemsetAsync(counts + n, 0, sizeof(uint32_t), st));
    size_t bytes = 0;
    KY_HIP(rocprim::exclusive_scan(nullptr, bytes, counts, offsets, (uint64_t)0, n + 1, rocprim::plus<uint64_t>(), st));
    tmp.reserve(std::max<size_t>(bytes, 1));
    KY_HIP(rocprim::exclusive_scan(tmp.ptr, bytes, counts, offsets, (uint64_t)0, n + 1, rocprim::plus<uint64_t>(), st));
}

}  // namespace ky
