// The input layer on the device: kyto's normalize_total(target_sum) then log1p on a
// counts CSR (crates/kyto-exec/src/elementwise.rs:1-24, normalize_divisor :262-269;
// Kernels::row_sums kernels.rs:130-138 and lane_sum; Kernels::divide_rows; log1p
// kernels/log1p.rs), in place on the resident values, so only the counts cross PCIe
// and the f32 layer never exists on the host.
//
// Per row: the total in f64 over kyto's 8 lanes (entry i of the row into lane i % 8,
// each lane summed in entry order, then the fixed tree (l0+l4)+(l2+l6) +
// ((l1+l5)+(l3+l7)), kernels.rs:290-299), divisor = f32(total) / f32(target_sum) (0
// becomes 1), each value f32(count) / divisor (a correctly rounded f32 division), then
// ln(1 + v) evaluated in f64 and rounded once to f32. One wavefront holds 8 rows: lane
// 8 r + L is the row's kyto lane L.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "device.hpp"
#include "layer.hpp"

namespace ky {

__global__ __launch_bounds__(256) void normalize_log1p_rows(const uint64_t *__restrict__ indptr, uint64_t n_rows,
                                                            float *__restrict__ values, double target_sum,
                                                            int normalize, int log1p_values) {
    const unsigned lane = lane_id(), L = lane % 8;
    const uint64_t r = ((uint64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64) * 8 + lane / 8;
    const bool live = r < n_rows;
    const uint64_t s = live ? indptr[r] : 0, e = live ? indptr[r + 1] : 0;
    float divisor = 1.0f;
    if (normalize) {
        double acc = 0.0;
        for (uint64_t i = s + L; i < e; i += 8) acc += (double)values[i];
        // The fixed tree over the row's 8 lanes (xor within the group of 8).
        const double h = acc + __shfl_xor(acc, 4, 8);   // L < 4: l_L + l_{L+4}
        const double q = h + __shfl_xor(h, 2, 8);       // L < 2: h_L + h_{L+2}
        const double total = q + __shfl_xor(q, 1, 8);  // L = 0: q0 + q1
        const double t0 = __shfl(total, 0, 8);
        divisor = (float)t0 / (float)target_sum;
        if (divisor == 0.0f) divisor = 1.0f;
    }
    if (!live) return;
    for (uint64_t i = s + L; i < e; i += 8) {
        float v = values[i];
        if (normalize) v = v / divisor;
        if (log1p_values) v = (float)log1p((double)v);
        values[i] = v;
    }
}

void normalize_log1p(hipStream_t st, const uint64_t *indptr, uint64_t n_rows, float *values, double target_sum,
                     bool normalize, bool log1p_values) {
    if (n_rows == 0 || (!normalize && !log1p_values)) return;
    const uint64_t waves = (n_rows + 7) / 8;
    normalize_log1p_rows<<<div_up(waves, 4), 256, 0, st>>>(indptr, n_rows, values, target_sum, normalize ? 1 : 0,
                                                           log1p_values ? 1 : 0);
    KY_HIP(hipGetLastError());
}

}  // namespace ky
