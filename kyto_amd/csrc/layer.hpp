// The input layer on the device (layer.hip).
#pragma once

#include "common.hpp"

namespace ky {

// In place on a resident CSR's values: normalize_total(target_sum) when `normalize`,
// then log1p when `log1p_values` (kyto's arithmetic, layer.hip).
void normalize_log1p(hipStream_t st, const uint64_t *indptr, uint64_t n_rows, float *values, double target_sum,
                     bool normalize, bool log1p_values);

}  // namespace ky
