// BF16 MFMA candidate search (to come).
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "knn.hpp"

namespace ky {

void bf16_candidates(Context &, const float *, const KnnPrep &, uint32_t, uint32_t *, float *, uint32_t *,
                     DevBuf<double> &, double &, double &) {
    throw Error(KY_E501, "neighbors: no BF16 kernel yet");
}

}  // namespace ky
