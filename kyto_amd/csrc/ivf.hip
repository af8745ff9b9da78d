// The learned IVF index: k-means lists over the embedding and a router MLP.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "ivf.hpp"

struct ky_ivf {
    ky::Context *ctx = nullptr;
};

namespace ky {

ky_ivf *ivf_build(Context &, const Dense &, const ky_ivf_params &, ky_ivf_info *) {
    throw Error(KY_E501, "ivf: not built yet");
}

void ivf_free(ky_ivf *ivf) { delete ivf; }

void ivf_search(Context &, const ky_ivf &, const Dense &, uint32_t, uint32_t, int, float, uint32_t *, float *,
                ky_ivf_search_info *) {
    throw Error(KY_E501, "ivf: not built yet");
}

}  // namespace ky
