// The learned IVF index (ivf.hip).
#pragma once

#include "common.hpp"

namespace ky {

ky_ivf *ivf_build(Context &ctx, const Dense &points, const ky_ivf_params &params, ky_ivf_info *info);
void ivf_free(ky_ivf *ivf);
// The index's k-means layout (null outputs skipped) and the router's trained alpha.
void ivf_layout(const ky_ivf &ivf, float *centroids, uint32_t *assign, uint32_t *list_ptr, uint32_t *list_rows,
                float *alpha);
void ivf_search(Context &ctx, const ky_ivf &ivf, const Dense &points, uint32_t n_neighbors, uint32_t n_probe,
                int use_router, float confidence, uint32_t *h_idx, float *h_dist, ky_ivf_search_info *info);

}  // namespace ky
