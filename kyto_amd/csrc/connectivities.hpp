// connectivities() on the device (connectivities.hip).
#pragma once

#include "common.hpp"

namespace ky {

// scanpy's connectivities of a neighbour graph given in host memory (n rows of
// n_neighbors - 1 entries): the symmetric weighted graph in `out`; each row's sigma
// and rho into h_sigma / h_rho when they are not null.
// The device memory of ky_connectivities (a model of its allocations).
size_t connectivities_device_bytes(uint64_t n, uint32_t n_neighbors);
void connectivities_run(Context &ctx, uint64_t n, uint32_t n_neighbors, const uint32_t *h_indices,
                        const float *h_dist, Graph &out, float *h_sigma, float *h_rho, ky_conn_info *info);

}  // namespace ky
