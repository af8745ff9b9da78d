// pca() on the device (pca.hip).
#pragma once

#include "common.hpp"

namespace ky {

// Rows of one chunk of the column-major copy (the transposed product's unit).
constexpr uint32_t CHUNK_ROWS = 2048;
// estimate.rs:153: the sketch's extra columns.
constexpr uint32_t PCA_OVERSAMPLES = 10;

void ordered_sum(hipStream_t st, const double *partial, uint32_t n_parts, uint64_t stride, double *tmp, double *out);
// Maps applied on the device to the uploaded values before the layer is checked and
// its columns built: kyto's normalize_total(target_sum) and log1p.
struct LayerMaps {
    bool normalize = false;
    double target_sum = 0.0;
    bool log1p = false;
};
void csr_build(Csr &csr, const uint64_t *indptr, const uint32_t *indices, const float *values,
               const LayerMaps &maps = LayerMaps());
// The layer read block by block through pinned double buffers (pca.hip).
void csr_stream(Csr &csr, uint64_t block_rows, ky_block_reader reader, void *user, uint64_t nnz_hint,
                const LayerMaps &maps, ky_stream_info *info);
void pca_sketch(Csr &csr, uint64_t seed, uint64_t cell_offset, uint32_t l, double *sketch, double *gene_sums,
                double *omega_sums);
void pca_product(Csr &csr, const double *q, uint32_t w, const double *block_means, double *product, double *sums);
void pca_project(Csr &csr, const double *v, uint32_t k, const double *mean_projection, const double *means,
                 float *host_embedding, Dense **embedding, double *squares, uint64_t *stored);
// The device memory of a resident layer and one pca() over it (a model of the
// allocations: the rows and the chunked columns, the pass buffers, the basis and its
// products, the embedding), for a memory limit.
size_t pca_device_bytes(uint64_t n_rows, uint32_t n_cols, uint64_t nnz, uint32_t n_components, uint32_t q);
// The context's pass buffers released (for a memory limit between stages).
void pca_release(Context &ctx);
void pca_run(Csr &csr, uint32_t n_components, uint64_t seed, uint32_t q, float *host_embedding, Dense **embedding,
             float *loadings_out, double *variance, double *variance_ratio, double *means_out, uint32_t *passes_out,
             ky_pca_timing *timing);

}  // namespace ky
