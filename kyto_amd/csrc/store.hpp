// kyto stores on disk: the tiled layer reader and the embedding / graph writers and
// readers (store.cpp).
#pragma once

#include <cstdint>
#include <string>

#include "../../include/kyto_gpu.h"

namespace ky {

ky_layer *layer_open(const std::string &store, const std::string &name);
void layer_shape(const ky_layer &layer, uint64_t *n_cells, uint64_t *n_genes, uint64_t *nnz, uint64_t *cell_block,
                 uint64_t *gene_block, bool *counts);
void layer_close(ky_layer *layer);
int64_t layer_read_rows(ky_layer &layer, uint64_t first, uint64_t n, uint64_t *indptr, uint32_t *indices, float *values,
                        uint64_t capacity);

void write_layer(const std::string &store, const std::string &name, const uint64_t *indptr, const uint32_t *indices,
                 const void *values, bool counts, uint64_t n_cells, uint64_t n_genes, uint64_t C, uint64_t G, uint64_t K,
                 uint64_t S, int level);
void write_embedding(const std::string &store, const std::string &name, int axis, const float *values, uint64_t rows,
                     uint32_t width, uint64_t chunk_rows, int level);
void read_embedding(const std::string &store, const std::string &name, uint64_t *rows, uint32_t *width, float *out,
                    uint64_t capacity);
void write_graph(const std::string &store, const std::string &name, const uint64_t *indptr, const uint32_t *indices,
                 const float *weights, uint64_t n_cells, uint64_t nnz_chunk, uint64_t nnz_shard, int level);
void read_graph(const std::string &store, const std::string &name, uint64_t *n_cells, uint64_t *nnz, uint64_t *indptr,
                uint32_t *indices, float *weights);

}  // namespace ky
