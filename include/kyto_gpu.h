/*
 * kyto_gpu.h — C ABI of libkyto_gpu.so, the MI355X (gfx950) embedding path behind
 * kyto's pca() and neighbors().
 *
 * Conventions (those of kyto-gpu's DeviceKernels, crates/kyto-gpu/src/launch.rs:42-93):
 * - every call returns an int status: KY_OK (0), or a kyto error code (KY_E215 for
 *   invalid arguments or data, KY_E501 when no kernel serves the request, KY_E503 when
 *   cancelled, KY_E504 when the device is not found, KY_E505 for a device failure,
 *   launch.rs:62-69, errors.toml:790-845); ky_last_error() gives the message of the
 *   calling thread's last failure;
 * - host buffers are borrowed and caller-owned; outputs are written in place;
 * - no C++ exception crosses the ABI;
 * - a context is used from one host thread at a time (calls on it are serialised on
 *   its stream; each call returns when its results are in the caller's buffers).
 */
#ifndef KYTO_GPU_H
#define KYTO_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KY_OK 0
/* Storage errors of the store reader and writers (errors.toml:108-445). */
#define KY_E102 102 /* a member name breaks format.md section 2.1 */
#define KY_E103 103 /* a member of that name (case-insensitive) already exists */
#define KY_E104 104 /* a writer parameter (chunk, shard, zstd level) is out of range */
#define KY_E105 105 /* a size is above a limit of the format (section 4.5) */
#define KY_E106 106 /* a CSR given to the layer writer is not valid (or stores a zero) */
#define KY_E107 107 /* a stored layer breaks the tiled sparse format (section 4) */
#define KY_E109 109 /* no layer, embedding or graph of that name in the store */
#define KY_E110 110 /* the file system failed, or a Zarr node cannot be read */
#define KY_E115 115 /* the store's format_version is not one this library reads */
#define KY_E116 116 /* the path is not a kyto store */
#define KY_E126 126 /* a stored embedding or graph breaks sections 6 and 7 */
#define KY_E215 215 /* invalid argument or data (KytoError::InvalidArgument) */
#define KY_E401 401 /* a plan needs more memory than its limit (KytoError::MemoryLimit) */
#define KY_E501 501 /* no kernel for the placed step */
#define KY_E503 503 /* cancelled */
#define KY_E504 504 /* device not found */
#define KY_E505 505 /* device failure (KytoError::DeviceFailure) */

#define KY_METRIC_EUCLIDEAN 0
#define KY_METRIC_COSINE 1

/* Key precision of the candidate search (the exact re-rank is always f64). */
#define KY_PRECISION_FP32 0 /* plain FP32 keys */
#define KY_PRECISION_BF16 1 /* BF16 MFMA keys, FP32 accumulation */

typedef struct ky_ctx ky_ctx;     /* one device, one stream */
typedef struct ky_csr ky_csr;     /* a CSR layer resident in device memory */
typedef struct ky_dense ky_dense; /* a row-major f32 matrix resident in device memory */
typedef struct ky_ivf ky_ivf;     /* an IVF index (k-means lists + router) on the device */
typedef struct ky_graph ky_graph; /* a weighted cell graph (CSR) resident in device memory */
typedef struct ky_layer ky_layer; /* a tiled layer of a kyto store on disk, open for reading */

/* The calling thread's last error message ("" if none). Replaces the
 * KytoError::DeviceFailure payload (launch.rs:62-69). */
const char *ky_last_error(void);

/* Number of visible devices (kyto_gpu::probe, crates/kyto-gpu/src/probe.rs:19-33). */
int ky_device_count(int *count);

/* Open / close a device context (DeviceKernels::open, launch.rs:81-93). */
int ky_open(int device, ky_ctx **out);
int ky_close(ky_ctx *ctx);
/* Device name and total memory, for explain()/provenance (device.rs:24-64). */
int ky_device_info(ky_ctx *ctx, char *name, size_t name_len, uint64_t *total_bytes);

/* -------------------------------------------------------------- run control ---- */
/* kyto's RunControl (crates/kyto-exec/src/run_control.rs:1-80) on a context: a run's
 * calls check a cancel flag before each pass, streamed block and round of queries and
 * stop with KY_E503; they count the cell rows they computed (summed over the passes).
 * ky_cancel and ky_progress may be called from any thread while another uses the
 * context. */
int ky_run_begin(ky_ctx *ctx, uint64_t total_rows); /* clears the flag, done = 0 */
int ky_cancel(ky_ctx *ctx);
int ky_progress(const ky_ctx *ctx, uint64_t *done, uint64_t *total);
/* Host threads of the library's worker pool (the run's `threads`, pipeline.rs:478;
 * 0: every hardware thread, at most 32). Results do not depend on it. */
int ky_set_threads(uint32_t threads);
/* Release the context's pass buffers and the calling thread's cached device blocks
 * (between the stages of a run under a memory limit). */
int ky_trim(ky_ctx *ctx);
/* Device memory models for a memory limit (crates/kyto-plan/src/memory_limit.rs):
 * a resident layer and one ky_pca over it; ky_connectivities of n_cells rows. */
int ky_pca_bytes(uint64_t n_rows, uint32_t n_cols, uint64_t nnz, uint32_t n_components, uint32_t power_iterations,
                 uint64_t *bytes);
int ky_connectivities_bytes(uint64_t n_cells, uint32_t n_neighbors, uint64_t *bytes);

/* ---------------------------------------------------------------- CSR layers ---- */

/* Upload a CSR layer (cells x genes, indices strictly increasing within a row) and
 * keep it resident for every pass. Replaces the per-call SparseRows copies of
 * launch_matrix.rs:31-84 and the host CSC of launch_matrix.rs:405-424: the device
 * keeps the rows and a chunked column-major copy built on the device.
 * KY_E215 if a value is not finite or outside 1e-10..1e10 in magnitude
 * (pca/gpu.rs:103-105,120: GPU_RANGE), an index is out of range, or indptr is not
 * a row pointer; KY_E505 past 2^32 stored values ("pass a smaller memory_limit",
 * pca/gpu.rs:62-70). */
int ky_csr_upload(ky_ctx *ctx, const uint64_t *indptr, const uint32_t *indices,
                  const float *values, uint64_t n_rows, uint32_t n_cols, ky_csr **out);
/* Upload a counts CSR and make the layer on the device: normalize_total(target_sum)
 * when target_sum > 0, then log1p when `log1p` is 1, in place of the host producing
 * the f32 layer (crates/kyto-exec/src/elementwise.rs:1-24, normalize_divisor
 * :262-269, kernels/log1p.rs): row totals over kyto's 8 f64 lanes, divisor
 * f32(total) / f32(target_sum) (0 becomes 1), f32 division, ln(1 + v) in f64 rounded
 * once. KY_E215 for a negative or non-finite target_sum, and as ky_csr_upload. */
int ky_csr_upload_counts(ky_ctx *ctx, const uint64_t *indptr, const uint32_t *indices, const float *counts,
                         uint64_t n_rows, uint32_t n_cols, double target_sum, int log1p, ky_csr **out);
/* The resident layer's values (nnz floats, in row order): the layer a
 * ky_csr_upload_counts made, for a sink of the layer or a check of its bits. */
int ky_csr_values(const ky_csr *csr, float *values);
/* A store's row blocks: fill rows [first_row, first_row + n_rows) of the layer as a
 * CSR of their own — indptr (n_rows + 1 entries, from 0), then, when the block's
 * values fit `capacity`, its indices and values — and return the block's number of
 * values (a number above capacity asks for the block again with larger buffers; a
 * negative number is a failure). The buffers are pinned host memory of the library. */
typedef int64_t (*ky_block_reader)(void *user, uint64_t first_row, uint64_t n_rows, uint64_t *indptr,
                                   uint32_t *indices, float *values, uint64_t capacity);

typedef struct ky_stream_info {
    uint64_t blocks;  /* blocks read */
    uint64_t nnz;     /* values of the layer */
    double read_ms;   /* time spent in the reader (host) */
    double total_ms;  /* the whole upload, the layer's maps and column copy included */
    double bytes;     /* bytes copied to the device */
} ky_stream_info;

/* The layer read block by block (kyto-exec's block stream over a store, crates/
 * kyto-exec/src/pipeline.rs): `block_rows` cells a block through two pinned buffer
 * sets, the reader filling one while the other is copied to the device on a stream of
 * its own and the previous block's maps (normalize_total(target_sum) when
 * target_sum > 0, then log1p when `log1p` is 1, as ky_csr_upload_counts) run, so
 * reading, transfer and device work overlap and the host holds at most two blocks
 * (memory_limit). nnz_hint (0: estimated from the first block) sizes the resident
 * arrays. KY_E215 as ky_csr_upload; KY_E505 when the reader fails. */
int ky_csr_stream(ky_ctx *ctx, uint64_t n_rows, uint32_t n_cols, uint64_t block_rows, ky_block_reader reader,
                  void *user, uint64_t nnz_hint, double target_sum, int log1p, ky_csr **out,
                  ky_stream_info *info);
int ky_csr_free(ky_csr *csr);
int ky_csr_shape(const ky_csr *csr, uint64_t *n_rows, uint32_t *n_cols, uint64_t *nnz);

/* ---- kyto stores on disk (the reference's docs/format.md, store format 0.1) ---- */

/* Open layer `layer` of the kyto store at directory `store` for reading (replaces
 * kyto_store's TileReader::open, crates/kyto-store/src/tile_reader.rs:94-108, and
 * layer_open.rs): the root's format_version (KY_E116 not a store, KY_E115 another
 * version), schema.json, the four Zarr v3 arrays of the layer checked against it
 * (shapes, dtypes, sharding; KY_E107), tile_offsets read and checked (from 0,
 * non-decreasing, ending at nnz), the shard indexes of indices and data read (crc32c
 * checked). KY_E109 if the store has no such layer. A ky_layer is read from one
 * thread at a time (its decode buffers are reused between reads). */
int ky_store_layer_open(const char *store, const char *layer, ky_layer **out);
/* The layer's axes, stored values and tiling; `counts` is 1 for uint32 values. */
int ky_store_layer_shape(const ky_layer *layer, uint64_t *n_cells, uint32_t *n_genes, uint64_t *nnz,
                         uint64_t *cell_block, uint64_t *gene_block, int *counts);
/* A ky_block_reader over the layer (user = the ky_layer): rows [first_row,
 * first_row + n_rows) as a CSR of their own, decoded from the zstd inner chunks of
 * their cell blocks' tiles, gene blocks in increasing order, global gene indices
 * (format.md sections 4.7 and 4.8; tile_decode.rs:74-130), values as f32 (uint32
 * counts converted), each tile's indptr and local indices checked (KY_E107). With one
 * gene block and whole cell blocks the chunks decode straight into the caller's
 * (pinned) buffers. Returns the number of values, or -1 with ky_last_error() set. */
int64_t ky_store_layer_read(void *layer, uint64_t first_row, uint64_t n_rows, uint64_t *indptr, uint32_t *indices,
                            float *values, uint64_t capacity);
/* The same read with a status (the store's error codes kept): *nnz is the rows'
 * number of values; above `capacity` only indptr is written. */
int ky_store_layer_rows(ky_layer *layer, uint64_t first_row, uint64_t n_rows, uint64_t *indptr, uint32_t *indices,
                        float *values, uint64_t capacity, uint64_t *nnz);
int ky_store_layer_close(ky_layer *layer);
/* ky_csr_stream over a store layer (kyto-exec's block stream over a store,
 * crates/kyto-exec/src/pipeline.rs with layer_stream.rs): block_rows cells a block
 * (0: 8 cell blocks), the store's own error codes kept. */
int ky_csr_stream_layer(ky_ctx *ctx, ky_layer *layer, uint64_t block_rows, double target_sum, int log1p,
                        ky_csr **out, ky_stream_info *info);

/* Write layer `name` (a CSR of n_cells x n_genes: indptr n_cells + 1, global gene
 * indices strictly increasing in a row, no stored zeros; values float32, or uint32
 * counts when `counts` is 1) into `<store>/layers/` as tiles of cell_block x
 * gene_block (format.md section 4; kyto_store's tile_encode.rs and the layer arrays of
 * zarr_arrays.rs:97-140): per-tile indptr sharded a cell block, tile_offsets, and
 * indices / data sharded (nnz_shard entries a shard, nnz_chunk an inner chunk; zstd,
 * crc32c index). KY_E104 / KY_E105 for sizes out of range, KY_E106 for an invalid CSR
 * or a stored zero. */
int ky_store_write_layer(const char *store, const char *name, uint64_t n_cells, uint64_t n_genes, const uint64_t *indptr,
                         const uint32_t *indices, const void *values, int counts, uint64_t cell_block,
                         uint64_t gene_block, uint64_t nnz_chunk, uint64_t nnz_shard, int zstd_level);
/* Write embedding `name` (rows x width float32, row-major; axis 0 = cells, 1 = genes)
 * into `<store>/embeddings/` as format.md section 6 gives it (embedding_io.rs:194-240):
 * chunks of [chunk_rows, width], bytes + zstd(zstd_level), fill value 0, chunks of
 * nothing but zeros not stored. KY_E102 for a bad name, KY_E103 if the name exists,
 * KY_E104 for a level outside 1..22, KY_E105 for a chunk above 2^31 bytes. The caller
 * lists the member in schema.json (a sink writes schema.json and the root last). */
int ky_store_write_embedding(const char *store, const char *name, int axis, const float *values, uint64_t rows,
                             uint32_t width, uint64_t chunk_rows, int zstd_level);
/* Read a float32 embedding whole (embedding_io.rs read_embedding): its shape from
 * schema.json, checked against the array (KY_E126); values may be null to ask the
 * shape. */
int ky_store_read_embedding(const char *store, const char *name, uint64_t *rows, uint32_t *width, float *values,
                            uint64_t capacity);
/* Write cell graph `name` (CSR: indptr n_cells + 1, indices strictly increasing in a
 * row and below n_cells, float32 weights) into `<store>/graphs/` as format.md section 7
 * gives it (graph_io.rs:189-243): indptr uint64 in unsharded chunks of 65,536, indices
 * and data sharded (nnz_shard entries a shard, nnz_chunk an inner chunk, zstd, crc32c
 * index at the end). KY_E215 for an invalid CSR (Graph::check, graph_io.rs:71-105). */
int ky_store_write_graph(const char *store, const char *name, uint64_t n_cells, const uint64_t *indptr,
                         const uint32_t *indices, const float *weights, uint64_t nnz_chunk, uint64_t nnz_shard,
                         int zstd_level);
/* Read a float32 cell graph whole (graph_io.rs read_graph, :119-187): indptr may be
 * null to ask the shape; KY_E126 if it breaks section 7 or disagrees with schema.json. */
int ky_store_read_graph(const char *store, const char *name, uint64_t *n_cells, uint64_t *edges, uint64_t *indptr,
                        uint32_t *indices, float *weights);

/* Device-resident dense matrices (the PCA embedding, X_pca). */
int ky_dense_upload(ky_ctx *ctx, const float *host, uint64_t rows, uint32_t cols, ky_dense **out);
int ky_dense_download(const ky_dense *m, float *host);
int ky_dense_shape(const ky_dense *m, uint64_t *rows, uint32_t *cols);
int ky_dense_free(ky_dense *m);

/* ------------------------------------------------------------- PCA passes ---- */
/* One pass each over the resident layer, as kyto's GPU fold computes a leaf
 * (crates/kyto-exec/src/pca/gpu.rs:149-219), here with exact f32 x f32 products
 * accumulated in f64 and fixed-order (deterministic) reductions. */

/* Pass 0 (pca/gpu.rs:159-178; pca.rs:588-622): X^T Omega (genes x l, row-major),
 * the gene sums and Omega's column sums, Omega drawn on the device per cell from
 * (seed, cell_offset + row) by kyto's SplitMix64 rows (pca.rs:936-950) and rounded
 * to f32 (its column sums from the rounded values). */
int ky_pca_sketch(ky_ctx *ctx, const ky_csr *csr, uint64_t seed, uint64_t cell_offset,
                  uint32_t l, double *sketch, double *gene_sums, double *omega_sums);

/* Passes 1..q (pca/gpu.rs:179-191; pca.rs:623-645): each row's centred projection
 * w = x Q - block_means (rounded to f32), then X^T W (genes x w) and the sum of the
 * rows w (f64). Q is genes x w row-major, w <= 64 x 16. Replaces sparse_dense +
 * sparse_transposed_dense (launch_matrix.rs:183-238). */
int ky_pca_product(ky_ctx *ctx, const ky_csr *csr, const double *q, uint32_t w,
                   const double *block_means, double *product, double *sums);

/* The last pass (pca/gpu.rs:192-214; pca.rs:649-677): the embedding rows
 * (x - mu) V = x V - mean_projection (n x k, f32, left on the device in *embedding
 * when it is not NULL, and copied to `host_embedding` when that is not NULL), each
 * gene's centred squares over its stored values and their count. */
int ky_pca_project(ky_ctx *ctx, const ky_csr *csr, const double *v, uint32_t k,
                   const double *mean_projection, const double *means, float *host_embedding,
                   ky_dense **embedding, double *squares, uint64_t *stored);

/* The whole pca() (PcaAccumulator with GPU folds, pca.rs:302-862): the passes above
 * with kyto's end_pass / extend_basis / rayleigh_ritz in FP64 on the host
 * (pca_dense.rs:105-174). n_components >= 1; power_iterations = q (kyto's constant
 * is 9, physical_plan.rs:73). Outputs: embedding (device, and host if not NULL),
 * loadings f32 genes x k, variance and variance_ratio f64 k, means f64 genes.
 * `passes_out` (may be NULL) receives the product passes run (fewer than q when the
 * basis spans the data early). KY_E215 for fewer than 2 cells or n_components above
 * min(cells, genes) (pca.rs:876-920). */
typedef struct ky_pca_timing {
    double upload_ms;   /* unused by ky_pca (the layer is already resident) */
    double pass_ms[16]; /* device time of each pass */
    double dense_ms;    /* host FP64 steps */
    double total_ms;
    uint32_t n_passes;
    uint32_t spmm_launches[2]; /* launches of the row product X M and the column product X^T W */
    double spmm_ms[2];         /* their summed device time (HIP events on the library's stream) */
    double spmm_bytes[2];      /* their summed algorithmic bytes: the sparse operand (8 B a
                                  stored value + pointers), the dense operand and the output */
} ky_pca_timing;

int ky_pca(ky_ctx *ctx, const ky_csr *csr, uint32_t n_components, uint64_t seed,
           uint32_t power_iterations, float *host_embedding, ky_dense **embedding,
           float *loadings, double *variance, double *variance_ratio, double *means,
           uint32_t *passes_out, ky_pca_timing *timing);

/* ------------------------------------------------------------- neighbours ---- */

/* Statistics of a search. */
typedef struct ky_knn_info {
    uint64_t candidates;    /* candidates kept per query (m) */
    uint64_t certified;     /* rows whose candidates were proven to hold the k nearest */
    uint64_t fallback;      /* rows searched again exactly against every row */
    double search_ms, rerank_ms, fallback_ms, total_ms;
    double scan_kernel_ms;  /* BF16: device time of the candidate scan launches (HIP events) */
    uint64_t rounds;        /* rounds of queries the candidate search ran in */
    uint64_t scratch_bytes; /* device scratch of the call (one region, every buffer of the search) */
    uint64_t rescanned;     /* BF16: rows left with fewer than m candidates (at least 256 of them),
                               scanned again from twice their starting threshold before any exact
                               search */
    double rescan_ms;
} ky_knn_info;

/* Exact neighbours of every row of `points` (its first `n_components` columns), as
 * kyto's exact search ranks them (neighbors.rs:226-312): for each row the
 * n_neighbors - 1 nearest other rows by (f64 key in kyto's 8-lane order, row), the
 * row itself excluded by position. Candidates come from the device's key search at
 * `precision`; each is re-ranked with the exact f64 key on the device, rows whose
 * candidates cannot be shown to hold their neighbours (neighbors_gpu.rs:50-54,
 * 132-159) are searched again exactly on the device. Output: CSR-ordered graph
 * (neighbors.rs:412-435): per row, indices increasing (n x (n_neighbors-1) u32) and
 * f32 distances (sqrt of the f64 key for Euclidean). Replaces nearest_keys
 * (launch_matrix.rs:253-286) and neighbors_gpu::exact (neighbors_gpu.rs:70-127). */
int ky_neighbors(ky_ctx *ctx, const ky_dense *points, uint32_t n_components,
                 uint32_t n_neighbors, int metric, int precision, uint32_t *indices,
                 float *distances, ky_knn_info *info);

/* ky_neighbors with the candidate search run in rounds of `round_rows` queries (0:
 * every row in one round; rounded down to the kernel's 256 or 64 queries), so its
 * device scratch (ky_neighbors_bytes) fits a memory limit. The same graph. */
int ky_neighbors_rounds(ky_ctx *ctx, const ky_dense *points, uint32_t n_components, uint32_t n_neighbors,
                        int metric, int precision, uint64_t round_rows, uint32_t *indices, float *distances,
                        ky_knn_info *info);
/* The device memory ky_neighbors_rounds takes beyond its input: its scratch region
 * (every buffer of the search and the graph) for `rows` rows of an embedding of `cols`
 * columns, `n_components` of them used (0: all). ctx may be NULL: the device's resident
 * blocks are then taken as an MI355X's. */
int ky_neighbors_bytes(ky_ctx *ctx, uint64_t rows, uint32_t cols, uint32_t n_components, uint32_t n_neighbors,
                       int metric, int precision, uint64_t round_rows, uint64_t *bytes);
/* The largest round whose bytes fit `limit` (memory_limit's search over block sizes,
 * crates/kyto-plan/src/memory_limit.rs:119-175): *round_rows and its *bytes; KY_E401
 * ("the plan needs at least ... bytes") when even the smallest round does not fit,
 * with *bytes the smallest. */
int ky_neighbors_plan(ky_ctx *ctx, uint64_t rows, uint32_t cols, uint32_t n_components, uint32_t n_neighbors,
                      int metric, int precision, uint64_t limit, uint64_t *round_rows, uint64_t *bytes);

/* Bytes of the fixed device region ky_neighbors_static works in (one a device). */
#define KY_STATIC_SCRATCH_BYTES (64u << 20)
/* No counterpart in the reference: ky_neighbors for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Every buffer of the search and the graph lives in a fixed device
 * array of KY_STATIC_SCRATCH_BYTES per device, set up by ky_open and held until the
 * process ends; the call allocates nothing (no hipMalloc, no pinned staging: the graph
 * is copied straight into the caller's host arrays) and runs its candidate search in
 * the largest round of queries that fits the array. KY_E215, with the bytes needed,
 * when even the smallest round does not fit (at 50 components and n_neighbors 15,
 * embeddings up to about 60,000 rows fit). The same graph as ky_neighbors. Calls that
 * overlap on one device (two threads, two contexts) share the array and race: one at
 * a time a device. */
int ky_neighbors_static(ky_ctx *ctx, const ky_dense *points, uint32_t n_components, uint32_t n_neighbors,
                        int metric, int precision, uint32_t *indices, float *distances, ky_knn_info *info);

/* Pairwise keys of ky_nearest_keys (launch_matrix.rs:150-161, PairwiseKey). */
#define KY_KEY_SQUARED_EUCLIDEAN 0 /* sum of squared differences in component order, f32 */
#define KY_KEY_COSINE 1            /* 1 - dot / (|q| |p|) in f32, clipped to [0, 2], 1 for a zero norm */

/* The candidate search alone (DeviceKernels::nearest_keys, launch_matrix.rs:240-286, the
 * keys of kernels_matrix.rs:140-257): for each of the n_queries rows of `queries` (host,
 * row-major, points' width a row), its m smallest keys against every row of `points`,
 * the point skip[i] left out (UINT32_MAX for none), ranked by (key, point), smallest
 * first: keys f32 and rows u32, n_queries x m (host). The cosine key takes each query's
 * and each point's f32 norm. KY_E215 for m outside 1..the points left after skip (as
 * the reference's panics), a missing norm or an unknown key; KY_E501 past 64 keys a
 * query or 64 components. */
int ky_nearest_keys(ky_ctx *ctx, const float *queries, uint64_t n_queries, const ky_dense *points, int key,
                    const float *query_norms, const float *point_norms, const uint32_t *skip, uint32_t m,
                    float *keys, uint32_t *rows);

/* --------------------------------------------------------- connectivities ---- */

typedef struct ky_conn_info {
    uint64_t edges;          /* stored entries of the symmetric graph */
    uint64_t zero_rho_rows;  /* rows whose distances are all 0 (sigma from the global floor) */
    double upload_ms;        /* the neighbour graph to the device */
    double scales_ms;        /* rho, sigma and the directed memberships */
    double union_ms;         /* transpose and fuzzy union */
    double total_ms;
} ky_conn_info;

/* scanpy's connectivities of a neighbour graph (UMAP's fuzzy simplicial set with
 * set_op_mix_ratio 1 and local_connectivity 1): connectivities (connectivities.rs:76-123),
 * smooth_knn_distance (:125-168), membership (:179-188), fuzzy_union (:190-256).
 * Input as kyto's NeighborGraph (neighbors.rs:52-175), checked as from_csr checks it
 * (:68-147, KY-E215): n_cells rows of n_neighbors - 1 entries (host arrays,
 * n_cells x (n_neighbors - 1)), columns increasing in each row and below n_cells,
 * distances finite and >= 0. Output: the symmetric weighted graph (columns increasing
 * in each row, float32 weights in (0, 1], no stored zeros) resident on the device;
 * each row's sigma and rho (float32, n_cells) when `sigma` / `rho` are not NULL. */
int ky_connectivities(ky_ctx *ctx, uint64_t n_cells, uint32_t n_neighbors, const uint32_t *indices,
                      const float *distances, ky_graph **out, float *sigma, float *rho,
                      ky_conn_info *info);
/* ConnectivityGraph's accessors (connectivities.rs:44-71): rows and stored edges, and
 * a copy of the CSR into caller buffers (indptr n_cells + 1, indices and weights
 * `edges` entries). */
int ky_graph_shape(const ky_graph *g, uint64_t *n_cells, uint64_t *edges);
int ky_graph_download(const ky_graph *g, uint64_t *indptr, uint32_t *indices, float *weights);
int ky_graph_free(ky_graph *g);

/* ------------------------------------------------------ learned IVF router ---- */
/* No counterpart in the reference (its approximate search is an HNSW index,
 * neighbors_hnsw.rs): an IVF index over the embedding with a learned router. */

typedef struct ky_ivf_params {
    uint32_t n_lists;       /* k-means lists (4096) */
    uint32_t kmeans_iters;  /* Lloyd iterations */
    uint64_t seed;
    uint32_t hidden;        /* router hidden width (256) */
    uint32_t train_queries; /* sampled training queries (200,000) */
    uint32_t batch;         /* 4096 */
    uint32_t epochs;
    float lr;               /* Adam learning rate (1e-3) */
    uint32_t label_k;       /* neighbours whose lists label a query (n_neighbors - 1) */
} ky_ivf_params;

typedef struct ky_ivf_info {
    double kmeans_ms, labels_ms, train_ms;
    float final_loss;
    float label_recall_top;  /* share of label mass in the router's top-probe lists */
} ky_ivf_info;

int ky_ivf_build(ky_ctx *ctx, const ky_dense *points, const ky_ivf_params *params,
                 ky_ivf **out, ky_ivf_info *info);
int ky_ivf_free(ky_ivf *ivf);
/* The index's k-means layout (any output may be NULL): centroids f32 n_lists x d,
 * each row's list (n), the lists' offsets (n_lists + 1) and rows (n, ascending within
 * a list), and the router's alpha (its logits are alpha (2 x.c - |c|^2) plus the MLP's
 * correction). */
int ky_ivf_layout(const ky_ivf *ivf, float *centroids, uint32_t *assign, uint32_t *list_ptr, uint32_t *list_rows,
                  float *alpha);

typedef struct ky_ivf_search_info {
    uint64_t fallback;      /* queries below the confidence threshold searched exactly */
    double route_ms;        /* router (or centroid) scores and the top lists */
    double search_ms;       /* exact keys over the probed lists */
    double fallback_ms;     /* the plain method on the fallback queries */
    double graph_ms;        /* graph assembly and download */
    double total_ms;
} ky_ivf_search_info;

/* Approximate neighbours: each row's top `n_probe` lists (from the router when
 * `use_router` is 1, else the nearest centroids), their points re-ranked exactly
 * (f64 keys, (key, row) order), rows whose router confidence (probability mass of
 * the probed lists) is below `confidence` searched exactly instead. Output as
 * ky_neighbors. */
int ky_ivf_search(ky_ctx *ctx, const ky_ivf *ivf, const ky_dense *points, uint32_t n_neighbors,
                  uint32_t n_probe, int use_router, float confidence, uint32_t *indices,
                  float *distances, ky_ivf_search_info *info);

/* One training step's pieces on the device, exposed so the router's kernels can be
 * checked against an FP64 CPU restatement (no counterpart in the reference).
 * params: [W1 (d_in x hidden), b1, W2 (hidden x hidden), b2, W3 (hidden x n_out), b3]
 * (f32, row-major); x: batch x d_in; targets as tk (list, weight) pairs a row.
 * Writes the gradients (same layout) of the mean cross-entropy and the mean loss. */
int ky_router_grad(ky_ctx *ctx, uint32_t d_in, uint32_t hidden, uint32_t n_out, const float *params,
                   const float *x, uint32_t batch, const uint32_t *t_idx, const float *t_w, uint32_t tk,
                   float *grads, float *loss);

/* One Adam step (lr, beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected at `step`)
 * on host arrays w, m, v (updated in place) with gradient g. */
int ky_adam_step(ky_ctx *ctx, float *w, const float *g, float *m, float *v, uint64_t count, float lr,
                 uint32_t step);

/* -------------------------------------------- host-only FP64 dense steps ---- */
/* The FP64 steps of the PCA on the host, exported so they can be checked without a
 * device. Row-major matrices. */

/* extend_basis (pca_dense.rs:105-129): the orthonormal directions `block` (rows x bw)
 * adds to `basis` (rows x bc); writes up to bw columns to `out` (rows x bw, the
 * first *kept columns used) and *kept. */
int ky_dense_extend_basis(const double *basis, uint32_t rows, uint32_t bc, const double *block,
                          uint32_t bw, double *out, uint32_t *kept);

/* rayleigh_ritz (pca_dense.rs:141-174): eigenvalues (k) and loadings (rows x k). */
int ky_dense_rayleigh_ritz(const double *basis, const double *products, uint32_t rows,
                           uint32_t cols, uint32_t k, double *eigenvalues, double *loadings);

#ifdef __cplusplus
}
#endif

#endif /* KYTO_GPU_H */
