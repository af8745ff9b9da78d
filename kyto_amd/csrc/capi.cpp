// The extern "C" boundary of libkyto_gpu.so (include/kyto_gpu.h): status codes,
// thread-local messages, no exception crossing the ABI.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "connectivities.hpp"
#include "dense.hpp"
#include "ivf.hpp"
#include "knn.hpp"
#include "pca.hpp"
#include "router.hpp"
#include "store.hpp"

namespace ky {

static thread_local std::string g_last_error;
// The stream of the context the current call works on (set by use_device): every call
// synchronises it before it returns, so no work of a call is pending afterwards (the
// caching allocator relies on it, memory.cpp).
static thread_local hipStream_t t_call_stream = nullptr;

void set_last_error(const std::string &message) { g_last_error = message; }
std::string last_error_text() { return g_last_error; }

template <typename F>
static int run_guarded(F &&f) {
    try {
        f();
        g_last_error.clear();
        return KY_OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "host memory exhausted";
        return KY_E505;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return KY_E505;
    } catch (...) {
        g_last_error = "unknown failure";
        return KY_E505;
    }
}

template <typename F>
static int guarded(F &&f) {
    int status = run_guarded(f);
    if (t_call_stream) {
        hipError_t e = hipStreamSynchronize(t_call_stream);
        t_call_stream = nullptr;
        if (e != hipSuccess && status == KY_OK) {
            g_last_error = std::string("device failure in hipStreamSynchronize: ") + hipGetErrorString(e);
            status = KY_E505;
        }
    }
    return status;
}

static void need(const void *p, const char *what) {
    if (!p) throw Error(KY_E215, std::string(what) + " is null");
}

static void use_device(const Context &ctx) {
    KY_HIP(hipSetDevice(ctx.device));
    t_call_stream = ctx.stream;
}

}  // namespace ky

using namespace ky;

extern "C" {

const char *ky_last_error(void) { return g_last_error.c_str(); }

int ky_device_count(int *count) {
    return guarded([&] {
        need(count, "count");
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        *count = e == hipSuccess ? n : 0;
    });
}

int ky_open(int device, ky_ctx **out) {
    return guarded([&] {
        need(out, "out");
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
            throw Error(KY_E504, "device " + std::to_string(device) + " was not found (" + std::to_string(n) +
                                     " visible)");
        ky_ctx *ctx = new ky_ctx();
        try {
            ctx->device = device;
            KY_HIP(hipSetDevice(device));
            KY_HIP(hipGetDeviceProperties(&ctx->props, device));
            ctx->n_cu = ctx->props.multiProcessorCount;
            KY_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
            KY_HIP(hipEventCreate(&ctx->timer[0]));
            KY_HIP(hipEventCreate(&ctx->timer[1]));
            // The device address of ky_neighbors_static's fixed region (once a device).
            static_scratch_init(device);
        } catch (...) {
            delete ctx;
            throw;
        }
        *out = ctx;
    });
}

int ky_close(ky_ctx *ctx) {
    return guarded([&] {
        if (!ctx) return;
        (void)hipSetDevice(ctx->device);
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        hipStream_t st = ctx->stream;
        delete ctx;
        if (st) (void)hipStreamDestroy(st);
        device_cache_trim();
    });
}

int ky_run_begin(ky_ctx *ctx, uint64_t total_rows) {
    return guarded([&] {
        need(ctx, "ctx");
        ctx->cancelled.store(false);
        ctx->rows_done.store(0);
        ctx->rows_total.store(total_rows);
    });
}

int ky_cancel(ky_ctx *ctx) {
    return guarded([&] {
        need(ctx, "ctx");
        ctx->cancelled.store(true);
    });
}

int ky_progress(const ky_ctx *ctx, uint64_t *done, uint64_t *total) {
    return guarded([&] {
        need(ctx, "ctx");
        if (done) *done = ctx->rows_done.load();
        if (total) *total = ctx->rows_total.load();
    });
}

int ky_set_threads(uint32_t threads) {
    return guarded([&] { host_threads(threads); });
}

int ky_trim(ky_ctx *ctx) {
    return guarded([&] {
        need(ctx, "ctx");
        KY_HIP(hipSetDevice(ctx->device));
        KY_HIP(hipStreamSynchronize(ctx->stream));
        pca_release(*ctx);
        device_cache_trim();
    });
}

int ky_pca_bytes(uint64_t n_rows, uint32_t n_cols, uint64_t nnz, uint32_t n_components, uint32_t power_iterations,
                 uint64_t *bytes) {
    return guarded([&] {
        need(bytes, "bytes");
        *bytes = pca_device_bytes(n_rows, n_cols, nnz, n_components, power_iterations);
    });
}

int ky_connectivities_bytes(uint64_t n_cells, uint32_t n_neighbors, uint64_t *bytes) {
    return guarded([&] {
        need(bytes, "bytes");
        KY_CHECK(n_neighbors >= 2, KY_E215, "connectivities: n_neighbors counts the cell itself (at least 2)");
        *bytes = connectivities_device_bytes(n_cells, n_neighbors);
    });
}

int ky_device_info(ky_ctx *ctx, char *name, size_t name_len, uint64_t *total_bytes) {
    return guarded([&] {
        need(ctx, "ctx");
        if (name && name_len) {
            std::strncpy(name, ctx->props.gcnArchName, name_len - 1);
            name[name_len - 1] = 0;
        }
        if (total_bytes) *total_bytes = ctx->props.totalGlobalMem;
    });
}

int ky_csr_upload(ky_ctx *ctx, const uint64_t *indptr, const uint32_t *indices, const float *values, uint64_t n_rows,
                  uint32_t n_cols, ky_csr **out) {
    return guarded([&] {
        need(ctx, "ctx");
        need(indptr, "indptr");
        need(out, "out");
        use_device(*ctx);
        if (indptr[n_rows] > 0) {
            need(indices, "indices");
            need(values, "values");
        }
        ky_csr *csr = new ky_csr();
        try {
            csr->ctx = ctx;
            csr->n_rows = n_rows;
            csr->n_cols = n_cols;
            csr_build(*csr, indptr, indices, values);
        } catch (...) {
            delete csr;
            throw;
        }
        *out = csr;
    });
}

int ky_csr_upload_counts(ky_ctx *ctx, const uint64_t *indptr, const uint32_t *indices, const float *counts,
                         uint64_t n_rows, uint32_t n_cols, double target_sum, int log1p, ky_csr **out) {
    return guarded([&] {
        need(ctx, "ctx");
        need(indptr, "indptr");
        need(out, "out");
        KY_CHECK(target_sum >= 0.0 && std::isfinite(target_sum), KY_E215,
                 "normalize_total: target_sum must be a finite number above 0 (or 0: no normalize_total)");
        use_device(*ctx);
        if (indptr[n_rows] > 0) {
            need(indices, "indices");
            need(counts, "counts");
        }
        ky_csr *csr = new ky_csr();
        try {
            csr->ctx = ctx;
            csr->n_rows = n_rows;
            csr->n_cols = n_cols;
            LayerMaps maps;
            maps.normalize = target_sum > 0.0;
            maps.target_sum = target_sum;
            maps.log1p = log1p != 0;
            csr_build(*csr, indptr, indices, counts, maps);
        } catch (...) {
            delete csr;
            throw;
        }
        *out = csr;
    });
}

int ky_csr_stream(ky_ctx *ctx, uint64_t n_rows, uint32_t n_cols, uint64_t block_rows, ky_block_reader reader,
                  void *user, uint64_t nnz_hint, double target_sum, int log1p, ky_csr **out,
                  ky_stream_info *info) {
    return guarded([&] {
        need(ctx, "ctx");
        need(out, "out");
        KY_CHECK(target_sum >= 0.0 && std::isfinite(target_sum), KY_E215,
                 "normalize_total: target_sum must be a finite number above 0 (or 0: no normalize_total)");
        use_device(*ctx);
        ky_csr *csr = new ky_csr();
        try {
            csr->ctx = ctx;
            csr->n_rows = n_rows;
            csr->n_cols = n_cols;
            LayerMaps maps;
            maps.normalize = target_sum > 0.0;
            maps.target_sum = target_sum;
            maps.log1p = log1p != 0;
            csr_stream(*csr, block_rows, reader, user, nnz_hint, maps, info);
        } catch (...) {
            delete csr;
            throw;
        }
        *out = csr;
    });
}

int ky_csr_values(const ky_csr *csr, float *values) {
    return guarded([&] {
        need(csr, "csr");
        need(values, "values");
        use_device(*csr->ctx);
        if (csr->nnz) download(*csr->ctx, values, csr->values.ptr, sizeof(float) * csr->nnz);
    });
}

int ky_csr_free(ky_csr *csr) {
    return guarded([&] {
        if (!csr) return;
        (void)hipSetDevice(csr->ctx->device);
        delete csr;
    });
}

int ky_csr_shape(const ky_csr *csr, uint64_t *n_rows, uint32_t *n_cols, uint64_t *nnz) {
    return guarded([&] {
        need(csr, "csr");
        if (n_rows) *n_rows = csr->n_rows;
        if (n_cols) *n_cols = csr->n_cols;
        if (nnz) *nnz = csr->nnz;
    });
}

int ky_dense_upload(ky_ctx *ctx, const float *host, uint64_t rows, uint32_t cols, ky_dense **out) {
    return guarded([&] {
        need(ctx, "ctx");
        need(host, "host");
        need(out, "out");
        use_device(*ctx);
        ky_dense *m = new ky_dense();
        try {
            m->ctx = ctx;
            m->rows = rows;
            m->cols = cols;
            m->data.alloc(rows * cols);
            upload_parts(*ctx, {{m->data.ptr, host, sizeof(float) * rows * cols}});
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
    });
}

int ky_dense_download(const ky_dense *m, float *host) {
    return guarded([&] {
        need(m, "matrix");
        need(host, "host");
        use_device(*m->ctx);
        download(*m->ctx, host, m->data.ptr, sizeof(float) * m->rows * m->cols);
    });
}

int ky_dense_shape(const ky_dense *m, uint64_t *rows, uint32_t *cols) {
    return guarded([&] {
        need(m, "matrix");
        if (rows) *rows = m->rows;
        if (cols) *cols = m->cols;
    });
}

int ky_dense_free(ky_dense *m) {
    return guarded([&] {
        if (!m) return;
        (void)hipSetDevice(m->ctx->device);
        delete m;
    });
}

int ky_pca_sketch(ky_ctx *ctx, const ky_csr *csr, uint64_t seed, uint64_t cell_offset, uint32_t l, double *sketch,
                  double *gene_sums, double *omega_sums) {
    return guarded([&] {
        need(ctx, "ctx");
        need(csr, "csr");
        need(sketch, "sketch");
        need(gene_sums, "gene_sums");
        need(omega_sums, "omega_sums");
        use_device(*ctx);
        pca_sketch(const_cast<ky_csr &>(*csr), seed, cell_offset, l, sketch, gene_sums, omega_sums);
    });
}

int ky_pca_product(ky_ctx *ctx, const ky_csr *csr, const double *q, uint32_t w, const double *block_means,
                   double *product, double *sums) {
    return guarded([&] {
        need(ctx, "ctx");
        need(csr, "csr");
        need(q, "q");
        need(block_means, "block_means");
        need(product, "product");
        need(sums, "sums");
        use_device(*ctx);
        pca_product(const_cast<ky_csr &>(*csr), q, w, block_means, product, sums);
    });
}

int ky_pca_project(ky_ctx *ctx, const ky_csr *csr, const double *v, uint32_t k, const double *mean_projection,
                   const double *means, float *host_embedding, ky_dense **embedding, double *squares,
                   uint64_t *stored) {
    return guarded([&] {
        need(ctx, "ctx");
        need(csr, "csr");
        need(v, "v");
        need(mean_projection, "mean_projection");
        need(means, "means");
        need(squares, "squares");
        need(stored, "stored");
        use_device(*ctx);
        Dense *emb = nullptr;
        pca_project(const_cast<ky_csr &>(*csr), v, k, mean_projection, means, host_embedding,
                    embedding ? &emb : nullptr, squares, stored);
        if (embedding) *embedding = static_cast<ky_dense *>(emb);
    });
}

int ky_pca(ky_ctx *ctx, const ky_csr *csr, uint32_t n_components, uint64_t seed, uint32_t power_iterations,
           float *host_embedding, ky_dense **embedding, float *loadings, double *variance, double *variance_ratio,
           double *means, uint32_t *passes_out, ky_pca_timing *timing) {
    return guarded([&] {
        need(ctx, "ctx");
        need(csr, "csr");
        need(variance, "variance");
        need(variance_ratio, "variance_ratio");
        use_device(*ctx);
        Dense *emb = nullptr;
        pca_run(const_cast<ky_csr &>(*csr), n_components, seed, power_iterations, host_embedding,
                embedding ? &emb : nullptr, loadings, variance, variance_ratio, means, passes_out, timing);
        if (embedding) *embedding = static_cast<ky_dense *>(emb);
    });
}

int ky_neighbors(ky_ctx *ctx, const ky_dense *points, uint32_t n_components, uint32_t n_neighbors, int metric,
                 int precision, uint32_t *indices, float *distances, ky_knn_info *info) {
    return guarded([&] {
        need(ctx, "ctx");
        need(points, "points");
        need(indices, "indices");
        need(distances, "distances");
        use_device(*ctx);
        neighbors_run(*ctx, *points, n_components, n_neighbors, metric, precision, 0, indices, distances, info);
    });
}

int ky_neighbors_rounds(ky_ctx *ctx, const ky_dense *points, uint32_t n_components, uint32_t n_neighbors, int metric,
                        int precision, uint64_t round_rows, uint32_t *indices, float *distances, ky_knn_info *info) {
    return guarded([&] {
        need(ctx, "ctx");
        need(points, "points");
        need(indices, "indices");
        need(distances, "distances");
        use_device(*ctx);
        neighbors_run(*ctx, *points, n_components, n_neighbors, metric, precision, round_rows, indices, distances,
                      info);
    });
}

int ky_neighbors_static(ky_ctx *ctx, const ky_dense *points, uint32_t n_components, uint32_t n_neighbors, int metric,
                        int precision, uint32_t *indices, float *distances, ky_knn_info *info) {
    return guarded([&] {
        need(ctx, "ctx");
        need(points, "points");
        need(indices, "indices");
        need(distances, "distances");
        use_device(*ctx);
        neighbors_static(*ctx, *points, n_components, n_neighbors, metric, precision, indices, distances, info);
    });
}

// The shape of a search over rows x cols (the checks of check_search on the sizes).
static KnnShape search_shape(ky_ctx *ctx, uint64_t rows, uint32_t cols, uint32_t n_components, uint32_t n_neighbors,
                             int metric, int precision, uint64_t round_rows, bool &packed) {
    Dense probe;
    probe.rows = rows;
    probe.cols = cols;
    const uint32_t d = check_search(probe, n_components, n_neighbors, metric);
    KY_CHECK(precision == KY_PRECISION_FP32 || precision == KY_PRECISION_BF16, KY_E215, "neighbors: unknown precision");
    packed = d != cols;
    if (ctx) use_device(*ctx);
    return knn_shape(ctx, rows, d, n_neighbors - 1, metric, precision, false, rows, round_rows);
}

int ky_neighbors_bytes(ky_ctx *ctx, uint64_t rows, uint32_t cols, uint32_t n_components, uint32_t n_neighbors,
                       int metric, int precision, uint64_t round_rows, uint64_t *bytes) {
    return guarded([&] {
        need(bytes, "bytes");
        bool packed = false;
        const KnnShape s = search_shape(ctx, rows, cols, n_components, n_neighbors, metric, precision, round_rows, packed);
        *bytes = knn_scratch_bytes(s, true, packed);
    });
}

int ky_neighbors_plan(ky_ctx *ctx, uint64_t rows, uint32_t cols, uint32_t n_components, uint32_t n_neighbors,
                      int metric, int precision, uint64_t limit, uint64_t *round_rows, uint64_t *bytes) {
    return guarded([&] {
        need(round_rows, "round_rows");
        need(bytes, "bytes");
        bool packed = false;
        KnnShape s = search_shape(ctx, rows, cols, n_components, n_neighbors, metric, precision, 0, packed);
        *round_rows = knn_plan_round(ctx, s.n, s.d, s.k, metric, precision, packed, limit, bytes);
        if (*round_rows == 0)
            throw Error(KY_E401, "the plan needs at least " + std::to_string(*bytes) + " bytes but memory_limit is " +
                                     std::to_string(limit) + " bytes (neighbors of " + std::to_string(rows) +
                                     " rows)");
    });
}

int ky_connectivities(ky_ctx *ctx, uint64_t n_cells, uint32_t n_neighbors, const uint32_t *indices,
                      const float *distances, ky_graph **out, float *sigma, float *rho, ky_conn_info *info) {
    return guarded([&] {
        need(ctx, "ctx");
        need(indices, "indices");
        need(distances, "distances");
        need(out, "out");
        use_device(*ctx);
        auto g = std::make_unique<ky_graph>();
        connectivities_run(*ctx, n_cells, n_neighbors, indices, distances, *g, sigma, rho, info);
        *out = g.release();
    });
}

int ky_graph_shape(const ky_graph *g, uint64_t *n_cells, uint64_t *edges) {
    return guarded([&] {
        need(g, "graph");
        if (n_cells) *n_cells = g->n;
        if (edges) *edges = g->nnz;
    });
}

int ky_graph_download(const ky_graph *g, uint64_t *indptr, uint32_t *indices, float *weights) {
    return guarded([&] {
        need(g, "graph");
        need(indptr, "indptr");
        need(indices, "indices");
        need(weights, "weights");
        use_device(*g->ctx);
        download_parts(*g->ctx, {{indptr, g->indptr.ptr, sizeof(uint64_t) * (g->n + 1)},
                                 {indices, g->indices.ptr, sizeof(uint32_t) * g->nnz},
                                 {weights, g->weights.ptr, sizeof(float) * g->nnz}});
    });
}

int ky_graph_free(ky_graph *g) {
    return guarded([&] {
        if (!g) return;
        (void)hipSetDevice(g->ctx->device);
        delete g;
    });
}

int ky_nearest_keys(ky_ctx *ctx, const float *queries, uint64_t n_queries, const ky_dense *points, int key,
                    const float *query_norms, const float *point_norms, const uint32_t *skip, uint32_t m, float *keys,
                    uint32_t *rows) {
    return guarded([&] {
        need(ctx, "ctx");
        need(points, "points");
        if (n_queries) {
            need(queries, "queries");
            need(skip, "skip");
            need(keys, "keys");
            need(rows, "rows");
        }
        use_device(*ctx);
        nearest_keys_run(*ctx, queries, n_queries, *points, key, query_norms, point_norms, skip, m, keys, rows);
    });
}

int ky_ivf_build(ky_ctx *ctx, const ky_dense *points, const ky_ivf_params *params, ky_ivf **out, ky_ivf_info *info) {
    return guarded([&] {
        need(ctx, "ctx");
        need(points, "points");
        need(params, "params");
        need(out, "out");
        use_device(*ctx);
        *out = ivf_build(*ctx, *points, *params, info);
    });
}

int ky_ivf_layout(const ky_ivf *ivf, float *centroids, uint32_t *assign, uint32_t *list_ptr, uint32_t *list_rows,
                  float *alpha) {
    return guarded([&] {
        need(ivf, "ivf");
        ivf_layout(*ivf, centroids, assign, list_ptr, list_rows, alpha);
    });
}

int ky_ivf_free(ky_ivf *ivf) {
    return guarded([&] { ivf_free(ivf); });
}

int ky_ivf_search(ky_ctx *ctx, const ky_ivf *ivf, const ky_dense *points, uint32_t n_neighbors, uint32_t n_probe,
                  int use_router, float confidence, uint32_t *indices, float *distances, ky_ivf_search_info *info) {
    return guarded([&] {
        need(ctx, "ctx");
        need(ivf, "ivf");
        need(points, "points");
        need(indices, "indices");
        need(distances, "distances");
        use_device(*ctx);
        ivf_search(*ctx, *ivf, *points, n_neighbors, n_probe, use_router, confidence, indices, distances, info);
    });
}

int ky_router_grad(ky_ctx *ctx, uint32_t d_in, uint32_t hidden, uint32_t n_out, const float *params, const float *x,
                   uint32_t batch, const uint32_t *t_idx, const float *t_w, uint32_t tk, float *grads, float *loss) {
    return guarded([&] {
        need(ctx, "ctx");
        need(params, "params");
        need(x, "x");
        need(t_idx, "t_idx");
        need(t_w, "t_w");
        need(grads, "grads");
        KY_CHECK(batch >= 1 && tk >= 1 && tk <= 64 && d_in >= 1 && hidden >= 1 && n_out >= 1, KY_E215,
                 "router: sizes out of range");
        use_device(*ctx);
        hipStream_t st = ctx->stream;
        Router R;
        R.init(*ctx, d_in, hidden, n_out, 0);
        KY_HIP(hipMemcpyAsync(R.params.ptr, params, sizeof(float) * R.n_params, hipMemcpyHostToDevice, st));
        DevBuf<float> xb((size_t)batch * d_in), h1((size_t)batch * hidden), h2((size_t)batch * hidden),
            lg((size_t)batch * n_out), dh1((size_t)batch * hidden), dh2((size_t)batch * hidden), ls(batch), tw((size_t)batch * tk);
        DevBuf<uint32_t> ti((size_t)batch * tk);
        KY_HIP(hipMemcpyAsync(xb.ptr, x, sizeof(float) * batch * d_in, hipMemcpyHostToDevice, st));
        KY_HIP(hipMemcpyAsync(ti.ptr, t_idx, sizeof(uint32_t) * batch * tk, hipMemcpyHostToDevice, st));
        KY_HIP(hipMemcpyAsync(tw.ptr, t_w, sizeof(float) * batch * tk, hipMemcpyHostToDevice, st));
        R.forward(st, xb.ptr, batch, h1.ptr, h2.ptr, lg.ptr);
        softmax_xent(st, lg.ptr, batch, (int)n_out, ti.ptr, tw.ptr, (int)tk, 1.0f / (float)batch, ls.ptr);
        R.backward(st, xb.ptr, batch, h1.ptr, h2.ptr, lg.ptr, dh2.ptr, dh1.ptr);
        KY_HIP(hipMemcpyAsync(grads, R.grads.ptr, sizeof(float) * R.n_params, hipMemcpyDeviceToHost, st));
        std::vector<float> hl(batch);
        KY_HIP(hipMemcpyAsync(hl.data(), ls.ptr, sizeof(float) * batch, hipMemcpyDeviceToHost, st));
        KY_HIP(hipStreamSynchronize(st));
        if (loss) {
            double s = 0.0;
            for (float v : hl) s += v;
            *loss = (float)(s / batch);
        }
    });
}

int ky_adam_step(ky_ctx *ctx, float *w, const float *g, float *m, float *v, uint64_t count, float lr, uint32_t step) {
    return guarded([&] {
        need(ctx, "ctx");
        need(w, "w");
        need(g, "g");
        need(m, "m");
        need(v, "v");
        KY_CHECK(step >= 1, KY_E215, "adam: step counts from 1");
        use_device(*ctx);
        hipStream_t st = ctx->stream;
        DevBuf<float> dw(count), dg(count), dm(count), dv(count);
        KY_HIP(hipMemcpyAsync(dw.ptr, w, sizeof(float) * count, hipMemcpyHostToDevice, st));
        KY_HIP(hipMemcpyAsync(dg.ptr, g, sizeof(float) * count, hipMemcpyHostToDevice, st));
        KY_HIP(hipMemcpyAsync(dm.ptr, m, sizeof(float) * count, hipMemcpyHostToDevice, st));
        KY_HIP(hipMemcpyAsync(dv.ptr, v, sizeof(float) * count, hipMemcpyHostToDevice, st));
        adam(st, dw.ptr, dg.ptr, dm.ptr, dv.ptr, (int64_t)count, lr, (int)step);
        KY_HIP(hipMemcpyAsync(w, dw.ptr, sizeof(float) * count, hipMemcpyDeviceToHost, st));
        KY_HIP(hipMemcpyAsync(m, dm.ptr, sizeof(float) * count, hipMemcpyDeviceToHost, st));
        KY_HIP(hipMemcpyAsync(v, dv.ptr, sizeof(float) * count, hipMemcpyDeviceToHost, st));
        KY_HIP(hipStreamSynchronize(st));
    });
}

int ky_dense_extend_basis(const double *basis, uint32_t rows, uint32_t bc, const double *block, uint32_t bw,
                          double *out, uint32_t *kept) {
    return guarded([&] {
        need(block, "block");
        need(out, "out");
        need(kept, "kept");
        Mat b(rows, bc), blk(rows, bw);
        if (bc) {
            need(basis, "basis");
            std::memcpy(b.v.data(), basis, sizeof(double) * rows * bc);
        }
        std::memcpy(blk.v.data(), block, sizeof(double) * rows * bw);
        Mat r = extend_basis(b, blk);
        *kept = (uint32_t)r.cols;
        for (uint32_t i = 0; i < rows; i++)
            for (uint32_t j = 0; j < r.cols; j++) out[(size_t)i * bw + j] = r.at(i, j);
    });
}

int ky_dense_rayleigh_ritz(const double *basis, const double *products, uint32_t rows, uint32_t cols, uint32_t k,
                           double *eigenvalues, double *loadings) {
    return guarded([&] {
        need(eigenvalues, "eigenvalues");
        need(loadings, "loadings");
        Mat b(rows, cols), p(rows, cols);
        if (cols) {
            need(basis, "basis");
            need(products, "products");
            std::memcpy(b.v.data(), basis, sizeof(double) * rows * cols);
            std::memcpy(p.v.data(), products, sizeof(double) * rows * cols);
        }
        std::vector<double> ev;
        Mat l;
        rayleigh_ritz(b, p, k, ev, l);
        for (uint32_t j = 0; j < k; j++) eigenvalues[j] = ev[j];
        std::memcpy(loadings, l.v.data(), sizeof(double) * rows * k);
    });
}

// ---- kyto stores on disk ----

int ky_store_layer_open(const char *store, const char *layer, ky_layer **out) {
    return guarded([&] {
        need(store, "store");
        need(layer, "layer");
        need(out, "out");
        *out = layer_open(store, layer);
    });
}

int ky_store_layer_shape(const ky_layer *layer, uint64_t *n_cells, uint32_t *n_genes, uint64_t *nnz,
                         uint64_t *cell_block, uint64_t *gene_block, int *counts) {
    return guarded([&] {
        need(layer, "layer");
        uint64_t n, g, z, c, gb;
        bool u32;
        layer_shape(*layer, &n, &g, &z, &c, &gb, &u32);
        KY_CHECK(g < (1ull << 32), KY_E215, "store: the layer has 2^32 or more genes");
        if (n_cells) *n_cells = n;
        if (n_genes) *n_genes = (uint32_t)g;
        if (nnz) *nnz = z;
        if (cell_block) *cell_block = c;
        if (gene_block) *gene_block = gb;
        if (counts) *counts = u32 ? 1 : 0;
    });
}

int64_t ky_store_layer_read(void *layer, uint64_t first_row, uint64_t n_rows, uint64_t *indptr, uint32_t *indices,
                            float *values, uint64_t capacity) {
    int64_t got = -1;
    int status = run_guarded([&] {
        need(layer, "layer");
        need(indptr, "indptr");
        got = layer_read_rows(*static_cast<ky_layer *>(layer), first_row, n_rows, indptr, indices, values, capacity);
    });
    return status == KY_OK ? got : -1;
}

int ky_store_layer_rows(ky_layer *layer, uint64_t first_row, uint64_t n_rows, uint64_t *indptr, uint32_t *indices,
                        float *values, uint64_t capacity, uint64_t *nnz) {
    return guarded([&] {
        need(layer, "layer");
        need(indptr, "indptr");
        need(nnz, "nnz");
        *nnz = (uint64_t)layer_read_rows(*layer, first_row, n_rows, indptr, indices, values, capacity);
    });
}

int ky_store_layer_close(ky_layer *layer) {
    return guarded([&] { layer_close(layer); });
}

int ky_csr_stream_layer(ky_ctx *ctx, ky_layer *layer, uint64_t block_rows, double target_sum, int log1p,
                        ky_csr **out, ky_stream_info *info) {
    return guarded([&] {
        need(ctx, "ctx");
        need(layer, "layer");
        need(out, "out");
        KY_CHECK(target_sum >= 0.0 && std::isfinite(target_sum), KY_E215,
                 "normalize_total: target_sum must be a finite number above 0 (or 0: no normalize_total)");
        uint64_t n, g, nnz, cb, gb;
        bool u32;
        layer_shape(*layer, &n, &g, &nnz, &cb, &gb, &u32);
        KY_CHECK(g >= 1 && g < (1ull << 32), KY_E215, "store: the layer must have 1 to 2^32 - 1 genes");
        if (block_rows == 0) block_rows = 8 * cb;
        use_device(*ctx);
        // The store's failure is kept and rethrown with its own code.
        struct Reader {
            ky_layer *layer;
            std::exception_ptr failure;
        } reader{layer, nullptr};
        ky_block_reader read = [](void *user, uint64_t first, uint64_t count, uint64_t *ip, uint32_t *ix, float *v,
                                  uint64_t cap) -> int64_t {
            Reader *r = static_cast<Reader *>(user);
            try {
                return layer_read_rows(*r->layer, first, count, ip, ix, v, cap);
            } catch (...) {
                r->failure = std::current_exception();
                return -1;
            }
        };
        ky_csr *csr = new ky_csr();
        try {
            csr->ctx = ctx;
            csr->n_rows = n;
            csr->n_cols = (uint32_t)g;
            LayerMaps maps;
            maps.normalize = target_sum > 0.0;
            maps.target_sum = target_sum;
            maps.log1p = log1p != 0;
            csr_stream(*csr, block_rows, read, &reader, nnz, maps, info);
        } catch (...) {
            delete csr;
            if (reader.failure) std::rethrow_exception(reader.failure);
            throw;
        }
        *out = csr;
    });
}

int ky_store_write_layer(const char *store, const char *name, uint64_t n_cells, uint64_t n_genes, const uint64_t *indptr,
                         const uint32_t *indices, const void *values, int counts, uint64_t cell_block,
                         uint64_t gene_block, uint64_t nnz_chunk, uint64_t nnz_shard, int zstd_level) {
    return guarded([&] {
        need(store, "store");
        need(name, "name");
        write_layer(store, name, indptr, indices, values, counts != 0, n_cells, n_genes, cell_block, gene_block, nnz_chunk,
                    nnz_shard, zstd_level);
    });
}

int ky_store_write_embedding(const char *store, const char *name, int axis, const float *values, uint64_t rows,
                             uint32_t width, uint64_t chunk_rows, int zstd_level) {
    return guarded([&] {
        need(store, "store");
        need(name, "name");
        write_embedding(store, name, axis, values, rows, width, chunk_rows, zstd_level);
    });
}

int ky_store_read_embedding(const char *store, const char *name, uint64_t *rows, uint32_t *width, float *values,
                            uint64_t capacity) {
    return guarded([&] {
        need(store, "store");
        need(name, "name");
        need(rows, "rows");
        need(width, "width");
        read_embedding(store, name, rows, width, values, capacity);
    });
}

int ky_store_write_graph(const char *store, const char *name, uint64_t n_cells, const uint64_t *indptr,
                         const uint32_t *indices, const float *weights, uint64_t nnz_chunk, uint64_t nnz_shard,
                         int zstd_level) {
    return guarded([&] {
        need(store, "store");
        need(name, "name");
        write_graph(store, name, indptr, indices, weights, n_cells, nnz_chunk, nnz_shard, zstd_level);
    });
}

int ky_store_read_graph(const char *store, const char *name, uint64_t *n_cells, uint64_t *edges, uint64_t *indptr,
                        uint32_t *indices, float *weights) {
    return guarded([&] {
        need(store, "store");
        need(name, "name");
        need(n_cells, "n_cells");
        need(edges, "edges");
        read_graph(store, name, n_cells, edges, indptr, indices, weights);
    });
}

}  // extern "C"
