// neighbors() on the device (knn.hip, knn_bf16.hip).
#pragma once

#include "common.hpp"

namespace ky {

// The embedding prepared for the candidate search (neighbors_exact.rs:76-180).
struct KnnPrep {
    uint64_t n = 0;
    uint32_t d = 0, dp = 0;
    DevBuf<double> mean;   // d
    DevBuf<float> cent;    // n x dp, f32(x - mean), zero padded
    DevBuf<float> offs;    // n, f32(|a|^2)
    DevBuf<double> sqn;    // n, |a|^2 in f64
    DevBuf<double> slack;  // n, E_q
    double largest = 0.0;  // max |a|
    bool finite = true;
    bool filter_ok = true; // norms small enough for f32 keys
};

uint32_t candidates_for(uint32_t k, uint64_t n, int precision);
void knn_prepare(Context &ctx, const float *x, uint64_t n, uint32_t d, KnnPrep &prep);
void knn_candidates(Context &ctx, const KnnPrep &prep, uint64_t q_first, uint64_t q_count, uint32_t m,
                    uint32_t *d_rows, float *d_keys, uint32_t *d_count);
void knn_rerank(Context &ctx, const float *pts, const KnnPrep &prep, const uint32_t *q_rows, uint64_t q_first,
                uint64_t q_count, uint32_t m, uint32_t k, const uint32_t *cand_rows, const float *cand_keys,
                const uint32_t *cand_count, const float *cand_thr, int floor_mode, const double *bf_err,
                double bf_err_max, double bf_acc_rel, double *d_keys, uint32_t *d_rows, uint8_t *d_cert);
void exact_search(Context &ctx, const float *pts, uint64_t n, uint32_t d, uint32_t k, int precision,
                  const uint32_t *q_rows, uint64_t nq, double *keys, uint32_t *rows, ky_knn_info *info);
uint32_t check_search(const Dense &points, uint32_t n_components, uint32_t n_neighbors, int metric);
// seeded: d_keys / d_rows already hold, at each listed row, k exact (key, row) pairs
// ascending (a re-ranked candidate set, INFINITY where missing); the search keeps only
// rows at or below the k-th of them.
void knn_exact_rows(Context &ctx, const float *pts, uint64_t n, uint32_t d, uint32_t k, const uint32_t *d_list,
                    uint64_t count, double *d_keys, uint32_t *d_rows, bool seeded = false);
void knn_graph(Context &ctx, const double *d_keys, const uint32_t *d_rows, uint64_t n, uint32_t k, uint32_t *g_idx,
               float *g_dist);
void knn_pack(Context &ctx, const float *x, uint64_t n, uint32_t d, uint32_t ld, float *out);

void neighbors_run(Context &ctx, const Dense &points, uint32_t n_components, uint32_t n_neighbors, int metric,
                   int precision, uint32_t *h_idx, float *h_dist, ky_knn_info *info);
void nearest_keys_run(Context &ctx, const Dense &points, uint32_t n_components, uint64_t first, uint64_t count,
                      uint32_t m, int precision, float *h_keys, uint32_t *h_rows);

}  // namespace ky
