// neighbors() on the device (knn.hip, knn_bf16.hip).
//
// Every device buffer of a search comes from one Arena, taken in a fixed order, so
// the bytes a search needs are known before it runs (knn_scratch_bytes): the points'
// buffers for the whole call, then, one phase at a time, the temporaries of the
// preparation, of one round of queries, of one chunk of the exact search, and of the
// graph. Queries run in rounds of `round` rows, so a memory limit bounds the call by
// choosing the round.
#pragma once

#include "common.hpp"

namespace ky {

// What a search is over (its device memory follows from these alone).
struct KnnShape {
    uint64_t n = 0;        // points (rows of the embedding)
    uint32_t d = 0;        // components used
    uint32_t dp = 0;       // components padded for the FP32 kernel
    uint32_t k = 0;        // neighbours a row (n_neighbors - 1)
    uint32_t m = 0;        // candidates a row
    int metric = 0;        // KY_METRIC_EUCLIDEAN or KY_METRIC_COSINE
    int precision = 0;     // candidate keys (after the width switch)
    uint64_t nq = 0;       // queries
    bool list = false;     // queries are a list of rows (else every row, in order)
    uint64_t round = 0;    // queries a round
    uint32_t resident = 0; // search_bf16 blocks resident on the device at once
    uint64_t slots = 0;    // a BF16 launch's most query slots (queries x point segments; 0: any)
};

// The embedding prepared for the candidate search (neighbors_exact.rs:76-180).
struct KnnPrep {
    double *mean = nullptr;   // d (Euclidean)
    float *cent = nullptr;    // n x dp: f32(x - mean) (Euclidean) or f32(x / |x|) (cosine), zero padded
    float *offs = nullptr;    // n: f32(|a|^2)
    double *sqn = nullptr;    // n: |a|^2 in f64
    double *slack = nullptr;  // n: E_q
    double *norms = nullptr;  // n: |x| (cosine), else null
    unsigned long long *flags = nullptr;  // largest |a|^2 (bits), not-finite flag
    double largest = 0.0;     // max |a|
    bool finite = true;
    bool filter_ok = true;    // norms small enough for f32 keys
};

// The BF16 search's buffers over the points (knn_bf16.hip).
struct Bf16Points {
    uint32_t *perm = nullptr;  // scan order -> row
    void *P = nullptr, *Q = nullptr;  // augmented rows in scan order
    double *err = nullptr;     // 2n: err_r, then delta_r
    float *thr_est = nullptr;  // nq starting thresholds (null below EST_MIN_ROWS)
    uint32_t *qpos = nullptr;  // list: the queries' scan positions, ascending
    uint32_t *qorig = nullptr; // list: their rows
    unsigned long long *small = nullptr;  // maxima and per-component absolute maxima
    double err_max = 0.0, acc_rel = 0.0;
};

// One round's candidate buffers.
struct KnnRound {
    uint32_t *rows = nullptr, *count = nullptr;
    float *keys = nullptr, *thr = nullptr;
    void *buf = nullptr;  // BF16 append buffers
    uint32_t *seg_rows = nullptr, *seg_count = nullptr;
    float *seg_keys = nullptr, *seg_thr = nullptr;
    unsigned main_blocks = 0, n_seg = 1, seg_tiles = 0;
    // The symmetric BF16 scan: each row's list of keys from earlier rows.
    bool sym = false;
    uint32_t *tcount = nullptr;
    void *tlist = nullptr;
    float *thr_max = nullptr;  // the largest starting threshold of each super-block of points
};

// The BF16 rescan's buffers (knn_bf16.hip): the listed rows' scan positions, sorted.
struct Bf16Rescan {
    uint32_t *inv = nullptr, *k2 = nullptr, *v2 = nullptr, *qpos = nullptr, *qorig = nullptr;
    void *tmp = nullptr;
    float *thr = nullptr;  // starting thresholds (null: none)
};

uint32_t candidates_for(uint32_t k, uint64_t n, int precision);
int padded_width(uint32_t d);
// The shape of a search on this device (precision switched to FP32 past 58
// components; round 0 = every query in one round).
KnnShape knn_shape(const Context *ctx, uint64_t n, uint32_t d, uint32_t k, int metric, int precision, bool list,
                   uint64_t nq, uint64_t round);
// Device scratch of exact_search (and, with graph, of neighbors_run: its keys, rows,
// packed columns and graph) for a shape.
size_t knn_scratch_bytes(const KnnShape &s, bool graph, bool packed);

template <class A>
void knn_take_prep(A &a, const KnnShape &s, KnnPrep &p);
void knn_prepare(Context &ctx, const KnnShape &s, const float *x, KnnPrep &prep, double *partial);

// BF16 (knn_bf16.hip).
template <class A>
void bf16_take_points(A &a, const KnnShape &s, Bf16Points &b);
size_t bf16_prepare_bytes(const KnnShape &s);
void bf16_prepare(Context &ctx, const KnnShape &s, const float *pts, const KnnPrep &prep, const uint32_t *q_rows,
                  Bf16Points &b, void *temp, uint32_t k);
template <class A>
void bf16_take_round(A &a, const KnnShape &s, uint64_t count, KnnRound &r);
void bf16_round(Context &ctx, const KnnShape &s, const Bf16Points &b, uint64_t lo, uint64_t count,
                const KnnRound &r, double *scan_ms);
// The candidates of `count` listed rows (device list `rows`) again, from twice their
// starting thresholds, into r (rows qorig[i] of x); buffers from bf16_take_rescan.
template <class A>
void bf16_take_rescan(A &a, const KnnShape &s, uint64_t count, Bf16Rescan &x, KnnRound &r);
// The most device memory a rescan in rounds of `chunk` rows takes.
size_t bf16_rescan_bytes(const KnnShape &s, uint64_t chunk);
void bf16_rescan(Context &ctx, const KnnShape &s, const Bf16Points &b, const uint32_t *rows, uint64_t count,
                 const Bf16Rescan &x, const KnnRound &r);

// Exact neighbours of every row (or of the rows q_rows[0..nq)) of the packed n x d
// embedding, with its scratch from `a`: candidates at `precision`, exact re-rank and
// certificate, exact search again of the rows the certificate does not cover. keys /
// rows are n x k, filled at the queried rows.
void exact_search(Context &ctx, Arena &a, const float *pts, uint64_t n, uint32_t d, uint32_t k, int metric,
                  int precision, const uint32_t *q_rows, uint64_t nq, uint64_t round, double *keys, uint32_t *rows,
                  ky_knn_info *info, bool timed);
// The same with a scratch region of its own from the caching allocator.
void exact_search(Context &ctx, const float *pts, uint64_t n, uint32_t d, uint32_t k, int precision,
                  const uint32_t *q_rows, uint64_t nq, double *keys, uint32_t *rows, ky_knn_info *info);

uint32_t check_search(const Dense &points, uint32_t n_components, uint32_t n_neighbors, int metric);
void knn_graph(Context &ctx, const double *d_keys, const uint32_t *d_rows, uint64_t n, uint32_t k, int metric,
               uint32_t *g_idx, float *g_dist);
void knn_pack(Context &ctx, const float *x, uint64_t n, uint32_t d, uint32_t ld, float *out);

// neighbors(): the graph on the host. region / region_bytes: a caller's device region
// (nothing allocated, the graph downloaded straight into the host arrays), or null for
// a region of the call's own.
void neighbors_run(Context &ctx, const Dense &points, uint32_t n_components, uint32_t n_neighbors, int metric,
                   int precision, uint64_t round, uint32_t *h_idx, float *h_dist, ky_knn_info *info,
                   void *region = nullptr, size_t region_bytes = 0);
// The largest round of queries whose scratch (with the graph) fits `limit` bytes (every
// row: n), or 0 when none does; *bytes: that round's, or the smallest round's.
uint64_t knn_plan_round(const Context *ctx, uint64_t n, uint32_t d, uint32_t k, int metric, int precision, bool packed,
                        uint64_t limit, uint64_t *bytes);
// The fixed scratch region of the current device (set up by ky_open).
void static_scratch_init(int device);
// neighbors_run in the fixed region, in the largest round that fits it.
void neighbors_static(Context &ctx, const Dense &points, uint32_t n_components, uint32_t n_neighbors, int metric,
                      int precision, uint32_t *h_idx, float *h_dist, ky_knn_info *info);
// DeviceKernels::nearest_keys (launch_matrix.rs:253-286): each query's m smallest f32
// keys against every point, the point skip[i] left out, ranked by (key, point).
void nearest_keys_run(Context &ctx, const float *h_queries, uint64_t nq, const Dense &points, int key,
                      const float *h_qn, const float *h_pn, const uint32_t *h_skip, uint32_t m, float *h_keys,
                      uint32_t *h_rows);

}  // namespace ky
