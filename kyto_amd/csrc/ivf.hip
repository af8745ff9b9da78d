// The learned IVF index over an embedding (no counterpart in kyto, whose approximate
// search is an HNSW index, neighbors_hnsw.rs):
//
// - k-means (Lloyd) on a seeded sample of the rows: scores 2 x.c - |c|^2 by the BF16
//   GEMM (router.hip), each row's best list by row_top, centroids as the f64 means of
//   their rows in row order (a stable sort groups them: the same bits on every run);
//   then every row assigned once and the lists laid out contiguously (rows ascending
//   within a list, their points copied in list order);
// - labels: the exact k nearest neighbours (the exact search of knn.hip, itself
//   kyto's graph) of a seeded sample of queries; a query's target distribution puts
//   weight 1/k on the list of each of its neighbours;
// - the router: list logits alpha (2 x.c - |c|^2) + MLP((x - mean) / rms), the MLP
//   d -> hidden -> hidden -> n_lists a learned correction of the nearest-centroid
//   score (its output layer starts at zero, so the router starts as centroid probing;
//   alpha = exp(gamma) starts at the scale that best fits one batch's labels), trained
//   with the softmax cross-entropy against those targets by its own forward,
//   backward and Adam kernels (router.hip), gamma with them;
// - search: each row's top n_probe lists (router logits, or nearest centroids), the
//   probed lists' points scored with kyto's exact f64 key (8-lane order) and ranked
//   by (key, row); rows whose router confidence (the softmax mass of the probed lists)
//   is below a threshold, or whose lists hold fewer than k other rows, are searched
//   exactly instead.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "ivf.hpp"
#include "knn.hpp"
#include "router.hpp"
#include "sort.hpp"

struct ky_ivf {
    ky::Context *ctx = nullptr;
    uint64_t n = 0;
    uint32_t d = 0, n_lists = 0, label_k = 0;
    ky::DevBuf<float> centroids;  // n_lists x d
    ky::DevBuf<float> cent2;      // 2 c, n_lists x d
    ky::DevBuf<float> cnorm;      // -|c|^2
    ky::DevBuf<uint32_t> assign;  // n: each row's list
    ky::DevBuf<uint32_t> list_ptr;   // n_lists + 1
    ky::DevBuf<uint32_t> list_rows;  // n
    ky::DevBuf<float> list_pts;      // n x d, in list order
    ky::DevBuf<float> mean;          // d (router input centring)
    float inv_scale = 1.0f;
    ky::Router router;
    ky::DevBuf<float> gamma;         // [gamma, its gradient, Adam m, Adam v]: alpha = exp(gamma)
    float alpha = 0.0f;              // the trained alpha (for info)
};

namespace ky {

static uint64_t mix64h(uint64_t v) {
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ULL;
    v = (v ^ (v >> 27)) * 0x94D049BB133111EBULL;
    return v ^ (v >> 31);
}

__device__ __forceinline__ uint64_t mix64d(uint64_t v) {
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ULL;
    v = (v ^ (v >> 27)) * 0x94D049BB133111EBULL;
    return v ^ (v >> 31);
}

// ------------------------------------------------------------------ kernels ----

// keys[i] = a seeded hash of i (the sample order), vals[i] = i.
__global__ void hash_rows(uint64_t n, uint64_t seed, uint64_t stream, uint32_t *__restrict__ keys,
                          uint32_t *__restrict__ vals) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = (uint32_t)(mix64d(seed ^ mix64d(stream * 0x9E3779B97F4A7C15ULL + i)) >> 32);
    vals[i] = (uint32_t)i;
}

// out[i] = x[rows[i]] (d columns).
__global__ void gather_rows(const float *__restrict__ x, uint32_t d, const uint32_t *__restrict__ rows, uint64_t count,
                            float *__restrict__ out) {
    uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= count * d) return;
    uint64_t i = t / d;
    uint32_t j = (uint32_t)(t % d);
    out[t] = x[(uint64_t)rows[i] * d + j];
}

// out[i] = (x[rows ? rows[i] : i] - mean) * inv_scale.
__global__ void normalize_rows(const float *__restrict__ x, uint32_t d, const uint32_t *__restrict__ rows,
                               uint64_t count, const float *__restrict__ mean, float inv_scale,
                               float *__restrict__ out) {
    uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= count * d) return;
    uint64_t i = t / d;
    uint32_t j = (uint32_t)(t % d);
    uint64_t r = rows ? rows[i] : i;
    out[t] = (x[r * d + j] - mean[j]) * inv_scale;
}

// cent2 = 2 c and cnorm = -|c|^2 (f64 sum, rounded) for each list.
__global__ void centroid_prep(const float *__restrict__ c, uint32_t n_lists, uint32_t d, float *__restrict__ cent2,
                              float *__restrict__ cnorm) {
    uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n_lists) return;
    double s = 0.0;
    for (uint32_t j = 0; j < d; j++) {
        float v = c[(uint64_t)l * d + j];
        cent2[(uint64_t)l * d + j] = 2.0f * v;
        s += (double)v * (double)v;
    }
    cnorm[l] = (float)(-s);
}

// Segment bounds of sorted keys: ptr[key] = first position, ptr_end[key] = last + 1.
__global__ void segment_bounds(const uint32_t *__restrict__ keys, uint64_t n, uint32_t *__restrict__ start,
                               uint32_t *__restrict__ end) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t k = keys[i];
    if (i == 0 || keys[i - 1] != k) start[k] = (uint32_t)i;
    if (i == n - 1 || keys[i + 1] != k) end[k] = (uint32_t)(i + 1);
}

// Each list's centroid: the f64 mean of its rows (in sorted order); an empty list
// keeps its centroid. One block per list, one thread per component.
__global__ void centroid_update(const float *__restrict__ x, uint32_t d, const uint32_t *__restrict__ sorted_rows,
                                const uint32_t *__restrict__ start, const uint32_t *__restrict__ end,
                                float *__restrict__ c) {
    const uint32_t l = blockIdx.x, j = threadIdx.x;
    if (j >= d) return;
    const uint32_t s = start[l], e = end[l];
    if (e <= s) return;
    double acc = 0.0;
    for (uint32_t i = s; i < e; i++) acc += (double)x[(uint64_t)sorted_rows[i] * d + j];
    c[(uint64_t)l * d + j] = (float)(acc / (double)(e - s));
}

// list_ptr from segment bounds (exclusive: empty lists get start = end = previous end).
__global__ void list_pointer(const uint32_t *__restrict__ start, const uint32_t *__restrict__ end, uint32_t n_lists,
                             uint32_t *__restrict__ ptr) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    uint32_t at = 0;
    for (uint32_t l = 0; l < n_lists; l++) {
        ptr[l] = at;
        if (end[l] > start[l]) at = end[l];
    }
    ptr[n_lists] = at;
}

// Targets of a labelled query: the list of each of its k neighbours, weight 1/k.
__global__ void make_targets(const uint32_t *__restrict__ q_rows, uint64_t count, const uint32_t *__restrict__ nbr,
                             uint32_t k, const uint32_t *__restrict__ assign, uint32_t *__restrict__ t_idx,
                             float *__restrict__ t_w) {
    uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= count * k) return;
    uint64_t i = t / k;
    uint32_t e = (uint32_t)(t % k);
    uint32_t row = nbr[(uint64_t)q_rows[i] * k + e];
    t_idx[t] = assign[row];
    t_w[t] = 1.0f / (float)k;
}

// L += exp(gamma) S (the router's correction added to the scaled centroid scores).
__global__ void add_centroid_scores(float *__restrict__ L, const float *__restrict__ S, uint64_t count,
                                    const float *__restrict__ gamma) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < count) L[i] += expf(gamma[0]) * S[i];
}

// Per row (one wavefront): sum over the lists of dL S, lane j summing lists j, j + 64,
// ... in order, then a fixed butterfly.
__global__ __launch_bounds__(256) void row_dot(const float *__restrict__ dL, const float *__restrict__ S,
                                               uint64_t rows, uint32_t cols, float *__restrict__ out) {
    const uint64_t r = (uint64_t)blockIdx.x * 4 + threadIdx.x / 64;
    if (r >= rows) return;
    const unsigned lane = lane_id();
    float acc = 0.0f;
    for (uint32_t j = lane; j < cols; j += 64) acc += dL[r * cols + j] * S[r * cols + j];
    for (int o = 32; o >= 1; o /= 2) acc += __shfl_xor(acc, o);
    if (lane == 0) out[r] = acc;
}

// gamma's gradient: alpha times the rows' sum. One block of 256 threads: thread t sums
// its contiguous run of rows in order, then the runs' sums are added by a fixed
// binary tree (the same bits on every run).
__global__ __launch_bounds__(256) void gamma_grad(const float *__restrict__ row_sums, uint64_t rows,
                                                  float *__restrict__ gamma) {
    __shared__ float part[256];
    const unsigned t = threadIdx.x;
    const uint64_t per = (rows + 255) / 256, r0 = min<uint64_t>(rows, t * per), r1 = min<uint64_t>(rows, r0 + per);
    float s = 0.0f;
    for (uint64_t r = r0; r < r1; r++) s += row_sums[r];
    part[t] = s;
    __syncthreads();
    for (unsigned h = 128; h > 0; h >>= 1) {
        if (t < h) part[t] += part[t + h];
        __syncthreads();
    }
    if (t == 0) gamma[1] = expf(gamma[0]) * part[0];
}

// confidence = sum over the probed lists of exp(logit - lse).
__global__ void probe_confidence(const float *__restrict__ top_val, const float *__restrict__ lse, uint64_t rows,
                                 uint32_t P, float *__restrict__ conf) {
    uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (r >= rows) return;
    float s = 0.0f;
    for (uint32_t p = 0; p < P; p++) s += expf(top_val[r * P + p] - lse[r]);
    conf[r] = s;
}

// Label mass the router's top P lists hold, per query (sparse targets).
__global__ void label_mass(const uint32_t *__restrict__ top_idx, uint64_t rows, uint32_t P,
                           const uint32_t *__restrict__ t_idx, const float *__restrict__ t_w, uint32_t tk,
                           float *__restrict__ mass) {
    uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (r >= rows) return;
    float s = 0.0f;
    for (uint32_t e = 0; e < tk; e++) {
        uint32_t l = t_idx[r * tk + e];
        for (uint32_t p = 0; p < P; p++)
            if (top_idx[r * P + p] == l) {
                s += t_w[r * tk + e];
                break;
            }
    }
    mass[r] = s;
}

// ------------------------------------------------------------ list-major scan ----
//
// The probed lists' exact keys, list by list: the (query, probe) pairs are sorted by
// list, and a block takes up to PAIRS_A_BLOCK pairs of one list and stages the list's
// points in LDS once (a chunk of LIST_CHUNK points at a time), so a list's points
// cross from L2 once per block instead of once per query. Each wavefront takes the
// block's pairs in turn: lane = point, kyto's exact f64 key, the hits below the pair's
// current k-th pair appended to the wave's LDS buffer and compacted (device.hpp), the
// query itself left out; each pair's k smallest (key, row) pairs, ascending, go to
// out[(q * P + p) * k]. merge_probes then keeps each query's k smallest of its P
// lists' (every probed point at most once: the lists are disjoint).
constexpr unsigned LIST_CHUNK = 256;     // points staged in LDS at a time
constexpr unsigned PAIRS_A_BLOCK = 128;  // pairs a block (32 a wavefront)
constexpr unsigned PAIR_CAP = 128;       // a wavefront's candidate buffer

// kyto's exact key (kernels.rs:235-247, 290-299, 422-448; exact_sq_distance) of the
// query qd (f64, zero-padded to 8 DC components) and the point p (f32, zero-padded):
// a padding component adds (0 - 0)^2 = +0.0 to its lane, which changes no sum.
template <int DC>
__device__ __forceinline__ double padded_sq_distance(const double *qd, const float *p) {
    double l[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < DC; c++)
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const double t = qd[8 * c + u] - (double)p[8 * c + u];
            l[u] += t * t;
        }
    const double h0 = l[0] + l[4], h1 = l[1] + l[5], h2 = l[2] + l[6], h3 = l[3] + l[7];
    return (h0 + h2) + (h1 + h3);
}

// units[b] = (list, first pair, pair count) of block b. Lists of more than LIST_CHUNK
// points are staged a chunk at a time, again for every 4 pairs.
template <int DC>
__global__ __launch_bounds__(256) void ivf_pair_scan(const float *__restrict__ pts, uint32_t d,
                                                     const uint32_t *__restrict__ list_ptr,
                                                     const uint32_t *__restrict__ list_rows,
                                                     const float *__restrict__ list_pts,
                                                     const uint32_t *__restrict__ pair_code,
                                                     const uint4 *__restrict__ units, uint32_t P, uint32_t k,
                                                     double *__restrict__ out_keys, uint32_t *__restrict__ out_rows,
                                                     uint32_t *__restrict__ out_count) {
    constexpr unsigned DP = 8 * DC, PS = DP + 1;  // padded width; odd row stride (lanes' rows on different banks)
    __shared__ float chunk[LIST_CHUNK * PS];
    __shared__ double qd[4][DP];
    __shared__ unsigned long long bkey[4][PAIR_CAP];
    __shared__ uint32_t brow[4][PAIR_CAP];
    __shared__ uint32_t bcnt[4], trow[4];
    __shared__ unsigned long long tkey[4];
    const unsigned wave = threadIdx.x / 64, lane = lane_id();
    const uint4 u = units[blockIdx.x];
    const uint32_t L = u.x, first = u.y, cnt = u.z;
    const uint32_t s = list_ptr[L], e = list_ptr[L + 1], np = e - s;
    const uint32_t n_chunks = (np + LIST_CHUNK - 1) / LIST_CHUNK;
    // The chunk's points staged by the whole block, 8 loads a thread in flight.
    auto stage = [&](uint32_t c) {
        const uint32_t c0 = c * LIST_CHUNK, m = min(LIST_CHUNK, np - c0);
        const float *src = list_pts + (uint64_t)(s + c0) * d;
        const uint32_t total = m * DP;
        for (uint32_t t0 = threadIdx.x; t0 < total; t0 += 8 * blockDim.x) {
            float v[8];
#pragma unroll
            for (int w = 0; w < 8; w++) {
                const uint32_t t = t0 + w * blockDim.x, r = t / DP, j = t % DP;
                v[w] = t < total && j < d ? src[(uint64_t)r * d + j] : 0.0f;
            }
#pragma unroll
            for (int w = 0; w < 8; w++) {
                const uint32_t t = t0 + w * blockDim.x;
                if (t < total) chunk[(t / DP) * PS + t % DP] = v[w];
            }
        }
    };
    if (n_chunks == 1) stage(0);
    __syncthreads();
    const uint32_t slots = (cnt + 3) / 4;
    for (uint32_t slot = 0; slot < slots; slot++) {
        const uint32_t pi = slot * 4 + wave;
        const bool live = pi < cnt;
        const uint32_t code = live ? pair_code[first + pi] : 0u;
        const uint64_t q = code / P;
        if (live) {
            if (lane < DP) qd[wave][lane] = lane < d ? (double)pts[q * d + lane] : 0.0;
            if (lane == 0) {
                bcnt[wave] = 0;
                tkey[wave] = ~0ull;
                trow[wave] = ~0u;
            }
        }
        wave_sync();
        bool started = false;  // the pair's first 64 points went straight to its k smallest
        for (uint32_t c = 0; c < n_chunks; c++) {
            if (n_chunks > 1) {
                __syncthreads();
                stage(c);
                __syncthreads();
            }
            if (!live) continue;
            const uint32_t c0 = c * LIST_CHUNK, m = min(LIST_CHUNK, np - c0);
            for (uint32_t e0 = 0; e0 < m; e0 += 64) {
                const uint32_t j = e0 + lane;
                uint32_t row = 0xFFFFFFFFu;
                unsigned long long kb = ~0ull;
                if (j < m) {
                    row = list_rows[s + c0 + j];
                    kb = (unsigned long long)__double_as_longlong(padded_sq_distance<DC>(qd[wave], chunk + j * PS));
                }
                const bool valid = j < m && row != (uint32_t)q;
                if (!started) {
                    // The first round: its k smallest pairs sorted in registers.
                    if (!valid) {
                        kb = ~0ull;
                        row = 0xFFFFFFFFu;
                    }
                    const unsigned n_valid = (unsigned)__popcll(ballot(valid));
                    wave_sort64(kb, row);
                    const unsigned kept = min(k, n_valid);
                    if (lane < kept) {
                        bkey[wave][lane] = kb;
                        brow[wave][lane] = row;
                    }
                    const unsigned long long tk = __shfl(kb, (int)k - 1);
                    const uint32_t tr = __shfl(row, (int)k - 1);
                    if (lane == 0) {
                        bcnt[wave] = kept;
                        if (kept == k) {
                            tkey[wave] = tk;
                            trow[wave] = tr;
                        }
                    }
                    wave_sync();
                    started = true;
                    continue;
                }
                if (bcnt[wave] > PAIR_CAP - 64)
                    compact<unsigned long long, PAIR_CAP>(bkey[wave], brow[wave], &bcnt[wave], &tkey[wave],
                                                          &trow[wave], k);
                const unsigned long long tk = tkey[wave];
                const uint32_t tr = trow[wave];
                const bool hit = valid && (kb < tk || (kb == tk && row < tr));
                const uint64_t mask = ballot(hit);
                if (mask) {
                    const unsigned base = bcnt[wave];
                    if (hit) {
                        const unsigned pos = base + mbcnt(mask);
                        bkey[wave][pos] = kb;
                        brow[wave][pos] = row;
                    }
                    wave_sync();
                    if (lane == 0) bcnt[wave] = base + (unsigned)__popcll(mask);
                    wave_sync();
                }
            }
        }
        if (live) {
            // The pair's k smallest, sorted: in registers when the buffer holds at most
            // 64, else by the LDS sort of the whole buffer.
            unsigned have = bcnt[wave];
            unsigned long long kb = ~0ull;
            uint32_t row = 0xFFFFFFFFu;
            if (have <= 64) {
                if (lane < have) {
                    kb = bkey[wave][lane];
                    row = brow[wave][lane];
                }
                wave_sort64(kb, row);
            } else {
                compact<unsigned long long, PAIR_CAP>(bkey[wave], brow[wave], &bcnt[wave], &tkey[wave], &trow[wave],
                                                      k);
                have = bcnt[wave];
                if (lane < have) {
                    kb = bkey[wave][lane];
                    row = brow[wave][lane];
                }
            }
            const unsigned kept = min(k, have);
            if (lane < k) {
                out_keys[(uint64_t)code * k + lane] = lane < kept ? __longlong_as_double((long long)kb) : INFINITY;
                out_rows[(uint64_t)code * k + lane] = lane < kept ? row : 0xFFFFFFFFu;
            }
            if (lane == 0) out_count[code] = kept;
        }
        wave_sync();
    }
}

// Each query's k smallest (key, row) pairs of its P lists' sorted lists (one thread a
// query, a P-way merge).
__global__ void merge_probes(const double *__restrict__ pk, const uint32_t *__restrict__ pr,
                             const uint32_t *__restrict__ pc, uint64_t n, uint32_t P, uint32_t k,
                             double *__restrict__ keys, uint32_t *__restrict__ rows, uint32_t *__restrict__ count) {
    const uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (q >= n) return;
    uint32_t head[8];
    for (uint32_t p = 0; p < P; p++) head[p] = 0;
    uint32_t out = 0;
    for (; out < k; out++) {
        int best = -1;
        double bk = 0.0;
        uint32_t br = 0;
        for (uint32_t p = 0; p < P; p++) {
            const uint64_t code = q * P + p;
            if (head[p] >= pc[code]) continue;
            const double key = pk[code * k + head[p]];
            const uint32_t row = pr[code * k + head[p]];
            if (best < 0 || key < bk || (key == bk && row < br)) {
                best = (int)p;
                bk = key;
                br = row;
            }
        }
        if (best < 0) break;
        head[best]++;
        keys[q * k + out] = bk;
        rows[q * k + out] = br;
    }
    for (uint32_t t = out; t < k; t++) {
        keys[q * k + t] = INFINITY;
        rows[q * k + t] = 0xFFFFFFFFu;
    }
    count[q] = out;
}

// ==================================================================== host ====

struct Timer {
    hipStream_t st;
    hipEvent_t a, b;
    explicit Timer(hipStream_t s) : st(s) {
        KY_HIP(hipEventCreate(&a));
        KY_HIP(hipEventCreate(&b));
    }
    ~Timer() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    void start() { KY_HIP(hipEventRecord(a, st)); }
    double stop() {
        KY_HIP(hipEventRecord(b, st));
        KY_HIP(hipEventSynchronize(b));
        float ms = 0.0f;
        KY_HIP(hipEventElapsedTime(&ms, a, b));
        return ms;
    }
};

static unsigned bits_for(uint32_t v) {
    unsigned b = 1;
    while ((1ull << b) < v) b++;
    return b;
}

// Scores of `rows` rows of X (row stride d) against the centroids: 2 x.c - |c|^2.
static void centroid_scores(hipStream_t st, const ky_ivf &ivf, const float *X, int rows, float *S) {
    gemm(st, X, ivf.d, 1, ivf.cent2.ptr, 1, ivf.d, S, ivf.n_lists, rows, (int)ivf.n_lists, (int)ivf.d,
         ivf.cnorm.ptr, 1, false);
}

// Each of `count` rows of X: its best list (into assign), in chunks.
static void assign_rows(hipStream_t st, const ky_ivf &ivf, const float *X, uint64_t count, uint32_t *assign,
                        DevBuf<float> &scores, DevBuf<float> &topv) {
    const uint64_t chunk = std::max<uint64_t>(1, std::min<uint64_t>(count, (256ull << 20) / ivf.n_lists));
    scores.reserve(chunk * ivf.n_lists);
    topv.reserve(chunk);
    for (uint64_t r0 = 0; r0 < count; r0 += chunk) {
        const int rows = (int)std::min<uint64_t>(chunk, count - r0);
        centroid_scores(st, ivf, X + r0 * ivf.d, rows, scores.ptr);
        row_top(st, scores.ptr, rows, (int)ivf.n_lists, 1, assign + r0, topv.ptr, nullptr);
    }
}

ky_ivf *ivf_build(Context &ctx, const Dense &points, const ky_ivf_params &p, ky_ivf_info *info) {
    const uint64_t n = points.rows;
    const uint32_t d = points.cols;
    KY_CHECK(d >= 1 && d <= 64, KY_E501, "ivf: no device kernel for more than 64 components");
    KY_CHECK(p.n_lists >= 1 && p.n_lists <= 65536 && p.n_lists <= n, KY_E215,
             "ivf: n_lists must be from 1 to min(65536, rows)");
    KY_CHECK(p.label_k >= 1 && p.label_k <= 64 && p.label_k < n, KY_E215, "ivf: label_k out of range");
    KY_CHECK(p.batch >= 1 && p.hidden >= 1 && p.hidden <= 4096, KY_E215, "ivf: batch and hidden must be positive");
    hipStream_t st = ctx.stream;
    Timer timer(st);
    auto ivf = std::make_unique<ky_ivf>();
    ivf->ctx = &ctx;
    ivf->n = n;
    ivf->d = d;
    ivf->n_lists = p.n_lists;
    ivf->label_k = p.label_k;
    const float *x = points.data.ptr;
    DevBuf<unsigned char> sort_tmp;

    // ---- k-means on a seeded sample of 64 rows a list.
    timer.start();
    DevBuf<uint32_t> hkeys(n), hvals(n), skeys(n), order(n);
    hash_rows<<<div_up(n, 256), 256, 0, st>>>(n, p.seed, 1, hkeys.ptr, hvals.ptr);
    KY_HIP(hipGetLastError());
    sort_pairs(st, hkeys.ptr, skeys.ptr, hvals.ptr, order.ptr, n, 32, sort_tmp);  // the sample order
    const uint64_t S = std::min<uint64_t>(n, 64ull * p.n_lists);
    DevBuf<float> sample(S * d);
    gather_rows<<<div_up(S * d, 256), 256, 0, st>>>(x, d, order.ptr, S, sample.ptr);
    ivf->centroids.alloc((size_t)p.n_lists * d);
    ivf->cent2.alloc((size_t)p.n_lists * d);
    ivf->cnorm.alloc(p.n_lists);
    // The first n_lists sampled rows start the centroids.
    KY_HIP(hipMemcpyAsync(ivf->centroids.ptr, sample.ptr, sizeof(float) * p.n_lists * d, hipMemcpyDeviceToDevice, st));
    DevBuf<uint32_t> sa(S), sa_sorted(S), idx(S), idx_sorted(S), seg_start(p.n_lists), seg_end(p.n_lists);
    DevBuf<float> scores, topv;
    std::vector<uint32_t> iota_h(S);
    for (uint64_t i = 0; i < S; i++) iota_h[i] = (uint32_t)i;
    KY_HIP(hipMemcpyAsync(idx.ptr, iota_h.data(), sizeof(uint32_t) * S, hipMemcpyHostToDevice, st));
    const unsigned lbits = bits_for(p.n_lists);
    for (uint32_t it = 0; it < p.kmeans_iters; it++) {
        centroid_prep<<<div_up(p.n_lists, 256), 256, 0, st>>>(ivf->centroids.ptr, p.n_lists, d, ivf->cent2.ptr,
                                                              ivf->cnorm.ptr);
        assign_rows(st, *ivf, sample.ptr, S, sa.ptr, scores, topv);
        sort_pairs(st, sa.ptr, sa_sorted.ptr, idx.ptr, idx_sorted.ptr, S, lbits, sort_tmp);
        KY_HIP(hipMemsetAsync(seg_start.ptr, 0, sizeof(uint32_t) * p.n_lists, st));
        KY_HIP(hipMemsetAsync(seg_end.ptr, 0, sizeof(uint32_t) * p.n_lists, st));
        segment_bounds<<<div_up(S, 256), 256, 0, st>>>(sa_sorted.ptr, S, seg_start.ptr, seg_end.ptr);
        centroid_update<<<p.n_lists, 64, 0, st>>>(sample.ptr, d, idx_sorted.ptr, seg_start.ptr, seg_end.ptr,
                                                   ivf->centroids.ptr);
        KY_HIP(hipGetLastError());
    }
    centroid_prep<<<div_up(p.n_lists, 256), 256, 0, st>>>(ivf->centroids.ptr, p.n_lists, d, ivf->cent2.ptr,
                                                          ivf->cnorm.ptr);
    // Every row's list, and the lists laid out.
    ivf->assign.alloc(n);
    assign_rows(st, *ivf, x, n, ivf->assign.ptr, scores, topv);
    DevBuf<uint32_t> a_sorted(n);
    ivf->list_rows.alloc(n);
    sort_pairs(st, ivf->assign.ptr, a_sorted.ptr, hvals.ptr, ivf->list_rows.ptr, n, lbits, sort_tmp);
    KY_HIP(hipMemsetAsync(seg_start.ptr, 0, sizeof(uint32_t) * p.n_lists, st));
    KY_HIP(hipMemsetAsync(seg_end.ptr, 0, sizeof(uint32_t) * p.n_lists, st));
    segment_bounds<<<div_up(n, 256), 256, 0, st>>>(a_sorted.ptr, n, seg_start.ptr, seg_end.ptr);
    ivf->list_ptr.alloc(p.n_lists + 1);
    list_pointer<<<1, 1, 0, st>>>(seg_start.ptr, seg_end.ptr, p.n_lists, ivf->list_ptr.ptr);
    ivf->list_pts.alloc(n * d);
    gather_rows<<<div_up(n * d, 256), 256, 0, st>>>(x, d, ivf->list_rows.ptr, n, ivf->list_pts.ptr);
    KY_HIP(hipGetLastError());
    const double kmeans_ms = timer.stop();

    // ---- labels: the exact neighbours of a seeded sample of queries.
    timer.start();
    const uint64_t T = std::min<uint64_t>(n, std::max<uint32_t>(p.train_queries, 1));
    hash_rows<<<div_up(n, 256), 256, 0, st>>>(n, p.seed, 2, hkeys.ptr, hvals.ptr);
    sort_pairs(st, hkeys.ptr, skeys.ptr, hvals.ptr, order.ptr, n, 32, sort_tmp);
    const uint32_t k = p.label_k;
    DevBuf<double> nkeys(n * k);
    DevBuf<uint32_t> nrows(n * k);
    exact_search(ctx, x, n, d, k, KY_PRECISION_BF16, order.ptr, T, nkeys.ptr, nrows.ptr, nullptr);
    DevBuf<uint32_t> t_idx(T * k);
    DevBuf<float> t_w(T * k);
    make_targets<<<div_up(T * k, 256), 256, 0, st>>>(order.ptr, T, nrows.ptr, k, ivf->assign.ptr, t_idx.ptr, t_w.ptr);
    KY_HIP(hipGetLastError());
    const double labels_ms = timer.stop();

    // ---- router: input normalisation, then Adam on mini-batches of the labelled queries.
    timer.start();
    {
        // mean and rms of the rows (host f64 from a device copy of the sample: small)
        std::vector<float> hs(S * d);
        KY_HIP(hipMemcpyAsync(hs.data(), sample.ptr, sizeof(float) * S * d, hipMemcpyDeviceToHost, st));
        KY_HIP(hipStreamSynchronize(st));
        std::vector<double> mu(d, 0.0);
        for (uint64_t i = 0; i < S; i++)
            for (uint32_t j = 0; j < d; j++) mu[j] += hs[i * d + j];
        for (uint32_t j = 0; j < d; j++) mu[j] /= (double)S;
        double ss = 0.0;
        for (uint64_t i = 0; i < S; i++)
            for (uint32_t j = 0; j < d; j++) {
                double t = hs[i * d + j] - mu[j];
                ss += t * t;
            }
        double rms = std::sqrt(ss / (double)S);
        ivf->inv_scale = rms > 0 ? (float)(1.0 / rms) * std::sqrt((float)d) : 1.0f;
        std::vector<float> muf(d);
        for (uint32_t j = 0; j < d; j++) muf[j] = (float)mu[j];
        ivf->mean.alloc(d);
        KY_HIP(hipMemcpyAsync(ivf->mean.ptr, muf.data(), sizeof(float) * d, hipMemcpyHostToDevice, st));
    }
    Router &R = ivf->router;
    R.init(ctx, d, p.hidden, p.n_lists, p.seed, true);
    const uint32_t B = (uint32_t)std::min<uint64_t>(p.batch, T);
    DevBuf<float> xb((size_t)B * d), h1((size_t)B * p.hidden), h2((size_t)B * p.hidden), lg((size_t)B * p.n_lists),
        dh1((size_t)B * p.hidden), dh2((size_t)B * p.hidden), loss(B), xr((size_t)B * d),
        sc((size_t)B * p.n_lists), rsum(B);
    const uint64_t steps_per_epoch = std::max<uint64_t>(1, T / B);
    std::vector<float> hloss(B);
    double last_loss = 0.0;
    // alpha's start: the scale of the centroid scores whose softmax best fits the first
    // batch's labels (the MLP's output is zero at the start), over powers of two.
    ivf->gamma.alloc(4);
    {
        gather_rows<<<div_up((uint64_t)B * d, 256), 256, 0, st>>>(x, d, order.ptr, B, xr.ptr);
        centroid_scores(st, *ivf, xr.ptr, (int)B, sc.ptr);
        const double rms2 = 1.0 / ((double)ivf->inv_scale * ivf->inv_scale) * d;  // mean squared centred norm
        double best_loss = INFINITY, best_alpha = 1.0 / rms2;
        for (int e = -4; e <= 12; e++) {
            const double alpha = std::ldexp(1.0 / rms2, e);
            const float hg[4] = {(float)std::log(alpha), 0.0f, 0.0f, 0.0f};
            KY_HIP(hipMemcpyAsync(ivf->gamma.ptr, hg, sizeof(hg), hipMemcpyHostToDevice, st));
            fill_words(st, lg.ptr, (size_t)B * p.n_lists, 0u);
            add_centroid_scores<<<div_up((uint64_t)B * p.n_lists, 256), 256, 0, st>>>(lg.ptr, sc.ptr,
                                                                                   (uint64_t)B * p.n_lists,
                                                                                   ivf->gamma.ptr);
            softmax_xent(st, lg.ptr, B, (int)p.n_lists, t_idx.ptr, t_w.ptr, (int)k, 1.0f / (float)B, loss.ptr);
            KY_HIP(hipMemcpyAsync(hloss.data(), loss.ptr, sizeof(float) * B, hipMemcpyDeviceToHost, st));
            KY_HIP(hipStreamSynchronize(st));
            double l = 0.0;
            for (float v : hloss) l += v;
            if (l < best_loss) {
                best_loss = l;
                best_alpha = alpha;
            }
        }
        const float hg[4] = {(float)std::log(best_alpha), 0.0f, 0.0f, 0.0f};
        KY_HIP(hipMemcpyAsync(ivf->gamma.ptr, hg, sizeof(hg), hipMemcpyHostToDevice, st));
    }
    for (uint32_t ep = 0; ep < p.epochs; ep++) {
        double ep_loss = 0.0;
        for (uint64_t sidx = 0; sidx < steps_per_epoch; sidx++) {
            const uint64_t r0 = sidx * B;
            normalize_rows<<<div_up((uint64_t)B * d, 256), 256, 0, st>>>(x, d, order.ptr + r0, B, ivf->mean.ptr,
                                                                          ivf->inv_scale, xb.ptr);
            R.forward(st, xb.ptr, B, h1.ptr, h2.ptr, lg.ptr);
            gather_rows<<<div_up((uint64_t)B * d, 256), 256, 0, st>>>(x, d, order.ptr + r0, B, xr.ptr);
            centroid_scores(st, *ivf, xr.ptr, (int)B, sc.ptr);
            add_centroid_scores<<<div_up((uint64_t)B * p.n_lists, 256), 256, 0, st>>>(lg.ptr, sc.ptr,
                                                                                   (uint64_t)B * p.n_lists,
                                                                                   ivf->gamma.ptr);
            softmax_xent(st, lg.ptr, B, (int)p.n_lists, t_idx.ptr + r0 * k, t_w.ptr + r0 * k, (int)k, 1.0f / (float)B,
                         loss.ptr);
            R.backward(st, xb.ptr, B, h1.ptr, h2.ptr, lg.ptr, dh2.ptr, dh1.ptr);
            // gamma: dLoss/dgamma = alpha sum dL S.
            row_dot<<<div_up(B, 4), 256, 0, st>>>(lg.ptr, sc.ptr, B, p.n_lists, rsum.ptr);
            gamma_grad<<<1, 256, 0, st>>>(rsum.ptr, B, ivf->gamma.ptr);
            KY_HIP(hipGetLastError());
            R.adam_step(st, p.lr);
            adam(st, ivf->gamma.ptr, ivf->gamma.ptr + 1, ivf->gamma.ptr + 2, ivf->gamma.ptr + 3, 1, p.lr, R.step);
            if (ep + 1 == p.epochs) {
                KY_HIP(hipMemcpyAsync(hloss.data(), loss.ptr, sizeof(float) * B, hipMemcpyDeviceToHost, st));
                KY_HIP(hipStreamSynchronize(st));
                for (float v : hloss) ep_loss += v;
            }
        }
        if (ep + 1 == p.epochs) last_loss = ep_loss / ((double)steps_per_epoch * B);
    }
    // Label mass of the top 8 lists on (up to) one batch of training queries.
    float top_mass = 0.0f;
    {
        normalize_rows<<<div_up((uint64_t)B * d, 256), 256, 0, st>>>(x, d, order.ptr, B, ivf->mean.ptr,
                                                                      ivf->inv_scale, xb.ptr);
        R.forward(st, xb.ptr, B, h1.ptr, h2.ptr, lg.ptr);
        gather_rows<<<div_up((uint64_t)B * d, 256), 256, 0, st>>>(x, d, order.ptr, B, xr.ptr);
        centroid_scores(st, *ivf, xr.ptr, (int)B, sc.ptr);
        add_centroid_scores<<<div_up((uint64_t)B * p.n_lists, 256), 256, 0, st>>>(lg.ptr, sc.ptr,
                                                                               (uint64_t)B * p.n_lists,
                                                                               ivf->gamma.ptr);
        const uint32_t P = std::min<uint32_t>(8, p.n_lists);
        DevBuf<uint32_t> ti((size_t)B * 8);
        DevBuf<float> tv((size_t)B * 8), mass(B);
        row_top(st, lg.ptr, B, (int)p.n_lists, P == 8 ? 8 : 1, ti.ptr, tv.ptr, nullptr);
        label_mass<<<div_up(B, 256), 256, 0, st>>>(ti.ptr, B, P == 8 ? 8 : 1, t_idx.ptr, t_w.ptr, k, mass.ptr);
        std::vector<float> hm(B);
        KY_HIP(hipMemcpyAsync(hm.data(), mass.ptr, sizeof(float) * B, hipMemcpyDeviceToHost, st));
        KY_HIP(hipStreamSynchronize(st));
        double s = 0.0;
        for (float v : hm) s += v;
        top_mass = (float)(s / B);
        float hg = 0.0f;
        KY_HIP(hipMemcpyAsync(&hg, ivf->gamma.ptr, sizeof(float), hipMemcpyDeviceToHost, st));
        KY_HIP(hipStreamSynchronize(st));
        ivf->alpha = std::exp(hg);
    }
    const double train_ms = timer.stop();
    if (info) {
        info->kmeans_ms = kmeans_ms;
        info->labels_ms = labels_ms;
        info->train_ms = train_ms;
        info->final_loss = (float)last_loss;
        info->label_recall_top = top_mass;
    }
    return ivf.release();
}

void ivf_layout(const ky_ivf &ivf, float *centroids, uint32_t *assign, uint32_t *list_ptr, uint32_t *list_rows,
                float *alpha) {
    Context &ctx = *ivf.ctx;
    KY_HIP(hipSetDevice(ctx.device));
    std::vector<CopyPart> parts;
    if (centroids) parts.push_back({centroids, ivf.centroids.ptr, sizeof(float) * ivf.n_lists * ivf.d});
    if (assign) parts.push_back({assign, ivf.assign.ptr, sizeof(uint32_t) * ivf.n});
    if (list_ptr) parts.push_back({list_ptr, ivf.list_ptr.ptr, sizeof(uint32_t) * (ivf.n_lists + 1)});
    if (list_rows) parts.push_back({list_rows, ivf.list_rows.ptr, sizeof(uint32_t) * ivf.n});
    if (!parts.empty()) download_parts(ctx, parts);
    if (alpha) *alpha = ivf.alpha;
}

void ivf_free(ky_ivf *ivf) {
    if (!ivf) return;
    (void)hipSetDevice(ivf->ctx->device);
    delete ivf;
}

void ivf_search(Context &ctx, const ky_ivf &ivf, const Dense &points, uint32_t n_neighbors, uint32_t n_probe,
                int use_router, float confidence, uint32_t *h_idx, float *h_dist, ky_ivf_search_info *info) {
    const double t_start = now_ms();
    KY_CHECK(points.rows == ivf.n && points.cols == ivf.d, KY_E215,
             "ivf: search the embedding the index was built on (same rows and components)");
    KY_CHECK(n_neighbors >= 2 && (uint64_t)n_neighbors <= ivf.n && n_neighbors <= 65, KY_E215,
             "ivf: n_neighbors out of range");
    KY_CHECK(n_probe == 8 || n_probe == 1, KY_E501, "ivf: n_probe must be 1 or 8");
    KY_CHECK(n_probe <= ivf.n_lists, KY_E215, "ivf: n_probe above n_lists");
    hipStream_t st = ctx.stream;
    Timer timer(st);
    const uint64_t n = ivf.n;
    const uint32_t d = ivf.d, k = n_neighbors - 1, P = n_probe;
    const float *x = points.data.ptr;
    // Route every row.
    timer.start();
    DevBuf<uint32_t> probes(n * P);
    DevBuf<float> conf(n);
    {
        const uint64_t chunk = std::max<uint64_t>(1, std::min<uint64_t>(n, (256ull << 20) / ivf.n_lists));
        DevBuf<float> xb(chunk * d), S(chunk * ivf.n_lists), tv(chunk * P), lse(chunk);
        DevBuf<float> h1, h2, C;
        if (use_router) {
            h1.alloc(chunk * ivf.router.hidden);
            h2.alloc(chunk * ivf.router.hidden);
            C.alloc(chunk * ivf.n_lists);
        }
        for (uint64_t r0 = 0; r0 < n; r0 += chunk) {
            const uint64_t rows = std::min<uint64_t>(chunk, n - r0);
            if (use_router) {
                // alpha times the centroid scores plus the MLP's correction.
                normalize_rows<<<div_up(rows * d, 256), 256, 0, st>>>(x + r0 * d, d, nullptr, rows, ivf.mean.ptr,
                                                                       ivf.inv_scale, xb.ptr);
                const_cast<Router &>(ivf.router).forward(st, xb.ptr, rows, h1.ptr, h2.ptr, S.ptr);
                centroid_scores(st, ivf, x + r0 * d, (int)rows, C.ptr);
                add_centroid_scores<<<div_up(rows * ivf.n_lists, 256), 256, 0, st>>>(S.ptr, C.ptr,
                                                                                     rows * ivf.n_lists,
                                                                                     ivf.gamma.ptr);
            } else {
                centroid_scores(st, ivf, x + r0 * d, (int)rows, S.ptr);
            }
            row_top(st, S.ptr, rows, (int)ivf.n_lists, (int)P, probes.ptr + r0 * P, tv.ptr, lse.ptr);
            probe_confidence<<<div_up(rows, 256), 256, 0, st>>>(tv.ptr, lse.ptr, rows, P, conf.ptr + r0);
            KY_HIP(hipGetLastError());
        }
    }
    const double route_ms = timer.stop();
    timer.start();
    DevBuf<double> keys(n * k);
    DevBuf<uint32_t> rows(n * k), count(n);
    {
        // The (query, probe) pairs sorted by list, a block's worth of one list a unit.
        const uint64_t np = n * P;
        DevBuf<uint32_t> codes(np), s_lists(np), s_codes(np), start(ivf.n_lists), end(ivf.n_lists);
        DevBuf<unsigned char> sort_tmp;
        iota_words(st, codes.ptr, np, 0u);
        sort_pairs(st, probes.ptr, s_lists.ptr, codes.ptr, s_codes.ptr, np, bits_for(ivf.n_lists), sort_tmp);
        fill_words(st, start.ptr, ivf.n_lists, 0u);
        fill_words(st, end.ptr, ivf.n_lists, 0u);
        segment_bounds<<<div_up(np, 256), 256, 0, st>>>(s_lists.ptr, np, start.ptr, end.ptr);
        KY_HIP(hipGetLastError());
        std::vector<uint32_t> hs(ivf.n_lists), he(ivf.n_lists);
        KY_HIP(hipMemcpyAsync(hs.data(), start.ptr, sizeof(uint32_t) * ivf.n_lists, hipMemcpyDeviceToHost, st));
        KY_HIP(hipMemcpyAsync(he.data(), end.ptr, sizeof(uint32_t) * ivf.n_lists, hipMemcpyDeviceToHost, st));
        KY_HIP(hipStreamSynchronize(st));
        std::vector<uint4> hunits;
        for (uint32_t l = 0; l < ivf.n_lists; l++)
            for (uint32_t f = hs[l]; f < he[l]; f += PAIRS_A_BLOCK)
                hunits.push_back(make_uint4(l, f, std::min<uint32_t>(PAIRS_A_BLOCK, he[l] - f), 0));
        DevBuf<uint4> units(std::max<size_t>(hunits.size(), 1));
        KY_HIP(hipMemcpyAsync(units.ptr, hunits.data(), sizeof(uint4) * hunits.size(), hipMemcpyHostToDevice, st));
        DevBuf<double> pk(np * k);
        DevBuf<uint32_t> pr(np * k), pc(np);
        if (!hunits.empty()) {
            const unsigned dc = (d + 7) / 8;
            const auto scan = dc == 1 ? ivf_pair_scan<1> : dc == 2 ? ivf_pair_scan<2> : dc == 3 ? ivf_pair_scan<3>
                            : dc == 4 ? ivf_pair_scan<4> : dc == 5 ? ivf_pair_scan<5> : dc == 6 ? ivf_pair_scan<6>
                            : dc == 7 ? ivf_pair_scan<7> : ivf_pair_scan<8>;
            scan<<<(unsigned)hunits.size(), 256, 0, st>>>(x, d, ivf.list_ptr.ptr, ivf.list_rows.ptr, ivf.list_pts.ptr,
                                                          s_codes.ptr, units.ptr, P, k, pk.ptr, pr.ptr, pc.ptr);
        }
        KY_HIP(hipGetLastError());
        merge_probes<<<div_up(n, 256), 256, 0, st>>>(pk.ptr, pr.ptr, pc.ptr, n, P, k, keys.ptr, rows.ptr, count.ptr);
        KY_HIP(hipGetLastError());
        KY_HIP(hipStreamSynchronize(st));
    }
    std::vector<float> hconf(n);
    std::vector<uint32_t> hcount(n);
    KY_HIP(hipMemcpyAsync(hconf.data(), conf.ptr, sizeof(float) * n, hipMemcpyDeviceToHost, st));
    KY_HIP(hipMemcpyAsync(hcount.data(), count.ptr, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, st));
    KY_HIP(hipStreamSynchronize(st));
    const double scan_ms = timer.stop();
    timer.start();
    std::vector<uint32_t> list;
    for (uint64_t q = 0; q < n; q++)
        if (hcount[q] < k || (use_router && hconf[q] < confidence)) list.push_back((uint32_t)q);
    if (!list.empty()) {
        DevBuf<uint32_t> d_list(list.size());
        KY_HIP(hipMemcpyAsync(d_list.ptr, list.data(), sizeof(uint32_t) * list.size(), hipMemcpyHostToDevice, st));
        // The plain method: BF16 candidates, exact re-rank and certificate.
        exact_search(ctx, x, n, d, k, KY_PRECISION_BF16, d_list.ptr, list.size(), keys.ptr, rows.ptr, nullptr);
    }
    const double fallback_ms = timer.stop();
    timer.start();
    DevBuf<uint32_t> g_idx(n * k);
    DevBuf<float> g_dist(n * k);
    knn_graph(ctx, keys.ptr, rows.ptr, n, k, KY_METRIC_EUCLIDEAN, g_idx.ptr, g_dist.ptr);
    download_parts(ctx, {{h_idx, g_idx.ptr, sizeof(uint32_t) * n * k}, {h_dist, g_dist.ptr, sizeof(float) * n * k}});
    const double graph_ms = timer.stop();
    if (info) {
        info->fallback = list.size();
        info->route_ms = route_ms;
        info->search_ms = scan_ms;
        info->fallback_ms = fallback_ms;
        info->graph_ms = graph_ms;
        info->total_ms = now_ms() - t_start;
    }
}

}  // namespace ky
