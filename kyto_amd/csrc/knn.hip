// neighbors() on the device: exact k nearest neighbours of every row of an embedding,
// the graph kyto's exact search gives (crates/kyto-exec/src/neighbors.rs:226-312),
// computed as candidates + exact re-rank + certification + exact fallback:
//
// 1. prepare (neighbors_exact.rs:76-180, `Prepared`): rows centred on the f64 mean
//    and rounded to f32, offsets fl32(|a|^2), each row's slack E_q.
// 2. scan_points<false>: each row's m candidates, the m smallest approximate keys
//    (f32: fl(fl(o_q + o_p) - 2 q.p); or BF16 MFMA, knn_bf16.hip) by (key, row):
//    one block per 64 queries, each wavefront owning 16 queries; lane = point of a
//    64-point tile, so a query's threshold is uniform across the wave. A query keeps
//    every point below its threshold (the m-th smallest (key, row) so far) in an LDS
//    buffer that a wave-level bitonic sort compacts back to m.
// 3. rerank: each candidate's exact key (kyto's f64 8-lane key, device.hpp), the k
//    smallest by (key, row) (`Nearest`, neighbors.rs:463-474); the row is certified
//    when every non-candidate's exact key is proven above the k-th: the m-th
//    candidate's approximate key less the key's error bound exceeds it
//    (neighbors_gpu.rs:50-54,149-158).
// 4. scan_points<true>: uncertified rows searched again with exact keys against
//    every row (`search_row`, neighbors.rs:521-536).
// 5. graph: each row's k sorted by index, distances f32(sqrt(key)) (neighbors.rs:412-435).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "knn.hpp"
#include "pca.hpp"

namespace ky {

constexpr int QB = 64;  // queries per block
constexpr int TP = 64;  // points per tile (one per lane)

// ---------------------------------------------------------------- preparation ----

// Column sums of rows [c*CH, (c+1)*CH) in row order (f64): one block per chunk.
__global__ void embed_col_sums(const float *__restrict__ x, uint64_t n, uint32_t d, uint32_t ld, uint32_t chunk,
                               double *__restrict__ partial) {
    const unsigned c = blockIdx.x, j = threadIdx.x;
    if (j >= d) return;
    const uint64_t r0 = (uint64_t)c * chunk, r1 = min(r0 + chunk, n);
    double s = 0.0;
    for (uint64_t r = r0; r < r1; r++) s += (double)x[r * ld + j];
    partial[(size_t)c * d + j] = s;
}

__global__ void embed_reduce(const double *__restrict__ partial, uint32_t d, uint32_t n_chunks, double inv_n,
                             double *__restrict__ mean) {
    const unsigned j = threadIdx.x;
    if (j >= d) return;
    double s = 0.0;
    for (uint32_t c = 0; c < n_chunks; c++) s += partial[(size_t)c * d + j];
    mean[j] = s * inv_n;
}

// Per row: a = f32(x - mean) (zero padded to dp <= 64), sq = sum f64(a)^2 (lane j
// holds component j; a fixed butterfly adds the lanes), offset f32(sq), the finite
// flag, and the largest sq (as bits: non-negative doubles order as their bits). One
// wavefront a row, EMBED_ROWS rows a block, one atomic a block.
constexpr unsigned EMBED_ROWS = 64;

__global__ __launch_bounds__(256) void embed_centre(const float *__restrict__ x, uint64_t n, uint32_t d, uint32_t ld,
                                                    const double *__restrict__ mean, uint32_t dp,
                                                    float *__restrict__ cent, float *__restrict__ offs,
                                                    double *__restrict__ sqn, unsigned long long *__restrict__ largest,
                                                    unsigned *__restrict__ bad) {
    __shared__ double wave_max[4];
    __shared__ unsigned wave_bad[4];
    const unsigned lane = lane_id(), wave = threadIdx.x / 64;
    const double mu = lane < d ? mean[lane] : 0.0;
    double mx = 0.0;
    bool ok = true;
    for (unsigned i = wave; i < EMBED_ROWS; i += 4) {
        const uint64_t r = (uint64_t)blockIdx.x * EMBED_ROWS + i;
        if (r >= n) break;
        float a = 0.0f;
        if (lane < d) {
            const float v = x[r * ld + lane];
            ok = ok && isfinite(v);
            a = (float)((double)v - mu);
        }
        if (lane < dp) cent[r * dp + lane] = a;
        double sq = (double)a * (double)a;
#pragma unroll
        for (int o = 32; o >= 1; o /= 2) sq += __shfl_xor(sq, o);
        if (lane == 0) {
            offs[r] = (float)sq;
            sqn[r] = sq;
        }
        mx = fmax(mx, sq);
    }
    const bool wave_ok = __ballot(!ok) == 0;
    if (lane == 0) {
        wave_max[wave] = mx;
        wave_bad[wave] = wave_ok ? 0u : 1u;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = fmax(fmax(wave_max[0], wave_max[1]), fmax(wave_max[2], wave_max[3]));
        atomicMax(largest, (unsigned long long)__double_as_longlong(m));
        if (wave_bad[0] | wave_bad[1] | wave_bad[2] | wave_bad[3]) atomicOr(bad, 1u);
    }
}

// slack[r] = factor (sqrt(sq_r) + largest)^2 + underflow (neighbors_exact.rs:76-89).
__global__ void embed_slack(const double *__restrict__ sqn, uint64_t n, double largest, double factor,
                            double underflow, double *__restrict__ slack) {
    uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (r >= n) return;
    double t = sqrt(sqn[r]) + largest;
    slack[r] = factor * t * t + underflow;
}

// Copy the first d columns of x (ld columns a row) into a packed n x d matrix.
__global__ void pack_columns(const float *__restrict__ x, uint64_t n, uint32_t d, uint32_t ld, float *__restrict__ out) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n * d) return;
    uint64_t r = i / d;
    uint32_t j = (uint32_t)(i % d);
    out[i] = x[r * ld + j];
}

// ------------------------------------------------------------- the point scan ----

// EXACT = false: approximate f32 keys on centred rows `a` (dp floats a row) with
// offsets; keys kept as order_f32 bits. EXACT = true: kyto's exact f64 keys on the
// original rows `a` (d floats a row); keys kept as f64 bits (non-negative).
// Queries: q_list[i] when q_list is not null, else q_first + i, for i < q_count.
// Output per query i: `keep` (key, row) pairs ascending and the count.
//
// EXACT with seed_keys / seed_rows (n x keep, at the query's row, ascending): the
// query starts at the threshold of its seed's keep-th pair (inclusive), so a search
// that already holds keep exact candidates keeps only the rows that can replace them.
//
// With n_seg > 1 the points are cut into n_seg segments of whole tiles: block b scans
// segment b % n_seg for the queries of group b / n_seg, and writes its `keep` pairs at
// slot i * n_seg + segment (merge_segments then keeps each query's smallest), so a
// short list of queries still fills the device.
template <bool EXACT, int DP, int CAP>
__global__ __launch_bounds__(256) void scan_points(const float *__restrict__ a, const float *__restrict__ offs,
                                                   uint64_t n, uint32_t d, const uint32_t *__restrict__ q_list,
                                                   uint64_t q_first, uint64_t q_count, uint32_t keep, uint32_t n_seg,
                                                   const double *__restrict__ seed_keys,
                                                   const uint32_t *__restrict__ seed_rows,
                                                   uint32_t *__restrict__ out_rows, void *__restrict__ out_keys,
                                                   uint32_t *__restrict__ out_count) {
    using K = typename std::conditional<EXACT, unsigned long long, uint32_t>::type;
    constexpr int QS = DP + 4;
    __shared__ __attribute__((aligned(16))) float sq[EXACT ? 1 : QB * QS];
    // EXACT: the queries' rows widened to f64 (DP a row, zero past d).
    __shared__ __attribute__((aligned(16))) double sqd[EXACT ? QB * DP : 1];
    __shared__ float sqo[QB];
    __shared__ K bkey[QB * CAP];
    __shared__ uint32_t brow[QB * CAP];
    __shared__ uint32_t bcnt[QB], trow[QB], qrow[QB];
    __shared__ K tkey[QB];
    const unsigned tid = threadIdx.x, lane = lane_id(), wave = tid / 64;
    const uint64_t i0 = (uint64_t)(blockIdx.x / n_seg) * QB;
    const unsigned seg = blockIdx.x % n_seg;
    const uint64_t seg_len = ((n + n_seg - 1) / n_seg + TP - 1) / TP * TP;
    const uint64_t p_begin = min<uint64_t>((uint64_t)seg * seg_len, n), p_end = min<uint64_t>(p_begin + seg_len, n);
    const unsigned stride = EXACT ? d : DP;
    if (tid < QB) {
        uint64_t i = i0 + tid;
        uint32_t qg = 0xFFFFFFFFu;
        if (i < q_count) qg = q_list ? q_list[i] : (uint32_t)(q_first + i);
        qrow[tid] = qg;
        bcnt[tid] = 0;
        tkey[tid] = ~K(0);
        trow[tid] = ~0u;
        if constexpr (EXACT) {
            // A seed (the keep-th (key, row) of a known set of rows) bounds the
            // query's keep-th pair from above: only pairs at or below it are kept.
            if (seed_keys && qg != 0xFFFFFFFFu) {
                const double sk = seed_keys[(uint64_t)qg * keep + keep - 1];
                const uint32_t sr = seed_rows[(uint64_t)qg * keep + keep - 1];
                if (isfinite(sk) && sr != 0xFFFFFFFFu) {
                    tkey[tid] = (K)__double_as_longlong(sk);
                    trow[tid] = sr + 1;
                }
            }
        }
        sqo[tid] = (!EXACT && qg != 0xFFFFFFFFu) ? offs[qg] : 0.0f;
    }
    __syncthreads();
    for (unsigned i = tid; i < QB * DP; i += 256) {
        unsigned r = i / DP, c = i % DP;
        uint32_t qg = qrow[r];
        float v = 0.0f;
        if (qg != 0xFFFFFFFFu && c < stride) v = a[(uint64_t)qg * stride + c];
        if constexpr (EXACT) sqd[r * DP + c] = (double)v;
        else sq[r * QS + c] = v;
    }
    __syncthreads();
    unsigned nq = 0;  // this wave's live queries
    for (unsigned qi = 0; qi < 16; qi++)
        if (qrow[wave * 16 + qi] != 0xFFFFFFFFu) nq = qi + 1;
    for (uint64_t t0 = p_begin; t0 < p_end; t0 += TP) {
        const uint64_t p = t0 + lane;
        const bool valid = p < p_end;
        float pr[EXACT ? 1 : DP];
        double pd[EXACT ? DP : 1];  // EXACT: the point widened to f64, zero past d
        if constexpr (EXACT) {
            // Unconditional loads (a lane past the segment reads its first point,
            // components past d read component d - 1, then zeroed), so no load waits
            // on a branch.
            const uint64_t pp = valid ? p : p_begin;
#pragma unroll
            for (int c = 0; c < DP; c++) {
                const float v = a[pp * d + min((unsigned)c, d - 1)];
                pd[c] = c < (int)d ? (double)v : 0.0;
            }
        } else if (valid) {
            const float4 *src = reinterpret_cast<const float4 *>(a + p * DP);
#pragma unroll
            for (int c = 0; c < DP / 4; c++) {
                float4 v = src[c];
                pr[4 * c] = v.x;
                pr[4 * c + 1] = v.y;
                pr[4 * c + 2] = v.z;
                pr[4 * c + 3] = v.w;
            }
        } else {
#pragma unroll
            for (int c = 0; c < DP; c++) pr[c] = 0.0f;
        }
        const float po = (!EXACT && valid) ? offs[p] : 0.0f;
        for (unsigned qi = 0; qi < nq; qi++) {
            const unsigned ql = wave * 16 + qi;
            const uint32_t qg = qrow[ql];
            K *bk = bkey + ql * CAP;
            uint32_t *br = brow + ql * CAP;
            if (bcnt[ql] > CAP - TP) compact<K, CAP>(bk, br, &bcnt[ql], &tkey[ql], &trow[ql], keep);
            K key;
            if constexpr (!EXACT) {
                float dot = 0.0f;
                const float4 *qv = reinterpret_cast<const float4 *>(sq + ql * QS);
#pragma unroll
                for (int c = 0; c < DP / 4; c++) {
                    float4 v = qv[c];
                    dot = __builtin_fmaf(v.x, pr[4 * c], dot);
                    dot = __builtin_fmaf(v.y, pr[4 * c + 1], dot);
                    dot = __builtin_fmaf(v.z, pr[4 * c + 2], dot);
                    dot = __builtin_fmaf(v.w, pr[4 * c + 3], dot);
                }
                float kf = (sqo[ql] + po) - 2.0f * dot;
                key = order_f32(kf);
            } else {
                double l[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                // The query's row: every LDS read first, then the arithmetic.
                double qv[DP];
                const double2 *qd = reinterpret_cast<const double2 *>(sqd + ql * DP);
#pragma unroll
                for (int c = 0; c < DP / 2; c++) {
                    const double2 v = qd[c];
                    qv[2 * c] = v.x;
                    qv[2 * c + 1] = v.y;
                }
#pragma unroll
                for (int c = 0; c < DP; c++) {
                    // Past d both sides are 0, and adding t * t = +0.0 leaves the sum as it is.
                    const double t = qv[c] - pd[c];
                    l[c & 7] += t * t;
                }
                double h0 = l[0] + l[4], h1 = l[1] + l[5], h2 = l[2] + l[6], h3 = l[3] + l[7];
                double kd = (h0 + h2) + (h1 + h3);
                key = (K)__double_as_longlong(kd);
            }
            const K tk = tkey[ql];
            const uint32_t tr = trow[ql];
            const bool hit = valid && p != qg && (key < tk || (key == tk && (uint32_t)p < tr));
            const uint64_t mask = ballot(hit);
            if (mask) {
                const unsigned base = bcnt[ql];
                if (hit) {
                    unsigned pos = base + mbcnt(mask);
                    bk[pos] = key;
                    br[pos] = (uint32_t)p;
                }
                wave_sync();
                if (lane == 0) bcnt[ql] = base + (unsigned)__popcll(mask);
                wave_sync();
            }
        }
    }
    for (unsigned qi = 0; qi < nq; qi++) {
        const unsigned ql = wave * 16 + qi;
        K *bk = bkey + ql * CAP;
        uint32_t *br = brow + ql * CAP;
        compact<K, CAP>(bk, br, &bcnt[ql], &tkey[ql], &trow[ql], keep);
        const unsigned cnt = bcnt[ql];
        const uint64_t i = (i0 + ql) * n_seg + seg;
        // Output slot: the query's position in the list (or range), times n_seg, plus
        // the segment.
        for (unsigned e = lane; e < keep; e += 64) {
            uint32_t row = e < cnt ? br[e] : 0xFFFFFFFFu;
            out_rows[i * keep + e] = row;
            if constexpr (!EXACT)
                static_cast<float *>(out_keys)[i * keep + e] = e < cnt ? unorder_f32(bk[e]) : INFINITY;
            else
                static_cast<double *>(out_keys)[i * keep + e] = e < cnt ? __longlong_as_double((long long)bk[e]) : INFINITY;
        }
        if (lane == 0) out_count[i] = cnt;
    }
}

// Per query (one wavefront): the smallest `keep` (key, row) pairs of its n_seg
// segments' lists (f64 keys, ascending), in order. N (a power of two, at least
// n_seg * keep) entries of LDS a wave.
template <int N>
__global__ __launch_bounds__(64) void merge_segments(const uint32_t *__restrict__ seg_rows,
                                                     const double *__restrict__ seg_keys,
                                                     const uint32_t *__restrict__ seg_count, uint64_t q_count,
                                                     uint32_t n_seg, uint32_t keep, uint32_t *__restrict__ out_rows,
                                                     double *__restrict__ out_keys, uint32_t *__restrict__ out_count) {
    __shared__ unsigned long long key[N];
    __shared__ uint32_t row[N];
    const uint64_t q = blockIdx.x;
    if (q >= q_count) return;
    const unsigned lane = lane_id();
    for (unsigned e = lane; e < N; e += 64) {
        const unsigned s = e / keep, j = e % keep;
        bool ok = s < n_seg && j < seg_count[q * n_seg + s];
        key[e] = ok ? (unsigned long long)__double_as_longlong(seg_keys[(q * n_seg + s) * keep + j]) : ~0ull;
        row[e] = ok ? seg_rows[(q * n_seg + s) * keep + j] : ~0u;
    }
    wave_sync();
    wave_bitonic<unsigned long long>(key, row, N);
    unsigned total = 0;
    for (unsigned s = 0; s < n_seg; s++) total += seg_count[q * n_seg + s];
    const unsigned cnt = min(total, keep);
    for (unsigned e = lane; e < keep; e += 64) {
        out_rows[q * keep + e] = e < cnt ? row[e] : 0xFFFFFFFFu;
        out_keys[q * keep + e] = e < cnt ? __longlong_as_double((long long)key[e]) : INFINITY;
    }
    if (lane == 0) out_count[q] = cnt;
}

// ------------------------------------------------------------------- re-rank ----

// One wavefront per query: exact keys of its candidates, the k smallest by (key, row),
// and the certificate. Eight lanes share a candidate: lane t of the group sums kyto's
// lane t (components t, t + 8, ...) in order, and the fixed tree combines the eight
// through shuffles, so the key has the bits of Kernels::squared_distances while the
// group reads the candidate's row in contiguous pieces. floor_mode 0: f32 keys of
// the centred expanded form, floor = key - slack[q] (neighbors_exact.rs:76-89);
// 1: BF16 keys (knn_bf16.hip). m <= 64.
__global__ __launch_bounds__(256) void rerank(const float *__restrict__ pts, uint64_t n, uint32_t d,
                                              const uint32_t *__restrict__ cand_rows,
                                              const float *__restrict__ cand_keys,
                                              const uint32_t *__restrict__ cand_count,
                                              const float *__restrict__ cand_thr,
                                              const uint32_t *__restrict__ q_rows, uint64_t q_first,
                                              uint64_t q_count, uint32_t m, uint32_t k,
                                              const double *__restrict__ slack, const double *__restrict__ bf_err,
                                              double bf_err_max, double bf_acc_rel, int floor_mode,
                                              double *__restrict__ out_keys, uint32_t *__restrict__ out_rows,
                                              uint8_t *__restrict__ certified) {
    const unsigned wave = threadIdx.x / 64, lane = lane_id();
    const uint64_t i = (uint64_t)blockIdx.x * 4 + wave;
    if (i >= q_count) return;
    const uint64_t q = q_rows ? (uint64_t)q_rows[i] : q_first + i;
    const unsigned cnt = cand_count[i];
    const unsigned t = lane & 7, g = lane >> 3;  // kyto lane t, candidate group g
    const float *qrow = pts + q * d;
    // Lane t sums components t, t + 8, ... of each of its group's 8 candidates
    // (8r + g, r < 8); each 64-component chunk issues all its loads first. Loads are
    // unconditional (a missing candidate reads the query's row, components past d
    // read component d - 1) so that none waits on a branch; those values are unused.
    uint32_t rowv[8], lrow[8];
    double part[8];
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const unsigned c = 8 * r + g;
        const uint32_t cr = cand_rows[i * m + (c < cnt ? c : 0)];
        rowv[r] = c < cnt ? cr : 0xFFFFFFFFu;
        lrow[r] = c < cnt ? cr : (uint32_t)q;
        part[r] = 0.0;
    }
    for (unsigned j0 = 0; j0 < d; j0 += 64) {
        float qv[8], pv[8][8];
#pragma unroll
        for (int u = 0; u < 8; u++) qv[u] = qrow[min(j0 + t + 8 * u, d - 1)];
#pragma unroll
        for (int r = 0; r < 8; r++)
#pragma unroll
            for (int u = 0; u < 8; u++) pv[r][u] = pts[(uint64_t)lrow[r] * d + min(j0 + t + 8 * u, d - 1)];
#pragma unroll
        for (int r = 0; r < 8; r++)
#pragma unroll
            for (int u = 0; u < 8; u++) {
                // Past d the difference is 0 and adds +0.0 (no change to the sum).
                const float p = j0 + t + 8 * u < d ? pv[r][u] : qv[u];
                const double diff = (double)qv[u] - (double)p;
                part[r] += diff * diff;
            }
    }
    // Lane 8g + t keeps the key of candidate 8t + g; every lane of a group ends the
    // reduction with the same sum (the additions commute), so it can pick its own.
    unsigned long long mine = ~0ull;
    uint32_t mine_row = 0xFFFFFFFFu;
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const unsigned c = 8 * r + g;
        // ((l0 + l4) + (l2 + l6)) + ((l1 + l5) + (l3 + l7)) over the group's lanes.
        const double h = part[r] + __shfl_xor(part[r], 4, 8);
        const double hq = h + __shfl_xor(h, 2, 8);
        const double key = hq + __shfl_xor(hq, 1, 8);
        if (t == (unsigned)r) {
            mine = c < cnt ? (unsigned long long)__double_as_longlong(key) : ~0ull;
            mine_row = rowv[r];
        }
    }
    // Lane c takes candidate c (= 8r + g) from lane 8g + r.
    const int src = 8 * (lane & 7) + (lane >> 3);
    unsigned long long skey = __shfl(mine, src);
    uint32_t srow = __shfl(mine_row, src);
    wave_sort64(skey, srow);
    bool cert = false;
    if (cnt >= k) {
        const double kth = __longlong_as_double((long long)__shfl(skey, (int)k - 1));
        if ((uint64_t)cnt >= n - 1) {
            cert = true;  // every other row is a candidate
        } else {
            // Every row whose key is at most the scan's threshold is a candidate: with
            // fewer than m candidates that threshold bounds the rest (cand_thr), else
            // the m-th candidate's key does.
            const double far = cand_thr && cnt < m ? (double)cand_thr[i] : (double)cand_keys[i * m + cnt - 1];
            double floor;
            if (floor_mode == 0) {
                floor = far - slack[q];
            } else {
                // BF16 (knn_bf16.hip): keys >= far give |b_q - b_p| >= sqrt(far - delta_q),
                // so the exact distance is at least that less err_q + max err.
                const double s = far - bf_err[n + q];
                const double r = s > 0.0 ? sqrt(s) - bf_err[q] - bf_err_max : -1.0;
                floor = r > 0.0 ? r * r * (1.0 - 1e-12) : -1.0;
            }
            cert = isfinite(far) && floor > kth;
        }
    }
    if (lane < k) {
        out_keys[q * k + lane] = __longlong_as_double((long long)skey);
        out_rows[q * k + lane] = srow;
    }
    if (lane == 0) certified[q] = cert ? 1 : 0;
}

// Exact search results for listed rows: copy the scan's output (per list position)
// to the rows' places.
__global__ void scatter_exact(const uint32_t *__restrict__ list, uint64_t count, uint32_t k,
                              const uint32_t *__restrict__ rows_in, const double *__restrict__ keys_in,
                              double *__restrict__ out_keys, uint32_t *__restrict__ out_rows) {
    uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= count * k) return;
    uint64_t i = t / k;
    uint32_t e = (uint32_t)(t % k);
    uint64_t q = list[i];
    out_rows[q * k + e] = rows_in[i * k + e];
    out_keys[q * k + e] = keys_in[i * k + e];
}

// ------------------------------------------------------------------- graph ----

// Per row: its k entries sorted by row index; distance f32(sqrt(key)) (Euclidean).
__global__ void graph_rows(const double *__restrict__ keys, const uint32_t *__restrict__ rows, uint64_t n, uint32_t k,
                           uint32_t *__restrict__ g_idx, float *__restrict__ g_dist) {
    __shared__ uint32_t sr[64][65];
    __shared__ double sk[64][65];
    const unsigned t = threadIdx.x;
    const uint64_t q = (uint64_t)blockIdx.x * 64 + t;
    if (q >= n) return;
    for (uint32_t e = 0; e < k; e++) {
        uint32_t r = rows[q * k + e];
        double key = keys[q * k + e];
        uint32_t at = e;
        while (at > 0 && sr[t][at - 1] > r) {
            sr[t][at] = sr[t][at - 1];
            sk[t][at] = sk[t][at - 1];
            at--;
        }
        sr[t][at] = r;
        sk[t][at] = key;
    }
    for (uint32_t e = 0; e < k; e++) {
        g_idx[q * k + e] = sr[t][e];
        g_dist[q * k + e] = (float)sqrt(sk[t][e]);
    }
}

// ================================================================== host side ====

template <bool EXACT, int CAP>
static void launch_scan(int dp, hipStream_t st, unsigned blocks, const float *a, const float *offs, uint64_t n,
                        uint32_t d, const uint32_t *list, uint64_t q_first, uint64_t q_count, uint32_t keep,
                        uint32_t *out_rows, void *out_keys, uint32_t *out_count, uint32_t n_seg = 1,
                        const double *seed_keys = nullptr, const uint32_t *seed_rows = nullptr) {
#define KY_SCAN(DPV)                                                                                   \
    scan_points<EXACT, DPV, CAP><<<blocks * n_seg, 256, 0, st>>>(a, offs, n, d, list, q_first, q_count, keep, \
                                                                 n_seg, seed_keys, seed_rows, out_rows, out_keys, \
                                                                 out_count)
    switch (dp) {
        case 8: KY_SCAN(8); break;
        case 16: KY_SCAN(16); break;
        case 32: KY_SCAN(32); break;
        case 52: KY_SCAN(52); break;
        default: KY_SCAN(64); break;
    }
#undef KY_SCAN
    KY_HIP(hipGetLastError());
}

static int padded_width(uint32_t d) {
    if (d <= 8) return 8;
    if (d <= 16) return 16;
    if (d <= 32) return 32;
    if (d <= 52) return 52;
    return 64;
}

// f32's unit roundoff and the Prepared slack's constants (neighbors_exact.rs:40-43).
static const double UNIT_ROUNDOFF = 1.0 / 16777216.0;
static const double LARGEST_NORM = 1e15;

uint32_t candidates_for(uint32_t k, uint64_t n, int precision) {
    // neighbors_gpu.rs:42-44 (2k + 16, at most every other row); BF16 keeps at least 64.
    uint64_t m = 2ull * k + 16;
    if (precision == KY_PRECISION_BF16) m = std::max<uint64_t>(m, 64);
    m = std::min<uint64_t>(m, 64);
    return (uint32_t)std::min<uint64_t>(m, n - 1);
}

void knn_prepare(Context &ctx, const float *x, uint64_t n, uint32_t d, KnnPrep &prep) {
    hipStream_t st = ctx.stream;
    prep.n = n;
    prep.d = d;
    prep.dp = padded_width(d);
    const uint32_t chunk = 256;
    const uint32_t n_chunks = div_up(n, chunk);
    DevBuf<double> partial((size_t)n_chunks * d * 2 + 2 * d);
    prep.mean.alloc(d);
    embed_col_sums<<<n_chunks, 64, 0, st>>>(x, n, d, d, chunk, partial.ptr);
    double *total = partial.ptr + (size_t)n_chunks * d * 2 + d;
    ordered_sum(st, partial.ptr, n_chunks, d, partial.ptr + (size_t)n_chunks * d, total);
    embed_reduce<<<1, 64, 0, st>>>(total, d, 1, 1.0 / (double)n, prep.mean.ptr);
    prep.cent.alloc(n * prep.dp);
    prep.offs.alloc(n);
    prep.sqn.alloc(n);
    prep.slack.alloc(n);
    DevBuf<unsigned long long> largest(1);
    DevBuf<unsigned> bad(1);
    KY_HIP(hipMemsetAsync(largest.ptr, 0, sizeof(unsigned long long), st));
    KY_HIP(hipMemsetAsync(bad.ptr, 0, sizeof(unsigned), st));
    embed_centre<<<div_up(n, EMBED_ROWS), 256, 0, st>>>(x, n, d, d, prep.mean.ptr, prep.dp, prep.cent.ptr, prep.offs.ptr,
                                                  prep.sqn.ptr, largest.ptr, bad.ptr);
    KY_HIP(hipGetLastError());
    unsigned long long lb = 0;
    unsigned hbad = 0;
    KY_HIP(hipMemcpyAsync(&lb, largest.ptr, sizeof(lb), hipMemcpyDeviceToHost, st));
    KY_HIP(hipMemcpyAsync(&hbad, bad.ptr, sizeof(hbad), hipMemcpyDeviceToHost, st));
    KY_HIP(hipStreamSynchronize(st));
    prep.finite = hbad == 0;
    double lsq;
    std::memcpy(&lsq, &lb, sizeof(lsq));
    prep.largest = std::sqrt(lsq);
    prep.filter_ok = std::isfinite(prep.largest) && prep.largest <= LARGEST_NORM;
    const double factor = 2.0 * ((double)d + 8.0) * UNIT_ROUNDOFF;
    const double underflow = 4.0 * ((double)d + 8.0) * (double)FLT_MIN;
    embed_slack<<<div_up(n, 256), 256, 0, st>>>(prep.sqn.ptr, n, prep.largest, factor, underflow, prep.slack.ptr);
    KY_HIP(hipGetLastError());
}

// Exact search of the listed rows against every row; results into keys/rows (n x k).
void knn_exact_rows(Context &ctx, const float *pts, uint64_t n, uint32_t d, uint32_t k, const uint32_t *d_list,
                    uint64_t count, double *d_keys, uint32_t *d_rows, bool seeded) {
    if (count == 0) return;
    hipStream_t st = ctx.stream;
    DevBuf<uint32_t> rows_tmp(count * k), cnt_tmp(count);
    DevBuf<double> keys_tmp(count * k);
    const unsigned blocks = div_up(count, QB);
    // Segments of the points: enough blocks to fill the device (about 2048), each
    // segment at least 16 tiles, n_seg * k within the merge's 4096 entries.
    uint32_t n_seg = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(div_up(2048, blocks), n / (16 * TP)));
    n_seg = std::min<uint32_t>(n_seg, 4096 / k);
    const double *sk = seeded ? d_keys : nullptr;
    const uint32_t *sr = seeded ? d_rows : nullptr;
    if (n_seg <= 1) {
        launch_scan<true, 128>(padded_width(d), st, blocks, pts, nullptr, n, d, d_list, 0, count, k, rows_tmp.ptr,
                               keys_tmp.ptr, cnt_tmp.ptr, 1, sk, sr);
    } else {
        DevBuf<uint32_t> seg_rows(count * n_seg * k), seg_cnt(count * n_seg);
        DevBuf<double> seg_keys(count * n_seg * k);
        launch_scan<true, 128>(padded_width(d), st, blocks, pts, nullptr, n, d, d_list, 0, count, k, seg_rows.ptr,
                               seg_keys.ptr, seg_cnt.ptr, n_seg, sk, sr);
        if (n_seg * k <= 1024)
            merge_segments<1024><<<(unsigned)count, 64, 0, st>>>(seg_rows.ptr, seg_keys.ptr, seg_cnt.ptr, count,
                                                                  n_seg, k, rows_tmp.ptr, keys_tmp.ptr, cnt_tmp.ptr);
        else
            merge_segments<4096><<<(unsigned)count, 64, 0, st>>>(seg_rows.ptr, seg_keys.ptr, seg_cnt.ptr, count,
                                                                  n_seg, k, rows_tmp.ptr, keys_tmp.ptr, cnt_tmp.ptr);
        KY_HIP(hipGetLastError());
    }
    scatter_exact<<<div_up(count * k, 256), 256, 0, st>>>(d_list, count, k, rows_tmp.ptr, keys_tmp.ptr, d_keys, d_rows);
    KY_HIP(hipGetLastError());
}

void knn_candidates(Context &ctx, const KnnPrep &prep, uint64_t q_first, uint64_t q_count, uint32_t m,
                    uint32_t *d_rows, float *d_keys, uint32_t *d_count) {
    const unsigned blocks = div_up(q_count, QB);
    KY_CHECK(m <= 64, KY_E215, "neighbors: more than 64 candidates per row");
    launch_scan<false, 128>(prep.dp, ctx.stream, blocks, prep.cent.ptr, prep.offs.ptr, prep.n, prep.d, nullptr,
                            q_first, q_count, m, d_rows, d_keys, d_count);
}

void knn_rerank(Context &ctx, const float *pts, const KnnPrep &prep, const uint32_t *q_rows, uint64_t q_first,
                uint64_t q_count, uint32_t m, uint32_t k, const uint32_t *cand_rows, const float *cand_keys,
                const uint32_t *cand_count, const float *cand_thr, int floor_mode, const double *bf_err,
                double bf_err_max, double bf_acc_rel, double *d_keys, uint32_t *d_rows, uint8_t *d_cert) {
    rerank<<<div_up(q_count, 4), 256, 0, ctx.stream>>>(pts, prep.n, prep.d, cand_rows, cand_keys, cand_count, cand_thr, q_rows,
                                                       q_first, q_count, m, k, prep.slack.ptr, bf_err, bf_err_max,
                                                       bf_acc_rel, floor_mode, d_keys, d_rows, d_cert);
    KY_HIP(hipGetLastError());
}

void knn_graph(Context &ctx, const double *d_keys, const uint32_t *d_rows, uint64_t n, uint32_t k, uint32_t *g_idx,
               float *g_dist) {
    KY_CHECK(k <= 64, KY_E215, "neighbors: n_neighbors above 65");
    graph_rows<<<div_up(n, 64), 64, 0, ctx.stream>>>(d_keys, d_rows, n, k, g_idx, g_dist);
    KY_HIP(hipGetLastError());
}

void knn_pack(Context &ctx, const float *x, uint64_t n, uint32_t d, uint32_t ld, float *out) {
    pack_columns<<<div_up(n * d, 256), 256, 0, ctx.stream>>>(x, n, d, ld, out);
    KY_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ driver ----

struct EventTimer {
    hipStream_t st;
    hipEvent_t a, b;
    explicit EventTimer(hipStream_t s) : st(s) {
        KY_HIP(hipEventCreate(&a));
        KY_HIP(hipEventCreate(&b));
    }
    ~EventTimer() {
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

void bf16_candidates(Context &ctx, const float *pts, const KnnPrep &prep, const uint32_t *q_rows, uint64_t nq,
                     uint32_t m, uint32_t k, uint32_t *d_rows, float *d_keys, uint32_t *d_count, float *d_thr,
                     DevBuf<double> &err, double &err_max, double &acc_rel, double &scan_ms);

// The checks of neighbors::search (neighbors.rs:226-312) that precede the search.
uint32_t check_search(const Dense &points, uint32_t n_components, uint32_t n_neighbors, int metric) {
    KY_CHECK(points.cols >= 1 && points.rows >= 1, KY_E215, "neighbors: an embedding without rows or components");
    uint32_t used = n_components == 0 ? points.cols : n_components;
    KY_CHECK(used >= 1 && used <= points.cols, KY_E215,
             "neighbors: n_components " + std::to_string(n_components) + " is not an int from 1 to " +
                 std::to_string(points.cols));
    KY_CHECK(n_neighbors >= 2 && (uint64_t)n_neighbors <= points.rows, KY_E215,
             "neighbors: n_neighbors " + std::to_string(n_neighbors) + " counts the cell itself, and " +
                 std::to_string(points.rows) + " cells reach neighbors (an int from 2 to " +
                 std::to_string(std::max<uint64_t>(points.rows, 2)) + ")");
    KY_CHECK(points.rows - 1 <= 0xFFFFFFFEull, KY_E215, "neighbors: graph indices are uint32 (at most 2^32 cells)");
    KY_CHECK(metric == KY_METRIC_EUCLIDEAN || metric == KY_METRIC_COSINE, KY_E215, "neighbors: unknown metric");
    KY_CHECK(metric == KY_METRIC_EUCLIDEAN, KY_E501, "neighbors: no device kernel for metric='cosine'");
    KY_CHECK(used <= 64, KY_E501, "neighbors: no device kernel for more than 64 components");
    KY_CHECK(n_neighbors <= 65, KY_E501, "neighbors: no device kernel for n_neighbors above 65");
    return used;
}

// Exact neighbours (kyto's (key, row) order) of the rows q_rows[0..nq) (every row
// when q_rows is null, nq = n) of the packed n x d embedding: candidates at
// `precision`, exact re-rank and certificate, exact search again of the rows the
// certificate does not cover. keys / rows are n x k, filled at the queried rows.
void exact_search(Context &ctx, const float *pts, uint64_t n, uint32_t d, uint32_t k, int precision,
                  const uint32_t *q_rows, uint64_t nq, double *keys, uint32_t *rows, ky_knn_info *info) {
    hipStream_t st = ctx.stream;
    EventTimer timer(st);
    KnnPrep prep;
    timer.start();
    knn_prepare(ctx, pts, n, d, prep);
    KY_CHECK(prep.finite, KY_E215, "neighbors: the embedding holds a value that is not finite");
    DevBuf<uint8_t> cert(n);
    uint32_t m = 0;
    double search_ms = 0, rerank_ms = 0, fallback_ms = 0, scan_ms = 0;
    std::vector<uint32_t> list;
    std::vector<uint32_t> hq;
    if (q_rows) {
        hq.resize(nq);
        KY_HIP(hipMemcpyAsync(hq.data(), q_rows, sizeof(uint32_t) * nq, hipMemcpyDeviceToHost, st));
    }
    // The BF16 kernel's augmented rows hold d + 6 <= 64 components; wider embeddings
    // take their candidates from the FP32 kernel (the graph is the same).
    if (precision == KY_PRECISION_BF16 && d > 58) precision = KY_PRECISION_FP32;
    if (prep.filter_ok) {
        m = candidates_for(k, n, precision);
        DevBuf<uint32_t> c_rows(nq * m), c_count(nq);
        DevBuf<float> c_keys(nq * m), c_thr;
        DevBuf<double> err;
        double err_max = 0.0, acc_rel = 0.0;
        if (precision == KY_PRECISION_FP32) {
            KY_CHECK(q_rows == nullptr, KY_E501, "neighbors: FP32 candidates for a row subset");
            knn_candidates(ctx, prep, 0, n, m, c_rows.ptr, c_keys.ptr, c_count.ptr);
        } else {
            c_thr.alloc(nq);
            bf16_candidates(ctx, pts, prep, q_rows, nq, m, k, c_rows.ptr, c_keys.ptr, c_count.ptr, c_thr.ptr, err,
                            err_max, acc_rel, scan_ms);
        }
        search_ms = timer.stop();
        timer.start();
        knn_rerank(ctx, pts, prep, q_rows, 0, nq, m, k, c_rows.ptr, c_keys.ptr, c_count.ptr, c_thr.ptr,
                   precision == KY_PRECISION_FP32 ? 0 : 1, err.ptr, err_max, acc_rel, keys, rows, cert.ptr);
        std::vector<uint8_t> hcert(n);
        KY_HIP(hipMemcpyAsync(hcert.data(), cert.ptr, n, hipMemcpyDeviceToHost, st));
        rerank_ms = timer.stop();
        for (uint64_t i = 0; i < nq; i++) {
            uint32_t q = q_rows ? hq[i] : (uint32_t)i;
            if (!hcert[q]) list.push_back(q);
        }
    } else {
        search_ms = timer.stop();
        for (uint64_t i = 0; i < nq; i++) list.push_back(q_rows ? hq[i] : (uint32_t)i);
    }
    timer.start();
    if (!list.empty()) {
        DevBuf<uint32_t> d_list(list.size());
        KY_HIP(hipMemcpyAsync(d_list.ptr, list.data(), sizeof(uint32_t) * list.size(), hipMemcpyHostToDevice, st));
        // Rows that went through the re-rank hold their candidates' k smallest exact
        // pairs: an upper bound of their k-th pair that seeds the exact search.
        knn_exact_rows(ctx, pts, n, d, k, d_list.ptr, list.size(), keys, rows, prep.filter_ok);
        KY_HIP(hipStreamSynchronize(st));
    }
    fallback_ms = timer.stop();
    if (info) {
        info->candidates = m;
        info->fallback = list.size();
        info->certified = nq - list.size();
        info->search_ms = search_ms;
        info->rerank_ms = rerank_ms;
        info->fallback_ms = fallback_ms;
        info->scan_kernel_ms = scan_ms;
    }
}

void neighbors_run(Context &ctx, const Dense &points, uint32_t n_components, uint32_t n_neighbors, int metric,
                   int precision, uint32_t *h_idx, float *h_dist, ky_knn_info *info) {
    const double t_start = now_ms();
    const uint32_t d = check_search(points, n_components, n_neighbors, metric);
    KY_CHECK(precision == KY_PRECISION_FP32 || precision == KY_PRECISION_BF16, KY_E215, "neighbors: unknown precision");
    hipStream_t st = ctx.stream;
    const uint64_t n = points.rows;
    const uint32_t k = n_neighbors - 1;
    DevBuf<float> packed;
    const float *pts = points.data.ptr;
    if (d != points.cols) {
        packed.alloc(n * d);
        knn_pack(ctx, points.data.ptr, n, d, points.cols, packed.ptr);
        pts = packed.ptr;
    }
    DevBuf<double> keys(n * k);
    DevBuf<uint32_t> rows(n * k);
    exact_search(ctx, pts, n, d, k, precision, nullptr, n, keys.ptr, rows.ptr, info);
    DevBuf<uint32_t> g_idx(n * k);
    DevBuf<float> g_dist(n * k);
    knn_graph(ctx, keys.ptr, rows.ptr, n, k, g_idx.ptr, g_dist.ptr);
    download_parts(ctx, {{h_idx, g_idx.ptr, sizeof(uint32_t) * n * k}, {h_dist, g_dist.ptr, sizeof(float) * n * k}});
    if (info) info->total_ms = now_ms() - t_start;
}

void nearest_keys_run(Context &ctx, const Dense &points, uint32_t n_components, uint64_t first, uint64_t count,
                      uint32_t m, int precision, float *h_keys, uint32_t *h_rows) {
    const uint32_t d = check_search(points, n_components, 2, KY_METRIC_EUCLIDEAN);
    KY_CHECK(first + count <= points.rows, KY_E215, "nearest_keys: rows out of range");
    KY_CHECK(m >= 1 && m <= 64 && (uint64_t)m <= points.rows - 1, KY_E215, "nearest_keys: m out of range");
    KY_CHECK(precision == KY_PRECISION_FP32, KY_E501, "nearest_keys: only FP32 keys are exposed");
    hipStream_t st = ctx.stream;
    const uint64_t n = points.rows;
    DevBuf<float> packed;
    const float *pts = points.data.ptr;
    if (d != points.cols) {
        packed.alloc(n * d);
        knn_pack(ctx, points.data.ptr, n, d, points.cols, packed.ptr);
        pts = packed.ptr;
    }
    KnnPrep prep;
    knn_prepare(ctx, pts, n, d, prep);
    KY_CHECK(prep.finite, KY_E215, "nearest_keys: the embedding holds a value that is not finite");
    KY_CHECK(prep.filter_ok, KY_E501, "nearest_keys: row norms too large for f32 keys");
    DevBuf<uint32_t> c_rows(count * m), c_count(count);
    DevBuf<float> c_keys(count * m);
    knn_candidates(ctx, prep, first, count, m, c_rows.ptr, c_keys.ptr, c_count.ptr);
    KY_HIP(hipMemcpyAsync(h_keys, c_keys.ptr, sizeof(float) * count * m, hipMemcpyDeviceToHost, st));
    KY_HIP(hipMemcpyAsync(h_rows, c_rows.ptr, sizeof(uint32_t) * count * m, hipMemcpyDeviceToHost, st));
    KY_HIP(hipStreamSynchronize(st));
}

}  // namespace ky
