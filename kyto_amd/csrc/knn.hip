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
#include <mutex>
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

// Cosine rows (neighbors_exact.rs:76-89, `Prepared` for cosine; neighbors.rs:336-350
// for the norms): each row's norm |x| = sqrt of kyto's 8-lane f64 dot product (lane t
// sums components t, t + 8, ... in order, then the fixed tree), its unit row
// u = f32(x / |x|) (zero for a zero row), sq = |u|^2 in f64, offset f32(sq), the finite
// flag and the largest sq, as embed_centre. One wavefront a row.
__global__ __launch_bounds__(256) void embed_unit(const float *__restrict__ x, uint64_t n, uint32_t d, uint32_t dp,
                                                  float *__restrict__ cent, float *__restrict__ offs,
                                                  double *__restrict__ sqn, double *__restrict__ norms,
                                                  unsigned long long *__restrict__ largest,
                                                  unsigned *__restrict__ bad) {
    __shared__ double wave_max[4];
    __shared__ unsigned wave_bad[4];
    const unsigned lane = lane_id(), wave = threadIdx.x / 64;
    double mx = 0.0;
    bool ok = true;
    for (unsigned i = wave; i < EMBED_ROWS; i += 4) {
        const uint64_t r = (uint64_t)blockIdx.x * EMBED_ROWS + i;
        if (r >= n) break;
        const float v = lane < d ? x[r * d + lane] : 0.0f;
        ok = ok && isfinite(v);
        const double prod = (double)v * (double)v;
        // kyto's lane t (t < 8) on lane t: components t, t + 8, ... added in order.
        double acc = 0.0;
#pragma unroll
        for (int j0 = 0; j0 < 64; j0 += 8) {
            const double term = __shfl(prod, j0 + (int)(lane & 7));
            if ((unsigned)j0 + (lane & 7) < d) acc += term;
        }
        const double h = acc + __shfl(acc, (int)(lane & 3) + 4);   // lanes 0..3: l_t + l_{t+4}
        const double hq = h + __shfl(h, (int)(lane & 1) + 2);      // lanes 0, 1: h_t + h_{t+2}
        const double total = __shfl(hq, 0) + __shfl(hq, 1);
        const double norm = sqrt(total);
        const float u = (lane < d && norm != 0.0) ? (float)((double)v / norm) : 0.0f;
        if (lane < dp) cent[r * dp + lane] = u;
        double sq = (double)u * (double)u;
#pragma unroll
        for (int o = 32; o >= 1; o /= 2) sq += __shfl_xor(sq, o);
        if (lane == 0) {
            offs[r] = (float)sq;
            sqn[r] = sq;
            norms[r] = norm;
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
// The approximate keys of nearest_keys (crates/kyto-gpu/src/kernels_matrix.rs:140-257):
// MODE 0 the expanded key fl(fl(o_q + o_p) - 2 q.p) of the search; MODE 1 the squared
// differences summed in component order in f32; MODE 2 the cosine key
// 1 - dot / (|q| |p|) (dot in component order, the caller's f32 norms) clipped to
// [0, 2], 1 when a norm is 0. Queries are rows of `qa` (null: rows of the points); the
// point skip[i] is left out for query i (null: the query's own row).
struct Pairwise {
    const float *qa = nullptr;
    const uint32_t *skip = nullptr;
    const float *qn = nullptr, *pn = nullptr;
};

template <bool EXACT, bool COS, int DP, int CAP, int MODE = 0>
__global__ __launch_bounds__(256) void scan_points(const float *__restrict__ a, const float *__restrict__ offs,
                                                   const double *__restrict__ norms, const Pairwise pw,
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
    __shared__ uint32_t bcnt[QB], trow[QB], qrow[QB], qskip[QB];
    __shared__ float qnorm[QB];
    __shared__ K tkey[QB];
    // EXACT: a tile's points, staged by the block's coalesced loads (global_load_lds),
    // each lane then reading its point's row from here.
    __shared__ __attribute__((aligned(16))) float spts[EXACT ? TP * DP : 1];
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
        qskip[tid] = (pw.skip && i < q_count) ? pw.skip[i] : qg;
        qnorm[tid] = (MODE == 2 && qg != 0xFFFFFFFFu) ? pw.qn[qg] : 0.0f;
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
        sqo[tid] = (!EXACT && MODE == 0 && qg != 0xFFFFFFFFu) ? offs[qg] : 0.0f;
    }
    __syncthreads();
    const float *qsrc = (MODE != 0 && pw.qa) ? pw.qa : a;
    for (unsigned i = tid; i < QB * DP; i += 256) {
        unsigned r = i / DP, c = i % DP;
        uint32_t qg = qrow[r];
        float v = 0.0f;
        if (qg != 0xFFFFFFFFu && c < stride) v = qsrc[(uint64_t)qg * stride + c];
        if constexpr (EXACT) sqd[r * DP + c] = (double)v;
        else sq[r * QS + c] = v;
    }
    __syncthreads();
    // The block's queries are dealt to its waves in turn (query ql to wave ql % 4), so a
    // short list of queries keeps every wave busy.
    unsigned nq = 0;  // this wave's live queries
    for (unsigned qi = 0; qi < 16; qi++)
        if (qrow[qi * 4 + wave] != 0xFFFFFFFFu) nq = qi + 1;
    // The points [t0, min(t0 + TP, p_end)) into spts, 64 floats a wave instruction (a
    // lane past them reads the tile's first float; those slots belong to no point).
    const auto stage_tile = [&](uint64_t t0) {
        const uint32_t count = (uint32_t)(min<uint64_t>(TP, p_end - t0) * d);
        const float *src = a + t0 * d;
        for (uint32_t i = wave; i * 64 < count; i += 4) {
            const uint32_t f = i * 64 + lane;
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(src + (f < count ? f : 0u)),
                                             (__attribute__((address_space(3))) void *)(spts + i * 64), 4, 0, 0);
        }
    };
    if constexpr (EXACT)
        if (p_begin < p_end) stage_tile(p_begin);
    for (uint64_t t0 = p_begin; t0 < p_end; t0 += TP) {
        const uint64_t p = t0 + lane;
        const bool valid = p < p_end;
        float pr[EXACT ? 1 : DP];
        double pd[EXACT ? DP : 1];  // EXACT: the point widened to f64, zero past d
        if constexpr (EXACT) {
            // The tile's loads landed (every wave's), the rows read into registers
            // (components past d read component d - 1, then zeroed; a lane past the
            // segment reads slots of no point, its keys never kept), then the next
            // tile's loads are issued into the same slots while this one is searched.
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
            __syncthreads();
            const float *row = spts + lane * d;
#pragma unroll
            for (int c = 0; c < DP; c++) {
                const float v = row[min((unsigned)c, d - 1)];
                pd[c] = c < (int)d ? (double)v : 0.0;
            }
            __syncthreads();
            if (t0 + TP < p_end) stage_tile(t0 + TP);
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
        const float po = (!EXACT && MODE == 0 && valid) ? offs[p] : 0.0f;
        const float pnv = (MODE == 2 && valid) ? pw.pn[p] : 0.0f;
        for (unsigned qi = 0; qi < nq; qi++) {
            const unsigned ql = qi * 4 + wave;
            const uint32_t qg = qrow[ql];
            K *bk = bkey + ql * CAP;
            uint32_t *br = brow + ql * CAP;
            if (bcnt[ql] > CAP - TP) compact<K, CAP>(bk, br, &bcnt[ql], &tkey[ql], &trow[ql], keep);
            K key;
            if constexpr (!EXACT && MODE == 0) {
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
            } else if constexpr (!EXACT) {
                // Component order in f32, each product rounded before its add (this
                // file is compiled without contraction); past the width both rows are 0.
                const float *qv = sq + ql * QS;
                float acc = 0.0f;
#pragma unroll
                for (int c = 0; c < DP; c++) {
                    if constexpr (MODE == 1) {
                        const float t = qv[c] - pr[c];
                        acc += t * t;
                    } else {
                        acc += qv[c] * pr[c];
                    }
                }
                if constexpr (MODE == 2) {
                    const float qn = qnorm[ql];
                    float kc = 1.0f;
                    if (qn != 0.0f && pnv != 0.0f) {
                        kc = 1.0f - acc / (qn * pnv);
                        kc = kc < 0.0f ? 0.0f : (kc > 2.0f ? 2.0f : kc);
                    }
                    acc = kc;
                }
                key = order_f32(acc);
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
                    // Past d both sides are 0, and adding +0.0 leaves the sum as it is.
                    if constexpr (COS) {
                        l[c & 7] += qv[c] * pd[c];
                    } else {
                        const double t = qv[c] - pd[c];
                        l[c & 7] += t * t;
                    }
                }
                double h0 = l[0] + l[4], h1 = l[1] + l[5], h2 = l[2] + l[6], h3 = l[3] + l[7];
                double kd = (h0 + h2) + (h1 + h3);
                if constexpr (COS) kd = cosine_key(kd, norms[qg], norms[valid ? p : p_begin]);
                key = (K)__double_as_longlong(kd);
            }
            const K tk = tkey[ql];
            const uint32_t tr = trow[ql];
            const bool hit = valid && p != qskip[ql] && (key < tk || (key == tk && (uint32_t)p < tr));
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
        const unsigned ql = qi * 4 + wave;
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
template <bool COS>
__global__ __launch_bounds__(256) void rerank(const float *__restrict__ pts, uint64_t n, uint32_t d,
                                              const uint32_t *__restrict__ cand_rows,
                                              const float *__restrict__ cand_keys,
                                              const uint32_t *__restrict__ cand_count,
                                              const float *__restrict__ cand_thr,
                                              const uint32_t *__restrict__ q_rows, uint64_t q_first,
                                              uint64_t q_count, uint32_t m, uint32_t k,
                                              const double *__restrict__ slack, const double *__restrict__ bf_err,
                                              double bf_err_max, double bf_acc_rel, int floor_mode,
                                              const double *__restrict__ norms,
                                              double *__restrict__ out_keys, uint32_t *__restrict__ out_rows,
                                              uint32_t *__restrict__ unc_list, uint32_t *__restrict__ unc_count,
                                              uint32_t *__restrict__ short_list, uint32_t *__restrict__ short_count) {
    const unsigned wave = threadIdx.x / 64, lane = lane_id();
    const uint64_t i = (uint64_t)blockIdx.x * 4 + wave;
    if (i >= q_count) return;
    const uint64_t q = q_rows ? (uint64_t)q_rows[i] : q_first + i;
    const unsigned cnt = cand_count[i];
    const unsigned t = lane & 7, g = lane >> 3;  // kyto lane t, candidate group g
    const float *qrow = pts + q * d;
    // BF16 candidates (Euclidean; ascending keys): candidate c can be among the k
    // nearest only if the least exact key its BF16 key allows, (sqrt(key_c - delta_q)
    // - e)^2 with e = err_q + max err, is at most the most the k-th candidate's allows,
    // (sqrt(key_k + delta_q) + e)^2 (the first k candidates' exact keys are all at or
    // below it). The others' rows are not read: their exact keys would exceed the k
    // smallest. `need`: one past the last candidate that can be among them.
    unsigned need = cnt;
    if (!COS && floor_mode == 1 && cnt > k) {
        const double e = bf_err[q] + bf_err_max, dq = bf_err[n + q];
        const double hi = sqrt((double)cand_keys[i * m + k - 1] + dq) + e;
        const double upper = hi * hi * (1.0 + 1e-12);
        bool may = false;
        if (lane < cnt) {
            const double s = (double)cand_keys[i * m + lane] - dq;
            const double r = s > 0.0 ? sqrt(s) - e : -1.0;
            may = !(r > 0.0 && r * r * (1.0 - 1e-12) > upper);
        }
        const uint64_t mask = ballot(may);
        need = mask ? 64u - (unsigned)__clzll(mask) : cnt;
    }
    // Lane t sums components t, t + 8, ... of each of its group's 8 candidates
    // (8r + g, r < 8); each 64-component chunk issues all its loads first. Loads are
    // unconditional (a missing or dropped candidate reads the query's row, components
    // past d read component d - 1) so that none waits on a branch; those values are
    // unused.
    uint32_t rowv[8], lrow[8];
    double part[8];
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const unsigned c = 8 * r + g;
        const uint32_t cr = cand_rows[i * m + (c < need ? c : 0)];
        rowv[r] = c < need ? cr : 0xFFFFFFFFu;
        lrow[r] = c < need ? cr : (uint32_t)q;
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
                if constexpr (COS) {
                    // Cosine: kyto's dot product lanes (past d the product is +0.0).
                    const float p = j0 + t + 8 * u < d ? pv[r][u] : 0.0f;
                    part[r] += (double)qv[u] * (double)p;
                } else {
                    const float p = j0 + t + 8 * u < d ? pv[r][u] : qv[u];
                    const double diff = (double)qv[u] - (double)p;
                    part[r] += diff * diff;
                }
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
        double key = hq + __shfl_xor(hq, 1, 8);
        if constexpr (COS) key = cosine_key(key, norms[q], norms[lrow[r]]);
        if (t == (unsigned)r) {
            mine = c < need ? (unsigned long long)__double_as_longlong(key) : ~0ull;
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
            // Cosine: the keys are half the squared distance between unit rows (u =
            // x / |x|, a zero row 0), so the cosine floor is half the Euclidean one, less
            // the f64 rounding of kyto's 1 - dot / (|q| |p|); a zero row's key, 1, lies
            // above half its distance to any row (0.5 or 0), so the floor still holds.
            if constexpr (COS) floor = 0.5 * floor - 1e-15;
            cert = isfinite(far) && floor > kth;
        }
    }
    if (lane < k) {
        out_keys[q * k + lane] = __longlong_as_double((long long)skey);
        out_rows[q * k + lane] = srow;
    }
    // An uncertified row with fewer than m candidates goes to the short list when one is
    // given (it is scanned again from a higher starting threshold), any other to the exact
    // search.
    if (lane == 0 && !cert) {
        if (short_list && cnt < m) short_list[atomicAdd(short_count, 1u)] = (uint32_t)q;
        else unc_list[atomicAdd(unc_count, 1u)] = (uint32_t)q;
    }
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
                           int cosine, uint32_t *__restrict__ g_idx, float *__restrict__ g_dist) {
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
        g_dist[q * k + e] = cosine ? (float)sk[t][e] : (float)sqrt(sk[t][e]);
    }
}

// ================================================================== host side ====

template <bool EXACT, bool COS, int CAP, int MODE = 0>
static void launch_scan(int dp, hipStream_t st, unsigned blocks, const float *a, const float *offs,
                        const double *norms, uint64_t n, uint32_t d, const uint32_t *list, uint64_t q_first,
                        uint64_t q_count, uint32_t keep, uint32_t *out_rows, void *out_keys, uint32_t *out_count,
                        uint32_t n_seg = 1, const double *seed_keys = nullptr, const uint32_t *seed_rows = nullptr,
                        const Pairwise &pw = Pairwise()) {
#define KY_SCAN(DPV)                                                                                         \
    scan_points<EXACT, COS, DPV, CAP, MODE><<<blocks * n_seg, 256, 0, st>>>(a, offs, norms, pw, n, d, list, q_first, \
                                                                            q_count, keep, n_seg, seed_keys,       \
                                                                            seed_rows, out_rows, out_keys, out_count)
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

int padded_width(uint32_t d) {
    if (d <= 8) return 8;
    if (d <= 16) return 16;
    if (d <= 32) return 32;
    if (d <= 52) return 52;
    return 64;
}

// f32's unit roundoff and the Prepared slack's constants (neighbors_exact.rs:40-43).
static const double UNIT_ROUNDOFF = 1.0 / 16777216.0;
static const double LARGEST_NORM = 1e15;
// Chunks of rows for the f64 column sums of the mean.
static const uint32_t MEAN_CHUNK = 256;

uint32_t candidates_for(uint32_t k, uint64_t n, int precision) {
    // neighbors_gpu.rs:42-44 (2k + 16, at most every other row); BF16 keeps at least 64.
    uint64_t m = 2ull * k + 16;
    if (precision == KY_PRECISION_BF16) m = std::max<uint64_t>(m, 64);
    m = std::min<uint64_t>(m, 64);
    return (uint32_t)std::min<uint64_t>(m, n - 1);
}

unsigned bf16_resident(const Context &ctx);
static const unsigned MI355X_BF16_RESIDENT = 3 * 256;  // 3 search blocks on each of 256 CUs

KnnShape knn_shape(const Context *ctx, uint64_t n, uint32_t d, uint32_t k, int metric, int precision, bool list,
                   uint64_t nq, uint64_t round) {
    KnnShape s;
    s.n = n;
    s.d = d;
    s.dp = (uint32_t)padded_width(d);
    s.k = k;
    s.metric = metric;
    // The BF16 kernel's augmented rows hold d + 6 <= 64 components; wider embeddings
    // take their candidates from the FP32 kernel (the graph is the same).
    s.precision = precision == KY_PRECISION_BF16 && d > 58 ? KY_PRECISION_FP32 : precision;
    s.m = candidates_for(k, n, s.precision);
    s.nq = nq;
    s.list = list;
    const uint64_t granule = s.precision == KY_PRECISION_BF16 ? 256 : QB;
    s.round = round == 0 || round >= nq ? nq : std::max<uint64_t>(granule, round / granule * granule);
    s.resident = ctx ? bf16_resident(*ctx) : MI355X_BF16_RESIDENT;
    return s;
}

template <class A>
void knn_take_prep(A &a, const KnnShape &s, KnnPrep &p) {
    p.mean = a.template take<double>(s.d);
    p.cent = a.template take<float>(s.n * s.dp);
    p.offs = a.template take<float>(s.n);
    p.sqn = a.template take<double>(s.n);
    p.slack = a.template take<double>(s.n);
    p.norms = s.metric == KY_METRIC_COSINE ? a.template take<double>(s.n) : nullptr;
    p.flags = a.template take<unsigned long long>(2);
}

static size_t prep_partial_count(const KnnShape &s) {
    return (size_t)div_up(s.n, MEAN_CHUNK) * s.d * 2 + 2 * s.d;
}

void knn_prepare(Context &ctx, const KnnShape &s, const float *x, KnnPrep &prep, double *partial) {
    hipStream_t st = ctx.stream;
    const uint64_t n = s.n;
    const uint32_t d = s.d;
    fill_words(st, prep.flags, 4, 0u);
    unsigned long long *largest = prep.flags;
    unsigned *bad = reinterpret_cast<unsigned *>(prep.flags + 1);
    if (s.metric == KY_METRIC_COSINE) {
        fill_words(st, prep.mean, 2 * (size_t)d, 0u);
        embed_unit<<<div_up(n, EMBED_ROWS), 256, 0, st>>>(x, n, d, s.dp, prep.cent, prep.offs, prep.sqn, prep.norms,
                                                         largest, bad);
    } else {
        const uint32_t n_chunks = div_up(n, MEAN_CHUNK);
        embed_col_sums<<<n_chunks, 64, 0, st>>>(x, n, d, d, MEAN_CHUNK, partial);
        double *total = partial + (size_t)n_chunks * d * 2 + d;
        ordered_sum(st, partial, n_chunks, d, partial + (size_t)n_chunks * d, total);
        embed_reduce<<<1, 64, 0, st>>>(total, d, 1, 1.0 / (double)n, prep.mean);
        embed_centre<<<div_up(n, EMBED_ROWS), 256, 0, st>>>(x, n, d, d, prep.mean, s.dp, prep.cent, prep.offs,
                                                           prep.sqn, largest, bad);
    }
    KY_HIP(hipGetLastError());
    unsigned long long flags[2] = {0, 0};
    KY_HIP(hipMemcpyAsync(flags, prep.flags, sizeof(flags), hipMemcpyDeviceToHost, st));
    KY_HIP(hipStreamSynchronize(st));
    prep.finite = (flags[1] & 0xFFFFFFFFull) == 0;
    double lsq;
    std::memcpy(&lsq, &flags[0], sizeof(lsq));
    prep.largest = std::sqrt(lsq);
    prep.filter_ok = std::isfinite(prep.largest) && prep.largest <= LARGEST_NORM;
    const double factor = 2.0 * ((double)d + 8.0) * UNIT_ROUNDOFF;
    const double underflow = 4.0 * ((double)d + 8.0) * (double)FLT_MIN;
    embed_slack<<<div_up(n, 256), 256, 0, st>>>(prep.sqn, n, prep.largest, factor, underflow, prep.slack);
    KY_HIP(hipGetLastError());
}

// The exact search of `count` listed rows: point segments that fill the device (about
// 2048 blocks), each at least 16 tiles, n_seg * k within the merge's 4096 entries.
static uint32_t exact_segments(uint64_t n, uint32_t k, uint64_t count) {
    const unsigned blocks = div_up(count, QB);
    uint32_t n_seg = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(div_up(2048, blocks), n / (16 * TP)));
    return std::min<uint32_t>(n_seg, 4096 / k);
}

struct ExactBufs {
    uint32_t *rows = nullptr, *cnt = nullptr, *seg_rows = nullptr, *seg_cnt = nullptr;
    double *keys = nullptr, *seg_keys = nullptr;
};

template <class A>
static void take_exact(A &a, const KnnShape &s, uint64_t count, uint32_t n_seg, ExactBufs &e) {
    e.rows = a.template take<uint32_t>(count * s.k);
    e.cnt = a.template take<uint32_t>(count);
    e.keys = a.template take<double>(count * s.k);
    if (n_seg > 1) {
        e.seg_rows = a.template take<uint32_t>(count * n_seg * s.k);
        e.seg_cnt = a.template take<uint32_t>(count * n_seg);
        e.seg_keys = a.template take<double>(count * n_seg * s.k);
    }
}

// The most bytes an exact-search chunk of at most `most` rows takes (its segments
// change with the count at multiples of QB).
static size_t exact_chunk_bytes(const KnnShape &s, uint64_t most) {
    size_t worst = 0;
    ExactBufs e;
    auto at = [&](uint64_t count) {
        SizeArena a;
        take_exact(a, s, count, exact_segments(s.n, s.k, count), e);
        worst = std::max(worst, a.high);
    };
    for (uint64_t c = QB; c < most; c += QB) at(c);
    if (most) at(most);
    return worst;
}

template <class A>
static void take_round(A &a, const KnnShape &s, uint64_t count, KnnRound &r) {
    r.rows = a.template take<uint32_t>(count * s.m);
    r.keys = a.template take<float>(count * s.m);
    r.count = a.template take<uint32_t>(count);
    r.thr = a.template take<float>(count);
    if (s.precision == KY_PRECISION_BF16) bf16_take_round(a, s, count, r);
}

size_t knn_scratch_bytes(const KnnShape &s, bool graph, bool packed) {
    SizeArena a;
    if (graph) {
        if (packed) a.take<float>(s.n * s.d);
        a.take<double>(s.n * s.k);
        a.take<uint32_t>(s.n * s.k);
    }
    const size_t m0 = a.mark();
    // exact_search's layout, phase by phase.
    KnnPrep p;
    knn_take_prep(a, s, p);
    a.take<uint32_t>(s.nq);
    a.take<uint32_t>(2);
    if (s.precision == KY_PRECISION_BF16) a.take<uint32_t>(s.nq);
    size_t m1 = a.mark();
    if (s.metric != KY_METRIC_COSINE) a.take<double>(prep_partial_count(s));
    a.release(m1);
    if (s.precision == KY_PRECISION_BF16) {
        Bf16Points b;
        bf16_take_points(a, s, b);
        m1 = a.mark();
        a.take<unsigned char>(bf16_prepare_bytes(s));
        a.release(m1);
    }
    KnnRound r;
    take_round(a, s, s.round, r);
    a.release(m1);
    if (s.nq % s.round) {
        take_round(a, s, s.nq % s.round, r);
        a.release(m1);
    }
    if (s.precision == KY_PRECISION_BF16) {
        a.take<unsigned char>(bf16_rescan_bytes(s, std::min(s.round, s.nq)));
        a.release(m1);
    }
    a.take<unsigned char>(exact_chunk_bytes(s, std::min(s.round, s.nq)));
    a.release(m0);
    if (graph) {
        a.take<uint32_t>(s.n * s.k);
        a.take<float>(s.n * s.k);
    }
    return a.high;
}

// Exact search of the listed rows against every row; results into keys/rows (n x k).
static void knn_exact_rows(Context &ctx, Arena &a, const KnnShape &s, const float *pts, const double *norms,
                           const uint32_t *d_list, uint64_t count, double *d_keys, uint32_t *d_rows, bool seeded) {
    if (count == 0) return;
    hipStream_t st = ctx.stream;
    const uint64_t n = s.n;
    const uint32_t d = s.d, k = s.k;
    const unsigned blocks = div_up(count, QB);
    const uint32_t n_seg = exact_segments(n, k, count);
    ExactBufs e;
    take_exact(a, s, count, n_seg, e);
    const double *sk = seeded ? d_keys : nullptr;
    const uint32_t *sr = seeded ? d_rows : nullptr;
    const bool cos = norms != nullptr;
    auto scan = [&](uint32_t *rows, double *keys, uint32_t *cnt, uint32_t segs) {
        if (cos)
            launch_scan<true, true, 128>(padded_width(d), st, blocks, pts, nullptr, norms, n, d, d_list, 0, count, k,
                                         rows, keys, cnt, segs, sk, sr);
        else
            launch_scan<true, false, 128>(padded_width(d), st, blocks, pts, nullptr, nullptr, n, d, d_list, 0, count,
                                          k, rows, keys, cnt, segs, sk, sr);
    };
    if (n_seg <= 1) {
        scan(e.rows, e.keys, e.cnt, 1);
    } else {
        scan(e.seg_rows, e.seg_keys, e.seg_cnt, n_seg);
        if (n_seg * k <= 1024)
            merge_segments<1024><<<(unsigned)count, 64, 0, st>>>(e.seg_rows, e.seg_keys, e.seg_cnt, count, n_seg, k,
                                                                  e.rows, e.keys, e.cnt);
        else
            merge_segments<4096><<<(unsigned)count, 64, 0, st>>>(e.seg_rows, e.seg_keys, e.seg_cnt, count, n_seg, k,
                                                                  e.rows, e.keys, e.cnt);
        KY_HIP(hipGetLastError());
    }
    scatter_exact<<<div_up(count * k, 256), 256, 0, st>>>(d_list, count, k, e.rows, e.keys, d_keys, d_rows);
    KY_HIP(hipGetLastError());
}

void knn_graph(Context &ctx, const double *d_keys, const uint32_t *d_rows, uint64_t n, uint32_t k, int metric,
               uint32_t *g_idx, float *g_dist) {
    KY_CHECK(k <= 64, KY_E215, "neighbors: n_neighbors above 65");
    graph_rows<<<div_up(n, 64), 64, 0, ctx.stream>>>(d_keys, d_rows, n, k, metric == KY_METRIC_COSINE ? 1 : 0, g_idx,
                                                     g_dist);
    KY_HIP(hipGetLastError());
}

void knn_pack(Context &ctx, const float *x, uint64_t n, uint32_t d, uint32_t ld, float *out) {
    pack_columns<<<div_up(n * d, 256), 256, 0, ctx.stream>>>(x, n, d, ld, out);
    KY_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ driver ----

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
    KY_CHECK(used <= 64, KY_E501, "neighbors: no device kernel for more than 64 components");
    KY_CHECK(n_neighbors <= 65, KY_E501, "neighbors: no device kernel for n_neighbors above 65");
    return used;
}

// Fewest short rows a BF16 rescan takes (fewer go to the exact search).
constexpr uint32_t RESCAN_MIN_ROWS = 256;

void exact_search(Context &ctx, Arena &a, const float *pts, uint64_t n, uint32_t d, uint32_t k, int metric,
                  int precision, const uint32_t *q_rows, uint64_t nq, uint64_t round, double *keys, uint32_t *rows,
                  ky_knn_info *info, bool timed) {
    hipStream_t st = ctx.stream;
    const KnnShape s = knn_shape(&ctx, n, d, k, metric, precision, q_rows != nullptr, nq, round);
    KY_CHECK(s.precision == KY_PRECISION_BF16 || q_rows == nullptr, KY_E501,
             "neighbors: FP32 candidates for a row subset");
    const size_t m0 = a.mark();
    const double t_start = now_ms();
    KnnPrep prep;
    knn_take_prep(a, s, prep);
    uint32_t *unc = a.take<uint32_t>(nq);
    uint32_t *unc_count = a.take<uint32_t>(2);  // [0]: uncertified rows, [1]: short rows
    uint32_t *shrt = s.precision == KY_PRECISION_BF16 ? a.take<uint32_t>(nq) : nullptr;
    size_t m1 = a.mark();
    knn_prepare(ctx, s, pts, prep, s.metric == KY_METRIC_COSINE ? nullptr : a.take<double>(prep_partial_count(s)));
    a.release(m1);
    KY_CHECK(prep.finite, KY_E215, "neighbors: the embedding holds a value that is not finite");
    double search_ms = 0, rerank_ms = 0, fallback_ms = 0, scan_ms = 0, rescan_ms = 0;
    uint64_t rounds = 0;
    uint32_t total = 0, n_short = 0;
    fill_words(st, unc_count, 2, 0u);
    if (prep.filter_ok) {
        Bf16Points b;
        if (s.precision == KY_PRECISION_BF16) {
            bf16_take_points(a, s, b);
            m1 = a.mark();
            bf16_prepare(ctx, s, pts, prep, q_rows, b, a.take<unsigned char>(bf16_prepare_bytes(s)), k);
            a.release(m1);
        }
        for (uint64_t lo = 0; lo < nq; lo += s.round) {
            const uint64_t count = std::min<uint64_t>(s.round, nq - lo);
            ctx.check_cancel();
            KnnRound r;
            take_round(a, s, count, r);
            double t0 = now_ms();
            const uint32_t *round_rows = nullptr;
            if (s.precision == KY_PRECISION_FP32) {
                launch_scan<false, false, 128>(s.dp, st, div_up(count, QB), prep.cent, prep.offs, nullptr, n, d,
                                               nullptr, lo, count, s.m, r.rows, r.keys, r.count);
            } else {
                bf16_round(ctx, s, b, lo, count, r, timed ? &scan_ms : nullptr);
                round_rows = s.list ? b.qorig + lo : b.perm + lo;
            }
            if (timed) KY_HIP(hipStreamSynchronize(st));
            const double t1 = now_ms();
            search_ms += t1 - t0;
            const auto rr = prep.norms ? rerank<true> : rerank<false>;
            rr<<<div_up(count, 4), 256, 0, st>>>(
                pts, n, d, r.rows, r.keys, r.count, s.precision == KY_PRECISION_BF16 ? r.thr : nullptr, round_rows,
                round_rows ? 0 : lo, count, s.m, k, prep.slack, b.err, b.err_max, b.acc_rel,
                s.precision == KY_PRECISION_FP32 ? 0 : 1, prep.norms, keys, rows, unc, unc_count, shrt,
                unc_count + 1);
            KY_HIP(hipGetLastError());
            // The round's buffers are released for the next: wait for its work.
            KY_HIP(hipStreamSynchronize(st));
            rerank_ms += now_ms() - t1;
            a.release(m1);
            rounds++;
            ctx.advance(count);
        }
        uint32_t counts[2];
        KY_HIP(hipMemcpyAsync(counts, unc_count, sizeof(counts), hipMemcpyDeviceToHost, st));
        KY_HIP(hipStreamSynchronize(st));
        n_short = counts[1];
        const double t3 = now_ms();
        // Fewer short rows than one query block take the exact search at once (a rescan's
        // fixed costs, a sort and a launch over the device, are then above the exact
        // search's).
        if (n_short > 0 && n_short < RESCAN_MIN_ROWS) {
            KY_HIP(hipMemcpyAsync(unc + counts[0], shrt, sizeof(uint32_t) * n_short, hipMemcpyDeviceToDevice, st));
            counts[0] += n_short;
            n_short = 0;
        }
        // The short rows again, from higher starting thresholds (in rounds of the same size);
        // their re-rank adds any still uncertified to the exact search's list.
        const uint64_t chunk = std::min(s.round, s.nq);
        for (uint64_t off = 0; off < n_short; off += chunk) {
            const uint64_t count = std::min<uint64_t>(chunk, n_short - off);
            ctx.check_cancel();
            Bf16Rescan x;
            KnnRound r;
            bf16_take_rescan(a, s, count, x, r);
            bf16_rescan(ctx, s, b, shrt + off, count, x, r);
            const auto rr = prep.norms ? rerank<true> : rerank<false>;
            rr<<<div_up(count, 4), 256, 0, st>>>(pts, n, d, r.rows, r.keys, r.count, r.thr, x.qorig, 0, count, s.m,
                                                 k, prep.slack, b.err, b.err_max, b.acc_rel, 1, prep.norms, keys,
                                                 rows, unc, unc_count, nullptr, nullptr);
            KY_HIP(hipGetLastError());
            KY_HIP(hipStreamSynchronize(st));
            a.release(m1);
        }
        if (n_short) {
            KY_HIP(hipMemcpyAsync(&total, unc_count, sizeof(total), hipMemcpyDeviceToHost, st));
            KY_HIP(hipStreamSynchronize(st));
        } else {
            total = counts[0];
        }
        rescan_ms = now_ms() - t3;
    } else {
        // Norms too large for f32 keys: every query by the exact search.
        if (q_rows)
            KY_HIP(hipMemcpyAsync(unc, q_rows, sizeof(uint32_t) * nq, hipMemcpyDeviceToDevice, st));
        else
            iota_words(st, unc, nq, 0u);
        KY_HIP(hipGetLastError());
        total = (uint32_t)nq;
        search_ms = now_ms() - t_start;
    }
    const double t2 = now_ms();
    // Rows that went through the re-rank hold their candidates' k smallest exact pairs:
    // an upper bound of their k-th pair that seeds the exact search.
    const uint64_t chunk = std::min(s.round, s.nq);
    for (uint64_t off = 0; off < total; off += chunk) {
        const size_t mk = a.mark();
        knn_exact_rows(ctx, a, s, pts, prep.norms, unc + off, std::min<uint64_t>(chunk, total - off), keys, rows,
                       prep.filter_ok);
        KY_HIP(hipStreamSynchronize(st));
        a.release(mk);
    }
    fallback_ms = now_ms() - t2;
    a.release(m0);
    if (info) {
        info->candidates = prep.filter_ok ? s.m : 0;
        info->fallback = total;
        info->certified = nq - total;
        info->rescanned = n_short;
        info->rescan_ms = rescan_ms;
        info->search_ms = search_ms;
        info->rerank_ms = rerank_ms;
        info->fallback_ms = fallback_ms;
        info->scan_kernel_ms = scan_ms;
        info->rounds = rounds;
    }
}

void exact_search(Context &ctx, const float *pts, uint64_t n, uint32_t d, uint32_t k, int precision,
                  const uint32_t *q_rows, uint64_t nq, double *keys, uint32_t *rows, ky_knn_info *info) {
    const KnnShape s = knn_shape(&ctx, n, d, k, KY_METRIC_EUCLIDEAN, precision, q_rows != nullptr, nq, 0);
    const size_t need = knn_scratch_bytes(s, false, false);
    DevBuf<unsigned char> region(need);
    Arena a(region.ptr, need);
    exact_search(ctx, a, pts, n, d, k, KY_METRIC_EUCLIDEAN, precision, q_rows, nq, 0, keys, rows, info, true);
}

uint64_t knn_plan_round(const Context *ctx, uint64_t n, uint32_t d, uint32_t k, int metric, int precision, bool packed,
                        uint64_t limit, uint64_t *bytes) {
    const KnnShape s = knn_shape(ctx, n, d, k, metric, precision, false, n, 0);
    const uint64_t granule = s.precision == KY_PRECISION_BF16 ? 256 : QB;
    auto bytes_at = [&](uint64_t round) {
        return (uint64_t)knn_scratch_bytes(knn_shape(ctx, n, d, k, metric, precision, false, n, round), true, packed);
    };
    // Every row in one round, else the largest round that fits (bisection over
    // multiples of the granule; every round it settles on fits).
    *bytes = bytes_at(0);
    if (*bytes <= limit) return n;
    *bytes = bytes_at(granule);
    if (*bytes > limit) return 0;
    uint64_t lo = 1, hi = (n + granule - 1) / granule;  // lo fits, hi does not
    while (hi - lo > 1) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (bytes_at(mid * granule) <= limit) lo = mid;
        else hi = mid;
    }
    *bytes = bytes_at(lo * granule);
    return lo * granule;
}

// ------------------------------------------------- the fixed scratch region ----
// The scratch of ky_neighbors_static: one array a device, held from the library's
// first use on that device until the process ends. Host code never passes the array's
// name to a kernel (on the host it stands for no device address): its device address
// is asked for once per device, with that device current, and kept.
static __device__ __align__(256) unsigned char scratch[KY_STATIC_SCRATCH_BYTES];
constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mutex;
static void *g_scratch_ptr[MAX_DEVICES] = {};

void static_scratch_init(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    KY_CHECK(device >= 0 && device < MAX_DEVICES, KY_E504, "device " + std::to_string(device) + " out of range");
    if (g_scratch_ptr[device]) return;
    void *p = nullptr;
    KY_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    KY_CHECK(p != nullptr, KY_E505, "device scratch: no device address for the fixed array");
    g_scratch_ptr[device] = p;
}

static void *static_scratch(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    KY_CHECK(device >= 0 && device < MAX_DEVICES && g_scratch_ptr[device], KY_E505,
             "device scratch: the context's device has no fixed array (ky_open sets it up)");
    return g_scratch_ptr[device];
}

void neighbors_static(Context &ctx, const Dense &points, uint32_t n_components, uint32_t n_neighbors, int metric,
                      int precision, uint32_t *h_idx, float *h_dist, ky_knn_info *info) {
    const uint32_t d = check_search(points, n_components, n_neighbors, metric);
    KY_CHECK(precision == KY_PRECISION_FP32 || precision == KY_PRECISION_BF16, KY_E215, "neighbors: unknown precision");
    uint64_t bytes = 0;
    const uint64_t round = knn_plan_round(&ctx, points.rows, d, n_neighbors - 1, metric, precision,
                                          d != points.cols, KY_STATIC_SCRATCH_BYTES, &bytes);
    KY_CHECK(round > 0, KY_E215,
             "neighbors: " + std::to_string(points.rows) + " rows of " + std::to_string(d) +
                 " components need at least " + std::to_string(bytes) + " bytes of device scratch, above the " +
                 std::to_string((uint64_t)KY_STATIC_SCRATCH_BYTES) + " of the fixed region (use ky_neighbors)");
    neighbors_run(ctx, points, n_components, n_neighbors, metric, precision, round, h_idx, h_dist, info,
                  static_scratch(ctx.device), KY_STATIC_SCRATCH_BYTES);
}

void neighbors_run(Context &ctx, const Dense &points, uint32_t n_components, uint32_t n_neighbors, int metric,
                   int precision, uint64_t round, uint32_t *h_idx, float *h_dist, ky_knn_info *info, void *region,
                   size_t region_bytes) {
    const double t_start = now_ms();
    const uint32_t d = check_search(points, n_components, n_neighbors, metric);
    KY_CHECK(precision == KY_PRECISION_FP32 || precision == KY_PRECISION_BF16, KY_E215, "neighbors: unknown precision");
    hipStream_t st = ctx.stream;
    const uint64_t n = points.rows;
    const uint32_t k = n_neighbors - 1;
    const bool packed = d != points.cols;
    const KnnShape s = knn_shape(&ctx, n, d, k, metric, precision, false, n, round);
    const size_t need = knn_scratch_bytes(s, true, packed);
    DevBuf<unsigned char> own;
    if (region) {
        KY_CHECK(need <= region_bytes, KY_E215,
                 "neighbors: " + std::to_string(n) + " rows of " + std::to_string(d) + " components need " +
                     std::to_string(need) + " bytes of device scratch, above the " + std::to_string(region_bytes) +
                     " this entry point holds");
    } else {
        own.alloc(need);
        region = own.ptr;
    }
    Arena a(region, need);
    const float *pts = points.data.ptr;
    if (packed) {
        float *p = a.take<float>(n * d);
        knn_pack(ctx, points.data.ptr, n, d, points.cols, p);
        pts = p;
    }
    double *keys = a.take<double>(n * k);
    uint32_t *rows = a.take<uint32_t>(n * k);
    exact_search(ctx, a, pts, n, d, k, metric, precision, nullptr, n, s.round, keys, rows, info, true);
    uint32_t *g_idx = a.take<uint32_t>(n * k);
    float *g_dist = a.take<float>(n * k);
    knn_graph(ctx, keys, rows, n, k, metric, g_idx, g_dist);
    if (own.ptr) {
        download_parts(ctx, {{h_idx, g_idx, sizeof(uint32_t) * n * k}, {h_dist, g_dist, sizeof(float) * n * k}});
    } else {
        // A caller's scratch: nothing allocated, not even the pinned staging buffer.
        KY_HIP(hipMemcpyAsync(h_idx, g_idx, sizeof(uint32_t) * n * k, hipMemcpyDeviceToHost, st));
        KY_HIP(hipMemcpyAsync(h_dist, g_dist, sizeof(float) * n * k, hipMemcpyDeviceToHost, st));
        KY_HIP(hipStreamSynchronize(st));
    }
    if (info) {
        info->total_ms = now_ms() - t_start;
        info->scratch_bytes = need;
    }
}

// Pad rows of w floats (stride ld) to dp floats, zeros past w.
__global__ void pad_rows_f32(const float *__restrict__ x, uint64_t n, uint32_t w, uint32_t ld, uint32_t dp,
                             float *__restrict__ out) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n * dp) return;
    const uint64_t r = i / dp;
    const uint32_t c = (uint32_t)(i % dp);
    out[i] = c < w ? x[r * ld + c] : 0.0f;
}

void nearest_keys_run(Context &ctx, const float *h_queries, uint64_t nq, const Dense &points, int key,
                      const float *h_qn, const float *h_pn, const uint32_t *h_skip, uint32_t m, float *h_keys,
                      uint32_t *h_rows) {
    const uint64_t n = points.rows;
    const uint32_t w = points.cols;
    KY_CHECK(w >= 1 && n >= 1, KY_E215, "nearest_keys: points without rows or components");
    KY_CHECK(n - 1 <= 0xFFFFFFFEull, KY_E505, "nearest_keys: 2^32 points or more");
    KY_CHECK(key == KY_KEY_SQUARED_EUCLIDEAN || key == KY_KEY_COSINE, KY_E215, "nearest_keys: unknown key");
    KY_CHECK(key != KY_KEY_COSINE || (h_qn && h_pn), KY_E215, "nearest_keys: the cosine key needs both norms");
    if (nq == 0) return;
    bool skips_none = true;
    for (uint64_t i = 0; i < nq && skips_none; i++) skips_none = h_skip[i] == 0xFFFFFFFFu;
    KY_CHECK(m >= 1 && ((uint64_t)m < n || ((uint64_t)m == n && skips_none)), KY_E215,
             "nearest_keys: 1 <= m <= the points left after skip");
    KY_CHECK(m <= 64, KY_E501, "nearest_keys: no device kernel for more than 64 keys a query");
    KY_CHECK(w <= 64, KY_E501, "nearest_keys: no device kernel for more than 64 components");
    hipStream_t st = ctx.stream;
    const uint32_t dp = (uint32_t)padded_width(w);
    SizeArena z;
    auto layout = [&](auto &a, float *&pts, float *&qs, float *&raw, uint32_t *&skip, float *&qn, float *&pn,
                      uint32_t *&rows, float *&keys, uint32_t *&count) {
        pts = a.template take<float>(n * dp);
        qs = a.template take<float>(nq * dp);
        raw = a.template take<float>(nq * w);
        skip = a.template take<uint32_t>(nq);
        qn = a.template take<float>(key == KY_KEY_COSINE ? nq : 0);
        pn = a.template take<float>(key == KY_KEY_COSINE ? n : 0);
        rows = a.template take<uint32_t>(nq * m);
        keys = a.template take<float>(nq * m);
        count = a.template take<uint32_t>(nq);
    };
    float *pts, *qs, *raw, *qn, *pn, *keys;
    uint32_t *skip, *rows, *count;
    layout(z, pts, qs, raw, skip, qn, pn, rows, keys, count);
    DevBuf<unsigned char> region(z.high);
    Arena a(region.ptr, z.high);
    layout(a, pts, qs, raw, skip, qn, pn, rows, keys, count);
    pad_rows_f32<<<div_up(n * dp, 256), 256, 0, st>>>(points.data.ptr, n, w, w, dp, pts);
    KY_HIP(hipGetLastError());
    std::vector<CopyPart> up = {{raw, h_queries, sizeof(float) * nq * w}, {skip, h_skip, sizeof(uint32_t) * nq}};
    if (key == KY_KEY_COSINE) {
        up.push_back({qn, h_qn, sizeof(float) * nq});
        up.push_back({pn, h_pn, sizeof(float) * n});
    }
    upload_parts(ctx, up);
    pad_rows_f32<<<div_up(nq * dp, 256), 256, 0, st>>>(raw, nq, w, w, dp, qs);
    KY_HIP(hipGetLastError());
    Pairwise pw;
    pw.qa = qs;
    pw.skip = skip;
    pw.qn = qn;
    pw.pn = pn;
    const unsigned blocks = div_up(nq, QB);
    if (key == KY_KEY_SQUARED_EUCLIDEAN)
        launch_scan<false, false, 128, 1>(dp, st, blocks, pts, nullptr, nullptr, n, w, nullptr, 0, nq, m, rows, keys,
                                          count, 1, nullptr, nullptr, pw);
    else
        launch_scan<false, false, 128, 2>(dp, st, blocks, pts, nullptr, nullptr, n, w, nullptr, 0, nq, m, rows, keys,
                                          count, 1, nullptr, nullptr, pw);
    download_parts(ctx, {{h_keys, keys, sizeof(float) * nq * m}, {h_rows, rows, sizeof(uint32_t) * nq * m}});
}

}  // namespace ky
