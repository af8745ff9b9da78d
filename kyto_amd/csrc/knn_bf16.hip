// BF16 MFMA candidate search for neighbors() (gfx950, v_mfma_f32_16x16x32_bf16).
//
// Each row's centred f32 row a~ (knn.hip, prepare) is rounded to bf16, b = bf16(a~),
// and written twice with 64 bf16 components:
//   point  P'_p = [-2 b_p (d values) | o_p split in three bf16 | 1 1 1 | 0 ...]
//   query  Q'_q = [   b_q (d values) | 1 1 1 | o_q split in three bf16 | 0 ...]
// where o = |b|^2 (f64, split hi + mid + lo, each bf16), so one MFMA dot product
// over K = 64 is the whole key o_q + o_p - 2 b_q.b_p = |b_q - b_p|^2 (products of
// bf16 values are exact in the f32 accumulator), and the epilogue is a comparison
// only. d <= 58.
//
// Scan order: the rows are sorted by the Morton code of their first three (principal)
// components, and each block, owning 256 consecutive queries of that order, starts
// its scan a few tiles before its own position and wraps around. Its queries' near
// neighbours then come first, the thresholds are tight after the first tiles, and
// the appends that follow are rare. The order changes nothing in the result: each
// query's candidates are its m smallest (key, row) pairs over all rows.
//
// Block: 4 wavefronts, each owning 16*NB queries (NB MFMA column blocks, lane = query
// l % 16 of its block); the block streams the points in tiles of TILE rows through
// two LDS buffers (padded rows, conflict-free 16-byte fragment reads; the next tile
// loaded into registers during the current one, one barrier a tile). The M-blocks of
// a tile are software-pipelined: the next block's MFMAs are issued before this one's
// epilogue, a min of each lane's 4 keys against its query's threshold, one branch an
// M-block. A query keeps every point whose key is at most its threshold in a buffer
// in device memory (BCAP entries): the first INIT points of its scan directly, then
// appends; radix selection compacts a buffer back to its m smallest (key, row) pairs
// when the next tile could overflow it, and sets the threshold to the m-th key.
//
// The certificate (knn.hip, rerank, floor_mode 1) uses, per row, err = |b - a| (f64,
// a = x - mean exactly) and the accumulation bound delta_q = c (|b_q| + max|b|)^2:
// a point outside the candidates has a key >= T (the m-th candidate's), so
// |b_q - b_p| >= sqrt(T - delta_q), and its exact distance is at least that less
// err_q + max err.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "knn.hpp"
#include "sort.hpp"

namespace ky {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int KDIM = 64;        // augmented components
constexpr int TILE = 128;       // points per tile
constexpr int ROW_BYTES = 144;  // LDS row stride (128 data bytes + 16 padding)
constexpr int BCAP = 512;       // buffer entries per query (power of two)
constexpr int NB = 4;           // column blocks (16 queries each) per wavefront
constexpr int QPW = 16 * NB;    // queries per wavefront
constexpr int QPB = 4 * QPW;    // queries per block

// ---------------------------------------------------------------- preparation ----

// Row i of the scan order is original row perm[i]: P', Q' written at i, err and |b|
// at the original row.
__global__ void bf16_rows(const float *__restrict__ x, const double *__restrict__ mean,
                          const float *__restrict__ cent, const uint32_t *__restrict__ perm, uint64_t n, uint32_t d,
                          uint32_t dp, __bf16 *__restrict__ P, __bf16 *__restrict__ Q, double *__restrict__ err,
                          double *__restrict__ nb, unsigned long long *__restrict__ err_max,
                          unsigned long long *__restrict__ nb_max) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t r = perm[i];
    double o = 0.0, e2 = 0.0;
    __bf16 *pr = P + i * KDIM, *qr = Q + i * KDIM;
    for (uint32_t j = 0; j < d; j++) {
        __bf16 b = (__bf16)cent[r * dp + j];
        double bv = (double)(float)b;
        double exact = (double)x[r * d + j] - mean[j];
        double t = bv - exact;
        e2 += t * t;
        o += bv * bv;
        pr[j] = (__bf16)(-2.0f * (float)b);
        qr[j] = b;
    }
    __bf16 hi = (__bf16)(float)o;
    double r1 = o - (double)(float)hi;
    __bf16 mid = (__bf16)(float)r1;
    double r2 = r1 - (double)(float)mid;
    __bf16 lo = (__bf16)(float)r2;
    const __bf16 one = (__bf16)1.0f, zero = (__bf16)0.0f;
    pr[d] = hi; pr[d + 1] = mid; pr[d + 2] = lo;
    pr[d + 3] = one; pr[d + 4] = one; pr[d + 5] = one;
    qr[d] = one; qr[d + 1] = one; qr[d + 2] = one;
    qr[d + 3] = hi; qr[d + 4] = mid; qr[d + 5] = lo;
    for (uint32_t j = d + 6; j < KDIM; j++) {
        pr[j] = zero;
        qr[j] = zero;
    }
    double e = sqrt(e2), nrm = sqrt(o);
    err[r] = e;
    nb[r] = nrm;
    atomicMax(err_max, (unsigned long long)__double_as_longlong(e));
    atomicMax(nb_max, (unsigned long long)__double_as_longlong(nrm));
}

// Largest |value| of each of the first `c` components (bits of non-negative floats
// order as the floats).
__global__ void col_absmax(const float *__restrict__ cent, uint64_t n, uint32_t dp, uint32_t c,
                           unsigned *__restrict__ out) {
    uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (r >= n) return;
    for (uint32_t j = 0; j < c; j++) atomicMax(out + j, __float_as_uint(fabsf(cent[r * dp + j])));
}

// The Morton code of a row's first three components, each quantised to 10 bits over
// [-max, max]; vals = the row.
__global__ void morton_keys(const float *__restrict__ cent, uint64_t n, uint32_t dp, uint32_t c,
                            const unsigned *__restrict__ absmax, uint32_t *__restrict__ keys,
                            uint32_t *__restrict__ vals) {
    uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (r >= n) return;
    uint32_t code = 0;
    for (uint32_t j = 0; j < 3; j++) {
        uint32_t qv = 0;
        if (j < c) {
            const float mx = __uint_as_float(absmax[j]);
            const float v = mx > 0.0f ? cent[r * dp + j] / mx : 0.0f;
            qv = (uint32_t)fminf(fmaxf((v + 1.0f) * 512.0f, 0.0f), 1023.0f);
        }
        for (uint32_t bit = 0; bit < 10; bit++) code |= ((qv >> bit) & 1u) << (3 * bit + j);
    }
    keys[r] = code;
    vals[r] = (uint32_t)r;
}

// inv[perm[i]] = i.
__global__ void invert_perm(const uint32_t *__restrict__ perm, uint64_t n, uint32_t *__restrict__ inv) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < n) inv[perm[i]] = (uint32_t)i;
}

// keys[i] = inv[q_rows[i]] (the query's scan-order position), vals[i] = i.
__global__ void query_positions(const uint32_t *__restrict__ q_rows, uint64_t nq, const uint32_t *__restrict__ inv,
                                uint32_t *__restrict__ keys, uint32_t *__restrict__ vals) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= nq) return;
    keys[i] = inv[q_rows[i]];
    vals[i] = (uint32_t)i;
}

__global__ void bf16_delta(const double *__restrict__ nb, uint64_t n, double nb_max, double c,
                           double *__restrict__ delta) {
    uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (r >= n) return;
    double t = nb[r] + nb_max;
    delta[r] = c * t * t;
}

// ------------------------------------------------------------------- search ----

constexpr int INIT = 4 * TILE;   // the first points, kept without a threshold
constexpr int EPL = BCAP / 64;   // buffer entries per lane during a compaction

// Number of entries (over the wave's registers) with key <= c.
__device__ __forceinline__ unsigned count_le(const uint32_t (&k)[EPL], unsigned E, uint32_t c) {
    unsigned s = 0;
#pragma unroll
    for (int e = 0; e < EPL; e++)
        if ((unsigned)e < E) s += (unsigned)__popcll(__ballot(k[e] <= c));
    return s;
}

// Compact a query's buffer (cnt entries in device memory) to its `keep` smallest
// (key, row) pairs by radix selection in registers (no sort): the keep-th pair
// (T, R) is found bit by bit, then the entries at or below it are written back in
// lane order. When cnt >= keep, *thr becomes T's key. With `sorted_out`, the kept
// entries are also sorted into the wave's LDS scratch (keep <= 64). Whole wavefront.
__device__ unsigned compact_buffer(float *__restrict__ bk, uint32_t *__restrict__ br, unsigned cnt, unsigned keep,
                                   float *thr, bool sorted_out, uint32_t *skey, uint32_t *srow) {
    const unsigned lane = lane_id();
    wave_sync_memory();
    const unsigned E = (cnt + 63) / 64;
    uint32_t k[EPL], r[EPL];
#pragma unroll
    for (int e = 0; e < EPL; e++) {
        unsigned i = lane + 64 * e;
        bool valid = (unsigned)e < E && i < cnt;
        k[e] = valid ? order_f32(bk[i]) : 0xFFFFFFFFu;
        r[e] = valid ? br[i] : 0xFFFFFFFFu;
    }
    uint32_t T = 0xFFFFFFFFu, R = 0xFFFFFFFFu;
    unsigned kept = cnt;
    if (cnt >= keep) {
        T = 0;
        for (int b = 31; b >= 0; b--) {
            uint32_t cand = T | ((1u << b) - 1u);
            if (count_le(k, E, cand) < keep) T |= 1u << b;
        }
        const unsigned less = T ? count_le(k, E, T - 1u) : 0u;
        const unsigned need = keep - less;  // >= 1 entries of key T
        unsigned eq = 0;
#pragma unroll
        for (int e = 0; e < EPL; e++)
            if ((unsigned)e < E) eq += (unsigned)__popcll(__ballot(k[e] == T));
        if (eq == 1) {
#pragma unroll
            for (int e = 0; e < EPL; e++) {
                uint64_t msk = (unsigned)e < E ? __ballot(k[e] == T) : 0ull;
                if (msk) R = __builtin_amdgcn_readlane((int)r[e], __builtin_ctzll(msk));
            }
        } else {
            R = 0;
            for (int b = 31; b >= 0; b--) {
                uint32_t cand = R | ((1u << b) - 1u);
                unsigned c = 0;
#pragma unroll
                for (int e = 0; e < EPL; e++)
                    if ((unsigned)e < E) c += (unsigned)__popcll(__ballot(k[e] == T && r[e] <= cand));
                if (c < need) R |= 1u << b;
            }
        }
        *thr = unorder_f32(T);
        kept = keep;
    }
    unsigned base = 0;
#pragma unroll
    for (int e = 0; e < EPL; e++) {
        if ((unsigned)e >= E) break;
        bool sel = (lane + 64 * e < cnt) && (k[e] < T || (k[e] == T && r[e] <= R));
        uint64_t msk = __ballot(sel);
        if (sel) {
            unsigned pos = base + mbcnt(msk);
            if (sorted_out) {
                skey[pos] = k[e];
                srow[pos] = r[e];
            } else {
                bk[pos] = unorder_f32(k[e]);
                br[pos] = r[e];
            }
        }
        base += (unsigned)__popcll(msk);
    }
    if (sorted_out) {
        for (unsigned i = kept + lane; i < 64; i += 64) {
            skey[i] = 0xFFFFFFFFu;
            srow[i] = 0xFFFFFFFFu;
        }
        wave_sync();
        wave_bitonic<uint32_t>(skey, srow, 64);
    }
    wave_sync_memory();
    return kept;
}

constexpr int INIT_TILES = INIT / TILE;
constexpr int LEAD_TILES = 16;  // tiles scanned before a block's own position

// Query i (< nq) is row qpos[i] of the scan order (ascending in i; row i when qpos is
// null); its outputs go to slot qout[i]; point rows are reported as the original rows
// perm[p].
__global__ __launch_bounds__(256) void search_bf16(const __bf16 *__restrict__ P, const __bf16 *__restrict__ Q,
                                                   uint64_t n, const uint32_t *__restrict__ qpos,
                                                   const uint32_t *__restrict__ qout, uint64_t nq,
                                                   const uint32_t *__restrict__ perm, uint32_t m,
                                                   float *__restrict__ buf_key, uint32_t *__restrict__ buf_row,
                                                   uint32_t *__restrict__ out_rows, float *__restrict__ out_keys,
                                                   uint32_t *__restrict__ out_count) {
    __shared__ __attribute__((aligned(16))) unsigned char tile[2][TILE * ROW_BYTES];
    __shared__ uint32_t scratch_key[4][64];
    __shared__ uint32_t scratch_row[4][64];
    __shared__ unsigned qcnt[QPB];
    __shared__ float qthr[QPB];
    const unsigned tid = threadIdx.x, lane = lane_id(), wave = tid / 64;
    const uint64_t qbase = (uint64_t)blockIdx.x * QPB + (uint64_t)wave * QPW;  // this wave's first query slot
    unsigned *cnt = qcnt + wave * QPW;
    float *thr_lds = qthr + wave * QPW;
    for (unsigned i = lane; i < QPW; i += 64) {
        cnt[i] = 0;
        thr_lds[i] = INFINITY;
    }
    // B fragments: the wave's queries, loaded once.
    bf16x8 bq[NB][2];
    float thr[NB];
    uint32_t qrow[NB];  // scan-order row of each of this lane's query slots (n for none)
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const uint64_t i = qbase + 16 * b + (lane % 16);
        qrow[b] = i < nq ? (qpos ? qpos[i] : (uint32_t)i) : (uint32_t)n;
        const bf16x8 *src = reinterpret_cast<const bf16x8 *>(Q + (uint64_t)min<uint32_t>(qrow[b], (uint32_t)(n - 1)) * KDIM);
#pragma unroll
        for (int ks = 0; ks < 2; ks++) bq[b][ks] = src[4 * ks + lane / 16];
        thr[b] = INFINITY;
    }
    const bool wave_live = qbase < nq;
    const uint64_t n_tiles = (n + TILE - 1) / TILE;
    const uint64_t q0 = min<uint64_t>((uint64_t)blockIdx.x * QPB, nq - 1);
    const uint64_t own = (qpos ? (uint64_t)qpos[q0] : q0) / TILE;
    const uint64_t lead = min<uint64_t>(LEAD_TILES, n_tiles - 1);
    const uint64_t first_tile = (own + n_tiles - lead) % n_tiles;
    // Tile staging through registers, one tile ahead: 128 rows x 128 bytes, 4 x 16 bytes a thread.
    uint4 pre[TILE * 8 / 256];
    auto fetch = [&](uint64_t t0) {
#pragma unroll
        for (int s = 0; s < TILE * 8 / 256; s++) {
            unsigned i = tid + 256 * s, row = i / 8, chunk = i % 8;
            uint64_t p = t0 + row;
            pre[s] = p < n ? reinterpret_cast<const uint4 *>(P + p * KDIM)[chunk] : make_uint4(0, 0, 0, 0);
        }
    };
    auto stage = [&](int buf) {
#pragma unroll
        for (int s = 0; s < TILE * 8 / 256; s++) {
            unsigned i = tid + 256 * s, row = i / 8, chunk = i % 8;
            *reinterpret_cast<uint4 *>(tile[buf] + row * ROW_BYTES + chunk * 16) = pre[s];
        }
    };
    fetch(first_tile * TILE);
    stage(0);
    __syncthreads();
    for (uint64_t it = 0; it < n_tiles; it++) {
        const int cur = (int)(it & 1);
        const uint64_t t0 = ((first_tile + it) % n_tiles) * TILE;
        if (it + 1 < n_tiles) fetch(((first_tile + it + 1) % n_tiles) * TILE);
        const bool init = it < INIT_TILES;
        if (wave_live) {
            const unsigned char *tb = tile[cur];
            f32x4 acc[2][NB];
            auto mfmas = [&](int mb, f32x4 (&out)[NB]) {
                const unsigned prow = mb * 16 + (lane % 16);
                const bf16x8 a0 = *reinterpret_cast<const bf16x8 *>(tb + prow * ROW_BYTES + (lane / 16) * 16);
                const bf16x8 a1 = *reinterpret_cast<const bf16x8 *>(tb + prow * ROW_BYTES + (4 + lane / 16) * 16);
#pragma unroll
                for (int b = 0; b < NB; b++) {
                    f32x4 z = {0.0f, 0.0f, 0.0f, 0.0f};
                    z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, bq[b][0], z, 0, 0, 0);
                    out[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, bq[b][1], z, 0, 0, 0);
                }
            };
            mfmas(0, acc[0]);
#pragma unroll
            for (int mb = 0; mb < TILE / 16; mb++) {
                const f32x4(&a)[NB] = acc[mb & 1];
                if (mb + 1 < TILE / 16) mfmas(mb + 1, acc[(mb + 1) & 1]);
                // Lane l, slot j: scan row t0 + 16 mb + 4 (l / 16) + j, query 16 b + l % 16.
                const uint64_t p0 = t0 + mb * 16 + 4 * (lane / 16);
                if (init) {
                    // The first INIT rows of the scan go to buffer slot (tile, row); rows past
                    // the end and the query itself as (+inf, none).
#pragma unroll
                    for (int b = 0; b < NB; b++) {
                        const uint64_t i = qbase + 16 * b + (lane % 16);
#pragma unroll
                        for (int j = 0; j < 4; j++) {
                            const uint64_t p = p0 + j;
                            if (i < nq) {
                                const bool valid = p < n && p != qrow[b];
                                const uint64_t slot = it * TILE + mb * 16 + 4 * (lane / 16) + j;
                                buf_key[i * BCAP + slot] = valid ? a[b][j] : INFINITY;
                                buf_row[i * BCAP + slot] = valid ? (uint32_t)p : 0xFFFFFFFFu;
                            }
                        }
                    }
                    continue;
                }
                bool any = false;
#pragma unroll
                for (int b = 0; b < NB; b++) {
                    const float lo = fminf(fminf(a[b][0], a[b][1]), fminf(a[b][2], a[b][3]));
                    any = any || lo <= thr[b];
                }
                if (__ballot(any)) {
#pragma unroll
                    for (int b = 0; b < NB; b++) {
                        const unsigned ql = 16 * b + (lane % 16);
                        const uint64_t i = qbase + ql;
#pragma unroll
                        for (int j = 0; j < 4; j++) {
                            const uint64_t p = p0 + j;
                            if (a[b][j] <= thr[b] && p < n && p != qrow[b] && i < nq) {
                                const unsigned slot = atomicAdd(&cnt[ql], 1u);
                                buf_key[i * BCAP + slot] = a[b][j];
                                buf_row[i * BCAP + slot] = (uint32_t)p;
                            }
                        }
                    }
                }
            }
            if (init) {
                if (it + 1 == INIT_TILES || it + 1 == n_tiles) {
                    // The seed set is complete: every live query holds `seeded` slots.
                    const unsigned seeded = (unsigned)((it + 1) * TILE);
                    for (unsigned ql = 0; ql < QPW; ql++) {
                        const uint64_t i = qbase + ql;
                        if (i >= nq) break;
                        float t = INFINITY;
                        const unsigned kept = compact_buffer(buf_key + i * BCAP, buf_row + i * BCAP, seeded, m, &t,
                                                             false, scratch_key[wave], scratch_row[wave]);
                        if (lane == 0) {
                            cnt[ql] = kept;
                            thr_lds[ql] = t;
                        }
                        wave_sync();
                    }
                }
            } else {
                // Compact the buffers that the next tile could overflow.
                wave_sync();
                const bool need = lane < QPW && cnt[lane] > BCAP - TILE;
                uint64_t mask = __ballot(need);
                while (mask) {
                    const unsigned ql = __builtin_ctzll(mask);
                    mask &= mask - 1;
                    const uint64_t i = qbase + ql;
                    float t = thr_lds[ql];
                    const unsigned kept = compact_buffer(buf_key + i * BCAP, buf_row + i * BCAP, cnt[ql], m, &t,
                                                         false, scratch_key[wave], scratch_row[wave]);
                    if (lane == 0) {
                        cnt[ql] = kept;
                        thr_lds[ql] = t;
                    }
                    wave_sync();
                }
            }
#pragma unroll
            for (int b = 0; b < NB; b++) thr[b] = thr_lds[16 * b + (lane % 16)];
        }
        if (it + 1 < n_tiles) stage(cur ^ 1);
        __syncthreads();
    }
    if (!wave_live) return;
    for (unsigned ql = 0; ql < QPW; ql++) {
        const uint64_t i = qbase + ql;
        if (i >= nq) break;
        float t = thr_lds[ql];
        const unsigned kept = compact_buffer(buf_key + i * BCAP, buf_row + i * BCAP, cnt[ql], m, &t, true,
                                             scratch_key[wave], scratch_row[wave]);
        const uint64_t o = qout[i];
        for (unsigned e = lane; e < m; e += 64) {
            const uint32_t row = scratch_row[wave][e];
            out_rows[o * m + e] = e < kept && row < n ? perm[row] : 0xFFFFFFFFu;
            out_keys[o * m + e] = e < kept ? unorder_f32(scratch_key[wave][e]) : INFINITY;
        }
        if (lane == 0) out_count[o] = kept;
        wave_sync();
    }
}

// ================================================================== host side ====

void bf16_candidates(Context &ctx, const float *pts, const KnnPrep &prep, const uint32_t *q_rows, uint64_t nq,
                     uint32_t m, uint32_t *d_rows, float *d_keys, uint32_t *d_count, DevBuf<double> &err,
                     double &err_max, double &acc_rel) {
    KY_CHECK(prep.d + 6 <= KDIM, KY_E501, "neighbors: no BF16 kernel for more than 58 components");
    KY_CHECK(m >= 1 && m <= 64, KY_E215, "neighbors: candidates out of range");
    hipStream_t st = ctx.stream;
    const uint64_t n = prep.n;
    // The scan order: rows sorted by the Morton code of their first three components.
    DevBuf<unsigned> absmax(3);
    KY_HIP(hipMemsetAsync(absmax.ptr, 0, 3 * sizeof(unsigned), st));
    const uint32_t c = std::min<uint32_t>(3, prep.d);
    col_absmax<<<div_up(n, 256), 256, 0, st>>>(prep.cent.ptr, n, prep.dp, c, absmax.ptr);
    DevBuf<uint32_t> keys(n), rows(n), skeys(n), perm(n);
    morton_keys<<<div_up(n, 256), 256, 0, st>>>(prep.cent.ptr, n, prep.dp, c, absmax.ptr, keys.ptr, rows.ptr);
    KY_HIP(hipGetLastError());
    DevBuf<unsigned char> sort_tmp;
    sort_pairs(st, keys.ptr, skeys.ptr, rows.ptr, perm.ptr, n, 30, sort_tmp);
    // Rows in scan order.
    DevBuf<__bf16> P(n * KDIM), Q(n * KDIM);
    err.alloc(2 * n);  // err, then delta (original rows)
    DevBuf<double> nb(n);
    DevBuf<unsigned long long> maxima(2);
    KY_HIP(hipMemsetAsync(maxima.ptr, 0, 2 * sizeof(unsigned long long), st));
    bf16_rows<<<div_up(n, 256), 256, 0, st>>>(pts, prep.mean.ptr, prep.cent.ptr, perm.ptr, n, prep.d, prep.dp, P.ptr,
                                               Q.ptr, err.ptr, nb.ptr, maxima.ptr, maxima.ptr + 1);
    KY_HIP(hipGetLastError());
    unsigned long long hm[2];
    KY_HIP(hipMemcpyAsync(hm, maxima.ptr, sizeof(hm), hipMemcpyDeviceToHost, st));
    KY_HIP(hipStreamSynchronize(st));
    std::memcpy(&err_max, &hm[0], sizeof(double));
    double nb_max;
    std::memcpy(&nb_max, &hm[1], sizeof(double));
    // Accumulation bound of a K-term f32 dot product in any order, doubled:
    // 2 (K + 8) u over (|b_q| + |b_p|)^2 (the terms' absolute sum).
    const double cb = 2.0 * (KDIM + 8.0) / 16777216.0;
    bf16_delta<<<div_up(n, 256), 256, 0, st>>>(nb.ptr, n, nb_max, cb, err.ptr + n);
    KY_HIP(hipGetLastError());
    acc_rel = cb;
    // Queries in scan order: every row (positions 0..n, outputs at the original row),
    // or the listed rows sorted by their position (outputs at their list index).
    DevBuf<uint32_t> qpos_buf, qout_buf;
    const uint32_t *qpos, *qout;
    if (!q_rows) {
        KY_CHECK(nq == n, KY_E215, "neighbors: every row or a list");
        qpos = nullptr;
        qout = perm.ptr;
    } else {
        DevBuf<uint32_t> inv(n), k2(nq), v2(nq);
        invert_perm<<<div_up(n, 256), 256, 0, st>>>(perm.ptr, n, inv.ptr);
        query_positions<<<div_up(nq, 256), 256, 0, st>>>(q_rows, nq, inv.ptr, k2.ptr, v2.ptr);
        KY_HIP(hipGetLastError());
        qpos_buf.alloc(nq);
        qout_buf.alloc(nq);
        sort_pairs(st, k2.ptr, qpos_buf.ptr, v2.ptr, qout_buf.ptr, nq, 32, sort_tmp);
        KY_HIP(hipStreamSynchronize(st));
        qpos = qpos_buf.ptr;
        qout = qout_buf.ptr;
    }
    DevBuf<float> bk(nq * BCAP);
    DevBuf<uint32_t> br(nq * BCAP);
    const unsigned blocks = div_up(nq, QPB);
    search_bf16<<<blocks, 256, 0, st>>>(P.ptr, Q.ptr, n, qpos, qout, nq, perm.ptr, m, bk.ptr, br.ptr, d_rows, d_keys,
                                        d_count);
    KY_HIP(hipGetLastError());
    KY_HIP(hipStreamSynchronize(st));
}

}  // namespace ky
