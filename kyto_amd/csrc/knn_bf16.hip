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
// three LDS buffers (global_load_lds two tiles ahead, XOR-swizzled rows for
// conflict-free 16-byte fragment reads, one barrier a tile). Within a tile, super-
// blocks of SPB points: the next super-block's fragments are read from LDS while this
// one's MFMAs and hit test run; the hit test is a min tree of each lane's keys of a
// query against its threshold, one scalar branch a super-block. A query keeps every
// point whose key is at most its threshold in a buffer in device memory (BCAP
// entries), appended through a per-wave LDS stage; radix selection compacts a buffer
// back to its m smallest (key, row) pairs when the next tile could overflow it, and
// sets the threshold to the m-th key.
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
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x8 __attribute__((ext_vector_type(8)));

constexpr int KDIM = 64;        // augmented components
constexpr int TILE = 128;       // points per tile
constexpr int ROW_BYTES = 128;  // LDS row stride (chunks XOR-swizzled by row)
constexpr int BCAP = 512;       // buffer entries per query (power of two)
constexpr int NB = 4;           // column blocks (16 queries each) per wavefront
constexpr int MB = 2;           // row blocks (16 points each) per super-block
constexpr int SPB = 16 * MB;    // points per super-block
constexpr int QPW = 16 * NB;    // queries per wavefront
constexpr int QPB = 4 * QPW;    // queries per block
constexpr int GMAX_PER_TILE = 4 * (TILE / SPB);  // symmetric scan: lane groups x super-blocks a tile

// ---------------------------------------------------------------- preparation ----

// Original row r is row inv[r] of the scan order: its P', Q' are written there, err and
// |b| at r (the rows read in order, each 128-byte row of P' and Q' written whole). 16
// lanes a row (lane t: components 4t .. 4t + 3; fixed butterflies over the 16 add o =
// |b|^2 and err^2), 4 rows a wavefront at a time, BROWS rows a block, one atomic a
// block for the largest err and |b|.
constexpr unsigned BROWS = 64;

__global__ __launch_bounds__(256) void bf16_rows(const float *__restrict__ x, const double *__restrict__ mean,
                                                 const double *__restrict__ norms,
                                                 const float *__restrict__ cent, const uint32_t *__restrict__ inv,
                                                 uint64_t n, uint32_t d, uint32_t dp, __bf16 *__restrict__ P,
                                                 __bf16 *__restrict__ Q, double *__restrict__ err,
                                                 double *__restrict__ nb, unsigned long long *__restrict__ err_max,
                                                 unsigned long long *__restrict__ nb_max) {
    __shared__ double wmax[2][4];
    const unsigned lane = lane_id(), wave = threadIdx.x / 64, t = lane % 16;
    double mu[4];
#pragma unroll
    for (int u = 0; u < 4; u++) mu[u] = (!norms && 4 * t + u < d) ? mean[4 * t + u] : 0.0;
    double emx = 0.0, nmx = 0.0;
    for (unsigned q0 = wave * 4; q0 < BROWS; q0 += 16) {
        const uint64_t r = (uint64_t)blockIdx.x * BROWS + q0 + lane / 16;
        const bool live = r < n;
        const uint64_t i = live ? inv[r] : 0;
        __bf16 b[4];
        double o = 0.0, e2 = 0.0;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const unsigned c = 4 * t + u;
            b[u] = (__bf16)0.0f;
            if (live && c < d) {
                b[u] = (__bf16)cent[r * dp + c];
                const double bv = (double)(float)b[u];
                // The row b stands for: x - mean (Euclidean), or x / |x| (cosine; 0 for
                // a zero row), in f64.
                const double exact = norms ? (norms[r] != 0.0 ? (double)x[r * d + c] / norms[r] : 0.0)
                                           : (double)x[r * d + c] - mu[u];
                const double tt = bv - exact;
                e2 += tt * tt;
                o += bv * bv;
            }
        }
#pragma unroll
        for (int off = 8; off >= 1; off /= 2) {
            o += __shfl_xor(o, off, 16);
            e2 += __shfl_xor(e2, off, 16);
        }
        const __bf16 hi = (__bf16)(float)o;
        const double r1 = o - (double)(float)hi;
        const __bf16 mid = (__bf16)(float)r1;
        const double r2 = r1 - (double)(float)mid;
        const __bf16 lo = (__bf16)(float)r2;
        const __bf16 one = (__bf16)1.0f, zero = (__bf16)0.0f;
        __bf16 pv[4], qv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const unsigned c = 4 * t + u;
            pv[u] = zero;
            qv[u] = zero;
            if (c < d) {
                pv[u] = (__bf16)(-2.0f * (float)b[u]);
                qv[u] = b[u];
            } else if (c < d + 3) {
                pv[u] = c == d ? hi : c == d + 1 ? mid : lo;
                qv[u] = one;
            } else if (c < d + 6) {
                const unsigned v = c - (d + 3);
                pv[u] = one;
                qv[u] = v == 0 ? hi : v == 1 ? mid : lo;
            }
        }
        if (live) {
            *reinterpret_cast<uint2 *>(P + i * KDIM + 4 * t) = __builtin_bit_cast(uint2, pv);
            *reinterpret_cast<uint2 *>(Q + i * KDIM + 4 * t) = __builtin_bit_cast(uint2, qv);
            if (t == 0) {
                const double e = sqrt(e2), nrm = sqrt(o);
                err[r] = e;
                nb[r] = nrm;
                emx = fmax(emx, e);
                nmx = fmax(nmx, nrm);
            }
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off /= 2) {
        emx = fmax(emx, __shfl_xor(emx, off));
        nmx = fmax(nmx, __shfl_xor(nmx, off));
    }
    if (lane == 0) {
        wmax[0][wave] = emx;
        wmax[1][wave] = nmx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double e = fmax(fmax(wmax[0][0], wmax[0][1]), fmax(wmax[0][2], wmax[0][3]));
        const double m = fmax(fmax(wmax[1][0], wmax[1][1]), fmax(wmax[1][2], wmax[1][3]));
        atomicMax(err_max, (unsigned long long)__double_as_longlong(e));
        atomicMax(nb_max, (unsigned long long)__double_as_longlong(m));
    }
}

// Padding rows of P' (after the last row): key 1e30 against every query.
__global__ void pad_rows(__bf16 *__restrict__ P, uint64_t first, uint64_t count, uint32_t d) {
    uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (t >= count * KDIM) return;
    const uint32_t k = (uint32_t)(t % KDIM);
    P[first * KDIM + t] = (__bf16)(k == d ? 1e30f : 0.0f);
}

// Largest |value| of each of the first `c` (<= 3) components (bits of non-negative
// floats order as the floats): a grid-stride loop, a block maximum, one atomic a block
// and component.
__global__ __launch_bounds__(256) void col_absmax(const float *__restrict__ cent, uint64_t n, uint32_t dp, uint32_t c,
                                                  unsigned *__restrict__ out) {
    __shared__ unsigned part[3][256];
    unsigned m[3] = {0u, 0u, 0u};
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < n; r += (uint64_t)gridDim.x * blockDim.x)
        for (uint32_t j = 0; j < c; j++) m[j] = max(m[j], __float_as_uint(fabsf(cent[r * dp + j])));
    for (uint32_t j = 0; j < 3; j++) part[j][threadIdx.x] = m[j];
    __syncthreads();
    for (unsigned s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (uint32_t j = 0; j < 3; j++) part[j][threadIdx.x] = max(part[j][threadIdx.x], part[j][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x < c) atomicMax(out + threadIdx.x, part[threadIdx.x][0]);
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

// qorig[i] = q_rows[qorig[i]] (list indices made rows).
__global__ void list_rows(const uint32_t *__restrict__ q_rows, uint64_t nq, uint32_t *__restrict__ qorig) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < nq) qorig[i] = q_rows[qorig[i]];
}

// The rescan's starting thresholds: `factor` times the first round's, by scan position.
__global__ void rescan_thresholds(const uint32_t *__restrict__ qpos, uint64_t nq, const float *__restrict__ thr,
                                  float factor, float *__restrict__ out) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < nq) out[i] = factor * thr[qpos[i]];
}

__global__ void bf16_delta(const double *__restrict__ nb, uint64_t n, double nb_max, double c, double absolute,
                           double *__restrict__ delta) {
    uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (r >= n) return;
    double t = nb[r] + nb_max;
    delta[r] = c * t * t + absolute;
}

// ------------------------------------------------------------------- search ----

constexpr int EPL = BCAP / 64;  // buffer entries per lane during a compaction

// A buffer entry: the key's f32 bits and the point's scan-order row.
typedef uint2 Entry;

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
__device__ unsigned compact_buffer(Entry *__restrict__ buf, unsigned cnt, unsigned keep, float *thr, bool sorted_out,
                                   uint32_t *skey, uint32_t *srow) {
    const unsigned lane = lane_id();
    wave_sync_memory();
    const unsigned E = (cnt + 63) / 64;
    uint32_t k[EPL], r[EPL];
#pragma unroll
    for (int e = 0; e < EPL; e++) {
        unsigned i = lane + 64 * e;
        bool valid = (unsigned)e < E && i < cnt;
        Entry v = valid ? buf[i] : make_uint2(0, 0);
        k[e] = valid ? order_f32(__uint_as_float(v.x)) : 0xFFFFFFFFu;
        r[e] = valid ? v.y : 0xFFFFFFFFu;
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
                buf[pos] = make_uint2(__float_as_uint(unorder_f32(k[e])), r[e]);
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

// The lane's index as an opaque value the compiler cannot hoist out of a loop or keep
// live across it (ALU only).
__device__ __forceinline__ uint32_t fresh_lane() {
    uint32_t l;
    asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(l));
    return l;
}

constexpr int LEAD_TILES = 16;        // tiles scanned before a block's own position
constexpr uint64_t EST_MIN_ROWS = 1 << 17;  // threshold estimates from a sample above this size
// The sample: every 16th row of the scan order. With 32 subsets of n / 512 sampled
// rows and the p-th smallest minimum as the estimate, about -ln(1 - p/32) * 512 rows
// fall at or below it on average; p is the smallest giving 5 (k + 1) (p = 5 and 87
// rows for k = 14, p = 9 and 169 rows for k = 29). Measured at 1M (k = 14): every
// 12th row costs 2.1 ms more estimate than it saves in rows scanned again; every
// 20th leaves 1,628 rows short and scans slower. Aiming at 5 (k + 1) rows in place of
// 7 (k + 1) keeps fewer keys (scan -3.2 ms at 1M, -26 ms at 4M) for more rows scanned
// again at 1M (3,858 rows in 2.9 ms, 560 in 1.3) and fewer at 4M (1,835 against
// 7,497: fewer lists overflow); 4.5 (k + 1) (p = 4) leaves 22,441 rows short.
constexpr uint32_t EST_STRIDE = 16;

constexpr unsigned SCAP = 64;  // staged appends per wavefront
constexpr unsigned MAX_TAIL_SEG = 64;                // point segments of a tail query block
constexpr unsigned MERGE_CAP = MAX_TAIL_SEG * 64;    // a tail query's merged lists (m <= 64)

// The candidate scan. Points: the np rows of the scan order (P'). Queries: i < nq, query i being row qpos[i] of the scan order (row
// qoff + i when qpos is null; ascending in i), outputs at slot o = i (a round of
// queries has slots of its own).
//
// Each query keeps in its buffer (BCAP entries in device memory) every point whose
// key is at most its threshold: thr_init[i] at the start (+inf when null), then,
// after a compaction, the m-th smallest key so far. Outputs per slot: the m smallest
// (key, row) sorted (rows as original rows perm[.], when out_rows is not null), their
// count, and out_thr[o]: the m-th key when m were kept, else the threshold every
// point above which was left out.
//
// A query's count and threshold live in registers of the 4 lanes that hold its keys
// (lane l % 16 of each column block); an append takes its slot from a ballot, and is
// staged in LDS (SCAP entries a wave) and written out when the stage fills, before a
// compaction and at the end. Global stores are then rare, so the wait for the next
// tile at the end of each tile seldom waits for stores as well.
//
// Blocks are mapped to queries XCD by XCD (block b runs on XCD b % 8): the blocks
// resident on one XCD hold consecutive queries, start their scans at nearby tiles
// and read the same tiles through that XCD's L2 at about the same time.
template <bool SYM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) void search_bf16(
    const __bf16 *__restrict__ P, const __bf16 *__restrict__ Q, uint64_t np,
    const uint32_t *__restrict__ qpos, uint64_t qoff, uint64_t nq,
    const uint32_t *__restrict__ perm, uint32_t m, const float *__restrict__ thr_init, Entry *__restrict__ buf,
    uint32_t *__restrict__ out_rows, float *__restrict__ out_keys, uint32_t *__restrict__ out_count,
    float *__restrict__ out_thr, uint32_t main_blocks, uint32_t n_seg, uint32_t seg_tiles,
    uint32_t *__restrict__ seg_rows, float *__restrict__ seg_keys, uint32_t *__restrict__ seg_count,
    float *__restrict__ seg_thr, uint32_t *__restrict__ tcount, Entry *__restrict__ tlist, uint32_t tcap,
    const float *__restrict__ thr_max) {
    __shared__ __attribute__((aligned(16))) unsigned char tile[3][TILE * ROW_BYTES];
    // Symmetric scan: the starting thresholds of the points of each tile buffer's tile,
    // then the largest of each of its super-blocks'.
    __shared__ __attribute__((aligned(16))) float tile_thr[SYM ? 3 : 1][SYM ? TILE + GMAX_PER_TILE : 4];
    // Per wave: the append stage, which a compaction's sort scratch overlays.
    __shared__ __attribute__((aligned(16))) uint32_t stage_key[4][SCAP];
    __shared__ __attribute__((aligned(16))) uint32_t stage_row[4][SCAP];
    __shared__ uint16_t stage_dst[4][SCAP];
    __shared__ uint32_t qrow_lds[SYM ? 1 : QPB];
    const unsigned tid = threadIdx.x, lane = lane_id(), wave = tid / 64;
    // Blocks below main_blocks scan every point for query block lblk (XCD-aware order);
    // the others are the tail: query block main_blocks + t / n_seg over point segment
    // t % n_seg (seg_tiles tiles), with its own buffers and outputs. The symmetric scan
    // has no tail; block b is query block b (the longest first).
    const bool tail = !SYM && blockIdx.x >= main_blocks;
    unsigned lblk, seg = 0;
    if (SYM) {
        lblk = blockIdx.x;
    } else if (!tail) {
        const unsigned nblk = main_blocks, xcd = blockIdx.x % 8, per = nblk / 8, extra = nblk % 8;
        lblk = xcd * per + min(xcd, extra) + blockIdx.x / 8;
    } else {
        const unsigned t = blockIdx.x - main_blocks;
        lblk = main_blocks + t / n_seg;
        seg = t % n_seg;
    }
    const uint64_t main_q = (uint64_t)main_blocks * QPB;
    // Buffer (and tail output) slot of query i.
    auto slot = [&](uint64_t i) { return tail ? main_q + (i - main_q) * n_seg + seg : i; };
    const uint64_t qbase = (uint64_t)lblk * QPB + (uint64_t)wave * QPW;  // this wave's first query
    // The scan-order row of each of the wave's queries (the symmetric scan's are
    // consecutive).
    uint32_t *qrow = qrow_lds + (SYM ? 0 : wave * QPW);
    auto query_row = [&](unsigned j) { return SYM ? (uint32_t)(qbase + j) : qrow[j]; };
    // B fragments (the wave's queries, loaded once), and per column block b the
    // threshold and buffer count of this lane's query 16 b + l % 16.
    bf16x8 bq[NB][2];
    float thr[NB];
    unsigned cb[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const uint64_t i = qbase + 16 * b + (lane % 16);
        const uint64_t ic = min<uint64_t>(i, nq - 1);
        const uint32_t r = qpos ? qpos[ic] : (uint32_t)(qoff + ic);
        if (!SYM && lane < 16) qrow[16 * b + lane] = r;
        const bf16x8 *src = reinterpret_cast<const bf16x8 *>(Q + (uint64_t)r * KDIM);
#pragma unroll
        for (int ks = 0; ks < 2; ks++) bq[b][ks] = src[4 * ks + lane / 16];
        // Slots past the last query take no point.
        thr[b] = i < nq ? (thr_init ? thr_init[i] : INFINITY) : -INFINITY;
        cb[b] = 0;
    }
    unsigned staged = 0;  // wave-uniform
    // Whether the wave appended to its queries' buffers since the last compaction check
    // (wave-uniform): without appends the counts and thresholds are unchanged, and so
    // is the check's answer (none needs compacting).
    bool appended = true;
    const bool wave_live = qbase < nq;
    // The staged appends to the buffers (one store instruction). In the symmetric scan
    // a stage entry with bit 15 of its destination set is a key for the point's own row
    // (row = the point, low bits = the wave's query): it takes a slot of the point's
    // list by an atomic (slots past tcap are counted, not written).
    auto flush = [&](const uint32_t *sk, const uint32_t *sr, const uint16_t *sd) {
        wave_sync();
        if (lane < staged) {
            const unsigned dst = sd[lane];
            if (SYM && (dst & 0x8000u)) {
                const uint32_t p = sr[lane];
                const unsigned at = atomicAdd(tcount + p, 1u);
                if (at < tcap) tlist[(uint64_t)p * tcap + at] = make_uint2(sk[lane], query_row(dst & (QPW - 1)));
            } else {
                buf[slot(qbase + (dst >> 9)) * BCAP + (dst & (BCAP - 1))] = make_uint2(sk[lane], sr[lane]);
            }
        }
        staged = 0;
        wave_sync();
    };
    // Compact query 16 b + l of the wave (whole wavefront; uniform b, l).
    auto compact_query = [&](int b, unsigned l, bool sorted, uint32_t *sk, uint32_t *sr) {
        const unsigned c = __builtin_amdgcn_readlane(cb[b], l);
        float t = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(thr[b]), l));
        const unsigned kept = compact_buffer(buf + slot(qbase + 16 * b + l) * BCAP, c, m, &t, sorted, sk, sr);
        if (lane % 16 == l) {
            cb[b] = kept;
            thr[b] = t;
        }
        return kept;
    };
    const uint64_t n_tiles = (np + TILE - 1) / TILE;
    const uint64_t q0 = min<uint64_t>((uint64_t)lblk * QPB, nq - 1);
    const uint64_t own = (qpos ? (uint64_t)qpos[q0] : qoff + q0) / TILE;
    const uint64_t lead = min<uint64_t>(LEAD_TILES, n_tiles - 1);
    // Without a starting threshold a block starts a few tiles before its own queries'
    // position (their near neighbours first tighten the thresholds early); with one,
    // every block starts at tile 0, so the blocks resident on an XCD read the same
    // tiles at about the same time through its L2.
    // A tail block scans its segment only.
    // The symmetric scan: query block b scans from its own tiles to the last (each pair
    // of rows once, by the block of the earlier row).
    static_assert(QPB % TILE == 0, "a query block spans whole tiles");
    const uint64_t first_tile = SYM    ? (uint64_t)lblk * (QPB / TILE)
                                : tail ? (uint64_t)seg * seg_tiles
                                : thr_init ? 0
                                           : (min<uint64_t>(own, n_tiles - 1) + n_tiles - lead) % n_tiles;
    const uint64_t n_scan = SYM ? n_tiles - first_tile : tail ? min<uint64_t>(seg_tiles, n_tiles - first_tile) : n_tiles;
    // Tiles go straight from global memory to LDS (global_load_lds_dwordx4) into three
    // buffers: tile t + 2 is requested once tile t's keys are done, so it has the
    // whole of tile t + 1 to arrive. A tile is 128 rows of 128 bytes with the 16-byte
    // chunk c of row r stored at chunk c ^ (r % 8) of its row, so the 16 rows an MFMA
    // fragment read touches fall on different banks. Each wave instruction fills 8 rows
    // (lane l: row l / 8, stored chunk l % 8). Rows past the last point read the padding
    // rows after P' (keys 1e30, never kept).
    // Each lane's source offset in the tile is one 32-bit value (row % 8 = lane / 8 for
    // every instruction, whose rows are 8 apart), instruction s s KiB further. The
    // offset is computed from a fresh lane index at each fetch: an address kept live
    // across the scan loop is spilled, and a spill's reload waits for every load in
    // flight, the next tile's included.
    auto fetch = [&](uint64_t t0, unsigned char *__restrict__ dst) {
        const uint32_t fl = fresh_lane();
        const uint32_t fetch_off = ((wave * (TILE / 32) * 8 + fl / 8) * KDIM + (((fl % 8) ^ (fl / 8)) * 8)) * 2;
        const unsigned char *tp = reinterpret_cast<const unsigned char *>(P + t0 * KDIM) + fetch_off;
#pragma unroll
        for (int s = 0; s < TILE / 32; s++)
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(tp + s * 8 * ROW_BYTES),
                                             (__attribute__((address_space(3))) void *)(dst + (wave * (TILE / 32) + s) * 1024),
                                             16, 0, 0);
    };
    // The symmetric scan's thresholds of a tile's points (thr_init holds -inf past the
    // last point), 64 by each of waves 0 and 1, and the largest of each lane group's
    // points in each super-block (thr_max) by 16 lanes of wave 2, requested with the
    // tile (one more load in flight for those waves).
    auto fetch_thr = [&](uint64_t t0, float *__restrict__ dthr) {
        if (!SYM) return;
        const uint32_t fl = fresh_lane();
        if (wave < 2)
            __builtin_amdgcn_global_load_lds(
                (const __attribute__((address_space(1))) void *)(reinterpret_cast<const unsigned char *>(thr_init + t0 + 64 * wave) + 4 * fl),
                (__attribute__((address_space(3))) void *)(dthr + 64 * wave), 4, 0, 0);
        else if (wave == 2 && fl < GMAX_PER_TILE)
            __builtin_amdgcn_global_load_lds(
                (const __attribute__((address_space(1))) void *)(reinterpret_cast<const unsigned char *>(thr_max + t0 / TILE * GMAX_PER_TILE) + 4 * fl),
                (__attribute__((address_space(3))) void *)(dthr + TILE), 4, 0, 0);
    };
    // Tile indices advance by one and wrap (no division in the loop).
    auto wrap_next = [&](uint32_t t) { return t + 1 == (uint32_t)n_tiles ? 0u : t + 1; };
    uint32_t tcur = (uint32_t)first_tile, tfetch = wrap_next(tcur);
    fetch((uint64_t)tcur * TILE, tile[0]);
    fetch_thr((uint64_t)tcur * TILE, tile_thr[0]);
    if (n_scan > 1) {
        fetch((uint64_t)tfetch * TILE, tile[1]);
        fetch_thr((uint64_t)tfetch * TILE, tile_thr[SYM ? 1 : 0]);
    }
    tfetch = wrap_next(tfetch);
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    __syncthreads();
    int cur = 0;
    for (uint64_t it = 0; it < n_scan; it++) {
        const uint64_t t0 = (uint64_t)tcur * TILE;
        const int nxt = cur == 2 ? 0 : cur + 1, after = nxt == 2 ? 0 : nxt + 1;
        // tb: this tile, nx: the buffer tile t + 2 goes to; sk, sr, sd: the stage. The
        // restrict qualifiers tell the compiler that the stage and the tile reads do
        // not touch the buffer in flight, so it does not wait for it before them.
        auto body = [&](const unsigned char *__restrict__ tb, unsigned char *__restrict__ nx,
                        const float *__restrict__ tt, float *__restrict__ ntt, uint32_t *__restrict__ sk,
                        uint32_t *__restrict__ sr, uint16_t *__restrict__ sd) {
            // Past the block's own two tiles, the symmetric scan tests each key against the
            // point's threshold as well.
            const bool cross = SYM && it >= QPB / TILE;
            if (wave_live) {
                // The point fragments of super-block sb (row blocks g of 16 points, the two
                // halves of K); those of the next super-block are read from LDS while this
                // one's MFMAs and hit test run.
                bf16x8 a0[MB], a1[MB];
                // Symmetric scan: the largest threshold of the super-block's points whose
                // keys this lane holds, read with its fragments.
                float pmax_next = 0.0f;
                auto fragments = [&](int sb) {
                    if (SYM) pmax_next = tt[TILE + sb * 4 + lane / 16];
#pragma unroll
                    for (int g = 0; g < MB; g++) {
                        const unsigned prow = sb * SPB + g * 16 + (lane % 16);
                        const unsigned char *rp = tb + prow * ROW_BYTES;
                        a0[g] = *reinterpret_cast<const bf16x8 *>(rp + (((lane / 16) ^ (prow % 8)) * 16));
                        a1[g] = *reinterpret_cast<const bf16x8 *>(rp + (((4 + lane / 16) ^ (prow % 8)) * 16));
                    }
                };
                fragments(0);
#pragma nounroll
                for (int sb = 0; sb < TILE / SPB; sb++) {
                    // Keys of the super-block: acc[b][4 g + j] is point 16 g + 4 (lane / 16) + j
                    // of the super-block, query 16 b + lane % 16 of the wave.
                    f32x16 acc[NB];
                    // The first halves of K of every (g, b), then the second halves: each
                    // accumulator's two MFMAs are MB * NB apart.
                    f32x4 z[MB][NB];
                    const float pmax = pmax_next;
                    // Raised wave priority while this wave issues its MFMAs (and reads the
                    // next fragments), back to normal for the hit test: the SIMD's MFMA
                    // pipe is fed first (0.9 ms a launch, A/B on one box).
                    __builtin_amdgcn_s_setprio(1);
#pragma unroll
                    for (int g = 0; g < MB; g++)
#pragma unroll
                        for (int b = 0; b < NB; b++) {
                            const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
                            z[g][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0[g], bq[b][0], zero, 0, 0, 0);
                        }
                    // The second halves column block by column block, each block's hit test
                    // placed after the next block's MFMAs: the test's VALU instructions issue
                    // while this wave's own MFMAs run (scheduling barriers keep the order).
                    // The test: the smallest of each lane's 4 MB keys of a query (a tree of
                    // 3-input mins) against the query's threshold; the column blocks' ballots
                    // OR'ed in scalar registers, one scalar branch.
                    bool hit[NB];
                    float lo_b[NB];
                    uint64_t hits = 0;
                    auto test_block = [&](int b) {
                        float part[(4 * MB + 2) / 3];
#pragma unroll
                        for (int e = 0; e < 4 * MB; e += 3)
                            part[e / 3] = e + 2 < 4 * MB ? fminf(fminf(acc[b][e], acc[b][e + 1]), acc[b][e + 2])
                                                         : e + 1 < 4 * MB ? fminf(acc[b][e], acc[b][e + 1])
                                                                          : acc[b][e];
                        float lo = part[0];
#pragma unroll
                        for (int e = 1; e < (4 * MB + 2) / 3; e += 2)
                            lo = e + 1 < (4 * MB + 2) / 3 ? fminf(fminf(lo, part[e]), part[e + 1]) : fminf(lo, part[e]);
                        lo_b[b] = lo;
                        hit[b] = lo <= thr[b];
                        hits |= __ballot(hit[b]);
                    };
#pragma unroll
                    for (int b = 0; b < NB; b++) {
#pragma unroll
                        for (int g = 0; g < MB; g++) {
                            const f32x4 r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1[g], bq[b][1], z[g][b], 0, 0, 0);
                            if (g == 0) acc[b].s0123 = r;
                            else if (g == 1) acc[b].s4567 = r;
                            else if (g == 2) acc[b].s89ab = r;
                            else acc[b].scdef = r;
                        }
                        if (b == NB - 1 && sb + 1 < TILE / SPB) fragments(sb + 1);
                        if (b > 0) test_block(b - 1);
                        __builtin_amdgcn_sched_barrier(0);
                    }
                    __builtin_amdgcn_s_setprio(0);
                    test_block(NB - 1);
                    // The symmetric scan's second test: the smallest of the lane's NB keys of
                    // each of its points against the point's threshold, made only when the
                    // smallest of all the lane's keys (from the minima above) is at or below
                    // the largest threshold of its points in some lane; most super-blocks
                    // end at that one comparison.
                    // psites: bit e set when some lane's key e is at or below its point's threshold.
                    uint32_t psites = 0;
                    auto point_thresholds = [&](float (&pt)[4 * MB]) {
#pragma unroll
                        for (int g = 0; g < MB; g++) {
                            const float4 t4 = *reinterpret_cast<const float4 *>(tt + sb * SPB + 16 * g + 4 * (lane / 16));
                            pt[4 * g] = t4.x;
                            pt[4 * g + 1] = t4.y;
                            pt[4 * g + 2] = t4.z;
                            pt[4 * g + 3] = t4.w;
                        }
                    };
                    static_assert(NB == 4, "the point tests take the minimum of 4 column blocks");
                    if (cross && __ballot(fminf(fminf(lo_b[0], lo_b[1]), fminf(lo_b[2], lo_b[3])) <= pmax)) {
                        float pthr[4 * MB];
                        point_thresholds(pthr);
                        auto point_min = [&](int e) {
                            return fminf(fminf(fminf(acc[0][e], acc[1][e]), acc[2][e]), acc[3][e]);
                        };
                        // First whether any point has a key at or below its threshold: the
                        // smallest of (key minimum - threshold) over the lane's points
                        // against 0 (x - t <= 0 iff x <= t for finite keys; a flushed
                        // difference only lets a point through), then, rarely, which.
                        float dmin = point_min(0) - pthr[0];
#pragma unroll
                        for (int e = 1; e < 4 * MB; e++) dmin = fminf(dmin, point_min(e) - pthr[e]);
                        if (__ballot(dmin <= 0.0f)) {
#pragma unroll
                            for (int e = 0; e < 4 * MB; e++)
                                if (__ballot(point_min(e) <= pthr[e])) psites |= 1u << e;
                        }
                    }
                    if (!hits && !psites) continue;
                    // Appends: the 4 lanes of a query take buffer slots in lane order after
                    // its count, and stage slots after the wave's staged entries. The keys of
                    // a column block are walked with a uniform index into a 16-register
                    // vector (VGPR indexing), not unrolled.
                    const uint64_t pbase = t0 + sb * SPB + 4 * (lane / 16);
                    if (hits) appended = true;
#pragma unroll
                    for (int b = 0; b < NB; b++) {
                        if (!__ballot(hit[b])) continue;
                        // The keys with a hit in some lane, one ballot a key. (No partial mins
                        // here: the compiler would share them with the test above and build
                        // that tree from 2-input mins, a VALU instruction more a column block
                        // on every super-block.)
                        uint32_t sites = 0;
#pragma unroll
                        for (int e = 0; e < 4 * MB; e++)
                            if (__ballot(acc[b][e] <= thr[b])) sites |= 1u << e;
                        const uint32_t self = query_row(16 * b + (lane % 16));
                        const uint64_t same = 0x0001000100010001ull << (lane % 16);
                        while (sites) {
                            const int e = __builtin_ctz(sites);
                            sites &= sites - 1;
                            const float x = acc[b][e];
                            const uint64_t p = pbase + 16 * (e / 4) + (e % 4);
                            const bool h = x <= thr[b] && p < np && p != self;
                            const uint64_t mh = __ballot(h);
                            if (!mh) continue;
                            if (staged + (unsigned)__popcll(mh) > SCAP) flush(sk, sr, sd);
                            const uint64_t mq = mh & same;
                            if (h) {
                                const unsigned slot = staged + mbcnt(mh);
                                sk[slot] = __float_as_uint(x);
                                sr[slot] = (uint32_t)p;
                                sd[slot] = (uint16_t)(((16 * b + (lane % 16)) << 9) | (cb[b] + mbcnt(mq)));
                            }
                            staged += (unsigned)__popcll(mh);
                            cb[b] += (unsigned)__popcll(mq);
                        }
                    }
                    // Keys at or below their point's threshold, staged for the points' lists
                    // (walked key by key, then column block by column block).
                    // The keys of a point are walked with a uniform index into the lanes'
                    // key vectors and thresholds (VGPR indexing), not unrolled: one copy of
                    // the append code a column block (the kernel's code stays small enough
                    // for the instruction cache).
                    if (psites) {
                        f32x8 pthr;
                        {
                            float pt[4 * MB];
                            point_thresholds(pt);
#pragma unroll
                            for (int e = 0; e < 4 * MB; e++) pthr[e] = pt[e];
                        }
                        while (psites) {
                            const int e = __builtin_ctz(psites);
                            psites &= psites - 1;
                            const uint32_t p = (uint32_t)(pbase + 16 * (e / 4) + (e % 4));
                            const float pt = pthr[e];
#pragma unroll
                            for (int b = 0; b < NB; b++) {
                                const float x = acc[b][e];
                                const bool h = x <= pt;
                                const uint64_t mh = __ballot(h);
                                if (!mh) continue;
                                if (staged + (unsigned)__popcll(mh) > SCAP) flush(sk, sr, sd);
                                if (h) {
                                    const unsigned at = staged + mbcnt(mh);
                                    sk[at] = __float_as_uint(x);
                                    sr[at] = p;
                                    sd[at] = (uint16_t)(0x8000u | (16 * b + (lane % 16)));
                                }
                                staged += (unsigned)__popcll(mh);
                            }
                        }
                    }
                }
                // Compact the buffers that the next tile could overflow, and those holding
                // m entries and either no threshold yet or 2m + 64 entries (the threshold
                // then tightens to their m-th key).
                uint64_t need[NB];
                bool any_need = false;
                if (appended) {
#pragma unroll
                    for (int b = 0; b < NB; b++) {
                        const unsigned c = cb[b];
                        need[b] = __ballot(c > BCAP - TILE || (c >= m && (thr[b] == INFINITY || c >= 2 * m + 64))) & 0xFFFFull;
                        any_need = any_need || need[b];
                    }
                    appended = false;
                }
                if (any_need) {
                    flush(sk, sr, sd);
                    wave_sync_memory();
#pragma unroll
                    for (int b = 0; b < NB; b++) {
                        while (need[b]) {
                            const unsigned l = __builtin_ctzll(need[b]);
                            need[b] &= need[b] - 1;
                            compact_query(b, l, false, sk, sr);
                        }
                    }
                }
            }
            if (it + 2 < n_scan) {
                // The thresholds first: any wait their addresses need then waits for tile
                // t + 1's loads only.
                fetch_thr((uint64_t)tfetch * TILE, ntt);
                fetch((uint64_t)tfetch * TILE, nx);
                // All but tile t + 2's loads (tile t + 1 is in LDS).
                if (SYM && wave < 3)
                    __builtin_amdgcn_s_waitcnt(0x0F75);  // vmcnt(5)
                else
                    __builtin_amdgcn_s_waitcnt(0x0F74);  // vmcnt(4)
            } else {
                __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
            }
        };
        body(tile[cur], tile[after], tile_thr[SYM ? cur : 0], tile_thr[SYM ? after : 0], stage_key[wave],
             stage_row[wave], stage_dst[wave]);
        // A bare barrier: each wave has waited for its own part of tile t + 1 above,
        // and the LDS the block shares in this loop is written by those loads only (a
        // __syncthreads() fence would also wait for tile t + 2). The signal fences keep
        // the compiler from moving memory operations across it.
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_barrier();
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        cur = nxt;
        tcur = wrap_next(tcur);
        tfetch = wrap_next(tfetch);
    }
    if (!wave_live) return;
    flush(stage_key[wave], stage_row[wave], stage_dst[wave]);
    wave_sync_memory();
    for (unsigned ql = 0; ql < QPW; ql++) {
        const uint64_t i = qbase + ql;
        if (i >= nq) break;
        unsigned kept = 0;
        float t = 0.0f;
        // ql / 16 is uniform; the unrolled test keeps the register arrays statically indexed.
#pragma unroll
        for (int b = 0; b < NB; b++)
            if ((unsigned)b == ql / 16) {
                kept = compact_query(b, ql % 16, true, stage_key[wave], stage_row[wave]);
                t = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(thr[b]), ql % 16));
            }
        const uint64_t o = tail ? (i - main_q) * n_seg + seg : i;
        uint32_t *const o_rows = tail ? seg_rows : out_rows;
        float *const o_keys = tail ? seg_keys : out_keys;
        if (o_rows) {
            // The symmetric scan leaves scan-order rows (its merge maps them).
            for (unsigned e = lane; e < m; e += 64) {
                const uint32_t row = stage_row[wave][e];
                o_rows[o * m + e] = e < kept && row < 0xFFFFFFFFu ? (SYM ? row : perm[row]) : 0xFFFFFFFFu;
                o_keys[o * m + e] = e < kept ? unorder_f32(stage_key[wave][e]) : INFINITY;
            }
        }
        if (lane == 0) {
            (tail ? seg_count : out_count)[o] = kept;
            (tail ? seg_thr : out_thr)[o] = t;
        }
        wave_sync();
    }
}

// The tail's segment lists merged (one wavefront a query): every segment kept all its
// points at or below its threshold (its m-th key, or the starting one), so below the
// smallest of them, T, the union is whole; its m smallest pairs at or below T are the
// query's candidates, as a scan of every point would leave them (the threshold the
// m-th key when m were kept, else T). Output at slot slot0 + j.
__global__ __launch_bounds__(64) void merge_tail(const uint32_t *__restrict__ seg_rows,
                                                 const float *__restrict__ seg_keys,
                                                 const uint32_t *__restrict__ seg_count,
                                                 const float *__restrict__ seg_thr, uint64_t nq, uint32_t n_seg,
                                                 uint32_t m, uint64_t slot0,
                                                 uint32_t *__restrict__ out_rows, float *__restrict__ out_keys,
                                                 uint32_t *__restrict__ out_count, float *__restrict__ out_thr) {
    __shared__ uint32_t skey[MERGE_CAP];
    __shared__ uint32_t srow[MERGE_CAP];
    const uint64_t j = blockIdx.x;
    if (j >= nq) return;
    const unsigned lane = lane_id();
    float T = INFINITY;
    for (uint32_t s = 0; s < n_seg; s++) T = fminf(T, seg_thr[j * n_seg + s]);
    unsigned total = 0;  // wave-uniform
    for (uint32_t s = 0; s < n_seg; s++) {
        const uint64_t sl = j * n_seg + s;
        const unsigned cnt = seg_count[sl];
        for (unsigned e0 = 0; e0 < cnt; e0 += 64) {
            const unsigned e = e0 + lane;
            const float key = e < cnt ? seg_keys[sl * m + e] : INFINITY;
            const bool keep = e < cnt && key <= T;
            const uint64_t mask = __ballot(keep);
            if (keep) {
                const unsigned at = total + mbcnt(mask);
                skey[at] = order_f32(key);
                srow[at] = seg_rows[sl * m + e];
            }
            total += (unsigned)__popcll(mask);
        }
    }
    unsigned N = 64;
    while (N < total) N <<= 1;
    for (unsigned e = total + lane; e < N; e += 64) {
        skey[e] = 0xFFFFFFFFu;
        srow[e] = 0xFFFFFFFFu;
    }
    wave_sync();
    wave_bitonic<uint32_t>(skey, srow, N);
    const unsigned kept = min(total, m);
    const uint64_t o = slot0 + j;
    for (unsigned e = lane; e < m; e += 64) {
        out_rows[o * m + e] = e < kept ? srow[e] : 0xFFFFFFFFu;
        out_keys[o * m + e] = e < kept ? unorder_f32(skey[e]) : INFINITY;
    }
    if (lane == 0) {
        out_count[o] = kept;
        out_thr[o] = total >= m ? unorder_f32(skey[m - 1]) : T;
    }
}

// The largest starting threshold of each lane group's points in each super-block:
// entry 4 s + j holds the largest over the points of super-block s whose keys lanes
// 16 j .. 16 j + 15 hold (point 16 g + 4 j + i, g < MB, i < 4; thr holds -inf past
// the last point, through the last tile).
__global__ __launch_bounds__(256) void group_max(const float *__restrict__ thr, uint64_t n_g,
                                                 float *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n_g) return;
    const uint64_t p0 = (i / 4) * SPB + 4 * (i % 4);
    float v = -INFINITY;
#pragma unroll
    for (int g = 0; g < MB; g++)
#pragma unroll
        for (int j = 0; j < 4; j++) v = fmaxf(v, thr[p0 + 16 * g + j]);
    out[i] = v;
}

// The symmetric scan's lists joined (one wavefront a row, four a block): row r's own
// scan kept every point of the later tiles at or below its threshold T (its m-th key,
// or the starting one), and its list holds every earlier row at or below the starting
// threshold, which is at least T; below T the union is whole. Its m smallest (key,
// scan-order row) pairs at or below T replace the row's candidates, as rows (perm[.]).
// A list that overflowed tcap leaves the row without candidates (the exact search takes
// it). m <= 64 (candidates_for).
constexpr uint32_t SYM_TCAP = 320;     // list entries a row

__global__ __launch_bounds__(256) void merge_sym(const uint32_t *__restrict__ perm, uint64_t n, uint32_t m,
                                                 const uint32_t *__restrict__ tcount, const Entry *__restrict__ tlist,
                                                 uint32_t *__restrict__ rows, float *__restrict__ keys,
                                                 uint32_t *__restrict__ count, float *__restrict__ thr) {
    __shared__ uint32_t skey[4][64];
    __shared__ uint32_t srow[4][64];
    const unsigned lane = lane_id(), wave = threadIdx.x / 64;
    const uint64_t r = (uint64_t)blockIdx.x * 4 + wave;
    if (r >= n) return;
    uint32_t *sk = skey[wave], *sr = srow[wave];
    const unsigned own = count[r], listed = tcount[r];
    const float T = thr[r];
    if (listed > SYM_TCAP) {
        for (unsigned e = lane; e < m; e += 64) {
            rows[r * m + e] = 0xFFFFFFFFu;
            keys[r * m + e] = INFINITY;
        }
        if (lane == 0) count[r] = 0;
        return;
    }
    {
        // In registers: slot 0 the row's own candidates (at most m), slots 1 .. the
        // listed keys at or below T (a free slot holds (~0, ~0)). The m smallest (key,
        // row) pairs are those at or below the m-th, found bit by bit (its key, then,
        // among the pairs of that key, its row) by counting ballots; they are gathered
        // one a lane through LDS and sorted in registers.
        constexpr int SL = 1 + (int)(SYM_TCAP / 64);
        const unsigned ns = 1 + (listed + 63) / 64;  // slots in use (wave-uniform)
        uint32_t kk[SL], rr[SL];
        bool ok[SL];
        ok[0] = lane < own;
        kk[0] = ok[0] ? order_f32(keys[r * m + lane]) : 0xFFFFFFFFu;
        rr[0] = ok[0] ? rows[r * m + lane] : 0xFFFFFFFFu;
        unsigned total = own;
#pragma unroll
        for (int sl = 1; sl < SL; sl++) {
            ok[sl] = false;
            kk[sl] = rr[sl] = 0xFFFFFFFFu;
            if ((unsigned)sl < ns) {
                const unsigned e = 64 * (sl - 1) + lane;
                const Entry v = e < listed ? tlist[r * SYM_TCAP + e] : make_uint2(0, 0);
                ok[sl] = e < listed && __uint_as_float(v.x) <= T;
                if (ok[sl]) {
                    kk[sl] = order_f32(__uint_as_float(v.x));
                    rr[sl] = v.y;
                }
                total += (unsigned)__popcll(__ballot(ok[sl]));
            }
        }
        // At most 64 listed pairs kept (the usual case): they are sorted in registers and
        // merged with the row's own candidates, already ascending by (key, row) (a
        // bitonic merge: one list reversed, the smaller of each pair, then a
        // half-cleaner cascade), so no selection is needed.
        const unsigned listed_kept = total - own;
        if (listed_kept <= 64) {
            unsigned at0 = 0;  // wave-uniform
#pragma unroll
            for (int sl = 1; sl < SL; sl++) {
                if ((unsigned)sl >= ns) continue;
                const uint64_t mask = __ballot(ok[sl]);
                if (ok[sl]) {
                    const unsigned at = at0 + mbcnt(mask);
                    sk[at] = kk[sl];
                    sr[at] = rr[sl];
                }
                at0 += (unsigned)__popcll(mask);
            }
            wave_sync();
            unsigned long long b = lane < listed_kept ? (unsigned long long)sk[lane] : ~0ull;
            uint32_t rb = lane < listed_kept ? sr[lane] : 0xFFFFFFFFu;
            wave_sort64(b, rb);
            unsigned long long a = ok[0] ? (unsigned long long)kk[0] : ~0ull;
            uint32_t ra = rr[0];
            const unsigned long long br = __shfl(b, 63 - (int)lane);
            const uint32_t rbr = (uint32_t)__shfl((int)rb, 63 - (int)lane);
            if (br < a || (br == a && rbr < ra)) {
                a = br;
                ra = rbr;
            }
#pragma unroll
            for (unsigned stride = 32; stride > 0; stride >>= 1) {
                const unsigned long long ok2 = __shfl_xor(a, stride);
                const uint32_t orow = (uint32_t)__shfl_xor((int)ra, stride);
                const bool other_less = ok2 < a || (ok2 == a && orow < ra);
                if (other_less == ((lane & stride) == 0)) {
                    a = ok2;
                    ra = orow;
                }
            }
            const unsigned kept2 = min(total, m);
            if (lane < m) {
                rows[r * m + lane] = lane < kept2 ? perm[ra] : 0xFFFFFFFFu;
                keys[r * m + lane] = lane < kept2 ? unorder_f32((uint32_t)a) : INFINITY;
            }
            const uint32_t mth2 = (uint32_t)__shfl(a, (int)m - 1);
            if (lane == 0) {
                count[r] = kept2;
                thr[r] = total >= m ? unorder_f32(mth2) : T;
            }
            return;
        }
        // Keep (key, row) <= (K, R); every listed pair when there are at most m.
        uint32_t K = 0xFFFFFFFFu, R = 0xFFFFFFFFu;
        if (total > m) {
            const auto count_if = [&](auto pred) {
                unsigned c = 0;
#pragma unroll
                for (int sl = 0; sl < SL; sl++)
                    if ((unsigned)sl < ns) c += (unsigned)__popcll(__ballot(ok[sl] && pred(kk[sl], rr[sl])));
                return c;
            };
            // K: the largest key with fewer than m pairs below it (the m-th key).
            K = 0;
            for (int bit = 31; bit >= 0; bit--) {
                const uint32_t t = K | (1u << bit);
                if (count_if([&](uint32_t kv, uint32_t) { return kv < t; }) < m) K = t;
            }
            const unsigned need = m - count_if([&](uint32_t kv, uint32_t) { return kv < K; });
            if (count_if([&](uint32_t kv, uint32_t) { return kv == K; }) > need) {
                R = 0;
                for (int bit = 31; bit >= 0; bit--) {
                    const uint32_t t = R | (1u << bit);
                    if (count_if([&](uint32_t kv, uint32_t rv) { return kv == K && rv < t; }) < need) R = t;
                }
            }
        }
        unsigned kept = 0;  // wave-uniform
#pragma unroll
        for (int sl = 0; sl < SL; sl++) {
            if ((unsigned)sl >= ns) continue;
            const bool keep = ok[sl] && (kk[sl] < K || (kk[sl] == K && rr[sl] <= R));
            const uint64_t mask = __ballot(keep);
            if (keep) {
                const unsigned at = kept + mbcnt(mask);
                sk[at] = kk[sl];
                sr[at] = rr[sl];
            }
            kept += (unsigned)__popcll(mask);
        }
        wave_sync();
        unsigned long long key = lane < kept ? (unsigned long long)sk[lane] : ~0ull;
        uint32_t row = lane < kept ? sr[lane] : 0xFFFFFFFFu;
        wave_sort64(key, row);
        if (lane < m) {
            rows[r * m + lane] = lane < kept ? perm[row] : 0xFFFFFFFFu;
            keys[r * m + lane] = lane < kept ? unorder_f32((uint32_t)key) : INFINITY;
        }
        const uint32_t mth = (uint32_t)__shfl(key, (int)m - 1);
        if (lane == 0) {
            count[r] = kept;
            thr[r] = total >= m ? unorder_f32(mth) : T;
        }
    }
}

// ----------------------------------------------------------- threshold estimate ----

constexpr int EST_NB = 4;                     // column blocks (16 queries each) per wavefront
constexpr int EST_MB = 2;                     // row blocks (16 points each) per super-block
constexpr int EST_SPB = 16 * EST_MB;          // points per super-block
constexpr int EST_QPW = 16 * EST_NB;          // queries per wavefront
constexpr int EST_QPB = 4 * EST_QPW;          // queries per block
constexpr int EST_SUBSETS = 4 * 4 * EST_MB;   // per query: 4 lanes x 4 EST_MB keys of a super-block

// A starting threshold for each query: the points (every pstride-th row of the scan
// order) fall into EST_SUBSETS disjoint subsets by their position within a
// super-block (the lane and register that hold the key), each lane keeps the running
// minimum of its keys per subset in registers, and the pick-th smallest of a query's
// subset minima is its threshold: pick subsets hold a point at or below it, so at
// least pick - 1 rows other than the query itself do. No appends and no
// compactions: the scan is the MFMA keys and one v_min a key. A starting threshold
// only steers how much the candidate scan keeps; the candidates do not depend on it.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) void estimate_bf16(
    const __bf16 *__restrict__ P, const __bf16 *__restrict__ Q, uint64_t np, uint32_t pstride,
    const uint32_t *__restrict__ qpos, uint64_t qoff, uint64_t nq, uint32_t pick,
    float *__restrict__ out_thr) {
    __shared__ __attribute__((aligned(16))) unsigned char tile[3][TILE * ROW_BYTES];
    const unsigned tid = threadIdx.x, lane = lane_id(), wave = tid / 64;
    const unsigned nblk = gridDim.x, xcd = blockIdx.x % 8, per = nblk / 8, extra = nblk % 8;
    const unsigned lblk = xcd * per + min(xcd, extra) + blockIdx.x / 8;
    const uint64_t qbase = (uint64_t)lblk * EST_QPB + (uint64_t)wave * EST_QPW;
    bf16x8 bq[EST_NB][2];
#pragma unroll
    for (int b = 0; b < EST_NB; b++) {
        const uint64_t ic = min<uint64_t>(qbase + 16 * b + (lane % 16), nq - 1);
        const uint32_t r = qpos ? qpos[ic] : (uint32_t)(qoff + ic);
        const bf16x8 *src = reinterpret_cast<const bf16x8 *>(Q + (uint64_t)r * KDIM);
#pragma unroll
        for (int ks = 0; ks < 2; ks++) bq[b][ks] = src[4 * ks + lane / 16];
    }
    float rmin[EST_NB][4 * EST_MB];
#pragma unroll
    for (int b = 0; b < EST_NB; b++)
#pragma unroll
        for (int e = 0; e < 4 * EST_MB; e++) rmin[b][e] = INFINITY;
    const uint64_t n_tiles = (np + TILE - 1) / TILE;
    const uint64_t q0 = min<uint64_t>((uint64_t)lblk * EST_QPB, nq - 1);
    const uint64_t own = ((qpos ? (uint64_t)qpos[q0] : qoff + q0) / pstride) / TILE;
    // Tiles as in search_bf16: global_load_lds into three buffers, 16-byte chunks
    // XOR-swizzled by row; rows past the last point read P's padding rows.
    auto fetch = [&](uint64_t t0, unsigned char *__restrict__ dst) {
        const __bf16 *tp = P + t0 * pstride * KDIM;
#pragma unroll
        for (int s = 0; s < TILE / 32; s++) {
            const unsigned row = (wave * (TILE / 32) + s) * 8 + lane / 8;
            const unsigned chunk = (lane % 8) ^ (row % 8);
            const __bf16 *src = tp + (row * pstride * KDIM + chunk * 8);
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)src,
                                             (__attribute__((address_space(3))) void *)(dst + (wave * (TILE / 32) + s) * 1024),
                                             16, 0, 0);
        }
    };
    auto wrap_next = [&](uint32_t t) { return t + 1 == (uint32_t)n_tiles ? 0u : t + 1; };
    uint32_t tcur = (uint32_t)min<uint64_t>(own, n_tiles - 1), tfetch = wrap_next(tcur);
    fetch((uint64_t)tcur * TILE, tile[0]);
    if (n_tiles > 1) fetch((uint64_t)tfetch * TILE, tile[1]);
    tfetch = wrap_next(tfetch);
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    __syncthreads();
    int cur = 0;
    for (uint64_t it = 0; it < n_tiles; it++) {
        const int nxt = cur == 2 ? 0 : cur + 1, after = nxt == 2 ? 0 : nxt + 1;
        auto body = [&](const unsigned char *__restrict__ tb, unsigned char *__restrict__ nx) {
#pragma nounroll
            for (int sb = 0; sb < TILE / EST_SPB; sb++) {
                bf16x8 a0[EST_MB], a1[EST_MB];
#pragma unroll
                for (int g = 0; g < EST_MB; g++) {
                    const unsigned prow = sb * EST_SPB + g * 16 + (lane % 16);
                    const unsigned char *rp = tb + prow * ROW_BYTES;
                    a0[g] = *reinterpret_cast<const bf16x8 *>(rp + (((lane / 16) ^ (prow % 8)) * 16));
                    a1[g] = *reinterpret_cast<const bf16x8 *>(rp + (((4 + lane / 16) ^ (prow % 8)) * 16));
                }
                f32x4 z[EST_MB][EST_NB];
#pragma unroll
                for (int g = 0; g < EST_MB; g++)
#pragma unroll
                    for (int b = 0; b < EST_NB; b++) {
                        const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
                        z[g][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0[g], bq[b][0], zero, 0, 0, 0);
                    }
#pragma unroll
                for (int g = 0; g < EST_MB; g++)
#pragma unroll
                    for (int b = 0; b < EST_NB; b++) {
                        const f32x4 r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1[g], bq[b][1], z[g][b], 0, 0, 0);
#pragma unroll
                        for (int j = 0; j < 4; j++) rmin[b][4 * g + j] = fminf(rmin[b][4 * g + j], r[j]);
                    }
            }
            if (it + 2 < n_tiles) {
                fetch((uint64_t)tfetch * TILE, nx);
                __builtin_amdgcn_s_waitcnt(0x0F74);  // vmcnt(4): all but tile t + 2
            } else {
                __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
            }
        };
        body(tile[cur], tile[after]);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_barrier();
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        cur = nxt;
        tcur = wrap_next(tcur);
        tfetch = wrap_next(tfetch);
    }
    __syncthreads();
    // The subset minima through LDS (over the tiles, no longer read): subset s of the
    // wave's query q at mins[s * (EST_QPW + 1) + q].
    static_assert(4 * EST_SUBSETS * (EST_QPW + 1) * sizeof(float) <= sizeof(tile), "subset minima fit the tiles");
    float *mins = reinterpret_cast<float *>(&tile[0][0]) + wave * EST_SUBSETS * (EST_QPW + 1);
#pragma unroll
    for (int b = 0; b < EST_NB; b++)
#pragma unroll
        for (int e = 0; e < 4 * EST_MB; e++) mins[((lane / 16) * 4 * EST_MB + e) * (EST_QPW + 1) + 16 * b + (lane % 16)] = rmin[b][e];
    wave_sync();
    // Lane q: the pick-th smallest of its query's minima (remove the smallest pick - 1
    // times).
    float v[EST_SUBSETS];
#pragma unroll
    for (int s = 0; s < EST_SUBSETS; s++) v[s] = mins[s * (EST_QPW + 1) + lane];
    float t = INFINITY;
    for (uint32_t r = 0; r < pick; r++) {
        t = INFINITY;
        int at = 0;
#pragma unroll
        for (int s = 0; s < EST_SUBSETS; s++)
            if (v[s] < t) {
                t = v[s];
                at = s;
            }
#pragma unroll
        for (int s = 0; s < EST_SUBSETS; s++)
            if (s == at) v[s] = INFINITY;
    }
    const uint64_t i = qbase + lane;
    if (i < nq) out_thr[i] = t;
}

// ================================================================== host side ====

unsigned bf16_resident(const Context &ctx) {
    int per_cu = 0;
    KY_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, search_bf16<false>, 256, 0));
    return (unsigned)std::max(per_cu, 1) * (unsigned)std::max(ctx.n_cu, 1);
}

// P' has PAD_ROWS padding rows: a scan reads whole tiles, up to row
// (ceil(n / stride) + TILE) * stride.
constexpr uint64_t PAD_ROWS = (uint64_t)(TILE + 1) * EST_STRIDE;

// The BF16 buffers over the points, kept for the whole search (the list mode's sorted
// query positions and rows among them).
template <class A>
void bf16_take_points(A &a, const KnnShape &s, Bf16Points &b) {
    b.perm = a.template take<uint32_t>(s.n);
    b.P = a.template take<__bf16>((s.n + PAD_ROWS) * KDIM);
    b.Q = a.template take<__bf16>(s.n * KDIM);
    b.err = a.template take<double>(2 * s.n);
    // The thresholds, then -inf for a tile's worth past the last (the symmetric scan
    // reads whole tiles of them).
    b.thr_est = s.n >= EST_MIN_ROWS ? a.template take<float>(s.nq + TILE) : nullptr;
    b.small = a.template take<unsigned long long>(4);
    if (s.list) {
        b.qpos = a.template take<uint32_t>(s.nq);
        b.qorig = a.template take<uint32_t>(s.nq);
    }
}
template void bf16_take_points<Arena>(Arena &, const KnnShape &, Bf16Points &);
template void bf16_take_points<SizeArena>(SizeArena &, const KnnShape &, Bf16Points &);

// The preparation's temporaries: Morton keys and values, the sort's scratch, |b| per
// row, the inverse order; for a list, the queries' keys and values.
template <class A>
static void bf16_take_temp(A &a, const KnnShape &s, uint32_t *&keys, uint32_t *&vals, uint32_t *&skeys, void *&tmp,
                           double *&nb, uint32_t *&inv, uint32_t *&k2, uint32_t *&v2) {
    keys = a.template take<uint32_t>(s.n);
    vals = a.template take<uint32_t>(s.n);
    skeys = a.template take<uint32_t>(s.n);
    tmp = a.template take<unsigned char>(sort_pairs_bytes(std::max(s.n, s.nq)));
    nb = a.template take<double>(s.n);
    inv = a.template take<uint32_t>(s.n);
    k2 = v2 = nullptr;
    if (s.list) {
        k2 = a.template take<uint32_t>(s.nq);
        v2 = a.template take<uint32_t>(s.nq);
    }
}

size_t bf16_prepare_bytes(const KnnShape &s) {
    SizeArena a;
    uint32_t *k, *v, *sk, *inv, *k2, *v2;
    void *tmp;
    double *nb;
    bf16_take_temp(a, s, k, v, sk, tmp, nb, inv, k2, v2);
    return a.high;
}

// The split of one launch's query blocks (see bf16_scan): the blocks that scan every
// point, and the tail's point segments.
static void bf16_split(const KnnShape &s, uint64_t count, unsigned &main_blocks, unsigned &n_seg,
                       unsigned &seg_tiles) {
    const unsigned blocks = div_up(count, QPB);
    const unsigned resident = std::max<unsigned>(s.resident, 1);
    const uint64_t n_tiles = div_up(s.n, TILE);
    main_blocks = blocks;
    n_seg = 1;
    seg_tiles = 0;
    // A rescan (a launch with a slot budget) of fewer query blocks than fill half the
    // device: every block is split the same way, within the budget.
    const unsigned rem = blocks > resident ? blocks % resident : s.slots && blocks < resident / 2 ? blocks : 0;
    if (rem != 0 && 2 * rem <= resident) {
        main_blocks = blocks - rem;
        n_seg = (unsigned)std::min<uint64_t>(std::min<unsigned>(MAX_TAIL_SEG, resident / rem), n_tiles);
        // Within the launch's query slots (each segment of a query holds a buffer).
        if (s.slots) n_seg = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(n_seg, s.slots / count));
        seg_tiles = div_up(n_tiles, n_seg);
        n_seg = div_up(n_tiles, seg_tiles);  // no empty segment
    }
}

// The symmetric scan (each pair of rows keyed once, for both rows) runs when a round
// holds every row, with starting thresholds; it needs a list of SYM_TCAP entries a row.
static bool bf16_symmetric(const KnnShape &s, uint64_t count) {
    return !s.list && count == s.n && s.n >= EST_MIN_ROWS;
}

template <class A>
void bf16_take_round(A &a, const KnnShape &s, uint64_t count, KnnRound &r) {
    r.sym = bf16_symmetric(s, count);
    if (r.sym) {
        r.main_blocks = div_up(count, QPB);
        r.n_seg = 1;
        r.seg_tiles = 0;
        r.buf = a.template take<Entry>(count * BCAP);
        r.tcount = a.template take<uint32_t>(count);
        r.tlist = a.template take<Entry>(count * SYM_TCAP);
        r.thr_max = a.template take<float>(div_up(count, TILE) * GMAX_PER_TILE);
        return;
    }
    bf16_split(s, count, r.main_blocks, r.n_seg, r.seg_tiles);
    const uint64_t main_q = std::min<uint64_t>(count, (uint64_t)r.main_blocks * QPB);
    const uint64_t tail_q = count - main_q;
    r.buf = a.template take<Entry>((main_q + tail_q * r.n_seg) * BCAP);
    r.seg_rows = a.template take<uint32_t>(tail_q * r.n_seg * s.m);
    r.seg_keys = a.template take<float>(tail_q * r.n_seg * s.m);
    r.seg_count = a.template take<uint32_t>(tail_q * r.n_seg);
    r.seg_thr = a.template take<float>(tail_q * r.n_seg);
}
template void bf16_take_round<Arena>(Arena &, const KnnShape &, uint64_t, KnnRound &);
template void bf16_take_round<SizeArena>(SizeArena &, const KnnShape &, uint64_t, KnnRound &);

// The scan order, the augmented rows, the certificate's per-row terms and the starting
// thresholds of every query (temporaries in `temp`, bf16_prepare_bytes of them).
void bf16_prepare(Context &ctx, const KnnShape &s, const float *pts, const KnnPrep &prep, const uint32_t *q_rows,
                  Bf16Points &b, void *temp, uint32_t k) {
    KY_CHECK(s.d + 6 <= KDIM, KY_E501, "neighbors: no BF16 kernel for more than 58 components");
    KY_CHECK(s.m >= 1 && s.m <= 64, KY_E215, "neighbors: candidates out of range");
    hipStream_t st = ctx.stream;
    const uint64_t n = s.n;
    Arena a(temp, bf16_prepare_bytes(s));
    uint32_t *keys, *vals, *skeys, *inv, *k2, *v2;
    void *sort_tmp;
    double *nb;
    bf16_take_temp(a, s, keys, vals, skeys, sort_tmp, nb, inv, k2, v2);
    // small: [0] err max, [1] |b| max (f64 bits), [2] the first three components'
    // absolute maxima (f32 bits, two a word).
    fill_words(st, b.small, 8, 0u);
    unsigned *absmax = reinterpret_cast<unsigned *>(b.small + 2);
    // The scan order: rows sorted by the Morton code of their first three components.
    const uint32_t c = std::min<uint32_t>(3, s.d);
    col_absmax<<<std::min<unsigned>(1024, div_up(n, 256)), 256, 0, st>>>(prep.cent, n, s.dp, c, absmax);
    morton_keys<<<div_up(n, 256), 256, 0, st>>>(prep.cent, n, s.dp, c, absmax, keys, vals);
    KY_HIP(hipGetLastError());
    sort_pairs(st, keys, skeys, vals, b.perm, n, 30, sort_tmp);
    __bf16 *P = static_cast<__bf16 *>(b.P), *Q = static_cast<__bf16 *>(b.Q);
    pad_rows<<<div_up(PAD_ROWS * KDIM, 256), 256, 0, st>>>(P, n, PAD_ROWS, s.d);
    invert_perm<<<div_up(n, 256), 256, 0, st>>>(b.perm, n, inv);
    bf16_rows<<<div_up(n, BROWS), 256, 0, st>>>(pts, prep.mean, prep.norms, prep.cent, inv, n, s.d, s.dp, P, Q,
                                               b.err, nb, b.small, b.small + 1);
    KY_HIP(hipGetLastError());
    unsigned long long hm[2];
    KY_HIP(hipMemcpyAsync(hm, b.small, sizeof(hm), hipMemcpyDeviceToHost, st));
    KY_HIP(hipStreamSynchronize(st));
    std::memcpy(&b.err_max, &hm[0], sizeof(double));
    double nb_max;
    std::memcpy(&nb_max, &hm[1], sizeof(double));
    // Accumulation bound of a K-term f32 dot product in any order, doubled:
    // 2 (K + 8) u over (|b_q| + |b_p|)^2 (the terms' absolute sum), plus an absolute
    // term for what the relative bound cannot see below f32's normal range: a product
    // or partial sum under FLT_MIN flushed or rounded to the subnormal grid (less than
    // FLT_MIN each, 2 K of them), the three bf16 parts of o_q and of o_p likewise, and
    // a subnormal component flushed (at most FLT_MIN max|b|, below the relative term
    // once max|b| > FLT_MIN / c, below FLT_MIN otherwise). A tiny embedding's keys are
    // then all within the bound and its rows go to the exact search.
    const double cb = 2.0 * (KDIM + 8.0) / 16777216.0;
    const double absolute = 4.0 * (KDIM + 8.0) * (double)FLT_MIN;
    bf16_delta<<<div_up(n, 256), 256, 0, st>>>(nb, n, nb_max, cb, absolute, b.err + n);
    KY_HIP(hipGetLastError());
    b.acc_rel = cb;
    // A list of queries: their scan positions, ascending, and their rows in that order.
    if (s.list) {
        query_positions<<<div_up(s.nq, 256), 256, 0, st>>>(q_rows, s.nq, inv, k2, v2);
        KY_HIP(hipGetLastError());
        sort_pairs(st, k2, b.qpos, v2, b.qorig, s.nq, 32, sort_tmp);
        // qorig holds list indices: make them rows.
        list_rows<<<div_up(s.nq, 256), 256, 0, st>>>(q_rows, s.nq, b.qorig);
        KY_HIP(hipGetLastError());
    }
    // Each query's starting threshold from the subset minima over every EST_STRIDE-th
    // row of the scan order (estimate_bf16); the scan starts every query at it, so its
    // buffer takes only the rows at or below it (any starting threshold gives the same
    // candidates; a low one only leaves fewer than m).
    if (b.thr_est) {
        const uint64_t np = div_up(n, EST_STRIDE);
        uint32_t pick = 1;
        while (pick < EST_SUBSETS - 1 &&
               -std::log1p(-(double)pick / EST_SUBSETS) * EST_SUBSETS * EST_STRIDE < 5.0 * (k + 1))
            pick++;
        estimate_bf16<<<div_up(s.nq, EST_QPB), 256, 0, st>>>(P, Q, np, EST_STRIDE, b.qpos, 0, s.nq, pick, b.thr_est);
        KY_HIP(hipGetLastError());
    }
    // The temporaries are released when this returns: finish with them first.
    KY_HIP(hipStreamSynchronize(st));
}

// One round: the candidates of the queries at sorted positions [lo, lo + count) (rows
// perm[lo + i], or the list's qorig[lo + i]) into the round's slots 0..count.
//
// The last round of blocks: when the query blocks fill the device's resident blocks r
// whole times and a remainder of at most half of them, the remainder would run almost
// alone at the end. Its query blocks are split instead over n_seg point segments
// (blocks of their own in the same launch, filling the free slots), and merge_tail
// joins their lists.
// The queries: at sorted positions qpos[0..count) (ascending), or at qoff + i when qpos
// is null; thr: their starting thresholds, or null for none.
static void bf16_scan(Context &ctx, const KnnShape &s, const Bf16Points &b, const uint32_t *qpos, uint64_t lo,
                      uint64_t count, const float *thr, const KnnRound &r, double *scan_ms) {
    hipStream_t st = ctx.stream;
    const unsigned blocks = div_up(count, QPB);
    const uint64_t main_q = std::min<uint64_t>(count, (uint64_t)r.main_blocks * QPB);
    const uint64_t tail_q = count - main_q;
    const unsigned grid = r.main_blocks + (blocks - r.main_blocks) * r.n_seg;
    hipEvent_t ev[2] = {ctx.timer[0], ctx.timer[1]};
    if (r.sym) {
        KY_CHECK(thr && lo == 0 && count == s.n, KY_E505, "neighbors: a symmetric scan of a partial round (internal)");
        fill_words(st, r.tcount, count, 0u);
        fill_words(st, reinterpret_cast<uint32_t *>(b.thr_est + count), TILE, 0xFF800000u);  // -inf
        const uint64_t n_g = div_up(count, TILE) * GMAX_PER_TILE;
        group_max<<<div_up(n_g, 256), 256, 0, st>>>(thr, n_g, r.thr_max);
        if (scan_ms) KY_HIP(hipEventRecord(ev[0], st));
        search_bf16<true><<<blocks, 256, 0, st>>>(
            static_cast<const __bf16 *>(b.P), static_cast<const __bf16 *>(b.Q), s.n, nullptr, 0, count, b.perm, s.m,
            thr, static_cast<Entry *>(r.buf), r.rows, r.keys, r.count, r.thr, blocks, 1, 0, nullptr, nullptr, nullptr,
            nullptr, r.tcount, static_cast<Entry *>(r.tlist), SYM_TCAP, r.thr_max);
        KY_HIP(hipGetLastError());
        if (scan_ms) KY_HIP(hipEventRecord(ev[1], st));
        KY_CHECK(s.m <= 64, KY_E505, "knn: more than 64 candidates a row in the symmetric scan (internal)");
        merge_sym<<<div_up(count, 4), 256, 0, st>>>(b.perm, count, s.m, r.tcount, static_cast<const Entry *>(r.tlist),
                                                     r.rows, r.keys, r.count, r.thr);
        KY_HIP(hipGetLastError());
    } else {
        if (scan_ms) KY_HIP(hipEventRecord(ev[0], st));
        search_bf16<false><<<grid, 256, 0, st>>>(
            static_cast<const __bf16 *>(b.P), static_cast<const __bf16 *>(b.Q), s.n, qpos, lo, count, b.perm, s.m,
            thr, static_cast<Entry *>(r.buf), r.rows, r.keys, r.count, r.thr, r.main_blocks, r.n_seg, r.seg_tiles,
            r.seg_rows, r.seg_keys, r.seg_count, r.seg_thr, nullptr, nullptr, 0, nullptr);
        KY_HIP(hipGetLastError());
        if (scan_ms) KY_HIP(hipEventRecord(ev[1], st));
    }
    if (tail_q) {
        merge_tail<<<(unsigned)tail_q, 64, 0, st>>>(r.seg_rows, r.seg_keys, r.seg_count, r.seg_thr, tail_q, r.n_seg,
                                                    s.m, main_q, r.rows, r.keys, r.count, r.thr);
        KY_HIP(hipGetLastError());
    }
    // A query whose starting threshold left fewer than k rows at or below it ends with
    // fewer than k candidates: the re-rank leaves it uncertified, and the search takes it
    // again (bf16_rescan, then the exact search; knn.hip).
    if (scan_ms) {
        KY_HIP(hipEventSynchronize(ev[1]));
        float ms = 0.0f;
        KY_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
        *scan_ms += ms;
    }
}

void bf16_round(Context &ctx, const KnnShape &s, const Bf16Points &b, uint64_t lo, uint64_t count,
                const KnnRound &r, double *scan_ms) {
    bf16_scan(ctx, s, b, s.list ? b.qpos + lo : nullptr, lo, count, b.thr_est ? b.thr_est + lo : nullptr, r, scan_ms);
}

// ------------------------------------------------------------------- rescan ----
// The short rows' starting thresholds times this: 1.15 at 1M x 50 (k = 14) rescans the
// 3,858 short rows in 1.7 ms with none left short; 1.10 1.4 ms but 3 rows left for the
// exact search, 1.25 2.9 ms, 1.35 4.1 ms (the keys kept grow fast with the threshold).
constexpr float RESCAN_FACTOR = 1.15f;
// The rows a round's re-rank left with fewer than m candidates (a starting threshold
// below their m-th key, or the symmetric scan's list of the row overflowed) are
// scanned again as a list, from RESCAN_FACTOR times the first round's starting
// threshold (none below 2^17 rows): each keeps its m smallest pairs at or below it;
// a row still short goes to the exact search.

template <class A>
void bf16_take_rescan(A &a, const KnnShape &s, uint64_t count, Bf16Rescan &x, KnnRound &r) {
    KnnShape ls = s;
    ls.list = true;
    ls.nq = count;
    ls.round = count;
    ls.slots = std::min(s.round, s.nq);  // the first rounds' buffers bound the rescan's
    x.inv = a.template take<uint32_t>(s.n);
    x.k2 = a.template take<uint32_t>(count);
    x.v2 = a.template take<uint32_t>(count);
    x.qpos = a.template take<uint32_t>(count);
    x.qorig = a.template take<uint32_t>(count);
    x.tmp = a.template take<unsigned char>(sort_pairs_bytes(count));
    x.thr = s.n >= EST_MIN_ROWS ? a.template take<float>(count) : nullptr;
    r.rows = a.template take<uint32_t>(count * s.m);
    r.keys = a.template take<float>(count * s.m);
    r.count = a.template take<uint32_t>(count);
    r.thr = a.template take<float>(count);
    bf16_take_round(a, ls, count, r);
}
template void bf16_take_rescan<Arena>(Arena &, const KnnShape &, uint64_t, Bf16Rescan &, KnnRound &);
template void bf16_take_rescan<SizeArena>(SizeArena &, const KnnShape &, uint64_t, Bf16Rescan &, KnnRound &);

// The largest of the rescan's bytes for a round of `chunk` rows and for the largest
// round that is split over point segments (whose buffers hold a query per segment).
size_t bf16_rescan_bytes(const KnnShape &s, uint64_t chunk) {
    size_t worst = 0;
    const uint64_t split = std::min<uint64_t>(chunk, (uint64_t)(std::max<uint32_t>(s.resident, 1) / 2) * QPB);
    for (const uint64_t count : {chunk, split}) {
        if (count == 0) continue;
        SizeArena a;
        Bf16Rescan x;
        KnnRound r;
        bf16_take_rescan(a, s, count, x, r);
        worst = std::max(worst, a.high);
    }
    return worst;
}

void bf16_rescan(Context &ctx, const KnnShape &s, const Bf16Points &b, const uint32_t *rows, uint64_t count,
                 const Bf16Rescan &x, const KnnRound &r) {
    hipStream_t st = ctx.stream;
    KnnShape ls = s;
    ls.list = true;
    ls.nq = count;
    ls.round = count;
    ls.slots = std::min(s.round, s.nq);
    // The rows' scan positions, ascending, and the rows in that order.
    invert_perm<<<div_up(s.n, 256), 256, 0, st>>>(b.perm, s.n, x.inv);
    query_positions<<<div_up(count, 256), 256, 0, st>>>(rows, count, x.inv, x.k2, x.v2);
    KY_HIP(hipGetLastError());
    sort_pairs(st, x.k2, x.qpos, x.v2, x.qorig, count, 32, x.tmp);
    list_rows<<<div_up(count, 256), 256, 0, st>>>(rows, count, x.qorig);
    KY_HIP(hipGetLastError());
    // RESCAN_FACTOR times the first round's starting threshold (a key): a row its list
    // could not hold keeps everything the first round would have, one whose threshold
    // left fewer than m rows gains the rows up to the larger key (most reach m; the
    // others go on to the exact search).
    if (x.thr && b.thr_est) {
        rescan_thresholds<<<div_up(count, 256), 256, 0, st>>>(x.qpos, count, b.thr_est, RESCAN_FACTOR, x.thr);
        KY_HIP(hipGetLastError());
    }
    bf16_scan(ctx, ls, b, x.qpos, 0, count, x.thr && b.thr_est ? x.thr : nullptr, r, nullptr);
}

}  // namespace ky
