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
// Block: 4 wavefronts, each owning 16*NB queries (NB MFMA column blocks, lane = query
// l % 16 of its block); the block streams the points in tiles of TILE rows through
// LDS (padded rows, conflict-free 16-byte fragment reads). A query keeps every point
// whose key is at most its threshold in a buffer in device memory (CAP entries); a
// wave-level bitonic sort in LDS compacts a buffer back to its m smallest (key, row)
// pairs when it could overflow in the next tile, and sets the threshold to the m-th.
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

__global__ void bf16_rows(const float *__restrict__ x, const double *__restrict__ mean,
                          const float *__restrict__ cent, uint64_t n, uint32_t d, uint32_t dp,
                          __bf16 *__restrict__ P, __bf16 *__restrict__ Q, double *__restrict__ err,
                          double *__restrict__ nb, unsigned long long *__restrict__ err_max,
                          unsigned long long *__restrict__ nb_max) {
    uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (r >= n) return;
    double o = 0.0, e2 = 0.0;
    __bf16 *pr = P + r * KDIM, *qr = Q + r * KDIM;
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

__global__ void bf16_delta(const double *__restrict__ nb, uint64_t n, double nb_max, double c,
                           double *__restrict__ delta) {
    uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (r >= n) return;
    double t = nb[r] + nb_max;
    delta[r] = c * t * t;
}

// ------------------------------------------------------------------- search ----

// Compact query q's buffer (cnt entries in device memory) to its `keep` smallest
// (key, row) pairs, sorted, through the wave's LDS scratch; returns the new count and
// sets *thr to the keep-th key when cnt >= keep. Whole wavefront.
__device__ unsigned compact_buffer(float *__restrict__ bk, uint32_t *__restrict__ br, unsigned cnt, unsigned keep,
                                   uint32_t *skey, uint32_t *srow, float *thr) {
    const unsigned lane = lane_id();
    wave_sync_memory();
    for (unsigned i = lane; i < BCAP; i += 64) {
        if (i < cnt) {
            skey[i] = order_f32(bk[i]);
            srow[i] = br[i];
        } else {
            skey[i] = 0xFFFFFFFFu;
            srow[i] = 0xFFFFFFFFu;
        }
    }
    wave_sync();
    wave_bitonic<uint32_t>(skey, srow, BCAP);
    const unsigned kept = min(cnt, keep);
    for (unsigned i = lane; i < kept; i += 64) {
        bk[i] = unorder_f32(skey[i]);
        br[i] = srow[i];
    }
    if (cnt >= keep) *thr = unorder_f32(skey[keep - 1]);
    wave_sync_memory();
    return kept;
}

__global__ __launch_bounds__(256) void search_bf16(const __bf16 *__restrict__ P, const __bf16 *__restrict__ Q,
                                                   uint64_t n, uint32_t m, float *__restrict__ buf_key,
                                                   uint32_t *__restrict__ buf_row, uint32_t *__restrict__ out_rows,
                                                   float *__restrict__ out_keys, uint32_t *__restrict__ out_count) {
    __shared__ __attribute__((aligned(16))) unsigned char tile[TILE * ROW_BYTES];
    __shared__ uint32_t scratch_key[4][BCAP];
    __shared__ uint32_t scratch_row[4][BCAP];
    __shared__ unsigned qcnt[QPB];
    __shared__ float qthr[QPB];
    const unsigned tid = threadIdx.x, lane = lane_id(), wave = tid / 64;
    const uint64_t qbase = (uint64_t)blockIdx.x * QPB + (uint64_t)wave * QPW;  // this wave's first query
    unsigned *cnt = qcnt + wave * QPW;
    float *thr_lds = qthr + wave * QPW;
    for (unsigned i = lane; i < QPW; i += 64) {
        cnt[i] = 0;
        thr_lds[i] = INFINITY;
    }
    // B fragments: the wave's queries, loaded once.
    bf16x8 bq[NB][2];
    float thr[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) {
        uint64_t q = qbase + 16 * b + (lane % 16);
        const bf16x8 *src = reinterpret_cast<const bf16x8 *>(Q + min(q, n - 1) * KDIM);
#pragma unroll
        for (int ks = 0; ks < 2; ks++) bq[b][ks] = src[4 * ks + lane / 16];
        thr[b] = INFINITY;
    }
    const bool wave_live = qbase < n;
    for (uint64_t t0 = 0; t0 < n; t0 += TILE) {
        __syncthreads();
        // Stage the tile: 128 rows x 128 bytes, 16 bytes a thread per step.
        for (unsigned i = tid; i < TILE * 8; i += 256) {
            unsigned row = i / 8, chunk = i % 8;
            uint64_t p = t0 + row;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (p < n) v = reinterpret_cast<const uint4 *>(P + p * KDIM)[chunk];
            *reinterpret_cast<uint4 *>(tile + row * ROW_BYTES + chunk * 16) = v;
        }
        __syncthreads();
        if (!wave_live) continue;
#pragma unroll 2
        for (int mb = 0; mb < TILE / 16; mb++) {
            const unsigned prow = mb * 16 + (lane % 16);
            bf16x8 a0 = *reinterpret_cast<const bf16x8 *>(tile + prow * ROW_BYTES + (lane / 16) * 16);
            bf16x8 a1 = *reinterpret_cast<const bf16x8 *>(tile + prow * ROW_BYTES + (4 + lane / 16) * 16);
            f32x4 acc[NB];
#pragma unroll
            for (int b = 0; b < NB; b++) {
                f32x4 z = {0.0f, 0.0f, 0.0f, 0.0f};
                z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, bq[b][0], z, 0, 0, 0);
                acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, bq[b][1], z, 0, 0, 0);
            }
            // Lane l, slot j: point t0 + 16 mb + 4 (l / 16) + j, query 16 b + l % 16.
            const uint64_t p0 = t0 + mb * 16 + 4 * (lane / 16);
#pragma unroll
            for (int b = 0; b < NB; b++) {
                float lo = fminf(fminf(acc[b][0], acc[b][1]), fminf(acc[b][2], acc[b][3]));
                bool any = lo <= thr[b];
                if (__ballot(any)) {
                    const unsigned ql = 16 * b + (lane % 16);
                    const uint64_t q = qbase + ql;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        uint64_t p = p0 + j;
                        if (acc[b][j] <= thr[b] && p < n && p != q && q < n) {
                            unsigned slot = atomicAdd(&cnt[ql], 1u);
                            buf_key[q * BCAP + slot] = acc[b][j];
                            buf_row[q * BCAP + slot] = (uint32_t)p;
                        }
                    }
                }
            }
        }
        // Compact the buffers that the next tile could overflow.
        wave_sync();
        uint64_t need = 0;
        for (unsigned i = lane; i < QPW; i += 64)
            if (cnt[i] > BCAP - TILE) need |= 1ull << i;
        // QPW <= 64: one bit per query lane.
        uint64_t mask = __ballot(need != 0);
        while (mask) {
            unsigned ql = __builtin_ctzll(mask);
            mask &= mask - 1;
            const uint64_t q = qbase + ql;
            unsigned c = cnt[ql];
            float t = thr_lds[ql];
            unsigned kept = compact_buffer(buf_key + q * BCAP, buf_row + q * BCAP, c, m, scratch_key[wave],
                                           scratch_row[wave], &t);
            if (lane == 0) {
                cnt[ql] = kept;
                thr_lds[ql] = t;
            }
            wave_sync();
        }
#pragma unroll
        for (int b = 0; b < NB; b++) thr[b] = thr_lds[16 * b + (lane % 16)];
    }
    if (!wave_live) return;
    for (unsigned ql = 0; ql < QPW; ql++) {
        const uint64_t q = qbase + ql;
        if (q >= n) break;
        float t = thr_lds[ql];
        unsigned kept = compact_buffer(buf_key + q * BCAP, buf_row + q * BCAP, cnt[ql], m, scratch_key[wave],
                                       scratch_row[wave], &t);
        for (unsigned e = lane; e < m; e += 64) {
            out_rows[q * m + e] = e < kept ? scratch_row[wave][e] : 0xFFFFFFFFu;
            out_keys[q * m + e] = e < kept ? unorder_f32(scratch_key[wave][e]) : INFINITY;
        }
        if (lane == 0) out_count[q] = kept;
        wave_sync();
    }
}

// ================================================================== host side ====

void bf16_candidates(Context &ctx, const float *pts, const KnnPrep &prep, uint32_t m, uint32_t *d_rows,
                     float *d_keys, uint32_t *d_count, DevBuf<double> &err, double &err_max, double &acc_rel) {
    KY_CHECK(prep.d + 6 <= KDIM, KY_E501, "neighbors: no BF16 kernel for more than 58 components");
    KY_CHECK(m >= 1 && m <= 64, KY_E215, "neighbors: candidates out of range");
    hipStream_t st = ctx.stream;
    const uint64_t n = prep.n;
    DevBuf<__bf16> P(n * KDIM), Q(n * KDIM);
    err.alloc(2 * n);  // err, then delta
    DevBuf<double> nb(n);
    DevBuf<unsigned long long> maxima(2);
    KY_HIP(hipMemsetAsync(maxima.ptr, 0, 2 * sizeof(unsigned long long), st));
    bf16_rows<<<div_up(n, 256), 256, 0, st>>>(pts, prep.mean.ptr, prep.cent.ptr, n, prep.d, prep.dp, P.ptr, Q.ptr,
                                               err.ptr, nb.ptr, maxima.ptr, maxima.ptr + 1);
    KY_HIP(hipGetLastError());
    unsigned long long hm[2];
    KY_HIP(hipMemcpyAsync(hm, maxima.ptr, sizeof(hm), hipMemcpyDeviceToHost, st));
    KY_HIP(hipStreamSynchronize(st));
    std::memcpy(&err_max, &hm[0], sizeof(double));
    double nb_max;
    std::memcpy(&nb_max, &hm[1], sizeof(double));
    // Accumulation bound of a K-term f32 dot product in any order, doubled:
    // 2 (K + 8) u over (|b_q| + |b_p|)^2 (the terms' absolute sum).
    const double c = 2.0 * (KDIM + 8.0) / 16777216.0;
    bf16_delta<<<div_up(n, 256), 256, 0, st>>>(nb.ptr, n, nb_max, c, err.ptr + n);
    KY_HIP(hipGetLastError());
    acc_rel = c;
    DevBuf<float> bk(n * BCAP);
    DevBuf<uint32_t> br(n * BCAP);
    const unsigned blocks = div_up(n, QPB);
    search_bf16<<<blocks, 256, 0, st>>>(P.ptr, Q.ptr, n, m, bk.ptr, br.ptr, d_rows, d_keys, d_count);
    KY_HIP(hipGetLastError());
    KY_HIP(hipStreamSynchronize(st));
}

}  // namespace ky
