// connectivities() on the device: scanpy's UMAP fuzzy weights on a neighbour graph
// (crates/kyto-exec/src/connectivities.rs:76-123), one symmetric weighted CSR out.
//
// 1. fuzzy_scales, a thread a row: the row's distances ranked (a sorting network in
//    registers over f32 total-order keys), rho (the first distance above 0), sigma by
//    umap's 64-step f64 bisection (:125-168), the row-mean floor, and, for rows with
//    rho > 0, the memberships (:179-188). Rows with rho = 0 wait for the global floor.
// 2. Only when some row has rho = 0: the mean of every ranked distance, a float32
//    sum in row order (:170-177), on the host (one sequential pass), then
//    zero_rho_rows applies the floor and weighs those rows.
// 3. The transpose: a stable radix sort of the entries by column (rows then increase
//    within a column, as the reference's counting transpose leaves them), column
//    offsets from the sorted columns.
// 4. The fuzzy union w + w^T - w o w^T in float32, merged by column, zeros dropped
//    (:190-256): union_rows counts each row's edges, a scan gives the offsets, and
//    union_rows again writes them.
//
// Arithmetic kept as the reference keeps it: f32 differences and quotients, f64 exp
// and sums in the bisection, f32 membership exp (the correctly rounded value:
// exp in f64, rounded once), no contraction (-ffp-contract=off).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "connectivities.hpp"
#include "device.hpp"
#include "sort.hpp"

namespace ky {

constexpr int SMOOTH_K_ITERATIONS = 64;      // umap's n_iter (connectivities.rs:26)
constexpr double SMOOTH_K_TOLERANCE = 1e-5;  // :28
constexpr double MIN_K_DIST_SCALE = 1e-3;    // :30

// Rust's total_cmp order of f32 values as an unsigned key, and back.
static __device__ __forceinline__ uint32_t total_key(float v) {
    const uint32_t b = __float_as_uint(v);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
static __device__ __forceinline__ float from_total_key(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

// f32 exp, correctly rounded: exp in f64 rounded once to f32.
static __device__ __forceinline__ float exp32(float x) { return (float)exp((double)x); }

// compute_membership_strengths for one entry (connectivities.rs:179-188).
static __device__ __forceinline__ float membership(uint64_t row, uint32_t column, float distance, float sigma, float rho) {
    if ((uint64_t)column == row) return 0.0f;
    const float diff = distance - rho;
    if (diff <= 0.0f || sigma == 0.0f) return 1.0f;
    return exp32(-(diff / sigma));
}

// flags bit 0: an invalid row (columns not increasing below n, or a distance that is
// not finite and >= 0); zero_rows counts the rows with rho = 0.
template <int KMAX>
__global__ __launch_bounds__(256) void fuzzy_scales(const uint32_t *__restrict__ indices,
                                                    const float *__restrict__ dist, uint64_t n, uint32_t k,
                                                    uint32_t n_neighbors, double target, float *__restrict__ ranked,
                                                    float *__restrict__ sigma_out, float *__restrict__ rho_out,
                                                    float *__restrict__ w, unsigned *__restrict__ flags,
                                                    unsigned long long *__restrict__ zero_rows) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const uint64_t base = r * k;
    uint32_t key[KMAX];
    bool bad = false;
    uint32_t prev = 0;
#pragma unroll
    for (int j = 0; j < KMAX; j++) {
        if (j < (int)k) {
            const float d = dist[base + j];
            const uint32_t c = indices[base + j];
            bad = bad || !(isfinite(d) && d >= 0.0f) || (uint64_t)c >= n || (j > 0 && c <= prev);
            prev = c;
            key[j] = total_key(d);
        } else {
            key[j] = 0xFFFFFFFFu;  // after every distance
        }
    }
    // Ascending bitonic network over the registers (static indices).
#pragma unroll
    for (int size = 2; size <= KMAX; size <<= 1)
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1)
#pragma unroll
            for (int i = 0; i < KMAX; i++) {
                const int j = i ^ stride;
                if (j > i) {
                    const bool asc = (i & size) == 0;
                    const uint32_t a = key[i], b = key[j];
                    if ((a > b) == asc) {
                        key[i] = b;
                        key[j] = a;
                    }
                }
            }
    float rk[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; j++) rk[j] = from_total_key(key[j]);
    // local_connectivity 1: the first distance above 0 (connectivities.rs:128).
    float rho = 0.0f;
#pragma unroll
    for (int j = KMAX - 1; j >= 0; j--)
        if (j < (int)k && rk[j] > 0.0f) rho = rk[j];
    // smooth_knn_dist's bisection in f64 (:129-156).
    const double ceiling = (double)FLT_MAX;
    double lo = 0.0, hi = ceiling, mid = 1.0;
    for (int it = 0; it < SMOOTH_K_ITERATIONS; it++) {
        double sum = 0.0;
#pragma unroll
        for (int j = 0; j < KMAX; j++) {
            if (j < (int)k) {
                const double d = (double)(rk[j] - rho);  // a f32 difference
                sum += d > 0.0 ? exp(-(d / mid)) : 1.0;
            }
        }
        if (fabs(sum - target) < SMOOTH_K_TOLERANCE) break;
        if (sum > target) {
            hi = mid;
            mid = (lo + hi) / 2.0;
        } else {
            lo = mid;
            mid = hi >= ceiling ? mid * 2.0 : (lo + hi) / 2.0;
        }
    }
    float sigma = (float)mid;
    if (rho > 0.0f) {
        // The row's mean as umap's numpy gives it: a f32 sum in ranked order, divided in
        // f64 by n_neighbors (the cell's own 0 counted) (:170-177).
        float s = 0.0f;
#pragma unroll
        for (int j = 0; j < KMAX; j++)
            if (j < (int)k) s += rk[j];
        const double floor = MIN_K_DIST_SCALE * ((double)s / (double)n_neighbors);
        if ((double)sigma < floor) sigma = (float)floor;
    }
#pragma unroll
    for (int j = 0; j < KMAX; j++)
        if (j < (int)k) ranked[base + j] = rk[j];
    sigma_out[r] = sigma;
    rho_out[r] = rho;
    if (rho > 0.0f) {
        for (uint32_t j = 0; j < k; j++) w[base + j] = membership(r, indices[base + j], dist[base + j], sigma, rho);
    } else {
        atomicAdd(zero_rows, 1ull);
    }
    if (bad) atomicOr(flags, 1u);
}

// Rows with rho = 0 (every distance 0): sigma raised to the global floor, then their
// memberships (connectivities.rs:103-110).
__global__ void zero_rho_rows(const uint32_t *__restrict__ indices, const float *__restrict__ dist, uint64_t n,
                              uint32_t k, double floor, float *__restrict__ sigma, const float *__restrict__ rho,
                              float *__restrict__ w) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n || rho[r] != 0.0f) return;
    float s = sigma[r];
    if ((double)s < floor) s = (float)floor;
    sigma[r] = s;
    const uint64_t base = r * k;
    for (uint32_t j = 0; j < k; j++) w[base + j] = membership(r, indices[base + j], dist[base + j], s, 0.0f);
}

// Column offsets of the entries sorted by column: tptr[c] is the first sorted position
// of column c (tptr[n] = count). Position i writes the columns in (col[i-1], col[i]].
__global__ void column_offsets(const uint32_t *__restrict__ cols, uint64_t count, uint64_t n,
                               uint64_t *__restrict__ tptr) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > count) return;
    const uint64_t c0 = i == 0 ? 0 : (uint64_t)cols[i - 1] + 1;
    const uint64_t c1 = i == count ? n : (uint64_t)cols[i];
    for (uint64_t c = c0; c <= c1; c++) tptr[c] = i;
}

// The fuzzy union of row r: its own entries (columns increasing) merged with the
// entries naming it (sorted entry ids, rows increasing), w + w^T - w o w^T in f32
// (scipy's order, (w + w^T) - w o w^T), zeros dropped (connectivities.rs:190-256).
// count_only: edges per row into count; else the edges written from indptr[r].
__global__ void union_rows(const uint32_t *__restrict__ indices, const float *__restrict__ w, uint64_t n,
                           uint32_t k, const uint64_t *__restrict__ tptr, const uint32_t *__restrict__ tentry,
                           int count_only, uint32_t *__restrict__ count, const uint64_t *__restrict__ indptr,
                           uint32_t *__restrict__ out_cols, float *__restrict__ out_w) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    uint64_t f = r * k;
    const uint64_t fe = f + k;
    uint64_t b = tptr[r];
    const uint64_t be = tptr[r + 1];
    uint64_t at = count_only ? 0 : indptr[r];
    uint32_t edges = 0;
    while (f < fe || b < be) {
        const uint32_t fc = f < fe ? indices[f] : 0xFFFFFFFFu;
        const uint32_t bentry = b < be ? tentry[b] : 0u;
        const uint32_t bc = b < be ? (uint32_t)(bentry / k) : 0xFFFFFFFFu;
        uint32_t col;
        float a = 0.0f, t = 0.0f;
        if (b >= be || (f < fe && fc < bc)) {
            col = fc;
            a = w[f++];
        } else if (f >= fe || bc < fc) {
            col = bc;
            t = w[bentry];
            b++;
        } else {
            col = fc;
            a = w[f++];
            t = w[bentry];
            b++;
        }
        const float sum = a + t;
        const float prod = a * t;
        const float weight = sum - prod;
        if (weight != 0.0f) {
            if (!count_only) {
                out_cols[at] = col;
                out_w[at] = weight;
                at++;
            }
            edges++;
        }
    }
    if (count_only) count[r] = edges;
}

template <int KMAX>
static void launch_scales(hipStream_t st, const uint32_t *idx, const float *dist, uint64_t n, uint32_t k, uint32_t nn,
                   double target, float *ranked, float *sigma, float *rho, float *w, unsigned *flags,
                   unsigned long long *zero_rows) {
    fuzzy_scales<KMAX><<<div_up(n, 256), 256, 0, st>>>(idx, dist, n, k, nn, target, ranked, sigma, rho, w, flags,
                                                       zero_rows);
}

void connectivities_run(Context &ctx, uint64_t n, uint32_t n_neighbors, const uint32_t *h_indices,
                        const float *h_dist, Graph &out, float *h_sigma, float *h_rho, ky_conn_info *info) {
    const double t0 = now_ms();
    KY_CHECK(n_neighbors >= 2, KY_E215, "connectivities: n_neighbors " + std::to_string(n_neighbors) +
                                            " counts the cell itself (an int of at least 2)");
    KY_CHECK(n >= 1, KY_E215, "connectivities: a graph of at least one row");
    KY_CHECK(n - 1 <= 0xFFFFFFFFull, KY_E215, "connectivities: graph indices are uint32 (at most 2^32 rows)");
    KY_CHECK(n_neighbors <= 65, KY_E501, "connectivities: no device kernel for n_neighbors above 65");
    const uint32_t k = n_neighbors - 1;
    const uint64_t ne = n * k;
    KY_CHECK(ne <= 0xFFFFFFFFull, KY_E501, "connectivities: at most 2^32 entries");
    hipStream_t st = ctx.stream;
    DevBuf<uint32_t> idx(ne);
    DevBuf<float> dist(ne);
    upload_parts(ctx, {{idx.ptr, h_indices, sizeof(uint32_t) * ne}, {dist.ptr, h_dist, sizeof(float) * ne}});
    const double t_up = now_ms();

    DevBuf<float> ranked(ne), w(ne), sigma(n), rho(n);
    DevBuf<unsigned> flags(1);
    DevBuf<unsigned long long> zero_rows(1);
    KY_HIP(hipMemsetAsync(flags.ptr, 0, sizeof(unsigned), st));
    KY_HIP(hipMemsetAsync(zero_rows.ptr, 0, sizeof(unsigned long long), st));
    const double target = std::log2((double)n_neighbors);
    if (k <= 16)
        launch_scales<16>(st, idx.ptr, dist.ptr, n, k, n_neighbors, target, ranked.ptr, sigma.ptr, rho.ptr, w.ptr,
                          flags.ptr, zero_rows.ptr);
    else if (k <= 32)
        launch_scales<32>(st, idx.ptr, dist.ptr, n, k, n_neighbors, target, ranked.ptr, sigma.ptr, rho.ptr, w.ptr,
                          flags.ptr, zero_rows.ptr);
    else
        launch_scales<64>(st, idx.ptr, dist.ptr, n, k, n_neighbors, target, ranked.ptr, sigma.ptr, rho.ptr, w.ptr,
                          flags.ptr, zero_rows.ptr);
    KY_HIP(hipGetLastError());
    unsigned hflags = 0;
    unsigned long long hzero = 0;
    KY_HIP(hipMemcpyAsync(&hflags, flags.ptr, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    KY_HIP(hipMemcpyAsync(&hzero, zero_rows.ptr, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    KY_HIP(hipStreamSynchronize(st));
    KY_CHECK(!(hflags & 1u), KY_E215,
             "connectivities: a row whose columns do not increase below n_cells, or a distance that is not "
             "finite and at least 0 (NeighborGraph::from_csr)");
    if (hzero) {
        // The global floor: umap's mean of every ranked distance, a float32 sum in row
        // order over n * k values divided in f64 by n * n_neighbors (one host pass: the
        // sum is sequential by definition).
        std::vector<float> hr(ne);
        download(ctx, hr.data(), ranked.ptr, sizeof(float) * ne);
        float s = 0.0f;
        for (uint64_t i = 0; i < ne; i++) s += hr[i];
        const double floor = MIN_K_DIST_SCALE * ((double)s / ((double)n * (double)n_neighbors));
        zero_rho_rows<<<div_up(n, 256), 256, 0, st>>>(idx.ptr, dist.ptr, n, k, floor, sigma.ptr, rho.ptr, w.ptr);
        KY_HIP(hipGetLastError());
    }
    KY_HIP(hipStreamSynchronize(st));
    const double t_scales = now_ms();

    // The transpose: entries sorted by column, stable (rows increase within a column).
    DevBuf<uint32_t> ent(ne), scol(ne), sent(ne);
    DevBuf<unsigned char> sort_tmp;
    iota_words(st, ent.ptr, ne, 0u);
    KY_HIP(hipGetLastError());
    unsigned bits = 1;
    while (bits < 32 && (1ull << bits) < n) bits++;
    sort_pairs(st, idx.ptr, scol.ptr, ent.ptr, sent.ptr, ne, bits, sort_tmp);
    DevBuf<uint64_t> tptr(n + 1);
    column_offsets<<<div_up(ne + 1, 256), 256, 0, st>>>(scol.ptr, ne, n, tptr.ptr);
    KY_HIP(hipGetLastError());
    DevBuf<uint32_t> count(n + 1);
    union_rows<<<div_up(n, 256), 256, 0, st>>>(idx.ptr, w.ptr, n, k, tptr.ptr, sent.ptr, 1, count.ptr, nullptr,
                                               nullptr, nullptr);
    KY_HIP(hipGetLastError());
    out.ctx = &ctx;
    out.n = n;
    out.indptr.alloc(n + 1);
    DevBuf<unsigned char> scan_tmp;
    scan_counts(st, count.ptr, n, out.indptr.ptr, scan_tmp);
    uint64_t nnz = 0;
    KY_HIP(hipMemcpyAsync(&nnz, out.indptr.ptr + n, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    KY_HIP(hipStreamSynchronize(st));
    out.nnz = nnz;
    out.indices.alloc(nnz);
    out.weights.alloc(nnz);
    union_rows<<<div_up(n, 256), 256, 0, st>>>(idx.ptr, w.ptr, n, k, tptr.ptr, sent.ptr, 0, nullptr,
                                               out.indptr.ptr, out.indices.ptr, out.weights.ptr);
    KY_HIP(hipGetLastError());
    if (h_sigma) download(ctx, h_sigma, sigma.ptr, sizeof(float) * n);
    if (h_rho) download(ctx, h_rho, rho.ptr, sizeof(float) * n);
    KY_HIP(hipStreamSynchronize(st));
    const double t_end = now_ms();
    if (info) {
        info->edges = nnz;
        info->zero_rho_rows = hzero;
        info->upload_ms = t_up - t0;
        info->scales_ms = t_scales - t_up;
        info->union_ms = t_end - t_scales;
        info->total_ms = t_end - t0;
    }
}

size_t connectivities_device_bytes(uint64_t n, uint32_t n_neighbors) {
    auto b = [](size_t bytes) { return Arena::piece(std::max<size_t>(bytes, 1)); };
    const uint64_t ne = n * (uint64_t)(n_neighbors - 1);
    // The graph (indices, distances), ranked distances and memberships, sigma and rho;
    // the entries sorted by column (entries, keys, values, the sort's scratch); the
    // transposed offsets and counts; the union (indptr, at most 2 ne indices and weights).
    return 4 * b(4 * ne) + 2 * b(4 * n) + 3 * b(4 * ne) + b(sort_pairs_bytes(ne)) + b(8 * (n + 1)) +
           b(4 * (n + 1)) + b(scan_counts_bytes(n)) + b(8 * (n + 1)) + 2 * b(4 * 2 * ne);
}

}  // namespace ky
