// The FP64 products of the PCA's dense steps on the device: the Krylov basis B and
// its products C = X^T (X B) stay resident (genes x columns), and the steps' large
// products run here — CGS2's B^T R and R - B (B^T R) (pca_dense.rs:111-117), the
// Rayleigh-Ritz matrix B^T C and the loadings B U (pca_dense.rs:141-160) — while the
// small factorisations (Cholesky QR, the SVD drop rule, the symmetric eigensolver)
// stay on the host (dense.cpp). Every output element is summed in one fixed order
// (row chunks in order, then the chunks' partials in order): the same bits on every
// run.
#include <hip/hip_runtime.h>

#include "dense.hpp"
#include "device.hpp"

namespace ky {

namespace {

constexpr unsigned TN_TILE = 16;     // output tile of A^T B: 16 x 16 threads
constexpr unsigned TN_CHUNK = 64;    // rows of A and B a partial sums

// partial[c][i][j] = sum over rows r of chunk c of A[r][i] B[r][j] (fused multiply-
// adds in row order). A: m x p (lda), B: m x q (ldb).
__global__ __launch_bounds__(256) void tn_partial(const double *__restrict__ A, uint64_t lda,
                                                  const double *__restrict__ B, uint64_t ldb, uint64_t m, uint32_t p,
                                                  uint32_t q, double *__restrict__ partial) {
    const unsigned j = blockIdx.x * TN_TILE + threadIdx.x % TN_TILE;
    const unsigned i = blockIdx.y * TN_TILE + threadIdx.x / TN_TILE;
    const uint64_t r0 = (uint64_t)blockIdx.z * TN_CHUNK, r1 = min<uint64_t>(m, r0 + TN_CHUNK);
    if (i >= p || j >= q) return;
    double acc = 0.0;
    for (uint64_t r = r0; r < r1; r++) acc = fma(A[r * lda + i], B[r * ldb + j], acc);
    partial[((uint64_t)blockIdx.z * p + i) * q + j] = acc;
}

// out[e] = sum of the chunks' partials in chunk order.
__global__ void sum_chunks(const double *__restrict__ partial, uint32_t chunks, uint64_t count,
                           double *__restrict__ out) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    double s = 0.0;
    for (uint32_t c = 0; c < chunks; c++) s += partial[(uint64_t)c * count + e];
    out[e] = s;
}

// C[r][j] = (subtract ? C[r][j] - s : s), s = sum over k of A[r][k] S[k][j] (fused
// multiply-adds in k order). A: m x p (lda), S: p x q (dense), C: m x q (ldc).
__global__ __launch_bounds__(256) void nn_rows(const double *__restrict__ A, uint64_t lda,
                                               const double *__restrict__ S, uint64_t m, uint32_t p, uint32_t q,
                                               double *__restrict__ C, uint64_t ldc, int subtract) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * q) return;
    const uint64_t r = e / q;
    const uint32_t j = (uint32_t)(e % q);
    double s = 0.0;
    for (uint32_t k = 0; k < p; k++) s = fma(A[r * lda + k], S[(uint64_t)k * q + j], s);
    C[r * ldc + j] = subtract ? C[r * ldc + j] - s : s;
}

// a <- a r^-1 for each row of a (m x n, n <= 64; r upper triangular, n x n dense):
// one wavefront a row, lane j holding column j; column k is final once the columns
// before it are subtracted, and is then broadcast to the lanes after it (forward
// substitution in k order, fused multiply-adds).
__global__ __launch_bounds__(256) void solve_rows_upper(double *__restrict__ a, uint64_t m, uint32_t n,
                                                        const double *__restrict__ r) {
    const uint64_t row = (uint64_t)blockIdx.x * 4 + threadIdx.x / 64;
    if (row >= m) return;
    const unsigned j = threadIdx.x % 64;
    double x = j < n ? a[row * n + j] : 0.0;
    for (uint32_t k = 0; k < n; k++) {
        const double vk = __shfl(x, (int)k) / r[(uint64_t)k * n + k];
        if (j == k) x = vk;
        if (j > k && j < n) x = fma(-vk, r[(uint64_t)k * n + j], x);
    }
    if (j < n) a[row * n + j] = x;
}

// The larger of a and m, NaN once either is NaN.
__device__ __forceinline__ double max_nan(double m, double a) { return (a > m || a != a) ? a : m; }

// out[0] = max |x| and out[1] = sum of (x / max)^2 over count values (one block of
// 1024 threads: strided runs in order, 8 loads in flight a thread, then a fixed tree;
// 0 and 0 for an all-zero x; a value that is not finite makes the norm out[0]
// sqrt(out[1]) not finite).
constexpr unsigned SQ_THREADS = 1024, SQ_U = 8;
__global__ __launch_bounds__(SQ_THREADS) void scaled_squares(const double *__restrict__ x, uint64_t count,
                                                             double *__restrict__ out) {
    __shared__ double part[SQ_THREADS];
    const unsigned t = threadIdx.x;
    double m = 0.0;
    for (uint64_t i0 = t; i0 < count; i0 += SQ_THREADS * SQ_U) {
        double v[SQ_U];
#pragma unroll
        for (unsigned u = 0; u < SQ_U; u++) {
            const uint64_t i = i0 + (uint64_t)u * SQ_THREADS;
            v[u] = i < count ? fabs(x[i]) : 0.0;
        }
#pragma unroll
        for (unsigned u = 0; u < SQ_U; u++) m = max_nan(m, v[u]);
    }
    part[t] = m;
    __syncthreads();
    for (unsigned o = SQ_THREADS / 2; o > 0; o >>= 1) {
        if (t < o) part[t] = max_nan(part[t], part[t + o]);
        __syncthreads();
    }
    const double big = part[0];
    __syncthreads();
    double s = 0.0;
    if (big > 0.0)
        for (uint64_t i0 = t; i0 < count; i0 += SQ_THREADS * SQ_U) {
            double v[SQ_U];
#pragma unroll
            for (unsigned u = 0; u < SQ_U; u++) {
                const uint64_t i = i0 + (uint64_t)u * SQ_THREADS;
                v[u] = i < count ? x[i] / big : 0.0;
            }
#pragma unroll
            for (unsigned u = 0; u < SQ_U; u++) s += v[u] * v[u];
        }
    part[t] = s;
    __syncthreads();
    for (unsigned o = SQ_THREADS / 2; o > 0; o >>= 1) {
        if (t < o) part[t] += part[t + o];
        __syncthreads();
    }
    if (t == 0) {
        out[0] = big;
        out[1] = part[0];
    }
}

}  // namespace

DeviceBasis::DeviceBasis(Context &c, size_t rows_, size_t cap_) : ctx(&c), rows(rows_), cap(cap_) {
    basis.alloc(rows * cap);
    products.alloc(rows * cap);
}

// A host matrix into `tmp` (contiguous) through the pinned staging buffer.
static void upload_mat(Context &ctx, DevBuf<double> &tmp, const Mat &m) {
    tmp.reserve(m.v.size());
    upload_parts(ctx, {CopyPart{tmp.ptr, m.v.data(), m.v.size() * sizeof(double)}});
}

// Columns [at, at + w) of a rows x cap row-major array from a host rows x w Mat: one
// contiguous upload, then a strided copy on the device (a strided copy from pageable
// host memory goes row by row).
static void put_columns(Context &ctx, DevBuf<double> &tmp, DevBuf<double> &dst, size_t cap, size_t at, const Mat &m) {
    if (m.cols == 0) return;
    upload_mat(ctx, tmp, m);
    KY_HIP(hipMemcpy2DAsync(dst.ptr + at, cap * sizeof(double), tmp.ptr, m.cols * sizeof(double),
                            m.cols * sizeof(double), m.rows, hipMemcpyDeviceToDevice, ctx.stream));
    KY_HIP(hipStreamSynchronize(ctx.stream));
}

void DeviceBasis::append(const Mat &block, const Mat &product) {
    KY_CHECK(block.rows == rows && product.rows == rows && block.cols == product.cols && cols + block.cols <= cap,
             KY_E505, "pca: a Krylov block that does not fit the device basis (internal)");
    put_columns(*ctx, work_r, basis, cap, cols, block);
    put_columns(*ctx, work_r, products, cap, cols, product);
    cols += block.cols;
}

// out (p x q, dense on the device) = A^T B over `m` rows.
static void tn(Context &ctx, const double *A, uint64_t lda, const double *B, uint64_t ldb, uint64_t m, uint32_t p,
               uint32_t q, DevBuf<double> &partial, double *out) {
    const unsigned chunks = div_up(m, TN_CHUNK);
    partial.reserve((size_t)chunks * p * q);
    dim3 grid(div_up(q, TN_TILE), div_up(p, TN_TILE), chunks);
    tn_partial<<<grid, TN_TILE * TN_TILE, 0, ctx.stream>>>(A, lda, B, ldb, m, p, q, partial.ptr);
    KY_HIP(hipGetLastError());
    const uint64_t count = (uint64_t)p * q;
    sum_chunks<<<div_up(count, 256), 256, 0, ctx.stream>>>(partial.ptr, chunks, count, out);
    KY_HIP(hipGetLastError());
}

void DeviceBasis::project_out(Mat &rest) {
    if (cols == 0 || rest.cols == 0) return;
    const uint32_t w = (uint32_t)rest.cols, bc = (uint32_t)cols;
    work_p.reserve((size_t)bc * w);
    upload_mat(*ctx, work_r, rest);
    for (int pass = 0; pass < 2; pass++) {
        tn(*ctx, basis.ptr, cap, work_r.ptr, w, rows, bc, w, partial, work_p.ptr);
        nn_rows<<<div_up(rows * w, 256), 256, 0, ctx->stream>>>(basis.ptr, cap, work_p.ptr, rows, bc, w, work_r.ptr,
                                                                w, 1);
        KY_HIP(hipGetLastError());
    }
    download(*ctx, rest.v.data(), work_r.ptr, rows * w * sizeof(double));
}

void DeviceBasis::project_out_resident(const Mat &rest) {
    upload_mat(*ctx, work_r, rest);
    project_resident(rest.cols);
}

void DeviceBasis::project_resident(size_t width) {
    const uint32_t w = (uint32_t)width, bc = (uint32_t)cols;
    if (cols == 0 || w == 0) return;
    work_p.reserve((size_t)bc * w);
    for (int pass = 0; pass < 2; pass++) {
        tn(*ctx, basis.ptr, cap, work_r.ptr, w, rows, bc, w, partial, work_p.ptr);
        nn_rows<<<div_up(rows * w, 256), 256, 0, ctx->stream>>>(basis.ptr, cap, work_p.ptr, rows, bc, w, work_r.ptr,
                                                                w, 1);
        KY_HIP(hipGetLastError());
    }
}

void DeviceBasis::take_block(const double *src, size_t width) {
    work_r.reserve(std::max<size_t>(rows * width, 1));
    if (width)
        KY_HIP(hipMemcpyAsync(work_r.ptr, src, rows * width * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
}

void DeviceBasis::put_block(const Mat &block) {
    work_r.reserve(std::max<size_t>(block.v.size(), 1));
    if (!block.v.empty()) upload_parts(*ctx, {CopyPart{work_r.ptr, block.v.data(), block.v.size() * sizeof(double)}});
}

Mat DeviceBasis::get_block(size_t width) {
    Mat out(rows, width);
    if (width) download(*ctx, out.v.data(), work_r.ptr, rows * width * sizeof(double));
    return out;
}

double DeviceBasis::block_norm(size_t width) {
    if (width == 0) return 0.0;
    work_p.reserve(2);
    scaled_squares<<<1, SQ_THREADS, 0, ctx->stream>>>(work_r.ptr, rows * width, work_p.ptr);
    KY_HIP(hipGetLastError());
    double h[2];
    download(*ctx, h, work_p.ptr, sizeof(h));
    return h[0] * std::sqrt(h[1]);
}

void DeviceBasis::append_resident(const double *block, const double *product, size_t width) {
    KY_CHECK(cols + width <= cap, KY_E505, "pca: a Krylov block that does not fit the device basis (internal)");
    if (width == 0) return;
    KY_HIP(hipMemcpy2DAsync(basis.ptr + cols, cap * sizeof(double), block, width * sizeof(double),
                            width * sizeof(double), rows, hipMemcpyDeviceToDevice, ctx->stream));
    KY_HIP(hipMemcpy2DAsync(products.ptr + cols, cap * sizeof(double), product, width * sizeof(double),
                            width * sizeof(double), rows, hipMemcpyDeviceToDevice, ctx->stream));
    cols += width;
}

bool DeviceBasis::cholesky_qr2(size_t w, Mat &q, Mat &r, bool fetch_q) {
    KY_CHECK(w >= 1 && w <= 64, KY_E505, "pca: a device Cholesky QR of more than 64 columns (internal)");
    const uint32_t n = (uint32_t)w;
    q = Mat(rows, w);
    work_p.reserve(w * w);
    Mat g(w, w), factor[2];
    for (int pass = 0; pass < 2; pass++) {
        // The Gram matrix of the resident block, its Cholesky factor on the host, the
        // block times the factor's inverse on the device.
        tn(*ctx, work_r.ptr, n, work_r.ptr, n, rows, n, n, partial, work_p.ptr);
        download(*ctx, g.v.data(), work_p.ptr, w * w * sizeof(double));
        if (!cholesky_factor(g, factor[pass]) || (pass == 0 && condition_bound_of(factor[0]) > 1e6)) {
            download(*ctx, q.v.data(), work_r.ptr, rows * w * sizeof(double));
            return false;
        }
        upload_mat(*ctx, work_p, factor[pass]);
        solve_rows_upper<<<div_up(rows, 4), 256, 0, ctx->stream>>>(work_r.ptr, rows, n, work_p.ptr);
        KY_HIP(hipGetLastError());
    }
    if (fetch_q) download(*ctx, q.v.data(), work_r.ptr, rows * w * sizeof(double));
    // r = r2 r1 (both upper triangular).
    r = Mat(w, w);
    for (size_t i = 0; i < w; i++)
        for (size_t j = i; j < w; j++) {
            double v = 0.0;
            for (size_t k = i; k <= j; k++) v += factor[1].at(i, k) * factor[0].at(k, j);
            r.at(i, j) = v;
        }
    return true;
}

Mat DeviceBasis::rayleigh_matrix() {
    Mat t(cols, cols);
    if (cols == 0) return t;
    work_p.reserve(cols * cols);
    tn(*ctx, basis.ptr, cap, products.ptr, cap, rows, (uint32_t)cols, (uint32_t)cols, partial, work_p.ptr);
    download(*ctx, t.v.data(), work_p.ptr, cols * cols * sizeof(double));
    return t;
}

Mat DeviceBasis::times(const Mat &small) {
    KY_CHECK(small.rows == cols, KY_E505, "pca: a Ritz vector block of the wrong height (internal)");
    Mat out(rows, small.cols);
    if (cols == 0 || small.cols == 0) return out;
    const uint32_t q = (uint32_t)small.cols;
    work_r.reserve(rows * q);
    upload_mat(*ctx, work_p, small);
    nn_rows<<<div_up(rows * q, 256), 256, 0, ctx->stream>>>(basis.ptr, cap, work_p.ptr, rows, (uint32_t)cols, q,
                                                            work_r.ptr, q, 0);
    KY_HIP(hipGetLastError());
    download(*ctx, out.v.data(), work_r.ptr, rows * q * sizeof(double));
    return out;
}

}  // namespace ky
