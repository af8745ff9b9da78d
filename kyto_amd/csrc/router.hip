// The router MLP's kernels (gfx950): a BF16 MFMA GEMM with strided operands and a
// bias / ReLU epilogue, row top-T with log-sum-exp, the softmax cross-entropy
// gradient, the ReLU backward mask, bias gradients and Adam. Also used by k-means
// (queries x centroids scores) and centroid probing.
//
// GEMM: C[M x N] (f32, row stride ldc) = op(A)[M x K] op(B)[K x N], with A(i, k) at
// A[i * sai + k * sak] and B(k, j) at B[k * sbk + j * sbj] (f32 in memory, rounded to
// bf16 when staged in LDS: BF16 weights and activations, FP32 accumulation). Block
// tile 64 x 64, K step 32, four wavefronts each computing 32 x 32 with four
// v_mfma_f32_16x16x32_bf16 per K step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "device.hpp"
#include "router.hpp"

namespace ky {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int GT = 64;     // block tile (rows and columns)
constexpr int GK = 32;     // K step
constexpr int GKP = 40;    // LDS row stride in bf16 (32 + 8 padding: 80 bytes)

__global__ __launch_bounds__(256) void gemm_bf16(const float *__restrict__ A, int64_t sai, int64_t sak,
                                                 const float *__restrict__ B, int64_t sbk, int64_t sbj,
                                                 float *__restrict__ C, int64_t ldc, int M, int N, int K,
                                                 const float *__restrict__ bias, int bias_mode, int relu) {
    __shared__ __attribute__((aligned(16))) __bf16 As[GT * GKP];  // [row][k]
    __shared__ __attribute__((aligned(16))) __bf16 Bs[GT * GKP];  // [col][k]
    const int tid = threadIdx.x, lane = lane_id(), wave = tid / 64;
    const int row0 = blockIdx.y * GT, col0 = blockIdx.x * GT;
    const int wr = (wave / 2) * 32, wc = (wave % 2) * 32;  // the wave's 32 x 32 sub-tile
    f32x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // Loads walk the operand's contiguous dimension across threads.
    const bool a_rows_fast = sai == 1 && sak != 1;
    const bool b_cols_fast = sbj == 1 && sbk != 1;
    for (int k0 = 0; k0 < K; k0 += GK) {
        for (int e = tid; e < GT * GK; e += 256) {
            int r, kk;
            if (a_rows_fast) {
                r = e % GT;
                kk = e / GT;
            } else {
                kk = e % GK;
                r = e / GK;
            }
            const int gi = row0 + r, gk = k0 + kk;
            float v = (gi < M && gk < K) ? A[gi * sai + gk * sak] : 0.0f;
            As[r * GKP + kk] = (__bf16)v;
        }
        for (int e = tid; e < GT * GK; e += 256) {
            int c, kk;
            if (b_cols_fast) {
                c = e % GT;
                kk = e / GT;
            } else {
                kk = e % GK;
                c = e / GK;
            }
            const int gj = col0 + c, gk = k0 + kk;
            float v = (gj < N && gk < K) ? B[gk * sbk + gj * sbj] : 0.0f;
            Bs[c * GKP + kk] = (__bf16)v;
        }
        __syncthreads();
        bf16x8 af[2], bfr[2];
#pragma unroll
        for (int a = 0; a < 2; a++)
            af[a] = *reinterpret_cast<const bf16x8 *>(&As[(wr + 16 * a + lane % 16) * GKP + 8 * (lane / 16)]);
#pragma unroll
        for (int b = 0; b < 2; b++)
            bfr[b] = *reinterpret_cast<const bf16x8 *>(&Bs[(wc + 16 * b + lane % 16) * GKP + 8 * (lane / 16)]);
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
        __syncthreads();
    }
    // D layout: lane l, slot j -> row 4 (l / 16) + j, column l % 16 of the 16 x 16 tile.
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int gi = row0 + wr + 16 * a + 4 * (lane / 16) + j;
                const int gj = col0 + wc + 16 * b + lane % 16;
                if (gi < M && gj < N) {
                    float v = acc[a][b][j];
                    if (bias_mode == 1) v += bias[gj];      // per column
                    else if (bias_mode == 2) v += bias[gi]; // per row
                    if (relu) v = fmaxf(v, 0.0f);
                    C[gi * ldc + gj] = v;
                }
            }
}

void gemm(hipStream_t st, const float *A, int64_t sai, int64_t sak, const float *B, int64_t sbk, int64_t sbj,
          float *C, int64_t ldc, int M, int N, int K, const float *bias, int bias_mode, bool relu) {
    if (M <= 0 || N <= 0) return;
    dim3 grid(div_up(N, GT), div_up(M, GT));
    gemm_bf16<<<grid, 256, 0, st>>>(A, sai, sak, B, sbk, sbj, C, ldc, M, N, K, bias, bias_mode, relu ? 1 : 0);
    KY_HIP(hipGetLastError());
}

// ------------------------------------------------------------ row reductions ----

// Per row of a rows x cols score matrix: the T largest (value, column) by (value
// descending, column ascending), and log-sum-exp of the row. One wavefront a row.
template <int T>
__global__ __launch_bounds__(256) void row_top(const float *__restrict__ S, int64_t rows, int cols,
                                               uint32_t *__restrict__ top_idx, float *__restrict__ top_val,
                                               float *__restrict__ lse) {
    const int wave = threadIdx.x / 64, lane = lane_id();
    const int64_t r = (int64_t)blockIdx.x * 4 + wave;
    if (r >= rows) return;
    const float *row = S + r * cols;
    // Each lane keeps its own top T, sorted.
    float v[T];
    int c[T];
#pragma unroll
    for (int t = 0; t < T; t++) {
        v[t] = -INFINITY;
        c[t] = 0x7FFFFFFF;
    }
    float mx = -INFINITY;
    for (int j = lane; j < cols; j += 64) {
        float x = row[j];
        mx = fmaxf(mx, x);
        if (x > v[T - 1]) {
            // insert (x, j); columns increase along the lane, so ties keep the earlier
            float cx = x;
            int cj = j;
#pragma unroll
            for (int t = 0; t < T; t++) {
                if (cx > v[t]) {
                    float tv = v[t];
                    int tc = c[t];
                    v[t] = cx;
                    c[t] = cj;
                    cx = tv;
                    cj = tc;
                }
            }
        }
    }
    // Row max and sum of exp (fixed lane tree).
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    float s = 0.0f;
    for (int j = lane; j < cols; j += 64) s += expf(row[j] - mx);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    // Merge the lanes' lists: T rounds of a wave-wide arg-max over the lanes' heads.
    int head = 0;
    for (int t = 0; t < T; t++) {
        float hv = head < T ? v[0] : -INFINITY;
        int hc = head < T ? c[0] : 0x7FFFFFFF;
        // Lanes shift their lists instead of indexing: v[0] is the lane's current head.
        float bv = hv;
        int bc = hc;
        for (int o = 32; o > 0; o >>= 1) {
            float ov = __shfl_xor(bv, o);
            int oc = __shfl_xor(bc, o);
            if (ov > bv || (ov == bv && oc < bc)) {
                bv = ov;
                bc = oc;
            }
        }
        if (lane == 0) {
            // A row without a finite entry (or fewer than T columns) names column 0.
            top_idx[r * T + t] = bc < cols ? (uint32_t)bc : 0u;
            top_val[r * T + t] = bv;
        }
        if (hc == bc && hv == bv && head < T) {
            // this lane's head was taken: shift
#pragma unroll
            for (int u = 0; u + 1 < T; u++) {
                v[u] = v[u + 1];
                c[u] = c[u + 1];
            }
            v[T - 1] = -INFINITY;
            c[T - 1] = 0x7FFFFFFF;
            head++;
        }
    }
    if (lane == 0 && lse) lse[r] = mx + logf(s);
}

void row_top(hipStream_t st, const float *S, int64_t rows, int cols, int T, uint32_t *top_idx, float *top_val,
             float *lse) {
    if (rows <= 0) return;
    const unsigned blocks = div_up(rows, 4);
    switch (T) {
        case 1: row_top<1><<<blocks, 256, 0, st>>>(S, rows, cols, top_idx, top_val, lse); break;
        case 8: row_top<8><<<blocks, 256, 0, st>>>(S, rows, cols, top_idx, top_val, lse); break;
        default: throw Error(KY_E215, "row_top: T must be 1 or 8");
    }
    KY_HIP(hipGetLastError());
}

// dL = (softmax(L) - T) / batch for one row per wavefront; T given sparse as
// (list, weight) pairs (k per row); also the row's cross-entropy -sum T log p.
__global__ __launch_bounds__(256) void softmax_xent(float *__restrict__ L, int64_t rows, int cols,
                                                    const uint32_t *__restrict__ t_idx, const float *__restrict__ t_w,
                                                    int tk, float inv_batch, float *__restrict__ loss) {
    const int wave = threadIdx.x / 64, lane = lane_id();
    const int64_t r = (int64_t)blockIdx.x * 4 + wave;
    if (r >= rows) return;
    float *row = L + r * cols;
    float mx = -INFINITY;
    for (int j = lane; j < cols; j += 64) mx = fmaxf(mx, row[j]);
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    float s = 0.0f;
    for (int j = lane; j < cols; j += 64) s += expf(row[j] - mx);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    const float lse = mx + logf(s);
    float ce = 0.0f;
    if (lane < tk) {
        uint32_t j = t_idx[r * tk + lane];
        ce = -t_w[r * tk + lane] * (row[j] - lse);
    }
    for (int o = 32; o > 0; o >>= 1) ce += __shfl_xor(ce, o);
    wave_sync_memory();
    for (int j = lane; j < cols; j += 64) row[j] = expf(row[j] - lse) * inv_batch;
    wave_sync_memory();
    // Subtract the targets lane by lane (targets may repeat a list: serialise).
    for (int t = 0; t < tk; t++) {
        if (lane == 0) {
            uint32_t j = t_idx[r * tk + t];
            row[j] -= t_w[r * tk + t] * inv_batch;
        }
        wave_sync_memory();
    }
    if (lane == 0 && loss) loss[r] = ce;
}

void softmax_xent(hipStream_t st, float *L, int64_t rows, int cols, const uint32_t *t_idx, const float *t_w, int tk,
                  float inv_batch, float *loss) {
    softmax_xent<<<div_up(rows, 4), 256, 0, st>>>(L, rows, cols, t_idx, t_w, tk, inv_batch, loss);
    KY_HIP(hipGetLastError());
}

// G *= (H > 0), elementwise.
__global__ void relu_mask(float *__restrict__ G, const float *__restrict__ H, int64_t count) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < count && !(H[i] > 0.0f)) G[i] = 0.0f;
}

void relu_mask(hipStream_t st, float *G, const float *H, int64_t count) {
    relu_mask<<<div_up(count, 256), 256, 0, st>>>(G, H, count);
    KY_HIP(hipGetLastError());
}

// out[j] = sum over rows i (in order) of G[i * cols + j].
__global__ void col_sum(const float *__restrict__ G, int64_t rows, int cols, float *__restrict__ out) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    float s = 0.0f;
    for (int64_t i = 0; i < rows; i++) s += G[i * cols + j];
    out[j] = s;
}

void col_sum(hipStream_t st, const float *G, int64_t rows, int cols, float *out) {
    col_sum<<<div_up(cols, 256), 256, 0, st>>>(G, rows, cols, out);
    KY_HIP(hipGetLastError());
}

// Adam (Kingma and Ba 2015): m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2,
// w -= lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps).
__global__ void adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                     float *__restrict__ v, int64_t count, float lr, float b1, float b2, float eps, float c1,
                     float c2) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= count) return;
    float gi = g[i];
    float mi = b1 * m[i] + (1.0f - b1) * gi;
    float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

void adam(hipStream_t st, float *w, const float *g, float *m, float *v, int64_t count, float lr, int step) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    adam<<<div_up(count, 256), 256, 0, st>>>(w, g, m, v, count, lr, b1, b2, eps, c1, c2);
    KY_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ the MLP ----

void Router::init(Context &ctx, uint32_t in, uint32_t hid, uint32_t out, uint64_t seed) {
    d_in = in;
    hidden = hid;
    n_out = out;
    const size_t sizes[6] = {(size_t)in * hid, hid, (size_t)hid * hid, hid, (size_t)hid * out, out};
    size_t total = 0;
    for (size_t s : sizes) total += s;
    n_params = total;
    params.alloc(total);
    grads.alloc(total);
    m1.alloc(total);
    m2.alloc(total);
    // He-uniform weights from the seed, zero biases (host, then one copy).
    std::vector<float> h(total, 0.0f);
    size_t off = 0;
    const uint32_t fan_in[3] = {in, hid, hid};
    for (int layer = 0; layer < 3; layer++) {
        const size_t cnt = sizes[2 * layer];
        const float bound = std::sqrt(6.0f / (float)fan_in[layer]);
        for (size_t i = 0; i < cnt; i++) {
            uint64_t z = seed * 0x9E3779B97F4A7C15ULL + (uint64_t)layer * 0xD1B54A32D192ED03ULL + i;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
            z ^= z >> 31;
            float u = (float)((double)(z >> 11) * 0x1p-53);
            h[off + i] = (2.0f * u - 1.0f) * bound;
        }
        off += cnt + sizes[2 * layer + 1];
    }
    KY_HIP(hipMemcpyAsync(params.ptr, h.data(), sizeof(float) * total, hipMemcpyHostToDevice, ctx.stream));
    KY_HIP(hipMemsetAsync(m1.ptr, 0, sizeof(float) * total, ctx.stream));
    KY_HIP(hipMemsetAsync(m2.ptr, 0, sizeof(float) * total, ctx.stream));
    step = 0;
}

float *Router::W(int layer, DevBuf<float> &buf) const {
    size_t off = 0;
    if (layer >= 1) off += (size_t)d_in * hidden + hidden;
    if (layer >= 2) off += (size_t)hidden * hidden + hidden;
    return buf.ptr + off;
}

float *Router::b(int layer, DevBuf<float> &buf) const {
    const size_t w = layer == 0 ? (size_t)d_in * hidden : layer == 1 ? (size_t)hidden * hidden : (size_t)hidden * n_out;
    return W(layer, buf) + w;
}

void Router::forward(hipStream_t st, const float *X, int64_t rows, float *H1, float *H2, float *L) {
    DevBuf<float> &p = const_cast<DevBuf<float> &>(params);
    gemm(st, X, d_in, 1, W(0, p), hidden, 1, H1, hidden, (int)rows, hidden, d_in, b(0, p), 1, true);
    gemm(st, H1, hidden, 1, W(1, p), hidden, 1, H2, hidden, (int)rows, hidden, hidden, b(1, p), 1, true);
    gemm(st, H2, hidden, 1, W(2, p), n_out, 1, L, n_out, (int)rows, n_out, hidden, b(2, p), 1, false);
}

void Router::backward(hipStream_t st, const float *X, int64_t rows, const float *H1, const float *H2, const float *dL,
                      float *dH2, float *dH1) {
    DevBuf<float> &p = const_cast<DevBuf<float> &>(params);
    // Layer 3: dW3 = H2^T dL, db3 = colsum dL, dH2 = dL W3^T (masked by H2 > 0).
    gemm(st, H2, 1, hidden, dL, n_out, 1, W(2, grads), n_out, hidden, n_out, (int)rows, nullptr, 0, false);
    col_sum(st, dL, rows, n_out, b(2, grads));
    gemm(st, dL, n_out, 1, W(2, p), 1, n_out, dH2, hidden, (int)rows, hidden, n_out, nullptr, 0, false);
    relu_mask(st, dH2, H2, rows * hidden);
    // Layer 2.
    gemm(st, H1, 1, hidden, dH2, hidden, 1, W(1, grads), hidden, hidden, hidden, (int)rows, nullptr, 0, false);
    col_sum(st, dH2, rows, hidden, b(1, grads));
    gemm(st, dH2, hidden, 1, W(1, p), 1, hidden, dH1, hidden, (int)rows, hidden, hidden, nullptr, 0, false);
    relu_mask(st, dH1, H1, rows * hidden);
    // Layer 1.
    gemm(st, X, 1, d_in, dH1, hidden, 1, W(0, grads), hidden, d_in, hidden, (int)rows, nullptr, 0, false);
    col_sum(st, dH1, rows, hidden, b(0, grads));
}

void Router::adam_step(hipStream_t st, float lr) {
    step++;
    adam(st, params.ptr, grads.ptr, m1.ptr, m2.ptr, (int64_t)n_params, lr, step);
}

}  // namespace ky
