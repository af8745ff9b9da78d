// The router MLP's kernels (gfx950): a BF16 MFMA GEMM with strided operands and a
// bias / ReLU epilogue, row top-T with log-sum-exp, the softmax cross-entropy
// gradient, the ReLU backward mask, bias gradients and Adam. Also used by k-means
// (queries x centroids scores) and centroid probing.
//
// GEMM: C[M x N] (f32, row stride ldc) = op(A)[M x K] op(B)[K x N], with A(i, k) at
// A[i * sai + k * sak] and B(k, j) at B[k * sbk + j * sbj] (f32 in memory, rounded to
// bf16 when staged in LDS: BF16 weights and activations, FP32 accumulation). Block
// tile 128 x 128, K step 32, four wavefronts each computing 64 x 64 with sixteen
// v_mfma_f32_16x16x32_bf16 a K step.
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

constexpr int GT = 128;    // block tile (rows and columns)
constexpr int GK = 32;     // K step
constexpr int GKP = 40;    // LDS row stride in bf16 (32 + 8 padding: 80 bytes)

// One operand tile (GT rows x GK of K, "row" = i for A, j for B) from f32 memory into
// LDS as bf16 [row][k]. KFAST: the operand is contiguous along k (thread t reads 16
// consecutive k of row t / 2). Else it is contiguous along the rows: with VEC (rows
// and the k stride 16-byte aligned) thread t reads a 4 x 4 block (rows 4 (t % 32)..,
// k 4 (t / 32)..) as four 16-byte loads, coalesced across the threads, and writes it
// transposed as four 8-byte rows; without, 16 consecutive rows at k = t / 8.
// Out-of-range elements are 0.
template <bool KFAST>
__device__ __forceinline__ void load_tile(const float *__restrict__ X, int64_t s_row, int64_t s_k, int row0, int rows,
                                          int k0, int K, bool vec, __bf16 *__restrict__ lds) {
    const int t = threadIdx.x;
    if constexpr (KFAST) {
        const int r = t / 2, kb = (t % 2) * 16;
        const int gr = row0 + r;
        __bf16 v[16];
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int gk = k0 + kb + u;
            v[u] = (__bf16)((gr < rows && gk < K) ? X[gr * s_row + gk * s_k] : 0.0f);
        }
        bf16x8 *dst = reinterpret_cast<bf16x8 *>(lds + r * GKP + kb);
        dst[0] = bf16x8{v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]};
        dst[1] = bf16x8{v[8], v[9], v[10], v[11], v[12], v[13], v[14], v[15]};
    } else if (vec) {
        const int rq = t % 32, kq = t / 32;
        const int gr = row0 + 4 * rq;
        float4 blk[4];
#pragma unroll
        for (int kk = 0; kk < 4; kk++) {
            const int gk = k0 + 4 * kq + kk;
            if (gk < K && gr + 3 < rows) {
                blk[kk] = *reinterpret_cast<const float4 *>(X + (int64_t)gk * s_k + gr);
            } else {
                float e[4];
#pragma unroll
                for (int u = 0; u < 4; u++) e[u] = (gk < K && gr + u < rows) ? X[(int64_t)gk * s_k + gr + u] : 0.0f;
                blk[kk] = make_float4(e[0], e[1], e[2], e[3]);
            }
        }
        typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
        *reinterpret_cast<bf16x4 *>(lds + (4 * rq + 0) * GKP + 4 * kq) =
            bf16x4{(__bf16)blk[0].x, (__bf16)blk[1].x, (__bf16)blk[2].x, (__bf16)blk[3].x};
        *reinterpret_cast<bf16x4 *>(lds + (4 * rq + 1) * GKP + 4 * kq) =
            bf16x4{(__bf16)blk[0].y, (__bf16)blk[1].y, (__bf16)blk[2].y, (__bf16)blk[3].y};
        *reinterpret_cast<bf16x4 *>(lds + (4 * rq + 2) * GKP + 4 * kq) =
            bf16x4{(__bf16)blk[0].z, (__bf16)blk[1].z, (__bf16)blk[2].z, (__bf16)blk[3].z};
        *reinterpret_cast<bf16x4 *>(lds + (4 * rq + 3) * GKP + 4 * kq) =
            bf16x4{(__bf16)blk[0].w, (__bf16)blk[1].w, (__bf16)blk[2].w, (__bf16)blk[3].w};
    } else {
        const int kk = t / 8, rb = (t % 8) * 16;
        const int gk = k0 + kk;
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int gr = row0 + rb + u;
            lds[(rb + u) * GKP + kk] = (__bf16)((gr < rows && gk < K) ? X[gr * s_row + gk * s_k] : 0.0f);
        }
    }
}

// Block tile GT x GT, four wavefronts each computing 64 x 64 with sixteen
// v_mfma_f32_16x16x32_bf16 a K step. Split K: block z sums K steps [z * kc, (z + 1) kc)
// and, with n_split > 1, writes its raw partial product to C + z * M * N (row stride
// N); gemm_split_sum then adds the parts in order and applies the epilogue.
template <bool A_KFAST, bool B_KFAST>
__global__ __launch_bounds__(256) void gemm_bf16(const float *__restrict__ A, int64_t sai, int64_t sak,
                                                 const float *__restrict__ B, int64_t sbk, int64_t sbj,
                                                 float *__restrict__ C, int64_t ldc, int M, int N, int K, int kc,
                                                 int n_split, bool a_vec, bool b_vec,
                                                 const float *__restrict__ bias, int bias_mode, int relu) {
    __shared__ __attribute__((aligned(16))) __bf16 As[GT * GKP];  // [row][k]
    __shared__ __attribute__((aligned(16))) __bf16 Bs[GT * GKP];  // [col][k]
    const int tid = threadIdx.x, lane = lane_id(), wave = tid / 64;
    const int row0 = blockIdx.y * GT, col0 = blockIdx.x * GT;
    const int wr = (wave / 2) * 64, wc = (wave % 2) * 64;  // the wave's 64 x 64 sub-tile
    const int kbeg = blockIdx.z * kc, kend = min(K, kbeg + kc);
    f32x4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = kbeg; k0 < kend; k0 += GK) {
        load_tile<A_KFAST>(A, sai, sak, row0, M, k0, kend, a_vec, As);
        load_tile<B_KFAST>(B, sbj, sbk, col0, N, k0, kend, b_vec, Bs);
        __syncthreads();
        bf16x8 af[4], bfr[4];
#pragma unroll
        for (int a = 0; a < 4; a++)
            af[a] = *reinterpret_cast<const bf16x8 *>(&As[(wr + 16 * a + lane % 16) * GKP + 8 * (lane / 16)]);
#pragma unroll
        for (int b = 0; b < 4; b++)
            bfr[b] = *reinterpret_cast<const bf16x8 *>(&Bs[(wc + 16 * b + lane % 16) * GKP + 8 * (lane / 16)]);
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
        __syncthreads();
    }
    float *out = n_split > 1 ? C + (int64_t)blockIdx.z * M * N : C;
    const int64_t ld = n_split > 1 ? N : ldc;
    // D layout: lane l, slot j -> row 4 (l / 16) + j, column l % 16 of the 16 x 16 tile.
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int gi = row0 + wr + 16 * a + 4 * (lane / 16) + j;
                const int gj = col0 + wc + 16 * b + lane % 16;
                if (gi < M && gj < N) {
                    float v = acc[a][b][j];
                    if (n_split == 1) {
                        if (bias_mode == 1) v += bias[gj];      // per column
                        else if (bias_mode == 2) v += bias[gi]; // per row
                        if (relu) v = fmaxf(v, 0.0f);
                    }
                    out[gi * ld + gj] = v;
                }
            }
}

// C = sum over the n_split parts (in order) + bias, ReLU.
__global__ void gemm_split_sum(const float *__restrict__ parts, int n_split, int M, int N, float *__restrict__ C,
                               int64_t ldc, const float *__restrict__ bias, int bias_mode, int relu) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)M * N) return;
    const int gi = (int)(t / N), gj = (int)(t % N);
    float v = 0.0f;
    for (int z = 0; z < n_split; z++) v += parts[(int64_t)z * M * N + t];
    if (bias_mode == 1) v += bias[gj];
    else if (bias_mode == 2) v += bias[gi];
    if (relu) v = fmaxf(v, 0.0f);
    C[gi * ldc + gj] = v;
}

void gemm(hipStream_t st, const float *A, int64_t sai, int64_t sak, const float *B, int64_t sbk, int64_t sbj,
          float *C, int64_t ldc, int M, int N, int K, const float *bias, int bias_mode, bool relu,
          DevBuf<float> *split_scratch) {
    if (M <= 0 || N <= 0) return;
    const bool a_kfast = sak == 1, b_kfast = sbk == 1;
    auto aligned = [](const float *p, int64_t s) { return ((uintptr_t)p % 16 == 0) && s % 4 == 0; };
    const bool a_vec = !a_kfast && sai == 1 && aligned(A, sak);
    const bool b_vec = !b_kfast && sbj == 1 && aligned(B, sbk);
    // Split K when the output tiles alone would leave most of the device idle.
    const unsigned tiles = div_up(N, GT) * div_up(M, GT);
    int n_split = 1;
    if (split_scratch && tiles < 256 && K >= 8 * GK)
        n_split = (int)std::min<int64_t>({8, div_up(256, tiles), K / (2 * GK)});
    const int kc = n_split > 1 ? (int)div_up(div_up(K, n_split), GK) * GK : K;
    n_split = n_split > 1 ? (int)div_up(K, kc) : 1;
    float *out = C;
    if (n_split > 1) {
        split_scratch->reserve((size_t)n_split * M * N);
        out = split_scratch->ptr;
    }
    dim3 grid(div_up(N, GT), div_up(M, GT), n_split);
    const int r = relu ? 1 : 0;
#define KY_GEMM(AK, BK)                                                                                          \
    gemm_bf16<AK, BK><<<grid, 256, 0, st>>>(A, sai, sak, B, sbk, sbj, out, ldc, M, N, K, kc, n_split, a_vec, b_vec, \
                                           bias, bias_mode, r)
    if (a_kfast && b_kfast) KY_GEMM(true, true);
    else if (a_kfast) KY_GEMM(true, false);
    else if (b_kfast) KY_GEMM(false, true);
    else KY_GEMM(false, false);
#undef KY_GEMM
    KY_HIP(hipGetLastError());
    if (n_split > 1) {
        gemm_split_sum<<<div_up((int64_t)M * N, 256), 256, 0, st>>>(out, n_split, M, N, C, ldc, bias, bias_mode, r);
        KY_HIP(hipGetLastError());
    }
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

// out[j] = sum over rows i of G[i * cols + j]: partial sums over chunks of CS_ROWS
// rows (each in row order), then the chunks' sums in order (the same bits on every
// run). Block (64 columns x 4 chunks): lanes read consecutive columns.
constexpr int CS_ROWS = 64;

__global__ __launch_bounds__(256) void col_sum_chunks(const float *__restrict__ G, int64_t rows, int cols,
                                                      float *__restrict__ partial) {
    const int j = blockIdx.x * 64 + lane_id();
    const int64_t chunk = (int64_t)blockIdx.y * 4 + threadIdx.x / 64;
    const int64_t r0 = chunk * CS_ROWS;
    if (j >= cols || r0 >= rows) return;
    const int64_t r1 = min<int64_t>(r0 + CS_ROWS, rows);
    float s = 0.0f;
    for (int64_t i = r0; i < r1; i++) s += G[i * cols + j];
    partial[chunk * cols + j] = s;
}

__global__ void col_sum_total(const float *__restrict__ partial, int64_t chunks, int cols, float *__restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    float s = 0.0f;
    for (int64_t c = 0; c < chunks; c++) s += partial[c * cols + j];
    out[j] = s;
}

void col_sum(hipStream_t st, const float *G, int64_t rows, int cols, float *out, DevBuf<float> &scratch) {
    const int64_t chunks = (rows + CS_ROWS - 1) / CS_ROWS;
    scratch.reserve((size_t)chunks * cols);
    dim3 grid(div_up(cols, 64), div_up(chunks, 4));
    col_sum_chunks<<<grid, 256, 0, st>>>(G, rows, cols, scratch.ptr);
    col_sum_total<<<div_up(cols, 256), 256, 0, st>>>(scratch.ptr, chunks, cols, out);
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

void Router::init(Context &ctx, uint32_t in, uint32_t hid, uint32_t out, uint64_t seed, bool zero_last) {
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
        const float bound = zero_last && layer == 2 ? 0.0f : std::sqrt(6.0f / (float)fan_in[layer]);
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
    gemm(st, H2, 1, hidden, dL, n_out, 1, W(2, grads), n_out, hidden, n_out, (int)rows, nullptr, 0, false, &split_scratch);
    col_sum(st, dL, rows, n_out, b(2, grads), colsum_scratch);
    gemm(st, dL, n_out, 1, W(2, p), 1, n_out, dH2, hidden, (int)rows, hidden, n_out, nullptr, 0, false);
    relu_mask(st, dH2, H2, rows * hidden);
    // Layer 2.
    gemm(st, H1, 1, hidden, dH2, hidden, 1, W(1, grads), hidden, hidden, hidden, (int)rows, nullptr, 0, false, &split_scratch);
    col_sum(st, dH2, rows, hidden, b(1, grads), colsum_scratch);
    gemm(st, dH2, hidden, 1, W(1, p), 1, hidden, dH1, hidden, (int)rows, hidden, hidden, nullptr, 0, false);
    relu_mask(st, dH1, H1, rows * hidden);
    // Layer 1.
    gemm(st, X, 1, d_in, dH1, hidden, 1, W(0, grads), hidden, d_in, hidden, (int)rows, nullptr, 0, false, &split_scratch);
    col_sum(st, dH1, rows, hidden, b(0, grads), colsum_scratch);
}

void Router::adam_step(hipStream_t st, float lr) {
    step++;
    adam(st, params.ptr, grads.ptr, m1.ptr, m2.ptr, (int64_t)n_params, lr, step);
}

}  // namespace ky
