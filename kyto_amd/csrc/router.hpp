// The router MLP and its kernels (router.hip).
#pragma once

#include "common.hpp"

namespace ky {

// split_scratch (optional): device scratch for splitting K when the output is small.
void gemm(hipStream_t st, const float *A, int64_t sai, int64_t sak, const float *B, int64_t sbk, int64_t sbj,
          float *C, int64_t ldc, int M, int N, int K, const float *bias, int bias_mode, bool relu,
          DevBuf<float> *split_scratch = nullptr);
void row_top(hipStream_t st, const float *S, int64_t rows, int cols, int T, uint32_t *top_idx, float *top_val,
             float *lse);
void softmax_xent(hipStream_t st, float *L, int64_t rows, int cols, const uint32_t *t_idx, const float *t_w, int tk,
                  float inv_batch, float *loss);
void relu_mask(hipStream_t st, float *G, const float *H, int64_t count);
void col_sum(hipStream_t st, const float *G, int64_t rows, int cols, float *out, DevBuf<float> &scratch);
void adam(hipStream_t st, float *w, const float *g, float *m, float *v, int64_t count, float lr, int step);

// d_in -> hidden -> hidden -> n_out, ReLU between, f32 master parameters laid out
// [W1 (d_in x hidden), b1, W2 (hidden x hidden), b2, W3 (hidden x n_out), b3].
struct Router {
    uint32_t d_in = 0, hidden = 0, n_out = 0;
    size_t n_params = 0;
    int step = 0;
    DevBuf<float> params, grads, m1, m2;
    DevBuf<float> colsum_scratch;  // the bias gradients' chunk sums
    DevBuf<float> split_scratch;   // split-K parts of the weight gradients
    // He-uniform weights from the seed, zero biases; zero_last: the output layer's
    // weights zero too (the MLP starts as a zero correction of another score).
    void init(Context &ctx, uint32_t in, uint32_t hid, uint32_t out, uint64_t seed, bool zero_last = false);
    float *W(int layer, DevBuf<float> &buf) const;
    float *b(int layer, DevBuf<float> &buf) const;
    void forward(hipStream_t st, const float *X, int64_t rows, float *H1, float *H2, float *L);
    void backward(hipStream_t st, const float *X, int64_t rows, const float *H1, const float *H2, const float *dL,
                  float *dH2, float *dH1);
    void adam_step(hipStream_t st, float lr);
};

}  // namespace ky
