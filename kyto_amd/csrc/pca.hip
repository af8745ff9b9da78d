// pca() on the device: the resident CSR layer, its chunked column-major copy, and the
// three kinds of pass kyto's GPU fold runs (crates/kyto-exec/src/pca/gpu.rs:149-219),
// with the host FP64 steps of PcaAccumulator (pca.rs:302-862) around them.
//
// Kernels (gfx950, wave64):
// - build_columns: per chunk of CHUNK_ROWS rows, the chunk's entries reordered by
//   (gene, row) in place of kyto's host CSC (launch_matrix.rs:405-424). Deterministic:
//   per-wave integer counts in LDS, a scan, then each wave places its rows in order.
// - omega_rows: kyto's Omega rows (pca.rs:936-950) drawn per (cell, column) from the
//   integer SplitMix64 stream, rounded to f32, plus a column of ones (gene sums).
// - row_project: one wavefront per row, lane j = column j: w_j = -offset_j +
//   sum_e value_e * M[gene_e, j], exact f32 x f32 products accumulated in f64, the
//   row's entries read with wave-uniform (scalar) loads, M's rows read from L2.
//   Serves X Q - mu^T Q (product passes) and X V - mu^T V (the embedding).
// - column_product: X^T W per chunk, one wavefront per (chunk, gene), lane = column,
//   over the gene's run of the chunk (rows ascending) with W's rows of the chunk read
//   from L2; blocks of one chunk are mapped onto one XCD (blockIdx % 8) so they share
//   its W rows. f64 partials per chunk, then reduce_chunks adds the chunks in order.
// No float atomics anywhere: the same bits on every run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>

#include "common.hpp"
#include "dense.hpp"
#include "device.hpp"
#include "layer.hpp"
#include "pca.hpp"

namespace ky {

// The stream's SplitMix64 output function (pca.rs:946-950).
__device__ __forceinline__ uint64_t mix64(uint64_t v) {
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ULL;
    v = (v ^ (v >> 27)) * 0x94D049BB133111EBULL;
    return v ^ (v >> 31);
}

// ----------------------------------------------------------------- validation --

// Indices strictly increasing below n_cols in every row, values finite and 0 or within
// GPU_RANGE in magnitude (pca/gpu.rs:120). flags: bit 0 index, bit 1 value. A wavefront
// a row at a time (rows wave, wave + waves, ...), lane j on entries s + j, s + j + 64, ...,
// each compared with the entry before it in the row: the reads are coalesced.
__global__ __launch_bounds__(256) void check_rows(const uint64_t *__restrict__ indptr,
                                                  const uint32_t *__restrict__ indices,
                                                  const float *__restrict__ values, uint64_t n_rows, uint32_t n_cols,
                                                  unsigned *__restrict__ flags) {
    const unsigned lane = lane_id();
    const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / 64);
    unsigned bad = 0;
    for (uint64_t r = blockIdx.x * (uint64_t)(blockDim.x / 64) + threadIdx.x / 64; r < n_rows; r += waves) {
        const uint64_t s = indptr[r], e = indptr[r + 1];
        for (uint64_t i = s + lane; i < e; i += 64) {
            const uint32_t g = indices[i];
            if (g >= n_cols || (i > s && g <= indices[i - 1])) bad |= 1u;
            const float a = fabsf(values[i]);
            if (!(a == 0.0f || (a >= 1e-10f && a <= 1e10f))) bad |= 2u;
        }
    }
    const unsigned any = (__ballot(bad & 1u) ? 1u : 0u) | (__ballot(bad & 2u) ? 2u : 0u);
    if (lane == 0 && any) atomicOr(flags, any);
}

// --------------------------------------------------------- chunked columns ----

constexpr int BC_U = 8;  // steps of 64 entries whose loads build_columns issues together

// One block per chunk; wavefront w takes a contiguous run of the chunk's rows. Each
// wave counts its rows' genes (LDS counters of its own), the counts are scanned in
// (gene, wave) order, and each wave then places its entries 64 at a time in entry
// order, a row at a time within a step (genes are distinct within a row), so every
// gene's entries land in row order. LDS: waves * n_cols counters, blockDim partial
// sums, then each wave's row offsets (span + 1 of them, relative to the chunk's first
// entry).
__global__ __launch_bounds__(1024) void build_columns(const uint64_t *__restrict__ indptr,
                                                      const uint32_t *__restrict__ indices,
                                                      const float *__restrict__ values, uint64_t n_rows,
                                                      uint32_t n_cols, uint32_t chunk_rows,
                                                      uint32_t *__restrict__ col_ptr, uint32_t *__restrict__ col_row,
                                                      float *__restrict__ col_val) {
    extern __shared__ uint32_t lds[];
    const unsigned waves = blockDim.x / 64;
    uint32_t *cnt = lds;                            // [waves][n_cols]
    uint32_t *part = lds + (size_t)waves * n_cols;  // [blockDim]
    const unsigned c = blockIdx.x;
    const uint64_t r0 = (uint64_t)c * chunk_rows;
    const uint64_t r1 = min(r0 + chunk_rows, n_rows);
    const unsigned wave = threadIdx.x / 64, lane = lane_id();
    const uint32_t span = (uint32_t)((r1 - r0 + waves - 1) / waves);
    uint32_t *rp = part + blockDim.x + (size_t)wave * (span + 1);  // [waves][span + 1]
    for (size_t i = threadIdx.x; i < (size_t)waves * n_cols; i += blockDim.x) cnt[i] = 0;
    const uint64_t wr0 = min(r0 + (uint64_t)wave * span, r1), wr1 = min(wr0 + span, r1);
    const uint32_t nr = (uint32_t)(wr1 - wr0);
    const uint64_t base = indptr[r0];
    for (uint32_t i = lane; i <= nr; i += 64) rp[i] = (uint32_t)(indptr[wr0 + i] - base);
    __syncthreads();
    const uint32_t f0 = rp[0], f1 = rp[nr];
    uint32_t *mine = cnt + (size_t)wave * n_cols;
    const uint32_t *ix = indices + base;
    const float *vx = values + base;
    // Counts, BC_U steps of 64 entries loaded at a time (the loads do not wait on one
    // another).
    for (uint32_t f = f0; f < f1; f += 64 * BC_U) {
        uint32_t g[BC_U];
#pragma unroll
        for (int u = 0; u < BC_U; u++) {
            const uint32_t i = f + 64 * u + lane;
            g[u] = i < f1 ? ix[i] : 0xFFFFFFFFu;
        }
#pragma unroll
        for (int u = 0; u < BC_U; u++)
            if (g[u] != 0xFFFFFFFFu) atomicAdd(&mine[g[u]], 1u);
    }
    __syncthreads();
    // Exclusive scan over (gene, wave) in gene-major order: each thread sums a run of
    // genes, the runs' sums are scanned (Hillis-Steele), then each thread writes its run.
    const unsigned per = (n_cols + blockDim.x - 1) / blockDim.x;
    const unsigned g0 = min(threadIdx.x * per, n_cols), g1 = min(g0 + per, n_cols);
    uint32_t local = 0;
    for (unsigned g = g0; g < g1; g++)
        for (unsigned w = 0; w < waves; w++) local += cnt[(size_t)w * n_cols + g];
    part[threadIdx.x] = local;
    __syncthreads();
    for (unsigned off = 1; off < blockDim.x; off <<= 1) {
        const uint32_t add = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += add;
        __syncthreads();
    }
    uint32_t run = (uint32_t)base + part[threadIdx.x] - local;
    uint32_t *cp = col_ptr + (size_t)c * (n_cols + 1);
    for (unsigned g = g0; g < g1; g++) {
        cp[g] = run;
        for (unsigned w = 0; w < waves; w++) {
            uint32_t x = cnt[(size_t)w * n_cols + g];
            cnt[(size_t)w * n_cols + g] = run;
            run += x;
        }
    }
    if (threadIdx.x == 0) cp[n_cols] = (uint32_t)indptr[r1];
    __syncthreads();
    // Placement. Lane l holds entry f + l of the step; its row (an index into rp)
    // advances with the entry. The step's rows are taken in order, the lanes of a row
    // together.
    uint32_t row = 0;
    const uint32_t row_off = (uint32_t)(wr0 - r0);
    for (uint32_t f = f0; f < f1; f += 64 * BC_U) {
        uint32_t g[BC_U];
        float v[BC_U];
#pragma unroll
        for (int u = 0; u < BC_U; u++) {
            const uint32_t i = f + 64 * u + lane;
            g[u] = i < f1 ? ix[i] : 0u;
            v[u] = i < f1 ? vx[i] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < BC_U; u++) {
            const uint32_t fu = f + 64 * u;
            if (fu >= f1) break;
            const uint32_t i = fu + lane;
            const bool valid = i < f1;
            if (valid)
                while (rp[row + 1] <= i) row++;
            const uint32_t last = min(63u, f1 - fu - 1);
            const uint32_t rlo = (uint32_t)__builtin_amdgcn_readfirstlane((int)row);
            const uint32_t rhi = (uint32_t)__shfl((int)row, (int)last);
            for (uint32_t rr = rlo; rr <= rhi; rr++) {
                if (valid && row == rr) {
                    const uint32_t pos = mine[g[u]];
                    mine[g[u]] = pos + 1;
                    col_row[pos] = row_off + rr;
                    col_val[pos] = v[u];
                }
            }
        }
    }
}

// A gene's run in a chunk with at least this many entries is summed by a whole
// wavefront (wave_run), a shorter one by a lane group (slice_walk). With the runs
// taken longest first, measured at 1M x 2,000 (column product a launch): 256 1.55 ms,
// 768 1.39, 1,024 1.37, 1,536 1.37, every run by a lane group 1.45.
constexpr uint32_t LONG_RUN = 1024;

// One block per chunk: the chunk's genes ordered by the length of their runs, longest
// first (a counting sort over lengths 0 .. chunk_rows: LDS counts, a descending scan,
// then each gene placed at its length's next slot), and the number of runs of at
// least LONG_RUN entries after them: order[c * (n_cols + 1) + i] for i < n_cols, then
// the count. Genes of equal length land in any order among themselves; the column
// product sums every run the same way wherever it stands, so its sums do not depend
// on that order. LDS: chunk_rows + 1 counters.
__global__ __launch_bounds__(1024) void order_runs(const uint32_t *__restrict__ col_ptr, uint32_t n_cols,
                                                   uint32_t chunk_rows, uint32_t *__restrict__ order) {
    extern __shared__ uint32_t slot[];
    const unsigned c = blockIdx.x;
    const uint32_t *cp = col_ptr + (size_t)c * (n_cols + 1);
    uint32_t *out = order + (size_t)c * (n_cols + 1);
    const uint32_t bins = chunk_rows + 1;
    for (uint32_t i = threadIdx.x; i < bins; i += blockDim.x) slot[i] = 0;
    __syncthreads();
    for (uint32_t g = threadIdx.x; g < n_cols; g += blockDim.x) atomicAdd(&slot[min(cp[g + 1] - cp[g], chunk_rows)], 1u);
    __syncthreads();
    // Descending exclusive scan: slot[len] = genes with longer runs. Each thread takes
    // a run of bins from the top, the runs' sums are scanned (Hillis-Steele), then each
    // thread writes its run.
    __shared__ uint32_t part[1024];
    const uint32_t per = (bins + blockDim.x - 1) / blockDim.x;
    const uint32_t b0 = min(threadIdx.x * per, bins), b1 = min(b0 + per, bins);
    uint32_t local = 0;
    for (uint32_t b = b0; b < b1; b++) local += slot[bins - 1 - b];
    part[threadIdx.x] = local;
    __syncthreads();
    for (unsigned off = 1; off < blockDim.x; off <<= 1) {
        const uint32_t add = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += add;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - local;
    for (uint32_t b = b0; b < b1; b++) {
        const uint32_t x = slot[bins - 1 - b];
        slot[bins - 1 - b] = run;
        run += x;
    }
    __syncthreads();
    if (threadIdx.x == 0) out[n_cols] = LONG_RUN <= chunk_rows ? slot[LONG_RUN - 1] : 0u;
    __syncthreads();
    for (uint32_t g = threadIdx.x; g < n_cols; g += blockDim.x)
        out[atomicAdd(&slot[min(cp[g + 1] - cp[g], chunk_rows)], 1u)] = g;
}

// ------------------------------------------------------------------- Omega ----

// W[r, j] = f32(Omega[cell_offset + r, j]) for j < l, and W[r, l] = 1.
__global__ void omega_rows(uint64_t n_rows, uint64_t seed, uint64_t cell_offset, uint32_t l,
                           float *__restrict__ W, uint32_t ldw) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    uint64_t total = n_rows * (uint64_t)(l + 1);
    if (i >= total) return;
    uint64_t r = i / (l + 1);
    uint32_t j = (uint32_t)(i % (l + 1));
    float out = 1.0f;
    if (j < l) {
        const uint64_t phi = 0x9E3779B97F4A7C15ULL;
        uint64_t s0 = mix64(seed ^ mix64(cell_offset + r + phi));
        uint64_t state = s0 + (uint64_t)(j + 1) * phi;
        double v = (double)(mix64(state) >> 11) * 0x1p-52 - 1.0;
        out = (float)v;
    }
    W[r * ldw + j] = out;
}

// --------------------------------------------------------------- row project ----

// Both sparse products walk a run of entries (gene or row index, value) that is the
// same for the whole wavefront. The run is read through wave-uniform pointers, which
// the compiler turns into scalar loads (8 entries a load), and each entry's dense row
// is one buffer load whose row offset is a scalar register and whose lane offset is
// fixed, so an entry costs one vector load, two conversions and one f64 fused
// multiply-add, with UNROLL dense rows in flight. An f32 x f32 product is exact in
// f64, so the fused multiply-add rounds exactly as the separate product and sum
// would. Every lane loads (a branch around the loads would wait for each one): a lane
// past the last column reads into the next dense row, or into the DENSE_PAD floats
// after the last one, and its sum is never stored.
constexpr int UNROLL = 16;
// Floats allocated after every dense operand of the products (>= 64 * max CPL): the
// lanes past the last column of the last row read them.
constexpr size_t DENSE_PAD = 256;

// The wave's index within its block as a wave-uniform value.
__device__ __forceinline__ unsigned wave_in_block() {
    return (unsigned)__builtin_amdgcn_readfirstlane((int)(threadIdx.x / 64));
}

// A buffer resource over `bytes` bytes at p (loads past the end return 0).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t dense_rsrc(const float *p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(p), 0, (int)bytes, 0x00020000);
}

__device__ __forceinline__ float dense_load(__amdgpu_buffer_rsrc_t r, uint32_t lane_bytes, uint32_t row_bytes) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)lane_bytes, (int)row_bytes, 0));
}

// One wavefront per row; lane j holds columns j, j+64, ... (CPL of them, w <= 64*CPL).
// out[r, j] = f32(-offsets[j] + sum_e f64(value_e) * f64(M[gene_e, j])), the entries
// added in row order. M: m_rows x w, rows ldm floats apart.
template <int CPL>
__global__ __launch_bounds__(256) void row_project(const uint64_t *__restrict__ indptr,
                                                   const uint32_t *__restrict__ indices,
                                                   const float *__restrict__ values, uint64_t n_rows,
                                                   const float *__restrict__ M, uint32_t m_rows, uint32_t w,
                                                   uint32_t ldm, const double *__restrict__ offsets,
                                                   float *__restrict__ out, uint32_t ldo) {
    const uint64_t r = (uint64_t)blockIdx.x * (blockDim.x / 64) + wave_in_block();
    if (r >= n_rows) return;
    const unsigned lane = lane_id();
    const __amdgpu_buffer_rsrc_t mr = dense_rsrc(M, (m_rows * ldm + (uint32_t)DENSE_PAD) * 4u);
    const uint32_t row_bytes = ldm * 4u;
    double acc[CPL];
#pragma unroll
    for (int c = 0; c < CPL; c++) {
        const unsigned j = lane + 64 * c;
        acc[c] = j < w ? -offsets[j] : 0.0;
    }
    const uint64_t s = indptr[r], e = indptr[r + 1];
    const uint32_t *gi = indices + s, *ge = indices + e;
    const float *vi = values + s;
    for (; gi + UNROLL <= ge; gi += UNROLL, vi += UNROLL) {
        float m[UNROLL][CPL];
#pragma unroll
        for (int u = 0; u < UNROLL; u++)
#pragma unroll
            for (int c = 0; c < CPL; c++) m[u][c] = dense_load(mr, 4u * (lane + 64 * c), gi[u] * row_bytes);
#pragma unroll
        for (int u = 0; u < UNROLL; u++)
#pragma unroll
            for (int c = 0; c < CPL; c++) acc[c] = fma((double)vi[u], (double)m[u][c], acc[c]);
    }
    for (; gi < ge; gi++, vi++) {
#pragma unroll
        for (int c = 0; c < CPL; c++)
            acc[c] = fma((double)*vi, (double)dense_load(mr, 4u * (lane + 64 * c), *gi * row_bytes), acc[c]);
    }
#pragma unroll
    for (int c = 0; c < CPL; c++) {
        const unsigned j = lane + 64 * c;
        if (j < w) out[r * ldo + j] = (float)acc[c];
    }
}

// ------------------------------------------------------------ column product ----

// partial[(chunk * n_cols + g) * w + j] = sum over gene g's entries of the chunk (rows
// ascending) of f64(value) * f64(W[row, j]). One wavefront per gene, 4 per block,
// `genes_per_block` genes per block; blocks of one chunk land on one XCD, so the
// chunk's rows of W are read through that XCD's L2. W: n_rows x ldw.
template <int CPL>
__global__ __launch_bounds__(256) void column_product(const uint32_t *__restrict__ col_ptr,
                                                      const uint32_t *__restrict__ col_row,
                                                      const float *__restrict__ col_val, uint32_t n_cols,
                                                      uint32_t n_chunks, uint32_t chunk_rows, uint64_t n_rows,
                                                      const float *__restrict__ W, uint32_t w, uint32_t ldw,
                                                      uint32_t genes_per_block, double *__restrict__ partial) {
    const unsigned groups = (n_cols + genes_per_block - 1) / genes_per_block;
    const unsigned xcd = blockIdx.x % 8, local = blockIdx.x / 8;
    const unsigned slot = local / groups, group = local % groups;
    const unsigned chunk = slot * 8 + xcd;
    if (chunk >= n_chunks) return;
    const unsigned lane = lane_id();
    const unsigned wave = wave_in_block(), waves = blockDim.x / 64;
    const unsigned gend = min((group + 1) * genes_per_block, n_cols);
    const uint32_t *cp = col_ptr + (size_t)chunk * (n_cols + 1);
    const uint64_t r0 = (uint64_t)chunk * chunk_rows;
    const uint32_t rows = (uint32_t)min<uint64_t>(chunk_rows, n_rows - r0);
    const __amdgpu_buffer_rsrc_t wr = dense_rsrc(W + r0 * ldw, (rows * ldw + (uint32_t)DENSE_PAD) * 4u);
    const uint32_t row_bytes = ldw * 4u;
    for (unsigned g = group * genes_per_block + wave; g < gend; g += waves) {
        double acc[CPL];
#pragma unroll
        for (int c = 0; c < CPL; c++) acc[c] = 0.0;
        const uint32_t *ri = col_row + cp[g], *re = col_row + cp[g + 1];
        const float *vi = col_val + cp[g];
        for (; ri + UNROLL <= re; ri += UNROLL, vi += UNROLL) {
            float m[UNROLL][CPL];
#pragma unroll
            for (int u = 0; u < UNROLL; u++)
#pragma unroll
                for (int c = 0; c < CPL; c++) m[u][c] = dense_load(wr, 4u * (lane + 64 * c), ri[u] * row_bytes);
#pragma unroll
            for (int u = 0; u < UNROLL; u++)
#pragma unroll
                for (int c = 0; c < CPL; c++) acc[c] = fma((double)vi[u], (double)m[u][c], acc[c]);
        }
        for (; ri < re; ri++, vi++) {
#pragma unroll
            for (int c = 0; c < CPL; c++)
                acc[c] = fma((double)*vi, (double)dense_load(wr, 4u * (lane + 64 * c), *ri * row_bytes), acc[c]);
        }
        double *out = partial + ((size_t)chunk * n_cols + g) * w;
#pragma unroll
        for (int c = 0; c < CPL; c++) {
            const unsigned j = lane + 64 * c;
            if (j < w) out[j] = acc[c];
        }
    }
}

// ------------------------------------------- the products with an LDS-staged slice ----
//
// The dense operand sliced by columns: a block holds one slice of SW (8 or 16) columns
// in LDS (Q's slice for every gene, or W's slice for one chunk's rows), so an entry
// reads its SW dense values from LDS instead of gathering a 256-byte row from L2.
// Each run (a row's entries, or a gene's in a chunk) belongs to LPR = SW / 4 lanes of
// one quad, each adding the exact products of 4 columns in f64 in the run's order, so
// the sums are the bits of the kernels above. A wavefront holds 64 / LPR runs and walks
// them 16 entries a step: a run's lanes load its next step's entries (index, value)
// while the current step is added, and each entry's LDS offset and value reach the
// run's other lanes by a DPP broadcast inside the quad. A lane group whose run ends
// takes the block's next run at once (its bounds are loaded a run ahead). The n_slices
// blocks of a row block (or chunk) follow one another on one XCD, so the entries they
// all stream come from HBM about once and from that XCD's L2 after.
constexpr unsigned LDS_FLOATS = 32768;  // 128 KB a block

template <int SW>
struct SliceShape {
    static constexpr int LPR = SW / 4;      // lanes a run
    static constexpr int RPW = 64 / LPR;    // runs a wavefront holds
    static constexpr int STEP = 16;         // entries of a run a step
    static constexpr int EPL = STEP / LPR;  // entries a lane loads a step
};

// x of lane SRC of this lane's run (the run's LPR lanes of its quad). Every source lane
// exists, so bound_ctrl (which lets the broadcast fold into the add using it) changes
// nothing.
template <int LPR, int SRC>
__device__ __forceinline__ int run_bcast(int x) {
    if constexpr (LPR == 4)
        return __builtin_amdgcn_mov_dpp(x, SRC * 0x55, 0xf, 0xf, true);
    else
        return __builtin_amdgcn_mov_dpp(x, SRC | SRC << 2 | (2 + SRC) << 4 | (2 + SRC) << 6, 0xf, 0xf, true);
}

// A lane's EPL entries of a step as loaded: the step's entry LPR * k + (the lane's
// place in the run) is the lane's k-th.
template <int SW>
struct SliceStep {
    uint32_t g[SliceShape<SW>::EPL];  // dense row in LDS
    float v[SliceShape<SW>::EPL];     // value
};

// The step at entries p, p + 1, ... of a run, read through buffer resources over the
// block's entries (an entry past them reads index 0 and value 0; one past the run's
// end but inside them belongs to a run of this block, so its index is a dense row in
// LDS).
template <int SW>
__device__ __forceinline__ void slice_step_load(__amdgpu_buffer_rsrc_t ri, __amdgpu_buffer_rsrc_t rv, uint32_t p,
                                                unsigned place, SliceStep<SW> &st) {
    using S = SliceShape<SW>;
#pragma unroll
    for (int k = 0; k < S::EPL; k++) {
        const int at = (int)(4u * (p + S::LPR * k + place));
        st.g[k] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(ri, at, 0, 0);
        st.v[k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rv, at, 0, 0));
    }
}

// Adds the step's entries below e, 8 LDS reads in flight at a time.
template <int SW, int H, int... J>
__device__ __forceinline__ void slice_half(const char *qs, uint32_t quad_bytes, const uint32_t (&a)[SliceShape<SW>::EPL],
                                           const float (&v)[SliceShape<SW>::EPL], double (&acc)[4],
                                           std::integer_sequence<int, J...>) {
    constexpr int LPR = SliceShape<SW>::LPR;
    const float4 q[] = {*reinterpret_cast<const float4 *>(
        qs + (uint32_t)run_bcast<LPR, (H + J) % LPR>((int)a[(H + J) / LPR]) + quad_bytes)...};
    const float x[] = {__builtin_bit_cast(float, run_bcast<LPR, (H + J) % LPR>(__builtin_bit_cast(int, v[(H + J) / LPR])))...};
#pragma unroll
    for (int u = 0; u < (int)sizeof...(J); u++) {
        const double dv = (double)x[u];
        acc[0] = fma(dv, (double)q[u].x, acc[0]);
        acc[1] = fma(dv, (double)q[u].y, acc[1]);
        acc[2] = fma(dv, (double)q[u].z, acc[2]);
        acc[3] = fma(dv, (double)q[u].w, acc[3]);
    }
}

template <int SW>
__device__ __forceinline__ void slice_add(const char *qs, uint32_t quad_bytes, const SliceStep<SW> &st, uint32_t p,
                                          uint32_t e, unsigned place, double (&acc)[4]) {
    using S = SliceShape<SW>;
    uint32_t a[S::EPL];
    float v[S::EPL];
#pragma unroll
    for (int k = 0; k < S::EPL; k++) {
        a[k] = st.g[k] * (uint32_t)(SW * 4);
        v[k] = p + S::LPR * k + place < e ? st.v[k] : 0.0f;
    }
    slice_half<SW, 0>(qs, quad_bytes, a, v, acc, std::make_integer_sequence<int, S::STEP / 2>{});
    slice_half<SW, S::STEP / 2>(qs, quad_bytes, a, v, acc, std::make_integer_sequence<int, S::STEP / 2>{});
}

// The lane group's next run: the group's first lane takes it from the block's counter
// in LDS and the group's other lanes receive it by the broadcast.
template <int LPR>
__device__ __forceinline__ uint32_t grab_run(uint32_t *counter, unsigned place) {
    uint32_t t = 0;
    if (place == 0) t = atomicAdd(counter, 1u);
    return (uint32_t)run_bcast<LPR, 0>((int)t);
}

// The block's runs 0 .. n_runs-1 over its lane groups: group `grp` (of `groups`) starts
// with run grp and then takes runs from `counter` (set to `groups` before) until none is
// left, so the groups end together whatever the runs' lengths. bounds(run, s, e) gives a
// run's entries [s, e) relative to the resources and whether this walk sums it (a run it
// leaves to another engine has e = s); done(run, acc) takes its 4 sums (started from
// init). Each step issues its loads first: the bounds of the run after the next (when
// this step ends a run), then the next step's entries, so the step's own work waits only
// for loads issued a step or more before.
template <int SW, typename Bounds, typename Done>
__device__ __forceinline__ void slice_walk(const float *qs, __amdgpu_buffer_rsrc_t ri, __amdgpu_buffer_rsrc_t rv,
                                           uint32_t n_runs, uint32_t *counter, uint32_t grp, unsigned place,
                                           const double (&init)[4], Bounds bounds, Done done) {
    using S = SliceShape<SW>;
    const char *qb = reinterpret_cast<const char *>(qs);
    const uint32_t quad_bytes = 16u * place;
    uint32_t run = grp, p = 0, e = 0, ns = 0, ne = 0;
    bool own = false, nown = false;
    if (run < n_runs) own = bounds(run, p, e);
    uint32_t nrun = grab_run<S::LPR>(counter, place);
    if (nrun < n_runs) nown = bounds(nrun, ns, ne);
    double acc[4] = {init[0], init[1], init[2], init[3]};
    // One step: returns false once no group of the wavefront has a run. Two step
    // buffers in turn, so no loaded value is copied (a copy would wait for its load).
    const auto step = [&](SliceStep<SW> &cur, SliceStep<SW> &next) {
        if (!__builtin_amdgcn_ballot_w64(run < n_runs)) return false;
        const bool ends = p + S::STEP >= e;
        uint32_t np = p + S::STEP, nend = e, after = nrun;
        bool aown = nown;
        if (ends) {
            np = ns;
            nend = ne;
            after = grab_run<S::LPR>(counter, place);
            aown = false;
            ns = ne = 0;
            if (after < n_runs) aown = bounds(after, ns, ne);
        }
        slice_step_load<SW>(ri, rv, np, place, next);
        slice_add<SW>(qb, quad_bytes, cur, p, e, place, acc);
        if (ends) {
            if (run < n_runs && own) done(run, acc);
#pragma unroll
            for (int c = 0; c < 4; c++) acc[c] = init[c];
            run = nrun;
            own = nown;
            nrun = after;
            nown = aown;
        }
        p = np;
        e = nend;
        return true;
    };
    SliceStep<SW> a, b;
    slice_step_load<SW>(ri, rv, p, place, a);
    while (step(a, b) && step(b, a)) {
    }
}

// One long run [s, e) by a whole wavefront: lane (group eg, place) adds entries s + eg,
// s + eg + EG, ... (EG = 64 / LPR groups, U of them a step, the next step's loaded while
// this one is added; only the last step is masked), then the groups' sums are added in
// a fixed butterfly, so every lane ends with its place's 4 sums.
template <int SW>
__device__ __forceinline__ void wave_run(const char *qb, __amdgpu_buffer_rsrc_t ri, __amdgpu_buffer_rsrc_t rv,
                                         uint32_t s, uint32_t e, unsigned eg, unsigned place, double (&acc)[4]) {
    constexpr int LPR = SW / 4, EG = 64 / LPR, U = 4, STEP = EG * U;
    const uint32_t quad_bytes = 16u * place;
#pragma unroll
    for (int c = 0; c < 4; c++) acc[c] = 0.0;
    uint32_t ag[U], bg[U];
    float av[U], bv[U];
    const auto load = [&](uint32_t p, uint32_t (&g)[U], float (&v)[U]) {
        const int at = (int)(4u * (p + eg));
#pragma unroll
        for (int u = 0; u < U; u++) {
            g[u] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(ri, at + 4 * EG * u, 0, 0);
            v[u] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rv, at + 4 * EG * u, 0, 0));
        }
    };
    const auto add = [&](uint32_t p, const uint32_t (&g)[U], float (&v)[U]) {
        if (p + STEP > e) {
#pragma unroll
            for (int u = 0; u < U; u++)
                if (p + eg + EG * u >= e) v[u] = 0.0f;
        }
        float4 q[U];
#pragma unroll
        for (int u = 0; u < U; u++)
            q[u] = *reinterpret_cast<const float4 *>(qb + g[u] * (uint32_t)(SW * 4) + quad_bytes);
#pragma unroll
        for (int u = 0; u < U; u++) {
            const double dv = (double)v[u];
            acc[0] = fma(dv, (double)q[u].x, acc[0]);
            acc[1] = fma(dv, (double)q[u].y, acc[1]);
            acc[2] = fma(dv, (double)q[u].z, acc[2]);
            acc[3] = fma(dv, (double)q[u].w, acc[3]);
        }
    };
    // Two step buffers in turn: a step's loads are issued before the step before it is
    // added, and no loaded value is copied (a copy would wait for its load).
    load(s, ag, av);
    for (uint32_t p = s;;) {
        load(p + STEP, bg, bv);
        add(p, ag, av);
        p += STEP;
        if (p >= e) break;
        load(p + STEP, ag, av);
        add(p, bg, bv);
        p += STEP;
        if (p >= e) break;
    }
#pragma unroll
    for (int o = LPR; o < 64; o <<= 1)
#pragma unroll
        for (int c = 0; c < 4; c++) acc[c] += __shfl_xor(acc[c], o);
}

// Block (row block or chunk `unit`, slice): the n_slices blocks of a unit are
// consecutive on XCD blockIdx % 8.
__device__ __forceinline__ void slice_block(unsigned n_slices, unsigned &unit, unsigned &slice) {
    const unsigned xcd = blockIdx.x % 8, local = blockIdx.x / 8;
    slice = local % n_slices;
    unit = (local / n_slices) * 8 + xcd;
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t entry_rsrc(const void *p, uint32_t count) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, (int)(4u * count), 0x00020000);
}

// row_project with M's column slice in LDS (m_rows * SW <= LDS_FLOATS, w <= ldm, M's
// columns up to n_slices * SW readable, rows 16-byte aligned).
template <int SW>
__global__ __launch_bounds__(1024) void row_project_lds(const uint64_t *__restrict__ indptr,
                                                        const uint32_t *__restrict__ indices,
                                                        const float *__restrict__ values, uint64_t n_rows,
                                                        const float *__restrict__ M, uint32_t m_rows, uint32_t w,
                                                        uint32_t ldm, const double *__restrict__ offsets,
                                                        float *__restrict__ out, uint32_t ldo, uint32_t n_slices,
                                                        uint32_t rows_per_block) {
    using S = SliceShape<SW>;
    __shared__ __attribute__((aligned(16))) float qs[LDS_FLOATS];
    __shared__ uint32_t counter;
    unsigned rb, slice;
    slice_block(n_slices, rb, slice);
    const uint64_t r0 = (uint64_t)rb * rows_per_block;
    if (r0 >= n_rows) return;
    const uint32_t groups = (blockDim.x / 64) * S::RPW;
    if (threadIdx.x == 0) counter = groups;
    const uint32_t nr = (uint32_t)min<uint64_t>(rows_per_block, n_rows - r0);
    const unsigned c0 = slice * SW;
    for (unsigned i = threadIdx.x; i < m_rows * (SW / 4); i += blockDim.x)
        reinterpret_cast<float4 *>(qs)[i] =
            *reinterpret_cast<const float4 *>(M + (size_t)(i / (SW / 4)) * ldm + c0 + 4 * (i % (SW / 4)));
    __syncthreads();
    const uint64_t base = indptr[r0];
    const uint32_t count = (uint32_t)(indptr[r0 + nr] - base);
    const __amdgpu_buffer_rsrc_t ri = entry_rsrc(indices + base, count), rv = entry_rsrc(values + base, count);
    const unsigned lane = lane_id(), place = lane % S::LPR;
    const unsigned col = c0 + 4 * place;
    double init[4];
#pragma unroll
    for (int c = 0; c < 4; c++) init[c] = col + c < w ? -offsets[col + c] : 0.0;
    slice_walk<SW>(
        qs, ri, rv, nr, &counter, wave_in_block() * S::RPW + lane / S::LPR, place, init,
        [&](uint32_t run, uint32_t &s, uint32_t &e) {
            s = (uint32_t)(indptr[r0 + run] - base);
            e = (uint32_t)(indptr[r0 + run + 1] - base);
            return true;
        },
        [&](uint32_t run, const double (&acc)[4]) {
            float *o = out + (r0 + run) * ldo + col;
#pragma unroll
            for (int c = 0; c < 4; c++)
                if (col + c < w) o[c] = (float)acc[c];
        });
}

// column_product with W's column slice of the chunk's rows in LDS (chunk_rows * SW <=
// LDS_FLOATS, W's columns up to n_slices * SW readable, rows 16-byte aligned). The
// chunk's runs are taken longest first (order, from order_runs): the wavefronts first
// sum the long runs one at a time, whole, then sum the short runs as lane groups, so
// the runs left when the block's groups run out of work are its shortest. partial as
// column_product's. With col_sums, the slice's columns of W summed over the chunk's
// rows too (f64, in a fixed order), into col_sums[chunk * w + column]. 1024 threads.
template <int SW>
__global__ __launch_bounds__(1024) void column_product_lds(const uint32_t *__restrict__ col_ptr,
                                                           const uint32_t *__restrict__ col_row,
                                                           const float *__restrict__ col_val,
                                                           const uint32_t *__restrict__ order, uint32_t n_cols,
                                                           uint32_t n_chunks, uint32_t chunk_rows, uint64_t n_rows,
                                                           const float *__restrict__ W, uint32_t w, uint32_t ldw,
                                                           uint32_t n_slices, double *__restrict__ partial,
                                                           double *__restrict__ col_sums) {
    using S = SliceShape<SW>;
    __shared__ __attribute__((aligned(16))) float ws[LDS_FLOATS];
    __shared__ uint32_t counter[2];  // next long run, next short run
    __shared__ double wave_sums[16][SW];
    unsigned chunk, slice;
    slice_block(n_slices, chunk, slice);
    if (chunk >= n_chunks) return;
    const uint32_t groups = (blockDim.x / 64) * S::RPW;
    if (threadIdx.x == 0) {
        counter[0] = 0;
        counter[1] = groups;
    }
    const uint64_t r0 = (uint64_t)chunk * chunk_rows;
    const uint32_t rows = (uint32_t)min<uint64_t>(chunk_rows, n_rows - r0);
    const unsigned c0 = slice * SW;
    for (unsigned i = threadIdx.x; i < rows * (SW / 4); i += blockDim.x)
        reinterpret_cast<float4 *>(ws)[i] =
            *reinterpret_cast<const float4 *>(W + (r0 + i / (SW / 4)) * ldw + c0 + 4 * (i % (SW / 4)));
    __syncthreads();
    if (col_sums) {
        // Wavefront v sums rows [v R, (v + 1) R) of the slice (R = rows / 16 rounded up),
        // lane l column l % SW of every (64 / SW)-th row from l / SW (each read a
        // contiguous 256 bytes of the slice), the lanes of a column then added in a
        // butterfly; the 16 wavefronts' sums in order by wavefront 0.
        const unsigned v = wave_in_block(), l = lane_id();
        const uint32_t R = (rows + 15) / 16, rend = min(rows, (v + 1) * R);
        double t = 0.0;
        for (uint32_t r = v * R + l / SW; r < rend; r += 64 / SW) t += (double)ws[r * SW + l % SW];
#pragma unroll
        for (int o = SW; o < 64; o <<= 1) t += __shfl_xor(t, o);
        if (l < SW) wave_sums[v][l] = t;
        __syncthreads();
        if (v == 0 && l < SW && c0 + l < w) {
            double total = 0.0;
            for (int u = 0; u < 16; u++) total += wave_sums[u][l];
            col_sums[(size_t)chunk * w + c0 + l] = total;
        }
    }
    const uint32_t *cp = col_ptr + (size_t)chunk * (n_cols + 1);
    const uint32_t *ord = order + (size_t)chunk * (n_cols + 1);
    const uint32_t n_long = ord[n_cols];
    const uint32_t base = cp[0], count = cp[n_cols] - base;
    const __amdgpu_buffer_rsrc_t ri = entry_rsrc(col_row + base, count), rv = entry_rsrc(col_val + base, count);
    const unsigned lane = lane_id(), place = lane % S::LPR;
    const unsigned col = c0 + 4 * place;
    const auto store = [&](uint32_t g, const double (&acc)[4]) {
        double *o = partial + ((size_t)chunk * n_cols + g) * w;
#pragma unroll
        for (int c = 0; c < 4; c++)
            if (col + c < w) o[col + c] = acc[c];
    };
    // The long runs, one a wavefront at a time.
    for (;;) {
        uint32_t i = 0;
        if (lane == 0) i = atomicAdd(&counter[0], 1u);
        i = (uint32_t)__builtin_amdgcn_readfirstlane((int)i);
        if (i >= n_long) break;
        const uint32_t g = ord[i];
        double acc[4];
        wave_run<SW>(reinterpret_cast<const char *>(ws), ri, rv, cp[g] - base, cp[g + 1] - base, lane / S::LPR,
                     place, acc);
        if (lane < S::LPR) store(g, acc);
    }
    // The short runs (run r of the walk is the chunk's (n_long + r)-th longest).
    const uint32_t *short_ord = ord + n_long;
    const double init[4] = {0.0, 0.0, 0.0, 0.0};
    slice_walk<SW>(
        ws, ri, rv, n_cols - n_long, &counter[1], wave_in_block() * S::RPW + lane / S::LPR, place, init,
        [&](uint32_t r, uint32_t &s, uint32_t &e) {
            const uint32_t g = short_ord[r];
            s = cp[g] - base;
            e = cp[g + 1] - base;
            return true;
        },
        [&](uint32_t r, const double (&acc)[4]) { store(short_ord[r], acc); });
}

// out[i] = sum over chunks c (in order) of partial[c * stride + i].
__global__ void reduce_chunks(const double *__restrict__ partial, uint64_t stride, uint32_t n_chunks,
                              double *__restrict__ out) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= stride) return;
    double s = 0.0;
    for (uint32_t c = 0; c < n_chunks; c++) s += partial[(size_t)c * stride + i];
    out[i] = s;
}

// out[g * stride + i] = sum over parts c in [g * group, (g + 1) * group) (in order)
// of partial[c * stride + i]: one level of a fixed reduction tree.
__global__ void reduce_groups(const double *__restrict__ partial, uint64_t stride, uint32_t n_parts, uint32_t group,
                              double *__restrict__ out) {
    uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint32_t n_groups = (n_parts + group - 1) / group;
    if (t >= stride * n_groups) return;
    const uint64_t g = t / stride, i = t % stride;
    const uint32_t c1 = min<uint32_t>((uint32_t)(g + 1) * group, n_parts);
    double s = 0.0;
    for (uint32_t c = (uint32_t)g * group; c < c1; c++) s += partial[(size_t)c * stride + i];
    out[g * stride + i] = s;
}

// seg_sums[s * w + j] = sum over the rows of segment s (SEG rows, in order) of
// f64(W[r, j]); the segments are then added in order (reduce_chunks).
constexpr uint32_t SEG = 256;
__global__ void column_sums(const float *__restrict__ W, uint64_t n_rows, uint32_t w, uint32_t ldw,
                            double *__restrict__ seg_sums) {
    const uint64_t s = blockIdx.x;
    const unsigned j = threadIdx.x;
    if (j >= w) return;
    const uint64_t r0 = s * SEG, r1 = min(r0 + SEG, n_rows);
    double acc = 0.0;
    for (uint64_t r = r0; r < r1; r++) acc += (double)W[r * ldw + j];
    seg_sums[s * w + j] = acc;
}

// Per (chunk, gene): the centred squares sum_e (f64(v_e) - mu_g)^2 over the chunk's
// stored values of the gene, and their count. One wavefront per (chunk, gene): lane l
// sums the run's entries l, l + 64, ... in order, then a fixed butterfly adds the
// lanes (the same bits on every run).
__global__ __launch_bounds__(256) void gene_squares(const uint32_t *__restrict__ col_ptr,
                                                    const float *__restrict__ col_val, uint32_t n_cols,
                                                    uint32_t n_chunks, const double *__restrict__ means,
                                                    double *__restrict__ sq_partial,
                                                    double *__restrict__ count_partial) {
    const uint64_t i = (uint64_t)blockIdx.x * (blockDim.x / 64) + wave_in_block();
    if (i >= (uint64_t)n_chunks * n_cols) return;
    const unsigned c = (unsigned)(i / n_cols), g = (unsigned)(i % n_cols);
    const unsigned lane = lane_id();
    const uint32_t *cp = col_ptr + (size_t)c * (n_cols + 1);
    const double mu = means[g];
    const uint32_t s = cp[g], e = cp[g + 1];
    double acc = 0.0;
    for (uint32_t k = s + lane; k < e; k += 64) {
        const double d = (double)col_val[k] - mu;
        acc += d * d;
    }
#pragma unroll
    for (int off = 32; off >= 1; off /= 2) acc += __shfl_xor(acc, off);
    if (lane == 0) {
        sq_partial[i] = acc;
        count_partial[i] = (double)(e - s);
    }
}

// ================================================================== host side ====

// Device time of the sparse products of one pca() call: HIP events around each launch
// on the library's stream, summed per kernel once the call has synchronised (kind 0:
// row_project, kind 1: column_product). Set for the duration of pca_run on its thread.
// bytes[kind] sums each launch's algorithmic bytes: the sparse operand it streams
// (8 bytes a stored value, plus row or column pointers), the dense operand read once
// and the output written once.
struct SpmmClock {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[2];
    double bytes[2] = {0.0, 0.0};
    void begin(int kind, hipStream_t st, double algorithmic_bytes) {
        bytes[kind] += algorithmic_bytes;
        hipEvent_t a, b;
        KY_HIP(hipEventCreate(&a));
        KY_HIP(hipEventCreate(&b));
        ev[kind].push_back({a, b});
        KY_HIP(hipEventRecord(a, st));
    }
    void end(int kind, hipStream_t st) { KY_HIP(hipEventRecord(ev[kind].back().second, st)); }
    double total(int kind) {
        double t = 0.0;
        for (auto &p : ev[kind]) {
            KY_HIP(hipEventSynchronize(p.second));
            float ms = 0.0f;
            KY_HIP(hipEventElapsedTime(&ms, p.first, p.second));
            t += ms;
        }
        return t;
    }
    ~SpmmClock() {
        for (auto &v : ev)
            for (auto &p : v) {
                (void)hipEventDestroy(p.first);
                (void)hipEventDestroy(p.second);
            }
    }
};
static thread_local SpmmClock *t_clock = nullptr;

static unsigned lds_limit(const Context &ctx) {
    return (unsigned)std::min<size_t>(ctx.props.sharedMemPerBlock, 64 * 1024);
}

// The layer's maps (when given), the row checks and the chunked column copy of a CSR
// whose rows are resident (indptr, indices, values at csr.nnz).
static void csr_finish(Csr &csr, const LayerMaps &maps) {
    Context &ctx = *csr.ctx;
    const uint64_t n = csr.n_rows;
    hipStream_t st = ctx.stream;
    // The input layer's maps on the uploaded values (normalize_total, log1p).
    normalize_log1p(st, csr.indptr.ptr, n, csr.values.ptr, maps.target_sum, maps.normalize, maps.log1p);
    DevBuf<unsigned> flags(1);
    KY_HIP(hipMemsetAsync(flags.ptr, 0, sizeof(unsigned), st));
    check_rows<<<(unsigned)std::min<uint64_t>(div_up(n, 4), 8u * 256u * 4u), 256, 0, st>>>(
        csr.indptr.ptr, csr.indices.ptr, csr.values.ptr, n, csr.n_cols,
                                              flags.ptr);
    KY_HIP(hipGetLastError());
    unsigned hflags = 0;
    KY_HIP(hipMemcpyAsync(&hflags, flags.ptr, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    KY_HIP(hipStreamSynchronize(st));
    KY_CHECK(!(hflags & 1u), KY_E215,
             "csr: a row's gene indices are not strictly increasing below the number of genes");
    KY_CHECK(!(hflags & 2u), KY_E215,
             "pca: the layer holds a value that is not finite or is outside the magnitudes pca's "
             "float32 products on a GPU keep exact (1e-10 to 1e10)");
    // Chunked columns.
    csr.chunk_rows = CHUNK_ROWS;
    csr.n_chunks = div_up(n, csr.chunk_rows);
    // Per wave: n_cols counters, 64 partial sums and two row offsets beyond its share of
    // the chunk's rows (chunk_rows offsets in all).
    const size_t fixed = sizeof(uint32_t) * csr.chunk_rows;
    const size_t limit = lds_limit(ctx);
    unsigned waves = std::min<unsigned>(
        16, std::max<unsigned>(1, (unsigned)((limit > fixed ? limit - fixed : 0) / (4u * csr.n_cols + 264))));
    size_t lds = sizeof(uint32_t) * ((size_t)waves * csr.n_cols + 64 * waves + csr.chunk_rows + 2 * waves);
    KY_CHECK(lds <= ctx.props.sharedMemPerBlock, KY_E215,
             "csr: too many genes for the device's column copy (" + std::to_string(csr.n_cols) + ")");
    csr.col_ptr.alloc((size_t)csr.n_chunks * (csr.n_cols + 1));
    csr.col_row.alloc(std::max<uint64_t>(csr.nnz, 1));
    csr.col_val.alloc(std::max<uint64_t>(csr.nnz, 1));
    build_columns<<<csr.n_chunks, 64 * waves, lds, st>>>(csr.indptr.ptr, csr.indices.ptr, csr.values.ptr, n,
                                                         csr.n_cols, csr.chunk_rows, csr.col_ptr.ptr,
                                                         csr.col_row.ptr, csr.col_val.ptr);
    KY_HIP(hipGetLastError());
    csr.col_order.alloc((size_t)csr.n_chunks * (csr.n_cols + 1));
    order_runs<<<csr.n_chunks, 1024, sizeof(uint32_t) * (csr.chunk_rows + 1), st>>>(csr.col_ptr.ptr, csr.n_cols,
                                                                                   csr.chunk_rows, csr.col_order.ptr);
    KY_HIP(hipGetLastError());
    KY_HIP(hipStreamSynchronize(st));
}

void csr_build(Csr &csr, const uint64_t *indptr, const uint32_t *indices, const float *values,
               const LayerMaps &maps) {
    Context &ctx = *csr.ctx;
    const uint64_t n = csr.n_rows;
    KY_CHECK(n >= 1, KY_E215, "csr: a layer without cells");
    KY_CHECK(indptr[0] == 0, KY_E215, "csr: indptr does not start at 0");
    for (uint64_t r = 0; r < n; r++)
        KY_CHECK(indptr[r + 1] >= indptr[r], KY_E215, "csr: indptr decreases at row " + std::to_string(r));
    csr.nnz = indptr[n];
    KY_CHECK(csr.nnz < (1ull << 32), KY_E505,
             "csr: a layer holds 2^32 or more values on one device; pass a smaller memory_limit");
    KY_CHECK(csr.n_cols >= 1, KY_E215, "csr: a layer without genes");
    csr.indptr.alloc(n + 1);
    csr.indices.alloc(std::max<uint64_t>(csr.nnz, 1));
    csr.values.alloc(std::max<uint64_t>(csr.nnz, 1));
    ctx.check_cancel();
    upload_parts(ctx, {{csr.indptr.ptr, indptr, sizeof(uint64_t) * (n + 1)},
                       {csr.indices.ptr, indices, sizeof(uint32_t) * csr.nnz},
                       {csr.values.ptr, values, sizeof(float) * csr.nnz}});
    csr_finish(csr, maps);
    ctx.advance(n);
}

// A pinned host buffer (hipHostMalloc), grown on demand.
template <typename T>
struct Pinned {
    T *ptr = nullptr;
    size_t n = 0;
    void reserve(size_t count) {
        if (count <= n) return;
        if (ptr) KY_HIP(hipHostFree(ptr));
        ptr = nullptr;
        n = 0;
        KY_HIP(hipHostMalloc(reinterpret_cast<void **>(&ptr), sizeof(T) * std::max<size_t>(count, 1),
                             hipHostMallocDefault));
        n = count;
    }
    ~Pinned() {
        if (ptr) (void)hipHostFree(ptr);
    }
};

// A device buffer grown to `count` keeping its first `keep` elements (the caller has
// synchronised every stream that uses it).
template <typename T>
static void grow_keep(DevBuf<T> &buf, size_t count, size_t keep, hipStream_t st) {
    if (count <= buf.n) return;
    DevBuf<T> bigger(count);
    if (keep) KY_HIP(hipMemcpyAsync(bigger.ptr, buf.ptr, sizeof(T) * keep, hipMemcpyDeviceToDevice, st));
    KY_HIP(hipStreamSynchronize(st));
    buf = std::move(bigger);
}

// The layer read block by block (a store's row blocks): the reader fills block b into
// one of two pinned buffer sets while the other set's copy to the device (its own
// stream) and the previous block's maps (the context's stream) run; the blocks' rows
// are appended to the resident CSR. The host never holds more than two blocks.
void csr_stream(Csr &csr, uint64_t block_rows, ky_block_reader reader, void *user, uint64_t nnz_hint,
                const LayerMaps &maps, ky_stream_info *info) {
    Context &ctx = *csr.ctx;
    const uint64_t n = csr.n_rows;
    KY_CHECK(n >= 1, KY_E215, "csr: a layer without cells");
    KY_CHECK(csr.n_cols >= 1, KY_E215, "csr: a layer without genes");
    KY_CHECK(block_rows >= 1, KY_E215, "csr: block_rows must be at least 1");
    KY_CHECK(reader != nullptr, KY_E215, "csr: reader is null");
    const double t_start = now_ms();
    double read_ms = 0.0;
    hipStream_t st = ctx.stream, cp = nullptr;
    KY_HIP(hipStreamCreateWithFlags(&cp, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        hipEvent_t e[2] = {nullptr, nullptr};
        ~StreamGuard() {
            if (s) {
                (void)hipStreamSynchronize(s);
                (void)hipStreamDestroy(s);
            }
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } guard{cp};
    KY_HIP(hipEventCreateWithFlags(&guard.e[0], hipEventDisableTiming));
    KY_HIP(hipEventCreateWithFlags(&guard.e[1], hipEventDisableTiming));
    Pinned<uint64_t> pin_ptr[2];
    Pinned<uint32_t> pin_idx[2];
    Pinned<float> pin_val[2];
    const uint64_t first_cap = std::max<uint64_t>(1024, std::min<uint64_t>(block_rows * csr.n_cols / 8, 1ull << 26));
    for (int k = 0; k < 2; k++) {
        pin_ptr[k].reserve(block_rows + 1);
        pin_idx[k].reserve(first_cap);
        pin_val[k].reserve(first_cap);
    }
    csr.indptr.alloc(n + 1);
    uint64_t total = 0, blocks = 0;
    bool used[2] = {false, false};
    for (uint64_t first = 0; first < n; first += block_rows, blocks++) {
        const int k = (int)(blocks % 2);
        const uint64_t cnt = std::min<uint64_t>(block_rows, n - first);
        if (ctx.cancelled.load()) {
            // Stop before the next block, with no copy from the pinned buffers in flight.
            KY_HIP(hipStreamSynchronize(cp));
            KY_HIP(hipStreamSynchronize(st));
            ctx.check_cancel();
        }
        if (used[k]) KY_HIP(hipEventSynchronize(guard.e[k]));  // the set's previous copy is done
        const double t0 = now_ms();
        int64_t got = reader(user, first, cnt, pin_ptr[k].ptr, pin_idx[k].ptr, pin_val[k].ptr, pin_idx[k].n);
        KY_CHECK(got >= 0, KY_E505,
                 "csr: the block reader failed at row " + std::to_string(first) +
                     (last_error_text().empty() ? std::string() : ": " + last_error_text()));
        if ((uint64_t)got > pin_idx[k].n) {
            pin_idx[k].reserve((uint64_t)got);
            pin_val[k].reserve((uint64_t)got);
            got = reader(user, first, cnt, pin_ptr[k].ptr, pin_idx[k].ptr, pin_val[k].ptr, pin_idx[k].n);
            KY_CHECK(got >= 0 && (uint64_t)got <= pin_idx[k].n, KY_E505,
                     "csr: the block reader failed at row " + std::to_string(first));
        }
        read_ms += now_ms() - t0;
        const uint64_t nnz = (uint64_t)got;
        uint64_t *ip = pin_ptr[k].ptr;
        KY_CHECK(ip[0] == 0 && ip[cnt] == nnz, KY_E215,
                 "csr: a block's indptr does not run from 0 to its number of values (rows " + std::to_string(first) +
                     ")");
        for (uint64_t r = 0; r < cnt; r++)
            KY_CHECK(ip[r + 1] >= ip[r], KY_E215, "csr: indptr decreases at row " + std::to_string(first + r));
        for (uint64_t r = 0; r <= cnt; r++) ip[r] += total;  // global offsets
        KY_CHECK(total + nnz < (1ull << 32), KY_E505,
                 "csr: a layer holds 2^32 or more values on one device; pass a smaller memory_limit");
        if (total + nnz > csr.values.n) {
            // Grow the resident arrays (first block: from the hint or an estimate).
            KY_HIP(hipStreamSynchronize(cp));
            KY_HIP(hipStreamSynchronize(st));
            uint64_t want = nnz_hint ? nnz_hint : (uint64_t)((double)(total + nnz) / (double)(first + cnt) * n * 1.05);
            want = std::max<uint64_t>(want, (total + nnz) + (total + nnz) / 4);
            grow_keep(csr.indices, want, total, st);
            grow_keep(csr.values, want, total, st);
        }
        // Copy the block (its indptr including the end offset) on the copy stream,
        // then the maps of its rows on the context's stream once it has arrived.
        KY_HIP(hipMemcpyAsync(csr.indptr.ptr + first, ip, sizeof(uint64_t) * (cnt + 1), hipMemcpyHostToDevice, cp));
        if (nnz) {
            KY_HIP(hipMemcpyAsync(csr.indices.ptr + total, pin_idx[k].ptr, sizeof(uint32_t) * nnz,
                                  hipMemcpyHostToDevice, cp));
            KY_HIP(hipMemcpyAsync(csr.values.ptr + total, pin_val[k].ptr, sizeof(float) * nnz,
                                  hipMemcpyHostToDevice, cp));
        }
        KY_HIP(hipEventRecord(guard.e[k], cp));
        used[k] = true;
        KY_HIP(hipStreamWaitEvent(st, guard.e[k], 0));
        normalize_log1p(st, csr.indptr.ptr + first, cnt, csr.values.ptr, maps.target_sum, maps.normalize, maps.log1p);
        total += nnz;
        ctx.advance(cnt);
    }
    KY_HIP(hipStreamSynchronize(cp));
    KY_HIP(hipStreamSynchronize(st));
    csr.nnz = total;
    csr_finish(csr, LayerMaps());
    if (info) {
        info->blocks = blocks;
        info->nnz = total;
        info->read_ms = read_ms;
        info->total_ms = now_ms() - t_start;
        info->bytes = (double)total * 8.0 + (double)(n + blocks) * 8.0;
    }
}

// The sum of n_parts arrays of `stride` values (part c at partial + c * stride) into
// out, by a fixed tree of groups of 64 parts (the same bits on every run). `tmp`
// holds at least stride * ceil(n_parts / 64) values.
void ordered_sum(hipStream_t st, const double *partial, uint32_t n_parts, uint64_t stride, double *tmp, double *out) {
    const uint32_t group = 64;
    while (n_parts > group) {
        const uint32_t n_groups = (n_parts + group - 1) / group;
        reduce_groups<<<div_up(stride * n_groups, 256), 256, 0, st>>>(partial, stride, n_parts, group, tmp);
        KY_HIP(hipGetLastError());
        partial = tmp;
        tmp = const_cast<double *>(partial) + stride * n_groups;  // the next level's output follows
        n_parts = n_groups;
    }
    reduce_chunks<<<div_up(stride, 256), 256, 0, st>>>(partial, stride, n_parts, out);
    KY_HIP(hipGetLastError());
}

// The next pass's operand from a resident f64 block q (genes x w, contiguous): its f32
// copy with rows ld floats apart (zero past w), and the rounded means bm[j] = sum over
// genes g in order of means[g] * f64(f32(q[g][j])) (one thread a column).
__global__ void block_f32(const double *__restrict__ q, uint32_t genes, uint32_t w, uint32_t ld,
                          float *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)genes * ld) return;
    const uint32_t g = (uint32_t)(i / ld), j = (uint32_t)(i % ld);
    out[i] = j < w ? (float)q[(uint64_t)g * w + j] : 0.0f;
}

// One block of 256 threads a column j: the column's terms means[g] * f64(f32(q[g][j]))
// staged in LDS by the block (RM_GENES at a time), then added in gene order by one
// thread.
constexpr uint32_t RM_GENES = 4096;
__global__ __launch_bounds__(256) void rounded_means(const double *__restrict__ q, const double *__restrict__ means,
                                                     uint32_t genes, uint32_t w, double *__restrict__ bm) {
    __shared__ double term[RM_GENES];
    const uint32_t j = blockIdx.x;
    double s = 0.0;
    for (uint32_t g0 = 0; g0 < genes; g0 += RM_GENES) {
        const uint32_t cnt = min(RM_GENES, genes - g0);
        for (uint32_t i = threadIdx.x; i < cnt; i += blockDim.x)
            term[i] = means[g0 + i] * (double)(float)q[(uint64_t)(g0 + i) * w + j];
        __syncthreads();
        if (threadIdx.x == 0)
            for (uint32_t i = 0; i < cnt; i++) s += term[i];
        __syncthreads();
    }
    if (threadIdx.x == 0) bm[j] = s;
}

// product[g][j] -= means[g] * sums[j] (the centring of X^T W).
__global__ void center_product(double *__restrict__ product, uint32_t genes, uint32_t w,
                               const double *__restrict__ means, const double *__restrict__ sums) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)genes * w) return;
    product[i] -= means[i / w] * sums[i % w];
}

// W (n x w floats, row stride ldw) is in ctx.wbuf. Computes X^T W into `product`
// (host, genes x w) and optionally the column sums of W into `sums` (host, w). With
// `resident`, both stay on the device (product in ctx.reduced, the sums at
// resident->sums) and nothing is downloaded.
struct ProductOnDevice {
    double *product = nullptr, *sums = nullptr;
};
static void transposed_product(Csr &csr, uint32_t w, uint32_t ldw, double *product, double *sums,
                               ProductOnDevice *resident = nullptr) {
    Context &ctx = *csr.ctx;
    hipStream_t st = ctx.stream;
    const size_t stride = (size_t)csr.n_cols * w;
    ctx.partials.reserve(stride * csr.n_chunks * 2);
    ctx.reduced.reserve(stride + w);
    // Without the LDS slice (more than 64 columns): 8 genes a block, and the blocks
    // resident on an XCD work on one or two chunks at a time, whose W rows stay in its L2.
    const uint32_t gpb = 8;
    const unsigned groups = div_up(csr.n_cols, gpb);
    const unsigned slots = div_up(csr.n_chunks, 8);
    const unsigned blocks = 8 * slots * groups;
    const bool lds16 = w <= 64 && ldw >= 64 && ldw % 4 == 0 && (size_t)csr.chunk_rows * 16 <= LDS_FLOATS;
    const bool lds8 = !lds16 && w <= 64 && ldw >= 64 && ldw % 4 == 0 && (size_t)csr.chunk_rows * 8 <= LDS_FLOATS;
    // The column sums of W: per chunk by the LDS-slice product (from its slice), else
    // per segment of SEG rows by column_sums; then the tree's levels (each at most 1/64
    // of the one before), then the total.
    const uint32_t sum_parts = lds16 || lds8 ? csr.n_chunks : div_up(csr.n_rows, SEG);
    const size_t sums_need = (size_t)sum_parts * w * 2 + 2 * w;
    const bool want_sums = sums || resident;
    if (want_sums) ctx.sums.reserve(sums_need);
    double *chunk_sums = want_sums && (lds16 || lds8) ? ctx.sums.ptr : nullptr;
    if (t_clock)
        t_clock->begin(1, st,
                       8.0 * csr.nnz + 4.0 * csr.n_chunks * (csr.n_cols + 1.0) + 4.0 * csr.n_rows * w +
                           8.0 * csr.n_chunks * csr.n_cols * w);
    if (lds16) {
        // W's chunk rows in LDS, 16 columns a slice.
        const uint32_t n_slices = div_up(w, 16);
        column_product_lds<16><<<8 * div_up(csr.n_chunks, 8) * n_slices, 1024, 0, st>>>(
            csr.col_ptr.ptr, csr.col_row.ptr, csr.col_val.ptr, csr.col_order.ptr, csr.n_cols, csr.n_chunks,
            csr.chunk_rows, csr.n_rows, ctx.wbuf.ptr, w, ldw, n_slices, ctx.partials.ptr, chunk_sums);
    } else if (lds8) {
        // W's chunk rows in LDS, 8 columns a slice.
        const uint32_t n_slices = div_up(w, 8);
        column_product_lds<8><<<8 * div_up(csr.n_chunks, 8) * n_slices, 1024, 0, st>>>(
            csr.col_ptr.ptr, csr.col_row.ptr, csr.col_val.ptr, csr.col_order.ptr, csr.n_cols, csr.n_chunks,
            csr.chunk_rows, csr.n_rows, ctx.wbuf.ptr, w, ldw, n_slices, ctx.partials.ptr, chunk_sums);
    } else if (w <= 64)
        column_product<1><<<blocks, 256, 0, st>>>(csr.col_ptr.ptr, csr.col_row.ptr, csr.col_val.ptr, csr.n_cols,
                                                  csr.n_chunks, csr.chunk_rows, csr.n_rows, ctx.wbuf.ptr, w, ldw, gpb,
                                                  ctx.partials.ptr);
    else if (w <= 128)
        column_product<2><<<blocks, 256, 0, st>>>(csr.col_ptr.ptr, csr.col_row.ptr, csr.col_val.ptr, csr.n_cols,
                                                  csr.n_chunks, csr.chunk_rows, csr.n_rows, ctx.wbuf.ptr, w, ldw, gpb,
                                                  ctx.partials.ptr);
    else
        column_product<4><<<blocks, 256, 0, st>>>(csr.col_ptr.ptr, csr.col_row.ptr, csr.col_val.ptr, csr.n_cols,
                                                  csr.n_chunks, csr.chunk_rows, csr.n_rows, ctx.wbuf.ptr, w, ldw, gpb,
                                                  ctx.partials.ptr);
    if (t_clock) t_clock->end(1, st);
    KY_HIP(hipGetLastError());
    ordered_sum(st, ctx.partials.ptr, csr.n_chunks, stride, ctx.partials.ptr + stride * csr.n_chunks,
                ctx.reduced.ptr);
    if (product) KY_HIP(hipMemcpyAsync(product, ctx.reduced.ptr, sizeof(double) * stride, hipMemcpyDeviceToHost, st));
    if (want_sums) {
        if (!chunk_sums) {
            column_sums<<<sum_parts, 64 * div_up(w, 64), 0, st>>>(ctx.wbuf.ptr, csr.n_rows, w, ldw, ctx.sums.ptr);
            KY_HIP(hipGetLastError());
        }
        double *tot = ctx.sums.ptr + sums_need - w;
        ordered_sum(st, ctx.sums.ptr, sum_parts, w, ctx.sums.ptr + (size_t)sum_parts * w, tot);
        KY_HIP(hipGetLastError());
        if (sums) KY_HIP(hipMemcpyAsync(sums, tot, sizeof(double) * w, hipMemcpyDeviceToHost, st));
        if (resident) resident->sums = tot;
    }
    if (resident) {
        resident->product = ctx.reduced.ptr;
        return;
    }
    KY_HIP(hipStreamSynchronize(st));
}

// Dense operands of the products keep their rows 64 floats (256 bytes: two whole
// 128-byte lines) apart or a multiple of that, so each gathered row touches exactly
// the lines it needs.
static uint32_t padded_row(uint32_t w) { return (w + 63) / 64 * 64; }

static void launch_row_project(Csr &csr, const float *M, uint32_t w, uint32_t ldm, const double *offsets, float *out,
                               uint32_t ldo) {
    Context &ctx = *csr.ctx;
    const unsigned rows_per_block = 4;
    const unsigned blocks = div_up(csr.n_rows, rows_per_block);
    KY_CHECK(w <= 256, KY_E215, "pca: more than 256 columns in one pass");
    if (t_clock)
        t_clock->begin(0, ctx.stream,
                       8.0 * csr.nnz + 8.0 * (csr.n_rows + 1.0) + 4.0 * csr.n_cols * w + 4.0 * csr.n_rows * w);
    // M's column slice in LDS: 16 columns a slice up to 2,048 genes. Past that the
    // gathers from L2 (row_project) beat 8-column slices (at 4M x 3,000: 13.5 against
    // 18.7 ms a launch; each entry is then streamed 8 times).
    const unsigned lds_rows_per_block = 8192;
    const unsigned lds_blocks = 8 * div_up(div_up(csr.n_rows, lds_rows_per_block), 8);
    if (w <= 64 && ldm >= 64 && ldm % 4 == 0 && (size_t)csr.n_cols * 16 <= LDS_FLOATS) {
        const uint32_t n_slices = div_up(w, 16);
        row_project_lds<16><<<lds_blocks * n_slices, 1024, 0, ctx.stream>>>(
            csr.indptr.ptr, csr.indices.ptr, csr.values.ptr, csr.n_rows, M, csr.n_cols, w, ldm, offsets, out, ldo,
            n_slices, lds_rows_per_block);
    } else if (w <= 64)
        row_project<1><<<blocks, 256, 0, ctx.stream>>>(csr.indptr.ptr, csr.indices.ptr, csr.values.ptr, csr.n_rows,
                                                       M, csr.n_cols, w, ldm, offsets, out, ldo);
    else if (w <= 128)
        row_project<2><<<blocks, 256, 0, ctx.stream>>>(csr.indptr.ptr, csr.indices.ptr, csr.values.ptr, csr.n_rows,
                                                       M, csr.n_cols, w, ldm, offsets, out, ldo);
    else
        row_project<4><<<blocks, 256, 0, ctx.stream>>>(csr.indptr.ptr, csr.indices.ptr, csr.values.ptr, csr.n_rows,
                                                       M, csr.n_cols, w, ldm, offsets, out, ldo);
    if (t_clock) t_clock->end(0, ctx.stream);
    KY_HIP(hipGetLastError());
}

void pca_sketch(Csr &csr, uint64_t seed, uint64_t cell_offset, uint32_t l, double *sketch, double *gene_sums,
                double *omega_sums) {
    Context &ctx = *csr.ctx;
    KY_CHECK(l >= 1 && l + 1 <= 256, KY_E215, "pca: sketch width out of range");
    const uint32_t w = l + 1, ldw = padded_row(w);
    ctx.wbuf.reserve((size_t)csr.n_rows * ldw + DENSE_PAD);
    uint64_t total = csr.n_rows * (uint64_t)w;
    omega_rows<<<div_up(total, 256), 256, 0, ctx.stream>>>(csr.n_rows, seed, cell_offset, l, ctx.wbuf.ptr, ldw);
    KY_HIP(hipGetLastError());
    std::vector<double> prod((size_t)csr.n_cols * w), sums(w);
    transposed_product(csr, w, ldw, prod.data(), sums.data());
    for (uint32_t g = 0; g < csr.n_cols; g++) {
        for (uint32_t j = 0; j < l; j++) sketch[(size_t)g * l + j] = prod[(size_t)g * w + j];
        gene_sums[g] = prod[(size_t)g * w + l];
    }
    for (uint32_t j = 0; j < l; j++) omega_sums[j] = sums[j];
}

void pca_product(Csr &csr, const double *q, uint32_t w, const double *block_means, double *product, double *sums) {
    Context &ctx = *csr.ctx;
    KY_CHECK(w >= 1 && w <= 256, KY_E215, "pca: block width out of range");
    const uint32_t ld = padded_row(w);
    std::vector<float> q32((size_t)csr.n_cols * ld, 0.0f);
    for (size_t g = 0; g < csr.n_cols; g++)
        for (size_t j = 0; j < w; j++) q32[g * ld + j] = (float)q[g * w + j];
    ctx.dense_f32.reserve(q32.size() + DENSE_PAD);
    ctx.dense_f64.reserve(w);
    KY_HIP(hipMemcpyAsync(ctx.dense_f32.ptr, q32.data(), sizeof(float) * q32.size(), hipMemcpyHostToDevice,
                          ctx.stream));
    KY_HIP(hipMemcpyAsync(ctx.dense_f64.ptr, block_means, sizeof(double) * w, hipMemcpyHostToDevice, ctx.stream));
    ctx.wbuf.reserve((size_t)csr.n_rows * ld + DENSE_PAD);
    launch_row_project(csr, ctx.dense_f32.ptr, w, ld, ctx.dense_f64.ptr, ctx.wbuf.ptr, ld);
    transposed_product(csr, w, ld, product, sums);
}

// pca_product over a resident block q (genes x w, f64, contiguous on the device) with
// the means on the device: the centred product (X - mu)^T (X Q32 - mu^T Q32) stays in
// ctx.reduced (genes x w); returns it. Queued on the stream, nothing synchronised.
static double *pca_product_resident(Csr &csr, const double *q, uint32_t w, const double *means) {
    Context &ctx = *csr.ctx;
    KY_CHECK(w >= 1 && w <= 256, KY_E215, "pca: block width out of range");
    hipStream_t st = ctx.stream;
    const uint32_t ld = padded_row(w);
    const uint64_t cells = (uint64_t)csr.n_cols * ld;
    ctx.dense_f32.reserve(cells + DENSE_PAD);
    ctx.dense_f64.reserve(w);
    block_f32<<<div_up(cells, 256), 256, 0, st>>>(q, csr.n_cols, w, ld, ctx.dense_f32.ptr);
    rounded_means<<<w, 256, 0, st>>>(q, means, csr.n_cols, w, ctx.dense_f64.ptr);
    KY_HIP(hipGetLastError());
    ctx.wbuf.reserve((size_t)csr.n_rows * ld + DENSE_PAD);
    launch_row_project(csr, ctx.dense_f32.ptr, w, ld, ctx.dense_f64.ptr, ctx.wbuf.ptr, ld);
    ProductOnDevice out;
    transposed_product(csr, w, ld, nullptr, nullptr, &out);
    center_product<<<div_up((uint64_t)csr.n_cols * w, 256), 256, 0, st>>>(out.product, csr.n_cols, w, means,
                                                                           out.sums);
    KY_HIP(hipGetLastError());
    return out.product;
}

void pca_project(Csr &csr, const double *v, uint32_t k, const double *mean_projection, const double *means,
                 float *host_embedding, Dense **embedding, double *squares, uint64_t *stored) {
    Context &ctx = *csr.ctx;
    KY_CHECK(k >= 1 && k <= 256, KY_E215, "pca: n_components out of range");
    const uint32_t ld = padded_row(k);
    std::vector<float> v32((size_t)csr.n_cols * ld, 0.0f);
    for (size_t g = 0; g < csr.n_cols; g++)
        for (size_t j = 0; j < k; j++) v32[g * ld + j] = (float)v[g * k + j];
    ctx.dense_f32.reserve(v32.size() + DENSE_PAD);
    ctx.dense_f64.reserve(k + csr.n_cols);
    hipStream_t st = ctx.stream;
    KY_HIP(hipMemcpyAsync(ctx.dense_f32.ptr, v32.data(), sizeof(float) * v32.size(), hipMemcpyHostToDevice, st));
    KY_HIP(hipMemcpyAsync(ctx.dense_f64.ptr, mean_projection, sizeof(double) * k, hipMemcpyHostToDevice, st));
    KY_HIP(hipMemcpyAsync(ctx.dense_f64.ptr + k, means, sizeof(double) * csr.n_cols, hipMemcpyHostToDevice, st));
    Dense *emb = new ky_dense();
    try {
        emb->ctx = &ctx;
        emb->rows = csr.n_rows;
        emb->cols = k;
        emb->data.alloc((size_t)csr.n_rows * k);
        launch_row_project(csr, ctx.dense_f32.ptr, k, ld, ctx.dense_f64.ptr, emb->data.ptr, k);
        // Squares and counts per gene.
        const size_t cells = (size_t)csr.n_chunks * csr.n_cols;
        ctx.partials.reserve(2 * cells);
        ctx.reduced.reserve(2 * (size_t)csr.n_cols);
        gene_squares<<<div_up(cells, 4), 256, 0, st>>>(csr.col_ptr.ptr, csr.col_val.ptr, csr.n_cols, csr.n_chunks,
                                                         ctx.dense_f64.ptr + k, ctx.partials.ptr,
                                                         ctx.partials.ptr + cells);
        KY_HIP(hipGetLastError());
        reduce_chunks<<<div_up(csr.n_cols, 256), 256, 0, st>>>(ctx.partials.ptr, csr.n_cols, csr.n_chunks,
                                                                ctx.reduced.ptr);
        reduce_chunks<<<div_up(csr.n_cols, 256), 256, 0, st>>>(ctx.partials.ptr + cells, csr.n_cols, csr.n_chunks,
                                                                ctx.reduced.ptr + csr.n_cols);
        KY_HIP(hipGetLastError());
        std::vector<double> counts(csr.n_cols);
        KY_HIP(hipMemcpyAsync(squares, ctx.reduced.ptr, sizeof(double) * csr.n_cols, hipMemcpyDeviceToHost, st));
        KY_HIP(hipMemcpyAsync(counts.data(), ctx.reduced.ptr + csr.n_cols, sizeof(double) * csr.n_cols,
                              hipMemcpyDeviceToHost, st));
        KY_HIP(hipStreamSynchronize(st));
        if (host_embedding) download(ctx, host_embedding, emb->data.ptr, sizeof(float) * csr.n_rows * k);
        for (uint32_t g = 0; g < csr.n_cols; g++) stored[g] = (uint64_t)counts[g];
    } catch (...) {
        delete static_cast<ky_dense *>(emb);
        throw;
    }
    if (embedding)
        *embedding = emb;
    else
        delete static_cast<ky_dense *>(emb);
}

size_t pca_device_bytes(uint64_t n, uint32_t g, uint64_t nnz, uint32_t k, uint32_t q) {
    auto b = [](size_t bytes) { return Arena::piece(std::max<size_t>(bytes, 1)); };
    const uint64_t chunks = (n + CHUNK_ROWS - 1) / CHUNK_ROWS;
    const size_t width = std::min<size_t>((size_t)k + PCA_OVERSAMPLES, g), w1 = width + 1;
    const size_t ld = padded_row((uint32_t)w1);
    const size_t cap = width * (q + 1);
    size_t total = 0;
    // The layer: rows (indptr, indices, values) and chunked columns with their run order.
    total += b(8 * (n + 1)) + 2 * b(4 * nnz) + 2 * b(4 * chunks * (g + 1)) + 2 * b(4 * nnz);
    // Pass buffers: W (n rows of ld floats), the chunks' f64 partials and their
    // reduction, the column sums, the dense operand.
    total += b(4 * (n * ld + DENSE_PAD)) + b(8 * g * w1 * chunks * 2) + b(8 * (g * w1 + w1)) + b(8 * 4096 * w1);
    total += b(4 * (g * ld + DENSE_PAD)) + b(8 * ((size_t)k + g + w1));
    // The basis and its products on the device, the dense steps' partials and work.
    total += 2 * b(8 * g * cap) + b(8 * ((g + 63) / 64) * cap * cap) + b(8 * cap * cap) + b(8 * g * cap);
    // The embedding.
    total += b(4 * n * k);
    return total;
}

void pca_release(Context &ctx) {
    ctx.partials.release();
    ctx.reduced.release();
    ctx.dense_f32.release();
    ctx.dense_f64.release();
    ctx.wbuf.release();
    ctx.sums.release();
}

// The whole pca(): PcaAccumulator's passes (pca.rs:302-862) with GPU folds.
void pca_run(Csr &csr, uint32_t n_components, uint64_t seed, uint32_t q, float *host_embedding, Dense **embedding,
             float *loadings_out, double *variance, double *variance_ratio, double *means_out, uint32_t *passes_out,
             ky_pca_timing *timing) {
    const double t_start = now_ms();
    double dense_ms = 0.0;
    SpmmClock clock;
    t_clock = &clock;
    struct Reset {
        ~Reset() { t_clock = nullptr; }
    } reset;
    const uint64_t n = csr.n_rows;
    const size_t genes = csr.n_cols;
    // check_cells (pca.rs:876-920).
    KY_CHECK(n >= 2, KY_E215,
             std::to_string(n) + " cells reach pca (the filters before it keep fewer than 2)");
    KY_CHECK(n_components >= 1 && n_components <= std::min<uint64_t>(n, genes), KY_E215,
             "pca: n_components " + std::to_string(n_components) + " is above min(cells, genes) = " +
                 std::to_string(std::min<uint64_t>(n, genes)));
    const size_t k = n_components;
    const size_t width = std::min<size_t>(k + PCA_OVERSAMPLES, genes);
    uint32_t pass_idx = 0;
    auto pass_time = [&](double t0) {
        if (timing && pass_idx < 16) timing->pass_ms[pass_idx] = now_ms() - t0;
        pass_idx++;
    };
    // Pass 0: the sketch.
    double t0 = now_ms();
    Mat sketch(genes, width);
    std::vector<double> sums(genes), omega_sums(width);
    Context &run = *csr.ctx;
    run.check_cancel();
    pca_sketch(csr, seed, 0, (uint32_t)width, sketch.v.data(), sums.data(), omega_sums.data());
    run.advance(n);
    pass_time(t0);
    double td = now_ms();
    std::vector<double> means(genes);
    for (size_t g = 0; g < genes; g++) {
        means[g] = sums[g] / (double)n;
        KY_CHECK(std::isfinite(means[g]), KY_E215,
                 "pca: gene " + std::to_string(g) + " holds a value that is not finite (NaN or infinity)");
    }
    for (size_t g = 0; g < genes; g++)
        for (size_t j = 0; j < width; j++) sketch.at(g, j) -= means[g] * omega_sums[j];
    // The basis and its products resident on the device only, where the dense steps'
    // large products run; so is each Krylov block (dev_basis.work_r) from its
    // orthonormalisation to the next pass, and the pass's product.
    DeviceBasis dev_basis(*csr.ctx, genes, width * (q + 1));
    dev_basis.work_r.reserve(genes * width);  // every block fits: no reallocation under queued copies
    DevBuf<double> means_dev(genes);
    upload_parts(run, {CopyPart{means_dev.ptr, means.data(), sizeof(double) * genes}});
    dev_basis.put_block(sketch);
    size_t block_cols = extend_basis_resident(dev_basis, sketch.cols, block_scale_of(sketch));
    std::vector<double> eigenvalues;
    Mat loadings;
    bool finished = false;
    // mu^T M for the f32-rounded operand M the device multiplies by: each row's
    // projection x M32 - mu^T M32 is then (x - mu) M32 to f64 accuracy whatever the
    // means (with mu^T M in f64 the rounding of M would leave mu (M32 - M) in every
    // row, as large as the data's spread at means 1e6 times it).
    auto rounded_means_of = [&](const Mat &b) {
        std::vector<double> bm(b.cols, 0.0);
        for (size_t g = 0; g < genes; g++)
            for (size_t j = 0; j < b.cols; j++) bm[j] += means[g] * (double)(float)b.at(g, j);
        return bm;
    };
    auto finish_basis = [&]() {
        rayleigh_ritz(dev_basis, k, eigenvalues, loadings);
        finished = true;
    };
    if (block_cols == 0) finish_basis();
    dense_ms += now_ms() - td;
    uint32_t passes = 0;
    for (uint32_t p = 1; p <= q && !finished; p++) {
        const size_t w = block_cols;
        t0 = now_ms();
        run.check_cancel();
        const double *product = pca_product_resident(csr, dev_basis.work_r.ptr, (uint32_t)w, means_dev.ptr);
        KY_HIP(hipStreamSynchronize(run.stream));
        run.advance(n);
        pass_time(t0);
        passes++;
        td = now_ms();
        dev_basis.append_resident(dev_basis.work_r.ptr, product, w);
        if (p == q) {
            finish_basis();
        } else {
            dev_basis.take_block(product, w);
            block_cols = extend_basis_resident(dev_basis, w, dev_basis.block_norm(w));
            if (block_cols == 0) finish_basis();
        }
        dense_ms += now_ms() - td;
    }
    if (!finished) {
        td = now_ms();
        finish_basis();
        dense_ms += now_ms() - td;
    }
    // mean_projection = mu^T V32 (V as the device rounds it, above).
    const std::vector<double> mean_projection = rounded_means_of(loadings);
    std::vector<double> squares(genes);
    std::vector<uint64_t> stored(genes);
    t0 = now_ms();
    run.check_cancel();
    pca_project(csr, loadings.v.data(), (uint32_t)k, mean_projection.data(), means.data(), host_embedding, embedding,
                squares.data(), stored.data());
    run.advance(n);
    pass_time(t0);
    // Total variance (pca.rs:776-791) and finish (pca.rs:804-845).
    double total = 0.0;
    for (size_t g = 0; g < genes; g++)
        total += squares[g] + (double)(n - stored[g]) * means[g] * means[g];
    const double total_variance = total / ((double)n - 1.0);
    const bool constant = total_variance == 0.0;
    for (size_t j = 0; j < k; j++) {
        variance[j] = constant ? 0.0 : eigenvalues[j] / ((double)n - 1.0);
        variance_ratio[j] = variance[j] / total_variance;
    }
    if (loadings_out)
        for (size_t i = 0; i < genes * k; i++) loadings_out[i] = (float)loadings.v[i];
    if (means_out)
        for (size_t g = 0; g < genes; g++) means_out[g] = means[g];
    if (passes_out) *passes_out = passes;
    if (timing) {
        timing->dense_ms = dense_ms;
        timing->total_ms = now_ms() - t_start;
        timing->n_passes = pass_idx;
        for (int kind = 0; kind < 2; kind++) {
            timing->spmm_ms[kind] = clock.total(kind);
            timing->spmm_launches[kind] = (uint32_t)clock.ev[kind].size();
            timing->spmm_bytes[kind] = clock.bytes[kind];
        }
    }
}

}  // namespace ky
