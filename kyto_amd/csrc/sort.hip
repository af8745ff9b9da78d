// Device sorting and scanning primitives of the library (no external library):
//
// - sort_pairs: a stable LSD radix sort of (u32 key, u32 value) pairs by the low
//   end_bit bits of the key, 8 bits a pass. Each pass: a histogram of the digit per
//   tile of 4,096 keys (LDS counters), an exclusive scan of the digit-major counts
//   (digit d's tiles in order, then digit d + 1's), and a scatter in which every tile
//   walks its keys in order, 256 at a time: a key's place is its digit's offset for
//   the tile, plus the keys of that digit earlier in the tile (earlier rounds, earlier
//   wavefronts of this round, and earlier lanes of its own wavefront, found with 8
//   ballots). Equal keys keep their order, so every run gives the same permutation.
// - scan_counts: u32 counts into u64 exclusive offsets.
//
// The temporary memory both need is known from n alone (sort_pairs_bytes,
// scan_counts_bytes), so a caller can carve it from a fixed scratch region.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "common.hpp"
#include "sort.hpp"

namespace ky {

constexpr unsigned RS_THREADS = 256;
constexpr unsigned RS_ITEMS = 16;
constexpr unsigned RS_TILE = RS_THREADS * RS_ITEMS;
constexpr unsigned RS_BINS = 256;

static unsigned rs_tiles(size_t n) { return (unsigned)((n + RS_TILE - 1) / RS_TILE); }

// Digit counts of one tile, digit-major: hist[d * tiles + tile].
__global__ __launch_bounds__(RS_THREADS) void radix_hist(const uint32_t *__restrict__ keys, size_t n, unsigned shift,
                                                         uint32_t mask, unsigned tiles, uint32_t *__restrict__ hist) {
    __shared__ uint32_t cnt[RS_BINS];
    const unsigned t = threadIdx.x;
    cnt[t] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * RS_TILE;
    for (unsigned j = 0; j < RS_ITEMS; j++) {
        const size_t i = base + (size_t)j * RS_THREADS + t;
        if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & mask], 1u);
    }
    __syncthreads();
    hist[(size_t)t * tiles + blockIdx.x] = cnt[t];
}

// ------------------------------------------------- the digit counts' scan ----
//
// In-place exclusive scan of m u32 values (m a multiple of 256): per tile of HS_TILE
// values its sum, then one block scans the tiles' sums, then every tile scans itself
// from its sum's offset. Each thread holds 16 consecutive values (four 16-byte loads).
constexpr unsigned HS_THREADS = 256, HS_ITEMS = 16, HS_TILE = HS_THREADS * HS_ITEMS;

// Exclusive scan of one value a thread over the block; *total gets the block's sum.
// Called by every thread of the block.
__device__ uint32_t block_scan(uint32_t x, uint32_t *total) {
    __shared__ uint32_t wsum[HS_THREADS / 64];
    const unsigned lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t s = x;
#pragma unroll
    for (unsigned o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(s, o);
        if (lane >= o) s += y;
    }
    if (lane == 63) wsum[wave] = s;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (unsigned w = 0; w < HS_THREADS / 64; w++) {
        if (w < wave) before += wsum[w];
        all += wsum[w];
    }
    __syncthreads();
    *total = all;
    return before + s - x;
}

// The thread's 16 values of tile blockIdx.x (zeros past m; m is a multiple of 16).
__device__ __forceinline__ void tile_items(const uint32_t *v, size_t m, uint32_t (&x)[HS_ITEMS]) {
    const size_t i = (size_t)blockIdx.x * HS_TILE + (size_t)threadIdx.x * HS_ITEMS;
#pragma unroll
    for (unsigned q = 0; q < HS_ITEMS / 4; q++) {
        const uint4 u = i < m ? reinterpret_cast<const uint4 *>(v + i)[q] : make_uint4(0, 0, 0, 0);
        x[4 * q] = u.x;
        x[4 * q + 1] = u.y;
        x[4 * q + 2] = u.z;
        x[4 * q + 3] = u.w;
    }
}

__global__ __launch_bounds__(HS_THREADS) void tile_sums_u32(const uint32_t *__restrict__ v, size_t m,
                                                           uint32_t *__restrict__ sums) {
    uint32_t x[HS_ITEMS];
    tile_items(v, m, x);
    uint32_t s = 0;
#pragma unroll
    for (unsigned k = 0; k < HS_ITEMS; k++) s += x[k];
    uint32_t total;
    block_scan(s, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// One block: the nb tile sums into exclusive offsets, in place.
__global__ __launch_bounds__(HS_THREADS) void scan_sums_u32(uint32_t *__restrict__ sums, size_t nb) {
    uint32_t carry = 0;
    for (size_t r = 0; r < nb; r += HS_THREADS) {
        const size_t i = r + threadIdx.x;
        const uint32_t x = i < nb ? sums[i] : 0u;
        uint32_t total;
        const uint32_t ex = block_scan(x, &total);
        if (i < nb) sums[i] = carry + ex;
        carry += total;
    }
}

__global__ __launch_bounds__(HS_THREADS) void tile_scan_u32(uint32_t *__restrict__ v, size_t m,
                                                           const uint32_t *__restrict__ sums) {
    uint32_t x[HS_ITEMS];
    tile_items(v, m, x);
    uint32_t s = 0;
#pragma unroll
    for (unsigned k = 0; k < HS_ITEMS; k++) s += x[k];
    uint32_t total;
    uint32_t run = sums[blockIdx.x] + block_scan(s, &total);
    const size_t i = (size_t)blockIdx.x * HS_TILE + (size_t)threadIdx.x * HS_ITEMS;
    if (i >= m) return;
#pragma unroll
    for (unsigned k = 0; k < HS_ITEMS; k++) {
        const uint32_t t = x[k];
        x[k] = run;
        run += t;
    }
#pragma unroll
    for (unsigned q = 0; q < HS_ITEMS / 4; q++)
        reinterpret_cast<uint4 *>(v + i)[q] = make_uint4(x[4 * q], x[4 * q + 1], x[4 * q + 2], x[4 * q + 3]);
}

// The stable scatter of one tile (see the file comment). offs: the scanned counts.
__global__ __launch_bounds__(RS_THREADS) void radix_scatter(const uint32_t *__restrict__ kin,
                                                            const uint32_t *__restrict__ vin,
                                                            uint32_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                                            size_t n, unsigned shift, uint32_t mask, unsigned tiles,
                                                            const uint32_t *__restrict__ offs) {
    __shared__ uint32_t base[RS_BINS];
    __shared__ uint32_t wcnt[RS_THREADS / 64][RS_BINS];
    const unsigned t = threadIdx.x, lane = t & 63, wave = t >> 6;
    base[t] = offs[(size_t)t * tiles + blockIdx.x];
#pragma unroll
    for (unsigned w = 0; w < RS_THREADS / 64; w++) wcnt[w][t] = 0;
    __syncthreads();
    const size_t tile0 = (size_t)blockIdx.x * RS_TILE;
    for (unsigned j = 0; j < RS_ITEMS; j++) {
        const size_t i = tile0 + (size_t)j * RS_THREADS + t;
        const bool valid = i < n;
        const uint32_t key = valid ? kin[i] : 0u, val = valid ? vin[i] : 0u;
        const unsigned d = (key >> shift) & mask;
        // The lanes of this wavefront holding a key of the same digit.
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const unsigned rank = (unsigned)__popcll(peers & ((1ull << lane) - 1ull));
        if (valid && rank == 0) wcnt[wave][d] = (unsigned)__popcll(peers);
        __syncthreads();
        if (valid) {
            unsigned at = base[d] + rank;
            for (unsigned w = 0; w < wave; w++) at += wcnt[w][d];
            kout[at] = key;
            vout[at] = val;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (unsigned w = 0; w < RS_THREADS / 64; w++) {
            add += wcnt[w][t];
            wcnt[w][t] = 0;
        }
        base[t] += add;
        __syncthreads();
    }
}

size_t sort_pairs_bytes(size_t n) {
    // Alternate keys and values, the digit counts, then their tiles' sums (each piece
    // 256-byte aligned).
    const size_t a = (sizeof(uint32_t) * n + 255) & ~size_t(255);
    const size_t m = (size_t)RS_BINS * std::max<size_t>(rs_tiles(n), 1);
    const size_t h = (sizeof(uint32_t) * m + 255) & ~size_t(255);
    return 2 * a + h + sizeof(uint32_t) * ((m + HS_TILE - 1) / HS_TILE);
}

void sort_pairs(hipStream_t st, const uint32_t *keys_in, uint32_t *keys_out, const uint32_t *vals_in,
                uint32_t *vals_out, size_t n, unsigned end_bit, void *tmp) {
    if (n == 0) return;
    KY_CHECK(n < (1ull << 32), KY_E215, "sort: 2^32 or more pairs");
    end_bit = std::max(1u, std::min(end_bit, 32u));
    const unsigned passes = (end_bit + 7) / 8, tiles = rs_tiles(n);
    const size_t a = (sizeof(uint32_t) * n + 255) & ~size_t(255);
    uint32_t *alt_k = static_cast<uint32_t *>(tmp);
    uint32_t *alt_v = reinterpret_cast<uint32_t *>(static_cast<char *>(tmp) + a);
    uint32_t *hist = reinterpret_cast<uint32_t *>(static_cast<char *>(tmp) + 2 * a);
    const size_t m = (size_t)RS_BINS * tiles, n_sums = (m + HS_TILE - 1) / HS_TILE;
    uint32_t *sums = reinterpret_cast<uint32_t *>(static_cast<char *>(tmp) + 2 * a +
                                                  ((sizeof(uint32_t) * m + 255) & ~size_t(255)));
    // Ping-pong so that the last pass writes the outputs.
    const uint32_t *sk = keys_in, *sv = vals_in;
    for (unsigned p = 0; p < passes; p++) {
        const bool to_out = ((passes - 1 - p) % 2) == 0;
        uint32_t *dk = to_out ? keys_out : alt_k, *dv = to_out ? vals_out : alt_v;
        const unsigned shift = 8 * p;
        const unsigned bits = std::min(8u, end_bit - shift);
        const uint32_t mask = bits == 32 ? 0xFFFFFFFFu : ((1u << bits) - 1u);
        radix_hist<<<tiles, RS_THREADS, 0, st>>>(sk, n, shift, mask, tiles, hist);
        tile_sums_u32<<<n_sums, HS_THREADS, 0, st>>>(hist, m, sums);
        scan_sums_u32<<<1, HS_THREADS, 0, st>>>(sums, n_sums);
        tile_scan_u32<<<n_sums, HS_THREADS, 0, st>>>(hist, m, sums);
        radix_scatter<<<tiles, RS_THREADS, 0, st>>>(sk, sv, dk, dv, n, shift, mask, tiles, hist);
        KY_HIP(hipGetLastError());
        sk = dk;
        sv = dv;
    }
}

void sort_pairs(hipStream_t st, const uint32_t *keys_in, uint32_t *keys_out, const uint32_t *vals_in,
                uint32_t *vals_out, size_t n, unsigned end_bit, DevBuf<unsigned char> &tmp) {
    tmp.reserve(std::max<size_t>(sort_pairs_bytes(n), 1));
    sort_pairs(st, keys_in, keys_out, vals_in, vals_out, n, end_bit, tmp.ptr);
}

// ----------------------------------------------------------------------- scan ----

constexpr unsigned SC_TILE = 4096;

// Per tile of SC_TILE counts: their sum (u64), in tile order.
__global__ __launch_bounds__(256) void count_tile_sums(const uint32_t *__restrict__ counts, size_t n,
                                                       uint64_t *__restrict__ sums) {
    __shared__ uint64_t part[256];
    const size_t base = (size_t)blockIdx.x * SC_TILE;
    uint64_t s = 0;
    for (unsigned j = threadIdx.x; j < SC_TILE; j += 256)
        if (base + j < n) s += counts[base + j];
    part[threadIdx.x] = s;
    __syncthreads();
    for (unsigned off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) part[threadIdx.x] += part[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = part[0];
}

// Exclusive scan of the tile sums (one block, sequential runs per thread).
__global__ __launch_bounds__(256) void scan_tile_sums(uint64_t *__restrict__ sums, size_t m) {
    __shared__ uint64_t part[256];
    const unsigned t = threadIdx.x;
    const size_t per = (m + 255) / 256;
    const size_t lo = std::min(m, (size_t)t * per), hi = std::min(m, lo + per);
    uint64_t s = 0;
    for (size_t i = lo; i < hi; i++) s += sums[i];
    part[t] = s;
    __syncthreads();
    for (unsigned off = 1; off < 256; off <<= 1) {
        const uint64_t add = t >= off ? part[t - off] : 0ull;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    uint64_t run = t ? part[t - 1] : 0ull;
    for (size_t i = lo; i < hi; i++) {
        const uint64_t x = sums[i];
        sums[i] = run;
        run += x;
    }
}

// Offsets within each tile (a block scan of its counts), plus the tile's offset;
// offsets[n] = the total.
__global__ __launch_bounds__(256) void count_offsets(const uint32_t *__restrict__ counts, size_t n,
                                                     const uint64_t *__restrict__ tile_offs,
                                                     uint64_t *__restrict__ offsets) {
    __shared__ uint64_t part[256];
    const unsigned t = threadIdx.x;
    const size_t base = (size_t)blockIdx.x * SC_TILE;
    constexpr unsigned PER = SC_TILE / 256;
    const size_t lo = base + (size_t)t * PER;
    uint64_t s = 0;
    for (unsigned j = 0; j < PER; j++)
        if (lo + j < n) s += counts[lo + j];
    part[t] = s;
    __syncthreads();
    for (unsigned off = 1; off < 256; off <<= 1) {
        const uint64_t add = t >= off ? part[t - off] : 0ull;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    uint64_t run = tile_offs[blockIdx.x] + (t ? part[t - 1] : 0ull);
    for (unsigned j = 0; j < PER; j++) {
        if (lo + j > n) break;
        offsets[lo + j] = run;
        if (lo + j < n) run += counts[lo + j];
    }
}

size_t scan_counts_bytes(size_t n) { return sizeof(uint64_t) * (n / SC_TILE + 1); }

void scan_counts(hipStream_t st, const uint32_t *counts, size_t n, uint64_t *offsets, void *tmp) {
    const unsigned tiles = (unsigned)(n / SC_TILE + 1);  // covers offsets[0..n]
    uint64_t *sums = static_cast<uint64_t *>(tmp);
    count_tile_sums<<<tiles, 256, 0, st>>>(counts, n, sums);
    scan_tile_sums<<<1, 256, 0, st>>>(sums, tiles);
    count_offsets<<<tiles, 256, 0, st>>>(counts, n, sums, offsets);
    KY_HIP(hipGetLastError());
}

void scan_counts(hipStream_t st, uint32_t *counts, size_t n, uint64_t *offsets, DevBuf<unsigned char> &tmp) {
    tmp.reserve(std::max<size_t>(scan_counts_bytes(n), 1));
    scan_counts(st, counts, n, offsets, tmp.ptr);
}

}  // namespace ky

namespace ky {

__global__ void fill_words_kernel(uint32_t *__restrict__ dst, size_t count, uint32_t value) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < count) dst[i] = value;
}

void fill_words(hipStream_t st, void *dst, size_t count, uint32_t value) {
    if (count == 0) return;
    fill_words_kernel<<<div_up(count, 256), 256, 0, st>>>(static_cast<uint32_t *>(dst), count, value);
    KY_HIP(hipGetLastError());
}

__global__ void iota_words_kernel(uint32_t *__restrict__ dst, size_t count, uint32_t first) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < count) dst[i] = first + (uint32_t)i;
}

void iota_words(hipStream_t st, uint32_t *dst, size_t count, uint32_t first) {
    if (count == 0) return;
    iota_words_kernel<<<div_up(count, 256), 256, 0, st>>>(dst, count, first);
    KY_HIP(hipGetLastError());
}

}  // namespace ky
