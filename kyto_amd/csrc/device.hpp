// Device helpers shared by the kernels (wave64, gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ky {

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// Order of one wavefront's LDS accesses made visible to its other lanes.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// As wave_sync, for device-memory writes of one lane read by another lane of the
// same wavefront (waits for the writes to complete).
__device__ __forceinline__ void wave_sync_memory() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Lanes below this one whose bit is set in mask.
__device__ __forceinline__ unsigned mbcnt(uint64_t mask) {
    unsigned lo = __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u);
    return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), lo);
}

// Broadcast lane `src` (uniform) of v to every lane.
__device__ __forceinline__ uint32_t bcast_u32(uint32_t v, int src) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, src);
}
__device__ __forceinline__ float bcast_f32(float v, int src) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), src));
}

// f32 -> u32 whose unsigned order is the float order (negative values included).
__device__ __forceinline__ uint32_t order_f32(float x) {
    uint32_t b = __float_as_uint(x);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float unorder_f32(uint32_t u) {
    uint32_t b = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    return __uint_as_float(b);
}

// kyto's exact key (kernels.rs:235-247, pair_sum :422-448, combine_lanes :290-299):
// component j's f64 difference squared into lane j % 8, then the fixed tree. This
// translation unit is compiled without contraction, so no product is fused.
// The components go eight at a time so that each lane sum stays in a register (a
// lane index that is not a compile-time constant would put l in scratch memory).
__device__ __forceinline__ double exact_sq_distance(const float *a, const float *b, unsigned d) {
    double l[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (unsigned j0 = 0; j0 < d; j0 += 8) {
#pragma unroll
        for (unsigned u = 0; u < 8; u++) {
            if (j0 + u < d) {
                double t = (double)a[j0 + u] - (double)b[j0 + u];
                l[u] += t * t;
            }
        }
    }
    double h0 = l[0] + l[4], h1 = l[1] + l[5], h2 = l[2] + l[6], h3 = l[3] + l[7];
    return (h0 + h2) + (h1 + h3);
}

// kyto's cosine distance from a dot product and the two row norms
// (crates/kyto-exec/src/neighbors.rs:440-445): 1 when either norm is 0, else
// 1 - dot / (|q| |p|) clipped to [0, 2], in f64.
__device__ __forceinline__ double cosine_key(double dot, double left, double right) {
    if (left == 0.0 || right == 0.0) return 1.0;
    const double v = 1.0 - dot / (left * right);
    return v < 0.0 ? 0.0 : (v > 2.0 ? 2.0 : v);
}

// Ascending bitonic sort of N (power of two, >= 2) (key, row) pairs in LDS by one
// wavefront, ties broken by row.
template <typename K>
__device__ __forceinline__ void wave_bitonic(K *keys, uint32_t *rows, unsigned N) {
    const unsigned lane = lane_id();
    for (unsigned size = 2; size <= N; size <<= 1) {
        for (unsigned stride = size >> 1; stride > 0; stride >>= 1) {
            for (unsigned i = lane; i < N / 2; i += 64) {
                unsigned pos = 2 * stride * (i / stride) + (i % stride);
                unsigned par = pos + stride;
                bool asc = (pos & size) == 0;
                K ka = keys[pos], kb = keys[par];
                uint32_t ra = rows[pos], rb = rows[par];
                bool b_less = kb < ka || (kb == ka && rb < ra);
                if (b_less == asc) {
                    keys[pos] = kb;
                    keys[par] = ka;
                    rows[pos] = rb;
                    rows[par] = ra;
                }
            }
            wave_sync();
        }
    }
}

// Ascending sort of 64 (key, row) pairs held one per lane, in registers: lane l ends
// with the l-th smallest pair, ties broken by row. The bitonic network of the LDS sort
// above, each compare-exchange done by a lane and its partner lane.
__device__ __forceinline__ void wave_sort64(unsigned long long &key, uint32_t &row) {
    const unsigned lane = lane_id();
#pragma unroll
    for (unsigned size = 2; size <= 64; size <<= 1) {
#pragma unroll
        for (unsigned stride = size >> 1; stride > 0; stride >>= 1) {
            const unsigned long long ok = __shfl_xor(key, stride);
            const uint32_t orow = __shfl_xor(row, stride);
            const bool other_less = ok < key || (ok == key && orow < row);
            // The lower position of an ascending pair keeps the smaller, and so on.
            const bool keep_small = ((lane & stride) == 0) == ((lane & size) == 0);
            if (other_less == keep_small) {
                key = ok;
                row = orow;
            }
        }
    }
}

// Compact query slot `ql`'s buffer (count entries) to its `keep` smallest (key, row)
// pairs, sorted ascending; when it held at least `keep`, the keep-th pair becomes the
// threshold. Called by the whole wavefront.
template <typename K, int CAP>
__device__ void compact(K *bkey, uint32_t *brow, uint32_t *bcnt, K *tkey, uint32_t *trow, unsigned keep) {
    const unsigned lane = lane_id();
    const unsigned cnt = *bcnt;
    for (unsigned i = lane; i < CAP; i += 64)
        if (i >= cnt) {
            bkey[i] = ~K(0);
            brow[i] = ~0u;
        }
    wave_sync();
    wave_bitonic<K>(bkey, brow, CAP);
    if (lane == 0) {
        if (cnt >= keep) {
            *tkey = bkey[keep - 1];
            *trow = brow[keep - 1];
        }
        *bcnt = min(cnt, keep);
    }
    wave_sync();
}

}  // namespace ky
