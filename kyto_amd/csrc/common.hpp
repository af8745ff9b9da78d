// Shared plumbing of libkyto_gpu.so: the context, device buffers, error reporting.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <mutex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/kyto_gpu.h"

namespace ky {

// A failure carrying a kyto error code; converted to a status at the C ABI.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &message);
// The calling thread's last error message (set by a failed C-ABI call).
std::string last_error_text();

#define KY_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw ::ky::Error(KY_E505, std::string("device failure in ") + #call + ": " + \
                                           hipGetErrorString(e_));                        \
    } while (0)

#define KY_CHECK(cond, code, msg)                   \
    do {                                            \
        if (!(cond)) throw ::ky::Error(code, msg);  \
    } while (0)

// The caching allocator under DevBuf (memory.cpp).
void *device_alloc(size_t bytes);
void device_free(void *p);
void device_cache_trim();

// A device allocation owned by one object.
template <typename T>
struct DevBuf {
    T *ptr = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), n(o.n) { o.ptr = nullptr; o.n = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); ptr = o.ptr; n = o.n; o.ptr = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) ptr = static_cast<T *>(device_alloc(sizeof(T) * count));
    }
    // Grow only (keeps the allocation when it is already large enough).
    void reserve(size_t count) {
        if (count > n) alloc(count);
    }
    void release() {
        if (ptr) device_free(ptr);
        ptr = nullptr;
        n = 0;
    }
    size_t bytes() const { return sizeof(T) * n; }
};

// A bump allocator over one device region (a call's scratch): pieces taken in
// order, each rounded up to 256 bytes, released back to a mark. The same sequence of
// takes on a SizeArena (no memory) gives the region's size before it is allocated.
struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0, high = 0;
    Arena() = default;
    Arena(void *b, size_t c) : base(static_cast<char *>(b)), cap(c) {}
    static size_t piece(size_t bytes) { return (bytes + 255) & ~size_t(255); }
    template <typename T>
    T *take(size_t count) {
        const size_t bytes = piece(sizeof(T) * count);
        if (used + bytes > cap)
            throw Error(KY_E505, "device scratch: a call needed more than the " + std::to_string(cap) +
                                     " bytes planned for it");
        T *p = reinterpret_cast<T *>(base + used);
        used += bytes;
        if (used > high) high = used;
        return p;
    }
    size_t mark() const { return used; }
    void release(size_t m) { used = m; }
};

// The sizes of an Arena's takes, for planning (the pointers are null).
struct SizeArena {
    size_t used = 0, high = 0;
    template <typename T>
    T *take(size_t count) {
        used += Arena::piece(sizeof(T) * count);
        if (used > high) high = used;
        return nullptr;
    }
    size_t mark() const { return used; }
    void release(size_t m) { used = m; }
};

// Fill `count` 32-bit words with `value` (a kernel: usable where a memset is not).
void fill_words(hipStream_t st, void *dst, size_t count, uint32_t value);
// dst[i] = first + i for i < count.
void iota_words(hipStream_t st, uint32_t *dst, size_t count, uint32_t first);

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipDeviceProp_t props{};
    int n_cu = 0;
    // Scratch reused between calls (grown on demand).
    DevBuf<double> partials;
    DevBuf<double> reduced;
    DevBuf<float> dense_f32;
    DevBuf<double> dense_f64;
    DevBuf<float> wbuf;
    DevBuf<double> sums;
    // Pinned staging for downloads (memory.cpp) and the events of its two halves.
    void *staging = nullptr;
    hipEvent_t staged[2] = {nullptr, nullptr};
    // Events for kernel timings, made when the context opens.
    hipEvent_t timer[2] = {nullptr, nullptr};
    // The run the context's calls belong to (kyto's RunControl, run_control.rs:1-32): a
    // cancel flag any thread may set, checked between passes, blocks and rounds of
    // queries, and the cell rows computed (summed over the passes).
    std::atomic<bool> cancelled{false};
    std::atomic<uint64_t> rows_done{0}, rows_total{0};
    void check_cancel() const {
        if (cancelled.load(std::memory_order_relaxed))
            throw Error(KY_E503, "the run was cancelled after " + std::to_string(rows_done.load()) + " of " +
                                     std::to_string(rows_total.load()) + " cell rows were streamed");
    }
    void advance(uint64_t rows) { rows_done.fetch_add(rows, std::memory_order_relaxed); }
    ~Context() {
        if (staged[0]) (void)hipEventDestroy(staged[0]);
        if (staged[1]) (void)hipEventDestroy(staged[1]);
        if (timer[0]) (void)hipEventDestroy(timer[0]);
        if (timer[1]) (void)hipEventDestroy(timer[1]);
        if (staging) (void)hipHostFree(staging);
    }
};

// Device-to-host copy of `bytes` into pageable memory through the pinned staging
// buffer; returns once the data is in dst.
void download(Context &ctx, void *dst, const void *src, size_t bytes);

// One device-to-host copy of several arrays, pipelined as one stream of chunks.
struct CopyPart {
    void *dst;
    const void *src;
    size_t bytes;
};
void download_parts(Context &ctx, const std::vector<CopyPart> &parts);
// Host-to-device copies of several arrays (dst on the device, src in pageable host
// memory) through the same staging buffer; returns once the data is on the device.
void upload_parts(Context &ctx, const std::vector<CopyPart> &parts);

// f(t) for t in [0, tasks) on the persistent host worker pool (dense.cpp).
void host_parallel(size_t tasks, const std::function<void(size_t)> &f);
// Threads the pool's runs use (0: every worker); results do not depend on it.
void host_threads(unsigned threads);
unsigned host_threads_active();

struct Csr {
    Context *ctx = nullptr;
    uint64_t n_rows = 0;
    uint32_t n_cols = 0;
    uint64_t nnz = 0;
    DevBuf<uint64_t> indptr;   // n_rows + 1
    DevBuf<uint32_t> indices;  // nnz
    DevBuf<float> values;      // nnz
    // Chunked column-major copy: for chunk c (rows [c*CH, (c+1)*CH)), the chunk's
    // entries (which occupy [indptr[c*CH], indptr[min((c+1)*CH, n)]) in the rows)
    // reordered by (gene, row). col_ptr[c*(n_cols+1) + g] is the global offset of
    // gene g's run in chunk c.
    uint32_t chunk_rows = 0;
    uint32_t n_chunks = 0;
    DevBuf<uint32_t> col_ptr;
    DevBuf<uint32_t> col_row;  // row within the chunk
    DevBuf<float> col_val;
    // For chunk c: its genes by the length of their runs, longest first, then the
    // number of runs of at least 1,024 entries (col_order[c*(n_cols+1) + n_cols]).
    DevBuf<uint32_t> col_order;
};

struct Dense {
    Context *ctx = nullptr;
    uint64_t rows = 0;
    uint32_t cols = 0;
    DevBuf<float> data;
};

// A weighted cell graph in CSR: columns increasing in each row.
struct Graph {
    Context *ctx = nullptr;
    uint64_t n = 0;
    uint64_t nnz = 0;
    DevBuf<uint64_t> indptr;  // n + 1
    DevBuf<uint32_t> indices;  // nnz
    DevBuf<float> weights;     // nnz
};

// Host wall clock in milliseconds.
inline double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

inline unsigned div_up(uint64_t a, uint64_t b) { return static_cast<unsigned>((a + b - 1) / b); }

}  // namespace ky

struct ky_ctx : ky::Context {};
struct ky_csr : ky::Csr {};
struct ky_dense : ky::Dense {};
struct ky_graph : ky::Graph {};
