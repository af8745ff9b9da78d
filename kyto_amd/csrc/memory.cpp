// Device memory and host transfers of libkyto_gpu.so (host code).
//
// - A caching allocator under every DevBuf: a freed block goes to a per-thread,
//   per-device cache and serves a later request of at most its size and at least half
//   of it, so the passes and searches that allocate the same temporaries on every call
//   stop paying hipMalloc/hipFree (hipFree also waits for the whole device). Blocks
//   are reused only by the thread that freed them: a context is used by one thread at
//   a time, every call runs on its context's one stream and synchronises it before it
//   returns, so a reused block's earlier work is ordered before its new work.
// - Downloads through a pinned staging buffer of the context: the arrays of one
//   download form one stream of 8 MB chunks, and the device-to-host copy of a chunk
//   overlaps the host copy of the chunk before it (8 slices on the persistent host
//   pool), in place of a copy into pageable memory, which the runtime stages at a few
//   GB/s.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <thread>
#include <unordered_map>

#include "common.hpp"

namespace ky {

namespace {

struct Block {
    void *ptr;
    size_t bytes;
    int device;
};

// Cached bytes per thread above which the largest cached blocks are released.
constexpr size_t CACHE_LIMIT = 96ull << 30;

struct Cache {
    std::multimap<size_t, Block> free_blocks;  // by size
    size_t cached = 0;
    ~Cache() { trim(); }
    void trim() {
        for (auto &kv : free_blocks) {
            (void)hipSetDevice(kv.second.device);
            (void)hipFree(kv.second.ptr);
        }
        free_blocks.clear();
        cached = 0;
    }
};

thread_local Cache t_cache;

// The size and device of every block this library allocated (a block goes to the
// cache of the thread that frees it).
std::mutex g_sizes_mutex;
std::unordered_map<void *, Block> g_sizes;

}  // namespace

void *device_alloc(size_t bytes) {
    if (bytes == 0) return nullptr;
    int dev = 0;
    KY_HIP(hipGetDevice(&dev));
    Cache &c = t_cache;
    // The smallest cached block of this device that fits, if it is at most twice the size.
    for (auto it = c.free_blocks.lower_bound(bytes); it != c.free_blocks.end() && it->first <= 2 * bytes; ++it) {
        if (it->second.device != dev) continue;
        Block b = it->second;
        c.free_blocks.erase(it);
        c.cached -= b.bytes;
        return b.ptr;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        c.trim();
        KY_HIP(hipMalloc(&p, bytes));
    }
    std::lock_guard<std::mutex> lock(g_sizes_mutex);
    g_sizes[p] = Block{p, bytes, dev};
    return p;
}

void device_free(void *p) {
    if (!p) return;
    Block b;
    {
        std::lock_guard<std::mutex> lock(g_sizes_mutex);
        auto it = g_sizes.find(p);
        if (it == g_sizes.end()) return;
        b = it->second;
    }
    Cache &c = t_cache;
    c.free_blocks.emplace(b.bytes, b);
    c.cached += b.bytes;
    while (c.cached > CACHE_LIMIT && !c.free_blocks.empty()) {
        auto big = std::prev(c.free_blocks.end());
        Block v = big->second;
        c.free_blocks.erase(big);
        c.cached -= v.bytes;
        {
            std::lock_guard<std::mutex> lock(g_sizes_mutex);
            g_sizes.erase(v.ptr);
        }
        int cur = 0;
        (void)hipGetDevice(&cur);
        (void)hipSetDevice(v.device);
        (void)hipFree(v.ptr);
        (void)hipSetDevice(cur);
    }
}

// Releases the calling thread's cached blocks.
void device_cache_trim() {
    Cache &c = t_cache;
    {
        std::lock_guard<std::mutex> lock(g_sizes_mutex);
        for (auto &kv : c.free_blocks) g_sizes.erase(kv.second.ptr);
    }
    c.trim();
}

// Copies `bytes` from device memory to pageable host memory through the context's
// pinned staging buffer, CHUNK bytes at a time: the device copy of chunk i + 1 runs
// while host threads copy chunk i out of the other half of the buffer.
void download(Context &ctx, void *dst, const void *src, size_t bytes) { download_parts(ctx, {{dst, src, bytes}}); }

namespace {
constexpr size_t STAGE_CHUNK = 8ull << 20;
constexpr size_t STAGE_SLICES = 8;

void ensure_staging(Context &ctx) {
    if (ctx.staging) return;
    KY_HIP(hipHostMalloc(&ctx.staging, 2 * STAGE_CHUNK, hipHostMallocDefault));
    KY_HIP(hipEventCreateWithFlags(&ctx.staged[0], hipEventDisableTiming));
    KY_HIP(hipEventCreateWithFlags(&ctx.staged[1], hipEventDisableTiming));
}
}  // namespace

void upload_parts(Context &ctx, const std::vector<CopyPart> &parts) {
    constexpr size_t SMALL = 1ull << 20;
    hipStream_t st = ctx.stream;
    size_t total = 0;
    for (const CopyPart &p : parts) total += p.bytes;
    if (total <= SMALL) {
        for (const CopyPart &p : parts)
            if (p.bytes) KY_HIP(hipMemcpyAsync(p.dst, p.src, p.bytes, hipMemcpyHostToDevice, st));
        KY_HIP(hipStreamSynchronize(st));
        return;
    }
    ensure_staging(ctx);
    auto *stage = static_cast<unsigned char *>(ctx.staging);
    // Chunk i is copied into staging half i % 2 (once that half's previous transfer has
    // ended) while chunk i - 1 is in flight.
    size_t i = 0;
    for (const CopyPart &p : parts) {
        for (size_t off = 0; off < p.bytes; off += STAGE_CHUNK, i++) {
            const size_t len = std::min(STAGE_CHUNK, p.bytes - off);
            if (i >= 2) KY_HIP(hipEventSynchronize(ctx.staged[i % 2]));
            unsigned char *to = stage + (i % 2) * STAGE_CHUNK;
            const unsigned char *from = static_cast<const unsigned char *>(p.src) + off;
            const size_t per = (len + STAGE_SLICES - 1) / STAGE_SLICES;
            host_parallel(STAGE_SLICES, [&](size_t t) {
                const size_t a = t * per, b = std::min(len, a + per);
                if (a < b) std::memcpy(to + a, from + a, b - a);
            });
            KY_HIP(hipMemcpyAsync(static_cast<unsigned char *>(p.dst) + off, to, len, hipMemcpyHostToDevice, st));
            KY_HIP(hipEventRecord(ctx.staged[i % 2], st));
        }
    }
    KY_HIP(hipStreamSynchronize(st));
}

void download_parts(Context &ctx, const std::vector<CopyPart> &parts) {
    constexpr size_t SMALL = 1ull << 20, CHUNK = STAGE_CHUNK, SLICES = STAGE_SLICES;
    hipStream_t st = ctx.stream;
    size_t total = 0;
    for (const CopyPart &p : parts) total += p.bytes;
    if (total <= SMALL) {
        for (const CopyPart &p : parts)
            if (p.bytes) KY_HIP(hipMemcpyAsync(p.dst, p.src, p.bytes, hipMemcpyDeviceToHost, st));
        KY_HIP(hipStreamSynchronize(st));
        return;
    }
    ensure_staging(ctx);
    // The chunks of every part in order: chunk i goes to staging half i % 2; its host
    // copy (SLICES slices on the host pool) runs while chunk i + 1 is in flight.
    struct Chunk {
        unsigned char *dst;
        const unsigned char *src;
        size_t len;
    };
    std::vector<Chunk> chunks;
    for (const CopyPart &p : parts)
        for (size_t off = 0; off < p.bytes; off += CHUNK)
            chunks.push_back({static_cast<unsigned char *>(p.dst) + off, static_cast<const unsigned char *>(p.src) + off,
                              std::min(CHUNK, p.bytes - off)});
    auto *stage = static_cast<unsigned char *>(ctx.staging);
    auto copy_out = [&](size_t i) {
        const Chunk &c = chunks[i];
        KY_HIP(hipEventSynchronize(ctx.staged[i % 2]));
        const unsigned char *from = stage + (i % 2) * CHUNK;
        const size_t per = (c.len + SLICES - 1) / SLICES;
        host_parallel(SLICES, [&](size_t t) {
            const size_t a = t * per, b = std::min(c.len, a + per);
            if (a < b) std::memcpy(c.dst + a, from + a, b - a);
        });
    };
    for (size_t i = 0; i < chunks.size(); i++) {
        KY_HIP(hipMemcpyAsync(stage + (i % 2) * CHUNK, chunks[i].src, chunks[i].len, hipMemcpyDeviceToHost, st));
        KY_HIP(hipEventRecord(ctx.staged[i % 2], st));
        if (i > 0) copy_out(i - 1);
    }
    copy_out(chunks.size() - 1);
}

}  // namespace ky
