// Device memory and stream cache of the device layer.
//
// A design loop creates and destroys hundreds of small grids a second (one per candidate
// geometry).  hipFree synchronises the device and unmaps memory, and a stream costs a queue
// set-up, so both are kept for reuse instead of being returned to the runtime:
//  * blocks are rounded up to a size class (powers of two up to 1 MiB, then eight classes per
//    octave, so at most 12.5 % is wasted) and cached per device on release;
//  * a released block may still be read or written by work already enqueued on any stream, so it
//    first waits in a pending list; the first allocation that would reuse a pending block
//    synchronises the device once and then moves every pending block to the free list;
//  * the cache holds at most kCacheCap bytes; beyond that a released block goes back to the
//    runtime; an allocation that fails empties the cache and tries once more;
//  * fdtd_dev_release_cached_memory() returns everything cached to the runtime;
//  * releasing a block the cache does not hold as live (a second release) is refused with a
//    message, never passed to hipFree: that would unmap memory the cache may hand out again.
// Streams (non-blocking, default priority) are pooled the same way; a stream is released only
// after its owner has synchronised it.
#include <map>
#include <mutex>
#include <unordered_map>

#include "dev_internal.h"

namespace fdtd {
namespace {

constexpr size_t kCacheCap = (size_t)16 << 30;

struct Block {
    size_t bytes;
    int device;
};

struct Cache {
    std::mutex mu;
    std::unordered_map<void *, Block> live;              // handed out
    std::multimap<std::pair<int, size_t>, void *> idle;  // (device, class) -> block, safe to reuse
    std::multimap<std::pair<int, size_t>, void *> pending;
    size_t cached = 0;  // bytes in idle + pending
    std::vector<std::pair<int, hipStream_t>> streams;
};

Cache &cache() {
    static Cache *c = new Cache();  // never destroyed: blocks may be released during exit
    return *c;
}

size_t size_class(size_t bytes) {
    if (bytes <= 256) return 256;
    size_t p = 256;
    while (p < bytes && p < ((size_t)1 << 20)) p <<= 1;
    if (p >= bytes) return p;
    size_t oct = (size_t)1 << 20;
    while (oct * 2 <= bytes) oct <<= 1;  // oct <= bytes < 2 oct
    size_t step = oct / 8;
    return (bytes + step - 1) / step * step;
}

int current_device() {
    int dev = 0;
    (void)hipGetDevice(&dev);
    return dev;
}

// Caller holds c.mu.  Returns every cached block of every device to the runtime.
void release_all(Cache &c) {
    if (!c.pending.empty()) (void)hipDeviceSynchronize();
    for (auto *m : {&c.idle, &c.pending}) {
        for (auto &kv : *m) (void)hipFree(kv.second);
        m->clear();
    }
    c.cached = 0;
}

}  // namespace

hipError_t dev_malloc_bytes(void **p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) return hipSuccess;
    Cache &c = cache();
    std::lock_guard<std::mutex> lock(c.mu);
    const int dev = current_device();
    const size_t cls = size_class(bytes);
    const auto key = std::make_pair(dev, cls);
    auto it = c.idle.find(key);
    if (it == c.idle.end() && c.pending.count(key)) {
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) return e;
        for (auto &kv : c.pending) c.idle.insert(kv);
        c.pending.clear();
        it = c.idle.find(key);
    }
    if (it != c.idle.end()) {
        *p = it->second;
        c.idle.erase(it);
        c.cached -= cls;
        c.live[*p] = Block{cls, dev};
        return hipSuccess;
    }
    hipError_t e = hipMalloc(p, cls);
    if (e != hipSuccess && c.cached > 0) {
        (void)hipGetLastError();
        release_all(c);
        e = hipMalloc(p, cls);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        return e;
    }
    c.live[*p] = Block{cls, dev};
    return hipSuccess;
}

// True when p already waits in the idle or pending list (caller holds c.mu).
bool cached_block(const Cache &c, void *p) {
    for (const auto *m : {&c.idle, &c.pending})
        for (const auto &kv : *m)
            if (kv.second == p) return true;
    return false;
}

hipError_t dev_free(void *p) {
    if (!p) return hipSuccess;
    Cache &c = cache();
    std::lock_guard<std::mutex> lock(c.mu);
    auto it = c.live.find(p);
    if (it == c.live.end()) {
        // Every device block of this library comes from dev_malloc, so this is a second release
        // (or a pointer that was never ours).  Returning it to the runtime here would unmap a
        // block the cache may already have handed to another owner, whose next access (a kernel,
        // or the runtime's host-side copy into device memory) would then fault; refuse loudly.
        std::fprintf(stderr, "fdtd3d: device block %p released twice or not allocated here (%s); ignored\n", p,
                     cached_block(c, p) ? "already cached" : "unknown");
        return hipErrorInvalidValue;
    }
    const Block b = it->second;
    c.live.erase(it);
    if (c.cached + b.bytes > kCacheCap) return hipFree(p);
    c.pending.insert({std::make_pair(b.device, b.bytes), p});
    c.cached += b.bytes;
    return hipSuccess;
}

hipError_t dev_stream_acquire(hipStream_t *s) {
    Cache &c = cache();
    {
        std::lock_guard<std::mutex> lock(c.mu);
        const int dev = current_device();
        for (size_t q = 0; q < c.streams.size(); q++)
            if (c.streams[q].first == dev) {
                *s = c.streams[q].second;
                c.streams.erase(c.streams.begin() + (long)q);
                return hipSuccess;
            }
    }
    return hipStreamCreateWithFlags(s, hipStreamNonBlocking);
}

hipError_t dev_stream_release(hipStream_t s) {
    if (!s) return hipSuccess;
    Cache &c = cache();
    std::lock_guard<std::mutex> lock(c.mu);
    if (c.streams.size() >= 256) return hipStreamDestroy(s);
    hipDevice_t dev = 0;
    if (hipStreamGetDevice(s, &dev) != hipSuccess) return hipStreamDestroy(s);
    c.streams.emplace_back((int)dev, s);
    return hipSuccess;
}

}  // namespace fdtd

extern "C" int fdtd_dev_release_cached_memory(void) {
    fdtd::Cache &c = fdtd::cache();
    std::lock_guard<std::mutex> lock(c.mu);
    fdtd::release_all(c);
    return 0;
}
