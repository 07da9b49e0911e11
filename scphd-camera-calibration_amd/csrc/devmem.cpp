/*
 * devmem.cpp -- process-wide device block cache and allocation floor.
 *
 * Released device blocks are kept, per device and size, and handed out
 * again, so that steady-state steps and rig turnover call neither hipMalloc
 * nor hipFree: hipFree waits for the whole device (stalling every other
 * rig's stream), and hipMalloc of large blocks becomes orders of magnitude
 * slower when most of the device memory is in use.  A block enters the cache
 * only when no queued work can still touch it (the filter retires a
 * handle's replaced blocks and hands them over once its streams have
 * drained).
 *
 * The allocation floor: no allocation may leave less than 1/16 of the
 * device's memory free.  The runtime allocates memory of its own on demand
 * (scratch for kernels that spill, queues) and aborts the process when it
 * cannot; a failed allocation here is an error the caller sees instead.
 * The free-memory check and the allocation it admits are one critical
 * section (alloc_mu), so allocations from concurrent handles -- the batch
 * workers grow several rigs' generations at once -- are admitted one after
 * another against the free memory each leaves, and cannot pass the check
 * together.  When an allocation would cross the floor, the device's cached
 * blocks are freed first and the check is made again.
 *
 * The cache holds at most 1/32 of the device's memory, and nothing while
 * less than 1/16 of it is free (a block is then freed instead).  It lives
 * as long as the process (its blocks are not freed at exit, when the
 * runtime may already be gone).
 */
#include "devmem.h"

#include <map>
#include <mutex>
#include <vector>

namespace scphd {

namespace {

hipError_t hip_get_device(int* d) { return hipGetDevice(d); }
hipError_t hip_mem_get_info(size_t* f, size_t* t) { return hipMemGetInfo(f, t); }
hipError_t hip_malloc(void** p, size_t b) { return hipMalloc(p, b); }
hipError_t hip_free(void* p) { return hipFree(p); }
void hip_clear_error() { (void)hipGetLastError(); }

const MemBackend kHip = {hip_get_device, hip_mem_get_info, hip_malloc, hip_free, hip_clear_error};
const MemBackend* g_mem = &kHip;

struct BlockCache {
    std::mutex mu;
    std::multimap<std::pair<int, size_t>, void*> blocks; /* (device, bytes) -> block */
    std::map<int, size_t> held;                          /* cached bytes per device */
};

BlockCache& block_cache() {
    static BlockCache* c = new BlockCache();
    return *c;
}

/* serialises the floor check with the allocation it admits */
std::mutex& alloc_mu() {
    static std::mutex* m = new std::mutex();
    return *m;
}

/* sizes: powers of two from 4 KiB below 1 MiB, whole MiB above */
size_t block_size(size_t bytes) {
    if (bytes >= (1u << 20)) return (bytes + (1u << 20) - 1) & ~(size_t)((1u << 20) - 1);
    size_t b = 4096;
    while (b < bytes) b <<= 1;
    return b;
}

/* the device's total memory and whether less than 1/16 of it is free */
bool memory_low(size_t* total) {
    size_t fr = 0, tot = 0;
    if (g_mem->mem_get_info(&fr, &tot) != hipSuccess) return true;
    if (total) *total = tot;
    return fr < tot / 16;
}

/* whether an allocation of want bytes keeps 1/16 of the memory free */
bool fits_floor(size_t want) {
    size_t fr = 0, tot = 0;
    if (g_mem->mem_get_info(&fr, &tot) != hipSuccess) return false;
    return fr >= want && fr - want >= tot / 16;
}

}  // namespace

void set_mem_backend(const MemBackend* b) { g_mem = b ? b : &kHip; }

void cache_trim(int dev) {
    std::vector<void*> drop;
    {
        BlockCache& c = block_cache();
        std::lock_guard<std::mutex> lk(c.mu);
        for (auto it = c.blocks.begin(); it != c.blocks.end();) {
            if (it->first.first == dev) {
                drop.push_back(it->second);
                c.held[dev] -= it->first.second;
                it = c.blocks.erase(it);
            } else {
                ++it;
            }
        }
    }
    for (void* p : drop) (void)g_mem->free(p);
}

hipError_t cache_alloc(void** p, size_t* bytes) {
    *p = nullptr;
    const size_t want = block_size(*bytes);
    int dev = 0;
    hipError_t e = g_mem->get_device(&dev);
    if (e != hipSuccess) return e;
    {
        BlockCache& c = block_cache();
        std::lock_guard<std::mutex> lk(c.mu);
        /* a cached block up to twice the size for small requests, a quarter
         * more for large ones (64 MiB and up) */
        const size_t most = want >= ((size_t)64 << 20) ? want + want / 4 : 2 * want;
        auto it = c.blocks.lower_bound({dev, want});
        if (it != c.blocks.end() && it->first.first == dev && it->first.second <= most) {
            *p = it->second;
            *bytes = it->first.second;
            c.held[dev] -= it->first.second;
            c.blocks.erase(it);
            return hipSuccess;
        }
    }
    std::lock_guard<std::mutex> lk(alloc_mu());
    if (memory_low(nullptr)) cache_trim(dev);
    if (!fits_floor(want)) {
        cache_trim(dev);
        if (!fits_floor(want)) return hipErrorOutOfMemory;
    }
    e = g_mem->malloc(p, want);
    if (e == hipErrorOutOfMemory) {
        g_mem->clear_error();
        cache_trim(dev);
        if (!fits_floor(want)) {
            *p = nullptr;
            return hipErrorOutOfMemory;
        }
        e = g_mem->malloc(p, want);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        return e;
    }
    *bytes = want;
    return hipSuccess;
}

void cache_put(const Block& b) {
    int dev = 0;
    if (g_mem->get_device(&dev) != hipSuccess) return;
    size_t total = 0;
    const bool low = memory_low(&total);
    if (!low) {
        BlockCache& c = block_cache();
        std::lock_guard<std::mutex> lk(c.mu);
        if (c.held[dev] + b.bytes <= total / 32) {
            c.blocks.insert({{dev, b.bytes}, b.p});
            c.held[dev] += b.bytes;
            return;
        }
    }
    (void)g_mem->free(b.p);
}

}  // namespace scphd
