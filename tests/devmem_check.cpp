/*
 * devmem_check.cpp -- host-only harness for the device block cache
 * (scphd-camera-calibration_amd/csrc/devmem.cpp): a simulated device
 * behind scphd::MemBackend, and worker threads that grow and release
 * sample-plane sets the way concurrent filter handles do.  Built with g++
 * by tests/test_devmem.py; needs no GPU.
 */
#include <atomic>
#include <chrono>
#include <cstdint>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "devmem.h"

namespace {

struct SimDevice {
    std::mutex mu;
    size_t total = 0, used = 0, peak = 0;
    uintptr_t next = 0x100000;
    std::vector<std::pair<uintptr_t, size_t>> live;
    std::atomic<int64_t> violations{0}, mallocs{0};
    int delay_us = 0;
} g_dev;

hipError_t sim_get_device(int* d) {
    *d = 0;
    return hipSuccess;
}

hipError_t sim_info(size_t* f, size_t* t) {
    std::this_thread::sleep_for(std::chrono::microseconds(g_dev.delay_us));
    std::lock_guard<std::mutex> lk(g_dev.mu);
    *f = g_dev.total - g_dev.used;
    *t = g_dev.total;
    return hipSuccess;
}

/* like hipMalloc: succeeds whenever the bytes are free, floor or not */
hipError_t sim_malloc(void** p, size_t b) {
    std::this_thread::sleep_for(std::chrono::microseconds(g_dev.delay_us));
    std::lock_guard<std::mutex> lk(g_dev.mu);
    if (b > g_dev.total - g_dev.used) return hipErrorOutOfMemory;
    g_dev.used += b;
    if (g_dev.used > g_dev.peak) g_dev.peak = g_dev.used;
    g_dev.live.push_back({g_dev.next, b});
    *p = (void*)g_dev.next;
    g_dev.next += b + 4096;
    g_dev.mallocs++;
    if (g_dev.total - g_dev.used < g_dev.total / 16) g_dev.violations++;
    return hipSuccess;
}

hipError_t sim_free(void* p) {
    std::lock_guard<std::mutex> lk(g_dev.mu);
    for (size_t i = 0; i < g_dev.live.size(); ++i)
        if (g_dev.live[i].first == (uintptr_t)p) {
            g_dev.used -= g_dev.live[i].second;
            g_dev.live[i] = g_dev.live.back();
            g_dev.live.pop_back();
            return hipSuccess;
        }
    return hipErrorInvalidValue;
}

void sim_clear() {}

const scphd::MemBackend kSim = {sim_get_device, sim_info, sim_malloc, sim_free, sim_clear};

}  // namespace

/* n_threads workers, each growing `rounds` seven-plane sets of random size
 * (1/40 to 1/16 of the device each plane), holding them briefly, then
 * releasing them through the cache.  out[0] = allocations that left less
 * than 1/16 of the device free, out[1] = device allocations made,
 * out[2] = plane allocations refused, out[3] = plane allocations granted,
 * out[4] = peak bytes in use on the device. */
extern "C" void devmem_stress(int n_threads, int rounds, size_t total_bytes, int delay_us, int64_t* out) {
    g_dev.total = total_bytes;
    g_dev.used = 0;
    g_dev.peak = 0;
    g_dev.delay_us = delay_us;
    g_dev.violations = 0;
    g_dev.mallocs = 0;
    scphd::set_mem_backend(&kSim);
    std::atomic<int64_t> refused{0}, granted{0};
    std::vector<std::thread> th;
    for (int t = 0; t < n_threads; ++t)
        th.emplace_back([&, t] {
            std::mt19937_64 rng(1234 + t);
            for (int r = 0; r < rounds; ++r) {
                const size_t plane = total_bytes / 40 + rng() % (total_bytes / 16 - total_bytes / 40);
                std::vector<scphd::Block> got;
                for (int c = 0; c < 7; ++c) {
                    void* p = nullptr;
                    size_t b = plane;
                    if (scphd::cache_alloc(&p, &b) == hipSuccess) {
                        got.push_back({p, b});
                        granted++;
                    } else {
                        refused++;
                    }
                }
                std::this_thread::sleep_for(std::chrono::microseconds(rng() % 300));
                for (const auto& b : got) scphd::cache_put(b);
            }
        });
    for (auto& x : th) x.join();
    scphd::cache_trim(0);
    scphd::set_mem_backend(nullptr);
    out[0] = g_dev.violations;
    out[1] = g_dev.mallocs;
    out[2] = refused;
    out[3] = granted;
    out[4] = (int64_t)g_dev.peak;
}
