/*
 * devmem.h -- process-wide device block cache and allocation floor of the
 * calibration library.  Every device buffer of a filter handle comes from
 * cache_alloc and goes back through cache_put.
 */
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>

namespace scphd {

struct Block {
    void* p;
    size_t bytes;
};

/* a block of at least *bytes on the current device (a cached one a little
 * larger, else a new one); *bytes returns its size.  Fails with
 * hipErrorOutOfMemory rather than leave less than 1/16 of the device's
 * memory free. */
hipError_t cache_alloc(void** p, size_t* bytes);

/* hand a block no queued work touches to the cache of the current device,
 * or free it when the cache is full or memory is low */
void cache_put(const Block& b);

/* free every cached block of device dev */
void cache_trim(int dev);

/* The device-memory calls the cache makes.  The library uses the HIP
 * runtime's; a host-only test installs a simulated device to check the
 * allocation floor under concurrent allocation without a GPU. */
struct MemBackend {
    hipError_t (*get_device)(int* dev);
    hipError_t (*mem_get_info)(size_t* free_bytes, size_t* total_bytes);
    hipError_t (*malloc)(void** p, size_t bytes);
    hipError_t (*free)(void* p);
    void (*clear_error)();
};

/* nullptr restores the HIP runtime backend.  Not thread-safe against
 * concurrent allocation: install before any allocation. */
void set_mem_backend(const MemBackend* b);

}  // namespace scphd
