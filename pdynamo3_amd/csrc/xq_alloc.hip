// Device and pinned host memory of libxcquad: process-wide caching allocators over hipMalloc and hipHostMalloc.
//
// hipFree synchronises the whole device, so a request loop that sets up and releases handles while other handles'
// quadratures run would stall them at every release.  Freed blocks are kept instead and handed out again: a request
// takes the smallest cached block of at least its size and at most a quarter larger (sizes are rounded up to 512 bytes,
// and to 2 MiB above 1 MiB), so a loop whose jobs come and go reuses device memory without runtime calls.  When
// hipMalloc fails, every cached block is released (hipFree) and the allocation is tried once more.  A block is cached
// only after the work that used it has completed: every caller synchronises the owning stream before it frees.
// Pinned host blocks (the staging of P and V) are cached the same way, since hipHostFree synchronises the device too.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <map>
#include <mutex>
#include <unordered_map>
#include <string>
#include <utility>

#include "xq_internal.h"

namespace xq {

namespace {

std::mutex g_mutex;
std::multimap<size_t, void *> g_cached;            // size -> idle block
std::unordered_map<void *, std::pair<size_t, size_t>> g_live;   // block -> (size, bytes requested)
size_t g_cachedBytes = 0;
long long g_failAfter = -1;                        // test hook: allocations left before one fails (-1: none fails)
std::multimap<size_t, void *> g_hostCached;        // pinned host blocks, as above
std::unordered_map<void *, size_t> g_hostLive;

size_t rounded(size_t bytes) {
    if (bytes == 0) bytes = 1;
    const size_t q = bytes > (size_t(1) << 20) ? (size_t(2) << 20) : size_t(512);
    return (bytes + q - 1) / q * q;
}

void release_cached_locked() {
    for (auto &kv : g_cached) (void)hipFree(kv.second);
    g_cached.clear();
    g_cachedBytes = 0;
}

}  // namespace

hipError_t cache_alloc(void **ptr, size_t bytes) {
    const size_t size = rounded(bytes);
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_failAfter >= 0 && g_failAfter-- == 0) {
        *ptr = nullptr;
        return hipErrorOutOfMemory;
    }
    auto it = g_cached.lower_bound(size);
    if (it != g_cached.end() && it->first <= size + size / 4) {
        *ptr = it->second;
        g_live[it->second] = std::make_pair(it->first, bytes);
        g_cachedBytes -= it->first;
        g_cached.erase(it);
        return hipSuccess;
    }
    hipError_t e = hipMalloc(ptr, size);
    if (e != hipSuccess && !g_cached.empty()) {
        (void)hipGetLastError();
        release_cached_locked();
        e = hipMalloc(ptr, size);
    }
    if (e != hipSuccess) {
        *ptr = nullptr;
        return e;
    }
    g_live[*ptr] = std::make_pair(size, bytes);
    return hipSuccess;
}

size_t cache_free(void *ptr) {
    if (ptr == nullptr) return 0;
    std::lock_guard<std::mutex> lock(g_mutex);
    auto it = g_live.find(ptr);
    if (it == g_live.end()) return 0;
    const size_t requested = it->second.second;
    g_cached.emplace(it->second.first, ptr);
    g_cachedBytes += it->second.first;
    g_live.erase(it);
    return requested;
}

hipError_t cache_host_alloc(void **ptr, size_t bytes) {
    const size_t size = rounded(bytes);
    std::lock_guard<std::mutex> lock(g_mutex);
    auto it = g_hostCached.lower_bound(size);
    if (it != g_hostCached.end() && it->first <= size + size / 4) {
        *ptr = it->second;
        g_hostLive[it->second] = it->first;
        g_hostCached.erase(it);
        return hipSuccess;
    }
    hipError_t e = hipHostMalloc(ptr, size, hipHostMallocDefault);
    if (e != hipSuccess && !g_hostCached.empty()) {
        (void)hipGetLastError();
        for (auto &kv : g_hostCached) (void)hipHostFree(kv.second);
        g_hostCached.clear();
        e = hipHostMalloc(ptr, size, hipHostMallocDefault);
    }
    if (e != hipSuccess) {
        *ptr = nullptr;
        return e;
    }
    g_hostLive[*ptr] = size;
    return hipSuccess;
}

void cache_host_free(void *ptr) {
    if (ptr == nullptr) return;
    std::lock_guard<std::mutex> lock(g_mutex);
    auto it = g_hostLive.find(ptr);
    if (it == g_hostLive.end()) return;
    g_hostCached.emplace(it->second, ptr);
    g_hostLive.erase(it);
}

size_t cache_idle_bytes() {
    std::lock_guard<std::mutex> lock(g_mutex);
    return g_cachedBytes;
}

void cache_fail_after(long long count) {
    std::lock_guard<std::mutex> lock(g_mutex);
    g_failAfter = count;
}

void cache_release() {
    std::lock_guard<std::mutex> lock(g_mutex);
    release_cached_locked();
}

int ensure_free_device_memory(size_t bytes, const char *what) {
    size_t freeB = 0, totalB = 0;
    if (hipMemGetInfo(&freeB, &totalB) != hipSuccess) {
        set_error(std::string(what) + ": hipMemGetInfo failed");
        return kAlgorithmError;
    }
    if (freeB >= bytes) return kOK;
    cache_release();   // idle blocks of the cache go back to the device
    if (hipMemGetInfo(&freeB, &totalB) != hipSuccess) {
        set_error(std::string(what) + ": hipMemGetInfo failed");
        return kAlgorithmError;
    }
    if (freeB >= bytes) return kOK;
    set_error(std::string(what) + ": " + std::to_string(freeB >> 20) + " MiB of device memory free, " +
              std::to_string(bytes >> 20) + " MiB needed for the kernel's private memory");
    return kOutOfMemory;
}

}  // namespace xq
