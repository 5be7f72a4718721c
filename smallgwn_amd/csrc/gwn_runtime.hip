// gwn_runtime.hip - stream-ordered device memory and revision identities.
//
// Allocation is stream-ordered only, as in the reference (include/gwn/gwn_utils.cuh:364-392,
// AGENTS.md "stream-ordered allocation"): hipMallocAsync / hipFreeAsync on the device's default
// memory pool.  The pool keeps freed blocks cached (release threshold = max) so repeated builds
// and query batches do not return memory to the driver between calls.
#include "gwn_core.hpp"

#include <atomic>
#include <limits>
#include <mutex>

namespace gwn {

namespace {

void configure_pool_once(int device) {
    static std::mutex mutex;
    static bool configured[64] = {};
    if (device < 0 || device >= 64)
        return;
    std::lock_guard<std::mutex> lock(mutex);
    if (configured[device])
        return;
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool != nullptr) {
        std::uint64_t threshold = std::numeric_limits<std::uint64_t>::max();
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &threshold);
    }
    configured[device] = true;
}

} // namespace

void *device_alloc(std::size_t bytes, hipStream_t stream) {
    if (bytes == 0)
        return nullptr;
    int device = 0;
    GWN_HIP(hipGetDevice(&device));
    configure_pool_once(device);
    void *ptr = nullptr;
    GWN_HIP(hipMallocAsync(&ptr, bytes, stream));
    return ptr;
}

void device_free(void *ptr, hipStream_t stream) noexcept {
    if (ptr != nullptr)
        (void)hipFreeAsync(ptr, stream);
}

// Process-unique nonzero identity per successful build or refit, so a moment built for an
// earlier BVH state can never validate against a later one even if allocations are reused
// (include/gwn/detail/gwn_bvh_revision.cuh:10-24).
std::uint64_t reserve_revision() noexcept {
    static std::atomic<std::uint64_t> next{1};
    std::uint64_t r = next.load(std::memory_order_relaxed);
    for (;;) {
        if (r == 0 || r == std::numeric_limits<std::uint64_t>::max())
            return 0;
        if (next.compare_exchange_weak(r, r + 1, std::memory_order_relaxed))
            return r;
    }
}

std::uint64_t next_storage_generation() noexcept {
    static std::atomic<std::uint64_t> next{1};
    return next.fetch_add(1, std::memory_order_relaxed);
}

} // namespace gwn
