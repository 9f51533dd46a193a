// Device context, memory and error handling shared by the C API.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>

#include "deck.hpp"
#include "tape.hpp"

namespace libfive {
namespace amd {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char* what);

// One library stream per process.  Calls through the C API serialise on
// `mu`; work enqueued by the *_device entry points goes to the caller's
// stream instead.
struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    float* d_small = nullptr;     // device staging for single-point calls
    float* h_small = nullptr;     // pinned host staging
    // Device address of the allocation-free entry point's staging array
    // on this context's device (nullptr if it could not be looked up, with
    // the reason in scratch_error).
    float* scratch = nullptr;
    std::string scratch_error;
    static constexpr size_t kSmall = 64;
};

// Throws Error when no GPU is usable.
Context& context();
bool gpu_available();

// Owning device buffer.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { reset(); }
    void alloc(size_t count) {
        reset();
        if (count) {
            hip_check(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)),
                      "hipMalloc");
        }
        n = count;
    }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// Device memory cache for the mesher's per-render buffers: a released
// block is kept and handed to a later request of the same size class
// (powers of two from 4 KiB), so that repeated renders do not pay for
// hipMalloc / hipFree.  Blocks are returned only after the render's stream
// has finished with them.  At most kPoolKeep bytes stay cached.
void* pool_alloc(size_t bytes, size_t* granted);
void pool_free(void* p, size_t granted);

// Between pool_defer_begin() and pool_defer_end(s) on one thread, released
// blocks are held back; pool_defer_end waits for stream s and then returns
// them to the cache, so that no block is reused while a kernel of that
// stream may still use it (also when the render ends by an exception).
void pool_defer_begin();
void pool_defer_end(hipStream_t s);
constexpr size_t kPoolKeep = size_t(48) << 30;

template <typename T>
struct PoolBuf {
    T* p = nullptr;
    size_t n = 0;
    size_t granted = 0;
    PoolBuf() = default;
    explicit PoolBuf(size_t count) { alloc(count); }
    PoolBuf(const PoolBuf&) = delete;
    PoolBuf& operator=(const PoolBuf&) = delete;
    ~PoolBuf() { reset(); }
    void alloc(size_t count) {
        reset();
        if (count) p = static_cast<T*>(pool_alloc(count * sizeof(T), &granted));
        n = count;
    }
    void reset() {
        if (p) pool_free(p, granted);
        p = nullptr;
        n = 0;
        granted = 0;
    }
    void swap(PoolBuf& o) {
        std::swap(p, o.p);
        std::swap(n, o.n);
        std::swap(granted, o.granted);
    }
};

// Pinned host staging blocks (hipHostMalloc), cached like pool_alloc's.
// Page-locked host blocks (power-of-two size classes from 1 MiB), pooled
// like the device blocks; at most kPinnedKeep bytes stay cached.
constexpr size_t kPinnedKeep = size_t(8) << 30;
void* pinned_alloc(size_t bytes, size_t* granted);
void pinned_free(void* p, size_t granted);

// memcpy split over host threads for large copies.
void parallel_memcpy(void* dst, const void* src, size_t bytes);

// A compiled tape resident in device memory.
struct GpuTape {
    DeviceTape host;
    DevBuf<tape::Clause> d;
    explicit GpuTape(DeviceTape t);
    uint32_t size() const { return uint32_t(host.clauses.size()); }
};

std::unique_ptr<GpuTape> compile_tree(const Tree& t);

}   // namespace amd
}   // namespace libfive
