// Device context, memory and error handling shared by the C API.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>

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
