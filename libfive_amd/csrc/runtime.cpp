#include "runtime.hpp"

#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>

#include "eval_kernels.hpp"

namespace libfive {
namespace amd {

namespace {
std::mutex g_pool_mu;
std::multimap<size_t, void*> g_pool;     // size class -> free block
size_t g_pool_bytes = 0;                 // bytes held in g_pool
}

void* pool_alloc(size_t bytes, size_t* granted) {
    size_t c = 4096;
    while (c < bytes) c <<= 1;
    *granted = c;
    {
        std::lock_guard<std::mutex> lock(g_pool_mu);
        auto it = g_pool.find(c);
        if (it != g_pool.end()) {
            void* p = it->second;
            g_pool.erase(it);
            g_pool_bytes -= c;
            return p;
        }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, c);
    if (e == hipErrorOutOfMemory) {
        // Give the cached blocks back and try once more.
        (void)hipGetLastError();
        std::lock_guard<std::mutex> lock(g_pool_mu);
        for (auto& kv : g_pool) (void)hipFree(kv.second);
        g_pool.clear();
        g_pool_bytes = 0;
        e = hipMalloc(&p, c);
    }
    hip_check(e, "hipMalloc");
    return p;
}

namespace {
thread_local bool t_defer = false;
thread_local std::vector<std::pair<void*, size_t>> t_pending;
}

void pool_defer_begin() { t_defer = true; }

void pool_defer_end(hipStream_t s) {
    (void)hipStreamSynchronize(s);
    t_defer = false;
    std::vector<std::pair<void*, size_t>> pending;
    pending.swap(t_pending);
    for (auto& b : pending) pool_free(b.first, b.second);
}

void pool_free(void* p, size_t granted) {
    if (t_defer) {
        t_pending.emplace_back(p, granted);
        return;
    }
    std::lock_guard<std::mutex> lock(g_pool_mu);
    if (g_pool_bytes + granted > kPoolKeep) {
        (void)hipFree(p);
        return;
    }
    g_pool.emplace(granted, p);
    g_pool_bytes += granted;
}

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        throw Error(std::string(what) + ": " + hipGetErrorString(e));
    }
}

static void apply_env_settings() {
    if (const char* p = std::getenv("LIBFIVE_POINTS_PER_LANE")) {
        set_points_per_lane(std::atoi(p));
    }
}

static Context* make_context() {
    apply_env_settings();
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return nullptr;
    }
    auto* c = new Context;
    hip_check(hipGetDevice(&c->device), "hipGetDevice");
    hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking),
              "hipStreamCreate");
    hip_check(hipMalloc(reinterpret_cast<void**>(&c->d_small),
                        Context::kSmall * sizeof(float)), "hipMalloc");
    hip_check(hipHostMalloc(reinterpret_cast<void**>(&c->h_small),
                            Context::kSmall * sizeof(float),
                            hipHostMallocDefault), "hipHostMalloc");
    return c;
}

static Context* context_or_null() {
    static Context* ctx = [] {
        try {
            return make_context();
        } catch (const Error& e) {
            std::fprintf(stderr, "libfive: GPU initialisation failed: %s\n",
                         e.what());
            return static_cast<Context*>(nullptr);
        }
    }();
    return ctx;
}

Context& context() {
    Context* c = context_or_null();
    if (!c) {
        throw Error("no usable GPU: the evaluation and meshing path runs only "
                    "on the GPU");
    }
    return *c;
}

bool gpu_available() { return context_or_null() != nullptr; }

GpuTape::GpuTape(DeviceTape t) : host(std::move(t)) {
    Context& c = context();
    // Padded with NOP clauses (all-zero words) to whole fetch batches plus
    // one batch for the interpreters' prefetch.
    std::vector<tape::Clause> buf(tape::padded_length(size()),
                                  tape::Clause{0, 0, 0, 0});
    std::copy(host.clauses.begin(), host.clauses.end(), buf.begin());
    d.alloc(buf.size());
    hip_check(hipMemcpyAsync(d.p, buf.data(), buf.size() * sizeof(tape::Clause),
                             hipMemcpyHostToDevice, c.stream),
              "upload tape");
    hip_check(hipStreamSynchronize(c.stream), "upload tape");
}

std::unique_ptr<GpuTape> compile_tree(const Tree& t) {
    Deck deck(t);
    return std::make_unique<GpuTape>(compile_deck(deck));
}

}   // namespace amd
}   // namespace libfive
