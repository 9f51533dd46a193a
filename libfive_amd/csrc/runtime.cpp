#include "runtime.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <thread>
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
struct Pending {
    void* p;
    size_t n;
    bool pinned;
};
thread_local std::vector<Pending> t_pending;
}

void pool_defer_begin() { t_defer = true; }

void pool_defer_end(hipStream_t s) {
    (void)hipStreamSynchronize(s);
    t_defer = false;
    std::vector<Pending> pending;
    pending.swap(t_pending);
    for (auto& b : pending) {
        if (b.pinned) pinned_free(b.p, b.n);
        else pool_free(b.p, b.n);
    }
}

void pool_free(void* p, size_t granted) {
    if (t_defer) {
        t_pending.push_back({p, granted, false});
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

namespace {
std::mutex g_pinned_mu;
std::multimap<size_t, void*> g_pinned;
size_t g_pinned_bytes = 0;               // bytes held in g_pinned
}

void* pinned_alloc(size_t bytes, size_t* granted) {
    size_t c = size_t(1) << 20;
    while (c < bytes) c <<= 1;
    *granted = c;
    {
        std::lock_guard<std::mutex> lock(g_pinned_mu);
        auto it = g_pinned.find(c);
        if (it != g_pinned.end()) {
            void* p = it->second;
            g_pinned.erase(it);
            g_pinned_bytes -= c;
            return p;
        }
    }
    void* p = nullptr;
    hip_check(hipHostMalloc(&p, c, hipHostMallocDefault), "hipHostMalloc");
    return p;
}

void pinned_free(void* p, size_t granted) {
    if (t_defer) {
        t_pending.push_back({p, granted, true});
        return;
    }
    std::lock_guard<std::mutex> lock(g_pinned_mu);
    if (g_pinned_bytes + granted > kPinnedKeep) {
        (void)hipHostFree(p);
        return;
    }
    g_pinned.emplace(granted, p);
    g_pinned_bytes += granted;
}

void parallel_memcpy(void* dst, const void* src, size_t bytes) {
    const size_t chunk = size_t(8) << 20;
    const size_t parts = std::min<size_t>(8, (bytes + chunk - 1) / chunk);
    if (parts <= 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const size_t per = (bytes + parts - 1) / parts;
    std::vector<std::thread> ts;
    for (size_t i = 1; i < parts; ++i) {
        const size_t off = i * per;
        const size_t len = std::min(per, bytes - off);
        ts.emplace_back([=] {
            std::memcpy(static_cast<char*>(dst) + off,
                        static_cast<const char*>(src) + off, len);
        });
    }
    std::memcpy(dst, src, std::min(per, bytes));
    for (auto& t : ts) t.join();
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
    // With c->device current: the per-device address of the static staging
    // array used by libfive_tape_eval_values_noalloc.
    const hipError_t se = noalloc_scratch_address(&c->scratch);
    if (se != hipSuccess) {
        c->scratch = nullptr;
        c->scratch_error = std::string("hipGetSymbolAddress: ") + hipGetErrorString(se);
        (void)hipGetLastError();
    }
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
