#include "runtime.hpp"

#include <cstdio>
#include <cstdlib>

#include "eval_kernels.hpp"

namespace libfive {
namespace amd {

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
