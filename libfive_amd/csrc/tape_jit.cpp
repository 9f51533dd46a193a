// Tape specialisation with hipRTC (see tape_jit.hpp).
#include "tape_jit.hpp"

#include <atomic>

#include <hip/hiprtc.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "deck.hpp"
#include "runtime.hpp"
#include "tape.hpp"

// kJitHeaderNames / kJitHeaderSources: the text of fmath.hpp, opcode.hpp,
// tape.hpp and clause_ops.hpp, generated at build time.
#include "jit_headers.inc"

namespace libfive {
namespace amd {

using namespace tape;

SpecializedTape::~SpecializedTape() {
    if (module) (void)hipModuleUnload(module);
}

namespace {

std::string fimm(float f) {
    char buf[48];
    std::snprintf(buf, sizeof(buf), "fmath::u2f(0x%08xu)", fmath::f2u(f));
    return buf;
}

// Register -> expression of its current value (and of its partials).
struct Names {
    std::vector<std::string> v;
    std::vector<std::string> g[3];
    explicit Names(uint32_t nregs) : v(nregs, "0.0f") {
        for (auto& gi : g) gi.assign(nregs, "0.0f");
        v[kRegX] = "px";
        v[kRegY] = "py";
        v[kRegZ] = "pz";
        g[0][kRegX] = "1.0f";
        g[1][kRegY] = "1.0f";
        g[2][kRegZ] = "1.0f";
    }
};

// Operands of a clause exactly as the interpreter loop forms them
// (LF_VALUE_STEP and LF_DERIV_STEP in eval_kernels.hip).  After each clause
// an empty asm statement takes its results as in-out register operands
// (or after each group of clauses, see pin_group):
// the clause is computed there, in tape order, instead of being sunk
// towards its uses (which keeps all of its inputs alive; on the
// 809-clause benchmark tape the derivative kernel otherwise needs all 512
// VGPRs and spills, and with the barriers it needs 168).
// Clauses per pinning statement (LIBFIVE_AMD_JIT_GROUP, default 2): the
// results of a group are pinned together after its last clause.  On the
// benchmark tape, groups of 2 give the derivative kernel 168 VGPRs (three
// waves per SIMD) and 3.21 ms; 1 gives 170 (3.25 ms once recompiled for
// three waves), 4 gives 174 (3.57 ms).
static size_t pin_group() {
    const char* e = std::getenv("LIBFIVE_AMD_JIT_GROUP");
    const long g = e ? std::atol(e) : 2;
    return size_t(g < 1 ? 1 : g > 4 ? 4 : g);
}

static void pin(std::string& out, const std::vector<std::string>& vars) {
    if (vars.empty()) return;
    out += "    asm volatile(\"\" :";
    for (size_t i = 0; i < vars.size(); ++i) {
        out += (i ? ", \"+v\"(" : " \"+v\"(") + vars[i] + ")";
    }
    out += ");\n";
}

static bool is_sincos(uint32_t o) {
    return o == Opcode::OP_SIN || o == Opcode::OP_COS;
}

// Call-free form (`fast`: the tape has only the common opcodes): sin and cos
// take no out-of-line path, and a point with an argument of 2^17 or more
// (or a NaN the accumulation cannot skip, or an infinity) gets
// tape::kRedoMark as its value, for the interpreter to evaluate again.
// Leaving out the calls takes the benchmark tape's derivative kernel from
// 3.39 to 3.26 ms (with the occupancy request below).
void value_body(const DeviceTape& t, bool fast, std::string& out) {
    Names n(t.num_regs);
    char buf[256];
    size_t k = 0;
    const size_t group = pin_group();
    std::vector<std::string> pending;
    if (fast) out += "    float amax_ = 0.0f;\n";
    for (const Clause& c : t.clauses) {
        const uint32_t o = op(c);
        if (o == OP_NOP) continue;
        const uint32_t m = mode(c);
        const std::string im = fimm(imm(c));
        const std::string a = (m == MODE_A_IMM) ? im : n.v[ra(c)];
        const std::string b = !is_binary(c) ? std::string("0.0f")
                              : (m == MODE_B_IMM) ? im : n.v[rb(c)];
        if (fast && is_sincos(o)) {
            std::snprintf(buf, sizeof(buf),
                          "    float v%zu = value_sincos_fast(%uu, %s, amax_);\n",
                          k, o, a.c_str());
        } else {
            std::snprintf(buf, sizeof(buf),
                          "    float v%zu; { const float a_ = %s, b_ = %s; "
                          "apply_value<1>(%uu, &a_, &b_, %s, &v%zu); }\n",
                          k, a.c_str(), b.c_str(), o, im.c_str(), k);
        }
        out += buf;
        pending.push_back("v" + std::to_string(k));
        if (pending.size() >= group) {
            pin(out, pending);
            pending.clear();
        }
        n.v[dst(c)] = "v" + std::to_string(k);
        ++k;
    }
    pin(out, pending);
    if (fast) {
        out += "    out[i] = (amax_ < fmath::kCwLimit) ? " + n.v[t.root_reg] +
               " : fmath::u2f(tape::kRedoMark);\n";
    } else {
        out += "    out[i] = " + n.v[t.root_reg] + ";\n";
    }
}

void deriv_body(const DeviceTape& t, bool fast, std::string& out) {
    Names n(t.num_regs);
    char buf[512];
    size_t k = 0;
    const size_t group = pin_group();
    std::vector<std::string> pending;
    if (fast) out += "    float amax_ = 0.0f;\n";
    for (const Clause& c : t.clauses) {
        const uint32_t o = op(c);
        if (o == OP_NOP) continue;
        const uint32_t m = mode(c);
        const std::string im = fimm(imm(c));
        std::string av = im, ad[3] = {"0.0f", "0.0f", "0.0f"};
        std::string bv = "0.0f", bd[3] = {"0.0f", "0.0f", "0.0f"};
        if (m != MODE_A_IMM && o != OP_LOADI) {
            av = n.v[ra(c)];
            for (int j = 0; j < 3; ++j) ad[j] = n.g[j][ra(c)];
        }
        if (is_binary(c)) {
            if (m == MODE_B_IMM) {
                bv = im;
            } else {
                bv = n.v[rb(c)];
                for (int j = 0; j < 3; ++j) bd[j] = n.g[j][rb(c)];
            }
        }
        if (fast && is_sincos(o)) {
            std::snprintf(buf, sizeof(buf),
                          "    float v%zu, d%zu[3]; { const float ad_[3] = {%s, %s, %s}; "
                          "deriv_sincos_fast(%uu, %s, ad_, v%zu, d%zu, amax_); }\n",
                          k, k, ad[0].c_str(), ad[1].c_str(), ad[2].c_str(), o,
                          av.c_str(), k, k);
        } else {
            std::snprintf(buf, sizeof(buf),
                          "    float v%zu, d%zu[3]; { const float ad_[3] = {%s, %s, %s}; "
                          "const float bd_[3] = {%s, %s, %s}; "
                          "apply_deriv(%uu, %s, ad_, %s, bd_, %s, v%zu, d%zu); }\n",
                          k, k, ad[0].c_str(), ad[1].c_str(), ad[2].c_str(),
                          bd[0].c_str(), bd[1].c_str(), bd[2].c_str(), o,
                          av.c_str(), bv.c_str(), im.c_str(), k, k);
        }
        out += buf;
        const std::string ks = std::to_string(k);
        for (const char* sfx : {"", "[0]", "[1]", "[2]"}) {
            pending.push_back((*sfx ? "d" : "v") + ks + sfx);
        }
        if (pending.size() >= 4 * group) {
            // the accumulator is pinned with every group: left free, it
            // takes its operands' registers out of the allocation and
            // costs the kernel 50 VGPRs on the benchmark tape
            if (fast) pending.push_back("amax_");
            pin(out, pending);
            pending.clear();
        }
        const uint32_t d = dst(c);
        n.v[d] = "v" + std::to_string(k);
        for (int j = 0; j < 3; ++j) {
            n.g[j][d] = "d" + std::to_string(k) + "[" + std::to_string(j) + "]";
        }
        ++k;
    }
    if (fast && !pending.empty()) pending.push_back("amax_");
    pin(out, pending);
    const uint32_t r = t.root_reg;
    if (fast) {
        out += "    ov[i] = (amax_ < fmath::kCwLimit) ? " + n.v[r] +
               " : fmath::u2f(tape::kRedoMark);\n";
    } else {
        out += "    ov[i] = " + n.v[r] + ";\n";
    }
    out += "    odx[i] = " + n.g[0][r] + ";\n";
    out += "    ody[i] = " + n.g[1][r] + ";\n";
    out += "    odz[i] = " + n.g[2][r] + ";\n";
}

}   // namespace

std::string specialized_source(const DeviceTape& t, int deriv_waves) {
    const bool fast = tape_is_fast(t);
    const std::string occupancy = deriv_waves > 0
        ? "__attribute__((amdgpu_waves_per_eu(" + std::to_string(deriv_waves) + ")))\n"
        : std::string();
    std::string s =
        "#include \"clause_ops.hpp\"\n"
        "using namespace libfive;\n"
        "using namespace libfive::amd;\n"
        "extern \"C\" __global__ __launch_bounds__(256)\n"
        "void lf_values(const float* __restrict__ xs, const float* __restrict__ ys,\n"
        "               const float* __restrict__ zs, float* __restrict__ out,\n"
        "               unsigned long long n) {\n"
        "    const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;\n"
        "    if (i >= n) return;\n"
        "    const float px = xs[i], py = ys[i], pz = zs[i];\n";
    value_body(t, fast, s);
    s += "}\n\n"
         "extern \"C\" __global__ __launch_bounds__(256)\n" + occupancy +
         "void lf_derivs(const float* __restrict__ xs, const float* __restrict__ ys,\n"
         "               const float* __restrict__ zs, float* __restrict__ ov,\n"
         "               float* __restrict__ odx, float* __restrict__ ody,\n"
         "               float* __restrict__ odz, unsigned long long n) {\n"
         "    const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;\n"
         "    if (i >= n) return;\n"
         "    const float px = xs[i], py = ys[i], pz = zs[i];\n";
    deriv_body(t, fast, s);
    s += "}\n";
    return s;
}

static void rtc_check(hiprtcResult r, const char* what) {
    if (r != HIPRTC_SUCCESS) {
        throw Error(std::string(what) + ": " + hiprtcGetErrorString(r));
    }
}

std::vector<char> compile_specialized(const DeviceTape& t, int deriv_waves) {
    const std::string src = specialized_source(t, deriv_waves);
    // LIBFIVE_AMD_JIT_DUMP=<path>: write the generated source there.
    if (const char* dump = std::getenv("LIBFIVE_AMD_JIT_DUMP")) {
        if (FILE* f = std::fopen(dump, "w")) {
            std::fwrite(src.data(), 1, src.size(), f);
            std::fclose(f);
        }
    }
    hiprtcProgram prog = nullptr;
    rtc_check(hiprtcCreateProgram(&prog, src.c_str(), "lf_tape.hip",
                                  int(sizeof(kJitHeaderNames) / sizeof(kJitHeaderNames[0])),
                                  kJitHeaderSources, kJitHeaderNames),
              "hiprtcCreateProgram");
    const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17",
                          "-ffp-contract=off", "-Xclang", "-target-feature",
                          "-Xclang", "-packed-fp32-ops"};
    const hiprtcResult cr = hiprtcCompileProgram(prog, int(sizeof(opts) / sizeof(opts[0])), opts);
    if (cr != HIPRTC_SUCCESS) {
        size_t ls = 0;
        hiprtcGetProgramLogSize(prog, &ls);
        std::string log(ls, '\0');
        if (ls) hiprtcGetProgramLog(prog, &log[0]);
        hiprtcDestroyProgram(&prog);
        throw Error("tape specialisation failed to compile: " + log.substr(0, 2000));
    }
    size_t code_size = 0;
    rtc_check(hiprtcGetCodeSize(prog, &code_size), "hiprtcGetCodeSize");
    std::vector<char> code(code_size);
    rtc_check(hiprtcGetCode(prog, code.data()), "hiprtcGetCode");
    hiprtcDestroyProgram(&prog);
    return code;
}

static std::unique_ptr<SpecializedTape> load_specialized(const std::vector<char>& code) {
    auto k = std::make_unique<SpecializedTape>();
    hip_check(hipModuleLoadData(&k->module, code.data()), "hipModuleLoadData");
    hip_check(hipModuleGetFunction(&k->values, k->module, "lf_values"),
              "hipModuleGetFunction(lf_values)");
    hip_check(hipModuleGetFunction(&k->derivs, k->module, "lf_derivs"),
              "hipModuleGetFunction(lf_derivs)");
    k->code_bytes = code.size();
    return k;
}

std::unique_ptr<SpecializedTape> specialize(const DeviceTape& t) {
    context();     // device selected and initialised
    const auto t0 = std::chrono::steady_clock::now();
    auto k = load_specialized(compile_specialized(t, 0));
    // A derivative kernel a few registers above the three-waves-per-SIMD
    // line (168 VGPRs) runs at two waves; compiled again with a request for
    // three, it usually fits with at most a few spilled registers, which
    // costs less than the lost occupancy (benchmark tape: 3.82 -> 3.26 ms).
    int regs = 0;
    if (hipFuncGetAttribute(&regs, HIP_FUNC_ATTRIBUTE_NUM_REGS, k->derivs) == hipSuccess &&
        regs > 168 && regs <= 184) {
        auto k3 = load_specialized(compile_specialized(t, 3));
        int spill = 0;
        if (hipFuncGetAttribute(&spill, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, k3->derivs) ==
                hipSuccess && spill <= 64) {
            k3->deriv_waves = 3;
            k = std::move(k3);
        }
    }
    static std::atomic<uint64_t> serial{0};
    k->serial = ++serial;
    k->redo = tape_is_fast(t);
    k->compile_ms = std::chrono::duration<double, std::milli>(
                        std::chrono::steady_clock::now() - t0).count();
    return k;
}

static bool valid_block(unsigned block) {
    return block == 64 || block == 128 || block == 256;
}

hipError_t launch_specialized_values(const SpecializedTape& k, const float* x,
                                     const float* y, const float* z,
                                     float* out, size_t n, hipStream_t s, unsigned block) {
    if (n == 0) return hipSuccess;
    if (!valid_block(block)) return hipErrorInvalidValue;
    unsigned long long nn = n;
    void* args[] = {&x, &y, &z, &out, &nn};
    const unsigned grid = unsigned((n + block - 1) / block);
    return hipModuleLaunchKernel(k.values, grid, 1, 1, block, 1, 1, 0, s, args, nullptr);
}

hipError_t launch_specialized_derivs(const SpecializedTape& k, const float* x,
                                     const float* y, const float* z, float* v,
                                     float* dx, float* dy, float* dz, size_t n,
                                     hipStream_t s, unsigned block) {
    if (n == 0) return hipSuccess;
    if (!valid_block(block)) return hipErrorInvalidValue;
    unsigned long long nn = n;
    void* args[] = {&x, &y, &z, &v, &dx, &dy, &dz, &nn};
    const unsigned grid = unsigned((n + block - 1) / block);
    return hipModuleLaunchKernel(k.derivs, grid, 1, 1, block, 1, 1, 0, s, args, nullptr);
}

}   // namespace amd
}   // namespace libfive
