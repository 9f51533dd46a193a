// Tape specialisation: a device tape turned into straight-line HIP source,
// compiled at run time with hipRTC for gfx950 and loaded as a module.
//
// The interpreters (eval_kernels.hip) walk any tape with one wave-uniform
// clause stream, paying for decode, dispatch and register-file indexing on
// every clause.  For a tape that is evaluated over many points (the
// ArrayEvaluator / DerivArrayEvaluator workloads of
// libfive/src/eval/eval_array.cpp:107-118 and
// libfive/src/eval/eval_deriv_array.cpp:102-117) this module emits one
// kernel per tape in which every clause is a plain expression on named
// values, so the compiler allocates registers and nothing is interpreted.
// Each clause is written with exactly the operations, operand order and
// rounding of the interpreter (fmath.hpp, contraction off), so the two
// paths produce identical bits.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "deck.hpp"

namespace libfive {
namespace amd {

struct SpecializedTape {
    hipModule_t module = nullptr;
    hipFunction_t values = nullptr;    // (x, y, z, out, n)
    hipFunction_t derivs = nullptr;    // (x, y, z, v, dx, dy, dz, n)
    // The kernels are call-free and mark points they cannot evaluate with
    // tape::kRedoMark: every launch is followed by launch_eval_redo.
    bool redo = false;
    int deriv_waves = 0;               // occupancy requested for lf_derivs (0: none)
    double compile_ms = 0;
    size_t code_bytes = 0;
    uint64_t serial = 0;               // distinct for every specialised tape
    ~SpecializedTape();
};

// HIP source of the two kernels for tape `t` (exposed for tests).
// deriv_waves > 0 asks the compiler for that many waves per SIMD in the
// derivative kernel.
std::string specialized_source(const DeviceTape& t, int deriv_waves = 0);

// Compiles the kernels for gfx950 (no device needed); throws Error.
std::vector<char> compile_specialized(const DeviceTape& t, int deriv_waves = 0);

// Compiles and loads the kernels on the current device; throws Error.
std::unique_ptr<SpecializedTape> specialize(const DeviceTape& t);

// block: workgroup size (64, 128 or 256; one point per lane), a launch
// schedule the autotuner picks per tape.
hipError_t launch_specialized_values(const SpecializedTape& k, const float* x,
                                     const float* y, const float* z,
                                     float* out, size_t n, hipStream_t s,
                                     unsigned block = 256);
hipError_t launch_specialized_derivs(const SpecializedTape& k, const float* x,
                                     const float* y, const float* z, float* v,
                                     float* dx, float* dy, float* dz, size_t n,
                                     hipStream_t s, unsigned block = 256);

}   // namespace amd
}   // namespace libfive
