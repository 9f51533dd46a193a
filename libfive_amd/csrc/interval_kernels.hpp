// Host-side launcher for interval evaluation over boxes (device pointers).
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#include "tape.hpp"

namespace libfive {
namespace amd {

hipError_t launch_eval_intervals(const tape::Clause* t, uint32_t n,
                                 uint32_t root, uint32_t nregs,
                                 const float* xlo, const float* xhi,
                                 const float* ylo, const float* yhi,
                                 const float* zlo, const float* zhi,
                                 float* olo, float* ohi, uint8_t* onan,
                                 size_t nbox, hipStream_t s);

}   // namespace amd
}   // namespace libfive
