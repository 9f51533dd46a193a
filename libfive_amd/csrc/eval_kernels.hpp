// Host-side launchers for the point evaluation kernels.  All pointers are
// device pointers; launches are asynchronous on `s`.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <vector>

#include "tape.hpp"

namespace libfive {
namespace amd {

// fast: the tape has only tape::fast_opcode opcodes (tape_is_fast): a
// call-free kernel, then a redo launch for the points it marked.
hipError_t launch_eval_values(const tape::Clause* t, uint32_t n, uint32_t root,
                              uint32_t nregs, bool fast, const float* x,
                              const float* y, const float* z, float* out,
                              size_t npts, hipStream_t s);

// The same with p points per lane (1, 2 or 4; 4 only for tapes of one
// register bank, 2 for at most two); value_points_options() lists the
// counts a tape allows, default_value_points() the one launch_eval_values
// uses (libfive_set_points_per_lane's value when set, value_points_forced).
hipError_t launch_eval_values_ppl(const tape::Clause* t, uint32_t n, uint32_t root,
                                  uint32_t nregs, bool fast, const float* x,
                                  const float* y, const float* z, float* out,
                                  size_t npts, hipStream_t s, int p);
std::vector<int> value_points_options(uint32_t nregs);
int default_value_points(uint32_t nregs);
bool value_points_forced();

// fast: the tape has only tape::fast_opcode opcodes (tape_is_fast).
hipError_t launch_eval_derivs(const tape::Clause* t, uint32_t n, uint32_t root,
                              uint32_t nregs, bool fast, const float* x,
                              const float* y, const float* z, float* v,
                              float* dx, float* dy, float* dz, size_t npts,
                              hipStream_t s);

// Evaluates again, with the general rules, the points that a call-free
// kernel (the fast derivative interpreter, or a specialised kernel of a
// tape of common opcodes) marked with tape::kRedoMark in `v`: values only
// when dx is null, else values and the three partials.
hipError_t launch_eval_redo(const tape::Clause* t, uint32_t n, uint32_t root,
                            uint32_t nregs, const float* x, const float* y,
                            const float* z, float* v, float* dx, float* dy,
                            float* dz, size_t npts, hipStream_t s);

// Per point: 1 where a MIN/MAX clause has equal operands (getAmbiguous), or
// with derivs = true equal operands with different partials
// (getAmbiguousDerivs).
hipError_t launch_eval_ambiguous(const tape::Clause* t, uint32_t n, uint32_t root,
                                 uint32_t nregs, bool derivs, const float* x,
                                 const float* y, const float* z, uint8_t* out,
                                 size_t npts, hipStream_t s);

// Largest point count of one libfive_tape_eval_values_noalloc call, and
// the device address of its staging array (x, y, z, values at strides of
// kNoAllocMaxPoints floats) on the current device.
constexpr size_t kNoAllocMaxPoints = size_t(1) << 20;
hipError_t noalloc_scratch_address(float** out);

// Points per lane for the value kernel (0: automatic).
void set_points_per_lane(int p);
// Largest register count each interpreter keeps in VGPRs.
int max_value_regs();
int max_deriv_regs();

}   // namespace amd
}   // namespace libfive
