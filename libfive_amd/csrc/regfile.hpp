// Per-lane register file for the tape interpreters.
//
// Tape clauses are the same for every lane of a wavefront, so register
// indices are wave-uniform.  The file is NV named banks of 32 floats, each
// an ext_vector_type(32) local variable; a uniform index selects the bank
// through a compile-time tree of selects and the element through relative
// VGPR addressing (s_set_gpr_idx_on + v_mov), so a read or write costs one
// VALU move and the whole file stays in VGPRs.
//
// The banks must be plain local variables: wrapped in a struct or an array
// they are not promoted out of scratch memory, and vectors wider than 32
// elements are lowered through scratch as well.  Hence the macros.  Kernels
// using them need __launch_bounds__(256) so that more than 128 VGPRs are
// available per lane.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace libfive {
namespace amd {
typedef float f32x32 __attribute__((ext_vector_type(32)));
}
}

// Declares banks P0..P7 (unused banks are eliminated by the compiler).
#define LF_RF_DECL(P)                                                        \
    ::libfive::amd::f32x32 P##0 = 0.0f, P##1 = 0.0f, P##2 = 0.0f,            \
        P##3 = 0.0f, P##4 = 0.0f, P##5 = 0.0f, P##6 = 0.0f, P##7 = 0.0f;     \
    (void)P##0; (void)P##1; (void)P##2; (void)P##3;                          \
    (void)P##4; (void)P##5; (void)P##6; (void)P##7

// Bank selection is written as uniform branches; the empty volatile asm
// statements keep the compiler from if-converting them into "read every
// bank and select", which costs a VALU move per bank.
#define LF_RF_BARRIER() __asm__ volatile("")

#define LF_RF_SEL2(P, a, b, hi, lo, r)                                       \
    do {                                                                     \
        if ((hi) & 1u) { LF_RF_BARRIER(); r = P##b[lo]; }                    \
        else { LF_RF_BARRIER(); r = P##a[lo]; }                              \
    } while (0)
#define LF_RF_SEL4(P, a, b, c, d, hi, lo, r)                                 \
    do {                                                                     \
        if ((hi) & 2u) LF_RF_SEL2(P, c, d, hi, lo, r);                       \
        else LF_RF_SEL2(P, a, b, hi, lo, r);                                 \
    } while (0)

// Value of register i (uniform) in a file of NV banks (NV in {1,2,4,8}).
#define LF_RF_GET(P, NV, i)                                                  \
    ({                                                                       \
        const uint32_t lo_ = uint32_t(i) & 31u, hi_ = uint32_t(i) >> 5;      \
        float r_;                                                            \
        if ((NV) == 1) r_ = P##0[lo_];                                       \
        else if ((NV) == 2) LF_RF_SEL2(P, 0, 1, hi_, lo_, r_);               \
        else if ((NV) == 4) LF_RF_SEL4(P, 0, 1, 2, 3, hi_, lo_, r_);         \
        else if (hi_ & 4u) LF_RF_SEL4(P, 4, 5, 6, 7, hi_, lo_, r_);          \
        else LF_RF_SEL4(P, 0, 1, 2, 3, hi_, lo_, r_);                        \
        r_;                                                                  \
    })

#define LF_RF_PUT2(P, a, b, hi, lo, x)                                       \
    do {                                                                     \
        if ((hi) & 1u) { LF_RF_BARRIER(); P##b[lo] = (x); }                  \
        else { LF_RF_BARRIER(); P##a[lo] = (x); }                            \
    } while (0)
#define LF_RF_PUT4(P, a, b, c, d, hi, lo, x)                                 \
    do {                                                                     \
        if ((hi) & 2u) LF_RF_PUT2(P, c, d, hi, lo, x);                       \
        else LF_RF_PUT2(P, a, b, hi, lo, x);                                 \
    } while (0)

// Writes x to register i (uniform).
#define LF_RF_SET(P, NV, i, x)                                               \
    do {                                                                     \
        const uint32_t lo_ = uint32_t(i) & 31u, hi_ = uint32_t(i) >> 5;      \
        const float x_ = (x);                                                \
        if ((NV) == 1) P##0[lo_] = x_;                                       \
        else if ((NV) == 2) LF_RF_PUT2(P, 0, 1, hi_, lo_, x_);               \
        else if ((NV) == 4) LF_RF_PUT4(P, 0, 1, 2, 3, hi_, lo_, x_);         \
        else if (hi_ & 4u) LF_RF_PUT4(P, 4, 5, 6, 7, hi_, lo_, x_);          \
        else LF_RF_PUT4(P, 0, 1, 2, 3, hi_, lo_, x_);                        \
    } while (0)
