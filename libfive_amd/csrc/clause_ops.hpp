// Per-clause evaluation rules shared by the tape interpreters
// (eval_kernels.hip) and the specialised kernels built at run time
// (tape_jit.cpp), so that both produce identical bits.
//
// Values follow ArrayEvaluator (libfive/src/eval/eval_array.cpp:221-369),
// partial derivatives DerivArrayEvaluator
// (libfive/src/eval/eval_deriv_array.cpp:119-256).  Build without
// floating-point contraction.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <cstdint>
#endif

#include "fmath.hpp"
#include "opcode.hpp"
#include "tape.hpp"

namespace libfive {
namespace amd {

namespace OC = Opcode;
using tape::OP_COPY;
using tape::OP_LOADI;

// Values of one clause for P points (b is ignored by unary opcodes).  The
// common opcodes are inline; the rest go through the out-of-line scalar
// table so that their register demand does not lower occupancy.
// SIN and COS without the out-of-line large-argument path, for kernels
// that call no function: the results of apply_value / apply_deriv for
// arguments below 2^17 in magnitude.  `amax` accumulates the largest
// argument magnitude (NaN arguments are skipped by fmaxf, except that a
// signalling NaN makes it NaN); the caller redoes the point with the
// general rules unless amax < 2^17 at the end.
__device__ __forceinline__ float value_sincos_fast(uint32_t o, float a,
                                                   float& amax) {
    float s, c;
    fmath::sincos_core(a, &s, &c);
    amax = fmaxf(amax, fabsf(a));
    return (o == OC::OP_SIN) ? s : c;
}

__device__ __forceinline__ void deriv_sincos_fast(uint32_t o, float av,
                                                  const float ad[3], float& ov,
                                                  float od[3], float& amax) {
    float s, c;
    fmath::sincos_core(av, &s, &c);
    amax = fmaxf(amax, fabsf(av));
    if (o == OC::OP_SIN) {
        ov = s;
        for (int j = 0; j < 3; ++j) od[j] = ad[j] * c;
    } else {
        ov = c;
        const float ns = -s;
        for (int j = 0; j < 3; ++j) od[j] = ad[j] * ns;
    }
}

// FAST: the tape has only the common opcodes (tape::fast_opcode); sin and
// cos take only the single-precision path and accumulate the argument
// magnitude in amax[p] (see value_sincos_fast), and nothing is called.
template <int P, bool FAST = false>
__device__ __forceinline__ void apply_value(uint32_t o, const float* a,
                                            const float* b, float im,
                                            float* out, float* amax = nullptr) {
#define LF_EACH for (int p = 0; p < P; ++p)
    switch (o) {
        case OC::OP_ADD: LF_EACH out[p] = a[p] + b[p]; break;
        case OC::OP_MUL: LF_EACH out[p] = a[p] * b[p]; break;
        case OC::OP_MIN: LF_EACH out[p] = (b[p] < a[p]) ? b[p] : a[p]; break;
        case OC::OP_MAX: LF_EACH out[p] = (a[p] < b[p]) ? b[p] : a[p]; break;
        case OC::OP_SUB: LF_EACH out[p] = a[p] - b[p]; break;
        case OC::OP_NEG: LF_EACH out[p] = -a[p]; break;
        case OC::OP_SQUARE: LF_EACH out[p] = a[p] * a[p]; break;
        case OC::OP_ABS: LF_EACH out[p] = fabsf(a[p]); break;
        case OC::OP_SQRT: LF_EACH out[p] = sqrtf(a[p]); break;
        case OC::OP_SIN:
        case OC::OP_COS:
            if (FAST) {
                LF_EACH out[p] = value_sincos_fast(o, a[p], amax[p]);
            } else if (o == OC::OP_SIN) {
                LF_EACH out[p] = fmath::sin_f32(a[p]);
            } else {
                LF_EACH out[p] = fmath::cos_f32(a[p]);
            }
            break;
        case OC::OP_DIV: LF_EACH out[p] = a[p] / b[p]; break;
        case OP_LOADI: LF_EACH out[p] = im; break;
        case OP_COPY:
        case OC::CONST_VAR: LF_EACH out[p] = a[p]; break;
        default:
            if (FAST) {     // not reached: a fast tape has no other opcode
                LF_EACH out[p] = 0.0f;
            } else {
                LF_EACH out[p] = fmath::eval_op(int(o), a[p], b[p]);
            }
            break;
    }
#undef LF_EACH
}

// Value and partials of one clause, returned in registers (a result
// array passed by pointer would give the kernels a stack frame).
struct ColdDeriv {
    float v, d0, d1, d2;
};

// Value and partial derivatives of one clause for the opcodes outside the
// common set (eval_deriv_array.cpp:119-256); out of line.
static __device__ __noinline__ ColdDeriv cold_deriv(uint32_t o, float av, float ad0,
                                        float ad1, float ad2, float bv,
                                        float bd0, float bd1, float bd2) {
    const float ad[3] = {ad0, ad1, ad2};
    const float bd[3] = {bd0, bd1, bd2};
    float ov;
    float od[3];
    switch (o) {
        case OC::OP_ATAN2: {
            ov = fmath::atan2_f32(av, bv);
            const float a2 = av * av;
            const float b2 = bv * bv;
            const float den = a2 + b2;
            for (int j = 0; j < 3; ++j) {
                const float p = ad[j] * bv;
                const float q = bd[j] * av;
                od[j] = (p - q) / den;
            }
            break;
        }
        case OC::OP_POW: {
            ov = fmath::pow_f32(av, bv);
            const float f = bv * fmath::pow_f32(av, bv - 1.0f);
            for (int j = 0; j < 3; ++j) od[j] = ad[j] * f;
            break;
        }
        case OC::OP_NTH_ROOT: {
            ov = fmath::nth_root_f32(av, bv);
            const float f = fmath::pow_f32(av, 1.0f / bv - 1.0f) / bv;
            for (int j = 0; j < 3; ++j) od[j] = (ad[j] == 0.0f) ? 0.0f : ad[j] * f;
            break;
        }
        case OC::OP_MOD:
            ov = fmath::mod_f32(av, bv);
            for (int j = 0; j < 3; ++j) od[j] = ad[j];
            break;
        case OC::OP_NANFILL: {
            const bool n_ = fmath::isnan_f(av);
            ov = n_ ? bv : av;
            for (int j = 0; j < 3; ++j) od[j] = n_ ? bd[j] : ad[j];
            break;
        }
        case OC::OP_COMPARE:
            ov = (av < bv) ? -1.0f : (av > bv) ? 1.0f : 0.0f;
            for (int j = 0; j < 3; ++j) od[j] = 0.0f;
            break;
        case OC::OP_TAN: {
            ov = fmath::tan_f32(av);
            const float r = 1.0f / fmath::cos_f32(av);
            const float f = r * r;
            for (int j = 0; j < 3; ++j) od[j] = ad[j] * f;
            break;
        }
        case OC::OP_ASIN: {
            ov = fmath::asin_f32(av);
            const float s = sqrtf(1.0f - av * av);
            for (int j = 0; j < 3; ++j) od[j] = ad[j] / s;
            break;
        }
        case OC::OP_ACOS: {
            ov = fmath::acos_f32(av);
            const float s = -sqrtf(1.0f - av * av);
            for (int j = 0; j < 3; ++j) od[j] = ad[j] / s;
            break;
        }
        case OC::OP_ATAN: {
            ov = fmath::atan_f32(av);
            const float s = av * av + 1.0f;
            for (int j = 0; j < 3; ++j) od[j] = ad[j] / s;
            break;
        }
        case OC::OP_LOG:
            ov = fmath::log_f32(av);
            for (int j = 0; j < 3; ++j) od[j] = ad[j] / av;
            break;
        case OC::OP_EXP:
            ov = fmath::exp_f32(av);
            for (int j = 0; j < 3; ++j) od[j] = ad[j] * ov;
            break;
        case OC::OP_RECIP: {
            ov = 1.0f / av;
            const float s = -(av * av);
            for (int j = 0; j < 3; ++j) od[j] = ad[j] / s;
            break;
        }
        default:
            ov = av;
            for (int j = 0; j < 3; ++j) od[j] = ad[j];
            break;
    }
    return {ov, od[0], od[1], od[2]};
}

// Value and partial derivatives of one clause (eval_deriv_array.cpp:
// 119-256); the value is the same as apply_value's.
__device__ __forceinline__ void apply_deriv(uint32_t o, float av,
                                            const float ad[3], float bv,
                                            const float bd[3], float im,
                                            float& ov, float od[3]) {
    switch (o) {
        case OC::OP_ADD:
            ov = av + bv;
            for (int j = 0; j < 3; ++j) od[j] = ad[j] + bd[j];
            break;
        case OC::OP_SUB:
            ov = av - bv;
            for (int j = 0; j < 3; ++j) od[j] = ad[j] - bd[j];
            break;
        case OC::OP_MUL:
            ov = av * bv;
            for (int j = 0; j < 3; ++j) {
                const float p = bd[j] * av;
                const float q = ad[j] * bv;
                od[j] = p + q;
            }
            break;
        case OC::OP_MIN: {
            ov = (bv < av) ? bv : av;
            const bool s = av < bv;
            for (int j = 0; j < 3; ++j) od[j] = s ? ad[j] : bd[j];
            break;
        }
        case OC::OP_MAX: {
            ov = (av < bv) ? bv : av;
            const bool s = av < bv;
            for (int j = 0; j < 3; ++j) od[j] = s ? bd[j] : ad[j];
            break;
        }
        case OC::OP_DIV: {
            ov = av / bv;
            const float b2 = bv * bv;
            for (int j = 0; j < 3; ++j) {
                const float p = ad[j] * bv;
                const float q = bd[j] * av;
                od[j] = (p - q) / b2;
            }
            break;
        }
        case OC::OP_SQUARE:
            ov = av * av;
            for (int j = 0; j < 3; ++j) od[j] = (ad[j] * av) * 2.0f;
            break;
        case OC::OP_SQRT: {
            ov = sqrtf(av);
            const float den = 2.0f * ov;
            for (int j = 0; j < 3; ++j) {
                od[j] = (av < 0.0f || ad[j] == 0.0f) ? 0.0f : ad[j] / den;
            }
            break;
        }
        case OC::OP_NEG:
            ov = -av;
            for (int j = 0; j < 3; ++j) od[j] = -ad[j];
            break;
        case OC::OP_ABS:
            ov = fabsf(av);
            for (int j = 0; j < 3; ++j) od[j] = (av > 0.0f) ? ad[j] : -ad[j];
            break;
        case OC::OP_SIN: {
            float s, co;
            fmath::sin_cos_f32(av, &s, &co);
            ov = s;
            for (int j = 0; j < 3; ++j) od[j] = ad[j] * co;
            break;
        }
        case OC::OP_COS: {
            float s, co;
            fmath::sin_cos_f32(av, &s, &co);
            ov = co;
            const float ns = -s;
            for (int j = 0; j < 3; ++j) od[j] = ad[j] * ns;
            break;
        }
        case OP_LOADI:
            ov = im;
            for (int j = 0; j < 3; ++j) od[j] = 0.0f;
            break;
        case OP_COPY:
        case OC::CONST_VAR:
            ov = av;
            for (int j = 0; j < 3; ++j) od[j] = ad[j];
            break;
        default: {
            const ColdDeriv r = cold_deriv(o, av, ad[0], ad[1], ad[2], bv, bd[0], bd[1], bd[2]);
            ov = r.v;
            od[0] = r.d0;
            od[1] = r.d1;
            od[2] = r.d2;
            break;
        }
    }
}

}   // namespace amd
}   // namespace libfive
