// Wave-level tape interpreters for the mesher: every lane evaluates its own
// point with the same (wave-uniform) tape, as the point kernels do, and
// additionally reports, per lane, whether some MIN/MAX clause saw
//   - equal operands (bit 0: ArrayEvaluator::getAmbiguous,
//     libfive/src/eval/eval_array.cpp:198-217), or
//   - operands of which neither is strictly greater, NaN included (bit 1:
//     the clauses a SPECIALIZED push keeps, eval_array.cpp:121-170).
// Device-only header; include from .hip files compiled with
// -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>

#include "fmath.hpp"
#include "opcode.hpp"
#include "regfile.hpp"
#include "tape.hpp"

namespace libfive {
namespace amd {

enum : uint32_t { PT_AMBIGUOUS = 1u, PT_TIE = 2u, PT_SLOW = 4u };

// FAST = true: the evaluators call no function, so that the kernels using
// them need no stack.  They then handle only the common opcodes (the
// caller checks that the tape has no other) and sin / cos arguments below
// 2^17; anything else sets PT_SLOW in the flags (with an unspecified value)
// and the caller redoes the point with FAST = false.  tape::fast_opcode
// lists the common opcodes.


struct Clause4x {
    tape::Clause c[tape::kFetch];
};

// Clause group g of a tape, read through the constant address space: no
// kernel writes a tape it evaluates, and the tape pointer is wave-uniform,
// so the reads are scalar loads (through a generic pointer the compiler
// must assume the kernel's own stores may alias the tape, and reads it
// with vector loads).  Tapes are followed by at least kFetch clauses of
// padding, so group g + 1 of a tape can always be read ahead.
typedef const __attribute__((address_space(4))) uint32_t* ConstWordPtr;
__device__ __forceinline__ Clause4x pe_fetch4(const tape::Clause* t, uint32_t g) {
    const ConstWordPtr w = (ConstWordPtr)(t) + 4u * tape::kFetch * g;
    Clause4x r;
#pragma unroll
    for (uint32_t k = 0; k < tape::kFetch; ++k) {
        r.c[k].w0 = w[4u * k];
        r.c[k].w1 = w[4u * k + 1u];
        r.c[k].w2 = w[4u * k + 2u];
        r.c[k].w3 = w[4u * k + 3u];
    }
    return r;
}

template <bool FAST = false>
__device__ __forceinline__ float pe_value_op(uint32_t o, float a, float b,
                                             float im, uint32_t& flags) {
    namespace OC = Opcode;
    if (FAST) {
        switch (o) {
            case OC::OP_ADD: return a + b;
            case OC::OP_MUL: return a * b;
            case OC::OP_MIN: return (b < a) ? b : a;
            case OC::OP_MAX: return (a < b) ? b : a;
            case OC::OP_SUB: return a - b;
            case OC::OP_NEG: return -a;
            case OC::OP_SQUARE: return a * a;
            case OC::OP_ABS: return fabsf(a);
            case OC::OP_SQRT: return sqrtf(a);
            case OC::OP_SIN:
            case OC::OP_COS: {
                float sv, cv;
                if (!fmath::sincos_fast(a, &sv, &cv)) flags |= PT_SLOW;
                return (o == OC::OP_SIN) ? sv : cv;
            }
            case OC::OP_DIV: return a / b;
            case tape::OP_LOADI: return im;
            case tape::OP_COPY:
            case OC::CONST_VAR: return a;
            default:
                flags |= PT_SLOW;
                return 0.0f;
        }
    }
    switch (o) {
        case OC::OP_ADD: return a + b;
        case OC::OP_MUL: return a * b;
        case OC::OP_MIN: return (b < a) ? b : a;
        case OC::OP_MAX: return (a < b) ? b : a;
        case OC::OP_SUB: return a - b;
        case OC::OP_NEG: return -a;
        case OC::OP_SQUARE: return a * a;
        case OC::OP_ABS: return fabsf(a);
        case OC::OP_SQRT: return sqrtf(a);
        case OC::OP_SIN: return fmath::sin_f32(a);
        case OC::OP_COS: return fmath::cos_f32(a);
        case OC::OP_DIV: return a / b;
        case tape::OP_LOADI: return im;
        case tape::OP_COPY:
        case OC::CONST_VAR: return a;
        default: return fmath::eval_op(int(o), a, b);
    }
}

__device__ __forceinline__ float pe_value_op(uint32_t o, float a, float b,
                                             float im) {
    uint32_t unused = 0;
    return pe_value_op<false>(o, a, b, im, unused);
}

__device__ __forceinline__ uint32_t pe_minmax_flags(uint32_t o, float a,
                                                    float b) {
    if (o != Opcode::OP_MIN && o != Opcode::OP_MAX) return 0u;
    return ((a == b) ? PT_AMBIGUOUS : 0u) | ((!(a > b) && !(b > a)) ? PT_TIE : 0u);
}

#define LF_PE_VALUE_STEP(C)                                                  \
    do {                                                                     \
        const tape::Clause c_ = (C);                                         \
        const uint32_t o_ = tape::op(c_);                                    \
        if (o_ != tape::OP_NOP) {                                            \
            const uint32_t m_ = tape::mode(c_);                              \
            const float im_ = tape::imm(c_);                                 \
            const float a_ = (m_ == tape::MODE_A_IMM || o_ == tape::OP_LOADI) \
                ? im_ : LF_RF_GET(r, NV, tape::ra(c_));                      \
            float b_ = 0.0f;                                                 \
            if (tape::is_binary(c_)) {                                       \
                b_ = (m_ == tape::MODE_B_IMM) ? im_                          \
                                              : LF_RF_GET(r, NV, tape::rb(c_)); \
            }                                                                \
            flags |= pe_minmax_flags(o_, a_, b_);                            \
            LF_RF_SET(r, NV, tape::dst(c_),                                  \
                      pe_value_op<FAST>(o_, a_, b_, im_, flags));            \
        }                                                                    \
    } while (0)

// Value of the tape at (x, y, z); *flags gets PT_* bits.
template <int NV, bool FAST = false>
__device__ __forceinline__ float pe_value(const tape::Clause* __restrict__ t,
                                          uint32_t n, uint32_t root, float x,
                                          float y, float z, uint32_t* out_flags)
{
    LF_RF_DECL(r);
    LF_RF_SET(r, NV, tape::kRegX, x);
    LF_RF_SET(r, NV, tape::kRegY, y);
    LF_RF_SET(r, NV, tape::kRegZ, z);
    uint32_t flags = 0;
    Clause4x cur = pe_fetch4(t, 0);
    for (uint32_t k = 0; k < n; k += tape::kFetch) {
        const Clause4x nxt = pe_fetch4(t, k / tape::kFetch + 1);
        LF_PE_VALUE_STEP(cur.c[0]);
        LF_PE_VALUE_STEP(cur.c[1]);
        LF_PE_VALUE_STEP(cur.c[2]);
        LF_PE_VALUE_STEP(cur.c[3]);
        cur = nxt;
    }
    *out_flags = flags;
    return LF_RF_GET(r, NV, root);
}

// Value and partials of one clause, returned in registers (a result
// array passed by pointer would give the kernels a stack frame).
struct PeColdDeriv {
    float v, d0, d1, d2;
};

// Out-of-line derivative rules of the less common opcodes
// (eval_deriv_array.cpp:119-256).
static __device__ __noinline__ PeColdDeriv pe_cold_deriv(uint32_t o, float av, float ad0,
                                           float ad1, float ad2, float bv,
                                           float bd0, float bd1, float bd2) {
    namespace OC = Opcode;
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

// Value and partial derivatives of one clause (the rules of
// DerivArrayEvaluator, libfive/src/eval/eval_deriv_array.cpp:119-256).
template <bool FAST = false>
__device__ __forceinline__ void pe_deriv_op(uint32_t o, float av,
                                            const float ad[3], float bv,
                                            const float bd[3], float im,
                                            float& ov, float od[3],
                                            uint32_t& flags) {
    namespace OC = Opcode;
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
            if (FAST) {
                if (!fmath::sincos_fast(av, &s, &co)) flags |= PT_SLOW;
            } else {
                fmath::sin_cos_f32(av, &s, &co);
            }
            ov = s;
            for (int j = 0; j < 3; ++j) od[j] = ad[j] * co;
            break;
        }
        case OC::OP_COS: {
            float s, co;
            if (FAST) {
                if (!fmath::sincos_fast(av, &s, &co)) flags |= PT_SLOW;
            } else {
                fmath::sin_cos_f32(av, &s, &co);
            }
            ov = co;
            const float ns = -s;
            for (int j = 0; j < 3; ++j) od[j] = ad[j] * ns;
            break;
        }
        case tape::OP_LOADI:
            ov = im;
            for (int j = 0; j < 3; ++j) od[j] = 0.0f;
            break;
        case tape::OP_COPY:
        case OC::CONST_VAR:
            ov = av;
            for (int j = 0; j < 3; ++j) od[j] = ad[j];
            break;
        default: {
            if (FAST) {
                flags |= PT_SLOW;
                ov = 0.0f;
                for (int j = 0; j < 3; ++j) od[j] = 0.0f;
                break;
            }
            const PeColdDeriv r = pe_cold_deriv(o, av, ad[0], ad[1], ad[2], bv, bd[0], bd[1], bd[2]);
            ov = r.v;
            od[0] = r.d0;
            od[1] = r.d1;
            od[2] = r.d2;
            break;
        }
    }
}

#define LF_PE_DERIV_STEP(C)                                                  \
    do {                                                                     \
        const tape::Clause c_ = (C);                                         \
        const uint32_t o_ = tape::op(c_);                                    \
        if (o_ != tape::OP_NOP) {                                            \
            const uint32_t m_ = tape::mode(c_);                              \
            const float im_ = tape::imm(c_);                                 \
            float av_ = im_, ad_[3] = {0.0f, 0.0f, 0.0f};                    \
            float bv_ = 0.0f, bd_[3] = {0.0f, 0.0f, 0.0f};                   \
            if (m_ != tape::MODE_A_IMM && o_ != tape::OP_LOADI) {            \
                const uint32_t a_ = tape::ra(c_);                            \
                av_ = LF_RF_GET(v, NV, a_);                                  \
                ad_[0] = LF_RF_GET(gx, NV, a_);                              \
                ad_[1] = LF_RF_GET(gy, NV, a_);                              \
                ad_[2] = LF_RF_GET(gz, NV, a_);                              \
            }                                                                \
            if (tape::is_binary(c_)) {                                       \
                if (m_ == tape::MODE_B_IMM) {                                \
                    bv_ = im_;                                               \
                } else {                                                     \
                    const uint32_t b_ = tape::rb(c_);                        \
                    bv_ = LF_RF_GET(v, NV, b_);                              \
                    bd_[0] = LF_RF_GET(gx, NV, b_);                          \
                    bd_[1] = LF_RF_GET(gy, NV, b_);                          \
                    bd_[2] = LF_RF_GET(gz, NV, b_);                          \
                }                                                            \
            }                                                                \
            flags |= pe_minmax_flags(o_, av_, bv_);                          \
            float ov_, od_[3];                                               \
            pe_deriv_op<FAST>(o_, av_, ad_, bv_, bd_, im_, ov_, od_, flags);              \
            const uint32_t d_ = tape::dst(c_);                               \
            LF_RF_SET(v, NV, d_, ov_);                                       \
            LF_RF_SET(gx, NV, d_, od_[0]);                                   \
            LF_RF_SET(gy, NV, d_, od_[1]);                                   \
            LF_RF_SET(gz, NV, d_, od_[2]);                                   \
        }                                                                    \
    } while (0)

// Value and gradient of the tape at (x, y, z).
template <int NV, bool FAST = false>
__device__ __forceinline__ float pe_deriv(const tape::Clause* __restrict__ t,
                                          uint32_t n, uint32_t root, float x,
                                          float y, float z, float g[3],
                                          uint32_t* out_flags)
{
    LF_RF_DECL(v);
    LF_RF_DECL(gx);
    LF_RF_DECL(gy);
    LF_RF_DECL(gz);
    LF_RF_SET(v, NV, tape::kRegX, x);
    LF_RF_SET(v, NV, tape::kRegY, y);
    LF_RF_SET(v, NV, tape::kRegZ, z);
    LF_RF_SET(gx, NV, tape::kRegX, 1.0f);
    LF_RF_SET(gy, NV, tape::kRegY, 1.0f);
    LF_RF_SET(gz, NV, tape::kRegZ, 1.0f);
    uint32_t flags = 0;
    Clause4x cur = pe_fetch4(t, 0);
    for (uint32_t k = 0; k < n; k += tape::kFetch) {
        const Clause4x nxt = pe_fetch4(t, k / tape::kFetch + 1);
        LF_PE_DERIV_STEP(cur.c[0]);
        LF_PE_DERIV_STEP(cur.c[1]);
        LF_PE_DERIV_STEP(cur.c[2]);
        LF_PE_DERIV_STEP(cur.c[3]);
        cur = nxt;
    }
    *out_flags = flags;
    g[0] = LF_RF_GET(gx, NV, root);
    g[1] = LF_RF_GET(gy, NV, root);
    g[2] = LF_RF_GET(gz, NV, root);
    return LF_RF_GET(v, NV, root);
}

// Value and one partial derivative (axis 0, 1, 2 for x, y, z): the same
// rules with one gradient component carried, so the register file is two
// banks instead of four; the component is bit-identical to pe_deriv's.
#define LF_PE_DERIV1_STEP(C)                                                 \
    do {                                                                     \
        const tape::Clause c_ = (C);                                         \
        const uint32_t o_ = tape::op(c_);                                    \
        if (o_ != tape::OP_NOP) {                                            \
            const uint32_t m_ = tape::mode(c_);                              \
            const float im_ = tape::imm(c_);                                 \
            float av_ = im_, ad_[3] = {0.0f, 0.0f, 0.0f};                    \
            float bv_ = 0.0f, bd_[3] = {0.0f, 0.0f, 0.0f};                   \
            if (m_ != tape::MODE_A_IMM && o_ != tape::OP_LOADI) {            \
                const uint32_t a_ = tape::ra(c_);                            \
                av_ = LF_RF_GET(v, NV, a_);                                  \
                ad_[0] = LF_RF_GET(gd, NV, a_);                              \
            }                                                                \
            if (tape::is_binary(c_)) {                                       \
                if (m_ == tape::MODE_B_IMM) {                                \
                    bv_ = im_;                                               \
                } else {                                                     \
                    const uint32_t b_ = tape::rb(c_);                        \
                    bv_ = LF_RF_GET(v, NV, b_);                              \
                    bd_[0] = LF_RF_GET(gd, NV, b_);                          \
                }                                                            \
            }                                                                \
            flags |= pe_minmax_flags(o_, av_, bv_);                          \
            float ov_, od_[3];                                               \
            pe_deriv_op<FAST>(o_, av_, ad_, bv_, bd_, im_, ov_, od_, flags);              \
            const uint32_t d_ = tape::dst(c_);                               \
            LF_RF_SET(v, NV, d_, ov_);                                       \
            LF_RF_SET(gd, NV, d_, od_[0]);                                   \
        }                                                                    \
    } while (0)

template <int NV, bool FAST = false>
__device__ __forceinline__ float pe_deriv1(const tape::Clause* __restrict__ t,
                                           uint32_t n, uint32_t root, float x,
                                           float y, float z, int axis,
                                           float* g, uint32_t* out_flags)
{
    LF_RF_DECL(v);
    LF_RF_DECL(gd);
    LF_RF_SET(v, NV, tape::kRegX, x);
    LF_RF_SET(v, NV, tape::kRegY, y);
    LF_RF_SET(v, NV, tape::kRegZ, z);
    LF_RF_SET(gd, NV, tape::kRegX, axis == 0 ? 1.0f : 0.0f);
    LF_RF_SET(gd, NV, tape::kRegY, axis == 1 ? 1.0f : 0.0f);
    LF_RF_SET(gd, NV, tape::kRegZ, axis == 2 ? 1.0f : 0.0f);
    uint32_t flags = 0;
    Clause4x cur = pe_fetch4(t, 0);
    for (uint32_t k = 0; k < n; k += tape::kFetch) {
        const Clause4x nxt = pe_fetch4(t, k / tape::kFetch + 1);
        LF_PE_DERIV1_STEP(cur.c[0]);
        LF_PE_DERIV1_STEP(cur.c[1]);
        LF_PE_DERIV1_STEP(cur.c[2]);
        LF_PE_DERIV1_STEP(cur.c[3]);
        cur = nxt;
    }
    *out_flags = flags;
    *g = LF_RF_GET(gd, NV, root);
    return LF_RF_GET(v, NV, root);
}

// Value and gradient as three single-component passes (pe_deriv's bits).
template <int NV, bool FAST = false>
__device__ __forceinline__ float pe_grad3(const tape::Clause* __restrict__ t,
                                          uint32_t n, uint32_t root, float x,
                                          float y, float z, float g[3],
                                          uint32_t* out_flags)
{
    float v = 0.0f;
#pragma unroll 1
    for (int axis = 0; axis < 3; ++axis) {
        float ga;
        v = pe_deriv1<NV, FAST>(t, n, root, x, y, z, axis, &ga, out_flags);
        if (axis == 0) g[0] = ga;
        else if (axis == 1) g[1] = ga;
        else g[2] = ga;
    }
    return v;
}

}   // namespace amd
}   // namespace libfive
