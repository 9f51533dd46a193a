// Scalar float semantics of every tape opcode, shared by the HIP kernels and
// by host-side constant folding so that both produce identical bits.
//
// Value semantics follow ArrayEvaluator::operator()
// (libfive/src/eval/eval_array.cpp:221-369):
//   - ADD/SUB/MUL/DIV/SQRT are single IEEE operations (round to nearest);
//   - MIN is (b < a) ? b : a and MAX is (a < b) ? b : a, i.e. std::min /
//     std::max, so a NaN in the first operand propagates and a NaN in the
//     second one does not (libfive/include/libfive/eval/interval.hpp:85-93);
//   - MOD follows Python semantics with the reference's clamps
//     (eval_array.cpp:269-296);
//   - NANFILL, COMPARE, CONST_VAR, ABS, NEG, SQUARE, RECIP as written there.
//
// Transcendentals.  The reference takes sin/cos/tan/... from Eigen's
// vectorised kernels, whose bit patterns depend on an unpinned Eigen
// version.  Here sin and cos use the routine below (documented accuracy:
// at most 1 ulp from the correctly rounded result, checked against float64
// numpy in the tests); every other transcendental is evaluated in double
// precision and rounded once to float.
//
// This header compiles as HIP device code and as plain C++ (host folding,
// tests).  Nothing here may be built with floating-point contraction: all
// fused multiply-adds are written explicitly with ::fma().
#pragma once

#ifndef __HIPCC_RTC__
#include <cstdint>
#include <cmath>
#endif

#if defined(__HIPCC_RTC__)
// Run-time compilation: hipRTC supplies the fixed-width integer types.
using __hip_internal::int32_t;
using __hip_internal::int64_t;
using __hip_internal::uint8_t;
using __hip_internal::uint16_t;
using __hip_internal::uint32_t;
using __hip_internal::uint64_t;
#define LF_HD __host__ __device__ inline __attribute__((always_inline))
#define LF_HD_COLD __host__ __device__ inline __attribute__((noinline))
#elif defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define LF_HD __host__ __device__ inline __attribute__((always_inline))
// Rarely executed helpers are kept out of line on the device so that their
// register demand does not set the occupancy of the interpreter loops.
#define LF_HD_COLD __host__ __device__ inline __attribute__((noinline))
#else
#define LF_HD inline
#define LF_HD_COLD inline
#endif

namespace libfive {
namespace fmath {

LF_HD uint32_t f2u(float f) { uint32_t u; __builtin_memcpy(&u, &f, 4); return u; }
LF_HD float u2f(uint32_t u) { float f; __builtin_memcpy(&f, &u, 4); return f; }
LF_HD uint64_t d2u(double f) { uint64_t u; __builtin_memcpy(&u, &f, 8); return u; }

LF_HD bool isnan_f(float f) { return (f2u(f) & 0x7fffffffu) > 0x7f800000u; }
LF_HD bool isinf_f(float f) { return (f2u(f) & 0x7fffffffu) == 0x7f800000u; }

// Next float toward -infinity / +infinity (NaN and the matching infinity
// pass through unchanged).
LF_HD float pred_f(float f) {
    uint32_t u = f2u(f);
    if (isnan_f(f) || u == 0xff800000u) return f;
    if (u == 0x00000000u || u == 0x80000000u) return u2f(0x80000001u);
    return (u & 0x80000000u) ? u2f(u + 1) : u2f(u - 1);
}
LF_HD float succ_f(float f) {
    uint32_t u = f2u(f);
    if (isnan_f(f) || u == 0x7f800000u) return f;
    if (u == 0x00000000u || u == 0x80000000u) return u2f(0x00000001u);
    return (u & 0x80000000u) ? u2f(u - 1) : u2f(u + 1);
}

////////////////////////////////////////////////////////////////////////////////
// sin / cos
//
// Method (Cody & Waite, "Software Manual for the Elementary Functions",
// 1980, additive range reduction; Payne & Hanek, SIGNUM Newsletter 18(1),
// 1983, for large arguments):
//   x = k*(pi/2) + r, |r| <= pi/4, evaluated in double precision;
//   sin(r) = r + r^3 p(r^2), cos(r) = 1 + r^2 q(r^2), with p of degree 3
//   and q of degree 4 interpolating (sin(r) - r) / r^3 and (cos(r) - 1) / r^2
//   at 2000 Chebyshev nodes of r^2 in [0, (pi/4)^2] (near-minimax; relative
//   error at most 1.9e-11 for sin and 3.2e-13 for cos on |r| <= pi/4, below
//   2^-12 of a float ulp);
//   the quadrant k mod 4 selects +-sin(r) or +-cos(r); one rounding to float.
// For |x| < 2^19 the reduction uses a two-double split of pi/2 (the first
// product is exact because k*PIO2_HI and x share the 2^-52 grid).  Larger
// arguments use an integer multiply by 96 bits of 2/pi taken at the
// exponent's offset.

// Bits of 2/pi after the binary point, most significant first (computed
// from Machin's formula with integer arithmetic; regenerated and checked by
// the tests).
#if defined(__HIPCC__) || defined(__HIP__) || defined(__HIPCC_RTC__)
__device__ __constant__
#endif
static const uint32_t kTwoOverPi[10] = {
    0xA2F9836Eu, 0x4E441529u, 0xFC2757D1u, 0xF534DDC0u, 0xDB629599u,
    0x3C439041u, 0xFE5163ABu, 0xDEBBC561u, 0xB7246E3Au, 0x424DD2E0u,
};

// Host copy for host-side callers (the __constant__ array above is a device
// symbol when compiled by hipcc).
static const uint32_t kTwoOverPiHost[10] = {
    0xA2F9836Eu, 0x4E441529u, 0xFC2757D1u, 0xF534DDC0u, 0xDB629599u,
    0x3C439041u, 0xFE5163ABu, 0xDEBBC561u, 0xB7246E3Au, 0x424DD2E0u,
};

constexpr double kPio2Hi = 1.5707963267948966;       // RN(pi/2)
constexpr double kPio2Lo = 6.123233995736766e-17;    // RN(pi/2 - kPio2Hi)
constexpr double kTwoOverPiD = 0.6366197723675814;   // RN(2/pi)

LF_HD uint32_t two_over_pi_word(const uint32_t* tbl, int bit) {
    // 32 bits of 2/pi starting at fractional bit `bit` (bit 1 is the first
    // bit after the binary point).
    int p = bit - 1;
    int w = p >> 5, s = p & 31;
    uint32_t hi = tbl[w];
    uint32_t lo = tbl[w + 1];
    return s ? ((hi << s) | (lo >> (32 - s))) : hi;
}

// Large-argument reduction for finite |x| >= 2^19: r (double, in
// [-pi/4, pi/4]) and the quadrant q (mod 4), returned by value so that the
// out-of-line call needs no stack memory.
struct Reduced {
    double r;
    int q;
};
LF_HD_COLD Reduced reduce_large(float ax, const uint32_t* tbl) {
    const uint32_t u = f2u(ax);
    const int e = int((u >> 23) & 0xff) - 150;          // ax = m * 2^e
    const uint64_t m = (u & 0x7fffffu) | 0x800000u;
    // Window of 96 bits of 2/pi starting at fractional bit i0.
    const int i0 = (e >= 2) ? (e - 1) : 1;
    const uint64_t w2 = two_over_pi_word(tbl, i0);
    const uint64_t w1 = two_over_pi_word(tbl, i0 + 32);
    const uint64_t w0 = two_over_pi_word(tbl, i0 + 64);
    // P = m * (w2:w1:w0), a 120-bit value held in (hi:lo) = 64:64 bits.
    const uint64_t p0 = m * w0;                 // < 2^56
    const uint64_t p1 = m * w1;
    const uint64_t p2 = m * w2;
    uint64_t lo = p0 + (p1 << 32);
    uint64_t carry = (lo < p0) ? 1u : 0u;
    uint64_t hi = (p1 >> 32) + p2 + carry;
    // Binary point sits s bits above the bottom of P.
    const int s = i0 + 95 - e;                  // 94 .. 100
    // Integer part mod 4: bits s and s+1 of P (s >= 64 always).
    int quad = int((hi >> (s - 64)) & 3u);
    // Fraction: bits [0, s) of P, scaled by 2^-s.
    const uint64_t fhi = hi & ((uint64_t(1) << (s - 64)) - 1);
    double f = double(fhi) * ::ldexp(1.0, 64 - s)
             + double(lo) * ::ldexp(1.0, -s);
    if (f >= 0.5) { f -= 1.0; quad += 1; }
    return {::fma(f, kPio2Hi, f * kPio2Lo), quad & 3};
}

// Coefficients of p (kS*) and q (kC*), constant term first.
constexpr double kS3 = -0.16666666663858068;
constexpr double kS5 = 0.008333331875520718;
constexpr double kS7 = -0.00019840087085699887;
constexpr double kS9 = 2.724996365903507e-06;
constexpr double kC2 = -0.49999999999963946;
constexpr double kC4 = 0.04166666663741224;
constexpr double kC6 = -0.0013888885092715444;
constexpr double kC8 = 2.4799862532640472e-05;
constexpr double kC10 = -2.7237168254380787e-07;

LF_HD void sincos_reduced(double r, double* s, double* c) {
    const double r2 = r * r;
    double ps = kS9;
    ps = ::fma(ps, r2, kS7);
    ps = ::fma(ps, r2, kS5);
    ps = ::fma(ps, r2, kS3);
    *s = ::fma(ps * r2, r, r);
    double pc = kC10;
    pc = ::fma(pc, r2, kC8);
    pc = ::fma(pc, r2, kC6);
    pc = ::fma(pc, r2, kC4);
    pc = ::fma(pc, r2, kC2);
    *c = ::fma(pc, r2, 1.0);
}

// sin(x) and cos(x) from one reduction, each within half an ulp (plus
// 2^-12 ulp) of the exact value: the bounds used by interval arithmetic,
// which widens them by one ulp to enclose the true range.
LF_HD void sincos_pair_accurate(float x, float* so, float* co,
                                const uint32_t* tbl) {
    if (isnan_f(x) || isinf_f(x)) {
        *so = x - x;    // NaN
        *co = x - x;
        return;
    }
    const float ax = ::fabsf(x);
    double r;
    int q;
    if (ax < 524288.0f) {
        const double xd = double(ax);
        const double k = ::rint(xd * kTwoOverPiD);
        r = ::fma(-k, kPio2Hi, xd);
        r = ::fma(-k, kPio2Lo, r);
        q = int(k) & 3;                 // 0 <= k < 2^19
    } else {
        const Reduced red = reduce_large(ax, tbl);
        r = red.r;
        q = red.q;
    }
    double s, c;
    sincos_reduced(r, &s, &c);
    // sin / cos of k*pi/2 + r: odd quadrants swap the two, quadrants 2 and
    // 3 negate sin, quadrants 1 and 2 negate cos; sin is odd in x.  Written
    // as two-way selects so that no branches are generated.
    const bool swap = (q & 1) != 0;
    double sv = swap ? c : s;
    double cv = swap ? s : c;
    const bool neg_s = ((q & 2) != 0) != ((f2u(x) & 0x80000000u) != 0);
    const bool neg_c = ((q + 1) & 2) != 0;
    sv = neg_s ? -sv : sv;
    cv = neg_c ? -cv : cv;
    *so = float(sv);
    *co = float(cv);
}

////////////////////////////////////////////////////////////////////////////////
// sin / cos of the tape evaluators, in single precision.
//
// Method (Cody & Waite 1980, additive reduction with a three-part constant;
// the products are exact or singly rounded by the fused multiply-add):
//   k = rint(|x| * 2/pi), r = |x| - k*P1 - k*P2 - k*P3 (three fmaf), where
//   P1 + P2 + P3 is pi/2 to 72 bits; for |x| < 2^17 the first step is
//   exact (|x| - k*P1 is a multiple of 2^-23 below 1), and the reduced
//   argument carries about one rounding of its own size;
//   sin(r) = r + r^3 (S1 + r^2 (S2 + r^2 S3)),
//   cos(r) = 1 + r^2 (C1 + r^2 (C2 + r^2 (C3 + r^2 C4))), coefficients
//   fitted to the relative error by iterated weighted least squares on
//   4000 Chebyshev nodes of [0, pi/4 + 1e-3] (fit error 3.9e-9 and 6.5e-11
//   before rounding to float); the quadrant k mod 4 swaps and negates.
// Arguments of 2^17 and above, and infinities and NaNs, take the double
// precision path above (Payne-Hanek reduction).  Checked against float64
// sin and cos on every float in [0, 2^17) and on samples beyond.
constexpr float kTwoOverPiF = 0.63661975f;            // RN(2/pi)
constexpr float kCwP1 = 1.5707963705062866f;          // RN(pi/2)
constexpr float kCwP2 = -4.371138828673793e-08f;      // RN(pi/2 - P1)
constexpr float kCwP3 = -1.7151245100058819e-15f;     // RN(pi/2 - P1 - P2)
constexpr float kSf1 = -0.16666653752326965f;
constexpr float kSf2 = 0.008332151919603348f;
constexpr float kSf3 = -0.0001951410959009081f;
constexpr float kCf1 = -0.5f;
constexpr float kCf2 = 0.04166661947965622f;
constexpr float kCf3 = -0.0013886665692552924f;
constexpr float kCf4 = 2.4382114133914e-05f;
constexpr float kCwLimit = 131072.0f;                 // 2^17

struct SinCos {
    float s, c;
};

// Arguments outside the single-precision reduction's range: |x| >= 2^17,
// infinities and NaNs (sincos_pair_accurate's results).
LF_HD_COLD SinCos sincos_large(float x, const uint32_t* tbl) {
    SinCos r;
    sincos_pair_accurate(x, &r.s, &r.c, tbl);
    return r;
}

// The single-precision path for |x| < 2^17 (for infinities and NaNs it
// gives NaNs).
LF_HD void sincos_core(float x, float* so, float* co) {
    const float ax = ::fabsf(x);
    const float k = ::rintf(ax * kTwoOverPiF);
    float r = ::fmaf(-k, kCwP1, ax);
    r = ::fmaf(-k, kCwP2, r);
    r = ::fmaf(-k, kCwP3, r);
    const uint32_t q = uint32_t(int32_t(k));
    const float r2 = r * r;
    float ps = ::fmaf(kSf3, r2, kSf2);
    ps = ::fmaf(ps, r2, kSf1);
    const float r3 = r2 * r;
    const float s = ::fmaf(r3, ps, r);
    float pc = ::fmaf(kCf4, r2, kCf3);
    pc = ::fmaf(pc, r2, kCf2);
    pc = ::fmaf(pc, r2, kCf1);
    const float c = ::fmaf(pc, r2, 1.0f);
    // sin / cos of k*pi/2 + r: odd quadrants swap the two, quadrants 2 and
    // 3 negate sin, quadrants 1 and 2 negate cos; sin is odd in x.  The
    // signs are applied as bit flips.
    const bool swap = (q & 1u) != 0;
    const float sv = swap ? c : s;
    const float cv = swap ? s : c;
    const uint32_t sneg = ((q << 30) ^ f2u(x)) & 0x80000000u;
    const uint32_t cneg = ((q + 1u) << 30) & 0x80000000u;
    *so = u2f(f2u(sv) ^ sneg);
    *co = u2f(f2u(cv) ^ cneg);
}

LF_HD void sincos_pair(float x, float* so, float* co, const uint32_t* tbl) {
    if (!(::fabsf(x) < kCwLimit)) {
        const SinCos r = sincos_large(x, tbl);
        *so = r.s;
        *co = r.c;
        return;
    }
    sincos_core(x, so, co);
}

// sincos_pair without the out-of-line large-argument path, for kernels
// that must not call functions: the same results, or false (results
// unspecified) for finite |x| >= 2^17.  Infinities and NaNs give NaNs.
LF_HD bool sincos_fast(float x, float* so, float* co) {
    sincos_core(x, so, co);
    const float ax = ::fabsf(x);
    return !(ax >= kCwLimit && ax <= 3.40282347e38f);
}

// which = 0 for sin, 1 for cos.
LF_HD float sincos_f32(float x, int which, const uint32_t* tbl) {
    float s, c;
    sincos_pair(x, &s, &c, tbl);
    return which ? c : s;
}

#if defined(__HIP_DEVICE_COMPILE__)
#define LF_TWO_OVER_PI kTwoOverPi
#else
#define LF_TWO_OVER_PI kTwoOverPiHost
#endif

LF_HD float sin_f32(float x) { return sincos_f32(x, 0, LF_TWO_OVER_PI); }
LF_HD float cos_f32(float x) { return sincos_f32(x, 1, LF_TWO_OVER_PI); }
LF_HD void sin_cos_f32(float x, float* s, float* c) {
    sincos_pair(x, s, c, LF_TWO_OVER_PI);
}
// The accurate routine, for interval bounds.
LF_HD float cos_f32_accurate(float x) {
    float s, c;
    sincos_pair_accurate(x, &s, &c, LF_TWO_OVER_PI);
    return c;
}

////////////////////////////////////////////////////////////////////////////////
// Remaining transcendentals: double evaluation, one rounding to float.
LF_HD float tan_f32(float a)  { return float(::tan(double(a))); }
LF_HD float asin_f32(float a) { return float(::asin(double(a))); }
LF_HD float acos_f32(float a) { return float(::acos(double(a))); }
LF_HD float atan_f32(float a) { return float(::atan(double(a))); }
LF_HD float exp_f32(float a)  { return float(::exp(double(a))); }
LF_HD float log_f32(float a)  { return float(::log(double(a))); }
LF_HD float atan2_f32(float a, float b) {
    return float(::atan2(double(a), double(b)));
}
LF_HD float pow_f32(float a, float b) {
    return float(::pow(double(a), double(b)));
}
// nth_root(a, b): b is an integer constant.  For a < 0 the reference uses
// Boost's interval nth_root on the point [a, a] (eval_array.cpp:255-266),
// which gives the real odd root and NaN for even roots.
LF_HD float nth_root_f32(float a, float b) {
    if (a < 0.0f) {
        const int n = int(b);
        if (!(n & 1)) return u2f(0x7fc00000u);
        return -float(::pow(-double(a), 1.0 / double(b)));
    }
    return float(::pow(double(a), 1.0 / double(b)));
}

// Python-style modulo with the reference's clamps (eval_array.cpp:269-296).
LF_HD float mod_f32(float a, float b) {
    float d = ::fabsf(a / b);
    if ((a < 0) != (b < 0)) {
        d = -::ceilf(d);
    } else {
        d = ::floorf(d);
    }
    float out = a - b * d;
    if ((b > 0 && out > b) || (b < 0 && out < b)) out = b;
    if ((b > 0.0f && out < 0.0f) || (b < 0.0f && out > 0.0f)) out = 0.0f;
    return out;
}

// One clause at one point (ArrayEvaluator::operator(), eval_array.cpp:221-369).
// Opcodes without a value (X/Y/Z, constants, variables, oracles) return NaN;
// they never appear as tape clauses.
LF_HD_COLD float eval_op(int op, float a, float b) {
    switch (op) {
        case 17: return a + b;                          // OP_ADD
        case 18: return a * b;                          // OP_MUL
        case 19: return (b < a) ? b : a;                // OP_MIN
        case 20: return (a < b) ? b : a;                // OP_MAX
        case 21: return a - b;                          // OP_SUB
        case 22: return a / b;                          // OP_DIV
        case 23: return atan2_f32(a, b);                // OP_ATAN2
        case 24: return pow_f32(a, b);                  // OP_POW
        case 25: return nth_root_f32(a, b);             // OP_NTH_ROOT
        case 26: return mod_f32(a, b);                  // OP_MOD
        case 27: return isnan_f(a) ? b : a;             // OP_NANFILL
        case 31: return (a < b) ? -1.0f : (a > b) ? 1.0f : 0.0f;  // OP_COMPARE
        case 7:  return a * a;                          // OP_SQUARE
        case 8:  return ::sqrtf(a);                       // OP_SQRT
        case 9:  return -a;                             // OP_NEG
        case 10: return sin_f32(a);                     // OP_SIN
        case 11: return cos_f32(a);                     // OP_COS
        case 12: return tan_f32(a);                     // OP_TAN
        case 13: return asin_f32(a);                    // OP_ASIN
        case 14: return acos_f32(a);                    // OP_ACOS
        case 15: return atan_f32(a);                    // OP_ATAN
        case 16: return exp_f32(a);                     // OP_EXP
        case 28: return ::fabsf(a);                       // OP_ABS
        case 29: return 1.0f / a;                       // OP_RECIP
        case 30: return log_f32(a);                     // OP_LOG
        case 6:  return a;                              // CONST_VAR
        default: return u2f(0x7fc00000u);
    }
}

}   // namespace fmath
}   // namespace libfive
