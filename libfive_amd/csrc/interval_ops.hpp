// Interval arithmetic of the IntervalEvaluator, device side.
//
// The reference wraps boost::numeric::interval<float> with policies
// save_state<rounded_transc_std<float>> + checking_base<float>
// (libfive/include/libfive/eval/interval.hpp:383-387) and adds a maybe_nan
// flag with its own propagation rules (interval.hpp:31-444).  Here:
//   - bounds of + - * are rounded outward with the hardware's directed
//     rounding modes (the ocml *_rtn / *_rtp entry points), and / and sqrt
//     by an exact double-precision method (rnd::dbl_dn/dbl_up), matching
//     what Boost obtains with fesetround;
//   - each operation follows Boost's case analysis (arith.hpp, arith2.hpp,
//     transc.hpp of Boost.Interval, restated below); an "empty" interval
//     is [NaN, NaN] and any empty input gives an empty result, as with
//     checking_base;
//   - Boost obtains transcendental bounds from std::sin/std::cos/... under
//     a directed rounding mode, whose results depend on the C library.
//     Here they are the round-to-nearest value moved one ulp outward,
//     which encloses the true value because the point routines are within
//     one ulp (fmath.hpp).
// The maybe_nan rules are the reference's, verbatim in meaning.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "fmath.hpp"

namespace libfive {
namespace amd {

struct Iv {
    float lo, hi;
    bool nan;
};

#define LF_DEV __device__ __forceinline__

namespace rnd {
// Quotients and square roots of floats are computed correctly rounded in
// double precision and then rounded to float in the wanted direction.  This
// equals directed rounding of the exact result: for float operands, a
// quotient or root that is not itself a float lies at least ~2^-48
// (relative) away from every float, so the double result is on the same
// side of each float boundary as the exact one.
LF_DEV float dbl_dn(double v) {
    const float f = float(v);
    return (double(f) > v) ? fmath::pred_f(f) : f;
}
LF_DEV float dbl_up(double v) {
    const float f = float(v);
    return (double(f) < v) ? fmath::succ_f(f) : f;
}
LF_DEV float add_dn(float a, float b) { return __ocml_add_rtn_f32(a, b); }
LF_DEV float add_up(float a, float b) { return __ocml_add_rtp_f32(a, b); }
LF_DEV float sub_dn(float a, float b) { return __ocml_sub_rtn_f32(a, b); }
LF_DEV float sub_up(float a, float b) { return __ocml_sub_rtp_f32(a, b); }
LF_DEV float mul_dn(float a, float b) { return __ocml_mul_rtn_f32(a, b); }
LF_DEV float mul_up(float a, float b) { return __ocml_mul_rtp_f32(a, b); }
LF_DEV float div_dn(float a, float b) { return dbl_dn(double(a) / double(b)); }
LF_DEV float div_up(float a, float b) { return dbl_up(double(a) / double(b)); }
LF_DEV float sqrt_dn(float a) { return dbl_dn(::sqrt(double(a))); }
LF_DEV float sqrt_up(float a) { return dbl_up(::sqrt(double(a))); }
// Outward-moved round-to-nearest transcendental bounds.
LF_DEV float dn(float v) { return fmath::pred_f(v); }
LF_DEV float up(float v) { return fmath::succ_f(v); }
}   // namespace rnd

constexpr float kInf = __builtin_huge_valf();

LF_DEV float nanv() { return fmath::u2f(0x7fc00000u); }
LF_DEV float fmin_std(float a, float b) { return (b < a) ? b : a; }  // std::min
LF_DEV float fmax_std(float a, float b) { return (a < b) ? b : a; }  // std::max

// checking_base::is_empty(l, u) == !(l <= u)
LF_DEV bool bad(float lo, float hi) { return !(lo <= hi); }
LF_DEV Iv empty_iv(bool nan) { return {nanv(), nanv(), nan}; }

// Boost constants (boost/numeric/interval/constants.hpp): float pi bounds
// 13176794/2^22 and 13176795/2^22.
constexpr float kPiLo = 13176794.0f / 4194304.0f;
constexpr float kPiHi = 13176795.0f / 4194304.0f;
constexpr float kPi2Lo = kPiLo * 2.0f, kPi2Hi = kPiHi * 2.0f;
constexpr float kPiHalfLo = kPiLo / 2.0f, kPiHalfHi = kPiHi / 2.0f;

// Interval(float lo, float hi): checked construction, maybe_nan from the
// arguments (interval.hpp:33-35).
LF_DEV Iv make_checked(float lo, float hi) {
    const bool n = fmath::isnan_f(lo) || fmath::isnan_f(hi);
    if (n || !(lo <= hi)) return {nanv(), nanv(), n};
    return {lo, hi, n};
}

LF_DEV Iv iv_add(Iv a, Iv b) {
    const bool u = a.nan || b.nan || (a.lo == -kInf && b.hi == kInf) ||
                   (b.lo == -kInf && a.hi == kInf);
    if (bad(a.lo, a.hi) || bad(b.lo, b.hi)) return empty_iv(u);
    return {rnd::add_dn(a.lo, b.lo), rnd::add_up(a.hi, b.hi), u};
}

LF_DEV Iv iv_sub(Iv a, Iv b) {
    const bool u = a.nan || b.nan || (a.lo == -kInf && b.lo == -kInf) ||
                   (a.hi == -kInf && b.hi == -kInf);
    if (bad(a.lo, a.hi) || bad(b.lo, b.hi)) return empty_iv(u);
    return {rnd::sub_dn(a.lo, b.hi), rnd::sub_up(a.hi, b.lo), u};
}

// Boost operator*(interval, interval): sign-case analysis (M: straddles 0,
// N: non-positive with negative lower, P: non-negative with positive
// upper, Z: [0,0]).
LF_DEV void mul_bounds(float xl, float xu, float yl, float yu, float& lo,
                       float& hi) {
    using namespace rnd;
    if (xl < 0) {
        if (xu > 0) {
            if (yl < 0) {
                if (yu > 0) {   // M * M
                    lo = fmin_std(mul_dn(xl, yu), mul_dn(xu, yl));
                    hi = fmax_std(mul_up(xl, yl), mul_up(xu, yu));
                } else {        // M * N
                    lo = mul_dn(xu, yl); hi = mul_up(xl, yl);
                }
            } else {
                if (yu > 0) {   // M * P
                    lo = mul_dn(xl, yu); hi = mul_up(xu, yu);
                } else {        // M * Z
                    lo = 0.0f; hi = 0.0f;
                }
            }
        } else {
            if (yl < 0) {
                if (yu > 0) {   // N * M
                    lo = mul_dn(xl, yu); hi = mul_up(xl, yl);
                } else {        // N * N
                    lo = mul_dn(xu, yu); hi = mul_up(xl, yl);
                }
            } else {
                if (yu > 0) {   // N * P
                    lo = mul_dn(xl, yu); hi = mul_up(xu, yl);
                } else {        // N * Z
                    lo = 0.0f; hi = 0.0f;
                }
            }
        }
    } else {
        if (xu > 0) {
            if (yl < 0) {
                if (yu > 0) {   // P * M
                    lo = mul_dn(xu, yl); hi = mul_up(xu, yu);
                } else {        // P * N
                    lo = mul_dn(xu, yl); hi = mul_up(xl, yu);
                }
            } else {
                if (yu > 0) {   // P * P
                    lo = mul_dn(xl, yl); hi = mul_up(xu, yu);
                } else {        // P * Z
                    lo = 0.0f; hi = 0.0f;
                }
            }
        } else {                // Z * ?
            lo = 0.0f; hi = 0.0f;
        }
    }
}

LF_DEV Iv iv_mul(Iv a, Iv b) {
    const bool u = a.nan || b.nan ||
        ((a.lo == -kInf || a.hi == kInf) && b.lo <= 0.0f && b.hi >= 0.0f) ||
        ((b.lo == -kInf || b.hi == kInf) && a.lo <= 0.0f && a.hi >= 0.0f);
    if (bad(a.lo, a.hi) || bad(b.lo, b.hi)) return empty_iv(u);
    Iv r;
    mul_bounds(a.lo, a.hi, b.lo, b.hi, r.lo, r.hi);
    r.nan = u;
    return r;
}

// Boost div_non_zero(interval, interval) (y does not contain 0).
LF_DEV void div_nz(float xl, float xu, float yl, float yu, float& lo,
                   float& hi) {
    using namespace rnd;
    if (xu < 0) {
        if (yu < 0) { lo = div_dn(xu, yl); hi = div_up(xl, yu); }
        else        { lo = div_dn(xl, yl); hi = div_up(xu, yu); }
    } else if (xl < 0) {
        if (yu < 0) { lo = div_dn(xu, yu); hi = div_up(xl, yu); }
        else        { lo = div_dn(xl, yl); hi = div_up(xu, yl); }
    } else {
        if (yu < 0) { lo = div_dn(xu, yu); hi = div_up(xl, yl); }
        else        { lo = div_dn(xl, yu); hi = div_up(xu, yl); }
    }
}

LF_DEV Iv iv_div(Iv a, Iv b) {
    const bool u = a.nan || b.nan ||
        ((a.lo == -kInf || a.hi == kInf) && (b.lo == -kInf || b.hi == kInf)) ||
        (a.lo <= 0.0f && a.hi >= 0.0f && b.lo <= 0.0f && b.hi >= 0.0f);
    if (b.lo <= 0.0f && b.hi >= 0.0f) return {-kInf, kInf, u};
    if (bad(a.lo, a.hi) || bad(b.lo, b.hi)) return empty_iv(u);
    Iv r;
    div_nz(a.lo, a.hi, b.lo, b.hi, r.lo, r.hi);
    r.nan = u;
    return r;
}

// hull (Boost): an empty operand is ignored.
LF_DEV void hull(float& lo, float& hi, float blo, float bhi) {
    const bool bx = bad(lo, hi), by = bad(blo, bhi);
    if (bx) {
        if (by) { lo = nanv(); hi = nanv(); }
        else { lo = blo; hi = bhi; }
        return;
    }
    if (by) return;
    lo = fmin_std(lo, blo);
    hi = fmax_std(hi, bhi);
}

// Interval::min / max (interval.hpp:82-131): bounds from std::min/std::max
// of the bounds; if b may be NaN, a's range is included; the flag is a's.
LF_DEV Iv iv_min(Iv a, Iv b) {
    float lo, hi;
    if (bad(a.lo, a.hi) || bad(b.lo, b.hi)) { lo = nanv(); hi = nanv(); }
    else { lo = fmin_std(a.lo, b.lo); hi = fmin_std(a.hi, b.hi); }
    if (b.nan) hull(lo, hi, a.lo, a.hi);
    return {lo, hi, a.nan};
}

LF_DEV Iv iv_max(Iv a, Iv b) {
    float lo, hi;
    if (bad(a.lo, a.hi) || bad(b.lo, b.hi)) { lo = nanv(); hi = nanv(); }
    else { lo = fmax_std(a.lo, b.lo); hi = fmax_std(a.hi, b.hi); }
    if (b.nan) hull(lo, hi, a.lo, a.hi);
    return {lo, hi, a.nan};
}

LF_DEV Iv iv_neg(Iv a) {
    if (bad(a.lo, a.hi)) return empty_iv(a.nan);
    return {-a.hi, -a.lo, a.nan};
}

LF_DEV Iv iv_square(Iv a) {
    if (bad(a.lo, a.hi)) return empty_iv(a.nan);
    using namespace rnd;
    if (a.hi < 0) return {mul_dn(a.hi, a.hi), mul_up(a.lo, a.lo), a.nan};
    if (a.lo > 0) return {mul_dn(a.lo, a.lo), mul_up(a.hi, a.hi), a.nan};
    return {0.0f, (-a.lo > a.hi) ? mul_up(a.lo, a.lo) : mul_up(a.hi, a.hi),
            a.nan};
}

LF_DEV Iv iv_sqrt(Iv a) {
    const bool u = a.nan || a.lo < 0.0f;
    if (bad(a.lo, a.hi) || a.hi < 0) return empty_iv(u);
    const float l = (a.lo > 0) ? rnd::sqrt_dn(a.lo) : 0.0f;
    return {l, rnd::sqrt_up(a.hi), u};
}

LF_DEV Iv iv_abs(Iv a) {
    if (bad(a.lo, a.hi)) return empty_iv(a.nan);
    if (!(a.lo < 0)) return a;
    if (!(a.hi > 0)) return {-a.hi, -a.lo, a.nan};
    return {0.0f, fmax_std(-a.lo, a.hi), a.nan};
}

// cos_down / cos_up of rounded_transc_std, substituted as described above.
LF_DEV float cos_dn(float x) { return fmax_std(-1.0f, rnd::dn(fmath::cos_f32_accurate(x))); }
LF_DEV float cos_up(float x) { return fmin_std(1.0f, rnd::up(fmath::cos_f32_accurate(x))); }

// Boost cos(interval) (transc.hpp): reduce by fmod against [2pi], split
// into monotone pieces; "tmp.lower() >= pi_upper" recurses on tmp - pi
// with a sign flip, restated here as a loop.
LF_DEV Iv iv_cos_bounds(float xl, float xu) {
    using namespace rnd;
    bool negate = false;
    float lo = xl, hi = xu;
    for (int iter = 0; iter < 64; ++iter) {
        // tmp = fmod(x, pi2): n = floor(div_dn(x.lo, yb)),
        // tmp = x - n * [pi2lo, pi2hi]
        const float yb = (lo < 0) ? kPi2Lo : kPi2Hi;
        const float n = floorf(div_dn(lo, yb));
        float nl, nu;   // n * pi2 (Boost scalar * interval)
        if (n < 0) { nl = mul_dn(n, kPi2Hi); nu = mul_up(n, kPi2Lo); }
        else if (n == 0) { nl = 0.0f; nu = 0.0f; }
        else { nl = mul_dn(n, kPi2Lo); nu = mul_up(n, kPi2Hi); }
        const float tl = sub_dn(lo, nu), tu = sub_up(hi, nl);
        if (sub_up(tu, tl) >= kPi2Lo) {
            return {-1.0f, 1.0f, false};
        }
        if (tl >= kPiHi) {
            // -cos(tmp - pi)
            negate = !negate;
            lo = sub_dn(tl, kPiHi);
            hi = sub_up(tu, kPiLo);
            continue;
        }
        float rl, ru;
        if (tu <= kPiLo) {
            rl = cos_dn(tu); ru = cos_up(tl);
        } else if (tu <= kPi2Lo) {
            rl = -1.0f; ru = cos_up(fmin_std(sub_dn(kPi2Lo, tu), tl));
        } else {
            rl = -1.0f; ru = 1.0f;
        }
        if (negate) return {-ru, -rl, false};
        return {rl, ru, false};
    }
    return {-1.0f, 1.0f, false};
}

LF_DEV Iv iv_cos(Iv a) {
    if (bad(a.lo, a.hi)) return empty_iv(a.nan);
    Iv r = iv_cos_bounds(a.lo, a.hi);
    r.nan = a.nan;
    return r;
}

// Boost sin(x) = cos(x - [pi/2]).
LF_DEV Iv iv_sin(Iv a) {
    if (bad(a.lo, a.hi)) return empty_iv(a.nan);
    Iv r = iv_cos_bounds(rnd::sub_dn(a.lo, kPiHalfHi),
                         rnd::sub_up(a.hi, kPiHalfLo));
    r.nan = a.nan;
    return r;
}

LF_DEV Iv iv_exp(Iv a) {
    if (bad(a.lo, a.hi)) return empty_iv(a.nan);
    return {fmax_std(0.0f, rnd::dn(fmath::exp_f32(a.lo))),
            rnd::up(fmath::exp_f32(a.hi)), a.nan};
}

LF_DEV Iv iv_log(Iv a) {
    const bool u = a.nan || a.lo < 0.0f;
    if (bad(a.lo, a.hi) || !(a.hi > 0)) return empty_iv(u);
    const float l = !(a.lo > 0) ? -kInf : rnd::dn(fmath::log_f32(a.lo));
    return {l, rnd::up(fmath::log_f32(a.hi)), u};
}

// Boost tan: reduce by fmod against [pi]; whole line across a pole.
LF_DEV Iv iv_tan(Iv a) {
    if (bad(a.lo, a.hi)) return empty_iv(a.nan);
    using namespace rnd;
    const float yb = (a.lo < 0) ? kPiLo : kPiHi;
    const float n = floorf(div_dn(a.lo, yb));
    float nl, nu;
    if (n < 0) { nl = mul_dn(n, kPiHi); nu = mul_up(n, kPiLo); }
    else if (n == 0) { nl = 0.0f; nu = 0.0f; }
    else { nl = mul_dn(n, kPiLo); nu = mul_up(n, kPiHi); }
    float tl = sub_dn(a.lo, nu), tu = sub_up(a.hi, nl);
    if (tl >= kPiHalfLo) {
        const float l2 = sub_dn(tl, kPiHi), u2 = sub_up(tu, kPiLo);
        tl = l2;
        tu = u2;
    }
    if (tl <= -kPiHalfLo || tu >= kPiHalfLo) return {-kInf, kInf, a.nan};
    return {dn(fmath::tan_f32(tl)), up(fmath::tan_f32(tu)), a.nan};
}

LF_DEV Iv iv_asin(Iv a) {
    const bool u = a.nan || a.lo < -1.0f || a.hi > 1.0f;
    if (bad(a.lo, a.hi) || a.hi < -1.0f || a.lo > 1.0f) return empty_iv(u);
    const float l = (a.lo <= -1.0f) ? -kPiHalfHi : rnd::dn(fmath::asin_f32(a.lo));
    const float h = (a.hi >= 1.0f) ? kPiHalfHi : rnd::up(fmath::asin_f32(a.hi));
    return {l, h, u};
}

LF_DEV Iv iv_acos(Iv a) {
    const bool u = a.nan || a.lo < -1.0f || a.hi > 1.0f;
    if (bad(a.lo, a.hi) || a.hi < -1.0f || a.lo > 1.0f) return empty_iv(u);
    const float l = (a.hi >= 1.0f) ? 0.0f : rnd::dn(fmath::acos_f32(a.hi));
    const float h = (a.lo <= -1.0f) ? kPiHi : rnd::up(fmath::acos_f32(a.lo));
    return {l, h, u};
}

LF_DEV Iv iv_atan(Iv a) {
    if (bad(a.lo, a.hi)) return empty_iv(a.nan);
    if (fmath::isinf_f(a.lo) || fmath::isinf_f(a.hi)) {
        // I(-M_PI/2, M_PI/2) built from doubles (interval.hpp:357-360)
        return {float(-1.5707963267948966), float(1.5707963267948966), a.nan};
    }
    return {rnd::dn(fmath::atan_f32(a.lo)), rnd::up(fmath::atan_f32(a.hi)),
            a.nan};
}

// Interval::atan2 (interval.hpp:133-205): quadrant cases on float atan2
// values (no outward rounding in the reference).
LF_DEV Iv iv_atan2(Iv y, Iv x) {
    const bool u = y.nan || x.nan ||
        (x.lo <= 0.0f && x.hi >= 0.0f && y.lo <= 0.0f && y.hi >= 0.0f);
    using fmath::atan2_f32;
    const float pi = float(3.14159265358979323846);
    float l, h;
    if (x.lo > 0) {
        if (y.lo > 0) { l = atan2_f32(y.lo, x.hi); h = atan2_f32(y.hi, x.lo); }
        else if (y.hi < 0) { l = atan2_f32(y.lo, x.lo); h = atan2_f32(y.hi, x.hi); }
        else { l = atan2_f32(y.lo, x.lo); h = atan2_f32(y.hi, x.lo); }
    } else if (x.hi < 0) {
        if (y.lo > 0) { l = atan2_f32(y.hi, x.hi); h = atan2_f32(y.lo, x.lo); }
        else if (y.hi < 0) { l = atan2_f32(y.hi, x.lo); h = atan2_f32(y.lo, x.hi); }
        else { l = -pi; h = pi; }
    } else {
        if (y.lo > 0) { l = atan2_f32(y.lo, x.hi); h = atan2_f32(y.lo, x.lo); }
        else if (y.hi < 0) { l = atan2_f32(y.hi, x.lo); h = atan2_f32(y.hi, x.hi); }
        else { l = -pi; h = pi; }
    }
    Iv r = make_checked(l, h);
    r.nan = u;
    return r;
}

// pow_dn / pow_up of Boost (binary exponentiation with directed products).
LF_DEV float pow_dn(float x_, int p) {
    float x = x_, y = (p & 1) ? x_ : 1.0f;
    p >>= 1;
    while (p > 0) {
        x = rnd::mul_dn(x, x);
        if (p & 1) y = rnd::mul_dn(x, y);
        p >>= 1;
    }
    return y;
}
LF_DEV float pow_up(float x_, int p) {
    float x = x_, y = (p & 1) ? x_ : 1.0f;
    p >>= 1;
    while (p > 0) {
        x = rnd::mul_up(x, x);
        if (p & 1) y = rnd::mul_up(x, y);
        p >>= 1;
    }
    return y;
}

// Boost scalar / interval (1.0f / a in Interval::recip).
LF_DEV Iv scalar_div(float x, Iv y) {
    using namespace rnd;
    if (fmath::isnan_f(x) || bad(y.lo, y.hi)) return empty_iv(false);
    if (y.lo <= 0 && 0 <= y.hi) {
        if (y.lo != 0) {
            if (y.hi != 0) {               // div_zero
                if (x == 0) return {0.0f, 0.0f, false};
                return {-kInf, kInf, false};
            }
            // div_negative: y = [yl, 0]
            if (x == 0) return {0.0f, 0.0f, false};
            if (x < 0) return {div_dn(x, y.lo), kInf, false};
            return {-kInf, div_up(x, y.lo), false};
        }
        if (y.hi != 0) {                   // div_positive: y = [0, yu]
            if (x == 0) return {0.0f, 0.0f, false};
            if (x < 0) return {-kInf, div_up(x, y.hi), false};
            return {div_dn(x, y.hi), kInf, false};
        }
        return empty_iv(false);
    }
    if (x < 0) return {div_dn(x, y.lo), div_up(x, y.hi), false};
    return {div_dn(x, y.hi), div_up(x, y.lo), false};
}

LF_DEV Iv iv_recip(Iv a) {
    Iv r = scalar_div(1.0f, a);
    r.nan = a.nan;
    return r;
}

// Boost pow(interval, int) + Interval::pow flags (interval.hpp:207-231).
LF_DEV Iv iv_pow(Iv a, Iv b) {
    const int p = int(b.lo);
    const bool a_zero = a.lo <= 0.0f && a.hi >= 0.0f;
    // pow(0, negative) is a pole (inf), pow(-1, p) is never NaN for int p
    const bool u = a.nan || b.nan || (a_zero && p == 0);
    Iv r;
    if (bad(a.lo, a.hi)) {
        r = empty_iv(false);
    } else if (p == 0) {
        if (a.lo == 0 && a.hi == 0) r = empty_iv(false);
        else r = {1.0f, 1.0f, false};
    } else {
        const int q = (p < 0) ? -p : p;
        Iv t;
        if (a.hi < 0) {
            const float yl = pow_dn(-a.hi, q), yu = pow_up(-a.lo, q);
            t = (q & 1) ? Iv{-yu, -yl, false} : Iv{yl, yu, false};
        } else if (a.lo < 0) {
            if (q & 1) t = {-pow_up(-a.lo, q), pow_up(a.hi, q), false};
            else t = {0.0f, pow_up(fmax_std(-a.lo, a.hi), q), false};
        } else {
            t = {pow_dn(a.lo, q), pow_up(a.hi, q), false};
        }
        r = (p < 0) ? scalar_div(1.0f, t) : t;
    }
    r.nan = u;
    return r;
}

// nth_root: Boost's interval Newton root is replaced by outward-moved
// double-precision roots (monotone in a); flags as interval.hpp:233-242.
LF_DEV Iv iv_nth_root(Iv a, Iv b) {
    const int n = int(b.lo);
    const bool u = a.nan || b.nan || (a.lo <= 0.0f && !(n & 2));
    if (bad(a.lo, a.hi)) return empty_iv(u);
    if (n == 1) return {a.lo, a.hi, u};
    auto root = [&](float v) { return fmath::nth_root_f32(v, float(n)); };
    if (!(a.hi > 0)) {
        if (a.hi == 0) {
            if (!(n & 1) || a.lo == 0) return {0.0f, 0.0f, u};
            return {rnd::dn(root(a.lo)), 0.0f, u};
        }
        if (!(n & 1)) return empty_iv(u);
        return {rnd::dn(root(a.lo)), rnd::up(root(a.hi)), u};
    }
    float l;
    if (a.lo > 0) l = rnd::dn(root(a.lo));
    else if (!(n & 1)) l = 0.0f;
    else l = rnd::dn(root(a.lo));
    return {l, rnd::up(root(a.hi)), u};
}

LF_DEV Iv iv_nanfill(Iv a, Iv b) {
    if (a.nan) {
        float lo = a.lo, hi = a.hi;
        hull(lo, hi, b.lo, b.hi);
        return {lo, hi, b.nan};
    }
    return {a.lo, a.hi, false};
}

LF_DEV Iv iv_compare(Iv a, Iv b) {
    if (a.hi < b.lo) return {-1.0f, -1.0f, false};
    if (a.lo > b.hi) return {1.0f, 1.0f, false};
    return {-1.0f, 1.0f, false};
}

// Interval::mod (interval.hpp:244-289).
LF_DEV Iv iv_mod(Iv a, Iv b) {
    using namespace rnd;
    float ol = fminf(b.lo, 0.0f), ou = fmaxf(0.0f, b.hi);
    if (!fmath::isinf_f(a.hi) && !fmath::isnan_f(a.hi) &&
        !fmath::isinf_f(a.lo) && !fmath::isnan_f(a.lo)) {
        const int position = (b.hi >= 0.0f) + 2 * (b.lo <= 0.0f);
        if (position == 1 || position == 2) {
            float ul = a.lo, uh = a.hi;
            if (position == 2) {            // usedA *= -1
                const float t = -uh;
                uh = -ul;
                ul = t;
            }
            // absB = abs(b)
            Iv ab = iv_abs(Iv{b.lo, b.hi, false});
            // quotients = usedA / absB
            float ql, qh;
            if (bad(ab.lo, ab.hi) || bad(ul, uh)) {
                ql = nanv(); qh = nanv();
            } else if (ab.lo <= 0 && 0 <= ab.hi) {
                ql = -kInf; qh = kInf;     // not reached: b excludes 0 here
            } else {
                div_nz(ul, uh, ab.lo, ab.hi, ql, qh);
            }
            const int q0 = int(floorf(ql));
            if (q0 == int(floorf(qh))) {
                // out = a - b * float(q)
                const float qf = float(q0);
                float ml, mu;
                if (qf < 0) { ml = mul_dn(qf, b.hi); mu = mul_up(qf, b.lo); }
                else if (qf == 0) { ml = 0.0f; mu = 0.0f; }
                else { ml = mul_dn(qf, b.lo); mu = mul_up(qf, b.hi); }
                ol = sub_dn(a.lo, mu);
                ou = sub_up(a.hi, ml);
            }
        } else if (position == 0) {
            ol = nanv();
            ou = nanv();
        }
    }
    const bool u = b.hi >= 0.0f && b.lo <= 0.0f;
    return {ol, ou, u};
}

// One interval clause.
// Opcodes outside the common set, out of line: the interval kernels are
// latency-bound walks over one tape, and keeping every rule inline makes
// the loop body large enough to miss in the instruction cache.
static __device__ __noinline__ Iv iv_apply_cold(uint32_t o, Iv a, Iv b) {
    switch (o) {
        case 22: return iv_div(a, b);
        case 23: return iv_atan2(a, b);
        case 24: return iv_pow(a, b);
        case 25: return iv_nth_root(a, b);
        case 26: return iv_mod(a, b);
        case 27: return iv_nanfill(a, b);
        case 31: return iv_compare(a, b);
        case 10: return iv_sin(a);
        case 11: return iv_cos(a);
        case 12: return iv_tan(a);
        case 13: return iv_asin(a);
        case 14: return iv_acos(a);
        case 15: return iv_atan(a);
        case 16: return iv_exp(a);
        case 29: return iv_recip(a);
        case 30: return iv_log(a);
        default: return a;
    }
}

LF_DEV Iv iv_apply(uint32_t o, Iv a, Iv b, float im) {
    switch (o) {
        case 17: return iv_add(a, b);
        case 18: return iv_mul(a, b);
        case 19: return iv_min(a, b);
        case 20: return iv_max(a, b);
        case 21: return iv_sub(a, b);
        case 7: return iv_square(a);
        case 8: return iv_sqrt(a);
        case 9: return iv_neg(a);
        case 28: return iv_abs(a);
        case 1: return {im, im, fmath::isnan_f(im)};     // LOADI
        case 6:
        case 40: return a;                                // CONST_VAR, COPY
        default: return iv_apply_cold(o, a, b);
    }
}

}   // namespace amd
}   // namespace libfive
