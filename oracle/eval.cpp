// Oracle (test infrastructure only): CPU restatement of the reference's
// tape evaluators and tape pruning.  Built with -ffp-contract=off
// -frounding-math; never linked into the product library.
//
// Follows:
//   libfive/src/eval/deck.cpp:19-94         slot layout (input format)
//   libfive/src/eval/eval_array.cpp:36-118, 198-369   ArrayEvaluator
//   libfive/src/eval/eval_deriv_array.cpp:28-256       DerivArrayEvaluator
//   libfive/src/eval/eval_interval.cpp:40-294          IntervalEvaluator
//   libfive/include/libfive/eval/interval.hpp:31-444   Interval + Boost ops
//   libfive/src/eval/tape.cpp:23-205                   Tape::push, getBase
#include "oracle_impl.hpp"

#include <cfenv>
#include <cmath>
#include <cstring>
#include <thread>

#include "../libfive_amd/csrc/fmath.hpp"

namespace oracle {

using libfive::fmath::isnan_f;

////////////////////////////////////////////////////////////////////////////////
// Deck

Deck::Deck(const uint32_t* clauses, size_t n, const uint32_t* cids,
           const float* cvals, size_t nc, uint32_t X_, uint32_t Y_,
           uint32_t Z_, uint32_t num_slots_)
    : X(X_), Y(Y_), Z(Z_), num_slots(num_slots_),
      disabled(num_slots_ + 1), remap(num_slots_ + 1)
{
    auto base = std::make_shared<Tape>();
    for (size_t i = 0; i < n; ++i) {
        base->t.push_back({clauses[4 * i], clauses[4 * i + 1],
                           clauses[4 * i + 2], clauses[4 * i + 3]});
    }
    base->i = 1;
    base->type = Tape::BASE;
    tape = base;
    for (size_t i = 0; i < nc; ++i) constants.push_back({cids[i], cvals[i]});
}

////////////////////////////////////////////////////////////////////////////////
// Tape::push (tape.cpp:23-158)

std::shared_ptr<Tape> push(Deck& deck, const std::shared_ptr<Tape>& tape,
                           const KeepFn& fn, Tape::Type type,
                           const float box[6]) {
    if (tape->terminal) return tape;
    std::fill(deck.disabled.begin(), deck.disabled.end(), 1);
    std::fill(deck.remap.begin(), deck.remap.end(), 0);
    deck.disabled[tape->i] = 0;
    bool terminal = true, changed = false;
    for (const auto& c : tape->t) {
        if (deck.disabled[c.id]) continue;
        switch (fn(c.op, c.id, c.a, c.b)) {
            case KEEP_A:
                deck.disabled[c.a] = 0;
                deck.remap[c.id] = c.a;
                changed = true;
                break;
            case KEEP_B:
                deck.disabled[c.b] = 0;
                deck.remap[c.id] = c.b;
                changed = true;
                break;
            case KEEP_BOTH:
                terminal = false;
                break;
            case KEEP_ALWAYS:
                break;
        }
        if (deck.remap[c.id]) {
            deck.disabled[c.id] = 1;
        } else {
            deck.disabled[c.a] = 0;
            deck.disabled[c.b] = 0;
        }
    }
    if (!changed) return tape;
    auto out = std::make_shared<Tape>();
    out->type = type;
    out->parent = tape;
    out->terminal = terminal;
    for (const auto& c : tape->t) {
        if (deck.disabled[c.id]) continue;
        uint32_t ra = c.a, rb = c.b;
        while (deck.remap[ra]) ra = deck.remap[ra];
        while (deck.remap[rb]) rb = deck.remap[rb];
        out->t.push_back({c.op, c.id, ra, rb});
    }
    out->i = tape->i;
    while (deck.remap[out->i]) out->i = deck.remap[out->i];
    if (box) std::memcpy(out->box, box, sizeof(out->box));
    return out;
}

// Tape::getBase (tape.cpp:168-205): walk up to the first INTERVAL tape whose
// box contains the point / region (inclusive), or the root tape.
std::shared_ptr<Tape> get_base(const std::shared_ptr<Tape>& t,
                               const double lo[3], const double hi[3]) {
    auto tape = t;
    while (tape->parent) {
        if (tape->type == Tape::INTERVAL &&
            lo[0] >= tape->box[0] && hi[0] <= tape->box[1] &&
            lo[1] >= tape->box[2] && hi[1] <= tape->box[3] &&
            lo[2] >= tape->box[4] && hi[2] <= tape->box[5]) {
            break;
        }
        tape = tape->parent;
    }
    return tape;
}

////////////////////////////////////////////////////////////////////////////////
// Scalar semantics (eval_array.cpp:221-369)

static float value_op(uint32_t op, float a, float b) {
    // The product's scalar table states the same rules; here they are
    // written out independently except for the transcendental routines.
    switch (op) {
        case 17: return a + b;
        case 18: return a * b;
        case 19: return (b < a) ? b : a;
        case 20: return (a < b) ? b : a;
        case 21: return a - b;
        case 22: return a / b;
        case 23: return float(std::atan2(double(a), double(b)));
        case 24: return float(std::pow(double(a), double(b)));
        case 25:
            if (a < 0) {
                const int n = int(b);
                if (!(n & 1)) return NAN;
                return -float(std::pow(-double(a), 1.0 / double(b)));
            }
            return float(std::pow(double(a), 1.0 / double(b)));
        case 26: {
            float d = std::fabs(a / b);
            if ((a < 0) ^ (b < 0)) d = -std::ceil(d);
            else d = std::floor(d);
            float out = a - b * d;
            if ((b > 0 && out > b) || (b < 0 && out < b)) out = b;
            if ((b > 0.0f && out < 0.0f) || (b < 0.0f && out > 0.0f)) out = 0.0f;
            return out;
        }
        case 27: return std::isnan(a) ? b : a;
        case 31: return (a < b) ? -1.0f : (a > b) ? 1.0f : 0.0f;
        case 7: return a * a;
        case 8: return std::sqrt(a);
        case 9: return -a;
        case 10: return libfive::fmath::sin_f32(a);
        case 11: return libfive::fmath::cos_f32(a);
        case 12: return float(std::tan(double(a)));
        case 13: return float(std::asin(double(a)));
        case 14: return float(std::acos(double(a)));
        case 15: return float(std::atan(double(a)));
        case 16: return float(std::exp(double(a)));
        case 28: return std::fabs(a);
        case 29: return 1 / a;
        case 30: return float(std::log(double(a)));
        case 6: return a;
        default: return NAN;
    }
}

////////////////////////////////////////////////////////////////////////////////
// ArrayEvaluator (eval_array.cpp:36-56: slots initialised to zero, then
// constants written into their rows; values() walks the tape leaves-first)

ArrayEval::ArrayEval(Deck& d) : deck(d), v((d.num_slots + 1) * N, 0.0f) {
    for (auto& c : d.constants) {
        float* row = &v[size_t(c.first) * N];
        for (size_t j = 0; j < N; ++j) row[j] = c.second;
    }
}

void ArrayEval::set(float x, float y, float z, size_t i) {
    v[size_t(deck.X) * N + i] = x;
    v[size_t(deck.Y) * N + i] = y;
    v[size_t(deck.Z) * N + i] = z;
}

const float* ArrayEval::values(size_t count, const Tape& tape) {
    for (auto it = tape.t.rbegin(); it != tape.t.rend(); ++it) {
        float* out = &v[size_t(it->id) * N];
        const float* a = &v[size_t(it->a) * N];
        const float* b = &v[size_t(it->b) * N];
        const uint32_t op = it->op;
        switch (op) {
            // The hot opcodes as separate loops so that they vectorise.
            case 17: for (size_t j = 0; j < count; ++j) out[j] = a[j] + b[j]; break;
            case 18: for (size_t j = 0; j < count; ++j) out[j] = a[j] * b[j]; break;
            case 19: for (size_t j = 0; j < count; ++j) out[j] = (b[j] < a[j]) ? b[j] : a[j]; break;
            case 20: for (size_t j = 0; j < count; ++j) out[j] = (a[j] < b[j]) ? b[j] : a[j]; break;
            case 21: for (size_t j = 0; j < count; ++j) out[j] = a[j] - b[j]; break;
            case 7:  for (size_t j = 0; j < count; ++j) out[j] = a[j] * a[j]; break;
            case 8:  for (size_t j = 0; j < count; ++j) out[j] = std::sqrt(a[j]); break;
            case 9:  for (size_t j = 0; j < count; ++j) out[j] = -a[j]; break;
            case 28: for (size_t j = 0; j < count; ++j) out[j] = std::fabs(a[j]); break;
            default:
                for (size_t j = 0; j < count; ++j) out[j] = value_op(op, a[j], b[j]);
                break;
        }
    }
    return &v[size_t(tape.i) * N];
}

// getAmbiguous (eval_array.cpp:198-217): lanes where a MIN/MAX clause sees
// equal operands.
void ArrayEval::ambiguous(size_t count, const Tape& tape, uint8_t* out) {
    std::fill(out, out + count, 0);
    for (auto it = tape.t.rbegin(); it != tape.t.rend(); ++it) {
        if (it->op == 19 || it->op == 20) {
            const float* a = &v[size_t(it->a) * N];
            const float* b = &v[size_t(it->b) * N];
            for (size_t j = 0; j < count; ++j) out[j] |= (a[j] == b[j]);
        }
    }
}

////////////////////////////////////////////////////////////////////////////////
// DerivArrayEvaluator (eval_deriv_array.cpp:28-256)

DerivEval::DerivEval(Deck& d)
    : ArrayEval(d), g(size_t(d.num_slots + 1) * 3 * N, 0.0f) {
    for (size_t j = 0; j < N; ++j) {
        g[(size_t(d.X) * 3 + 0) * N + j] = 1;
        g[(size_t(d.Y) * 3 + 1) * N + j] = 1;
        g[(size_t(d.Z) * 3 + 2) * N + j] = 1;
    }
}

void DerivEval::derivs(size_t count, const Tape& tape) {
    values(count, tape);
    for (auto it = tape.t.rbegin(); it != tape.t.rend(); ++it) {
        const uint32_t op = it->op;
        const float* ov = &v[size_t(it->id) * N];
        const float* av = &v[size_t(it->a) * N];
        const float* bv = &v[size_t(it->b) * N];
        for (int k = 0; k < 3; ++k) {
            float* od = &g[(size_t(it->id) * 3 + k) * N];
            const float* ad = &g[(size_t(it->a) * 3 + k) * N];
            const float* bd = &g[(size_t(it->b) * 3 + k) * N];
            for (size_t j = 0; j < count; ++j) {
                const float a = av[j], b = bv[j], da = ad[j], db = bd[j];
                float r;
                switch (op) {
                    case 17: r = da + db; break;
                    case 21: r = da - db; break;
                    case 18: { const float p = db * a; const float q = da * b; r = p + q; break; }
                    case 19: r = (a < b) ? da : db; break;
                    case 20: r = (a < b) ? db : da; break;
                    case 22: { const float p = da * b; const float q = db * a; r = (p - q) / (b * b); break; }
                    case 23: { const float p = da * b; const float q = db * a;
                               const float a2 = a * a; const float b2 = b * b;
                               r = (p - q) / (a2 + b2); break; }
                    case 24: { const float f = b * float(std::pow(double(a), double(b - 1.0f)));
                               r = da * f; break; }
                    case 25: { const float f = float(std::pow(double(a), double(1.0f / b - 1.0f))) / b;
                               r = (da == 0) ? 0.0f : da * f; break; }
                    case 26: r = da; break;
                    case 27: r = std::isnan(a) ? db : da; break;
                    case 31: r = 0.0f; break;
                    case 7: r = (da * a) * 2.0f; break;
                    case 8: r = (a < 0 || da == 0) ? 0.0f : da / (2.0f * ov[j]); break;
                    case 9: r = -da; break;
                    case 10: r = da * libfive::fmath::cos_f32(a); break;
                    case 11: r = da * -libfive::fmath::sin_f32(a); break;
                    case 12: { const float c = 1.0f / libfive::fmath::cos_f32(a); r = da * (c * c); break; }
                    case 13: r = da / std::sqrt(1.0f - a * a); break;
                    case 14: r = da / -std::sqrt(1.0f - a * a); break;
                    case 15: r = da / (a * a + 1.0f); break;
                    case 30: r = da / a; break;
                    case 16: r = da * float(std::exp(double(a))); break;
                    case 28: r = (a > 0) ? da : -da; break;
                    case 29: r = da / -(a * a); break;
                    case 6: r = da; break;
                    default: r = NAN; break;
                }
                od[j] = r;
            }
        }
    }
}

const float* DerivEval::grad(size_t k, const Tape& tape) const {
    return &g[(size_t(tape.i) * 3 + k) * N];
}

////////////////////////////////////////////////////////////////////////////////
// Interval arithmetic: Boost.Interval<float> with fesetround directed
// rounding (rounded_arith_std), checking_base (empty = [NaN, NaN]), plus
// the reference's maybe_nan flags (interval.hpp:31-444).  Transcendental
// bounds are the round-to-nearest routine moved one ulp outward (the
// reference's come from the C library under a directed mode).

namespace rnd {
static float add_dn(float a, float b) { std::fesetround(FE_DOWNWARD); volatile float r = a + b; std::fesetround(FE_TONEAREST); return r; }
static float add_up(float a, float b) { std::fesetround(FE_UPWARD); volatile float r = a + b; std::fesetround(FE_TONEAREST); return r; }
static float sub_dn(float a, float b) { std::fesetround(FE_DOWNWARD); volatile float r = a - b; std::fesetround(FE_TONEAREST); return r; }
static float sub_up(float a, float b) { std::fesetround(FE_UPWARD); volatile float r = a - b; std::fesetround(FE_TONEAREST); return r; }
static float mul_dn(float a, float b) { std::fesetround(FE_DOWNWARD); volatile float r = a * b; std::fesetround(FE_TONEAREST); return r; }
static float mul_up(float a, float b) { std::fesetround(FE_UPWARD); volatile float r = a * b; std::fesetround(FE_TONEAREST); return r; }
static float div_dn(float a, float b) { std::fesetround(FE_DOWNWARD); volatile float r = a / b; std::fesetround(FE_TONEAREST); return r; }
static float div_up(float a, float b) { std::fesetround(FE_UPWARD); volatile float r = a / b; std::fesetround(FE_TONEAREST); return r; }
static float sqrt_dn(float a) { std::fesetround(FE_DOWNWARD); volatile float r = std::sqrt(a); std::fesetround(FE_TONEAREST); return r; }
static float sqrt_up(float a) { std::fesetround(FE_UPWARD); volatile float r = std::sqrt(a); std::fesetround(FE_TONEAREST); return r; }
static float dn(float v) { return libfive::fmath::pred_f(v); }
static float up(float v) { return libfive::fmath::succ_f(v); }
static float int_down(float v) { return std::floor(v); }
}   // namespace rnd

static const float kInf = INFINITY;
static float smin(float a, float b) { return (b < a) ? b : a; }   // std::min
static float smax(float a, float b) { return (a < b) ? b : a; }   // std::max
static bool bad(const Iv& a) { return !(a.lo <= a.hi); }
static Iv empty(bool nan) { return {NAN, NAN, nan}; }

// Interval(float lower, float upper) (interval.hpp:33-35)
Iv iv_checked(float lo, float hi) {
    const bool n = std::isnan(lo) || std::isnan(hi);
    if (n || !(lo <= hi)) return {NAN, NAN, n};
    return {lo, hi, n};
}

// boost pi constants: 13176794/2^22, 13176795/2^22
static const float kPiLo = 13176794.0f / 4194304.0f;
static const float kPiHi = 13176795.0f / 4194304.0f;

static void scalar_mul(float n, float yl, float yu, float& lo, float& hi) {
    // Boost operator*(T, interval)
    if (n < 0) { lo = rnd::mul_dn(n, yu); hi = rnd::mul_up(n, yl); }
    else if (n == 0) { lo = 0; hi = 0; }
    else { lo = rnd::mul_dn(n, yl); hi = rnd::mul_up(n, yu); }
}

// boost fmod(interval x, interval y) followed by the caller's use.
static void fmod_iv(float xl, float xu, float yl, float yu, float& tl, float& tu) {
    const float yb = (xl < 0) ? yl : yu;
    const float n = rnd::int_down(rnd::div_dn(xl, yb));
    float nl, nu;
    scalar_mul(n, yl, yu, nl, nu);
    tl = rnd::sub_dn(xl, nu);
    tu = rnd::sub_up(xu, nl);
}

static float cos_dn(float x) { return smax(-1.0f, rnd::dn(libfive::fmath::cos_f32_accurate(x))); }
static float cos_up(float x) { return smin(1.0f, rnd::up(libfive::fmath::cos_f32_accurate(x))); }

// boost cos(interval) (transc.hpp), with its recursion on tmp - pi.
static void boost_cos(float xl, float xu, float& rl, float& ru, int depth) {
    float tl, tu;
    fmod_iv(xl, xu, 2 * kPiLo, 2 * kPiHi, tl, tu);
    if (rnd::sub_up(tu, tl) >= 2 * kPiLo) { rl = -1; ru = 1; return; }
    if (tl >= kPiHi) {
        if (depth > 64) { rl = -1; ru = 1; return; }
        float nl, nu;
        boost_cos(rnd::sub_dn(tl, kPiHi), rnd::sub_up(tu, kPiLo), nl, nu, depth + 1);
        rl = -nu; ru = -nl;
        return;
    }
    if (tu <= kPiLo) { rl = cos_dn(tu); ru = cos_up(tl); }
    else if (tu <= 2 * kPiLo) { rl = -1; ru = cos_up(smin(rnd::sub_dn(2 * kPiLo, tu), tl)); }
    else { rl = -1; ru = 1; }
}

static void hull(float& lo, float& hi, float blo, float bhi) {
    const bool bx = !(lo <= hi), by = !(blo <= bhi);
    if (bx) { if (by) { lo = NAN; hi = NAN; } else { lo = blo; hi = bhi; } return; }
    if (by) return;
    lo = smin(lo, blo);
    hi = smax(hi, bhi);
}

static void mul_bounds(float xl, float xu, float yl, float yu, float& lo, float& hi) {
    using namespace rnd;
    if (xl < 0) {
        if (xu > 0) {
            if (yl < 0) {
                if (yu > 0) { lo = smin(mul_dn(xl, yu), mul_dn(xu, yl)); hi = smax(mul_up(xl, yl), mul_up(xu, yu)); }
                else { lo = mul_dn(xu, yl); hi = mul_up(xl, yl); }
            } else {
                if (yu > 0) { lo = mul_dn(xl, yu); hi = mul_up(xu, yu); }
                else { lo = 0; hi = 0; }
            }
        } else {
            if (yl < 0) {
                if (yu > 0) { lo = mul_dn(xl, yu); hi = mul_up(xl, yl); }
                else { lo = mul_dn(xu, yu); hi = mul_up(xl, yl); }
            } else {
                if (yu > 0) { lo = mul_dn(xl, yu); hi = mul_up(xu, yl); }
                else { lo = 0; hi = 0; }
            }
        }
    } else {
        if (xu > 0) {
            if (yl < 0) {
                if (yu > 0) { lo = mul_dn(xu, yl); hi = mul_up(xu, yu); }
                else { lo = mul_dn(xu, yl); hi = mul_up(xl, yu); }
            } else {
                if (yu > 0) { lo = mul_dn(xl, yl); hi = mul_up(xu, yu); }
                else { lo = 0; hi = 0; }
            }
        } else { lo = 0; hi = 0; }
    }
}

static void div_nz(float xl, float xu, float yl, float yu, float& lo, float& hi) {
    using namespace rnd;
    if (xu < 0) {
        if (yu < 0) { lo = div_dn(xu, yl); hi = div_up(xl, yu); }
        else { lo = div_dn(xl, yl); hi = div_up(xu, yu); }
    } else if (xl < 0) {
        if (yu < 0) { lo = div_dn(xu, yu); hi = div_up(xl, yu); }
        else { lo = div_dn(xl, yl); hi = div_up(xu, yl); }
    } else {
        if (yu < 0) { lo = div_dn(xu, yu); hi = div_up(xl, yl); }
        else { lo = div_dn(xl, yu); hi = div_up(xu, yl); }
    }
}

static float pow_dir(float x_, int p, bool up) {
    float x = x_, y = (p & 1) ? x_ : 1.0f;
    p >>= 1;
    while (p > 0) {
        x = up ? rnd::mul_up(x, x) : rnd::mul_dn(x, x);
        if (p & 1) y = up ? rnd::mul_up(x, y) : rnd::mul_dn(x, y);
        p >>= 1;
    }
    return y;
}

static Iv scalar_div(float x, const Iv& y) {
    using namespace rnd;
    if (std::isnan(x) || bad(y)) return empty(false);
    if (y.lo <= 0 && 0 <= y.hi) {
        if (y.lo != 0) {
            if (y.hi != 0) { if (x == 0) return {0, 0, false}; return {-kInf, kInf, false}; }
            if (x == 0) return {0, 0, false};
            if (x < 0) return {div_dn(x, y.lo), kInf, false};
            return {-kInf, div_up(x, y.lo), false};
        }
        if (y.hi != 0) {
            if (x == 0) return {0, 0, false};
            if (x < 0) return {-kInf, div_up(x, y.hi), false};
            return {div_dn(x, y.hi), kInf, false};
        }
        return empty(false);
    }
    if (x < 0) return {div_dn(x, y.lo), div_up(x, y.hi), false};
    return {div_dn(x, y.hi), div_up(x, y.lo), false};
}

static float nth(float v, int n) {
    return (v < 0) ? ((n & 1) ? -float(std::pow(-double(v), 1.0 / n)) : NAN)
                   : float(std::pow(double(v), 1.0 / n));
}

Iv interval_op(uint32_t op, const Iv& a, const Iv& b) {
    using namespace rnd;
    switch (op) {
        case 17: {   // operator+ (interval.hpp:399-405)
            const bool u = a.nan || b.nan || (a.lo == -kInf && b.hi == kInf) || (b.lo == -kInf && a.hi == kInf);
            if (bad(a) || bad(b)) return empty(u);
            return {add_dn(a.lo, b.lo), add_up(a.hi, b.hi), u};
        }
        case 21: {   // operator- (interval.hpp:418-424)
            const bool u = a.nan || b.nan || (a.lo == -kInf && b.lo == -kInf) || (a.hi == -kInf && b.hi == -kInf);
            if (bad(a) || bad(b)) return empty(u);
            return {sub_dn(a.lo, b.hi), sub_up(a.hi, b.lo), u};
        }
        case 18: {   // operator* (interval.hpp:407-416)
            const bool u = a.nan || b.nan ||
                ((a.lo == -kInf || a.hi == kInf) && b.lo <= 0 && b.hi >= 0) ||
                ((b.lo == -kInf || b.hi == kInf) && a.lo <= 0 && a.hi >= 0);
            if (bad(a) || bad(b)) return empty(u);
            Iv r; mul_bounds(a.lo, a.hi, b.lo, b.hi, r.lo, r.hi); r.nan = u;
            return r;
        }
        case 22: {   // operator/ (interval.hpp:426-442)
            const bool u = a.nan || b.nan ||
                ((a.lo == -kInf || a.hi == kInf) && (b.lo == -kInf || b.hi == kInf)) ||
                (a.lo <= 0 && a.hi >= 0 && b.lo <= 0 && b.hi >= 0);
            if (b.lo <= 0 && b.hi >= 0) return {-kInf, kInf, u};
            if (bad(a) || bad(b)) return empty(u);
            Iv r; div_nz(a.lo, a.hi, b.lo, b.hi, r.lo, r.hi); r.nan = u;
            return r;
        }
        case 19:
        case 20: {   // Interval::min / max (interval.hpp:82-131)
            float lo, hi;
            if (bad(a) || bad(b)) { lo = NAN; hi = NAN; }
            else if (op == 19) { lo = smin(a.lo, b.lo); hi = smin(a.hi, b.hi); }
            else { lo = smax(a.lo, b.lo); hi = smax(a.hi, b.hi); }
            if (b.nan) hull(lo, hi, a.lo, a.hi);
            return {lo, hi, a.nan};
        }
        case 7: {    // square (interval.hpp:303-306)
            if (bad(a)) return empty(a.nan);
            if (a.hi < 0) return {mul_dn(a.hi, a.hi), mul_up(a.lo, a.lo), a.nan};
            if (a.lo > 0) return {mul_dn(a.lo, a.lo), mul_up(a.hi, a.hi), a.nan};
            return {0, (-a.lo > a.hi) ? mul_up(a.lo, a.lo) : mul_up(a.hi, a.hi), a.nan};
        }
        case 8: {    // sqrt (interval.hpp:308-312)
            const bool u = a.nan || a.lo < 0;
            if (bad(a) || a.hi < 0) return empty(u);
            return {(a.lo > 0) ? sqrt_dn(a.lo) : 0.0f, sqrt_up(a.hi), u};
        }
        case 9:      // unary minus (interval.hpp:314-317)
            if (bad(a)) return empty(a.nan);
            return {-a.hi, -a.lo, a.nan};
        case 28:     // abs (interval.hpp:371-374)
            if (bad(a)) return empty(a.nan);
            if (!(a.lo < 0)) return a;
            if (!(a.hi > 0)) return {-a.hi, -a.lo, a.nan};
            return {0, smax(-a.lo, a.hi), a.nan};
        case 11: {   // cos (interval.hpp:324-327)
            if (bad(a)) return empty(a.nan);
            Iv r; boost_cos(a.lo, a.hi, r.lo, r.hi, 0); r.nan = a.nan;
            return r;
        }
        case 10: {   // sin = cos(x - pi/2) (interval.hpp:319-322)
            if (bad(a)) return empty(a.nan);
            Iv r; boost_cos(sub_dn(a.lo, kPiHi / 2), sub_up(a.hi, kPiLo / 2), r.lo, r.hi, 0);
            r.nan = a.nan;
            return r;
        }
        case 16:     // exp
            if (bad(a)) return empty(a.nan);
            return {smax(0.0f, dn(float(std::exp(double(a.lo))))), up(float(std::exp(double(a.hi)))), a.nan};
        case 30: {   // log
            const bool u = a.nan || a.lo < 0;
            if (bad(a) || !(a.hi > 0)) return empty(u);
            return {!(a.lo > 0) ? -kInf : dn(float(std::log(double(a.lo)))), up(float(std::log(double(a.hi)))), u};
        }
        case 12: {   // tan
            if (bad(a)) return empty(a.nan);
            float tl, tu;
            fmod_iv(a.lo, a.hi, kPiLo, kPiHi, tl, tu);
            if (tl >= kPiLo / 2) { const float l2 = sub_dn(tl, kPiHi), u2 = sub_up(tu, kPiLo); tl = l2; tu = u2; }
            if (tl <= -(kPiLo / 2) || tu >= kPiLo / 2) return {-kInf, kInf, a.nan};
            return {dn(float(std::tan(double(tl)))), up(float(std::tan(double(tu)))), a.nan};
        }
        case 13: case 14: {   // asin, acos
            const bool u = a.nan || a.lo < -1 || a.hi > 1;
            if (bad(a) || a.hi < -1 || a.lo > 1) return empty(u);
            if (op == 13) {
                return {(a.lo <= -1) ? -(kPiHi / 2) : dn(float(std::asin(double(a.lo)))),
                        (a.hi >= 1) ? kPiHi / 2 : up(float(std::asin(double(a.hi)))), u};
            }
            return {(a.hi >= 1) ? 0.0f : dn(float(std::acos(double(a.hi)))),
                    (a.lo <= -1) ? kPiHi : up(float(std::acos(double(a.lo)))), u};
        }
        case 15:     // atan (interval.hpp:347-359)
            if (bad(a)) return empty(a.nan);
            if (std::isinf(a.lo) || std::isinf(a.hi)) return {float(-M_PI / 2), float(M_PI / 2), a.nan};
            return {dn(float(std::atan(double(a.lo)))), up(float(std::atan(double(a.hi)))), a.nan};
        case 29: {   // recip (interval.hpp:376-380)
            Iv r = scalar_div(1.0f, a); r.nan = a.nan; return r;
        }
        case 23: {   // atan2 (interval.hpp:133-205)
            const Iv& y = a; const Iv& x = b;
            const bool u = y.nan || x.nan || (x.lo <= 0 && x.hi >= 0 && y.lo <= 0 && y.hi >= 0);
            auto at = [](float p, float q) { return float(std::atan2(double(p), double(q))); };
            const float pi = float(M_PI);
            float l, h;
            if (x.lo > 0) {
                if (y.lo > 0) { l = at(y.lo, x.hi); h = at(y.hi, x.lo); }
                else if (y.hi < 0) { l = at(y.lo, x.lo); h = at(y.hi, x.hi); }
                else { l = at(y.lo, x.lo); h = at(y.hi, x.lo); }
            } else if (x.hi < 0) {
                if (y.lo > 0) { l = at(y.hi, x.hi); h = at(y.lo, x.lo); }
                else if (y.hi < 0) { l = at(y.hi, x.lo); h = at(y.lo, x.hi); }
                else { l = -pi; h = pi; }
            } else {
                if (y.lo > 0) { l = at(y.lo, x.hi); h = at(y.lo, x.lo); }
                else if (y.hi < 0) { l = at(y.hi, x.lo); h = at(y.hi, x.hi); }
                else { l = -pi; h = pi; }
            }
            Iv r = iv_checked(l, h); r.nan = u;
            return r;
        }
        case 24: {   // pow (interval.hpp:207-231)
            const int p = int(b.lo);
            const bool a_zero = a.lo <= 0 && a.hi >= 0;
            const bool u = a.nan || b.nan || (a_zero && p == 0);
            Iv r;
            if (bad(a)) r = empty(false);
            else if (p == 0) r = (a.lo == 0 && a.hi == 0) ? empty(false) : Iv{1, 1, false};
            else {
                const int q = (p < 0) ? -p : p;
                Iv t;
                if (a.hi < 0) {
                    const float yl = pow_dir(-a.hi, q, false), yu = pow_dir(-a.lo, q, true);
                    t = (q & 1) ? Iv{-yu, -yl, false} : Iv{yl, yu, false};
                } else if (a.lo < 0) {
                    if (q & 1) t = {-pow_dir(-a.lo, q, true), pow_dir(a.hi, q, true), false};
                    else t = {0, pow_dir(smax(-a.lo, a.hi), q, true), false};
                } else {
                    t = {pow_dir(a.lo, q, false), pow_dir(a.hi, q, true), false};
                }
                r = (p < 0) ? scalar_div(1.0f, t) : t;
            }
            r.nan = u;
            return r;
        }
        case 25: {   // nth_root (interval.hpp:233-242), bounds as on the GPU
            const int n = int(b.lo);
            const bool u = a.nan || b.nan || (a.lo <= 0 && !(n & 2));
            if (bad(a)) return empty(u);
            if (n == 1) return {a.lo, a.hi, u};
            if (!(a.hi > 0)) {
                if (a.hi == 0) {
                    if (!(n & 1) || a.lo == 0) return {0, 0, u};
                    return {dn(nth(a.lo, n)), 0, u};
                }
                if (!(n & 1)) return empty(u);
                return {dn(nth(a.lo, n)), up(nth(a.hi, n)), u};
            }
            float l;
            if (a.lo > 0) l = dn(nth(a.lo, n));
            else if (!(n & 1)) l = 0;
            else l = dn(nth(a.lo, n));
            return {l, up(nth(a.hi, n)), u};
        }
        case 27:     // nanfill (interval.hpp:291-297)
            if (a.nan) { float lo = a.lo, hi = a.hi; hull(lo, hi, b.lo, b.hi); return {lo, hi, b.nan}; }
            return {a.lo, a.hi, false};
        case 31:     // compare (interval.hpp:299-301)
            if (a.hi < b.lo) return {-1, -1, false};
            if (a.lo > b.hi) return {1, 1, false};
            return {-1, 1, false};
        case 26: {   // mod (interval.hpp:244-289)
            float ol = std::fmin(b.lo, 0.0f), ou = std::fmax(0.0f, b.hi);
            if (std::isfinite(a.hi) && std::isfinite(a.lo)) {
                const int position = (b.hi >= 0.0f) + 2 * (b.lo <= 0.0f);
                if (position == 1 || position == 2) {
                    float ul = a.lo, uh = a.hi;
                    if (position == 2) { const float t = -uh; uh = -ul; ul = t; }
                    const Iv ab = interval_op(28, Iv{b.lo, b.hi, false}, Iv{0, 0, false});
                    float ql, qh;
                    if (bad(ab) || !(ul <= uh)) { ql = NAN; qh = NAN; }
                    else div_nz(ul, uh, ab.lo, ab.hi, ql, qh);
                    const int q0 = int(std::floor(ql));
                    if (q0 == int(std::floor(qh))) {
                        float ml, mu;
                        scalar_mul(float(q0), b.lo, b.hi, ml, mu);
                        ol = sub_dn(a.lo, mu);
                        ou = sub_up(a.hi, ml);
                    }
                } else if (position == 0) {
                    ol = NAN; ou = NAN;
                }
            }
            return {ol, ou, b.hi >= 0.0f && b.lo <= 0.0f};
        }
        case 6:
            return a;
        default:
            return empty(true);
    }
}

IntervalEval::IntervalEval(Deck& d) : deck(d), i(d.num_slots + 1) {
    for (auto& c : d.constants) i[c.first] = iv_checked(c.second, c.second);
}

Iv IntervalEval::eval(const float lo[3], const float hi[3], const Tape& tape) {
    i[deck.X] = iv_checked(lo[0], hi[0]);
    i[deck.Y] = iv_checked(lo[1], hi[1]);
    i[deck.Z] = iv_checked(lo[2], hi[2]);
    for (auto it = tape.t.rbegin(); it != tape.t.rend(); ++it) {
        i[it->id] = interval_op(it->op, i[it->a], i[it->b]);
    }
    return i[tape.i];
}

// IntervalEvaluator::push keep function (eval_interval.cpp:116-167).
std::shared_ptr<Tape> IntervalEval::push(const std::shared_ptr<Tape>& tape) {
    const float box[6] = {i[deck.X].lo, i[deck.X].hi, i[deck.Y].lo,
                          i[deck.Y].hi, i[deck.Z].lo, i[deck.Z].hi};
    return oracle::push(deck, tape,
        [&](uint32_t op, uint32_t, uint32_t a, uint32_t b) -> Keep {
            if (op == 20) {
                if (a == b) return KEEP_A;
                if (i[a].lo > i[b].hi) return KEEP_A;
                if (i[b].lo > i[a].hi) return KEEP_B;
                return KEEP_BOTH;
            } else if (op == 19) {
                if (a == b) return KEEP_A;
                if (i[a].lo > i[b].hi) return KEEP_B;
                if (i[b].lo > i[a].hi) return KEEP_A;
                return KEEP_BOTH;
            }
            return KEEP_ALWAYS;
        }, Tape::INTERVAL, box);
}

////////////////////////////////////////////////////////////////////////////////
// FP64 twin (same tape, double arithmetic) for error measurement.

static double value_op_f64(uint32_t op, double a, double b) {
    switch (op) {
        case 17: return a + b;
        case 18: return a * b;
        case 19: return (b < a) ? b : a;
        case 20: return (a < b) ? b : a;
        case 21: return a - b;
        case 22: return a / b;
        case 23: return std::atan2(a, b);
        case 24: return std::pow(a, b);
        case 25: return (a < 0) ? ((int(b) & 1) ? -std::pow(-a, 1.0 / b) : NAN) : std::pow(a, 1.0 / b);
        case 26: {
            double d = std::fabs(a / b);
            if ((a < 0) ^ (b < 0)) d = -std::ceil(d); else d = std::floor(d);
            double out = a - b * d;
            if ((b > 0 && out > b) || (b < 0 && out < b)) out = b;
            if ((b > 0 && out < 0) || (b < 0 && out > 0)) out = 0;
            return out;
        }
        case 27: return std::isnan(a) ? b : a;
        case 31: return (a < b) ? -1.0 : (a > b) ? 1.0 : 0.0;
        case 7: return a * a;
        case 8: return std::sqrt(a);
        case 9: return -a;
        case 10: return std::sin(a);
        case 11: return std::cos(a);
        case 12: return std::tan(a);
        case 13: return std::asin(a);
        case 14: return std::acos(a);
        case 15: return std::atan(a);
        case 16: return std::exp(a);
        case 28: return std::fabs(a);
        case 29: return 1 / a;
        case 30: return std::log(a);
        case 6: return a;
        default: return NAN;
    }
}

// Scratch v holds one value per slot.  It is initialised (zeros, then the
// constants) only when its size does not match the deck: clauses write
// only their own slots, so constants and the dummy slot 0 keep their
// values from one call to the next.
double eval_point_f64(const Deck& deck, const Tape& tape, double x, double y,
                      double z, std::vector<double>& v) {
    if (v.size() != size_t(deck.num_slots) + 1) {
        v.assign(deck.num_slots + 1, 0.0);
        for (auto& c : deck.constants) v[c.first] = c.second;
    }
    v[deck.X] = x;
    v[deck.Y] = y;
    v[deck.Z] = z;
    for (auto it = tape.t.rbegin(); it != tape.t.rend(); ++it) {
        v[it->id] = value_op_f64(it->op, v[it->a], v[it->b]);
    }
    return v[tape.i];
}

// Forward-mode partials in double, with the rules of
// DerivArrayEvaluator (libfive/src/eval/eval_deriv_array.cpp:119-256).
static void deriv_op_f64(uint32_t op, double av, const double* ad, double bv,
                         const double* bd, double ov, double* od) {
    for (int j = 0; j < 3; ++j) {
        double r;
        switch (op) {
            case 17: r = ad[j] + bd[j]; break;
            case 21: r = ad[j] - bd[j]; break;
            case 18: r = bd[j] * av + ad[j] * bv; break;
            case 19: r = (av < bv) ? ad[j] : bd[j]; break;
            case 20: r = (av < bv) ? bd[j] : ad[j]; break;
            case 22: r = (ad[j] * bv - bd[j] * av) / (bv * bv); break;
            case 23: r = (ad[j] * bv - bd[j] * av) / (av * av + bv * bv); break;
            case 24: r = ad[j] * (bv * std::pow(av, bv - 1)); break;
            case 25: r = (ad[j] == 0) ? 0.0 : ad[j] * (std::pow(av, 1.0 / bv - 1) / bv); break;
            case 26: r = ad[j]; break;
            case 27: r = std::isnan(av) ? bd[j] : ad[j]; break;
            case 31: r = 0; break;
            case 7: r = ad[j] * av * 2; break;
            case 8: r = (av < 0 || ad[j] == 0) ? 0.0 : ad[j] / (2 * ov); break;
            case 9: r = -ad[j]; break;
            case 10: r = ad[j] * std::cos(av); break;
            case 11: r = ad[j] * -std::sin(av); break;
            case 12: r = ad[j] / (std::cos(av) * std::cos(av)); break;
            case 13: r = ad[j] / std::sqrt(1 - av * av); break;
            case 14: r = ad[j] / -std::sqrt(1 - av * av); break;
            case 15: r = ad[j] / (av * av + 1); break;
            case 16: r = ad[j] * ov; break;
            case 28: r = (av > 0) ? ad[j] : -ad[j]; break;
            case 29: r = ad[j] / -(av * av); break;
            case 30: r = ad[j] / av; break;
            default: r = ad[j]; break;
        }
        od[j] = r;
    }
}

double eval_deriv_f64(const Deck& deck, const Tape& tape, double x, double y,
                      double z, std::vector<double>& v, double grad[3]) {
    const size_t n = deck.num_slots + 1;
    double* g;
    if (v.size() != 4 * n) {
        // first call with this scratch (see eval_point_f64)
        v.assign(4 * n, 0.0);
        g = v.data() + n;
        for (auto& c : deck.constants) v[c.first] = c.second;
        g[3 * deck.X] = 1;
        g[3 * deck.Y + 1] = 1;
        g[3 * deck.Z + 2] = 1;
    }
    g = v.data() + n;
    v[deck.X] = x;
    v[deck.Y] = y;
    v[deck.Z] = z;
    for (auto it = tape.t.rbegin(); it != tape.t.rend(); ++it) {
        const double av = v[it->a], bv = v[it->b];
        const double ov = value_op_f64(it->op, av, bv);
        double od[3];
        deriv_op_f64(it->op, av, g + 3 * it->a, bv, g + 3 * it->b, ov, od);
        v[it->id] = ov;
        for (int j = 0; j < 3; ++j) g[3 * it->id + j] = od[j];
    }
    for (int j = 0; j < 3; ++j) grad[j] = g[3 * tape.i + j];
    return v[tape.i];
}

}   // namespace oracle

////////////////////////////////////////////////////////////////////////////////
// C entry points

using namespace oracle;


template <typename F>
static void parallel_batches(size_t n, int threads, F&& f) {
    const size_t nb = (n + ArrayEval::N - 1) / ArrayEval::N;
    if (threads <= 1 || nb < 2) {
        f(size_t(0), nb);
        return;
    }
    std::vector<std::thread> pool;
    const size_t per = (nb + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        const size_t b0 = t * per, b1 = std::min(nb, b0 + per);
        if (b0 >= b1) break;
        pool.emplace_back([&f, b0, b1] { f(b0, b1); });
    }
    for (auto& th : pool) th.join();
}

extern "C" {

oracle_deck* oracle_deck_new(const uint32_t* clauses, size_t n,
                             const uint32_t* cids, const float* cvals,
                             size_t nc, uint32_t X, uint32_t Y, uint32_t Z,
                             uint32_t num_slots) {
    return new oracle_deck{Deck(clauses, n, cids, cvals, nc, X, Y, Z, num_slots)};
}

void oracle_deck_delete(oracle_deck* d) { delete d; }

void oracle_eval_values(oracle_deck* d, const float* x, const float* y,
                        const float* z, size_t n, float* out, int threads) {
    parallel_batches(n, threads, [&](size_t b0, size_t b1) {
        ArrayEval e(d->deck);
        for (size_t b = b0; b < b1; ++b) {
            const size_t s = b * ArrayEval::N;
            const size_t cnt = std::min(ArrayEval::N, n - s);
            for (size_t j = 0; j < cnt; ++j) e.set(x[s + j], y[s + j], z[s + j], j);
            const float* r = e.values(cnt, *d->deck.tape);
            std::memcpy(out + s, r, cnt * sizeof(float));
        }
    });
}

void oracle_eval_derivs(oracle_deck* d, const float* x, const float* y,
                        const float* z, size_t n, float* v, float* dx,
                        float* dy, float* dz, int threads) {
    parallel_batches(n, threads, [&](size_t b0, size_t b1) {
        DerivEval e(d->deck);
        const Tape& t = *d->deck.tape;
        for (size_t b = b0; b < b1; ++b) {
            const size_t s = b * ArrayEval::N;
            const size_t cnt = std::min(ArrayEval::N, n - s);
            for (size_t j = 0; j < cnt; ++j) e.set(x[s + j], y[s + j], z[s + j], j);
            e.derivs(cnt, t);
            std::memcpy(v + s, e.value_row(t), cnt * sizeof(float));
            std::memcpy(dx + s, e.grad(0, t), cnt * sizeof(float));
            std::memcpy(dy + s, e.grad(1, t), cnt * sizeof(float));
            std::memcpy(dz + s, e.grad(2, t), cnt * sizeof(float));
        }
    });
}

void oracle_eval_values_f64(oracle_deck* d, const double* x, const double* y,
                            const double* z, size_t n, double* out) {
    std::vector<double> v;
    for (size_t i = 0; i < n; ++i) {
        out[i] = eval_point_f64(d->deck, *d->deck.tape, x[i], y[i], z[i], v);
    }
}

void oracle_eval_intervals(oracle_deck* d, const float* xlo, const float* xhi,
                           const float* ylo, const float* yhi,
                           const float* zlo, const float* zhi, size_t n,
                           float* lower, float* upper, uint8_t* maybe_nan) {
    IntervalEval e(d->deck);
    for (size_t k = 0; k < n; ++k) {
        const float lo[3] = {xlo[k], ylo[k], zlo[k]};
        const float hi[3] = {xhi[k], yhi[k], zhi[k]};
        const Iv r = e.eval(lo, hi, *d->deck.tape);
        lower[k] = r.lo;
        upper[k] = r.hi;
        if (maybe_nan) maybe_nan[k] = r.nan;
    }
}

size_t oracle_interval_push(oracle_deck* d, const float lo[3],
                            const float hi[3], float* out_lo, float* out_hi,
                            uint8_t* out_nan, int* changed) {
    IntervalEval e(d->deck);
    const Iv r = e.eval(lo, hi, *d->deck.tape);
    auto t = e.push(d->deck.tape);
    if (out_lo) *out_lo = r.lo;
    if (out_hi) *out_hi = r.hi;
    if (out_nan) *out_nan = r.nan;
    if (changed) *changed = (t != d->deck.tape);
    return t->t.size();
}

// Values of `tape` at n points, 256 at a time (ArrayEvaluator::values).
static void eval_on_tape(Deck& deck, const Tape& tape, const float* x,
                         const float* y, const float* z, size_t n, float* out) {
    ArrayEval e(deck);
    for (size_t s = 0; s < n; s += ArrayEval::N) {
        const size_t cnt = std::min(ArrayEval::N, n - s);
        for (size_t j = 0; j < cnt; ++j) e.set(x[s + j], y[s + j], z[s + j], j);
        std::memcpy(out + s, e.values(cnt, tape), cnt * sizeof(float));
    }
}

void oracle_eval_ambiguous(oracle_deck* d, const float* x, const float* y,
                           const float* z, size_t n, uint8_t* out) {
    ArrayEval e(d->deck);
    const Tape& t = *d->deck.tape;
    for (size_t s = 0; s < n; s += ArrayEval::N) {
        const size_t cnt = std::min(ArrayEval::N, n - s);
        for (size_t j = 0; j < cnt; ++j) e.set(x[s + j], y[s + j], z[s + j], j);
        e.values(cnt, t);
        e.ambiguous(cnt, t, out + s);
    }
}

// DerivArrayEvaluator::getAmbiguousDerivs (eval_deriv_array.cpp:58-81): a
// MIN/MAX clause with equal operand values and operand partials that
// differ in at least one component.
void oracle_eval_deriv_ambiguous(oracle_deck* d, const float* x,
                                 const float* y, const float* z, size_t n,
                                 uint8_t* out) {
    DerivEval e(d->deck);
    const Tape& t = *d->deck.tape;
    for (size_t s = 0; s < n; s += ArrayEval::N) {
        const size_t cnt = std::min(ArrayEval::N, n - s);
        for (size_t j = 0; j < cnt; ++j) e.set(x[s + j], y[s + j], z[s + j], j);
        e.derivs(cnt, t);
        uint8_t* o = out + s;
        std::fill(o, o + cnt, 0);
        for (auto it = t.t.rbegin(); it != t.t.rend(); ++it) {
            if (it->op != 19 && it->op != 20) continue;
            for (size_t j = 0; j < cnt; ++j) {
                bool differ = false;
                for (int k = 0; k < 3; ++k) {
                    differ |= e.gslot(it->a, k, j) != e.gslot(it->b, k, j);
                }
                o[j] |= (e.slot(it->a, j) == e.slot(it->b, j)) && differ;
            }
        }
    }
}

// ArrayEvaluator::valueAndPush (eval_array.cpp:121-170): the value at p,
// then a SPECIALIZED push that keeps only the strictly larger (MAX) or
// smaller (MIN) operand; the pushed tape is evaluated at the nq points q.
// Returns the pushed tape's clause count.
size_t oracle_value_and_push(oracle_deck* d, const float p[3],
                             const float* qx, const float* qy,
                             const float* qz, size_t nq, float* out,
                             float* value) {
    Deck& deck = d->deck;
    ArrayEval e(deck);
    e.set(p[0], p[1], p[2], 0);
    const float v0 = e.values(1, *deck.tape)[0];
    auto t = oracle::push(deck, deck.tape,
        [&](uint32_t op, uint32_t, uint32_t a, uint32_t b) -> Keep {
            const float va = e.slot(a, 0), vb = e.slot(b, 0);
            if (op == 20) return (va > vb) ? KEEP_A : (vb > va) ? KEEP_B : KEEP_BOTH;
            if (op == 19) return (va > vb) ? KEEP_B : (vb > va) ? KEEP_A : KEEP_BOTH;
            return KEEP_ALWAYS;
        }, Tape::SPECIALIZED, nullptr);
    if (value) *value = v0;
    if (nq) eval_on_tape(deck, *t, qx, qy, qz, nq, out);
    return t->t.size();
}

// IntervalEvaluator::intervalAndPush on [lo, hi], then the pushed tape's
// values at the nq points q.  Returns the pushed tape's clause count.
size_t oracle_interval_push_eval(oracle_deck* d, const float lo[3],
                                 const float hi[3], const float* qx,
                                 const float* qy, const float* qz, size_t nq,
                                 float* out) {
    IntervalEval e(d->deck);
    e.eval(lo, hi, *d->deck.tape);
    auto t = e.push(d->deck.tape);
    if (nq) eval_on_tape(d->deck, *t, qx, qy, qz, nq, out);
    return t->t.size();
}

void oracle_sincos(const float* x, size_t n, float* s, float* c) {
    for (size_t i = 0; i < n; ++i) {
        s[i] = libfive::fmath::sin_f32(x[i]);
        c[i] = libfive::fmath::cos_f32(x[i]);
    }
}

}   // extern "C"
