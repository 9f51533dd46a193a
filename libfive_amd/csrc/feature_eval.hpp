// Device evaluation of the exact-zero path: the feature sets of a point
// where the field is zero and some MIN/MAX ties (FeatureEvaluator and
// Feature, libfive/src/eval/eval_feature.cpp:56-372, libfive/src/eval/
// feature.cpp:59-282).
//
// A feature is a gradient plus the unit "epsilon" directions that select
// it; a point's features are found by a forward walk over the tape in which
// each MIN/MAX that is strictly decided at the point passes one operand's
// features on (the SPECIALIZED push, eval_array.cpp:121-170) and each tied
// one combines both operands' features under the epsilon cone test.
//
// The routines run on one lane, with their lists in a per-cell global
// workspace.  Fixed capacities replace the reference's dynamic vectors;
// exceeding one raises an error flag instead of truncating.
//
// Three-term float sums follow Eigen's fixed-size reduction, grouped
// x0 + (x1 + x2).  Compile with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>

#include "fmath.hpp"
#include "opcode.hpp"
#include "point_eval.hpp"
#include "tape.hpp"

namespace libfive {
namespace amd {
namespace feat {

constexpr int kMaxFeat = 48;     // features per register
constexpr int kMaxEps = 8;       // epsilons per feature
constexpr int kMaxCorners = 720; // polytope corners in the cone test (C(17, 3))
constexpr int kMaxOut = 16;      // distinct gradients returned by features()

struct Feature {
    float d[3];
    int ne;
    float e[kMaxEps][3];
};

struct List {
    int n;
    Feature f[kMaxFeat];
};

// Workspace layout for a tape of n clauses over nregs registers.
struct Workspace {
    uint8_t* dec;       // per clause: 0 keep both / not min-max, 1 a, 2 b
    uint8_t* live;      // per clause
    float* val;         // per register
    uint32_t* rslot;    // per register: reference slot of the held value
    List* lists;        // per register, then two scratch lists
    float* corners;     // kMaxCorners x 3
};

__host__ __device__ inline size_t workspace_bytes(uint32_t n, uint32_t nregs) {
    const size_t a = (2 * size_t(n) + 15) & ~size_t(15);
    return a + 8 * size_t(nregs) + sizeof(List) * (size_t(nregs) + 2) +
           sizeof(float) * 3 * kMaxCorners + 64;
}

__device__ inline Workspace carve(uint8_t* base, uint32_t n, uint32_t nregs) {
    Workspace w;
    w.dec = base;
    w.live = base + n;
    size_t off = (2 * size_t(n) + 15) & ~size_t(15);
    w.val = reinterpret_cast<float*>(base + off);
    off += 4 * size_t(nregs);
    w.rslot = reinterpret_cast<uint32_t*>(base + off);
    off += 4 * size_t(nregs);
    off = (off + 15) & ~size_t(15);
    w.lists = reinterpret_cast<List*>(base + off);
    off += sizeof(List) * (size_t(nregs) + 2);
    w.corners = reinterpret_cast<float*>(base + off);
    return w;
}

__device__ __forceinline__ float dot3(const float* a, const float* b) {
    return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]);
}

__device__ __forceinline__ bool veq(const float* a, const float* b) {
    return a[0] == b[0] && a[1] == b[1] && a[2] == b[2];
}

// Feature::check (feature.cpp:116-241): is the cone of the feature's
// epsilons plus e non-empty?
__device__ bool check(const Feature& f, const float* e, bool* dup, float* corners,
                      uint32_t* err) {
    for (int i = 0; i < f.ne; ++i) {
        if (double(dot3(e, f.e[i])) > 1 - 1e-8) {
            if (dup) *dup = true;
            return true;
        }
    }
    if (f.ne == 0) return true;
    if (f.ne == 1) return dot3(e, f.e[0]) != -1;

    // Planes: the epsilons and e through the origin, then the unit cube.
    const int np = f.ne + 1 + 6;
    auto plane = [&](int i, float* nrm) -> float {
        if (i < f.ne) { nrm[0] = f.e[i][0]; nrm[1] = f.e[i][1]; nrm[2] = f.e[i][2]; return 0.0f; }
        if (i == f.ne) { nrm[0] = e[0]; nrm[1] = e[1]; nrm[2] = e[2]; return 0.0f; }
        const int c = i - f.ne - 1;
        nrm[0] = nrm[1] = nrm[2] = 0.0f;
        nrm[c >> 1] = (c & 1) ? -1.0f : 1.0f;
        return 1.0f;
    };
    int nc = 0;
    for (int i = 0; i < np; ++i) {
        for (int j = i; j < np; ++j) {
            for (int k = j; k < np; ++k) {
                float a[3], b[3], c[3];
                const float da = plane(i, a), db = plane(j, b), dc = plane(k, c);
                if (da == 0.0f && db == 0.0f && dc == 0.0f) continue;
                // M = [a b c] as columns
                float m[3][3];
                for (int r = 0; r < 3; ++r) { m[r][0] = a[r]; m[r][1] = b[r]; m[r][2] = c[r]; }
                auto det3 = [&](int p, int q, int t) {
                    return m[0][p] * (m[1][q] * m[2][t] - m[1][t] * m[2][q]);
                };
                const float det = det3(0, 1, 2) - det3(1, 0, 2) + det3(2, 0, 1);
                if (fabsf(det) < 1e-6) continue;
                auto cof = [&](int i0, int j0) {
                    const int i1 = (i0 + 1) % 3, i2 = (i0 + 2) % 3;
                    const int j1 = (j0 + 1) % 3, j2 = (j0 + 2) % 3;
                    return m[i1][j1] * m[i2][j2] - m[i1][j2] * m[i2][j1];
                };
                const float c0[3] = {cof(0, 0), cof(1, 0), cof(2, 0)};
                const float idet = 1.0f / (c0[0] * m[0][0] + (c0[1] * m[1][0] + c0[2] * m[2][0]));
                const float d[3] = {da, db, dc};
                float v[3];
                for (int r = 0; r < 3; ++r) {
                    const float i0 = cof(0, r) * idet, i1 = cof(1, r) * idet, i2 = cof(2, r) * idet;
                    v[r] = i0 * d[0] + (i1 * d[1] + i2 * d[2]);
                }
                bool seen = false;
                for (int q = 0; q < nc && !seen; ++q) seen = veq(corners + 3 * q, v);
                if (!seen) {
                    if (nc == kMaxCorners) { atomicOr(err, 4u); return false; }
                    corners[3 * nc] = v[0];
                    corners[3 * nc + 1] = v[1];
                    corners[3 * nc + 2] = v[2];
                    ++nc;
                }
            }
        }
    }
    auto inside_all = [&](const float* p) {
        for (int i = 0; i < f.ne; ++i) if (dot3(p, f.e[i]) <= 0) return false;
        return dot3(p, e) > 0;
    };
    for (int q = 0; q < nc; ++q) if (inside_all(corners + 3 * q)) return true;
    for (int q = 0; q < nc; ++q) {
        for (int r = 0; r < nc; ++r) {
            const float p[3] = {(corners[3 * q] + corners[3 * r]) / 2.0f,
                                (corners[3 * q + 1] + corners[3 * r + 1]) / 2.0f,
                                (corners[3 * q + 2] + corners[3 * r + 2]) / 2.0f};
            if (inside_all(p)) return true;
        }
    }
    return false;
}

// Feature::push (feature.cpp:59-96)
__device__ bool push_eps(Feature& f, const float* e_raw, float* corners, uint32_t* err) {
    const float n = sqrtf(dot3(e_raw, e_raw));
    if (n == 0) return false;
    const float e[3] = {e_raw[0] / n, e_raw[1] / n, e_raw[2] / n};
    bool dup = false;
    if (check(f, e, &dup, corners, err)) {
        if (!dup) {
            if (f.ne == kMaxEps) { atomicOr(err, 4u); return false; }
            f.e[f.ne][0] = e[0];
            f.e[f.ne][1] = e[1];
            f.e[f.ne][2] = e[2];
            f.ne++;
        }
        return true;
    }
    return false;
}

__device__ bool push_feature(Feature& f, const Feature& other, float* corners, uint32_t* err) {
    for (int i = 0; i < other.ne; ++i) {
        if (!push_eps(f, other.e[i], corners, err)) return false;
    }
    return true;
}

__device__ __forceinline__ int vcmp(const float* a, const float* b) {
    for (int i = 0; i < 3; ++i) {
        if (a[i] < b[i]) return -1;
        if (a[i] > b[i]) return 1;
    }
    return 0;
}

// Feature ordering (feature.cpp:243-282)
__device__ bool less(const Feature& a, const Feature& b) {
    int i = 0;
    for (; i < a.ne && i < b.ne; ++i) {
        const int c = vcmp(a.e[i], b.e[i]);
        if (c != 0) return c < 0;
    }
    if (i < a.ne) return false;
    if (i < b.ne) return true;
    return vcmp(a.d, b.d) < 0;
}

__device__ bool same_eps(const Feature& a, const Feature& b) {
    if (a.ne != b.ne) return false;
    for (int i = 0; i < a.ne; ++i) if (!veq(a.e[i], b.e[i])) return false;
    return true;
}

// Derivative rule of one clause for one gradient (eval_deriv_array.cpp:
// 119-256, as FeatureEvaluator applies it per feature).
__device__ void deriv_rule(uint32_t op, float av, const float* ad, float bv,
                           const float* bd, float ov, float* out) {
    namespace OC = Opcode;
    for (int j = 0; j < 3; ++j) {
        const float da = ad[j], db = bd[j];
        float o;
        switch (op) {
            case OC::OP_ADD: o = da + db; break;
            case OC::OP_SUB: o = da - db; break;
            case OC::OP_MUL: { const float p = db * av; const float q = da * bv; o = p + q; break; }
            case OC::OP_DIV: { const float p = da * bv; const float q = db * av; o = (p - q) / (bv * bv); break; }
            case OC::OP_ATAN2: { const float p = da * bv; const float q = db * av;
                                 const float a2 = av * av; const float b2 = bv * bv;
                                 o = (p - q) / (a2 + b2); break; }
            case OC::OP_POW: o = da * (bv * fmath::pow_f32(av, bv - 1.0f)); break;
            case OC::OP_NTH_ROOT: o = (da == 0) ? 0.0f : da * (fmath::pow_f32(av, 1.0f / bv - 1.0f) / bv); break;
            case OC::OP_MOD: o = da; break;
            case OC::OP_NANFILL: o = fmath::isnan_f(av) ? db : da; break;
            case OC::OP_COMPARE: o = 0; break;
            case OC::OP_SQUARE: o = (da * av) * 2.0f; break;
            case OC::OP_SQRT: o = (av < 0 || da == 0) ? 0.0f : da / (2.0f * ov); break;
            case OC::OP_NEG: o = -da; break;
            case OC::OP_SIN: o = da * fmath::cos_f32(av); break;
            case OC::OP_COS: o = da * -fmath::sin_f32(av); break;
            case OC::OP_TAN: { const float c = 1.0f / fmath::cos_f32(av); o = da * (c * c); break; }
            case OC::OP_ASIN: o = da / sqrtf(1.0f - av * av); break;
            case OC::OP_ACOS: o = da / -sqrtf(1.0f - av * av); break;
            case OC::OP_ATAN: o = da / (av * av + 1.0f); break;
            case OC::OP_LOG: o = da / av; break;
            case OC::OP_EXP: o = da * fmath::exp_f32(av); break;
            case OC::OP_ABS: o = (av > 0) ? da : -da; break;
            case OC::OP_RECIP: o = da / -(av * av); break;
            default: o = da; break;
        }
        out[j] = o;
    }
}

// Sort, deduplicate and collapse an output list (eval_feature.cpp: the
// tail of FeatureEvaluator::operator()).
__device__ void normalise(List& l) {
    if (l.n <= 1) return;
    for (int i = 1; i < l.n; ++i) {          // insertion sort
        Feature x = l.f[i];
        int j = i - 1;
        while (j >= 0 && less(x, l.f[j])) {
            l.f[j + 1] = l.f[j];
            --j;
        }
        l.f[j + 1] = x;
    }
    int m = 1;
    for (int i = 1; i < l.n; ++i) {
        const Feature& last = l.f[m - 1];
        float dd[3] = {l.f[i].d[0] - last.d[0], l.f[i].d[1] - last.d[1], l.f[i].d[2] - last.d[2]};
        const bool same = double(dot3(dd, dd)) <= 1e-10 && same_eps(l.f[i], last);
        if (!same) l.f[m++] = l.f[i];
    }
    l.n = m;
    bool all = true;
    for (int i = 0; i < l.n && all; ++i) {
        float dd[3] = {l.f[i].d[0] - l.f[0].d[0], l.f[i].d[1] - l.f[0].d[1], l.f[i].d[2] - l.f[0].d[2]};
        all = double(dot3(dd, dd)) < 1e-10;
    }
    if (all) {
        l.n = 1;
        l.f[0].ne = 0;
    }
}

__device__ __forceinline__ bool append(List& l, const Feature& f, uint32_t* err) {
    if (l.n == kMaxFeat) { atomicOr(err, 4u); return false; }
    l.f[l.n++] = f;
    return true;
}

// Features of the tape's root at p (FeatureEvaluator::features_ with the
// SPECIALIZED push).  Returns false when a capacity was exceeded.
__device__ bool walk(const tape::Clause* t, uint32_t n, uint32_t nregs, uint32_t root,
                     uint32_t slot_x, uint32_t slot_y, uint32_t slot_z,
                     const float p[3], Workspace& w, float* value, uint32_t* err) {
    namespace OC = Opcode;
    using namespace tape;
    // Pass 1: values, and the SPECIALIZED keep decision of each MIN/MAX.
    w.val[kRegX] = p[0];
    w.val[kRegY] = p[1];
    w.val[kRegZ] = p[2];
    for (uint32_t k = 0; k < n; ++k) {
        const Clause c = t[k];
        const uint32_t o = op(c);
        w.dec[k] = 0;
        if (o == OP_NOP) continue;
        const uint32_t m = mode(c);
        const float im = imm(c);
        const float av = (m == MODE_A_IMM || o == OP_LOADI) ? im : w.val[ra(c)];
        const float bv = is_binary(c) ? ((m == MODE_B_IMM) ? im : w.val[rb(c)]) : 0.0f;
        if (o == OC::OP_MAX) w.dec[k] = (av > bv) ? 1 : (bv > av) ? 2 : 0;
        if (o == OC::OP_MIN) w.dec[k] = (av > bv) ? 2 : (bv > av) ? 1 : 0;
        w.val[dst(c)] = pe_value_op(o, av, bv, im);
    }
    *value = w.val[root];
    // Pass 2: clauses reachable from the root after the specialisation.
    for (uint32_t r = 0; r < nregs; ++r) w.rslot[r] = 0;      // used as need flags
    w.rslot[root] = 1;
    for (uint32_t k = n; k-- > 0;) {
        const Clause c = t[k];
        const uint32_t o = op(c);
        w.live[k] = 0;
        if (o == OP_NOP) continue;
        const uint32_t d = dst(c);
        if (!w.rslot[d]) continue;
        w.rslot[d] = 0;
        w.live[k] = 1;
        const uint32_t m = mode(c);
        const bool ua = nargs(o) >= 1 && m != MODE_A_IMM && o != OP_LOADI;
        const bool ub = is_binary(c) && m != MODE_B_IMM;
        if (w.dec[k] == 1) { if (ua) w.rslot[ra(c)] = 1; }
        else if (w.dec[k] == 2) { if (ub) w.rslot[rb(c)] = 1; }
        else {
            if (ua) w.rslot[ra(c)] = 1;
            if (ub) w.rslot[rb(c)] = 1;
        }
    }
    // Pass 3: features of the live clauses.
    auto unit = [&](uint32_t r, float x, float y, float z, uint32_t slot) {
        w.lists[r].n = 1;
        w.lists[r].f[0].d[0] = x;
        w.lists[r].f[0].d[1] = y;
        w.lists[r].f[0].d[2] = z;
        w.lists[r].f[0].ne = 0;
        w.rslot[r] = slot;
    };
    unit(kRegX, 1, 0, 0, slot_x);
    unit(kRegY, 0, 1, 0, slot_y);
    unit(kRegZ, 0, 0, 1, slot_z);
    w.val[kRegX] = p[0];
    w.val[kRegY] = p[1];
    w.val[kRegZ] = p[2];
    List& zero = w.lists[nregs];
    zero.n = 1;
    zero.f[0].d[0] = zero.f[0].d[1] = zero.f[0].d[2] = 0.0f;
    zero.f[0].ne = 0;
    List& out = w.lists[nregs + 1];
    for (uint32_t k = 0; k < n; ++k) {
        if (!w.live[k]) continue;
        const Clause c = t[k];
        const uint32_t o = op(c);
        const uint32_t m = mode(c);
        const float im = imm(c);
        const uint32_t d = dst(c);
        const bool a_imm = (m == MODE_A_IMM || o == OP_LOADI);
        const bool b_imm = (m == MODE_B_IMM);
        const float av = a_imm ? im : w.val[ra(c)];
        const float bv = is_binary(c) ? (b_imm ? im : w.val[rb(c)]) : 0.0f;
        // A MIN/MAX decided at the point is replaced by the chosen operand
        // (whose partner may not have been evaluated in this pass).
        const float ov = (w.dec[k] == 1) ? av : (w.dec[k] == 2) ? bv
                                              : pe_value_op(o, av, bv, im);
        const List& la = a_imm ? zero : w.lists[ra(c)];
        const List& lb = (!is_binary(c) || b_imm) ? zero : w.lists[rb(c)];
        const uint32_t sa = a_imm ? 0xffffffffu : w.rslot[ra(c)];
        const uint32_t sb = (!is_binary(c) || b_imm) ? 0xfffffffeu : w.rslot[rb(c)];
        out.n = 0;
        uint32_t oslot = slot(c);
        if (o == OP_LOADI) {
            out = zero;
        } else if (o == OP_COPY) {
            out = la;
            oslot = sa;
        } else if (o == OC::OP_MIN || o == OC::OP_MAX) {
            const bool is_min = (o == OC::OP_MIN);
            if (w.dec[k] == 1) { out = la; oslot = sa; }
            else if (w.dec[k] == 2) { out = lb; oslot = sb; }
            else if (av < bv || sa == sb) {
                out = is_min ? la : lb;
            } else if (av > bv) {
                out = is_min ? lb : la;
            } else {
                for (int i = 0; i < la.n; ++i) {
                    for (int j = 0; j < lb.n; ++j) {
                        const Feature& fa = la.f[i];
                        const Feature& fb = lb.f[j];
                        float eps[3];
                        for (int q = 0; q < 3; ++q) {
                            eps[q] = is_min ? (fb.d[q] - fa.d[q]) : (fa.d[q] - fb.d[q]);
                        }
                        if (sqrtf(dot3(eps, eps)) == 0) {
                            if (fa.ne) append(out, fa, err);
                            if (fb.ne) append(out, fb, err);
                            if (!fa.ne && !fb.ne) append(out, fa, err);
                        } else {
                            Feature comb = fa;
                            if (push_feature(comb, fb, w.corners, err)) {
                                Feature ga = comb;
                                for (int q = 0; q < 3; ++q) ga.d[q] = fa.d[q];
                                if (push_eps(ga, eps, w.corners, err)) append(out, ga, err);
                                Feature gb = comb;
                                for (int q = 0; q < 3; ++q) gb.d[q] = fb.d[q];
                                const float neg[3] = {-eps[0], -eps[1], -eps[2]};
                                if (push_eps(gb, neg, w.corners, err)) append(out, gb, err);
                            }
                        }
                    }
                }
            }
        } else if (tape::nargs(o) == 1) {
            const float zd[3] = {0, 0, 0};
            for (int i = 0; i < la.n; ++i) {
                Feature g;
                deriv_rule(o, av, la.f[i].d, 0.0f, zd, ov, g.d);
                g.ne = la.f[i].ne;
                for (int q = 0; q < g.ne; ++q) {
                    g.e[q][0] = la.f[i].e[q][0];
                    g.e[q][1] = la.f[i].e[q][1];
                    g.e[q][2] = la.f[i].e[q][2];
                }
                append(out, g, err);
            }
        } else {
            for (int i = 0; i < la.n; ++i) {
                for (int j = 0; j < lb.n; ++j) {
                    Feature g;
                    deriv_rule(o, av, la.f[i].d, bv, lb.f[j].d, ov, g.d);
                    g.ne = la.f[i].ne;
                    for (int q = 0; q < g.ne; ++q) {
                        g.e[q][0] = la.f[i].e[q][0];
                        g.e[q][1] = la.f[i].e[q][1];
                        g.e[q][2] = la.f[i].e[q][2];
                    }
                    for (int q = 0; q < lb.f[j].ne; ++q) {
                        bool found = false;
                        for (int r = 0; r < g.ne && !found; ++r) found = veq(g.e[r], lb.f[j].e[q]);
                        if (found) continue;
                        if (g.ne == kMaxEps) { atomicOr(err, 4u); break; }
                        g.e[g.ne][0] = lb.f[j].e[q][0];
                        g.e[g.ne][1] = lb.f[j].e[q][1];
                        g.e[g.ne][2] = lb.f[j].e[q][2];
                        g.ne++;
                    }
                    append(out, g, err);
                }
            }
        }
        normalise(out);
        w.lists[d] = out;
        w.val[d] = ov;
        w.rslot[d] = oslot;
    }
    return true;
}

// FeatureEvaluator::isInside (eval_feature.cpp:56-117)
__device__ bool is_inside(const tape::Clause* t, uint32_t n, uint32_t nregs, uint32_t root,
                          uint32_t sx, uint32_t sy, uint32_t sz, const float p[3],
                          Workspace& w, uint32_t* err) {
    float v;
    // The value alone decides when it is not zero.
    walk(t, n, nregs, root, sx, sy, sz, p, w, &v, err);
    if (v < 0) return true;
    if (v > 0) return false;
    const List& fs = w.lists[root];
    if (fs.n == 1) return sqrtf(dot3(fs.f[0].d, fs.f[0].d)) > 0;
    bool pos = false, neg = false;
    for (int i = 0; i < fs.n; ++i) {
        pos |= check(fs.f[i], fs.f[i].d, nullptr, w.corners, err);
        const float nd[3] = {-fs.f[i].d[0], -fs.f[i].d[1], -fs.f[i].d[2]};
        neg |= check(fs.f[i], nd, nullptr, w.corners, err);
    }
    return !(pos && !neg);
}

// FeatureEvaluator::features (eval_feature.cpp:125-169): the distinct
// gradients, in order.  Returns the count (at most kMaxOut).
__device__ int features(const tape::Clause* t, uint32_t n, uint32_t nregs, uint32_t root,
                        uint32_t sx, uint32_t sy, uint32_t sz, const float p[3],
                        Workspace& w, float out[kMaxOut][3], uint32_t* err) {
    float v;
    walk(t, n, nregs, root, sx, sy, sz, p, w, &v, err);
    const List& fs = w.lists[root];
    int m = 0;
    for (int i = 0; i < fs.n; ++i) {
        bool found = false;
        for (int j = 0; j < m && !found; ++j) found = veq(out[j], fs.f[i].d);
        if (found) continue;
        if (m == kMaxOut) { atomicOr(err, 4u); break; }
        out[m][0] = fs.f[i].d[0];
        out[m][1] = fs.f[i].d[1];
        out[m][2] = fs.f[i].d[2];
        ++m;
    }
    return m;
}

}   // namespace feat
}   // namespace amd
}   // namespace libfive
