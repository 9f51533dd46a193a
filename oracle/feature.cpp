// Oracle (test infrastructure only): FeatureEvaluator restatement, the
// exact-zero path of corner classification and intersection normals.
//
// Follows:
//   libfive/src/eval/eval_array.cpp:121-170      valueAndPush (SPECIALIZED)
//   libfive/src/eval/eval_feature.cpp:56-372     isInside, features, walk
//   libfive/src/eval/feature.cpp:15-290          Feature push/check/order
#include "feature.hpp"

#include <algorithm>
#include <cmath>

#include "../libfive_amd/csrc/fmath.hpp"

namespace oracle {

// Three-term float sums are grouped x0 + (x1 + x2): Eigen's unrolled
// reduction for fixed size 3 (Eigen/src/Core/Redux.h, the reference's
// dependency) splits the range in halves when no SIMD packet fits.
static float dot3(const V3& a, const V3& b) {
    return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]);
}
static float norm3(const V3& a) { return std::sqrt(dot3(a, a)); }

// Feature::check (feature.cpp:116-241)
bool Feature::check(const V3& e, bool* duplicate) const {
    for (const auto& i : eps) {
        if (dot3(e, i) > 1 - 1e-8) {
            if (duplicate) *duplicate = true;
            return true;
        }
    }
    if (eps.empty()) return true;
    if (eps.size() == 1) return dot3(e, eps.front()) != -1;

    std::vector<V3> es = eps;
    es.push_back(e);
    std::vector<std::pair<V3, float>> planes;
    for (const auto& x : es) planes.push_back({x, 0.0f});
    planes.push_back({{1, 0, 0}, 1});
    planes.push_back({{-1, 0, 0}, 1});
    planes.push_back({{0, 1, 0}, 1});
    planes.push_back({{0, -1, 0}, 1});
    planes.push_back({{0, 0, 1}, 1});
    planes.push_back({{0, 0, -1}, 1});

    std::vector<V3> corners;
    for (size_t i = 0; i < planes.size(); ++i) {
        for (size_t j = i; j < planes.size(); ++j) {
            for (size_t k = j; k < planes.size(); ++k) {
                const auto& a = planes[i];
                const auto& b = planes[j];
                const auto& c = planes[k];
                if (a.second == 0.0f && b.second == 0.0f && c.second == 0.0f) {
                    continue;
                }
                // M has the three plane normals as its columns
                float m[3][3];
                for (int r = 0; r < 3; ++r) {
                    m[r][0] = a.first[r];
                    m[r][1] = b.first[r];
                    m[r][2] = c.first[r];
                }
                // Matrix3f::determinant: expansion along the first row
                auto det3 = [&](int p, int q, int t) {
                    return m[0][p] * (m[1][q] * m[2][t] - m[1][t] * m[2][q]);
                };
                const float det = det3(0, 1, 2) - det3(1, 0, 2) + det3(2, 0, 1);
                if (std::fabs(det) < 1e-6) continue;
                // Matrix3f::inverse: cofactors, with its own determinant
                // from the first column
                auto cof = [&](int i0, int j0) {
                    const int i1 = (i0 + 1) % 3, i2 = (i0 + 2) % 3;
                    const int j1 = (j0 + 1) % 3, j2 = (j0 + 2) % 3;
                    return m[i1][j1] * m[i2][j2] - m[i1][j2] * m[i2][j1];
                };
                const float c0[3] = {cof(0, 0), cof(1, 0), cof(2, 0)};
                const float idet =
                    1.0f / (c0[0] * m[0][0] + (c0[1] * m[1][0] + c0[2] * m[2][0]));
                float inv[3][3];
                for (int r = 0; r < 3; ++r) {
                    for (int q = 0; q < 3; ++q) inv[r][q] = cof(q, r) * idet;
                }
                const float d[3] = {a.second, b.second, c.second};
                V3 vert;
                for (int r = 0; r < 3; ++r) {
                    vert[r] = inv[r][0] * d[0] + (inv[r][1] * d[1] + inv[r][2] * d[2]);
                }
                if (std::find(corners.begin(), corners.end(), vert) == corners.end()) {
                    corners.push_back(vert);
                }
            }
        }
    }
    for (const auto& c : corners) {
        bool okay = true;
        for (const auto& x : es) {
            if (dot3(c, x) <= 0) { okay = false; break; }
        }
        if (okay) return true;
    }
    for (const auto& c : corners) {
        for (const auto& d : corners) {
            const V3 p = {(c[0] + d[0]) / 2.0f, (c[1] + d[1]) / 2.0f,
                          (c[2] + d[2]) / 2.0f};
            bool okay = true;
            for (const auto& x : es) {
                if (dot3(p, x) <= 0) { okay = false; break; }
            }
            if (okay) return true;
        }
    }
    return false;
}

// Feature::push (feature.cpp:66-96)
bool Feature::push(const V3& e_) {
    const float n = norm3(e_);
    if (n == 0) return false;
    const V3 e = {e_[0] / n, e_[1] / n, e_[2] / n};
    bool dup = false;
    if (check(e, &dup)) {
        if (!dup) eps.push_back(e);
        return true;
    }
    return false;
}

bool Feature::push(const Feature& other) {
    for (const auto& e : other.eps) {
        if (!push(e)) return false;
    }
    return true;
}

// operator< (feature.cpp:243-282)
static int vcmp(const V3& a, const V3& b) {
    for (int i = 0; i < 3; ++i) {
        if (a[i] < b[i]) return -1;
        if (a[i] > b[i]) return 1;
    }
    return 0;
}
bool Feature::operator<(const Feature& o) const {
    size_t i = 0;
    for (; i < eps.size() && i < o.eps.size(); ++i) {
        const int c = vcmp(eps[i], o.eps[i]);
        if (c != 0) return c < 0;
    }
    if (i < eps.size()) return false;
    if (i < o.eps.size()) return true;
    return vcmp(deriv, o.deriv) < 0;
}

static Feature merged(const V3& d, const Feature& a, const Feature& b) {
    Feature f;
    f.deriv = d;
    f.eps = a.eps;
    for (const auto& e : b.eps) {
        if (std::find(f.eps.begin(), f.eps.end(), e) == f.eps.end()) {
            f.eps.push_back(e);
        }
    }
    return f;
}

// Derivative rule of one clause for one (value, partials) pair, as in
// DerivArrayEvaluator::operator() (eval_deriv_array.cpp:119-256).
static V3 deriv_rule(uint32_t op, float av, const V3& ad, float bv,
                     const V3& bd, float ov) {
    V3 r;
    for (int j = 0; j < 3; ++j) {
        const float da = ad[j], db = bd[j];
        float o;
        switch (op) {
            case 17: o = da + db; break;
            case 21: o = da - db; break;
            case 18: { const float p = db * av; const float q = da * bv; o = p + q; break; }
            case 22: { const float p = da * bv; const float q = db * av; o = (p - q) / (bv * bv); break; }
            case 23: { const float p = da * bv; const float q = db * av;
                       const float a2 = av * av; const float b2 = bv * bv;
                       o = (p - q) / (a2 + b2); break; }
            case 24: o = da * (bv * float(std::pow(double(av), double(bv - 1.0f)))); break;
            case 25: o = (da == 0) ? 0.0f
                       : da * (float(std::pow(double(av), double(1.0f / bv - 1.0f))) / bv);
                     break;
            case 26: o = da; break;
            case 27: o = std::isnan(av) ? db : da; break;
            case 31: o = 0; break;
            case 7: o = (da * av) * 2.0f; break;
            case 8: o = (av < 0 || da == 0) ? 0.0f : da / (2.0f * ov); break;
            case 9: o = -da; break;
            case 10: o = da * libfive::fmath::cos_f32(av); break;
            case 11: o = da * -libfive::fmath::sin_f32(av); break;
            case 12: { const float c = 1.0f / libfive::fmath::cos_f32(av); o = da * (c * c); break; }
            case 13: o = da / std::sqrt(1.0f - av * av); break;
            case 14: o = da / -std::sqrt(1.0f - av * av); break;
            case 15: o = da / (av * av + 1.0f); break;
            case 30: o = da / av; break;
            case 16: o = da * float(std::exp(double(av))); break;
            case 28: o = (av > 0) ? da : -da; break;
            case 29: o = da / -(av * av); break;
            default: o = da; break;
        }
        r[j] = o;
    }
    return r;
}

FeatureEval::FeatureEval(Deck& d) : deck(d), arr(d) {}

std::shared_ptr<Tape> FeatureEval::value_and_push(const float p[3],
                                                  const std::shared_ptr<Tape>& tape,
                                                  float* value) {
    arr.set(p[0], p[1], p[2], 0);
    *value = arr.values(1, *tape)[0];
    ArrayEval& A = arr;
    return push(deck, tape,
        [&](uint32_t op, uint32_t, uint32_t a, uint32_t b) -> Keep {
            const float va = A.slot(a, 0), vb = A.slot(b, 0);
            if (op == 20) {
                if (va > vb) return KEEP_A;
                if (vb > va) return KEEP_B;
                return KEEP_BOTH;
            } else if (op == 19) {
                if (va > vb) return KEEP_B;
                if (vb > va) return KEEP_A;
                return KEEP_BOTH;
            }
            return KEEP_ALWAYS;
        }, Tape::SPECIALIZED, nullptr);
}

// Feature walk over a (specialised) tape (eval_feature.cpp:171-372).  The
// output value of a clause is read from the point slot (column 0) for every
// feature.
std::vector<Feature> FeatureEval::walk(const Tape& tape) {
    std::vector<std::vector<Feature>> f(deck.num_slots + 1);
    auto one = [](float x, float y, float z) {
        Feature g;
        g.deriv = {x, y, z};
        return std::vector<Feature>{g};
    };
    f[deck.X] = one(1, 0, 0);
    f[deck.Y] = one(0, 1, 0);
    f[deck.Z] = one(0, 0, 1);
    for (const auto& c : deck.constants) f[c.first] = one(0, 0, 0);

    for (auto it = tape.t.rbegin(); it != tape.t.rend(); ++it) {
        const uint32_t op = it->op, id = it->id, a = it->a, b = it->b;
        std::vector<Feature> of;
        const float av = arr.slot(a, 0), bv = arr.slot(b, 0);
        if (op == 19 || op == 20) {
            const bool is_min = (op == 19);
            if (av < bv || a == b) {
                of = is_min ? f[a] : f[b];
            } else if (av > bv) {
                of = is_min ? f[b] : f[a];
            } else {
                for (const auto& fa : f[a]) {
                    for (const auto& fb : f[b]) {
                        V3 eps;
                        for (int j = 0; j < 3; ++j) {
                            eps[j] = is_min ? (fb.deriv[j] - fa.deriv[j])
                                            : (fa.deriv[j] - fb.deriv[j]);
                        }
                        if (norm3(eps) == 0) {
                            if (!fa.eps.empty()) of.push_back(fa);
                            if (!fb.eps.empty()) of.push_back(fb);
                            if (fa.eps.empty() && fb.eps.empty()) of.push_back(fa);
                        } else {
                            Feature combined = fa;
                            if (combined.push(fb)) {
                                Feature ga = combined;
                                ga.deriv = fa.deriv;
                                if (ga.push(eps)) of.push_back(ga);
                                Feature gb = combined;
                                gb.deriv = fb.deriv;
                                const V3 neg = {-eps[0], -eps[1], -eps[2]};
                                if (gb.push(neg)) of.push_back(gb);
                            }
                        }
                    }
                }
            }
        } else if (libfive_args(op) == 1) {
            const float ov = arr.slot(id, 0);
            for (const auto& fa : f[a]) {
                Feature g;
                g.deriv = deriv_rule(op, av, fa.deriv, 0, {0, 0, 0}, ov);
                g.eps = fa.eps;
                of.push_back(g);
            }
        } else {
            const float ov = arr.slot(id, 0);
            for (const auto& fa : f[a]) {
                for (const auto& fb : f[b]) {
                    of.push_back(merged(deriv_rule(op, av, fa.deriv, bv,
                                                   fb.deriv, ov), fa, fb));
                }
            }
        }
        if (of.size() > 1) {
            std::sort(of.begin(), of.end());
            auto same = [](const Feature& f1, const Feature& f2) {
                V3 d;
                for (int j = 0; j < 3; ++j) d[j] = f1.deriv[j] - f2.deriv[j];
                return dot3(d, d) <= 1e-10 && f1.eps == f2.eps;
            };
            of.erase(std::unique(of.begin(), of.end(), same), of.end());
            const V3 first = of.front().deriv;
            bool all = true;
            for (const auto& g : of) {
                V3 d;
                for (int j = 0; j < 3; ++j) d[j] = g.deriv[j] - first[j];
                if (!(dot3(d, d) < 1e-10)) { all = false; break; }
            }
            if (all) {
                Feature g;
                g.deriv = first;
                of = {g};
            }
        }
        f[id] = of;
    }
    return f[tape.i];
}

// FeatureEvaluator::isInside (eval_feature.cpp:56-117)
bool FeatureEval::is_inside(const float p[3], const std::shared_ptr<Tape>& tape) {
    float v;
    auto h = value_and_push(p, tape, &v);
    if (v < 0) return true;
    if (v > 0) return false;
    const auto fs = walk(*h);
    if (fs.size() == 1) return norm3(fs.front().deriv) > 0;
    bool pos = false, neg = false;
    for (const auto& g : fs) {
        pos |= g.check(g.deriv, nullptr);
        const V3 n = {-g.deriv[0], -g.deriv[1], -g.deriv[2]};
        neg |= g.check(n, nullptr);
    }
    return !(pos && !neg);
}

// FeatureEvaluator::features (eval_feature.cpp:125-169): unique derivatives.
std::vector<V3> FeatureEval::features(const float p[3],
                                      const std::shared_ptr<Tape>& tape) {
    float v;
    auto h = value_and_push(p, tape, &v);
    const auto fs = walk(*h);
    std::vector<V3> out;
    for (const auto& g : fs) {
        if (std::find(out.begin(), out.end(), g.deriv) == out.end()) {
            out.push_back(g.deriv);
        }
    }
    return out;
}

int libfive_args(uint32_t op) {
    switch (op) {
        case 7: case 8: case 9: case 10: case 11: case 12: case 13: case 14:
        case 15: case 16: case 28: case 29: case 30: case 6:
            return 1;
        case 17: case 18: case 19: case 20: case 21: case 22: case 23: case 24:
        case 25: case 26: case 27: case 31:
            return 2;
        default:
            return 0;
    }
}

}   // namespace oracle
