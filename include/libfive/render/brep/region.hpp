// Axis-aligned region: libfive::Region<N>
// (libfive/include/libfive/render/brep/region.hpp:20-269) with its lower
// and upper corners, level and withResolution().  Points are a small value
// type instead of Eigen arrays: Region<3>({-1, -1, -1}, {1, 1, 1}) and
// r.lower[0] / r.lower(0) read as in the reference.
#pragma once

#include <cmath>
#include <cstdint>
#include <initializer_list>

namespace libfive {

template <unsigned N>
class Region {
public:
    struct Pt {
        double v[N];
        Pt() {
            for (unsigned i = 0; i < N; ++i) v[i] = 0.0;
        }
        Pt(std::initializer_list<double> xs) {
            unsigned i = 0;
            for (double x : xs) {
                if (i < N) v[i++] = x;
            }
            for (; i < N; ++i) v[i] = 0.0;
        }
        double& operator[](unsigned i) { return v[i]; }
        double operator[](unsigned i) const { return v[i]; }
        double& operator()(unsigned i) { return v[i]; }
        double operator()(unsigned i) const { return v[i]; }
    };

    Region() : level(-1) {}
    Region(Pt lower, Pt upper) : lower(lower), upper(upper), level(-1) {}
    Region(Pt lower, Pt upper, int32_t level)
        : lower(lower), upper(upper), level(level) {}

    bool contains(const Pt& p, double epsilon = 1e-8) const {
        for (unsigned i = 0; i < N; ++i) {
            if (p[i] < lower[i] - epsilon || p[i] > upper[i] + epsilon) return false;
        }
        return true;
    }
    Pt center() const {
        Pt c;
        for (unsigned i = 0; i < N; ++i) c[i] = (lower[i] + upper[i]) / 2;
        return c;
    }

    // The same region with the subdivision depth that brings the smallest
    // edge below min_feature (region.hpp:226-231, computed in float).
    Region<N> withResolution(double min_feature) const {
        double md = upper[0] - lower[0];
        for (unsigned i = 1; i < N; ++i) {
            const double d = upper[i] - lower[i];
            md = d < md ? d : md;
        }
        const float r = logf(float(md / min_feature));
        const float m = fmaxf(0.0f, r);
        return Region<N>(lower, upper, int32_t(std::ceil(m / logf(2.0f))));
    }

    Pt lower, upper;
    int32_t level;
};

}   // namespace libfive
