// Oracle (test infrastructure only): occupancy raster of a Z slice,
// restating libfive_tree_render_pixels.
//
// Follows:
//   libfive/src/libfive.cpp:442-465                  region -> Voxels, pixels
//   libfive/src/render/discrete/voxels.cpp:23-45     voxel grid, centres
//   libfive/include/libfive/render/discrete/voxels.hpp:71-104   View::split
//   libfive/src/render/discrete/heightmap.cpp:106-246   pixels, fill, recurse
//   libfive/src/render/discrete/heightmap.cpp:275-319   initial XY regions
//
// The recursion is the reference's, depth first, one region after another
// (the reference gives each of its up to 8 regions a thread; the regions
// are disjoint, so the order does not change the image).  Every interval
// evaluation pushes a pruned tape that the region's children use, as the
// reference's intervalAndPush does.  The slice has one voxel in Z, so the
// depth test and the early exit down a Z column reduce to "set the pixel".
#include <algorithm>
#include <cmath>
#include <deque>
#include <memory>
#include <vector>

#include "oracle.h"
#include "oracle_impl.hpp"

namespace oracle {
namespace {

struct Grid {
    float lower[3], upper[3];
    std::vector<float> pts[3];
};

// Voxels::Voxels (voxels.cpp:23-45)
Grid voxel_grid(const float lo[3], const float hi[3], float res) {
    Grid g;
    for (int a = 0; a < 3; ++a) {
        const float d = hi[a] - lo[a];
        const int n = int(std::max(std::ceil(res * d), 1.0f));
        const float extra = res > 0 ? float(n) / res - d : 0.0f;
        g.lower[a] = lo[a] - extra / 2;
        g.upper[a] = hi[a] + extra / 2;
        for (int k = 0; k < n; ++k) {
            const float f = float((k + 0.5) / n);
            g.pts[a].push_back(g.lower[a] * (1 - f) + g.upper[a] * f);
        }
    }
    return g;
}

struct View {
    float lower[3], upper[3];
    int size[3], corner[3];
};

// View::split<A> (voxels.hpp:71-104)
std::pair<View, View> split(const View& v, unsigned axes) {
    int m[3];
    for (int a = 0; a < 3; ++a) m[a] = (axes >> a) & 1 ? v.size[a] : 0;
    int axis = 0;                        // maxCoeff: the first largest
    for (int a = 1; a < 3; ++a) {
        if (m[a] > m[axis]) axis = a;
    }
    const int up = v.size[axis] / 2;
    const int low = v.size[axis] - up;
    const float frac = low / float(v.size[axis]);
    const float mid = (v.upper[axis] * frac) + (v.lower[axis] * (1 - frac));
    View a = v, b = v;
    a.upper[axis] = mid;
    a.size[axis] = low;
    b.lower[axis] = mid;
    b.size[axis] = up;
    b.corner[axis] += low;
    return {a, b};
}

struct Raster {
    Deck& deck;
    const Grid& g;
    uint32_t width;
    uint8_t* out;
    IntervalEval ie;
    ArrayEval ae;

    Raster(Deck& d, const Grid& grid, uint32_t w, uint8_t* o)
        : deck(d), g(grid), width(w), out(o), ie(d), ae(d) {}

    void set(const View& r, int i, int j) {
        out[size_t(r.corner[1] + j) * width + r.corner[0] + i] = 1;
    }

    // Heightmap::pixels (heightmap.cpp:108-155)
    void pixels(const std::shared_ptr<Tape>& tape, const View& r) {
        size_t n = 0;
        const float z = g.pts[2][r.corner[2] + r.size[2] - 1];
        for (int i = 0; i < r.size[0]; ++i) {
            for (int j = 0; j < r.size[1]; ++j) {
                ae.set(g.pts[0][r.corner[0] + i], g.pts[1][r.corner[1] + j], z, n++);
            }
        }
        const float* v = ae.values(n, *tape);
        n = 0;
        for (int i = 0; i < r.size[0]; ++i) {
            for (int j = 0; j < r.size[1]; ++j) {
                if (v[n++] < 0) set(r, i, j);
            }
        }
    }

    // Heightmap::recurse (heightmap.cpp:193-246)
    void recurse(const std::shared_ptr<Tape>& tape, const View& r) {
        const size_t voxels = size_t(r.size[0]) * r.size[1] * r.size[2];
        if (voxels == 0) return;         // an empty block is "all below"
        if (voxels <= ArrayEval::N) {
            pixels(tape, r);
            return;
        }
        const Iv out = ie.eval(r.lower, r.upper, *tape);
        auto sub = ie.push(tape);
        if (out.hi < 0) {                // Interval::isFilled
            for (int i = 0; i < r.size[0]; ++i) {
                for (int j = 0; j < r.size[1]; ++j) set(r, i, j);
            }
        } else if (!(out.lo > 0)) {      // not Interval::isEmpty
            auto rs = split(r, 7);
            recurse(sub, rs.second);
            recurse(sub, rs.first);
        }
    }
};

}   // namespace
}   // namespace oracle

extern "C" {

int oracle_render_pixels(oracle_deck* d, float xlo, float xhi, float ylo,
                         float yhi, float z, float res, uint8_t* out,
                         uint32_t* width, uint32_t* height) {
    using namespace oracle;
    const float lo[3] = {xlo, ylo, z}, hi[3] = {xhi, yhi, z};
    const Grid g = voxel_grid(lo, hi, res);
    const uint32_t w = uint32_t(g.pts[0].size()), h = uint32_t(g.pts[1].size());
    if (width) *width = w;
    if (height) *height = h;
    if (!out) return 0;
    std::fill(out, out + size_t(w) * h, uint8_t(0));

    View root;
    for (int a = 0; a < 3; ++a) {
        root.lower[a] = g.lower[a];
        root.upper[a] = g.upper[a];
        root.size[a] = int(g.pts[a].size());
        root.corner[a] = 0;
    }
    std::deque<View> rs{root};
    while (rs.size() < 8 && std::min(rs.front().size[0], rs.front().size[1]) > 1) {
        const View f = rs.front();
        rs.pop_front();
        auto p = split(f, 3);
        rs.push_back(p.first);
        rs.push_back(p.second);
    }
    Raster r(d->deck, g, w, out);
    for (const View& v : rs) r.recurse(d->deck.tape, v);
    return 0;
}

void oracle_voxel_pts(float lo, float hi, float res, float* out, uint32_t* n,
                      float* lower, float* upper) {
    using namespace oracle;
    const float l3[3] = {lo, lo, lo}, h3[3] = {hi, hi, hi};
    const Grid g = voxel_grid(l3, h3, res);
    if (n) *n = uint32_t(g.pts[0].size());
    if (lower) *lower = g.lower[0];
    if (upper) *upper = g.upper[0];
    if (out) std::copy(g.pts[0].begin(), g.pts[0].end(), out);
}

}   // extern "C"
