// Z-slice occupancy raster (see heightmap.hpp).
#include "heightmap.hpp"

#include <algorithm>
#include <cmath>
#include <deque>
#include <mutex>
#include <string>

#include "contour_kernels.hpp"
#include "deck.hpp"
#include "runtime.hpp"
#include "slice_tapes.hpp"

namespace libfive {
namespace amd {

namespace {

// Voxels (voxels.cpp:23-45): the region is widened to a whole number of
// voxels per axis and the sample points are voxel centres, in float.
struct Grid {
    float lower[3], upper[3];
    std::vector<float> pts[3];
};

Grid make_grid(const float lo[3], const float hi[3], float res) {
    Grid g;
    for (int i = 0; i < 3; ++i) {
        const float span = hi[i] - lo[i];
        const int size = int(std::max(std::ceil(res * span), 1.0f));
        const float extra = (res > 0.0f) ? float(size) / res - span : 0.0f;
        g.lower[i] = lo[i] - extra / 2;
        g.upper[i] = hi[i] + extra / 2;
        for (int index = 0; index < size; ++index) {
            const float frac = float((index + 0.5) / size);
            g.pts[i].push_back(g.lower[i] * (1 - frac) + g.upper[i] * frac);
        }
    }
    return g;
}

// Voxels::View (voxels.hpp:40-120): a box, its voxel counts, its first
// voxel.
struct View {
    float lower[3], upper[3];
    int size[3];
    int corner[3];
    size_t voxels() const { return size_t(size[0]) * size[1] * size[2]; }
};

// View::split<A> (voxels.hpp:71-104): halves the largest axis among those
// in mask A (the first one on ties); the lower half gets the extra voxel.
void split(const View& v, unsigned mask, View& a, View& b) {
    int axis = 0, best = -1;
    for (int i = 0; i < 3; ++i) {
        const int s = (mask & (1u << i)) ? v.size[i] : 0;
        if (s > best) {
            best = s;
            axis = i;
        }
    }
    const int size_upper = v.size[axis] / 2;
    const int size_lower = v.size[axis] - size_upper;
    const float frac = size_lower / float(v.size[axis]);
    const float middle = (v.upper[axis] * frac) + (v.lower[axis] * (1 - frac));
    a = v;
    b = v;
    a.upper[axis] = middle;
    a.size[axis] = size_lower;
    b.lower[axis] = middle;
    b.size[axis] = size_upper;
    b.corner[axis] += size_lower;
}

}   // namespace

PixelImage render_pixels(const Tree& t, float xlo, float xhi, float ylo,
                         float yhi, float z, float res) {
    Context& c = context();
    const float lo[3] = {xlo, ylo, z}, hi[3] = {xhi, yhi, z};
    const Grid g = make_grid(lo, hi, res);
    if (double(g.pts[0].size()) * double(g.pts[1].size()) >= 4294967296.0) {
        throw Error("render_pixels: the image has 2^32 pixels or more");
    }
    PixelImage img;
    img.width = uint32_t(g.pts[0].size());
    img.height = uint32_t(g.pts[1].size());
    img.filled.assign(size_t(img.width) * img.height, 0);
    const float zs = g.pts[2].back();    // the slice's single voxel centre

    const Deck deck(t);
    const DeviceTape base = compile_deck(deck);
    if (base.num_regs > 256) {
        throw Error("render_pixels: the tape needs " + std::to_string(base.num_regs) +
                    " registers; the raster supports 256");
    }
    std::lock_guard<std::mutex> lock(c.mu);
    SliceTapes tapes(base, deck.X, deck.Y, deck.Z, c.stream);

    // Initial regions: XY halves until there are 8 (Heightmap::render with
    // 8 workers, heightmap.cpp:285-294), all with the base tape.
    View root;
    for (int i = 0; i < 3; ++i) {
        root.lower[i] = g.lower[i];
        root.upper[i] = g.upper[i];
        root.size[i] = int(g.pts[i].size());
        root.corner[i] = 0;
    }
    std::deque<View> rs{root};
    while (rs.size() < 8 && std::min(rs.front().size[0], rs.front().size[1]) > 1) {
        const View f = rs.front();
        rs.pop_front();
        View a, b;
        split(f, 3u, a, b);
        rs.push_back(a);
        rs.push_back(b);
    }

    auto fill = [&](const View& v) {
        for (int j = 0; j < v.size[1]; ++j) {
            for (int i = 0; i < v.size[0]; ++i) {
                img.filled[size_t(v.corner[1] + j) * img.width + v.corner[0] + i] = 1;
            }
        }
    };

    // Heightmap::recurse (heightmap.cpp:193-246), one level at a time.  A
    // view is interval-evaluated with the tape its parent pushed
    // (intervalAndPush, heightmap.cpp:219-222), and both halves of an
    // ambiguous view, or its pixels, use the tape pushed for it.
    std::vector<View> frontier(rs.begin(), rs.end()), pixel_views;
    std::vector<uint32_t> ftape(frontier.size(), 0), pixel_tapes;
    std::vector<float> boxes, bounds;
    std::vector<uint8_t> nan;
    std::vector<uint32_t> btape, pushed;
    while (!frontier.empty()) {
        std::vector<View> big;
        btape.clear();
        for (size_t k = 0; k < frontier.size(); ++k) {
            const View& v = frontier[k];
            if (v.voxels() == 0) continue;
            if (v.voxels() <= 256) {
                pixel_views.push_back(v);
                pixel_tapes.push_back(ftape[k]);
            } else {
                big.push_back(v);
                btape.push_back(ftape[k]);
            }
        }
        frontier.clear();
        ftape.clear();
        const size_t n = big.size();
        if (n == 0) break;
        boxes.resize(6 * n);
        for (size_t k = 0; k < n; ++k) {
            for (int a = 0; a < 3; ++a) {
                boxes[6 * k + 2 * a] = big[k].lower[a];
                boxes[6 * k + 2 * a + 1] = big[k].upper[a];
            }
        }
        tapes.eval(boxes, btape, true, bounds, nan, pushed);
        for (size_t k = 0; k < n; ++k) {
            // Interval::isFilled / isEmpty on the bounds alone; the
            // may-be-NaN flag plays no part here (heightmap.cpp:221-233).
            if (bounds[2 * k + 1] < 0.0f) {
                fill(big[k]);
                continue;
            } else if (bounds[2 * k] > 0.0f) {
                continue;
            }
            View a, b;
            split(big[k], 7u, a, b);
            frontier.push_back(b);
            frontier.push_back(a);
            ftape.push_back(pushed[k]);
            ftape.push_back(pushed[k]);
        }
    }

    // Heightmap::pixels (heightmap.cpp:100-155): point values at the voxel
    // centres of the small views, one workgroup per view; filled where the
    // value is below zero.
    const size_t nv = pixel_views.size();
    if (nv) {
        std::vector<int32_t> vd(4 * nv);
        std::vector<uint32_t> offs(nv);
        uint32_t np = 0;
        for (size_t k = 0; k < nv; ++k) {
            const View& v = pixel_views[k];
            vd[4 * k] = v.corner[0];
            vd[4 * k + 1] = v.corner[1];
            vd[4 * k + 2] = v.size[0];
            vd[4 * k + 3] = v.size[1];
            offs[k] = np;
            np += uint32_t(v.size[0] * v.size[1]);
        }
        const size_t w = g.pts[0].size(), hgt = g.pts[1].size();
        PoolBuf<float> dp(w + hgt);
        PoolBuf<int32_t> dv(4 * nv);
        PoolBuf<uint32_t> doff(nv), dvt(nv);
        PoolBuf<uint8_t> dout(np);
        hip_check(hipMemcpyAsync(dp.p, g.pts[0].data(), 4 * w, hipMemcpyHostToDevice, c.stream), "copy x");
        hip_check(hipMemcpyAsync(dp.p + w, g.pts[1].data(), 4 * hgt, hipMemcpyHostToDevice, c.stream), "copy y");
        hip_check(hipMemcpyAsync(dv.p, vd.data(), 16 * nv, hipMemcpyHostToDevice, c.stream), "copy views");
        hip_check(hipMemcpyAsync(doff.p, offs.data(), 4 * nv, hipMemcpyHostToDevice, c.stream), "copy offsets");
        hip_check(hipMemcpyAsync(dvt.p, pixel_tapes.data(), 4 * nv, hipMemcpyHostToDevice, c.stream), "copy view tapes");
        const PoolTapes T{tapes.pool(), tapes.records(), tapes.root(), tapes.nregs(),
                          tapes.max_len()};
        hip_check(launch_raster_views(T, dvt.p, dp.p, dp.p + w, zs, dv.p, doff.p, dout.p, nv,
                                      c.stream),
                  "pixel values");
        std::vector<uint8_t> out(np);
        hip_check(hipMemcpyAsync(out.data(), dout.p, np, hipMemcpyDeviceToHost, c.stream),
                  "copy pixels");
        hip_check(hipStreamSynchronize(c.stream), "pixel values");
        for (size_t k = 0; k < nv; ++k) {
            const View& v = pixel_views[k];
            const uint8_t* o = out.data() + offs[k];
            for (int i = 0; i < v.size[0]; ++i) {
                for (int j = 0; j < v.size[1]; ++j) {
                    if (o[i * v.size[1] + j]) {
                        img.filled[size_t(v.corner[1] + j) * img.width + v.corner[0] + i] = 1;
                    }
                }
            }
        }
    }
    return img;
}

}   // namespace amd
}   // namespace libfive
