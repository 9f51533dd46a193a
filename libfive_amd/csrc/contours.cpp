// 2D dual contouring of a Z slice (see contours.hpp).
#include "contours.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <list>
#include <mutex>
#include <thread>

#include "contour_kernels.hpp"
#include "deck.hpp"
#include "qef.hpp"
#include "runtime.hpp"
#include "slice_tapes.hpp"

namespace libfive {
namespace amd {

namespace {

enum : uint8_t { EMPTY = 0, FILLED = 1, AMBIGUOUS = 2 };

// MarchingTable<2> (marching.cpp:71-98, 374-514): the four seed cases of
// marching squares closed under the quarter turn (x, y) -> (-y, x); edge
// ids numbered by first use over the masks in order.
struct Squares {
    int8_t v[16][2][4][2];
    int8_t e[4][4];
    int8_t p[16][8];
    Squares() {
        std::memset(v, -1, sizeof(v));
        std::memset(e, -1, sizeof(e));
        std::memset(p, -1, sizeof(p));
        const int seeds[][4] = {{1, 0, 0, 1}, {1, 0, 0, 2}, {3, 0, 1, 3}, {3, 0, 0, 2},
                                {9, 0, 0, 1}, {9, 0, 0, 2}, {9, 1, 3, 2}, {9, 1, 3, 1},
                                {7, 0, 1, 3}, {7, 0, 2, 3}};
        int fill[16][2] = {};
        for (const auto& s : seeds) {
            const int k = fill[s[0]][s[1]]++;
            v[s[0]][s[1]][k][0] = int8_t(s[2]);
            v[s[0]][s[1]][k][1] = int8_t(s[3]);
        }
        auto turn = [](int c) {                 // corner id after the turn
            const int x = (c & 1) ? 1 : -1, y = (c & 2) ? 1 : -1;
            return ((-y) > 0 ? 1 : 0) | (x > 0 ? 2 : 0);
        };
        bool pending[16];
        for (int i = 0; i < 16; ++i) pending[i] = v[i][0][0][0] >= 0;
        for (bool again = true; again;) {
            again = false;
            for (int i = 0; i < 16; ++i) {
                if (!pending[i]) continue;
                pending[i] = false;
                int j = 0;
                for (int c = 0; c < 4; ++c) j |= ((i >> c) & 1) << turn(c);
                if (v[j][0][0][0] >= 0) continue;
                pending[j] = again = true;
                for (int q = 0; q < 2 && v[i][q][0][0] >= 0; ++q) {
                    for (int k = 0; k < 4 && v[i][q][k][0] >= 0; ++k) {
                        v[j][q][k][0] = int8_t(turn(v[i][q][k][0]));
                        v[j][q][k][1] = int8_t(turn(v[i][q][k][1]));
                    }
                }
            }
        }
        int next = 0;
        for (int i = 0; i < 16; ++i) {
            for (int q = 0; q < 2 && v[i][q][0][0] >= 0; ++q) {
                for (int k = 0; k < 4 && v[i][q][k][0] >= 0; ++k) {
                    int8_t& id = e[v[i][q][k][0]][v[i][q][k][1]];
                    if (id < 0) id = int8_t(next++);
                    p[i][id] = int8_t(q);
                }
            }
        }
    }
};

const Squares& squares() {
    static const Squares s;
    return s;
}

struct Leaf {
    int level = 0;
    double vert[2][3] = {};
    int rank = 0;
    uint8_t mask = 0;
    int count = 0;
    bool manifold = false;
    double mass[4] = {0, 0, 0, 0};
    double ata[3][3] = {};
    double atb[3] = {0, 0, 0};
    double btb = 0;
    uint32_t index[2] = {0, 0};
};

struct Cell {
    double lo[2], hi[2];
    uint32_t key[2];
    int level;
    uint8_t type = AMBIGUOUS;
    int32_t kids = -1;        // first of four consecutive children
    int32_t leaf = -1;
    uint32_t tape_in = 0;     // tape the cell is evaluated with
    uint32_t tape_out = 0;    // tape pushed for it (its children's tape_in)
};

struct Quadtree {
    std::vector<Cell> cells;
    std::vector<Leaf> leaves;
    uint8_t corner(const Cell& c, int i) const {
        if (c.type == AMBIGUOUS) return (leaves[c.leaf].mask >> i) & 1 ? FILLED : EMPTY;
        return c.type;
    }
    const Cell& child(const Cell& c, int i) const {
        return c.kids >= 0 ? cells[c.kids + i] : c;
    }
    bool branch(const Cell& c) const { return c.kids >= 0; }
};

// One device round trip: points (x, y, z interleaved) and their tapes in,
// results out.
struct Device {
    Context& c;
    const SliceTapes& tp;
    uint32_t sx, sy, sz;
    PoolBuf<uint32_t> err{1};

    PoolTapes tapes() const {
        return {tp.pool(), tp.records(), tp.root(), tp.nregs(), tp.max_len()};
    }

    void values(const std::vector<float>& xyz, const std::vector<uint32_t>& tid,
                std::vector<float>& v, std::vector<uint8_t>& amb) {
        const size_t n = xyz.size() / 3;
        v.resize(n);
        amb.resize(n);
        if (!n) return;
        PoolBuf<float> d(4 * n);
        PoolBuf<uint8_t> da(n);
        PoolBuf<uint32_t> dt(n);
        hip_check(hipMemcpyAsync(d.p, xyz.data(), 12 * n, hipMemcpyHostToDevice, c.stream), "copy points");
        hip_check(hipMemcpyAsync(dt.p, tid.data(), 4 * n, hipMemcpyHostToDevice, c.stream), "copy tapes");
        hip_check(launch_point_values(tapes(), dt.p, d.p, d.p + 3 * n, da.p, n, c.stream),
                  "point values");
        hip_check(hipMemcpyAsync(v.data(), d.p + 3 * n, 4 * n, hipMemcpyDeviceToHost, c.stream), "copy values");
        hip_check(hipMemcpyAsync(amb.data(), da.p, n, hipMemcpyDeviceToHost, c.stream), "copy flags");
        hip_check(hipStreamSynchronize(c.stream), "point values");
    }

    void grads(const std::vector<float>& xyz, const std::vector<uint32_t>& tid,
               std::vector<float>& v, std::vector<float>& g, std::vector<uint8_t>& amb) {
        const size_t n = xyz.size() / 3;
        v.resize(n);
        g.resize(3 * n);
        amb.resize(n);
        if (!n) return;
        PoolBuf<float> d(7 * n);
        PoolBuf<uint8_t> da(n);
        PoolBuf<uint32_t> dt(n);
        hip_check(hipMemcpyAsync(d.p, xyz.data(), 12 * n, hipMemcpyHostToDevice, c.stream), "copy points");
        hip_check(hipMemcpyAsync(dt.p, tid.data(), 4 * n, hipMemcpyHostToDevice, c.stream), "copy tapes");
        hip_check(launch_point_grads(tapes(), dt.p, d.p, d.p + 3 * n, d.p + 4 * n, da.p, n,
                                     c.stream), "point partials");
        hip_check(hipMemcpyAsync(v.data(), d.p + 3 * n, 4 * n, hipMemcpyDeviceToHost, c.stream), "copy values");
        hip_check(hipMemcpyAsync(g.data(), d.p + 4 * n, 12 * n, hipMemcpyDeviceToHost, c.stream), "copy partials");
        hip_check(hipMemcpyAsync(amb.data(), da.p, n, hipMemcpyDeviceToHost, c.stream), "copy flags");
        hip_check(hipStreamSynchronize(c.stream), "point partials");
    }

    // isInside (want_inside) or the feature gradients of each point.
    void features(const std::vector<float>& xyz, const std::vector<uint32_t>& tid,
                  bool want_inside, std::vector<uint8_t>& inside, std::vector<float>& feats,
                  std::vector<int32_t>& nfeat) {
        const size_t n = xyz.size() / 3;
        inside.resize(n);
        feats.resize(n * 16 * 3);
        nfeat.resize(n);
        if (!n) return;
        const size_t wsb = point_feature_workspace(tp.max_len(), tp.nregs());
        const size_t waves = std::min<size_t>(n, kFeatureWaves);
        PoolBuf<float> d(3 * n + n * 16 * 3);
        PoolBuf<uint8_t> di(n), ws(waves * wsb);
        PoolBuf<int32_t> dn(n);
        PoolBuf<uint32_t> dt(n);
        hip_check(hipMemsetAsync(err.p, 0, 4, c.stream), "clear error");
        hip_check(hipMemcpyAsync(d.p, xyz.data(), 12 * n, hipMemcpyHostToDevice, c.stream), "copy points");
        hip_check(hipMemcpyAsync(dt.p, tid.data(), 4 * n, hipMemcpyHostToDevice, c.stream), "copy tapes");
        hip_check(launch_point_features(tapes(), dt.p, sx, sy, sz, d.p,
                                        want_inside ? di.p : nullptr,
                                        want_inside ? nullptr : d.p + 3 * n, dn.p, ws.p, wsb,
                                        err.p, n, c.stream), "point features");
        uint32_t e = 0;
        if (want_inside) {
            hip_check(hipMemcpyAsync(inside.data(), di.p, n, hipMemcpyDeviceToHost, c.stream), "copy inside");
        } else {
            hip_check(hipMemcpyAsync(feats.data(), d.p + 3 * n, 4 * n * 48, hipMemcpyDeviceToHost, c.stream), "copy features");
            hip_check(hipMemcpyAsync(nfeat.data(), dn.p, 4 * n, hipMemcpyDeviceToHost, c.stream), "copy counts");
        }
        hip_check(hipMemcpyAsync(&e, err.p, 4, hipMemcpyDeviceToHost, c.stream), "copy error");
        hip_check(hipStreamSynchronize(c.stream), "point features");
        if (e) {
            throw Error("render_slice: a point on a min/max feature has more "
                        "features than the GPU feature evaluation holds");
        }
    }
};

// DCTree<2>::leafsAreManifold (dc_tree2.cpp:26-47)
bool children_manifold(const Quadtree& q, const Cell* cs[4], const uint8_t corners[4]) {
    auto st = [&](int k, int i) { return q.corner(*cs[k], i); };
    auto either = [&](uint8_t s, int a, int b) { return s == corners[a] || s == corners[b]; };
    const uint8_t mid = st(0, 3);
    return either(st(0, 1), 0, 1) && either(st(0, 2), 0, 2) &&
           either(st(1, 3), 1, 3) && either(st(2, 3), 2, 3) &&
           (mid == corners[0] || mid == corners[2] || mid == corners[1] || mid == corners[3]);
}

struct Edge {
    int32_t leaf;
    int patch;
    int id;                  // marching-table edge id
    uint32_t tape;           // the leaf's tape
    double a[2], b[2];       // inside and outside bracket ends
};

}   // namespace

ContourResult render_contours(const Tree& tree, const double lo[2], const double hi[2],
                              float z, double min_feature, double max_err) {
    Context& c = context();
    const Deck deck(tree);
    const DeviceTape base = compile_deck(deck);
    if (base.num_regs > 256) {
        throw Error("render_slice: the tape needs " + std::to_string(base.num_regs) +
                    " registers; the contourer supports 256");
    }
    const float mind = float(std::min(hi[0] - lo[0], hi[1] - lo[1]) / min_feature);
    const int top = int(std::ceil(std::fmax(0.0f, std::log(mind)) / std::log(2.0f)));
    if (top > 24) throw Error("render_slice: resolution too fine (more than 24 levels)");

    std::lock_guard<std::mutex> lock(c.mu);
    SliceTapes tapes(base, deck.X, deck.Y, deck.Z, c.stream);
    Device dev{c, tapes, deck.X, deck.Y, deck.Z};
    const Squares& M = squares();
    Quadtree q;
    ContourResult out;
    {
        Cell root;
        for (int j = 0; j < 2; ++j) {
            root.lo[j] = lo[j];
            root.hi[j] = hi[j];
            root.key[j] = 0;
        }
        root.level = top;
        q.cells.push_back(root);
    }

    // Interval subdivision, one launch per level (DCTree::evalInterval,
    // dc_tree.inl:94-126): each cell is evaluated with its parent's pushed
    // tape and pushes its own, except where the result may be NaN.
    std::vector<std::vector<int32_t>> branches(top + 1);
    std::vector<int32_t> frontier{0};
    std::vector<float> box, r;
    std::vector<uint8_t> nan;
    std::vector<uint32_t> tin, pushed;
    for (int L = top; L > 0 && !frontier.empty(); --L) {
        const size_t n = frontier.size();
        box.resize(6 * n);
        tin.resize(n);
        for (size_t k = 0; k < n; ++k) {
            const Cell& cl = q.cells[frontier[k]];
            box[6 * k] = float(cl.lo[0]);
            box[6 * k + 1] = float(cl.hi[0]);
            box[6 * k + 2] = float(cl.lo[1]);
            box[6 * k + 3] = float(cl.hi[1]);
            box[6 * k + 4] = z;
            box[6 * k + 5] = z;
            tin[k] = cl.tape_in;
        }
        tapes.eval(box, tin, false, r, nan, pushed);
        std::vector<int32_t> next;
        for (size_t k = 0; k < n; ++k) {
            const int32_t ci = frontier[k];
            // Interval::state (interval.hpp:66-77)
            uint8_t ty = AMBIGUOUS;
            if (!nan[k]) {
                if (r[2 * k] > 0) ty = EMPTY;
                else if (r[2 * k + 1] < 0) ty = FILLED;
            }
            q.cells[ci].type = ty;
            q.cells[ci].tape_out = pushed[k];
            const Cell cl = q.cells[ci];
            out.cells.insert(out.cells.end(), {cl.key[0], cl.key[1], uint32_t(L), ty});
            if (ty != AMBIGUOUS) continue;
            const int32_t first = int32_t(q.cells.size());
            q.cells[ci].kids = first;
            branches[L].push_back(ci);
            for (int i = 0; i < 4; ++i) {
                Cell kid;
                for (int j = 0; j < 2; ++j) {
                    const double mid = (cl.lo[j] + cl.hi[j]) / 2;
                    const bool up = (i >> j) & 1;
                    kid.lo[j] = up ? mid : cl.lo[j];
                    kid.hi[j] = up ? cl.hi[j] : mid;
                    kid.key[j] = cl.key[j] + (up ? (1u << (L - 1)) : 0u);
                }
                kid.level = L - 1;
                kid.tape_in = pushed[k];
                q.cells.push_back(kid);
                next.push_back(first + i);
            }
        }
        frontier.swap(next);
    }

    // Leaf cells (DCTree::evalLeaf, dc_tree.inl:131-552).
    const size_t nl = frontier.size();
    if (nl) {
        std::vector<float> pts(12 * nl);
        std::vector<uint32_t> ptape(4 * nl);
        for (size_t k = 0; k < nl; ++k) {
            const Cell& cl = q.cells[frontier[k]];
            for (int i = 0; i < 4; ++i) {
                ptape[4 * k + i] = cl.tape_in;
                pts[3 * (4 * k + i)] = float((i & 1) ? cl.hi[0] : cl.lo[0]);
                pts[3 * (4 * k + i) + 1] = float((i & 2) ? cl.hi[1] : cl.lo[1]);
                pts[3 * (4 * k + i) + 2] = z;
            }
        }
        std::vector<float> v, g;
        std::vector<uint8_t> amb;
        dev.values(pts, ptape, v, amb);
        std::vector<uint8_t> state(4 * nl, EMPTY);
        std::vector<size_t> zeros, ties;
        for (size_t i = 0; i < 4 * nl; ++i) {
            if (v[i] > 0 || std::isnan(v[i])) state[i] = EMPTY;
            else if (v[i] < 0) state[i] = FILLED;
            else if (!amb[i]) zeros.push_back(i);
            else ties.push_back(i);
        }
        auto gather = [&](const std::vector<size_t>& idx) {
            std::vector<float> p(3 * idx.size());
            for (size_t j = 0; j < idx.size(); ++j) {
                std::memcpy(&p[3 * j], &pts[3 * idx[j]], 12);
            }
            return p;
        };
        auto gather_tapes = [&](const std::vector<size_t>& idx) {
            std::vector<uint32_t> t(idx.size());
            for (size_t j = 0; j < idx.size(); ++j) t[j] = ptape[idx[j]];
            return t;
        };
        if (!zeros.empty()) {
            std::vector<float> zv;
            std::vector<uint8_t> za;
            dev.grads(gather(zeros), gather_tapes(zeros), zv, g, za);
            for (size_t j = 0; j < zeros.size(); ++j) {
                const bool nz = g[3 * j] != 0 || g[3 * j + 1] != 0 || g[3 * j + 2] != 0;
                state[zeros[j]] = nz ? FILLED : EMPTY;
            }
        }
        if (!ties.empty()) {
            std::vector<uint8_t> ins;
            std::vector<float> fs;
            std::vector<int32_t> nf;
            dev.features(gather(ties), gather_tapes(ties), true, ins, fs, nf);
            for (size_t j = 0; j < ties.size(); ++j) state[ties[j]] = ins[j] ? FILLED : EMPTY;
        }

        // Mixed cells become leaves; their sign-change edges are searched
        // together.
        std::vector<Edge> edges;
        edges.reserve(2 * nl);
        q.leaves.reserve(q.leaves.size() + nl);
        out.cells.reserve(out.cells.size() + 4 * nl);
        for (size_t k = 0; k < nl; ++k) {
            Cell& cl = q.cells[frontier[k]];
            uint8_t mask = 0;
            for (int i = 0; i < 4; ++i) mask |= uint8_t((state[4 * k + i] == FILLED) << i);
            cl.type = mask == 0 ? EMPTY : mask == 15 ? FILLED : AMBIGUOUS;
            out.cells.insert(out.cells.end(), {cl.key[0], cl.key[1], 0u, uint32_t(cl.type)});
            if (cl.type != AMBIGUOUS) continue;
            cl.leaf = int32_t(q.leaves.size());
            Leaf lf;
            lf.mask = mask;
            lf.manifold = mask != 6 && mask != 9;     // dc_tree2.cpp:17-22
            q.leaves.push_back(lf);
            for (int p = 0; p < 2 && M.v[mask][p][0][0] >= 0; ++p) {
                for (int e = 0; e < 4 && M.v[mask][p][e][0] >= 0; ++e) {
                    const int a = M.v[mask][p][e][0], b = M.v[mask][p][e][1];
                    Edge ed;
                    ed.leaf = cl.leaf;
                    ed.patch = p;
                    ed.id = M.e[a][b];
                    ed.tape = cl.tape_in;
                    ed.a[0] = (a & 1) ? cl.hi[0] : cl.lo[0];
                    ed.a[1] = (a & 2) ? cl.hi[1] : cl.lo[1];
                    ed.b[0] = (b & 1) ? cl.hi[0] : cl.lo[0];
                    ed.b[1] = (b & 2) ? cl.hi[1] : cl.lo[1];
                    edges.push_back(ed);
                }
            }
        }

        // Edge search: 4 rounds of 16 samples per edge (dc_tree.inl:339-410),
        // on the device; an edge whose round meets an exact-zero sample
        // first is finished here from that round, with isInside on the
        // device.
        const size_t ne = edges.size();
        std::vector<uint8_t> stop(ne, 4);
        if (ne) {
            std::vector<double> ends(4 * ne);
            std::vector<uint32_t> etape(ne);
            for (size_t e = 0; e < ne; ++e) {
                etape[e] = edges[e].tape;
                ends[4 * e] = edges[e].a[0];
                ends[4 * e + 1] = edges[e].a[1];
                ends[4 * e + 2] = edges[e].b[0];
                ends[4 * e + 3] = edges[e].b[1];
            }
            PoolBuf<double> de(4 * ne);
            PoolBuf<uint8_t> dst(ne);
            PoolBuf<uint32_t> dt(ne);
            hip_check(hipMemcpyAsync(de.p, ends.data(), 32 * ne, hipMemcpyHostToDevice, c.stream),
                      "copy edges");
            hip_check(hipMemcpyAsync(dt.p, etape.data(), 4 * ne, hipMemcpyHostToDevice, c.stream),
                      "copy edge tapes");
            hip_check(launch_edge_search(dev.tapes(), dt.p, z, de.p, dst.p, ne, c.stream),
                      "edge search");
            hip_check(hipMemcpyAsync(ends.data(), de.p, 32 * ne, hipMemcpyDeviceToHost, c.stream),
                      "copy brackets");
            hip_check(hipMemcpyAsync(stop.data(), dst.p, ne, hipMemcpyDeviceToHost, c.stream),
                      "copy stops");
            hip_check(hipStreamSynchronize(c.stream), "edge search");
            for (size_t e = 0; e < ne; ++e) {
                edges[e].a[0] = ends[4 * e];
                edges[e].a[1] = ends[4 * e + 1];
                edges[e].b[0] = ends[4 * e + 2];
                edges[e].b[1] = ends[4 * e + 3];
            }
        }
        for (int round = 0; round < 4; ++round) {
            std::vector<size_t> act;
            for (size_t e = 0; e < ne; ++e) {
                if (stop[e] <= round) act.push_back(e);
            }
            if (act.empty()) continue;
            const size_t na = act.size();
            std::vector<double> sp(2 * 16 * na);
            std::vector<float> sx(3 * 16 * na);
            std::vector<uint32_t> st(16 * na);
            for (size_t k = 0; k < na; ++k) {
                const Edge& ed = edges[act[k]];
                for (int j = 0; j < 16; ++j) {
                    const double f = j / 15.0;
                    const size_t i = 16 * k + j;
                    st[i] = ed.tape;
                    for (int a = 0; a < 2; ++a) {
                        sp[2 * i + a] = ed.a[a] * (1 - f) + ed.b[a] * f;
                        sx[3 * i + a] = float(sp[2 * i + a]);
                    }
                    sx[3 * i + 2] = z;
                }
            }
            std::vector<float> sv;
            std::vector<uint8_t> sa;
            dev.values(sx, st, sv, sa);
            // isInside for the samples that are exactly zero
            std::vector<size_t> zs;
            for (size_t i = 0; i < 16 * na; ++i) {
                if ((i & 15) && sv[i] == 0) zs.push_back(i);
            }
            std::vector<uint8_t> zin;
            if (!zs.empty()) {
                std::vector<float> zp(3 * zs.size());
                std::vector<uint32_t> zt(zs.size());
                for (size_t j = 0; j < zs.size(); ++j) {
                    std::memcpy(&zp[3 * j], &sx[3 * zs[j]], 12);
                    zt[j] = st[zs[j]];
                }
                std::vector<float> fs;
                std::vector<int32_t> nf;
                dev.features(zp, zt, true, zin, fs, nf);
            }
            // First sample after the inside end that is outside (v > 0, or
            // v == 0 and not inside); else the last one, NaN included.
            size_t zk = 0;
            for (size_t k = 0; k < na; ++k) {
                Edge& ed = edges[act[k]];
                for (int j = 1; j < 16; ++j) {
                    const size_t i = 16 * k + j;
                    bool take;
                    if (sv[i] > 0) {
                        take = true;
                    } else if (sv[i] == 0) {
                        while (zs[zk] != i) ++zk;
                        take = !zin[zk];
                    } else {
                        take = j == 15;
                    }
                    if (take) {
                        for (int a = 0; a < 2; ++a) {
                            ed.a[a] = sp[2 * (i - 1) + a];
                            ed.b[a] = sp[2 * i + a];
                        }
                        break;
                    }
                }
            }
        }

        // Value and partials at both bracket ends (dc_tree.inl:412-471).
        std::vector<float> bx(6 * ne);
        std::vector<uint32_t> bt(2 * ne);
        for (size_t e = 0; e < ne; ++e) {
            bt[2 * e] = bt[2 * e + 1] = edges[e].tape;
            for (int a = 0; a < 2; ++a) {
                bx[6 * e + a] = float(edges[e].a[a]);
                bx[6 * e + 3 + a] = float(edges[e].b[a]);
            }
            bx[6 * e + 2] = z;
            bx[6 * e + 5] = z;
        }
        std::vector<float> bv, bg;
        std::vector<uint8_t> ba;
        dev.grads(bx, bt, bv, bg, ba);
        std::vector<size_t> amb_ends;
        for (size_t i = 0; i < 2 * ne; ++i) if (ba[i]) amb_ends.push_back(i);
        std::vector<uint8_t> fin;
        std::vector<float> fs;
        std::vector<int32_t> nf;
        if (!amb_ends.empty()) {
            std::vector<float> fp(3 * amb_ends.size());
            std::vector<uint32_t> ft(amb_ends.size());
            for (size_t j = 0; j < amb_ends.size(); ++j) {
                std::memcpy(&fp[3 * j], &bx[3 * amb_ends[j]], 12);
                ft[j] = bt[amb_ends[j]];
            }
            dev.features(fp, ft, false, fin, fs, nf);
        }

        // QEF per patch (dc_tree.inl:412-552): edges of one leaf and patch
        // are consecutive in `edges`; leaves are independent, so ranges of
        // whole leaves go to host threads.
        auto solve = [&](size_t ebegin, size_t eend) {
            size_t fk = size_t(std::lower_bound(amb_ends.begin(), amb_ends.end(), 2 * ebegin) -
                               amb_ends.begin());
            for (size_t e0 = ebegin; e0 < eend;) {
                size_t e1 = e0;
                while (e1 < ne && edges[e1].leaf == edges[e0].leaf &&
                       edges[e1].patch == edges[e0].patch) {
                    ++e1;
                }
                Leaf& lf = q.leaves[edges[e0].leaf];
                const int ec = int(e1 - e0);
                double acc[4][14] = {};
                bool used[4] = {false, false, false, false};
                for (int i = 0; i < 2 * ec; ++i) {
                    const size_t pi = 2 * e0 + i;          // bracket end index
                    const Edge& ed = edges[e0 + i / 2];
                    const double* p2 = (i & 1) ? ed.b : ed.a;
                    const double p[3] = {p2[0], p2[1], 0.0};
                    const double val = bv[pi];
                    if (!ba[pi]) {
                        const double d[3] = {bg[3 * pi], bg[3 * pi + 1], 0.0};
                        qef::push_sample(acc[i / 2], p, d, val);
                        used[i / 2] = true;
                    } else {
                        while (amb_ends[fk] != pi) ++fk;
                        for (int j = 0; j < nf[fk]; ++j) {
                            const double d[3] = {fs[(fk * 16 + j) * 3], fs[(fk * 16 + j) * 3 + 1], 0.0};
                            qef::push_sample(acc[i / 2], p, d, val);
                            used[i / 2] = true;
                        }
                    }
                }
                int max_rank = 0, ranks[4];
                for (int i = 0; i < ec; ++i) {
                    ranks[i] = -1;
                    if (!used[i]) continue;
                    double ata[3][3];
                    qef::unpack_ata(acc[i], ata);
                    ranks[i] = qef::small_eigenvalue_count(ata);
                    max_rank = std::max(max_rank, ranks[i]);
                }
                double mass[4] = {0, 0, 0, 0};
                for (int i = 0; i < ec; ++i) {
                    if (ranks[i] == max_rank) {
                        mass[0] = mass[0] + acc[i][0] / acc[i][3];
                        mass[1] = mass[1] + acc[i][1] / acc[i][3];
                        mass[2] = mass[2] + acc[i][2] / acc[i][3];
                        mass[3] = mass[3] + 1.0;
                    }
                }
                double sum[14] = {};
                for (int i = 0; i < ec; ++i) {
                    if (!used[i]) continue;
                    for (int k = 4; k < 14; ++k) sum[k] = sum[k] + acc[i][k];
                }
                std::memcpy(lf.mass, mass, sizeof(mass));
                qef::unpack_ata(sum, lf.ata);
                lf.atb[0] = sum[10];
                lf.atb[1] = sum[11];
                lf.atb[2] = sum[12];
                lf.btb = sum[13];
                int rank = 0;
                qef::find_vertex(lf.ata, lf.atb, lf.btb, mass, lf.vert[lf.count], &rank);
                lf.rank = rank;
                lf.count++;
                e0 = e1;
            }
        };
        const size_t parts = std::min<size_t>(16, std::max<size_t>(1, ne / 4096));
        std::vector<size_t> cut{0};
        for (size_t k = 1; k < parts; ++k) {
            size_t e = std::max(cut.back(), ne * k / parts);
            while (e < ne && e > 0 && edges[e].leaf == edges[e - 1].leaf) ++e;
            cut.push_back(e);
        }
        cut.push_back(ne);
        std::vector<std::thread> pool;
        for (size_t k = 1; k + 1 < cut.size(); ++k) {
            pool.emplace_back(solve, cut[k], cut[k + 1]);
        }
        solve(cut[0], cut[1]);
        for (auto& t : pool) t.join();
    }

    // Collapse, bottom up (DCTree::collectChildren, dc_tree.inl:581-735);
    // the vertex values of each level's candidates in one launch, each
    // with the tape Tape::getBase picks for the vertex (dc_tree.inl:716)
    // starting from the cell's pushed tape narrowed to the cell.
    tapes.refresh_records();
    for (int L = 1; L <= top; ++L) {
        struct Cand { int32_t cell; Leaf leaf; };
        std::vector<Cand> cands;
        for (int32_t ci : branches[L]) {
            Cell& cl = q.cells[ci];
            const Cell* cs[4];
            bool any_branch = false, all_empty = true, all_full = true;
            uint8_t corners[4];
            for (int i = 0; i < 4; ++i) {
                cs[i] = &q.cells[cl.kids + i];
                any_branch |= q.branch(*cs[i]);
            }
            if (any_branch) continue;
            for (int i = 0; i < 4; ++i) {
                corners[i] = q.corner(*cs[i], i);
                all_empty &= cs[i]->type == EMPTY;
                all_full &= cs[i]->type == FILLED;
            }
            if (all_empty || all_full) {
                cl.type = all_empty ? EMPTY : FILLED;
                cl.kids = -1;
                continue;
            }
            uint8_t mask = 0;
            for (int i = 0; i < 4; ++i) mask |= uint8_t((corners[i] == FILLED) << i);
            bool ok = mask != 6 && mask != 9;
            for (int i = 0; i < 4 && ok; ++i) {
                ok &= cs[i]->type != AMBIGUOUS || q.leaves[cs[i]->leaf].manifold;
            }
            if (!ok || !children_manifold(q, cs, corners)) continue;
            Leaf lf;
            lf.manifold = true;
            lf.mask = mask;
            for (int i = 0; i < 4; ++i) {
                if (cs[i]->type == AMBIGUOUS) lf.rank = std::max(lf.rank, q.leaves[cs[i]->leaf].rank);
            }
            for (int i = 0; i < 4; ++i) {
                if (cs[i]->type != AMBIGUOUS) continue;
                const Leaf& k = q.leaves[cs[i]->leaf];
                if (k.rank == lf.rank) {
                    for (int j = 0; j < 4; ++j) lf.mass[j] = lf.mass[j] + k.mass[j];
                }
                for (int a = 0; a < 3; ++a) {
                    for (int b = 0; b < 3; ++b) lf.ata[a][b] = lf.ata[a][b] + k.ata[a][b];
                    lf.atb[a] = lf.atb[a] + k.atb[a];
                }
                lf.btb = lf.btb + k.btb;
            }
            const double err = qef::find_vertex(lf.ata, lf.atb, lf.btb, lf.mass, lf.vert[0],
                                                nullptr);
            lf.count = 1;
            const double* vx = lf.vert[0];
            const double eps = 1e-6;
            const bool inside = vx[0] >= cl.lo[0] - eps && vx[1] >= cl.lo[1] - eps &&
                                vx[0] <= cl.hi[0] + eps && vx[1] <= cl.hi[1] + eps;
            if (err < max_err && inside) cands.push_back({ci, lf});
        }
        if (cands.empty()) continue;
        std::vector<float> cp(3 * cands.size());
        std::vector<uint32_t> ct(cands.size());
        for (size_t j = 0; j < cands.size(); ++j) {
            cp[3 * j] = float(cands[j].leaf.vert[0][0]);
            cp[3 * j + 1] = float(cands[j].leaf.vert[0][1]);
            cp[3 * j + 2] = z;
            const Cell& cl = q.cells[cands[j].cell];
            const double clo[3] = {cl.lo[0], cl.lo[1], double(z)};
            const double chi[3] = {cl.hi[0], cl.hi[1], double(z)};
            const double vd[3] = {cp[3 * j], cp[3 * j + 1], cp[3 * j + 2]};
            ct[j] = tapes.base_for(tapes.base_for(cl.tape_out, clo, chi), vd, vd);
        }
        std::vector<float> cv;
        std::vector<uint8_t> ca;
        dev.values(cp, ct, cv, ca);
        for (size_t j = 0; j < cands.size(); ++j) {
            if (!(std::fabs(cv[j]) < max_err)) continue;
            Cell& cl = q.cells[cands[j].cell];
            cands[j].leaf.level = L;
            cl.type = AMBIGUOUS;
            cl.leaf = int32_t(q.leaves.size());
            cl.kids = -1;
            q.leaves.push_back(cands[j].leaf);
        }
    }

    // Dual walk (dual.hpp:82-122) and segments (dc_contourer.cpp:16-98).
    out.verts.push_back({0.0f, 0.0f});
    out.keys.push_back({0, 0, 0, 0});
    auto vertex = [&](const Cell& t, int vi) {
        Leaf& lf = q.leaves[t.leaf];
        if (lf.index[vi] == 0) {
            lf.index[vi] = uint32_t(out.verts.size());
            out.verts.push_back({float(lf.vert[vi][0]), float(lf.vert[vi][1])});
            out.keys.push_back({t.key[0], t.key[1], uint32_t(t.level), uint32_t(vi)});
        }
        return lf.index[vi];
    };
    auto load = [&](int A, const Cell& t0, const Cell& t1) {
        const Cell* ts[2] = {&t0, &t1};
        const int idx = q.leaves[t1.leaf].level < q.leaves[t0.leaf].level ? 1 : 0;
        const int base_corner = idx == 0 ? (3 ^ A) : 0;
        const uint8_t a = q.corner(*ts[idx], base_corner);
        const uint8_t b = q.corner(*ts[idx], base_corner | A);
        if (a == b) return;
        const int D = ((a == FILLED && A == 2) || (b == FILLED && A == 1)) ? 0 : 1;
        int es[2];
        if (D ^ (A == 1)) {
            es[0] = M.e[3][3 ^ A];
            es[1] = M.e[A][0];
        } else {
            es[0] = M.e[3 ^ A][3];
            es[1] = M.e[0][A];
        }
        uint32_t vs[2];
        for (int i = 0; i < 2; ++i) {
            const Leaf& lf = q.leaves[ts[i]->leaf];
            const int vi = lf.level > 0 ? 0 : M.p[lf.mask][es[i]];
            vs[i] = vertex(*ts[i], vi);
        }
        out.segs.push_back({vs[1 - D], vs[D]});
    };
    std::function<void(int, const Cell&, const Cell&)> edge2 =
        [&](int A, const Cell& t0, const Cell& t1) {
            const int perp = 3 ^ A;
            if (q.branch(t0) || q.branch(t1)) {
                edge2(A, q.child(t0, perp), q.child(t1, 0));
                edge2(A, q.child(t0, A | perp), q.child(t1, A));
            } else if (t0.type == AMBIGUOUS && t1.type == AMBIGUOUS) {
                load(A, t0, t1);
            }
        };
    std::function<void(const Cell&)> walk = [&](const Cell& t) {
        if (!q.branch(t)) return;
        for (int i = 0; i < 4; ++i) walk(q.cells[t.kids + i]);
        edge2(2, q.child(t, 0), q.child(t, 1));
        edge2(2, q.child(t, 2), q.child(t, 3));
        edge2(1, q.child(t, 0), q.child(t, 2));
        edge2(1, q.child(t, 1), q.child(t, 3));
    };
    walk(q.cells[0]);

    // Contours::collect (contours.cpp:112-184): join segments head to tail.
    // The reference's head and tail maps, keyed by vertex index, as arrays
    // over the vertices (kNone: no entry); a segment joins two different
    // cells' vertices, so its ends differ.
    constexpr uint32_t kNone = 0xffffffffu;
    std::vector<uint32_t> heads(out.verts.size(), kNone), tails(out.verts.size(), kNone);
    std::vector<std::list<uint32_t>> runs;
    for (const auto& s : out.segs) {
        if (tails[s[0]] != kNone) {
            const uint32_t r = tails[s[0]];
            runs[r].push_back(s[1]);
            if (tails[s[1]] == kNone) tails[s[1]] = r;     // insert keeps an entry
            tails[s[0]] = kNone;
            continue;
        }
        if (heads[s[1]] != kNone) {
            const uint32_t r = heads[s[1]];
            runs[r].push_front(s[0]);
            if (heads[s[0]] == kNone) heads[s[0]] = r;
            heads[s[1]] = kNone;
            continue;
        }
        heads[s[0]] = uint32_t(runs.size());
        tails[s[1]] = uint32_t(runs.size());
        runs.push_back({s[0], s[1]});
    }
    std::vector<bool> used(runs.size(), false);
    for (size_t i = 0; i < runs.size(); ++i) {
        if (used[i]) continue;
        out.contours.emplace_back();
        for (size_t r = i;;) {
            auto& dst = out.contours.back();
            dst.insert(dst.end(), runs[r].begin(), runs[r].end());
            used[r] = true;
            const uint32_t h = heads[runs[r].back()];
            if (h == kNone || used[h]) break;
            dst.pop_back();
            r = h;
        }
    }
    return out;
}

}   // namespace amd
}   // namespace libfive
