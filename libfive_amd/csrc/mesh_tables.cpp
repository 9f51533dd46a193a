// Mesher lookup tables (see mesh_tables.hpp).
#include "mesh_tables.hpp"

#include <algorithm>
#include <cstring>
#include <set>
#include <utility>
#include <vector>

namespace libfive {
namespace meshing {

namespace {

// Corner i of the unit cube has coordinates (i & 1, i & 2, i & 4) along
// (X, Y, Z).  A quarter turn about an axis maps corners to corners.
int turn_corner(int corner, int axis) {
    int s[3];
    for (int k = 0; k < 3; ++k) s[k] = (corner >> k) & 1 ? 1 : -1;
    int t[3];
    switch (axis) {
        case 0: t[0] = s[0]; t[1] = -s[2]; t[2] = s[1]; break;   // about X
        case 1: t[0] = s[2]; t[1] = s[1]; t[2] = -s[0]; break;   // about Y
        default: t[0] = -s[1]; t[1] = s[0]; t[2] = s[2]; break;  // about Z
    }
    return (t[0] > 0) | ((t[1] > 0) << 1) | ((t[2] > 0) << 2);
}

struct Seed {
    std::vector<int> inside;
    std::vector<std::vector<std::pair<int, int>>> patches;
};

// The distinct sign configurations of the cube up to rotation, each with
// its vertex patches as (inside, outside) corner pairs in the order the
// reference lists them (marching.cpp:115-400).
const std::vector<Seed>& seeds() {
    static const std::vector<Seed> s = {
        {{0}, {{{0, 1}, {0, 2}, {0, 4}}}},
        {{0, 1}, {{{1, 3}, {0, 2}, {0, 4}, {1, 5}}}},
        {{0, 5}, {{{0, 1}, {0, 2}, {0, 4}}, {{5, 7}, {5, 4}, {5, 1}}}},
        {{0, 7}, {{{0, 1}, {0, 2}, {0, 4}}, {{7, 6}, {7, 3}, {7, 5}}}},
        {{1, 2, 3}, {{{1, 0}, {1, 5}, {3, 7}, {2, 6}, {2, 0}}}},
        {{0, 1, 7}, {{{1, 3}, {0, 2}, {0, 4}, {1, 5}}, {{7, 6}, {7, 3}, {7, 5}}}},
        {{1, 4, 7}, {{{1, 3}, {1, 0}, {1, 5}}, {{4, 6}, {4, 5}, {4, 0}},
                     {{7, 6}, {7, 3}, {7, 5}}}},
        {{0, 1, 2, 3}, {{{0, 4}, {1, 5}, {3, 7}, {2, 6}}}},
        {{0, 2, 3, 6}, {{{0, 1}, {3, 1}, {3, 7}, {6, 7}, {6, 4}, {0, 4}}}},
        {{0, 4, 3, 7}, {{{0, 1}, {0, 2}, {4, 6}, {4, 5}},
                        {{3, 2}, {3, 1}, {7, 5}, {7, 6}}}},
        {{0, 2, 3, 7}, {{{0, 1}, {3, 1}, {7, 5}, {7, 6}, {2, 6}, {0, 4}}}},
        {{1, 2, 3, 4}, {{{4, 6}, {4, 5}, {4, 0}},
                        {{1, 0}, {1, 5}, {3, 7}, {2, 6}, {2, 0}}}},
        {{0, 3, 6, 5}, {{{0, 2}, {0, 4}, {0, 1}}, {{3, 7}, {3, 2}, {3, 1}},
                        {{6, 4}, {6, 2}, {6, 7}}, {{5, 7}, {5, 1}, {5, 4}}}},
        {{1, 3, 2, 6}, {{{1, 5}, {3, 7}, {6, 7}, {6, 4}, {2, 0}, {1, 0}}}},
        {{0, 2, 3, 5, 6}, {{{0, 1}, {3, 1}, {3, 7}, {6, 7}, {6, 4}, {0, 4}},
                           {{5, 7}, {5, 1}, {5, 4}}}},
        {{2, 3, 4, 5, 6}, {{{3, 1}, {3, 7}, {6, 7}, {5, 7}, {5, 1}, {4, 0},
                            {2, 0}}}},
        {{0, 4, 5, 6, 7}, {{{0, 2}, {6, 2}, {7, 3}, {5, 1}, {0, 1}}}},
        {{1, 2, 3, 4, 5, 6}, {{{1, 0}, {4, 0}, {2, 0}}, {{6, 7}, {5, 7}, {3, 7}}}},
        {{1, 2, 3, 4, 6, 7}, {{{1, 0}, {1, 5}, {7, 5}, {4, 5}, {4, 0}, {2, 0}}}},
        {{2, 3, 4, 5, 6, 7}, {{{4, 0}, {5, 1}, {3, 1}, {2, 0}}}},
        {{1, 2, 3, 4, 5, 6, 7}, {{{1, 0}, {4, 0}, {2, 0}}}},
    };
    return s;
}

// Manifold test of a sign pattern: contract cube edges whose ends have
// equal signs; the pattern is manifold when at most one distinct edge
// remains (the generator printed in dc_tree3.cpp:114-142).
bool manifold_pattern(int mask) {
    // The twelve cube edges, in the order the generator walks them.
    static const int order[12][2] = {{0, 1}, {0, 2}, {2, 3}, {1, 3},
                                     {4, 5}, {4, 6}, {6, 7}, {5, 7},
                                     {0, 4}, {2, 6}, {1, 5}, {3, 7}};
    std::vector<std::pair<int, int>> edges;
    for (const auto& o : order) edges.push_back({o[0], o[1]});
    auto sign = [&](int c) { return (mask >> c) & 1; };
    for (bool merged = true; merged;) {
        merged = false;
        for (size_t i = 0; i < edges.size(); ++i) {
            const int a = edges[i].first, b = edges[i].second;
            if (sign(a) != sign(b)) continue;
            std::vector<std::pair<int, int>> kept;
            for (auto f : edges) {
                if (f.first == a) f.first = b;
                if (f.second == a) f.second = b;
                if (f.first != f.second) kept.push_back(f);
            }
            edges.swap(kept);
            merged = true;
            break;
        }
    }
    std::set<std::pair<int, int>> distinct;
    for (const auto& f : edges) {
        distinct.insert({std::min(f.first, f.second), std::max(f.first, f.second)});
    }
    return distinct.size() <= 1;
}

Tables build() {
    Tables t;
    std::memset(t.v, -1, sizeof(t.v));
    std::memset(t.e, -1, sizeof(t.e));
    std::memset(t.p, -1, sizeof(t.p));
    bool pending[256] = {};
    for (const auto& s : seeds()) {
        int mask = 0;
        for (int c : s.inside) mask |= 1 << c;
        for (size_t p = 0; p < s.patches.size(); ++p) {
            for (size_t k = 0; k < s.patches[p].size(); ++k) {
                t.v[mask][p][k][0] = int8_t(s.patches[p][k].first);
                t.v[mask][p][k][1] = int8_t(s.patches[p][k].second);
            }
        }
        pending[mask] = true;
    }
    // Fill every rotated configuration from the first one that reaches it,
    // sweeping masks in increasing order until nothing new appears.
    for (bool again = true; again;) {
        again = false;
        for (int m = 0; m < 256; ++m) {
            if (!pending[m]) continue;
            pending[m] = false;
            for (int axis = 0; axis < 3; ++axis) {
                int r = 0;
                for (int c = 0; c < 8; ++c) {
                    if (m & (1 << c)) r |= 1 << turn_corner(c, axis);
                }
                if (t.v[r][0][0][0] != -1) continue;
                pending[r] = true;
                again = true;
                for (int p = 0; p < 4 && t.v[m][p][0][0] != -1; ++p) {
                    for (int k = 0; k < 12 && t.v[m][p][k][0] != -1; ++k) {
                        t.v[r][p][k][0] = int8_t(turn_corner(t.v[m][p][k][0], axis));
                        t.v[r][p][k][1] = int8_t(turn_corner(t.v[m][p][k][1], axis));
                    }
                }
            }
        }
    }
    int next_id = 0;
    for (int m = 0; m < 256; ++m) {
        for (int p = 0; p < 4 && t.v[m][p][0][0] != -1; ++p) {
            for (int k = 0; k < 12 && t.v[m][p][k][0] != -1; ++k) {
                int8_t& id = t.e[t.v[m][p][k][0]][t.v[m][p][k][1]];
                if (id == -1) id = int8_t(next_id++);
                t.p[m][id] = int8_t(p);
            }
        }
    }
    for (int m = 0; m < 256; ++m) t.manifold[m] = manifold_pattern(m) ? 1 : 0;
    return t;
}

}   // namespace

const Tables& tables() {
    static const Tables t = build();
    return t;
}

}   // namespace meshing
}   // namespace libfive
