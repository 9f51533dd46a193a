// 2D dual contouring of a Z slice (libfive_tree_render_slice / _slice3,
// libfive/src/libfive.cpp:233-293 -> Contours::render,
// libfive/src/render/brep/contours.cpp:25-49).
//
// The quadtree is built level by level: every level's cells are classified
// by one interval launch; the level-0 cells' corner values, the edge
// search's sample values, the bracket ends' partials and the rare
// exact-zero feature evaluations are each one GPU launch over all cells of
// the pass.  The per-cell decisions, QEF solves (FP64), collapse, dual walk
// and segment chaining run on the host between those launches.
#pragma once

#include <array>
#include <cstdint>
#include <vector>

#include "tree.hpp"

namespace libfive {
namespace amd {

struct ContourResult {
    std::vector<std::array<float, 2>> verts;        // [0] is the dummy vertex
    std::vector<std::array<uint32_t, 4>> keys;      // x, y, level, patch
    std::vector<std::array<uint32_t, 2>> segs;
    std::vector<std::vector<uint32_t>> contours;    // chained vertex indices
    std::vector<uint32_t> cells;                    // x, y, level, state
};

ContourResult render_contours(const Tree& t, const double lo[2], const double hi[2],
                              float z, double min_feature, double max_err);

}   // namespace amd
}   // namespace libfive
