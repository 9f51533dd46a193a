// Occupancy raster of a Z slice (libfive_tree_render_pixels,
// libfive/src/libfive.cpp:442-465 -> Heightmap::render,
// libfive/src/render/discrete/heightmap.cpp:106-319, on the voxel grid of
// libfive/src/render/discrete/voxels.cpp:16-45).
//
// The reference recurses over views of the voxel grid: a view of more than
// 256 voxels is classified by interval arithmetic (filled: every pixel set;
// empty: none; ambiguous: split in two along its largest axis), a smaller
// one by point evaluation at its pixel centres.  Here the recursion runs
// level by level on the host with each level's interval boxes in one GPU
// launch; the small views are then evaluated in one launch with a
// workgroup per view.
#pragma once

#include <cstdint>
#include <vector>

#include "tree.hpp"

namespace libfive {
namespace amd {

struct PixelImage {
    uint32_t width = 0, height = 0;
    std::vector<uint8_t> filled;         // row-major, height rows of width
};

PixelImage render_pixels(const Tree& t, float xlo, float xhi, float ylo,
                         float yhi, float z, float res);

}   // namespace amd
}   // namespace libfive
