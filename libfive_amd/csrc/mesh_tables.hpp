// Lookup tables of the dual-contouring mesher, built on the host once and
// copied to device constant memory.
//
//   v[mask][patch][k] = (inside corner, outside corner) of the k-th edge of a
//       vertex patch, -1 terminated: Nielson's dual marching cubes cases
//       closed under quarter turns (MarchingTable<3>,
//       libfive/src/render/brep/dc/marching.cpp:115-557)
//   e[a][b]   = directed edge id (0..23) of corner pair a -> b, -1 if none
//   p[mask][edge] = patch of a directed edge for a corner mask, -1 if none
//   manifold[mask] = whether the sign pattern is manifold
//       (DCTree<3>::cornersAreManifold, dc_tree3.cpp:112-154)
#pragma once

#include <cstdint>

namespace libfive {
namespace meshing {

struct Tables {
    int8_t v[256][4][12][2];
    int8_t e[8][8];
    int8_t p[256][24];
    uint8_t manifold[256];
};

const Tables& tables();

}   // namespace meshing
}   // namespace libfive
