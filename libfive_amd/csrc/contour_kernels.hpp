// Host-side launchers for the point passes of the 2D contourer and the
// raster (device pointers, asynchronous on `s`).  One lane per point for
// values and partials; one wave per point for the feature evaluator.
// Points carry their own (interval-pruned) tapes.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#include "mesh_kernels.hpp"
#include "tape.hpp"

namespace libfive {
namespace amd {

// The tapes a pass reads: records of the clause pool (the mesher's layout,
// mesh_kernels.hpp); each point (edge, view) names its tape by record id.
// Every tape has the same root and register count; max_len bounds their
// lengths.
struct PoolTapes {
    const tape::Clause* pool;
    const mesh::TapeRec* recs;
    uint32_t root;
    uint32_t nregs;
    uint32_t max_len;
};

// Values and ambiguity (a min / max with equal operands on the point's
// path, ArrayEvaluator::getAmbiguous), point i with tape tid[i].
hipError_t launch_point_values(const PoolTapes& T, const uint32_t* tid, const float* xyz,
                               float* v, uint8_t* amb, size_t npts, hipStream_t s);

// Values, the three partials (g: 3 per point) and ambiguity.
hipError_t launch_point_grads(const PoolTapes& T, const uint32_t* tid, const float* xyz,
                              float* v, float* g, uint8_t* amb, size_t npts, hipStream_t s);

// Edge search over nedges edges (edge e with tape tid[e]), ends = (inside
// x, y, outside x, y) per edge in double, updated in place; stop[e] = 4
// when the four rounds ran on the device, else the round at which an
// exact-zero sample needs the feature evaluator (the bracket is left as at
// that round's start).
hipError_t launch_edge_search(const PoolTapes& T, const uint32_t* tid, float z, double* ends,
                              uint8_t* stop, size_t nedges, hipStream_t s);

// Raster pixel pass: views = (first x, first y, x count, y count) per view
// of at most 256 pixels, evaluated with tape vtape[v]; offs = each view's
// first output byte; out gets 1 where the value at the pixel's voxel
// centre (px[x], py[y], z) is below zero, in view order, x major.
hipError_t launch_raster_views(const PoolTapes& T, const uint32_t* vtape, const float* px,
                               const float* py, float z, const int32_t* views,
                               const uint32_t* offs, uint8_t* out, size_t nviews,
                               hipStream_t s);

// FeatureEvaluator::isInside (inside[i]) and ::features (feats: up to 16
// gradients of 3 floats per point, nfeat[i] of them); a null inside or
// feats skips that evaluation.  ws holds
// point_feature_workspace(T.max_len, T.nregs) bytes per wave in flight (at
// most kFeatureWaves); err gets a nonzero bit when a point has more
// features than the evaluation holds.
constexpr uint32_t kFeatureWaves = 256;
size_t point_feature_workspace(uint32_t max_len, uint32_t nregs);
hipError_t launch_point_features(const PoolTapes& T, const uint32_t* tid, uint32_t sx,
                                 uint32_t sy, uint32_t sz, const float* xyz, uint8_t* inside,
                                 float* feats, int32_t* nfeat, uint8_t* ws, size_t ws_bytes,
                                 uint32_t* err, size_t npts, hipStream_t s);

}   // namespace amd
}   // namespace libfive
