// Dual-contouring mesher entry (Mesh::render, libfive/src/render/brep/
// mesh.cpp:35-125, DC branch 53-72): builds the octree, collapses it and
// emits triangles on the GPU, then copies the mesh to the host.
#pragma once

#include <atomic>
#include <cstdint>
#include <functional>
#include <map>
#include <vector>

#include <hip/hip_runtime.h>

#include "tree.hpp"

namespace libfive {
namespace amd {

struct MeshSettings {
    double min_feature = 0.1;     // settings.hpp:34-43 defaults
    double max_err = 1e-8;
    // QEF accumulation: 0 in double (the reference's arithmetic), 1 as
    // compensated float pairs (qef::push_sample_f2); the eigensolve is in
    // double either way.
    uint32_t precision = 0;
    bool collect_stats = false;   // per-level cell counts and tape lengths
    volatile const int* cancel = nullptr;           // C API flag
    const std::atomic<bool>* cancel_flag = nullptr; // BRepSettings::cancel
    bool cancelled() const {
        return (cancel && *cancel) || (cancel_flag && cancel_flag->load());
    }
    // Values of free variables by tree id (libfive_tree_id); a variable
    // not listed is 0.  The tape is compiled with these values.
    const std::map<const void*, float>* vars = nullptr;
};

// Flat output: verts as xyz floats with the dummy vertex 0 first, tris as
// index triples, keys as 5 words per vertex (cell x, y, z in finest-cell
// units, cell level, vertex patch).
struct HostMesh {
    std::vector<float> verts;
    std::vector<uint32_t> tris;
    std::vector<uint32_t> keys;
    // When set, called once with the vertex count (dummy included) and the
    // triangle count; it returns the host buffers (3 floats per vertex,
    // 3 indices per triangle) the mesh is copied into, and verts / tris
    // stay empty.  It returns true when both buffers are page-locked host
    // memory, which the device then writes directly.
    std::function<bool(size_t, size_t, float**, uint32_t**)> alloc;
    bool want_keys = true;          // fill keys
};

// Statistics of the last render on this thread.
struct MeshStats {
    int top = 0;                                  // root level
    std::vector<uint64_t> level_counts;           // root first: empty, filled, ambiguous
    std::vector<uint32_t> tape_lengths;           // per ambiguous cell above level 0
    std::vector<uint32_t> cells;                  // x, y, z, level, state per cell
    uint64_t leaves = 0;                          // ambiguous leaf cells
    uint64_t collapsed = 0;
    double ms_build = 0, ms_collapse = 0, ms_emit = 0, ms_total = 0;
};

// Returns false when cancelled (the C API then returns NULL): the flag is
// polled between the octree levels of the build, after the leaf pass,
// between the collapse levels and between the emit steps, so a render ends
// within one level's kernels of the request.  Runs on the library stream,
// one render at a time.
bool render_mesh(const Tree& t, const double lo[3], const double hi[3],
                 const MeshSettings& s, HostMesh& out);

// The same on the caller's stream, without the library lock: renders on
// different streams (from different host threads) run concurrently.
bool render_mesh_on(const Tree& t, const double lo[3], const double hi[3],
                    const MeshSettings& s, HostMesh& out, hipStream_t stream);

// One model of a merged batch: a tree meshed over [lo, hi] with
// min_feature (default max_err, FP64 QEF sums).
struct BatchItem {
    const Tree* tree;
    double lo[3], hi[3];
    double min_feature;
};

// Meshes many models through one set of level-synchronous passes: every
// pass (interval and pruning, leaf corners, edges and solves, collapse,
// quads, vertex gather) is one launch over the cells of all models, and the
// host waits once per level for all of them instead of once per level per
// model.  A model's cells and tapes carry its index (ModelRec), so each
// model's mesh is the one render_mesh gives for it, bit for bit.  outs[i]
// receives model i (through its alloc callback, or its vectors; keys when
// want_keys).  Models whose root is a leaf (resolution level 0) are meshed
// by render_mesh_on on the same stream.
void render_meshes_merged(const BatchItem* items, size_t n, HostMesh* outs, hipStream_t s);

const MeshStats& last_mesh_stats();

// Region::withResolution (region.hpp:226-231), evaluated in float.
int resolution_level(const double lo[3], const double hi[3], double min_feature);

}   // namespace amd
}   // namespace libfive
