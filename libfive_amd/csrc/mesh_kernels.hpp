// Device data layout and kernel launchers of the GPU dual-contouring mesher.
//
// The octree is stored level by level (level 0 = finest cells).  Each level
// is a set of SoA arrays indexed by cell; the eight children of an
// ambiguous cell are contiguous in the next finer level, in child order
// (bit 0 = upper half along X, bit 1 = Y, bit 2 = Z, as Region::subdivide,
// libfive/include/libfive/render/brep/region.hpp:89-114).  Cell bounds are
// not stored: they are recomputed from the integer key by the same chain of
// double-precision midpoints the reference uses.
//
// Pruned tapes live in one clause pool; a tape record names its clauses,
// its parent tape and the box it was pruned for (Tape, libfive/include/
// libfive/eval/tape.hpp:28-114).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "tape.hpp"

namespace libfive {
namespace amd {
namespace mesh {

enum : uint8_t { EMPTY = 0, FILLED = 1, AMBIGUOUS = 2 };

constexpr uint32_t kNoTape = 0xffffffffu;

struct TapeRec {
    uint32_t off;        // first clause in the pool
    uint32_t len;        // clause count (padding excluded)
    uint32_t parent;     // parent tape id, kNoTape for the base tape
    uint32_t flags;      // TAPE_* bits
    float box[6];        // xlo, xhi, ylo, yhi, zlo, zhi of an interval tape
    uint32_t ref_len;    // clauses that are not COPY / LOADI
    uint32_t model;      // model the tape belongs to (index into Params::models)
};
enum : uint32_t { TAPE_INTERVAL = 1u, TAPE_TERMINAL = 2u };

// Per-cell dual-contouring data (DCLeaf, libfive/include/libfive/render/
// brep/dc/dc_tree.hpp:36-100); the QEF sums are those of the last vertex
// solved.
struct LeafRec {
    double vert[4][3];
    double mass[4];
    double ata[6];       // xx, xy, xz, yy, yz, zz
    double atb[3];
    double btb;
    uint32_t rank;
    uint32_t vcount;
    uint32_t mask;       // corner mask (bit i = corner i inside)
    uint32_t flags;      // bit 0 manifold; bits 8-15 level of the vertex
};

// A level's cells are ordered by model (a batch renders several models
// through the same passes); the global number of cell i of model m is
// cbase[m] + i, which numbers the cells model by model, and level by level
// within a model.
struct LevelView {
    uint32_t* key;       // x, y, z of the lower corner in finest-cell units
    uint8_t* type;
    int32_t* child;      // first child in the finer level, -1 if not a branch
    uint32_t* tape;      // tape of the cell's children (or of its leaf eval)
    LeafRec* leaf;       // valid when the cell is an ambiguous leaf
    uint32_t n;
    uint32_t pad;
    const uint32_t* cbase;   // per model: global cell number of index 0
};

// One model of a render: its region, octree depth and tape layout.  The
// base tape of model m is tape record m, and every pushed tape carries its
// parent's model.
struct ModelRec {
    double lo[3], hi[3];
    int32_t top;         // level of the model's root cell
    uint32_t root_reg;
    uint32_t slot_x, slot_y, slot_z;
    uint32_t root_cell;  // index of the root in the level-`top` arrays
    uint32_t slot0;      // first vertex slot of the model (4 x its first cell)
};

struct Counters {
    uint32_t pool_top;
    uint32_t tape_count;
    uint32_t deferred;   // leaf cells left for the feature pass
    uint32_t error;
    uint32_t max_len;    // longest tape of an ambiguous cell of the last level
    uint32_t edges;      // lattice edges searched by the leaf pass
    uint32_t tasks;      // edge-search tasks (four edges each)
    uint32_t leaves;     // ambiguous leaves solved by the leaf pass
    uint32_t pad;
    // Clauses the pushes of the next level may take: over the ambiguous
    // cells of the last level, 8 x (tape length + kPushExtra).
    unsigned long long push_need;
};
enum : uint32_t {
    ERR_POOL = 1u,       // clause pool exhausted
    ERR_TAPES = 2u,      // tape records exhausted
    ERR_FEATURE = 4u,    // a feature evaluation exceeded its fixed capacity
    ERR_LOCATE = 8u,     // point location inconsistency
    ERR_WIDE = 16u,      // the wide interval pass met an invalid tape or stalled
};

// lo, hi, top, root_reg, slot_*, root_cell and slot0 describe one model;
// a kernel that serves several models takes them from `models` (when not
// null) for the model of the tape it evaluates (model_view).
struct Params {
    double lo[3], hi[3];
    int top;                     // level of the root cell
    uint32_t root_cell;
    uint32_t slot0;
    double max_err;
    tape::Clause* pool;
    uint32_t pool_cap;
    TapeRec* tapes;
    uint32_t tape_cap;
    uint32_t nregs;
    uint32_t root_reg;
    uint32_t slot_x, slot_y, slot_z;
    uint32_t qef_f32c;           // QEF sums as compensated float pairs
    Counters* counters;
    const LevelView* levels;     // device copy, indexed by level
    const ModelRec* models;      // null: one model, described by the fields above
};

hipError_t upload_tables(hipStream_t s);

// A pushed tape takes at most its parent's length plus a root copy and 8
// clauses of padding; the pool is budgeted with the same figure.
constexpr uint32_t kPushExtra = 9;

// The interval passes keep two per-clause work arrays (3 bytes per clause)
// in LDS while they fit in kIntervalLdsMax bytes together with the
// per-register tables; longer tapes use a global workspace of
// interval_workspace_bytes() instead, and are launched over at most
// kIntervalWaves groups at a time (one workspace slot per workgroup).  No
// tape length is refused.
constexpr size_t kIntervalLdsMax = 65536;
constexpr uint32_t kIntervalWaves = 2048;
__host__ __device__ inline size_t interval_ws_stride(uint32_t max_len) {
    return (3u * size_t(max_len) + 255u) & ~size_t(255);
}
size_t interval_lds_bytes(uint32_t nregs, uint32_t max_len);
size_t interval_workspace_bytes(uint32_t nregs, uint32_t max_len, uint32_t ngroups);
// The same for launch_interval_boxes (which never uses the wide kernel).
size_t interval_boxes_workspace_bytes(uint32_t nregs, uint32_t max_len);
// Whether launch_interval runs a level of ngroups parents whose longest
// tape has max_len clauses with k_interval_wide (a workgroup of 8 waves
// per parent working on the tape in dependency order) instead of
// k_interval.
bool interval_uses_wide(uint32_t max_len, uint32_t ngroups);
size_t interval_wide_lds(uint32_t nregs, uint32_t max_len);

// Interval classification and pruning of the cells of one level: one wave
// per parent (8 children), or, when groups is null, the roots of models
// first_root .. first_root + ngroups - 1 (each evaluated with its base tape
// and stored at its ModelRec::root_cell).
// max_len bounds the length of every tape the pass reads; ws holds
// interval_workspace_bytes(p.nregs, max_len, ngroups) bytes (null when that
// is 0).
hipError_t launch_interval(const Params& p, int level, const uint32_t* groups,
                           uint32_t ngroups, const LevelView& parent,
                           const LevelView& cur, uint32_t max_len, uint8_t* ws,
                           size_t ws_bytes, hipStream_t s, uint32_t first_root = 0);

// Interval evaluation of listed boxes (xlo, xhi, ylo, yhi, zlo, zhi per
// box), grouped by the tape they are evaluated with: group g covers boxes
// first .. first + count - 1 (count <= 8) and reads tape `tape`.  Per box:
// bounds (lower, upper), nan (the may-be-NaN flag) and tapes (the tape for
// the box's subdivisions: the tape pushed for it, or the group's tape when
// nothing was pruned).  A box is pushed when its bounds are neither above
// nor below zero, and when its result may be NaN only if push_nan
// (Heightmap::recurse pushes regardless, heightmap.cpp:219-222;
// DCTree::evalInterval keeps the unpruned tape, dc_tree.inl:105-109).
// Counters::max_len gets the longest tape returned.
struct BoxGroup {
    uint32_t tape;
    uint32_t first;
    uint32_t count;
    uint32_t pad;
};
hipError_t launch_interval_boxes(const Params& p, const BoxGroup* groups, uint32_t ngroups,
                                 const float* boxes, float* bounds, uint8_t* nan,
                                 uint32_t* tapes, bool push_nan, uint32_t max_len,
                                 uint8_t* ws, size_t ws_bytes, hipStream_t s);

// Links an evaluated level to the next: child offsets and the list of
// ambiguous cells (8 children each).  base = exclusive scan of flags.
hipError_t launch_flags(const LevelView& cur, uint32_t* flags, hipStream_t s);
hipError_t launch_link(const LevelView& cur, const uint32_t* base,
                       uint32_t* groups, hipStream_t s);

// Leaf cells (level 0): corner states, vertex patches, edge search, QEF.
// groups lists cells of l1 whose per_group (8, or 1 for a root leaf)
// children in l0 are evaluated, one wave per child.  Cells that need the
// exact-zero feature evaluation are listed in `deferred` (counted in
// Counters::deferred) and evaluated by launch_leaf_features, each with
// leaf_feature_workspace() bytes of device workspace.
hipError_t launch_leaf(const Params& p, const uint32_t* groups,
                       uint32_t ngroups, uint32_t per_group,
                       const LevelView& l1, const LevelView& l0,
                       uint32_t* deferred, hipStream_t s);
hipError_t launch_leaf_features(const Params& p, const uint32_t* groups,
                                uint32_t ngroups, uint32_t per_group,
                                const LevelView& l1, const LevelView& l0,
                                const uint32_t* work, uint32_t nwork,
                                uint8_t* ws, size_t ws_bytes, hipStream_t s);
size_t leaf_feature_workspace(uint32_t max_len, uint32_t nregs);

// Phase-split leaf pass over the 8 children of each parent in `groups`
// (the children share the parent's tape):
//   corners  one wave per parent: values at the 27 points of the parent's
//            2x2x2 lattice of children, the children's corner masks and
//            cell records, the set of lattice edges the children's patches
//            cross (a 54-bit mask), the edge-search tasks and the list of
//            ambiguous leaves;
//   edges    one wave per task (4 lattice edges of one parent): the
//            4 x 16-sample bracket search, the partials at both bracket
//            ends and the edge's QEF and rank;
//   solve    one lane per ambiguous leaf: per patch, the mass point and QEF
//            sum over the patch's edges and the FP64 vertex solve.
// Each lattice edge is searched once per parent instead of once per cell
// that uses it; the values are the same, since the children evaluate the
// same tape at the same points.  A parent whose points need the feature
// evaluation (an exact zero on a tied MIN/MAX) is listed whole (its 8
// cells) in `deferred` for launch_leaf_features.
struct LeafGroupRec {
    uint64_t edge_mask;  // lattice edges used by the children's patches
    uint32_t edge_base;  // first of the parent's edges in the edge records
    uint32_t inside;     // bit i: lattice point i is inside
    uint32_t flags;      // bit 0: deferred to the feature pass
    uint32_t pad;        // bit i: child i is an ambiguous leaf
};
struct EdgeRec {
    double acc[14];      // Intersection sums (qef::push_sample layout)
    int32_t rank;        // eigenvalues below the cutoff; -1 if no sample
    uint32_t pad;
};
constexpr uint32_t kLeafTasksPerGroup = 14;   // ceil(54 / 4)

// fast: the tape has only the opcodes the call-free evaluators handle
// (pe_fast_opcode); points they cannot evaluate send the parent to the
// feature pass.
// Corners of groups g0 .. g0 + ngroups - 1 (a launch may cover part of a
// level); the edge kernel runs the tasks it is given.
// cnt: 3 * nall counts (edges, tasks, leaves per group, see k_leaf_corners);
// after their exclusive scans (base, same layout) launch_leaf_scatter fills
// the edge bases, the task list and the leaf list and sets
// Counters::edges, tasks and leaves.
// waves: waves per workgroup (1 or 4; one parent or task per wave), a
// launch schedule picked by the autotuner; tapes of 8 register banks use 1.
hipError_t launch_leaf_corners(const Params& p, bool fast, const uint32_t* groups,
                               uint32_t g0, uint32_t ngroups, uint32_t nall,
                               const LevelView& l1, const LevelView& l0, LeafGroupRec* grec,
                               uint32_t* cnt, uint32_t* deferred, hipStream_t s,
                               int waves = 1);
hipError_t launch_leaf_scatter(const Params& p, uint32_t nall, LeafGroupRec* grec,
                               const uint32_t* cnt, const uint32_t* base, uint64_t* tasks,
                               uint32_t* leaves, hipStream_t s);
hipError_t launch_leaf_edges(const Params& p, bool fast, const uint32_t* groups,
                             const LevelView& l1, LeafGroupRec* grec,
                             const uint64_t* tasks, uint32_t ntasks,
                             EdgeRec* edges, uint32_t* deferred, hipStream_t s,
                             int waves = 1);
// Register banks (1, 2, 4 or 8) of the kernels' register files for a tape
// of nregs registers; 0 if the tape has more than 256.
int register_banks(uint32_t nregs);
hipError_t launch_leaf_solve(const Params& p, const uint32_t* groups,
                             const LevelView& l1, const LevelView& l0,
                             const LeafGroupRec* grec, const uint32_t* leaves,
                             uint32_t nleaves, const EdgeRec* edges,
                             hipStream_t s);

// Bottom-up collapse of one level.
hipError_t launch_collapse(const Params& p, int level, const LevelView& cur,
                           const LevelView& below, hipStream_t s);

// Quad emission over the ambiguous leaves of one level.  Pass 0 counts
// triangles per cell and marks used vertex slots; pass 1 writes them.
hipError_t launch_quads(const Params& p, const LevelView& cur, int pass,
                        uint32_t* tri_count, uint32_t* vslot,
                        const uint32_t* tri_off, const uint32_t* vid,
                        uint32_t* tris, hipStream_t s);
// Vertex positions (and keys: x, y, z, level, patch) of the used slots.
// Model m's vertices go to vid + 1 + m: each model's block of the vertex
// array starts with its own dummy vertex (written as zeros here).
hipError_t launch_verts(const Params& p, const LevelView& cur, int level,
                        const uint32_t* vslot, const uint32_t* vid, float* verts,
                        uint32_t* keys, hipStream_t s);

// Per model m (and m = nmodels for the totals): first triangle (tri_off at
// the model's first cell, ModelRec::slot0 / 4) and first vertex (vid at its
// first slot), from the scans of the emit pass; out holds 2 (nmodels + 1)
// words, the triangle offsets first.
hipError_t launch_model_bases(const ModelRec* models, uint32_t nmodels,
                              const uint32_t* tri_off,
                              const uint32_t* tri_count, uint32_t ncells,
                              const uint32_t* vid, const uint32_t* vslot, uint32_t nslots,
                              uint32_t* out, hipStream_t s);

// Per-model cell ranges of a level ([begin, end) in the level's arrays),
// derived from the ranges of the level above and its scanned ambiguity
// flags: the children of model m's ambiguous cells, then, for models whose
// root is at this level, the root.  ranges: 2 nmodels words, begin then end.
hipError_t launch_child_ranges(const ModelRec* models, uint32_t nmodels, int level,
                               const uint32_t* upper, const uint32_t* upper_base,
                               const uint32_t* upper_flags, uint32_t upper_n,
                               uint32_t* ranges, hipStream_t s);

// Exclusive prefix sum of n uint32 values (device-wide), using
// scan_scratch_words(n) words of device scratch.
size_t scan_scratch_words(size_t n);
hipError_t exclusive_scan(const uint32_t* in, uint32_t* out, size_t n,
                          uint32_t* scratch, hipStream_t s);

}   // namespace mesh
}   // namespace amd
}   // namespace libfive
