/*
 * C ABI of the MI355X tape-evaluation and meshing library.
 *
 * The first part declares, with identical names, struct layouts and
 * semantics, the subset of the reference C API that sits on the evaluation
 * and meshing path (reference: libfive/include/libfive.h; implementation
 * libfive/src/libfive.cpp).  Each declaration cites the reference
 * declaration it replaces.  A caller of the reference library can link
 * against this one instead for these entry points.
 *
 * The second part holds additive entry points with no counterpart in the
 * reference: batched evaluation over arrays (host or device pointers),
 * compiled tapes, mesh settings, and introspection used by tests.
 *
 * Every evaluation runs on the GPU.  When no GPU is usable, evaluation
 * entry points return NaN (or a NULL mesh / non-zero status) and set the
 * message returned by libfive_last_error(); there is no CPU fallback.
 */
#ifndef LIBFIVE_AMD_H
#define LIBFIVE_AMD_H

#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* libfive.h:26-57 */
typedef struct libfive_interval  { float lower; float upper; } libfive_interval;
typedef struct libfive_region2   { libfive_interval X, Y; } libfive_region2;
typedef struct libfive_region3   { libfive_interval X, Y, Z; } libfive_region3;
typedef struct libfive_vec2      { float x, y; } libfive_vec2;
typedef struct libfive_vec3      { float x, y, z; } libfive_vec3;
typedef struct libfive_vec4      { float x, y, z, w; } libfive_vec4;
typedef struct libfive_tri       { uint32_t a, b, c; } libfive_tri;

/* libfive.h:99-104.  verts[0] is a dummy (0,0,0) and vert_count includes
 * it; triangle indices are >= 1. */
typedef struct libfive_mesh {
    libfive_vec3* verts;
    libfive_tri* tris;
    uint32_t tri_count;
    uint32_t vert_count;
} libfive_mesh;

/* libfive.h:106-118: an indexed mesh with the vertex indices of each
 * triangle followed by -1 (coord_index_count / 4 triangles). */
typedef struct libfive_mesh_coords {
    libfive_vec3* verts;
    uint32_t vert_count;
    int32_t* coord_indices;
    uint32_t coord_index_count;
} libfive_mesh_coords;

/* libfive.h:60-92: contours as point sequences (2D, or 3D with the
 * slice height as Z); a closed contour repeats its first point last. */
typedef struct libfive_contour {
    libfive_vec2* pts;
    uint32_t count;
} libfive_contour;
typedef struct libfive_contours {
    libfive_contour* cs;
    uint32_t count;
} libfive_contours;
typedef struct libfive_contour3 {
    libfive_vec3* pts;
    uint32_t count;
} libfive_contour3;
typedef struct libfive_contours3 {
    libfive_contour3* cs;
    uint32_t count;
} libfive_contours3;

/* libfive.h:120-128: an occupancy bitmap, width * height, row-major. */
typedef struct libfive_pixels {
    bool* pixels;
    uint32_t width;
    uint32_t height;
} libfive_pixels;

/* libfive.h:184-188: an opaque, intrusively refcounted tree node. */
typedef struct libfive_tree_ libfive_tree_;
typedef struct libfive_tree_* libfive_tree;

/* libfive.h:142 (impl libfive.cpp:51-56): frees a mesh returned by any
 * render entry point. */
void libfive_mesh_delete(libfive_mesh* m);

/* libfive.h:163 / libfive.cpp:78-82: opcode lookup by Scheme-style name
 * ("add", "nth-root"), -1 if unknown. */
int libfive_opcode_enum(const char* op);
/* libfive.h:168 / libfive.cpp:84-89: argument count, -1 if invalid. */
int libfive_opcode_args(int op);

/* libfive.h:196-256 / libfive.cpp:93-154: tree construction.  Returned
 * trees are owned by the caller and released with libfive_tree_delete.
 * Invalid opcodes, wrong arity or NULL children give NULL. */
libfive_tree libfive_tree_x(void);
libfive_tree libfive_tree_y(void);
libfive_tree libfive_tree_z(void);
libfive_tree libfive_tree_var(void);
bool libfive_tree_is_var(libfive_tree t);
libfive_tree libfive_tree_const(float f);
float libfive_tree_get_const(libfive_tree t, bool* success);
libfive_tree libfive_tree_nullary(int op);
libfive_tree libfive_tree_unary(int op, libfive_tree a);
libfive_tree libfive_tree_binary(int op, libfive_tree a, libfive_tree b);
const void* libfive_tree_id(libfive_tree t);
void libfive_tree_delete(libfive_tree ptr);
/* libfive.h:290-293 / libfive.cpp:186-189: q(x, y, z) = p(x', y', z').
 * The substitution is applied at once (the tree the reference's remap node
 * becomes after Tree::flatten, tree.cpp:389-520), so printing the result
 * shows the substituted expression rather than a "(remap ...)" form. */
libfive_tree libfive_tree_remap(libfive_tree p, libfive_tree x,
                                libfive_tree y, libfive_tree z);
/* libfive.h:306 / libfive.cpp:190-193 */
libfive_tree libfive_tree_optimized(libfive_tree t);
/* libfive.h:311-316 / libfive.cpp:216-231: prefix print, freed with
 * libfive_free_str. */
char* libfive_tree_print(libfive_tree t);
void libfive_free_str(char* ptr);

/* libfive.h:262 / libfive.cpp:195-199: value at one point. */
float libfive_tree_eval_f(libfive_tree t, libfive_vec3 p);
/* libfive.h:269 / libfive.cpp:201-207: interval bound over a box. */
libfive_interval libfive_tree_eval_r(libfive_tree t, libfive_region3 r);
/* libfive.h:275 / libfive.cpp:209-214: gradient at one point. */
libfive_vec3 libfive_tree_eval_d(libfive_tree t, libfive_vec3 p);

/* libfive.h:356 / libfive.cpp:341-345: dual-contouring mesh of the region
 * with min_feature = 1/res (default max_err 1e-8).  NULL on failure. */
libfive_mesh* libfive_tree_render_mesh(libfive_tree tree, libfive_region3 R,
                                       float res);
/* libfive.h:362 / libfive.cpp:347-352: the single-worker variant; on the
 * GPU the result is the same mesh. */
libfive_mesh* libfive_tree_render_mesh_st(libfive_tree tree,
                                          libfive_region3 R, float res);

/* ------------------------------------------------------------------------
 * Additive entry points (no counterpart in the reference).
 * --------------------------------------------------------------------- */

/* Message for the most recent failure on this thread ("" if none). */
const char* libfive_last_error(void);

/* 1 when a GPU is usable by this process, 0 otherwise. */
int libfive_gpu_available(void);

/* Values of the tree at n points given as host arrays x[], y[], z[];
 * writes out[n].  Returns 0 on success. */
int libfive_tree_eval_array(libfive_tree t, const float* x, const float* y,
                            const float* z, size_t n, float* out);

/* Values and gradients at n host-array points.  Returns 0 on success. */
int libfive_tree_eval_deriv_array(libfive_tree t, const float* x,
                                  const float* y, const float* z, size_t n,
                                  float* value, float* dx, float* dy,
                                  float* dz);

/* Interval bounds over n boxes (host arrays of lower/upper corners);
 * maybe_nan may be NULL.  Returns 0 on success. */
int libfive_tree_eval_interval_array(libfive_tree t, const float* xlo,
                                     const float* xhi, const float* ylo,
                                     const float* yhi, const float* zlo,
                                     const float* zhi, size_t n,
                                     float* lower, float* upper,
                                     uint8_t* maybe_nan);

/* Ambiguity flags at n host-array points: out[i] = 1 where some MIN/MAX
 * clause of the tree's tape has equal operands at point i (replaces
 * ArrayEvaluator::getAmbiguous, libfive/src/eval/eval_array.cpp:198-217).
 * Returns 0 on success. */
int libfive_tree_eval_ambiguous_array(libfive_tree t, const float* x,
                                      const float* y, const float* z,
                                      size_t n, uint8_t* out);

/* As above, but a clause counts only when its operands' partial
 * derivatives also differ (replaces
 * DerivArrayEvaluator::getAmbiguousDerivs,
 * libfive/src/eval/eval_deriv_array.cpp:58-81).  Returns 0 on success. */
int libfive_tree_eval_deriv_ambiguous_array(libfive_tree t, const float* x,
                                            const float* y, const float* z,
                                            size_t n, uint8_t* out);

/* A tree compiled once to a device-resident tape. */
typedef struct libfive_tape_ libfive_tape_;
typedef struct libfive_tape_* libfive_tape;

libfive_tape libfive_tape_new(libfive_tree t);
void libfive_tape_delete(libfive_tape t);
/* Clause count (reference meaning: tree operations, excluding constants and
 * inputs) and register-file size of the compiled tape. */
uint32_t libfive_tape_clauses(libfive_tape t);
uint32_t libfive_tape_registers(libfive_tape t);

/* Evaluation with device pointers, enqueued on `stream` (a hipStream_t,
 * NULL for the library's stream); asynchronous.  Returns 0 on success. */
int libfive_tape_eval_values_device(libfive_tape t, const float* x,
                                    const float* y, const float* z, size_t n,
                                    float* out, void* stream);
int libfive_tape_eval_derivs_device(libfive_tape t, const float* x,
                                    const float* y, const float* z, size_t n,
                                    float* value, float* dx, float* dy,
                                    float* dz, void* stream);

/* The same with host pointers (copies in, evaluates, copies out; blocks). */
int libfive_tape_eval_values(libfive_tape t, const float* x, const float* y,
                             const float* z, size_t n, float* out);
int libfive_tape_eval_derivs(libfive_tape t, const float* x, const float* y,
                             const float* z, size_t n, float* value, float* dx,
                             float* dy, float* dz);

/* libfive_tape_eval_values without any device allocation: no counterpart
 * in the reference; for callers that cannot allocate device memory (code
 * under graph capture, tight serving loops).  The points are staged
 * through a fixed 16 MiB device array that the library holds on each GPU
 * until the process ends, so n is at most 1048576 (larger n fails with an
 * error and nothing is evaluated).  Calls serialise on the library's lock;
 * the array keeps the last call's contents.  Returns 0 on success. */
int libfive_tape_eval_values_noalloc(libfive_tape t, const float* x,
                                     const float* y, const float* z, size_t n,
                                     float* out);

/* Mesh output forms (libfive.h:369-410, libfive.cpp:354-440): the same
 * mesh as libfive_tree_render_mesh, as -1-separated coordinate indices, or
 * written as binary STL (80-byte header, triangle count, per triangle a
 * zero normal, three vertices and a zero attribute word;
 * mesh.cpp:139-198).  save_meshes takes a NULL-terminated tree array and
 * uses max_err = 10^-quality. */
libfive_mesh_coords* libfive_tree_render_mesh_coords(libfive_tree tree,
                                                     libfive_region3 R,
                                                     float res);
void libfive_mesh_coords_delete(libfive_mesh_coords* m);
bool libfive_tree_save_mesh(libfive_tree tree, libfive_region3 R, float res,
                            const char* f);
bool libfive_tree_save_meshes(libfive_tree trees[], libfive_region3 R,
                              float res, float quality, const char* f);

/* 2D dual contouring of the slice at height z (libfive.h:318-335,
 * libfive.cpp:233-293 -> Contours::render, contours.cpp:25-49, with
 * min_feature = 1 / res and max_err = 1e-8): quadtree, leaf edge search and
 * QEF vertices as the 3D mesher, segments from the dual walk
 * (dc_contourer.cpp:16-98) joined head to tail into contours
 * (contours.cpp:112-184).  Free with libfive_contours_delete /
 * libfive_contours3_delete (libfive.h:132-150). */
libfive_contours* libfive_tree_render_slice(libfive_tree tree, libfive_region2 R,
                                            float z, float res);
libfive_contours3* libfive_tree_render_slice3(libfive_tree tree, libfive_region2 R,
                                              float z, float res);
void libfive_contours_delete(libfive_contours* cs);
void libfive_contours3_delete(libfive_contours3* cs);
/* The same with explicit settings, and the vertices, their keys (x, y of
 * the owning cell's lower corner in finest-cell units, level, patch), the
 * segments as vertex-index pairs and every cell whose state was evaluated
 * (interval or leaf; state 0 empty, 1 filled, 2 ambiguous), for tests and
 * tools: returns NULL
 * with libfive_last_error() set on failure. */
typedef struct libfive_slice_segments {
    float* verts;          /* 2 per vertex; vertex 0 is a reserved dummy */
    uint32_t* keys;        /* 4 per vertex */
    uint32_t vert_count;
    uint32_t* segs;        /* 2 per segment */
    uint32_t seg_count;
    uint32_t* cells;       /* 4 per evaluated cell: x, y, level, state */
    uint32_t cell_count;
} libfive_slice_segments;
libfive_slice_segments* libfive_tree_render_slice_segments(
    libfive_tree tree, libfive_region2 R, float z, double min_feature,
    double max_err);
void libfive_slice_segments_delete(libfive_slice_segments* s);

/* Occupancy raster of the slice at height z (libfive.h:408-415,
 * libfive.cpp:442-465 -> Heightmap::render, heightmap.cpp:106-319).  The
 * region is widened to whole voxels of size 1/res (voxels.cpp:23-45); a
 * pixel is set where the tree is below zero at the voxel centre, decided by
 * interval bounds for blocks of more than 256 voxels.  Free with
 * libfive_pixels_delete (libfive.h:155). */
libfive_pixels* libfive_tree_render_pixels(libfive_tree tree,
                                           libfive_region2 R,
                                           float z, float res);
void libfive_pixels_delete(libfive_pixels* ps);

/* Free variables (libfive.h:171-179, 389-430; libfive.cpp:71-76,
 * 408-496): an evaluator holds a tree and values for its free variables
 * (keyed by libfive_tree_id of each variable); update_vars returns true
 * when a value of a variable of the tree changed; save_mesh renders with
 * the current values and the default settings (min_feature 0.1) and writes
 * binary STL. */
typedef struct libfive_vars {
    void* const* vars;
    float* values;
    uint32_t size;
} libfive_vars;
void libfive_vars_delete(libfive_vars* j);
typedef struct libfive_evaluator_ libfive_evaluator_;
typedef struct libfive_evaluator_* libfive_evaluator;
libfive_evaluator libfive_tree_evaluator(libfive_tree tree, libfive_vars vars);
bool libfive_evaluator_update_vars(libfive_evaluator eval_tree, libfive_vars vars);
bool libfive_evaluator_save_mesh(libfive_evaluator evaluator, libfive_region3 R,
                                 const char* f);
void libfive_evaluator_delete(libfive_evaluator ptr);

/* Many independent meshes in one call (the many-model configuration; no
 * counterpart in the reference, whose callers loop over
 * libfive_tree_render_mesh, libfive.cpp:341-345).  Model i is meshed over
 * regions[i] at res[i] exactly as libfive_tree_render_mesh would (same
 * output), `streams` models at a time on their own GPU streams (<= 0: 8),
 * largest resolution first.  out[i] receives the mesh or NULL; returns the
 * number of models that failed (the first error is in
 * libfive_last_error()). */
int libfive_trees_render_mesh_batch(const libfive_tree* trees,
                                    const libfive_region3* regions,
                                    const float* res, size_t n, int streams,
                                    libfive_mesh** out);

/* The same batch through one set of level-synchronous GPU passes (no
 * counterpart in the reference): every octree level, the leaf pass, the
 * collapse and the triangle emission run as one launch over the cells of
 * all n models, so the host waits once per level for the whole batch
 * instead of once per level per model.  Model i's mesh is the one
 * libfive_tree_render_mesh(trees[i], regions[i], res[i]) returns, bit for
 * bit.  out[i] receives the mesh or NULL; returns the number of models
 * that failed (all of them when a GPU step fails; the error is in
 * libfive_last_error()). */
int libfive_trees_render_mesh_merged(const libfive_tree* trees,
                                     const libfive_region3* regions,
                                     const float* res, size_t n,
                                     libfive_mesh** out);

/* Launch-schedule autotuner (no counterpart in the reference, whose CPU
 * evaluators have no launch schedules).  The first large evaluation call of
 * a tape and the first leaf pass of each register-bank class time every
 * launch schedule the kernel offers (workgroup shape, points per lane,
 * waves per workgroup; all give the same results) on a slice of the real
 * work and keep the fastest for the process.  libfive_autotune_enable(0)
 * turns tuning off (default schedules); libfive_autotune_report writes the
 * decisions made so far as JSON text into buf (at most cap bytes, NUL
 * terminated) and returns the length needed including the NUL. */
void libfive_autotune_enable(int on);
size_t libfive_autotune_report(char* buf, size_t cap);
/* Forgets every decision; makes `choice` the decision for `key` (a key of
 * the report) without timing, so that each schedule can be exercised. */
void libfive_autotune_reset(void);
void libfive_autotune_force(const char* key, int choice);

/* Tape specialisation (no counterpart in the reference, whose evaluators
 * interpret the tape on the CPU: libfive/src/eval/eval_array.cpp:107-118).
 * Compiles the tape into straight-line value and derivative kernels with
 * hipRTC (about a second for a thousand clauses) and loads them; after
 * that, libfive_tape_eval_values_device / _derivs_device run the
 * specialised kernels, which give the same bits as the interpreter.
 * compile_ms (may be NULL) receives the compile and load time.  Returns 0
 * on success; on failure the interpreter stays in use. */
int libfive_tape_specialize(libfive_tape t, double* compile_ms);
/* Chooses the specialised kernels (on != 0, if compiled) or the
 * interpreter for the *_device evaluation calls of this tape. */
void libfive_tape_use_specialized(libfive_tape t, int on);
/* Compiles a tree's specialised kernels on the host without loading them
 * (no GPU needed); returns the code object size in bytes, or 0 on error. */
size_t libfive_tree_specialized_code_size(libfive_tree t);

/* Points evaluated per GPU lane by the value kernel (1, 2 or 4; 0 picks by
 * register count).  Also read from LIBFIVE_POINTS_PER_LANE at start-up. */
void libfive_set_points_per_lane(int p);

/* Compiles a tree on the host only (no GPU needed) and reports the tape's
 * clause count (reference meaning) and register count.  Returns 0 on
 * success. */
int libfive_tree_tape_info(libfive_tree t, uint32_t* clauses,
                           uint32_t* registers);

/* The reference-format clause list of a tree's Deck (deck.cpp:19-94), for
 * checking the front end: `clauses` receives 4 words (op, id, a, b) per
 * clause, root first; constants are (id, value) pairs.  Call with NULL
 * buffers to get the counts.  Returns 0 on success. */
int libfive_tree_deck(libfive_tree t, uint32_t* clauses, size_t* n_clauses,
                      uint32_t* const_ids, float* const_values,
                      size_t* n_consts, uint32_t* xyz_ids,
                      uint32_t* num_slots);
/* The free variables of a tree's Deck: slot_ids[k] is the slot of the
 * variable whose libfive_tree_id is var_ids[k], in the Deck's order
 * (evaluator.hpp:41-49 loads each variable's value into its slot).  Call
 * with NULL buffers to get the count in n.  Returns 0 on success. */
int libfive_tree_deck_vars(libfive_tree t, uint32_t* slot_ids, const void** var_ids,
                           size_t* n);

/* Mesh settings beyond the reference's two render entry points
 * (BRepSettings, libfive/include/libfive/render/brep/settings.hpp:27-74). */
typedef struct libfive_mesh_settings {
    double min_feature;     /* default 0.1 */
    double max_err;         /* default 1e-8; < 0 disables collapse */
    uint32_t precision;     /* QEF accumulation: 0 FP64 (default), 1
                               compensated FP32; FP64 eigensolve in both */
    uint32_t collect_stats; /* non-zero: record per-level statistics */
    volatile const int* cancel;   /* optional; non-zero requests cancel */
} libfive_mesh_settings;

void libfive_mesh_settings_default(libfive_mesh_settings* s);
libfive_mesh* libfive_tree_render_mesh_settings(libfive_tree tree,
                                                libfive_region3 R,
                                                const libfive_mesh_settings* s);

/* Per-vertex keys of a mesh returned by this library: 5 words per vertex
 * (x, y, z of the owning cell's lower corner in finest-cell units, the
 * cell's octree level, the vertex patch within the cell), starting with the
 * dummy vertex 0.  Kept for meshes rendered with collect_stats set; valid
 * until libfive_mesh_delete; NULL for other meshes.  Used to compare meshes
 * independently of vertex numbering. */
const uint32_t* libfive_mesh_vertex_keys(const libfive_mesh* m);

/* Statistics of the last render on the calling thread made with
 * collect_stats set: per-level cell counts (empty, filled, ambiguous; root
 * level first) and the clause count of the tape of every ambiguous cell
 * above the finest level (Tape::size, libfive/include/libfive/eval/
 * tape.hpp:46).  Both return the number of entries (levels, lengths) and
 * copy at most cap of them when out is not NULL. */
size_t libfive_last_mesh_level_counts(uint64_t* out, size_t cap);
size_t libfive_last_mesh_tape_lengths(uint32_t* out, size_t cap);
/* Every cell of the last render (with collect_stats) whose state was
 * evaluated, before collapse: 5 words per cell (x, y, z of the lower corner
 * in finest-cell units, level, state: 0 empty, 1 filled, 2 ambiguous).
 * Returns the count; copies at most cap cells when out is not NULL. */
size_t libfive_last_mesh_cells(uint32_t* out, size_t cap);
/* Wall-clock milliseconds of the last render: build, collapse, emit, total. */
void libfive_last_mesh_times(double* ms4);

#ifdef __cplusplus
}
#endif

#endif  /* LIBFIVE_AMD_H */
