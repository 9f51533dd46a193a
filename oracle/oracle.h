/*
 * Test infrastructure: CPU restatement of the reference's evaluation and
 * meshing algorithms, used only as the checker in tests/, by the smoke
 * test and as the CPU baseline in bench.py.  Nothing in the product library
 * links or calls it.
 *
 * Input is the reference-format Deck (root-first clauses of opcode and slot
 * ids, constants by slot, X/Y/Z slots), as exported by the product's
 * libfive_tree_deck(): the tape is the parity contract.
 */
#ifndef LIBFIVE_ORACLE_H
#define LIBFIVE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_deck oracle_deck;

oracle_deck* oracle_deck_new(const uint32_t* clauses, size_t n_clauses,
                             const uint32_t* const_ids,
                             const float* const_values, size_t n_consts,
                             uint32_t X, uint32_t Y, uint32_t Z,
                             uint32_t num_slots);
void oracle_deck_delete(oracle_deck* d);

/* ArrayEvaluator::values in batches of 256 points; threads > 1 splits the
 * points across std::threads, one evaluator each. */
void oracle_eval_values(oracle_deck* d, const float* x, const float* y,
                        const float* z, size_t n, float* out, int threads);

/* DerivArrayEvaluator::derivs. */
void oracle_eval_derivs(oracle_deck* d, const float* x, const float* y,
                        const float* z, size_t n, float* v, float* dx,
                        float* dy, float* dz, int threads);

/* Same tape evaluated in double precision (error reference). */
void oracle_eval_values_f64(oracle_deck* d, const double* x, const double* y,
                            const double* z, size_t n, double* out);

/* IntervalEvaluator::eval over boxes. */
void oracle_eval_intervals(oracle_deck* d, const float* xlo, const float* xhi,
                           const float* ylo, const float* yhi,
                           const float* zlo, const float* zhi, size_t n,
                           float* lower, float* upper, uint8_t* maybe_nan);

/* IntervalEvaluator::intervalAndPush on the base tape for one box; returns
 * the pushed tape's clause count and sets *changed (0 if push returned the
 * same tape). */
size_t oracle_interval_push(oracle_deck* d, const float lo[3],
                            const float hi[3], float* out_lo, float* out_hi,
                            uint8_t* out_nan, int* changed);

/* ArrayEvaluator::getAmbiguous: out[i] = 1 where a MIN/MAX clause sees equal
 * operand values at point i. */
void oracle_eval_ambiguous(oracle_deck* d, const float* x, const float* y,
                           const float* z, size_t n, uint8_t* out);
/* DerivArrayEvaluator::getAmbiguousDerivs: equal operand values and
 * different operand partials. */
void oracle_eval_deriv_ambiguous(oracle_deck* d, const float* x,
                                 const float* y, const float* z, size_t n,
                                 uint8_t* out);
/* ArrayEvaluator::valueAndPush at p; the pushed (SPECIALIZED) tape is then
 * evaluated at the nq points q.  Returns the pushed tape's clause count. */
size_t oracle_value_and_push(oracle_deck* d, const float p[3],
                             const float* qx, const float* qy,
                             const float* qz, size_t nq, float* out,
                             float* value);
/* intervalAndPush on [lo, hi], then the pushed tape evaluated at the nq
 * points q.  Returns the pushed tape's clause count. */
size_t oracle_interval_push_eval(oracle_deck* d, const float lo[3],
                                 const float hi[3], const float* qx,
                                 const float* qy, const float* qz, size_t nq,
                                 float* out);

/* Single-precision sin/cos routine shared with the product (for the
 * accuracy tests against float64 references). */
void oracle_sincos(const float* x, size_t n, float* s, float* c);

/* Dual-contouring mesh of the shape over [lo, hi] (Mesh::render with the
 * DC algorithm, one worker).  The result holds the vertices (index 0 is the
 * reserved dummy vertex), triangles, one key per vertex (x, y, z of the
 * owning cell's lower corner in finest-cell units, its level, its patch),
 * statistics (root level, ambiguous leaves, collapsed cells, then per level
 * from the root down: empty, filled, ambiguous counts) and the clause
 * count of every pushed interval tape. */
typedef struct oracle_mesh oracle_mesh;
oracle_mesh* oracle_render_mesh(oracle_deck* d, const double lo[3],
                                const double hi[3], double min_feature,
                                double max_err);
/* The same mesh built by `threads` worker threads (the top three octree
 * levels on the calling thread, the subtrees below them in parallel). */
oracle_mesh* oracle_render_mesh_mt(oracle_deck* d, const double lo[3],
                                   const double hi[3], double min_feature,
                                   double max_err, int threads);
/* fp64 bit 0: the FP64 error oracle.  The same octree algorithm, but
 * field values, partial derivatives and the collapse test's |f(v)| are
 * evaluated in double on the unpruned tape at double positions (points
 * whose double value is exactly zero use the float ambiguity and feature
 * path).  Its vertices in double come from oracle_mesh_verts64. */
oracle_mesh* oracle_render_mesh_ex(oracle_deck* d, const double lo[3],
                                   const double hi[3], double min_feature,
                                   double max_err, int threads, int fp64);
/* (fp64 bit 1: QEF sums as compensated float pairs, the product's
 * precision form.) */
/* The compensated-FP32 QEF primitives (qef.hpp, shared with the product):
 * state = 14 hi floats then 14 lo floats. */
void oracle_push_sample_f2(float* state, const float* pos, const float* deriv,
                           float value);
void oracle_f2_add_acc(float* state, const double* acc);
void oracle_f2_widen(const float* state, double* acc);
/* Emitted vertices in double (same order and count as the float ones). */
size_t oracle_mesh_verts64(const oracle_mesh* m, double* out);
void oracle_mesh_sizes(const oracle_mesh* m, size_t* n_verts, size_t* n_tris,
                       size_t* n_stats, size_t* n_tape_lengths);
void oracle_mesh_copy(const oracle_mesh* m, float* verts, uint32_t* tris,
                      uint32_t* keys, uint64_t* stats, uint32_t* tape_lengths);
/* All vertices of all ambiguous leaf cells (x, y, z in double); returns the
 * count, copies when out is not NULL. */
size_t oracle_mesh_leaf_verts(const oracle_mesh* m, double* out);
/* Every cell whose state was evaluated (interval or leaf evaluation), as
 * x, y, z (finest-cell units), level, state (0 empty, 1 filled, 2
 * ambiguous); returns the count, copies when out is not NULL. */
size_t oracle_mesh_cells(const oracle_mesh* m, uint32_t* out);
void oracle_mesh_delete(oracle_mesh* m);

/* Occupancy raster of the slice at height z (libfive_tree_render_pixels):
 * writes width * height bytes, row-major, 1 where the shape is filled.
 * With out == NULL only the sizes are returned. */
int oracle_render_pixels(oracle_deck* d, float xlo, float xhi, float ylo,
                         float yhi, float z, float res, uint8_t* out,
                         uint32_t* width, uint32_t* height);
/* The voxel grid of one axis over [lo, hi] at res (centres, count and the
 * widened bounds). */
void oracle_voxel_pts(float lo, float hi, float res, float* out, uint32_t* n,
                      float* lower, float* upper);

/* 2D dual contouring of the slice at height z over [lo, hi]
 * (libfive_tree_render_slice with min_feature and max_err).  Holds the
 * vertices (index 0 is the reserved dummy vertex) with one key each (x, y
 * of the owning cell's lower corner in finest-cell units, its level, its
 * patch), the segments as vertex pairs, the chained contours (point counts
 * and vertex indices) and every evaluated cell (x, y, level, state). */
typedef struct oracle_contours oracle_contours;
oracle_contours* oracle_render_contours(oracle_deck* d, const double lo[2],
                                        const double hi[2], float z,
                                        double min_feature, double max_err);
void oracle_contours_sizes(const oracle_contours* c, size_t* n_verts,
                           size_t* n_segs, size_t* n_contours,
                           size_t* n_points, size_t* n_cells);
void oracle_contours_copy(const oracle_contours* c, float* verts,
                          uint32_t* keys, uint32_t* segs, uint32_t* counts,
                          uint32_t* points, uint32_t* cells);
void oracle_contours_delete(oracle_contours* c);

/* Tables and building blocks, exposed for their own tests. */
int oracle_corner_manifold(int mask);
void oracle_marching_table(int8_t* v /* 256*4*12*2 */, int8_t* e /* 8*8 */,
                           int8_t* p /* 256*24 */);
int oracle_is_inside(oracle_deck* d, const float p[3]);
size_t oracle_features(oracle_deck* d, const float p[3], float* out,
                       size_t cap);
void oracle_eigen3(const double* a, double* w, double* v);
double oracle_find_vertex(const double* ata9, const double* atb,
                          double btb, const double* mass, double* vert,
                          int* rank);
void oracle_push_sample(double* acc /* 14 */, const double* pos,
                        const double* deriv, double value);
int oracle_resolution_level(const double lo[3], const double hi[3],
                            double min_feature);

#ifdef __cplusplus
}
#endif

#endif
