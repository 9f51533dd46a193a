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

/* Single-precision sin/cos routine shared with the product (for the
 * accuracy tests against float64 references). */
void oracle_sincos(const float* x, size_t n, float* s, float* c);

#ifdef __cplusplus
}
#endif

#endif
