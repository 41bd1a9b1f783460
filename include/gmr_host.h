/* gmr_host.h -- C ABI of the host side: seeded matrices, the RePair grammar
 * builder, the grammar files and the level scheduler.
 *
 * Every function returns GMR_OK (0) or an error code; gmr_last_error() gives
 * the message of the calling thread's last failure. Handles are opaque and
 * owned by the caller (free them with the matching *_free). No function here
 * touches a GPU.
 *
 * Reference interfaces these replace:
 *   gmr_matrix_*      the dataset generators' dense/edge-list output that
 *                     mm-repair consumes (reference tests/run_integration_tests.py
 *                     :110-128; reproduce.sh:135) -- here seeded and in memory
 *   gmr_grammar_repair  mm-repair's `matrepair` (absent submodule, call sites
 *                     tests/run_integration_tests.py:130-139, reproduce.sh:135)
 *   gmr_grammar_read/write  the .val/.vc.R/.vc.C readers
 *                     (gpu-engine/gpu_engine_test.cu:291-345)
 *   gmr_grammar_expand  reconstruct_csrv (gpu-engine/cpu_selftest.cpp:149-156)
 *   gmr_schedule_*    build_schedule + BuiltSchedule::gpu_schedule
 *                     (gpu-engine/build_schedule.h:72-77); the C++ API itself is
 *                     in build_schedule.h
 */
#ifndef GMR_HOST_H
#define GMR_HOST_H

#include <stdint.h>

#include "layered_types.h"

#ifdef __cplusplus
extern "C" {
#endif

enum gmr_status {
    GMR_OK = 0,
    GMR_ERR_INVALID = 1,  /* bad argument or malformed input */
    GMR_ERR_IO = 2,       /* missing, truncated or unwritable file */
    GMR_ERR_NOMEM = 3,    /* host or device allocation failed */
    GMR_ERR_DEVICE = 4,   /* HIP runtime error */
    GMR_ERR_INTERNAL = 5,
    GMR_ERR_BARRIER = 6   /* a persistent launch's grid barrier timed out: the Y it wrote is invalid */
};

enum gmr_layout { GMR_LAYOUT_REFCOMPAT = 0, GMR_LAYOUT_WAVE64 = 1, GMR_LAYOUT_RESIDENT = 2 };

typedef struct gmr_matrix gmr_matrix;
typedef struct gmr_grammar gmr_grammar;
typedef struct gmr_schedule gmr_schedule;

const char* gmr_last_error(void);

/* ---- matrices in CSRV form (row terminals then 0; terminal = 1 + v*cols + c) */
int gmr_matrix_haplotype(uint64_t seed, int rows, int cols, int founders, double density,
                         double seg_mean, double flip, gmr_matrix** out);
int gmr_matrix_genotype(uint64_t seed, int rows, int cols, int founders, int classes,
                        double seg_mean, gmr_matrix** out);
int gmr_matrix_rmat(uint64_t seed, int scale, long long edges, double a, double b, double c,
                    int nweights, gmr_matrix** out);
int gmr_matrix_from_dense(const double* M, int rows, int cols, gmr_matrix** out);
int gmr_matrix_from_csrv(int rows, int cols, const int* seq, long long len, const double* V,
                         int nvals, gmr_matrix** out);
/* info: rows, cols, nvals, alpha (1 + largest terminal); len = CSRV length. */
int gmr_matrix_info(const gmr_matrix* m, int* rows, int* cols, int* nvals, int* alpha,
                    long long* len);
int gmr_matrix_copy(const gmr_matrix* m, int* seq, double* V);
/* Rows [r0, r1) as a matrix of their own (same columns and values): the row
 * blocks mm-repair's `matrepair -b` compresses separately (run_benchmark_16.py:88-91). */
int gmr_matrix_row_block(const gmr_matrix* m, int r0, int r1, gmr_matrix** out);
void gmr_matrix_free(gmr_matrix* m);

/* ---- grammars */
int gmr_grammar_repair(const gmr_matrix* m, int max_rounds, int threads, gmr_grammar** out,
                       int* rounds, double* seconds);
/* kind[i] in {0 binary, 1 run-length}; right ignored and rep >= 1 for run-length. */
int gmr_grammar_from_arrays(int alpha, int nrules, const int* kind, const int* left,
                            const int* right, const int* rep, const int* C, long long clen,
                            const double* V, int nvals, gmr_grammar** out);
int gmr_grammar_read(const char* base, gmr_grammar** out);
int gmr_grammar_write(const gmr_grammar* g, const char* base);
int gmr_grammar_info(const gmr_grammar* g, int* alpha, int* nrules, long long* clen, int* nvals);
/* Any output pointer may be NULL. */
int gmr_grammar_copy(const gmr_grammar* g, int* kind, int* left, int* right, int* rep, int* C,
                     double* V);
/* Expand into the CSRV string: out may be NULL to query *len; cap = capacity. */
int gmr_grammar_expand(const gmr_grammar* g, int* out, long long cap, long long* len);
/* A new grammar of the same matrix in which every binary rule that no rule
 * references and that C uses at most max_emits times is replaced, at each of
 * its occurrences in C, by its two children (surviving rules keep their order
 * and are renumbered). With max_emits = 2 the stored size does not grow and a
 * sweep over many vectors does fewer node updates for the same row-gather
 * reads. No reference counterpart: a rewrite of this builder's output that the
 * sweep benefits from; *n_inlined (may be NULL) = rules removed. */
int gmr_grammar_inline_top(const gmr_grammar* g, int max_emits, gmr_grammar** out,
                           long long* n_inlined);
void gmr_grammar_free(gmr_grammar* g);

/* ---- schedules */
int gmr_schedule_build(const gmr_grammar* g, int rows, int cols, int layout, gmr_schedule** out);
/* info[0..10] = max_depth, max_width, num_raw_nt, num_layered, num_passthrough,
 * emit_total, num_term_emit, U (compacted alpha), rows, cols, layout. */
int gmr_schedule_info(const gmr_schedule* s, long long info[11]);
/* Any output pointer may be NULL. Sizes: rules num_layered; offsets/widths
 * max_depth+1; emits emit_total; terms num_term_emit; term_orig U. */
int gmr_schedule_copy(const gmr_schedule* s, LayeredRule* rules, int* level_offsets,
                      int* level_widths, int* emit_offsets, int* emit_pos, int* emit_row,
                      int* term_sym, int* term_row, int* term_orig);
/* Borrowed view (valid while s lives). */
int gmr_schedule_view(gmr_schedule* s, GPUSchedule* out);
/* y = M x on the host CPU (OpenMP; threads 0 = the OpenMP default) over the
 * schedule's own layout, with the values the GPU path uses: float y for
 * semiring 0 (+,x), int32 y for 1 (Boolean) and 2 (min/plus, 1e9 = +inf). x has
 * cols floats, V the grammar's nvals values. The same level sweep as the
 * reference driver's OpenMP CPU reference (reference gpu_engine_test.cu:49-104);
 * each level's row-sorted emissions are split at row boundaries, so no atomics
 * are needed. */
int gmr_schedule_cpu_sweep(const gmr_schedule* s, int semiring, const float* V, int nvals, const float* x, void* y,
                           int threads);
void gmr_schedule_free(gmr_schedule* s);

#ifdef __cplusplus
}
#endif

#endif /* GMR_HOST_H */
