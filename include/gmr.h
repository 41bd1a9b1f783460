/* gmr.h -- C ABI of the device engine: a grammar made resident in HBM once and
 * swept many times by hand-written HIP kernels for gfx950.
 *
 * Replaces the inside of the reference's run_gpu_right_multiply_layered /
 * run_gpu_right_multiply_batched (gpu-engine/gpu_engine.cu:135-587), which
 * allocate, upload, sweep and free on every call. Here the upload
 * (gmr_engine_create) and the sweep (gmr_engine_run) are separate, the
 * semiring is an argument instead of the SEMIRING environment variable
 * (gpu_engine.cu:379-396), inputs and outputs are device pointers on a caller
 * stream, and failures return a status instead of exit(1) (gpu_engine.cu:38-46).
 * The drop-in entry points with the reference's exact signatures are declared
 * in gpu_engine.h and are built on this ABI.
 *
 * Value layout (reference gpu_engine.h:116-120): X is cols x B with X[c*B + b],
 * Y is rows x B with Y[r*B + b]. X is float for every semiring (Boolean and
 * Tropical read it as the reference does: x != 0, (int)x); Y holds float for
 * (+,x) and int32 for Boolean and Tropical.
 */
#ifndef GMR_H
#define GMR_H

#include <stdint.h>

#include "gmr_host.h"
#include "layered_types.h"

#ifdef __cplusplus
extern "C" {
#endif

enum gmr_semiring { GMR_PLUS_TIMES = 0, GMR_BOOLEAN = 1, GMR_TROPICAL = 2 };

/* How a sweep is launched. */
enum gmr_variant {
    GMR_VARIANT_LEVELS = 0,     /* one launch per phase and level (the plain sweep) */
    GMR_VARIANT_GRAPH = 1,      /* the same launches replayed from a captured hipGraph */
    GMR_VARIANT_PERSISTENT = 2  /* one launch, the engine's own grid barrier between levels */
};

/* How top-sequence symbols reach y. */
enum gmr_emit {
    GMR_EMIT_SPOT = 0,   /* emit-on-the-spot: atomics right after each level */
    GMR_EMIT_GATHER = 1  /* Resident layout only: one row-ordered gather after the sweep */
};

/* Node-update kernel of a level. */
enum gmr_eval_kind {
    GMR_EVAL_GROUP = 0,   /* a lane group per rule, lanes over vectors (dwordx4), four rules in flight */
    GMR_EVAL_THREAD = 1,  /* one lane per (rule, 4-vector chunk) */
    GMR_EVAL_WAVE = 2,    /* one wave per rule */
    GMR_EVAL_LDS = 3      /* single vector: left-child window staged in LDS */
};

enum gmr_run_flags {
    GMR_RUN_TIMED = 1,        /* record device time (see gmr_engine_last_kernel_ms) */
    /* Row gather of one vector, or of 2-4 vectors per tile. By default the top-sequence
     * symbols are split into 8 x passes value ranges, 8 per pass (one per XCD), each swept in
     * 64-symbol chunks, so each XCD's L2 caches its share of the values; row pieces combine
     * atomically into a zeroed y (in a fused multi-grammar persistent launch the
     * single-vector gather uses the single list below and 2-4 vectors the row segments). */
    GMR_RUN_ROW_SEGMENTS = 2, /* lanes per row segment instead */
    GMR_RUN_SINGLE_LIST = 4,  /* one vector: the whole top sequence as one chunked list instead */
    /* Boolean semiring, one vector, gather emission, levels or graph variant: every terminal
     * and rule value is one bit of a bitset ((U + rules) / 8 bytes, small enough to stay in
     * L2), node updates and the gather read bits, and y gets 0/1 as int32. */
    GMR_RUN_BITS = 8,
    /* One vector, gather emission, levels or graph variant: the top-sequence symbols are
     * regrouped by value into ranges of lds_values values, one list per range; a workgroup
     * copies a range of the value array into LDS once and looks its symbols' values up
     * there, and every row piece combines atomically into a zeroed y. */
    GMR_RUN_LDS_RANGES = 16,
    /* With GMR_RUN_LDS_RANGES: the lists number the rows so that rows of equal top-sequence
     * length and similar symbols sit next to each other (length, then the least hash of the
     * row's symbols), so the pieces a wave combines land on nearby words of a row-ordered
     * copy of y that the engine holds (rows x 4 bytes, allocated with the first such lists);
     * a last pass writes that copy into y in the matrix's row order. */
    GMR_RUN_ROW_ORDER = 32,
    /* One vector, Resident layout, levels or graph variant: the top levels, from the lowest
     * one whose rules number at most 8,192 together with every level above it (at least two
     * levels), run in one workgroup that keeps their values in LDS (160 KiB), after a launch
     * that stages the values of their children below them: two launches instead of one per
     * level. Without such levels, or where a workgroup may not hold that much LDS, the sweep
     * runs as without the flag. Not with GMR_RUN_BITS. */
    GMR_RUN_FUSED_TAIL = 64,
    /* One vector, Resident layout, levels or graph variant: levels run two to a launch; a rule
     * of the upper level of a pair computes its children in the lower level from their own
     * kids instead of waiting for them (levels below the fused tail, with GMR_RUN_FUSED_TAIL;
     * with GMR_RUN_BITS every level, the bits of a child being the OR of its kids' bits). */
    GMR_RUN_PAIRED_LEVELS = 128,
    /* gmr_engine_traverse only: the engine holds the transposed matrix, so every step is the
     * right product y = M^T x, with the one-vector gathers of the right product (GMR_RUN_BITS,
     * GMR_RUN_LDS_RANGES and the flags above); the rest of opts is that product's. Other entry
     * points refuse the flag. */
    GMR_RUN_TRANSPOSED = 256
};

typedef struct gmr_run_opts {
    int variant;            /* gmr_variant */
    int emit;               /* gmr_emit */
    int b_tile;             /* vectors per pass over the grammar (0: all B at once) */
    int block;              /* threads per workgroup (0: 256) */
    int grid_cap;           /* persistent variant: max workgroups (0: 16 waves per CU) */
    int eval_kind;          /* gmr_eval_kind of every level not listed in level_kind */
    int flags;              /* gmr_run_flags */
    int xcd_passes;         /* one vector, XCD-partitioned gather: the values are cut into 8 x this
                               many ranges, swept in this many launches so each XCD's L2 holds the
                               values of one range at a time (0: sized from the value count) */
    const int* level_kind;  /* optional host array, max_depth + 1 entries ([0] unused) */
    /* Level-chunk sizes (0: the defaults below; the autotuner searches them):
     * lds_values  values a workgroup stages in LDS, a multiple of 1024: with GMR_RUN_LDS_RANGES
     *             the values per range (default 32768; at least 4096, below 65535, at most what
     *             one workgroup's LDS holds; ranges of which two fit one CU's LDS, up to 20480
     *             4-byte values on MI355X, run two workgroups per CU); with GMR_RUN_BITS the bits per range (default 524288,
     *             131072..1048576); otherwise the left-child window of GMR_EVAL_LDS levels in the
     *             levels and graph variants (default 4096, 1024..16384)
     * chunks      64-symbol chunks a wave of a chunked single-vector gather has in flight:
     *             the XCD-partitioned gather 1, 2 or 4 (default 2), GMR_RUN_LDS_RANGES 2, 4
     *             or 8 (default 4), GMR_RUN_BITS 2, 4 or 8 (default 8) */
    int lds_values;
    int chunks;
} gmr_run_opts;

typedef struct gmr_engine gmr_engine;

/* Upload a layered schedule. `layout` says how children are numbered:
 * GMR_LAYOUT_REFCOMPAT / GMR_LAYOUT_WAVE64 -- slots of the previous level;
 * GMR_LAYOUT_RESIDENT -- global slots of earlier levels. val_idx / col_idx
 * have U-1 entries (compact terminal q >= 1 is value V[val_idx[q-1]] in column
 * col_idx[q-1]). Every index is range-checked on the host before upload. */
int gmr_engine_create(int rows, int cols, int U, int num_rules, const LayeredRule* rules,
                      const int* val_idx, const int* col_idx, const float* V, int nvals,
                      const GPUSchedule* sched, int layout, gmr_engine** out);
/* Upload a schedule built by gmr_schedule_build. */
int gmr_engine_from_schedule(const gmr_schedule* s, const float* V, int nvals, gmr_engine** out);
/* info[0..12] = rows, cols, U, num_layered, max_depth, max_width, emit_total,
 * num_term_emit, layout, has_runlen, device bytes (structure), nvals, row-gather
 * segments (0 outside the Resident layout). */
int gmr_engine_info(const gmr_engine* e, long long info[16]);
/* Y = M X on `stream` (a hipStream_t; NULL = default stream). Asynchronous:
 * returns after enqueueing. opts may be NULL (levels variant, spot emit). */
int gmr_engine_run(gmr_engine* e, int semiring, const void* d_X, int B, void* d_Y,
                   const gmr_run_opts* opts, void* stream);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc. As gmr_engine_run, but the sweep's workspace (terminal
 * values and the rule-value array or frontier buffers: about (U + rules) x 4
 * bytes per vector of a tile in the Resident layout, plus 64 KiB of staging with
 * GMR_RUN_FUSED_TAIL) comes from a static device
 * array of GMR_NOALLOC_SCRATCH_BYTES per GPU, and the call allocates nothing
 * (no hipMalloc, hipFree, graph instantiation or table upload). A run whose
 * workspace exceeds that size is refused with GMR_ERR_INVALID (use a narrower
 * b_tile or gmr_engine_run), and so is any run that would have to build
 * something first: the graph variant, level_kind tables, and the one-vector
 * gathers' lists (XCD-partitioned, LDS ranges, bit mode) for options not yet
 * used with gmr_engine_run on this engine -- run it once with the same options
 * at set-up, or use GMR_RUN_SINGLE_LIST / GMR_RUN_ROW_SEGMENTS. Every call on a
 * GPU shares that GPU's array, so calls that overlap on one GPU (several
 * streams or threads, any engines) race; calls in stream order are safe,
 * since each sweep writes all of its workspace before reading it. */
#define GMR_NOALLOC_SCRATCH_BYTES (64u * 1024u * 1024u)
int gmr_engine_run_noalloc(gmr_engine* e, int semiring, const void* d_X, int B, void* d_Y,
                           const gmr_run_opts* opts, void* stream);
/* Resident layout: order in which the row gather visits its segments.
 * 0 = row order (default); 1 = segments sharing their highest rule together
 * (cache locality for the value rows they read). Results are unchanged. */
int gmr_engine_set_gather_order(gmr_engine* e, int order);
/* Left product Y^T = X^T M on `stream`: X is rows x B (X[r*B + b]), Y is
 * cols x B (Y[c*B + b]); same semirings and leaves as gmr_engine_run, with the
 * matrix value on the left. Resident layout only; opts->variant LEVELS or GRAPH
 * (b_tile and block honoured; for one vector the gather is XCD-partitioned
 * unless flags has GMR_RUN_SINGLE_LIST). The first call builds the
 * incoming-edge lists (synchronising the device). mm-repair computes this product beside Ax
 * (reference run_benchmark.py:138-141); the reference's GPU engine has none. */
int gmr_engine_run_left(gmr_engine* e, int semiring, const void* d_X, int B, void* d_Y,
                        const gmr_run_opts* opts, void* stream);
/* Fixpoint traversal over a square matrix: x <- x (+) x^T M until no entry
 * changes or max_iters products were computed. d_x: rows floats, in: the
 * start (Boolean 0/1 sources; Tropical distances, 1e9 = unreached), out: the
 * reachable set / shortest distances along row -> column edges. Synchronises
 * `stream` after every product; *iters_out = products computed. Each product
 * is the left product of the engine's matrix M, or, with GMR_RUN_TRANSPOSED in
 * opts->flags, the right product of the engine's matrix when it holds M^T
 * (the same x^T M, by the one-vector right-product gathers). Every product
 * after the first multiplies only the vertices the step before moved (both
 * semirings are idempotent, so the steps and their count are unchanged); the
 * engine holds that frontier beside y (2 x rows x 4 bytes, allocated on the
 * first call). The reference's fixpoint comparison is cuGraph BFS / SSSP
 * (reference cugraph_bench.py:53-61). */
int gmr_engine_traverse(gmr_engine* e, int semiring, void* d_x, int max_iters, int* iters_out,
                        const gmr_run_opts* opts, void* stream);
/* The persistent variant. Its grid barrier needs every workgroup of the launch
 * resident at once; no cooperative launch guarantees it, so the grid is sized
 * from the occupancy query with one workgroup per CU of headroom (the query
 * has reported one more than the dispatcher keeps resident): min(occupancy - 1,
 * 1024 / block) workgroups per CU, times the CU count, capped by grid_cap. A
 * block size at which fewer than two workgroups fit per CU is refused
 * (GMR_ERR_INVALID). Every barrier wait is bounded (2 s by default); a wait that
 * times out marks that launch failed, and the kernel still ends. A failure is
 * reported as GMR_ERR_BARRIER by the next gmr_engine_run on the engine (which
 * then does not run), by the same call when it was timed (GMR_RUN_TIMED
 * synchronises), and by gmr_engine_check. Later launches are not affected.
 *
 * Health check: *barrier_error = 1 if a grid-barrier wait of an earlier
 * persistent launch on this engine timed out since the last report (its
 * results are invalid; the barrier has been reset), else 0. Synchronises the
 * device. */
int gmr_engine_check(gmr_engine* e, int* barrier_error);
/* Limit of one grid-barrier wait in microseconds (0: the default of 2 s); for
 * tests of the failure path, which set a tiny limit so the waits give up. */
int gmr_engine_set_barrier_timeout(gmr_engine* e, long long microseconds);
/* Diagnostics of the persistent variant: out[0..6] = barrier arrivals on the
 * device, unreported barrier failure (0/1), arrivals the host expects, wall-clock rate (kHz),
 * CU count, workgroups of the last persistent launch, workgroups per CU,
 * captured graphs held. Synchronises the device. */
int gmr_engine_debug(gmr_engine* e, long long out[8]);
/* Device time of the last gmr_engine_run on this engine, when timing was
 * requested with GMR_RUN_TIMED in opts->flags (that run synchronises its stream). */
int gmr_engine_last_kernel_ms(const gmr_engine* e, double* ms);
/* Per-level device times: runs the levels variant `reps` times with events
 * around every level's node-update launch, using the eval kinds in opts (no
 * vector tiling: B is one tile). level_ms has max_depth + 2 entries: [0] from
 * the start to the first level (terminal values, terminal emits), [k] level k's
 * node update, [max_depth+1] from the last node update to the end (the last
 * level's emits, or the row gather). Synchronises the stream. */
int gmr_engine_time_levels(gmr_engine* e, int semiring, const void* d_X, int B, void* d_Y,
                           const gmr_run_opts* opts, int reps, float* level_ms, void* stream);
void gmr_engine_destroy(gmr_engine* e);
int gmr_device_sync(void);

/* ---- many grammars swept together (uneven sizes, own B each) ---- */
typedef struct gmr_batch gmr_batch;
/* The engines stay owned by the caller and must outlive the batch; an engine
 * may appear once. */
int gmr_batch_create(gmr_engine* const* engines, int n, gmr_batch** out);
/* Y_g = M_g X_g for every grammar g: d_X[g] is cols_g x B[g], d_Y[g] rows_g x B[g].
 * variant: GMR_VARIANT_LEVELS runs the grammars one after another;
 * GMR_VARIANT_GRAPH replays one hipGraph with a branch of per-level launches per
 * grammar; GMR_VARIANT_PERSISTENT runs every grammar in one launch, level by
 * level, with a grid barrier between levels. opts: n entries (emission, eval
 * kinds per grammar; block and grid_cap of entry 0 apply to the launch) or NULL. */
int gmr_batch_run(gmr_batch* b, int semiring, const void* const* d_X, const int* B, void* const* d_Y,
                  int variant, const gmr_run_opts* opts, void* stream);
/* As gmr_engine_check and gmr_engine_set_barrier_timeout, for the batch's
 * persistent launches (a failure is also reported by the next gmr_batch_run). */
int gmr_batch_check(gmr_batch* b, int* barrier_error);
int gmr_batch_set_barrier_timeout(gmr_batch* b, long long microseconds);
void gmr_batch_destroy(gmr_batch* b);

/* ---- the schedule proposer's network, trained on the GPU by its own kernels ----
 * in -> H -> H -> out, ReLU on both hidden layers, masked mean-squared error,
 * Adam. Parameters are one FP32 array: W1 [H x in], b1 [H], W2 [H x H], b2 [H],
 * W3 [out x H], b3 [out], row-major. d_X is N x in, d_T and d_M (0/1 mask) are
 * N x out, all device pointers; count = number of unmasked targets. No
 * counterpart in the reference: it chooses launch schedules only. */
typedef struct gmr_mlp gmr_mlp;
int gmr_mlp_create(int n_in, int n_hidden, int n_out, uint64_t seed, gmr_mlp** out);
long long gmr_mlp_num_params(const gmr_mlp* m);
int gmr_mlp_get_params(const gmr_mlp* m, float* host_params);
int gmr_mlp_set_params(gmr_mlp* m, const float* host_params);
int gmr_mlp_forward(gmr_mlp* m, const float* d_X, int N, float* d_out, void* stream);
/* Loss (mean over unmasked targets) and its gradient (host copy, may be NULL). */
int gmr_mlp_loss_grad(gmr_mlp* m, const float* d_X, const float* d_T, const float* d_M, int N, float count,
                      float* loss, float* host_grad);
/* `epochs` full-batch Adam steps at learning rate lr; losses[e] (may be NULL). */
int gmr_mlp_train(gmr_mlp* m, const float* d_X, const float* d_T, const float* d_M, int N, float count,
                  int epochs, float lr, float* losses);
void gmr_mlp_destroy(gmr_mlp* m);

#ifdef __cplusplus
}
#endif

#endif /* GMR_H */
