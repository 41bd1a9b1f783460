// The device engine object behind the gmr_engine handle.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <tuple>
#include <vector>

#include "gmr.h"
#include "layered_types.h"
#include "sweep_kernels.hpp"

namespace gmr {

// The grid barrier of the persistent launches (one engine's, or one batch's):
// the device counter and failed epoch, and a word of pinned host memory that a
// timed-out wait sets, so a failure is seen on the host without synchronising.
struct BarrierState {
    dev::GridBarrier* d = nullptr;
    int* h_fail = nullptr;           // pinned host word (mapped)
    int* d_fail = nullptr;           // its device address
    unsigned long long base = 0;     // arrivals made by earlier launches
    long long timeout_us = 0;        // a wait's limit (0: 2 s)
    int wall_khz = 100000;
    void init(int wall_khz);
    void release() noexcept;
    // The view a launch of `grid` workgroups passes to its kernel; then
    // advance() by the arrivals it makes.
    dev::BarrierRef ref() const;
    void advance(unsigned long long arrivals) { base += arrivals; }
    // A timed-out wait since the last call (clears the word; no device sync).
    bool take_failure();
    // Synchronises the device: 1 if a wait timed out or the arrivals disagree
    // with what the launches made (the state is then reset), else 0.
    int check();
};

struct Engine {
    // structure (host copies of what the launch loop needs)
    int rows = 0, cols = 0, U = 0, NL = 0, L = 0, nvals = 0, layout = 0;
    bool resident = false, has_runlen = false;
    int max_width = 0, E = 0, Et = 0;
    int gather_lanes = 1;  // lanes per row of the single-vector gather
    std::vector<int> level_off, emit_off;
    std::size_t struct_bytes = 0;

    // structure (device)
    int2* d_kids = nullptr;
    float* d_coeff = nullptr;
    int* d_tcol = nullptr;
    float* d_tval = nullptr;
    int* d_tref = nullptr;           // terminals some rule has as a child (increasing)
    int n_tref = 0;
    int2* d_emits = nullptr;
    int2* d_term = nullptr;
    int* d_crow = nullptr;
    int* d_csym = nullptr;
    int4* d_gseg = nullptr;          // gather segments (row, first, end, split)
    int n_gseg = 0;
    int gather_split = 0;            // some row is split over several segments
    int* d_zrows = nullptr;          // the split rows
    int n_zrows = 0;
    static constexpr int kGatherSeg = 128;
    int gather_order = 0;            // 0: row order, 1: by highest rule
    unsigned long long* d_fbits = nullptr;  // flat single-vector gather: row starts per 64 symbols
    int* d_ford = nullptr;           // rows started before each 64-symbol chunk
    int* d_frow = nullptr;           // non-empty rows in order
    int n_fchunk = 0;
    // XCD-partitioned single-vector gather, per pass count (built on first use)
    struct XParts {
        int* d_sym = nullptr;
        unsigned long long* d_bits = nullptr;
        int* d_ord = nullptr;
        int* d_row = nullptr;
        int* d_poff = nullptr;
        std::vector<int> poff;       // 8 x passes + 1 chunk offsets
        // LDS-range gather with packed entries: entry = (row - d_cbase[chunk]) << 16 |
        // index inside the range (0xFFFF: padding), so a piece's row needs no lookup;
        // d_ord / d_row are then not used
        bool packed = false;
        int* d_cbase = nullptr;      // per chunk: row of its first entry
    };
    std::map<int, XParts> xparts;
    std::map<int, XParts> bparts;    // Boolean bit mode: lists by LDS-sized bit range, per range size
    const XParts& bit_lists(int range_bits);
    // one vector, GMR_RUN_LDS_RANGES: lists by value range, keyed by 2 x range size + 1 when their
    // rows follow row_new (GMR_RUN_ROW_ORDER)
    std::map<int, XParts> lparts;
    const XParts& lds_lists(int span, bool row_order = false);
    // GMR_RUN_ROW_ORDER: rows numbered by (top-sequence length, min-hash of the row's symbols);
    // d_row_new[r] is the number of matrix row r, ws_yrow the row-ordered copy of y (rows values)
    int* d_row_new = nullptr;
    void* ws_yrow = nullptr;
    bool row_order_built = false;
    std::vector<int> row_order_host;  // new number -> matrix row (built with the first ordered lists)
    // GMR_RUN_FUSED_TAIL: staged child values of the fused top levels (2 x 8,192
    // values), allocated with the first such run
    void* ws_tail = nullptr;
    static constexpr std::size_t kTailStageBytes = 2 * 8192 * 4;
    int max_lds = 65536;             // LDS bytes a workgroup may use
    const XParts& xparts_for(int passes);
    int auto_xcd_passes(int Bs = 1, int Bt = 1) const;  // pass count of the XCD-partitioned gather
    std::vector<int> crow_host;      // [0] = number of top-sequence symbols
    int* d_level_off = nullptr;
    int* d_emit_off = nullptr;

    // value workspace, grown on demand
    void* ws_T = nullptr;
    void* ws_W0 = nullptr;
    void* ws_W1 = nullptr;
    std::size_t ws_T_bytes = 0, ws_W_bytes = 0;
    unsigned long long* ws_bits = nullptr;  // Boolean bit mode: terminal words, then rule words
    int bits_tw = 0;                 // terminal words of ws_bits: ceil(U / 64)
    unsigned long long* ensure_bits();

    // no-allocation runs (gmr_engine_run_noalloc): while set, the workspace is
    // the caller-side scratch array and every path that would allocate throws
    bool no_alloc = false;
    int device = 0;                  // the device the engine was uploaded to
    void forbid_alloc(const char* what) const;
    void run_noalloc(int semiring, const float* X, int B, void* Y, const gmr_run_opts& o, hipStream_t st,
                     void* scratch, std::size_t scratch_bytes);

    // device facts and timing
    int num_cus = 0, wall_khz = 0;
    BarrierState barrier;
    int last_grid = 0, last_per_cu = 0;
    int* d_kinds = nullptr;          // per-level eval kinds of the latest run
    std::vector<int> kinds_host;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipStream_t capture_stream = nullptr;
    double last_ms = 0.0;

    // captured per-level launch sequences, keyed by everything they bake in
    struct Graph {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    using GraphKey = std::tuple<int, uintptr_t, uintptr_t, int, int, int, int, int, uintptr_t>;
    std::map<GraphKey, Graph> graphs;

    // left product (built on first use, Resident layout only)
    bool left_ready = false;
    int* d_in_off = nullptr;         // nodes + 1 (rule slots, then terminals in column order)
    int* d_in_src = nullptr;         // >= 0 parent slot, < 0 row -e-1
    float* d_in_coeff = nullptr;     // per entry, only with run-length rules
    int* d_col_off = nullptr;        // cols + 1: column c's terminals are positions [off[c], off[c+1])
    float* d_col_val = nullptr;      // terminal value per position (position p is node NL + p)
    // one-vector fold: positions 1..U-1 in 64-position chunks, column-start
    // bitmap, starts before each chunk, the non-empty columns in order
    unsigned long long* d_lf_bits = nullptr;
    int* d_lf_ord = nullptr;
    int* d_lf_col = nullptr;
    long n_lf_chunk = 0;
    void* ws_A = nullptr;            // nodes x Bs values
    std::size_t ws_A_bytes = 0;
    void* ws_Xs = nullptr;           // (rows + 1) x Bs: x as semiring values, then a zero row
    std::size_t ws_Xs_bytes = 0;
    // the same incoming lists in row-gather form (grammars without run-length rules)
    int4* d_lseg = nullptr;          // (node, first, end, split) segments, node order
    int* d_lsplit = nullptr;         // nodes cut into several segments
    int n_lsplit = 0;
    int* d_lpsrc = nullptr;          // entries, each level's run padded to 64 (single vector)
    int* d_lpoff = nullptr;          // nodes + 1 offsets into d_lpsrc
    unsigned long long* d_lbits = nullptr;
    int* d_lord = nullptr;
    int* d_lrow = nullptr;
    int n_lpsrc = 0;
    std::vector<int> lseg_off;       // groups (levels 1..L, then terminals) -> segment ranges
    std::vector<int> lchunk_off;     // groups -> 64-entry chunk ranges
    // one vector, XCD-partitioned: per group 8 x lx_passes lists
    int* d_lxsym = nullptr;
    unsigned long long* d_lxbits = nullptr;
    int* d_lxord = nullptr;
    int* d_lxrow = nullptr;
    int* d_lxpoff = nullptr;         // per group 8 x lx_passes + 1 chunk offsets
    std::vector<int> lxpoff;
    int lx_passes = 1;
    std::map<GraphKey, Graph> left_graphs;

    Engine() = default;
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;
    ~Engine();

    void upload(int rows, int cols, int U, int num_rules, const LayeredRule* rules, const int* val_idx,
                const int* col_idx, const float* V, int nvals, const GPUSchedule& s, int layout);
    void ensure_workspace(int Bt, int vbytes);
    void run(int semiring, const float* X, int B, void* Y, const gmr_run_opts& o, hipStream_t st);
    // 1 if a persistent launch's barrier timed out since the last check (and
    // the barrier state was reset), else 0. Synchronises the device.
    int check_barrier() { return barrier.check(); }
    const int* device_kinds(const int* host_kinds);
    void set_gather_order(int order);
    void time_levels(int semiring, const float* X, int B, void* Y, const gmr_run_opts& o, int reps, float* out,
                     hipStream_t st);
    // Y^T = X^T M: X rows x B, Y cols x B (Resident layout).
    void run_left(int semiring, const float* X, int B, void* Y, const gmr_run_opts& o, hipStream_t st);
    // Fixpoint x <- x (+) x^T M (square matrix); returns the products computed.
    int traverse(int semiring, float* x, int max_iters, const gmr_run_opts& o, hipStream_t st);
    void* ws_trav = nullptr;         // rows values (the product) + a change flag
    std::size_t ws_trav_bytes = 0;
    void build_left();
};

}  // namespace gmr
