// The reference's two extern "C" entry points (gpu_engine.h) on top of the
// device-resident engine.
//
// Each call does what the reference's does (gpu-engine/gpu_engine.cu:135-587):
// take host buffers, run `iters` sweeps, time them with events, print the same
// lines, return the last y. The schedule it receives is the reference's
// double-buffered layout; before upload it is re-laid out for the engine:
// identity copies (pass-throughs: right child = the zero terminal, coeff 1) are
// resolved to the slot they copy, every remaining rule gets a slot in one
// resident value array, and the top sequence is regrouped by row so y is
// produced by one row gather with no atomics. The values are the same as the
// reference's sweep: an identity copy returns its input in every semiring (for
// Tropical any value >= 1e9 acts as +infinity either way).
#include <hip/hip_runtime.h>

#include <dlfcn.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "gmr.h"
#include "gpu_engine.h"

const char* gmr_engine_last_error_internal();

namespace {

[[noreturn]] void die(const std::string& what) {
    std::cerr << "GPU engine error: " << what << std::endl;
    std::exit(1);
}

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) die(std::string(what) + ": " + hipGetErrorString(e));
}

int semiring_from_env() {
    const char* s = std::getenv("SEMIRING");
    std::string v = s ? s : "plustimes";
    if (v == "boolean" || v == "bool") return GMR_BOOLEAN;
    if (v == "tropical" || v == "minplus") return GMR_TROPICAL;
    return GMR_PLUS_TIMES;
}

// Reference layout -> Resident layout (global slots, no identity copies).
struct Resident {
    std::vector<LayeredRule> rules;
    std::vector<int> level_off, emit_off, emit_pos, emit_row;
    std::vector<int> term_sym, term_row;  // terminal emissions (index into T, row)
    int L = 0, max_width = 0;
};

Resident to_resident(int total, const LayeredRule* r, const GPUSchedule& s) {
    Resident o;
    const int L = s.max_depth;
    if (L < 0) die("negative max_depth");
    std::vector<int> prev, cur;  // canonical global slot of each slot of level k-1 / k
    o.level_off.assign(1, 0);
    o.emit_off.assign(1, 0);
    for (int k = 1; k <= L; ++k) {
        const int lo = s.level_offsets[k - 1], hi = s.level_offsets[k];
        if (lo < 0 || hi < lo || hi > total) die("schedule level offsets out of range");
        cur.assign(hi - lo, -1);
        const int base = static_cast<int>(o.rules.size());
        auto canon = [&](int c) -> int {
            if (c < 0) return c;
            if (c >= static_cast<int>(prev.size())) die("rule child out of range");
            return prev[c];
        };
        for (int i = lo; i < hi; ++i) {
            const LayeredRule& x = r[i];
            if (x.right == -1 && x.coeff == 1.0f && x.left >= 0) {
                cur[i - lo] = canon(x.left);  // identity copy
            } else {
                LayeredRule y{canon(x.left), canon(x.right), x.coeff};
                cur[i - lo] = static_cast<int>(o.rules.size());
                o.rules.push_back(y);
            }
        }
        const int eb = s.emit_offsets[k - 1], ee = s.emit_offsets[k];
        if (eb < 0 || ee < eb || ee > s.emit_total) die("schedule emit offsets out of range");
        // Emissions of level k go to the level that computed the value.
        for (int e = eb; e < ee; ++e) {
            int p = s.emit_pos[e];
            if (p < 0 || p >= hi - lo) die("emission position out of range");
            o.emit_pos.push_back(cur[p]);
            o.emit_row.push_back(s.emit_row[e]);
        }
        const int w = static_cast<int>(o.rules.size()) - base;
        o.level_off.push_back(static_cast<int>(o.rules.size()));
        o.max_width = std::max(o.max_width, w);
        prev.swap(cur);
    }
    // An emission of an identity copy refers to a slot of an earlier level:
    // regroup emissions by the level that owns their slot.
    o.L = L;
    std::vector<int> owner_cnt(L + 2, 0);
    auto owner = [&](int slot) {
        int k = static_cast<int>(std::upper_bound(o.level_off.begin(), o.level_off.end(), slot) - o.level_off.begin());
        return k;  // level_off[k-1] <= slot < level_off[k]
    };
    std::vector<int> lv(o.emit_pos.size());
    for (std::size_t e = 0; e < o.emit_pos.size(); ++e) {
        lv[e] = owner(o.emit_pos[e]);
        ++owner_cnt[lv[e]];
    }
    o.emit_off.assign(L + 1, 0);
    for (int k = 1; k <= L; ++k) o.emit_off[k] = o.emit_off[k - 1] + owner_cnt[k];
    std::vector<int> fill(o.emit_off.begin(), o.emit_off.end() - 1), ep(o.emit_pos.size()), er(ep.size());
    for (std::size_t e = 0; e < o.emit_pos.size(); ++e) {
        int k = lv[e];
        int d = fill[k - 1]++;
        ep[d] = o.emit_pos[e];
        er[d] = o.emit_row[e];
    }
    o.emit_pos.swap(ep);
    o.emit_row.swap(er);
    return o;
}

// Top rules written out in the top sequence: a rule that no rule references,
// that C uses 1..max_emits times, with coeff 1 and no zero-terminal child, is
// removed, and each of its emissions becomes the emissions of its two children
// (a terminal child: a terminal emission). A row's result is unchanged in
// every semiring, whose (+) is associative and commutative: the emission of
// N -> A B added v(A) (+) v(B), the two emissions add the same. It trades a
// node update (read A and B, write N) and e reads of N for 2e reads of the
// children in the row gather; the thresholds are those the autotuner picks on
// the 500,000 x 10,000 genotype grammar (one vector: 8; two to four and 64 or
// more vectors: 2; between: 0, the grammar as the reference scheduled it).
void inline_top(Resident& o, int max_emits) {
    const int n = static_cast<int>(o.rules.size());
    if (max_emits <= 0 || n == 0) return;
    std::vector<int> refs(n, 0), emits(n, 0);
    for (const LayeredRule& r : o.rules) {
        if (r.left >= 0) ++refs[r.left];
        if (r.right >= 0) ++refs[r.right];
    }
    for (int p : o.emit_pos) ++emits[p];
    std::vector<char> sel(n, 0);
    std::vector<int> new_id(n, -1);
    int kept = 0;
    for (int i = 0; i < n; ++i) {
        const LayeredRule& r = o.rules[i];
        sel[i] = refs[i] == 0 && emits[i] >= 1 && emits[i] <= max_emits && r.coeff == 1.0f && r.left != -1 &&
                 r.right != -1;
        if (!sel[i]) new_id[i] = kept++;
    }
    if (kept == n) return;
    // levels of the old slots; children stay (a referenced rule is kept)
    std::vector<int> level(n);
    for (int k = 1; k <= o.L; ++k)
        for (int i = o.level_off[k - 1]; i < o.level_off[k]; ++i) level[i] = k;
    Resident t;
    t.rules.reserve(kept);
    t.level_off.assign(o.L + 1, 0);
    for (int i = 0; i < n; ++i) {
        if (sel[i]) continue;
        LayeredRule r = o.rules[i];
        if (r.left >= 0) r.left = new_id[r.left];
        if (r.right >= 0) r.right = new_id[r.right];
        t.rules.push_back(r);
        ++t.level_off[level[i]];
    }
    for (int k = 1; k <= o.L; ++k) t.level_off[k] += t.level_off[k - 1];
    // emissions, grouped by the level that owns their slot
    std::vector<int> ep, er, lv;
    ep.reserve(o.emit_pos.size() + o.emit_pos.size() / 4);
    auto emit = [&](int c, int row) {
        if (c >= 0) {
            ep.push_back(new_id[c]);
            er.push_back(row);
            lv.push_back(level[c]);
        } else {
            o.term_sym.push_back(-c - 1);
            o.term_row.push_back(row);
        }
    };
    for (std::size_t e = 0; e < o.emit_pos.size(); ++e) {
        const int p = o.emit_pos[e], row = o.emit_row[e];
        if (sel[p]) {
            emit(o.rules[p].left, row);
            emit(o.rules[p].right, row);
        } else {
            emit(p, row);
        }
    }
    t.emit_off.assign(o.L + 1, 0);
    for (int k : lv) ++t.emit_off[k];
    for (int k = 1; k <= o.L; ++k) t.emit_off[k] += t.emit_off[k - 1];
    std::vector<int> fill(t.emit_off.begin(), t.emit_off.end() - 1);
    t.emit_pos.resize(ep.size());
    t.emit_row.resize(ep.size());
    for (std::size_t e = 0; e < ep.size(); ++e) {
        const int d = fill[lv[e] - 1]++;
        t.emit_pos[d] = ep[e];
        t.emit_row[d] = er[e];
    }
    // only top levels can empty (a rule of level k + 1 references one of level k)
    t.L = o.L;
    while (t.L > 0 && t.level_off[t.L] == t.level_off[t.L - 1]) --t.L;
    t.level_off.resize(t.L + 1);
    t.emit_off.resize(t.L + 1);
    for (int k = 1; k <= t.L; ++k) t.max_width = std::max(t.max_width, t.level_off[k] - t.level_off[k - 1]);
    t.term_sym.swap(o.term_sym);
    t.term_row.swap(o.term_row);
    o = std::move(t);
}

// Energy per vector, as the reference's NVML loop (gpu_engine.cu:286-330):
// sweeps run back to back between two updates of the amdsmi energy counter at
// least GMR_ENERGY_SECONDS apart (default 1.5 s; the counter advances about
// once a second). libgmr_vendor.so reads the counter; it is loaded on first
// use, so this library needs neither amdsmi nor rocSPARSE. Returns false --
// the caller prints -1, as the reference does -- when the counter cannot be
// read or does not advance, or when GMR_ENERGY_SECONDS is 0.
// The drop-in's sweep: one hipGraph of the per-level launches, y by row gather.
gmr_run_opts gather_opts() {
    gmr_run_opts o{};
    o.variant = GMR_VARIANT_GRAPH;
    o.emit = GMR_EMIT_GATHER;
    return o;
}

// The schedule of the timed loop. Many vectors: gather_opts(). One vector: the
// fastest of three, each run once as set-up (graph capture, gather lists) and
// then timed over three sweeps: gather_opts() as is; with the top levels fused
// into one workgroup and small levels two to a launch; and with, in addition,
// the values gathered from LDS-staged ranges of 40,960 (160 KiB) in the rows'
// similarity order. A schedule the engine refuses for this grammar or device
// is skipped. On the 500,000 x 10,000 genotype grammar as built they take
// 665 / 627 / 456 us per vector; on small grammars the fused tail wins.
gmr_run_opts choose_opts(gmr_engine* eng, int sr, const float* dX, int B, void* dY, hipEvent_t e0, hipEvent_t e1) {
    gmr_run_opts cand[3] = {gather_opts(), gather_opts(), gather_opts()};
    const int n = B == 1 ? 3 : 1;
    cand[1].flags = GMR_RUN_FUSED_TAIL | GMR_RUN_PAIRED_LEVELS;
    cand[2].flags = cand[1].flags | GMR_RUN_LDS_RANGES | GMR_RUN_ROW_ORDER;
    cand[2].lds_values = 40960;
    cand[2].chunks = 4;
    int best = -1;
    float best_ms = 0.0f;
    for (int i = 0; i < n; ++i) {
        const int rc = gmr_engine_run(eng, sr, dX, B, dY, &cand[i], nullptr);
        if (rc == GMR_ERR_INVALID && i > 0) continue;
        if (rc != GMR_OK) die(gmr_engine_last_error_internal());
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        if (n == 1) return cand[0];
        hip_check(hipEventRecord(e0, nullptr), "hipEventRecord");
        for (int r = 0; r < 3; ++r)
            if (gmr_engine_run(eng, sr, dX, B, dY, &cand[i], nullptr) != GMR_OK)
                die(gmr_engine_last_error_internal());
        hip_check(hipEventRecord(e1, nullptr), "hipEventRecord");
        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        if (best < 0 || ms < best_ms) {
            best = i;
            best_ms = ms;
        }
    }
    return cand[best];
}

bool energy_loop(gmr_engine* eng, int sr, const float* dX, int B, void* dY, const gmr_run_opts& o, double* mj,
                 double* mw) {
    double secs_min = 1.5;
    if (const char* e = std::getenv("GMR_ENERGY_SECONDS")) secs_min = std::atof(e);
    if (!(secs_min > 0.0)) return false;
    using EnergyFn = int (*)(int, double*, unsigned long long*);
    static EnergyFn fn = [] {
        void* h = dlopen("libgmr_vendor.so", RTLD_NOW | RTLD_LOCAL);
        return h ? reinterpret_cast<EnergyFn>(dlsym(h, "gmr_energy_uj")) : nullptr;
    }();
    int dev = 0;
    double e = 0.0, e0 = 0.0;
    unsigned long long ts = 0, ts_prev = 0, ts0 = 0;
    if (!fn || hipGetDevice(&dev) != hipSuccess || fn(dev, &e, &ts_prev) != GMR_OK) return false;
    const auto start = std::chrono::steady_clock::now();
    auto batch = [&] {
        for (int i = 0; i < 10; ++i)
            if (gmr_engine_run(eng, sr, dX, B, dY, &o, nullptr) != GMR_OK) die(gmr_engine_last_error_internal());
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count() <= 2 * secs_min + 5.0;
    };
    do {  // up to the next counter update
        if (!batch() || fn(dev, &e0, &ts0) != GMR_OK) return false;
    } while (ts0 == ts_prev);
    long count = 0;
    do {  // to the first update at least secs_min later
        count += 10;
        if (!batch() || fn(dev, &e, &ts) != GMR_OK) return false;
    } while (ts == ts0 || static_cast<double>(ts - ts0) * 1e-9 < secs_min);
    *mj = (e - e0) * 1e-3 / static_cast<double>(count * B);
    *mw = (e - e0) * 1e-3 / (static_cast<double>(ts - ts0) * 1e-9);
    return true;
}

void dropin(int rows, int cols, int alpha, int total, const LayeredRule* h_rules, const int* h_val,
            const int* h_col, const float* h_V, int nvals, const float* h_X, const GPUSchedule& s,
            float* h_Y, int B, int iters, bool batched) {
    const int sr = semiring_from_env();
    if (!batched) {
        std::cout << "[GPU Setup Debug] rows=" << rows << ", cols=" << cols << ", alpha=" << alpha
                  << ", total_rules=" << total << ", distinct_vals=" << nvals
                  << ", sched_max_depth=" << s.max_depth << ", sched_max_width=" << s.max_width
                  << ", emit_total=" << s.emit_total << ", num_term_emit=" << s.num_term_emit << std::endl;
    }
    // Reference analytic footprint (gpu_engine.cu:174-184), with B-wide buffers.
    const std::size_t vb = 4;
    const std::size_t mem_analytic = static_cast<std::size_t>(alpha) * vb * B + static_cast<std::size_t>(total) * sizeof(LayeredRule) +
                                     static_cast<std::size_t>(alpha > 0 ? alpha - 1 : 0) * sizeof(int) * 2 +
                                     static_cast<std::size_t>(nvals) * sizeof(float) + static_cast<std::size_t>(cols) * vb * B +
                                     static_cast<std::size_t>(rows) * vb * B + static_cast<std::size_t>(s.max_width) * vb * 2 * B +
                                     static_cast<std::size_t>(s.emit_total) * sizeof(int) * 2 +
                                     static_cast<std::size_t>(s.num_term_emit) * sizeof(int) * 2;
    std::size_t free0 = 0, tot0 = 0;
    hip_check(hipMemGetInfo(&free0, &tot0), "hipMemGetInfo");

    Resident res = to_resident(total, h_rules, s);
    res.term_sym.assign(s.term_emit_sym, s.term_emit_sym + s.num_term_emit);
    res.term_row.assign(s.term_emit_row, s.term_emit_row + s.num_term_emit);
    inline_top(res, B == 1 ? 8 : (B <= 4 || B >= 64) ? 2 : 0);
    GPUSchedule rs{};
    rs.max_depth = res.L;
    rs.max_width = res.max_width;
    std::vector<int> widths(res.L + 1, 0);
    for (int k = 1; k <= res.L; ++k) widths[k] = res.level_off[k] - res.level_off[k - 1];
    rs.level_offsets = res.level_off.data();
    rs.level_widths = widths.data();
    rs.emit_offsets = res.emit_off.data();
    rs.emit_pos = res.emit_pos.data();
    rs.emit_row = res.emit_row.data();
    rs.emit_total = static_cast<int>(res.emit_pos.size());
    rs.term_emit_sym = res.term_sym.data();
    rs.term_emit_row = res.term_row.data();
    rs.num_term_emit = static_cast<int>(res.term_sym.size());
    gmr_engine* eng = nullptr;
    if (gmr_engine_create(rows, cols, alpha, static_cast<int>(res.rules.size()), res.rules.data(), h_val, h_col, h_V,
                          nvals, &rs, GMR_LAYOUT_RESIDENT, &eng) != GMR_OK)
        die(gmr_engine_last_error_internal());

    float* dX = nullptr;
    void* dY = nullptr;
    hip_check(hipMalloc(&dX, sizeof(float) * static_cast<std::size_t>(cols) * B), "hipMalloc(X)");
    hip_check(hipMalloc(&dY, vb * static_cast<std::size_t>(rows > 0 ? rows : 1) * B), "hipMalloc(Y)");
    hip_check(hipMemcpy(dX, h_X, sizeof(float) * static_cast<std::size_t>(cols) * B, hipMemcpyHostToDevice), "copy X");
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    // Untimed sweeps first: they capture the graph, build the gather lists and
    // (one vector) choose the schedule, set-up that the reference does before
    // its timed loop (uploads).
    const gmr_run_opts o = iters > 0 ? choose_opts(eng, sr, dX, B, dY, e0, e1) : gather_opts();
    hip_check(hipEventRecord(e0, nullptr), "hipEventRecord");
    for (int it = 0; it < iters; ++it) {
        if (gmr_engine_run(eng, sr, dX, B, dY, &o, nullptr) != GMR_OK) die(gmr_engine_last_error_internal());
    }
    hip_check(hipEventRecord(e1, nullptr), "hipEventRecord");
    hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    if (batched) {
        double per_batch = static_cast<double>(ms) / iters;
        std::cout << "GPU batched (B=" << B << "): " << per_batch << " ms/batch, " << per_batch / B
                  << " ms/vector over " << iters << " batches." << std::endl;
    } else {
        std::cout << "GPU kernel: " << ms << " ms total over " << iters << " vectors = " << (ms / iters)
                  << " ms/vector." << std::endl;
    }
    std::size_t free1 = 0, tot1 = 0;
    hip_check(hipMemGetInfo(&free1, &tot1), "hipMemGetInfo");
    if (!batched) {
        std::cout << "MEM analytic bytes: " << mem_analytic << std::endl;
        std::cout << "MEM peak bytes: " << (free0 > free1 ? free0 - free1 : 0) << std::endl;
        double mj = -1.0, mw = -1.0;
        if (iters <= 0 || !energy_loop(eng, sr, dX, B, dY, o, &mj, &mw)) mj = mw = -1.0;
        std::cout << "ENERGY mJ/vector: " << mj << std::endl;
        std::cout << "ENERGY avg power mW: " << mw << std::endl;
    }
    if (iters > 0 && rows > 0)
        hip_check(hipMemcpy(h_Y, dY, vb * static_cast<std::size_t>(rows) * B, hipMemcpyDeviceToHost), "copy Y");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(dX);
    (void)hipFree(dY);
    gmr_engine_destroy(eng);
}

}  // namespace

extern "C" void run_gpu_right_multiply_layered(int rows, int cols, int alpha, int total_layered_rules,
                                               const LayeredRule* h_rules, const int* h_val_indices,
                                               const int* h_col_indices, const float* h_V,
                                               int num_distinct_vals, const float* h_x,
                                               const GPUSchedule& h_sched, float* h_y, int iters) {
    dropin(rows, cols, alpha, total_layered_rules, h_rules, h_val_indices, h_col_indices, h_V, num_distinct_vals,
           h_x, h_sched, h_y, 1, iters, false);
}

extern "C" void run_gpu_right_multiply_batched(int rows, int cols, int alpha, int total_layered_rules,
                                               const LayeredRule* h_rules, const int* h_val_indices,
                                               const int* h_col_indices, const float* h_V,
                                               int num_distinct_vals, const float* h_X,
                                               const GPUSchedule& h_sched, float* h_Y, int B, int iters) {
    if (B < 1) die("B must be >= 1");
    dropin(rows, cols, alpha, total_layered_rules, h_rules, h_val_indices, h_col_indices, h_V, num_distinct_vals,
           h_X, h_sched, h_Y, B, iters, true);
}
