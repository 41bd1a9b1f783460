// extern "C" surface of the host library (declared in gmr_host.h).
#include <cstring>
#include <exception>
#include <new>
#include <stdexcept>
#include <string>

#include "build_schedule.h"
#include "gmr_host.h"
#include "host.hpp"

struct gmr_matrix {
    gmr::Csrv m;
};
struct gmr_grammar {
    grammar::Grammar g;
    std::vector<double> V;
};
struct gmr_schedule {
    BuiltSchedule s;
};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return GMR_OK;
    } catch (const std::bad_alloc&) {
        return fail(GMR_ERR_NOMEM, "host allocation failed");
    } catch (const std::invalid_argument& e) {
        return fail(GMR_ERR_INVALID, e.what());
    } catch (const std::runtime_error& e) {
        return fail(GMR_ERR_IO, e.what());
    } catch (const std::exception& e) {
        return fail(GMR_ERR_INTERNAL, e.what());
    }
}

}  // namespace

extern "C" {

const char* gmr_last_error(void) { return g_err.c_str(); }

int gmr_matrix_haplotype(uint64_t seed, int rows, int cols, int founders, double density,
                         double seg_mean, double flip, gmr_matrix** out) {
    if (!out) return fail(GMR_ERR_INVALID, "out is NULL");
    return guarded([&] {
        auto* h = new gmr_matrix;
        h->m = gmr::gen_haplotype(seed, rows, cols, founders, density, seg_mean, flip);
        *out = h;
    });
}

int gmr_matrix_genotype(uint64_t seed, int rows, int cols, int founders, int classes,
                        double seg_mean, gmr_matrix** out) {
    if (!out) return fail(GMR_ERR_INVALID, "out is NULL");
    return guarded([&] {
        auto* h = new gmr_matrix;
        h->m = gmr::gen_genotype(seed, rows, cols, founders, classes, seg_mean);
        *out = h;
    });
}

int gmr_matrix_rmat(uint64_t seed, int scale, long long edges, double a, double b, double c,
                    int nweights, gmr_matrix** out) {
    if (!out) return fail(GMR_ERR_INVALID, "out is NULL");
    return guarded([&] {
        auto* h = new gmr_matrix;
        h->m = gmr::gen_rmat(seed, scale, edges, a, b, c, nweights);
        *out = h;
    });
}

int gmr_matrix_from_dense(const double* M, int rows, int cols, gmr_matrix** out) {
    if (!out || (!M && rows > 0 && cols > 0) || rows < 0 || cols < 1)
        return fail(GMR_ERR_INVALID, "bad dense matrix");
    return guarded([&] {
        auto* h = new gmr_matrix;
        h->m = gmr::csrv_from_dense(M, rows, cols);
        *out = h;
    });
}

int gmr_matrix_from_csrv(int rows, int cols, const int* seq, long long len, const double* V,
                         int nvals, gmr_matrix** out) {
    if (!out || rows < 0 || cols < 1 || len < 0 || nvals < 0 || (len > 0 && !seq) ||
        (nvals > 0 && !V))
        return fail(GMR_ERR_INVALID, "bad CSRV arguments");
    return guarded([&] {
        long long zeros = 0;
        const long long lim = static_cast<long long>(nvals) * cols + 1;
        for (long long i = 0; i < len; ++i) {
            if (seq[i] < 0 || seq[i] >= lim) throw std::invalid_argument("CSRV terminal out of range");
            zeros += seq[i] == 0;
        }
        if (zeros != rows) throw std::invalid_argument("CSRV row count does not match rows");
        auto* h = new gmr_matrix;
        h->m.rows = rows;
        h->m.cols = cols;
        h->m.seq.assign(seq, seq + len);
        h->m.V.assign(V, V + nvals);
        *out = h;
    });
}

int gmr_matrix_row_block(const gmr_matrix* m, int r0, int r1, gmr_matrix** out) {
    if (!m || !out || r0 < 0 || r1 < r0 || r1 > m->m.rows) return fail(GMR_ERR_INVALID, "bad row block");
    return guarded([&] {
        // Row r's terminals start after the r-th 0 of the CSRV string (every
        // row, empty ones too, ends with its own 0).
        const auto& seq = m->m.seq;
        std::size_t a = 0, b = 0;
        int zeros = 0;
        for (std::size_t i = 0; i < seq.size() && zeros < r1; ++i) {
            if (seq[i] != 0) continue;
            ++zeros;
            if (zeros == r0) a = i + 1;
            if (zeros == r1) b = i + 1;
        }
        auto* h = new gmr_matrix;
        h->m.rows = r1 - r0;
        h->m.cols = m->m.cols;
        h->m.seq.assign(seq.begin() + static_cast<long>(a), seq.begin() + static_cast<long>(b));
        h->m.V = m->m.V;
        *out = h;
    });
}

int gmr_matrix_info(const gmr_matrix* m, int* rows, int* cols, int* nvals, int* alpha,
                    long long* len) {
    if (!m) return fail(GMR_ERR_INVALID, "matrix is NULL");
    if (rows) *rows = m->m.rows;
    if (cols) *cols = m->m.cols;
    if (nvals) *nvals = static_cast<int>(m->m.V.size());
    if (alpha) *alpha = m->m.alpha();
    if (len) *len = static_cast<long long>(m->m.seq.size());
    return GMR_OK;
}

int gmr_matrix_copy(const gmr_matrix* m, int* seq, double* V) {
    if (!m) return fail(GMR_ERR_INVALID, "matrix is NULL");
    if (seq) std::memcpy(seq, m->m.seq.data(), m->m.seq.size() * sizeof(int));
    if (V) std::memcpy(V, m->m.V.data(), m->m.V.size() * sizeof(double));
    return GMR_OK;
}

void gmr_matrix_free(gmr_matrix* m) { delete m; }

int gmr_grammar_repair(const gmr_matrix* m, int max_rounds, int threads, gmr_grammar** out,
                       int* rounds, double* seconds) {
    if (!m || !out) return fail(GMR_ERR_INVALID, "NULL argument");
    return guarded([&] {
        gmr::RepairOptions opt;
        if (max_rounds > 0) opt.max_rounds = max_rounds;
        opt.threads = threads;
        gmr::RepairStats st;
        auto* h = new gmr_grammar;
        h->g = gmr::repair(m->m, opt, &st);
        h->V = m->m.V;
        if (rounds) *rounds = st.rounds;
        if (seconds) *seconds = st.seconds;
        *out = h;
    });
}

int gmr_grammar_from_arrays(int alpha, int nrules, const int* kind, const int* left,
                            const int* right, const int* rep, const int* C, long long clen,
                            const double* V, int nvals, gmr_grammar** out) {
    if (!out || nrules < 0 || clen < 0 || nvals < 0 || (nrules > 0 && (!left || !right)) ||
        (clen > 0 && !C) || (nvals > 0 && !V))
        return fail(GMR_ERR_INVALID, "bad grammar arrays");
    return guarded([&] {
        auto* h = new gmr_grammar;
        h->g.alpha = alpha;
        h->g.rules.resize(nrules);
        for (int i = 0; i < nrules; ++i) {
            h->g.rules[i].kind = kind ? kind[i] : grammar::BINARY;
            h->g.rules[i].left = left[i];
            h->g.rules[i].right = right[i];
            h->g.rules[i].rep = rep ? rep[i] : 1;
        }
        h->g.C.assign(C, C + clen);
        h->V.assign(V, V + nvals);
        *out = h;
    });
}

int gmr_grammar_read(const char* base, gmr_grammar** out) {
    if (!base || !out) return fail(GMR_ERR_INVALID, "NULL argument");
    return guarded([&] {
        auto* h = new gmr_grammar;
        try {
            h->g = gmr::read_grammar_files(base, &h->V);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int gmr_grammar_write(const gmr_grammar* g, const char* base) {
    if (!g || !base) return fail(GMR_ERR_INVALID, "NULL argument");
    return guarded([&] { gmr::write_grammar_files(base, g->g, g->V); });
}

int gmr_grammar_info(const gmr_grammar* g, int* alpha, int* nrules, long long* clen, int* nvals) {
    if (!g) return fail(GMR_ERR_INVALID, "grammar is NULL");
    if (alpha) *alpha = g->g.alpha;
    if (nrules) *nrules = static_cast<int>(g->g.rules.size());
    if (clen) *clen = static_cast<long long>(g->g.C.size());
    if (nvals) *nvals = static_cast<int>(g->V.size());
    return GMR_OK;
}

int gmr_grammar_copy(const gmr_grammar* g, int* kind, int* left, int* right, int* rep, int* C,
                     double* V) {
    if (!g) return fail(GMR_ERR_INVALID, "grammar is NULL");
    const auto& rs = g->g.rules;
    for (std::size_t i = 0; i < rs.size(); ++i) {
        if (kind) kind[i] = rs[i].kind;
        if (left) left[i] = rs[i].left;
        if (right) right[i] = rs[i].right;
        if (rep) rep[i] = rs[i].rep;
    }
    if (C) std::memcpy(C, g->g.C.data(), g->g.C.size() * sizeof(int));
    if (V) std::memcpy(V, g->V.data(), g->V.size() * sizeof(double));
    return GMR_OK;
}

int gmr_grammar_expand(const gmr_grammar* g, int* out, long long cap, long long* len) {
    if (!g || !len) return fail(GMR_ERR_INVALID, "NULL argument");
    return guarded([&] {
        std::vector<int> v = gmr::expand_grammar(g->g);
        *len = static_cast<long long>(v.size());
        if (out) {
            if (cap < *len) throw std::invalid_argument("expand: output buffer too small");
            std::memcpy(out, v.data(), v.size() * sizeof(int));
        }
    });
}

int gmr_grammar_inline_top(const gmr_grammar* g, int max_emits, gmr_grammar** out,
                           long long* n_inlined) {
    if (!g || !out) return fail(GMR_ERR_INVALID, "NULL argument");
    return guarded([&] {
        auto* h = new gmr_grammar;
        try {
            h->g = gmr::inline_top_rules(g->g, max_emits, n_inlined);
        } catch (...) {
            delete h;
            throw;
        }
        h->V = g->V;
        *out = h;
    });
}

void gmr_grammar_free(gmr_grammar* g) { delete g; }

int gmr_schedule_build(const gmr_grammar* g, int rows, int cols, int layout, gmr_schedule** out) {
    if (!g || !out) return fail(GMR_ERR_INVALID, "NULL argument");
    if (layout < GMR_LAYOUT_REFCOMPAT || layout > GMR_LAYOUT_RESIDENT)
        return fail(GMR_ERR_INVALID, "unknown layout");
    return guarded([&] {
        ScheduleOptions opt;
        opt.layout = static_cast<ScheduleLayout>(layout);
        opt.keep_nested = false;
        auto* h = new gmr_schedule;
        h->s = build_schedule(g->g, rows, cols, opt);
        *out = h;
    });
}

int gmr_schedule_info(const gmr_schedule* s, long long info[11]) {
    if (!s || !info) return fail(GMR_ERR_INVALID, "NULL argument");
    const BuiltSchedule& b = s->s;
    info[0] = b.max_depth;
    info[1] = b.max_width;
    info[2] = static_cast<long long>(b.num_raw_nt);
    info[3] = static_cast<long long>(b.num_layered);
    info[4] = static_cast<long long>(b.num_passthrough);
    info[5] = static_cast<long long>(b.emit_pos.size());
    info[6] = static_cast<long long>(b.term_emit_sym.size());
    info[7] = b.alpha;
    info[8] = b.rows;
    info[9] = b.cols;
    info[10] = static_cast<long long>(b.layout);
    return GMR_OK;
}

int gmr_schedule_copy(const gmr_schedule* s, LayeredRule* rules, int* level_offsets,
                      int* level_widths, int* emit_offsets, int* emit_pos, int* emit_row,
                      int* term_sym, int* term_row, int* term_orig) {
    if (!s) return fail(GMR_ERR_INVALID, "schedule is NULL");
    const BuiltSchedule& b = s->s;
    auto cp = [](void* dst, const auto& v) {
        if (dst) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    cp(rules, b.flattened_rules);
    cp(level_offsets, b.level_offsets);
    cp(level_widths, b.level_widths);
    cp(emit_offsets, b.emit_offsets);
    cp(emit_pos, b.emit_pos);
    cp(emit_row, b.emit_row);
    cp(term_sym, b.term_emit_sym);
    cp(term_row, b.term_emit_row);
    cp(term_orig, b.term_orig);
    return GMR_OK;
}

int gmr_schedule_cpu_sweep(const gmr_schedule* s, int semiring, const float* V, int nvals, const float* x, void* y,
                           int threads) {
    if (!s || !V || !x || !y) return fail(GMR_ERR_INVALID, "NULL argument");
    return guarded([&] {
        const BuiltSchedule& b = s->s;
        const int U = b.alpha;
        std::vector<int> vi(U > 0 ? U - 1 : 0), ci(vi.size());
        for (int q = 1; q < U; ++q) {
            const int t = b.term_orig[q];
            vi[q - 1] = (t - 1) / b.cols;
            ci[q - 1] = (t - 1) % b.cols;
            if (vi[q - 1] >= nvals) throw std::invalid_argument("a terminal's value index is beyond nvals");
        }
        gmr::cpu_layered_sweep(semiring, b, vi.data(), ci.data(), V, x, y, threads);
    });
}

int gmr_schedule_view(gmr_schedule* s, GPUSchedule* out) {
    if (!s || !out) return fail(GMR_ERR_INVALID, "NULL argument");
    *out = s->s.gpu_schedule();
    return GMR_OK;
}

void gmr_schedule_free(gmr_schedule* s) { delete s; }

}  // extern "C"

// Internal hooks for the device engine library (not part of gmr_host.h).
void gmr_host_set_error(const char* msg) { g_err = msg ? msg : ""; }
const BuiltSchedule* gmr_schedule_internal(const gmr_schedule* s) { return s ? &s->s : nullptr; }
