// ref_shim.cpp -- C entry points around the reference's own build_schedule(),
// compiled together with the reference's gpu-engine/build_schedule.cpp (read in
// place under the reference checkout, never copied) into
// oracle/_ref/libref_schedule.so. TEST INFRASTRUCTURE ONLY: it lets the tests
// compare the oracle's and the engine's layouts with the reference's real one.
#include <cstring>
#include <vector>

#include "build_schedule.h"  // the reference header (reference gpu-engine/build_schedule.h)
#include "grammar.h"

extern "C" {

void* ref_build(int alpha, int nrules, const int* kind, const int* left, const int* right,
                const int* rep, const int* C, long long clen, int rows, int cols) {
    grammar::Grammar g;
    g.alpha = alpha;
    g.rules.resize(nrules);
    for (int i = 0; i < nrules; ++i) {
        g.rules[i].kind = kind ? kind[i] : grammar::BINARY;
        g.rules[i].left = left[i];
        g.rules[i].right = right[i];
        g.rules[i].rep = rep ? rep[i] : 1;
    }
    g.C.assign(C, C + clen);
    return new BuiltSchedule(build_schedule(g, rows, cols));
}

// info: max_depth, max_width, num_raw_nt, num_layered, num_passthrough,
// emit_total, num_term_emit, alpha (compacted)
void ref_info(void* h, long long* info) {
    auto* b = static_cast<BuiltSchedule*>(h);
    info[0] = b->max_depth;
    info[1] = b->max_width;
    info[2] = (long long)b->num_raw_nt;
    info[3] = (long long)b->num_layered;
    info[4] = (long long)b->num_passthrough;
    info[5] = (long long)b->emit_pos.size();
    info[6] = (long long)b->term_emit_sym.size();
    info[7] = b->alpha;
}

void ref_copy(void* h, int* left, int* right, float* coeff, int* level_offsets, int* level_widths,
              int* emit_offsets, int* emit_pos, int* emit_row, int* term_sym, int* term_row,
              int* term_orig) {
    auto* b = static_cast<BuiltSchedule*>(h);
    for (std::size_t i = 0; i < b->flattened_rules.size(); ++i) {
        left[i] = b->flattened_rules[i].left;
        right[i] = b->flattened_rules[i].right;
        coeff[i] = b->flattened_rules[i].coeff;
    }
    auto cp = [](int* d, const std::vector<int>& v) {
        if (!v.empty()) std::memcpy(d, v.data(), v.size() * sizeof(int));
    };
    cp(level_offsets, b->level_offsets);
    cp(level_widths, b->level_widths);
    cp(emit_offsets, b->emit_offsets);
    cp(emit_pos, b->emit_pos);
    cp(emit_row, b->emit_row);
    cp(term_sym, b->term_emit_sym);
    cp(term_row, b->term_emit_row);
    cp(term_orig, b->term_orig);
}

void ref_free(void* h) { delete static_cast<BuiltSchedule*>(h); }

}  // extern "C"
