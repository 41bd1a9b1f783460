// Host scheduler: grammar -> level schedule (RefCompat, Wave64, Resident).
//
// RefCompat reproduces the reference's build_schedule (gpu-engine/
// build_schedule.cpp:42-227) output exactly. The construction here is different:
// instead of a dense rep[R x (L+1)] slot table (O(L*R) memory, build_schedule.cpp
// :86-88) every rule owns a "chain" of consecutive slots, one per level from its
// own level up to the last level it is carried to, stored in one flat array of
// num_layered entries. Level k's slot list is built from level k-1's list in one
// pass, so time and memory are O(|R| + num_layered + |C|).
#include "build_schedule.h"

#include <algorithm>
#include <numeric>
#include <stdexcept>
#include <string>

using grammar::Grammar;
using grammar::Rule;

GPUSchedule BuiltSchedule::gpu_schedule() {
    GPUSchedule s;
    s.max_depth = max_depth;
    s.max_width = max_width;
    s.level_offsets = level_offsets.data();
    s.level_widths = level_widths.data();
    s.emit_offsets = emit_offsets.data();
    s.emit_pos = emit_pos.data();
    s.emit_row = emit_row.data();
    s.emit_total = static_cast<int>(emit_pos.size());
    s.term_emit_sym = term_emit_sym.data();
    s.term_emit_row = term_emit_row.data();
    s.num_term_emit = static_cast<int>(term_emit_sym.size());
    return s;
}

namespace {

void check_grammar(const Grammar& g, int rows) {
    if (g.alpha < 1) throw std::invalid_argument("grammar alpha must be >= 1");
    if (rows < 0) throw std::invalid_argument("rows must be >= 0");
    const long long alpha = g.alpha;
    for (std::size_t i = 0; i < g.rules.size(); ++i) {
        const Rule& r = g.rules[i];
        const long long lim = alpha + static_cast<long long>(i);  // defined symbols
        if (r.kind != grammar::BINARY && r.kind != grammar::RUNLEN)
            throw std::invalid_argument("rule " + std::to_string(i) + ": unknown kind");
        if (r.left < 0 || r.left >= lim)
            throw std::invalid_argument("rule " + std::to_string(i) + ": left child out of range");
        if (r.kind == grammar::BINARY && (r.right < 0 || r.right >= lim))
            throw std::invalid_argument("rule " + std::to_string(i) + ": right child out of range");
        if (r.kind == grammar::RUNLEN && r.rep < 1)
            throw std::invalid_argument("rule " + std::to_string(i) + ": run length < 1");
    }
    const long long nsym = alpha + static_cast<long long>(g.rules.size());
    long long row = 0;
    for (int c : g.C) {
        if (c < 0 || c >= nsym) throw std::invalid_argument("top sequence symbol out of range");
        if (c == 0) { ++row; continue; }
        if (row >= rows) throw std::invalid_argument("top sequence has more rows than `rows`");
    }
}

}  // namespace

BuiltSchedule build_schedule(const Grammar& g, int rows, int cols) {
    return build_schedule(g, rows, cols, ScheduleOptions{});
}

BuiltSchedule build_schedule(const Grammar& g, int rows, int cols, const ScheduleOptions& opt) {
    check_grammar(g, rows);
    BuiltSchedule out;
    out.alpha = g.alpha;
    out.rows = rows;
    out.cols = cols;
    out.layout = opt.layout;

    const int alpha = g.alpha;
    const int R = static_cast<int>(g.rules.size());
    out.num_raw_nt = static_cast<std::size_t>(R);
    const bool resident = opt.layout == ScheduleLayout::Resident;

    // Level of each rule: 1 + max level of its children (terminals are level 0).
    std::vector<int> lvl(R);
    int L = 0;
    for (int i = 0; i < R; ++i) {
        const Rule& r = g.rules[i];
        int a = r.left >= alpha ? lvl[r.left - alpha] : 0;
        int b = (r.kind == grammar::BINARY && r.right >= alpha) ? lvl[r.right - alpha] : 0;
        lvl[i] = 1 + std::max(a, b);
        L = std::max(L, lvl[i]);
    }
    out.max_depth = L;

    // Highest level of a rule that reads each symbol (top-sequence uses do not
    // count: a symbol of C is emitted at its own level).
    std::vector<int> maxref(lvl);
    for (int i = 0; i < R; ++i) {
        const Rule& r = g.rules[i];
        if (r.left >= alpha) maxref[r.left - alpha] = std::max(maxref[r.left - alpha], lvl[i]);
        if (r.kind == grammar::BINARY && r.right >= alpha)
            maxref[r.right - alpha] = std::max(maxref[r.right - alpha], lvl[i]);
    }

    // Last level each symbol is present in the frontier. A reader at level m
    // takes its children from frontier m-1, so the copy at level maxref is dead:
    // RefCompat keeps it (the reference does), Wave64 drops it, Resident never
    // copies at all.
    std::vector<int> last(R);
    for (int i = 0; i < R; ++i) {
        if (opt.layout == ScheduleLayout::RefCompat) last[i] = maxref[i];
        else if (opt.layout == ScheduleLayout::Wave64) last[i] = std::max(lvl[i], maxref[i] - 1);
        else last[i] = lvl[i];
    }
    std::vector<long long> chain(R + 1, 0);
    for (int i = 0; i < R; ++i) chain[i + 1] = chain[i] + (last[i] - lvl[i] + 1);
    const long long total = chain[R];
    if (total > 0x7fffffffLL) throw std::invalid_argument("layered rule count exceeds int32");
    std::vector<int> slot(static_cast<std::size_t>(total), -1);
    auto slot_at = [&](int sym, int k) -> int { return slot[chain[sym] + (k - lvl[sym])]; };

    // Rules grouped by level, index order inside a level (counting sort).
    std::vector<int> lstart(L + 2, 0);
    for (int i = 0; i < R; ++i) ++lstart[lvl[i] + 1];
    for (int k = 1; k <= L + 1; ++k) lstart[k] += lstart[k - 1];
    std::vector<int> by_level(R);
    {
        std::vector<int> fill(lstart.begin(), lstart.end());
        for (int i = 0; i < R; ++i) by_level[fill[lvl[i]]++] = i;
    }

    out.flattened_rules.reserve(static_cast<std::size_t>(total));
    out.level_offsets.assign(1, 0);
    out.level_widths.assign(1, 0);

    // Child reference as seen from level k (reading frontier k-1).
    auto child_ref = [&](int sym, int k) -> int {
        if (sym < alpha) return -sym - 1;
        return slot_at(sym - alpha, resident ? lvl[sym - alpha] : k - 1);
    };

    std::vector<int> prev_list, cur_list, own;
    int max_width = 0;
    for (int k = 1; k <= L; ++k) {
        own.assign(by_level.begin() + lstart[k], by_level.begin() + lstart[k + 1]);
        if (opt.layout != ScheduleLayout::RefCompat) {
            // Order a level's own rules by where their left child lives, so
            // neighbouring lanes gather from neighbouring addresses.
            auto key = [&](int i) -> long long {
                int c = g.rules[i].left;
                if (c < alpha) return c;
                return static_cast<long long>(alpha) +
                       slot_at(c - alpha, resident ? lvl[c - alpha] : k - 1);
            };
            std::vector<std::pair<long long, int>> keyed(own.size());
            for (std::size_t j = 0; j < own.size(); ++j) keyed[j] = {key(own[j]), own[j]};
            std::sort(keyed.begin(), keyed.end());
            for (std::size_t j = 0; j < own.size(); ++j) own[j] = keyed[j].second;
        }
        cur_list.clear();
        const int base = resident ? static_cast<int>(out.flattened_rules.size()) : 0;
        auto push_own = [&]() {
            for (int i : own) {
                const Rule& r = g.rules[i];
                LayeredRule lr;
                lr.left = child_ref(r.left, k);
                if (r.kind == grammar::RUNLEN) {
                    lr.right = -1;
                    lr.coeff = static_cast<float>(r.rep);
                } else {
                    lr.right = child_ref(r.right, k);
                    lr.coeff = 1.0f;
                }
                slot[chain[i]] = base + static_cast<int>(cur_list.size());
                cur_list.push_back(i);
                out.flattened_rules.push_back(lr);
            }
        };
        auto push_carried = [&]() {
            for (std::size_t j = 0; j < prev_list.size(); ++j) {
                int s = prev_list[j];
                if (last[s] < k) continue;
                LayeredRule lr;
                lr.left = static_cast<int>(j);  // slot in frontier k-1
                lr.right = -1;
                lr.coeff = 1.0f;
                slot[chain[s] + (k - lvl[s])] = static_cast<int>(cur_list.size());
                cur_list.push_back(s);
                out.flattened_rules.push_back(lr);
            }
        };
        if (opt.layout == ScheduleLayout::RefCompat) {
            push_own();
            push_carried();
        } else if (opt.layout == ScheduleLayout::Wave64) {
            push_carried();
            push_own();
        } else {
            push_own();
        }
        const int w = static_cast<int>(cur_list.size());
        out.level_widths.push_back(w);
        out.level_offsets.push_back(static_cast<int>(out.flattened_rules.size()));
        max_width = std::max(max_width, w);
        std::swap(prev_list, cur_list);
    }
    out.max_width = max_width;
    out.num_layered = out.flattened_rules.size();
    out.num_passthrough = out.num_layered - out.num_raw_nt;

    // Emission lists. Scanning C left to right visits rows in order, so every
    // level's list comes out sorted by row.
    std::vector<int> ecount(L + 2, 0);
    for (int c : g.C)
        if (c >= alpha) ++ecount[lvl[c - alpha] + 1];
    for (int k = 1; k <= L + 1; ++k) ecount[k] += ecount[k - 1];
    // ecount[k] = emissions of levels < k; emit_offsets[k] = emissions of levels <= k.
    out.emit_offsets.resize(L + 1);
    for (int k = 0; k <= L; ++k) out.emit_offsets[k] = ecount[k + 1];
    out.emit_pos.assign(ecount[L + 1], 0);
    out.emit_row.assign(ecount[L + 1], 0);
    {
        std::vector<int> fill(ecount.begin(), ecount.end());
        int row = 0;
        for (int c : g.C) {
            if (c == 0) { ++row; continue; }
            if (c < alpha) {
                out.term_emit_sym.push_back(c);
                out.term_emit_row.push_back(row);
                continue;
            }
            int n = c - alpha;
            int e = fill[lvl[n]]++;
            out.emit_pos[e] = slot[chain[n]];
            out.emit_row[e] = row;
        }
    }

    // Terminal compaction: keep the referenced terminals, renumbered in
    // increasing original id; terminal 0 (the semiring zero) keeps id 0.
    {
        std::vector<int> remap(static_cast<std::size_t>(alpha), -1);
        remap[0] = 0;
        for (const LayeredRule& lr : out.flattened_rules) {
            if (lr.left < 0) remap[-lr.left - 1] = 0;
            if (lr.right < 0) remap[-lr.right - 1] = 0;
        }
        for (int s : out.term_emit_sym) remap[s] = 0;
        out.term_orig.assign(1, 0);
        for (int t = 1; t < alpha; ++t) {
            if (remap[t] == 0) {
                remap[t] = static_cast<int>(out.term_orig.size());
                out.term_orig.push_back(t);
            }
        }
        for (LayeredRule& lr : out.flattened_rules) {
            if (lr.left < 0) lr.left = -remap[-lr.left - 1] - 1;
            if (lr.right < 0) lr.right = -remap[-lr.right - 1] - 1;
        }
        for (int& s : out.term_emit_sym) s = remap[s];
        out.alpha = static_cast<int>(out.term_orig.size());
    }

    if (opt.keep_nested) {
        out.layered_levels.assign(L + 1, {});
        for (int k = 1; k <= L; ++k) {
            auto first = out.flattened_rules.begin() + out.level_offsets[k - 1];
            auto lastr = out.flattened_rules.begin() + out.level_offsets[k];
            out.layered_levels[k].assign(first, lastr);
        }
        out.emit_by_level.assign(L + 1, {});
        for (int k = 1; k <= L; ++k) {
            auto& v = out.emit_by_level[k];
            v.reserve(out.emit_offsets[k] - out.emit_offsets[k - 1]);
            for (int e = out.emit_offsets[k - 1]; e < out.emit_offsets[k]; ++e)
                v.emplace_back(out.emit_pos[e], out.emit_row[e]);
        }
        out.term_emit.resize(out.term_emit_sym.size());
        for (std::size_t e = 0; e < out.term_emit_sym.size(); ++e)
            out.term_emit[e] = {out.term_emit_sym[e], out.term_emit_row[e]};
    }
    return out;
}
