// Rewrite of a grammar that trades stored rules for top-sequence symbols.
//
// A binary rule N -> A B that no rule references and that the top sequence C
// uses e times costs the sweep one node update (read A and B, write N) plus e
// reads of N by the row gather. Replacing each of its occurrences in C by
// "A B" costs 2e reads instead: for e <= 2 that is no more, and the reads move
// from the node update into the row gather, which streams them at a higher
// fraction of HBM bandwidth. The stored size is unchanged for e = 2 (two
// symbols leave the rule table, two enter C) and drops by one for e = 1, and
// the expansion of C -- the matrix -- is the same string.
#include <stdexcept>

#include "host.hpp"

namespace gmr {

grammar::Grammar inline_top_rules(const grammar::Grammar& g, int max_emits, long long* n_inlined) {
    const int alpha = g.alpha;
    const long long n = static_cast<long long>(g.rules.size());
    std::vector<int> refs(n, 0), emits(n, 0);
    auto child = [&](int s) {
        if (s < 1 || s - alpha >= n) throw std::invalid_argument("inline_top: rule child out of range");
        if (s >= alpha) ++refs[s - alpha];
    };
    for (const auto& r : g.rules) {
        child(r.left);
        if (r.kind == grammar::BINARY) child(r.right);
    }
    for (int s : g.C) {
        if (s < 0 || s - alpha >= n) throw std::invalid_argument("inline_top: symbol out of range in C");
        if (s >= alpha) ++emits[s - alpha];
    }
    // Selected rules are referenced by no rule, so none of them is the child
    // of another: one pass over C replaces every occurrence.
    std::vector<char> sel(n, 0);
    std::vector<int> new_id(n, -1);
    long long kept = 0, added = 0, count = 0;
    for (long long i = 0; i < n; ++i) {
        const auto& r = g.rules[i];
        sel[i] = max_emits > 0 && r.kind == grammar::BINARY && refs[i] == 0 && emits[i] >= 1 &&
                 emits[i] <= max_emits;
        if (sel[i]) {
            ++count;
            added += emits[i];
        } else {
            new_id[i] = static_cast<int>(kept++);
        }
    }
    auto map = [&](int s) { return s < alpha ? s : alpha + new_id[s - alpha]; };

    grammar::Grammar out;
    out.alpha = alpha;
    out.rules.reserve(kept);
    for (long long i = 0; i < n; ++i) {
        if (sel[i]) continue;
        grammar::Rule r = g.rules[i];
        r.left = map(r.left);
        if (r.kind == grammar::BINARY) r.right = map(r.right);
        out.rules.push_back(r);
    }
    out.C.reserve(g.C.size() + added);
    for (int s : g.C) {
        if (s >= alpha && sel[s - alpha]) {
            const auto& r = g.rules[s - alpha];
            out.C.push_back(map(r.left));
            out.C.push_back(map(r.right));
        } else {
            out.C.push_back(map(s));
        }
    }
    if (n_inlined) *n_inlined = count;
    return out;
}

}  // namespace gmr
