// RePair grammar builder (replaces the absent mm-repair `matrepair` call sites,
// reference tests/run_integration_tests.py:130-139 and reproduce.sh:135).
//
// RePair (Larsson and Moffat, "Off-line dictionary-based compression", 2000)
// repeatedly replaces the most frequent adjacent pair by a new rule. This
// builder runs the replacement in rounds so it scales to 5e8-symbol strings on
// many cores:
//   1. count every adjacent pair inside a row (runs a^n count floor(n/2));
//   2. rank the pairs seen at least twice by (count desc, pair asc);
//   3. scan each row left to right; the pair at i is taken unless the pair at
//      i+1 ranks strictly higher (then i is left alone and the scan moves on);
//   4. a ranked pair taken fewer than twice is not turned into a rule (its
//      occurrence stays two symbols); the others become rules in rank order.
// Rounds repeat until no pair occurs twice. Pairs never contain the row
// delimiter 0, rules are numbered in creation order (topological), and the
// result does not depend on the thread count.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <stdexcept>

#include "host.hpp"
#include "rng.hpp"

namespace gmr {

namespace {

constexpr int kPartBits = 8;
constexpr int kParts = 1 << kPartBits;

inline uint64_t pair_key(int a, int b) {
    return (static_cast<uint64_t>(static_cast<uint32_t>(a)) << 32) | static_cast<uint32_t>(b);
}

// Open-addressing count table for one partition (key 0 = empty; pair keys are
// never 0 because symbols in a row are >= 1).
struct CountTable {
    std::vector<uint64_t> keys;
    std::vector<uint32_t> counts;
    std::size_t used = 0;
    void init(std::size_t cap) {
        keys.assign(cap, 0);
        counts.assign(cap, 0);
        used = 0;
    }
    void grow() {
        std::vector<uint64_t> ok;
        std::vector<uint32_t> oc;
        ok.swap(keys);
        oc.swap(counts);
        init(ok.size() * 2);
        for (std::size_t i = 0; i < ok.size(); ++i)
            if (ok[i]) add(ok[i], oc[i]);
    }
    inline void add(uint64_t k, uint32_t n) {
        if ((used + 1) * 2 > keys.size()) grow();
        const std::size_t mask = keys.size() - 1;
        std::size_t h = (mix64(k) >> kPartBits) & mask;
        while (true) {
            if (keys[h] == k) { counts[h] += n; return; }
            if (keys[h] == 0) { keys[h] = k; counts[h] = n; ++used; return; }
            h = (h + 1) & mask;
        }
    }
};

// Read-only pair -> rank table.
struct RankTable {
    std::vector<uint64_t> keys;
    std::vector<int> rank;
    std::size_t mask = 0;
    void build(const std::vector<std::pair<uint32_t, uint64_t>>& ranked) {
        std::size_t cap = 16;
        while (cap < ranked.size() * 2 + 16) cap <<= 1;
        keys.assign(cap, 0);
        rank.assign(cap, -1);
        mask = cap - 1;
        for (std::size_t r = 0; r < ranked.size(); ++r) {
            uint64_t k = ranked[r].second;
            std::size_t h = mix64(k) & mask;
            while (keys[h] != 0) h = (h + 1) & mask;
            keys[h] = k;
            rank[h] = static_cast<int>(r);
        }
    }
    inline int find(uint64_t k) const {
        std::size_t h = mix64(k) & mask;
        while (true) {
            if (keys[h] == k) return rank[h];
            if (keys[h] == 0) return -1;
            h = (h + 1) & mask;
        }
    }
};

}  // namespace

grammar::Grammar repair(const Csrv& m, const RepairOptions& opt, RepairStats* stats) {
    auto t0 = std::chrono::steady_clock::now();
    if (opt.threads > 0) omp_set_num_threads(opt.threads);
    grammar::Grammar g;
    g.alpha = m.alpha();

    // Working rows (delimiters stripped): row r is cur[start[r], start[r]+len[r]).
    std::vector<long long> start, len;
    {
        long long s = 0;
        for (long long i = 0; i < static_cast<long long>(m.seq.size()); ++i) {
            if (m.seq[i] == 0) {
                start.push_back(s);
                len.push_back(i - s);
                s = i + 1;
            }
        }
        if (s < static_cast<long long>(m.seq.size())) {  // trailing row without delimiter
            start.push_back(s);
            len.push_back(static_cast<long long>(m.seq.size()) - s);
        }
    }
    const long long nrows = static_cast<long long>(start.size());
    const bool trailing_open = !m.seq.empty() && m.seq.back() != 0;
    std::vector<int> cur(m.seq);

    std::vector<CountTable> tables(kParts);
    std::vector<int> uses;
    std::vector<int> final_sym;
    int round = 0;
    const int nthreads = omp_get_max_threads();
    const long long batch_rows_target = 1 << 14;

    for (; round < opt.max_rounds; ++round) {
        // ---- 1. pair counts, in row batches, partitioned by pair hash.
        for (auto& t : tables) t.init(1024);
        std::vector<std::vector<uint64_t>> buf(static_cast<std::size_t>(nthreads) * kParts);
        for (long long r0 = 0; r0 < nrows;) {
            // Batch of rows holding roughly 2^26 pairs.
            long long r1 = r0, pairs = 0;
            while (r1 < nrows && (pairs < (1LL << 26) || r1 - r0 < batch_rows_target / 64)) pairs += len[r1++];
#pragma omp parallel
            {
                const int t = omp_get_thread_num();
                auto* mine = &buf[static_cast<std::size_t>(t) * kParts];
                for (int p = 0; p < kParts; ++p) mine[p].clear();
#pragma omp for schedule(dynamic, 256)
                for (long long r = r0; r < r1; ++r) {
                    const int* s = cur.data() + start[r];
                    const long long n = len[r];
                    bool run_open = false;  // previous (a,a) was counted
                    for (long long i = 0; i + 1 < n; ++i) {
                        int a = s[i], b = s[i + 1];
                        if (a == b) {
                            if (run_open) { run_open = false; continue; }
                            run_open = true;
                        } else {
                            run_open = false;
                        }
                        uint64_t k = pair_key(a, b);
                        mine[mix64(k) >> (64 - kPartBits)].push_back(k);
                    }
                }
#pragma omp for schedule(dynamic, 1)
                for (int p = 0; p < kParts; ++p) {
                    for (int tt = 0; tt < nthreads; ++tt) {
                        for (uint64_t k : buf[static_cast<std::size_t>(tt) * kParts + p]) tables[p].add(k, 1);
                    }
                }
            }
            r0 = r1;
        }
        buf.clear();
        buf.shrink_to_fit();

        // ---- 2. rank pairs seen at least twice.
        std::vector<std::pair<uint32_t, uint64_t>> ranked;
        {
            std::vector<std::vector<std::pair<uint32_t, uint64_t>>> part(kParts);
#pragma omp parallel for schedule(dynamic, 1)
            for (int p = 0; p < kParts; ++p) {
                const auto& t = tables[p];
                for (std::size_t i = 0; i < t.keys.size(); ++i)
                    if (t.keys[i] && t.counts[i] >= 2) part[p].emplace_back(~t.counts[i], t.keys[i]);
            }
            std::size_t n = 0;
            for (auto& v : part) n += v.size();
            ranked.reserve(n);
            for (auto& v : part) ranked.insert(ranked.end(), v.begin(), v.end());
        }
        for (auto& t : tables) t.init(16);
        if (ranked.empty()) break;
        std::sort(ranked.begin(), ranked.end());  // ~count ascending = count descending
        RankTable rt;
        rt.build(ranked);
        const int nsel = static_cast<int>(ranked.size());

        // ---- 3a. decide replacements, count uses per ranked pair.
        std::vector<std::atomic<int>> use(nsel);
        for (auto& u : use) u.store(0, std::memory_order_relaxed);
        auto decide = [&](const int* s, long long n, long long i, int& rk) -> bool {
            if (i + 1 >= n) return false;
            rk = rt.find(pair_key(s[i], s[i + 1]));
            if (rk < 0) return false;
            if (i + 2 < n) {
                int nx = rt.find(pair_key(s[i + 1], s[i + 2]));
                if (nx >= 0 && nx < rk) return false;
            }
            return true;
        };
#pragma omp parallel for schedule(dynamic, 256)
        for (long long r = 0; r < nrows; ++r) {
            const int* s = cur.data() + start[r];
            const long long n = len[r];
            for (long long i = 0; i < n;) {
                int rk;
                if (decide(s, n, i, rk)) {
                    use[rk].fetch_add(1, std::memory_order_relaxed);
                    i += 2;
                } else {
                    ++i;
                }
            }
        }
        // ---- 4. pairs used at least twice become rules, in rank order.
        final_sym.assign(nsel, -1);
        int made = 0;
        const long long base = static_cast<long long>(g.alpha) + static_cast<long long>(g.rules.size());
        for (int r = 0; r < nsel; ++r) {
            if (use[r].load(std::memory_order_relaxed) < 2) continue;
            if (base + made > 0x7fffffffLL) throw std::runtime_error("repair: symbol space exceeds int32");
            final_sym[r] = static_cast<int>(base + made);
            uint64_t k = ranked[r].second;
            g.rules.push_back({grammar::BINARY, static_cast<int>(k >> 32),
                               static_cast<int>(k & 0xffffffffULL), 1});
            ++made;
        }
        if (made == 0) break;

        // ---- 3b. rewrite rows in place with the same decisions.
        long long total = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : total)
        for (long long r = 0; r < nrows; ++r) {
            int* s = cur.data() + start[r];
            const long long n = len[r];
            long long o = 0;
            for (long long i = 0; i < n;) {
                int rk;
                if (decide(s, n, i, rk)) {
                    int a = s[i], b = s[i + 1];
                    if (final_sym[rk] >= 0) {
                        s[o++] = final_sym[rk];
                    } else {
                        s[o++] = a;
                        s[o++] = b;
                    }
                    i += 2;
                } else {
                    s[o++] = s[i];
                    ++i;
                }
            }
            len[r] = o;
            total += o;
        }
        if (stats) stats->length_after_round.push_back(total + nrows);

        // Compact rows to keep later rounds cache-friendly.
        std::vector<long long> ns(nrows + 1, 0);
        for (long long r = 0; r < nrows; ++r) ns[r + 1] = ns[r] + len[r];
        std::vector<int> nxt(static_cast<std::size_t>(ns[nrows]));
#pragma omp parallel for schedule(static)
        for (long long r = 0; r < nrows; ++r)
            std::memcpy(nxt.data() + ns[r], cur.data() + start[r], sizeof(int) * len[r]);
        cur.swap(nxt);
        for (long long r = 0; r < nrows; ++r) start[r] = ns[r];
    }

    // Top sequence: rows with their delimiters back.
    long long tot = 0;
    for (long long r = 0; r < nrows; ++r) tot += len[r] + 1;
    g.C.resize(static_cast<std::size_t>(tot - (trailing_open ? 1 : 0)));
    {
        long long o = 0;
        for (long long r = 0; r < nrows; ++r) {
            std::memcpy(g.C.data() + o, cur.data() + start[r], sizeof(int) * len[r]);
            o += len[r];
            if (!(trailing_open && r == nrows - 1)) g.C[o++] = 0;
        }
    }
    if (stats) {
        stats->rounds = round;
        stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return g;
}

std::vector<int> expand_grammar(const grammar::Grammar& g) {
    std::vector<int> out;
    std::vector<std::pair<int, int>> stack;  // (symbol, repetitions)
    for (int c : g.C) {
        if (c < g.alpha) { out.push_back(c); continue; }
        stack.push_back({c, 1});
        while (!stack.empty()) {
            auto [s, reps] = stack.back();
            stack.pop_back();
            if (s < g.alpha) {
                for (int i = 0; i < reps; ++i) out.push_back(s);
                continue;
            }
            const grammar::Rule& r = g.rules.at(static_cast<std::size_t>(s - g.alpha));
            for (int i = 0; i < reps; ++i) {
                if (r.kind == grammar::RUNLEN) {
                    stack.push_back({r.left, r.rep});
                } else {
                    stack.push_back({r.right, 1});
                    stack.push_back({r.left, 1});
                }
            }
        }
    }
    return out;
}

}  // namespace gmr
