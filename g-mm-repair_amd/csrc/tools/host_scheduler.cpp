// host_scheduler -- grammar inspector (counterpart of the reference's
// gpu-engine/host_scheduler.cpp:31-102).
//
// Usage: host_scheduler <base> [rows cols]
// With only <base> it reads <base>.vc.R and prints what the reference tool
// prints: alpha, the rule count, the depth L and the raw per-level widths
// (levels from the level equation, before any pass-through completion). With
// rows and cols it also reads <base>.vc.C and prints, for each schedule layout
// (RefCompat = the reference's, Wave64, Resident), the layered rule count,
// pass-throughs, max width, emissions, and how the levels map onto 64-lane
// waves (waves per level, lanes idle in each level's last wave).
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

#include "build_schedule.h"
#include "host.hpp"

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "Usage: " << argv[0] << " <matrix_base_name> [rows cols]" << std::endl;
        return 1;
    }
    const std::string base = argv[1];
    grammar::Grammar g;
    std::vector<double> V;
    try {
        g = gmr::read_grammar_files(base, &V);
    } catch (const std::exception& e) {
        // Only .vc.R is required for the raw statistics.
        FILE* f = std::fopen((base + ".vc.R").c_str(), "rb");
        if (!f) {
            std::cerr << "Error: Could not open rules file: " << base << ".vc.R" << std::endl;
            return 1;
        }
        int alpha = 0;
        if (std::fread(&alpha, sizeof(int), 1, f) != 1) {
            std::cerr << "Error: Could not read alphabet size (Alpha)." << std::endl;
            std::fclose(f);
            return 1;
        }
        g.alpha = alpha;
        int pr[2];
        while (std::fread(pr, sizeof(int), 2, f) == 2) g.rules.push_back({grammar::BINARY, pr[0], pr[1], 1});
        std::fclose(f);
    }
    const std::size_t R = g.rules.size();
    std::cout << "Alphabet size (Alpha): " << g.alpha << std::endl;
    std::cout << "Number of non-terminals (NTnum): " << R << std::endl;
    std::vector<int> lvl(R, 0);
    int L = 0;
    for (std::size_t i = 0; i < R; ++i) {
        const auto& r = g.rules[i];
        int a = r.left >= g.alpha ? lvl[r.left - g.alpha] : 0;
        int b = r.right >= g.alpha ? lvl[r.right - g.alpha] : 0;
        lvl[i] = 1 + std::max(a, b);
        L = std::max(L, lvl[i]);
    }
    std::cout << "Grammar depth (L): " << L << std::endl;
    std::vector<std::size_t> w(L + 1, 0);
    for (int v : lvl) ++w[v];
    std::cout << "\nLevel Width Distribution:" << std::endl << "Level\tWidth" << std::endl;
    for (int k = 1; k <= L; ++k) std::cout << k << "\t" << w[k] << std::endl;

    if (argc >= 4) {
        const int rows = std::stoi(argv[2]), cols = std::stoi(argv[3]);
        const char* names[] = {"RefCompat", "Wave64", "Resident"};
        for (int lay = 0; lay < 3; ++lay) {
            ScheduleOptions opt;
            opt.layout = static_cast<ScheduleLayout>(lay);
            opt.keep_nested = false;
            BuiltSchedule s;
            try {
                s = build_schedule(g, rows, cols, opt);
            } catch (const std::exception& e) {
                std::cerr << "Error: " << e.what() << std::endl;
                return 1;
            }
            long long waves = 0, idle = 0;
            for (int k = 1; k <= s.max_depth; ++k) {
                int wk = s.level_widths[k];
                waves += (wk + 63) / 64;
                idle += (64 - wk % 64) % 64;
            }
            std::cout << "\n[" << names[lay] << "] layered rules: " << s.num_layered
                      << " (raw NTs: " << s.num_raw_nt << ", +pt: " << s.num_passthrough << ")"
                      << ", max width: " << s.max_width << ", emissions: " << s.emit_pos.size() << "+"
                      << s.term_emit_sym.size() << " term, U: " << s.alpha << std::endl;
            std::cout << "  wave64: " << waves << " waves over " << s.max_depth << " levels, " << idle
                      << " idle lanes in level tails" << std::endl;
        }
    }
    return 0;
}
