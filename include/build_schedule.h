// build_schedule.h -- host scheduler: RePair grammar -> level schedule.
//
// Drop-in replacement for the reference's gpu-engine/build_schedule.h:38-77.
// `BuiltSchedule` keeps every member of the reference struct with the same name,
// type and meaning, and `build_schedule(g, rows, cols)` (build_schedule.h:77)
// returns the reference's exact layout: same depth, same rules in the same slot
// order, same pass-through count, same emission lists and the same terminal
// compaction. That layout is called RefCompat below.
//
// Two further layouts are laid out for 64-lane waves (ScheduleOptions):
//   Wave64    double-buffered like RefCompat, but drops every pass-through that is
//             never read (the copy of a symbol at its highest referencing level),
//             puts each level's pass-throughs first in source-slot order (a
//             monotone, near-contiguous copy) followed by the level's own rules
//             ordered by their left child, and keeps each level's emissions in
//             row order so a wave can pre-reduce runs of one row.
//   Resident  no pass-throughs at all: every value stays resident in one array
//             of num_layered slots (level k owns [level_offsets[k-1],
//             level_offsets[k])), and a child c >= 0 is a global slot. The sweep
//             is still level-synchronous; only the frontier is never recycled.
#ifndef GMR_BUILD_SCHEDULE_H
#define GMR_BUILD_SCHEDULE_H

#include <cstddef>
#include <utility>
#include <vector>

#include "grammar.h"
#include "layered_types.h"

enum class ScheduleLayout : int { RefCompat = 0, Wave64 = 1, Resident = 2 };

struct ScheduleOptions {
    ScheduleLayout layout = ScheduleLayout::RefCompat;
    // Fill the per-level nested views (layered_levels, emit_by_level, term_emit)
    // that CPU sweeps consume. The GPU path needs only the flattened arrays.
    bool keep_nested = true;
};

struct BuiltSchedule {
    int alpha = 0;  // after compaction: U, the number of kept terminals (incl. 0)
    int rows = 0;
    int cols = 0;
    int max_depth = 0;
    int max_width = 0;
    std::size_t num_raw_nt = 0;       // rules of the grammar
    std::size_t num_layered = 0;      // rules + pass-throughs
    std::size_t num_passthrough = 0;  // num_layered - num_raw_nt

    std::vector<LayeredRule> flattened_rules;
    std::vector<int> level_offsets;  // max_depth + 1
    std::vector<int> level_widths;   // max_depth + 1
    std::vector<int> emit_pos;
    std::vector<int> emit_row;
    std::vector<int> emit_offsets;   // max_depth + 1
    std::vector<int> term_emit_sym;
    std::vector<int> term_emit_row;

    // term_orig[q] = original terminal id of compact id q; term_orig[0] = 0.
    std::vector<int> term_orig;

    // Per-level views for CPU sweeps ([0] is empty).
    std::vector<std::vector<LayeredRule>> layered_levels;
    std::vector<std::vector<std::pair<int, int>>> emit_by_level;  // (pos, row)
    std::vector<std::pair<int, int>> term_emit;                   // (T index, row)

    // Not in the reference struct: which layout produced the arrays.
    ScheduleLayout layout = ScheduleLayout::RefCompat;

    // Non-owning view for the GPU entry points; *this must outlive it.
    GPUSchedule gpu_schedule();
};

// Reference layout (reference build_schedule.h:77, build_schedule.cpp:42-227).
BuiltSchedule build_schedule(const grammar::Grammar& g, int rows, int cols);

// Any layout. Throws std::invalid_argument on a malformed grammar (a child that
// is negative or not yet defined, a run-length count < 1, a top-sequence symbol
// out of range, or more rows in C than `rows`).
BuiltSchedule build_schedule(const grammar::Grammar& g, int rows, int cols,
                             const ScheduleOptions& opt);

#endif  // GMR_BUILD_SCHEDULE_H
