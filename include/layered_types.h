// layered_types.h -- the layered rule and the level schedule the sweep consumes.
//
// Drop-in replacement for the reference's gpu-engine/layered_types.h:45-75.
// LayeredRule is {int, int, float} = 12 bytes and GPUSchedule has the same
// members in the same order, so a schedule produced by either implementation can
// be handed to the other.
//
// Node update (the unified branch-free form):
//     W_k[i] = apply(cv(left), coeff)  (+)  cv(right)
//     cv(c)  = W_{k-1}[c]   if c >= 0   (slot of the previous level's frontier)
//            = T[-c-1]      if c <  0   (terminal value; T[0] is the semiring zero)
// binary N -> A B:   left = A, right = B,         coeff = 1
// pass-through N->M: left = M, right = -1 (T[0]), coeff = 1
// run-length N->B^t: left = B, right = -1 (T[0]), coeff = t
#ifndef GMR_LAYERED_TYPES_H
#define GMR_LAYERED_TYPES_H

struct LayeredRule {
    int left;
    int right;
    float coeff;
};

// Level k (1..max_depth) owns flattened rules [level_offsets[k-1], level_offsets[k]).
// Non-terminal occurrences of the top sequence that are computed at level k are
// emitted right after level k: entries [emit_offsets[k-1], emit_offsets[k]) add
// frontier_k[emit_pos[e]] into y[emit_row[e]]. Terminal occurrences of the top
// sequence add T[term_emit_sym[e]] into y[term_emit_row[e]] before the sweep.
// All pointers are borrowed; the owner must outlive the view.
struct GPUSchedule {
    int max_depth;
    int max_width;
    int* level_offsets;   // max_depth + 1 entries
    int* level_widths;    // max_depth + 1 entries

    int* emit_offsets;    // max_depth + 1 entries
    int* emit_pos;        // emit_total entries
    int* emit_row;        // emit_total entries
    int emit_total;

    int* term_emit_sym;   // num_term_emit entries (index into T)
    int* term_emit_row;   // num_term_emit entries
    int num_term_emit;
};

#endif  // GMR_LAYERED_TYPES_H
