// grammar.h -- in-memory RePair grammar over a matrix's CSRV string.
//
// Drop-in replacement for the reference's gpu-engine/grammar.h:43-61
// (namespace grammar, RuleKind, Rule, Grammar). Field names, order and sizes are
// kept identical (Rule is four int32, 16 bytes) so code written against the
// reference header compiles and links unchanged.
//
// Symbol encoding (same as the .vc.R / .vc.C files, reference grammar.h:31-40):
//   0 <= s < alpha       terminal; 0 is the row delimiter "$", a terminal p >= 1
//                        decodes to value index (p-1)/cols and column (p-1)%cols
//   s >= alpha           non-terminal, rule index s - alpha
// Rules are numbered topologically: every child of rule i is a terminal or a rule
// with a smaller index, so one forward pass over `rules` assigns levels.
#ifndef GMR_GRAMMAR_H
#define GMR_GRAMMAR_H

#include <cstdint>
#include <vector>

namespace grammar {

enum RuleKind : int {
    BINARY = 0,  // N -> A B            left = A, right = B, rep = 1
    RUNLEN = 1   // N -> B^t, t >= 2    left = B, right unused, rep = t
};

struct Rule {
    int kind;   // RuleKind
    int left;   // first child symbol
    int right;  // second child symbol (BINARY only)
    int rep;    // repetition count (RUNLEN), 1 for BINARY
};

struct Grammar {
    int alpha = 0;            // terminals are [0, alpha)
    std::vector<Rule> rules;  // rules[i] defines symbol alpha + i
    std::vector<int> C;       // top sequence, 0 ends each row
};

}  // namespace grammar

#endif  // GMR_GRAMMAR_H
