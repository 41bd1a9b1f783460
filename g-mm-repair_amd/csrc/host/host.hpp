// Host-side building blocks of the engine: CSRV matrices, seeded generators,
// the RePair grammar builder and the .val/.vc.R/.vc.C file formats.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "grammar.h"

struct BuiltSchedule;

namespace gmr {

// A matrix as its CSRV string: for every row, the terminals of its non-zeros in
// increasing column order followed by 0. A non-zero of value V[v] in column c is
// the terminal 1 + v*cols + c (the .vc terminal convention).
struct Csrv {
    int rows = 0;
    int cols = 0;
    std::vector<double> V;  // distinct non-zero values
    std::vector<int> seq;   // CSRV string
    long long nnz() const { return static_cast<long long>(seq.size()) - rows; }
    int alpha() const;      // 1 + largest terminal (at least 1)
};

// 0/1 haplotype mosaic: `founders` rows of Bernoulli(density) bits; every row
// walks the columns in segments of Geometric(seg_mean) length, copying one
// uniformly drawn founder per segment, then flips each cell with prob. `flip`.
Csrv gen_haplotype(uint64_t seed, int rows, int cols, int founders, double density,
                   double seg_mean, double flip);

// Genotype-like 0/1 mosaic with per-row allele count k, P(k) ~ 1/k on
// [1, cols-1]. `founders` founder rows are split evenly over `classes` density
// classes with log-spaced densities; a row copies founders of the class whose
// density is closest to k/cols (segments Geometric(seg_mean)), then random
// flips add or remove ones until the row holds exactly k.
Csrv gen_genotype(uint64_t seed, int rows, int cols, int founders, int classes,
                  double seg_mean);

// R-MAT directed graph on 2^scale vertices from `edges` draws with quadrant
// probabilities (a, b, c, 1-a-b-c); duplicate edges removed, rows sorted.
// nweights == 1: V = {1}; nweights > 1: every edge carries an integer weight
// in [1, nweights] fixed by (seed, u, v), V = {1, ..., nweights}.
Csrv gen_rmat(uint64_t seed, int scale, long long edges, double a, double b, double c,
              int nweights);

// Dense row-major matrix -> CSRV (zeros dropped, V = sorted distinct values).
Csrv csrv_from_dense(const double* M, int rows, int cols);

struct RepairOptions {
    int max_rounds = 1000;  // stop after this many replacement rounds
    int threads = 0;        // 0: OpenMP default
};

struct RepairStats {
    int rounds = 0;
    std::vector<long long> length_after_round;
    double seconds = 0.0;
};

// RePair grammar of a CSRV string (pairs never include the delimiter 0).
grammar::Grammar repair(const Csrv& m, const RepairOptions& opt, RepairStats* stats);

// Expand the grammar back into its CSRV string.
std::vector<int> expand_grammar(const grammar::Grammar& g);

// The same matrix with every binary rule that no rule references and that C
// uses 1..max_emits times written out in C as its two children; the other
// rules keep their order and are renumbered. max_emits <= 0 copies.
grammar::Grammar inline_top_rules(const grammar::Grammar& g, int max_emits, long long* n_inlined);

// Files: <base>.val float64[nvals]; <base>.vc.R int32 alpha then (left,right)
// int32 pairs; <base>.vc.C int32 top sequence. Binary rules only, as RePair
// writes them. Throws std::runtime_error on missing or truncated files.
void write_grammar_files(const std::string& base, const grammar::Grammar& g,
                         const std::vector<double>& V);
grammar::Grammar read_grammar_files(const std::string& base, std::vector<double>* V);

// CPU layered sweep (FP32 for semiring 0, int32 for 1 Boolean and 2 min/plus):
// y = M x over any layout of a built schedule; val_idx / col_idx decode compact
// terminal q >= 1 (value V[val_idx[q-1]] in column col_idx[q-1]). threads > 0
// sets the OpenMP thread count for the call. Throws std::invalid_argument on
// an unknown semiring.
void cpu_layered_sweep(int semiring, const BuiltSchedule& bs, const int* val_idx, const int* col_idx, const float* V,
                       const float* x, void* y, int threads);

}  // namespace gmr
