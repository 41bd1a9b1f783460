/* oracle.c -- CPU restatement of the reference's y = Mx path. TEST
 * INFRASTRUCTURE ONLY: tests, the smoke check and the bench's CPU-baseline leg
 * call it as the checker; the product never links, loads or calls it.
 *
 * Semiring ids: 0 = (+,x) FP64, 1 = Boolean (OR, AND) on int32, 2 = Tropical
 * (min, +) on int32 with 1e9 as +infinity. The Boolean and Tropical leaves follow
 * the engine's integer policies (reference gpu-engine/gpu_engine.h:59-97): x and
 * V travel as float, leaf = (V != 0 && x != 0) and leaf = (int)V + (int)x or 1e9
 * when either is >= 1e9.
 *
 * Pinning (what checks each function against something other than itself):
 *   orc_brute_force      hand-computed answers (unit-test grammar y = [4.5, 0],
 *                        README 8x7 example), dense matrix-vector products
 *                        computed by numpy in the tests, closed forms
 *   orc_expand           round trip: expand(RePair(s)) == s, and the reference's
 *                        own cpu_selftest binary (oracle/_ref) on the same files
 *   orc_layer_refcompat  the reference's build_schedule compiled from its own
 *                        sources (oracle/_ref/libref_schedule.so), the unit-test
 *                        counts (max_depth 4, 5 pass-throughs) and invariants
 *   orc_sweep            orc_brute_force on the same inputs (exact for Boolean and
 *                        Tropical, 1e-12 relative for FP64)
 *   orc_cpu_layered_f32  orc_sweep on the same schedule (FP32 tolerance)
 *   orc_mlp_loss_grad    central finite differences of its own loss (the
 *                        proposer network has no reference counterpart)
 *   orc_brute_force_left dense x^T M computed by numpy in the tests, and the
 *                        right product of the transposed matrix (x^T M = M^T x)
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_INF 1000000000

/* Leaves. (reference gpu-engine/gpu_engine.h:46-97; cpu_selftest.cpp:39-63) */
static inline int leaf_bool(float v, float x) { return (v != 0.0f && x != 0.0f) ? 1 : 0; }
static inline int leaf_trop(float v, float x) {
    int iv = (int)v, ix = (int)x;
    if (iv >= ORC_INF || ix >= ORC_INF) return ORC_INF;
    return iv + ix;
}

/* y = M x straight from the CSRV string, FP64 for (+,x) and exact integers for
 * Boolean / Tropical. den[r] = sum_j |M_rj x_j| for (+,x) (0 otherwise).
 * (reference gpu-engine/cpu_selftest.cpp:105-117) */
int orc_brute_force(int semiring, int rows, int cols, const int* seq, long long len,
                    const double* V, const float* x, double* y, double* den) {
    for (int r = 0; r < rows; ++r) {
        y[r] = semiring == 2 ? (double)ORC_INF : 0.0;
        if (den) den[r] = 0.0;
    }
    int row = 0;
    for (long long i = 0; i < len; ++i) {
        int t = seq[i];
        if (t == 0) { ++row; continue; }
        if (row >= rows) return 1;
        int v = (t - 1) / cols, c = (t - 1) % cols;
        if (semiring == 0) {
            double p = V[v] * (double)x[c];
            y[row] += p;
            if (den) den[row] += fabs(p);
        } else if (semiring == 1) {
            if (leaf_bool((float)V[v], x[c])) y[row] = 1.0;
        } else {
            int l = leaf_trop((float)V[v], x[c]);
            if (l < y[row]) y[row] = l;
        }
    }
    return 0;
}

/* Left product Y^T = X^T M straight from the CSRV string: X[r*B + b] (rows x B),
 * Y[c*B + b] (cols x B), y_c = (+)_r leaf(V, x_r) with the same leaves as the
 * right product; den[c*B + b] = sum_r |M_rc x_r| for (+,x). mm-repair computes
 * this product beside Ax (reference run_benchmark.py:138-141); its code is not
 * in the reference, so this restates the definition. */
int orc_brute_force_left(int semiring, int rows, int cols, const int* seq, long long len, const double* V,
                         const float* X, int B, double* Y, double* den) {
    for (long long i = 0; i < (long long)cols * B; ++i) {
        Y[i] = semiring == 2 ? (double)ORC_INF : 0.0;
        if (den) den[i] = 0.0;
    }
    int row = 0;
    for (long long i = 0; i < len; ++i) {
        int t = seq[i];
        if (t == 0) { ++row; continue; }
        if (row >= rows) return 1;
        int v = (t - 1) / cols, c = (t - 1) % cols;
        double* yc = Y + (long long)c * B;
        const float* xr = X + (long long)row * B;
        for (int b = 0; b < B; ++b) {
            if (semiring == 0) {
                double p = V[v] * (double)xr[b];
                yc[b] += p;
                if (den) den[(long long)c * B + b] += fabs(p);
            } else if (semiring == 1) {
                if (leaf_bool((float)V[v], xr[b])) yc[b] = 1.0;
            } else {
                int l = leaf_trop((float)V[v], xr[b]);
                if (l < yc[b]) yc[b] = l;
            }
        }
    }
    return 0;
}

/* Same with B right-hand sides, X[c*B + b] and Y[r*B + b].
 * (batched layout: reference gpu-engine/gpu_engine.h:116-120) */
int orc_brute_force_batched(int semiring, int rows, int cols, const int* seq, long long len,
                            const double* V, const float* X, int B, double* Y, double* den) {
    for (long long i = 0; i < (long long)rows * B; ++i) {
        Y[i] = semiring == 2 ? (double)ORC_INF : 0.0;
        if (den) den[i] = 0.0;
    }
    int row = 0;
    for (long long i = 0; i < len; ++i) {
        int t = seq[i];
        if (t == 0) { ++row; continue; }
        if (row >= rows) return 1;
        int v = (t - 1) / cols, c = (t - 1) % cols;
        double* yr = Y + (long long)row * B;
        const float* xc = X + (long long)c * B;
        for (int b = 0; b < B; ++b) {
            if (semiring == 0) {
                double p = V[v] * (double)xc[b];
                yr[b] += p;
                if (den) den[(long long)row * B + b] += fabs(p);
            } else if (semiring == 1) {
                if (leaf_bool((float)V[v], xc[b])) yr[b] = 1.0;
            } else {
                int l = leaf_trop((float)V[v], xc[b]);
                if (l < yr[b]) yr[b] = l;
            }
        }
    }
    return 0;
}

/* Expand a grammar (binary and run-length rules) into its CSRV string. Returns
 * the length, or -1 if `cap` is too small or a symbol is out of range.
 * (reference gpu-engine/cpu_selftest.cpp:119-156) */
long long orc_expand(int alpha, int nrules, const int* kind, const int* left, const int* right,
                     const int* rep, const int* C, long long clen, int* out, long long cap) {
    long long n = 0;
    long long scap = 64, sp = 0;
    long long* st = (long long*)malloc(sizeof(long long) * 2 * scap);
    for (long long j = 0; j < clen; ++j) {
        st[0] = C[j];
        st[1] = 1;
        sp = 1;
        while (sp > 0) {
            --sp;
            long long s = st[2 * sp], reps = st[2 * sp + 1];
            if (s < alpha) {
                for (long long q = 0; q < reps; ++q) {
                    if (n >= cap) { free(st); return -1; }
                    out[n++] = (int)s;
                }
                continue;
            }
            long long ri = s - alpha;
            if (ri >= nrules) { free(st); return -1; }
            for (long long q = 0; q < reps; ++q) {
                if (sp + 2 > scap) {
                    scap *= 2;
                    st = (long long*)realloc(st, sizeof(long long) * 2 * scap);
                }
                if (kind && kind[ri] == 1) {
                    st[2 * sp] = left[ri];
                    st[2 * sp + 1] = rep[ri];
                    ++sp;
                } else {
                    st[2 * sp] = right[ri];
                    st[2 * sp + 1] = 1;
                    ++sp;
                    st[2 * sp] = left[ri];
                    st[2 * sp + 1] = 1;
                    ++sp;
                }
            }
        }
    }
    free(st);
    return n;
}

/* Layered schedule exactly as the reference lays it out, by the reference's own
 * method: levels in one topological pass, the highest referencing rule level of
 * every symbol, then per level k the rules of level k in index order followed
 * by one pass-through per previous-level slot whose symbol is still referenced
 * above k-1, tracked in a dense rep[R x (L+1)] slot table; emissions in C
 * order; terminals compacted to the referenced ids in increasing order.
 * (reference gpu-engine/build_schedule.cpp:42-227)
 *
 * Two-call protocol: with out_rules == NULL only counts[] is filled:
 *   counts[0] max_depth, [1] max_width, [2] num_layered, [3] emit_total,
 *   [4] num_term_emit, [5] U (compacted alpha).
 * Then call again with buffers of those sizes (offsets/widths max_depth+1).
 * Returns 0, or 1 on a malformed grammar. */
int orc_layer_refcompat(int alpha, int nrules, const int* kind, const int* left,
                        const int* right, const int* rep, const int* C, long long clen,
                        long long counts[6], int* out_left, int* out_right, float* out_coeff,
                        int* level_offsets, int* level_widths, int* emit_offsets,
                        int* emit_pos, int* emit_row, int* term_sym, int* term_row,
                        int* term_orig, int* out_rules) {
    int R = nrules;
    int* lvl = (int*)calloc(R + 1, sizeof(int));
    int L = 0;
    for (int i = 0; i < R; ++i) {
        int a = left[i] >= alpha ? lvl[left[i] - alpha] : 0;
        int b = 0;
        if (!(kind && kind[i] == 1)) b = right[i] >= alpha ? lvl[right[i] - alpha] : 0;
        lvl[i] = 1 + (a > b ? a : b);
        if (lvl[i] > L) L = lvl[i];
    }
    int* mref = (int*)malloc(sizeof(int) * (R + 1));
    for (int i = 0; i < R; ++i) mref[i] = lvl[i];
    for (int i = 0; i < R; ++i) {
        if (left[i] >= alpha && lvl[i] > mref[left[i] - alpha]) mref[left[i] - alpha] = lvl[i];
        if (!(kind && kind[i] == 1) && right[i] >= alpha && lvl[i] > mref[right[i] - alpha])
            mref[right[i] - alpha] = lvl[i];
    }
    long long stride = L + 1;
    int* reptab = (int*)malloc(sizeof(int) * ((size_t)R * stride + 1));
    for (long long i = 0; i < (long long)R * stride; ++i) reptab[i] = -1;
    /* Per-level lists of (orig symbol) and the layered rules, grown level by level. */
    long long cap = R + 16, nl = 0;
    int* L_left = (int*)malloc(sizeof(int) * cap);
    int* L_right = (int*)malloc(sizeof(int) * cap);
    float* L_coeff = (float*)malloc(sizeof(float) * cap);
    int* L_orig = (int*)malloc(sizeof(int) * cap);
    long long* off = (long long*)calloc(L + 2, sizeof(long long));
    int maxw = 0;
    for (int k = 1; k <= L; ++k) {
        long long lvl_begin = nl;
        for (int i = 0; i < R; ++i) {
            if (lvl[i] != k) continue;
            if (nl + 1 > cap) {
                cap *= 2;
                L_left = realloc(L_left, sizeof(int) * cap);
                L_right = realloc(L_right, sizeof(int) * cap);
                L_coeff = realloc(L_coeff, sizeof(float) * cap);
                L_orig = realloc(L_orig, sizeof(int) * cap);
            }
            int lc = left[i];
            L_left[nl] = lc < alpha ? -lc - 1 : reptab[(long long)(lc - alpha) * stride + k - 1];
            if (kind && kind[i] == 1) {
                L_right[nl] = -1;
                L_coeff[nl] = (float)rep[i];
            } else {
                int rc = right[i];
                L_right[nl] = rc < alpha ? -rc - 1 : reptab[(long long)(rc - alpha) * stride + k - 1];
                L_coeff[nl] = 1.0f;
            }
            L_orig[nl] = i;
            reptab[(long long)i * stride + k] = (int)(nl - lvl_begin);
            ++nl;
        }
        if (k > 1) {
            long long pb = off[k - 2], pe = off[k - 1];
            for (long long j = pb; j < pe; ++j) {
                int o = L_orig[j];
                if (mref[o] > k - 1) {
                    if (nl + 1 > cap) {
                        cap *= 2;
                        L_left = realloc(L_left, sizeof(int) * cap);
                        L_right = realloc(L_right, sizeof(int) * cap);
                        L_coeff = realloc(L_coeff, sizeof(float) * cap);
                        L_orig = realloc(L_orig, sizeof(int) * cap);
                    }
                    L_left[nl] = (int)(j - pb);
                    L_right[nl] = -1;
                    L_coeff[nl] = 1.0f;
                    L_orig[nl] = o;
                    reptab[(long long)o * stride + k] = (int)(nl - lvl_begin);
                    ++nl;
                }
            }
        }
        off[k] = nl;
        if (nl - lvl_begin > maxw) maxw = (int)(nl - lvl_begin);
    }
    /* Emissions from C, in C order, grouped by level (stable). */
    long long nE = 0, nT = 0;
    for (long long j = 0; j < clen; ++j) {
        if (C[j] == 0) continue;
        if (C[j] < alpha) ++nT; else ++nE;
    }
    /* Terminal compaction. */
    int* remap = (int*)malloc(sizeof(int) * (size_t)(alpha > 0 ? alpha : 1));
    for (int t = 0; t < alpha; ++t) remap[t] = -1;
    remap[0] = 0;
    for (long long j = 0; j < nl; ++j) {
        if (L_left[j] < 0) remap[-L_left[j] - 1] = 0;
        if (L_right[j] < 0) remap[-L_right[j] - 1] = 0;
    }
    for (long long j = 0; j < clen; ++j)
        if (C[j] != 0 && C[j] < alpha) remap[C[j]] = 0;
    int U = 1;
    for (int t = 1; t < alpha; ++t)
        if (remap[t] == 0) remap[t] = U++;

    counts[0] = L;
    counts[1] = maxw;
    counts[2] = nl;
    counts[3] = nE;
    counts[4] = nT;
    counts[5] = U;
    if (out_rules) {
        for (long long j = 0; j < nl; ++j) {
            out_left[j] = L_left[j] < 0 ? -remap[-L_left[j] - 1] - 1 : L_left[j];
            out_right[j] = L_right[j] < 0 ? -remap[-L_right[j] - 1] - 1 : L_right[j];
            out_coeff[j] = L_coeff[j];
        }
        level_offsets[0] = 0;
        level_widths[0] = 0;
        for (int k = 1; k <= L; ++k) {
            level_offsets[k] = (int)off[k];
            level_widths[k] = (int)(off[k] - off[k - 1]);
        }
        /* emissions grouped by level, C order inside a level */
        long long* ecnt = (long long*)calloc(L + 2, sizeof(long long));
        for (long long j = 0; j < clen; ++j)
            if (C[j] >= alpha) ++ecnt[lvl[C[j] - alpha]];
        emit_offsets[0] = 0;
        long long acc = 0;
        long long* efill = (long long*)calloc(L + 2, sizeof(long long));
        for (int k = 1; k <= L; ++k) {
            efill[k] = acc;
            acc += ecnt[k];
            emit_offsets[k] = (int)acc;
        }
        int row = 0;
        long long te = 0;
        for (long long j = 0; j < clen; ++j) {
            int c = C[j];
            if (c == 0) { ++row; continue; }
            if (c < alpha) {
                term_sym[te] = remap[c];
                term_row[te] = row;
                ++te;
            } else {
                int n = c - alpha, k = lvl[n];
                long long e = efill[k]++;
                emit_pos[e] = reptab[(long long)n * stride + k];
                emit_row[e] = row;
            }
        }
        int q = 0;
        for (int t = 0; t < alpha; ++t)
            if (remap[t] >= 0) term_orig[q++] = t;
        free(ecnt);
        free(efill);
        *out_rules = (int)nl;
    }
    free(lvl);
    free(mref);
    free(reptab);
    free(L_left);
    free(L_right);
    free(L_coeff);
    free(L_orig);
    free(off);
    free(remap);
    return 0;
}

/* Double-buffered layered sweep with emit-on-the-spot in FP64 / exact integers,
 * one vector. rules (left, right, coeff) flattened by level; children c >= 0 are
 * slots of the previous level, c < 0 the terminal -c-1 whose value is
 * leaf(V[(t-1)/cols], x[(t-1)%cols]) with t = term_orig[-c-1] (t = 0: zero).
 * (reference gpu-engine/cpu_selftest.cpp:67-102) */
int orc_sweep(int semiring, int rows, int cols, int U, const int* term_orig, const double* V,
              const float* x, int L, const int* rleft, const int* rright, const float* rcoeff,
              const int* level_offsets, const int* emit_offsets, const int* emit_pos,
              const int* emit_row, long long n_term, const int* term_sym, const int* term_row,
              double* y) {
    const double zero = semiring == 2 ? (double)ORC_INF : 0.0;
    double* T = (double*)malloc(sizeof(double) * (size_t)(U > 0 ? U : 1));
    T[0] = zero;
    for (int p = 1; p < U; ++p) {
        int t = term_orig[p], v = (t - 1) / cols, c = (t - 1) % cols;
        if (semiring == 0) T[p] = V[v] * (double)x[c];
        else if (semiring == 1) T[p] = leaf_bool((float)V[v], x[c]);
        else T[p] = leaf_trop((float)V[v], x[c]);
    }
    for (int r = 0; r < rows; ++r) y[r] = zero;
#define OPLUS(a, b) (semiring == 0 ? (a) + (b) : semiring == 1 ? (((a) != 0 || (b) != 0) ? 1.0 : 0.0) : ((a) < (b) ? (a) : (b)))
    for (long long e = 0; e < n_term; ++e) y[term_row[e]] = OPLUS(y[term_row[e]], T[term_sym[e]]);
    int maxw = 1;
    for (int k = 1; k <= L; ++k)
        if (level_offsets[k] - level_offsets[k - 1] > maxw) maxw = level_offsets[k] - level_offsets[k - 1];
    double* prev = (double*)malloc(sizeof(double) * maxw);
    double* cur = (double*)malloc(sizeof(double) * maxw);
    for (int k = 1; k <= L; ++k) {
        int b = level_offsets[k - 1], e = level_offsets[k];
        for (int i = b; i < e; ++i) {
            double lv = rleft[i] >= 0 ? prev[rleft[i]] : T[-rleft[i] - 1];
            double rv = rright[i] >= 0 ? prev[rright[i]] : T[-rright[i] - 1];
            double av = semiring == 0 ? (double)rcoeff[i] * lv : lv;
            cur[i - b] = OPLUS(av, rv);
        }
        for (int q = emit_offsets[k - 1]; q < emit_offsets[k]; ++q)
            y[emit_row[q]] = OPLUS(y[emit_row[q]], cur[emit_pos[q]]);
        double* t = prev;
        prev = cur;
        cur = t;
    }
#undef OPLUS
    free(T);
    free(prev);
    free(cur);
    return 0;
}

/* The reference's OpenMP CPU sweep in the engine's value types (FP32 for (+,x),
 * int32 otherwise): terminals, then terminal emits, then per level the node
 * update and the level's emits, each loop `parallel for` above 1000 items;
 * emits use `omp atomic` for + and OR and `omp critical` for min, and the
 * frontier is copied after every level. This is the CPU baseline the engine is
 * timed against. y holds float bits for (+,x) and int32 bits otherwise.
 * (reference gpu-engine/gpu_engine_test.cu:49-104; policies gpu_engine.h:46-97) */
/* Threads of the OpenMP sweep below (the CPU baseline is timed at a stated
 * thread count). */
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

void orc_cpu_layered_f32(int semiring, int rows, int U, const int* val_idx, const int* col_idx,
                         const float* V, const float* x, int L, const int* rleft,
                         const int* rright, const float* rcoeff, const int* level_offsets,
                         const int* emit_offsets, const int* emit_pos, const int* emit_row,
                         long long n_term, const int* term_sym, const int* term_row, void* yv) {
    int maxw = 1;
    for (int k = 1; k <= L; ++k)
        if (level_offsets[k] - level_offsets[k - 1] > maxw) maxw = level_offsets[k] - level_offsets[k - 1];
    if (semiring == 0) {
        float* y = (float*)yv;
        float* T = (float*)malloc(sizeof(float) * (size_t)(U > 0 ? U : 1));
        float* rd = (float*)calloc(maxw, sizeof(float));
        float* wr = (float*)calloc(maxw, sizeof(float));
        T[0] = 0.0f;
#pragma omp parallel for if (U > 1000)
        for (int p = 1; p < U; ++p) T[p] = V[val_idx[p - 1]] * x[col_idx[p - 1]];
#pragma omp parallel for if (rows > 1000)
        for (int r = 0; r < rows; ++r) y[r] = 0.0f;
#pragma omp parallel for if (n_term > 1000)
        for (long long e = 0; e < n_term; ++e) {
#pragma omp atomic
            y[term_row[e]] += T[term_sym[e]];
        }
        for (int k = 1; k <= L; ++k) {
            int b = level_offsets[k - 1], n = level_offsets[k] - b;
#pragma omp parallel for if (n > 1000)
            for (int i = 0; i < n; ++i) {
                int l = rleft[b + i], r = rright[b + i];
                float lv = l >= 0 ? rd[l] : T[-l - 1];
                float rv = r >= 0 ? rd[r] : T[-r - 1];
                wr[i] = rcoeff[b + i] * lv + rv;
            }
            int eb = emit_offsets[k - 1], ne = emit_offsets[k] - eb;
#pragma omp parallel for if (ne > 1000)
            for (int q = 0; q < ne; ++q) {
#pragma omp atomic
                y[emit_row[eb + q]] += wr[emit_pos[eb + q]];
            }
            memcpy(rd, wr, sizeof(float) * (size_t)n);
        }
        free(T);
        free(rd);
        free(wr);
        return;
    }
    int* y = (int*)yv;
    const int zero = semiring == 2 ? ORC_INF : 0;
    int* T = (int*)malloc(sizeof(int) * (size_t)(U > 0 ? U : 1));
    int* rd = (int*)calloc(maxw, sizeof(int));
    int* wr = (int*)calloc(maxw, sizeof(int));
    T[0] = zero;
#pragma omp parallel for if (U > 1000)
    for (int p = 1; p < U; ++p)
        T[p] = semiring == 1 ? leaf_bool(V[val_idx[p - 1]], x[col_idx[p - 1]])
                             : leaf_trop(V[val_idx[p - 1]], x[col_idx[p - 1]]);
#pragma omp parallel for if (rows > 1000)
    for (int r = 0; r < rows; ++r) y[r] = zero;
#pragma omp parallel for if (n_term > 1000)
    for (long long e = 0; e < n_term; ++e) {
        int v = T[term_sym[e]];
        if (semiring == 1) {
#pragma omp atomic
            y[term_row[e]] |= v;
        } else {
#pragma omp critical(orc_tropical_emit)
            if (v < y[term_row[e]]) y[term_row[e]] = v;
        }
    }
    for (int k = 1; k <= L; ++k) {
        int b = level_offsets[k - 1], n = level_offsets[k] - b;
#pragma omp parallel for if (n > 1000)
        for (int i = 0; i < n; ++i) {
            int l = rleft[b + i], r = rright[b + i];
            int lv = l >= 0 ? rd[l] : T[-l - 1];
            int rv = r >= 0 ? rd[r] : T[-r - 1];
            wr[i] = semiring == 1 ? (lv | rv) : (lv < rv ? lv : rv);
        }
        int eb = emit_offsets[k - 1], ne = emit_offsets[k] - eb;
#pragma omp parallel for if (ne > 1000)
        for (int q = 0; q < ne; ++q) {
            int v = wr[emit_pos[eb + q]];
            if (semiring == 1) {
#pragma omp atomic
                y[emit_row[eb + q]] |= v;
            } else {
#pragma omp critical(orc_tropical_emit)
                if (v < y[emit_row[eb + q]]) y[emit_row[eb + q]] = v;
            }
        }
        memcpy(rd, wr, sizeof(int) * (size_t)n);
    }
    free(T);
    free(rd);
    free(wr);
}

/* The schedule proposer's network in FP64, for checking the GPU kernels:
 * in -> H -> H -> out, ReLU on the hidden layers, loss = sum of
 * mask*(out - target)^2 / count. Parameters as in the engine (W1 [H x in], b1,
 * W2 [H x H], b2, W3 [out x H], b3). Writes the loss and, if grad != NULL,
 * d loss / d params. There is no reference counterpart; pinned by central
 * finite differences in the tests. */
double orc_mlp_loss_grad(int nin, int nh, int nout, const double* p, const double* X, const double* T,
                         const double* M, int N, double count, double* grad, double* out) {
    const double *W1 = p, *b1 = W1 + (long)nh * nin, *W2 = b1 + nh, *b2 = W2 + (long)nh * nh,
                 *W3 = b2 + nh, *b3 = W3 + (long)nout * nh;
    double* z1 = (double*)malloc(sizeof(double) * nh);
    double* a1 = (double*)malloc(sizeof(double) * nh);
    double* z2 = (double*)malloc(sizeof(double) * nh);
    double* a2 = (double*)malloc(sizeof(double) * nh);
    double* o = (double*)malloc(sizeof(double) * nout);
    double* dO = (double*)malloc(sizeof(double) * nout);
    double* d2 = (double*)malloc(sizeof(double) * nh);
    double* d1 = (double*)malloc(sizeof(double) * nh);
    long np = (long)nh * nin + nh + (long)nh * nh + nh + (long)nout * nh + nout;
    if (grad) memset(grad, 0, sizeof(double) * np);
    double *gW1 = grad, *gb1 = grad ? gW1 + (long)nh * nin : 0, *gW2 = grad ? gb1 + nh : 0,
           *gb2 = grad ? gW2 + (long)nh * nh : 0, *gW3 = grad ? gb2 + nh : 0, *gb3 = grad ? gW3 + (long)nout * nh : 0;
    double loss = 0.0;
    for (int n = 0; n < N; ++n) {
        const double* x = X + (long)n * nin;
        for (int h = 0; h < nh; ++h) {
            double z = b1[h];
            for (int i = 0; i < nin; ++i) z += W1[(long)h * nin + i] * x[i];
            z1[h] = z;
            a1[h] = z > 0 ? z : 0;
        }
        for (int h = 0; h < nh; ++h) {
            double z = b2[h];
            for (int i = 0; i < nh; ++i) z += W2[(long)h * nh + i] * a1[i];
            z2[h] = z;
            a2[h] = z > 0 ? z : 0;
        }
        for (int k = 0; k < nout; ++k) {
            double z = b3[k];
            for (int i = 0; i < nh; ++i) z += W3[(long)k * nh + i] * a2[i];
            o[k] = z;
            if (out) out[(long)n * nout + k] = z;
            double d = (z - T[(long)n * nout + k]) * M[(long)n * nout + k];
            loss += d * d;
            dO[k] = 2.0 * d / count;
        }
        if (!grad) continue;
        for (int k = 0; k < nout; ++k) {
            gb3[k] += dO[k];
            for (int i = 0; i < nh; ++i) gW3[(long)k * nh + i] += dO[k] * a2[i];
        }
        for (int i = 0; i < nh; ++i) {
            double g = 0;
            for (int k = 0; k < nout; ++k) g += dO[k] * W3[(long)k * nh + i];
            d2[i] = z2[i] > 0 ? g : 0;
        }
        for (int h = 0; h < nh; ++h) {
            gb2[h] += d2[h];
            for (int i = 0; i < nh; ++i) gW2[(long)h * nh + i] += d2[h] * a1[i];
        }
        for (int i = 0; i < nh; ++i) {
            double g = 0;
            for (int h = 0; h < nh; ++h) g += d2[h] * W2[(long)h * nh + i];
            d1[i] = z1[i] > 0 ? g : 0;
        }
        for (int h = 0; h < nh; ++h) {
            gb1[h] += d1[h];
            for (int i = 0; i < nin; ++i) gW1[(long)h * nin + i] += d1[h] * x[i];
        }
    }
    free(z1); free(a1); free(z2); free(a2); free(o); free(dO); free(d2); free(d1);
    return loss / count;
}
