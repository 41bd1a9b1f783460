// Device code of the level sweep (gfx950, wave64).
//
// Work is written as "phase" functions over a range of items, parameterised by
// a group index and a group count, so the same code runs as one launch per
// level (the plain sweep), inside captured graphs, inside the persistent
// kernels (one grammar, or many as tiles) where phases are separated by the
// bounded grid barrier below, and for the left product's node gathers.
//
// Lane mapping: a group of G lanes (G = 1, 2, 4, ..., 64, a divisor of the
// wave) owns one item (a rule, an emission chunk, a row); its lanes stride over
// the B vectors of that item. With vectors innermost (W[slot*B + b]) a G = 64
// group reads and writes 256-byte contiguous runs, so a wave moves whole cache
// lines for every child gather and every emit.
#pragma once

#include <hip/hip_runtime.h>

namespace gmr {
namespace dev {

constexpr int kTropicalInf = 1000000000;

// Semiring policies (same algebra as the reference's gpu_engine.h:46-97).
struct PlusTimesOp {
    using T = float;
    static __device__ __forceinline__ T zero() { return 0.0f; }
    static __device__ __forceinline__ T oplus(T a, T b) { return a + b; }
    static __device__ __forceinline__ T leaf(float v, float x) { return v * x; }
    static __device__ __forceinline__ T apply(T v, float c) { return c * v; }
    static __device__ __forceinline__ bool skip(T) { return false; }
    // A whole row's sum stored into y: what (+)-combining it into a y that
    // starts at zero() gives (0 + v = v; only the sign of a zero differs).
    static __device__ __forceinline__ T fin(T v) { return v; }
    // Hardware global_atomic_add_f32 (no compare-and-swap loop).
    static __device__ __forceinline__ void emit(T* y, T v) { unsafeAtomicAdd(y, v); }
};

struct BooleanOp {
    using T = int;
    static __device__ __forceinline__ T zero() { return 0; }
    static __device__ __forceinline__ T oplus(T a, T b) { return a | b; }
    static __device__ __forceinline__ T leaf(float v, float x) { return (v != 0.0f && x != 0.0f) ? 1 : 0; }
    static __device__ __forceinline__ T apply(T v, float) { return v; }
    static __device__ __forceinline__ bool skip(T v) { return v == 0; }  // OR with 0 is a no-op
    static __device__ __forceinline__ T fin(T v) { return v; }            // 0 | v
    // Values are 0 or 1 and y starts at 0, so OR-ing a 1 in is storing 1: any
    // number of writers store the same value, with no read-modify-write (a
    // device-scope atomic would travel to the memory side; a store stays in L2).
    static __device__ __forceinline__ void emit(T* y, T v) {
        if (v) *y = 1;
    }
};

struct TropicalOp {
    using T = int;
    static __device__ __forceinline__ T zero() { return kTropicalInf; }
    static __device__ __forceinline__ T oplus(T a, T b) { return a < b ? a : b; }
    static __device__ __forceinline__ T leaf(float v, float x) {
        int iv = static_cast<int>(v), ix = static_cast<int>(x);
        if (iv >= kTropicalInf || ix >= kTropicalInf) return kTropicalInf;
        return iv + ix;
    }
    static __device__ __forceinline__ T apply(T v, float) { return v; }
    // y starts at +inf, so min with anything >= +inf leaves it unchanged.
    static __device__ __forceinline__ bool skip(T v) { return v >= kTropicalInf; }
    // min(+inf, v): y starts at +inf, so a row whose finite leaf sums all exceed
    // 1e9 still reads +inf, as with atomicMin into the filled y.
    static __device__ __forceinline__ T fin(T v) { return v < kTropicalInf ? v : kTropicalInf; }
    static __device__ __forceinline__ void emit(T* y, T v) { atomicMin(y, v); }
};

// Everything one sweep needs, passed by value to every kernel.
struct SweepArgs {
    // grammar structure (device)
    const int2* kids;        // (left, right) per layered rule
    const float* coeff;      // per layered rule, or nullptr when every coeff is 1
    const int* tcol;         // column of compact terminal p (p >= 1)
    const float* tval;       // value of compact terminal p (p >= 1)
    const int2* emits;       // (pos, row) per emission, grouped by level, row-sorted
    const int2* term_emits;  // (sym, row) per terminal occurrence, row-sorted
    const int* crow;         // gather mode: row r's symbols are csym[crow[r], crow[r+1])
    const int* csym;         // >= 0 global slot of a rule, < 0 terminal -c-1
    const int4* gseg;        // gather segments: (row, first symbol, end, row is split)
    int n_gseg;
    int gsplit;              // some row is split: its y row is zeroed and the pieces add atomically
    const int* zrows;        // the split rows (zeroed by the init phase of a segment gather)
    int n_zrows;
    const unsigned long long* fbits;  // flat gather: bit i of word c set = csym[64c+i] starts a row
    const int* ford;         // flat gather: rows started before csym[64c]
    const int* frow;         // flat gather: row id of the k-th non-empty row
    int n_fchunk;            // flat gather: 64-symbol chunks of csym
    int n_csym;              // top-sequence symbols (csym entries)
    int flat;                // one vector: gather over 64-symbol chunks instead of row segments
    // One vector, XCD-partitioned gather: the symbols split into 8 x passes
    // value ranges (terminals then rule slots, equal use counts), each a
    // row-ordered list padded to 64-symbol chunks with the zero terminal.
    const int* xsym;         // the lists, concatenated
    const unsigned long long* xbits;  // per chunk: which symbols start a row of their list
    const int* xord;         // per chunk: index into xrow of the first row started before it
    const int* xrow;         // per list: its non-empty rows in order
    const int* xpoff;        // chunk range of list p: [xpoff[p], xpoff[p+1])
    int xpass0, xpass1;      // passes [xpass0, xpass1) of this launch; pass k owns lists 8k .. 8k+7
    int xparts;              // use the partitioned gather
    const int* level_off;    // max_depth + 1
    const int* emit_off;     // max_depth + 1
    // values (device)
    const float* X;          // cols x B (this tile's column base folded in)
    void* Y;                 // rows x B
    void* Tv;                // U x Bs terminal values
    void* W0;                // frontier buffer 0 (or the resident value array)
    void* W1;                // frontier buffer 1 (double-buffered layouts)
    // sizes
    int rows, cols, U, L;
    int n_term;              // terminal emissions
    int B;                   // row stride of X and Y
    int Bt;                  // vectors in this tile
    int Bs;                  // stride of T and W rows: Bt rounded up to a multiple of 4 (1 for one vector)
    int G;                   // lanes per item
    int resident;            // children index one resident array
    int gather;              // emit by a final row gather
    int vec;                 // 1: four vectors per lane (Bt > 1; T and W rows padded to Bs)
    int yvec;                // Y rows are 16-byte aligned (B and the tile's first column multiples of 4)
    int GR;                  // lanes per row in the single-vector gather
    int eval_kind;           // EvalKind for levels without an entry in level_kind
    const int* level_kind;   // device, max_depth + 1 entries (or nullptr)
    // Boolean, one vector, bit mode: terminal q is bit q of vbits, rule slot s
    // bit 64*tw + s (tw = terminal words); levels and gather read bits.
    unsigned long long* vbits;
    long nvw;                // words of vbits that hold values
    int tw;
    int bits;
    int nbl;                 // bit mode, LDS ranges: lists (in xsym ... xpoff)
    // One vector, values staged in LDS by range (GMR_RUN_LDS_RANGES): list p of
    // xsym ... xpoff holds the symbols whose value key (terminal t -> t, rule
    // slot s -> U + s) lies in [p * lspan, (p + 1) * lspan), stored as the key's
    // index inside the range.
    int lspan;               // values per range (0: not this gather)
    long nkey;               // U + rule slots
    int lds_window;          // EV_LDS left-child window of the per-level launches (values)
    int bit_range;           // bit mode: bits per LDS range
    int chunks;              // 64-symbol chunks per wave in flight (chunked single-vector gathers)
    // LDS-range gather (one vector): the init phase computes only the terminal
    // values rules read (tref, n_tref of them); the range fills compute the rest
    const int* tref;
    int n_tref;
    // LDS-range gather: entry = (row - cbase[chunk]) << 16 | index inside the
    // range (0xFFFF: padding); xord / xrow are not used
    const int* cbase;
    // LDS-range gather, rows renumbered (GMR_RUN_ROW_ORDER, lorder = 1): the
    // lists' rows are new numbers, pieces combine into Y (then the row-ordered
    // copy) and a last pass writes Yout[r] = Y[ynew[r]]
    int lorder;
    const int* ynew;
    void* Yout;
    // GMR_RUN_FUSED_TAIL: 2 x kTailRules values staging the tail's children that
    // lie below it
    void* tail_stage;
};

// Per-level node-update kernels (one is chosen for each level):
//   EV_GROUP   a group of G lanes per rule, lanes over vectors (dwordx4 when
//              possible); a group takes four rules at once, their kids first,
//              then all eight child rows (eight loads in flight per lane)
//   EV_THREAD  one lane per (rule, 4-vector chunk): no idle lanes for narrow B
//   EV_WAVE    a whole wave per rule, lanes over single vectors
//   EV_LDS     single vector only: a workgroup stages the window of the
//              previous frontier its rules' left children fall in into LDS
//              (the Wave64/Resident layouts sort a level's rules by left
//              child), then gathers left children from LDS
enum EvalKind : int { EV_GROUP = 0, EV_THREAD = 1, EV_WAVE = 2, EV_LDS = 3 };
constexpr int kEvalKinds = 4;
constexpr int kLdsWindow = 4096;  // values staged per workgroup (16 KiB)

template <class S>
__device__ __forceinline__ typename S::T* tp(void* p) {
    return static_cast<typename S::T*>(p);
}

// Four vectors per lane: float4 / int4 loads and stores (dwordx4). T and W
// rows are padded to Bs (a multiple of 4), so they are always 16-byte aligned;
// the padding lanes carry the semiring zero and are never written to Y.
template <class T> struct Vec4;
template <> struct Vec4<float> { using type = float4; };
template <> struct Vec4<int> { using type = int4; };

template <class S>
struct Op4 {
    using T = typename S::T;
    using T4 = typename Vec4<T>::type;
    static __device__ __forceinline__ T4 zero() {
        T4 r;
        r.x = r.y = r.z = r.w = S::zero();
        return r;
    }
    static __device__ __forceinline__ T4 oplus(T4 a, T4 b) {
        T4 r;
        r.x = S::oplus(a.x, b.x);
        r.y = S::oplus(a.y, b.y);
        r.z = S::oplus(a.z, b.z);
        r.w = S::oplus(a.w, b.w);
        return r;
    }
    static __device__ __forceinline__ T4 apply(T4 a, float c) {
        T4 r;
        r.x = S::apply(a.x, c);
        r.y = S::apply(a.y, c);
        r.z = S::apply(a.z, c);
        r.w = S::apply(a.w, c);
        return r;
    }
    static __device__ __forceinline__ T4 ld(const T* p) { return *reinterpret_cast<const T4*>(p); }
    static __device__ __forceinline__ void st(T* p, T4 v) { *reinterpret_cast<T4*>(p) = v; }
    // The first n (<= 4) components into y, atomically / by plain stores.
    static __device__ __forceinline__ void emit_n(T* y, T4 v, int n) {
        if (n > 0 && !S::skip(v.x)) S::emit(y + 0, v.x);
        if (n > 1 && !S::skip(v.y)) S::emit(y + 1, v.y);
        if (n > 2 && !S::skip(v.z)) S::emit(y + 2, v.z);
        if (n > 3 && !S::skip(v.w)) S::emit(y + 3, v.w);
    }
    static __device__ __forceinline__ T4 fin(T4 a) {
        T4 r;
        r.x = S::fin(a.x);
        r.y = S::fin(a.y);
        r.z = S::fin(a.z);
        r.w = S::fin(a.w);
        return r;
    }
    static __device__ __forceinline__ void st_n(T* y, T4 v, int n, int aligned) {
        if (n >= 4 && aligned) {
            st(y, v);
            return;
        }
        if (n > 0) y[0] = v.x;
        if (n > 1) y[1] = v.y;
        if (n > 2) y[2] = v.z;
        if (n > 3) y[3] = v.w;
    }
};

// Segmented inclusive scan of a wave by DPP: lane l ends up with the oplus of
// lanes [l - dist, l], where dist is its distance to the first lane of its
// piece (0 at a piece's first lane). Four row_shr steps scan each row of 16
// lanes; row_bcast:15 then carries lane 15's (47's) partial into rows 1 (3),
// and row_bcast:31 lane 31's into rows 2 and 3, each taken only by lanes whose
// piece reaches back that far. Every lane of the wave must be active.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int dpp_move(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROW_MASK, 0xF, false);
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_move(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xF, false));
}
template <class S>
__device__ __forceinline__ typename S::T seg_scan(typename S::T v, int dist, int lane) {
    using T = typename S::T;
    const int r = lane & 15;
    const int m = min(dist, r);  // how far the piece reaches back inside the row
    T t = dpp_move<0x111, 0xF>(v);  // row_shr:1
    if (m >= 1) v = S::oplus(v, t);
    t = dpp_move<0x112, 0xF>(v);  // row_shr:2
    if (m >= 2) v = S::oplus(v, t);
    t = dpp_move<0x114, 0xF>(v);  // row_shr:4
    if (m >= 4) v = S::oplus(v, t);
    t = dpp_move<0x118, 0xF>(v);  // row_shr:8
    if (m >= 8) v = S::oplus(v, t);
    t = dpp_move<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    if ((lane & 16) && dist > r) v = S::oplus(v, t);
    t = dpp_move<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
    if (lane >= 32 && dist > lane - 32) v = S::oplus(v, t);
    return v;
}

// Lanes per item for rows of Bs values: one lane per value (vec = 0) or per
// four values (vec = 1), capped at the wave width.
__host__ __device__ __forceinline__ int lanes_for(int Bs, int vec) {
    int need = vec ? Bs / 4 : Bs;
    int g = 1;
    while (g < need && g < 64) g <<= 1;
    return g;
}

// Whether the init phase zeroes only the listed split rows: a row-segment
// gather stores every other row whole.
__host__ __device__ __forceinline__ bool zero_list(const SweepArgs& a) {
    return a.gather && !a.flat && !a.xparts && !a.lspan;
}

// T[p*Bs + b] = leaf(V, X[col*B + b]) for b < Bt, zero for p = 0 and for the
// padding lanes; Y = zero for spot emits (and gathers with split rows).
template <class S>
__device__ void ph_init_range(const SweepArgs& a, long i0, long i1, long tid, long nth) {
    using T = typename S::T;
    T* Tt = tp<S>(a.Tv);
    T* Y = tp<S>(a.Y);
    const long nT = a.lspan ? static_cast<long>(a.n_tref) : static_cast<long>(a.U) * a.Bs;
    for (long i = i0 + tid; i < i1; i += nth) {
        if (a.lspan && i < nT) {  // one vector: only the terminals rules read
            const int p = a.tref[i];
            Tt[p] = S::leaf(a.tval[p], a.X[static_cast<long>(a.tcol[p]) * a.B]);
        } else if (i < nT) {
            int p = static_cast<int>(i / a.Bs);
            int b = static_cast<int>(i - static_cast<long>(p) * a.Bs);
            Tt[i] = (p == 0 || b >= a.Bt) ? S::zero()
                                          : S::leaf(a.tval[p], a.X[static_cast<long>(a.tcol[p]) * a.B + b]);
        } else {
            const long j = i - nT;
            const long q = j / a.Bt;
            const long r = zero_list(a) ? a.zrows[q] : q;
            Y[r * a.B + (j - q * a.Bt)] = S::zero();
        }
    }
}

// Items of the init phase: U*Bs terminal values, then the y values that are
// combined atomically: every row for spot emission and the chunked gathers,
// the split rows for the row-segment gather.
__host__ __device__ __forceinline__ long init_items(const SweepArgs& a) {
    const long ny = zero_list(a) ? (a.gsplit ? static_cast<long>(a.n_zrows) : 0L) : static_cast<long>(a.rows);
    const long nt = a.lspan ? static_cast<long>(a.n_tref) : static_cast<long>(a.U) * a.Bs;
    return nt + ny * a.Bt;
}

// Work items of the gather phase: 64-symbol chunks (flat) or row segments.
__host__ __device__ __forceinline__ long gather_items(const SweepArgs& a) {
    return a.flat ? static_cast<long>(a.n_fchunk) : static_cast<long>(a.n_gseg);
}

template <class S>
__device__ void ph_init(const SweepArgs& a, long tid, long nth) {
    ph_init_range<S>(a, 0, init_items(a), tid, nth);
}

// Rules [i0, i1) of one level. src is the previous frontier (or the resident
// array), dst this level's frontier slice.
template <class S>
__device__ void ph_eval(const SweepArgs& a, int i0, int i1, const typename S::T* src,
                        typename S::T* dst, long gid, long ng, int lane) {
    using T = typename S::T;
    const T* Tt = tp<S>(a.Tv);
    const int Bt = a.Bs, G = a.G;
    for (long i = i0 + gid; i < i1; i += ng) {
        int2 c = a.kids[i];
        float cf = a.coeff ? a.coeff[i] : 1.0f;
        const T* lp = c.x >= 0 ? src + static_cast<long>(c.x) * Bt : Tt + static_cast<long>(-c.x - 1) * Bt;
        const T* rp = c.y >= 0 ? src + static_cast<long>(c.y) * Bt : Tt + static_cast<long>(-c.y - 1) * Bt;
        T* o = dst + (i - i0) * Bt;
        for (int b = lane; b < Bt; b += G) o[b] = S::oplus(S::apply(lp[b], cf), rp[b]);
    }
}

// Emit-on-the-spot for one vector: 64 consecutive (pos, row) pairs per wave,
// pre-reduced by a segmented scan over equal rows (rows are sorted), then one
// atomic per row run.
template <class S>
__device__ void ph_emit_wave(const int2* em, int n, const typename S::T* src, typename S::T* Y,
                             long wid, long nw, int lane) {
    using T = typename S::T;
    for (long base = wid * 64; base < n; base += nw * 64) {
        long e = base + lane;
        int row = -1;
        T v = S::zero();
        if (e < n) {
            int2 pr = em[e];
            row = pr.y;
            v = src[pr.x];
        }
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            T ov = __shfl_up(v, d, 64);
            int orow = __shfl_up(row, d, 64);
            if (lane >= d && orow == row) v = S::oplus(v, ov);
        }
        int nrow = __shfl_down(row, 1, 64);
        bool tail = lane == 63 || nrow != row;
        if (row >= 0 && tail && !S::skip(v)) S::emit(Y + row, v);
    }
}

// Emit-on-the-spot, B-wide: a group takes CH consecutive emissions, keeps a
// running sum per vector while the row repeats and issues one atomic per run.
template <class S, int CH>
__device__ void ph_emit_chunk(const int2* em, int n, const typename S::T* src, typename S::T* Y,
                              int B, int Bt, int Bs, int G, long gid, long ng, int lane) {
    using T = typename S::T;
    const long nch = (static_cast<long>(n) + CH - 1) / CH;
    for (long ch = gid; ch < nch; ch += ng) {
        const int e0 = static_cast<int>(ch * CH);
        const int e1 = min(n, e0 + CH);
        for (int b = lane; b < Bt; b += G) {
            int row = -1;
            T acc = S::zero();
            for (int e = e0; e < e1; ++e) {
                int2 pr = em[e];
                T v = src[static_cast<long>(pr.x) * Bs + b];
                if (pr.y != row) {
                    if (row >= 0 && !S::skip(acc)) S::emit(Y + static_cast<long>(row) * B + b, acc);
                    row = pr.y;
                    acc = v;
                } else {
                    acc = S::oplus(acc, v);
                }
            }
            if (row >= 0 && !S::skip(acc)) S::emit(Y + static_cast<long>(row) * B + b, acc);
        }
    }
}

template <class S, int CH>
__device__ void ph_emit_chunk_v4(const int2* em, int n, const typename S::T* src, typename S::T* Y,
                                 int B, int Bt, int Bs, int G, long gid, long ng, int lane) {
    using O = Op4<S>;
    using T4 = typename O::T4;
    const long nch = (static_cast<long>(n) + CH - 1) / CH;
    for (long ch = gid; ch < nch; ch += ng) {
        const int e0 = static_cast<int>(ch * CH);
        const int e1 = min(n, e0 + CH);
        for (int b = 4 * lane; b < Bs; b += 4 * G) {
            const int nb = min(4, Bt - b);
            if (nb <= 0) break;
            int row = -1;
            T4 acc = O::zero();
            for (int e = e0; e < e1; ++e) {
                int2 pr = em[e];
                T4 v = O::ld(src + static_cast<long>(pr.x) * Bs + b);
                if (pr.y != row) {
                    if (row >= 0) O::emit_n(Y + static_cast<long>(row) * B + b, acc, nb);
                    row = pr.y;
                    acc = v;
                } else {
                    acc = O::oplus(acc, v);
                }
            }
            if (row >= 0) O::emit_n(Y + static_cast<long>(row) * B + b, acc, nb);
        }
    }
}

constexpr int kEmitChunk = 8;

template <class S>
__device__ __forceinline__ void ph_emit(const SweepArgs& a, const int2* em, int n,
                                        const typename S::T* src, long tid, long nth, int lane) {
    if (a.Bt == 1 && a.B == 1) {
        ph_emit_wave<S>(em, n, src, tp<S>(a.Y), tid / 64, nth / 64, static_cast<int>(tid & 63));
    } else if (a.vec) {
        ph_emit_chunk_v4<S, kEmitChunk>(em, n, src, tp<S>(a.Y), a.B, a.Bt, a.Bs, a.G, tid / a.G, nth / a.G, lane);
    } else {
        ph_emit_chunk<S, kEmitChunk>(em, n, src, tp<S>(a.Y), a.B, a.Bt, a.Bs, a.G, tid / a.G, nth / a.G, lane);
    }
}

// Resident layout, gather emission: Y[r] = (+) over row r's top-sequence
// symbols, computed per segment (a run of at most 128 symbols of one row). A
// segment that is its row's only one stores the sum; the segments of a split
// row add into the zeroed row with the semiring's atomic.
template <class S>
__device__ __forceinline__ void gather_out(typename S::T* y, typename S::T v, int split) {
    if (!split) *y = S::fin(v);
    else if (!S::skip(v)) S::emit(y, v);
}

// Scalar B-wide form: a lane group per segment, lanes over vectors.
template <class S>
__device__ void ph_gather_rows(const SweepArgs& a, long q0, long q1, long gid, long ng, int lane) {
    using T = typename S::T;
    const T* W = tp<S>(a.W0);
    const T* Tt = tp<S>(a.Tv);
    T* Y = tp<S>(a.Y);
    const int Bt = a.Bt, Bs = a.Bs;
    for (long q = q0 + gid; q < q1; q += ng) {
        const int4 sg = a.gseg[q];
        for (int b = lane; b < Bt; b += a.G) {
            T acc = S::zero();
            for (int j = sg.y; j < sg.z; ++j) {
                int c = a.csym[j];
                T v = c >= 0 ? W[static_cast<long>(c) * Bs + b] : Tt[static_cast<long>(-c - 1) * Bs + b];
                acc = S::oplus(acc, v);
            }
            gather_out<S>(Y + static_cast<long>(sg.x) * a.B + b, acc, sg.w);
        }
    }
}

// Four vectors per lane, symbols unrolled by 8 so each lane has eight dwordx4
// gathers in flight; the next eight symbol ids are loaded while those gathers
// are outstanding, so the id latency stays off the critical path.
template <class S>
__device__ void ph_gather_rows_v4(const SweepArgs& a, long q0, long q1, long gid, long ng, int lane) {
    using T = typename S::T;
    using O = Op4<S>;
    using T4 = typename O::T4;
    constexpr int U = 8;
    const T* W = tp<S>(a.W0);
    const T* Tt = tp<S>(a.Tv);
    T* Y = tp<S>(a.Y);
    const int Bt = a.Bt, Bs = a.Bs;
    auto row = [&](int c) -> const T* {
        return c >= 0 ? W + static_cast<long>(c) * Bs : Tt + static_cast<long>(-c - 1) * Bs;
    };
    auto store = [&](const int4& sg, int b, int nb, const T4& v) {
        T* y = Y + static_cast<long>(sg.x) * a.B + b;
        if (!sg.w) O::st_n(y, O::fin(v), nb, a.yvec);
        else O::emit_n(y, v, nb);
    };
    auto one = [&](const int4& sg) {
        for (int b = 4 * lane; b < Bs; b += 4 * a.G) {
            const int nb = min(4, Bt - b);
            if (nb <= 0) break;
            T4 acc0 = O::zero(), acc1 = O::zero();
            int j = sg.y;
            int c[U];
#pragma unroll
            for (int u = 0; u < U; ++u) c[u] = j + u < sg.z ? a.csym[j + u] : 0;
            while (j + U <= sg.z) {
                T4 v[U];
#pragma unroll
                for (int u = 0; u < U; ++u) v[u] = O::ld(row(c[u]) + b);
                j += U;
#pragma unroll
                for (int u = 0; u < U; ++u) c[u] = j + u < sg.z ? a.csym[j + u] : 0;
                acc0 = O::oplus(acc0, O::oplus(O::oplus(v[0], v[1]), O::oplus(v[2], v[3])));
                acc1 = O::oplus(acc1, O::oplus(O::oplus(v[4], v[5]), O::oplus(v[6], v[7])));
            }
            // Fewer than U symbols left; their ids are already in c.
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (j + u < sg.z) acc0 = O::oplus(acc0, O::ld(row(c[u]) + b));
            store(sg, b, nb, O::oplus(acc0, acc1));
        }
    };
    // Two segments of at most four symbols each: their eight reads together.
    auto two_short = [&](const int4& s1, const int4& s2) {
        constexpr int H = U / 2;
        int c[U];
#pragma unroll
        for (int u = 0; u < H; ++u) {
            c[u] = s1.y + u < s1.z ? a.csym[s1.y + u] : 0;
            c[H + u] = s2.y + u < s2.z ? a.csym[s2.y + u] : 0;
        }
        for (int b = 4 * lane; b < Bs; b += 4 * a.G) {
            const int nb = min(4, Bt - b);
            if (nb <= 0) break;
            T4 v[U];
#pragma unroll
            for (int u = 0; u < H; ++u) {
                v[u] = s1.y + u < s1.z ? O::ld(row(c[u]) + b) : O::zero();
                v[H + u] = s2.y + u < s2.z ? O::ld(row(c[H + u]) + b) : O::zero();
            }
            store(s1, b, nb, O::oplus(O::oplus(v[0], v[1]), O::oplus(v[2], v[3])));
            store(s2, b, nb, O::oplus(O::oplus(v[4], v[5]), O::oplus(v[6], v[7])));
        }
    };
    for (long q = q0 + gid; q < q1; q += 2 * ng) {
        const int4 s1 = a.gseg[q];
        if (q + ng < q1) {
            const int4 s2 = a.gseg[q + ng];
            if (s1.z - s1.y <= U / 2 && s2.z - s2.y <= U / 2) {
                two_short(s1, s2);
            } else {
                one(s1);
                one(s2);
            }
        } else {
            one(s1);
        }
    }
}

// One vector: GR lanes per segment stride over its symbols, then a shuffle
// tree reduces the GR partial results.
template <class S>
__device__ void ph_gather_rows_1(const SweepArgs& a, long q0, long q1, long tid, long nth) {
    using T = typename S::T;
    const T* W = tp<S>(a.W0);
    const T* Tt = tp<S>(a.Tv);
    T* Y = tp<S>(a.Y);
    const int GR = a.GR;
    const int sub = static_cast<int>(tid % GR);
    // All lanes of a wave run the same number of iterations (shuffles below).
    const long span = q1 - q0;
    const long pad = ((span + (nth / GR) - 1) / (nth / GR)) * (nth / GR);
    for (long i = tid / GR; i < pad; i += nth / GR) {
        T acc = S::zero();
        int4 sg = make_int4(0, 0, 0, 0);
        if (i < span) {
            sg = a.gseg[q0 + i];
            for (int j = sg.y + sub; j < sg.z; j += GR) {
                int c = a.csym[j];
                acc = S::oplus(acc, c >= 0 ? W[c] : Tt[-c - 1]);
            }
        }
        for (int d = GR / 2; d >= 1; d >>= 1) acc = S::oplus(acc, __shfl_xor(acc, d, 64));
        if (sub == 0 && i < span) gather_out<S>(Y + sg.x, acc, sg.w);
    }
}

// One vector, flat: a wave takes 64 consecutive top-sequence symbols, one per
// lane (so the ids are read coalesced), gathers their values and reduces them
// per row with a segmented shuffle scan; the chunk's row-start bitmap marks the
// segments. A row that starts and ends inside the chunk is stored; the pieces
// of a row that crosses a chunk boundary combine atomically into y, which the
// init phase zeroed (as it did every empty row).
template <class S>
__device__ void ph_gather_flat(const SweepArgs& a, long c0, long c1, long wid, long nw, int lane) {
    using T = typename S::T;
    const T* W = tp<S>(a.W0);
    const T* Tt = tp<S>(a.Tv);
    T* Y = tp<S>(a.Y);
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);  // bits 0..lane
    for (long c = c0 + wid; c < c1; c += nw) {
        const long p = c * 64 + lane;
        const bool valid = p < a.n_csym;
        T v = S::zero();
        if (valid) {
            const int s = a.csym[p];
            v = s >= 0 ? W[s] : Tt[-s - 1];
        }
        const unsigned long long bits = a.fbits[c];
        // Inclusive scan inside segments: lane takes lane-d's partial sum when
        // no segment starts in (lane-d, lane].
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const T t = __shfl_up(v, d, 64);
            if (lane >= d && ((bits >> (lane - d + 1)) & ((1ull << d) - 1ull)) == 0ull) v = S::oplus(v, t);
        }
        const bool last = valid && (lane == 63 || p + 1 == a.n_csym || ((bits >> (lane + 1)) & 1ull) != 0ull);
        if (last) {
            const int r = a.frow[a.ford[c] - 1 + __popcll(bits & upto)];
            const bool starts_here = (bits & upto) != 0ull;
            const bool ends_here = lane < 63 || a.crow[r + 1] == p + 1;
            if (starts_here && ends_here) Y[r] = S::fin(v);
            else if (!S::skip(v)) S::emit(Y + r, v);
        }
    }
}

// Chunks a wave of the XCD-partitioned gather takes at once (two: 15% faster
// than one on the 500,000 x 10,000 genotype grammar; four: slower than one).
constexpr int kPartChunks = 2;

// One vector, XCD-partitioned: workgroups are dispatched to the 8 XCDs round-
// robin, so in pass k workgroup b works on list 8k + b % 8 and every XCD's L2
// only holds the values of one of the 8 x passes value ranges. Each list is swept like
// the chunked gather; every row piece is combined atomically into the zeroed y,
// since the row's other symbols live in other lists. `blk`/`nblk` are the raw
// workgroup index and count (not renumbered).
template <class S, int K>
__device__ void ph_gather_parts(const SweepArgs& a, long blk, long nblk, int tid_in_blk, int bdim) {
    using T = typename S::T;
    const T* W = tp<S>(a.W0);
    const T* Tt = tp<S>(a.Tv);
    T* Y = tp<S>(a.Y);
    const int wpb = bdim >> 6, w = tid_in_blk >> 6, lane = tid_in_blk & 63;
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    long p0, pstep, j, nj;
    if (nblk >= 8) {
        p0 = blk & 7;
        pstep = 8;
        j = blk >> 3;
        nj = (nblk >> 3) + ((blk & 7) < (nblk & 7) ? 1 : 0);
    } else {
        p0 = blk;
        pstep = nblk;
        j = 0;
        nj = 1;
    }
    // Segmented scan of one chunk's values and the row pieces it completes.
    auto finish = [&](long c, T v) {
        const unsigned long long bits = a.xbits[c];
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const T t = __shfl_up(v, d, 64);
            if (lane >= d && ((bits >> (lane - d + 1)) & ((1ull << d) - 1ull)) == 0ull) v = S::oplus(v, t);
        }
        if (lane == 63 || ((bits >> (lane + 1)) & 1ull) != 0ull) {
            const int r = a.xrow[a.xord[c] - 1 + __popcll(bits & upto)];
            if (!S::skip(v)) S::emit(Y + r, v);
        }
    };
    // K chunks per wave at once: all symbol ids, then all values, then the
    // scans (the later chunks' loads hide behind the first scan).
    for (long pass = a.xpass0; pass < a.xpass1; ++pass)
    for (long q = p0; q < 8; q += pstep) {
        const long p = pass * 8 + q;
        const long end = a.xpoff[p + 1], stride = nj * wpb;
        for (long c = a.xpoff[p] + j * wpb + w; c < end; c += K * stride) {
            int sy[K];
            T v[K];
#pragma unroll
            for (int u = 0; u < K; ++u) sy[u] = c + u * stride < end ? a.xsym[(c + u * stride) * 64 + lane] : -1;
#pragma unroll
            for (int u = 0; u < K; ++u) v[u] = sy[u] >= 0 ? W[sy[u]] : Tt[-sy[u] - 1];
#pragma unroll
            for (int u = 0; u < K; ++u)
                if (c + u * stride < end) finish(c + u * stride, v[u]);  // the same for the whole wave
        }
    }
}

// Several vectors (Bs = 4 NV), XCD-partitioned: the same lists as one vector,
// with a 16 NV-byte value row per symbol (NV dwordx4 per lane) and the
// segmented scan and the atomics on its 4 NV components. Used for Bs = 4; at
// Bs = 8 (32-byte rows) it measured no faster than the row segments.
template <class S, int NV, int K>
__device__ void ph_gather_parts4(const SweepArgs& a, long blk, long nblk, int tid_in_blk, int bdim) {
    using T = typename S::T;
    using O = Op4<S>;
    using T4 = typename O::T4;
    constexpr int BS = 4 * NV;
    const T* W = tp<S>(a.W0);
    const T* Tt = tp<S>(a.Tv);
    T* Y = tp<S>(a.Y);
    const int wpb = bdim >> 6, w = tid_in_blk >> 6, lane = tid_in_blk & 63;
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    long p0, pstep, j, nj;
    if (nblk >= 8) {
        p0 = blk & 7;
        pstep = 8;
        j = blk >> 3;
        nj = (nblk >> 3) + ((blk & 7) < (nblk & 7) ? 1 : 0);
    } else {
        p0 = blk;
        pstep = nblk;
        j = 0;
        nj = 1;
    }
    auto finish = [&](long c, T4* v) {
        const unsigned long long bits = a.xbits[c];
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const bool take = lane >= d && ((bits >> (lane - d + 1)) & ((1ull << d) - 1ull)) == 0ull;
#pragma unroll
            for (int n = 0; n < NV; ++n) {
                T4 t;
                t.x = __shfl_up(v[n].x, d, 64);
                t.y = __shfl_up(v[n].y, d, 64);
                t.z = __shfl_up(v[n].z, d, 64);
                t.w = __shfl_up(v[n].w, d, 64);
                if (take) v[n] = O::oplus(v[n], t);
            }
        }
        if (lane == 63 || ((bits >> (lane + 1)) & 1ull) != 0ull) {
            const long r = a.xrow[a.xord[c] - 1 + __popcll(bits & upto)];
#pragma unroll
            for (int n = 0; n < NV; ++n)
                if (4 * n < a.Bt) O::emit_n(Y + r * a.B + 4 * n, v[n], a.Bt - 4 * n);
        }
    };
    for (long pass = a.xpass0; pass < a.xpass1; ++pass)
    for (long q = p0; q < 8; q += pstep) {
        const long p = pass * 8 + q;
        const long end = a.xpoff[p + 1], stride = nj * wpb;
        for (long c = a.xpoff[p] + j * wpb + w; c < end; c += K * stride) {
            int sy[K];
            T4 v[K][NV];
#pragma unroll
            for (int u = 0; u < K; ++u) sy[u] = c + u * stride < end ? a.xsym[(c + u * stride) * 64 + lane] : -1;
#pragma unroll
            for (int u = 0; u < K; ++u) {
                const T* row = sy[u] >= 0 ? W + static_cast<long>(sy[u]) * BS : Tt + static_cast<long>(-sy[u] - 1) * BS;
#pragma unroll
                for (int n = 0; n < NV; ++n) v[u][n] = O::ld(row + 4 * n);
            }
#pragma unroll
            for (int u = 0; u < K; ++u)
                if (c + u * stride < end) finish(c + u * stride, v[u]);  // the same for the whole wave
        }
    }
}

// Chunks per wave in flight: a.chunks of 1, 2 or 4 (0: kPartChunks).
template <class S>
__device__ __forceinline__ void ph_gather_parts_any(const SweepArgs& a, long blk, long nblk, int tid_in_blk,
                                                    int bdim) {
    const int k = a.chunks > 0 ? a.chunks : kPartChunks;
    if (a.vec) {  // Bs = 4 only
        if (k == 1) ph_gather_parts4<S, 1, 1>(a, blk, nblk, tid_in_blk, bdim);
        else if (k == 4) ph_gather_parts4<S, 1, 4>(a, blk, nblk, tid_in_blk, bdim);
        else ph_gather_parts4<S, 1, 2>(a, blk, nblk, tid_in_blk, bdim);
    } else {
        if (k == 1) ph_gather_parts<S, 1>(a, blk, nblk, tid_in_blk, bdim);
        else if (k == 4) ph_gather_parts<S, 4>(a, blk, nblk, tid_in_blk, bdim);
        else ph_gather_parts<S, 2>(a, blk, nblk, tid_in_blk, bdim);
    }
}

template <class S>
__device__ __forceinline__ void ph_gather_range(const SweepArgs& a, long r0, long r1, long tid, long nth) {
    if (a.flat) {
        ph_gather_flat<S>(a, r0, r1, tid / 64, nth / 64, static_cast<int>(tid % 64));
    } else if (a.Bt == 1 && a.B == 1) {
        ph_gather_rows_1<S>(a, r0, r1, tid, nth);
    } else if (a.vec) {
        ph_gather_rows_v4<S>(a, r0, r1, tid / a.G, nth / a.G, static_cast<int>(tid % a.G));
    } else {
        ph_gather_rows<S>(a, r0, r1, tid / a.G, nth / a.G, static_cast<int>(tid % a.G));
    }
}

template <class S>
__device__ __forceinline__ void ph_gather(const SweepArgs& a, long tid, long nth) {
    ph_gather_range<S>(a, 0, gather_items(a), tid, nth);
}

// EV_GROUP, four vectors per lane: rules i, i + ng, ..., i + (U-1)ng per
// step, all kids first, then all 2U child rows, then the U stores. The node
// update is latency-bound (two dependent loads per rule), so four rules in
// flight beat one (-31% at B = 64 on the 500,000 x 10,000 genotype grammar)
// and eight (register pressure).
template <class S, int U = 4>
__device__ void ph_eval_v4x4(const SweepArgs& a, int i0, int i1, const typename S::T* src, typename S::T* dst,
                             long gid, long ng, int lane) {
    using T = typename S::T;
    using O = Op4<S>;
    using T4 = typename O::T4;
    const T* Tt = tp<S>(a.Tv);
    const int Bt = a.Bs, G = a.G;
    auto row = [&](int c) -> const T* {
        return c >= 0 ? src + static_cast<long>(c) * Bt : Tt + static_cast<long>(-c - 1) * Bt;
    };
    for (long i = i0 + gid; i < i1; i += U * ng) {
        int2 c[U];
        float f[U];
        bool on[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long r = i + u * ng;
            on[u] = r < i1;
            c[u] = on[u] ? a.kids[r] : make_int2(-1, -1);
            f[u] = (on[u] && a.coeff) ? a.coeff[r] : 1.0f;
        }
        for (int b = 4 * lane; b < Bt; b += 4 * G) {
            T4 l[U], r[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (on[u]) {
                    l[u] = O::ld(row(c[u].x) + b);
                    r[u] = O::ld(row(c[u].y) + b);
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (on[u]) O::st(dst + (i + u * ng - i0) * Bt + b, O::oplus(O::apply(l[u], f[u]), r[u]));
        }
    }
}

// EV_GROUP, one vector: U rules per lane at once (rules i, i + n, ...,
// i + (U-1)n), all kids first, then all 2U child values, then the U stores.
template <class S, int U>
__device__ void ph_eval_1xU(const SweepArgs& a, int i0, int i1, const typename S::T* src, typename S::T* dst,
                            long tid, long nth) {
    using T = typename S::T;
    const T* Tt = tp<S>(a.Tv);
    for (long i = i0 + tid; i < i1; i += U * nth) {
        int2 c[U];
        float f[U];
        bool on[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long r = i + u * nth;
            on[u] = r < i1;
            c[u] = on[u] ? a.kids[r] : make_int2(-1, -1);
            f[u] = (on[u] && a.coeff) ? a.coeff[r] : 1.0f;
        }
        T l[U], rv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (on[u]) {
                l[u] = c[u].x >= 0 ? src[c[u].x] : Tt[-c[u].x - 1];
                rv[u] = c[u].y >= 0 ? src[c[u].y] : Tt[-c[u].y - 1];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (on[u]) dst[i + u * nth - i0] = S::oplus(S::apply(l[u], f[u]), rv[u]);
    }
}

// EV_THREAD: flattened (rule, chunk) items, chunk = 4 vectors (vec) or 1.
template <class S>
__device__ void ph_eval_thread(const SweepArgs& a, int i0, int i1, const typename S::T* src,
                               typename S::T* dst, long tid, long nth) {
    using T = typename S::T;
    using O = Op4<S>;
    const T* Tt = tp<S>(a.Tv);
    const int Bt = a.Bs;
    const int c = a.vec ? Bt / 4 : Bt;
    const long n = static_cast<long>(i1 - i0) * c;
    for (long j = tid; j < n; j += nth) {
        const long r = j / c;
        const int ch = static_cast<int>(j - r * c);
        const long i = i0 + r;
        const int2 k = a.kids[i];
        const float cf = a.coeff ? a.coeff[i] : 1.0f;
        const T* lp = k.x >= 0 ? src + static_cast<long>(k.x) * Bt : Tt + static_cast<long>(-k.x - 1) * Bt;
        const T* rp = k.y >= 0 ? src + static_cast<long>(k.y) * Bt : Tt + static_cast<long>(-k.y - 1) * Bt;
        T* o = dst + r * Bt;
        if (a.vec) {
            const int b = 4 * ch;
            O::st(o + b, O::oplus(O::apply(O::ld(lp + b), cf), O::ld(rp + b)));
        } else {
            o[ch] = S::oplus(S::apply(lp[ch], cf), rp[ch]);
        }
    }
}

// EV_WAVE: 64 lanes per rule.
template <class S>
__device__ void ph_eval_wave(const SweepArgs& a, int i0, int i1, const typename S::T* src,
                             typename S::T* dst, long wid, long nw, int lane) {
    using T = typename S::T;
    const T* Tt = tp<S>(a.Tv);
    const int Bt = a.Bs;
    for (long i = i0 + wid; i < i1; i += nw) {
        const int2 k = a.kids[i];
        const float cf = a.coeff ? a.coeff[i] : 1.0f;
        const T* lp = k.x >= 0 ? src + static_cast<long>(k.x) * Bt : Tt + static_cast<long>(-k.x - 1) * Bt;
        const T* rp = k.y >= 0 ? src + static_cast<long>(k.y) * Bt : Tt + static_cast<long>(-k.y - 1) * Bt;
        T* o = dst + (i - i0) * Bt;
        for (int b = lane; b < Bt; b += 64) o[b] = S::oplus(S::apply(lp[b], cf), rp[b]);
    }
}

// Workgroup-wide min and max (every thread of the workgroup must call it).
__device__ __forceinline__ void block_minmax(int& lo, int& hi) {
    __shared__ int red[2][16];
    for (int d = 32; d >= 1; d >>= 1) {
        lo = min(lo, __shfl_xor(lo, d, 64));
        hi = max(hi, __shfl_xor(hi, d, 64));
    }
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = lo;
        red[1][w] = hi;
    }
    __syncthreads();
    lo = red[0][0];
    hi = red[1][0];
    for (int i = 1; i < nw; ++i) {
        lo = min(lo, red[0][i]);
        hi = max(hi, red[1][i]);
    }
}

// EV_LDS (single vector): one rule per thread; blocks walk the level in
// blockDim-sized chunks, all threads of a block run the same iterations.
template <class S>
__device__ void ph_eval_lds(const SweepArgs& a, int i0, int i1, const typename S::T* src,
                            typename S::T* dst, long bid, long nb, typename S::T* win, int wsize) {
    using T = typename S::T;
    const T* Tt = tp<S>(a.Tv);
    for (long base = i0 + bid * blockDim.x; base < i1; base += nb * blockDim.x) {
        const long i = base + threadIdx.x;
        int2 k = make_int2(-1, -1);
        float cf = 1.0f;
        if (i < i1) {
            k = a.kids[i];
            if (a.coeff) cf = a.coeff[i];
        }
        int lo = k.x >= 0 ? k.x : 0x7fffffff, hi = k.x >= 0 ? k.x : -1;
        block_minmax(lo, hi);
        const bool staged = hi >= lo && hi - lo < wsize;
        if (staged)
            for (int t = threadIdx.x; t <= hi - lo; t += blockDim.x) win[t] = src[lo + t];
        __syncthreads();
        if (i < i1) {
            T lv = k.x >= 0 ? (staged ? win[k.x - lo] : src[k.x]) : Tt[-k.x - 1];
            T rv = k.y >= 0 ? src[k.y] : Tt[-k.y - 1];
            dst[i - i0] = S::oplus(S::apply(lv, cf), rv);
        }
        __syncthreads();
    }
}

// The level's kernel: kind as given, EV_LDS only for a single vector. `win`
// (wsize values of LDS) is the EV_LDS window: dynamic LDS sized by the launch
// in the per-level kernels, a fixed kLdsWindow array in the persistent ones.
template <class S>
__device__ __forceinline__ void ph_eval_kind(const SweepArgs& a, int kind, int i0, int i1,
                                             const typename S::T* src, typename S::T* dst, long tid,
                                             long nth, typename S::T* win, int wsize) {
    if (kind == EV_LDS && a.Bt != 1) kind = EV_GROUP;
    switch (kind) {
        case EV_THREAD: ph_eval_thread<S>(a, i0, i1, src, dst, tid, nth); break;
        case EV_WAVE: ph_eval_wave<S>(a, i0, i1, src, dst, tid / 64, nth / 64, static_cast<int>(tid & 63)); break;
        case EV_LDS: ph_eval_lds<S>(a, i0, i1, src, dst, tid / blockDim.x, nth / blockDim.x, win, wsize); break;
        default:
            if (a.vec) ph_eval_v4x4<S, 4>(a, i0, i1, src, dst, tid / a.G, nth / a.G, static_cast<int>(tid % a.G));
            else if (a.Bs == 1) ph_eval_1xU<S, 4>(a, i0, i1, src, dst, tid, nth);
            else ph_eval<S>(a, i0, i1, src, dst, tid / a.G, nth / a.G, static_cast<int>(tid % a.G));
    }
}

template <class S>
__device__ __forceinline__ int level_kind(const SweepArgs& a, int k) {
    return a.level_kind ? a.level_kind[k] : a.eval_kind;
}

// Frontier that level k writes (k >= 1) and the source it reads.
template <class S>
__device__ __forceinline__ typename S::T* level_dst(const SweepArgs& a, int k) {
    if (a.resident) return tp<S>(a.W0) + static_cast<long>(a.level_off[k - 1]) * a.Bs;
    return (k & 1) ? tp<S>(a.W1) : tp<S>(a.W0);
}
template <class S>
__device__ __forceinline__ const typename S::T* level_src(const SweepArgs& a, int k) {
    if (a.resident) return tp<S>(a.W0);
    return (k & 1) ? tp<S>(a.W0) : tp<S>(a.W1);  // frontier k-1
}

// Global thread index with an XCD-aware block order. Workgroups are dealt to
// the 8 XCDs round-robin (workgroup b runs on XCD b mod 8) and each XCD has
// its own L2; renumbering so XCD x owns the contiguous block range
// [x*n/8, (x+1)*n/8) keeps neighbouring items (rules sorted by left child,
// gather segments sorted by shared rule) on one L2. Grids are multiples of 8.
__device__ __forceinline__ long xcd_tid() {
    unsigned nb = gridDim.x, b = blockIdx.x;
    if ((nb & 7u) == 0u) b = (b & 7u) * (nb >> 3) + (b >> 3);
    return static_cast<long>(b) * blockDim.x + threadIdx.x;
}

// ---- one launch per phase (levels / graph variants) ----

template <class S>
__global__ void k_init(SweepArgs a) {
    long tid = xcd_tid();
    ph_init<S>(a, tid, static_cast<long>(gridDim.x) * blockDim.x);
}

// Dynamic LDS of the per-level launches (the EV_LDS window, the range of the
// LDS gathers); the launch gives its size.
extern __shared__ __align__(16) unsigned char dyn_lds[];

template <class S>
__global__ void k_eval(SweepArgs a, int k, int kind) {
    long tid = xcd_tid();
    long nth = static_cast<long>(gridDim.x) * blockDim.x;
    ph_eval_kind<S>(a, kind, a.level_off[k - 1], a.level_off[k], level_src<S>(a, k), level_dst<S>(a, k), tid,
                    nth, reinterpret_cast<typename S::T*>(dyn_lds), a.lds_window);
}

// Two levels of a one-vector sweep in one launch (Resident layout,
// GMR_RUN_PAIRED_LEVELS): the rules of levels k and k + 1 together. A level
// k + 1 rule computes each child that lies in level k from that child's own
// kids, whose values earlier launches wrote, so it need not wait for level k's
// stores: three rounds of loads (kids; the children's kids and values; the
// grandchildren's values) in one launch instead of two rounds in each of two.
// U rules per lane; every load is unconditional, and a load whose value is not
// used reads terminal 0 (the semiring zero, written by the init phase).
constexpr int kPairRules = 4;
// Pairs hold at most this many rules: larger levels are bound by their loads, not by the
// launch, and their recomputed children would only add loads.
constexpr int kPairMaxRules = 1 << 18;

template <class S>
__global__ void k_eval_pair(SweepArgs a, int k) {
    using T = typename S::T;
    constexpr int U = kPairRules;
    const T* Tt = tp<S>(a.Tv);
    T* W = tp<S>(a.W0);
    const int lo = a.level_off[k - 1], mid = a.level_off[k], hi = a.level_off[k + 1];
    const long tid = xcd_tid(), nth = static_cast<long>(gridDim.x) * blockDim.x;
    auto at = [&](int c) -> const T* { return c >= 0 ? W + c : Tt + (-c - 1); };
    for (long i0 = lo + tid; i0 < hi; i0 += U * nth) {
        int2 kc[U];
        float cf[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long r = i0 + u * nth;
            const long rc = r < hi ? r : hi - 1;
            kc[u] = a.kids[rc];
            cf[u] = a.coeff ? a.coeff[rc] : 1.0f;
        }
        int2 gl[U], gr[U];
        float fl[U], fr[U];
        T vl[U], vr[U];
        bool il[U], ir[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool up = i0 + u * nth >= mid;  // a level k + 1 rule
            il[u] = up && kc[u].x >= lo && kc[u].x < mid;
            ir[u] = up && kc[u].y >= lo && kc[u].y < mid;
            gl[u] = a.kids[il[u] ? kc[u].x : 0];
            gr[u] = a.kids[ir[u] ? kc[u].y : 0];
            fl[u] = a.coeff ? a.coeff[il[u] ? kc[u].x : 0] : 1.0f;
            fr[u] = a.coeff ? a.coeff[ir[u] ? kc[u].y : 0] : 1.0f;
            vl[u] = *(il[u] ? Tt : at(kc[u].x));
            vr[u] = *(ir[u] ? Tt : at(kc[u].y));
        }
        T al[U], bl[U], ar[U], br[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            al[u] = *(il[u] ? at(gl[u].x) : Tt);
            bl[u] = *(il[u] ? at(gl[u].y) : Tt);
            ar[u] = *(ir[u] ? at(gr[u].x) : Tt);
            br[u] = *(ir[u] ? at(gr[u].y) : Tt);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long r = i0 + u * nth;
            if (r < hi) {
                const T l = il[u] ? S::oplus(S::apply(al[u], fl[u]), bl[u]) : vl[u];
                const T rv = ir[u] ? S::oplus(S::apply(ar[u], fr[u]), br[u]) : vr[u];
                W[r] = S::oplus(S::apply(l, cf[u]), rv);
            }
        }
    }
}

// The top levels of a one-vector sweep in one workgroup (Resident layout,
// GMR_RUN_FUSED_TAIL): levels k0 .. L, at most kTailRules rules together, each
// a launch of its own otherwise. One workgroup alone would fetch the tail's
// children from below the tail slowly (one CU's L1 looks up about one line
// per clock, and these lookups are random), so a staging launch over many
// workgroups first copies them into stage[] in the tail's rule order; the
// tail workgroup then reads kids and staged words coalesced into LDS, runs the
// levels out of LDS with a workgroup barrier between them and writes the
// tail's values to the resident array. Two launches replace one per level.
constexpr int kTailBlock = 1024;
constexpr int kTailRules = 8192;
constexpr int kTailLdsBytes = kTailRules * 20;  // per rule: its value, kids and two staged words

// The 32-bit word of a value or a coefficient in the staging buffer.
__device__ __forceinline__ unsigned tail_word(float v) { return __float_as_uint(v); }
__device__ __forceinline__ unsigned tail_word(int v) { return static_cast<unsigned>(v); }
template <class T> __device__ __forceinline__ T tail_value(unsigned w);
template <> __device__ __forceinline__ float tail_value<float>(unsigned w) { return __uint_as_float(w); }
template <> __device__ __forceinline__ int tail_value<int>(unsigned w) { return static_cast<int>(w); }

// Two words per tail rule i, at 2(i - base): the left one holds the left
// child's value with the rule's coefficient applied when that child lies below
// the tail (terminal or lower rule), else the coefficient itself; the right one
// the right child's value when it lies below the tail (else unused).
template <class S>
__global__ void k_tail_stage(SweepArgs a, int k0, unsigned* stage) {
    using T = typename S::T;
    const T* Tt = tp<S>(a.Tv);
    const T* W = tp<S>(a.W0);
    const int base = a.level_off[k0 - 1], end = a.level_off[a.L];
    for (long i = base + static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < end;
         i += static_cast<long>(gridDim.x) * blockDim.x) {
        const int2 c = a.kids[i];
        const float cf = a.coeff ? a.coeff[i] : 1.0f;
        const T l = *(c.x < 0 ? Tt + (-c.x - 1) : W + (c.x < base ? c.x : 0));
        const T r = *(c.y < 0 ? Tt + (-c.y - 1) : W + (c.y < base ? c.y : 0));
        stage[2 * (i - base)] = c.x >= base ? __float_as_uint(cf) : tail_word(S::apply(l, cf));
        stage[2 * (i - base) + 1] = tail_word(r);
    }
}

template <class S>
__global__ __launch_bounds__(kTailBlock) void k_eval_tail(SweepArgs a, int k0, const unsigned* stage) {
    using T = typename S::T;
    const int base = a.level_off[k0 - 1], end = a.level_off[a.L], n = end - base;
    T* tv = reinterpret_cast<T*>(dyn_lds);                        // n values
    int4* meta = reinterpret_cast<int4*>(dyn_lds + 4 * kTailRules);  // n x (left, right, left word, right word)
    T* W = tp<S>(a.W0);
    // kids and staged words into LDS: eight rules per lane in flight
    for (int t0 = 0; t0 < n; t0 += 8 * kTailBlock) {
        int2 c[8];
        uint2 w[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int t = t0 + u * kTailBlock + static_cast<int>(threadIdx.x);
            const int tc = t < n ? t : n - 1;
            c[u] = a.kids[base + tc];
            w[u] = *reinterpret_cast<const uint2*>(stage + 2 * tc);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int t = t0 + u * kTailBlock + static_cast<int>(threadIdx.x);
            if (t < n) meta[t] = make_int4(c[u].x, c[u].y, static_cast<int>(w[u].x), static_cast<int>(w[u].y));
        }
    }
    __syncthreads();
    for (int k = k0; k <= a.L; ++k) {
        const int lo = a.level_off[k - 1], hi = a.level_off[k];
        for (int i = lo + static_cast<int>(threadIdx.x); i < hi; i += kTailBlock) {
            const int4 m = meta[i - base];
            const T l = tv[m.x >= base ? m.x - base : 0];
            const T r = tv[m.y >= base ? m.y - base : 0];
            const unsigned lw = static_cast<unsigned>(m.z), rw = static_cast<unsigned>(m.w);
            const T lv = m.x >= base ? S::apply(l, __uint_as_float(lw)) : tail_value<T>(lw);
            const T rv = m.y >= base ? r : tail_value<T>(rw);
            tv[i - base] = S::oplus(lv, rv);
        }
        __syncthreads();
    }
    for (int t = threadIdx.x; t < n; t += kTailBlock) W[base + t] = tv[t];
}

// k = 0: terminal emissions; k >= 1: level k's emissions.
template <class S>
__global__ void k_emit(SweepArgs a, int k) {
    long tid = xcd_tid();
    long nth = static_cast<long>(gridDim.x) * blockDim.x;
    int lane = static_cast<int>(tid % a.G);
    if (k == 0) {
        ph_emit<S>(a, a.term_emits, a.n_term, tp<S>(a.Tv), tid, nth, lane);
    } else {
        const int e0 = a.emit_off[k - 1], e1 = a.emit_off[k];
        // Resident positions are global slots; double-buffered ones are local.
        const typename S::T* src = a.resident ? tp<S>(a.W0) : level_dst<S>(a, k);
        ph_emit<S>(a, a.emits + e0, e1 - e0, src, tid, nth, lane);
    }
}

// Items [r0, r1) of the gather phase (the left product runs it level by level).
template <class S>
__global__ void k_gather_range(SweepArgs a, long r0, long r1) {
    const long tid = xcd_tid();
    const long nth = static_cast<long>(gridDim.x) * blockDim.x;
    ph_gather_range<S>(a, r0, r1, tid, nth);
}

template <class S>
__global__ void k_gather(SweepArgs a) {
    if (a.xparts) {
        ph_gather_parts_any<S>(a, blockIdx.x, gridDim.x, threadIdx.x, blockDim.x);
        return;
    }
    long tid = xcd_tid();
    long nth = static_cast<long>(gridDim.x) * blockDim.x;
    ph_gather<S>(a, tid, nth);
}

// ---- Boolean, one vector: values as bits ----
//
// Every value of the Boolean sweep is 0 or 1, so the terminal and rule values
// fit in a bitset of (U + rules) bits: about 1 MB for a grammar of 8 M symbols,
// which stays in every XCD's L2 while the gather streams the top sequence from
// HBM. A wave computes 64 consecutive values, packs them with a ballot and
// writes one 64-bit word; loops advance whole waves together (thread counts
// and loop bounds are multiples of 64), so every ballot sees its full wave.

// Bit of a child or top-sequence symbol: c >= 0 is a rule slot, c < 0 the
// terminal -c-1 (the convention of kids and csym).
__device__ __forceinline__ bool value_bit(const unsigned long long* vb, int tw, int c) {
    const long k = c >= 0 ? (static_cast<long>(tw) << 6) + c : static_cast<long>(-c - 1);
    return ((vb[k >> 6] >> (k & 63)) & 1ull) != 0ull;
}

// Terminal bits (leaf of V and x), zeroed rule words, zeroed y rows.
__device__ void ph_bits_init(const SweepArgs& a, long tid, long nth) {
    const int lane = static_cast<int>(tid & 63);
    for (long i = tid; i < static_cast<long>(a.tw) * 64; i += nth) {
        const int q = static_cast<int>(i);
        const bool b = q >= 1 && q < a.U && BooleanOp::leaf(a.tval[q], a.X[a.tcol[q]]) != 0;
        const unsigned long long m = __ballot(b);
        if (lane == 0) a.vbits[i >> 6] = m;
    }
    const long rw = (static_cast<long>(a.level_off[a.L]) + 63) >> 6;
    for (long i = tid; i < rw; i += nth) a.vbits[a.tw + i] = 0ull;
    // y zeroed four rows per store where it is 16-byte aligned
    int* Y = static_cast<int*>(a.Y);
    const long head = (reinterpret_cast<uintptr_t>(Y) & 15u) ? static_cast<long>((16u - (reinterpret_cast<uintptr_t>(Y) & 15u)) >> 2) : 0L;
    const long h = head < a.rows ? head : static_cast<long>(a.rows);
    for (long i = tid; i < h; i += nth) Y[i] = 0;
    int4* Y4 = reinterpret_cast<int4*>(Y + h);
    const long n4 = (a.rows - h) >> 2;
    for (long i = tid; i < n4; i += nth) Y4[i] = make_int4(0, 0, 0, 0);
    for (long i = h + 4 * n4 + tid; i < a.rows; i += nth) Y[i] = 0;
}

// Levels k and k + 1 of the bit mode in one launch (GMR_RUN_PAIRED_LEVELS): a
// level k + 1 rule computes each child in level k as the OR of that child's
// kids' bits (written by earlier launches) instead of reading its bit, so the
// two levels need no barrier between them. Words wholly inside the pair are
// stored, words shared with the levels around it ORed in.
__device__ void ph_bits_eval_pair(const SweepArgs& a, int k, long tid, long nth) {
    const int lane = static_cast<int>(tid & 63);
    const long lo = a.level_off[k - 1], mid = a.level_off[k], hi = a.level_off[k + 1];
    const long s0 = (lo >> 6) << 6, s1 = (hi + 63) >> 6 << 6;
    for (long s = s0 + tid; s < s1; s += nth) {
        // every lane loads (slots outside the pair read the nearest one in it)
        const long sc = s < lo ? lo : (s < hi ? s : hi - 1);
        const int2 c = a.kids[sc];
        const bool il = s >= mid && c.x >= lo && c.x < mid, ir = s >= mid && c.y >= lo && c.y < mid;
        const int2 gl = a.kids[il ? c.x : sc], gr = a.kids[ir ? c.y : sc];
        const int p0 = il ? gl.x : c.x, p1 = il ? gl.y : c.x, q0 = ir ? gr.x : c.y, q1 = ir ? gr.y : c.y;
        const int v = static_cast<int>(value_bit(a.vbits, a.tw, p0)) | static_cast<int>(value_bit(a.vbits, a.tw, p1)) |
                      static_cast<int>(value_bit(a.vbits, a.tw, q0)) | static_cast<int>(value_bit(a.vbits, a.tw, q1));
        const unsigned long long m = __ballot(s >= lo && s < hi && v != 0);
        if (lane == 0) {
            const long w = s >> 6;
            unsigned long long* dst = a.vbits + a.tw + w;
            if ((w << 6) >= lo && (w << 6) + 64 <= hi) *dst = m;
            else if (m) atomicOr(dst, m);
        }
    }
}

// Level k: rule slots [lo, hi), swept as the 64-slot words that hold them. A
// word wholly inside the level is stored; one shared with a neighbouring level
// is ORed in (the init phase zeroed it, and only nonzero words are sent).
__device__ void ph_bits_eval(const SweepArgs& a, int k, long tid, long nth) {
    const int lane = static_cast<int>(tid & 63);
    const long lo = a.level_off[k - 1], hi = a.level_off[k];
    const long s0 = (lo >> 6) << 6, s1 = (hi + 63) >> 6 << 6;
    for (long s = s0 + tid; s < s1; s += nth) {
        bool b = false;
        if (s >= lo && s < hi) {
            const int2 c = a.kids[s];
            b = value_bit(a.vbits, a.tw, c.x) || value_bit(a.vbits, a.tw, c.y);
        }
        const unsigned long long m = __ballot(b);
        if (lane == 0) {
            const long w = s >> 6;
            unsigned long long* dst = a.vbits + a.tw + w;
            if ((w << 6) >= lo && (w << 6) + 64 <= hi) *dst = m;
            else if (m) atomicOr(dst, m);
        }
    }
}

// Gather: 64-symbol chunks of the row-ordered top sequence, kBitChunks
// consecutive ones per wave in flight. The ballot of the symbols' bits is a
// chunk's values; a row piece ending at lane l is 1 iff its bits from the
// piece's first lane to l meet the ballot. Rows wholly inside a chunk are
// stored, pieces of longer rows ORed in. Everything but the bit words and the
// row ids is loaded up front, so a chunk costs two dependent round trips.
// Build-time tuning knobs of the bit mode (defaults: the measured best).
#ifndef GMR_BIT_RANGE_LOG2
#define GMR_BIT_RANGE_LOG2 19
#endif
#ifndef GMR_BIT_BLOCK
#define GMR_BIT_BLOCK 1024
#endif
#ifndef GMR_BIT_WG_PER_CU
#define GMR_BIT_WG_PER_CU 2
#endif
#ifndef GMR_BIT_NT
#define GMR_BIT_NT 1
#endif
constexpr int kBitChunks = 4;  // chunks per wave of the flat bit gather

__device__ void ph_bits_gather(const SweepArgs& a, long tid, long nth) {
    const int lane = static_cast<int>(tid & 63);
    int* Y = static_cast<int*>(a.Y);
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);  // bits 0..lane
    const long nch = a.n_fchunk;
    const long wid = tid >> 6, nw = nth >> 6;
    for (long c0 = wid * kBitChunks; c0 < nch; c0 += nw * kBitChunks) {
        int sym[kBitChunks], ord[kBitChunks];
        unsigned long long fb[kBitChunks], fnext[kBitChunks];
#pragma unroll
        for (int j = 0; j < kBitChunks; ++j) {
            const long c = c0 + j, p = c * 64 + lane;
            sym[j] = p < a.n_csym ? a.csym[p] : -1;  // terminal 0: always 0
            fb[j] = c < nch ? a.fbits[c] : 0ull;
            fnext[j] = c + 1 < nch ? a.fbits[c + 1] : 1ull;
            ord[j] = c < nch ? a.ford[c] : 0;
        }
        bool b[kBitChunks];
        int row[kBitChunks];
#pragma unroll
        for (int j = 0; j < kBitChunks; ++j) {
            b[j] = value_bit(a.vbits, a.tw, sym[j]);
            const unsigned long long st = fb[j] & upto;
            const long p = (c0 + j) * 64 + lane;
            const bool last = p < a.n_csym && (lane == 63 || p + 1 == a.n_csym || ((fb[j] >> (lane + 1)) & 1ull) != 0ull);
            row[j] = last ? a.frow[ord[j] - 1 + __popcll(st)] : 0;
        }
#pragma unroll
        for (int j = 0; j < kBitChunks; ++j) {
            if (c0 + j >= nch) break;  // wave-uniform
            const unsigned long long m = __ballot(b[j]);
            const long p = (c0 + j) * 64 + lane;
            const bool last = p < a.n_csym && (lane == 63 || p + 1 == a.n_csym || ((fb[j] >> (lane + 1)) & 1ull) != 0ull);
            if (last) {
                const unsigned long long st = fb[j] & upto;
                const int first = st ? 63 - __clzll(st) : 0;
                const bool v = (m & upto & (~0ull << first)) != 0ull;
                // lane 63 ends its row iff the next symbol starts one (or none follows)
                const bool ends_here = lane < 63 || p + 1 == a.n_csym || (fnext[j] & 1ull) != 0ull;
                if (st != 0ull && ends_here) Y[row[j]] = v ? 1 : 0;
                else if (v) Y[row[j]] = 1;  // a piece: OR into the zeroed row (BooleanOp::emit)
            }
        }
    }
}

// The list that holds chunk c of range lists (xpoff: first chunk of every
// list, non-decreasing, c < xpoff[nbl]): the number of lists that end at or
// before c, counted by the whole workgroup at once -- a walk from list 0 would
// be one dependent load per list. `cnt` is a word of LDS, free until the
// caller's next barrier.
__device__ __forceinline__ int list_of_chunk(const SweepArgs& a, long c, unsigned* cnt) {
    if (threadIdx.x == 0) *cnt = 0u;
    __syncthreads();
    unsigned mine = 0u;
    for (int i = threadIdx.x; i < a.nbl; i += blockDim.x) mine += a.xpoff[i + 1] <= c ? 1u : 0u;
    if (mine) atomicAdd(cnt, mine);
    __syncthreads();
    const int p = static_cast<int>(*cnt);
    __syncthreads();  // every lane has read the word before LDS is refilled
    return p;
}

// Lane j's value of v, as a wave-uniform value.
__device__ __forceinline__ unsigned long long readlane64(unsigned long long v, int j) {
    const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v), j));
    const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v >> 32), j));
    return (static_cast<unsigned long long>(hi) << 32) | lo;
}

// Loads issued per lane before any is used in the LDS range fills below: a
// fill then waits for one round trip (two for terminal leaves, whose column
// is loaded first) per kFillBatch values a lane writes, not per value.
constexpr int kFillBatch = 16;

// Gather from LDS: list p holds the symbols whose bits lie in range p
// ([p * a.bit_range, (p + 1) * a.bit_range), 64 KiB by default). Workgroup w
// takes the w-th equal share of all lists' chunks; for each list its share
// touches it copies that range of the bitset into LDS (dynamic, a.bit_range / 8
// bytes), then its waves sweep the chunks (K per wave in flight) looking bits
// up in LDS. A row's symbols sit in several lists, so every row piece that is
// 1 is ORed into the zeroed y (by storing 1, as BooleanOp::emit does).
constexpr int kBitRangeBits = 1 << GMR_BIT_RANGE_LOG2;
constexpr int kBitGatherBlock = GMR_BIT_BLOCK;
constexpr int kBitWgPerCu = GMR_BIT_WG_PER_CU;
// Waves per SIMD of kBitWgPerCu workgroups per CU (two of 1,024 lanes: eight,
// which leaves 64 VGPRs a lane). The kernel is held to that register count, so
// eight chunks per wave in flight still run that many workgroups per CU.
constexpr int kBitWavesPerSimd = kBitWgPerCu * kBitGatherBlock / 64 / 4;

template <int K>
__global__ __launch_bounds__(kBitGatherBlock, kBitWavesPerSimd) void k_bits_gather_lds(SweepArgs a) {
    unsigned* lds = reinterpret_cast<unsigned*>(dyn_lds);
    const int range = a.bit_range;
    const int lane = threadIdx.x & 63;
    // wave-uniform, so a chunk's words (row starts, row order) are read at one uniform address
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    const int nwave = blockDim.x >> 6;
    int* Y = static_cast<int*>(a.Y);
    const unsigned* loc = reinterpret_cast<const unsigned*>(a.xsym);  // bit index inside the list's range
    const long N = a.xpoff[a.nbl];
    const long g0 = N * blockIdx.x / gridDim.x, g1 = N * (blockIdx.x + 1) / gridDim.x;
    int p = g0 < g1 ? list_of_chunk(a, g0, lds) : 0;
    for (long c0 = g0; c0 < g1;) {
        while (a.xpoff[p + 1] <= c0) ++p;  // ends below nbl, since c0 < N
        const long c1 = a.xpoff[p + 1] < g1 ? static_cast<long>(a.xpoff[p + 1]) : g1;
        __syncthreads();  // the previous range's lookups are done
        const long w0 = static_cast<long>(p) * (range / 64);
        const int nw = range / 64;
        for (int i0 = threadIdx.x; i0 < nw; i0 += kFillBatch * blockDim.x) {
            unsigned long long v[kFillBatch];
#pragma unroll
            for (int u = 0; u < kFillBatch; ++u) {
                const long w = w0 + i0 + static_cast<long>(u) * blockDim.x;
                v[u] = a.vbits[w < a.nvw ? w : a.nvw - 1];  // clamped: unconditional
            }
#pragma unroll
            for (int u = 0; u < kFillBatch; ++u) {
                const int i = i0 + u * static_cast<int>(blockDim.x);
                if (i < nw) {
                    const unsigned long long x = w0 + i < a.nvw ? v[u] : 0ull;
                    lds[2 * i] = static_cast<unsigned>(x);
                    lds[2 * i + 1] = static_cast<unsigned>(x >> 32);
                }
            }
        }
        __syncthreads();
        // Every load is unconditional (chunk indices clamped into the segment,
        // row indices to a valid entry), so a stage's loads are in flight
        // together: a conditional load would end in a wait for everything
        // outstanding. Software-pipelined: an iteration issues its row-id loads,
        // then the next iteration's symbol and chunk-word loads, so the two
        // round trips overlap.
        const long step = static_cast<long>(nwave) * K;
        long c = c0 + static_cast<long>(wave) * K;
        unsigned l[K];
        // lane j < K holds chunk c + j's row-start word and row order (one load
        // for the K chunks; read out per chunk where they are used)
        unsigned long long fw;
        int ow;
        auto load_chunks = [&](long cc) {
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const long cj = cc + j < c1 ? cc + j : c1 - 1;
#if GMR_BIT_NT
                l[j] = __builtin_nontemporal_load(loc + cj * 64 + lane);
#else
                l[j] = loc[cj * 64 + lane];
#endif
            }
            const long cl = cc + (lane < K ? lane : 0);
            const long cw = cl < c1 ? cl : c1 - 1;
            fw = a.xbits[cw];
            ow = a.xord[cw];
        };
        load_chunks(c);  // (past the segment: its last chunk again, unused)
        while (c < c1) {
            bool hit[K];
            int ri[K];
            unsigned long long fb[K];
            int ord[K];
#pragma unroll
            for (int j = 0; j < K; ++j) {
                fb[j] = readlane64(fw, j);
                ord[j] = __builtin_amdgcn_readlane(ow, j);
            }
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const bool in_range = l[j] < static_cast<unsigned>(range);
                const unsigned li = in_range ? l[j] : 0u;
                const bool b = (((lds[li >> 5] >> (li & 31)) & 1u) & (in_range ? 1u : 0u)) != 0u;
                const unsigned long long m = __ballot(b);
                // Which row pieces hold a 1, for the whole wave in a few scalar
                // operations: add m to x = ~starts | m, so a carry starts at
                // every 1 and runs up to the next piece start that is not a 1;
                // the carry into lane l then says whether l's piece holds a 1
                // below l. A piece ending at l holds one if m has l, or if l is
                // not a start and that carry reaches it.
                const unsigned long long S = fb[j];
                const unsigned long long x = ~S | m;
                const unsigned long long cin = (x + m) ^ x ^ m;
                const unsigned long long ends = (S >> 1) | (1ull << 63);  // lane l ends a row piece
                const unsigned long long H = c + j < c1 ? (m | (cin & ~S)) & ends : 0ull;
                hit[j] = ((H >> lane) & 1ull) != 0ull;
                // the piece's row: its order is that of the last start at or below l
                const int below = static_cast<int>(__builtin_amdgcn_mbcnt_hi(
                    static_cast<unsigned>(S >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(S), 0u)));
                ri[j] = hit[j] ? ord[j] - 1 + below + static_cast<int>((S >> lane) & 1ull) : 0;
            }
            int row[K];
#pragma unroll
            for (int j = 0; j < K; ++j) row[j] = a.xrow[ri[j]];
            const long cn = c + step;
            load_chunks(cn);  // unconditional: a branch here would make the stores wait for it
            // The row ids are used here, after the prefetch, so the compiler cannot
            // sink their loads into the stores' branches (where the wait for them
            // would also wait for the prefetch). An empty statement: no code.
#pragma unroll
            for (int j = 0; j < K; ++j) asm volatile("" : : "v"(row[j]));
#pragma unroll
            for (int j = 0; j < K; ++j)
                if (hit[j]) Y[row[j]] = 1;
            c = cn;
        }
        c0 = c1;
    }
}

// One vector, 4-byte values staged in LDS by range (GMR_RUN_LDS_RANGES): the
// same structure as the bit-mode gather above with a value per entry. List p
// holds the symbols whose value key lies in [p * lspan, (p + 1) * lspan); a
// workgroup fills LDS (dynamic, lspan values) with that range once per list
// its share of chunks touches -- terminal values computed from V and x on the
// spot, rule values copied from the resident array -- looks
// every symbol's value up there, reduces each 64-symbol chunk per row piece
// with the segmented shuffle scan, and combines every row piece atomically
// into the zeroed y (a row's symbols sit in many lists). No value lookup
// reaches the L1/L2 caches; the cost moves to one atomic per row piece.
constexpr int kLdsGatherBlock = 1024;

// Fills lds[0, span) with the values of keys k0 .. k0 + span - 1 of a
// one-vector LDS range: terminal leaves computed from their value and x
// (terminal 0 is the semiring zero), rule values copied from W, zero past the
// last key. Every load is clamped into its array, so none is conditional, and
// a lane issues FB of each before using any.
// A range of rule values only (no terminal key) is filled RF values per lane
// at a time: its loads do not depend on other loads, so with RF = 40 one round
// trip fills a 40,960-value range of a 1,024-lane workgroup. Terminal ranges
// take FB values of each array per lane per round trip.
template <class S, int FB, int RF>
__device__ __forceinline__ void fill_range(const SweepArgs& a, typename S::T* lds, long k0, int span) {
    using T = typename S::T;
    const T* W = tp<S>(a.W0);
    const long U = a.U, nw = a.nkey - a.U;  // terminal keys; rule values held in W
    if (k0 >= U && nw > 0) {
        for (int i0 = threadIdx.x; i0 < span; i0 += RF * blockDim.x) {
            T wv[RF];
#pragma unroll
            for (int u = 0; u < RF; ++u) {
                const long kw = k0 + i0 + static_cast<long>(u) * blockDim.x - U;
                wv[u] = W[kw < nw ? kw : nw - 1];
            }
#pragma unroll
            for (int u = 0; u < RF; ++u) {
                const int i = i0 + u * static_cast<int>(blockDim.x);
                if (i < span) lds[i] = k0 + i < a.nkey ? wv[u] : S::zero();
            }
        }
        return;
    }
    if (k0 + span <= U) {
        // terminal keys only: no rule value to load
        for (int i0 = threadIdx.x; i0 < span; i0 += FB * blockDim.x) {
            float tv[FB];
            int tc[FB];
#pragma unroll
            for (int u = 0; u < FB; ++u) {
                const long k = k0 + i0 + static_cast<long>(u) * blockDim.x;
                const long kt = k < U ? k : U - 1;
                tv[u] = a.tval[kt];
                tc[u] = a.tcol[kt];
            }
            float xv[FB];
#pragma unroll
            for (int u = 0; u < FB; ++u) xv[u] = a.X[static_cast<long>(tc[u]) * a.B];
#pragma unroll
            for (int u = 0; u < FB; ++u) {
                const int i = i0 + u * static_cast<int>(blockDim.x);
                if (i < span) lds[i] = k0 + i == 0 ? S::zero() : S::leaf(tv[u], xv[u]);
            }
        }
        return;
    }
    for (int i0 = threadIdx.x; i0 < span; i0 += FB * blockDim.x) {
        float tv[FB];
        int tc[FB];
        T wv[FB];
#pragma unroll
        for (int u = 0; u < FB; ++u) {
            const long k = k0 + i0 + static_cast<long>(u) * blockDim.x;
            const long kt = k < U ? k : U - 1;
            tv[u] = a.tval[kt];
            tc[u] = a.tcol[kt];
            const long kw = k - U;
            wv[u] = nw > 0 ? W[kw < 0 ? 0 : (kw < nw ? kw : nw - 1)] : S::zero();
        }
        float xv[FB];
#pragma unroll
        for (int u = 0; u < FB; ++u) xv[u] = a.X[static_cast<long>(tc[u]) * a.B];
#pragma unroll
        for (int u = 0; u < FB; ++u) {
            const int i = i0 + u * static_cast<int>(blockDim.x);
            const long k = k0 + i;
            if (i < span) lds[i] = k < U ? (k == 0 ? S::zero() : S::leaf(tv[u], xv[u])) : (k < a.nkey ? wv[u] : S::zero());
        }
    }
}

// Every entry carries its row as an offset from its chunk's first row (the high
// 16 bits; the index inside the range is the low 16 bits), so the row of a
// piece comes from the entry that ends it, with no dependent load.
// WPS: waves per SIMD the launch bounds ask for. 4 is one 1,024-lane
// workgroup per CU (ranges above 80 KiB), with fills of 16 terminal or 40 rule
// values per lane per round trip; 8 is two workgroups per CU (ranges up to 80
// KiB each), which caps a lane at 64 VGPRs and so takes 8 and 24.
template <class S, int K, int WPS>
__global__ __launch_bounds__(kLdsGatherBlock, WPS) void k_gather_lds(SweepArgs a) {
    static_assert(WPS == 4 || WPS == 8, "one or two workgroups per CU");
    using T = typename S::T;
    T* lds = reinterpret_cast<T*>(dyn_lds);
    const int span = a.lspan;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    const int nwave = blockDim.x >> 6;
    T* Y = tp<S>(a.Y);
    const unsigned* loc = reinterpret_cast<const unsigned*>(a.xsym);  // key index inside the list's range
    const long N = a.xpoff[a.nbl];
    const long g0 = N * blockIdx.x / gridDim.x, g1 = N * (blockIdx.x + 1) / gridDim.x;
    int p = g0 < g1 ? list_of_chunk(a, g0, reinterpret_cast<unsigned*>(dyn_lds)) : 0;
    for (long c0 = g0; c0 < g1;) {
        while (a.xpoff[p + 1] <= c0) ++p;  // ends below nbl, since c0 < N
        const long c1 = a.xpoff[p + 1] < g1 ? static_cast<long>(a.xpoff[p + 1]) : g1;
        __syncthreads();  // the previous range's lookups are done
        const long k0 = static_cast<long>(p) * span;
        // (the init phase computed only the terminal values rules read)
        fill_range<S, WPS == 4 ? 16 : 8, WPS == 4 ? 40 : 24>(a, lds, k0, span);
        __syncthreads();
        // As in the bit-mode gather: unconditional, clamped loads, and the next
        // iteration's chunk loads issued before this one's emits.
        const long step = static_cast<long>(nwave) * K;
        long c = c0 + static_cast<long>(wave) * K;
        unsigned l[K];
        // lane j < K holds chunk c + j's row-start word and first row: one load
        // for the K chunks, read out per chunk where used
        unsigned long long fw;
        int ow;
        auto load_chunks = [&](long cc) {
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const long cj = cc + j < c1 ? cc + j : c1 - 1;
                l[j] = __builtin_nontemporal_load(loc + cj * 64 + lane);
            }
            const long cl = cc + (lane < K ? lane : 0);
            const long cw = cl < c1 ? cl : c1 - 1;
            fw = a.xbits[cw];
            ow = a.cbase[cw];
        };
        load_chunks(c);
        while (c < c1) {
            T v[K];
            bool end[K];
            unsigned long long fb[K];
            int base[K];  // the row of chunk j's first entry
#pragma unroll
            for (int j = 0; j < K; ++j) {
                fb[j] = readlane64(fw, j);
                base[j] = __builtin_amdgcn_readlane(ow, j);
            }
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const unsigned li = l[j] & 0xFFFFu;
                v[j] = lds[li < static_cast<unsigned>(span) ? li : 0u];
            }
#pragma unroll
            for (int j = 0; j < K; ++j) {
                if ((l[j] & 0xFFFFu) >= static_cast<unsigned>(span)) v[j] = S::zero();  // padding
                end[j] = false;
                // A chunk whose values are all ones the emission skips (the
                // semiring zero of min/plus and Boolean) has no piece to emit.
                if (__ballot(!S::skip(v[j])) == 0ull) continue;
                const unsigned long long bits = fb[j];
                // Segmented scan over the lanes of each row piece (DPP; the K
                // chunks' steps interleave). dist: lanes back to the last piece
                // start at or below this lane (all of them when there is none).
                const unsigned long long sh = bits << (63 - lane);
                const int dist = sh ? static_cast<int>(__clzll(sh)) : lane;
                v[j] = seg_scan<S>(v[j], dist, lane);
                const bool ends = lane == 63 || ((bits >> (lane + 1)) & 1ull) != 0ull;
                end[j] = c + j < c1 && ends && !S::skip(v[j]);
            }
            int row[K];
#pragma unroll
            for (int j = 0; j < K; ++j) row[j] = base[j] + static_cast<int>(l[j] >> 16);
            const long cn = c + step;
            load_chunks(cn);
#pragma unroll
            for (int j = 0; j < K; ++j) asm volatile("" : : "v"(row[j]));
#pragma unroll
            for (int j = 0; j < K; ++j)
                if (end[j]) S::emit(Y + row[j], v[j]);
            c = cn;
        }
        c0 = c1;
    }
}

// Y of the lists' row order into the caller's Y in matrix row order (one vector).
template <class S>
__global__ void k_rows_to_matrix_order(SweepArgs a) {
    using T = typename S::T;
    const T* Yr = tp<S>(a.Y);
    T* Y = tp<S>(a.Yout);
    const long n = a.rows, nth = static_cast<long>(gridDim.x) * blockDim.x;
    for (long r = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; r < n; r += nth) Y[r] = Yr[a.ynew[r]];
}

__global__ void k_bits_init(SweepArgs a) { ph_bits_init(a, xcd_tid(), static_cast<long>(gridDim.x) * blockDim.x); }

__global__ void k_bits_eval(SweepArgs a, int k) {
    ph_bits_eval(a, k, xcd_tid(), static_cast<long>(gridDim.x) * blockDim.x);
}

__global__ void k_bits_eval_pair(SweepArgs a, int k) {
    ph_bits_eval_pair(a, k, xcd_tid(), static_cast<long>(gridDim.x) * blockDim.x);
}

__global__ void k_bits_gather(SweepArgs a) {
    ph_bits_gather(a, xcd_tid(), static_cast<long>(gridDim.x) * blockDim.x);
}

// ---- persistent variant: the whole sweep in one launch ----
//
// Grid-wide barrier of our own. Every workgroup's first thread publishes the
// workgroup's writes (agent-scope release), adds one arrival to a monotonic
// 64-bit counter and waits for the counter to reach `target` (all workgroups of
// this barrier), then acquires. Every wait is bounded: after `deadline` ticks
// of the 100 MHz wall clock it stops waiting, records the launch's epoch as
// failed and sets a word in pinned host memory (so the host sees the failure
// without synchronising); waits of the same launch see the failed epoch and
// stop too, so the kernel always ends. Every workgroup still makes all of its
// arrivals, so the counter stays consistent across launches, and a later
// launch (another epoch) is not affected by an earlier one's failure.
struct GridBarrier {
    unsigned long long arrived;  // monotonic across launches
    unsigned long long failed;   // epoch of the last launch whose wait timed out (0: none)
};

// One launch's view of the barrier.
struct BarrierRef {
    GridBarrier* gb;
    int* fail_host;               // device address of the pinned host word set on a timeout
    unsigned long long base;      // arrivals made by earlier launches
    unsigned long long deadline;  // wall-clock ticks a wait may last
};

__device__ __forceinline__ void grid_barrier(const BarrierRef& br, unsigned long long target) {
    __syncthreads();
    if (threadIdx.x == 0) {
        GridBarrier* gb = br.gb;
        const unsigned long long epoch = br.base + 1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(&gb->arrived, 1ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long t0 = wall_clock64();
        while (__hip_atomic_load(&gb->arrived, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(&gb->failed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch) break;
            if (wall_clock64() - t0 > br.deadline) {
                __hip_atomic_store(&gb->failed, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(br.fail_host, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

// Phases (L + 1 barriers per launch):
//   phase 0        T (and Y for spot emission) initialised            | barrier
//   phase k        level k; spot: also level k-1's emits (terminal    | barrier
//                  emits for k = 1) -- both only read frontier k-1
//   after level L  spot: level L's emits; gather: the row gather
// `base` is the arrival count before this launch.
// One sweep of the tile described by `a`; `nbar` counts the barriers passed.
template <class S>
__device__ void persistent_sweep(const SweepArgs& a, const BarrierRef& br, unsigned long long& nbar) {
    const long tid = xcd_tid();
    const long nth = static_cast<long>(gridDim.x) * blockDim.x;
    const int lane = static_cast<int>(tid % a.G);
    const unsigned long long nb = gridDim.x;
    ph_init<S>(a, tid, nth);
    grid_barrier(br, br.base + nb * (++nbar));
    for (int k = 1; k <= a.L; ++k) {
        if (!a.gather) {
            if (k == 1) {
                ph_emit<S>(a, a.term_emits, a.n_term, tp<S>(a.Tv), tid, nth, lane);
            } else {
                const int e0 = a.emit_off[k - 2], e1 = a.emit_off[k - 1];
                const typename S::T* src = a.resident ? tp<S>(a.W0) : level_dst<S>(a, k - 1);
                ph_emit<S>(a, a.emits + e0, e1 - e0, src, tid, nth, lane);
            }
        }
        __shared__ typename S::T win[kLdsWindow];
        ph_eval_kind<S>(a, level_kind<S>(a, k), a.level_off[k - 1], a.level_off[k], level_src<S>(a, k),
                        level_dst<S>(a, k), tid, nth, win, kLdsWindow);
        grid_barrier(br, br.base + nb * (++nbar));
    }
    if (a.gather && a.xparts) {
        ph_gather_parts_any<S>(a, blockIdx.x, gridDim.x, threadIdx.x, blockDim.x);
    } else if (a.gather) {
        ph_gather<S>(a, tid, nth);
    } else if (a.L == 0) {
        ph_emit<S>(a, a.term_emits, a.n_term, tp<S>(a.Tv), tid, nth, lane);
    } else {
        const int e0 = a.emit_off[a.L - 1], e1 = a.emit_off[a.L];
        const typename S::T* src = a.resident ? tp<S>(a.W0) : level_dst<S>(a, a.L);
        ph_emit<S>(a, a.emits + e0, e1 - e0, src, tid, nth, lane);
    }
}

// The whole product in one launch: `tiles` vector tiles of width `tile_w`
// (the last may be narrower), one after another through the same workspace,
// with a barrier between a tile's emission and the next tile's init. `a`
// describes the first tile; the others differ only in their column base and
// width. Barriers per launch: tiles * (L + 2) - 1.
template <class S>
__global__ void k_persistent(SweepArgs a, BarrierRef br, int tiles, int tile_w, int vbytes) {
    unsigned long long nbar = 0;
    for (int t = 0; t < tiles; ++t) {
        if (t > 0) {
            const unsigned long long nb = gridDim.x;
            grid_barrier(br, br.base + nb * (++nbar));
            const int b0 = t * tile_w;
            SweepArgs at = a;
            at.X = a.X + b0;
            at.Y = static_cast<char*>(a.Y) + static_cast<long>(b0) * vbytes;
            at.Bt = min(tile_w, a.B - b0);
            at.Bs = at.Bt <= 1 ? 1 : (at.Bt + 3) / 4 * 4;
            at.vec = at.Bt > 1 ? 1 : 0;
            at.yvec = (a.B % 4 == 0 && b0 % 4 == 0) ? 1 : 0;
            at.G = lanes_for(at.Bs, at.vec);
            persistent_sweep<S>(at, br, nbar);
        } else {
            persistent_sweep<S>(a, br, nbar);
        }
    }
}

// ---- many grammars in one persistent launch ----
//
// The host lists, phase by phase, workgroup-sized tiles of every grammar's
// work; a workgroup runs whole tiles (so a tile may use the grammar's own
// per-level kernel, LDS staging included) and the bounded grid barrier
// separates the phases. Phase 0 initialises every grammar; phase p >= 1 runs
// level p of every grammar that has one together with the emissions of its
// level p-1 (terminal emissions for p = 1); phase L_g + 1 holds grammar g's
// last emissions or its row gather.
enum TileTask : int { TT_INIT = 0, TT_EVAL = 1, TT_EMIT = 2, TT_GATHER = 3 };

struct Tile {
    int g;      // grammar
    int task;   // TileTask
    int level;  // TT_EVAL: level; TT_EMIT: level whose emissions (0 = terminals)
    int kind;   // TT_EVAL: EvalKind
    long begin, end;  // item range: init items, rules (absolute), emissions (within level), gather segments
};

template <class S>
__device__ void run_tile(const SweepArgs& a, const Tile& t) {
    const long tid = threadIdx.x, nth = blockDim.x;
    switch (t.task) {
        case TT_INIT: ph_init_range<S>(a, t.begin, t.end, tid, nth); break;
        case TT_EVAL: {
            __shared__ typename S::T win[kLdsWindow];
            ph_eval_kind<S>(a, t.kind, static_cast<int>(t.begin), static_cast<int>(t.end), level_src<S>(a, t.level),
                            level_dst<S>(a, t.level) + (t.begin - a.level_off[t.level - 1]) * a.Bs, tid, nth, win,
                            kLdsWindow);
            break;
        }
        case TT_EMIT: {
            const int lane = static_cast<int>(tid % a.G);
            if (t.level == 0) {
                ph_emit<S>(a, a.term_emits + t.begin, static_cast<int>(t.end - t.begin), tp<S>(a.Tv), tid, nth, lane);
            } else {
                const int e0 = a.emit_off[t.level - 1];
                const typename S::T* src = a.resident ? tp<S>(a.W0) : level_dst<S>(a, t.level);
                ph_emit<S>(a, a.emits + e0 + t.begin, static_cast<int>(t.end - t.begin), src, tid, nth, lane);
            }
            break;
        }
        default: ph_gather_range<S>(a, t.begin, t.end, tid, nth);
    }
}

template <class S>
__global__ void k_multi(const SweepArgs* args, const Tile* tiles, const long* phase_off, int phases, BarrierRef br) {
    const unsigned long long nb = gridDim.x;
    for (int p = 0; p < phases; ++p) {
        for (long t = phase_off[p] + blockIdx.x; t < phase_off[p + 1]; t += gridDim.x) {
            const Tile tile = tiles[t];
            run_tile<S>(args[tile.g], tile);
        }
        if (p + 1 < phases) grid_barrier(br, br.base + nb * static_cast<unsigned long long>(p + 1));
    }
}

}  // namespace dev
}  // namespace gmr
