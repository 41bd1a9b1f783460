// Device code of the left product Y^T = X^T M over a Resident-layout grammar
// (gfx950, wave64).
//
// The right product pushes values bottom-up (terminals -> rules -> rows). The
// left product runs the same DAG top-down, as a pull: a node's value is the
// semiring sum of everything that points at it --
//   A[s] = (+)_{occurrences of s in row r of the top sequence} x_r
//        (+) (+)_{rules p with s as a child} coeff(p -> s) * A[p]
// -- and every parent of a rule sits on a higher level, so levels L..1 run in
// order, then the terminals. A column then collects its terminals:
// y_c = (+)_{terminals t in column c} leaf(V_t, A[t]).
// Incoming lists are CSR over nodes (rule slots, then the terminals grouped
// by column, so the fold reads them in sequence); an entry
// >= 0 is a parent slot, < 0 is a row -e-1. That is the shape of the right
// product's row gather (nodes for rows, entries for top-sequence symbols, x as
// a semiring value for the terminal rows), so without run-length rules the
// levels run on the row-gather kernels (k_gather_range); with them, on
// k_left_nodes below, which applies the coefficients. Vectors are innermost
// as in the right product: X[r*B + b], A[node*Bs + b], Y[c*B + b].
#pragma once

#include <hip/hip_runtime.h>

#include "sweep_kernels.hpp"

namespace gmr {
namespace dev {

// x as a semiring value, and the leaf with the value on the left, for each
// policy (the same algebra as the right product's leaf(V, x)).
template <class S> struct LeftOps;
template <> struct LeftOps<PlusTimesOp> {
    static __device__ __forceinline__ float from_x(float x) { return x; }
    static __device__ __forceinline__ float leaf(float v, float a) { return v * a; }
};
template <> struct LeftOps<BooleanOp> {
    static __device__ __forceinline__ int from_x(float x) { return x != 0.0f ? 1 : 0; }
    static __device__ __forceinline__ int leaf(float v, int a) { return (v != 0.0f && a != 0) ? 1 : 0; }
};
template <> struct LeftOps<TropicalOp> {
    static __device__ __forceinline__ int from_x(float x) {
        return x >= static_cast<float>(kTropicalInf) ? kTropicalInf : static_cast<int>(x);
    }
    static __device__ __forceinline__ int leaf(float v, int a) {
        const int iv = static_cast<int>(v);
        return (iv >= kTropicalInf || a >= kTropicalInf) ? kTropicalInf : iv + a;
    }
};

struct LeftArgs {
    const int* in_off;     // nodes + 1: incoming entries of node n are [in_off[n], in_off[n+1])
    const int* in_src;     // >= 0 parent slot, < 0 row -e-1
    const float* in_coeff; // per entry (nullptr when every coefficient is 1)
    const int* col_off;    // cols + 1: column c's terminals are positions [col_off[c], col_off[c+1])
    const float* col_val;  // terminal value per position
    const unsigned long long* fbits;  // one vector: column starts per 64-position chunk (positions from 1)
    const int* ford;       // column starts before each chunk
    const int* fcol;       // non-empty columns in order
    long n_fchunk;
    int npos;              // positions 0..npos-1 (npos = U; 0 is the zero terminal)
    const float* X;        // rows x B (this tile's column base folded in)
    void* Y;               // cols x B
    void* A;               // nodes x Bs node values
    int NL;                // rule slots; terminal position p is node NL + p
    int B, Bt, Bs, G;
    int xparts;            // one vector: the XCD-partitioned lists (else one chunked list)
};

// Nodes [n0, n1) with coefficients: a group of G lanes per node, lanes over
// the vectors, entries one after another.
template <class S>
__device__ void ph_left_nodes(const LeftArgs& a, int n0, int n1, long gid, long ng, int lane) {
    using T = typename S::T;
    using L = LeftOps<S>;
    T* A = tp<S>(a.A);
    const int Bs = a.Bs;
    for (long n = n0 + gid; n < n1; n += ng) {
        const int e0 = a.in_off[n], e1 = a.in_off[n + 1];
        for (int b = lane; b < Bs; b += a.G) {
            T acc = S::zero();
            if (b < a.Bt) {
                for (int e = e0; e < e1; ++e) {
                    const int s = a.in_src[e];
                    T v;
                    if (s >= 0) v = S::apply(A[static_cast<long>(s) * Bs + b], a.in_coeff ? a.in_coeff[e] : 1.0f);
                    else v = L::from_x(a.X[static_cast<long>(-s - 1) * a.B + b]);
                    acc = S::oplus(acc, v);
                }
            }
            A[n * Bs + b] = acc;
        }
    }
}

template <class S>
__device__ void ph_left_cols(const LeftArgs& a, int cols, long gid, long ng, int lane) {
    using T = typename S::T;
    using L = LeftOps<S>;
    const T* A = tp<S>(a.A);
    T* Y = tp<S>(a.Y);
    for (long c = gid; c < cols; c += ng) {
        const int p0 = a.col_off[c], p1 = a.col_off[c + 1];
        for (int b = lane; b < a.Bt; b += a.G) {
            T acc = S::zero();
            for (int p = p0; p < p1; ++p)
                acc = S::oplus(acc, L::leaf(a.col_val[p], A[static_cast<long>(a.NL + p) * a.Bs + b]));
            Y[c * a.B + b] = acc;
        }
    }
}

// One vector: the column fold as a flat segmented reduction. A wave takes 64
// consecutive terminal positions (one per lane, so every load is coalesced),
// applies the leaf, and reduces per column with the segmented shuffle scan of
// the row gather (the chunk's column-start bitmap marks the segments). A column
// that starts and ends inside the chunk is stored; the pieces of one that
// crosses a chunk boundary combine atomically into y, which k_left_yzero
// cleared (as it did every column without terminals).
template <class S>
__global__ void k_left_fold(LeftArgs a) {
    using T = typename S::T;
    using L = LeftOps<S>;
    const T* A = tp<S>(a.A);
    T* Y = tp<S>(a.Y);
    const long tid = xcd_tid();
    const int lane = static_cast<int>(tid & 63);
    const long nw = (static_cast<long>(gridDim.x) * blockDim.x) >> 6;
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);  // bits 0..lane
    for (long c = tid >> 6; c < a.n_fchunk; c += nw) {
        const long p = 1 + c * 64 + lane;
        const bool valid = p < a.npos;
        T v = S::zero();
        if (valid) v = L::leaf(a.col_val[p], A[(a.NL + p) * a.Bs]);
        const unsigned long long bits = a.fbits[c];
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const T t = __shfl_up(v, d, 64);
            if (lane >= d && ((bits >> (lane - d + 1)) & ((1ull << d) - 1ull)) == 0ull) v = S::oplus(v, t);
        }
        const bool last = valid && (lane == 63 || p + 1 == a.npos || ((bits >> (lane + 1)) & 1ull) != 0ull);
        if (last) {
            const long col = a.fcol[a.ford[c] - 1 + __popcll(bits & upto)];
            const bool starts_here = (bits & upto) != 0ull;
            const bool ends_here = lane < 63 || a.col_off[col + 1] == p + 1;
            if (starts_here && ends_here) Y[col * a.B] = S::fin(v);
            else if (!S::skip(v)) S::emit(Y + col * a.B, v);
        }
    }
}

// One vector: y = the semiring zero on every column (before k_left_fold).
template <class S>
__global__ void k_left_yzero(LeftArgs a, int cols) {
    typename S::T* Y = tp<S>(a.Y);
    const long tid = xcd_tid();
    const long nth = static_cast<long>(gridDim.x) * blockDim.x;
    for (long c = tid; c < cols; c += nth) Y[c * a.B] = S::zero();
}

// Before a sweep without run-length rules (which runs on the row-gather
// kernels): Xs[r*Bs + b] = x as a semiring value for b < Bt, the zero elsewhere
// and on the extra row `rows` (the target of padding entries) -- skipped when
// Xs is null because the gathers can read X itself; then the node rows that
// are combined atomically are zeroed -- every node (one vector) or the listed
// split nodes.
template <class S>
__global__ void k_left_init(const float* X, int B, int Bt, int Bs, int rows, void* Xs_, void* A_, const int* zero_nodes,
                            long n_zero, long zero_all) {
    using T = typename S::T;
    T* Xs = tp<S>(Xs_);
    T* A = tp<S>(A_);
    const long tid = xcd_tid();
    const long nth = static_cast<long>(gridDim.x) * blockDim.x;
    const long nx = Xs ? static_cast<long>(rows + 1) * Bs : 0;  // Xs = nullptr: the gathers read X itself
    for (long i = tid; i < nx; i += nth) {
        const long r = i / Bs;
        const int b = static_cast<int>(i - r * Bs);
        Xs[i] = (r < rows && b < Bt) ? LeftOps<S>::from_x(X[r * B + b]) : S::zero();
    }
    if (zero_all) {
        for (long i = tid; i < zero_all; i += nth) A[i] = S::zero();
    } else {
        for (long i = tid; i < n_zero * Bs; i += nth) A[static_cast<long>(zero_nodes[i / Bs]) * Bs + i % Bs] = S::zero();
    }
}

template <class S>
__global__ void k_left_nodes(LeftArgs a, int n0, int n1) {
    const long tid = xcd_tid();
    const long nth = static_cast<long>(gridDim.x) * blockDim.x;
    ph_left_nodes<S>(a, n0, n1, tid / a.G, nth / a.G, static_cast<int>(tid % a.G));
}

template <class S>
__global__ void k_left_cols(LeftArgs a, int cols) {
    const long tid = xcd_tid();
    const long nth = static_cast<long>(gridDim.x) * blockDim.x;
    ph_left_cols<S>(a, cols, tid / a.G, nth / a.G, static_cast<int>(tid % a.G));
}

// One relaxation step of a fixpoint traversal: x <- x (+) y for every vertex,
// *changed set (plain stores of 1) when some x moved, and delta the frontier:
// each vertex that moved (its new value) and the semiring zero elsewhere (0,
// or +inf for min/plus), so the next product runs over what changed only. With
// P the step's product and an idempotent (+), x_k (+) P(x_(k-1)) = x_k and
// P(x_k) = P(x_(k-1)) (+) P(delta_k), hence x_k (+) P(x_k) = x_k (+) P(delta_k):
// the steps and their count are those of the traversal over all of x.
template <class S>
__global__ void k_relax_frontier(float* x, float* delta, const void* y_, int n, int* changed) {
    using T = typename S::T;
    const T* y = static_cast<const T*>(y_);
    const float zero_x = static_cast<float>(S::zero());
    const long tid = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long nth = static_cast<long>(gridDim.x) * blockDim.x;
    bool moved = false;
    for (long i = tid; i < n; i += nth) {
        const T old = LeftOps<S>::from_x(x[i]);
        const T nv = S::oplus(old, y[i]);
        float d = zero_x;
        if (nv != old) {
            d = static_cast<float>(nv);
            x[i] = d;
            moved = true;
        }
        delta[i] = d;
    }
    if (moved) *changed = 1;
}

static __global__ void k_clear_flag(int* flag) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *flag = 0;
}

}  // namespace dev
}  // namespace gmr
