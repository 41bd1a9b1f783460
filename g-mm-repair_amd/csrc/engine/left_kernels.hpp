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
// Incoming lists are CSR over nodes (rule slots, then terminals); an entry
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
    const int* col_off;    // cols + 1
    const int* col_term;   // compact terminals grouped by column
    const float* tval;     // value of compact terminal t
    const float* X;        // rows x B (this tile's column base folded in)
    void* Y;               // cols x B
    void* A;               // nodes x Bs node values
    int NL;                // rule slots; terminal t is node NL + t
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
        const int t0 = a.col_off[c], t1 = a.col_off[c + 1];
        for (int b = lane; b < a.Bt; b += a.G) {
            T acc = S::zero();
            for (int i = t0; i < t1; ++i) {
                const int t = a.col_term[i];
                acc = S::oplus(acc, L::leaf(a.tval[t], A[static_cast<long>(a.NL + t) * a.Bs + b]));
            }
            Y[c * a.B + b] = acc;
        }
    }
}

// Before a sweep without run-length rules (which runs on the row-gather
// kernels): Xs[r*Bs + b] = x as a semiring value for b < Bt, the zero elsewhere
// and on the extra row `rows` (the target of padding entries); then the node
// rows that are combined atomically are zeroed -- every node (one vector) or
// the listed split nodes.
template <class S>
__global__ void k_left_init(const float* X, int B, int Bt, int Bs, int rows, void* Xs_, void* A_, const int* zero_nodes,
                            long n_zero, long zero_all) {
    using T = typename S::T;
    T* Xs = tp<S>(Xs_);
    T* A = tp<S>(A_);
    const long tid = xcd_tid();
    const long nth = static_cast<long>(gridDim.x) * blockDim.x;
    const long nx = static_cast<long>(rows + 1) * Bs;
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

// One relaxation step of a fixpoint traversal: x <- x (+) y for every vertex;
// *changed is set (plain stores of 1) when some x moved.
template <class S>
__global__ void k_relax(float* x, const void* y_, int n, int* changed) {
    using T = typename S::T;
    const T* y = static_cast<const T*>(y_);
    const long tid = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long nth = static_cast<long>(gridDim.x) * blockDim.x;
    bool moved = false;
    for (long i = tid; i < n; i += nth) {
        const T old = LeftOps<S>::from_x(x[i]);
        const T nv = S::oplus(old, y[i]);
        if (nv != old) {
            x[i] = static_cast<float>(nv);
            moved = true;
        }
    }
    if (moved) *changed = 1;
}

static __global__ void k_clear_flag(int* flag) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *flag = 0;
}

}  // namespace dev
}  // namespace gmr
