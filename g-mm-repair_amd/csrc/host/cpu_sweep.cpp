// CPU layered sweep of the host library: y = M x over a built schedule, with
// the values the GPU path uses (FP32 for (+,x), int32 for Boolean and min/plus)
// and the semiring policies of gpu_engine.h. It is what bin/gpu_test checks
// the device result against in the same run, as the reference's driver does
// with its OpenMP sweep (reference gpu-engine/gpu_engine_test.cu:49-104,
// :143-176, :251-274).
//
// Level by level: the rules of a level are split over the threads; every rule
// reads its children from the previous frontier (double-buffered layouts) or
// from the one resident array (Resident layout), or a terminal value. A level's
// emissions (and the terminal emissions) are row-sorted, so the threads split
// them at row boundaries: every run of one row is folded by one thread and
// combined into y once, with no atomics or critical sections, and the result
// does not depend on the thread count except for the order of FP32 additions
// inside a run (none: a run is folded left to right by one thread).
#include <omp.h>

#include <algorithm>
#include <stdexcept>
#include <vector>

#include "build_schedule.h"
#include "gpu_engine.h"
#include "host.hpp"

namespace gmr {

namespace {

// Splits [0, n) into `parts` slices whose boundaries fall where row[] changes,
// then calls f(b, e) for every slice (in parallel).
template <class F>
void by_row_runs(const int* row, long n, F&& f) {
    if (n <= 0) return;
    const int parts = n > 4096 ? omp_get_max_threads() : 1;
    std::vector<long> cut(static_cast<std::size_t>(parts) + 1, n);
    cut[0] = 0;
    for (int p = 1; p < parts; ++p) {
        long c = n * p / parts;
        c = std::max(c, cut[p - 1]);
        while (c > 0 && c < n && row[c] == row[c - 1]) ++c;  // start at a row boundary
        cut[p] = c;
    }
#pragma omp parallel for schedule(static) if (parts > 1)
    for (int p = 0; p < parts; ++p) f(cut[p], cut[p + 1]);
}

template <class S>
void sweep(const BuiltSchedule& bs, const int* val_idx, const int* col_idx, const float* V, const float* x,
           typename S::T* y) {
    using T = typename S::T;
    const bool resident = bs.layout == ScheduleLayout::Resident;
    const int U = bs.alpha, L = bs.max_depth;
    std::vector<T> Tv(static_cast<std::size_t>(std::max(U, 1)), S::zero());
#pragma omp parallel for if (U > 4096)
    for (int q = 1; q < U; ++q) Tv[q] = S::leaf(V[val_idx[q - 1]], x[col_idx[q - 1]]);
#pragma omp parallel for if (bs.rows > 4096)
    for (int r = 0; r < bs.rows; ++r) y[r] = S::zero();

    // Folds runs of equal rows of (src value, row) pairs into y.
    auto emit = [&](const int* pos, const int* row, long n, const T* src) {
        by_row_runs(row, n, [&](long b, long e) {
            for (long i = b; i < e;) {
                const int r = row[i];
                T acc = src[pos[i]];
                for (++i; i < e && row[i] == r; ++i) acc = S::oplus(acc, src[pos[i]]);
                y[r] = S::oplus(y[r], acc);
            }
        });
    };
    emit(bs.term_emit_sym.data(), bs.term_emit_row.data(), static_cast<long>(bs.term_emit_sym.size()), Tv.data());

    const std::size_t wn = resident ? bs.num_layered : static_cast<std::size_t>(std::max(bs.max_width, 0));
    std::vector<T> W0(std::max<std::size_t>(wn, 1), S::zero()), W1(resident ? 1 : std::max<std::size_t>(wn, 1));
    for (int k = 1; k <= L; ++k) {
        const long lo = bs.level_offsets[k - 1], hi = bs.level_offsets[k];
        const T* src = resident ? W0.data() : ((k & 1) ? W0.data() : W1.data());        // frontier k-1
        T* dst = resident ? W0.data() + lo : ((k & 1) ? W1.data() : W0.data());        // frontier k
#pragma omp parallel for if (hi - lo > 4096)
        for (long i = lo; i < hi; ++i) {
            const LayeredRule& r = bs.flattened_rules[static_cast<std::size_t>(i)];
            const T l = r.left >= 0 ? src[r.left] : Tv[-r.left - 1];
            const T rv = r.right >= 0 ? src[r.right] : Tv[-r.right - 1];
            dst[i - lo] = S::oplus(S::apply(l, r.coeff), rv);
        }
        const long e0 = bs.emit_offsets[k - 1], e1 = bs.emit_offsets[k];
        const T* esrc = resident ? W0.data() : dst;  // resident positions are global slots
        emit(bs.emit_pos.data() + e0, bs.emit_row.data() + e0, e1 - e0, esrc);
    }
}

}  // namespace

void cpu_layered_sweep(int semiring, const BuiltSchedule& bs, const int* val_idx, const int* col_idx, const float* V,
                       const float* x, void* y, int threads) {
    const int saved = omp_get_max_threads();
    if (threads > 0) omp_set_num_threads(threads);
    try {
        switch (semiring) {
            case 0: sweep<PlusTimes>(bs, val_idx, col_idx, V, x, static_cast<float*>(y)); break;
            case 1: sweep<Boolean>(bs, val_idx, col_idx, V, x, static_cast<int*>(y)); break;
            case 2: sweep<Tropical>(bs, val_idx, col_idx, V, x, static_cast<int*>(y)); break;
            default: throw std::invalid_argument("unknown semiring");
        }
    } catch (...) {
        omp_set_num_threads(saved);
        throw;
    }
    omp_set_num_threads(saved);
}

}  // namespace gmr
