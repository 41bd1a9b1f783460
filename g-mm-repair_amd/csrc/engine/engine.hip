// Device engine: upload once, sweep many times (the gmr.h C ABI).
//
// The grammar structure lives in HBM for the life of the engine:
//   kids      int2 (left, right) per layered rule, level by level (8 B/rule;
//             the reference's 12-byte AoS LayeredRule minus the coefficient)
//   coeff     float per rule, only when some rule is run-length (coeff != 1)
//   tcol/tval column and value of every compact terminal
//   emits     int2 (pos, row) per top-sequence non-terminal, level by level,
//             row-sorted inside a level
//   term_emits int2 (sym, row) per top-sequence terminal
//   crow/csym row-ordered top sequence (Resident layout, gather emission)
// Values (per run, sized by the vector tile Bt): T = U x Bt terminal values and
// the frontier: two w* x Bt buffers for the double-buffered layouts, or one
// num_layered x Bt resident array.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <initializer_list>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "build_schedule.h"
#include "engine_internal.hpp"
#include "gmr.h"
#include "left_kernels.hpp"
#include "sweep_kernels.hpp"

const BuiltSchedule* gmr_schedule_internal(const gmr_schedule* s);
void gmr_host_set_error(const char* msg);

namespace gmr {

namespace {

thread_local std::string g_dev_err;

struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define GMR_HIP(call)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw ::gmr::DeviceError(std::string(#call) + ": " + hipGetErrorString(e_));   \
    } while (0)

template <class T>
T* dev_upload(const T* h, std::size_t n, std::size_t* acc) {
    if (n == 0) return nullptr;
    T* d = nullptr;
    GMR_HIP(hipMalloc(&d, n * sizeof(T)));
    GMR_HIP(hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice));
    if (acc) *acc += n * sizeof(T);
    return d;
}

}  // namespace

Engine::~Engine() {
    for (auto* m : {&graphs, &left_graphs})
        for (auto& kv : *m) {
            if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
            if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
        }
    for (void* p : {static_cast<void*>(d_in_off), static_cast<void*>(d_in_src), static_cast<void*>(d_in_coeff),
                    static_cast<void*>(d_col_off), static_cast<void*>(d_col_val), static_cast<void*>(d_lf_bits),
                    static_cast<void*>(d_lf_ord), static_cast<void*>(d_lf_col), ws_A, ws_Xs,
                    static_cast<void*>(d_lseg), static_cast<void*>(d_lsplit), static_cast<void*>(d_lpsrc),
                    static_cast<void*>(d_lpoff), static_cast<void*>(d_lbits), static_cast<void*>(d_lord),
                    static_cast<void*>(d_lrow), ws_trav, static_cast<void*>(d_lxsym), static_cast<void*>(d_lxbits),
                    static_cast<void*>(d_lxord), static_cast<void*>(d_lxrow), static_cast<void*>(d_lxpoff)})
        if (p) (void)hipFree(p);
    auto free_lists = [](const XParts& x) {
        for (void* p : {static_cast<void*>(x.d_sym), static_cast<void*>(x.d_bits), static_cast<void*>(x.d_ord),
                        static_cast<void*>(x.d_row), static_cast<void*>(x.d_poff), static_cast<void*>(x.d_cbase)})
            if (p) (void)hipFree(p);
    };
    for (auto* m : {&xparts, &bparts, &lparts})
        for (auto& kv : *m) free_lists(kv.second);
    void* ptrs[] = {d_kids, d_coeff, d_tcol, d_tval, d_tref, d_emits, d_term, d_crow, d_csym, d_gseg, d_zrows, d_fbits, d_ford,
                    d_frow, d_level_off, d_emit_off, ws_T, ws_W0, ws_W1, d_kinds,
                    static_cast<void*>(ws_bits), d_row_new, ws_yrow, ws_tail};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    barrier.release();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (capture_stream) (void)hipStreamDestroy(capture_stream);
}

void BarrierState::init(int khz) {
    wall_khz = khz > 0 ? khz : 100000;
    GMR_HIP(hipMalloc(&d, sizeof(dev::GridBarrier)));
    GMR_HIP(hipMemset(d, 0, sizeof(dev::GridBarrier)));
    GMR_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_fail), sizeof(int), hipHostMallocMapped | hipHostMallocCoherent));
    *h_fail = 0;
    GMR_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&d_fail), h_fail, 0));
    GMR_HIP(hipDeviceSynchronize());
    base = 0;
}

void BarrierState::release() noexcept {
    if (d) (void)hipFree(d);
    if (h_fail) (void)hipHostFree(h_fail);
    d = nullptr;
    h_fail = d_fail = nullptr;
}

dev::BarrierRef BarrierState::ref() const {
    dev::BarrierRef r;
    r.gb = d;
    r.fail_host = d_fail;
    r.base = base;
    const unsigned long long us = timeout_us > 0 ? static_cast<unsigned long long>(timeout_us) : 2000000ULL;
    r.deadline = static_cast<unsigned long long>(wall_khz) * us / 1000ULL;  // ticks
    return r;
}

bool BarrierState::take_failure() {
    if (!h_fail) return false;
    if (__atomic_load_n(h_fail, __ATOMIC_ACQUIRE) == 0) return false;
    __atomic_store_n(h_fail, 0, __ATOMIC_RELEASE);
    return true;
}

int BarrierState::check() {
    dev::GridBarrier h{};
    GMR_HIP(hipDeviceSynchronize());
    GMR_HIP(hipMemcpy(&h, d, sizeof(h), hipMemcpyDeviceToHost));
    const bool failed = take_failure();
    if (failed || h.arrived != base) {
        // Restart the barrier from a clean state.
        GMR_HIP(hipMemset(d, 0, sizeof(dev::GridBarrier)));
        GMR_HIP(hipDeviceSynchronize());
        base = 0;
        return 1;
    }
    return 0;
}

void Engine::upload(int rows_, int cols_, int U_, int num_rules, const LayeredRule* rules,
                    const int* val_idx, const int* col_idx, const float* V, int nvals_,
                    const GPUSchedule& s, int layout_) {
    if (rows_ < 0 || cols_ < 1 || U_ < 1 || num_rules < 0 || nvals_ < 0 || s.max_depth < 0)
        throw std::invalid_argument("bad engine dimensions");
    if (layout_ < 0 || layout_ > 2) throw std::invalid_argument("unknown layout");
    rows = rows_;
    cols = cols_;
    U = U_;
    NL = num_rules;
    L = s.max_depth;
    nvals = nvals_;
    layout = layout_;
    resident = layout == 2;
    level_off.assign(L + 1, 0);
    emit_off.assign(L + 1, 0);
    if (L > 0 && (!s.level_offsets || !s.emit_offsets))
        throw std::invalid_argument("schedule offsets are NULL");
    for (int k = 0; k <= L; ++k) {
        level_off[k] = s.level_offsets ? s.level_offsets[k] : 0;
        emit_off[k] = s.emit_offsets ? s.emit_offsets[k] : 0;
    }
    if (level_off[0] != 0 || emit_off[0] != 0) throw std::invalid_argument("offsets must start at 0");
    max_width = 0;
    for (int k = 1; k <= L; ++k) {
        if (level_off[k] < level_off[k - 1] || emit_off[k] < emit_off[k - 1])
            throw std::invalid_argument("schedule offsets decrease");
        max_width = std::max(max_width, level_off[k] - level_off[k - 1]);
    }
    if (level_off[L] != NL) throw std::invalid_argument("level offsets do not cover the rules");
    E = emit_off[L];
    if (E != s.emit_total) throw std::invalid_argument("emit offsets do not cover the emissions");
    Et = s.num_term_emit;
    if (Et < 0 || (Et > 0 && (!s.term_emit_sym || !s.term_emit_row)) || (E > 0 && (!s.emit_pos || !s.emit_row)))
        throw std::invalid_argument("schedule emission arrays are NULL");
    if (NL > 0 && !rules) throw std::invalid_argument("rules are NULL");
    if (U > 1 && (!val_idx || !col_idx)) throw std::invalid_argument("terminal arrays are NULL");
    if (nvals > 0 && !V) throw std::invalid_argument("V is NULL");

    // Range checks: a bad index in any of these would be a wild device access.
    std::vector<int2> kids(NL);
    std::vector<float> coeff;
    has_runlen = false;
    for (int k = 1; k <= L; ++k) {
        const int lo = level_off[k - 1], hi = level_off[k];
        const int prev_w = k > 1 ? level_off[k - 1] - level_off[k - 2] : 0;
        const int lim = resident ? lo : prev_w;
        for (int i = lo; i < hi; ++i) {
            const LayeredRule& r = rules[i];
            for (int c : {r.left, r.right}) {
                if (c >= 0 ? c >= lim : -static_cast<long long>(c) - 1 >= U)
                    throw std::invalid_argument("rule " + std::to_string(i) + " child out of range");
            }
            kids[i] = make_int2(r.left, r.right);
            if (r.coeff != 1.0f) has_runlen = true;
        }
    }
    if (has_runlen) {
        coeff.resize(NL);
        for (int i = 0; i < NL; ++i) coeff[i] = rules[i].coeff;
    }
    std::vector<int2> em(E), te(Et);
    for (int k = 1; k <= L; ++k) {
        const int w = level_off[k] - level_off[k - 1];
        for (int e = emit_off[k - 1]; e < emit_off[k]; ++e) {
            int p = s.emit_pos[e], r = s.emit_row[e];
            bool ok = resident ? (p >= level_off[k - 1] && p < level_off[k]) : (p >= 0 && p < w);
            if (!ok || r < 0 || r >= rows) throw std::invalid_argument("emission out of range");
            em[e] = make_int2(p, r);
        }
    }
    for (int e = 0; e < Et; ++e) {
        int sy = s.term_emit_sym[e], r = s.term_emit_row[e];
        if (sy < 0 || sy >= U || r < 0 || r >= rows) throw std::invalid_argument("terminal emission out of range");
        te[e] = make_int2(sy, r);
    }
    // Emissions must be row-sorted inside a level for the run pre-reduction;
    // the order only changes which atomics combine first, so sort when needed.
    auto row_sort = [](int2* b, int2* e) {
        if (!std::is_sorted(b, e, [](const int2& x, const int2& y) { return x.y < y.y; }))
            std::stable_sort(b, e, [](const int2& x, const int2& y) { return x.y < y.y; });
    };
    for (int k = 1; k <= L; ++k) row_sort(em.data() + emit_off[k - 1], em.data() + emit_off[k]);
    row_sort(te.data(), te.data() + Et);

    std::vector<int> tcol(U, 0);
    std::vector<float> tval(U, 0.0f);
    for (int q = 1; q < U; ++q) {
        int v = val_idx[q - 1], c = col_idx[q - 1];
        if (v < 0 || v >= nvals || c < 0 || c >= cols) throw std::invalid_argument("terminal decode out of range");
        tcol[q] = c;
        tval[q] = V[v];
    }
    // Renumber the compact terminals in column order (terminal 0 stays first;
    // a column's terminals keep their order): consecutive terminals then read
    // neighbouring x entries, so the LDS-range fills and the per-terminal leaf
    // computations read x in runs instead of scattered. Only the engine's own
    // arrays hold compact terminal numbers, and all of them are renumbered here.
    {
        std::vector<int> coff(static_cast<std::size_t>(cols) + 1, 0), pos(static_cast<std::size_t>(U), 0);
        for (int q = 1; q < U; ++q) ++coff[static_cast<std::size_t>(tcol[q]) + 1];
        coff[0] = 1;
        for (int c = 0; c < cols; ++c) coff[static_cast<std::size_t>(c) + 1] += coff[static_cast<std::size_t>(c)];
        bool moved = false;
        for (int q = 1; q < U; ++q) {
            pos[q] = coff[static_cast<std::size_t>(tcol[q])]++;
            moved |= pos[q] != q;
        }
        if (moved) {
            std::vector<int> tc2(static_cast<std::size_t>(U), 0);
            std::vector<float> tv2(static_cast<std::size_t>(U), 0.0f);
            for (int q = 1; q < U; ++q) {
                tc2[static_cast<std::size_t>(pos[q])] = tcol[q];
                tv2[static_cast<std::size_t>(pos[q])] = tval[q];
            }
            tcol.swap(tc2);
            tval.swap(tv2);
            auto rn = [&](int c) { return c < 0 ? -pos[static_cast<std::size_t>(-c - 1)] - 1 : c; };
            for (int2& k : kids) k = make_int2(rn(k.x), rn(k.y));
            for (int2& e : te) e.x = pos[static_cast<std::size_t>(e.x)];
        }
    }

    // Resident layout: the row-ordered top sequence for gather emission.
    std::vector<int> crow, csym;
    std::vector<int4> gseg;
    std::vector<int> zrows;
    gather_split = 0;
    if (resident) {
        crow.assign(static_cast<std::size_t>(rows) + 1, 0);
        for (const int2& p : em) ++crow[p.y + 1];
        for (const int2& p : te) ++crow[p.y + 1];
        for (int r = 0; r < rows; ++r) crow[r + 1] += crow[r];
        csym.resize(static_cast<std::size_t>(E) + Et);
        std::vector<int> fill(crow.begin(), crow.end() - 1);
        // Terminals first, then levels in order (C order is not needed).
        for (const int2& p : te) csym[fill[p.y]++] = -p.x - 1;
        for (const int2& p : em) csym[fill[p.y]++] = p.x;
        // Gather work items: rows cut into segments of at most kGatherSeg
        // symbols, so a few long rows do not serialise the gather. A segment
        // that covers its whole row stores its sum; the pieces of a split row
        // combine atomically into a zeroed y row.
        for (int r = 0; r < rows; ++r) {
            const int a0 = crow[r], a1 = crow[r + 1];
            const int split = (a1 - a0) > kGatherSeg ? 1 : 0;
            gather_split |= split;
            if (split) zrows.push_back(r);
            if (a1 == a0) gseg.push_back(make_int4(r, a0, a1, 0));
            for (int q = a0; q < a1; q += kGatherSeg) gseg.push_back(make_int4(r, q, std::min(a1, q + kGatherSeg), split));
        }
    }
    // Single-vector flat gather: which symbols start a row, per 64-symbol chunk.
    std::vector<unsigned long long> fbits;
    std::vector<int> ford, frow;
    if (resident) {
        const long nsym = static_cast<long>(E) + Et;
        fbits.assign(static_cast<std::size_t>((nsym + 63) / 64), 0ull);
        for (int r = 0; r < rows; ++r) {
            if (crow[r + 1] == crow[r]) continue;
            const int p = crow[r];
            fbits[p >> 6] |= 1ull << (p & 63);
            frow.push_back(r);
        }
        ford.resize(fbits.size());
        int before = 0;
        for (std::size_t c = 0; c < fbits.size(); ++c) {
            ford[c] = before;
            before += __builtin_popcountll(fbits[c]);
        }
    }
    // Single-vector gather: lanes per row ~ a quarter of the mean row length.
    {
        const double mean = !gseg.empty() ? static_cast<double>(E + Et) / gseg.size() : 0.0;
        gather_lanes = 1;
        while (gather_lanes < 64 && gather_lanes * 4 < mean) gather_lanes <<= 1;
    }
    // Terminals that rules have as children: the only terminal values a sweep
    // needs outside the row gather (the LDS-range gather computes the others
    // as it fills its ranges).
    std::vector<int> tref;
    {
        std::vector<char> used(static_cast<std::size_t>(std::max(U, 1)), 0);
        for (const int2& k : kids) {
            if (k.x < 0) used[static_cast<std::size_t>(-k.x - 1)] = 1;
            if (k.y < 0) used[static_cast<std::size_t>(-k.y - 1)] = 1;
        }
        for (int q = 1; q < U; ++q)
            if (used[q]) tref.push_back(q);
    }
    n_tref = static_cast<int>(tref.size());
    struct_bytes = 0;
    d_tref = dev_upload(tref.data(), tref.size(), &struct_bytes);
    d_kids = dev_upload(kids.data(), kids.size(), &struct_bytes);
    d_coeff = has_runlen ? dev_upload(coeff.data(), coeff.size(), &struct_bytes) : nullptr;
    d_tcol = dev_upload(tcol.data(), tcol.size(), &struct_bytes);
    d_tval = dev_upload(tval.data(), tval.size(), &struct_bytes);
    d_emits = dev_upload(em.data(), em.size(), &struct_bytes);
    d_term = dev_upload(te.data(), te.size(), &struct_bytes);
    if (resident) {
        d_crow = dev_upload(crow.data(), crow.size(), &struct_bytes);
        d_csym = dev_upload(csym.data(), csym.size(), &struct_bytes);
        d_gseg = dev_upload(gseg.data(), gseg.size(), &struct_bytes);
        d_zrows = dev_upload(zrows.data(), zrows.size(), &struct_bytes);
        n_zrows = static_cast<int>(zrows.size());
        n_gseg = static_cast<int>(gseg.size());
        crow_host.assign(1, static_cast<int>(csym.size()));
        d_fbits = dev_upload(fbits.data(), fbits.size(), &struct_bytes);
        d_ford = dev_upload(ford.data(), ford.size(), &struct_bytes);
        d_frow = dev_upload(frow.data(), frow.size(), &struct_bytes);
        n_fchunk = static_cast<int>(fbits.size());
    }
    d_level_off = dev_upload(level_off.data(), level_off.size(), &struct_bytes);
    d_emit_off = dev_upload(emit_off.data(), emit_off.size(), &struct_bytes);
    GMR_HIP(hipEventCreate(&ev0));
    GMR_HIP(hipEventCreate(&ev1));
    int dev = 0;
    GMR_HIP(hipGetDevice(&dev));
    device = dev;
    GMR_HIP(hipDeviceGetAttribute(&num_cus, hipDeviceAttributeMultiprocessorCount, dev));
    GMR_HIP(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, dev));
    if (wall_khz <= 0) wall_khz = 100000;
    GMR_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    barrier.init(wall_khz);
}

// Passes that keep each XCD's share of one value range near 3 MB (a 4 MB L2
// per XCD), at most 8.
int Engine::auto_xcd_passes(int Bs, int Bt) const {
    const double bytes = (static_cast<double>(U) + NL) * 4.0 * Bs;
    const double per = 8.0 * 3.0 * 1024 * 1024;
    if (Bs == 1) return std::max(1, std::min(8, static_cast<int>(std::ceil(bytes / per))));
    // Several vectors: every row piece costs Bt atomics, so fewer passes than
    // the cache size alone asks for (measured on the 500,000 x 10,000 genotype
    // grammar: 4 passes best at Bt = 2, 2 at Bt = 4).
    return std::max(1, std::min(8, static_cast<int>(std::ceil(bytes / (per * Bt)))));
}

namespace {

// Host lists of the XCD-partitioned single-vector gather, appended group by
// group: the entries of rows [r0, r1) of a CSR (off, sym) are cut by value key
// into P ranges of equal use counts; each range keeps its entries in row order,
// padded to whole 64-entry chunks with `pad` (an entry that reads the semiring
// zero). poff gets the P + 1 chunk offsets of this group.
struct PartLists {
    std::vector<int> sym, ord, row, poff;
    std::vector<unsigned long long> bits;
};

// With max_rows > 0, a row piece that would start max_rows or more rows after
// the row of its chunk's first entry starts a new chunk instead (the rest of
// the chunk is padding, which belongs to the piece before it), so every chunk
// spans fewer than max_rows rows.
template <class Off, class Key>
void build_part_lists(long r0, long r1, const Off& off, const std::vector<int>& sym, Key key, long nkey, int P, int pad,
                      PartLists& out, long span = 0, long max_rows = 0) {
    const long e0 = off[r0], e1 = off[r1], nent = e1 - e0;
    std::vector<long> cut(P + 1);
    cut[0] = 0;
    cut[P] = nkey;
    if (span > 0) {
        // fixed key ranges [p * span, (p + 1) * span)
        for (int p = 1; p < P; ++p) cut[p] = std::min(nkey, static_cast<long>(p) * span);
    } else {
        std::vector<long> uses(static_cast<std::size_t>(nkey) + 1, 0);
        for (long q = e0; q < e1; ++q) ++uses[key(sym[q]) + 1];
        for (long k = 0; k < nkey; ++k) uses[k + 1] += uses[k];
        for (int p = 1; p < P; ++p)
            cut[p] = std::max(cut[p - 1],
                              static_cast<long>(std::lower_bound(uses.begin(), uses.end(), nent * p / P) - uses.begin()));
    }
    auto part = [&](int c) -> int {
        return static_cast<int>(std::upper_bound(cut.begin() + 1, cut.begin() + P, key(c)) - (cut.begin() + 1));
    };
    std::vector<std::vector<int>> lsym(P), lrow(P);
    std::vector<std::vector<long>> lstart(P);  // positions (in the list) that start a row
    std::vector<long> seen(P, -1), row0(P, 0);  // row0: row of the current chunk's first entry
    for (long r = r0; r < r1; ++r) {
        for (long q = off[r]; q < off[r + 1]; ++q) {
            const int p = part(sym[q]);
            if (seen[p] != r) {
                seen[p] = r;
                if (max_rows > 0 && (lsym[p].size() & 63) != 0 && r - row0[p] >= max_rows)
                    lsym[p].resize((lsym[p].size() + 63) & ~static_cast<std::size_t>(63), pad);
                lstart[p].push_back(static_cast<long>(lsym[p].size()));
                lrow[p].push_back(static_cast<int>(r));
            }
            if ((lsym[p].size() & 63) == 0) row0[p] = r;
            lsym[p].push_back(sym[q]);
        }
    }
    for (int p = 0; p < P; ++p) {
        out.poff.push_back(static_cast<int>(out.bits.size()));
        const long n = static_cast<long>(lsym[p].size());
        const long nc = (n + 63) / 64;
        std::vector<unsigned long long> bits(static_cast<std::size_t>(nc), 0ull);
        for (long st : lstart[p]) bits[st >> 6] |= 1ull << (st & 63);
        long before = static_cast<long>(out.row.size());
        for (long c = 0; c < nc; ++c) {
            out.bits.push_back(bits[c]);
            out.ord.push_back(static_cast<int>(before));
            before += __builtin_popcountll(bits[c]);
        }
        for (long q = 0; q < nc * 64; ++q) out.sym.push_back(q < n ? lsym[p][q] : pad);
        out.row.insert(out.row.end(), lrow[p].begin(), lrow[p].end());
        std::vector<int>().swap(lsym[p]);
    }
    out.poff.push_back(static_cast<int>(out.bits.size()));
    if (static_cast<long>(out.sym.size()) > 0x7fffffffL || static_cast<long>(out.row.size()) > 0x7fffffffL)
        throw std::invalid_argument("partitioned gather: list too long for int32 offsets");
}

}  // namespace

// The XCD-partitioned single-vector gather's lists for `passes` passes: value
// keys (terminal t -> t, rule slot s -> U + s) cut into 8 x passes ranges;
// padding reads the zero terminal. Built from the uploaded top sequence on
// first use.
const Engine::XParts& Engine::xparts_for(int passes) {
    auto it = xparts.find(passes);
    if (it != xparts.end()) return it->second;
    forbid_alloc("the XCD-partitioned gather's lists for this pass count");
    if (!resident) throw std::invalid_argument("the partitioned gather needs the Resident layout");
    const long nsym = crow_host[0];
    std::vector<int> crow(static_cast<std::size_t>(rows) + 1), csym(static_cast<std::size_t>(nsym));
    GMR_HIP(hipDeviceSynchronize());
    GMR_HIP(hipMemcpy(crow.data(), d_crow, sizeof(int) * crow.size(), hipMemcpyDeviceToHost));
    if (nsym) GMR_HIP(hipMemcpy(csym.data(), d_csym, sizeof(int) * nsym, hipMemcpyDeviceToHost));
    auto key = [&](int c) -> long { return c < 0 ? static_cast<long>(-c - 1) : static_cast<long>(U) + c; };
    PartLists L;
    build_part_lists(0, rows, crow, csym, key, static_cast<long>(U) + NL, 8 * passes, -1, L);
    XParts x;
    x.poff = L.poff;
    x.d_sym = dev_upload(L.sym.data(), L.sym.size(), &struct_bytes);
    x.d_bits = dev_upload(L.bits.data(), L.bits.size(), &struct_bytes);
    x.d_ord = dev_upload(L.ord.data(), L.ord.size(), &struct_bytes);
    x.d_row = dev_upload(L.row.data(), L.row.size(), &struct_bytes);
    x.d_poff = dev_upload(x.poff.data(), x.poff.size(), &struct_bytes);
    return xparts.emplace(passes, std::move(x)).first->second;
}

namespace {

// Packs every entry of range lists (already indices inside their range, -1
// for padding) as (row - first row of its chunk) << 16 | index (0xFFFF for
// padding), with the chunks' first rows in cbase. Fails (false, L unchanged)
// when a range is 65,535 keys or more or a chunk spans 65,536 rows or more
// (lists built with max_rows = 65,536 never do).
bool pack_entries(PartLists& L, long span, std::vector<int>& cbase) {
    if (span >= 0xFFFF) return false;
    const long nch = static_cast<long>(L.bits.size());
    std::vector<int> packed(L.sym.size());
    cbase.assign(static_cast<std::size_t>(nch), 0);
    for (long c = 0; c < nch; ++c) {
        int base = 0;
        for (int lane = 0; lane < 64; ++lane) {
            const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
            // the row of this entry: the last row started at or before it (padding
            // belongs to the list's last piece)
            const int row = L.row[static_cast<std::size_t>(L.ord[c] - 1 + __builtin_popcountll(L.bits[c] & upto))];
            if (lane == 0) base = cbase[c] = row;
            const long d = static_cast<long>(row) - base;
            if (d < 0 || d > 0xFFFF) return false;
            const int idx = L.sym[static_cast<std::size_t>(c * 64 + lane)];
            const unsigned e = (static_cast<unsigned>(d) << 16) | (idx < 0 ? 0xFFFFu : static_cast<unsigned>(idx));
            packed[static_cast<std::size_t>(c * 64 + lane)] = static_cast<int>(e);
        }
    }
    L.sym.swap(packed);
    return true;
}

// A row order that puts similar rows next to each other: rows by their
// top-sequence length, then by the least hash (splitmix64 finaliser) of their
// symbols, then by row. Returns new number -> matrix row.
std::vector<int> similar_row_order(const std::vector<int>& crow, const std::vector<int>& csym, int rows) {
    std::vector<unsigned long long> mh(static_cast<std::size_t>(rows), ~0ull);
    for (int r = 0; r < rows; ++r)
        for (int q = crow[r]; q < crow[r + 1]; ++q) {
            unsigned long long z = static_cast<unsigned long long>(static_cast<unsigned>(csym[q])) + 0x9E3779B97F4A7C15ull;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            if (z < mh[r]) mh[r] = z;
        }
    std::vector<int> order(static_cast<std::size_t>(rows));
    for (int r = 0; r < rows; ++r) order[r] = r;
    std::sort(order.begin(), order.end(), [&](int a, int b) {
        const int la = crow[a + 1] - crow[a], lb = crow[b + 1] - crow[b];
        if (la != lb) return la < lb;
        if (mh[a] != mh[b]) return mh[a] < mh[b];
        return a < b;
    });
    return order;
}

// Lists by fixed key ranges of `span` keys (one list per range, each entry
// stored as its key's index inside the range; padding -1), uploaded. With
// `pack`, entries are packed with their row (pack_entries; ranges must be
// below 65,535 keys). With `row_order` (new number -> matrix row) the lists
// number rows by it.
Engine::XParts range_lists(Engine& en, long nkey, long span, const std::function<long(int)>& key,
                           bool pack = false, const std::vector<int>* row_order = nullptr) {
    const long nsym = en.crow_host[0];
    std::vector<int> crow(static_cast<std::size_t>(en.rows) + 1), csym(static_cast<std::size_t>(nsym));
    GMR_HIP(hipDeviceSynchronize());
    GMR_HIP(hipMemcpy(crow.data(), en.d_crow, sizeof(int) * crow.size(), hipMemcpyDeviceToHost));
    if (nsym) GMR_HIP(hipMemcpy(csym.data(), en.d_csym, sizeof(int) * nsym, hipMemcpyDeviceToHost));
    if (row_order) {
        // the same rows, renumbered: new row j holds matrix row (*row_order)[j]
        std::vector<int> crow2(crow.size()), csym2(csym.size());
        crow2[0] = 0;
        for (int j = 0; j < en.rows; ++j) {
            const int r = (*row_order)[static_cast<std::size_t>(j)];
            const int n = crow[r + 1] - crow[r];
            std::copy(csym.begin() + crow[r], csym.begin() + crow[r] + n, csym2.begin() + crow2[j]);
            crow2[j + 1] = crow2[j] + n;
        }
        crow.swap(crow2);
        csym.swap(csym2);
    }
    const int P = static_cast<int>(std::max(1L, (nkey + span - 1) / span));
    PartLists L;
    // packed entries hold a 16-bit row offset: chunks then span fewer than 65,536 rows
    build_part_lists(0, en.rows, crow, csym, key, nkey, P, -1, L, span, pack ? 65536 : 0);
    // Entries become indices inside their list's range; padding (-1) becomes an
    // index past every range (it reads the semiring zero).
    for (int p = 0; p < P; ++p)
        for (long q = static_cast<long>(L.poff[p]) * 64; q < static_cast<long>(L.poff[p + 1]) * 64; ++q) {
            const int c = L.sym[q];
            L.sym[q] = c == -1 ? -1 : static_cast<int>(key(c) - static_cast<long>(p) * span);
        }
    Engine::XParts x;
    std::vector<int> cbase;
    if (pack && !pack_entries(L, span, cbase))
        throw std::invalid_argument("range lists: a range of 65,535 keys or more cannot be packed");
    x.packed = pack;
    x.poff = L.poff;
    x.d_sym = dev_upload(L.sym.data(), L.sym.size(), &en.struct_bytes);
    x.d_bits = dev_upload(L.bits.data(), L.bits.size(), &en.struct_bytes);
    if (x.packed) {
        x.d_cbase = dev_upload(cbase.data(), cbase.size(), &en.struct_bytes);
    } else {
        x.d_ord = dev_upload(L.ord.data(), L.ord.size(), &en.struct_bytes);
        x.d_row = dev_upload(L.row.data(), L.row.size(), &en.struct_bytes);
    }
    x.d_poff = dev_upload(x.poff.data(), x.poff.size(), &en.struct_bytes);
    return x;
}

}  // namespace

// Lists of the Boolean bit-mode gather: the top-sequence symbols regrouped by
// bit index (terminal q -> q, rule slot s -> 64 * ceil(U / 64) + s) into ranges
// of `range_bits` bits, one range per list, so a workgroup can hold its list's
// bits in LDS. The zero terminal's bit is 0, so it reads 0 like the padding.
const Engine::XParts& Engine::bit_lists(int range_bits) {
    auto it = bparts.find(range_bits);
    if (it != bparts.end()) return it->second;
    forbid_alloc("the bit-mode lists for this range size");
    if (!resident) throw std::invalid_argument("the bit-mode gather needs the Resident layout");
    ensure_bits();
    const long tbits = static_cast<long>(bits_tw) * 64;
    auto key = [tbits](int c) -> long { return c < 0 ? static_cast<long>(-c - 1) : tbits + c; };
    return bparts.emplace(range_bits, range_lists(*this, tbits + NL, range_bits, key)).first->second;
}

// Lists of the LDS-range gather of one vector: value keys (terminal t -> t,
// rule slot s -> U + s) in ranges of `span` values. With `row_order` the lists
// number rows by similar_row_order, and the first such call also uploads the
// matrix row -> new number map and allocates the row-ordered copy of y.
const Engine::XParts& Engine::lds_lists(int span, bool row_order) {
    const int mk = 2 * span + (row_order ? 1 : 0);
    auto it = lparts.find(mk);
    if (it != lparts.end()) return it->second;
    forbid_alloc("the LDS-range gather's lists for this range size");
    if (!resident) throw std::invalid_argument("the LDS-range gather needs the Resident layout");
    const long u = U;
    auto key = [u](int c) -> long { return c < 0 ? static_cast<long>(-c - 1) : u + c; };
    if (row_order && !row_order_built) {
        const long nsym = crow_host[0];
        std::vector<int> crow(static_cast<std::size_t>(rows) + 1), csym(static_cast<std::size_t>(nsym));
        GMR_HIP(hipDeviceSynchronize());
        GMR_HIP(hipMemcpy(crow.data(), d_crow, sizeof(int) * crow.size(), hipMemcpyDeviceToHost));
        if (nsym) GMR_HIP(hipMemcpy(csym.data(), d_csym, sizeof(int) * nsym, hipMemcpyDeviceToHost));
        row_order_host = similar_row_order(crow, csym, rows);
        std::vector<int> inv(static_cast<std::size_t>(rows));
        for (int j = 0; j < rows; ++j) inv[static_cast<std::size_t>(row_order_host[j])] = j;
        d_row_new = dev_upload(inv.data(), inv.size(), &struct_bytes);
        GMR_HIP(hipMalloc(&ws_yrow, sizeof(float) * std::max(rows, 1)));
        row_order_built = true;
    }
    return lparts.emplace(mk, range_lists(*this, u + NL, span, key, true, row_order ? &row_order_host : nullptr))
        .first->second;
}

void Engine::ensure_workspace(int Bt, int vbytes) {
    if (no_alloc) return;  // the workspace is the scratch array
    const std::size_t tb = static_cast<std::size_t>(U) * Bt * vbytes;
    const std::size_t wn = resident ? static_cast<std::size_t>(NL) : static_cast<std::size_t>(max_width);
    const std::size_t wb = std::max<std::size_t>(wn, 1) * Bt * vbytes;
    if (tb > ws_T_bytes || wb > ws_W_bytes) {
        // Buffers move: captured graphs hold the old addresses.
        for (auto& kv : graphs) {
            if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
            if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
        }
        graphs.clear();
        GMR_HIP(hipDeviceSynchronize());
    }
    if (tb > ws_T_bytes) {
        if (ws_T) GMR_HIP(hipFree(ws_T));
        ws_T = nullptr;
        GMR_HIP(hipMalloc(&ws_T, tb));
        ws_T_bytes = tb;
    }
    if (wb > ws_W_bytes) {
        if (ws_W0) GMR_HIP(hipFree(ws_W0));
        if (ws_W1) GMR_HIP(hipFree(ws_W1));
        ws_W0 = ws_W1 = nullptr;
        GMR_HIP(hipMalloc(&ws_W0, wb));
        if (!resident) GMR_HIP(hipMalloc(&ws_W1, wb));
        ws_W_bytes = wb;
    }
}

// The bitset of the Boolean bit mode: ceil(U / 64) terminal words, then one
// word per 64 rule slots. Its size is fixed by the grammar, so it never moves.
unsigned long long* Engine::ensure_bits() {
    if (!ws_bits) {
        forbid_alloc("the bit-mode bitset");
        bits_tw = (U + 63) / 64;
        const std::size_t words = static_cast<std::size_t>(bits_tw) + (static_cast<std::size_t>(NL) + 63) / 64 + 1;
        GMR_HIP(hipMalloc(&ws_bits, words * sizeof(unsigned long long)));
    }
    return ws_bits;
}

namespace {

using dev::SweepArgs;

// Row width of T and W for a tile of Bt vectors: a multiple of 4 so rows are
// 16-byte aligned and every lane moves four values (one vector stays 1).
// 128-byte lines a row of w floats touches, averaged over the offsets rows of
// that stride take in an array with a line-aligned base.
inline double row_lines(int w) {
    const int bytes = 4 * w;
    int g = 128;
    while (bytes % g) g >>= 1;  // rows start at the multiples of gcd(bytes, 128)
    double sum = 0.0;
    for (int off = 0; off < 128; off += g) sum += (off + bytes + 127) / 128;
    return sum * g / 128.0;
}

// Row width of the value arrays for tiles of Bt vectors: a multiple of 4
// (dwordx4 lanes), or 8, 16 or a multiple of 32 floats where that touches at
// least 15% fewer 128-byte lines per row (every kernel reads and writes whole
// padded rows, so the lines a row touches are what a random row access costs:
// 27 floats in rows of 28 touch 1.75 lines on average, in rows of 32 one).
inline int padded_width(int Bt) {
    if (Bt <= 1) return 1;
    int best = (Bt + 3) / 4 * 4;
    for (int w : {8, 16, (Bt + 31) / 32 * 32})
        if (w > best && row_lines(w) * 1.15 <= row_lines(best)) best = w;
    return best;
}

// Blocks for `work` threads, rounded up to a multiple of 8 (one share per XCD;
// the kernels' grid-stride loops leave surplus threads idle).
inline unsigned grid_for(long work, int block, int cap = 1 << 20) {
    long g = (work + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    if (g > 8) g = (g + 7) / 8 * 8;
    return static_cast<unsigned>(g);
}

// The plain sweep: one launch per phase and level, in the reference's order
// (init, terminal emits, then per level the node update and its emits).
inline int kinds_hash(const Engine& en, const gmr_run_opts& o) {
    unsigned h = 2166136261u ^ static_cast<unsigned>(o.eval_kind);
    if (o.level_kind)
        for (int k = 1; k <= en.L; ++k) h = (h ^ static_cast<unsigned>(o.level_kind[k])) * 16777619u;
    return static_cast<int>(h & 0x7fffff);
}

// Grid for one level's node-update launch of the given kind.
inline unsigned eval_grid(const SweepArgs& a, int kind, long w, int block) {
    if (kind == dev::EV_LDS && a.Bt != 1) kind = dev::EV_GROUP;
    switch (kind) {
        case dev::EV_THREAD: return grid_for(w * (a.vec ? a.Bs / 4 : a.Bs), block);
        case dev::EV_WAVE: return grid_for(w, block / 64);
        case dev::EV_LDS: return grid_for(w, block);
        default: return grid_for((a.vec || a.Bs == 1) ? (w + 3) / 4 : w, block / a.G);  // four rules per group
    }
}

inline int host_kind(const gmr_run_opts& o, int k) { return o.level_kind ? o.level_kind[k] : o.eval_kind; }

// Workgroups of an LDS-range gather (1,024 lanes, `lds` bytes of LDS each):
// two per CU when a range takes at most 64 KiB, else one, and no more than
// one per 64 chunks.
inline unsigned lds_gather_grid(const Engine& en, long nch, std::size_t lds) {
    const long per_cu = lds <= 65536 ? 2 : 1;
    return static_cast<unsigned>(std::max(1L, std::min<long>(per_cu * en.num_cus, (nch + 63) / 64)));
}

static_assert(Engine::kTailStageBytes == 2 * dev::kTailRules * 4, "the fused tail's staging buffer");

// First level of the fused tail (GMR_RUN_FUSED_TAIL): the lowest k whose
// levels k .. L hold at most kTailRules rules, if they are two levels or more;
// L + 1 (no tail) otherwise or without the flag.
inline int tail_start(const Engine& en, const SweepArgs& a, const gmr_run_opts& o) {
    if (!(o.flags & GMR_RUN_FUSED_TAIL) || !a.resident || a.Bs != 1 || a.bits ||
        en.max_lds < dev::kTailLdsBytes)
        return en.L + 1;
    int k = en.L + 1;
    while (k > 1 && en.level_off[en.L] - en.level_off[k - 2] <= dev::kTailRules) --k;
    return en.L - k + 1 >= 2 ? k : en.L + 1;
}

template <class S>
void launch_levels(const Engine& en, const SweepArgs& a, const gmr_run_opts& o, int block, hipStream_t st,
                   hipEvent_t* level_ev = nullptr) {
    const int G = a.G;
    if (level_ev) GMR_HIP(hipEventRecord(level_ev[0], st));
    if (a.bits) {
        const long init = std::max<long>(static_cast<long>(a.tw) * 64, std::max<long>(a.rows, (en.NL + 63) / 64));
        hipLaunchKernelGGL(dev::k_bits_init, dim3(grid_for(init, block, 4096)), dim3(block), 0, st, a);
        // GMR_RUN_PAIRED_LEVELS: levels two to a launch (each pair's time is its lower level's)
        const bool paired = (o.flags & GMR_RUN_PAIRED_LEVELS) != 0;
        for (int k = 1; k <= en.L; ++k) {
            const long lo = en.level_off[k - 1], hi = en.level_off[k];
            if (level_ev) GMR_HIP(hipEventRecord(level_ev[2 * k - 1], st));
            if (paired && k + 1 <= en.L) {
                const long hi2 = en.level_off[k + 1];
                if (hi2 > lo)
                    hipLaunchKernelGGL(dev::k_bits_eval_pair, dim3(grid_for(((hi2 + 63) / 64 - lo / 64) * 64, block)),
                                       dim3(block), 0, st, a, k);
                if (level_ev) {
                    GMR_HIP(hipEventRecord(level_ev[2 * k], st));
                    GMR_HIP(hipEventRecord(level_ev[2 * k + 1], st));
                    GMR_HIP(hipEventRecord(level_ev[2 * k + 2], st));
                }
                ++k;
                continue;
            }
            if (hi > lo)
                hipLaunchKernelGGL(dev::k_bits_eval, dim3(grid_for(((hi + 63) / 64 - lo / 64) * 64, block)), dim3(block),
                                   0, st, a, k);
            if (level_ev) GMR_HIP(hipEventRecord(level_ev[2 * k], st));
        }
        if (!a.flat) {
            // one LDS range per list; workgroups take equal shares of all chunks
            const long nch = en.bparts.at(a.bit_range).poff.back();
            const std::size_t lds = static_cast<std::size_t>(a.bit_range) / 8;
            const unsigned g = lds_gather_grid(en, nch, lds);
            if (nch > 0) {
                if (a.chunks == 2)
                    hipLaunchKernelGGL(dev::k_bits_gather_lds<2>, dim3(g), dim3(dev::kBitGatherBlock), lds, st, a);
                else if (a.chunks == 8)
                    hipLaunchKernelGGL(dev::k_bits_gather_lds<8>, dim3(g), dim3(dev::kBitGatherBlock), lds, st, a);
                else
                    hipLaunchKernelGGL(dev::k_bits_gather_lds<4>, dim3(g), dim3(dev::kBitGatherBlock), lds, st, a);
            }
        } else if (en.n_fchunk > 0) {
            hipLaunchKernelGGL(dev::k_bits_gather,
                               dim3(grid_for((static_cast<long>(en.n_fchunk) + dev::kBitChunks - 1) / dev::kBitChunks * 64,
                                             block)),
                               dim3(block), 0, st, a);
        }
        if (level_ev) GMR_HIP(hipEventRecord(level_ev[2 * en.L + 1], st));
        GMR_HIP(hipGetLastError());
        return;
    }
    hipLaunchKernelGGL(dev::k_init<S>, dim3(grid_for(dev::init_items(a), block, 4096)),
                       dim3(block), 0, st, a);
    if (!a.gather && en.Et > 0) {
        long work = (a.Bt == 1 && a.B == 1) ? en.Et : ((static_cast<long>(en.Et) + dev::kEmitChunk - 1) / dev::kEmitChunk) * G;
        hipLaunchKernelGGL(dev::k_emit<S>, dim3(grid_for(work, block)), dim3(block), 0, st, a, 0);
    }
    const int k_tail = tail_start(en, a, o);
    const bool paired = (o.flags & GMR_RUN_PAIRED_LEVELS) && a.resident && a.Bs == 1 && !a.bits;
    int pair_hi = 0;  // a level below this one was computed with the pair launch before it
    for (int k = 1; k <= en.L; ++k) {
        const int w = en.level_off[k] - en.level_off[k - 1];
        const int kind = host_kind(o, k);
        if (level_ev) GMR_HIP(hipEventRecord(level_ev[2 * k - 1], st));
        if (paired && k + 1 < k_tail && k > pair_hi &&
            en.level_off[k + 1] - en.level_off[k - 1] <= dev::kPairMaxRules) {
            // levels k and k + 1 in one launch (their time is level k's)
            const long n2 = en.level_off[k + 1] - en.level_off[k - 1];
            hipLaunchKernelGGL(dev::k_eval_pair<S>, dim3(grid_for((n2 + dev::kPairRules - 1) / dev::kPairRules, block)),
                               dim3(block), 0, st, a, k);
            pair_hi = k + 1;
        } else if (k <= pair_hi) {
            // computed by the pair launch of level k - 1
        } else if (k == k_tail) {
            // levels k .. L at once (their time is level k's)
            auto* stage = static_cast<unsigned*>(a.tail_stage);
            const long n = en.level_off[en.L] - en.level_off[k - 1];
            hipLaunchKernelGGL(dev::k_tail_stage<S>, dim3(grid_for(n, block)), dim3(block), 0, st, a, k, stage);
            hipLaunchKernelGGL(dev::k_eval_tail<S>, dim3(1), dim3(dev::kTailBlock),
                               static_cast<std::size_t>(dev::kTailLdsBytes), st, a, k,
                               static_cast<const unsigned*>(stage));
        } else if (w > 0 && k < k_tail) {
            const bool lds = kind == dev::EV_LDS && a.Bt == 1;
            hipLaunchKernelGGL(dev::k_eval<S>, dim3(eval_grid(a, kind, w, block)), dim3(block),
                               lds ? static_cast<std::size_t>(a.lds_window) * sizeof(typename S::T) : 0, st, a, k, kind);
        }
        if (level_ev) GMR_HIP(hipEventRecord(level_ev[2 * k], st));
        const int ne = en.emit_off[k] - en.emit_off[k - 1];
        if (!a.gather && ne > 0) {
            long work = (a.Bt == 1 && a.B == 1) ? ne : ((static_cast<long>(ne) + dev::kEmitChunk - 1) / dev::kEmitChunk) * G;
            hipLaunchKernelGGL(dev::k_emit<S>, dim3(grid_for(work, block)), dim3(block), 0, st, a, k);
        }
    }
    if (a.gather && a.xparts) {
        // One launch per pass: the lists of a pass share the XCDs' L2s only
        // with each other.
        const int passes = a.xpass1;  // make_args sets [0, passes)
        const std::vector<int>& poff = en.xparts.at(passes).poff;
        for (int pass = 0; pass < passes; ++pass) {
            SweepArgs ap = a;
            ap.xpass0 = pass;
            ap.xpass1 = pass + 1;
            long most = 0;
            for (int q = 0; q < 8; ++q) most = std::max<long>(most, poff[8 * pass + q + 1] - poff[8 * pass + q]);
            const int K = a.chunks > 0 ? a.chunks : dev::kPartChunks;
            if (most > 0)  // K chunks per wave
                hipLaunchKernelGGL(dev::k_gather<S>, dim3(grid_for((most + K - 1) / K * 8 * 64, block)), dim3(block), 0,
                                   st, ap);
        }
    } else if (a.gather && a.lspan) {
        // one LDS range per list; workgroups take equal shares of all chunks
        const long nch = en.lparts.at(2 * a.lspan + a.lorder).poff.back();
        const std::size_t lds = static_cast<std::size_t>(a.lspan) * sizeof(typename S::T);
        // two workgroups per CU where two ranges fit in one CU's LDS
        const bool two = 2 * lds <= static_cast<std::size_t>(en.max_lds);
        const unsigned g = static_cast<unsigned>(
            std::max(1L, std::min<long>((two ? 2 : 1) * static_cast<long>(en.num_cus), (nch + 63) / 64)));
        if (nch > 0) {
            const dim3 gr(g), bl(dev::kLdsGatherBlock);
            if (two) {
                if (a.chunks == 2) hipLaunchKernelGGL((dev::k_gather_lds<S, 2, 8>), gr, bl, lds, st, a);
                else if (a.chunks == 8) hipLaunchKernelGGL((dev::k_gather_lds<S, 8, 8>), gr, bl, lds, st, a);
                else hipLaunchKernelGGL((dev::k_gather_lds<S, 4, 8>), gr, bl, lds, st, a);
            } else {
                if (a.chunks == 2) hipLaunchKernelGGL((dev::k_gather_lds<S, 2, 4>), gr, bl, lds, st, a);
                else if (a.chunks == 8) hipLaunchKernelGGL((dev::k_gather_lds<S, 8, 4>), gr, bl, lds, st, a);
                else hipLaunchKernelGGL((dev::k_gather_lds<S, 4, 4>), gr, bl, lds, st, a);
            }
        }
        // rows in the lists' order: y into the caller's Y in matrix row order
        if (a.lorder) hipLaunchKernelGGL(dev::k_rows_to_matrix_order<S>, dim3(grid_for(a.rows, block)), dim3(block), 0, st, a);
    } else if (a.gather) {
        // (the B-wide row gather takes two segments per lane group at once)
        const long threads = a.flat ? static_cast<long>(en.n_fchunk) * 64
                             : (a.Bt == 1 && a.B == 1) ? static_cast<long>(en.n_gseg) * a.GR
                                                       : (static_cast<long>(en.n_gseg) + 1) / 2 * G;
        hipLaunchKernelGGL(dev::k_gather<S>, dim3(grid_for(threads, block)), dim3(block), 0, st, a);
    }
    if (level_ev) GMR_HIP(hipEventRecord(level_ev[2 * en.L + 1], st));
    GMR_HIP(hipGetLastError());
}

// Workgroups per CU of a persistent launch. The barrier needs every workgroup
// resident at once, and the occupancy query can promise one workgroup per CU
// more than the dispatcher actually keeps resident (measured: 7/CU reported,
// fewer held, at 71 VGPRs x 256 lanes), so one slot of the reported occupancy
// stays free, and at most 16 waves per CU are aimed for. A block size at which
// only one workgroup fits per CU leaves no such slot and is refused.
inline int persistent_per_cu(int occ, int block) {
    if (occ < 2)
        throw std::invalid_argument(
            "persistent variant: only " + std::to_string(occ) + " workgroup(s) of " + std::to_string(block) +
            " lanes fit per CU, which leaves no residency headroom for the grid barrier; use a smaller block");
    return std::min(occ - 1, std::max(1, 1024 / block));
}

// One launch for the whole sweep, sized to one resident wave of workgroups
// (occupancy x CUs with the headroom above, capped) so every workgroup is on
// the chip while the barrier waits; a wait that outlives its deadline marks
// the launch failed (reported by the next gmr_engine_run or gmr_engine_check).
template <class S>
void launch_persistent(Engine& en, const SweepArgs& a, int block, int grid_cap, hipStream_t st, int tiles = 1,
                       int tile_w = 0, int vbytes = 4) {
    int occ = 0;
    GMR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, dev::k_persistent<S>, block, 0));
    const int per_cu = persistent_per_cu(occ, block);
    int grid = per_cu * en.num_cus;
    if (grid_cap > 0) grid = std::min(grid, grid_cap);
    en.last_grid = grid;
    en.last_per_cu = per_cu;
    hipLaunchKernelGGL(dev::k_persistent<S>, dim3(grid), dim3(block), 0, st, a, en.barrier.ref(), tiles,
                       tile_w > 0 ? tile_w : a.Bt, vbytes);
    GMR_HIP(hipGetLastError());
    en.barrier.advance(static_cast<unsigned long long>(grid) *
                       (static_cast<unsigned long long>(tiles) * static_cast<unsigned long long>(a.L + 2) - 1ULL));
}

template <class S>
void run_tile(Engine& en, int semiring, SweepArgs a, const gmr_run_opts& o, hipStream_t st) {
    const int block = o.block > 0 ? o.block : 256;
    if (o.variant == GMR_VARIANT_PERSISTENT) {
        launch_persistent<S>(en, a, block, o.grid_cap, st);
    } else if (o.variant == GMR_VARIANT_GRAPH) {
        // Everything the captured launches bake in: arguments, kernels per
        // level, gather form and pass count.
        auto key = std::make_tuple(semiring * 1000003 + kinds_hash(en, o), reinterpret_cast<uintptr_t>(a.X),
                                   reinterpret_cast<uintptr_t>(a.Y), a.B, a.Bt,
                                   a.gather | (o.flags << 1) | (o.xcd_passes << 8) | (o.chunks << 12),
                                   o.lds_values, block,
                                   reinterpret_cast<uintptr_t>(st));
        auto it = en.graphs.find(key);
        if (it == en.graphs.end()) {
            en.forbid_alloc("a captured graph");
            // Capture on the engine's own stream (the caller's may be the
            // NULL stream, which cannot be captured); replay on the caller's.
            if (!en.capture_stream) GMR_HIP(hipStreamCreateWithFlags(&en.capture_stream, hipStreamNonBlocking));
            Engine::Graph g;
            GMR_HIP(hipStreamBeginCapture(en.capture_stream, hipStreamCaptureModeThreadLocal));
            try {
                launch_levels<S>(en, a, o, block, en.capture_stream);
            } catch (...) {
                hipGraph_t dummy;
                (void)hipStreamEndCapture(en.capture_stream, &dummy);
                throw;
            }
            GMR_HIP(hipStreamEndCapture(en.capture_stream, &g.graph));
            GMR_HIP(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
            it = en.graphs.emplace(key, g).first;
        }
        GMR_HIP(hipGraphLaunch(it->second.exec, st));
    } else {
        launch_levels<S>(en, a, o, block, st);
    }
}

// Arguments of one tile of vectors [b0, b0 + Bt) of one engine.
SweepArgs make_args(Engine& en, const float* X, int B, void* Y, int vbytes, int b0, int Bt, bool gather,
                    const gmr_run_opts& o) {
    SweepArgs a{};
    a.kids = en.d_kids;
    a.coeff = en.d_coeff;
    a.tcol = en.d_tcol;
    a.tval = en.d_tval;
    a.emits = en.d_emits;
    a.term_emits = en.d_term;
    a.crow = en.d_crow;
    a.csym = en.d_csym;
    a.gseg = en.d_gseg;
    a.n_gseg = en.n_gseg;
    a.gsplit = en.gather_split;
    a.zrows = en.d_zrows;
    a.n_zrows = en.n_zrows;
    a.fbits = en.d_fbits;
    a.ford = en.d_ford;
    a.frow = en.d_frow;
    a.n_fchunk = en.n_fchunk;
    a.n_csym = en.resident ? en.crow_host[0] : 0;
    a.level_off = en.d_level_off;
    a.emit_off = en.d_emit_off;
    a.X = X + b0;
    a.Y = static_cast<char*>(Y) + static_cast<std::size_t>(b0) * vbytes;
    a.Tv = en.ws_T;
    a.W0 = en.ws_W0;
    a.W1 = en.ws_W1;
    a.rows = en.rows;
    a.cols = en.cols;
    a.U = en.U;
    a.L = en.L;
    a.n_term = en.Et;
    a.B = B;
    a.Bt = std::min(Bt, B - b0);
    a.Bs = padded_width(a.Bt);
    a.vec = a.Bt > 1 ? 1 : 0;
    a.yvec = (B % 4 == 0 && b0 % 4 == 0) ? 1 : 0;
    a.G = dev::lanes_for(a.Bs, a.vec);
    a.resident = en.resident ? 1 : 0;
    a.gather = gather ? 1 : 0;
    a.flat = (gather && a.Bt == 1 && B == 1 && !(o.flags & GMR_RUN_ROW_SEGMENTS)) ? 1 : 0;
    const bool narrow = gather && en.resident && a.vec && a.Bs == 4 && !(o.flags & GMR_RUN_ROW_SEGMENTS);
    a.xparts = ((a.flat && en.resident && !(o.flags & GMR_RUN_SINGLE_LIST)) || narrow) ? 1 : 0;
    if (a.xparts) {
        if (o.xcd_passes < 0 || o.xcd_passes > 8) throw std::invalid_argument("xcd_passes must be in [0, 8]");
        const int passes = o.xcd_passes > 0 ? o.xcd_passes : en.auto_xcd_passes(a.Bs, a.Bt);
        const Engine::XParts& x = en.xparts_for(passes);
        a.xsym = x.d_sym;
        a.xbits = x.d_bits;
        a.xord = x.d_ord;
        a.xrow = x.d_row;
        a.xpoff = x.d_poff;
        a.xpass0 = 0;
        a.xpass1 = passes;
    }
    const bool lds_gather = (o.flags & (GMR_RUN_BITS | GMR_RUN_LDS_RANGES)) != 0;
    if ((o.flags & GMR_RUN_BITS) && (o.flags & GMR_RUN_LDS_RANGES))
        throw std::invalid_argument("GMR_RUN_BITS and GMR_RUN_LDS_RANGES exclude each other");
    if (o.lds_values != 0 && (o.lds_values < 0 || (o.lds_values % 1024) != 0))
        throw std::invalid_argument("lds_values must be 0 or a multiple of 1024");
    if (o.chunks != 0) {
        const bool ok = lds_gather ? (o.chunks == 2 || o.chunks == 4 || o.chunks == 8)
                                   : (o.chunks == 1 || o.chunks == 2 || o.chunks == 4);
        if (!ok) throw std::invalid_argument(lds_gather ? "chunks must be 0, 2, 4 or 8 for the LDS-range gathers"
                                                        : "chunks must be 0, 1, 2 or 4");
    }
    a.chunks = o.chunks > 0 ? o.chunks : ((o.flags & GMR_RUN_BITS) ? 8 : lds_gather ? 4 : dev::kPartChunks);
    // EV_LDS window of the per-level launches (lds_values, unless it sizes an LDS-range gather)
    a.lds_window = dev::kLdsWindow;
    if (!lds_gather && o.lds_values > 0) {
        // (the kernel also holds a small static array for its block min/max)
        if (o.lds_values < 1024 || o.lds_values > 16384 ||
            static_cast<long>(o.lds_values) * 4 + 1024 > static_cast<long>(en.max_lds))
            throw std::invalid_argument("lds_values (the EV_LDS window) must be in [1024, 16384]");
        a.lds_window = o.lds_values;
    }
    if (o.flags & GMR_RUN_TRANSPOSED)
        throw std::invalid_argument("GMR_RUN_TRANSPOSED applies to gmr_engine_traverse only");
    if ((o.flags & (GMR_RUN_FUSED_TAIL | GMR_RUN_PAIRED_LEVELS)) &&
        (!en.resident || B != 1 || o.variant == GMR_VARIANT_PERSISTENT))
        throw std::invalid_argument(
            "GMR_RUN_FUSED_TAIL and GMR_RUN_PAIRED_LEVELS need one vector, the Resident layout and the levels or graph "
            "variant");
    if ((o.flags & GMR_RUN_FUSED_TAIL) && (o.flags & GMR_RUN_BITS))
        throw std::invalid_argument("GMR_RUN_FUSED_TAIL does not apply to the bit mode (GMR_RUN_PAIRED_LEVELS does)");
    if ((o.flags & GMR_RUN_FUSED_TAIL) && !en.ws_tail) {
        en.forbid_alloc("the fused tail's staging buffer");
        GMR_HIP(hipMalloc(&en.ws_tail, Engine::kTailStageBytes));
    }
    a.tail_stage = en.ws_tail;
    if ((o.flags & GMR_RUN_ROW_ORDER) && !(o.flags & GMR_RUN_LDS_RANGES))
        throw std::invalid_argument("GMR_RUN_ROW_ORDER orders the rows of GMR_RUN_LDS_RANGES lists; set both");
    if (lds_gather && (!gather || !en.resident || B != 1 || o.variant == GMR_VARIANT_PERSISTENT))
        throw std::invalid_argument(
            "GMR_RUN_BITS and GMR_RUN_LDS_RANGES need one vector, gather emission (Resident layout) and the levels or "
            "graph variant");
    if (o.flags & GMR_RUN_BITS) {
        a.bits = 1;
        a.vbits = en.ensure_bits();
        a.tw = en.bits_tw;
        a.nvw = static_cast<long>(en.bits_tw) + (static_cast<long>(en.NL) + 63) / 64;
        a.flat = 1;
        a.xparts = 0;
        if (!(o.flags & GMR_RUN_SINGLE_LIST)) {
            // the LDS-range lists (the default); GMR_RUN_SINGLE_LIST keeps the flat list
            const int range = o.lds_values > 0 ? o.lds_values : dev::kBitRangeBits;
            if (range < 131072 || range > 1048576 || range % 64 != 0 ||
                static_cast<long>(range / 8) > static_cast<long>(en.max_lds))
                throw std::invalid_argument("lds_values (bits per range of GMR_RUN_BITS) must be in [131072, 1048576]");
            const Engine::XParts& x = en.bit_lists(range);
            a.bit_range = range;
            a.xsym = x.d_sym;
            a.xbits = x.d_bits;
            a.xord = x.d_ord;
            a.xrow = x.d_row;
            a.xpoff = x.d_poff;
            a.nbl = static_cast<int>(x.poff.size()) - 1;
            a.flat = 0;
        }
    }
    if (o.flags & GMR_RUN_LDS_RANGES) {
        const int span = o.lds_values > 0 ? o.lds_values : 32768;
        if (span < 4096 || span >= 0xFFFF || static_cast<long>(span) * vbytes > static_cast<long>(en.max_lds))
            throw std::invalid_argument(
                "lds_values (values per range of GMR_RUN_LDS_RANGES) must be at least 4096, below 65,535 and fit "
                "the LDS of a workgroup (" + std::to_string(en.max_lds) + " bytes)");
        const bool ordered = (o.flags & GMR_RUN_ROW_ORDER) != 0;
        const Engine::XParts& x = en.lds_lists(span, ordered);
        if (ordered) {
            // pieces combine in the row-ordered copy, which the init zeroes
            a.lorder = 1;
            a.ynew = en.d_row_new;
            a.Yout = a.Y;
            a.Y = en.ws_yrow;
        }
        a.lspan = span;
        a.cbase = x.d_cbase;
        // the range fills compute the terminal values; the init computes only
        // those the rules read
        a.tref = en.d_tref;
        a.n_tref = en.n_tref;
        a.nkey = static_cast<long>(en.U) + en.NL;
        a.xsym = x.d_sym;
        a.xbits = x.d_bits;
        a.xord = x.d_ord;
        a.xrow = x.d_row;
        a.xpoff = x.d_poff;
        a.nbl = static_cast<int>(x.poff.size()) - 1;
        a.flat = 0;
        a.xparts = 0;
    }
    a.GR = en.gather_lanes;
    a.eval_kind = o.eval_kind;
    a.level_kind = o.level_kind ? en.device_kinds(o.level_kind) : nullptr;
    return a;
}

template <class S>
void run_all(Engine& en, int semiring, const float* X, int B, void* Y, const gmr_run_opts& o, hipStream_t st) {
    using T = typename S::T;
    const int Bt = (o.b_tile > 0 && o.b_tile < B) ? o.b_tile : B;
    const bool gather = o.emit == GMR_EMIT_GATHER;
    if (gather && !en.resident) throw std::invalid_argument("gather emission needs the Resident layout");
    if ((o.flags & GMR_RUN_BITS) && !std::is_same<S, dev::BooleanOp>::value)
        throw std::invalid_argument("GMR_RUN_BITS runs the Boolean semiring");
    en.ensure_workspace(padded_width(Bt), static_cast<int>(sizeof(T)));
    if (o.variant == GMR_VARIANT_PERSISTENT && Bt < B) {
        // every vector tile in one launch
        const SweepArgs a = make_args(en, X, B, Y, static_cast<int>(sizeof(T)), 0, Bt, gather, o);
        launch_persistent<S>(en, a, o.block > 0 ? o.block : 256, o.grid_cap, st, (B + Bt - 1) / Bt, Bt,
                             static_cast<int>(sizeof(T)));
        return;
    }
    for (int b0 = 0; b0 < B; b0 += Bt) {
        SweepArgs a = make_args(en, X, B, Y, static_cast<int>(sizeof(T)), b0, Bt, gather, o);
        run_tile<S>(en, semiring, a, o, st);
    }
}

}  // namespace

void Engine::run(int semiring, const float* X, int B, void* Y, const gmr_run_opts& o, hipStream_t st) {
    if (B < 1) throw std::invalid_argument("B must be >= 1");
    if (!X || !Y) throw std::invalid_argument("X or Y is NULL");
    if (o.block != 0 && (o.block < 64 || o.block > 1024 || o.block % 64))
        throw std::invalid_argument("block must be a multiple of 64 in [64, 1024]");
    if (o.variant < 0 || o.variant > 2) throw std::invalid_argument("unknown variant");
    if (o.eval_kind < 0 || o.eval_kind >= dev::kEvalKinds) throw std::invalid_argument("unknown eval kind");
    if (o.level_kind)
        for (int k = 1; k <= L; ++k)
            if (o.level_kind[k] < 0 || o.level_kind[k] >= dev::kEvalKinds) throw std::invalid_argument("unknown eval kind");
    const bool timed = (o.flags & GMR_RUN_TIMED) != 0;
    if (timed) GMR_HIP(hipEventRecord(ev0, st));
    switch (semiring) {
        case GMR_PLUS_TIMES: run_all<dev::PlusTimesOp>(*this, semiring, X, B, Y, o, st); break;
        case GMR_BOOLEAN: run_all<dev::BooleanOp>(*this, semiring, X, B, Y, o, st); break;
        case GMR_TROPICAL: run_all<dev::TropicalOp>(*this, semiring, X, B, Y, o, st); break;
        default: throw std::invalid_argument("unknown semiring");
    }
    if (timed) {
        GMR_HIP(hipEventRecord(ev1, st));
        GMR_HIP(hipEventSynchronize(ev1));
        float ms = 0;
        GMR_HIP(hipEventElapsedTime(&ms, ev0, ev1));
        last_ms = ms;
    }
}

void Engine::forbid_alloc(const char* what) const {
    if (no_alloc)
        throw std::invalid_argument(std::string("no-allocation run: ") + what +
                                    " would be allocated now; run gmr_engine_run once with the same options first "
                                    "(it builds them), or choose options that need none");
}

// The sweep with its workspace (terminal values, then the value array or the
// two frontier buffers, each 256-byte aligned) carved from `scratch`.
void Engine::run_noalloc(int semiring, const float* X, int B, void* Y, const gmr_run_opts& o, hipStream_t st,
                         void* scratch, std::size_t scratch_bytes) {
    if (B < 1) throw std::invalid_argument("B must be >= 1");
    if (o.variant == GMR_VARIANT_GRAPH)
        throw std::invalid_argument("no-allocation run: the levels or persistent variant (a captured graph allocates)");
    int dev = 0;
    GMR_HIP(hipGetDevice(&dev));
    if (dev != device) throw std::invalid_argument("no-allocation run: the engine belongs to another device");
    const int Bt = (o.b_tile > 0 && o.b_tile < B) ? o.b_tile : B;
    const std::size_t Bs = static_cast<std::size_t>(Bt <= 1 ? 1 : (Bt + 3) / 4 * 4);
    auto up = [](std::size_t n) { return (n + 255) / 256 * 256; };
    const std::size_t tb = up(static_cast<std::size_t>(U) * Bs * 4);
    const std::size_t wn = resident ? static_cast<std::size_t>(NL) : static_cast<std::size_t>(max_width);
    const std::size_t wb = up(std::max<std::size_t>(wn, 1) * Bs * 4);
    const std::size_t sb = (o.flags & GMR_RUN_FUSED_TAIL) ? kTailStageBytes : 0;
    const std::size_t need = tb + wb * (resident ? 1 : 2) + sb;
    if (need > scratch_bytes)
        throw std::invalid_argument("no-allocation run: the sweep needs " + std::to_string(need) +
                                    " bytes of workspace (vector tile of " + std::to_string(Bt) +
                                    "), more than the static scratch array's " + std::to_string(scratch_bytes) +
                                    "; use a narrower b_tile or gmr_engine_run");
    void* const saved[4] = {ws_T, ws_W0, ws_W1, ws_tail};
    const std::size_t saved_b[2] = {ws_T_bytes, ws_W_bytes};
    char* base = static_cast<char*>(scratch);
    ws_T = base;
    ws_W0 = base + tb;
    ws_W1 = resident ? nullptr : base + tb + wb;
    if (sb) ws_tail = base + tb + wb * (resident ? 1 : 2);
    ws_T_bytes = tb;
    ws_W_bytes = wb;
    no_alloc = true;
    auto restore = [&] {
        no_alloc = false;
        ws_T = saved[0];
        ws_W0 = saved[1];
        ws_W1 = saved[2];
        ws_tail = saved[3];
        ws_T_bytes = saved_b[0];
        ws_W_bytes = saved_b[1];
    };
    try {
        run(semiring, X, B, Y, o, st);
    } catch (...) {
        restore();
        throw;
    }
    restore();
}

void Engine::set_gather_order(int order) {
    if (!resident || n_gseg == 0) return;
    std::vector<int4> seg(n_gseg);
    std::vector<int> sym(crow_host.back());
    GMR_HIP(hipDeviceSynchronize());
    GMR_HIP(hipMemcpy(seg.data(), d_gseg, sizeof(int4) * n_gseg, hipMemcpyDeviceToHost));
    if (!sym.empty()) GMR_HIP(hipMemcpy(sym.data(), d_csym, sizeof(int) * sym.size(), hipMemcpyDeviceToHost));
    if (order == 0) {
        std::stable_sort(seg.begin(), seg.end(), [](const int4& a, const int4& b) {
            return a.x != b.x ? a.x < b.x : a.y < b.y;
        });
    } else {
        // Segments that share their highest rule (the longest substring they
        // copy) run next to each other, so their value rows are read while
        // still in the caches.
        std::vector<std::pair<long long, int>> key(n_gseg);
        for (int q = 0; q < n_gseg; ++q) {
            int hi = -1;
            for (int j = seg[q].y; j < seg[q].z; ++j) hi = std::max(hi, sym[j]);
            key[q] = {static_cast<long long>(hi) * (1LL << 31) + seg[q].x, q};
        }
        std::sort(key.begin(), key.end());
        std::vector<int4> out(n_gseg);
        for (int q = 0; q < n_gseg; ++q) out[q] = seg[key[q].second];
        seg.swap(out);
    }
    GMR_HIP(hipMemcpy(d_gseg, seg.data(), sizeof(int4) * n_gseg, hipMemcpyHostToDevice));
    for (auto& kv : graphs) {
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
        if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
    }
    graphs.clear();
    gather_order = order;
}

const int* Engine::device_kinds(const int* host_kinds) {
    std::vector<int> v(host_kinds, host_kinds + L + 1);
    if (v != kinds_host || !d_kinds) {
        forbid_alloc("a per-level kernel table (level_kind)");
        if (!d_kinds) GMR_HIP(hipMalloc(&d_kinds, sizeof(int) * (L + 1)));
        // The previous table may still be read by queued work.
        GMR_HIP(hipDeviceSynchronize());
        GMR_HIP(hipMemcpy(d_kinds, v.data(), sizeof(int) * (L + 1), hipMemcpyHostToDevice));
        kinds_host = v;
    }
    return d_kinds;
}

namespace {
template <class S>
void time_levels_t(Engine& en, const float* X, int B, void* Y, const gmr_run_opts& o, int reps, float* out,
                   hipStream_t st) {
    using T = typename S::T;
    if ((o.flags & GMR_RUN_BITS) && !std::is_same<S, dev::BooleanOp>::value)
        throw std::invalid_argument("GMR_RUN_BITS runs the Boolean semiring");
    en.ensure_workspace(padded_width(B), static_cast<int>(sizeof(T)));
    const SweepArgs a = make_args(en, X, B, Y, static_cast<int>(sizeof(T)), 0, B, o.emit == GMR_EMIT_GATHER, o);
    const int block = o.block > 0 ? o.block : 256;
    std::vector<hipEvent_t> ev(2 * en.L + 2);
    for (auto& e : ev) GMR_HIP(hipEventCreate(&e));
    // out[0]: start to the first level (init, terminal emits); out[k]: level k's
    // node update; out[L+1]: after the last node update to the end (the last
    // level's emits, or the row gather).
    std::vector<double> acc(en.L + 2, 0.0);
    for (int r = 0; r < reps; ++r) {
        launch_levels<S>(en, a, o, block, st, ev.data());
        GMR_HIP(hipStreamSynchronize(st));
        float ms = 0;
        GMR_HIP(hipEventElapsedTime(&ms, ev[0], en.L > 0 ? ev[1] : ev[2 * en.L + 1]));
        acc[0] += ms;
        for (int k = 1; k <= en.L; ++k) {
            GMR_HIP(hipEventElapsedTime(&ms, ev[2 * k - 1], ev[2 * k]));
            acc[k] += ms;
        }
        if (en.L > 0) {
            GMR_HIP(hipEventElapsedTime(&ms, ev[2 * en.L], ev[2 * en.L + 1]));
            acc[en.L + 1] += ms;
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    for (int k = 0; k <= en.L + 1; ++k) out[k] = static_cast<float>(acc[k] / std::max(reps, 1));
}
}  // namespace

void Engine::time_levels(int semiring, const float* X, int B, void* Y, const gmr_run_opts& o, int reps,
                         float* out, hipStream_t st) {
    if (B < 1 || !X || !Y || !out || reps < 1) throw std::invalid_argument("bad time_levels arguments");
    switch (semiring) {
        case GMR_PLUS_TIMES: time_levels_t<dev::PlusTimesOp>(*this, X, B, Y, o, reps, out, st); break;
        case GMR_BOOLEAN: time_levels_t<dev::BooleanOp>(*this, X, B, Y, o, reps, out, st); break;
        case GMR_TROPICAL: time_levels_t<dev::TropicalOp>(*this, X, B, Y, o, reps, out, st); break;
        default: throw std::invalid_argument("unknown semiring");
    }
}


// ---------------------------------------------------------------- left product
// Incoming lists of every node, built once from the uploaded structure.
void Engine::build_left() {
    if (left_ready) return;
    if (!resident) throw std::invalid_argument("the left product needs the Resident layout");
    GMR_HIP(hipDeviceSynchronize());
    std::vector<int2> kids(NL), em(E), te(Et);
    std::vector<float> coeff(has_runlen ? NL : 0);
    std::vector<int> tcol(U);
    std::vector<float> tv(U);
    if (NL) GMR_HIP(hipMemcpy(kids.data(), d_kids, sizeof(int2) * NL, hipMemcpyDeviceToHost));
    if (has_runlen && NL) GMR_HIP(hipMemcpy(coeff.data(), d_coeff, sizeof(float) * NL, hipMemcpyDeviceToHost));
    if (E) GMR_HIP(hipMemcpy(em.data(), d_emits, sizeof(int2) * E, hipMemcpyDeviceToHost));
    if (Et) GMR_HIP(hipMemcpy(te.data(), d_term, sizeof(int2) * Et, hipMemcpyDeviceToHost));
    if (U) GMR_HIP(hipMemcpy(tcol.data(), d_tcol, sizeof(int) * U, hipMemcpyDeviceToHost));
    if (U) GMR_HIP(hipMemcpy(tv.data(), d_tval, sizeof(float) * U, hipMemcpyDeviceToHost));
    const long nodes = static_cast<long>(NL) + U;
    // Terminal nodes in column order: compact terminal t is node NL + tpos[t];
    // position 0 is terminal 0, then each column's terminals together, so the
    // column fold reads node values and terminal values in sequence.
    std::vector<int> coff(static_cast<std::size_t>(cols) + 1, 0), tpos(U, 0);
    std::vector<float> cval(U, 0.0f);
    for (int t = 1; t < U; ++t) ++coff[tcol[t] + 1];
    coff[0] = 1;
    for (int c = 0; c < cols; ++c) coff[c + 1] += coff[c];
    {
        std::vector<int> next(coff.begin(), coff.end() - 1);
        for (int t = 1; t < U; ++t) tpos[t] = next[tcol[t]]++;
        for (int t = 0; t < U; ++t) cval[tpos[t]] = tv[t];
    }
    // A child c >= 0 is a slot; c < 0 the compact terminal -c-1.
    // Terminal 0 is the semiring zero: nothing flows into it that matters.
    auto node_of = [&](int c) -> long { return c >= 0 ? c : static_cast<long>(NL) + tpos[-c - 1]; };
    auto wanted = [&](int c) { return c >= 0 || c < -1; };
    std::vector<long> off(static_cast<std::size_t>(nodes) + 1, 0);
    for (int p = 0; p < NL; ++p) {
        if (wanted(kids[p].x)) ++off[node_of(kids[p].x) + 1];
        if (wanted(kids[p].y)) ++off[node_of(kids[p].y) + 1];
    }
    for (const int2& e : em) ++off[e.x + 1];
    for (const int2& e : te)
        if (e.x > 0) ++off[static_cast<long>(NL) + tpos[e.x] + 1];
    for (long n = 0; n < nodes; ++n) off[n + 1] += off[n];
    if (off[nodes] > 0x7fffffffL) throw std::invalid_argument("left product: more than 2^31 incoming entries");
    std::vector<int> src(static_cast<std::size_t>(off[nodes]));
    std::vector<float> cf(has_runlen ? src.size() : 0);
    std::vector<long> fill(off.begin(), off.end() - 1);
    // Parents in slot order, then the top-sequence uses in row order.
    for (int p = 0; p < NL; ++p) {
        if (wanted(kids[p].x)) {
            const long q = fill[node_of(kids[p].x)]++;
            src[q] = p;
            if (has_runlen) cf[q] = coeff[p];
        }
        if (wanted(kids[p].y)) {
            const long q = fill[node_of(kids[p].y)]++;
            src[q] = p;
            if (has_runlen) cf[q] = 1.0f;
        }
    }
    for (const int2& e : em) {
        const long q = fill[e.x]++;
        src[q] = -e.y - 1;
        if (has_runlen) cf[q] = 1.0f;
    }
    for (const int2& e : te) {
        if (e.x <= 0) continue;
        const long q = fill[static_cast<long>(NL) + tpos[e.x]]++;
        src[q] = -e.y - 1;
        if (has_runlen) cf[q] = 1.0f;
    }
    std::vector<int> off32(off.begin(), off.end());
    // One-vector fold: 64-position chunks of positions 1..U-1 with their
    // column-start bitmap (as the chunked row gather).
    n_lf_chunk = U > 1 ? (static_cast<long>(U) - 1 + 63) / 64 : 0;
    std::vector<unsigned long long> lfbits(static_cast<std::size_t>(n_lf_chunk), 0ull);
    std::vector<int> lford(static_cast<std::size_t>(n_lf_chunk), 0), lfcol;
    for (int c = 0; c < cols; ++c) {
        if (coff[c + 1] == coff[c]) continue;
        const long q = coff[c] - 1;
        lfbits[q >> 6] |= 1ull << (q & 63);
        lfcol.push_back(c);
    }
    for (long c = 0, before = 0; c < n_lf_chunk; ++c) {
        lford[c] = static_cast<int>(before);
        before += __builtin_popcountll(lfbits[c]);
    }
    if (!has_runlen) {
        // Row-gather form. Groups: levels 1..L, then the terminals; node order
        // inside a group (nodes of a level are contiguous slots).
        std::vector<long> gnode(L + 2);
        for (int k = 0; k <= L; ++k) gnode[k] = level_off[k];
        gnode[L + 1] = nodes;
        // B > 1: segments of at most kGatherSeg entries per node.
        std::vector<int4> seg;
        std::vector<int> split;
        lseg_off.assign(L + 2, 0);
        for (int gi = 0; gi <= L; ++gi) {
            lseg_off[gi] = static_cast<int>(seg.size());
            for (long n = gnode[gi]; n < gnode[gi + 1]; ++n) {
                const int a0 = static_cast<int>(off[n]), a1 = static_cast<int>(off[n + 1]);
                const int sp = (a1 - a0) > kGatherSeg ? 1 : 0;
                if (sp) split.push_back(static_cast<int>(n));
                if (a1 == a0) seg.push_back(make_int4(static_cast<int>(n), a0, a1, 0));
                for (int q = a0; q < a1; q += kGatherSeg)
                    seg.push_back(make_int4(static_cast<int>(n), q, std::min(a1, q + kGatherSeg), sp));
            }
        }
        lseg_off[L + 1] = static_cast<int>(seg.size());
        // One vector: every group's entries padded to whole 64-entry chunks
        // with the zero row of Xs (row `rows`), row-start bitmap per chunk.
        const int pad = -rows - 1;
        std::vector<int> psrc, poff(static_cast<std::size_t>(nodes) + 1, 0), prow, pord;
        std::vector<unsigned long long> pbits;
        lchunk_off.assign(L + 2, 0);
        for (int gi = 0; gi <= L; ++gi) {
            lchunk_off[gi] = static_cast<int>(psrc.size() / 64);
            for (long n = gnode[gi]; n < gnode[gi + 1]; ++n) {
                poff[n] = static_cast<int>(psrc.size());
                if (off[n + 1] > off[n]) {
                    const std::size_t p0 = psrc.size();
                    if ((p0 >> 6) >= pbits.size()) pbits.resize((p0 >> 6) + 1, 0ull);
                    pbits[p0 >> 6] |= 1ull << (p0 & 63);
                    prow.push_back(static_cast<int>(n));
                }
                psrc.insert(psrc.end(), src.begin() + off[n], src.begin() + off[n + 1]);
            }
            while (psrc.size() % 64) psrc.push_back(pad);
        }
        poff[nodes] = static_cast<int>(psrc.size());
        lchunk_off[L + 1] = static_cast<int>(psrc.size() / 64);
        pbits.resize(psrc.size() / 64, 0ull);
        pord.resize(pbits.size());
        int before = 0;
        for (std::size_t c = 0; c < pbits.size(); ++c) {
            pord[c] = before;
            before += __builtin_popcountll(pbits[c]);
        }
        if (psrc.size() > 0x7fffffffUL) throw std::invalid_argument("left product: too many entries");
        // One vector, XCD-partitioned: value keys (row r of x -> r, the zero
        // row -> rows, rule slot s -> rows + 1 + s), 8 x passes lists per group.
        {
            const double bytes = (static_cast<double>(rows) + 1 + NL) * 4.0;
            lx_passes = std::max(1, std::min(8, static_cast<int>(std::ceil(bytes / (8.0 * 3.0 * 1024 * 1024)))));
            auto key = [&](int c) -> long { return c < 0 ? static_cast<long>(-c - 1) : static_cast<long>(rows) + 1 + c; };
            PartLists PL;
            for (int gi = 0; gi <= L; ++gi)
                build_part_lists(gnode[gi], gnode[gi + 1], off, src, key, static_cast<long>(rows) + 1 + NL, 8 * lx_passes,
                                 pad, PL);
            lxpoff = PL.poff;
            d_lxsym = dev_upload(PL.sym.data(), PL.sym.size(), &struct_bytes);
            d_lxbits = dev_upload(PL.bits.data(), PL.bits.size(), &struct_bytes);
            d_lxord = dev_upload(PL.ord.data(), PL.ord.size(), &struct_bytes);
            d_lxrow = dev_upload(PL.row.data(), PL.row.size(), &struct_bytes);
            d_lxpoff = dev_upload(lxpoff.data(), lxpoff.size(), &struct_bytes);
        }
        d_lseg = dev_upload(seg.data(), seg.size(), &struct_bytes);
        d_lsplit = dev_upload(split.data(), split.size(), &struct_bytes);
        n_lsplit = static_cast<int>(split.size());
        d_lpsrc = dev_upload(psrc.data(), psrc.size(), &struct_bytes);
        d_lpoff = dev_upload(poff.data(), poff.size(), &struct_bytes);
        d_lbits = dev_upload(pbits.data(), pbits.size(), &struct_bytes);
        d_lord = dev_upload(pord.data(), pord.size(), &struct_bytes);
        d_lrow = dev_upload(prow.data(), prow.size(), &struct_bytes);
        n_lpsrc = static_cast<int>(psrc.size());
    }
    d_in_off = dev_upload(off32.data(), off32.size(), &struct_bytes);
    d_in_src = dev_upload(src.data(), src.size(), &struct_bytes);
    d_in_coeff = has_runlen ? dev_upload(cf.data(), cf.size(), &struct_bytes) : nullptr;
    d_col_off = dev_upload(coff.data(), coff.size(), &struct_bytes);
    d_col_val = dev_upload(cval.data(), cval.size(), &struct_bytes);
    d_lf_bits = dev_upload(lfbits.data(), lfbits.size(), &struct_bytes);
    d_lf_ord = dev_upload(lford.data(), lford.size(), &struct_bytes);
    d_lf_col = dev_upload(lfcol.data(), lfcol.size(), &struct_bytes);
    left_ready = true;
}

namespace {

// The row-gather arguments for one group of nodes: nodes play rows, incoming
// entries play top-sequence symbols, A plays both the rule values and Y (row
// stride Bs), Xs the terminal rows.
SweepArgs left_gather_args(Engine& en, int Bt, int Bs) {
    SweepArgs g{};
    g.W0 = en.ws_A;
    g.Tv = en.ws_Xs;
    g.Y = en.ws_A;
    g.B = Bs;
    g.Bt = Bt;
    g.Bs = Bs;
    g.vec = Bt > 1 ? 1 : 0;
    g.yvec = (Bs % 4 == 0) ? 1 : 0;
    g.G = dev::lanes_for(Bs, g.vec);
    g.gather = 1;
    g.resident = 1;
    g.rows = static_cast<int>(std::min<long>(static_cast<long>(en.NL) + en.U, 0x7fffffffL));
    if (Bt == 1) {
        g.flat = 1;
        g.csym = en.d_lpsrc;
        g.crow = en.d_lpoff;
        g.fbits = en.d_lbits;
        g.ford = en.d_lord;
        g.frow = en.d_lrow;
        g.n_csym = en.n_lpsrc;
        g.n_fchunk = en.n_lpsrc / 64;
    } else {
        g.csym = en.d_in_src;
        g.gseg = en.d_lseg;
        g.n_gseg = en.lseg_off.back();
        g.gsplit = en.n_lsplit > 0 ? 1 : 0;
    }
    return g;
}

// Groups of at least this many 64-entry chunks use the XCD-partitioned lists.
constexpr long kLeftPartMinChunks = 4096;

template <class S>
void launch_left(Engine& en, const dev::LeftArgs& a, int block, hipStream_t st) {
    const bool fast = !en.has_runlen;
    if (fast) {
        const long zero_all = a.Bt == 1 ? (static_cast<long>(en.NL) + en.U) * a.Bs : 0;
        // (+,x) over whole rows of X (the row segments of B > 1 never read the
        // padding row): the gathers read X in place of a converted copy.
        const bool x_direct = std::is_same<S, dev::PlusTimesOp>::value && a.Bt > 1 && a.Bt == a.B && a.Bs == a.B;
        const long work = std::max<long>(x_direct ? 0 : static_cast<long>(en.rows + 1) * a.Bs,
                                         zero_all ? zero_all : static_cast<long>(en.n_lsplit) * a.Bs);
        if (work > 0)
            hipLaunchKernelGGL(dev::k_left_init<S>, dim3(grid_for(work, block, 4096)), dim3(block), 0, st, a.X, a.B,
                               a.Bt, a.Bs, en.rows, x_direct ? nullptr : en.ws_Xs, en.ws_A, en.d_lsplit,
                               static_cast<long>(en.n_lsplit), zero_all);
        SweepArgs g = left_gather_args(en, a.Bt, a.Bs);
        if (x_direct) g.Tv = const_cast<float*>(a.X);
        const int P = 8 * en.lx_passes;
        // group gi = level gi + 1 for gi < L, the terminals for gi = L
        auto group = [&](int gi) {
            // Small groups (the top levels) are not worth a launch per pass:
            // they take the single chunked list.
            const long chunks = static_cast<long>(en.lchunk_off[gi + 1]) - en.lchunk_off[gi];
            if (g.flat && a.xparts && chunks >= kLeftPartMinChunks) {
                // one launch per pass, as in the right product's gather
                for (int pass = 0; pass < en.lx_passes; ++pass) {
                    SweepArgs gp = g;
                    gp.xparts = 1;
                    gp.xsym = en.d_lxsym;
                    gp.xbits = en.d_lxbits;
                    gp.xord = en.d_lxord;
                    gp.xrow = en.d_lxrow;
                    gp.xpoff = en.d_lxpoff + static_cast<long>(gi) * (P + 1);
                    gp.xpass0 = pass;
                    gp.xpass1 = pass + 1;
                    const int* po = en.lxpoff.data() + static_cast<long>(gi) * (P + 1);
                    long most = 0;
                    for (int q = 0; q < 8; ++q) most = std::max<long>(most, po[8 * pass + q + 1] - po[8 * pass + q]);
                    if (most > 0)  // kPartChunks chunks per wave
                        hipLaunchKernelGGL(dev::k_gather<S>,
                                           dim3(grid_for((most + dev::kPartChunks - 1) / dev::kPartChunks * 8 * 64, block)),
                                           dim3(block), 0, st, gp);
                }
                return;
            }
            long r0, r1, threads;
            if (g.flat) {
                r0 = en.lchunk_off[gi];
                r1 = en.lchunk_off[gi + 1];
                threads = (r1 - r0) * 64;
            } else {
                r0 = en.lseg_off[gi];
                r1 = en.lseg_off[gi + 1];
                threads = (r1 - r0 + 1) / 2 * g.G;  // two segments per lane group
            }
            if (r1 > r0)
                hipLaunchKernelGGL(dev::k_gather_range<S>, dim3(grid_for(threads, block)), dim3(block), 0, st, g, r0, r1);
        };
        for (int k = en.L; k >= 1; --k) group(k - 1);
        group(en.L);
    } else {
        for (int k = en.L; k >= 1; --k) {
            const int n0 = en.level_off[k - 1], n1 = en.level_off[k];
            if (n1 > n0)
                hipLaunchKernelGGL(dev::k_left_nodes<S>, dim3(grid_for(static_cast<long>(n1 - n0) * a.G, block)),
                                   dim3(block), 0, st, a, n0, n1);
        }
        if (en.U > 0)
            hipLaunchKernelGGL(dev::k_left_nodes<S>, dim3(grid_for(static_cast<long>(en.U) * a.G, block)), dim3(block),
                               0, st, a, en.NL, en.NL + en.U);
    }
    if (en.cols > 0 && a.Bt == 1) {
        hipLaunchKernelGGL(dev::k_left_yzero<S>, dim3(grid_for(en.cols, block, 4096)), dim3(block), 0, st, a, en.cols);
        if (en.n_lf_chunk > 0)
            hipLaunchKernelGGL(dev::k_left_fold<S>, dim3(grid_for(en.n_lf_chunk * 64, block)), dim3(block), 0, st, a);
    } else if (en.cols > 0) {
        hipLaunchKernelGGL(dev::k_left_cols<S>, dim3(grid_for(static_cast<long>(en.cols) * a.G, block)), dim3(block),
                           0, st, a, en.cols);
    }
    GMR_HIP(hipGetLastError());
}

template <class S>
void run_left_t(Engine& en, int semiring, const float* X, int B, void* Y, const gmr_run_opts& o, hipStream_t st) {
    using T = typename S::T;
    const int Bt = (o.b_tile > 0 && o.b_tile < B) ? o.b_tile : B;
    const int Bs = en.has_runlen ? Bt : padded_width(Bt);
    const int block = o.block > 0 ? o.block : 256;
    const std::size_t need = static_cast<std::size_t>(std::max<long>(static_cast<long>(en.NL) + en.U, 1)) * Bs * sizeof(T);
    const std::size_t need_x = static_cast<std::size_t>(en.rows + 1) * Bs * sizeof(T);
    if (need > en.ws_A_bytes || need_x > en.ws_Xs_bytes) {
        for (auto& kv : en.left_graphs) {
            if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
            if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
        }
        en.left_graphs.clear();
        GMR_HIP(hipDeviceSynchronize());
    }
    if (need > en.ws_A_bytes) {
        if (en.ws_A) GMR_HIP(hipFree(en.ws_A));
        en.ws_A = nullptr;
        GMR_HIP(hipMalloc(&en.ws_A, need));
        en.ws_A_bytes = need;
    }
    if (need_x > en.ws_Xs_bytes) {
        if (en.ws_Xs) GMR_HIP(hipFree(en.ws_Xs));
        en.ws_Xs = nullptr;
        GMR_HIP(hipMalloc(&en.ws_Xs, need_x));
        en.ws_Xs_bytes = need_x;
    }
    for (int b0 = 0; b0 < B; b0 += Bt) {
        dev::LeftArgs a{};
        a.in_off = en.d_in_off;
        a.in_src = en.d_in_src;
        a.in_coeff = en.d_in_coeff;
        a.col_off = en.d_col_off;
        a.col_val = en.d_col_val;
        a.fbits = en.d_lf_bits;
        a.ford = en.d_lf_ord;
        a.fcol = en.d_lf_col;
        a.n_fchunk = en.n_lf_chunk;
        a.npos = en.U;
        a.X = X + b0;
        a.Y = static_cast<char*>(Y) + static_cast<std::size_t>(b0) * sizeof(T);
        a.A = en.ws_A;
        a.NL = en.NL;
        a.B = B;
        a.Bt = std::min(Bt, B - b0);
        a.Bs = en.has_runlen ? a.Bt : padded_width(a.Bt);
        a.G = dev::lanes_for(a.Bt, 0);
        a.xparts = (a.Bt == 1 && !(o.flags & GMR_RUN_SINGLE_LIST)) ? 1 : 0;
        if (o.variant == GMR_VARIANT_GRAPH) {
            auto key = std::make_tuple(semiring, reinterpret_cast<uintptr_t>(a.X), reinterpret_cast<uintptr_t>(a.Y), a.B,
                                       a.Bt, a.xparts, 0, block, reinterpret_cast<uintptr_t>(st));
            auto it = en.left_graphs.find(key);
            if (it == en.left_graphs.end()) {
                if (!en.capture_stream) GMR_HIP(hipStreamCreateWithFlags(&en.capture_stream, hipStreamNonBlocking));
                Engine::Graph g;
                GMR_HIP(hipStreamBeginCapture(en.capture_stream, hipStreamCaptureModeThreadLocal));
                try {
                    launch_left<S>(en, a, block, en.capture_stream);
                } catch (...) {
                    hipGraph_t dummy;
                    (void)hipStreamEndCapture(en.capture_stream, &dummy);
                    throw;
                }
                GMR_HIP(hipStreamEndCapture(en.capture_stream, &g.graph));
                GMR_HIP(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
                it = en.left_graphs.emplace(key, g).first;
            }
            GMR_HIP(hipGraphLaunch(it->second.exec, st));
        } else {
            launch_left<S>(en, a, block, st);
        }
    }
}

}  // namespace

void Engine::run_left(int semiring, const float* X, int B, void* Y, const gmr_run_opts& o, hipStream_t st) {
    if (B < 1) throw std::invalid_argument("B must be >= 1");
    if (!X || !Y) throw std::invalid_argument("X or Y is NULL");
    if (o.block != 0 && (o.block < 64 || o.block > 1024 || o.block % 64))
        throw std::invalid_argument("block must be a multiple of 64 in [64, 1024]");
    if (o.variant != GMR_VARIANT_LEVELS && o.variant != GMR_VARIANT_GRAPH)
        throw std::invalid_argument("the left product runs as per-level launches or one hipGraph");
    if (o.flags & GMR_RUN_TRANSPOSED)
        throw std::invalid_argument("GMR_RUN_TRANSPOSED applies to gmr_engine_traverse only");
    build_left();
    switch (semiring) {
        case GMR_PLUS_TIMES: run_left_t<dev::PlusTimesOp>(*this, semiring, X, B, Y, o, st); break;
        case GMR_BOOLEAN: run_left_t<dev::BooleanOp>(*this, semiring, X, B, Y, o, st); break;
        case GMR_TROPICAL: run_left_t<dev::TropicalOp>(*this, semiring, X, B, Y, o, st); break;
        default: throw std::invalid_argument("unknown semiring");
    }
}

// Fixpoint traversal over a square matrix: x <- x (+) x^T M until nothing
// moves (Boolean: the set reachable from x along row -> column edges;
// Tropical: shortest distances from the x distances, Bellman-Ford style).
// The change flag is read back after every product.
int Engine::traverse(int semiring, float* x, int max_iters, const gmr_run_opts& o, hipStream_t st) {
    if (rows != cols) throw std::invalid_argument("traversal needs a square matrix");
    if (semiring != GMR_BOOLEAN && semiring != GMR_TROPICAL)
        throw std::invalid_argument("traversal runs in the Boolean or Tropical semiring");
    if (!x) throw std::invalid_argument("x is NULL");
    if (max_iters < 1) throw std::invalid_argument("max_iters must be >= 1");
    // y, then the frontier (the vertices the last step moved), then the flag
    const std::size_t vb = (static_cast<std::size_t>(std::max(rows, 1)) * sizeof(int) + 255) / 256 * 256;
    const std::size_t need = 2 * vb + 256;
    if (need > ws_trav_bytes) {
        GMR_HIP(hipStreamSynchronize(st));
        if (ws_trav) GMR_HIP(hipFree(ws_trav));
        ws_trav = nullptr;
        GMR_HIP(hipMalloc(&ws_trav, need));
        ws_trav_bytes = need;
    }
    void* y = ws_trav;
    float* delta = reinterpret_cast<float*>(static_cast<char*>(ws_trav) + vb);
    int* flag = reinterpret_cast<int*>(static_cast<char*>(ws_trav) + 2 * vb);
    const int block = 256;
    const unsigned grid = grid_for(std::max(rows, 1), block);
    // with GMR_RUN_TRANSPOSED the engine holds M^T: x^T M = (M^T x)^T, a right product
    const bool right = (o.flags & GMR_RUN_TRANSPOSED) != 0;
    gmr_run_opts op = o;
    op.flags &= ~GMR_RUN_TRANSPOSED;
    // Both semirings are idempotent, so a step's product needs only the vertices
    // the step before moved (the first step: all of x): a product over a sparse
    // frontier lets the gathers skip the chunks whose values are all zero.
    int it = 0;
    for (; it < max_iters;) {
        const float* in = it == 0 ? x : delta;
        if (right) run(semiring, in, 1, y, op, st);
        else run_left(semiring, in, 1, y, op, st);
        hipLaunchKernelGGL(dev::k_clear_flag, dim3(1), dim3(64), 0, st, flag);
        if (semiring == GMR_BOOLEAN)
            hipLaunchKernelGGL(dev::k_relax_frontier<dev::BooleanOp>, dim3(grid), dim3(block), 0, st, x, delta, y,
                               rows, flag);
        else
            hipLaunchKernelGGL(dev::k_relax_frontier<dev::TropicalOp>, dim3(grid), dim3(block), 0, st, x, delta, y,
                               rows, flag);
        GMR_HIP(hipGetLastError());
        int h = 0;
        GMR_HIP(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, st));
        GMR_HIP(hipStreamSynchronize(st));
        ++it;
        if (!h) break;
    }
    return it;
}

// ---------------------------------------------------------------- batches
// A fixed set of engines swept together (Y_g = M_g X_g, B_g vectors each).
struct Batch {
    std::vector<Engine*> eng;
    dev::SweepArgs* d_args = nullptr;
    dev::Tile* d_tiles = nullptr;
    long* d_phase_off = nullptr;
    std::size_t cap_args = 0, cap_tiles = 0, cap_phase = 0;
    BarrierState barrier;
    int num_cus = 0, wall_khz = 100000;
    hipStream_t capture_stream = nullptr;
    std::vector<hipStream_t> branch;
    std::vector<hipEvent_t> branch_ev;
    hipEvent_t fork_ev = nullptr;
    std::map<std::size_t, Engine::Graph> graphs;
    std::size_t tiles_key = 0;
    long last_tiles = 0;
    int last_phases = 0;
    ~Batch();
};

Batch::~Batch() {
    for (auto& kv : graphs) {
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
        if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
    }
    for (void* p : {static_cast<void*>(d_args), static_cast<void*>(d_tiles), static_cast<void*>(d_phase_off)})
        if (p) (void)hipFree(p);
    barrier.release();
    for (auto st : branch) (void)hipStreamDestroy(st);
    for (auto ev : branch_ev) (void)hipEventDestroy(ev);
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    if (capture_stream) (void)hipStreamDestroy(capture_stream);
}

namespace {

template <class T>
void grow_upload(T** d, std::size_t* cap, const std::vector<T>& h) {
    if (h.size() > *cap) {
        if (*d) GMR_HIP(hipFree(*d));
        *d = nullptr;
        GMR_HIP(hipMalloc(d, std::max<std::size_t>(h.size(), 1) * sizeof(T)));
        *cap = h.size();
    }
    if (!h.empty()) GMR_HIP(hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
}

inline std::size_t mix(std::size_t h, std::size_t v) { return (h ^ v) * 1099511628211ULL; }

// Tile sizes: enough work per workgroup visit to amortise the tile fetch.
inline long eval_tile_rules(const dev::SweepArgs& a, int kind, int block) {
    if (kind == dev::EV_LDS && a.Bt != 1) kind = dev::EV_GROUP;
    switch (kind) {
        case dev::EV_THREAD: return std::max(1L, static_cast<long>(block) * 4 / (a.vec ? a.Bs / 4 : a.Bs));
        case dev::EV_WAVE: return static_cast<long>(block / 64) * 4;
        case dev::EV_LDS: return block;
        default: return static_cast<long>(block / a.G) * 4;
    }
}
inline long emit_tile(const dev::SweepArgs& a, int block) {
    return (a.Bt == 1 && a.B == 1) ? static_cast<long>(block) * 4 : static_cast<long>(block / a.G) * dev::kEmitChunk * 2;
}
inline long gather_tile(const dev::SweepArgs& a, int block) {
    if (a.flat) return static_cast<long>(block / 64) * 2;
    return (a.Bt == 1 && a.B == 1) ? std::max(1L, static_cast<long>(block / a.GR) * 2) : static_cast<long>(block / a.G) * 2;
}

template <class S>
void batch_run_t(Batch& bt, int semiring, const float* const* X, const int* B, void* const* Y, int variant,
                 const gmr_run_opts* opts, hipStream_t st) {
    using T = typename S::T;
    const int n = static_cast<int>(bt.eng.size());
    std::vector<gmr_run_opts> o(n);
    for (int g = 0; g < n; ++g) o[g] = opts ? opts[g] : gmr_run_opts{};
    for (int g = 0; g < n; ++g) {
        if (B[g] < 1 || !X[g] || !Y[g]) throw std::invalid_argument("batch: bad X, Y or B");
        if (o[g].emit == GMR_EMIT_GATHER && !bt.eng[g]->resident)
            throw std::invalid_argument("batch: gather emission needs the Resident layout");
        if (o[g].eval_kind < 0 || o[g].eval_kind >= dev::kEvalKinds) throw std::invalid_argument("unknown eval kind");
        if (o[g].level_kind)
            for (int k = 1; k <= bt.eng[g]->L; ++k)
                if (o[g].level_kind[k] < 0 || o[g].level_kind[k] >= dev::kEvalKinds)
                    throw std::invalid_argument("unknown eval kind");
    }
    if (variant == GMR_VARIANT_LEVELS) {
        for (int g = 0; g < n; ++g) {
            gmr_run_opts og = o[g];
            og.variant = GMR_VARIANT_LEVELS;
            bt.eng[g]->run(semiring, X[g], B[g], Y[g], og, st);
        }
        return;
    }
    const int block = o[0].block > 0 ? o[0].block : 256;
    std::vector<dev::SweepArgs> args(n);
    for (int g = 0; g < n; ++g) {
        bt.eng[g]->ensure_workspace(padded_width(B[g]), static_cast<int>(sizeof(T)));
        args[g] = make_args(*bt.eng[g], X[g], B[g], Y[g], static_cast<int>(sizeof(T)), 0, B[g],
                            o[g].emit == GMR_EMIT_GATHER, o[g]);
        // Tiles of the fused kernel are taken by whichever workgroup is free, so
        // the XCD-partitioned gather (which needs workgroup b on XCD b % 8) is
        // replaced there by the single chunked list.
        if (variant == GMR_VARIANT_PERSISTENT) args[g].xparts = 0;
        if (args[g].bits || args[g].lspan)
            throw std::invalid_argument("GMR_RUN_BITS and GMR_RUN_LDS_RANGES run only in the levels variant of a batch");
        if (variant == GMR_VARIANT_PERSISTENT && (o[g].flags & (GMR_RUN_FUSED_TAIL | GMR_RUN_PAIRED_LEVELS)))
            throw std::invalid_argument(
                "GMR_RUN_FUSED_TAIL and GMR_RUN_PAIRED_LEVELS run in the levels and graph variants of a batch");
    }
    if (variant == GMR_VARIANT_GRAPH) {
        std::size_t key = mix(static_cast<std::size_t>(semiring), static_cast<std::size_t>(block));
        for (int g = 0; g < n; ++g) {
            key = mix(key, reinterpret_cast<uintptr_t>(X[g]));
            key = mix(key, reinterpret_cast<uintptr_t>(Y[g]));
            key = mix(key, static_cast<std::size_t>(B[g]) * 4 + static_cast<std::size_t>(o[g].emit));
            key = mix(key, static_cast<std::size_t>(kinds_hash(*bt.eng[g], o[g])));
            key = mix(key, static_cast<std::size_t>(o[g].flags) * 16 + static_cast<std::size_t>(o[g].xcd_passes));
            key = mix(key, static_cast<std::size_t>(o[g].lds_values) * 16 + static_cast<std::size_t>(o[g].chunks));
            // an engine's workspace may have moved since the graph was captured
            key = mix(key, reinterpret_cast<uintptr_t>(args[g].Tv));
            key = mix(key, reinterpret_cast<uintptr_t>(args[g].W0));
        }
        auto it = bt.graphs.find(key);
        if (it == bt.graphs.end()) {
            if (!bt.capture_stream) GMR_HIP(hipStreamCreateWithFlags(&bt.capture_stream, hipStreamNonBlocking));
            if (!bt.fork_ev) GMR_HIP(hipEventCreateWithFlags(&bt.fork_ev, hipEventDisableTiming));
            while (static_cast<int>(bt.branch.size()) < n) {
                hipStream_t s2;
                hipEvent_t e2;
                GMR_HIP(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
                GMR_HIP(hipEventCreateWithFlags(&e2, hipEventDisableTiming));
                bt.branch.push_back(s2);
                bt.branch_ev.push_back(e2);
            }
            // One branch per grammar, forked from and joined into the capture stream.
            Engine::Graph gr;
            GMR_HIP(hipStreamBeginCapture(bt.capture_stream, hipStreamCaptureModeThreadLocal));
            try {
                GMR_HIP(hipEventRecord(bt.fork_ev, bt.capture_stream));
                for (int g = 0; g < n; ++g) {
                    GMR_HIP(hipStreamWaitEvent(bt.branch[g], bt.fork_ev, 0));
                    launch_levels<S>(*bt.eng[g], args[g], o[g], block, bt.branch[g]);
                    GMR_HIP(hipEventRecord(bt.branch_ev[g], bt.branch[g]));
                    GMR_HIP(hipStreamWaitEvent(bt.capture_stream, bt.branch_ev[g], 0));
                }
            } catch (...) {
                hipGraph_t dummy;
                (void)hipStreamEndCapture(bt.capture_stream, &dummy);
                throw;
            }
            GMR_HIP(hipStreamEndCapture(bt.capture_stream, &gr.graph));
            GMR_HIP(hipGraphInstantiate(&gr.exec, gr.graph, nullptr, nullptr, 0));
            it = bt.graphs.emplace(key, gr).first;
        }
        GMR_HIP(hipGraphLaunch(it->second.exec, st));
        return;
    }
    // Persistent: tile lists per phase.
    int Lmax = 0;
    for (int g = 0; g < n; ++g) Lmax = std::max(Lmax, bt.eng[g]->L);
    const int phases = Lmax + 2;
    std::size_t key = mix(static_cast<std::size_t>(block), static_cast<std::size_t>(semiring));
    for (int g = 0; g < n; ++g) {
        key = mix(key, reinterpret_cast<uintptr_t>(X[g]));
        key = mix(key, reinterpret_cast<uintptr_t>(Y[g]));
        key = mix(key, static_cast<std::size_t>(B[g]) * 4 + static_cast<std::size_t>(o[g].emit));
        key = mix(key, static_cast<std::size_t>(kinds_hash(*bt.eng[g], o[g])));
        key = mix(key, static_cast<std::size_t>(o[g].flags) * 16 + static_cast<std::size_t>(o[g].xcd_passes));
        key = mix(key, reinterpret_cast<uintptr_t>(args[g].Tv));
        key = mix(key, reinterpret_cast<uintptr_t>(args[g].W0));
    }
    if (key != bt.tiles_key) {
        std::vector<std::vector<dev::Tile>> per(phases);
        for (int g = 0; g < n; ++g) {
            const Engine& en = *bt.eng[g];
            const dev::SweepArgs& a = args[g];
            const long ni = dev::init_items(a);
            for (long b0 = 0; b0 < ni; b0 += 8192) per[0].push_back({g, dev::TT_INIT, 0, 0, b0, std::min(ni, b0 + 8192)});
            for (int p = 1; p <= en.L + 1; ++p) {
                if (p <= en.L) {
                    const int kind = host_kind(o[g], p);
                    const long lo = en.level_off[p - 1], hi = en.level_off[p];
                    const long step = eval_tile_rules(a, kind, block);
                    for (long r = lo; r < hi; r += step) per[p].push_back({g, dev::TT_EVAL, p, kind, r, std::min(hi, r + step)});
                }
                if (!a.gather) {
                    const long ne = p == 1 ? en.Et : en.emit_off[p - 1] - en.emit_off[p - 2];
                    const long step = emit_tile(a, block);
                    for (long e = 0; e < ne; e += step) per[p].push_back({g, dev::TT_EMIT, p - 1, 0, e, std::min(ne, e + step)});
                } else if (p == en.L + 1) {
                    const long step = gather_tile(a, block);
                    const long ng = dev::gather_items(a);
                    for (long r = 0; r < ng; r += step) per[p].push_back({g, dev::TT_GATHER, 0, 0, r, std::min(ng, r + step)});
                }
            }
        }
        std::vector<dev::Tile> tiles;
        std::vector<long> off(phases + 1, 0);
        for (int p = 0; p < phases; ++p) {
            off[p] = static_cast<long>(tiles.size());
            tiles.insert(tiles.end(), per[p].begin(), per[p].end());
        }
        off[phases] = static_cast<long>(tiles.size());
        GMR_HIP(hipDeviceSynchronize());  // the old tables may still be in use
        grow_upload(&bt.d_tiles, &bt.cap_tiles, tiles);
        grow_upload(&bt.d_phase_off, &bt.cap_phase, off);
        bt.tiles_key = key;
        bt.last_tiles = static_cast<long>(tiles.size());
        bt.last_phases = phases;
    }
    GMR_HIP(hipDeviceSynchronize());  // d_args may still be read by a queued launch
    grow_upload(&bt.d_args, &bt.cap_args, args);
    int occ = 0;
    GMR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, dev::k_multi<S>, block, 0));
    const int per_cu = persistent_per_cu(occ, block);
    int grid = per_cu * bt.num_cus;
    if (o[0].grid_cap > 0) grid = std::min(grid, o[0].grid_cap);
    hipLaunchKernelGGL(dev::k_multi<S>, dim3(grid), dim3(block), 0, st, bt.d_args, bt.d_tiles, bt.d_phase_off,
                       bt.last_phases, bt.barrier.ref());
    GMR_HIP(hipGetLastError());
    bt.barrier.advance(static_cast<unsigned long long>(grid) * static_cast<unsigned long long>(bt.last_phases - 1));
}

}  // namespace
}  // namespace gmr

// ---------------------------------------------------------------- C ABI
struct gmr_engine {
    gmr::Engine e;
};

namespace {

int dev_fail(int code, const char* msg) {
    gmr::g_dev_err = msg;
    gmr_host_set_error(msg);
    return code;
}

template <class F>
int dev_guarded(F&& f) {
    try {
        f();
        return GMR_OK;
    } catch (const gmr::DeviceError& e) {
        return dev_fail(GMR_ERR_DEVICE, e.what());
    } catch (const std::bad_alloc&) {
        return dev_fail(GMR_ERR_NOMEM, "host allocation failed");
    } catch (const std::invalid_argument& e) {
        return dev_fail(GMR_ERR_INVALID, e.what());
    } catch (const std::exception& e) {
        return dev_fail(GMR_ERR_INTERNAL, e.what());
    }
}

}  // namespace

const char* gmr_engine_last_error_internal() { return gmr::g_dev_err.c_str(); }

extern "C" {

int gmr_engine_create(int rows, int cols, int U, int num_rules, const LayeredRule* rules,
                      const int* val_idx, const int* col_idx, const float* V, int nvals,
                      const GPUSchedule* sched, int layout, gmr_engine** out) {
    if (!out || !sched) return dev_fail(GMR_ERR_INVALID, "NULL argument");
    gmr_engine* h = new (std::nothrow) gmr_engine;
    if (!h) return GMR_ERR_NOMEM;
    int rc = dev_guarded([&] { h->e.upload(rows, cols, U, num_rules, rules, val_idx, col_idx, V, nvals, *sched, layout); });
    if (rc != GMR_OK) {
        delete h;
        return rc;
    }
    *out = h;
    return GMR_OK;
}

int gmr_engine_from_schedule(const gmr_schedule* s, const float* V, int nvals, gmr_engine** out) {
    const BuiltSchedule* b = gmr_schedule_internal(s);
    if (!b || !out) return dev_fail(GMR_ERR_INVALID, "NULL argument");
    std::vector<int> vi(b->alpha > 0 ? b->alpha - 1 : 0), ci(vi.size());
    for (int q = 1; q < b->alpha; ++q) {
        int t = b->term_orig[q];
        vi[q - 1] = (t - 1) / b->cols;
        ci[q - 1] = (t - 1) % b->cols;
    }
    GPUSchedule view = const_cast<BuiltSchedule*>(b)->gpu_schedule();
    return gmr_engine_create(b->rows, b->cols, b->alpha, static_cast<int>(b->num_layered),
                             b->flattened_rules.data(), vi.data(), ci.data(), V, nvals, &view,
                             static_cast<int>(b->layout), out);
}

int gmr_engine_info(const gmr_engine* h, long long info[16]) {
    if (!h || !info) return GMR_ERR_INVALID;
    const gmr::Engine& e = h->e;
    long long v[16] = {e.rows, e.cols, e.U, e.NL, e.L, e.max_width, e.E, e.Et, e.layout,
                       e.has_runlen ? 1 : 0, static_cast<long long>(e.struct_bytes), e.nvals,
                       e.n_gseg, 0, 0, 0};
    std::memcpy(info, v, sizeof(v));
    return GMR_OK;
}

namespace {
const char* const kBarrierMsg =
    "a persistent launch's grid barrier timed out (its workgroups were not all resident): the Y it wrote is "
    "invalid; this call did not run";
}

int gmr_engine_run(gmr_engine* h, int semiring, const void* d_X, int B, void* d_Y,
                   const gmr_run_opts* opts, void* stream) {
    if (!h) return dev_fail(GMR_ERR_INVALID, "engine is NULL");
    // A failure of an earlier launch (seen without synchronising) is reported
    // once, by the next call, which then does not run.
    if (h->e.barrier.take_failure()) return dev_fail(GMR_ERR_BARRIER, kBarrierMsg);
    gmr_run_opts o{};
    if (opts) o = *opts;
    int rc = dev_guarded([&] {
        h->e.run(semiring, static_cast<const float*>(d_X), B, d_Y, o, static_cast<hipStream_t>(stream));
    });
    // A timed run has synchronised its stream: its own failure is known now.
    if (rc == GMR_OK && (o.flags & GMR_RUN_TIMED) && h->e.barrier.take_failure())
        return dev_fail(GMR_ERR_BARRIER, "this persistent launch's grid barrier timed out: the Y it wrote is invalid");
    return rc;
}

// ---- no-allocation runs ----
// The workspace of gmr_engine_run_noalloc: a file-scope device array. Each
// GPU the library runs on holds its own copy (GMR_NOALLOC_SCRATCH_BYTES) until
// the process ends. In host code `scratch` names a host-side stand-in, not a
// device address: its address on a device comes from hipGetSymbolAddress with
// that device current, looked up once per device and kept in g_scratch_ptr,
// and only that pointer is handed to kernels.
static __device__ __align__(256) unsigned char scratch[GMR_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

void* scratch_on_current_device() {
    int dev = 0;
    GMR_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw std::invalid_argument("device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        GMR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw gmr::DeviceError("hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}
}  // namespace

int gmr_engine_run_noalloc(gmr_engine* h, int semiring, const void* d_X, int B, void* d_Y,
                           const gmr_run_opts* opts, void* stream) {
    if (!h) return dev_fail(GMR_ERR_INVALID, "engine is NULL");
    if (h->e.barrier.take_failure()) return dev_fail(GMR_ERR_BARRIER, kBarrierMsg);
    gmr_run_opts o{};
    if (opts) o = *opts;
    int rc = dev_guarded([&] {
        void* ws = scratch_on_current_device();
        h->e.run_noalloc(semiring, static_cast<const float*>(d_X), B, d_Y, o, static_cast<hipStream_t>(stream), ws,
                         GMR_NOALLOC_SCRATCH_BYTES);
    });
    if (rc == GMR_OK && (o.flags & GMR_RUN_TIMED) && h->e.barrier.take_failure())
        return dev_fail(GMR_ERR_BARRIER, "this persistent launch's grid barrier timed out: the Y it wrote is invalid");
    return rc;
}

int gmr_engine_set_barrier_timeout(gmr_engine* h, long long microseconds) {
    if (!h) return dev_fail(GMR_ERR_INVALID, "engine is NULL");
    if (microseconds < 0) return dev_fail(GMR_ERR_INVALID, "timeout must be >= 0");
    h->e.barrier.timeout_us = microseconds;
    return GMR_OK;
}

int gmr_engine_time_levels(gmr_engine* h, int semiring, const void* d_X, int B, void* d_Y,
                           const gmr_run_opts* opts, int reps, float* level_ms, void* stream) {
    if (!h) return dev_fail(GMR_ERR_INVALID, "engine is NULL");
    gmr_run_opts o{};
    if (opts) o = *opts;
    return dev_guarded([&] {
        h->e.time_levels(semiring, static_cast<const float*>(d_X), B, d_Y, o, reps, level_ms,
                         static_cast<hipStream_t>(stream));
    });
}

int gmr_engine_run_left(gmr_engine* h, int semiring, const void* d_X, int B, void* d_Y, const gmr_run_opts* opts,
                        void* stream) {
    if (!h) return dev_fail(GMR_ERR_INVALID, "engine is NULL");
    gmr_run_opts o{};
    if (opts) o = *opts;
    return dev_guarded([&] {
        h->e.run_left(semiring, static_cast<const float*>(d_X), B, d_Y, o, static_cast<hipStream_t>(stream));
    });
}

int gmr_engine_traverse(gmr_engine* h, int semiring, void* d_x, int max_iters, int* iters_out,
                        const gmr_run_opts* opts, void* stream) {
    if (!h) return dev_fail(GMR_ERR_INVALID, "engine is NULL");
    gmr_run_opts o{};
    if (opts) o = *opts;
    return dev_guarded([&] {
        const int n = h->e.traverse(semiring, static_cast<float*>(d_x), max_iters, o, static_cast<hipStream_t>(stream));
        if (iters_out) *iters_out = n;
    });
}

int gmr_engine_set_gather_order(gmr_engine* h, int order) {
    if (!h) return dev_fail(GMR_ERR_INVALID, "engine is NULL");
    if (order < 0 || order > 1) return dev_fail(GMR_ERR_INVALID, "unknown gather order");
    return dev_guarded([&] { h->e.set_gather_order(order); });
}

int gmr_engine_check(gmr_engine* h, int* barrier_error) {
    if (!h || !barrier_error) return dev_fail(GMR_ERR_INVALID, "NULL argument");
    return dev_guarded([&] { *barrier_error = h->e.check_barrier(); });
}

int gmr_engine_debug(gmr_engine* h, long long out[8]) {
    if (!h || !out) return dev_fail(GMR_ERR_INVALID, "NULL argument");
    return dev_guarded([&] {
        gmr::dev::GridBarrier b{};
        GMR_HIP(hipDeviceSynchronize());
        GMR_HIP(hipMemcpy(&b, h->e.barrier.d, sizeof(b), hipMemcpyDeviceToHost));
        const long long failed = h->e.barrier.h_fail ? *h->e.barrier.h_fail : 0;
        long long v[8] = {static_cast<long long>(b.arrived), failed, static_cast<long long>(h->e.barrier.base),
                          h->e.wall_khz, h->e.num_cus, h->e.last_grid, h->e.last_per_cu,
                          static_cast<long long>(h->e.graphs.size())};
        std::memcpy(out, v, sizeof(v));
    });
}

int gmr_engine_last_kernel_ms(const gmr_engine* h, double* ms) {
    if (!h || !ms) return GMR_ERR_INVALID;
    *ms = h->e.last_ms;
    return GMR_OK;
}

void gmr_engine_destroy(gmr_engine* h) { delete h; }

struct gmr_batch {
    gmr::Batch b;
};

int gmr_batch_create(gmr_engine* const* engines, int n, gmr_batch** out) {
    if (!engines || n < 1 || !out) return dev_fail(GMR_ERR_INVALID, "batch: need engines");
    for (int i = 0; i < n; ++i) {
        if (!engines[i]) return dev_fail(GMR_ERR_INVALID, "batch: NULL engine");
        for (int j = 0; j < i; ++j)
            if (engines[j] == engines[i]) return dev_fail(GMR_ERR_INVALID, "batch: an engine appears twice");
    }
    gmr_batch* h = new (std::nothrow) gmr_batch;
    if (!h) return GMR_ERR_NOMEM;
    int rc = dev_guarded([&] {
        for (int i = 0; i < n; ++i) h->b.eng.push_back(&engines[i]->e);
        int dev = 0;
        GMR_HIP(hipGetDevice(&dev));
        GMR_HIP(hipDeviceGetAttribute(&h->b.num_cus, hipDeviceAttributeMultiprocessorCount, dev));
        GMR_HIP(hipDeviceGetAttribute(&h->b.wall_khz, hipDeviceAttributeWallClockRate, dev));
        if (h->b.wall_khz <= 0) h->b.wall_khz = 100000;
        h->b.barrier.init(h->b.wall_khz);
    });
    if (rc != GMR_OK) {
        delete h;
        return rc;
    }
    *out = h;
    return GMR_OK;
}

int gmr_batch_run(gmr_batch* h, int semiring, const void* const* d_X, const int* B, void* const* d_Y, int variant,
                  const gmr_run_opts* opts, void* stream) {
    if (!h || !d_X || !B || !d_Y) return dev_fail(GMR_ERR_INVALID, "NULL argument");
    if (variant < 0 || variant > 2) return dev_fail(GMR_ERR_INVALID, "unknown variant");
    if (h->b.barrier.take_failure()) return dev_fail(GMR_ERR_BARRIER, kBarrierMsg);
    return dev_guarded([&] {
        const float* const* X = reinterpret_cast<const float* const*>(d_X);
        auto st = static_cast<hipStream_t>(stream);
        switch (semiring) {
            case GMR_PLUS_TIMES: gmr::batch_run_t<gmr::dev::PlusTimesOp>(h->b, semiring, X, B, d_Y, variant, opts, st); break;
            case GMR_BOOLEAN: gmr::batch_run_t<gmr::dev::BooleanOp>(h->b, semiring, X, B, d_Y, variant, opts, st); break;
            case GMR_TROPICAL: gmr::batch_run_t<gmr::dev::TropicalOp>(h->b, semiring, X, B, d_Y, variant, opts, st); break;
            default: throw std::invalid_argument("unknown semiring");
        }
    });
}

int gmr_batch_check(gmr_batch* h, int* barrier_error) {
    if (!h || !barrier_error) return dev_fail(GMR_ERR_INVALID, "NULL argument");
    return dev_guarded([&] { *barrier_error = h->b.barrier.check(); });
}

int gmr_batch_set_barrier_timeout(gmr_batch* h, long long microseconds) {
    if (!h) return dev_fail(GMR_ERR_INVALID, "batch is NULL");
    if (microseconds < 0) return dev_fail(GMR_ERR_INVALID, "timeout must be >= 0");
    h->b.barrier.timeout_us = microseconds;
    return GMR_OK;
}

void gmr_batch_destroy(gmr_batch* h) { delete h; }

int gmr_device_sync(void) {
    return hipDeviceSynchronize() == hipSuccess ? GMR_OK : GMR_ERR_DEVICE;
}

}  // extern "C"
