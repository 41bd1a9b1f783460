// Seeded synthetic matrices (haplotype mosaic, genotype mosaic, R-MAT) in CSRV
// form. Every row (or block of edges) draws from its own counter-seeded stream,
// so the output is the same for any thread count.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <parallel/algorithm>
#include <stdexcept>

#include "host.hpp"
#include "rng.hpp"

namespace gmr {

namespace {

// Stream ids (high bits) keep the generators' streams apart.
constexpr uint64_t kFounderStream = 1ULL << 60;
constexpr uint64_t kRowStream = 2ULL << 60;
constexpr uint64_t kEdgeStream = 3ULL << 60;

// Concatenate per-row column lists into a CSRV string (value index 0 only).
Csrv pack_rows(int rows, int cols, const std::vector<std::vector<int>>& row_cols) {
    Csrv m;
    m.rows = rows;
    m.cols = cols;
    m.V = {1.0};
    std::vector<long long> off(rows + 1, 0);
    for (int r = 0; r < rows; ++r) off[r + 1] = off[r] + static_cast<long long>(row_cols[r].size()) + 1;
    m.seq.resize(static_cast<std::size_t>(off[rows]));
#pragma omp parallel for schedule(static)
    for (int r = 0; r < rows; ++r) {
        int* dst = m.seq.data() + off[r];
        for (int c : row_cols[r]) *dst++ = 1 + c;
        *dst = 0;
    }
    return m;
}

}  // namespace

int Csrv::alpha() const {
    int mx = 0;
    for (int s : seq) mx = std::max(mx, s);
    return mx + 1;
}

Csrv gen_haplotype(uint64_t seed, int rows, int cols, int founders, double density,
                   double seg_mean, double flip) {
    if (rows < 0 || cols < 1 || founders < 1) throw std::invalid_argument("gen_haplotype: bad shape");
    std::vector<std::vector<uint8_t>> F(founders, std::vector<uint8_t>(cols));
    for (int f = 0; f < founders; ++f) {
        Rng g(seed, kFounderStream + f);
        for (int c = 0; c < cols; ++c) F[f][c] = g.uniform() < density;
    }
    std::vector<std::vector<int>> out(rows);
#pragma omp parallel
    {
        std::vector<uint8_t> bits(cols);
#pragma omp for schedule(dynamic, 64)
        for (int r = 0; r < rows; ++r) {
            Rng g(seed, kRowStream + r);
            for (long long c = 0; c < cols;) {
                long long len = g.geometric(seg_mean);
                int f = static_cast<int>(g.below(founders));
                long long end = std::min<long long>(cols, c + len);
                for (long long j = c; j < end; ++j) bits[j] = F[f][j];
                c = end;
            }
            if (flip > 0.0) {
                // Flip positions by geometric skips: each cell flips w.p. `flip`.
                for (long long c = g.geometric(1.0 / flip) - 1; c < cols; c += g.geometric(1.0 / flip))
                    bits[c] ^= 1;
            }
            auto& v = out[r];
            for (int c = 0; c < cols; ++c)
                if (bits[c]) v.push_back(c);
        }
    }
    return pack_rows(rows, cols, out);
}

Csrv gen_genotype(uint64_t seed, int rows, int cols, int founders, int classes,
                  double seg_mean) {
    if (rows < 0 || cols < 2 || founders < classes || classes < 1)
        throw std::invalid_argument("gen_genotype: bad shape");
    // Class c has density 10^(lo + (hi-lo)(c+0.5)/classes), spanning the
    // allele-count range [1, cols-1] in log space.
    const double lo = std::log10(1.0 / cols), hi = std::log10((cols - 1.0) / cols);
    std::vector<double> dens(classes);
    for (int c = 0; c < classes; ++c) dens[c] = std::pow(10.0, lo + (hi - lo) * (c + 0.5) / classes);
    const int per_class = founders / classes;
    std::vector<std::vector<uint8_t>> F(founders, std::vector<uint8_t>(cols));
    for (int f = 0; f < founders; ++f) {
        Rng g(seed, kFounderStream + f);
        double d = dens[std::min(classes - 1, f / per_class)];
        for (int c = 0; c < cols; ++c) F[f][c] = g.uniform() < d;
    }
    // P(k) ~ 1/k on [1, cols-1]: inverse CDF over the harmonic partial sums.
    const int kmax = cols - 1;
    std::vector<double> cdf(kmax + 1, 0.0);
    for (int k = 1; k <= kmax; ++k) cdf[k] = cdf[k - 1] + 1.0 / k;
    for (int k = 1; k <= kmax; ++k) cdf[k] /= cdf[kmax];

    std::vector<std::vector<int>> out(rows);
#pragma omp parallel
    {
        std::vector<uint8_t> bits(cols);
        std::vector<int> ones, zeros;
#pragma omp for schedule(dynamic, 256)
        for (int r = 0; r < rows; ++r) {
            Rng g(seed, kRowStream + r);
            double u = g.uniform();
            int k = static_cast<int>(std::upper_bound(cdf.begin() + 1, cdf.end(), u) - cdf.begin());
            k = std::min(std::max(k, 1), kmax);
            double lk = std::log10(static_cast<double>(k) / cols);
            int cls = static_cast<int>(std::floor((lk - lo) / (hi - lo) * classes));
            cls = std::min(std::max(cls, 0), classes - 1);
            for (long long c = 0; c < cols;) {
                long long len = g.geometric(seg_mean);
                int f = cls * per_class + static_cast<int>(g.below(per_class));
                long long end = std::min<long long>(cols, c + len);
                for (long long j = c; j < end; ++j) bits[j] = F[f][j];
                c = end;
            }
            ones.clear();
            zeros.clear();
            for (int c = 0; c < cols; ++c) (bits[c] ? ones : zeros).push_back(c);
            int p = static_cast<int>(ones.size());
            // Mutation flips until the row holds exactly k ones (partial
            // Fisher-Yates over the ones or the zeros).
            if (p > k) {
                for (int i = 0; i < p - k; ++i) {
                    int j = i + static_cast<int>(g.below(p - i));
                    std::swap(ones[i], ones[j]);
                    bits[ones[i]] = 0;
                }
            } else if (p < k) {
                int z = static_cast<int>(zeros.size());
                for (int i = 0; i < k - p; ++i) {
                    int j = i + static_cast<int>(g.below(z - i));
                    std::swap(zeros[i], zeros[j]);
                    bits[zeros[i]] = 1;
                }
            }
            auto& v = out[r];
            v.reserve(k);
            for (int c = 0; c < cols; ++c)
                if (bits[c]) v.push_back(c);
        }
    }
    return pack_rows(rows, cols, out);
}

Csrv gen_rmat(uint64_t seed, int scale, long long edges, double a, double b, double c,
              int nweights) {
    if (scale < 1 || scale > 30 || edges < 0 || nweights < 1)
        throw std::invalid_argument("gen_rmat: bad parameters");
    const long long n = 1LL << scale;
    if (static_cast<long long>(nweights) * n + 1 > 0x7fffffffLL)
        throw std::invalid_argument("gen_rmat: terminal alphabet exceeds int32");
    const double ab = a + b, abc = a + b + c;
    std::vector<uint64_t> key(static_cast<std::size_t>(edges));
    const long long block = 1 << 20;
    const long long nblocks = (edges + block - 1) / block;
#pragma omp parallel for schedule(dynamic, 1)
    for (long long blk = 0; blk < nblocks; ++blk) {
        Rng g(seed, kEdgeStream + static_cast<uint64_t>(blk));
        long long e1 = std::min(edges, (blk + 1) * block);
        for (long long e = blk * block; e < e1; ++e) {
            uint64_t u = 0, v = 0;
            for (int bit = scale - 1; bit >= 0; --bit) {
                double p = g.uniform();
                if (p >= abc) { u |= 1ULL << bit; v |= 1ULL << bit; }
                else if (p >= ab) u |= 1ULL << bit;
                else if (p >= a) v |= 1ULL << bit;
            }
            key[e] = (u << 32) | v;
        }
    }
    __gnu_parallel::sort(key.begin(), key.end());
    key.erase(std::unique(key.begin(), key.end()), key.end());

    Csrv m;
    m.rows = static_cast<int>(n);
    m.cols = static_cast<int>(n);
    m.V.resize(nweights);
    for (int w = 0; w < nweights; ++w) m.V[w] = w + 1.0;
    // Row starts: edges are sorted by (u, v).
    std::vector<long long> deg(n + 1, 0);
    for (uint64_t k : key) ++deg[(k >> 32) + 1];
    for (long long r = 0; r < n; ++r) deg[r + 1] += deg[r];
    m.seq.resize(static_cast<std::size_t>(key.size() + n));
#pragma omp parallel for schedule(static)
    for (long long r = 0; r < n; ++r) {
        int* dst = m.seq.data() + deg[r] + r;
        for (long long e = deg[r]; e < deg[r + 1]; ++e) {
            uint64_t v = key[e] & 0xffffffffULL;
            long long w = 0;
            if (nweights > 1) w = static_cast<long long>(mix64(seed ^ (key[e] * 0x9E3779B97F4A7C15ULL)) % nweights);
            *dst++ = static_cast<int>(1 + w * n + static_cast<long long>(v));
        }
        *dst = 0;
    }
    return m;
}

Csrv csrv_from_dense(const double* M, int rows, int cols) {
    Csrv m;
    m.rows = rows;
    m.cols = cols;
    const long long n = static_cast<long long>(rows) * cols;
    std::vector<double> vals;
    for (long long i = 0; i < n; ++i)
        if (M[i] != 0.0) vals.push_back(M[i]);
    std::sort(vals.begin(), vals.end());
    vals.erase(std::unique(vals.begin(), vals.end()), vals.end());
    if (static_cast<long long>(vals.size()) * cols + 1 > 0x7fffffffLL)
        throw std::invalid_argument("csrv_from_dense: terminal alphabet exceeds int32");
    m.V = vals;
    for (int r = 0; r < rows; ++r) {
        for (int c = 0; c < cols; ++c) {
            double x = M[static_cast<long long>(r) * cols + c];
            if (x == 0.0) continue;
            int v = static_cast<int>(std::lower_bound(vals.begin(), vals.end(), x) - vals.begin());
            m.seq.push_back(1 + v * cols + c);
        }
        m.seq.push_back(0);
    }
    return m;
}

}  // namespace gmr
