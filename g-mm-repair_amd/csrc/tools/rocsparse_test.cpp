// rocsparse_test -- the vendor CSR baseline on MI355X, with the command line of
// the reference's cuSPARSE driver (gpu-engine/cusparse_test.cu:73-510).
//
// Usage: rocsparse_test <base> <rows> <cols> [iters=100] [mode=repair] [B=1]
// Reads the grammar files <base>.val/.vc.R/.vc.C, expands them back into the
// matrix (CSR, FP32 values), checks rocSPARSE's SpMV against a CPU CSR SpMV in
// double, times `iters` SpMVs, reads the energy of a >= 1.5 s SpMV loop from
// the amdsmi counter, and for B > 1 times SpMM under each CSR algorithm.
// SEMIRING and CROSSCHECK behave as in the reference driver: with
// CROSSCHECK=<prefix> the engine's dumped <prefix>.<semiring>.x is the input
// and its <prefix>.<semiring>.y is checked (Boolean: nonzero pattern of the
// (+,x) product; Tropical: a CPU min-plus CSR product, since no vendor min-plus
// kernel exists), then the driver stops.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "gmr_host.h"
#include "gmr_vendor.h"

namespace {

std::string semiring_name() {
    const char* e = std::getenv("SEMIRING");
    std::string s = e ? e : "plustimes";
    if (s == "bool" || s == "boolean") return "boolean";
    if (s == "tropical" || s == "minplus") return "tropical";
    return "plustimes";
}

bool read_floats(const std::string& path, std::vector<float>& v, std::size_t n) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    v.resize(n);
    f.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(n * sizeof(float)));
    return static_cast<std::size_t>(f.gcount()) == n * sizeof(float);
}

bool read_doubles(const std::string& path, std::vector<double>& v, std::size_t n) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    v.resize(n);
    f.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(n * sizeof(double)));
    return static_cast<std::size_t>(f.gcount()) == n * sizeof(double);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 4) {
        std::cerr << "Usage: " << argv[0] << " <matrix_base_name> <rows> <cols> [iters] [mode] [B]" << std::endl;
        return 1;
    }
    const std::string base = argv[1];
    const int rows = std::atoi(argv[2]);
    const int cols = std::atoi(argv[3]);
    const int iters = argc >= 5 ? std::atoi(argv[4]) : 100;
    const std::string mode = argc >= 6 ? argv[5] : "repair";
    const int B = argc >= 7 ? std::atoi(argv[6]) : 1;
    if (mode == "recompress" || mode == "partII" || mode == "II") {
        std::cerr << "Error: Recompression mode (Part II) is no longer supported/needed." << std::endl;
        return 1;
    }
    if (rows < 1 || cols < 1 || iters < 1 || B < 1) {
        std::cerr << "Error: rows, cols, iters and B must be positive." << std::endl;
        return 1;
    }

    // 1-3. Grammar files -> the matrix in CSR form (values as float).
    gmr_grammar* g = nullptr;
    if (gmr_grammar_read(base.c_str(), &g) != GMR_OK) {
        std::cerr << "Error: " << gmr_last_error() << std::endl;
        return 1;
    }
    int alpha = 0, nrules = 0, nvals = 0;
    long long clen = 0, len = 0;
    gmr_grammar_info(g, &alpha, &nrules, &clen, &nvals);
    std::vector<double> Vd(static_cast<std::size_t>(nvals));
    gmr_grammar_copy(g, nullptr, nullptr, nullptr, nullptr, nullptr, Vd.data());
    if (gmr_grammar_expand(g, nullptr, 0, &len) != GMR_OK) {
        std::cerr << "Error: " << gmr_last_error() << std::endl;
        return 1;
    }
    std::vector<int> seq(static_cast<std::size_t>(len));
    gmr_grammar_expand(g, seq.data(), len, &len);
    gmr_grammar_free(g);
    std::vector<int> row_ptr{0}, col_idx;
    std::vector<float> vals;
    col_idx.reserve(static_cast<std::size_t>(len));
    vals.reserve(static_cast<std::size_t>(len));
    for (int s : seq) {
        if (s == 0) {
            row_ptr.push_back(static_cast<int>(col_idx.size()));
            continue;
        }
        const long long t = s - 1;
        const long long v = t / cols, c = t % cols;
        if (v >= nvals) {
            std::cerr << "Error: terminal " << s << " decodes to value " << v << " of " << nvals << std::endl;
            return 1;
        }
        col_idx.push_back(static_cast<int>(c));
        vals.push_back(static_cast<float>(Vd[static_cast<std::size_t>(v)]));
    }
    if (row_ptr.size() > static_cast<std::size_t>(rows) + 1) {
        std::cerr << "Error: the grammar holds more than " << rows << " rows." << std::endl;
        return 1;
    }
    while (row_ptr.size() <= static_cast<std::size_t>(rows)) row_ptr.push_back(static_cast<int>(col_idx.size()));
    const long long nnz = static_cast<long long>(col_idx.size());
    std::cout << "Reconstructed CSR Matrix Stats:" << std::endl;
    std::cout << "Rows: " << rows << ", Cols: " << cols << ", NNZ: " << nnz << std::endl;

    // 4. Input vector: the engine's dumped x under CROSSCHECK, else seeded U[0,1).
    const char* cc = std::getenv("CROSSCHECK");
    const std::string sr = semiring_name();
    std::vector<float> x(static_cast<std::size_t>(cols));
    if (cc && cc[0]) {
        const std::string xf = std::string(cc) + "." + sr + ".x";
        if (!read_floats(xf, x, x.size())) {
            std::cerr << "CROSSCHECK: cannot open " << xf << std::endl;
            return 1;
        }
        std::cout << "CROSSCHECK: loaded engine x from " << xf << std::endl;
    } else {
        unsigned long long st = 20261204ull;
        for (auto& xi : x) {
            st = st * 6364136223846793005ull + 1442695040888963407ull;
            xi = static_cast<float>((st >> 40) * (1.0 / 16777216.0));
        }
    }
    auto load_engine_y = [&](std::vector<double>& ey) {
        const std::string yf = std::string(cc) + "." + sr + ".y";
        if (!read_doubles(yf, ey, static_cast<std::size_t>(rows))) {
            std::cerr << "CROSSCHECK: cannot open " << yf << std::endl;
            return false;
        }
        return true;
    };

    // Tropical: no vendor min-plus product; check the engine against a CPU
    // min-plus over the same CSR (1e9 is +infinity, as in the engine).
    if (cc && cc[0] && sr == "tropical") {
        const double INF = 1e9;
        std::vector<double> ey;
        if (!load_engine_y(ey)) return 1;
        double maxd = 0.0;
        long mism = 0;
        for (int r = 0; r < rows; ++r) {
            double best = INF;
            for (int i = row_ptr[r]; i < row_ptr[r + 1]; ++i) {
                const double mv = vals[i], xv = x[col_idx[i]];
                const double cand = (mv >= INF || xv >= INF) ? INF : mv + xv;
                if (cand < best) best = cand;
            }
            const double d = std::fabs(best - ey[r]);
            if (d > maxd) maxd = d;
            if (d > 0.5) ++mism;
        }
        std::cout << "CROSSCHECK cpu_csr_minplus vs engine: max_abs_diff=" << maxd << ", mismatches=" << mism
                  << (mism == 0 ? "  SUCCESS" : "  FAILURE") << std::endl;
        return 0;
    }

    std::cout << "Running CPU reference SpMV..." << std::endl;
    std::vector<float> y_ref(static_cast<std::size_t>(rows));
    for (int r = 0; r < rows; ++r) {
        double sum = 0.0;
        for (int i = row_ptr[r]; i < row_ptr[r + 1]; ++i) sum += static_cast<double>(vals[i]) * x[col_idx[i]];
        y_ref[r] = static_cast<float>(sum);
    }

    // 5-6. rocSPARSE.
    gmr_csr* m = nullptr;
    if (gmr_csr_create(rows, cols, nnz, row_ptr.data(), col_idx.data(), vals.data(), &m) != GMR_OK) {
        std::cerr << "Error: " << gmr_last_error() << std::endl;
        return 1;
    }
    std::vector<float> y(static_cast<std::size_t>(rows));
    double ms = 0.0;
    if (gmr_csr_bench(m, x.data(), 1, y.data(), GMR_CSR_ALG_DEFAULT, 1, &ms) != GMR_OK) {
        std::cerr << "Error: " << gmr_last_error() << std::endl;
        return 1;
    }
    float max_abs = 0.0f, max_rel = 0.0f;
    for (int r = 0; r < rows; ++r) {
        const float d = std::fabs(y[r] - y_ref[r]);
        max_abs = std::max(max_abs, d);
        max_rel = std::max(max_rel, d / (1.0f + std::fabs(y_ref[r])));
    }
    std::cout << "Verification: max absolute diff = " << max_abs << ", max relative diff = " << max_rel << std::endl;
    if (max_rel > 1e-4f) std::cerr << "WARNING: Verification mismatch!" << std::endl;
    else std::cout << "SUCCESS: Verification passed!" << std::endl;

    if (cc && cc[0]) {
        std::vector<double> ey;
        if (load_engine_y(ey)) {
            if (sr == "boolean") {
                long mc = 0, mg = 0;
                for (int r = 0; r < rows; ++r) {
                    const int eb = ey[r] != 0.0 ? 1 : 0;
                    if ((y_ref[r] != 0.0f ? 1 : 0) != eb) ++mc;
                    if ((y[r] != 0.0f ? 1 : 0) != eb) ++mg;
                }
                std::cout << "CROSSCHECK cpu_csr_spmv(bool) vs engine: mismatches=" << mc
                          << (mc == 0 ? "  SUCCESS" : "  FAILURE") << std::endl;
                std::cout << "CROSSCHECK rocsparse(bool) vs engine: mismatches=" << mg
                          << (mg == 0 ? "  SUCCESS" : "  FAILURE") << std::endl;
            } else {
                double rc = 0.0, rg = 0.0;
                for (int r = 0; r < rows; ++r) {
                    const double a = std::fabs(ey[r]);
                    rc = std::max(rc, std::fabs(y_ref[r] - ey[r]) / (1.0 + a));
                    rg = std::max(rg, std::fabs(y[r] - ey[r]) / (1.0 + a));
                }
                std::cout << "CROSSCHECK cpu_csr_spmv vs engine: max_rel_diff=" << rc
                          << (rc < 1e-4 ? "  SUCCESS" : "  FAILURE") << std::endl;
                std::cout << "CROSSCHECK rocsparse vs engine: max_rel_diff=" << rg
                          << (rg < 1e-4 ? "  SUCCESS" : "  FAILURE") << std::endl;
            }
        }
        gmr_csr_free(m);
        return 0;
    }

    std::cout << "Running " << iters << " iterations of rocSPARSE SpMV..." << std::endl;
    if (gmr_csr_bench(m, x.data(), 1, y.data(), GMR_CSR_ALG_DEFAULT, iters, &ms) != GMR_OK) {
        std::cerr << "Error: " << gmr_last_error() << std::endl;
        return 1;
    }
    std::cout << "rocSPARSE SpMV completed. Average time: " << ms << " ms/vector" << std::endl;
    double mj = -1.0, mw = -1.0;
    if (gmr_csr_energy(m, 1, GMR_CSR_ALG_DEFAULT, 1.5, &mj, &mw) != GMR_OK) {
        std::cerr << "energy counter unavailable: " << gmr_last_error() << std::endl;
        mj = mw = -1.0;
    }
    const long long analytic = gmr_csr_device_bytes(m) + 4LL * (rows + cols);
    std::cout << "MEM analytic bytes: " << analytic << std::endl;
    std::cout << "MEM peak bytes: " << analytic << std::endl;
    std::cout << "ENERGY mJ/vector: " << mj << std::endl;
    std::cout << "ENERGY avg power mW: " << mw << std::endl;

    if (B > 1) {
        std::vector<float> X(static_cast<std::size_t>(cols) * B), Y(static_cast<std::size_t>(rows) * B);
        for (int j = 0; j < cols; ++j)
            for (int b = 0; b < B; ++b) X[static_cast<std::size_t>(j) * B + b] = x[j];
        const struct {
            const char* name;
            int alg;
        } algs[] = {{"default", GMR_CSR_ALG_DEFAULT},
                    {"csr", GMR_CSR_ALG_ROW},
                    {"csr_row_split", GMR_CSR_ALG_ROW_SPLIT},
                    {"csr_merge", GMR_CSR_ALG_NNZ_SPLIT}};
        for (const auto& a : algs) {
            double mm = 0.0;
            if (gmr_csr_bench(m, X.data(), B, Y.data(), a.alg, iters, &mm) != GMR_OK) {
                std::cout << "rocSPARSE SpMM (B=" << B << ", " << a.name << "): UNSUPPORTED (" << gmr_last_error()
                          << ")" << std::endl;
                continue;
            }
            float rel = 0.0f;
            const int check = rows < 2000 ? rows : 2000;
            for (int r = 0; r < check; ++r)
                for (int b = 0; b < B; ++b)
                    rel = std::max(rel, std::fabs(Y[static_cast<std::size_t>(r) * B + b] - y_ref[r]) /
                                            (1.0f + std::fabs(y_ref[r])));
            std::cout << "rocSPARSE SpMM (B=" << B << ", " << a.name << "): " << mm << " ms/batch, " << mm / B
                      << " ms/vector  rel=" << rel << (rel < 1e-4f ? " SUCCESS" : " FAILURE") << std::endl;
        }
    }
    gmr_csr_free(m);
    return 0;
}
