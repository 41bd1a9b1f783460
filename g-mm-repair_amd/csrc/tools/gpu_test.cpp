// gpu_test -- command-line driver with the reference's interface
// (reference gpu-engine/gpu_engine_test.cu:277-470):
//
//   gpu_test <matrix_base_name> <rows> <cols> [iters] [mode] [B]
//   env SEMIRING = plustimes | boolean | bool | tropical | minplus
//       XVEC     = file of exactly `cols` float32 values used as x
//       CROSSCHECK = <prefix>: write <prefix>.<semiring>.x (float32[cols]) and
//                    <prefix>.<semiring>.y (float64[rows]) after the run
//       FORCE_CPU_VERIFY = 1: run the CPU check also for more than 30M rows
//
// It reads <base>.val/.vc.R/.vc.C, builds the reference-layout schedule, makes
// x the way the reference driver does (C rand() from the default seed: x in
// [0,1] for (+,x), 0/1 for Boolean, 0..99 with 10% 1e9 for Tropical) unless XVEC
// is given, and checks the drop-in entry points in the same run as the
// reference driver does (gpu_engine_test.cu:135-275): the host library's CPU
// layered sweep with all OpenMP threads and with one (the two must agree:
// a disagreement would be a race in the parallel sweep), then
// run_gpu_right_multiply_layered, and for B > 1 run_gpu_right_multiply_batched
// with x in every column; every result is compared with the CPU y as
// |gpu - cpu| / (1 + |cpu|) < 1e-4, and the exit status is 1 on a mismatch.
#include <sys/resource.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include <omp.h>

#include "build_schedule.h"
#include "gpu_engine.h"
#include "grammar.h"
#include "host.hpp"

namespace {

template <class T>
void dump(const std::string& prefix, const std::string& sr, const std::vector<float>& x, const std::vector<T>& y) {
    std::ofstream fx(prefix + "." + sr + ".x", std::ios::binary);
    fx.write(reinterpret_cast<const char*>(x.data()), static_cast<std::streamsize>(x.size() * sizeof(float)));
    std::vector<double> yd(y.begin(), y.end());
    std::ofstream fy(prefix + "." + sr + ".y", std::ios::binary);
    fy.write(reinterpret_cast<const char*>(yd.data()), static_cast<std::streamsize>(yd.size() * sizeof(double)));
    std::cout << "CROSSCHECK: dumped engine x/y to " << prefix << "." << sr << ".{x,y}" << std::endl;
}

// max_i |a_i - b_i| / (1 + |b_i|), and max |a_i - b_i|, max |b_i| (reference gpu_engine_test.cu:256-263).
template <class T>
double max_rel(const std::vector<T>& a, const std::vector<T>& b, double* max_abs = nullptr, double* max_y = nullptr) {
    double rel = 0.0, ab = 0.0, my = 0.0;
    for (std::size_t i = 0; i < a.size(); ++i) {
        const double d = std::abs(static_cast<double>(a[i]) - static_cast<double>(b[i]));
        const double m = std::abs(static_cast<double>(b[i]));
        rel = std::max(rel, d / (1.0 + m));
        ab = std::max(ab, d);
        my = std::max(my, m);
    }
    if (max_abs) *max_abs = ab;
    if (max_y) *max_y = my;
    return rel;
}

template <class T>
int run(int semiring, int rows, int cols, int iters, int B, bool skip_cpu, const std::string& srname,
        const std::vector<float>& x, const std::vector<float>& V, const std::vector<int>& vi,
        const std::vector<int>& ci, BuiltSchedule& bs) {
    GPUSchedule sched = bs.gpu_schedule();
    std::vector<T> y_cpu;
    if (!skip_cpu) {
        const int threads = omp_get_max_threads();
        std::vector<T> y_omp(rows);
        y_cpu.resize(rows);
        std::cout << "Running CPU reference with OpenMP (" << threads << " threads)..." << std::endl;
        auto c0 = std::chrono::high_resolution_clock::now();
        gmr::cpu_layered_sweep(semiring, bs, vi.data(), ci.data(), V.data(), x.data(), y_omp.data(), threads);
        std::chrono::duration<double, std::milli> dc = std::chrono::high_resolution_clock::now() - c0;
        std::cout << "CPU reference (OpenMP): " << dc.count() << " ms/vector (" << threads << " threads)." << std::endl;
        std::cout << "Running CPU reference sequentially (forced 1 thread)..." << std::endl;
        c0 = std::chrono::high_resolution_clock::now();
        gmr::cpu_layered_sweep(semiring, bs, vi.data(), ci.data(), V.data(), x.data(), y_cpu.data(), 1);
        dc = std::chrono::high_resolution_clock::now() - c0;
        std::cout << "CPU reference (sequential): " << dc.count() << " ms/vector." << std::endl;
        const double d = max_rel(y_omp, y_cpu);
        std::cout << "CPU OpenMP vs sequential: max_rel_diff=" << d << (d < 1e-4 ? "  SUCCESS" : "  FAILURE")
                  << std::endl;
    }
    std::vector<T> y(rows);
    std::cout << "Running GPU engine (double-buffered proper layered, emit-on-the-spot)..." << std::endl;
    auto t0 = std::chrono::high_resolution_clock::now();
    run_gpu_right_multiply_layered(rows, cols, bs.alpha, static_cast<int>(bs.flattened_rules.size()),
                                   bs.flattened_rules.data(), vi.data(), ci.data(), V.data(),
                                   static_cast<int>(V.size()), x.data(), sched, reinterpret_cast<float*>(y.data()),
                                   iters);
    std::chrono::duration<double, std::milli> dt = std::chrono::high_resolution_clock::now() - t0;
    std::cout << "GPU multiplication completed in " << dt.count() << " ms." << std::endl;
    bool batched_ok = true;
    if (B > 1) {
        std::cout << "Batched run B=" << B << "..." << std::endl;
        std::vector<float> X(static_cast<std::size_t>(cols) * B);
        for (int j = 0; j < cols; ++j)
            for (int b = 0; b < B; ++b) X[static_cast<std::size_t>(j) * B + b] = x[j];
        std::vector<T> Y(static_cast<std::size_t>(rows) * B);
        run_gpu_right_multiply_batched(rows, cols, bs.alpha, static_cast<int>(bs.flattened_rules.size()),
                                       bs.flattened_rules.data(), vi.data(), ci.data(), V.data(),
                                       static_cast<int>(V.size()), X.data(), sched, reinterpret_cast<float*>(Y.data()),
                                       B, iters);
        // Every column against the CPU y (or, without the CPU run, the GPU's single-vector y).
        const std::vector<T>& ref = skip_cpu ? y : y_cpu;
        double worst = 0.0;
        std::vector<T> col(rows);
        for (int b = 0; b < B; ++b) {
            for (int r = 0; r < rows; ++r) col[r] = Y[static_cast<std::size_t>(r) * B + b];
            worst = std::max(worst, max_rel(col, ref));
        }
        batched_ok = worst < 1e-4;
        std::cout << "Batched verification: max_rel_diff=" << worst << (batched_ok ? "  SUCCESS" : "  FAILURE")
                  << std::endl;
    }
    if (const char* cc = std::getenv("CROSSCHECK"))
        if (cc[0]) dump(cc, srname, x, y);
    if (skip_cpu) {
        std::cout << "Skipped CPU reference verification for large dataset." << std::endl;
        return batched_ok ? 0 : 1;
    }
    double max_abs = 0.0, max_y = 0.0;
    const double rel = max_rel(y, y_cpu, &max_abs, &max_y);
    std::cout << "Verification: max_abs_diff=" << max_abs << ", max_rel_diff=" << rel << ", max|y|=" << max_y
              << std::endl;
    if (rel < 1e-4 && batched_ok) {
        std::cout << "SUCCESS: GPU and CPU reference results match (within float precision)!" << std::endl;
        return 0;
    }
    std::cerr << "FAILURE: Results mismatch!" << std::endl;
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 4) {
        std::cerr << "Usage: " << argv[0] << " <matrix_base_name> <rows> <cols> [iters] [mode] [B]" << std::endl;
        std::cerr << "  mode: 'repair' (default; the RePair construction)" << std::endl;
        std::cerr << "  B:    optional batch width for the batched (SpMM) path" << std::endl;
        std::cerr << "  semiring is chosen by the SEMIRING env var: plustimes | boolean | tropical" << std::endl;
        return 1;
    }
    const std::string base = argv[1];
    const int rows = std::stoi(argv[2]);
    const int cols = std::stoi(argv[3]);
    const int iters = argc >= 5 ? std::stoi(argv[4]) : 1;
    const int B = argc >= 7 ? std::stoi(argv[6]) : 1;
    std::vector<double> Vd;
    grammar::Grammar g;
    try {
        g = gmr::read_grammar_files(base, &Vd);
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << std::endl;
        return 1;
    }
    std::vector<float> V(Vd.begin(), Vd.end());
    std::cout << "RePair construction: alpha=" << g.alpha << ", NTs=" << g.rules.size()
              << ", distinct values=" << V.size() << std::endl;
    const int alpha_full = g.alpha;
    auto t0 = std::chrono::high_resolution_clock::now();
    BuiltSchedule bs;
    try {
        bs = build_schedule(g, rows, cols);
    } catch (const std::exception& e) {
        std::cerr << "Error: " << e.what() << std::endl;
        return 1;
    }
    std::chrono::duration<double, std::milli> dt = std::chrono::high_resolution_clock::now() - t0;
    struct rusage ru;
    getrusage(RUSAGE_SELF, &ru);
    std::cout << "Schedule builder construction time: " << dt.count() << " ms" << std::endl;
    std::cout << "BUILD schedule ms: " << dt.count() << " peak host MB: " << ru.ru_maxrss / 1024.0 << std::endl;
    const int U = bs.alpha;
    std::vector<int> vi(U > 0 ? U - 1 : 0), ci(vi.size());
    for (int q = 1; q < U; ++q) {
        const int t = bs.term_orig[q];
        vi[q - 1] = (t - 1) / cols;
        ci[q - 1] = (t - 1) % cols;
    }
    std::cout << "Terminal compaction: alpha " << alpha_full << " -> " << U << " ("
              << (alpha_full > 0 ? 100.0 * U / alpha_full : 0.0) << "% kept)" << std::endl;
    std::cout << "Proper Layered Completed Grammar Stats:" << std::endl;
    std::cout << "Max depth (L): " << bs.max_depth << ", Max width (w*): " << bs.max_width
              << ", Total layered rules: " << bs.num_layered << " (raw NTs: " << bs.num_raw_nt
              << ", +pt: " << bs.num_passthrough << ")"
              << ", emissions: " << bs.emit_pos.size() << "+" << bs.term_emit_sym.size() << " term" << std::endl;

    const char* se = std::getenv("SEMIRING");
    const std::string sr = se ? se : "plustimes";
    const bool is_bool = sr == "boolean" || sr == "bool";
    const bool is_trop = sr == "tropical" || sr == "minplus";
    std::vector<float> x(cols);
    for (int j = 0; j < cols; ++j) {
        if (is_bool) x[j] = (std::rand() % 2 == 0) ? 0.0f : 1.0f;
        else if (is_trop) x[j] = (std::rand() % 10 == 0) ? 1000000000.0f : static_cast<float>(std::rand() % 100);
        else x[j] = static_cast<float>(std::rand()) / RAND_MAX;
    }
    if (const char* xf = std::getenv("XVEC")) {
        std::ifstream fx(xf, std::ios::binary | std::ios::ate);
        if (!fx.is_open()) {
            std::cerr << "XVEC: cannot open " << xf << std::endl;
            return 1;
        }
        const std::size_t bytes = static_cast<std::size_t>(fx.tellg());
        if (bytes % sizeof(float) != 0 || bytes / sizeof(float) != static_cast<std::size_t>(cols)) {
            std::cerr << "XVEC: " << xf << " holds " << bytes / sizeof(float) << " float32 values, expected cols="
                      << cols << std::endl;
            return 1;
        }
        fx.seekg(0, std::ios::beg);
        fx.read(reinterpret_cast<char*>(x.data()), static_cast<std::streamsize>(bytes));
        std::cout << "XVEC: loaded fixed x (" << cols << " float32) from " << xf << std::endl;
    }
    // The CPU check is skipped above 30M rows unless FORCE_CPU_VERIFY is set
    // (reference gpu_engine_test.cu:436-441).
    bool skip_cpu = rows > 30000000;
    if (const char* fv = std::getenv("FORCE_CPU_VERIFY"))
        if (fv[0] && fv[0] != '0') skip_cpu = false;
    if (is_bool) {
        std::cout << "--- SEMIRING: Boolean ---" << std::endl;
        return run<int>(1, rows, cols, iters, B, skip_cpu, "boolean", x, V, vi, ci, bs);
    }
    if (is_trop) {
        std::cout << "--- SEMIRING: Tropical (min,+) ---" << std::endl;
        return run<int>(2, rows, cols, iters, B, skip_cpu, "tropical", x, V, vi, ci, bs);
    }
    std::cout << "--- SEMIRING: PlusTimes (+,*) ---" << std::endl;
    return run<float>(0, rows, cols, iters, B, skip_cpu, "plustimes", x, V, vi, ci, bs);
}
