// .val / .vc.R / .vc.C files (the format the reference's readers expect,
// gpu-engine/gpu_engine_test.cu:291-345 and cpu_selftest.cpp:164-221):
//   <base>.val   float64[nvals], the distinct non-zero values
//   <base>.vc.R  int32 alpha, then int32 (left, right) per rule
//   <base>.vc.C  int32 top sequence, 0 ends each row
// Unlike the reference readers, a truncated .vc.R (a dangling half rule) or a
// file whose size is not a whole number of elements is an error here.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>

#include "host.hpp"

namespace gmr {

namespace {

std::vector<char> slurp(const std::string& path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f.is_open()) throw std::runtime_error("cannot open " + path);
    std::streamsize n = f.tellg();
    f.seekg(0, std::ios::beg);
    std::vector<char> b(static_cast<std::size_t>(n));
    if (n > 0 && !f.read(b.data(), n)) throw std::runtime_error("cannot read " + path);
    return b;
}

void spill(const std::string& path, const void* p, std::size_t n) {
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f.is_open()) throw std::runtime_error("cannot create " + path);
    if (n > 0) f.write(static_cast<const char*>(p), static_cast<std::streamsize>(n));
    if (!f) throw std::runtime_error("cannot write " + path);
}

}  // namespace

void write_grammar_files(const std::string& base, const grammar::Grammar& g,
                         const std::vector<double>& V) {
    spill(base + ".val", V.data(), V.size() * sizeof(double));
    std::vector<int> r;
    r.reserve(1 + 2 * g.rules.size());
    r.push_back(g.alpha);
    for (const auto& rule : g.rules) {
        if (rule.kind != grammar::BINARY)
            throw std::invalid_argument("the .vc.R format holds binary rules only");
        r.push_back(rule.left);
        r.push_back(rule.right);
    }
    spill(base + ".vc.R", r.data(), r.size() * sizeof(int));
    spill(base + ".vc.C", g.C.data(), g.C.size() * sizeof(int));
}

grammar::Grammar read_grammar_files(const std::string& base, std::vector<double>* V) {
    grammar::Grammar g;
    auto val = slurp(base + ".val");
    if (val.size() % sizeof(double)) throw std::runtime_error(base + ".val: size is not a multiple of 8");
    if (V) {
        V->resize(val.size() / sizeof(double));
        std::memcpy(V->data(), val.data(), val.size());
    }
    auto rb = slurp(base + ".vc.R");
    if (rb.size() < sizeof(int) || (rb.size() - sizeof(int)) % (2 * sizeof(int)))
        throw std::runtime_error(base + ".vc.R: truncated");
    std::memcpy(&g.alpha, rb.data(), sizeof(int));
    const std::size_t nr = (rb.size() - sizeof(int)) / (2 * sizeof(int));
    g.rules.resize(nr);
    const int* p = reinterpret_cast<const int*>(rb.data() + sizeof(int));
    for (std::size_t i = 0; i < nr; ++i) g.rules[i] = {grammar::BINARY, p[2 * i], p[2 * i + 1], 1};
    auto cb = slurp(base + ".vc.C");
    if (cb.size() % sizeof(int)) throw std::runtime_error(base + ".vc.C: size is not a multiple of 4");
    g.C.resize(cb.size() / sizeof(int));
    std::memcpy(g.C.data(), cb.data(), cb.size());
    return g;
}

}  // namespace gmr
