// gather_ceiling -- what rate can a gather of random 256-byte rows reach on
// this GPU, as a function of the size of the array the rows come from?
//
// The B = 64 row gather reads one 256-byte value row (64 floats) per top-
// sequence symbol from anywhere in a resident value array of about 1.7 GB.
// This tool runs the same access shape on synthetic data: lane groups of
// row_bytes / 16 lanes (one dwordx4 each), 64 uniformly random row indices per group, eight
// row reads in flight, one result row per group, over arrays from
// 2 MB (inside one XCD's L2) to 1.69 GB. The printed rate counts row bytes
// read, so it is the practical ceiling the sweep's gather is measured against.
//
// Usage: gather_ceiling [reads_in_millions (default 64)] [row_bytes: 64, 128 or 256 (default 256)]
// (64- and 128-byte rows are the B = 16 and B = 32 gathers' rows.)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#define CK(x)                                                                        \
    do {                                                                             \
        hipError_t e_ = (x);                                                         \
        if (e_ != hipSuccess) {                                                      \
            std::fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), \
                         __LINE__);                                                  \
            std::exit(1);                                                            \
        }                                                                            \
    } while (0)

namespace {

constexpr int kPerGroup = 64;  // row indices per lane group
constexpr int kInFlight = 8;   // row reads in flight per lane

template <int LANES>
__global__ void k_random_rows(const float4* __restrict__ W, const int* __restrict__ idx,
                              float4* __restrict__ out, long groups) {
    const long tid = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long g = tid / LANES;
    const int lane = static_cast<int>(tid % LANES);
    if (g >= groups) return;
    const int* ix = idx + g * kPerGroup;
    float4 s0 = make_float4(0.f, 0.f, 0.f, 0.f), s1 = s0;
    for (int j = 0; j < kPerGroup; j += kInFlight) {
        int c[kInFlight];
#pragma unroll
        for (int u = 0; u < kInFlight; ++u) c[u] = ix[j + u];
        float4 v[kInFlight];
#pragma unroll
        for (int u = 0; u < kInFlight; ++u) v[u] = W[static_cast<long>(c[u]) * LANES + lane];
#pragma unroll
        for (int u = 0; u < kInFlight / 2; ++u) {
            s0.x += v[u].x;
            s0.y += v[u].y;
            s0.z += v[u].z;
            s0.w += v[u].w;
            s1.x += v[u + kInFlight / 2].x;
            s1.y += v[u + kInFlight / 2].y;
            s1.z += v[u + kInFlight / 2].z;
            s1.w += v[u + kInFlight / 2].w;
        }
    }
    out[g * LANES + lane] = make_float4(s0.x + s1.x, s0.y + s1.y, s0.z + s1.z, s0.w + s1.w);
}

template <int LANES>
void launch(unsigned grid, int block, const float4* W, const int* idx, float4* out, long groups) {
    hipLaunchKernelGGL(k_random_rows<LANES>, dim3(grid), dim3(block), 0, 0, W, idx, out, groups);
}

}  // namespace

int main(int argc, char** argv) {
    const long reads = (argc > 1 ? std::atol(argv[1]) : 64) << 20;
    const int row_bytes = argc > 2 ? std::atoi(argv[2]) : 256;
    if (reads <= 0 || reads % kPerGroup || (row_bytes != 64 && row_bytes != 128 && row_bytes != 256)) {
        std::fprintf(stderr, "usage: gather_ceiling [reads_in_millions] [64|128|256]\n");
        return 2;
    }
    const int lanes = row_bytes / 16;
    const long groups = reads / kPerGroup;
    const long max_bytes = 6600000L * 256;  // 1.69 GB
    float4* W = nullptr;
    int* idx = nullptr;
    float4* out = nullptr;
    CK(hipMalloc(&W, max_bytes));
    CK(hipMemset(W, 0, max_bytes));
    CK(hipMalloc(&idx, reads * sizeof(int)));
    CK(hipMalloc(&out, groups * row_bytes));
    std::vector<int> h(static_cast<std::size_t>(reads));
    std::mt19937_64 rng(20261204);
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    const int block = 256;
    const unsigned grid = static_cast<unsigned>((groups * lanes + block - 1) / block);
    auto run = [&] {
        if (lanes == 4) launch<4>(grid, block, W, idx, out, groups);
        else if (lanes == 8) launch<8>(grid, block, W, idx, out, groups);
        else launch<16>(grid, block, W, idx, out, groups);
    };
    std::printf("# %ld random %d-byte row reads per launch, %d lanes x dwordx4 per row, %d in flight\n", reads,
                row_bytes, lanes, kInFlight);
    for (long bytes : {2L << 20, 16L << 20, 64L << 20, 256L << 20, 512L << 20, max_bytes}) {
        const long rows = bytes / row_bytes;
        for (long i = 0; i < reads; ++i) h[static_cast<std::size_t>(i)] = static_cast<int>(rng() % rows);
        CK(hipMemcpy(idx, h.data(), reads * sizeof(int), hipMemcpyHostToDevice));
        run();
        CK(hipGetLastError());
        CK(hipDeviceSynchronize());
        const int reps = 5;
        CK(hipEventRecord(e0));
        for (int r = 0; r < reps; ++r) run();
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        float ms = 0.f;
        CK(hipEventElapsedTime(&ms, e0, e1));
        ms /= reps;
        std::printf("array %9ld rows (%7.1f MB): %.3f ms per launch, %.2f TB/s of row reads\n", rows,
                    rows * static_cast<double>(row_bytes) / 1e6, ms,
                    reads * static_cast<double>(row_bytes) / (ms * 1e-3) / 1e12);
        std::fflush(stdout);
    }
    CK(hipFree(W));
    CK(hipFree(idx));
    CK(hipFree(out));
    return 0;
}
