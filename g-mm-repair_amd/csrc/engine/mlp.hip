// A small multilayer perceptron trained and evaluated on the GPU by its own
// kernels: in -> H -> H -> out, ReLU after the two hidden layers, masked
// mean-squared-error loss, Adam. The schedule proposer uses it to predict the
// log device time of each node-update kernel for a level from the level's
// features; it never touches the sweep's arithmetic.
//
// Parameters, one flat FP32 array: W1 [H x in], b1 [H], W2 [H x H], b2 [H],
// W3 [out x H], b3 [out] (row-major, W[o][i] at o*in + i). Every kernel writes
// each output element from exactly one thread in a fixed order, so training is
// deterministic (no floating-point atomics).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "gmr.h"
#include "rng.hpp"

void gmr_host_set_error(const char* msg);

namespace gmr {
namespace mlp {

__global__ void k_dense_fwd(const float* __restrict__ X, const float* __restrict__ W, const float* __restrict__ b,
                            float* __restrict__ Z, float* __restrict__ A, int N, int In, int Out, int relu) {
    const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<long>(N) * Out) return;
    const int n = static_cast<int>(t / Out), o = static_cast<int>(t % Out);
    float z = b[o];
    const float* x = X + static_cast<long>(n) * In;
    const float* w = W + static_cast<long>(o) * In;
    for (int i = 0; i < In; ++i) z = fmaf(w[i], x[i], z);
    Z[t] = z;
    if (A) A[t] = relu ? fmaxf(z, 0.0f) : z;
}

// dOut = 2 * mask * (out - target) / count; one thread per element.
__global__ void k_mse_grad(const float* __restrict__ O, const float* __restrict__ T, const float* __restrict__ M,
                           float* __restrict__ dO, long n, float inv_count) {
    const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float d = (O[t] - T[t]) * M[t];
    dO[t] = 2.0f * d * inv_count;
}

// Sum of mask * (out - target)^2 in one workgroup (fixed order per thread,
// then a fixed tree): deterministic.
__global__ void k_mse_loss(const float* __restrict__ O, const float* __restrict__ T, const float* __restrict__ M,
                           long n, float* __restrict__ loss) {
    __shared__ double part[1024];
    double acc = 0.0;
    for (long i = threadIdx.x; i < n; i += blockDim.x) {
        const double d = (static_cast<double>(O[i]) - T[i]) * M[i];
        acc += d * d;
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss[0] = static_cast<float>(part[0]);
}

// dA_prev[n][i] = sum_o dZ[n][o] * W[o][i], then (relu) * (Zprev[n][i] > 0).
__global__ void k_dense_bwd_in(const float* __restrict__ dZ, const float* __restrict__ W,
                               const float* __restrict__ Zprev, float* __restrict__ dZprev, int N, int In, int Out) {
    const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<long>(N) * In) return;
    const int n = static_cast<int>(t / In), i = static_cast<int>(t % In);
    float g = 0.0f;
    for (int o = 0; o < Out; ++o) g = fmaf(dZ[static_cast<long>(n) * Out + o], W[static_cast<long>(o) * In + i], g);
    dZprev[t] = Zprev[t] > 0.0f ? g : 0.0f;
}

// dW[o][i] = sum_n dZ[n][o] X[n][i]; thread (o, In) computes db[o].
__global__ void k_dense_bwd_w(const float* __restrict__ dZ, const float* __restrict__ X, float* __restrict__ dW,
                              float* __restrict__ db, int N, int In, int Out) {
    const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<long>(Out) * (In + 1)) return;
    const int o = static_cast<int>(t / (In + 1)), i = static_cast<int>(t % (In + 1));
    float g = 0.0f;
    if (i < In) {
        for (int n = 0; n < N; ++n) g = fmaf(dZ[static_cast<long>(n) * Out + o], X[static_cast<long>(n) * In + i], g);
        dW[static_cast<long>(o) * In + i] = g;
    } else {
        for (int n = 0; n < N; ++n) g += dZ[static_cast<long>(n) * Out + o];
        db[o] = g;
    }
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, long n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gi = g[t];
    const float mi = b1 * m[t] + (1.0f - b1) * gi;
    const float vi = b2 * v[t] + (1.0f - b2) * gi * gi;
    m[t] = mi;
    v[t] = vi;
    p[t] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace mlp

struct Mlp {
    int nin = 0, nh = 0, nout = 0;
    long nparam = 0;
    float* d_p = nullptr;  // parameters
    float* d_g = nullptr;  // gradients
    float* d_m = nullptr;  // Adam first moment
    float* d_v = nullptr;  // Adam second moment
    int step = 0;
    // activations, sized for cap samples
    int cap = 0;
    float *Z1 = nullptr, *A1 = nullptr, *Z2 = nullptr, *A2 = nullptr, *O = nullptr;
    float *dO = nullptr, *dZ2 = nullptr, *dZ1 = nullptr, *loss = nullptr;

    long off_W1() const { return 0; }
    long off_b1() const { return static_cast<long>(nh) * nin; }
    long off_W2() const { return off_b1() + nh; }
    long off_b2() const { return off_W2() + static_cast<long>(nh) * nh; }
    long off_W3() const { return off_b2() + nh; }
    long off_b3() const { return off_W3() + static_cast<long>(nout) * nh; }

    ~Mlp() {
        for (float* q : {d_p, d_g, d_m, d_v, Z1, A1, Z2, A2, O, dO, dZ2, dZ1, loss})
            if (q) (void)hipFree(q);
    }
};

namespace {

thread_local std::string g_mlp_err;

void chk(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

inline unsigned blocks(long n, int b = 256) { return static_cast<unsigned>((n + b - 1) / b); }

void ensure_cap(Mlp& m, int N) {
    if (N <= m.cap) return;
    for (float** q : {&m.Z1, &m.A1, &m.Z2, &m.A2, &m.O, &m.dO, &m.dZ2, &m.dZ1}) {
        if (*q) chk(hipFree(*q), "hipFree");
        *q = nullptr;
    }
    const long h = static_cast<long>(N) * m.nh, o = static_cast<long>(N) * m.nout;
    chk(hipMalloc(&m.Z1, h * 4), "hipMalloc");
    chk(hipMalloc(&m.A1, h * 4), "hipMalloc");
    chk(hipMalloc(&m.Z2, h * 4), "hipMalloc");
    chk(hipMalloc(&m.A2, h * 4), "hipMalloc");
    chk(hipMalloc(&m.O, o * 4), "hipMalloc");
    chk(hipMalloc(&m.dO, o * 4), "hipMalloc");
    chk(hipMalloc(&m.dZ2, h * 4), "hipMalloc");
    chk(hipMalloc(&m.dZ1, h * 4), "hipMalloc");
    m.cap = N;
}

void forward(Mlp& m, const float* X, int N, hipStream_t st) {
    ensure_cap(m, N);
    const float* p = m.d_p;
    hipLaunchKernelGGL(mlp::k_dense_fwd, dim3(blocks(static_cast<long>(N) * m.nh)), dim3(256), 0, st, X,
                       p + m.off_W1(), p + m.off_b1(), m.Z1, m.A1, N, m.nin, m.nh, 1);
    hipLaunchKernelGGL(mlp::k_dense_fwd, dim3(blocks(static_cast<long>(N) * m.nh)), dim3(256), 0, st, m.A1,
                       p + m.off_W2(), p + m.off_b2(), m.Z2, m.A2, N, m.nh, m.nh, 1);
    hipLaunchKernelGGL(mlp::k_dense_fwd, dim3(blocks(static_cast<long>(N) * m.nout)), dim3(256), 0, st, m.A2,
                       p + m.off_W3(), p + m.off_b3(), m.O, nullptr, N, m.nh, m.nout, 0);
    chk(hipGetLastError(), "mlp forward");
}

// Loss and gradients into m.d_g (and m.loss[0] = sum of squared masked errors).
void loss_grad(Mlp& m, const float* X, const float* T, const float* M, int N, float count, hipStream_t st) {
    forward(m, X, N, st);
    const long no = static_cast<long>(N) * m.nout;
    float* g = m.d_g;
    hipLaunchKernelGGL(mlp::k_mse_grad, dim3(blocks(no)), dim3(256), 0, st, m.O, T, M, m.dO, no, 1.0f / count);
    hipLaunchKernelGGL(mlp::k_mse_loss, dim3(1), dim3(1024), 0, st, m.O, T, M, no, m.loss);
    // layer 3
    hipLaunchKernelGGL(mlp::k_dense_bwd_w, dim3(blocks(static_cast<long>(m.nout) * (m.nh + 1))), dim3(256), 0, st,
                       m.dO, m.A2, g + m.off_W3(), g + m.off_b3(), N, m.nh, m.nout);
    hipLaunchKernelGGL(mlp::k_dense_bwd_in, dim3(blocks(static_cast<long>(N) * m.nh)), dim3(256), 0, st, m.dO,
                       m.d_p + m.off_W3(), m.Z2, m.dZ2, N, m.nh, m.nout);
    // layer 2
    hipLaunchKernelGGL(mlp::k_dense_bwd_w, dim3(blocks(static_cast<long>(m.nh) * (m.nh + 1))), dim3(256), 0, st,
                       m.dZ2, m.A1, g + m.off_W2(), g + m.off_b2(), N, m.nh, m.nh);
    hipLaunchKernelGGL(mlp::k_dense_bwd_in, dim3(blocks(static_cast<long>(N) * m.nh)), dim3(256), 0, st, m.dZ2,
                       m.d_p + m.off_W2(), m.Z1, m.dZ1, N, m.nh, m.nh);
    // layer 1
    hipLaunchKernelGGL(mlp::k_dense_bwd_w, dim3(blocks(static_cast<long>(m.nh) * (m.nin + 1))), dim3(256), 0, st,
                       m.dZ1, X, g + m.off_W1(), g + m.off_b1(), N, m.nin, m.nh);
    chk(hipGetLastError(), "mlp backward");
}

void adam(Mlp& m, float lr, hipStream_t st) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    ++m.step;
    const float c1 = 1.0f - std::pow(b1, static_cast<float>(m.step));
    const float c2 = 1.0f - std::pow(b2, static_cast<float>(m.step));
    hipLaunchKernelGGL(mlp::k_adam, dim3(blocks(m.nparam)), dim3(256), 0, st, m.d_p, m.d_g, m.d_m, m.d_v,
                       m.nparam, lr, b1, b2, eps, c1, c2);
    chk(hipGetLastError(), "mlp adam");
}

template <class F>
int mlp_guarded(F&& f) {
    try {
        f();
        return GMR_OK;
    } catch (const std::invalid_argument& e) {
        g_mlp_err = e.what();
        gmr_host_set_error(e.what());
        return GMR_ERR_INVALID;
    } catch (const std::bad_alloc&) {
        gmr_host_set_error("host allocation failed");
        return GMR_ERR_NOMEM;
    } catch (const std::exception& e) {
        g_mlp_err = e.what();
        gmr_host_set_error(e.what());
        return GMR_ERR_DEVICE;
    }
}

}  // namespace
}  // namespace gmr

struct gmr_mlp {
    gmr::Mlp m;
};

extern "C" {

int gmr_mlp_create(int n_in, int n_hidden, int n_out, uint64_t seed, gmr_mlp** out) {
    if (!out || n_in < 1 || n_hidden < 1 || n_out < 1) {
        gmr_host_set_error("mlp: bad shape");
        return GMR_ERR_INVALID;
    }
    gmr_mlp* h = new (std::nothrow) gmr_mlp;
    if (!h) return GMR_ERR_NOMEM;
    int rc = gmr::mlp_guarded([&] {
        gmr::Mlp& m = h->m;
        m.nin = n_in;
        m.nh = n_hidden;
        m.nout = n_out;
        m.nparam = m.off_b3() + n_out;
        // He-uniform weights from the seed, zero biases.
        std::vector<float> p(static_cast<std::size_t>(m.nparam), 0.0f);
        gmr::Rng r(seed, 0x4d4c50ULL);
        auto fill = [&](long off, int fan_out, int fan_in) {
            const double lim = std::sqrt(6.0 / fan_in);
            for (long i = 0; i < static_cast<long>(fan_out) * fan_in; ++i)
                p[off + i] = static_cast<float>((2.0 * r.uniform() - 1.0) * lim);
        };
        fill(m.off_W1(), m.nh, m.nin);
        fill(m.off_W2(), m.nh, m.nh);
        fill(m.off_W3(), m.nout, m.nh);
        gmr::chk(hipMalloc(&m.d_p, m.nparam * 4), "hipMalloc");
        gmr::chk(hipMalloc(&m.d_g, m.nparam * 4), "hipMalloc");
        gmr::chk(hipMalloc(&m.d_m, m.nparam * 4), "hipMalloc");
        gmr::chk(hipMalloc(&m.d_v, m.nparam * 4), "hipMalloc");
        gmr::chk(hipMalloc(&m.loss, 4), "hipMalloc");
        gmr::chk(hipMemcpy(m.d_p, p.data(), m.nparam * 4, hipMemcpyHostToDevice), "hipMemcpy");
        gmr::chk(hipMemset(m.d_g, 0, m.nparam * 4), "hipMemset");
        gmr::chk(hipMemset(m.d_m, 0, m.nparam * 4), "hipMemset");
        gmr::chk(hipMemset(m.d_v, 0, m.nparam * 4), "hipMemset");
        gmr::chk(hipDeviceSynchronize(), "sync");
    });
    if (rc != GMR_OK) {
        delete h;
        return rc;
    }
    *out = h;
    return GMR_OK;
}

long long gmr_mlp_num_params(const gmr_mlp* h) { return h ? h->m.nparam : -1; }

int gmr_mlp_get_params(const gmr_mlp* h, float* host_params) {
    if (!h || !host_params) return GMR_ERR_INVALID;
    return gmr::mlp_guarded([&] {
        gmr::chk(hipMemcpy(host_params, h->m.d_p, h->m.nparam * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    });
}

int gmr_mlp_set_params(gmr_mlp* h, const float* host_params) {
    if (!h || !host_params) return GMR_ERR_INVALID;
    return gmr::mlp_guarded([&] {
        gmr::chk(hipMemcpy(h->m.d_p, host_params, h->m.nparam * 4, hipMemcpyHostToDevice), "hipMemcpy");
        gmr::chk(hipMemset(h->m.d_m, 0, h->m.nparam * 4), "hipMemset");
        gmr::chk(hipMemset(h->m.d_v, 0, h->m.nparam * 4), "hipMemset");
        h->m.step = 0;
        gmr::chk(hipDeviceSynchronize(), "sync");
    });
}

int gmr_mlp_forward(gmr_mlp* h, const float* d_X, int N, float* d_out, void* stream) {
    if (!h || !d_X || !d_out || N < 1) return GMR_ERR_INVALID;
    return gmr::mlp_guarded([&] {
        auto st = static_cast<hipStream_t>(stream);
        gmr::forward(h->m, d_X, N, st);
        gmr::chk(hipMemcpyAsync(d_out, h->m.O, static_cast<long>(N) * h->m.nout * 4, hipMemcpyDeviceToDevice, st),
                 "hipMemcpyAsync");
    });
}

int gmr_mlp_loss_grad(gmr_mlp* h, const float* d_X, const float* d_T, const float* d_M, int N, float count,
                      float* loss, float* host_grad) {
    if (!h || !d_X || !d_T || !d_M || N < 1 || !(count > 0)) return GMR_ERR_INVALID;
    return gmr::mlp_guarded([&] {
        gmr::loss_grad(h->m, d_X, d_T, d_M, N, count, nullptr);
        gmr::chk(hipDeviceSynchronize(), "sync");
        if (loss) {
            float l = 0;
            gmr::chk(hipMemcpy(&l, h->m.loss, 4, hipMemcpyDeviceToHost), "hipMemcpy");
            *loss = l / count;
        }
        if (host_grad) gmr::chk(hipMemcpy(host_grad, h->m.d_g, h->m.nparam * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    });
}

int gmr_mlp_train(gmr_mlp* h, const float* d_X, const float* d_T, const float* d_M, int N, float count, int epochs,
                  float lr, float* losses) {
    if (!h || !d_X || !d_T || !d_M || N < 1 || epochs < 0 || !(count > 0)) return GMR_ERR_INVALID;
    return gmr::mlp_guarded([&] {
        for (int e = 0; e < epochs; ++e) {
            gmr::loss_grad(h->m, d_X, d_T, d_M, N, count, nullptr);
            gmr::adam(h->m, lr, nullptr);
            if (losses) {
                float l = 0;
                gmr::chk(hipMemcpy(&l, h->m.loss, 4, hipMemcpyDeviceToHost), "hipMemcpy");
                losses[e] = l / count;
            }
        }
        gmr::chk(hipDeviceSynchronize(), "sync");
    });
}

void gmr_mlp_destroy(gmr_mlp* h) { delete h; }

}  // extern "C"
