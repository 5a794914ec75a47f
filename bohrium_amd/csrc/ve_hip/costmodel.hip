/* The learned kernel cost model: a two-hidden-layer ReLU perceptron trained and
 * evaluated by hand-written HIP kernels (C-ABI in include/bh_costmodel.h).
 *
 * Layout of the flat parameter vector P (fp32):
 *   W1[F][H] | b1[H] | W2[H][H] | b2[H] | w3[H] | b3
 * Forward, per sample x (standardised as (x - mean) * inv_std):
 *   h1 = relu(x W1 + b1), h2 = relu(h1 W2 + b2), y = h2 . w3 + b3
 * Loss = mean((y - t)^2).  Backward: one thread per sample computes dy, dh2, dh1;
 * then one thread per parameter sums its gradient over the samples in sample order
 * (deterministic), and one thread per parameter applies Adam.
 */
#include "../../../include/bh_costmodel.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace {

constexpr int MAXF = 32;
constexpr int MAXH = 64;
constexpr int WG = 256;

thread_local std::string g_err;

struct Layout {
    int F, H;
    int64_t oW1, ob1, oW2, ob2, ow3, ob3, np;
};

Layout make_layout(int F, int H) {
    Layout L;
    L.F = F;
    L.H = H;
    L.oW1 = 0;
    L.ob1 = L.oW1 + static_cast<int64_t>(F) * H;
    L.oW2 = L.ob1 + H;
    L.ob2 = L.oW2 + static_cast<int64_t>(H) * H;
    L.ow3 = L.ob2 + H;
    L.ob3 = L.ow3 + H;
    L.np = L.ob3 + 1;
    return L;
}

__device__ __forceinline__ void forward_one(const float *__restrict__ P, const Layout L, const float *__restrict__ x_in,
                                            const float *__restrict__ mean, const float *__restrict__ istd,
                                            float *x, float *h1, float *h2, float *y) {
#pragma unroll
    for (int f = 0; f < MAXF; ++f) x[f] = f < L.F ? (x_in[f] - mean[f]) * istd[f] : 0.0f;
#pragma unroll
    for (int j = 0; j < MAXH; ++j) {
        if (j < L.H) {
            float s = P[L.ob1 + j];
#pragma unroll
            for (int f = 0; f < MAXF; ++f)
                if (f < L.F) s += x[f] * P[L.oW1 + f * L.H + j];
            h1[j] = s > 0.0f ? s : 0.0f;
        } else {
            h1[j] = 0.0f;
        }
    }
#pragma unroll
    for (int k = 0; k < MAXH; ++k) {
        if (k < L.H) {
            float s = P[L.ob2 + k];
#pragma unroll
            for (int j = 0; j < MAXH; ++j)
                if (j < L.H) s += h1[j] * P[L.oW2 + j * L.H + k];
            h2[k] = s > 0.0f ? s : 0.0f;
        } else {
            h2[k] = 0.0f;
        }
    }
    float o = P[L.ob3];
#pragma unroll
    for (int k = 0; k < MAXH; ++k)
        if (k < L.H) o += h2[k] * P[L.ow3 + k];
    *y = o;
}

__global__ void __launch_bounds__(WG) k_forward(const float *__restrict__ P, const Layout L, const float *__restrict__ mean,
                                                const float *__restrict__ istd, const float *__restrict__ X, int64_t n,
                                                float *__restrict__ Y, float *__restrict__ H1, float *__restrict__ H2) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * WG + threadIdx.x;
    if (i >= n) return;
    float x[MAXF], h1[MAXH], h2[MAXH], y;
    forward_one(P, L, X + i * L.F, mean, istd, x, h1, h2, &y);
    Y[i] = y;
    if (H1 != nullptr) {
        for (int j = 0; j < L.H; ++j) {
            H1[i * L.H + j] = h1[j];
            H2[i * L.H + j] = h2[j];
        }
    }
}

// Per sample: dL/dy, dL/dh2 (pre-activation), dL/dh1 (pre-activation) and the squared error.
__global__ void __launch_bounds__(WG) k_backward_sample(const float *__restrict__ P, const Layout L, const float *__restrict__ Y,
                                                        const float *__restrict__ T, const float *__restrict__ H1,
                                                        const float *__restrict__ H2, int64_t n, float *__restrict__ DY,
                                                        float *__restrict__ DH1, float *__restrict__ DH2,
                                                        float *__restrict__ SQ) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * WG + threadIdx.x;
    if (i >= n) return;
    const float e = Y[i] - T[i];
    SQ[i] = e * e;
    const float dy = 2.0f * e / static_cast<float>(n);
    DY[i] = dy;
    float dh2[MAXH];
#pragma unroll
    for (int k = 0; k < MAXH; ++k) {
        if (k < L.H) {
            dh2[k] = H2[i * L.H + k] > 0.0f ? dy * P[L.ow3 + k] : 0.0f;
            DH2[i * L.H + k] = dh2[k];
        } else {
            dh2[k] = 0.0f;
        }
    }
    for (int j = 0; j < L.H; ++j) {
        float s = 0.0f;
#pragma unroll
        for (int k = 0; k < MAXH; ++k)
            if (k < L.H) s += P[L.oW2 + j * L.H + k] * dh2[k];
        DH1[i * L.H + j] = H1[i * L.H + j] > 0.0f ? s : 0.0f;
    }
}

// One thread per parameter: its gradient summed over the samples in order.
__global__ void __launch_bounds__(WG) k_param_grad(const Layout L, const float *__restrict__ mean,
                                                   const float *__restrict__ istd, const float *__restrict__ X,
                                                   const float *__restrict__ H1, const float *__restrict__ H2,
                                                   const float *__restrict__ DY, const float *__restrict__ DH1,
                                                   const float *__restrict__ DH2, int64_t n, float *__restrict__ G) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * WG + threadIdx.x;
    if (p >= L.np) return;
    float s = 0.0f;
    if (p < L.ob1) {
        const int f = static_cast<int>((p - L.oW1) / L.H), j = static_cast<int>((p - L.oW1) % L.H);
        for (int64_t i = 0; i < n; ++i) s += (X[i * L.F + f] - mean[f]) * istd[f] * DH1[i * L.H + j];
    } else if (p < L.oW2) {
        const int j = static_cast<int>(p - L.ob1);
        for (int64_t i = 0; i < n; ++i) s += DH1[i * L.H + j];
    } else if (p < L.ob2) {
        const int j = static_cast<int>((p - L.oW2) / L.H), k = static_cast<int>((p - L.oW2) % L.H);
        for (int64_t i = 0; i < n; ++i) s += H1[i * L.H + j] * DH2[i * L.H + k];
    } else if (p < L.ow3) {
        const int k = static_cast<int>(p - L.ob2);
        for (int64_t i = 0; i < n; ++i) s += DH2[i * L.H + k];
    } else if (p < L.ob3) {
        const int k = static_cast<int>(p - L.ow3);
        for (int64_t i = 0; i < n; ++i) s += DY[i] * H2[i * L.H + k];
    } else {
        for (int64_t i = 0; i < n; ++i) s += DY[i];
    }
    G[p] = s;
}

// Sum of n values by one workgroup (fixed order per thread, then a fixed tree).
__global__ void __launch_bounds__(WG) k_sum(const float *__restrict__ v, int64_t n, float *__restrict__ out) {
    __shared__ float part[WG];
    float s = 0.0f;
    for (int64_t i = threadIdx.x; i < n; i += WG) s += v[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = WG / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = part[0] / static_cast<float>(n);
}

__global__ void __launch_bounds__(WG) k_adam(float *__restrict__ P, float *__restrict__ M, float *__restrict__ V,
                                             const float *__restrict__ G, int64_t np, float lr, float c1, float c2) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * WG + threadIdx.x;
    if (p >= np) return;
    const float g = G[p];
    const float m = 0.9f * M[p] + 0.1f * g;
    const float v = 0.999f * V[p] + 0.001f * g * g;
    M[p] = m;
    V[p] = v;
    P[p] -= lr * (m / c1) / (sqrtf(v / c2) + 1e-8f);
}

#define CHK(expr, what)                                                               \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) {                                                       \
            g_err = std::string(what) + ": " + hipGetErrorString(e_);                 \
            return 1;                                                                 \
        }                                                                             \
    } while (0)

unsigned blocks(int64_t n) { return static_cast<unsigned>((n + WG - 1) / WG); }

} // namespace

struct bhcm_model {
    Layout L;
    float *P = nullptr, *M = nullptr, *V = nullptr, *G = nullptr, *mean = nullptr, *istd = nullptr, *loss = nullptr;
    int64_t adam_t = 0;
    // Per-sample workspaces, grown on demand.
    int64_t cap = 0;
    float *X = nullptr, *T = nullptr, *Y = nullptr, *H1 = nullptr, *H2 = nullptr, *DY = nullptr, *DH1 = nullptr,
          *DH2 = nullptr, *SQ = nullptr;
    std::mutex mu;
};

namespace {
bhcm_model *g_installed = nullptr;
bhcm_model *g_installed_launch = nullptr;  // the launch-shape model (block and launch features)
std::mutex g_install_mu;

int reserve(bhcm_model *m, int64_t n) {
    if (n <= m->cap) return 0;
    float **bufs[] = {&m->X, &m->T, &m->Y, &m->H1, &m->H2, &m->DY, &m->DH1, &m->DH2, &m->SQ};
    for (float **b : bufs)
        if (*b != nullptr) {
            (void)hipFree(*b);
            *b = nullptr;
        }
    const int64_t F = m->L.F, H = m->L.H;
    CHK(hipMalloc(&m->X, sizeof(float) * n * F), "hipMalloc");
    CHK(hipMalloc(&m->T, sizeof(float) * n), "hipMalloc");
    CHK(hipMalloc(&m->Y, sizeof(float) * n), "hipMalloc");
    CHK(hipMalloc(&m->H1, sizeof(float) * n * H), "hipMalloc");
    CHK(hipMalloc(&m->H2, sizeof(float) * n * H), "hipMalloc");
    CHK(hipMalloc(&m->DY, sizeof(float) * n), "hipMalloc");
    CHK(hipMalloc(&m->DH1, sizeof(float) * n * H), "hipMalloc");
    CHK(hipMalloc(&m->DH2, sizeof(float) * n * H), "hipMalloc");
    CHK(hipMalloc(&m->SQ, sizeof(float) * n), "hipMalloc");
    m->cap = n;
    return 0;
}

// Forward + backward over n device-resident samples (X, T already copied); G and *loss on the device.
int gradient_on_device(bhcm_model *m, int64_t n) {
    const Layout &L = m->L;
    hipLaunchKernelGGL(k_forward, dim3(blocks(n)), dim3(WG), 0, 0, m->P, L, m->mean, m->istd, m->X, n, m->Y, m->H1, m->H2);
    hipLaunchKernelGGL(k_backward_sample, dim3(blocks(n)), dim3(WG), 0, 0, m->P, L, m->Y, m->T, m->H1, m->H2, n, m->DY,
                       m->DH1, m->DH2, m->SQ);
    hipLaunchKernelGGL(k_param_grad, dim3(blocks(L.np)), dim3(WG), 0, 0, L, m->mean, m->istd, m->X, m->H1, m->H2, m->DY,
                       m->DH1, m->DH2, n, m->G);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(WG), 0, 0, m->SQ, n, m->loss);
    CHK(hipGetLastError(), "cost model kernels");
    return 0;
}

uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double normal(uint64_t &s) {
    const double u1 = (static_cast<double>(splitmix(s) >> 11) + 0.5) / 9007199254740992.0;
    const double u2 = (static_cast<double>(splitmix(s) >> 11) + 0.5) / 9007199254740992.0;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}
} // namespace

extern "C" {

const char *bhcm_last_error(void) { return g_err.c_str(); }

int bhcm_create(int nfeat, int hidden, uint64_t seed, bhcm_model **out) {
    if (nfeat < 1 || nfeat > MAXF || hidden < 1 || hidden > MAXH || out == nullptr) {
        g_err = "bhcm_create: need 1 <= nfeat <= 32 and 1 <= hidden <= 64";
        return 1;
    }
    bhcm_model *m = new bhcm_model();
    m->L = make_layout(nfeat, hidden);
    const int64_t np = m->L.np;
    CHK(hipMalloc(&m->P, sizeof(float) * np), "hipMalloc");
    CHK(hipMalloc(&m->M, sizeof(float) * np), "hipMalloc");
    CHK(hipMalloc(&m->V, sizeof(float) * np), "hipMalloc");
    CHK(hipMalloc(&m->G, sizeof(float) * np), "hipMalloc");
    CHK(hipMalloc(&m->mean, sizeof(float) * nfeat), "hipMalloc");
    CHK(hipMalloc(&m->istd, sizeof(float) * nfeat), "hipMalloc");
    CHK(hipMalloc(&m->loss, sizeof(float)), "hipMalloc");
    CHK(hipMemset(m->M, 0, sizeof(float) * np), "hipMemset");
    CHK(hipMemset(m->V, 0, sizeof(float) * np), "hipMemset");
    std::vector<float> p(static_cast<size_t>(np), 0.0f), mean(nfeat, 0.0f), istd(nfeat, 1.0f);
    uint64_t s = seed;
    const Layout &L = m->L;
    for (int64_t k = 0; k < static_cast<int64_t>(nfeat) * hidden; ++k)
        p[L.oW1 + k] = static_cast<float>(normal(s) * std::sqrt(2.0 / nfeat));
    for (int64_t k = 0; k < static_cast<int64_t>(hidden) * hidden; ++k)
        p[L.oW2 + k] = static_cast<float>(normal(s) * std::sqrt(2.0 / hidden));
    for (int k = 0; k < hidden; ++k) p[L.ow3 + k] = static_cast<float>(normal(s) * std::sqrt(1.0 / hidden));
    CHK(hipMemcpy(m->P, p.data(), sizeof(float) * np, hipMemcpyHostToDevice), "hipMemcpy");
    CHK(hipMemcpy(m->mean, mean.data(), sizeof(float) * nfeat, hipMemcpyHostToDevice), "hipMemcpy");
    CHK(hipMemcpy(m->istd, istd.data(), sizeof(float) * nfeat, hipMemcpyHostToDevice), "hipMemcpy");
    *out = m;
    return 0;
}

int bhcm_destroy(bhcm_model *m) {
    if (m == nullptr) return 0;
    {
        std::lock_guard<std::mutex> lock(g_install_mu);
        if (g_installed == m) g_installed = nullptr;
        if (g_installed_launch == m) g_installed_launch = nullptr;
    }
    (void)hipDeviceSynchronize();
    float *bufs[] = {m->P, m->M, m->V, m->G, m->mean, m->istd, m->loss, m->X, m->T, m->Y, m->H1, m->H2, m->DY, m->DH1,
                     m->DH2, m->SQ};
    for (float *b : bufs)
        if (b != nullptr) (void)hipFree(b);
    delete m;
    return 0;
}

int64_t bhcm_num_params(const bhcm_model *m) { return m == nullptr ? 0 : m->L.np; }

int bhcm_get_params(bhcm_model *m, float *out) {
    std::lock_guard<std::mutex> lock(m->mu);
    CHK(hipMemcpy(out, m->P, sizeof(float) * m->L.np, hipMemcpyDeviceToHost), "hipMemcpy");
    return 0;
}

int bhcm_set_params(bhcm_model *m, const float *in) {
    std::lock_guard<std::mutex> lock(m->mu);
    CHK(hipMemcpy(m->P, in, sizeof(float) * m->L.np, hipMemcpyHostToDevice), "hipMemcpy");
    CHK(hipMemset(m->M, 0, sizeof(float) * m->L.np), "hipMemset");
    CHK(hipMemset(m->V, 0, sizeof(float) * m->L.np), "hipMemset");
    m->adam_t = 0;
    return 0;
}

int bhcm_set_normalization(bhcm_model *m, const float *mean, const float *inv_std) {
    std::lock_guard<std::mutex> lock(m->mu);
    CHK(hipMemcpy(m->mean, mean, sizeof(float) * m->L.F, hipMemcpyHostToDevice), "hipMemcpy");
    CHK(hipMemcpy(m->istd, inv_std, sizeof(float) * m->L.F, hipMemcpyHostToDevice), "hipMemcpy");
    return 0;
}

int bhcm_predict(bhcm_model *m, const float *x, int64_t n, float *y) {
    if (n <= 0) return 0;
    std::lock_guard<std::mutex> lock(m->mu);
    if (reserve(m, n) != 0) return 1;
    CHK(hipMemcpy(m->X, x, sizeof(float) * n * m->L.F, hipMemcpyHostToDevice), "hipMemcpy");
    hipLaunchKernelGGL(k_forward, dim3(blocks(n)), dim3(WG), 0, 0, m->P, m->L, m->mean, m->istd, m->X, n, m->Y,
                       static_cast<float *>(nullptr), static_cast<float *>(nullptr));
    CHK(hipGetLastError(), "k_forward");
    CHK(hipMemcpy(y, m->Y, sizeof(float) * n, hipMemcpyDeviceToHost), "hipMemcpy");
    return 0;
}

int bhcm_gradient(bhcm_model *m, const float *x, const float *t, int64_t n, float *grad, float *loss) {
    if (n <= 0) {
        g_err = "bhcm_gradient: no samples";
        return 1;
    }
    std::lock_guard<std::mutex> lock(m->mu);
    if (reserve(m, n) != 0) return 1;
    CHK(hipMemcpy(m->X, x, sizeof(float) * n * m->L.F, hipMemcpyHostToDevice), "hipMemcpy");
    CHK(hipMemcpy(m->T, t, sizeof(float) * n, hipMemcpyHostToDevice), "hipMemcpy");
    if (gradient_on_device(m, n) != 0) return 1;
    CHK(hipMemcpy(grad, m->G, sizeof(float) * m->L.np, hipMemcpyDeviceToHost), "hipMemcpy");
    if (loss != nullptr) CHK(hipMemcpy(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
    return 0;
}

int bhcm_train(bhcm_model *m, const float *x, const float *t, int64_t n, int steps, float lr, float *loss) {
    if (n <= 0 || steps < 0) {
        g_err = "bhcm_train: bad arguments";
        return 1;
    }
    std::lock_guard<std::mutex> lock(m->mu);
    if (reserve(m, n) != 0) return 1;
    CHK(hipMemcpy(m->X, x, sizeof(float) * n * m->L.F, hipMemcpyHostToDevice), "hipMemcpy");
    CHK(hipMemcpy(m->T, t, sizeof(float) * n, hipMemcpyHostToDevice), "hipMemcpy");
    for (int k = 0; k < steps; ++k) {
        if (gradient_on_device(m, n) != 0) return 1;
        ++m->adam_t;
        const float c1 = 1.0f - std::pow(0.9f, static_cast<float>(m->adam_t));
        const float c2 = 1.0f - std::pow(0.999f, static_cast<float>(m->adam_t));
        hipLaunchKernelGGL(k_adam, dim3(blocks(m->L.np)), dim3(WG), 0, 0, m->P, m->M, m->V, m->G, m->L.np, lr, c1, c2);
        CHK(hipGetLastError(), "k_adam");
        if (loss != nullptr) CHK(hipMemcpy(loss + k, m->loss, sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
    }
    CHK(hipDeviceSynchronize(), "hipDeviceSynchronize");
    return 0;
}

int bhcm_install(bhcm_model *m) {
    std::lock_guard<std::mutex> lock(g_install_mu);
    g_installed = m;
    return 0;
}

int bhcm_install_launch(bhcm_model *m) {
    std::lock_guard<std::mutex> lock(g_install_mu);
    g_installed_launch = m;
    return 0;
}

} // extern "C"

namespace bohrium {
namespace vehip {
/* Predicted log-times of n feature rows with the installed model; false when none is installed. */
bool costmodel_predict(const std::vector<float> &x, int64_t n, std::vector<float> &y) {
    std::lock_guard<std::mutex> lock(g_install_mu);
    if (g_installed == nullptr) return false;
    y.resize(static_cast<size_t>(n));
    if (n == 0) return true;
    if (static_cast<int64_t>(x.size()) != n * g_installed->L.F) return false;
    return bhcm_predict(g_installed, x.data(), n, y.data()) == 0;
}

int costmodel_features() {
    std::lock_guard<std::mutex> lock(g_install_mu);
    return g_installed == nullptr ? 0 : g_installed->L.F;
}

bool launchmodel_predict(const std::vector<float> &x, int64_t n, std::vector<float> &y) {
    std::lock_guard<std::mutex> lock(g_install_mu);
    if (g_installed_launch == nullptr) return false;
    y.resize(static_cast<size_t>(n));
    if (n == 0) return true;
    if (static_cast<int64_t>(x.size()) != n * g_installed_launch->L.F) return false;
    return bhcm_predict(g_installed_launch, x.data(), n, y.data()) == 0;
}

int launchmodel_features() {
    std::lock_guard<std::mutex> lock(g_install_mu);
    return g_installed_launch == nullptr ? 0 : g_installed_launch->L.F;
}
} // namespace vehip
} // namespace bohrium
