// Learned first iterate of the strong fluid-structure coupling loop: a small
// MLP (no counterpart in the reference, whose coupling is loose; the
// iterated coupling the reference prepares for is declarations.F90:250,
// 271-278) that maps the recent history of the body (displacement, velocity,
// loads) to the next step's loads.  The coupling iteration still decides the result:
// the network only proposes where it starts.
//
// Network: NI inputs -> H tanh -> H tanh -> NO linear outputs (H = 64), FP32
// parameters laid out as [W1 (H x NI, row-major), b1 (H), W2 (H x H), b2 (H),
// W3 (NO x H), b3 (NO)].
//
// Training (FP32): one step over a minibatch of B samples is three launches:
//   k_mlp_fwd   one workgroup per sample: the two hidden layers and the
//               outputs, kept for the backward pass;
//   k_mlp_bwd   one workgroup per sample: output error of the mean squared
//               loss and the back-propagated errors of both hidden layers;
//   k_mlp_adam  one thread per parameter: its gradient summed over the
//               minibatch in sample order (deterministic) and the Adam update.
// Inference (BF16): k_mlp_infer rounds inputs, weights and hidden
// activations to bfloat16 and accumulates in FP32; one workgroup per sample.
#include "fribo_internal.h"

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#define MLP_H 64

struct MlpDims {
    int ni, no;
    int w1, b1, w2, b2, w3, b3, np;   // parameter offsets and count
};

static MlpDims mlp_dims(int ni, int no)
{
    MlpDims d;
    d.ni = ni;
    d.no = no;
    d.w1 = 0;
    d.b1 = d.w1 + MLP_H * ni;
    d.w2 = d.b1 + MLP_H;
    d.b2 = d.w2 + MLP_H * MLP_H;
    d.w3 = d.b2 + MLP_H;
    d.b3 = d.w3 + no * MLP_H;
    d.np = d.b3 + no;
    return d;
}

// bfloat16 rounding (round to nearest even) of a finite float
__device__ inline float bf16r(float x)
{
    unsigned u = __float_as_uint(x);
    u = u + 0x7fffu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xffff0000u);
}

__global__ __launch_bounds__(MLP_H) void k_mlp_fwd(MlpDims d, const float* __restrict__ p, const float* __restrict__ X,
                                                   float* __restrict__ H1, float* __restrict__ H2,
                                                   float* __restrict__ Y)
{
    __shared__ float xs[256], h1[MLP_H], h2[MLP_H];
    const int b = blockIdx.x, u = threadIdx.x;
    for (int i = u; i < d.ni; i += MLP_H) xs[i] = X[(size_t)b * d.ni + i];
    __syncthreads();
    float s = p[d.b1 + u];
    for (int i = 0; i < d.ni; ++i) s += p[d.w1 + u * d.ni + i] * xs[i];
    h1[u] = tanhf(s);
    H1[(size_t)b * MLP_H + u] = h1[u];
    __syncthreads();
    s = p[d.b2 + u];
    for (int i = 0; i < MLP_H; ++i) s += p[d.w2 + u * MLP_H + i] * h1[i];
    h2[u] = tanhf(s);
    H2[(size_t)b * MLP_H + u] = h2[u];
    __syncthreads();
    if (u < d.no) {
        s = p[d.b3 + u];
        for (int i = 0; i < MLP_H; ++i) s += p[d.w3 + u * MLP_H + i] * h2[i];
        Y[(size_t)b * d.no + u] = s;
    }
}

// Errors of the loss L = (1 / (B NO)) sum (Y - T)^2 with respect to the
// pre-activations: D3 = dL/dY, D2 and D1 of the hidden layers.
__global__ __launch_bounds__(MLP_H) void k_mlp_bwd(MlpDims d, const float* __restrict__ p, int B,
                                                   const float* __restrict__ T, const float* __restrict__ H1,
                                                   const float* __restrict__ H2, const float* __restrict__ Y,
                                                   float* __restrict__ D1, float* __restrict__ D2,
                                                   float* __restrict__ D3)
{
    __shared__ float d3[16], d2[MLP_H];
    const int b = blockIdx.x, u = threadIdx.x;
    const float sc = 2.0f / (float)(B * d.no);
    if (u < d.no) {
        d3[u] = sc * (Y[(size_t)b * d.no + u] - T[(size_t)b * d.no + u]);
        D3[(size_t)b * d.no + u] = d3[u];
    }
    __syncthreads();
    float s = 0.0f;
    for (int o = 0; o < d.no; ++o) s += p[d.w3 + o * MLP_H + u] * d3[o];
    const float h2 = H2[(size_t)b * MLP_H + u];
    d2[u] = s * (1.0f - h2 * h2);
    D2[(size_t)b * MLP_H + u] = d2[u];
    __syncthreads();
    s = 0.0f;
    for (int o = 0; o < MLP_H; ++o) s += p[d.w2 + o * MLP_H + u] * d2[o];
    const float h1 = H1[(size_t)b * MLP_H + u];
    D1[(size_t)b * MLP_H + u] = s * (1.0f - h1 * h1);
}

// Gradient of parameter q (summed over the B samples in order) and, when
// lr > 0, the Adam step t (bias-corrected moments); grad (optional) receives
// the gradient.
__device__ inline float mlp_grad(const MlpDims& d, int q, int B, const float* X, const float* H1, const float* H2,
                                 const float* D1, const float* D2, const float* D3)
{
    float g = 0.0f;
    if (q < d.b1) {
        const int u = (q - d.w1) / d.ni, i = (q - d.w1) % d.ni;
        for (int b = 0; b < B; ++b) g += D1[(size_t)b * MLP_H + u] * X[(size_t)b * d.ni + i];
    } else if (q < d.w2) {
        const int u = q - d.b1;
        for (int b = 0; b < B; ++b) g += D1[(size_t)b * MLP_H + u];
    } else if (q < d.b2) {
        const int u = (q - d.w2) / MLP_H, i = (q - d.w2) % MLP_H;
        for (int b = 0; b < B; ++b) g += D2[(size_t)b * MLP_H + u] * H1[(size_t)b * MLP_H + i];
    } else if (q < d.w3) {
        const int u = q - d.b2;
        for (int b = 0; b < B; ++b) g += D2[(size_t)b * MLP_H + u];
    } else if (q < d.b3) {
        const int o = (q - d.w3) / MLP_H, i = (q - d.w3) % MLP_H;
        for (int b = 0; b < B; ++b) g += D3[(size_t)b * d.no + o] * H2[(size_t)b * MLP_H + i];
    } else {
        const int o = q - d.b3;
        for (int b = 0; b < B; ++b) g += D3[(size_t)b * d.no + o];
    }
    return g;
}

__global__ __launch_bounds__(256) void k_mlp_adam(MlpDims d, float* __restrict__ p, float* __restrict__ m1,
                                                  float* __restrict__ m2, int B, const float* __restrict__ X,
                                                  const float* __restrict__ H1, const float* __restrict__ H2,
                                                  const float* __restrict__ D1, const float* __restrict__ D2,
                                                  const float* __restrict__ D3, float lr, float c1, float c2,
                                                  float* __restrict__ grad)
{
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= d.np) return;
    const float g = mlp_grad(d, q, B, X, H1, H2, D1, D2, D3);
    if (grad) grad[q] = g;
    if (lr <= 0.0f) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float a = b1 * m1[q] + (1.0f - b1) * g;
    const float v = b2 * m2[q] + (1.0f - b2) * g * g;
    m1[q] = a;
    m2[q] = v;
    p[q] -= lr * (a * c1) / (sqrtf(v * c2) + eps);
}

__global__ __launch_bounds__(MLP_H) void k_mlp_infer(MlpDims d, const float* __restrict__ p,
                                                     const float* __restrict__ X, float* __restrict__ Y)
{
    __shared__ float xs[256], h1[MLP_H], h2[MLP_H];
    const int b = blockIdx.x, u = threadIdx.x;
    for (int i = u; i < d.ni; i += MLP_H) xs[i] = bf16r(X[(size_t)b * d.ni + i]);
    __syncthreads();
    float s = bf16r(p[d.b1 + u]);
    for (int i = 0; i < d.ni; ++i) s += bf16r(p[d.w1 + u * d.ni + i]) * xs[i];
    h1[u] = bf16r(tanhf(s));
    __syncthreads();
    s = bf16r(p[d.b2 + u]);
    for (int i = 0; i < MLP_H; ++i) s += bf16r(p[d.w2 + u * MLP_H + i]) * h1[i];
    h2[u] = bf16r(tanhf(s));
    __syncthreads();
    if (u < d.no) {
        s = bf16r(p[d.b3 + u]);
        for (int i = 0; i < MLP_H; ++i) s += bf16r(p[d.w3 + u * MLP_H + i]) * h2[i];
        Y[(size_t)b * d.no + u] = s;
    }
}

// ---------------------------------------------------------------------------
// C ABI
struct fribo_mlp_ctx {
    MlpDims d;
    hipStream_t st = nullptr;
    float *p = nullptr, *m1 = nullptr, *m2 = nullptr;
    int cap = 0;    // samples the work buffers hold
    float *X = nullptr, *T = nullptr, *H1 = nullptr, *H2 = nullptr, *Y = nullptr, *D1 = nullptr, *D2 = nullptr,
          *D3 = nullptr, *G = nullptr;
    long step = 0;  // Adam steps taken
};
typedef fribo_mlp_ctx* fribo_mlp;

static void mlp_free_work(fribo_mlp m)
{
    float** w[] = {&m->X, &m->T, &m->H1, &m->H2, &m->Y, &m->D1, &m->D2, &m->D3};
    for (auto q : w) {
        if (*q) hipFree(*q);
        *q = nullptr;
    }
    m->cap = 0;
}

static int mlp_reserve(fribo_mlp m, int n)
{
    if (n <= m->cap) return 0;
    mlp_free_work(m);
    const size_t ni = m->d.ni, no = m->d.no, H = MLP_H;
    const size_t sz[] = {ni, no, H, H, no, H, H, no};
    float** w[] = {&m->X, &m->T, &m->H1, &m->H2, &m->Y, &m->D1, &m->D2, &m->D3};
    for (int i = 0; i < 8; ++i)
        if (hipMalloc(w[i], (size_t)n * sz[i] * sizeof(float)) != hipSuccess) {
            mlp_free_work(m);
            return FRIBO_E_HIP;
        }
    m->cap = n;
    return 0;
}

// Network with ni inputs (<= 256) and no outputs (<= 16); parameters from
// init (np floats, as fribo_mlp_size reports) or, when init is NULL,
// Glorot-uniform weights from the seed and zero biases.
extern "C" int fribo_mlp_create(int32_t ni, int32_t no, const float* init, uint64_t seed, fribo_mlp* out)
{
    if (!out || ni < 1 || ni > 256 || no < 1 || no > 16) return FRIBO_E_ARG;
    fribo_mlp m = new fribo_mlp_ctx();
    m->d = mlp_dims(ni, no);
    std::vector<float> p0(m->d.np, 0.0f);
    if (init) {
        std::copy(init, init + m->d.np, p0.begin());
    } else {
        uint64_t s = seed ? seed : 0x9e3779b97f4a7c15ull;
        auto uni = [&s]() {   // splitmix64 -> [-1, 1)
            uint64_t z = (s += 0x9e3779b97f4a7c15ull);
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            z = z ^ (z >> 31);
            return (float)((double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0);
        };
        const float a1 = std::sqrt(6.0f / (ni + MLP_H)), a2 = std::sqrt(6.0f / (2 * MLP_H)),
                    a3 = std::sqrt(6.0f / (MLP_H + no));
        for (int q = m->d.w1; q < m->d.b1; ++q) p0[q] = a1 * uni();
        for (int q = m->d.w2; q < m->d.b2; ++q) p0[q] = a2 * uni();
        for (int q = m->d.w3; q < m->d.b3; ++q) p0[q] = a3 * uni();
    }
    const size_t nb = (size_t)m->d.np * sizeof(float);
    if (hipStreamCreate(&m->st) != hipSuccess || hipMalloc(&m->p, nb) != hipSuccess ||
        hipMalloc(&m->m1, nb) != hipSuccess || hipMalloc(&m->m2, nb) != hipSuccess ||
        hipMalloc(&m->G, nb) != hipSuccess || hipMemcpy(m->p, p0.data(), nb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(m->m1, 0, nb) != hipSuccess || hipMemset(m->m2, 0, nb) != hipSuccess) {
        if (m->p) hipFree(m->p);
        if (m->m1) hipFree(m->m1);
        if (m->m2) hipFree(m->m2);
        if (m->G) hipFree(m->G);
        if (m->st) hipStreamDestroy(m->st);
        delete m;
        return FRIBO_E_HIP;
    }
    *out = m;
    return 0;
}

extern "C" int fribo_mlp_destroy(fribo_mlp m)
{
    if (!m) return FRIBO_E_ARG;
    hipStreamSynchronize(m->st);
    mlp_free_work(m);
    hipFree(m->p);
    hipFree(m->m1);
    hipFree(m->m2);
    hipFree(m->G);
    hipStreamDestroy(m->st);
    delete m;
    return 0;
}

extern "C" int fribo_mlp_size(fribo_mlp m)
{
    return m ? m->d.np : -1;
}

extern "C" int fribo_mlp_get_params(fribo_mlp m, float* out)
{
    if (!m || !out) return FRIBO_E_ARG;
    if (hipMemcpy(out, m->p, (size_t)m->d.np * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        return FRIBO_E_HIP;
    return 0;
}

// One pass over n samples (X n x ni, T n x no, row-major): forward, backward
// and, with lr > 0, one Adam step of the whole set as the minibatch.  grad
// (optional, np floats) receives the gradient of the mean squared loss,
// loss (optional) its value.
static int mlp_pass(fribo_mlp m, int n, const float* X, const float* T, float lr, float* grad, float* loss)
{
    if (!m || n < 1 || !X || !T) return FRIBO_E_ARG;
    if (mlp_reserve(m, n)) return FRIBO_E_HIP;
    const MlpDims& d = m->d;
    if (hipMemcpyAsync(m->X, X, (size_t)n * d.ni * sizeof(float), hipMemcpyHostToDevice, m->st) != hipSuccess ||
        hipMemcpyAsync(m->T, T, (size_t)n * d.no * sizeof(float), hipMemcpyHostToDevice, m->st) != hipSuccess)
        return FRIBO_E_HIP;
    hipLaunchKernelGGL(k_mlp_fwd, dim3(n), dim3(MLP_H), 0, m->st, d, (const float*)m->p, (const float*)m->X, m->H1,
                       m->H2, m->Y);
    hipLaunchKernelGGL(k_mlp_bwd, dim3(n), dim3(MLP_H), 0, m->st, d, (const float*)m->p, n, (const float*)m->T,
                       (const float*)m->H1, (const float*)m->H2, (const float*)m->Y, m->D1, m->D2, m->D3);
    float c1 = 1.0f, c2 = 1.0f;
    if (lr > 0.0f) {
        ++m->step;
        c1 = 1.0f / (1.0f - std::pow(0.9f, (float)m->step));
        c2 = 1.0f / (1.0f - std::pow(0.999f, (float)m->step));
    }
    hipLaunchKernelGGL(k_mlp_adam, dim3((d.np + 255) / 256), dim3(256), 0, m->st, d, m->p, m->m1, m->m2, n,
                       (const float*)m->X, (const float*)m->H1, (const float*)m->H2, (const float*)m->D1,
                       (const float*)m->D2, (const float*)m->D3, lr, c1, c2, grad ? m->G : (float*)nullptr);
    if (hipGetLastError() != hipSuccess) return FRIBO_E_HIP;
    if (grad && hipMemcpyAsync(grad, m->G, (size_t)d.np * sizeof(float), hipMemcpyDeviceToHost, m->st) != hipSuccess)
        return FRIBO_E_HIP;
    std::vector<float> y;
    if (loss) {
        y.resize((size_t)n * d.no);
        if (hipMemcpyAsync(y.data(), m->Y, y.size() * sizeof(float), hipMemcpyDeviceToHost, m->st) != hipSuccess)
            return FRIBO_E_HIP;
    }
    if (hipStreamSynchronize(m->st) != hipSuccess) return FRIBO_E_HIP;
    if (loss) {
        double s = 0.0;
        for (size_t i = 0; i < y.size(); ++i) {
            const double e = (double)y[i] - (double)T[i];
            s += e * e;
        }
        *loss = (float)(s / (double)y.size());
    }
    return 0;
}

// Gradient of the mean squared loss over the n samples (no update).
extern "C" int fribo_mlp_grad(fribo_mlp m, int32_t n, const float* X, const float* T, float* grad, float* loss)
{
    return mlp_pass(m, n, X, T, 0.0f, grad, loss);
}

// Adam training: epochs passes over the n samples in minibatches of
// `batch` samples (in order; the caller shuffles), learning rate lr.
// loss (optional) gets the mean squared loss of the last minibatch.
extern "C" int fribo_mlp_train(fribo_mlp m, int32_t n, const float* X, const float* T, int32_t epochs,
                               int32_t batch, float lr, float* loss)
{
    if (!m || n < 1 || epochs < 1 || batch < 1 || !(lr > 0.0f)) return FRIBO_E_ARG;
    for (int e = 0; e < epochs; ++e)
        for (int s = 0; s < n; s += batch) {
            const int b = std::min(batch, n - s);
            const bool last = e == epochs - 1 && s + batch >= n;
            int rc = mlp_pass(m, b, X + (size_t)s * m->d.ni, T + (size_t)s * m->d.no, lr, nullptr,
                              last ? loss : nullptr);
            if (rc) return rc;
        }
    return 0;
}

// Inference in BF16 (FP32 accumulation) of n samples: Y = net(X).
extern "C" int fribo_mlp_predict(fribo_mlp m, int32_t n, const float* X, float* Y)
{
    if (!m || n < 1 || !X || !Y) return FRIBO_E_ARG;
    if (mlp_reserve(m, n)) return FRIBO_E_HIP;
    const MlpDims& d = m->d;
    if (hipMemcpyAsync(m->X, X, (size_t)n * d.ni * sizeof(float), hipMemcpyHostToDevice, m->st) != hipSuccess)
        return FRIBO_E_HIP;
    hipLaunchKernelGGL(k_mlp_infer, dim3(n), dim3(MLP_H), 0, m->st, d, (const float*)m->p, (const float*)m->X, m->Y);
    if (hipGetLastError() != hipSuccess) return FRIBO_E_HIP;
    if (hipMemcpyAsync(Y, m->Y, (size_t)n * d.no * sizeof(float), hipMemcpyDeviceToHost, m->st) != hipSuccess ||
        hipStreamSynchronize(m->st) != hipSuccess)
        return FRIBO_E_HIP;
    return 0;
}
