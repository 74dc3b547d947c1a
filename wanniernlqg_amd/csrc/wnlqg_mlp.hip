// Learned refinement proposal: an 8 -> 64 -> 64 -> 1 SiLU multilayer perceptron in FP32 with
// hand-written training and inference kernels.  It has no counterpart in the reference (which has no
// learned component); it ranks integration cells of a k mesh for adaptive refinement from per-cell
// features of the base mesh, trained on the refinement gains of cheaper coarse runs.  The integrand
// itself is always computed exactly.
//
// Parameter vector (float, WNLQG_MLP_PARAMS = 4801):
//   W1[64][8] | b1[64] | W2[64][64] | b2[64] | W3[64] | b3
// Loss: mean squared error over the samples.  Gradients are deterministic: every wave owns a fixed
// contiguous sample range and keeps its partial gradient in registers; partial gradients are summed
// in wave order by a second kernel.  Training uses Adam.
#include "../../include/wnlqg.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#define MLP_H 64
#define MLP_IN WNLQG_MLP_INPUTS
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + MLP_H * MLP_IN)
#define OFF_W2 (OFF_B1 + MLP_H)
#define OFF_B2 (OFF_W2 + MLP_H * MLP_H)
#define OFF_W3 (OFF_B2 + MLP_H)
#define OFF_B3 (OFF_W3 + MLP_H)
#define MLP_NP (OFF_B3 + 1)
static_assert(MLP_NP == WNLQG_MLP_PARAMS, "parameter count");

#define MLP_WAVES 4
#define MLP_GRAD_BLOCKS 128

namespace {

__device__ inline float sigmoidf_(float z) { return 1.0f / (1.0f + expf(-z)); }

}  // namespace

// Per-wave partial gradient of sum_s scale * (y_s - t_s)^2 over the wave's sample range.
// partial[wave][MLP_NP], loss[wave] = sum (y - t)^2 over the range.
__global__ __launch_bounds__(64 * MLP_WAVES) void k_mlp_grad(const float *__restrict__ P, const float *__restrict__ x,
                                                             const float *__restrict__ t, long long n, long long per_wave,
                                                             float scale, float *__restrict__ partial,
                                                             float *__restrict__ loss) {
    __shared__ float W2s[MLP_H * (MLP_H + 1)];
    for (int e = threadIdx.x; e < MLP_H * MLP_H; e += blockDim.x) {
        const int j = e / MLP_H, i = e - (e / MLP_H) * MLP_H;
        W2s[j * (MLP_H + 1) + i] = P[OFF_W2 + e];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const long long wave = (long long)blockIdx.x * MLP_WAVES + (threadIdx.x >> 6);
    float w1[MLP_IN], gw1[MLP_IN];
#pragma unroll
    for (int c = 0; c < MLP_IN; ++c) {
        w1[c] = P[OFF_W1 + lane * MLP_IN + c];
        gw1[c] = 0.0f;
    }
    const float b1 = P[OFF_B1 + lane], b2 = P[OFF_B2 + lane], w3 = P[OFF_W3 + lane], b3 = P[OFF_B3];
    float gb1 = 0.0f, gb2 = 0.0f, gw3 = 0.0f, gb3 = 0.0f, lsum = 0.0f;
    float gw2[MLP_H];
#pragma unroll
    for (int i = 0; i < MLP_H; ++i) gw2[i] = 0.0f;
    const long long s0 = wave * per_wave, s1 = s0 + per_wave < n ? s0 + per_wave : n;
    for (long long s = s0; s < s1; ++s) {
        float xs[MLP_IN];
        float z1 = b1;
#pragma unroll
        for (int c = 0; c < MLP_IN; ++c) {
            xs[c] = x[MLP_IN * s + c];
            z1 += w1[c] * xs[c];
        }
        const float s1g = sigmoidf_(z1);
        const float h1 = z1 * s1g;
        float z2 = b2;
#pragma unroll
        for (int i = 0; i < MLP_H; ++i) z2 += W2s[lane * (MLP_H + 1) + i] * __shfl(h1, i);
        const float s2g = sigmoidf_(z2);
        const float h2 = z2 * s2g;
        float y = w3 * h2;
        for (int o = 32; o > 0; o >>= 1) y += __shfl_xor(y, o);
        y += b3;
        const float r = y - t[s];
        lsum += r * r;
        const float d = scale * r;
        gb3 += d;
        gw3 += d * h2;
        const float dz2 = d * w3 * (s2g * (1.0f + z2 * (1.0f - s2g)));
        gb2 += dz2;
#pragma unroll
        for (int i = 0; i < MLP_H; ++i) gw2[i] += dz2 * __shfl(h1, i);
        float dh1 = 0.0f;
#pragma unroll
        for (int j = 0; j < MLP_H; ++j) dh1 += W2s[j * (MLP_H + 1) + lane] * __shfl(dz2, j);
        const float dz1 = dh1 * (s1g * (1.0f + z1 * (1.0f - s1g)));
        gb1 += dz1;
#pragma unroll
        for (int c = 0; c < MLP_IN; ++c) gw1[c] += dz1 * xs[c];
    }
    float *g = partial + wave * MLP_NP;
#pragma unroll
    for (int c = 0; c < MLP_IN; ++c) g[OFF_W1 + lane * MLP_IN + c] = gw1[c];
    g[OFF_B1 + lane] = gb1;
#pragma unroll
    for (int i = 0; i < MLP_H; ++i) g[OFF_W2 + lane * MLP_H + i] = gw2[i];
    g[OFF_B2 + lane] = gb2;
    g[OFF_W3 + lane] = gw3;
    if (lane == 0) {
        g[OFF_B3] = gb3;
        loss[wave] = lsum;
    }
}

// grad[p] = sum over waves in order of partial[wave][p]; grad[MLP_NP] = sum of the wave losses.
__global__ void k_mlp_reduce(const float *__restrict__ partial, const float *__restrict__ loss, int nwaves,
                             float *__restrict__ grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < MLP_NP) {
        float s = 0.0f;
        for (int w = 0; w < nwaves; ++w) s += partial[(size_t)w * MLP_NP + p];
        grad[p] = s;
    } else if (p == MLP_NP) {
        float s = 0.0f;
        for (int w = 0; w < nwaves; ++w) s += loss[w];
        grad[p] = s;
    }
}

// Adam step (Kingma & Ba 2015) with bias correction factors c1 = 1 - beta1^t, c2 = 1 - beta2^t.
__global__ void k_mlp_adam(float *__restrict__ P, float *__restrict__ m, float *__restrict__ v,
                           const float *__restrict__ grad, float lr, float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NP) return;
    const float g = grad[p];
    const float mm = 0.9f * m[p] + 0.1f * g;
    const float vv = 0.999f * v[p] + 0.001f * g * g;
    m[p] = mm;
    v[p] = vv;
    P[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + 1.0e-8f);
}

// Inference: one thread per sample, W2 in LDS.
__global__ __launch_bounds__(256) void k_mlp_predict(const float *__restrict__ P, const float *__restrict__ x,
                                                     long long n, float *__restrict__ y) {
    __shared__ float W2s[MLP_H * (MLP_H + 1)];
    __shared__ float Ps[OFF_W2];
    __shared__ float tail[MLP_NP - OFF_B2];
    for (int e = threadIdx.x; e < MLP_H * MLP_H; e += blockDim.x) {
        const int j = e / MLP_H, i = e - (e / MLP_H) * MLP_H;
        W2s[j * (MLP_H + 1) + i] = P[OFF_W2 + e];
    }
    for (int e = threadIdx.x; e < OFF_W2; e += blockDim.x) Ps[e] = P[e];
    for (int e = threadIdx.x; e < MLP_NP - OFF_B2; e += blockDim.x) tail[e] = P[OFF_B2 + e];
    __syncthreads();
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float xs[MLP_IN];
#pragma unroll
    for (int c = 0; c < MLP_IN; ++c) xs[c] = x[MLP_IN * s + c];
    float h1[MLP_H];
#pragma unroll
    for (int i = 0; i < MLP_H; ++i) {
        float z = Ps[OFF_B1 + i];
#pragma unroll
        for (int c = 0; c < MLP_IN; ++c) z += Ps[OFF_W1 + MLP_IN * i + c] * xs[c];
        h1[i] = z * sigmoidf_(z);
    }
    float out = tail[OFF_B3 - OFF_B2];
    for (int j = 0; j < MLP_H; ++j) {
        float z = tail[j];
#pragma unroll
        for (int i = 0; i < MLP_H; ++i) z += W2s[j * (MLP_H + 1) + i] * h1[i];
        out += tail[OFF_W3 - OFF_B2 + j] * (z * sigmoidf_(z));
    }
    y[s] = out;
}

namespace {

thread_local std::string g_mlp_error;

int mlp_fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_mlp_error = buf;
    return code;
}

#define MLP_HIP(expr)                                                                                  \
    do {                                                                                               \
        hipError_t _e = (expr);                                                                        \
        if (_e != hipSuccess)                                                                          \
            return mlp_fail(_e == hipErrorOutOfMemory ? WNLQG_E_OOM : WNLQG_E_HIP, "%s failed: %s", #expr, \
                            hipGetErrorString(_e));                                                    \
    } while (0)

template <typename T>
struct Buf {
    T *p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t count) {
        if (count <= n) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
        if (e == hipSuccess) n = count;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

}  // namespace

struct wnlqg_mlp {
    hipStream_t stream = nullptr;
    int device = 0;
    Buf<float> P, m, v, grad, partial, loss, x, t, y;
    long long step = 0;
};

// Scratch of wnlqg_mlp_gradient_noalloc: the per-wave partial gradients [MLP_GRAD_BLOCKS * MLP_WAVES][MLP_NP]
// followed by the per-wave losses (WNLQG_MLP_NOALLOC_SCRATCH_BYTES, about 9.2 MB).  Every GPU the library runs
// on holds its own copy until the process ends.  Host code never uses the name `scratch` as an address: the
// device address comes from hipGetSymbolAddress, once per device.
#define MLP_SCRATCH_FLOATS ((size_t)MLP_GRAD_BLOCKS * MLP_WAVES * (MLP_NP + 1))
static_assert(MLP_SCRATCH_FLOATS * sizeof(float) == WNLQG_MLP_NOALLOC_SCRATCH_BYTES, "scratch size");
static __device__ __align__(256) float scratch[MLP_SCRATCH_FLOATS];

namespace {

#define MLP_MAX_DEVICES 64
std::mutex g_scratch_mutex;
float *g_scratch_ptr[MLP_MAX_DEVICES] = {};

// device address of `scratch` on `device` (which must be current): resolved on the first call per device
int scratch_address(int device, float **out) {
    if (device < 0 || device >= MLP_MAX_DEVICES) return mlp_fail(WNLQG_E_UNSUPPORTED, "device %d out of range", device);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return mlp_fail(WNLQG_E_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", device,
                            hipGetErrorString(e));
        g_scratch_ptr[device] = (float *)ptr;
    }
    *out = g_scratch_ptr[device];
    return WNLQG_OK;
}

}  // namespace

namespace {

int upload_samples(wnlqg_mlp *g, const float *x, const float *t, long long n) {
    MLP_HIP(g->x.ensure((size_t)MLP_IN * n));
    MLP_HIP(hipMemcpy(g->x.p, x, (size_t)MLP_IN * n * sizeof(float), hipMemcpyHostToDevice));
    if (t) {
        MLP_HIP(g->t.ensure((size_t)n));
        MLP_HIP(hipMemcpy(g->t.p, t, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
    }
    return WNLQG_OK;
}

// full gradient of the mean squared error over samples [0, n) already on the device, into grad[MLP_NP + 1]
// (the sum of squared errors last); partial: [nwaves][MLP_NP] per-wave gradients, loss: [nwaves]
int launch_gradient_to(wnlqg_mlp *g, const float *x, const float *t, long long n, float *partial, float *loss,
                       float *grad, hipStream_t st) {
    const int nwaves = MLP_GRAD_BLOCKS * MLP_WAVES;
    const long long per_wave = (n + nwaves - 1) / nwaves;
    hipLaunchKernelGGL(k_mlp_grad, dim3(MLP_GRAD_BLOCKS), dim3(64 * MLP_WAVES), 0, st, g->P.p, x, t, n, per_wave,
                       2.0f / (float)n, partial, loss);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NP + 1 + 255) / 256), dim3(256), 0, st, partial, loss, nwaves, grad);
    MLP_HIP(hipGetLastError());
    return WNLQG_OK;
}

int launch_gradient(wnlqg_mlp *g, const float *x, const float *t, long long n) {
    return launch_gradient_to(g, x, t, n, g->partial.p, g->loss.p, g->grad.p, g->stream);
}

}  // namespace

extern "C" {

int32_t wnlqg_mlp_last_error(char *buf, size_t len) {
    if (buf && len) {
        size_t c = std::min(len - 1, g_mlp_error.size());
        memcpy(buf, g_mlp_error.data(), c);
        buf[c] = 0;
    }
    return (int32_t)g_mlp_error.size();
}

int32_t wnlqg_mlp_create(uint64_t seed, wnlqg_mlp **out) {
    if (!out) return mlp_fail(WNLQG_E_INVALID, "out is NULL");
    *out = nullptr;
    wnlqg_mlp *g = new wnlqg_mlp();
    auto bail = [&](int code) {
        wnlqg_mlp_destroy(g);
        return code;
    };
    if (hipGetDevice(&g->device) != hipSuccess) return bail(mlp_fail(WNLQG_E_HIP, "no HIP device"));
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(mlp_fail(WNLQG_E_HIP, "stream creation failed"));
    const int nwaves = MLP_GRAD_BLOCKS * MLP_WAVES;
    if (g->P.ensure(MLP_NP) || g->m.ensure(MLP_NP) || g->v.ensure(MLP_NP) || g->grad.ensure(MLP_NP + 1) ||
        g->partial.ensure((size_t)nwaves * MLP_NP) || g->loss.ensure(nwaves))
        return bail(mlp_fail(WNLQG_E_OOM, "device allocation failed"));
    // Glorot-uniform weights, zero biases, from a seeded 64-bit Mersenne twister
    std::vector<float> p(MLP_NP, 0.0f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int fan_in, int fan_out, int count) {
        const double a = std::sqrt(6.0 / (fan_in + fan_out));
        std::uniform_real_distribution<double> u(-a, a);
        for (int i = 0; i < count; ++i) p[off + i] = (float)u(rng);
    };
    fill(OFF_W1, MLP_IN, MLP_H, MLP_H * MLP_IN);
    fill(OFF_W2, MLP_H, MLP_H, MLP_H * MLP_H);
    fill(OFF_W3, MLP_H, 1, MLP_H);
    std::vector<float> zero(MLP_NP, 0.0f);
    if (hipMemcpy(g->P.p, p.data(), MLP_NP * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->m.p, zero.data(), MLP_NP * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(g->v.p, zero.data(), MLP_NP * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return bail(mlp_fail(WNLQG_E_HIP, "parameter upload failed"));
    *out = g;
    return WNLQG_OK;
}

int32_t wnlqg_mlp_destroy(wnlqg_mlp *g) {
    if (!g) return WNLQG_OK;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    g->P.release(); g->m.release(); g->v.release(); g->grad.release(); g->partial.release(); g->loss.release();
    g->x.release(); g->t.release(); g->y.release();
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
    return WNLQG_OK;
}

int32_t wnlqg_mlp_get_params(wnlqg_mlp *g, float *params) {
    if (!g || !params) return mlp_fail(WNLQG_E_INVALID, "invalid arguments");
    MLP_HIP(hipStreamSynchronize(g->stream));
    MLP_HIP(hipMemcpy(params, g->P.p, MLP_NP * sizeof(float), hipMemcpyDeviceToHost));
    return WNLQG_OK;
}

int32_t wnlqg_mlp_set_params(wnlqg_mlp *g, const float *params) {
    if (!g || !params) return mlp_fail(WNLQG_E_INVALID, "invalid arguments");
    for (int i = 0; i < MLP_NP; ++i)
        if (!std::isfinite(params[i])) return mlp_fail(WNLQG_E_NONFINITE, "non-finite parameter %d", i);
    MLP_HIP(hipStreamSynchronize(g->stream));
    MLP_HIP(hipMemcpy(g->P.p, params, MLP_NP * sizeof(float), hipMemcpyHostToDevice));
    std::vector<float> zero(MLP_NP, 0.0f);
    MLP_HIP(hipMemcpy(g->m.p, zero.data(), MLP_NP * sizeof(float), hipMemcpyHostToDevice));
    MLP_HIP(hipMemcpy(g->v.p, zero.data(), MLP_NP * sizeof(float), hipMemcpyHostToDevice));
    g->step = 0;
    return WNLQG_OK;
}

int32_t wnlqg_mlp_predict(wnlqg_mlp *g, const float *x, int64_t n, float *y) {
    if (!g || !x || !y || n < 0) return mlp_fail(WNLQG_E_INVALID, "invalid arguments");
    if (n == 0) return WNLQG_OK;
    int rc = upload_samples(g, x, nullptr, n);
    if (rc) return rc;
    MLP_HIP(g->y.ensure((size_t)n));
    hipLaunchKernelGGL(k_mlp_predict, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, g->stream, g->P.p, g->x.p,
                       (long long)n, g->y.p);
    MLP_HIP(hipGetLastError());
    MLP_HIP(hipStreamSynchronize(g->stream));
    MLP_HIP(hipMemcpy(y, g->y.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
    return WNLQG_OK;
}

int32_t wnlqg_mlp_gradient(wnlqg_mlp *g, const float *x, const float *t, int64_t n, float *grad, float *loss) {
    if (!g || !x || !t || !grad || n < 1) return mlp_fail(WNLQG_E_INVALID, "invalid arguments");
    int rc = upload_samples(g, x, t, n);
    if (rc) return rc;
    rc = launch_gradient(g, g->x.p, g->t.p, n);
    if (rc) return rc;
    std::vector<float> h(MLP_NP + 1);
    MLP_HIP(hipStreamSynchronize(g->stream));
    MLP_HIP(hipMemcpy(h.data(), g->grad.p, (MLP_NP + 1) * sizeof(float), hipMemcpyDeviceToHost));
    memcpy(grad, h.data(), MLP_NP * sizeof(float));
    if (loss) *loss = h[MLP_NP] / (float)n;
    return WNLQG_OK;
}

int32_t wnlqg_mlp_gradient_noalloc(wnlqg_mlp *g, const float *x, const float *t, int64_t n, float *grad,
                                   void *stream) {
    if (!g || !x || !t || !grad || n < 1) return mlp_fail(WNLQG_E_INVALID, "invalid arguments");
    if (n > WNLQG_MLP_NOALLOC_MAX_SAMPLES)
        return mlp_fail(WNLQG_E_INVALID, "%lld samples exceed the no-allocation limit of %lld", (long long)n,
                        (long long)WNLQG_MLP_NOALLOC_MAX_SAMPLES);
    MLP_HIP(hipSetDevice(g->device));
    float *scr = nullptr;
    int rc = scratch_address(g->device, &scr);
    if (rc) return rc;
    const size_t nwaves = (size_t)MLP_GRAD_BLOCKS * MLP_WAVES;
    hipStream_t st = stream ? (hipStream_t)stream : g->stream;
    // every per-wave partial and loss is written by k_mlp_grad before k_mlp_reduce reads it: nothing to clear
    rc = launch_gradient_to(g, x, t, n, scr, scr + nwaves * MLP_NP, grad, st);
    if (rc) return rc;
    if (!stream) MLP_HIP(hipStreamSynchronize(st));
    return WNLQG_OK;
}

int32_t wnlqg_mlp_train(wnlqg_mlp *g, const float *x, const float *t, int64_t n, int32_t steps, int32_t batch,
                        float lr, uint64_t seed, float *loss_history) {
    if (!g || !x || !t || n < 1 || steps < 0 || batch < 1 || !(lr > 0.0f))
        return mlp_fail(WNLQG_E_INVALID, "invalid arguments");
    for (int64_t i = 0; i < n; ++i) {
        bool ok = std::isfinite(t[i]);
        for (int c = 0; c < MLP_IN; ++c) ok = ok && std::isfinite(x[MLP_IN * i + c]);
        if (!ok) return mlp_fail(WNLQG_E_NONFINITE, "non-finite training sample %lld", (long long)i);
    }
    const long long b = std::min<long long>(batch, n);
    const long long per_epoch = std::max<long long>(1, n / b);
    std::vector<float> xs((size_t)MLP_IN * n), ts((size_t)n);
    std::vector<long long> perm((size_t)n);
    for (long long i = 0; i < n; ++i) perm[i] = i;
    std::mt19937_64 rng(seed);
    std::vector<float> hist;
    for (int s = 0; s < steps; ++s) {
        const long long e = s % per_epoch;
        if (e == 0) {
            // new epoch: reshuffle on the host (deterministic in the seed) and upload
            std::shuffle(perm.begin(), perm.end(), rng);
            for (long long i = 0; i < n; ++i) {
                for (int c = 0; c < MLP_IN; ++c) xs[MLP_IN * i + c] = x[MLP_IN * perm[i] + c];
                ts[i] = t[perm[i]];
            }
            MLP_HIP(hipStreamSynchronize(g->stream));
            int rc = upload_samples(g, xs.data(), ts.data(), n);
            if (rc) return rc;
        }
        int rc = launch_gradient(g, g->x.p + (size_t)MLP_IN * e * b, g->t.p + e * b, b);
        if (rc) return rc;
        ++g->step;
        const float c1 = 1.0f - std::pow(0.9f, (float)g->step), c2 = 1.0f - std::pow(0.999f, (float)g->step);
        hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NP + 255) / 256), dim3(256), 0, g->stream, g->P.p, g->m.p, g->v.p,
                           g->grad.p, lr, c1, c2);
        MLP_HIP(hipGetLastError());
        if (loss_history) {
            float l = 0.0f;
            MLP_HIP(hipMemcpyAsync(&l, g->grad.p + MLP_NP, sizeof(float), hipMemcpyDeviceToHost, g->stream));
            MLP_HIP(hipStreamSynchronize(g->stream));
            loss_history[s] = l / (float)b;
        }
    }
    MLP_HIP(hipStreamSynchronize(g->stream));
    return WNLQG_OK;
}

}  // extern "C"
