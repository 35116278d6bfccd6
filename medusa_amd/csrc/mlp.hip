// Learned per-superstep scheduler: a small MLP cost model trained on the GPU.
//
// No reference counterpart (the nearest analogues are the frontier counters of
// Algorithm/CPUFunctorHoster.cpp:13-15,21-30 and the per-operator timers of
// SSSP/UserOperators.h:198-262).  The network maps a superstep's features plus a
// candidate action to a predicted log-cost; the engine evaluates every candidate and
// takes the cheapest.  The plain engine's results never depend on the choice: every
// action (push/pull direction, bucket width, kernel variant) computes the same output.
//
// Architecture: x[16] -> 32 -> 32 -> 1, leaky ReLU (slope 1/64) on both hidden layers.
// Parameters, in order: W1[32][16], b1[32], W2[32][32], b2[32], W3[32], b3.
// Training: full-batch mean squared error, Adam.  Forward and backward run as HIP
// kernels written here; every reduction over samples has a fixed order (samples in
// order within chunks of 256, chunks in order), so gradients are deterministic.

#include <cmath>

#include "engine.h"
#include "mlp.h"

namespace medusa {

__device__ __forceinline__ float lrelu(float z) { return z > 0.f ? z : z * kMlpSlope; }
__device__ __forceinline__ float lrelu_d(float z) { return z > 0.f ? 1.f : kMlpSlope; }

// Forward for one sample per thread; stores hidden activations and pre-activation
// signs for the backward pass (when z1/z2 are non-null) and the output.
__global__ __launch_bounds__(256) void k_mlp_forward(const float* __restrict__ P, const float* __restrict__ X,
                                                     uint32_t n, float* __restrict__ out, float* __restrict__ Z1,
                                                     float* __restrict__ Z2) {
    __shared__ float sP[kMlpParams];
    for (uint32_t i = threadIdx.x; i < kMlpParams; i += blockDim.x) sP[i] = P[i];
    __syncthreads();
    const float* W1 = sP;
    const float* b1 = W1 + kMlpH * kMlpIn;
    const float* W2 = b1 + kMlpH;
    const float* b2 = W2 + kMlpH * kMlpH;
    const float* W3 = b2 + kMlpH;
    const float b3 = W3[kMlpH];
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        float x[kMlpIn], a1[kMlpH];
#pragma unroll
        for (int k = 0; k < kMlpIn; ++k) x[k] = X[(size_t)s * kMlpIn + k];
#pragma unroll
        for (int i = 0; i < kMlpH; ++i) {
            float z = b1[i];
#pragma unroll
            for (int k = 0; k < kMlpIn; ++k) z = fmaf(W1[i * kMlpIn + k], x[k], z);
            if (Z1) Z1[(size_t)s * kMlpH + i] = z;
            a1[i] = lrelu(z);
        }
        float o = b3;
        for (int i = 0; i < kMlpH; ++i) {
            float z = b2[i];
#pragma unroll
            for (int k = 0; k < kMlpH; ++k) z = fmaf(W2[i * kMlpH + k], a1[k], z);
            if (Z2) Z2[(size_t)s * kMlpH + i] = z;
            o = fmaf(W3[i], lrelu(z), o);
        }
        out[s] = o;
    }
}

// Per-sample deltas: g = 2(out - y)/n, dZ2 = g*W3 .* f'(Z2), dZ1 = (dZ2 W2) .* f'(Z1).
// Overwrites Z2 with dZ2 and Z1 with dZ1; writes g into out.
__global__ __launch_bounds__(256) void k_mlp_delta(const float* __restrict__ P, const float* __restrict__ y, uint32_t n,
                                                   float* __restrict__ out, float* __restrict__ Z1,
                                                   float* __restrict__ Z2) {
    __shared__ float sP[kMlpParams];
    for (uint32_t i = threadIdx.x; i < kMlpParams; i += blockDim.x) sP[i] = P[i];
    __syncthreads();
    const float* W2 = sP + kMlpH * kMlpIn + kMlpH;
    const float* W3 = W2 + kMlpH * kMlpH + kMlpH;
    const float scale = 2.f / (float)n;
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
        const float g = (out[s] - y[s]) * scale;
        out[s] = g;
        float d2[kMlpH];
#pragma unroll
        for (int i = 0; i < kMlpH; ++i) {
            d2[i] = g * W3[i] * lrelu_d(Z2[(size_t)s * kMlpH + i]);
            Z2[(size_t)s * kMlpH + i] = d2[i];
        }
        for (int k = 0; k < kMlpH; ++k) {
            float acc = 0.f;
#pragma unroll
            for (int i = 0; i < kMlpH; ++i) acc = fmaf(d2[i], W2[i * kMlpH + k], acc);
            const size_t at = (size_t)s * kMlpH + k;
            Z1[at] = acc * lrelu_d(Z1[at]);
        }
    }
}

constexpr uint32_t kGradChunk = 256;  // samples per partial gradient

// Partial gradients: thread p of column `chunk` sums parameter p's gradient over the
// chunk's samples in order.  Needs the activations A1 = f(Z1_pre), A2 = f(Z2_pre).
__global__ __launch_bounds__(256) void k_mlp_grad_part(const float* __restrict__ X, uint32_t n,
                                                       const float* __restrict__ G, const float* __restrict__ D1,
                                                       const float* __restrict__ D2, const float* __restrict__ A1,
                                                       const float* __restrict__ A2, float* __restrict__ part) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    const uint32_t s0 = blockIdx.y * kGradChunk, s1 = min(n, s0 + kGradChunk);
    float acc = 0.f;
    constexpr uint32_t ob1 = kMlpH * kMlpIn, oW2 = ob1 + kMlpH, ob2 = oW2 + kMlpH * kMlpH, oW3 = ob2 + kMlpH,
                       ob3 = oW3 + kMlpH;
    if (p < ob1) {
        const uint32_t i = p / kMlpIn, k = p % kMlpIn;
        for (uint32_t s = s0; s < s1; ++s) acc = fmaf(D1[(size_t)s * kMlpH + i], X[(size_t)s * kMlpIn + k], acc);
    } else if (p < oW2) {
        const uint32_t i = p - ob1;
        for (uint32_t s = s0; s < s1; ++s) acc += D1[(size_t)s * kMlpH + i];
    } else if (p < ob2) {
        const uint32_t i = (p - oW2) / kMlpH, k = (p - oW2) % kMlpH;
        for (uint32_t s = s0; s < s1; ++s) acc = fmaf(D2[(size_t)s * kMlpH + i], A1[(size_t)s * kMlpH + k], acc);
    } else if (p < oW3) {
        const uint32_t i = p - ob2;
        for (uint32_t s = s0; s < s1; ++s) acc += D2[(size_t)s * kMlpH + i];
    } else if (p < ob3) {
        const uint32_t i = p - oW3;
        for (uint32_t s = s0; s < s1; ++s) acc = fmaf(G[s], A2[(size_t)s * kMlpH + i], acc);
    } else {
        for (uint32_t s = s0; s < s1; ++s) acc += G[s];
    }
    part[(size_t)blockIdx.y * kMlpParams + p] = acc;
}

// Gradient = partials summed over chunks in chunk order (deterministic).
__global__ void k_mlp_grad_sum(const float* __restrict__ part, uint32_t nchunks, float* __restrict__ grad) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    float acc = 0.f;
    for (uint32_t c = 0; c < nchunks; ++c) acc += part[(size_t)c * kMlpParams + p];
    grad[p] = acc;
}

// Hidden activations from pre-activations (A = f(Z)), elementwise.
__global__ void k_mlp_act(const float* __restrict__ Z, float* __restrict__ A, size_t count) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        A[i] = lrelu(Z[i]);
}

__global__ void k_mlp_adam(float* __restrict__ P, const float* __restrict__ grad, float* __restrict__ M,
                           float* __restrict__ V, float lr, float bc1, float bc2) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    const float g = grad[p];
    const float m = 0.9f * M[p] + 0.1f * g;
    const float v = 0.999f * V[p] + 0.001f * g * g;
    M[p] = m;
    V[p] = v;
    P[p] -= lr * (m / bc1) / (sqrtf(v / bc2) + 1e-8f);
}

__global__ void k_mlp_loss(const float* __restrict__ out, const float* __restrict__ y, uint32_t n,
                           float* __restrict__ loss) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    float acc = 0.f;
    for (uint32_t s = 0; s < n; ++s) {
        const float e = out[s] - y[s];
        acc = fmaf(e, e, acc);
    }
    *loss = acc / (float)n;
}

// Device work buffers for a batch of n samples.
struct MlpWork {
    DBuf<float> X, y, out, Z1, Z2, A1, A2, grad, M, V, loss, part;
    uint32_t cap = 0;
    void ensure(uint32_t n) {
        if (n <= cap) return;
        part.alloc((size_t)((n + kGradChunk - 1) / kGradChunk) * kMlpParams);
        X.alloc((size_t)n * kMlpIn);
        y.alloc(n);
        out.alloc(n);
        Z1.alloc((size_t)n * kMlpH);
        Z2.alloc((size_t)n * kMlpH);
        A1.alloc((size_t)n * kMlpH);
        A2.alloc((size_t)n * kMlpH);
        cap = n;
    }
};

}  // namespace medusa

struct medusa_policy {
    uint32_t kind = 0;
    std::vector<float> params;  // host copy used for per-superstep inference
    medusa::DBuf<float> d_params;
    medusa::MlpWork w;
    hipStream_t stream = nullptr;
    int device = 0;
};

namespace medusa {

static void mlp_gradients(medusa_policy* pol, uint32_t n, float* loss_out) {
    hipStream_t s = pol->stream;
    MlpWork& w = pol->w;
    const uint32_t grid = grid_for(n, 256, 1024);
    hipLaunchKernelGGL(k_mlp_forward, dim3(grid), dim3(256), 0, s, (const float*)pol->d_params.p,
                       (const float*)w.X.p, n, w.out.p, w.Z1.p, w.Z2.p);
    MDS_LAUNCHED();
    if (loss_out) {
        hipLaunchKernelGGL(k_mlp_loss, dim3(1), dim3(1), 0, s, (const float*)w.out.p, (const float*)w.y.p, n, w.loss.p);
        MDS_LAUNCHED();
    }
    hipLaunchKernelGGL(k_mlp_act, dim3(grid_for((uint64_t)n * kMlpH, 256, 1024)), dim3(256), 0, s,
                       (const float*)w.Z1.p, w.A1.p, (size_t)n * kMlpH);
    MDS_LAUNCHED();
    hipLaunchKernelGGL(k_mlp_act, dim3(grid_for((uint64_t)n * kMlpH, 256, 1024)), dim3(256), 0, s,
                       (const float*)w.Z2.p, w.A2.p, (size_t)n * kMlpH);
    MDS_LAUNCHED();
    hipLaunchKernelGGL(k_mlp_delta, dim3(grid), dim3(256), 0, s, (const float*)pol->d_params.p, (const float*)w.y.p,
                       n, w.out.p, w.Z1.p, w.Z2.p);
    MDS_LAUNCHED();
    const uint32_t nchunks = (n + kGradChunk - 1) / kGradChunk;
    hipLaunchKernelGGL(k_mlp_grad_part, dim3((kMlpParams + 255) / 256, nchunks), dim3(256), 0, s,
                       (const float*)w.X.p, n, (const float*)w.out.p, (const float*)w.Z1.p, (const float*)w.Z2.p,
                       (const float*)w.A1.p, (const float*)w.A2.p, w.part.p);
    MDS_LAUNCHED();
    hipLaunchKernelGGL(k_mlp_grad_sum, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, (const float*)w.part.p,
                       nchunks, w.grad.p);
    MDS_LAUNCHED();
    if (loss_out) {
        MDS_HIP(hipMemcpyAsync(loss_out, w.loss.p, 4, hipMemcpyDeviceToHost, s));
        MDS_HIP(hipStreamSynchronize(s));
    }
}

static void upload_batch(medusa_policy* pol, const float* X, const float* y, uint32_t n) {
    pol->w.ensure(n);
    MDS_HIP(hipMemcpyAsync(pol->w.X.p, X, (size_t)n * kMlpIn * 4, hipMemcpyHostToDevice, pol->stream));
    if (y) MDS_HIP(hipMemcpyAsync(pol->w.y.p, y, (size_t)n * 4, hipMemcpyHostToDevice, pol->stream));
}

// Host evaluation with the same FP32 operations as k_mlp_forward.
float mlp_eval_host(const float* P, const float* x) {
    const float* W1 = P;
    const float* b1 = W1 + kMlpH * kMlpIn;
    const float* W2 = b1 + kMlpH;
    const float* b2 = W2 + kMlpH * kMlpH;
    const float* W3 = b2 + kMlpH;
    float a1[kMlpH];
    for (int i = 0; i < kMlpH; ++i) {
        float z = b1[i];
        for (int k = 0; k < kMlpIn; ++k) z = std::fmaf(W1[i * kMlpIn + k], x[k], z);
        a1[i] = z > 0.f ? z : z * kMlpSlope;
    }
    float o = W3[kMlpH];
    for (int i = 0; i < kMlpH; ++i) {
        float z = b2[i];
        for (int k = 0; k < kMlpH; ++k) z = std::fmaf(W2[i * kMlpH + k], a1[k], z);
        o = std::fmaf(W3[i], z > 0.f ? z : z * kMlpSlope, o);
    }
    return o;
}

const float* policy_params(const medusa_policy* p) { return p->params.data(); }
const float* policy_device_params(const medusa_policy* p) { return p->d_params.p; }
uint32_t policy_kind(const medusa_policy* p) { return p->kind; }

}  // namespace medusa

using namespace medusa;

extern "C" int medusa_policy_create(uint32_t kind, uint64_t seed, medusa_policy** out) {
    return guard([&] {
        MDS_CHECK(out, MEDUSA_ERR_ARG, "policy_create: null out");
        MDS_CHECK(kind == MEDUSA_POLICY_DIRECTION || kind == MEDUSA_POLICY_DELTA || kind == MEDUSA_POLICY_VARIANT,
                  MEDUSA_ERR_ARG,
                  "policy_create: unknown kind %u", kind);
        auto p = std::make_unique<medusa_policy>();
        p->kind = kind;
        p->params.assign(kMlpParams, 0.f);
        // He-style uniform init from a deterministic stream
        Rng rng(seed, 7);
        auto fill = [&](uint32_t off, uint32_t count, float fan_in) {
            const float lim = std::sqrt(6.f / fan_in);
            for (uint32_t i = 0; i < count; ++i) p->params[off + i] = (float)((2.0 * rng.uniform() - 1.0) * lim);
        };
        fill(0, kMlpH * kMlpIn, (float)kMlpIn);
        fill(kMlpH * kMlpIn + kMlpH, kMlpH * kMlpH, (float)kMlpH);
        fill(kMlpH * kMlpIn + kMlpH + kMlpH * kMlpH + kMlpH, kMlpH, (float)kMlpH);
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess) ndev = 0;
        (void)hipGetLastError();
        MDS_CHECK(ndev > 0, MEDUSA_ERR_HIP, "policy_create: no GPU visible");
        MDS_HIP(hipGetDevice(&p->device));
        MDS_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
        std::vector<float> padded(kMlpParamsPadded, 0.f);  // whole 16-byte vectors for LDS staging
        std::copy(p->params.begin(), p->params.end(), padded.begin());
        p->d_params.upload(padded.data(), kMlpParamsPadded, p->stream);
        p->w.grad.alloc(kMlpParams);
        p->w.M.alloc(kMlpParams);
        p->w.V.alloc(kMlpParams);
        p->w.loss.alloc(1);
        MDS_HIP(hipStreamSynchronize(p->stream));
        *out = p.release();
    });
}

extern "C" int medusa_policy_destroy(medusa_policy* p) {
    return guard([&] {
        if (!p) return;
        (void)hipSetDevice(p->device);
        hipStream_t s = p->stream;
        if (s) (void)hipStreamSynchronize(s);
        delete p;
        if (s) (void)hipStreamDestroy(s);
    });
}

extern "C" int medusa_policy_params(medusa_policy* p, float* get, const float* set, uint32_t count) {
    return guard([&] {
        MDS_CHECK(p, MEDUSA_ERR_ARG, "null policy");
        MDS_CHECK(count == kMlpParams, MEDUSA_ERR_ARG, "policy_params: count must be %u", kMlpParams);
        if (set) {
            p->params.assign(set, set + kMlpParams);
            MDS_HIP(hipSetDevice(p->device));
            MDS_HIP(hipMemcpy(p->d_params.p, set, kMlpParams * 4, hipMemcpyHostToDevice));
        }
        if (get) std::memcpy(get, p->params.data(), kMlpParams * 4);
    });
}

extern "C" int medusa_policy_gradients(medusa_policy* p, const float* X, const float* y, uint32_t n, float* grad,
                                       float* loss) {
    return guard([&] {
        MDS_CHECK(p && X && y && grad && n > 0, MEDUSA_ERR_ARG, "policy_gradients: bad arguments");
        MDS_CHECK(n <= 65535u * kGradChunk, MEDUSA_ERR_ARG, "policy_gradients: at most %u samples", 65535u * kGradChunk);
        MDS_HIP(hipSetDevice(p->device));
        upload_batch(p, X, y, n);
        mlp_gradients(p, n, loss);
        MDS_HIP(hipMemcpyAsync(grad, p->w.grad.p, kMlpParams * 4, hipMemcpyDeviceToHost, p->stream));
        MDS_HIP(hipStreamSynchronize(p->stream));
    });
}

extern "C" int medusa_policy_predict(medusa_policy* p, const float* X, uint32_t n, float* out) {
    return guard([&] {
        MDS_CHECK(p && X && out && n > 0, MEDUSA_ERR_ARG, "policy_predict: bad arguments");
        MDS_HIP(hipSetDevice(p->device));
        upload_batch(p, X, nullptr, n);
        hipLaunchKernelGGL(k_mlp_forward, dim3(grid_for(n, 256, 1024)), dim3(256), 0, p->stream,
                           (const float*)p->d_params.p, (const float*)p->w.X.p, n, p->w.out.p, (float*)nullptr,
                           (float*)nullptr);
        MDS_LAUNCHED();
        MDS_HIP(hipMemcpyAsync(out, p->w.out.p, (size_t)n * 4, hipMemcpyDeviceToHost, p->stream));
        MDS_HIP(hipStreamSynchronize(p->stream));
    });
}

extern "C" int medusa_policy_train(medusa_policy* p, const float* X, const float* y, uint32_t n, uint32_t iters,
                                   float lr, float* final_loss) {
    return guard([&] {
        MDS_CHECK(p && X && y && n > 0, MEDUSA_ERR_ARG, "policy_train: bad arguments");
        MDS_CHECK(n <= 65535u * kGradChunk, MEDUSA_ERR_ARG, "policy_train: at most %u samples", 65535u * kGradChunk);
        MDS_HIP(hipSetDevice(p->device));
        upload_batch(p, X, y, n);
        MDS_HIP(hipMemsetAsync(p->w.M.p, 0, kMlpParams * 4, p->stream));
        MDS_HIP(hipMemsetAsync(p->w.V.p, 0, kMlpParams * 4, p->stream));
        double b1t = 1.0, b2t = 1.0;
        for (uint32_t it = 0; it < iters; ++it) {
            mlp_gradients(p, n, nullptr);
            b1t *= 0.9;
            b2t *= 0.999;
            hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, p->stream, p->d_params.p,
                               (const float*)p->w.grad.p, p->w.M.p, p->w.V.p, lr, (float)(1.0 - b1t),
                               (float)(1.0 - b2t));
            MDS_LAUNCHED();
        }
        float loss = 0;
        mlp_gradients(p, n, &loss);
        MDS_HIP(hipMemcpy(p->params.data(), p->d_params.p, kMlpParams * 4, hipMemcpyDeviceToHost));
        if (final_loss) *final_loss = loss;
    });
}
