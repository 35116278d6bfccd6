// Shape of the scheduler MLP and its feature encoding (shared by the engine and the
// training kernels).
#pragma once

#include <cmath>
#include <cstdint>

#include "../../include/medusa_c.h"

namespace medusa {

constexpr int kMlpIn = MEDUSA_POLICY_INPUTS;  // 16
constexpr int kMlpH = 32;
constexpr uint32_t kMlpParams = kMlpH * kMlpIn + kMlpH + kMlpH * kMlpH + kMlpH + kMlpH + 1;
constexpr float kMlpSlope = 1.0f / 64.0f;

float mlp_eval_host(const float* params, const float* x);
const float* policy_params(const medusa_policy* p);
const float* policy_device_params(const medusa_policy* p);
uint32_t policy_kind(const medusa_policy* p);

}  // namespace medusa

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
namespace medusa {
// A workgroup evaluates the MLP on one input row held in LDS (x[16]): threads 0..31
// each compute one hidden unit, thread 0 accumulates the output in unit order.  The
// FP32 operations and their order are those of mlp_eval_host, so host and device pick
// the same action.  Every thread of the block must call it; thread 0 gets the value.
__device__ inline float mlp_eval_block(const float* __restrict__ P, const float* x, float* s_a1, float* s_a2) {
    const int t = threadIdx.x;
    const float* W1 = P;
    const float* b1 = W1 + kMlpH * kMlpIn;
    const float* W2 = b1 + kMlpH;
    const float* b2 = W2 + kMlpH * kMlpH;
    const float* W3 = b2 + kMlpH;
    if (t < kMlpH) {
        float z = b1[t];
        for (int k = 0; k < kMlpIn; ++k) z = fmaf(W1[t * kMlpIn + k], x[k], z);
        s_a1[t] = z > 0.f ? z : z * kMlpSlope;
    }
    __syncthreads();
    if (t < kMlpH) {
        float z = b2[t];
        for (int k = 0; k < kMlpH; ++k) z = fmaf(W2[t * kMlpH + k], s_a1[k], z);
        s_a2[t] = z > 0.f ? z : z * kMlpSlope;
    }
    __syncthreads();
    float o = W3[kMlpH];
    if (t == 0)
        for (int i = 0; i < kMlpH; ++i) o = fmaf(W3[i], s_a2[i], o);
    __syncthreads();
    return o;
}
}  // namespace medusa
#endif
