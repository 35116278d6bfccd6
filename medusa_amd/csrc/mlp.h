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
// The policy overrides the plain engine's choice only when it predicts the other action to
// be cheaper by more than a margin (log2 units of cost); otherwise the plain engine's choice
// stands, so an uncertain model falls back to it.  Direction and kernel variant: 0.25 (about
// 19 %; a 2x margin kept every BFS decision the heuristic's, while at 0.25 the direction policy
// pulls on about six more supersteps of the 2.4 M-vertex Chung-Lu graph, 1.07x faster); bucket
// width: 0.1 (7 %; the width model's fit error on its training runs is about 0.06, and the best
// width on RMAT-20 beats the default by 12-41 %).  The bucket width is chosen once per run, from the features at its first phase,
// whose training rows carry the cost of whole runs at each fixed width; switching widths at
// later phases on the network's predictions lost 5-22 % on 1400x1400 lattices (its training
// runs never switch).
constexpr float kPolicyMargin = 0.25f;
constexpr float kDirectionMargin = kPolicyMargin;
constexpr float kWidthMargin = 0.1f;

// Direction from the two predicted costs and the plain engine's choice.
inline __host__ __device__ bool policy_pull(float c_push, float c_pull, bool plain_pull) {
    if (c_pull + kDirectionMargin < c_push) return true;
    if (c_push + kDirectionMargin < c_pull) return false;
    return plain_pull;
}

float mlp_eval_host(const float* params, const float* x);
const float* policy_params(const medusa_policy* p);
const float* policy_device_params(const medusa_policy* p);
uint32_t policy_kind(const medusa_policy* p);

}  // namespace medusa

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
namespace medusa {
// Copies the parameters into LDS (no barrier).  With 256 threads every thread issues all
// of its loads before its first store, so the copy costs one memory round trip instead of
// one per element it moves.  A kernel may call it first thing, so that the copy overlaps
// its other loads, and then evaluate with staged = true.
constexpr int kStageBlock = 256;
constexpr int kStagePer = (int)((kMlpParams + kStageBlock - 1) / kStageBlock);
__device__ inline void mlp_stage(const float* __restrict__ P, float* sP) {
    const int t = threadIdx.x;
    if (blockDim.x == kStageBlock) {
        float r[kStagePer];
#pragma unroll
        for (int j = 0; j < kStagePer; ++j) {
            const int i = t + j * kStageBlock;
            r[j] = i < (int)kMlpParams ? P[i] : 0.f;
        }
#pragma unroll
        for (int j = 0; j < kStagePer; ++j) {
            const int i = t + j * kStageBlock;
            if (i < (int)kMlpParams) sP[i] = r[j];
        }
        return;
    }
    for (int i = t; i < (int)kMlpParams; i += blockDim.x) sP[i] = P[i];
}

// A workgroup evaluates the MLP on `rows` input rows held in LDS (x[rows][16]) with the
// parameters staged once into LDS (sP[kMlpParams]; copied here unless `staged`).  Hidden
// units of all rows are computed in parallel, one (row, unit) per thread; out[r] is written
// for every row.  The FP32 operations and their order per row are those of mlp_eval_host,
// so host and device pick the same action.  Every thread of the block must call it; out[]
// is valid afterwards.  The dot products are unrolled by 4 only: kernels that evaluate the
// policy in their last workgroup keep their register budget (and occupancy) for the main body.
__device__ inline void mlp_eval_rows(const float* __restrict__ P, float* sP, const float* x, int rows, float* s_a1,
                                     float* s_a2, float* out, bool staged = false) {
    const int t = threadIdx.x, nt = blockDim.x;
    if (!staged) mlp_stage(P, sP);
    __syncthreads();
    const float* W1 = sP;
    const float* b1 = W1 + kMlpH * kMlpIn;
    const float* W2 = b1 + kMlpH;
    const float* b2 = W2 + kMlpH * kMlpH;
    const float* W3 = b2 + kMlpH;
    for (int j = t; j < rows * kMlpH; j += nt) {
        const int r = j / kMlpH, i = j % kMlpH;
        float z = b1[i];
#pragma unroll 4
        for (int k = 0; k < kMlpIn; ++k) z = fmaf(W1[i * kMlpIn + k], x[r * kMlpIn + k], z);
        s_a1[j] = z > 0.f ? z : z * kMlpSlope;
    }
    __syncthreads();
    for (int j = t; j < rows * kMlpH; j += nt) {
        const int r = j / kMlpH, i = j % kMlpH;
        float z = b2[i];
#pragma unroll 4
        for (int k = 0; k < kMlpH; ++k) z = fmaf(W2[i * kMlpH + k], s_a1[r * kMlpH + k], z);
        s_a2[j] = z > 0.f ? z : z * kMlpSlope;
    }
    __syncthreads();
    for (int r = t; r < rows; r += nt) {
        float o = W3[kMlpH];
#pragma unroll 4
        for (int i = 0; i < kMlpH; ++i) o = fmaf(W3[i], s_a2[r * kMlpH + i], o);
        out[r] = o;
    }
    __syncthreads();
}

// Parameters copied into LDS by one wave: every lane issues its 16-byte loads before its first
// store (one memory round trip), and LDS keeps one wave's accesses in order, so the wave may
// read sP right after (no barrier; other waves must not use sP).  sP holds kMlpParamsPadded
// floats (whole 16-byte vectors).
constexpr int kMlpVec = (int)((kMlpParams + 3) / 4);           // 16-byte vectors of the parameters
constexpr int kMlpVecPerLane = (kMlpVec + 63) / 64;
constexpr int kMlpParamsPadded = 4 * kMlpVec;  // a policy's device copy is zero-padded to this
__device__ inline void mlp_stage_wave(const float* __restrict__ P, float* sP) {
    const int lane = threadIdx.x & 63;
    const float4* P4 = reinterpret_cast<const float4*>(P);
    float4 r[kMlpVecPerLane];
#pragma unroll
    for (int j = 0; j < kMlpVecPerLane; ++j) {
        const int v = lane + 64 * j;
        if (v < kMlpVec - 1) {
            r[j] = P4[v];
        } else if (v == kMlpVec - 1) {  // the last vector may be partial
            r[j] = make_float4(0.f, 0.f, 0.f, 0.f);
            for (int e = 0; e < 4 && 4 * v + e < (int)kMlpParams; ++e) (&r[j].x)[e] = P[4 * v + e];
        }
    }
#pragma unroll
    for (int j = 0; j < kMlpVecPerLane; ++j) {
        const int v = lane + 64 * j;
        if (v < kMlpVec) reinterpret_cast<float4*>(sP)[v] = r[j];
    }
    __builtin_amdgcn_wave_barrier();
}

// The same copy started by one wave with direct global-to-LDS loads, which hold no registers
// while in flight (lane l of load j writes the 16 bytes at sP + 256 j + 4 l); the wave must call
// mlp_stage_wave_wait() before it reads sP, and before it ends.  P is a policy's device copy
// (kMlpParamsPadded floats, so every vector is whole).
__device__ inline void mlp_stage_wave_begin(const float* __restrict__ P, float* sP) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int j = 0; j < kMlpVecPerLane; ++j) {
        const int v = lane + 64 * j;
        if (v < kMlpVec)
            __builtin_amdgcn_global_load_lds(P + 4 * v, (__attribute__((address_space(3))) void*)(sP + 256 * j), 16, 0,
                                             0);
    }
}
__device__ inline void mlp_stage_wave_wait() {
    __builtin_amdgcn_s_waitcnt(0);  // vmcnt(0) expcnt(0) lgkmcnt(0): the LDS writes have landed
    __builtin_amdgcn_wave_barrier();
}

// LDS of a one-wave evaluation: the parameters (mlp_stage_wave), then 64 floats through which
// the activations cross the wave.
constexpr int kMlpWaveLds = kMlpParamsPadded + 2 * kMlpH;

// One wave evaluates the MLP on two input rows with the parameters in sP (kMlpWaveLds floats of
// LDS, staged by mlp_stage_wave): lanes 0-31 pass row 0 in x, lanes 32-63 row 1, and lane
// 32 * r + i computes hidden unit i of row r.  A layer's activations are written to LDS and
// every lane reads its row's 32 (the same addresses across a half-wave: broadcasts); one wave's
// LDS accesses complete in order, so there is no barrier.  The FP32 operations and their order
// per row are those of mlp_eval_host.  Every lane gets both outputs; the calling wave must be
// converged.
__device__ inline void mlp_eval_wave2(float* sP, const float* x, float& out0, float& out1) {
    const int lane = threadIdx.x & 63, r = lane >> 5, i = lane & 31;
    const float* W1 = sP;
    const float* b1 = W1 + kMlpH * kMlpIn;
    const float* W2 = b1 + kMlpH;
    const float* b2 = W2 + kMlpH * kMlpH;
    const float* W3 = b2 + kMlpH;
    float* act = sP + kMlpParamsPadded;
    const float* ar = act + r * kMlpH;
    float z = b1[i];
#pragma unroll
    for (int k = 0; k < kMlpIn; ++k) z = fmaf(W1[i * kMlpIn + k], x[k], z);
    act[lane] = z > 0.f ? z : z * kMlpSlope;
    __builtin_amdgcn_wave_barrier();
    float z2 = b2[i];
#pragma unroll 8
    for (int k = 0; k < kMlpH; ++k) z2 = fmaf(W2[i * kMlpH + k], ar[k], z2);
    const float a2 = z2 > 0.f ? z2 : z2 * kMlpSlope;
    __builtin_amdgcn_wave_barrier();
    act[lane] = a2;  // after every lane's reads of the first layer (in order in LDS)
    __builtin_amdgcn_wave_barrier();
    float o = W3[kMlpH];
#pragma unroll 8
    for (int k = 0; k < kMlpH; ++k) o = fmaf(W3[k], ar[k], o);
    out0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, o), 0));
    out1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, o), kMlpH));
}
}  // namespace medusa
#endif
