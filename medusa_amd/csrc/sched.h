// Per-superstep scheduling: frontier features, the plain engine's heuristics,
// the learned policy hook, and the per-superstep trace that trains it.
#pragma once

#include <algorithm>
#include <cmath>

#include "engine.h"
#include "mlp.h"

namespace medusa {

__host__ __device__ inline float lg1(double x) { return (float)log2(1.0 + (x > 0.0 ? x : 0.0)); }

// Direction rows (record = step, n_f, m_f, n_u, m_u, n, m, ...): frontier and unvisited
// vertex/edge shares on log scales, mean degree, step, size, and the candidate action
// (0 push, 1 pull) with its interactions.  The seven logarithms do not depend on the
// action: direction_log(r, i) is logarithm i (of n, m, n_f, m_f, n_u, m_u, m/n) and
// direction_fill builds a row from them, so a workgroup can compute them in parallel.
constexpr int kDirectionLogs = 7;
__host__ __device__ inline float direction_log(const double* r, int i) {
    const double n = r[5] > 1.0 ? r[5] : 1.0, m = r[6] > 1.0 ? r[6] : 1.0;
    if (i == 0) return lg1(n);
    if (i == 1) return lg1(m);
    if (i == 6) return lg1(m / n);
    return lg1(r[i - 1]);  // i = 2..5: n_f, m_f, n_u, m_u
}
__host__ __device__ inline void direction_fill(const double* r, const float* L, double action, float* x) {
    for (int i = 0; i < MEDUSA_POLICY_INPUTS; ++i) x[i] = 0.f;
    const float ln = L[0], lm = L[1];
    const float a = action != 0.0 ? 1.f : 0.f;
    x[0] = L[2] / ln;
    x[1] = L[3] / lm;
    x[2] = L[4] / ln;
    x[3] = L[5] / lm;
    x[4] = L[6] / 8.f;
    x[5] = (float)(r[0] < 64.0 ? r[0] : 64.0) / 64.f;
    x[6] = ln / 32.f;
    x[8] = a;
    x[9] = a * x[0];
    x[10] = a * x[1];
    x[11] = a * x[2];
    x[12] = a * x[3];
    x[13] = 1.f - a;
}
__host__ __device__ inline void direction_row(const double* r, double action, float* x) {
    float L[kDirectionLogs];
    for (int i = 0; i < kDirectionLogs; ++i) L[i] = direction_log(r, i);
    direction_fill(r, L, action, x);
}

// Delta rows (record = phase, far entries, mean weight, threshold, 0, n, m, ...): pending far
// entries, threshold in units of the mean weight, mean degree, size, phase index, and the
// candidate width in units of the mean weight with its square and interactions.
__host__ __device__ inline void delta_row(const double* r, double action, float* x) {
    for (int i = 0; i < MEDUSA_POLICY_INPUTS; ++i) x[i] = 0.f;
    const double n = r[5] > 1.0 ? r[5] : 1.0, m = r[6] > 1.0 ? r[6] : 1.0;
    const float ln = lg1(n);
    const double wmean = r[2] > 1.0 ? r[2] : 1.0;
    const float la = (float)log2((action > 1.0 ? action : 1.0) / wmean);
    x[0] = lg1(r[1]) / ln;
    x[1] = lg1(r[3] / wmean) / 16.f;
    x[2] = lg1(m / n) / 8.f;
    x[3] = ln / 32.f;
    x[4] = (float)(r[0] < 4096.0 ? r[0] : 4096.0) / 4096.f;
    x[5] = la / 8.f;
    x[6] = (la / 8.f) * (la / 8.f);
    x[7] = x[0] * la / 8.f;
    x[8] = x[1] * la / 8.f;
}

// Kernel-variant rows (PageRank record = step, dangling, wave-run vertices, lane-run vertices,
// 0, n, m, ...): size, mean degree, the shares of dangling and high in-degree vertices, and
// the candidate variant (push, binned or pull) as indicators with their interactions.
__host__ __device__ inline void variant_row(const double* r, double action, float* x) {
    for (int i = 0; i < MEDUSA_POLICY_INPUTS; ++i) x[i] = 0.f;
    const double n = r[5] > 1.0 ? r[5] : 1.0, m = r[6] > 1.0 ? r[6] : 1.0;
    const float push = action == (double)MEDUSA_PR_PUSH ? 1.f : 0.f;
    const float bin = action == (double)MEDUSA_PR_BINNED ? 1.f : 0.f;
    x[0] = lg1(n) / 32.f;
    x[1] = lg1(m / n) / 8.f;
    x[2] = (float)(r[1] / n);
    x[3] = (float)(r[2] / n);
    x[4] = (float)(r[3] / n);
    // x[5] stays 0: PageRank supersteps repeat the same work, so the row carries no step
    // index; one (training runs cover 6 supersteps) let the policy switch variants mid-run
    // outside that range, 2.5x slower on a held-out Chung-Lu graph
    x[6] = push;
    x[7] = push * x[0];
    x[8] = push * x[1];
    x[9] = push * x[2];
    x[10] = push * x[3];
    x[11] = 1.f - push - bin;
    x[12] = bin;
    x[13] = bin * x[0];
    x[14] = bin * x[1];
    x[15] = bin * x[2];
}

#if defined(__HIPCC__)
// LDS working set of a policy evaluation inside a decide kernel (up to 12 candidate rows).
struct MlpShared {
    float P[kMlpParams];
    float x[12 * MEDUSA_POLICY_INPUTS];
    float a1[12 * kMlpH];
    float a2[12 * kMlpH];
    float out[12];
};

// direction_log in FP32 with the hardware logarithm: within about 2e-6 of it, far below any
// decision margin, and a few instructions instead of an FP64 logarithm's long dependent chain.
__device__ inline float direction_log_f32(const double* r, int i) {
    const float n = fmaxf((float)r[5], 1.f), m = fmaxf((float)r[6], 1.f);
    auto lg = [](float v) { return __log2f(1.f + fmaxf(v, 0.f)); };
    if (i == 0) return lg(n);
    if (i == 1) return lg(m);
    if (i == 6) return lg(m / n);
    return lg((float)r[i - 1]);  // i = 2..5: n_f, m_f, n_u, m_u
}

// Predicted costs of push and pull for the superstep described by `rec`, by one converged wave
// with `sP` (kMlpWaveLds floats of LDS) to itself: the parameters are staged (or, if `staged`,
// were started by mlp_stage_wave_begin) while every lane computes the seven logarithms in
// FP32, then lanes 0-31 evaluate the push row and lanes 32-63 the pull row (direction_row's
// rows, up to the logarithms' rounding).
__device__ inline void policy_direction_costs_wave(const double* rec, const float* __restrict__ policy, float* sP,
                                                   float& c_push, float& c_pull, bool staged = false) {
    const int lane = threadIdx.x & 63;
    if (!staged) mlp_stage_wave(policy, sP);
    float L[kDirectionLogs];
#pragma unroll
    for (int i = 0; i < kDirectionLogs; ++i) L[i] = direction_log_f32(rec, i);
    float x[MEDUSA_POLICY_INPUTS];
    direction_fill(rec, L, (double)(lane >> 5), x);
    if (staged) mlp_stage_wave_wait();
    mlp_eval_wave2(sP, x, c_push, c_pull);
}

// Predicted costs of push (out[0]) and pull (out[1]) for the superstep described by `rec`.
// Every thread of the block calls it; `staged`: the kernel already ran mlp_stage into m.P.
// Threads 0 and 1 build the push and pull rows with FP32 logarithms (direction_row's rows, up
// to their rounding); mlp_eval_rows' first barrier publishes them.
__device__ inline void policy_direction_costs(const double* rec, const float* __restrict__ policy, MlpShared& m,
                                              bool staged = false) {
    float x[MEDUSA_POLICY_INPUTS];
    if (threadIdx.x < 2) {
        float L[kDirectionLogs];  // in FP32, as the one-wave form does
#pragma unroll
        for (int i = 0; i < kDirectionLogs; ++i) L[i] = direction_log_f32(rec, i);
        direction_fill(rec, L, (double)threadIdx.x, x);
        for (int i = 0; i < MEDUSA_POLICY_INPUTS; ++i) m.x[threadIdx.x * MEDUSA_POLICY_INPUTS + i] = x[i];
    }
    mlp_eval_rows(policy, m.P, m.x, 2, m.a1, m.a2, m.out, staged);  // its first barrier publishes the rows
}
#endif

// Frontier features at the start of a superstep.
struct Frontier {
    uint32_t step = 0;
    uint32_t n = 0;
    uint64_t m = 0;
    uint64_t n_f = 0;  // frontier vertices
    uint64_t m_f = 0;  // their out-edges
    uint64_t n_u = 0;  // unvisited vertices
    uint64_t m_u = 0;  // their out-edges
};

// BFS and MBFS consult the learned direction policy only near the heuristic's switching
// point, where its decision ratio (alpha * m_f / m_u while pushing, beta * n_f / n while
// pulling) lies within a factor kPolicyBand of 1; farther out the heuristic's choice is taken
// without evaluating the network, which saves its cost on most supersteps.
constexpr double kPolicyBand = 8.0;
__host__ __device__ inline bool bfs_policy_consulted(bool cur_pull, double n_f, double m_f, double m_u, double n,
                                                     double alpha, double beta) {
    const double num = cur_pull ? beta * n_f : alpha * m_f, den = cur_pull ? n : m_u;
    return num * kPolicyBand > den && num < den * kPolicyBand;
}

// Largest bucket width a policy may pick: one below the value that turns buckets off.
constexpr double kMaxWidth = 4294967294.0;

struct Sched {
    enum Algo { kBfs, kSssp, kMbfs };
    // Direction-optimising thresholds (Beamer, Asanovic, Patterson, SC'12): pull once the
    // frontier's out-edges exceed the unvisited out-edges / alpha, push again once the frontier
    // falls below n / beta.  The paper uses alpha = 14.  Set from per-superstep push and pull
    // times measured on Erdos-Renyi, RMAT, Chung-Lu and lattice graphs: alpha = 3 for the
    // single-source engine, whose pull costs relatively more on this GPU, and alpha = 100 for
    // the 64-source mask engine, whose pull (over coalesced ELL slabs, reading nearly every
    // in-edge) beats the push much earlier.
    static constexpr double kAlphaBfs = 3.0, kAlphaMbfs = 100.0, kBeta = 24.0;
    const medusa_opts& o;
    Algo algo;
    Sched(const medusa_opts& opts, Algo a) : o(opts), algo(a) {}

    // Plain engine: switch to pull when the frontier's out-edges exceed the
    // unvisited out-edges / alpha; back to push when the frontier shrinks below n / beta.
    bool heuristic_pull(const Frontier& f, bool cur_pull) const {
        const double alpha = algo == kMbfs ? kAlphaMbfs : kAlphaBfs;
        if (!cur_pull) return (double)f.m_f > (double)f.m_u / alpha;
        return !((double)f.n_f < (double)f.n / kBeta);
    }
    // Direction for the next superstep: the learned policy's cheaper action when one is
    // given (and of the right kind), else the heuristic.
    bool bfs_pull(const Frontier& f, bool cur_pull) const {
        if (!o.policy || policy_kind(o.policy) != MEDUSA_POLICY_DIRECTION) return heuristic_pull(f, cur_pull);
        if (!bfs_policy_consulted(cur_pull, (double)f.n_f, (double)f.m_f, (double)f.m_u, (double)f.n,
                                  algo == kMbfs ? kAlphaMbfs : kAlphaBfs, kBeta))
            return heuristic_pull(f, cur_pull);
        double rec[MEDUSA_TRACE_FIELDS];
        to_record(f, cur_pull ? 1.0 : 0.0, 0.0, rec);
        float x[MEDUSA_POLICY_INPUTS];
        policy_row(MEDUSA_POLICY_DIRECTION, rec, 0.0, x);
        const float c_push = mlp_eval_host(policy_params(o.policy), x);
        policy_row(MEDUSA_POLICY_DIRECTION, rec, 1.0, x);
        const float c_pull = mlp_eval_host(policy_params(o.policy), x);
        return policy_pull(c_push, c_pull, heuristic_pull(f, cur_pull));
    }
    // SSSP bucket width of a run, from the features at its first phase: the learned policy
    // picks among wmean * 2^k (k = -2..8) the width with the lowest predicted cost to the end
    // of the run per vertex, if it beats the predicted cost of the default width by
    // kWidthMargin.
    uint32_t sssp_delta(const Frontier& f, uint32_t dflt, uint32_t wmean, uint32_t thr) const {
        if (!o.policy || policy_kind(o.policy) != MEDUSA_POLICY_DELTA || o.delta) return dflt;
        double rec[MEDUSA_TRACE_FIELDS];
        to_sssp_record(f, wmean, thr, 0.0, 0.0, rec);
        float x[MEDUSA_POLICY_INPUTS];
        policy_row(MEDUSA_POLICY_DELTA, rec, (double)dflt, x);
        const float c_dflt = mlp_eval_host(policy_params(o.policy), x);
        float best = 0.f;
        uint32_t pick = dflt;
        for (int k = -2; k <= 8; ++k) {
            // clamped below the "no buckets" value, so the cast cannot wrap (wmean may reach 2^32-1)
            const double d = std::min(kMaxWidth, std::max(1.0, std::ldexp((double)wmean, k)));
            policy_row(MEDUSA_POLICY_DELTA, rec, d, x);
            const float c = mlp_eval_host(policy_params(o.policy), x);
            if (k == -2 || c < best) {
                best = c;
                pick = (uint32_t)d;
            }
        }
        return best + kWidthMargin < c_dflt ? pick : dflt;
    }
    // PageRank kernel variant for one superstep (record as in variant_row): the policy's
    // cheapest variant (the binned one only where its layout exists) if it predicts a clear gain
    // over `dflt`, else `dflt`.
    int pr_variant(const double* rec, int dflt, bool binned_ok) const {
        if (!o.policy || policy_kind(o.policy) != MEDUSA_POLICY_VARIANT) return dflt;
        float x[MEDUSA_POLICY_INPUTS];
        const int cand[3] = {MEDUSA_PR_PULL, MEDUSA_PR_PUSH, MEDUSA_PR_BINNED};
        float cost[3];
        for (int i = 0; i < 3; ++i) {
            policy_row(MEDUSA_POLICY_VARIANT, rec, (double)cand[i], x);
            cost[i] = mlp_eval_host(policy_params(o.policy), x);
        }
        int best = -1;
        float c_dflt = 0.f;
        for (int i = 0; i < 3; ++i) {
            if (cand[i] == MEDUSA_PR_BINNED && !binned_ok) continue;
            if (cand[i] == dflt) c_dflt = cost[i];
            if (best < 0 || cost[i] < cost[best]) best = i;
        }
        return cost[best] + kPolicyMargin < c_dflt ? cand[best] : dflt;
    }
    static void to_record(const Frontier& f, double action, double kms, double* r) {
        r[0] = f.step;
        r[1] = (double)f.n_f;
        r[2] = (double)f.m_f;
        r[3] = (double)f.n_u;
        r[4] = (double)f.m_u;
        r[5] = f.n;
        r[6] = (double)f.m;
        r[7] = action;
        r[8] = kms;
    }
    // SSSP phase record: phase index, far-queue entries, mean weight, threshold, n, m, delta, ms.
    static void to_sssp_record(const Frontier& f, uint32_t wmean, uint32_t thr, double delta, double ms, double* r) {
        r[0] = f.step;
        r[1] = (double)f.n_f;
        r[2] = wmean;
        r[3] = thr;
        r[4] = 0;
        r[5] = f.n;
        r[6] = (double)f.m;
        r[7] = delta;
        r[8] = ms;
    }
    static void policy_row(uint32_t kind, const double* rec, double action, float* x);
    void record_sssp_phase(medusa_graph* g, const Frontier& f, uint32_t wmean, uint32_t thr, double delta,
                           double ms) const {
        if (!o.trace) return;
        TraceRec r;
        to_sssp_record(f, wmean, thr, delta, ms, r.f);
        g->trace.push_back(r);
    }

    // Appends one trace record when tracing is on.
    void record(medusa_graph* g, const Frontier& f, double action, double kernel_ms) const {
        if (!o.trace) return;
        TraceRec r;
        to_record(f, action, kernel_ms, r.f);
        g->trace.push_back(r);
    }
};

}  // namespace medusa
