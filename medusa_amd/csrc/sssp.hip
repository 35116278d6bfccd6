// SSSP superstep engine (uint32 weights and distances, exact).
//
// Reference: each superstep fills all |E| message slots with FLT_MAX
// (InitMessageBuffer, MultipleGPU/MultiPublicAPI.h:344-376), every updated vertex
// writes dist+w into the slot msgDstID of each out-edge (ssspEdgeProcAA,
// SSSP/UserOperators.h:10-29), a CUDPP segmented MIN scan reduces each vertex's
// slots (SSSP/MsgCombiner.h:73-165) and UpdateVertex keeps the minimum and sets
// `updated` (SSSP/UserOperators.h:151-168); loop at SSSP/main.cu:369-382.
//
// Here a superstep relaxes only the current frontier: dist[v] = atomicMin(dist[v],
// dist[u] + w) on uint32, so the "combine" is the atomic itself and no message
// array or scan exists.  Frontier vertices are bucketed near/far by a threshold
// that advances by `delta` when the near frontier empties (delta-stepping in the
// near-far form); delta = 0xFFFFFFFF degenerates to frontier Bellman-Ford.  Any
// delta gives the same (shortest) distances; only the number of supersteps and
// relaxations change.

#include <algorithm>
#include <chrono>

#include "engine.h"
#include "sched.h"

namespace medusa {

struct SsspCounters {
    uint32_t n_near;   // entries in the next near queue
    uint32_t n_far;    // entries in the far queue
    uint32_t far_min;  // smallest distance among far entries kept by a split
    uint32_t pad;
    unsigned long long near_deg;  // out-edges of the next near queue
    unsigned long long relax;     // edges relaxed this superstep
};

__global__ void k_sssp_init(uint32_t n, uint32_t* __restrict__ dist, uint32_t* __restrict__ stamp, uint32_t src,
                            uint32_t* __restrict__ near_q) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        dist[v] = v == src ? 0u : 0xFFFFFFFFu;
        stamp[v] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) near_q[0] = src;
}

__global__ void k_sssp_zero(SsspCounters* c, uint32_t keep_far) {
    c->n_near = 0;
    if (!keep_far) c->n_far = 0;
    c->far_min = 0xFFFFFFFFu;
    c->pad = 0;
    c->near_deg = 0;
    c->relax = 0;
}

// Appends v (where `take`) to queue q through one atomic per wave.
__device__ __forceinline__ void wave_append(bool take, uint32_t v, uint32_t* __restrict__ q, uint32_t* counter,
                                            uint32_t cap) {
    const unsigned long long mask = __ballot(take);
    if (!mask) return;
    uint32_t base = 0;
    const uint32_t leader = __builtin_ctzll(mask);
    if (lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader, kWave);
    if (take) {
        const uint32_t pos = base + lanes_below(mask);
        if (pos < cap) q[pos] = v;
    }
}

// One relaxation superstep over the near queue (wave-cooperative over out-edges).
__global__ __launch_bounds__(256) void k_sssp_relax(const uint32_t* __restrict__ q, uint32_t qn,
                                                    const uint32_t* __restrict__ row_ptr,
                                                    const uint32_t* __restrict__ col, const uint32_t* __restrict__ w,
                                                    uint32_t* __restrict__ dist, uint32_t* __restrict__ stamp,
                                                    uint32_t epoch, uint32_t thr, uint32_t* __restrict__ near_next,
                                                    uint32_t near_cap, uint32_t* __restrict__ far_q, uint32_t far_cap,
                                                    SsspCounters* __restrict__ ctr) {
    __shared__ uint32_t s_beg[4][kWave], s_off[4][kWave], s_du[4][kWave];
    const uint32_t lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const uint32_t total_waves = gridDim.x * 4;
    unsigned long long ndeg = 0, relax = 0;
    for (uint32_t base = (blockIdx.x * 4 + wv) * kWave; base < qn; base += total_waves * kWave) {
        const uint32_t i = base + lane;
        uint32_t beg = 0, deg = 0, du = 0;
        if (i < qn) {
            const uint32_t u = q[i];
            beg = row_ptr[u];
            deg = row_ptr[u + 1] - beg;
            du = dist[u];
        }
        const uint32_t incl = wave_incl_scan(deg);
        const uint32_t total = __shfl(incl, kWave - 1, kWave);
        s_beg[wv][lane] = beg;
        s_off[wv][lane] = incl - deg;
        s_du[wv][lane] = du;
        __builtin_amdgcn_wave_barrier();
        relax += total;
        for (uint32_t e0 = 0; e0 < total; e0 += kWave) {
            const uint32_t e = e0 + lane;
            bool to_near = false, to_far = false;
            uint32_t v = 0;
            if (e < total) {
                uint32_t lo = 0;
#pragma unroll
                for (uint32_t step = kWave / 2; step; step >>= 1)
                    if (s_off[wv][lo + step] <= e) lo += step;
                const uint32_t k = s_beg[wv][lo] + (e - s_off[wv][lo]);
                v = col[k];
                const unsigned long long nd64 = (unsigned long long)s_du[wv][lo] + w[k];
                if (nd64 < dist[v]) {
                    const uint32_t nd = (uint32_t)nd64;
                    const uint32_t old = atomicMin(&dist[v], nd);
                    if (nd < old) {
                        if (nd < thr) {
                            to_near = atomicMax(&stamp[v], epoch) < epoch;
                        } else {
                            to_far = true;
                        }
                    }
                }
            }
            wave_append(to_near, v, near_next, &ctr->n_near, near_cap);
            wave_append(to_far, v, far_q, &ctr->n_far, far_cap);
            if (to_near) ndeg += row_ptr[v + 1] - row_ptr[v];
        }
        __builtin_amdgcn_wave_barrier();
    }
    ndeg = wave_sum_u64(ndeg);
    if (lane == 0) {
        if (ndeg) atomicAdd(&ctr->near_deg, ndeg);
        if (relax) atomicAdd(&ctr->relax, relax);
    }
}

// Splits the far queue at a new threshold: entries in [old_thr, thr) move to the
// near queue (once, by stamp), entries >= thr stay far (compacted into far_out),
// entries below old_thr were already settled through the near queue.
__global__ __launch_bounds__(256) void k_sssp_split(const uint32_t* __restrict__ far_in, uint32_t nfar,
                                                    const uint32_t* __restrict__ dist, uint32_t* __restrict__ stamp,
                                                    uint32_t epoch, uint32_t old_thr, uint32_t thr,
                                                    const uint32_t* __restrict__ row_ptr,
                                                    uint32_t* __restrict__ near_q, uint32_t near_cap,
                                                    uint32_t* __restrict__ far_out, uint32_t far_cap,
                                                    SsspCounters* __restrict__ ctr) {
    unsigned long long ndeg = 0;
    uint32_t fmin = 0xFFFFFFFFu;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < nfar; i0 += stride) {  // uniform per wave
        const uint32_t i = i0 + threadIdx.x;
        bool to_near = false, keep = false;
        uint32_t v = 0;
        if (i < nfar) {
            v = far_in[i];
            const uint32_t d = dist[v];
            if (d >= old_thr && d < thr) to_near = atomicMax(&stamp[v], epoch) < epoch;
            else if (d >= thr) {
                keep = atomicMax(&stamp[v], epoch) < epoch;  // drop duplicates
                if (keep) fmin = min(fmin, d);
            }
        }
        wave_append(to_near, v, near_q, &ctr->n_near, near_cap);
        wave_append(keep, v, far_out, &ctr->n_far, far_cap);
        if (to_near) ndeg += row_ptr[v + 1] - row_ptr[v];
    }
    ndeg = wave_sum_u64(ndeg);
#pragma unroll
    for (int off = kWave / 2; off >= 1; off >>= 1) fmin = min(fmin, (uint32_t)__shfl_xor(fmin, off, kWave));
    if ((threadIdx.x & (kWave - 1)) == 0) {
        if (ndeg) atomicAdd(&ctr->near_deg, ndeg);
        if (fmin != 0xFFFFFFFFu) atomicMin(&ctr->far_min, fmin);
    }
}

// Rebuilds the far queue from scratch after an overflow: every reached vertex
// with distance >= thr.
__global__ __launch_bounds__(256) void k_sssp_refill(uint32_t n, const uint32_t* __restrict__ dist, uint32_t thr,
                                                     uint32_t* __restrict__ far_q, uint32_t far_cap,
                                                     SsspCounters* __restrict__ ctr) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v0 = blockIdx.x * blockDim.x; v0 < n; v0 += stride) {
        const uint32_t v = v0 + threadIdx.x;
        const bool take = v < n && dist[v] != 0xFFFFFFFFu && dist[v] >= thr;
        wave_append(take, v, far_q, &ctr->n_far, far_cap);
    }
}

// ---- device-driven loop ----------------------------------------------------------------
// Two kernels per superstep, both ending with a decision by their last workgroup:
//   relax+decide: relax the current near queue; when it comes out empty, the phase ends:
//                 pick the next bucket width (default, fixed, or the learned policy on
//                 the GPU), raise the threshold and mark a split as pending;
//   split+decide: when pending, move far entries below the new threshold to the near
//                 queue (a full rescan of the distances if the far queue overflowed);
//                 an empty bucket raises the threshold past the smallest far distance.
// kSsspBatch supersteps are captured once as a hipGraph and replayed until done.

constexpr uint32_t kSsspBatch = 32;
constexpr uint32_t kSsspRelaxGrid = 512;
constexpr uint32_t kSsspSplitGrid = 512;

struct SsspCtl {
    uint32_t done, epoch, thr, old_thr, delta, n_near, n_far, split_pending;
    uint32_t overflow, phase, supersteps, spare, mode, wmean, dflt_delta, near_sel;
    uint32_t far_sel, pad[3];
    unsigned long long n, m, relaxed, kernel_bytes;
};

struct SsspDevCounters {
    uint32_t n_near;   // near entries written by this kernel
    uint32_t n_far;    // far queue length (relax appends after the current entries)
    uint32_t far_out;  // far entries kept by a split
    uint32_t far_min;  // smallest distance kept by a split
    unsigned long long relax;
};

// Bucket width at a threshold advance; all threads of the block call it, thread 0's value counts.
__device__ uint32_t sssp_pick_delta(const SsspCtl& c, uint32_t n_far, const float* __restrict__ policy, float* s_x,
                                    float* s_a1, float* s_a2, float* s_best, uint32_t* s_pick) {
    if (c.mode != 2 || !policy) return c.dflt_delta;
    const double rec[MEDUSA_TRACE_FIELDS] = {(double)c.phase, (double)n_far, (double)c.wmean, (double)c.thr, 0.0,
                                             (double)c.n, (double)c.m, 0.0, 0.0};
    for (int k = -2; k <= 8; ++k) {
        const double d = fmax(1.0, ldexp((double)c.wmean, k));
        if (threadIdx.x == 0) {
            float x[MEDUSA_POLICY_INPUTS];
            delta_row(rec, d, x);
            for (int i = 0; i < MEDUSA_POLICY_INPUTS; ++i) s_x[i] = x[i];
        }
        __syncthreads();
        const float cost = mlp_eval_block(policy, s_x, s_a1, s_a2);
        if (threadIdx.x == 0 && (k == -2 || cost < *s_best)) {
            *s_best = cost;
            *s_pick = (uint32_t)d;
        }
        __syncthreads();
    }
    return *s_pick;
}

__device__ __forceinline__ uint32_t sat_add(uint32_t a, uint32_t b) {
    const unsigned long long s = (unsigned long long)a + b;
    return s >= 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)s;
}

__global__ __launch_bounds__(256) void k_sssp_relax_dev(SsspCtl* __restrict__ ctl, uint32_t* __restrict__ near0,
                                                        uint32_t* __restrict__ near1, uint32_t* __restrict__ far0,
                                                        uint32_t* __restrict__ far1, uint32_t far_cap,
                                                        const uint32_t* __restrict__ row_ptr,
                                                        const uint32_t* __restrict__ col,
                                                        const uint32_t* __restrict__ w, uint32_t* __restrict__ dist,
                                                        uint32_t* __restrict__ stamp,
                                                        SsspDevCounters* __restrict__ ctr) {
    __shared__ uint32_t s_beg[4][kWave], s_off[4][kWave], s_du[4][kWave];
    if (ctl->done) return;
    const uint32_t N = ctl->n_near, n = (uint32_t)ctl->n;
    const uint32_t lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    if (N > 0) {
        const uint32_t* q = ctl->near_sel ? near1 : near0;
        uint32_t* next = ctl->near_sel ? near0 : near1;
        uint32_t* farq = ctl->far_sel ? far1 : far0;
        const uint32_t epoch = ctl->epoch, thr = ctl->thr;
        unsigned long long relax = 0;
        for (uint32_t base = (blockIdx.x * 4 + wv) * kWave; base < N; base += gridDim.x * 4 * kWave) {
            const uint32_t i = base + lane;
            uint32_t beg = 0, deg = 0, du = 0;
            if (i < N) {
                const uint32_t u = q[i];
                beg = row_ptr[u];
                deg = row_ptr[u + 1] - beg;
                du = dist[u];
            }
            const uint32_t incl = wave_incl_scan(deg);
            const uint32_t total = __shfl(incl, kWave - 1, kWave);
            s_beg[wv][lane] = beg;
            s_off[wv][lane] = incl - deg;
            s_du[wv][lane] = du;
            __builtin_amdgcn_wave_barrier();
            relax += total;
            for (uint32_t e0 = 0; e0 < total; e0 += kWave) {
                const uint32_t e = e0 + lane;
                bool to_near = false, to_far = false;
                uint32_t v = 0;
                if (e < total) {
                    uint32_t lo = 0;
#pragma unroll
                    for (uint32_t step = kWave / 2; step; step >>= 1)
                        if (s_off[wv][lo + step] <= e) lo += step;
                    const uint32_t k = s_beg[wv][lo] + (e - s_off[wv][lo]);
                    v = col[k];
                    const unsigned long long nd64 = (unsigned long long)s_du[wv][lo] + w[k];
                    if (nd64 < dist[v]) {
                        const uint32_t nd = (uint32_t)nd64;
                        const uint32_t old = atomicMin(&dist[v], nd);
                        if (nd < old) {
                            if (nd < thr) to_near = atomicMax(&stamp[v], epoch) < epoch;
                            else to_far = true;
                        }
                    }
                }
                wave_append(to_near, v, next, &ctr->n_near, n);
                wave_append(to_far, v, farq, &ctr->n_far, far_cap);
            }
            __builtin_amdgcn_wave_barrier();
        }
        relax = wave_sum_u64(relax);
        if (lane == 0 && relax) atomicAdd(&ctr->relax, relax);
    }
}

// One workgroup, after the relaxation kernel: reads its counters (visible at the kernel
// boundary, so no fence is needed) and decides the next superstep.
__global__ __launch_bounds__(256) void k_sssp_relax_decide(SsspCtl* __restrict__ ctl,
                                                           SsspDevCounters* __restrict__ ctr, uint32_t far_cap,
                                                           const float* __restrict__ policy) {
    __shared__ uint32_t s_pick;
    __shared__ float s_x[MEDUSA_POLICY_INPUTS], s_a1[kMlpH], s_a2[kMlpH], s_best;
    if (ctl->done) return;
    const uint32_t N = ctl->n_near;
    const uint32_t n_next = ctr->n_near;
    uint32_t n_far = ctr->n_far;
    const unsigned long long relaxed = ctr->relax;
    SsspCtl c = *ctl;
    if (N > 0) {
        c.supersteps += 1;
        c.near_sel ^= 1u;
        c.relaxed += relaxed;
        c.kernel_bytes += 16ull * N + 12ull * relaxed + 12ull * n_next;
    }
    if (n_far > far_cap) {
        c.overflow = 1;
        n_far = far_cap;
    }
    c.n_near = n_next;
    c.n_far = n_far;
    c.epoch += 1;
    bool advance = false;
    if (n_next == 0 && !c.split_pending) {
        if (n_far == 0 && !c.overflow) c.done = 1;
        else advance = true;
    }
    if (advance) {
        const uint32_t d = sssp_pick_delta(c, n_far, policy, s_x, s_a1, s_a2, &s_best, &s_pick);
        c.old_thr = c.thr;
        c.delta = d;
        c.thr = sat_add(c.thr, d);
        c.split_pending = 1;
        c.phase += 1;
    }
    __syncthreads();  // every wave has read the control block and counters
    if (threadIdx.x == 0) {
        *ctl = c;
        ctr->n_near = 0;
        ctr->n_far = n_far;
        ctr->relax = 0;
        ctr->far_out = 0;
        ctr->far_min = 0xFFFFFFFFu;
    }
}

__global__ __launch_bounds__(256) void k_sssp_split_dev(SsspCtl* __restrict__ ctl, uint32_t* __restrict__ near0,
                                                        uint32_t* __restrict__ near1, uint32_t* __restrict__ far0,
                                                        uint32_t* __restrict__ far1, uint32_t far_cap,
                                                        const uint32_t* __restrict__ dist,
                                                        uint32_t* __restrict__ stamp,
                                                        SsspDevCounters* __restrict__ ctr) {
    if (ctl->done || !ctl->split_pending) return;
    const uint32_t n = (uint32_t)ctl->n, epoch = ctl->epoch, old_thr = ctl->old_thr, thr = ctl->thr;
    const bool rescan = ctl->overflow != 0;
    const uint32_t nin = rescan ? n : ctl->n_far;
    const uint32_t* far_in = ctl->far_sel ? far1 : far0;
    uint32_t* far_out = ctl->far_sel ? far0 : far1;
    uint32_t* near_q = ctl->near_sel ? near1 : near0;
    uint32_t fmin = 0xFFFFFFFFu;
    for (uint32_t i0 = blockIdx.x * blockDim.x; i0 < nin; i0 += gridDim.x * blockDim.x) {  // uniform per block
        const uint32_t i = i0 + threadIdx.x;
        bool to_near = false, keep = false;
        uint32_t v = 0;
        if (i < nin) {
            v = rescan ? i : far_in[i];
            const uint32_t d = dist[v];
            // on a rescan every unsettled vertex (distance >= old threshold) is a far entry
            if (d != 0xFFFFFFFFu && d >= old_thr) {
                if (d < thr) to_near = atomicMax(&stamp[v], epoch) < epoch;
                else {
                    keep = atomicMax(&stamp[v], epoch) < epoch;
                    if (keep) fmin = min(fmin, d);
                }
            }
        }
        wave_append(to_near, v, near_q, &ctr->n_near, n);
        wave_append(keep, v, far_out, &ctr->far_out, far_cap);
    }
#pragma unroll
    for (int off = kWave / 2; off >= 1; off >>= 1) fmin = min(fmin, (uint32_t)__shfl_xor(fmin, off, kWave));
    if ((threadIdx.x & (kWave - 1)) == 0 && fmin != 0xFFFFFFFFu) atomicMin(&ctr->far_min, fmin);
}

// One workgroup, after the split kernel (which ran iff a split was pending).
__global__ __launch_bounds__(256) void k_sssp_split_decide(SsspCtl* __restrict__ ctl,
                                                           SsspDevCounters* __restrict__ ctr,
                                                           const float* __restrict__ policy) {
    __shared__ uint32_t s_pick;
    __shared__ float s_x[MEDUSA_POLICY_INPUTS], s_a1[kMlpH], s_a2[kMlpH], s_best;
    if (ctl->done || !ctl->split_pending) return;
    const uint32_t n_near = ctr->n_near;
    const uint32_t kept = ctr->far_out;
    const uint32_t far_min = ctr->far_min;
    SsspCtl c = *ctl;
    c.far_sel ^= 1u;
    c.n_far = kept;
    c.overflow = 0;
    c.epoch += 1;
    c.n_near = n_near;
    bool again = false;
    if (n_near > 0) {
        c.split_pending = 0;
    } else if (kept == 0) {
        c.done = 1;
    } else {
        again = true;  // empty bucket: restart from the smallest far distance
    }
    if (again) {
        const uint32_t d = sssp_pick_delta(c, kept, policy, s_x, s_a1, s_a2, &s_best, &s_pick);
        c.old_thr = c.thr;
        c.delta = d;
        c.thr = sat_add(far_min > c.thr ? far_min : c.thr, d);
        c.split_pending = 1;
    }
    __syncthreads();  // every wave has read the control block and counters
    if (threadIdx.x == 0) {
        *ctl = c;
        ctr->n_near = 0;
        ctr->n_far = kept;
        ctr->far_out = 0;
        ctr->far_min = 0xFFFFFFFFu;
        ctr->relax = 0;
    }
}

__global__ void k_sssp_seed_ctl(SsspCtl* ctl, SsspDevCounters* c, uint32_t n, unsigned long long m, uint32_t thr,
                                uint32_t mode, uint32_t wmean, uint32_t dflt) {
    SsspCtl z{};
    z.epoch = 1;
    z.thr = thr;
    z.delta = thr;
    z.n_near = 1;
    z.mode = mode;
    z.wmean = wmean;
    z.dflt_delta = dflt;
    z.n = n;
    z.m = m;
    *ctl = z;
    c->n_near = 0;
    c->n_far = 0;
    c->far_out = 0;
    c->far_min = 0xFFFFFFFFu;
    c->relax = 0;
}

// Mean weight (sampled) and the default bucket width, a small multiple of it.
static void sssp_widths(const medusa_graph* g, uint32_t& wmean, uint32_t& dflt) {
    uint64_t wsum = 0, wcnt = 0;
    const uint64_t m = g->m;
    const uint64_t wstride = std::max<uint64_t>(1, m / 4096);
    for (uint64_t e = 0; e < m; e += wstride, ++wcnt) wsum += g->h_w[e];
    wmean = (uint32_t)std::max<uint64_t>(1, wcnt ? wsum / wcnt : 1);
    dflt = (uint32_t)std::min<uint64_t>(0xFFFFFFFEull, 2ull * wmean);
}

static void sssp_run_device(medusa_graph* g, uint32_t src, uint32_t* out, const medusa_opts& o, medusa_stats& st) {
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    const size_t far_cap = (size_t)std::max<uint64_t>(m, n) + n;
    hipStream_t s = g->stream;
    uint32_t wmean, dflt;
    sssp_widths(g, wmean, dflt);
    const bool learned = o.policy && policy_kind(o.policy) == MEDUSA_POLICY_DELTA && !o.delta;
    const uint32_t mode = learned ? 2u : (o.delta ? 1u : 0u);
    const uint32_t width = o.delta ? o.delta : dflt;
    const float* policy = learned ? policy_device_params(o.policy) : nullptr;
    Sched sched(o, Sched::kSssp);
    Frontier f{};
    f.n = n;
    f.m = m;
    f.n_f = 1;
    const uint32_t d0 = sched.sssp_delta(f, width, wmean, 0);
    const uint32_t thr0 = d0 >= 0xFFFFFFFFu - 1 ? 0xFFFFFFFFu : d0;
    g->sssp_ctl.ensure(1);
    SsspCtl* ctl = reinterpret_cast<SsspCtl*>(g->sssp_ctl.p);
    static_assert(sizeof(SsspCtl) <= sizeof(CtlStorage), "control block storage");
    SsspDevCounters* ctr = dev_counters<SsspDevCounters>(g);
    uint32_t* far0 = g->q_aux.p;
    uint32_t* far1 = g->q_aux.p + far_cap;

    GraphCache& gc = g->sssp_graph;
    const uintptr_t key[4] = {(uintptr_t)g->dist.p, (uintptr_t)policy, (uintptr_t)g->q_aux.p, (uintptr_t)g->w.p};
    if (!gc.exec || !std::equal(key, key + 4, gc.key)) {
        gc.reset();
        MDS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (uint32_t k = 0; k < kSsspBatch; ++k) {
            hipLaunchKernelGGL(k_sssp_relax_dev, dim3(kSsspRelaxGrid), dim3(256), 0, s, ctl, g->q_small.p,
                               g->q_large.p, far0, far1, (uint32_t)far_cap, g->row_ptr.p, g->col.p, g->w.p, g->dist.p,
                               g->stamp.p, ctr);
            hipLaunchKernelGGL(k_sssp_relax_decide, dim3(1), dim3(256), 0, s, ctl, ctr, (uint32_t)far_cap, policy);
            hipLaunchKernelGGL(k_sssp_split_dev, dim3(kSsspSplitGrid), dim3(256), 0, s, ctl, g->q_small.p,
                               g->q_large.p, far0, far1, (uint32_t)far_cap, (const uint32_t*)g->dist.p, g->stamp.p,
                               ctr);
            hipLaunchKernelGGL(k_sssp_split_decide, dim3(1), dim3(256), 0, s, ctl, ctr, policy);
        }
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(s, &graph);
        if (e == hipSuccess) e = hipGraphInstantiate(&gc.exec, graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        MDS_HIP(e);
        std::copy(key, key + 4, gc.key);
    }

    MDS_HIP(hipEventRecord(g->ev_begin, s));
    hipLaunchKernelGGL(k_sssp_init, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, n, g->dist.p, g->stamp.p, src,
                       g->q_small.p);
    MDS_LAUNCHED();
    hipLaunchKernelGGL(k_sssp_seed_ctl, dim3(1), dim3(1), 0, s, ctl, ctr, n, (unsigned long long)m, thr0, mode, wmean,
                       width);
    MDS_LAUNCHED();
    SsspCtl h{};
    // every replay either relaxes a non-empty queue or raises the threshold
    const uint64_t max_batches = 4ull * n / kSsspBatch + 1024;
    for (uint64_t batch = 0;; ++batch) {
        MDS_HIP(hipGraphLaunch(gc.exec, s));
        MDS_HIP(hipMemcpyAsync(g->h_ctr_raw.p, ctl, sizeof(SsspCtl), hipMemcpyDeviceToHost, s));
        MDS_HIP(hipStreamSynchronize(s));
        std::memcpy(&h, g->h_ctr_raw.p, sizeof(SsspCtl));
        if (h.done) break;
        MDS_CHECK(batch < max_batches, MEDUSA_ERR_INTERNAL, "sssp: superstep loop did not end");
    }
    MDS_HIP(hipEventRecord(g->ev_end, s));
    MDS_HIP(hipStreamSynchronize(s));
    float ms = 0;
    MDS_HIP(hipEventElapsedTime(&ms, g->ev_begin, g->ev_end));
    st.device_ms = ms;
    st.supersteps = h.supersteps;
    st.push_steps = h.supersteps;
    st.kernel_bytes = h.kernel_bytes;
    st.alg_bytes = h.kernel_bytes + 8ull * n;
    st.reserved[0] = (uint32_t)std::min<uint64_t>(h.relaxed, 0xFFFFFFFFull);
    st.reserved[1] = h.phase;
    copy_out(g, out, g->dist.p, (size_t)n * 4);
}

static void sssp_run_host(medusa_graph* g, uint32_t src, uint32_t* out, const medusa_opts& o, medusa_stats& st) {
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    const size_t far_cap = (size_t)std::max<uint64_t>(m, n) + n;
    g->dist.ensure(n);
    g->stamp.ensure(n);
    g->q_small.ensure(n);  // near (current)
    g->q_large.ensure(n);  // near (next)
    g->q_aux.ensure(2 * far_cap);  // far in / far out
    SsspCounters* C = dev_counters<SsspCounters>(g);
    hipStream_t s = g->stream;
    g->trace.clear();
    // default bucket width: a small multiple of the mean weight (sampled)
    uint64_t wsum = 0, wcnt = 0;
    const uint64_t wstride = std::max<uint64_t>(1, m / 4096);
    for (uint64_t e = 0; e < m; e += wstride, ++wcnt) wsum += g->h_w[e];
    const uint32_t wmean = (uint32_t)std::max<uint64_t>(1, wcnt ? wsum / wcnt : 1);
    const uint32_t delta_default = (uint32_t)std::min<uint64_t>(0xFFFFFFFEull, 2ull * wmean);
    Sched sched(o, Sched::kSssp);

    uint32_t* near_cur = g->q_small.p;
    uint32_t* near_nxt = g->q_large.p;
    uint32_t* far_a = g->q_aux.p;
    uint32_t* far_b = g->q_aux.p + far_cap;

    MDS_HIP(hipEventRecord(g->ev_begin, s));
    hipLaunchKernelGGL(k_sssp_init, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, n, g->dist.p, g->stamp.p, src,
                       near_cur);
    MDS_LAUNCHED();
    hipLaunchKernelGGL(k_sssp_zero, dim3(1), dim3(1), 0, s, C, 0u);
    MDS_LAUNCHED();
    uint32_t n_near = 1, n_far = 0, epoch = 1;
    uint64_t near_deg = g->h_row_ptr[src + 1] - g->h_row_ptr[src];
    uint64_t relaxed = 0;
    const bool timing = o.timing != 0;
    if (timing) ensure_events(g, 2);
    Frontier f{};
    f.n = n;
    f.m = m;
    f.n_f = 1;
    // a phase runs from one threshold advance until the near queue empties; its record
    // (features at the advance, chosen width, wall time) is what the delta policy learns from
    uint32_t phase = 0;
    Frontier phase_f = f;
    uint32_t phase_thr = 0;
    auto phase_t0 = std::chrono::steady_clock::now();
    uint32_t delta = sched.sssp_delta(f, o.delta ? o.delta : delta_default, wmean, 0);
    uint32_t thr = delta >= 0xFFFFFFFFu - 1 ? 0xFFFFFFFFu : delta;
    uint32_t phase_delta = delta;
    auto close_phase = [&]() {
        const double ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - phase_t0).count();
        sched.record_sssp_phase(g, phase_f, wmean, phase_thr, (double)phase_delta, ms);
        ++phase;
    };
    for (uint32_t step = 0;; ++step) {
        if (n_near == 0) {
            if (n_far == 0) break;
            close_phase();
            // advance the threshold and split the far queue; skip buckets that come out empty
            uint32_t old_thr = thr;
            f.step = phase;
            f.n_f = n_far;
            f.m_f = 0;
            phase_f = f;
            phase_thr = thr;
            phase_t0 = std::chrono::steady_clock::now();
            delta = sched.sssp_delta(f, o.delta ? o.delta : delta_default, wmean, thr);
            phase_delta = delta;
            for (;;) {
                const uint64_t nt = (uint64_t)thr + delta;
                thr = nt >= 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)nt;
                ++epoch;
                hipLaunchKernelGGL(k_sssp_zero, dim3(1), dim3(1), 0, s, C, 0u);
                MDS_LAUNCHED();
                hipLaunchKernelGGL(k_sssp_split, dim3(grid_for(n_far, 256, 4096)), dim3(256), 0, s, far_a, n_far,
                                   g->dist.p, g->stamp.p, epoch, old_thr, thr, g->row_ptr.p, near_cur, n, far_b,
                                   (uint32_t)far_cap, C);
                MDS_LAUNCHED();
                const SsspCounters sc = read_counters<SsspCounters>(g);
                std::swap(far_a, far_b);
                n_near = sc.n_near;
                n_far = std::min<uint32_t>(sc.n_far, (uint32_t)far_cap);
                near_deg = sc.near_deg;
                if (n_near || n_far == 0) break;
                old_thr = thr;
                if (sc.far_min != 0xFFFFFFFFu && sc.far_min > thr) thr = sc.far_min;
            }
            if (n_near == 0) break;
        }
        f.step = step;
        f.n_f = n_near;
        f.m_f = near_deg;
        ++epoch;
        if (timing) MDS_HIP(hipEventRecord(g->events[0], s));
        hipLaunchKernelGGL(k_sssp_zero, dim3(1), dim3(1), 0, s, C, 1u);
        MDS_LAUNCHED();
        const uint32_t waves = (n_near + kWave - 1) / kWave;
        // the far queue keeps growing at far_a[n_far..)
        hipLaunchKernelGGL(k_sssp_relax, dim3(grid_for(waves, 4, 16384)), dim3(256), 0, s, near_cur, n_near,
                           g->row_ptr.p, g->col.p, g->w.p, g->dist.p, g->stamp.p, epoch, thr, near_nxt, n, far_a,
                           (uint32_t)far_cap, C);
        MDS_LAUNCHED();
        if (timing) MDS_HIP(hipEventRecord(g->events[1], s));
        const SsspCounters c = read_counters<SsspCounters>(g);
        if (timing) {
            float ms = 0;
            MDS_HIP(hipEventElapsedTime(&ms, g->events[0], g->events[1]));
            st.kernel_ms += ms;
            ++st.kernel_launches;
        }
        relaxed += c.relax;
        // byte model: queue + offsets + dist of frontier, col + w + dist probe per edge,
        // atomic (8 B) + enqueue per improvement (upper bound: every enqueued entry)
        st.kernel_bytes += 16ull * n_near + 12ull * c.relax + 12ull * (c.n_near);
        ++st.supersteps;
        ++st.push_steps;
        std::swap(near_cur, near_nxt);
        n_near = c.n_near;
        near_deg = c.near_deg;
        if (c.n_far > far_cap) {
            // far queue overflowed: rebuild it from the distances
            hipLaunchKernelGGL(k_sssp_zero, dim3(1), dim3(1), 0, s, C, 0u);
            MDS_LAUNCHED();
            hipLaunchKernelGGL(k_sssp_refill, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, n, g->dist.p, thr,
                               far_a, (uint32_t)far_cap, C);
            MDS_LAUNCHED();
            n_far = read_counters<SsspCounters>(g).n_far;
        } else {
            n_far = c.n_far;
        }
    }
    close_phase();
    st.reserved[1] = phase;
    MDS_HIP(hipEventRecord(g->ev_end, s));
    MDS_HIP(hipStreamSynchronize(s));
    float ms = 0;
    MDS_HIP(hipEventElapsedTime(&ms, g->ev_begin, g->ev_end));
    st.device_ms = ms;
    copy_out(g, out, g->dist.p, (size_t)n * 4);
    st.alg_bytes = st.kernel_bytes + 8ull * n;
    st.reserved[0] = (uint32_t)std::min<uint64_t>(relaxed, 0xFFFFFFFFull);
}

// GTEPS numerator on the GPU: reached vertices and the out-edges they own (the same
// for every bucket width).
__global__ __launch_bounds__(256) void k_sssp_reached(uint32_t n, const uint32_t* __restrict__ dist,
                                                      const uint32_t* __restrict__ row_ptr,
                                                      unsigned long long* __restrict__ out) {
    __shared__ unsigned long long s_e[4], s_v[4];
    unsigned long long e = 0, c = 0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        if (dist[v] != 0xFFFFFFFFu) {
            e += row_ptr[v + 1] - row_ptr[v];
            ++c;
        }
    e = wave_sum_u64(e);
    c = wave_sum_u64(c);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        s_e[threadIdx.x / kWave] = e;
        s_v[threadIdx.x / kWave] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&out[0], s_e[0] + s_e[1] + s_e[2] + s_e[3]);
        atomicAdd(&out[1], s_v[0] + s_v[1] + s_v[2] + s_v[3]);
    }
}

__global__ void k_zero_u64x2(unsigned long long* p) {
    p[0] = 0;
    p[1] = 0;
}

}  // namespace medusa

using namespace medusa;

extern "C" int medusa_sssp(medusa_graph* g, uint32_t source, uint32_t* dist, const medusa_opts* opts,
                           medusa_stats* stats) {
    return guard([&] {
        enter(g);
        MDS_CHECK(dist, MEDUSA_ERR_ARG, "sssp: null output");
        MDS_CHECK(source < g->n, MEDUSA_ERR_ARG, "sssp: source %u out of range (n=%u)", source, g->n);
        MDS_CHECK(g->weighted, MEDUSA_ERR_STATE, "sssp: graph has no weights");
        medusa_opts o{};
        if (opts) o = *opts;
        const uint32_t n = g->n;
        const size_t far_cap = (size_t)std::max<uint64_t>(g->m, n) + n;
        g->dist.ensure(n);
        g->stamp.ensure(n);
        g->q_small.ensure(n);          // near queue A
        g->q_large.ensure(n);          // near queue B
        g->q_aux.ensure(2 * far_cap);  // far queues A and B
        medusa_stats st{};
        if (o.timing || o.trace) sssp_run_host(g, source, dist, o, st);
        else sssp_run_device(g, source, dist, o, st);
        unsigned long long* red = reinterpret_cast<unsigned long long*>(g->sssp_ctl.p) + 30;
        if (!g->sssp_ctl.p) {
            g->sssp_ctl.ensure(1);
            red = reinterpret_cast<unsigned long long*>(g->sssp_ctl.p) + 30;
        }
        hipLaunchKernelGGL(k_zero_u64x2, dim3(1), dim3(1), 0, g->stream, red);
        MDS_LAUNCHED();
        hipLaunchKernelGGL(k_sssp_reached, dim3(grid_for(n, 256, 1024)), dim3(256), 0, g->stream, n,
                           (const uint32_t*)g->dist.p, g->row_ptr.p, red);
        MDS_LAUNCHED();
        unsigned long long hr[2];
        MDS_HIP(hipMemcpyAsync(hr, red, sizeof hr, hipMemcpyDeviceToHost, g->stream));
        MDS_HIP(hipStreamSynchronize(g->stream));
        st.edges_traversed = hr[0];
        st.reached = (uint32_t)hr[1];
        if (stats) *stats = st;
    });
}
