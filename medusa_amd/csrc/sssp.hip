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

static void sssp_run(medusa_graph* g, uint32_t src, uint32_t* out, const medusa_opts& o, medusa_stats& st) {
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
    uint32_t delta = o.delta ? o.delta : delta_default;
    uint32_t thr = delta >= 0xFFFFFFFFu - 1 ? 0xFFFFFFFFu : delta;
    const bool timing = o.timing != 0;
    if (timing) ensure_events(g, 2);
    Frontier f{};
    f.n = n;
    f.m = m;
    for (uint32_t step = 0;; ++step) {
        if (n_near == 0) {
            if (n_far == 0) break;
            // advance the threshold past the smallest far distance and split
            uint32_t old_thr = thr;
            for (;;) {
                f.step = step;
                f.n_f = n_far;
                f.m_f = 0;
                delta = sched.sssp_delta(f, o.delta ? o.delta : delta_default);
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
                if (n_near) break;
                if (n_far == 0) break;
                // skip empty buckets: restart just below the smallest remaining far distance
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
        double kms = 0;
        if (timing) {
            float ms = 0;
            MDS_HIP(hipEventElapsedTime(&ms, g->events[0], g->events[1]));
            kms = ms;
            st.kernel_ms += ms;
            ++st.kernel_launches;
        }
        relaxed += c.relax;
        // byte model: queue + offsets + dist of frontier, col + w + dist probe per edge,
        // atomic (8 B) + enqueue per improvement (upper bound: every enqueued entry)
        st.kernel_bytes += 16ull * n_near + 12ull * c.relax + 12ull * (c.n_near);
        sched.record(g, f, (double)delta, kms);
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
    MDS_HIP(hipEventRecord(g->ev_end, s));
    MDS_HIP(hipStreamSynchronize(s));
    float ms = 0;
    MDS_HIP(hipEventElapsedTime(&ms, g->ev_begin, g->ev_end));
    st.device_ms = ms;
    copy_out(g, out, g->dist.p, (size_t)n * 4);
    // GTEPS numerator: out-edges of reached vertices (same for every delta)
    std::vector<uint32_t> hd(n);
    MDS_HIP(hipMemcpy(hd.data(), g->dist.p, (size_t)n * 4, hipMemcpyDeviceToHost));
    uint64_t mr = 0, reached = 0;
    for (uint32_t v = 0; v < n; ++v)
        if (hd[v] != 0xFFFFFFFFu) {
            mr += g->h_row_ptr[v + 1] - g->h_row_ptr[v];
            ++reached;
        }
    st.edges_traversed = mr;
    st.reached = (uint32_t)reached;
    st.alg_bytes = st.kernel_bytes + 8ull * n;
    st.reserved[0] = (uint32_t)std::min<uint64_t>(relaxed, 0xFFFFFFFFull);
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
        medusa_stats st{};
        sssp_run(g, source, dist, o, st);
        if (stats) *stats = st;
    });
}
