// Multi-source BFS: up to 64 sources advance together, one bit per source in a
// 64-bit mask per vertex (frontier, visited, next).
//
// No reference counterpart: "MBFS" in the reference is multi-GPU BFS
// (HowToBuild:13); the semantics here are those of k independent BFS runs
// (Algorithm/UserOperators.h:17-36 per source), checked against exactly that.
//
// Superstep:
//   push: active vertices (frontier mask != 0) OR their mask into next[v] of every
//         out-neighbour that still lacks some of those bits (64-bit atomicOr); vertices
//         up to `large` out-degree go through the wave-cooperative queue, hubs above it
//         are split over whole workgroups (the threshold is autotunable);
//   pull: every vertex not yet visited by all sources ORs the frontier masks of its
//         in-neighbours, stopping early once every missing bit has been found; vertices are
//         visited in in-degree order over ELL slabs of source ids (the pull layout), so the
//         index loads are coalesced;
//   update: new = next & ~visited; visited |= new; frontier = new; level[s][v] =
//         step+1 for each new bit s; the two active queues and the counters are rebuilt
//         with ballot + in-workgroup prefix scan.  Seeding is the same update at level 0
//         after the sources' bits are placed in next[].
//
// Drivers as for BFS: device-driven (a one-workgroup decide kernel after each update
// writes the next superstep into a control block; batches of supersteps are captured as
// hipGraphs and replayed until done, the first batch sized by the previous run on the
// handle) and host-driven (timing or tracing) over the same kernels.

#include <algorithm>
#include <cstring>

#include "engine.h"
#include "sched.h"

namespace medusa {

constexpr uint32_t kMbfsLargeDefault = 1024;
constexpr uint32_t kMbfsGridDefault = 2048;
constexpr uint32_t kMbfsBatch = 8;
constexpr uint32_t kMbfsSplit = 8;  // minimum workgroup items per hub

struct MbfsCounters {
    MDS_COUNTER(uint32_t, n_small);    // active vertices with 1 <= outdeg <= large
    MDS_COUNTER(uint32_t, n_large);    // active vertices with outdeg > large
    MDS_COUNTER(uint32_t, n_active);   // all active vertices
    MDS_COUNTER(uint32_t, saturated);  // a level >= 255 was clamped to 0xFF
    MDS_COUNTER(unsigned long long, active_deg);
    MDS_COUNTER(unsigned long long, new_bits);
    MDS_COUNTER(unsigned long long, reached_edges);  // sum over new (source, vertex) pairs of outdeg(vertex)
    MDS_COUNTER(unsigned long long, probes);
    MDS_COUNTER(unsigned long long, deg_large);      // out-degree sum of the hub queue
};

struct MbfsCtl {
    int32_t step;
    uint32_t pull, done, n_small, n_large, hub_wgs, mode, saturated;
    uint32_t supersteps, push_steps, pull_steps, n_active;
    unsigned long long n, m, n_f, m_f, n_u, m_u, reached, reached_edges, kernel_bytes, alg_bytes, full;
};

__global__ void k_mbfs_zero(MbfsCounters* c) {
    c->n_small = 0;
    c->n_large = 0;
    c->n_active = 0;
    c->saturated = 0;
    c->active_deg = 0;
    c->new_bits = 0;
    c->reached_edges = 0;
    c->probes = 0;
    c->deg_large = 0;
}

__global__ void k_mbfs_fill(uint32_t n, uint64_t level_words, uint32_t* __restrict__ level32,
                            unsigned long long* __restrict__ visited, unsigned long long* __restrict__ frontier,
                            unsigned long long* __restrict__ next) {
    const uint64_t tid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x, stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = tid; i < level_words; i += stride) level32[i] = 0xFFFFFFFFu;
    for (uint64_t v = tid; v < n; v += stride) {
        visited[v] = 0;
        frontier[v] = 0;
        next[v] = 0;
    }
}

// Places the k source bits (k <= 64, one thread) into next[]; the level-0 update does the rest.
__global__ void k_mbfs_seed(const uint32_t* __restrict__ srcs, uint32_t k, unsigned long long* __restrict__ next) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (uint32_t s = 0; s < k; ++s) next[srcs[s]] |= 1ull << s;
}

// Workgroups of a push expansion given to the hub queue (as for BFS).
__host__ __device__ inline uint32_t mbfs_hub_wgs(uint32_t grid, uint32_t n_small, uint32_t n_large,
                                                 unsigned long long deg_large, unsigned long long deg_sum) {
    if (!n_large) return 0;
    if (!n_small) return grid;
    unsigned long long h = deg_sum ? (unsigned long long)grid * deg_large / deg_sum : 1ull;
    if (h < 1) h = 1;
    if (h > grid - 1) h = grid - 1;
    return (uint32_t)h;
}

// ---- superstep bodies ---------------------------------------------------------------

// Wave-cooperative push; wave w of the nwaves taking part holds queue entries
// w, w + nwaves, ... one per lane.
__device__ __forceinline__ void mbfs_push_small(const uint32_t* __restrict__ q, uint32_t qn,
                                                const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col,
                                                const unsigned long long* __restrict__ frontier,
                                                const unsigned long long* __restrict__ visited,
                                                unsigned long long* __restrict__ next, uint32_t wave0, uint32_t nwaves) {
    __shared__ uint32_t s_beg[4][kWave], s_off[4][kWave];
    __shared__ unsigned long long s_mask[4][kWave];
    const uint32_t lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    for (uint64_t r = 0; wave0 + r * kWave * nwaves < qn; ++r) {  // uniform across the wave
        const uint64_t i = wave0 + (r * kWave + lane) * nwaves;
        uint32_t beg = 0, deg = 0;
        unsigned long long fm = 0;
        if (i < qn) {
            const uint32_t u = q[i];
            beg = row_ptr[u];
            deg = row_ptr[u + 1] - beg;
            fm = frontier[u];
        }
        const uint32_t incl = wave_incl_scan(deg);
        const uint32_t total = __shfl(incl, kWave - 1, kWave);
        s_beg[wv][lane] = beg;
        s_off[wv][lane] = incl - deg;
        s_mask[wv][lane] = fm;
        __builtin_amdgcn_wave_barrier();
        for (uint32_t e = lane; e < total; e += kWave) {
            uint32_t lo = 0;
#pragma unroll
            for (uint32_t step = kWave / 2; step; step >>= 1)
                if (s_off[wv][lo + step] <= e) lo += step;
            const uint32_t v = col[s_beg[wv][lo] + (e - s_off[wv][lo])];
            const unsigned long long add = s_mask[wv][lo] & ~visited[v];
            if (add && (next[v] & add) != add) atomicOr(&next[v], add);
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// Hubs: each queue vertex split into `split` edge ranges, one per workgroup item.
__device__ __forceinline__ void mbfs_push_large(const uint32_t* __restrict__ q, uint32_t qn,
                                                const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col,
                                                const unsigned long long* __restrict__ frontier,
                                                const unsigned long long* __restrict__ visited,
                                                unsigned long long* __restrict__ next, uint32_t wg0, uint32_t nwg) {
    uint32_t split = nwg / qn;
    split = split < kMbfsSplit ? kMbfsSplit : (split > 4096u ? 4096u : split);
    for (uint64_t w = wg0; w < (uint64_t)qn * split; w += nwg) {
        const uint32_t u = q[w / split], part = (uint32_t)(w % split);
        const uint32_t b = row_ptr[u], deg = row_ptr[u + 1] - b;
        const unsigned long long fm = frontier[u];
        const uint32_t per = (deg + split - 1) / split;
        const uint32_t lo = b + min(deg, part * per), hi = b + min(deg, (part + 1) * per);
        for (uint32_t k = lo + threadIdx.x; k < hi; k += blockDim.x) {
            const uint32_t v = col[k];
            const unsigned long long add = fm & ~visited[v];
            if (add && (next[v] & add) != add) atomicOr(&next[v], add);
        }
    }
}

constexpr uint32_t kMbfsPullGroup = 4;  // in-edges a lane probes per round

// Bottom-up over the pull layout: a vertex missing some sources ORs its in-neighbours'
// frontier masks until every missing bit is found.  Work item w of the layout is 256
// lane-vertices of one equal-in-degree run (their k-th sources are one coalesced ELL load;
// kMbfsPullGroup in-edges per round, loads issued together, merged one by one with the same
// early exit and probe count as a plain loop) or 4 wave-vertices of in-degree > 64 (the wave
// scans 64 in-edges per step and ORs them by shuffles).  Workgroups stride over the items.
// `probes` counts in-edges examined up to and including the last one needed.
__device__ __forceinline__ void mbfs_pull_ell(unsigned long long full, const PrRun* __restrict__ runs,
                                              const uint32_t* __restrict__ wg_run, uint32_t items,
                                              const uint32_t* __restrict__ perm, const uint32_t* __restrict__ ell,
                                              const uint32_t* __restrict__ cptr, const uint32_t* __restrict__ csrc,
                                              const unsigned long long* __restrict__ frontier,
                                              const unsigned long long* __restrict__ visited,
                                              unsigned long long* __restrict__ next, MbfsCounters* __restrict__ ctr) {
    const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    unsigned long long probes = 0;
    for (uint32_t w = blockIdx.x; w < items; w += gridDim.x) {  // uniform per block
        const PrRun R = runs[wg_run[w]];
        const uint32_t local = w - R.wg_begin;
        if (R.kind == kPrWave) {
            const uint32_t p = R.begin + local * 4 + wave;
            if (p >= R.begin + R.count) continue;  // uniform across the wave
            const uint32_t v = perm[p];
            const unsigned long long vis = visited[v];
            if (vis == full) continue;
            const uint32_t vb = cptr[v], ve = cptr[v + 1];
            unsigned long long acc = 0;
            for (uint32_t k0 = vb; k0 < ve; k0 += kWave) {
                const uint32_t k = k0 + lane;
                unsigned long long x = k < ve ? frontier[csrc[k]] : 0ull;
#pragma unroll
                for (int off = kWave / 2; off >= 1; off >>= 1) x |= __shfl_xor(x, off, kWave);
                acc |= x;
                if (lane == 0) probes += min(ve - k0, (uint32_t)kWave);
                if ((acc | vis) == full) break;
            }
            acc &= ~vis;
            if (lane == 0 && acc) next[v] = acc;
            continue;
        }
        const uint32_t i = local * 256 + threadIdx.x;
        if (i >= R.count) continue;
        const uint32_t v = perm[R.begin + i];
        const unsigned long long vis = visited[v];
        if (vis == full) continue;
        const uint32_t* col = ell + R.ell_base + i;
        const uint32_t deg = R.deg, st = R.ell_stride;
        unsigned long long acc = 0;
        bool all = false;
        for (uint32_t k0 = 0; k0 < deg && !all; k0 += kMbfsPullGroup) {
            uint32_t u[kMbfsPullGroup];
            unsigned long long f[kMbfsPullGroup];
#pragma unroll
            for (uint32_t j = 0; j < kMbfsPullGroup; ++j) u[j] = k0 + j < deg ? col[(size_t)(k0 + j) * st] : 0u;
#pragma unroll
            for (uint32_t j = 0; j < kMbfsPullGroup; ++j) f[j] = k0 + j < deg ? frontier[u[j]] : 0ull;
#pragma unroll
            for (uint32_t j = 0; j < kMbfsPullGroup; ++j) {
                if (all || k0 + j >= deg) continue;
                acc |= f[j];
                ++probes;
                all = (acc | vis) == full;
            }
        }
        acc &= ~vis;
        if (acc) next[v] = acc;
    }
    // one atomic per workgroup (every wave of the block reaches this point)
    __shared__ unsigned long long s_probe[4];
    probes = wave_sum_u64(probes);
    if (lane == 0) s_probe[threadIdx.x / kWave] = probes;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long t = s_probe[0] + s_probe[1] + s_probe[2] + s_probe[3];
        if (t) atomicAdd(&ctr->probes, t);
    }
}

__global__ __launch_bounds__(256) void k_mbfs_expand(const MbfsCtl* __restrict__ ctl,
                                                     const uint32_t* __restrict__ q_small,
                                                     const uint32_t* __restrict__ q_large,
                                                     const uint32_t* __restrict__ row_ptr,
                                                     const uint32_t* __restrict__ col, const uint32_t* __restrict__ cptr,
                                                     const uint32_t* __restrict__ csrc,
                                                     const unsigned long long* __restrict__ frontier,
                                                     const unsigned long long* __restrict__ visited,
                                                     unsigned long long* __restrict__ next,
                                                     const PrRun* __restrict__ runs, const uint32_t* __restrict__ wg_run,
                                                     uint32_t items, const uint32_t* __restrict__ perm,
                                                     const uint32_t* __restrict__ ell, MbfsCounters* __restrict__ ctr) {
    if (ctl->done) return;
    if (ctl->pull) {
        mbfs_pull_ell(ctl->full, runs, wg_run, items, perm, ell, cptr, csrc, frontier, visited, next, ctr);
        return;
    }
    const uint32_t hub = ctl->hub_wgs;
    if (blockIdx.x < hub) {
        mbfs_push_large(q_large, ctl->n_large, row_ptr, col, frontier, visited, next, blockIdx.x, hub);
        return;
    }
    const uint32_t wg = blockIdx.x - hub, nwg = gridDim.x - hub;
    mbfs_push_small(q_small, ctl->n_small, row_ptr, col, frontier, visited, next, wg * 4 + threadIdx.x / kWave,
                    nwg * 4);
}

constexpr uint32_t kMbfsTile = 4 * kWave * 4;  // vertices per workgroup tile (4 chunks per wave)

// new = next & ~visited for every vertex: levels, masks, the two active queues, counters.
__global__ __launch_bounds__(256) void k_mbfs_update(const MbfsCtl* __restrict__ ctl, int32_t level_fixed,
                                                     uint32_t n, uint32_t large, const uint32_t* __restrict__ row_ptr,
                                                     unsigned long long* __restrict__ visited,
                                                     unsigned long long* __restrict__ frontier,
                                                     unsigned long long* __restrict__ next,
                                                     uint8_t* __restrict__ level, uint32_t* __restrict__ q_small,
                                                     uint32_t* __restrict__ q_large, MbfsCounters* __restrict__ ctr) {
    __shared__ uint32_t s_cnt[4][3];
    __shared__ unsigned long long s_sum[4][4];
    __shared__ uint32_t s_base[2];
    if (ctl && ctl->done) return;
    const int32_t level_val = ctl ? ctl->step + 1 : level_fixed;
    const uint32_t lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const uint8_t lv = level_val >= 255 ? (uint8_t)0xFF : (uint8_t)level_val;
    for (uint32_t tile = blockIdx.x; (uint64_t)tile * kMbfsTile < n; tile += gridDim.x) {
        unsigned long long msm[4], mlg[4];
        uint32_t csm = 0, clg = 0, cact = 0;
        unsigned long long adeg = 0, bits = 0, redges = 0, dlg = 0;
        // loads of the tile first (independent across chunks, so they overlap)
        unsigned long long raw[4], fr[4], vis[4];
        uint32_t dg[4];
#pragma unroll
        for (uint32_t c = 0; c < 4; ++c) {
            const uint32_t v = (tile * 16 + c * 4 + wv) * kWave + lane;
            raw[c] = v < n ? next[v] : 0ull;
            fr[c] = v < n ? frontier[v] : 0ull;
        }
#pragma unroll
        for (uint32_t c = 0; c < 4; ++c) {
            const uint32_t v = (tile * 16 + c * 4 + wv) * kWave + lane;
            vis[c] = raw[c] ? visited[v] : 0ull;
        }
#pragma unroll
        for (uint32_t c = 0; c < 4; ++c) {
            const uint32_t v = (tile * 16 + c * 4 + wv) * kWave + lane;
            dg[c] = (raw[c] & ~vis[c]) ? row_ptr[v + 1] - row_ptr[v] : 0u;
        }
#pragma unroll
        for (uint32_t c = 0; c < 4; ++c) {
            const uint32_t v = (tile * 16 + c * 4 + wv) * kWave + lane;
            unsigned long long nx = 0;
            if (v < n) {
                if (raw[c]) {
                    next[v] = 0;
                    nx = raw[c] & ~vis[c];
                }
                if (nx) visited[v] = vis[c] | nx;
                if (fr[c] != nx) frontier[v] = nx;
            }
            const unsigned long long amask = __ballot(nx != 0);
            msm[c] = 0;
            mlg[c] = 0;
            if (amask) {
                const uint32_t deg = nx ? dg[c] : 0u;
                msm[c] = __ballot(nx != 0 && deg > 0 && deg <= large);
                mlg[c] = __ballot(nx != 0 && deg > large);
                adeg += deg;
                dlg += deg > large ? deg : 0u;
                bits += __popcll(nx);
                redges += (unsigned long long)__popcll(nx) * deg;
                // level bytes: row s gets this wave's 64 consecutive vertices
                unsigned long long any = nx;
#pragma unroll
                for (int off = kWave / 2; off >= 1; off >>= 1) any |= __shfl_xor(any, off, kWave);
                while (any) {
                    const uint32_t s = __builtin_ctzll(any);
                    any &= any - 1;
                    if ((nx >> s) & 1ull) level[(size_t)s * n + v] = lv;
                }
                cact += __popcll(amask);
                csm += __popcll(msm[c]);
                clg += __popcll(mlg[c]);
            }
        }
        adeg = wave_sum_u64(adeg);
        bits = wave_sum_u64(bits);
        redges = wave_sum_u64(redges);
        dlg = wave_sum_u64(dlg);
        if (lane == 0) {
            s_cnt[wv][0] = csm;
            s_cnt[wv][1] = clg;
            s_cnt[wv][2] = cact;
            s_sum[wv][0] = adeg;
            s_sum[wv][1] = bits;
            s_sum[wv][2] = redges;
            s_sum[wv][3] = dlg;
        }
        __syncthreads();
        if (threadIdx.x == 0) {  // one set of atomics per tile
            uint32_t tsm = 0, tlg = 0, tact = 0;
            unsigned long long t[4] = {0, 0, 0, 0};
            for (int w = 0; w < 4; ++w) {
                tsm += s_cnt[w][0];
                tlg += s_cnt[w][1];
                tact += s_cnt[w][2];
                for (int j = 0; j < 4; ++j) t[j] += s_sum[w][j];
            }
            s_base[0] = tsm ? atomicAdd(&ctr->n_small, tsm) : 0u;
            s_base[1] = tlg ? atomicAdd(&ctr->n_large, tlg) : 0u;
            if (tact) {
                atomicAdd(&ctr->n_active, tact);
                atomicAdd(&ctr->active_deg, t[0]);
                atomicAdd(&ctr->new_bits, t[1]);
                atomicAdd(&ctr->reached_edges, t[2]);
                if (t[3]) atomicAdd(&ctr->deg_large, t[3]);
            }
            if (level_val >= 255 && tact) ctr->saturated = 1;
        }
        __syncthreads();
        uint32_t osm = s_base[0], olg = s_base[1];
        for (uint32_t w = 0; w < wv; ++w) {
            osm += s_cnt[w][0];
            olg += s_cnt[w][1];
        }
#pragma unroll
        for (uint32_t c = 0; c < 4; ++c) {
            const uint32_t v = (tile * 16 + c * 4 + wv) * kWave + lane;
            if ((msm[c] >> lane) & 1ull) q_small[osm + lanes_below(msm[c])] = v;
            if ((mlg[c] >> lane) & 1ull) q_large[olg + lanes_below(mlg[c])] = v;
            osm += __popcll(msm[c]);
            olg += __popcll(mlg[c]);
        }
        __syncthreads();
    }
}

// Direction of the next superstep: fixed, learned (two MLP evaluations by the block), or the
// direction-optimising heuristic.  All threads of the block call it.
__device__ bool mbfs_direction(const MbfsCtl& c, bool pull, const float* __restrict__ policy, MlpShared& m,
                               bool staged) {
    if (c.mode == 1) return false;
    if (c.mode == 2) return true;
    if (c.mode == 3 && policy &&
        bfs_policy_consulted(pull, (double)c.n_f, (double)c.m_f, (double)c.m_u, (double)c.n, Sched::kAlphaMbfs,
                             Sched::kBeta)) {
        const double rec[MEDUSA_TRACE_FIELDS] = {(double)c.step, (double)c.n_f, (double)c.m_f, (double)c.n_u,
                                                 (double)c.m_u, (double)c.n, (double)c.m, 0.0, 0.0};
        policy_direction_costs(rec, policy, m, staged);
        const bool plain = !pull ? (double)c.m_f > (double)c.m_u / Sched::kAlphaMbfs
                                 : !((double)c.n_f < (double)c.n / Sched::kBeta);
        return policy_pull(m.out[0], m.out[1], plain);
    }
    if (!pull) return (double)c.m_f > (double)c.m_u / Sched::kAlphaMbfs;
    return !((double)c.n_f < (double)c.n / Sched::kBeta);
}

// One workgroup after an update (or the seeding update when `first`): folds the counters
// into the control block and decides the next superstep.
__global__ __launch_bounds__(256) void k_mbfs_decide(MbfsCtl* __restrict__ ctl, MbfsCounters* __restrict__ ctr,
                                                     const float* __restrict__ policy, uint32_t grid, uint32_t first) {
    __shared__ MlpShared s_mlp;
    MbfsCounters k;
    k.n_small = load_now(&ctr->n_small);
    k.n_large = load_now(&ctr->n_large);
    k.n_active = load_now(&ctr->n_active);
    k.saturated = load_now(&ctr->saturated);
    k.active_deg = load_now(&ctr->active_deg);
    k.new_bits = load_now(&ctr->new_bits);
    k.reached_edges = load_now(&ctr->reached_edges);
    k.probes = load_now(&ctr->probes);
    k.deg_large = load_now(&ctr->deg_large);
    MbfsCtl c = load_now_block(ctl);
    if (c.done) return;
    const unsigned long long n = c.n;
    if (!first) {
        // byte model of the superstep just finished: push = queue + offsets + mask per active
        // vertex, col + visited + next probe per edge; pull = offsets + visited per vertex,
        // source + frontier mask per probe; update = 24n
        const unsigned long long kb = c.pull ? 4ull * (n + 1) + 8ull * n + 12ull * k.probes
                                             : 20ull * c.n_f + 20ull * c.m_f;
        c.kernel_bytes += kb;
        c.alg_bytes += kb + 24ull * n + k.new_bits;
        c.supersteps += 1;
        if (c.pull) c.pull_steps += 1; else c.push_steps += 1;
        c.step += 1;
    }
    c.saturated |= k.saturated;
    c.reached += k.new_bits;
    c.reached_edges += k.reached_edges;
    c.n_u -= k.new_bits;
    c.m_u -= k.reached_edges;
    c.n_f = k.n_active;
    c.m_f = k.active_deg;
    c.n_small = k.n_small;
    c.n_large = k.n_large;
    c.n_active = k.n_active;
    c.hub_wgs = mbfs_hub_wgs(grid, k.n_small, k.n_large, k.deg_large, k.active_deg);
    if (k.n_active == 0) c.done = 1;
    const bool pull = c.done ? false : mbfs_direction(c, first ? false : c.pull != 0, policy, s_mlp, false);
    c.pull = pull ? 1u : 0u;
    __syncthreads();  // every wave has read the control block and counters
    if (threadIdx.x == 0) {
        *ctl = c;
        ctr->n_small = 0;
        ctr->n_large = 0;
        ctr->n_active = 0;
        ctr->saturated = 0;
        ctr->active_deg = 0;
        ctr->new_bits = 0;
        ctr->reached_edges = 0;
        ctr->probes = 0;
        ctr->deg_large = 0;
    }
}

__global__ void k_mbfs_init_ctl(MbfsCtl* ctl, uint32_t n, unsigned long long m, uint32_t k, uint32_t mode) {
    MbfsCtl z{};
    z.mode = mode;
    z.n = n;
    z.m = m;
    z.n_u = (unsigned long long)n * k;
    z.m_u = m * k;
    z.full = k == 64 ? ~0ull : ((1ull << k) - 1);
    *ctl = z;
}

// Host-driven superstep: the host's decision written into the control block.
__global__ void k_mbfs_set_ctl(MbfsCtl* ctl, int32_t step, uint32_t pull, uint32_t n_small, uint32_t n_large,
                               uint32_t hub_wgs) {
    ctl->step = step;
    ctl->pull = pull;
    ctl->done = 0;
    ctl->n_small = n_small;
    ctl->n_large = n_large;
    ctl->hub_wgs = hub_wgs;
}

static uint32_t mbfs_large(const medusa_graph* g, const medusa_opts& o) {
    return o.large_degree ? o.large_degree : (g->tuned.mbfs_large ? g->tuned.mbfs_large : kMbfsLargeDefault);
}
static uint32_t mbfs_grid(const medusa_graph* g, const medusa_opts& o) {
    return o.grid ? o.grid : (g->tuned.mbfs_grid ? g->tuned.mbfs_grid : kMbfsGridDefault);
}

static void mbfs_run(medusa_graph* g, const uint32_t* srcs, uint32_t k, uint8_t* out, const medusa_opts& o,
                     medusa_stats& st) {
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    const uint32_t large = mbfs_large(g, o), grid = mbfs_grid(g, o);
    g->q_small.ensure(n);
    g->q_large.ensure(n);
    g->q_aux.ensure(64);
    DBuf<unsigned long long>& masks = g->bitmap;
    masks.ensure(3 * (size_t)n);
    unsigned long long* visited = masks.p;
    unsigned long long* frontier = masks.p + n;
    unsigned long long* next = masks.p + 2 * (size_t)n;
    const size_t level_bytes = (size_t)k * n;
    DBuf<uint32_t>& lvbuf = g->stamp;  // reused as byte storage for the levels
    lvbuf.ensure((level_bytes + 3) / 4);
    uint8_t* level = reinterpret_cast<uint8_t*>(lvbuf.p);
    MbfsCounters* C = dev_counters<MbfsCounters>(g);
    static_assert(sizeof(MbfsCtl) <= sizeof(CtlStorage), "control block storage");
    g->mbfs_ctl.ensure(1);
    MbfsCtl* ctl = reinterpret_cast<MbfsCtl*>(g->mbfs_ctl.p);
    hipStream_t s = g->stream;
    const bool host_driven = o.timing || o.trace;
    uint32_t mode = 0;
    if (o.direction == MEDUSA_DIR_PUSH) mode = 1;
    else if (o.direction == MEDUSA_DIR_PULL) mode = 2;
    else if (o.policy && policy_kind(o.policy) == MEDUSA_POLICY_DIRECTION) mode = 3;
    const float* policy = (mode == 3 && !host_driven) ? policy_device_params(o.policy) : nullptr;
    MDS_HIP(hipMemcpyAsync(g->q_aux.p, srcs, (size_t)k * 4, hipMemcpyHostToDevice, s));
    g->trace.clear();

    build_pull_layout(g);
    const PullLayout& PL = *g->pull;
    auto expand = [&]() {
        hipLaunchKernelGGL(k_mbfs_expand, dim3(grid), dim3(256), 0, s, (const MbfsCtl*)ctl,
                           (const uint32_t*)g->q_small.p, (const uint32_t*)g->q_large.p, g->row_ptr.p, g->col.p,
                           g->cptr.p, g->csrc.p, (const unsigned long long*)frontier,
                           (const unsigned long long*)visited, next, (const PrRun*)PL.d_runs.p,
                           (const uint32_t*)PL.wg_run.p, PL.total_wgs, (const uint32_t*)PL.perm.p,
                           (const uint32_t*)PL.ell.p, C);
        MDS_LAUNCHED();
    };
    auto update = [&](const MbfsCtl* c, int32_t lv) {
        hipLaunchKernelGGL(k_mbfs_update, dim3(grid_for(n, kMbfsTile, 4096)), dim3(256), 0, s, c, lv, n, large,
                           g->row_ptr.p, visited, frontier, next, level, g->q_small.p, g->q_large.p, C);
        MDS_LAUNCHED();
    };

    // capture (once per graph, launch shape, policy buffer and slot count) a batch of supersteps
    auto captured = [&](GraphCache& gc, uint32_t slots) -> hipGraphExec_t {
        const uintptr_t key[4] = {(uintptr_t)masks.p, (uintptr_t)policy, ((uintptr_t)grid << 32) | large,
                                  ((uintptr_t)lvbuf.p ^ (uintptr_t)n) ^ ((uintptr_t)slots << 48)};
        if (gc.exec && std::equal(key, key + 4, gc.key)) return gc.exec;
        gc.reset();
        MDS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (uint32_t b = 0; b < slots; ++b) {
            expand();
            update(ctl, 0);
            hipLaunchKernelGGL(k_mbfs_decide, dim3(1), dim3(256), 0, s, ctl, C, policy, grid, 0u);
        }
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(s, &graph);
        if (e == hipSuccess) e = hipGraphInstantiate(&gc.exec, graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        MDS_HIP(e);
        std::copy(key, key + 4, gc.key);
        return gc.exec;
    };
    const uint32_t head = head_slots(g->mbfs_last_steps, kMbfsBatch);
    hipGraphExec_t head_exec = nullptr;
    if (!host_driven) head_exec = captured(head == kMbfsBatch ? g->mbfs_graph : g->mbfs_head, head);

    MDS_HIP(hipEventRecord(g->ev_begin, s));
    hipLaunchKernelGGL(k_mbfs_fill, dim3(grid_for(std::max<uint64_t>(n, (level_bytes + 3) / 4), 256, 8192)),
                       dim3(256), 0, s, n, (uint64_t)((level_bytes + 3) / 4), lvbuf.p, visited, frontier, next);
    MDS_LAUNCHED();
    hipLaunchKernelGGL(k_mbfs_zero, dim3(1), dim3(1), 0, s, C);
    MDS_LAUNCHED();
    hipLaunchKernelGGL(k_mbfs_seed, dim3(1), dim3(1), 0, s, (const uint32_t*)g->q_aux.p, k, next);
    MDS_LAUNCHED();
    update(nullptr, 0);
    hipLaunchKernelGGL(k_mbfs_init_ctl, dim3(1), dim3(1), 0, s, ctl, n, (unsigned long long)m, k, mode);
    MDS_LAUNCHED();

    MbfsCtl h{};
    if (!host_driven) {
        hipLaunchKernelGGL(k_mbfs_decide, dim3(1), dim3(256), 0, s, ctl, C, policy, grid, 1u);
        MDS_LAUNCHED();
        for (uint32_t batch = 0;; ++batch) {
            MDS_HIP(hipGraphLaunch(batch == 0 ? head_exec : captured(g->mbfs_graph, kMbfsBatch), s));
            MDS_HIP(hipMemcpyAsync(g->h_ctr_raw.p, ctl, sizeof(MbfsCtl), hipMemcpyDeviceToHost, s));
            MDS_HIP(hipStreamSynchronize(s));
            std::memcpy(&h, g->h_ctr_raw.p, sizeof(MbfsCtl));
            if (h.done) break;
            MDS_CHECK(batch < (n + kMbfsBatch) / kMbfsBatch + 2, MEDUSA_ERR_INTERNAL,
                      "mbfs: superstep loop did not end");
        }
        g->mbfs_last_steps = h.supersteps;
    } else {
        Sched sched(o, Sched::kMbfs);
        const bool timing = o.timing != 0;
        if (timing) ensure_events(g, 2);
        MbfsCounters c = read_counters<MbfsCounters>(g);
        Frontier f{};
        f.n = n;
        f.m = m;
        h.n_u = (uint64_t)n * k - c.new_bits;
        h.m_u = (uint64_t)m * k - c.reached_edges;
        h.reached = c.new_bits;
        h.reached_edges = c.reached_edges;
        h.saturated = c.saturated;
        bool pull = false;
        for (uint32_t t = 0; c.n_active; ++t) {
            f.step = t;
            f.n_f = c.n_active;
            f.m_f = c.active_deg;
            f.n_u = h.n_u;
            f.m_u = h.m_u;
            if (o.direction == MEDUSA_DIR_PUSH) pull = false;
            else if (o.direction == MEDUSA_DIR_PULL) pull = true;
            else pull = sched.bfs_pull(f, pull);
            hipLaunchKernelGGL(k_mbfs_set_ctl, dim3(1), dim3(1), 0, s, ctl, (int32_t)t, pull ? 1u : 0u, c.n_small,
                               c.n_large, mbfs_hub_wgs(grid, c.n_small, c.n_large, c.deg_large, c.active_deg));
            MDS_LAUNCHED();
            hipLaunchKernelGGL(k_mbfs_zero, dim3(1), dim3(1), 0, s, C);
            MDS_LAUNCHED();
            if (timing) MDS_HIP(hipEventRecord(g->events[0], s));
            expand();
            if (timing) MDS_HIP(hipEventRecord(g->events[1], s));
            update(nullptr, (int32_t)(t + 1));
            const MbfsCounters prev = c;
            c = read_counters<MbfsCounters>(g);
            double kms = 0;
            if (timing) {
                float ms = 0;
                MDS_HIP(hipEventElapsedTime(&ms, g->events[0], g->events[1]));
                kms = ms;
                st.kernel_ms += ms;
                ++st.kernel_launches;
            }
            const uint64_t kb = pull ? 4ull * (n + 1) + 8ull * n + 12ull * c.probes
                                     : 20ull * prev.n_active + 20ull * prev.active_deg;
            h.kernel_bytes += kb;
            h.alg_bytes += kb + 24ull * n + c.new_bits;
            if (pull) ++h.pull_steps; else ++h.push_steps;
            sched.record(g, f, pull ? 1.0 : 0.0, kms);
            ++h.supersteps;
            h.saturated |= c.saturated;
            h.reached += c.new_bits;
            h.reached_edges += c.reached_edges;
            h.n_u -= c.new_bits;
            h.m_u -= c.reached_edges;
        }
    }
    MDS_HIP(hipEventRecord(g->ev_end, s));
    MDS_HIP(hipStreamSynchronize(s));
    float ms = 0;
    MDS_HIP(hipEventElapsedTime(&ms, g->ev_begin, g->ev_end));
    st.device_ms = ms;
    st.supersteps = h.supersteps;
    st.push_steps = h.push_steps;
    st.pull_steps = h.pull_steps;
    st.kernel_bytes = h.kernel_bytes;
    st.alg_bytes = h.alg_bytes + level_bytes + 24ull * n;
    st.edges_traversed = h.reached_edges;
    st.reached = (uint32_t)std::min<uint64_t>(h.reached, 0xFFFFFFFFull);
    st.reserved[0] = h.saturated;
    copy_out(g, out, level, level_bytes);
}

}  // namespace medusa

using namespace medusa;

extern "C" int medusa_mbfs(medusa_graph* g, const uint32_t* sources, uint32_t k, uint8_t* levels,
                           const medusa_opts* opts, medusa_stats* stats) {
    return guard([&] {
        const auto lock = enter(g);
        MDS_CHECK(levels && sources, MEDUSA_ERR_ARG, "mbfs: null argument");
        MDS_CHECK(k >= 1 && k <= 64, MEDUSA_ERR_ARG, "mbfs: k must be in [1, 64]");
        for (uint32_t s = 0; s < k; ++s)
            MDS_CHECK(sources[s] < g->n, MEDUSA_ERR_ARG, "mbfs: source %u out of range", sources[s]);
        medusa_opts o{};
        if (opts) o = *opts;
        MDS_CHECK(o.direction >= 0 && o.direction <= 2, MEDUSA_ERR_ARG, "mbfs: bad direction %d", o.direction);
        MDS_CHECK(o.grid == 0 || (o.grid >= 2 && o.grid <= 65535), MEDUSA_ERR_ARG, "mbfs: grid must be 0 or 2..65535");
        medusa_stats st{};
        mbfs_run(g, sources, k, levels, o, st);
        if (stats) *stats = st;
    });
}
