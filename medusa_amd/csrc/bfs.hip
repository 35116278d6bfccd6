// BFS superstep engine.
//
// Reference: every superstep FunctorHolder<EDGE>::Run(UpdateLevel) sweeps all |E|
// edges (MultipleGPU/MultiAPIHoster.h:657-678, MedusaRT/APIKernelWrappers.h:34-44):
// if level[src]==step and level[dst]==-1 then level[dst]=step+1 and vote to
// continue (Algorithm/UserOperators.h:17-36, MedusaRT/SystemLibGPU.h:19-22); the
// host loop resets and reads the vote each step (Algorithm/main.cu:283-302,
// MultipleGPU/MultiUtilities.h:48-69).  Levels are int32, -1 = unreached
// (Algorithm/Message.h:9-10).
//
// Any number of sources may be given (medusa_bfs_multi): all start at level 0, which on a
// packed batch of disjoint graphs with one source per graph gives every graph's own BFS
// levels in one run.  Here a superstep only touches the frontier:
//   push:  frontier queue -> wave-cooperative expansion (lanes share the out-edges
//          of 64 frontier vertices through an in-wave prefix scan), hubs above
//          `large` out-degree get split over whole workgroups;
//   pull:  every unvisited vertex scans its in-edges for a parent in the
//          frontier bitmap and stops at the first hit;
//   direction: Beamer's rule with thresholds measured for this GPU (Sched), or the
//          learned policy near the rule's switching point;
//   compaction: one pass builds the next frontier bitmap (ballot per 64 vertices)
//          and the small/large queues (ballot + in-workgroup prefix scan, one
//          atomic per workgroup per queue) and counts the frontier's out-edges.
// The claim `if (level[v] < 0) level[v] = step+1` is a plain store: every
// writer stores the same value, so the race is benign (as in the reference).
//
// Two drivers run the same kernels:
//   device-driven (default): a one-workgroup kernel after each compaction decides the
//     next superstep (direction by the heuristic or the learned policy, queue sizes, the
//     termination vote) into a control block in HBM; batches of [expand, compact, decide]
//     slots are captured once as hipGraphs and replayed until the control block says
//     done, so the host waits once per batch, not per superstep (the first batch has as
//     many slots as the previous run on the handle had supersteps, then kBfsBatch);
//   host-driven (timing or trace requested): the host reads the counters after every
//     superstep, times the expansion with hipEvents and logs the trace record.

#include <algorithm>

#include "engine.h"
#include "sched.h"

namespace medusa {

constexpr uint32_t kLargeDefault = 2048;
constexpr uint32_t kLargeSplit = 16;  // workgroups per hub in the large-vertex push
constexpr uint32_t kBfsBatch = 8;     // supersteps per captured graph
constexpr uint32_t kExpandGrid = 2048;

// Control block of the device-driven loop.
struct BfsCtl {
    int32_t step;       // superstep being run (= level of the frontier)
    uint32_t pull;      // direction of this superstep
    uint32_t done;      // the vote: no vertex joined the frontier
    uint32_t n_small, n_large;
    uint32_t supersteps, push_steps, pull_steps;
    uint32_t mode;      // 0 heuristic, 1 push, 2 pull, 3 learned policy
    uint32_t hub_wgs;   // push: workgroups of the expansion given to the hub queue
    uint32_t pad[2];
    unsigned long long n, m, n_f, m_f, n_u, m_u, reached, reached_edges, kernel_bytes, alg_bytes;
};

// Level reset (with `single` at level 0 when it is < n) and zeroed frontier counters.
__global__ void k_bfs_reset(uint32_t n, int32_t* __restrict__ level, uint32_t single,
                            FrontierCounters* __restrict__ ctr) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        level[v] = v == single ? 0 : -1;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctr->n_small = 0;
        ctr->n_large = 0;
        ctr->n_new = 0;
        ctr->deg_sum = 0;
        ctr->probes = 0;
        ctr->deg_large = 0;
    }
}

__global__ void k_bfs_mark(const uint32_t* __restrict__ src, uint32_t count, int32_t* __restrict__ level) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x)
        level[src[i]] = 0;
}

__global__ void k_zero_counters(FrontierCounters* c) {
    c->n_small = 0;
    c->n_large = 0;
    c->n_new = 0;
    c->deg_sum = 0;
    c->probes = 0;
    c->deg_large = 0;
}

// Workgroups of a push expansion given to the hub queue: in proportion to its share of
// the frontier's out-edges, at least one for each non-empty queue.
__host__ __device__ inline uint32_t bfs_hub_wgs(uint32_t grid, uint32_t n_small, uint32_t n_large,
                                                unsigned long long deg_large, unsigned long long deg_sum) {
    if (!n_large) return 0;
    if (!n_small) return grid;
    unsigned long long h = deg_sum ? (unsigned long long)grid * deg_large / deg_sum : 1ull;
    if (h < 1) h = 1;
    if (h > grid - 1) h = grid - 1;
    return (uint32_t)h;
}

// Host-driven superstep: the host's decision written into the control block.
__global__ void k_bfs_set_ctl(BfsCtl* ctl, int32_t step, uint32_t pull, uint32_t n_small, uint32_t n_large,
                              uint32_t hub_wgs) {
    ctl->step = step;
    ctl->pull = pull;
    ctl->done = 0;
    ctl->n_small = n_small;
    ctl->n_large = n_large;
    ctl->hub_wgs = hub_wgs;
}

// ---- superstep bodies (shared by both drivers) -----------------------------------

// Wave-cooperative push over a queue of vertices with 1 <= outdeg <= large.  Of the
// `nwaves` waves taking part (`wave0` is this wave's index among them), wave w holds queue
// entries w, w + nwaves, w + 2*nwaves, ... one per lane, so a short queue of high-degree
// vertices spreads over many waves instead of sharing one.
__device__ __forceinline__ void push_small(const uint32_t* __restrict__ q, uint32_t qn,
                                           const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col,
                                           int32_t* __restrict__ level, int32_t next, uint32_t wave0, uint32_t nwaves,
                                           uint32_t (*s_beg)[kWave], uint32_t (*s_off)[kWave]) {
    const uint32_t lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    for (uint64_t r = 0; wave0 + r * kWave * nwaves < qn; ++r) {  // uniform across the wave
        const uint64_t i = wave0 + (r * kWave + lane) * nwaves;
        uint32_t beg = 0, deg = 0;
        if (i < qn) {
            const uint32_t u = q[i];
            beg = row_ptr[u];
            deg = row_ptr[u + 1] - beg;
        }
        const uint32_t incl = wave_incl_scan(deg);
        const uint32_t total = __shfl(incl, kWave - 1, kWave);
        s_beg[wv][lane] = beg;
        s_off[wv][lane] = incl - deg;
        __builtin_amdgcn_wave_barrier();
        for (uint32_t e = lane; e < total; e += kWave) {
            // owner = last lane whose exclusive offset is <= e (it has deg > 0)
            uint32_t lo = 0;
#pragma unroll
            for (uint32_t step = kWave / 2; step; step >>= 1)
                if (s_off[wv][lo + step] <= e) lo += step;
            const uint32_t v = col[s_beg[wv][lo] + (e - s_off[wv][lo])];
            if (level[v] < 0) level[v] = next;
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// Hubs: each vertex of the queue is split into `split` contiguous edge ranges, one per
// workgroup item; the `nwg` workgroups taking part stride over the items.
__device__ __forceinline__ void push_large(const uint32_t* __restrict__ q, uint32_t qn,
                                           const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ col,
                                           int32_t* __restrict__ level, int32_t next, uint32_t wg0, uint32_t nwg) {
    uint32_t split = nwg / qn;
    split = split < kLargeSplit ? kLargeSplit : (split > 4096u ? 4096u : split);
    for (uint64_t w = wg0; w < (uint64_t)qn * split; w += nwg) {
        const uint32_t u = q[w / split], part = (uint32_t)(w % split);
        const uint32_t b = row_ptr[u], deg = row_ptr[u + 1] - b;
        const uint32_t per = (deg + split - 1) / split;
        const uint32_t lo = b + min(deg, part * per), hi = b + min(deg, (part + 1) * per);
        for (uint32_t k = lo + threadIdx.x; k < hi; k += blockDim.x) {
            const uint32_t v = col[k];
            if (level[v] < 0) level[v] = next;
        }
    }
}

constexpr uint32_t kPullLane = 32;  // in-degree up to which one lane scans a vertex alone
constexpr uint32_t kPullGroup = 4;  // in-edges a lane probes per round

// Bottom-up: an unvisited vertex joins if any in-neighbour is in the frontier bitmap.  A
// lane scans a vertex of in-degree <= kPullLane alone; above that the whole wave scans it
// 64 in-edges at a time and stops at the first hit (ballot), so one high in-degree vertex
// does not hold its wave for hundreds of sequential probes.  `probes` counts in-edges
// examined up to and including the first hit in either case.
__device__ __forceinline__ void pull_all(uint32_t n, const uint32_t* __restrict__ cptr,
                                         const uint32_t* __restrict__ csrc,
                                         const unsigned long long* __restrict__ bitmap, int32_t* __restrict__ level,
                                         int32_t next, FrontierCounters* __restrict__ ctr) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    unsigned long long probes = 0;
    for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x - lane); base < n; base += gridDim.x * blockDim.x) {
        const uint32_t v = base + lane;  // uniform loop: every lane of the wave iterates together
        uint32_t b = 0, e = 0;
        bool todo = false;
        if (v < n && level[v] < 0) {
            b = cptr[v];
            e = cptr[v + 1];
            todo = b < e;
        }
        const bool big = todo && e - b > kPullLane;
        if (todo && !big) {
            // kPullGroup in-edges per round: their source and bitmap loads are issued together,
            // so a round costs two memory round trips instead of two per in-edge; probes still
            // count the in-edges up to and including the first hit
            for (uint32_t k0 = b; k0 < e; k0 += kPullGroup) {
                uint32_t u[kPullGroup];
                unsigned long long word[kPullGroup];
#pragma unroll
                for (uint32_t j = 0; j < kPullGroup; ++j) u[j] = k0 + j < e ? csrc[k0 + j] : 0u;
#pragma unroll
                for (uint32_t j = 0; j < kPullGroup; ++j) word[j] = k0 + j < e ? bitmap[u[j] >> 6] : 0ull;
                uint32_t hit = kPullGroup;
#pragma unroll
                for (uint32_t j = kPullGroup; j-- > 0;)
                    if ((word[j] >> (u[j] & 63)) & 1ull) hit = j;
                if (hit < kPullGroup) {
                    probes += hit + 1;
                    level[v] = next;
                    break;
                }
                probes += min(e - k0, kPullGroup);
            }
        }
        for (unsigned long long bm = __ballot(big); bm; bm &= bm - 1) {
            const uint32_t owner = __builtin_ctzll(bm);
            const uint32_t vb = __shfl(b, owner, kWave), ve = __shfl(e, owner, kWave);
            for (uint32_t k0 = vb; k0 < ve; k0 += kWave) {
                const uint32_t k = k0 + lane;
                bool hit = false;
                if (k < ve) {
                    const uint32_t u = csrc[k];
                    hit = (bitmap[u >> 6] >> (u & 63)) & 1ull;
                }
                const unsigned long long hm = __ballot(hit);
                if (hm) {
                    if (lane == 0) {
                        probes += __builtin_ctzll(hm) + 1;
                        level[base + owner] = next;
                    }
                    break;
                }
                if (lane == 0) probes += min(ve - k0, (uint32_t)kWave);
            }
        }
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

// A compaction tile is 4 waves x K chunks of 64 vertices; every tile adds to the frontier
// counters with one atomic each, and atomics on one address serialise, so large graphs use
// large tiles while small graphs use small ones to keep the GPU busy (see compact_k).
template <uint32_t K>
constexpr uint32_t compact_tile() { return 4 * K * kWave; }

// Next frontier = {v : level[v] == want}: bitmap word per 64 vertices, queues by
// out-degree class, and frontier counters.
template <uint32_t kCompactK>
__device__ __forceinline__ void compact(uint32_t n, const int32_t* __restrict__ level, int32_t want,
                                        const uint32_t* __restrict__ row_ptr, uint32_t large,
                                        uint32_t* __restrict__ q_small, uint32_t* __restrict__ q_large,
                                        unsigned long long* __restrict__ bitmap, uint32_t nwords,
                                        FrontierCounters* __restrict__ ctr) {
    __shared__ uint32_t s_cnt[4][3];
    __shared__ unsigned long long s_deg[4], s_dlg[4];
    __shared__ uint32_t s_base[2];
    const uint32_t lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    for (uint32_t tile = blockIdx.x; (uint64_t)tile * compact_tile<kCompactK>() < n; tile += gridDim.x) {
        unsigned long long msm[kCompactK], mlg[kCompactK];
        uint32_t csm = 0, clg = 0, cnew = 0;
        unsigned long long dsum = 0, dlg = 0;
        // all loads of the tile first (independent, so they overlap), then the ballots
        int32_t lv[kCompactK];
        uint32_t deg_of[kCompactK];
#pragma unroll
        for (uint32_t k = 0; k < kCompactK; ++k) {
            const uint32_t v = (tile * (4 * kCompactK) + k * 4 + wv) * kWave + lane;
            lv[k] = v < n ? level[v] : -1;
        }
#pragma unroll
        for (uint32_t k = 0; k < kCompactK; ++k) {
            const uint32_t v = (tile * (4 * kCompactK) + k * 4 + wv) * kWave + lane;
            deg_of[k] = lv[k] == want ? row_ptr[v + 1] - row_ptr[v] : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < kCompactK; ++k) {
            const uint32_t chunk = tile * (4 * kCompactK) + k * 4 + wv;
            const bool pred = lv[k] == want;
            const unsigned long long mask = __ballot(pred);
            if (lane == 0 && chunk < nwords) bitmap[chunk] = mask;
            const uint32_t deg = deg_of[k];
            msm[k] = __ballot(pred && deg > 0 && deg <= large);
            mlg[k] = __ballot(pred && deg > large);
            cnew += __popcll(mask);
            csm += __popcll(msm[k]);
            clg += __popcll(mlg[k]);
            dsum += deg;
            dlg += deg > large ? deg : 0u;
        }
        dsum = wave_sum_u64(dsum);
        dlg = wave_sum_u64(dlg);
        if (lane == 0) {
            s_cnt[wv][0] = csm;
            s_cnt[wv][1] = clg;
            s_cnt[wv][2] = cnew;
            s_deg[wv] = dsum;
            s_dlg[wv] = dlg;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t tsm = 0, tlg = 0, tnew = 0;
            unsigned long long tdeg = 0, tdlg = 0;
            for (int w = 0; w < 4; ++w) {
                tsm += s_cnt[w][0];
                tlg += s_cnt[w][1];
                tnew += s_cnt[w][2];
                tdeg += s_deg[w];
                tdlg += s_dlg[w];
            }
            s_base[0] = tsm ? atomicAdd(&ctr->n_small, tsm) : 0u;
            s_base[1] = tlg ? atomicAdd(&ctr->n_large, tlg) : 0u;
            if (tnew) atomicAdd(&ctr->n_new, tnew);
            if (tdeg) atomicAdd(&ctr->deg_sum, tdeg);
            if (tdlg) atomicAdd(&ctr->deg_large, tdlg);
        }
        __syncthreads();
        uint32_t osm = s_base[0], olg = s_base[1];
        for (uint32_t w = 0; w < wv; ++w) {
            osm += s_cnt[w][0];
            olg += s_cnt[w][1];
        }
#pragma unroll
        for (uint32_t k = 0; k < kCompactK; ++k) {
            const uint32_t v = (tile * (4 * kCompactK) + k * 4 + wv) * kWave + lane;
            if ((msm[k] >> lane) & 1ull) q_small[osm + lanes_below(msm[k])] = v;
            if ((mlg[k] >> lane) & 1ull) q_large[olg + lanes_below(mlg[k])] = v;
            osm += __popcll(msm[k]);
            olg += __popcll(mlg[k]);
        }
        __syncthreads();
    }
}

// ---- kernels (both drivers) ---------------------------------------------------------

template <uint32_t K>
__global__ __launch_bounds__(256) void k_bfs_compact(uint32_t n, const int32_t* __restrict__ level, int32_t want,
                                                     const uint32_t* __restrict__ row_ptr, uint32_t large,
                                                     uint32_t* __restrict__ q_small, uint32_t* __restrict__ q_large,
                                                     unsigned long long* __restrict__ bitmap, uint32_t nwords,
                                                     FrontierCounters* __restrict__ ctr) {
    compact<K>(n, level, want, row_ptr, large, q_small, q_large, bitmap, nwords, ctr);
}

// Expansion of the current superstep as the control block describes it: pull over every
// vertex, or push where the first hub_wgs workgroups take the hub queue and the rest the
// wave-cooperative queue.
__global__ __launch_bounds__(256) void k_bfs_expand_dev(const BfsCtl* __restrict__ ctl,
                                                        const uint32_t* __restrict__ q_small,
                                                        const uint32_t* __restrict__ q_large,
                                                        const uint32_t* __restrict__ row_ptr,
                                                        const uint32_t* __restrict__ col,
                                                        const uint32_t* __restrict__ cptr,
                                                        const uint32_t* __restrict__ csrc,
                                                        const unsigned long long* __restrict__ bitmap,
                                                        int32_t* __restrict__ level, uint32_t n,
                                                        FrontierCounters* __restrict__ ctr) {
    __shared__ uint32_t s_beg[4][kWave], s_off[4][kWave];
    if (ctl->done) return;
    const int32_t next = ctl->step + 1;
    if (ctl->pull) {
        pull_all(n, cptr, csrc, bitmap, level, next, ctr);
        return;
    }
    const uint32_t hub = ctl->hub_wgs;
    if (blockIdx.x < hub) {
        push_large(q_large, ctl->n_large, row_ptr, col, level, next, blockIdx.x, hub);
        return;
    }
    const uint32_t wg = blockIdx.x - hub, nwg = gridDim.x - hub;
    push_small(q_small, ctl->n_small, row_ptr, col, level, next, wg * 4 + threadIdx.x / kWave, nwg * 4, s_beg, s_off);
}

// Compaction of the current superstep.
template <uint32_t K>
__global__ __launch_bounds__(256) void k_bfs_compact_dev(const BfsCtl* __restrict__ ctl, uint32_t n,
                                                         const int32_t* __restrict__ level,
                                                         const uint32_t* __restrict__ row_ptr, uint32_t large,
                                                         uint32_t* __restrict__ q_small, uint32_t* __restrict__ q_large,
                                                         unsigned long long* __restrict__ bitmap, uint32_t nwords,
                                                         FrontierCounters* __restrict__ ctr) {
    if (ctl->done) return;
    compact<K>(n, level, ctl->step + 1, row_ptr, large, q_small, q_large, bitmap, nwords, ctr);
}

// Chunks per wave of a compaction tile for a graph of n vertices: small tiles below 2^18
// vertices (enough workgroups to fill the GPU), large ones above 1.5 M (fewer counter
// atomics).  Measured on Erdos-Renyi 65 K, RMAT 1 M and Chung-Lu 2.4 M vertex graphs.
static uint32_t compact_k(uint32_t n) { return n <= (1u << 18) ? 2u : (n <= 3u * (1u << 19) ? 8u : 16u); }

// Launches the compaction of level `want` (host-driven and seeding: ctl == nullptr) or of
// the control block's superstep (device-driven).
template <uint32_t K>
static void launch_compact_k(hipStream_t s, const BfsCtl* ctl, uint32_t n, int32_t want, const int32_t* level,
                             const uint32_t* row_ptr, uint32_t large, uint32_t* q_small, uint32_t* q_large,
                             unsigned long long* bitmap, uint32_t nwords, FrontierCounters* ctr) {
    const dim3 grid(grid_for(n, compact_tile<K>(), 4096));
    if (ctl)
        hipLaunchKernelGGL(k_bfs_compact_dev<K>, grid, dim3(256), 0, s, ctl, n, level, row_ptr, large, q_small,
                           q_large, bitmap, nwords, ctr);
    else
        hipLaunchKernelGGL(k_bfs_compact<K>, grid, dim3(256), 0, s, n, level, want, row_ptr, large, q_small, q_large,
                           bitmap, nwords, ctr);
}
static void launch_compact(hipStream_t s, const BfsCtl* ctl, uint32_t n, int32_t want, const int32_t* level,
                           const uint32_t* row_ptr, uint32_t large, uint32_t* q_small, uint32_t* q_large,
                           unsigned long long* bitmap, uint32_t nwords, FrontierCounters* ctr) {
    switch (compact_k(n)) {
        case 2: launch_compact_k<2>(s, ctl, n, want, level, row_ptr, large, q_small, q_large, bitmap, nwords, ctr); break;
        case 8: launch_compact_k<8>(s, ctl, n, want, level, row_ptr, large, q_small, q_large, bitmap, nwords, ctr); break;
        default:
            launch_compact_k<16>(s, ctl, n, want, level, row_ptr, large, q_small, q_large, bitmap, nwords, ctr);
    }
}

// One workgroup after the compaction: its counters are visible at the kernel boundary,
// so the decision needs no fence.  Decides the next superstep into the control block.
__global__ __launch_bounds__(256) void k_bfs_decide(BfsCtl* __restrict__ ctl, uint32_t n, uint32_t nwords,
                                                    FrontierCounters* __restrict__ ctr,
                                                    const float* __restrict__ policy, uint32_t grid) {
    __shared__ MlpShared s_mlp;
    const uint32_t n_new = load_now(&ctr->n_new);
    const uint32_t n_small = load_now(&ctr->n_small);
    const uint32_t n_large = load_now(&ctr->n_large);
    const unsigned long long deg_sum = load_now(&ctr->deg_sum);
    const unsigned long long probes = load_now(&ctr->probes);
    const unsigned long long deg_large = load_now(&ctr->deg_large);
    BfsCtl c = load_now_block(ctl);
    if (c.done) return;
    // byte model of the superstep just finished (same as the host-driven driver): push =
    // 12 per frontier vertex + 8 per out-edge + 8 per claim; pull = 2 bitmap passes (n/4) +
    // 8 per unvisited vertex + 8 per probed in-edge + 4 per claim
    const unsigned long long kb = c.pull ? n / 4 + 8ull * c.n_u + 8ull * probes + 4ull * n_new
                                         : 12ull * c.n_f + 8ull * c.m_f + 8ull * n_new;
    c.kernel_bytes += kb;
    c.alg_bytes += kb + 4ull * n + nwords * 8ull + 12ull * n_new;
    if (n_new == 0) {
        c.done = 1;
    } else {
        c.step += 1;
        c.n_f = n_new;
        c.m_f = deg_sum;
        c.n_u -= n_new;
        c.m_u -= deg_sum;
        c.reached += n_new;
        c.reached_edges += deg_sum;
        c.n_small = n_small;
        c.n_large = n_large;
        c.hub_wgs = bfs_hub_wgs(grid, n_small, n_large, deg_large, deg_sum);
    }
    bool pull = c.pull != 0;
    if (!c.done) {
        if (c.mode == 1) {
            pull = false;
        } else if (c.mode == 2) {
            pull = true;
        } else if (c.mode == 3 && policy &&
                   bfs_policy_consulted(pull, (double)c.n_f, (double)c.m_f, (double)c.m_u, (double)c.n,
                                        Sched::kAlphaBfs, Sched::kBeta)) {
            const double rec[MEDUSA_TRACE_FIELDS] = {(double)c.step, (double)c.n_f, (double)c.m_f, (double)c.n_u,
                                                     (double)c.m_u, (double)c.n, (double)c.m, 0.0, 0.0};
            policy_direction_costs(rec, policy, s_mlp);
            // the heuristic's choice stands unless the policy predicts a clear gain
            const bool plain = !pull ? (double)c.m_f > (double)c.m_u / Sched::kAlphaBfs
                                     : !((double)c.n_f < (double)c.n / Sched::kBeta);
            pull = policy_pull(s_mlp.out[0], s_mlp.out[1], plain);
        } else {
            // direction-optimising heuristic (Beamer, Asanovic, Patterson, SC'12)
            if (!pull)
                pull = (double)c.m_f > (double)c.m_u / Sched::kAlphaBfs;
            else
                pull = !((double)c.n_f < (double)c.n / Sched::kBeta);
        }
        c.pull = pull ? 1u : 0u;
        c.supersteps += 1;
        if (pull)
            c.pull_steps += 1;
        else
            c.push_steps += 1;
    }
    __syncthreads();  // every wave has read the control block and counters
    if (threadIdx.x == 0) {
        *ctl = c;
        ctr->n_small = 0;
        ctr->n_large = 0;
        ctr->n_new = 0;
        ctr->deg_sum = 0;
        ctr->probes = 0;
        ctr->deg_large = 0;
    }
}

// First decision (superstep 0) from the seeded counters; one workgroup.
__global__ __launch_bounds__(256) void k_bfs_first_decide(BfsCtl* __restrict__ ctl, FrontierCounters* __restrict__ ctr,
                                                          const float* __restrict__ policy, uint32_t n,
                                                          unsigned long long m, uint32_t mode, uint32_t grid) {
    __shared__ MlpShared s_mlp;
    BfsCtl c{};
    c.mode = mode;
    c.n = n;
    c.m = m;
    c.n_u = n;
    c.m_u = m;
    c.step = 0;
    c.n_f = ctr->n_new;
    c.m_f = ctr->deg_sum;
    c.n_u -= c.n_f;
    c.m_u -= c.m_f;
    c.reached = c.n_f;
    c.reached_edges = c.m_f;
    c.n_small = ctr->n_small;
    c.n_large = ctr->n_large;
    c.hub_wgs = bfs_hub_wgs(grid, c.n_small, c.n_large, ctr->deg_large, ctr->deg_sum);
    bool pull = false;
    if (c.mode == 2) {
        pull = true;
    } else if (c.mode == 3 && policy &&
               bfs_policy_consulted(false, (double)c.n_f, (double)c.m_f, (double)c.m_u, (double)c.n,
                                    Sched::kAlphaBfs, Sched::kBeta)) {
        const double rec[MEDUSA_TRACE_FIELDS] = {0.0, (double)c.n_f, (double)c.m_f, (double)c.n_u, (double)c.m_u,
                                                 (double)c.n, (double)c.m, 0.0, 0.0};
        policy_direction_costs(rec, policy, s_mlp);
        pull = policy_pull(s_mlp.out[0], s_mlp.out[1], (double)c.m_f > (double)c.m_u / Sched::kAlphaBfs);
    } else if (c.mode == 0 || c.mode == 3) {  // heuristic, or the policy far from its switching point
        pull = (double)c.m_f > (double)c.m_u / Sched::kAlphaBfs;
    }
    c.pull = pull ? 1u : 0u;
    c.supersteps = 1;
    c.push_steps = pull ? 0 : 1;
    c.pull_steps = pull ? 1 : 0;
    __syncthreads();  // every wave has read the control block and counters
    if (threadIdx.x == 0) {
        *ctl = c;
        ctr->n_small = 0;
        ctr->n_large = 0;
        ctr->n_new = 0;
        ctr->deg_sum = 0;
        ctr->probes = 0;
        ctr->deg_large = 0;
    }
}

// Degree threshold of the hub queue and workgroups of the expansion: the caller's value,
// else the handle's autotuned one, else the default.
static uint32_t bfs_large(const medusa_graph* g, const medusa_opts& o) {
    return o.large_degree ? o.large_degree : (g->tuned.bfs_large ? g->tuned.bfs_large : kLargeDefault);
}
static uint32_t bfs_grid(const medusa_graph* g, const medusa_opts& o) {
    return o.grid ? o.grid : (g->tuned.bfs_grid ? g->tuned.bfs_grid : kExpandGrid);
}

static uint32_t bfs_mode(const medusa_opts& o) {
    if (o.direction == MEDUSA_DIR_PUSH) return 1;
    if (o.direction == MEDUSA_DIR_PULL) return 2;
    if (o.policy && policy_kind(o.policy) == MEDUSA_POLICY_DIRECTION) return 3;
    return 0;
}

// Work buffers of one BFS run: the handle's own (allocated on its first BFS), or, for
// medusa_bfs_noalloc, the caller's level array and the library's static scratch.
struct BfsBufs {
    int32_t* level;
    uint32_t *q_small, *q_large, *q_aux;
    unsigned long long* bitmap;
    BfsCtl* ctl;
};
static BfsBufs handle_bufs(medusa_graph* g) {
    g->bfs_ctl.ensure(1);
    static_assert(sizeof(BfsCtl) <= sizeof(BfsCtlStorage), "control block storage");
    return {g->level.p, g->q_small.p, g->q_large.p, g->q_aux.p, g->bitmap.p, reinterpret_cast<BfsCtl*>(g->bfs_ctl.p)};
}

// Seeding: every source at level 0 (one source by value, several from B.q_aux), then the
// compaction of level 0 builds the first frontier's bitmap, queues and counters exactly
// as a superstep would.
static void bfs_seed(medusa_graph* g, const BfsBufs& B, uint32_t nsrc, uint32_t single, uint32_t large) {
    const uint32_t n = g->n, nwords = (n + 63) / 64;
    hipStream_t s = g->stream;
    FrontierCounters* ctr = dev_counters<FrontierCounters>(g);
    hipLaunchKernelGGL(k_bfs_reset, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, n, B.level,
                       nsrc == 1 ? single : 0xFFFFFFFFu, ctr);
    MDS_LAUNCHED();
    if (nsrc > 1) {
        hipLaunchKernelGGL(k_bfs_mark, dim3(grid_for(nsrc, 256, 1024)), dim3(256), 0, s, (const uint32_t*)B.q_aux,
                           nsrc, B.level);
        MDS_LAUNCHED();
    }
    launch_compact(s, nullptr, n, 0, (const int32_t*)B.level, g->row_ptr.p, large, B.q_small, B.q_large, B.bitmap,
                   nwords, ctr);
    MDS_LAUNCHED();
}

// One device-driven superstep slot: expansion, compaction, decision.
static void bfs_enqueue_slot(medusa_graph* g, const BfsBufs& B, uint32_t large, uint32_t grid, const float* policy) {
    const uint32_t n = g->n, nwords = (n + 63) / 64;
    hipStream_t s = g->stream;
    FrontierCounters* ctr = dev_counters<FrontierCounters>(g);
    hipLaunchKernelGGL(k_bfs_expand_dev, dim3(grid), dim3(256), 0, s, (const BfsCtl*)B.ctl, (const uint32_t*)B.q_small,
                       (const uint32_t*)B.q_large, g->row_ptr.p, g->col.p, g->cptr.p, g->csrc.p,
                       (const unsigned long long*)B.bitmap, B.level, n, ctr);
    launch_compact(s, B.ctl, n, 0, (const int32_t*)B.level, g->row_ptr.p, large, B.q_small, B.q_large, B.bitmap,
                   nwords, ctr);
    hipLaunchKernelGGL(k_bfs_decide, dim3(1), dim3(256), 0, s, B.ctl, n, nwords, ctr, policy, grid);
}

// Device-driven run: seed, first decision, then batches of superstep slots until the control
// block says done.  capture: the batches are captured once as hipGraphs and replayed; without
// it (medusa_bfs_noalloc, which may not allocate) the same slots are launched one by one.
static void bfs_run_device(medusa_graph* g, const BfsBufs& B, uint32_t nsrc, uint32_t single, int32_t* out,
                           const medusa_opts& o, medusa_stats& st, bool capture = true) {
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    const uint32_t nwords = (n + 63) / 64;
    const uint32_t large = bfs_large(g, o), grid = bfs_grid(g, o);
    hipStream_t s = g->stream;
    const uint32_t mode = bfs_mode(o);
    const float* policy = mode == 3 ? policy_device_params(o.policy) : nullptr;
    BfsCtl* ctl = B.ctl;
    FrontierCounters* ctr = dev_counters<FrontierCounters>(g);

    // capture (once per graph, threshold, policy buffer and slot count) a batch of supersteps
    auto captured = [&](GraphCache& gc, uint32_t slots) -> hipGraphExec_t {
        const uintptr_t key[4] = {(uintptr_t)B.level, (uintptr_t)policy, ((uintptr_t)grid << 32) | large,
                                  (uintptr_t)B.bitmap ^ ((uintptr_t)slots << 48)};
        if (gc.exec && std::equal(key, key + 4, gc.key)) return gc.exec;
        gc.reset();
        MDS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (uint32_t k = 0; k < slots; ++k) bfs_enqueue_slot(g, B, large, grid, policy);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(s, &graph);
        if (e == hipSuccess) e = hipGraphInstantiate(&gc.exec, graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        MDS_HIP(e);
        std::copy(key, key + 4, gc.key);
        return gc.exec;
    };
    const uint32_t head = head_slots(g->bfs_last_steps, kBfsBatch);
    const hipGraphExec_t head_exec = capture ? captured(head == kBfsBatch ? g->bfs_graph : g->bfs_head, head) : nullptr;

    MDS_HIP(hipEventRecord(g->ev_begin, s));
    bfs_seed(g, B, nsrc, single, large);
    hipLaunchKernelGGL(k_bfs_first_decide, dim3(1), dim3(256), 0, s, ctl, ctr, policy, n, (unsigned long long)m,
                       mode, grid);
    MDS_LAUNCHED();
    BfsCtl h{};
    for (uint32_t batch = 0;; ++batch) {
        // the tail batch is captured after the timed region has started only on a handle's
        // first runs; later runs find it cached
        if (capture) {
            MDS_HIP(hipGraphLaunch(batch == 0 ? head_exec : captured(g->bfs_graph, kBfsBatch), s));
        } else {
            for (uint32_t k = 0; k < kBfsBatch; ++k) bfs_enqueue_slot(g, B, large, grid, policy);
            MDS_LAUNCHED();
        }
        MDS_HIP(hipMemcpyAsync(g->h_ctr_raw.p, ctl, sizeof(BfsCtl), hipMemcpyDeviceToHost, s));
        MDS_HIP(hipStreamSynchronize(s));
        std::memcpy(&h, g->h_ctr_raw.p, sizeof(BfsCtl));
        if (h.done) break;
        MDS_CHECK(batch < (n + kBfsBatch) / kBfsBatch + 2, MEDUSA_ERR_INTERNAL, "bfs: superstep loop did not end");
    }
    if (capture) g->bfs_last_steps = h.supersteps;
    MDS_HIP(hipEventRecord(g->ev_end, s));
    MDS_HIP(hipStreamSynchronize(s));
    float ms = 0;
    MDS_HIP(hipEventElapsedTime(&ms, g->ev_begin, g->ev_end));
    st.device_ms = ms;
    st.supersteps = h.supersteps;
    st.push_steps = h.push_steps;
    st.pull_steps = h.pull_steps;
    st.edges_traversed = h.reached_edges;
    st.reached = (uint32_t)h.reached;
    st.kernel_bytes = h.kernel_bytes;
    st.alg_bytes = h.alg_bytes + 8ull * n + nwords * 8ull + 8ull * nsrc;
    if (out != B.level) copy_out(g, out, B.level, (size_t)n * 4);
}

// Host-driven run: the same expansion and compaction kernels, with the host deciding each
// superstep from counters read back after it (timing and tracing).
static void bfs_run_host(medusa_graph* g, uint32_t nsrc, uint32_t single, int32_t* out, const medusa_opts& o,
                         medusa_stats& st) {
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    const uint32_t nwords = (n + 63) / 64;
    const uint32_t large = bfs_large(g, o), grid = bfs_grid(g, o);
    hipStream_t s = g->stream;
    const bool timing = o.timing != 0;
    FrontierCounters* ctr = dev_counters<FrontierCounters>(g);
    const BfsBufs B = handle_bufs(g);
    BfsCtl* ctl = B.ctl;
    g->trace.clear();

    MDS_HIP(hipEventRecord(g->ev_begin, s));
    bfs_seed(g, B, nsrc, single, large);
    FrontierCounters c = read_counters<FrontierCounters>(g);

    Frontier f{};
    f.n = n;
    f.m = m;
    f.n_f = c.n_new;
    f.m_f = c.deg_sum;
    f.n_u = n - c.n_new;
    f.m_u = m - c.deg_sum;
    uint64_t reached_edges = c.deg_sum, reached = c.n_new, bytes = 8ull * n + nwords * 8ull + 8ull * nsrc;
    bool pull = false;
    Sched sched(o, Sched::kBfs);
    if (timing) ensure_events(g, 2);
    for (uint32_t t = 0;; ++t) {
        f.step = t;
        if (o.direction == MEDUSA_DIR_PUSH) pull = false;
        else if (o.direction == MEDUSA_DIR_PULL) pull = true;
        else pull = sched.bfs_pull(f, pull);
        const uint32_t hub = bfs_hub_wgs(grid, c.n_small, c.n_large, c.deg_large, c.deg_sum);
        hipLaunchKernelGGL(k_bfs_set_ctl, dim3(1), dim3(1), 0, s, ctl, (int32_t)t, pull ? 1u : 0u, c.n_small,
                           c.n_large, hub);
        MDS_LAUNCHED();
        hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(1), 0, s, ctr);
        MDS_LAUNCHED();
        if (timing) MDS_HIP(hipEventRecord(g->events[0], s));
        hipLaunchKernelGGL(k_bfs_expand_dev, dim3(grid), dim3(256), 0, s, (const BfsCtl*)ctl,
                           (const uint32_t*)g->q_small.p, (const uint32_t*)g->q_large.p, g->row_ptr.p, g->col.p,
                           g->cptr.p, g->csrc.p, (const unsigned long long*)g->bitmap.p, g->level.p, n, ctr);
        MDS_LAUNCHED();
        if (timing) MDS_HIP(hipEventRecord(g->events[1], s));
        launch_compact(s, nullptr, n, (int32_t)(t + 1), (const int32_t*)g->level.p, g->row_ptr.p, large,
                       g->q_small.p, g->q_large.p, g->bitmap.p, nwords, ctr);
        MDS_LAUNCHED();
        c = read_counters<FrontierCounters>(g);
        double kms = 0;
        if (timing) {
            float ms = 0;
            MDS_HIP(hipEventElapsedTime(&ms, g->events[0], g->events[1]));
            kms = ms;
            st.kernel_ms += ms;
            ++st.kernel_launches;
        }
        // byte model: push = queue + offsets + col + level probe + claims;
        // pull = 2 bitmap passes + level and offsets of unvisited vertices + probed in-edges
        // (source + bitmap word) + claims
        uint64_t kb;
        if (!pull) {
            kb = 12ull * f.n_f + 8ull * f.m_f + 8ull * c.n_new;
            ++st.push_steps;
        } else {
            kb = n / 4 + 8ull * f.n_u + 8ull * c.probes + 4ull * c.n_new;
            ++st.pull_steps;
        }
        st.kernel_bytes += kb;
        bytes += kb + 4ull * n + nwords * 8ull + 12ull * c.n_new;
        sched.record(g, f, pull ? 1.0 : 0.0, kms);
        ++st.supersteps;
        if (c.n_new == 0) break;
        f.n_f = c.n_new;
        f.m_f = c.deg_sum;
        f.n_u -= c.n_new;
        f.m_u -= c.deg_sum;
        reached += c.n_new;
        reached_edges += c.deg_sum;
    }
    MDS_HIP(hipEventRecord(g->ev_end, s));
    MDS_HIP(hipStreamSynchronize(s));
    float ms = 0;
    MDS_HIP(hipEventElapsedTime(&ms, g->ev_begin, g->ev_end));
    st.device_ms = ms;
    st.edges_traversed = reached_edges;
    st.reached = (uint32_t)reached;
    st.alg_bytes = bytes;
    copy_out(g, out, g->level.p, (size_t)n * 4);
}

static void bfs_entry(medusa_graph* g, const uint32_t* sources, uint32_t count, int32_t* levels,
                      const medusa_opts* opts, medusa_stats* stats) {
    const auto lock = enter(g);
    MDS_CHECK(levels, MEDUSA_ERR_ARG, "bfs: null output");
    MDS_CHECK(sources && count >= 1, MEDUSA_ERR_ARG, "bfs: need at least one source");
    const uint32_t n = g->n;
    for (uint32_t i = 0; i < count; ++i)
        MDS_CHECK(sources[i] < n, MEDUSA_ERR_ARG, "bfs: source %u out of range (n=%u)", sources[i], n);
    medusa_opts o{};
    if (opts) o = *opts;
    MDS_CHECK(o.direction >= 0 && o.direction <= 2, MEDUSA_ERR_ARG, "bfs: bad direction %d", o.direction);
    MDS_CHECK(o.grid == 0 || (o.grid >= 2 && o.grid <= 65535), MEDUSA_ERR_ARG, "bfs: grid must be 0 or 2..65535");
    g->level.ensure(n);
    g->q_small.ensure(n);
    g->q_large.ensure(n);
    g->q_aux.ensure(std::max(n, count));
    g->bitmap.ensure((n + 63) / 64);
    if (count > 1)
        MDS_HIP(hipMemcpyAsync(g->q_aux.p, sources, (size_t)count * 4, hipMemcpyHostToDevice, g->stream));
    medusa_stats st{};
    if (o.timing || o.trace) bfs_run_host(g, count, sources[0], levels, o, st);
    else bfs_run_device(g, handle_bufs(g), count, sources[0], levels, o, st);
    if (stats) *stats = st;
}

// ---- BFS with no device allocation ------------------------------------------------
// Work buffers of medusa_bfs_noalloc: two frontier queues of kNoAllocMaxVertices entries,
// the frontier bitmap and a control block, in one array per GPU that lives as long as the
// process.  Host code must not pass `scratch` itself to a kernel (in host code the name is
// a stand-in, not a device address): the device address is looked up once per device with
// hipGetSymbolAddress and kept in s_scratch.
constexpr uint32_t kNoAllocMaxVertices = MEDUSA_NOALLOC_MAX_VERTICES;
constexpr size_t kNoAllocQueueBytes = (size_t)kNoAllocMaxVertices * 4;
constexpr size_t kNoAllocBitmapBytes = (size_t)kNoAllocMaxVertices / 8;
constexpr size_t kNoAllocBytes = 2 * kNoAllocQueueBytes + kNoAllocBitmapBytes + sizeof(CtlStorage);
static_assert(kNoAllocMaxVertices % 2048 == 0, "queue and bitmap regions stay 256-byte aligned");
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
static std::mutex s_scratch_mu;
static unsigned char* s_scratch[kMaxDevices] = {};

static unsigned char* scratch_on(int device) {
    MDS_CHECK(device >= 0 && device < kMaxDevices, MEDUSA_ERR_ARG, "bfs_noalloc: device %d out of range", device);
    const std::lock_guard<std::mutex> lock(s_scratch_mu);
    if (!s_scratch[device]) {
        void* p = nullptr;
        MDS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));  // the calling thread has `device` current
        MDS_CHECK(p != nullptr, MEDUSA_ERR_HIP, "bfs_noalloc: no address for the scratch array");
        s_scratch[device] = static_cast<unsigned char*>(p);
    }
    return s_scratch[device];
}

static void bfs_noalloc(medusa_graph* g, uint32_t source, int32_t* levels, medusa_stats* stats) {
    const auto lock = enter(g);  // makes the graph's device current
    const uint32_t n = g->n;
    MDS_CHECK(levels, MEDUSA_ERR_ARG, "bfs_noalloc: null output");
    MDS_CHECK(source < n, MEDUSA_ERR_ARG, "bfs_noalloc: source %u out of range (n=%u)", source, n);
    MDS_CHECK(n <= kNoAllocMaxVertices, MEDUSA_ERR_ARG,
              "bfs_noalloc: %u vertices exceed the scratch limit of %u (MEDUSA_NOALLOC_MAX_VERTICES)", n,
              kNoAllocMaxVertices);
    hipPointerAttribute_t attr{};
    const hipError_t e = hipPointerGetAttributes(&attr, levels);
    (void)hipGetLastError();
    MDS_CHECK(e == hipSuccess && attr.type == hipMemoryTypeDevice && attr.device == g->device, MEDUSA_ERR_ARG,
              "bfs_noalloc: levels must be device memory of the graph's GPU");
    unsigned char* base = scratch_on(g->device);
    BfsBufs B{};
    B.level = levels;
    B.q_small = reinterpret_cast<uint32_t*>(base);
    B.q_large = reinterpret_cast<uint32_t*>(base + kNoAllocQueueBytes);
    B.bitmap = reinterpret_cast<unsigned long long*>(base + 2 * kNoAllocQueueBytes);
    B.ctl = reinterpret_cast<BfsCtl*>(base + 2 * kNoAllocQueueBytes + kNoAllocBitmapBytes);
    B.q_aux = nullptr;  // one source: not used
    static_assert(sizeof(BfsCtl) <= sizeof(CtlStorage), "control block storage");
    medusa_opts o{};  // heuristic direction, default or tuned launch shape
    medusa_stats st{};
    bfs_run_device(g, B, 1, source, levels, o, st, /*capture=*/false);
    if (stats) *stats = st;
}

}  // namespace medusa

using namespace medusa;

extern "C" int medusa_bfs(medusa_graph* g, uint32_t source, int32_t* levels, const medusa_opts* opts,
                          medusa_stats* stats) {
    return guard([&] { bfs_entry(g, &source, 1, levels, opts, stats); });
}

extern "C" int medusa_bfs_multi(medusa_graph* g, const uint32_t* sources, uint32_t count, int32_t* levels,
                                const medusa_opts* opts, medusa_stats* stats) {
    return guard([&] { bfs_entry(g, sources, count, levels, opts, stats); });
}

extern "C" int medusa_bfs_noalloc(medusa_graph* g, uint32_t source, int32_t* levels, medusa_stats* stats) {
    return guard([&] { bfs_noalloc(g, source, levels, stats); });
}
