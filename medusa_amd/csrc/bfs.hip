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
// levels in one run.  Here a superstep is ONE kernel that only touches the frontier and
// builds the next one as it goes:
//   frontier: a bitmap (one bit per vertex, three buffers in rotation) plus a queue of its
//          hubs (out-degree above `large`);
//   push:  waves take bitmap words and share the out-edges of a word's frontier vertices
//          through an in-wave prefix scan; hubs are split over whole workgroups; a target is
//          claimed by the lane whose 64-bit atomic OR first sets its bit in the next bitmap,
//          so each joining vertex is counted exactly once;
//   pull:  every unvisited vertex scans its in-edges for a parent in the frontier bitmap and
//          stops at the first hit; each wave owns 64 consecutive vertices and writes their
//          word of the next bitmap whole (ballot), no atomics;
//   counters: joined vertices, their out-degrees, hub degrees and probed in-edges, summed
//          per workgroup into two packed 64-bit atomics; joined hubs are appended to the next
//          hub queue (one atomic per wave that has any);
//   direction: Beamer's rule with thresholds measured for this GPU (Sched), or the
//          learned policy near the rule's switching point.
// In the reference every writer stores the same level, so its claim is a plain store; here the
// first setter of the bit also writes the level, which gives the same levels.
//
// Two drivers run the same kernel:
//   device-driven (default): the last workgroup of each superstep (by a ticket taken after
//     its counters) decides the next one (direction, hub queue split, the termination vote)
//     into a control block in HBM; batches of superstep slots are captured once as hipGraphs
//     and replayed until the control block says done, so the host waits once per batch, not
//     per superstep (the first batch has as many slots as the previous run on the handle had
//     supersteps, then kBfsBatch);
//   host-driven (timing or trace requested): the host reads the counters after every
//     superstep, times it with hipEvents and logs the trace record.

#include <algorithm>

#include "engine.h"
#include "sched.h"

namespace medusa {

constexpr uint32_t kLargeDefault = 2048;
constexpr uint32_t kLargeSplit = 16;  // workgroups per hub in the large-vertex push
constexpr uint32_t kBfsBatch = 8;     // supersteps per captured graph
constexpr uint32_t kExpandGrid = 1024;

// Control block of the device-driven loop.
struct BfsCtl {
    int32_t step;       // superstep being run (= level of the frontier)
    uint32_t pull;      // direction of this superstep
    uint32_t done;      // the vote: no vertex joined the frontier
    uint32_t n_small, n_large;
    uint32_t supersteps, push_steps, pull_steps;
    uint32_t mode;      // 0 heuristic, 1 push, 2 pull, 3 learned policy
    uint32_t hub_wgs;   // push: workgroups of the expansion given to the hub queue
    uint32_t pending;   // a push with plain claims has run and its compaction has not
    uint32_t small;     // push with exact (atomic) claims: the frontier's out-edges fit one queue segment
    uint32_t slots;     // kernel launches the loop has used (a push with plain claims takes two)
    uint32_t pad;
    unsigned long long n, m, n_f, m_f, n_u, m_u, reached, reached_edges, kernel_bytes, alg_bytes;
};

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
__global__ void k_bfs_set_ctl(BfsCtl* ctl, int32_t step, uint32_t pull, uint32_t small, uint32_t n_small,
                              uint32_t n_large, uint32_t hub_wgs) {
    ctl->step = step;
    ctl->pull = pull;
    ctl->small = small;
    ctl->done = 0;
    ctl->n_small = n_small;
    ctl->n_large = n_large;
    ctl->hub_wgs = hub_wgs;
    ctl->pending = 0;
}

// Counters of one superstep.  Atomics on one address serialise (about 10 ns each on this
// GPU), so they are spread over kBfsSlots addresses, each on its own 128-byte line:
//   seg[p][s]: the frontier of parity p (superstep t's frontier has parity t % 2) is a queue
//     in kBfsSlots segments by vertex range (segment s holds vertices [s * span, (s + 1) *
//     span)); its counter packs the vertices appended (low half, also the append offset) and
//     the sum of their out-degrees (high half);
//   work[s]: workgroup b adds the out-degree sum of the hubs it found (low half) and the
//     in-edges it probed (high half) into slot b % kBfsSlots;
//   sub / top: two levels of tickets that find the last workgroup of a superstep.
// Every half's total is bounded by n or by m (both below 2^32), so no carry crosses into a
// high half.
constexpr uint32_t kBfsSlots = 32;
constexpr uint32_t kBfsRound = 1024;  // vertices per workgroup round of the compaction; segment spans are multiples
struct BfsLine {
    alignas(128) unsigned long long v;
};
struct BfsTicket {
    alignas(128) uint32_t v;
};
struct BfsCounters {
    BfsLine seg[2][kBfsSlots];
    BfsLine work[kBfsSlots];
    BfsTicket sub[kBfsSlots];      // workgroups of slot s that have finished the superstep
    MDS_COUNTER(uint32_t, top);    // slots whose workgroups have all finished
    MDS_COUNTER(uint32_t, n_hub);  // joined hubs, appended to the next hub queue
};

// Buffers and shape of a BFS run, passed by value to every superstep kernel.  Superstep t
// reads the frontier as bitmap bm[t % 3] (pull), queue q[t % 2] (push) and hub queue
// hub[t % 2], builds the next frontier's in bm[(t + 1) % 3], q[(t + 1) % 2] and
// hub[(t + 1) % 2], and clears bm[(t + 2) % 3] for superstep t + 1 (an exact-claim push sets
// bits in its next bitmap).
struct BfsArgs {
    const uint32_t *row_ptr, *col, *cptr, *csrc;
    int32_t* level;
    uint32_t* q[2];    // segmented queues, kBfsSlots * span entries each
    uint32_t* hub[2];  // hub queues, n entries each
    unsigned long long* bm[3];
    unsigned long long* visited;  // one bit per vertex that has a level (n / 8 bytes, stays in L2)
    BfsCounters* ctr;
    BfsCtl* ctl;
    const float* policy;  // learned direction policy (device parameters) or null
    uint32_t n, nwords, large, grid, span;
    unsigned long long exact_edges;  // frontier out-edges up to which a push claims exactly
};

// Segment span for n vertices: kBfsSlots segments cover [0, n) in whole compaction rounds.
__host__ __device__ constexpr uint32_t bfs_span(uint32_t n) {
    const uint64_t per = ((uint64_t)n + kBfsSlots - 1) / kBfsSlots;
    const uint64_t span = (per + kBfsRound - 1) / kBfsRound * kBfsRound;
    return (uint32_t)(span ? span : kBfsRound);
}

// A push claims exactly (compare-and-swap, no compaction) when the frontier has at most
// `exact_edges` out-edges: 1, or 2 when it also has at most 64 vertices (every wave then reads
// all of them and takes an equal share of their out-edges).  Its winners all go to segment 0 of
// the next queue, which may then run on over the other segments' space (kBfsSlots * span >= n
// entries, and no other segment receives any).
__host__ __device__ inline uint32_t bfs_small(bool pull, unsigned long long n_f, unsigned long long m_f,
                                              unsigned long long exact_edges) {
    return pull || m_f > exact_edges ? 0u : (n_f <= (unsigned long long)kWave ? 2u : 1u);
}

__device__ __forceinline__ unsigned long long bfs_pack(uint32_t lo, uint32_t hi) {
    return (unsigned long long)lo | ((unsigned long long)hi << 32);
}

// LDS of a workgroup's queue appends.
struct BfsAppendLds {
    uint32_t cnt[4];
    unsigned long long deg[4];
    uint32_t base;
};

constexpr uint32_t kRoundChunks = kBfsRound / 256;  // 64-vertex chunks per wave per workgroup round

// One workgroup round's joins (every thread calls it; the round covers kBfsRound consecutive
// vertices from `blk`, wave w holding chunks w * kRoundChunks + k, and all of them lie in one
// queue segment): jm[k] is this wave's join mask of its chunk k.  The joined vertices are
// appended to the next queue's segment with one atomic, whose packed increment also counts
// them and their out-degrees; joined hubs (out-degree above `large`) also go to the next hub
// queue (one atomic per wave that has any; hubs are few) and their degrees into `hacc`.
__device__ __forceinline__ void bfs_append_round(const BfsArgs& a, const unsigned long long* jm, uint64_t blk,
                                                 uint32_t* __restrict__ q_next, BfsLine* __restrict__ seg_next,
                                                 uint32_t* __restrict__ hub_next, unsigned long long& hacc,
                                                 BfsAppendLds& L) {
    const uint32_t lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const uint32_t sg = (uint32_t)(blk / a.span);
    const uint32_t v0 = (uint32_t)blk + wv * kRoundChunks * kWave + lane;  // this lane's vertex in chunk 0
    uint32_t deg[kRoundChunks], dsum = 0, cnt = 0;
#pragma unroll
    for (uint32_t k = 0; k < kRoundChunks; ++k) {
        const uint32_t v = v0 + k * kWave;
        deg[k] = (jm[k] >> lane) & 1ull ? a.row_ptr[v + 1] - a.row_ptr[v] : 0u;
    }
    unsigned long long hm[kRoundChunks], hany = 0;
    uint32_t hcnt = 0;
#pragma unroll
    for (uint32_t k = 0; k < kRoundChunks; ++k) {
        dsum += deg[k];
        cnt += (uint32_t)__popcll(jm[k]);
        hm[k] = __ballot(deg[k] > a.large);  // deg is 0 unless the lane's vertex joined
        hany |= hm[k];
        hcnt += (uint32_t)__popcll(hm[k]);
    }
    if (hany) {
        uint32_t hb = 0;
        if (lane == 0) hb = atomicAdd(&a.ctr->n_hub, hcnt);
        hb = __shfl(hb, 0, kWave);
#pragma unroll
        for (uint32_t k = 0; k < kRoundChunks; ++k) {
            if ((hm[k] >> lane) & 1ull) {
                hub_next[hb + lanes_below(hm[k])] = v0 + k * kWave;
                hacc += deg[k];
            }
            hb += (uint32_t)__popcll(hm[k]);
        }
    }
    const unsigned long long dtot = wave_sum_u64(dsum);
    if (lane == 0) {
        L.cnt[wv] = cnt;
        L.deg[wv] = dtot;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t c = L.cnt[0] + L.cnt[1] + L.cnt[2] + L.cnt[3];
        const unsigned long long d = L.deg[0] + L.deg[1] + L.deg[2] + L.deg[3];
        L.base = c ? (uint32_t)atomicAdd(&seg_next[sg].v, bfs_pack(c, (uint32_t)d)) : 0u;
    }
    __syncthreads();
    uint32_t off = L.base;
    for (uint32_t w = 0; w < wv; ++w) off += L.cnt[w];
#pragma unroll
    for (uint32_t k = 0; k < kRoundChunks; ++k) {
        if ((jm[k] >> lane) & 1ull) q_next[(size_t)sg * a.span + off + lanes_below(jm[k])] = v0 + k * kWave;
        off += (uint32_t)__popcll(jm[k]);
    }
    __syncthreads();  // L is reused by the next round
}

// End of a counting kernel: the workgroup's hub degrees and probes go into its slot (one
// atomic), then, with kDecide, it draws a ticket; returns true (to every thread) in the last
// workgroup to finish.  Tickets are drawn only after the counter atomics have returned (the
// asm makes the increment depend on their results), and every append of the block returned
// before a barrier above, so when the last ticket is drawn every workgroup's counts have been
// performed at the device's coherence point.  Device-scope atomics and atomic loads are
// coherent across the XCDs' L2 caches without a fence (a device-scope release fence would
// write the L2 back).  Two levels of tickets (per slot, then across slots) keep the atomics on
// any one address to about grid / kBfsSlots.
template <bool kDecide>
__device__ __forceinline__ bool bfs_finish(BfsCounters* __restrict__ ctr, unsigned long long wacc) {
    __shared__ unsigned long long s_sum[4];
    __shared__ uint32_t s_last;
    const uint32_t lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    wacc = wave_sum_u64(wacc);
    if (lane == 0) s_sum[wv] = wacc;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long tw = s_sum[0] + s_sum[1] + s_sum[2] + s_sum[3];
        const uint32_t s = blockIdx.x % kBfsSlots;
        unsigned long long ret = 0;
        if (tw) ret = atomicAdd(&ctr->work[s].v, tw);
        if (kDecide) {
            uint32_t inc = 1;
            asm volatile("" : "+v"(inc) : "v"(ret));
            const uint32_t members = (gridDim.x - s + kBfsSlots - 1) / kBfsSlots;  // workgroups of slot s
            const uint32_t nslots = gridDim.x < kBfsSlots ? gridDim.x : kBfsSlots;
            bool last = false;
            if (atomicAdd(&ctr->sub[s].v, inc) == members - 1) {
                uint32_t inc2 = 1;
                asm volatile("" : "+v"(inc2) : "v"(members));
                last = atomicAdd(&ctr->top, inc2) == nslots - 1;
            }
            s_last = last ? 1u : 0u;
        }
    }
    if constexpr (!kDecide) return false;
    __syncthreads();
    return s_last != 0;
}

// ---- superstep bodies ---------------------------------------------------------------

// Prefix of the frontier queue's segment counts in LDS (one wave scans them): entries
// [pre[s], pre[s + 1]) of the frontier live in segment s.
__device__ __forceinline__ void bfs_queue_prefix(const BfsLine* __restrict__ seg, uint32_t* pre) {
    if (threadIdx.x < kWave) {
        const uint32_t lane = threadIdx.x;
        const uint32_t c = lane < kBfsSlots ? (uint32_t)seg[lane].v : 0u;
        const uint32_t incl = wave_incl_scan(c);
        if (lane < kBfsSlots) pre[lane + 1] = incl;
        if (lane == 0) pre[0] = 0;
    }
    __syncthreads();
}

// Queue entry i of the frontier.
__device__ __forceinline__ uint32_t bfs_queue_at(const BfsArgs& a, const uint32_t* __restrict__ q, const uint32_t* pre,
                                                 uint32_t i) {
    uint32_t s = 0;  // last segment with pre[s] <= i
#pragma unroll
    for (uint32_t step = kBfsSlots / 2; step; step >>= 1)
        if (pre[s + step] <= i) s += step;
    return q[(size_t)s * a.span + (i - pre[s])];
}

// Output of an exact-claim push: its winners go to segment 0 of the next queue (which may run on
// over the space of the others), the next bitmap and the hub queue.
struct BfsExactOut {
    unsigned long long* nxt;
    uint32_t* q_next;
    BfsLine* seg_next;
    uint32_t* hub_next;
};

// A push claim of v for level `next`; a vertex whose visited bit is set is skipped (the bitmap
// is small enough to stay in L2, the level array is not).  Plain: `level[v] = next`, a plain
// store; every writer stores the same value, so the race is benign (as in the reference) and
// the compaction after the push finds the vertices that joined.  Exact: a compare-and-swap
// from -1 picks one winner per vertex, which sets its visited bit and is recorded at once
// (bfs_record_round).
template <bool kExact>
__device__ __forceinline__ bool bfs_claim(const BfsArgs& a, uint32_t v, int32_t next) {
    const unsigned long long bit = 1ull << (v & 63);
    if (a.visited[v >> 6] & bit) return false;
    if constexpr (!kExact) {
        a.level[v] = next;
        return false;
    } else {
        if (atomicCAS(&a.level[v], -1, next) != -1) return false;
        atomicOr(&a.visited[v >> 6], bit);
        return true;
    }
}

constexpr uint32_t kPushUnroll = 4;  // out-edges per lane per round of the push, loads issued together

// Winners of a round of an exact-claim push (the calling wave is converged; won[j] / v[j] for
// the round's kPushUnroll edges of this lane): one packed atomic per wave appends them to
// segment 0 of the next queue and counts them with their out-degrees; each sets its bit in the
// next bitmap; hubs also go to the next hub queue.
__device__ __forceinline__ void bfs_record_round(const BfsArgs& a, const BfsExactOut& o, const bool* won,
                                                 const uint32_t* v, unsigned long long& hacc) {
    unsigned long long wm[kPushUnroll], any = 0;
#pragma unroll
    for (uint32_t j = 0; j < kPushUnroll; ++j) {
        wm[j] = __ballot(won[j]);
        any |= wm[j];
    }
    if (!any) return;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    uint32_t deg[kPushUnroll], dsum = 0, cnt = 0;
#pragma unroll
    for (uint32_t j = 0; j < kPushUnroll; ++j) deg[j] = won[j] ? a.row_ptr[v[j] + 1] - a.row_ptr[v[j]] : 0u;
#pragma unroll
    for (uint32_t j = 0; j < kPushUnroll; ++j) {
        if (won[j]) atomicOr(&o.nxt[v[j] >> 6], 1ull << (v[j] & 63));
        dsum += deg[j];
        cnt += (uint32_t)__popcll(wm[j]);
    }
    const unsigned long long dtot = wave_sum_u64(dsum);
    uint32_t base = 0;
    if (lane == 0) base = (uint32_t)atomicAdd(&o.seg_next[0].v, bfs_pack(cnt, (uint32_t)dtot));
    base = __shfl(base, 0, kWave);
    unsigned long long hm[kPushUnroll], hany = 0;
    uint32_t hcnt = 0;
#pragma unroll
    for (uint32_t j = 0; j < kPushUnroll; ++j) {
        if (won[j]) o.q_next[base + lanes_below(wm[j])] = v[j];
        base += (uint32_t)__popcll(wm[j]);
        hm[j] = __ballot(won[j] && deg[j] > a.large);
        hany |= hm[j];
        hcnt += (uint32_t)__popcll(hm[j]);
    }
    if (hany) {
        uint32_t hb = 0;
        if (lane == 0) hb = atomicAdd(&a.ctr->n_hub, hcnt);
        hb = __shfl(hb, 0, kWave);
#pragma unroll
        for (uint32_t j = 0; j < kPushUnroll; ++j) {
            if ((hm[j] >> lane) & 1ull) {
                o.hub_next[hb + lanes_below(hm[j])] = v[j];
                hacc += deg[j];
            }
            hb += (uint32_t)__popcll(hm[j]);
        }
    }
}

// Wave-cooperative push over the frontier queue (hubs and vertices without out-edges are
// skipped: out-degree taken as 0).  Of the `nwaves` waves taking part (`wave0` is this wave's
// index among them), wave w holds queue entries w, w + nwaves, w + 2*nwaves, ... one per lane,
// so a run of high-degree vertices spreads over many waves instead of sharing one; the lanes
// then share the entries' out-edges through an in-wave prefix scan, kPushUnroll edges per lane
// per round.
template <bool kExact>
__device__ __forceinline__ void push_queue(const BfsArgs& a, const uint32_t* __restrict__ q, const uint32_t* pre,
                                           int32_t next, uint32_t wave0, uint32_t nwaves, uint32_t (*s_beg)[kWave],
                                           uint32_t (*s_off)[kWave], const BfsExactOut& o, unsigned long long& hacc) {
    const uint32_t lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const uint32_t qn = pre[kBfsSlots];
    for (uint64_t r = 0; wave0 + r * kWave * nwaves < qn; ++r) {  // uniform across the wave
        const uint64_t i = wave0 + (r * kWave + lane) * nwaves;
        uint32_t beg = 0, deg = 0;
        if (i < qn) {
            const uint32_t u = bfs_queue_at(a, q, pre, (uint32_t)i);
            beg = a.row_ptr[u];
            deg = a.row_ptr[u + 1] - beg;
            if (deg > a.large) deg = 0;  // a hub: expanded from the hub queue
        }
        const uint32_t incl = wave_incl_scan(deg);
        const uint32_t total = __shfl(incl, kWave - 1, kWave);
        s_beg[wv][lane] = beg;
        s_off[wv][lane] = incl - deg;
        __builtin_amdgcn_wave_barrier();
        for (uint32_t e0 = 0; e0 < total; e0 += kWave * kPushUnroll) {  // uniform across the wave
            uint32_t v[kPushUnroll];
#pragma unroll
            for (uint32_t j = 0; j < kPushUnroll; ++j) {
                const uint32_t e = e0 + j * kWave + lane;
                v[j] = 0xFFFFFFFFu;
                if (e < total) {
                    // owner = last lane whose exclusive offset is <= e (it has deg > 0)
                    uint32_t lo = 0;
#pragma unroll
                    for (uint32_t step = kWave / 2; step; step >>= 1)
                        if (s_off[wv][lo + step] <= e) lo += step;
                    v[j] = a.col[s_beg[wv][lo] + (e - s_off[wv][lo])];
                }
            }
            bool won[kPushUnroll];
#pragma unroll
            for (uint32_t j = 0; j < kPushUnroll; ++j)
                won[j] = v[j] != 0xFFFFFFFFu && bfs_claim<kExact>(a, v[j], next);
            if constexpr (kExact) bfs_record_round(a, o, won, v, hacc);
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// Exact-claim push of a frontier of at most 64 vertices (hubs included): every wave reads the
// whole frontier and its out-degree prefix, and wave w of the grid takes out-edges
// [w * 64 * kPushUnroll, (w + 1) * 64 * kPushUnroll) of their concatenation, so a frontier of
// one high-degree vertex still spreads over many waves.
__device__ __forceinline__ void push_tiny(const BfsArgs& a, const uint32_t* __restrict__ q, const uint32_t* pre,
                                          int32_t next, uint32_t (*s_beg)[kWave], uint32_t (*s_off)[kWave],
                                          const BfsExactOut& o, unsigned long long& hacc) {
    const uint32_t lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const uint32_t qn = pre[kBfsSlots];
    const uint64_t e_begin = ((uint64_t)blockIdx.x * (blockDim.x / kWave) + wv) * kWave * kPushUnroll;
    uint32_t beg = 0, deg = 0;
    if (lane < qn) {
        const uint32_t u = bfs_queue_at(a, q, pre, lane);
        beg = a.row_ptr[u];
        deg = a.row_ptr[u + 1] - beg;
    }
    const uint32_t incl = wave_incl_scan(deg);
    const uint32_t total = __shfl(incl, kWave - 1, kWave);
    if (e_begin >= total) return;  // uniform across the wave
    s_beg[wv][lane] = beg;
    s_off[wv][lane] = incl - deg;
    __builtin_amdgcn_wave_barrier();
    uint32_t v[kPushUnroll];
#pragma unroll
    for (uint32_t j = 0; j < kPushUnroll; ++j) {
        const uint64_t e = e_begin + j * kWave + lane;
        v[j] = 0xFFFFFFFFu;
        if (e < total) {
            uint32_t lo = 0;
#pragma unroll
            for (uint32_t step = kWave / 2; step; step >>= 1)
                if (s_off[wv][lo + step] <= e) lo += step;
            v[j] = a.col[s_beg[wv][lo] + ((uint32_t)e - s_off[wv][lo])];
        }
    }
    bool won[kPushUnroll];
#pragma unroll
    for (uint32_t j = 0; j < kPushUnroll; ++j) won[j] = v[j] != 0xFFFFFFFFu && bfs_claim<true>(a, v[j], next);
    bfs_record_round(a, o, won, v, hacc);
}

// Hubs: each vertex of the hub queue is split into `split` contiguous edge ranges, one per
// workgroup item; the `nwg` workgroups taking part stride over the items.
template <bool kExact>
__device__ __forceinline__ void push_hubs(const BfsArgs& a, const uint32_t* __restrict__ q, uint32_t qn, int32_t next,
                                          uint32_t wg0, uint32_t nwg, const BfsExactOut& o,
                                          unsigned long long& hacc) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    uint32_t split = nwg / qn;
    split = split < kLargeSplit ? kLargeSplit : (split > 4096u ? 4096u : split);
    for (uint64_t w = wg0; w < (uint64_t)qn * split; w += nwg) {
        const uint32_t u = q[w / split], part = (uint32_t)(w % split);
        const uint32_t b = a.row_ptr[u], deg = a.row_ptr[u + 1] - b;
        const uint32_t per = (deg + split - 1) / split;
        const uint32_t lo = b + min(deg, part * per), hi = b + min(deg, (part + 1) * per);
        for (uint32_t k0 = lo + (threadIdx.x - lane) * kPushUnroll; k0 < hi;
             k0 += blockDim.x * kPushUnroll) {  // uniform per wave
            uint32_t v[kPushUnroll];
            bool won[kPushUnroll];
#pragma unroll
            for (uint32_t j = 0; j < kPushUnroll; ++j) {
                const uint32_t k = k0 + j * kWave + lane;
                v[j] = k < hi ? a.col[k] : 0xFFFFFFFFu;
            }
#pragma unroll
            for (uint32_t j = 0; j < kPushUnroll; ++j)
                won[j] = v[j] != 0xFFFFFFFFu && bfs_claim<kExact>(a, v[j], next);
            if constexpr (kExact) bfs_record_round(a, o, won, v, hacc);
        }
    }
}

constexpr uint32_t kPullLane = 32;  // in-degree up to which one lane scans a vertex alone
constexpr uint32_t kPullGroup = 4;  // in-edges a lane probes per round

// Bottom-up over 64 consecutive vertices from `base` (one per lane; the wave is converged): an
// unvisited vertex (visited bitmap) joins if any in-neighbour is in the frontier bitmap.  A lane scans a vertex
// of in-degree <= kPullLane alone; above that the whole wave scans it 64 in-edges at a time and
// stops at the first hit (ballot), so one high in-degree vertex does not hold its wave for
// hundreds of sequential probes.  Returns the join mask, which is also written as the chunk's
// word of the next bitmap.  `probes` counts in-edges examined up to and including the first
// hit in either case.
__device__ __forceinline__ unsigned long long pull_chunk(const BfsArgs& a, const unsigned long long* __restrict__ cur,
                                                         unsigned long long* __restrict__ nxt, uint64_t base,
                                                         int32_t next, unsigned long long& probes) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t n = a.n;
    if (base >= n) return 0;  // uniform across the wave
    const uint32_t v = (uint32_t)base + lane;
    const unsigned long long vis = a.visited[base / kWave];
    uint32_t b = 0, e = 0;
    bool todo = false, joined = false;
    if (v < n && !((vis >> lane) & 1ull)) {
        b = a.cptr[v];
        e = a.cptr[v + 1];
        todo = b < e;
    }
    const bool big = todo && e - b > kPullLane;
    if (todo && !big) {
        // kPullGroup in-edges per round: their source and bitmap loads are issued together,
        // so a round costs two memory round trips instead of two per in-edge
        for (uint32_t k0 = b; k0 < e; k0 += kPullGroup) {
            uint32_t u[kPullGroup];
            unsigned long long word[kPullGroup];
#pragma unroll
            for (uint32_t j = 0; j < kPullGroup; ++j) u[j] = k0 + j < e ? a.csrc[k0 + j] : 0u;
#pragma unroll
            for (uint32_t j = 0; j < kPullGroup; ++j) word[j] = k0 + j < e ? cur[u[j] >> 6] : 0ull;
            uint32_t hit = kPullGroup;
#pragma unroll
            for (uint32_t j = kPullGroup; j-- > 0;)
                if ((word[j] >> (u[j] & 63)) & 1ull) hit = j;
            if (hit < kPullGroup) {
                probes += hit + 1;
                a.level[v] = next;
                joined = true;
                break;
            }
            probes += min(e - k0, kPullGroup);
        }
    }
    unsigned long long bigjoin = 0;
    for (unsigned long long bm = __ballot(big); bm; bm &= bm - 1) {
        const uint32_t owner = __builtin_ctzll(bm);
        const uint32_t vb = __shfl(b, owner, kWave), ve = __shfl(e, owner, kWave);
        for (uint32_t k0 = vb; k0 < ve; k0 += kWave) {
            const uint32_t k = k0 + lane;
            bool hit = false;
            if (k < ve) {
                const uint32_t u = a.csrc[k];
                hit = (cur[u >> 6] >> (u & 63)) & 1ull;
            }
            const unsigned long long hm = __ballot(hit);
            if (hm) {
                if (lane == 0) {
                    probes += __builtin_ctzll(hm) + 1;
                    a.level[(uint32_t)base + owner] = next;
                }
                bigjoin |= 1ull << owner;
                break;
            }
            if (lane == 0) probes += min(ve - k0, (uint32_t)kWave);
        }
    }
    const unsigned long long jm = __ballot(joined) | bigjoin;
    if (lane == 0) {
        nxt[base / kWave] = jm;
        a.visited[base / kWave] = vis | jm;  // the wave owns this word in a pull
    }
    return jm;
}

// Pull superstep: workgroup rounds of kBfsRound consecutive vertices, wave w taking chunks
// w * kRoundChunks + k; each wave writes its chunks' words of the next bitmap whole and the
// round's joins are appended to the next queue, so the pull needs no compaction.
__device__ __forceinline__ void pull_all(const BfsArgs& a, const unsigned long long* __restrict__ cur,
                                         unsigned long long* __restrict__ nxt, uint32_t* __restrict__ q_next,
                                         BfsLine* __restrict__ seg_next, uint32_t* __restrict__ hub_next, int32_t next,
                                         unsigned long long& hacc, unsigned long long& probes, BfsAppendLds& L) {
    const uint32_t wv = threadIdx.x / kWave;
    for (uint64_t blk = (uint64_t)blockIdx.x * kBfsRound; blk < a.n; blk += (uint64_t)gridDim.x * kBfsRound) {
        unsigned long long jm[kRoundChunks];
#pragma unroll 1
        for (uint32_t k = 0; k < kRoundChunks; ++k)
            jm[k] = pull_chunk(a, cur, nxt, blk + (wv * kRoundChunks + k) * kWave, next, probes);
        bfs_append_round(a, jm, blk, q_next, seg_next, hub_next, hacc, L);
    }
}

// Sums of the next frontier's segment counters and of the work slots (each wave reduces them
// itself, so every thread gets them without a barrier).
__device__ __forceinline__ void bfs_sum_counters(BfsCounters* __restrict__ ctr, uint32_t par_next,
                                                 unsigned long long& joined, unsigned long long& work) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const unsigned long long j = lane < kBfsSlots ? load_now(&ctr->seg[par_next][lane].v) : 0ull;
    const unsigned long long w = lane < kBfsSlots ? load_now(&ctr->work[lane].v) : 0ull;
    joined = wave_sum_u64(j);
    work = wave_sum_u64(w);
}

// Counters cleared for the next superstep (by one workgroup, after every thread has read
// them): the segment counters of parity `par_free` (a frontier already expanded), the work
// slots, the tickets and the hub count.
__device__ __forceinline__ void bfs_clear_counters(BfsCounters* __restrict__ ctr, uint32_t par_free) {
    if (threadIdx.x < kBfsSlots) {
        ctr->seg[par_free][threadIdx.x].v = 0;
        ctr->work[threadIdx.x].v = 0;
        ctr->sub[threadIdx.x].v = 0;
    }
    if (threadIdx.x == 0) {
        ctr->top = 0;
        ctr->n_hub = 0;
    }
}

// The next superstep from this one's counters (the deciding workgroup's first wave, after
// every workgroup of the superstep has added its counters): direction by the heuristic or the
// learned policy, hub queue split and the termination vote, written into the control block;
// the counters are cleared for the next superstep.  The updated control block waits in LDS
// while the policy is evaluated, so it holds no registers there.
__device__ __forceinline__ void bfs_decide(const BfsArgs& a) {
    __shared__ BfsCtl s_c;
    __shared__ float s_params[kMlpWaveLds];
    if (threadIdx.x >= kWave) return;
    // with a policy, its parameters start for LDS now, behind the counter loads, holding no
    // registers; a decision that does not consult it only waits for them at the end
    if (a.policy) mlp_stage_wave_begin(a.policy, s_params);
    BfsCounters* ctr = a.ctr;
    const uint32_t lane = threadIdx.x;
    uint32_t par;
    bool pull, ask = false, plain = false, done;
    double rec[MEDUSA_TRACE_FIELDS] = {};
    {
        BfsCtl c = load_now_block(a.ctl);
        par = (uint32_t)c.step & 1u;
        unsigned long long joined, work;
        bfs_sum_counters(ctr, par ^ 1u, joined, work);
        const uint32_t n_hub = load_now(&ctr->n_hub);
        const uint32_t n = a.n, n_new = (uint32_t)joined;
        const unsigned long long deg_sum = joined >> 32, deg_large = (uint32_t)work, probes = work >> 32;
        // bytes of the superstep just finished (every expansion clears one bitmap: 8 per word):
        // push = queue entry and out-offsets per frontier vertex (12) + target id and visited bit
        // per out-edge (8); exact claims then add compare-and-swap (4), two bit ORs (16),
        // out-offsets (8) and queue entry (4) per join, plain claims the store per join (4) and the compaction's level
        // pass (4n), next and visited bitmaps (16 per word), out-offsets and queue entry per join
        // (12); pull = visited word per 64 vertices and in-offsets of unvisited vertices (8) +
        // source id and bitmap word per probed in-edge (8) + next and visited bitmaps (16 per
        // word) + level, out-offsets and queue entry per join (16)
        const unsigned long long kb =
            c.pull    ? 8ull * c.n_u + 8ull * probes + 32ull * a.nwords + 16ull * n_new
            : c.small ? 12ull * c.n_f + 8ull * c.m_f + 8ull * a.nwords + 32ull * n_new
                      : 12ull * c.n_f + 8ull * c.m_f + 4ull * n + 24ull * a.nwords + 16ull * n_new;
        c.kernel_bytes += kb;
        c.alg_bytes += kb;
        c.slots += c.pending ? 2u : 1u;
        c.pending = 0;
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
            c.n_small = n_new - n_hub;
            c.n_large = n_hub;
            c.hub_wgs = bfs_hub_wgs(a.grid, c.n_small, n_hub, deg_large, deg_sum);
        }
        done = c.done != 0;
        pull = c.pull != 0;
        if (!done) {
            if (c.mode == 1) {
                pull = false;
            } else if (c.mode == 2) {
                pull = true;
            } else if (c.mode == 3 && a.policy &&
                       bfs_policy_consulted(pull, (double)c.n_f, (double)c.m_f, (double)c.m_u, (double)c.n,
                                            Sched::kAlphaBfs, Sched::kBeta)) {
                ask = true;
                // the heuristic's choice stands unless the policy predicts a clear gain
                plain = !pull ? (double)c.m_f > (double)c.m_u / Sched::kAlphaBfs
                              : !((double)c.n_f < (double)c.n / Sched::kBeta);
                rec[0] = (double)c.step;
                rec[1] = (double)c.n_f;
                rec[2] = (double)c.m_f;
                rec[3] = (double)c.n_u;
                rec[4] = (double)c.m_u;
                rec[5] = (double)c.n;
                rec[6] = (double)c.m;
            } else {
                // direction-optimising heuristic (Beamer, Asanovic, Patterson, SC'12)
                if (!pull)
                    pull = (double)c.m_f > (double)c.m_u / Sched::kAlphaBfs;
                else
                    pull = !((double)c.n_f < (double)c.n / Sched::kBeta);
            }
        }
        if (lane == 0) s_c = c;
    }
    if (ask) {  // uniform across the wave
        float c_push, c_pull;
        policy_direction_costs_wave(rec, a.policy, s_params, c_push, c_pull, true);
        pull = policy_pull(c_push, c_pull, plain);
    } else if (a.policy) {
        mlp_stage_wave_wait();
    }
    bfs_clear_counters(ctr, par);  // this wave read them above
    if (lane == 0) {
        BfsCtl c = s_c;
        if (!done) {
            c.pull = pull ? 1u : 0u;
            c.small = bfs_small(pull, c.n_f, c.m_f, a.exact_edges);
            c.supersteps += 1;
            if (pull)
                c.pull_steps += 1;
            else
                c.push_steps += 1;
        }
        *a.ctl = c;
    }
}

// ---- kernels (both drivers) ---------------------------------------------------------

// First decision (superstep 0) from the first frontier's counts, by the calling workgroup's
// first wave (the others return; every thread passes the same values): the control block is
// written and the work slots, tickets and hub count are cleared (queue 0's counters stay:
// superstep 0 reads them).
__device__ __forceinline__ void bfs_first_decision(const BfsArgs& a, unsigned long long joined,
                                                   unsigned long long work, uint32_t n_hub, unsigned long long m,
                                                   uint32_t mode) {
    if (threadIdx.x >= kWave) return;
    BfsCounters* ctr = a.ctr;
    BfsCtl c{};
    c.mode = mode;
    c.n = a.n;
    c.m = m;
    c.step = 0;
    c.n_f = (uint32_t)joined;
    c.m_f = joined >> 32;
    c.n_u = a.n - c.n_f;
    c.m_u = m - c.m_f;
    c.reached = c.n_f;
    c.reached_edges = c.m_f;
    c.n_small = (uint32_t)c.n_f - n_hub;
    c.n_large = n_hub;
    c.hub_wgs = bfs_hub_wgs(a.grid, c.n_small, c.n_large, (uint32_t)work, c.m_f);
    bool pull = false;
    if (c.mode == 2) {
        pull = true;
    } else if (c.mode == 3 && a.policy &&
               bfs_policy_consulted(false, (double)c.n_f, (double)c.m_f, (double)c.m_u, (double)c.n,
                                    Sched::kAlphaBfs, Sched::kBeta)) {
        const double rec[MEDUSA_TRACE_FIELDS] = {0.0, (double)c.n_f, (double)c.m_f, (double)c.n_u, (double)c.m_u,
                                                 (double)c.n, (double)c.m, 0.0, 0.0};
        __shared__ float s_params[kMlpWaveLds];
        float c_push, c_pull;
        policy_direction_costs_wave(rec, a.policy, s_params, c_push, c_pull);
        pull = policy_pull(c_push, c_pull, (double)c.m_f > (double)c.m_u / Sched::kAlphaBfs);
    } else if (c.mode == 0 || c.mode == 3) {  // heuristic, or the policy far from its switching point
        pull = (double)c.m_f > (double)c.m_u / Sched::kAlphaBfs;
    }
    c.pull = pull ? 1u : 0u;
    c.small = bfs_small(pull, c.n_f, c.m_f, a.exact_edges);
    c.supersteps = 1;
    c.push_steps = pull ? 0 : 1;
    c.pull_steps = pull ? 1 : 0;
    if (threadIdx.x < kBfsSlots) {
        ctr->work[threadIdx.x].v = 0;
        ctr->sub[threadIdx.x].v = 0;
    }
    if (threadIdx.x == 0) {
        ctr->top = 0;
        ctr->n_hub = 0;
        *a.ctl = c;
    }
}

// First decision after k_bfs_mark placed several sources; one workgroup.
__global__ __launch_bounds__(256) void k_bfs_first_decide(BfsArgs a, unsigned long long m, uint32_t mode) {
    unsigned long long joined, work;
    bfs_sum_counters(a.ctr, 0, joined, work);
    bfs_first_decision(a, joined, work, a.ctr->n_hub, m, mode);
}

// Level reset, counters and bitmap 1 cleared, and the first frontier when there is one source:
// with `single` < n that vertex is placed at level 0, in bitmap 0 and in queue 0 here, and with
// kDecide workgroup 0 also makes the first decision (the device-driven loop); otherwise
// k_bfs_mark places the sources.
template <bool kDecide>
__global__ __launch_bounds__(256) void k_bfs_seed(BfsArgs a, uint32_t single, unsigned long long m, uint32_t mode) {
    __shared__ uint32_t s_deg;
    const uint32_t n = a.n;
    const bool one = single < n;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride)
        a.level[v] = v == single ? 0 : -1;
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < a.nwords; w += stride) {
        a.bm[0][w] = a.visited[w] = one && w == single / kWave ? 1ull << (single & 63) : 0ull;
        a.bm[1][w] = 0;  // superstep 0's next bitmap (superstep 0 clears bitmap 2 itself)
    }
    if (blockIdx.x != 0) return;
    bfs_clear_counters(a.ctr, 0);
    bfs_clear_counters(a.ctr, 1);
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t deg = one ? a.row_ptr[single + 1] - a.row_ptr[single] : 0u;
        const bool hub = one && deg > a.large;
        s_deg = deg;
        if (one) {
            const uint32_t sg = single / a.span;
            a.ctr->seg[0][sg].v = bfs_pack(1u, deg);
            a.q[0][(size_t)sg * a.span] = single;
        }
        if (hub) a.hub[0][0] = single;
        if (!kDecide) {  // the host reads the first frontier's counters
            a.ctr->work[0].v = hub ? deg : 0ull;
            a.ctr->n_hub = hub ? 1u : 0u;
        }
    }
    if constexpr (kDecide) {
        __syncthreads();
        const uint32_t deg = s_deg;
        const bool hub = deg > a.large;
        bfs_first_decision(a, bfs_pack(1u, deg), hub ? deg : 0ull, hub ? 1u : 0u, m, mode);
    }
}

// Several sources: each is placed at level 0, in bitmap 0 and in queue 0; a vertex given more
// than once is counted once (the lane that sets its bit).  One queue atomic per source.
__global__ __launch_bounds__(256) void k_bfs_mark(BfsArgs a, const uint32_t* __restrict__ src, uint32_t count) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    unsigned long long hacc = 0;
    for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x - lane); i0 < count;
         i0 += (uint64_t)gridDim.x * blockDim.x) {  // uniform per wave
        const uint64_t i = i0 + lane;
        bool won = false;
        uint32_t v = 0, deg = 0;
        if (i < count) {
            v = src[i];
            const unsigned long long bit = 1ull << (v & 63);
            won = !(atomicOr(&a.bm[0][v >> 6], bit) & bit);
            if (won) atomicOr(&a.visited[v >> 6], bit);
        }
        if (won) {
            a.level[v] = 0;
            deg = a.row_ptr[v + 1] - a.row_ptr[v];
            const uint32_t sg = v / a.span;
            const uint32_t off = (uint32_t)atomicAdd(&a.ctr->seg[0][sg].v, bfs_pack(1u, deg));
            a.q[0][(size_t)sg * a.span + off] = v;
        }
        const bool hub = won && deg > a.large;
        const unsigned long long hm = __ballot(hub);
        if (hm) {
            uint32_t hb = 0;
            if (lane == 0) hb = atomicAdd(&a.ctr->n_hub, (uint32_t)__popcll(hm));
            hb = __shfl(hb, 0, kWave);
            if (hub) {
                a.hub[0][hb + lanes_below(hm)] = v;
                hacc += deg;
            }
        }
    }
    bfs_finish<false>(a.ctr, hacc);
}

// Host-driven superstep t: counters cleared except the frontier's queue segments (parity t % 2).
__global__ __launch_bounds__(64) void k_zero_counters(BfsCounters* c, uint32_t par_free) {
    bfs_clear_counters(c, par_free);
}

// After a push with plain claims: the next frontier = {v : level[v] == step + 1}, as its
// bitmap (one ballot per 64 vertices), queue, hub queue and counters, and the visited bitmap
// rebuilt from the levels, in workgroup rounds of
// kBfsRound consecutive vertices (wave w the chunks w * kRoundChunks + k), whose level loads are
// issued together.
__device__ __forceinline__ void compact_level(const BfsArgs& a, int32_t next, unsigned long long* __restrict__ nxt,
                                              uint32_t* __restrict__ q_next, BfsLine* __restrict__ seg_next,
                                              uint32_t* __restrict__ hub_next, unsigned long long& hacc,
                                              BfsAppendLds& L) {
    const uint32_t lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    for (uint64_t blk = (uint64_t)blockIdx.x * kBfsRound; blk < a.n; blk += (uint64_t)gridDim.x * kBfsRound) {
        const uint64_t c0 = blk / kWave + wv * kRoundChunks;  // first chunk of this wave
        int32_t lv[kRoundChunks];
#pragma unroll
        for (uint32_t k = 0; k < kRoundChunks; ++k) {
            const uint64_t v = (c0 + k) * kWave + lane;
            lv[k] = v < a.n ? a.level[v] : -1;
        }
        unsigned long long jm[kRoundChunks];
#pragma unroll
        for (uint32_t k = 0; k < kRoundChunks; ++k) {
            jm[k] = __ballot(lv[k] == next);
            const unsigned long long vm = __ballot(lv[k] >= 0);
            if (lane == 0 && c0 + k < a.nwords) {
                nxt[c0 + k] = jm[k];
                a.visited[c0 + k] = vm;
            }
        }
        bfs_append_round(a, jm, blk, q_next, seg_next, hub_next, hacc, L);
    }
}

// One launch of the superstep loop, doing what the control block says:
//   pending: the compaction after a push with plain claims;
//   pull:    every unvisited vertex; the next frontier's bitmap, queue, hub queue and
//            counters come out of the same pass;
//   small:   push with exact claims (the frontier's out-edges fit one queue segment); the
//            winners are recorded as they are claimed; with at most 64 frontier vertices every
//            wave takes an equal share of their out-edges (hubs included);
//   else:    push with plain claims; it sets `pending`, so the next launch compacts.
// The first hub_wgs workgroups of a push take the hub queue, the rest the frontier queue.  Every
// expansion clears the bitmap that superstep + 1 builds.  With kDecide the last workgroup to
// finish a counting phase decides the next superstep; otherwise the host reads the counters.
template <bool kDecide>
__global__ __launch_bounds__(256) void k_bfs_super(BfsArgs a) {
    __shared__ uint32_t s_beg[4][kWave], s_off[4][kWave];
    __shared__ uint32_t s_pre[kBfsSlots + 1];
    __shared__ BfsAppendLds s_app;
    const BfsCtl* ctl = a.ctl;
    if (ctl->done) return;
    const int32_t step = ctl->step, next = step + 1;
    const uint32_t par = (uint32_t)step & 1u, t3 = (uint32_t)step % 3u;
    unsigned long long* nxt = a.bm[t3 == 2 ? 0 : t3 + 1];
    uint32_t* q_next = a.q[par ^ 1u];
    BfsLine* seg_next = a.ctr->seg[par ^ 1u];
    uint32_t* hub_next = a.hub[par ^ 1u];
    unsigned long long hacc = 0;
    if (ctl->pending) {
        compact_level(a, next, nxt, q_next, seg_next, hub_next, hacc, s_app);
        if (bfs_finish<kDecide>(a.ctr, hacc)) bfs_decide(a);
        return;
    }
    unsigned long long* clr = a.bm[t3 == 0 ? 2 : t3 - 1];
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < a.nwords; w += stride) clr[w] = 0;
    if (ctl->pull) {
        unsigned long long probes = 0;
        pull_all(a, a.bm[t3], nxt, q_next, seg_next, hub_next, next, hacc, probes, s_app);
        if (bfs_finish<kDecide>(a.ctr, hacc + (probes << 32))) bfs_decide(a);
        return;
    }
    const BfsExactOut o{nxt, q_next, seg_next, hub_next};
    const uint32_t hub = ctl->hub_wgs, small = ctl->small;
    if (!small && blockIdx.x == 0 && threadIdx.x == 0) a.ctl->pending = 1;  // the only writer of ctl in this launch
    if (small == 2) {
        bfs_queue_prefix(a.ctr->seg[par], s_pre);
        push_tiny(a, a.q[par], s_pre, next, s_beg, s_off, o, hacc);
    } else if (blockIdx.x < hub) {
        if (small)
            push_hubs<true>(a, a.hub[par], ctl->n_large, next, blockIdx.x, hub, o, hacc);
        else
            push_hubs<false>(a, a.hub[par], ctl->n_large, next, blockIdx.x, hub, o, hacc);
    } else {
        bfs_queue_prefix(a.ctr->seg[par], s_pre);
        const uint32_t wg = blockIdx.x - hub, nwg = gridDim.x - hub, w0 = wg * 4 + threadIdx.x / kWave;
        if (small)
            push_queue<true>(a, a.q[par], s_pre, next, w0, nwg * 4, s_beg, s_off, o, hacc);
        else
            push_queue<false>(a, a.q[par], s_pre, next, w0, nwg * 4, s_beg, s_off, o, hacc);
    }
    if (small && bfs_finish<kDecide>(a.ctr, hacc)) bfs_decide(a);
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
    uint32_t* q[2];                 // frontier queues, kBfsSlots * bfs_span(n) entries each
    uint32_t* hub[2];               // hub queues, n entries each
    unsigned long long* bm[3];      // frontier bitmaps, (n + 63) / 64 words each
    unsigned long long* visited;    // visited bitmap, (n + 63) / 64 words
    const uint32_t* sources;        // several sources (device), or null
    BfsCtl* ctl;
    BfsCounters* ctr;
};
static size_t bfs_queue_cap(uint32_t n) { return (size_t)kBfsSlots * bfs_span(n); }
// Control block and counters share one allocation: the control block, then the counters.
constexpr size_t kBfsCtlUnits = 1 + (sizeof(BfsCounters) + sizeof(CtlStorage) - 1) / sizeof(CtlStorage);
static BfsBufs handle_bufs(medusa_graph* g) {
    g->bfs_ctl.ensure(kBfsCtlUnits);
    static_assert(sizeof(BfsCtl) <= sizeof(BfsCtlStorage), "control block storage");
    const size_t nwords = ((size_t)g->n + 63) / 64;
    const size_t qcap = bfs_queue_cap(g->n);
    return {g->level.p,
            {g->q_small.p, g->q_small.p + qcap},
            {g->q_large.p, g->q_large.p + g->n},
            {g->bitmap.p, g->bitmap.p + nwords, g->bitmap.p + 2 * nwords},
            g->bitmap.p + 3 * nwords,
            g->q_aux.p,
            reinterpret_cast<BfsCtl*>(g->bfs_ctl.p),
            reinterpret_cast<BfsCounters*>(g->bfs_ctl.p + 1)};
}

// Frontier out-edges up to which a push claims exactly instead of compacting after it.
static unsigned long long bfs_exact_edges(uint32_t n) { return bfs_span(n); }

static BfsArgs bfs_args(medusa_graph* g, const BfsBufs& B, uint32_t large, uint32_t grid, const float* policy) {
    BfsArgs a{};
    a.row_ptr = g->row_ptr.p;
    a.col = g->col.p;
    a.cptr = g->cptr.p;
    a.csrc = g->csrc.p;
    a.level = B.level;
    a.q[0] = B.q[0];
    a.q[1] = B.q[1];
    a.hub[0] = B.hub[0];
    a.hub[1] = B.hub[1];
    for (int i = 0; i < 3; ++i) a.bm[i] = B.bm[i];
    a.visited = B.visited;
    a.ctr = B.ctr;
    a.ctl = B.ctl;
    a.policy = policy;
    a.n = g->n;
    a.nwords = (g->n + 63) / 64;
    a.large = large;
    a.grid = grid;
    a.span = bfs_span(g->n);
    a.exact_edges = bfs_exact_edges(g->n);
    return a;
}

// Seeding: every source at level 0 and in bitmap 0 (one source by value, several from
// B.sources), the first queue, hub queue and counters; with `decide` also the first decision
// (inside the seed kernel when there is one source).
static void bfs_seed(medusa_graph* g, const BfsBufs& B, const BfsArgs& a, uint32_t nsrc, uint32_t single, bool decide,
                     uint32_t mode) {
    hipStream_t s = g->stream;
    const dim3 grid(grid_for(g->n, 256, 4096));
    const unsigned long long m = g->m;
    if (nsrc == 1 && decide)
        hipLaunchKernelGGL(k_bfs_seed<true>, grid, dim3(256), 0, s, a, single, m, mode);
    else
        hipLaunchKernelGGL(k_bfs_seed<false>, grid, dim3(256), 0, s, a, nsrc == 1 ? single : 0xFFFFFFFFu, m, mode);
    MDS_LAUNCHED();
    if (nsrc > 1) {
        hipLaunchKernelGGL(k_bfs_mark, dim3(grid_for(nsrc, 256, 1024)), dim3(256), 0, s, a, B.sources, nsrc);
        MDS_LAUNCHED();
        if (decide) {
            hipLaunchKernelGGL(k_bfs_first_decide, dim3(1), dim3(256), 0, s, a, m, mode);
            MDS_LAUNCHED();
        }
    }
}

// One device-driven slot: one launch of the superstep kernel (a push with plain claims takes two
// slots, its expansion and its compaction; every other superstep one).
static void bfs_enqueue_slot(medusa_graph* g, const BfsArgs& a) {
    hipLaunchKernelGGL(k_bfs_super<true>, dim3(a.grid), dim3(256), 0, g->stream, a);
}

// Bytes of the seeding: the level reset (4 per vertex), bitmaps 0 and 1 and the visited bitmap
// (24 per word) and the sources (16 each: id, level, bit OR).
static uint64_t bfs_seed_bytes(uint32_t n, uint32_t nwords, uint32_t nsrc) {
    return 4ull * n + 24ull * nwords + 16ull * nsrc;
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
    const BfsArgs a = bfs_args(g, B, large, grid, policy);

    // capture (once per graph, threshold, policy buffer and slot count) a batch of supersteps
    auto captured = [&](GraphCache& gc, uint32_t slots) -> hipGraphExec_t {
        const uintptr_t key[4] = {(uintptr_t)B.level ^ ((uintptr_t)B.hub[0] << 16), (uintptr_t)policy,
                                  ((uintptr_t)grid << 32) | large, (uintptr_t)B.bm[0] ^ ((uintptr_t)slots << 48)};
        if (gc.exec && std::equal(key, key + 4, gc.key)) return gc.exec;
        gc.reset();
        MDS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (uint32_t k = 0; k < slots; ++k) bfs_enqueue_slot(g, a);
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

    // A device output is copied behind every batch, so the usual run (one batch) waits for the
    // stream once; a run that takes more batches copies again after the next one.
    hipPointerAttribute_t attr{};
    const bool out_dev = out != B.level && hipPointerGetAttributes(&attr, out) == hipSuccess &&
                         attr.type == hipMemoryTypeDevice;
    (void)hipGetLastError();  // host pointers unknown to HIP report an error

    MDS_HIP(hipEventRecord(g->ev_begin, s));
    bfs_seed(g, B, a, nsrc, single, true, mode);
    BfsCtl h{};
    for (uint32_t batch = 0;; ++batch) {
        // the tail batch is captured after the timed region has started only on a handle's
        // first runs; later runs find it cached
        if (capture) {
            MDS_HIP(hipGraphLaunch(batch == 0 ? head_exec : captured(g->bfs_graph, kBfsBatch), s));
        } else {
            for (uint32_t k = 0; k < kBfsBatch; ++k) bfs_enqueue_slot(g, a);
            MDS_LAUNCHED();
        }
        MDS_HIP(hipEventRecord(g->ev_end, s));
        if (out_dev) MDS_HIP(hipMemcpyAsync(out, B.level, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
        MDS_HIP(hipMemcpyAsync(g->h_ctr_raw.p, B.ctl, sizeof(BfsCtl), hipMemcpyDeviceToHost, s));
        MDS_HIP(hipStreamSynchronize(s));
        std::memcpy(&h, g->h_ctr_raw.p, sizeof(BfsCtl));
        if (h.done) break;
        MDS_CHECK(batch < (2ull * n + 2) / kBfsBatch + 2, MEDUSA_ERR_INTERNAL, "bfs: superstep loop did not end");
    }
    if (capture) g->bfs_last_steps = h.slots;
    float ms = 0;
    MDS_HIP(hipEventElapsedTime(&ms, g->ev_begin, g->ev_end));
    st.device_ms = ms;
    st.supersteps = h.supersteps;
    st.push_steps = h.push_steps;
    st.pull_steps = h.pull_steps;
    st.edges_traversed = h.reached_edges;
    st.reached = (uint32_t)h.reached;
    st.kernel_bytes = h.kernel_bytes;
    st.alg_bytes = h.alg_bytes + bfs_seed_bytes(n, nwords, nsrc);
    if (out != B.level && !out_dev) copy_out(g, out, B.level, (size_t)n * 4);
}

// Sums of the counter slots, read back to the host (waits for the stream).
struct BfsSums {
    uint32_t n_new = 0, n_hub = 0;
    unsigned long long deg_sum = 0, deg_large = 0, probes = 0;
};
static BfsSums read_bfs_counters(medusa_graph* g, const BfsCounters* ctr, uint32_t par_next) {
    g->h_ctr_raw.ensure((sizeof(BfsCounters) + 7) / 8);
    MDS_HIP(hipMemcpyAsync(g->h_ctr_raw.p, ctr, sizeof(BfsCounters), hipMemcpyDeviceToHost, g->stream));
    MDS_HIP(hipStreamSynchronize(g->stream));
    BfsCounters h;
    std::memcpy(&h, g->h_ctr_raw.p, sizeof(BfsCounters));
    unsigned long long joined = 0, work = 0;
    for (uint32_t i = 0; i < kBfsSlots; ++i) {
        joined += h.seg[par_next][i].v;
        work += h.work[i].v;
    }
    BfsSums r;
    r.n_new = (uint32_t)joined;
    r.deg_sum = joined >> 32;
    r.deg_large = (uint32_t)work;
    r.probes = work >> 32;
    r.n_hub = h.n_hub;
    return r;
}

// Host-driven run: the same kernels, with the host deciding each superstep from the counters
// read back after it (timing and tracing); a superstep's time covers its expansion and, after
// a push, the compaction.
static void bfs_run_host(medusa_graph* g, uint32_t nsrc, uint32_t single, int32_t* out, const medusa_opts& o,
                         medusa_stats& st) {
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    const uint32_t nwords = (n + 63) / 64;
    const uint32_t large = bfs_large(g, o), grid = bfs_grid(g, o);
    hipStream_t s = g->stream;
    const bool timing = o.timing != 0;
    const BfsBufs B = handle_bufs(g);
    const BfsArgs a = bfs_args(g, B, large, grid, nullptr);
    g->trace.clear();

    MDS_HIP(hipEventRecord(g->ev_begin, s));
    bfs_seed(g, B, a, nsrc, single, false, 0);
    BfsSums c = read_bfs_counters(g, B.ctr, 0);

    Frontier f{};
    f.n = n;
    f.m = m;
    f.n_f = c.n_new;
    f.m_f = c.deg_sum;
    f.n_u = n - f.n_f;
    f.m_u = m - f.m_f;
    uint64_t reached_edges = f.m_f, reached = f.n_f, bytes = bfs_seed_bytes(n, nwords, nsrc);
    bool pull = false;
    Sched sched(o, Sched::kBfs);
    if (timing) ensure_events(g, 2);
    for (uint32_t t = 0;; ++t) {
        f.step = t;
        if (o.direction == MEDUSA_DIR_PUSH) pull = false;
        else if (o.direction == MEDUSA_DIR_PULL) pull = true;
        else pull = sched.bfs_pull(f, pull);
        const uint32_t n_small = c.n_new - c.n_hub;
        const uint32_t hub = bfs_hub_wgs(grid, n_small, c.n_hub, c.deg_large, c.deg_sum);
        const uint32_t small = bfs_small(pull, f.n_f, f.m_f, a.exact_edges);
        hipLaunchKernelGGL(k_bfs_set_ctl, dim3(1), dim3(1), 0, s, B.ctl, (int32_t)t, pull ? 1u : 0u, small, n_small,
                           c.n_hub, hub);
        MDS_LAUNCHED();
        hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(kWave), 0, s, a.ctr, (t & 1u) ^ 1u);
        MDS_LAUNCHED();
        if (timing) MDS_HIP(hipEventRecord(g->events[0], s));
        hipLaunchKernelGGL(k_bfs_super<false>, dim3(grid), dim3(256), 0, s, a);
        MDS_LAUNCHED();
        if (!pull && !small) {  // the compaction (the push set `pending`)
            hipLaunchKernelGGL(k_bfs_super<false>, dim3(grid), dim3(256), 0, s, a);
            MDS_LAUNCHED();
        }
        if (timing) MDS_HIP(hipEventRecord(g->events[1], s));
        c = read_bfs_counters(g, B.ctr, (t & 1u) ^ 1u);
        double kms = 0;
        if (timing) {
            float ms = 0;
            MDS_HIP(hipEventElapsedTime(&ms, g->events[0], g->events[1]));
            kms = ms;
            st.kernel_ms += ms;
            ++st.kernel_launches;
        }
        // the byte model of the device-driven loop's decision
        uint64_t kb;
        if (pull) {
            kb = 8ull * f.n_u + 8ull * c.probes + 32ull * nwords + 16ull * c.n_new;
            ++st.pull_steps;
        } else if (small) {
            kb = 12ull * f.n_f + 8ull * f.m_f + 8ull * nwords + 32ull * c.n_new;
            ++st.push_steps;
        } else {
            kb = 12ull * f.n_f + 8ull * f.m_f + 4ull * n + 24ull * nwords + 16ull * c.n_new;
            ++st.push_steps;
        }
        st.kernel_bytes += kb;
        bytes += kb;
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
    g->q_small.ensure(2 * bfs_queue_cap(n));
    g->q_large.ensure(2 * (size_t)n);
    g->q_aux.ensure(std::max(n, count));
    g->bitmap.ensure(4 * (((size_t)n + 63) / 64));
    if (count > 1)
        MDS_HIP(hipMemcpyAsync(g->q_aux.p, sources, (size_t)count * 4, hipMemcpyHostToDevice, g->stream));
    medusa_stats st{};
    if (o.timing || o.trace) bfs_run_host(g, count, sources[0], levels, o, st);
    else bfs_run_device(g, handle_bufs(g), count, sources[0], levels, o, st);
    if (stats) *stats = st;
}

// ---- BFS with no device allocation ------------------------------------------------
// Work buffers of medusa_bfs_noalloc: two frontier queues and two hub queues of
// kNoAllocMaxVertices entries each, three frontier bitmaps and the visited bitmap, the control
// block and the counters, in one array per GPU that lives as long as the process.  Host code
// must not pass `scratch` itself to a kernel (in host code the name is a stand-in, not a device
// address): the device address is looked up once per device with hipGetSymbolAddress and kept
// in s_scratch.
constexpr uint32_t kNoAllocMaxVertices = MEDUSA_NOALLOC_MAX_VERTICES;
constexpr size_t kNoAllocQueueBytes = (size_t)kNoAllocMaxVertices * 4;
constexpr size_t kNoAllocBitmapBytes = (size_t)kNoAllocMaxVertices / 8;
static_assert((size_t)kBfsSlots * bfs_span(kNoAllocMaxVertices) == kNoAllocMaxVertices,
              "the segmented queue of the largest graph fills kNoAllocMaxVertices entries exactly");
constexpr size_t kNoAllocBytes = 4 * kNoAllocQueueBytes + 4 * kNoAllocBitmapBytes + kBfsCtlUnits * sizeof(CtlStorage);
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
    // a graph of n <= kNoAllocMaxVertices vertices needs at most kNoAllocMaxVertices entries per queue
    MDS_CHECK(bfs_queue_cap(n) <= kNoAllocMaxVertices, MEDUSA_ERR_INTERNAL, "bfs_noalloc: queue layout");
    for (int i = 0; i < 2; ++i) {
        B.q[i] = reinterpret_cast<uint32_t*>(base + i * kNoAllocQueueBytes);
        B.hub[i] = reinterpret_cast<uint32_t*>(base + (2 + i) * kNoAllocQueueBytes);
    }
    for (int i = 0; i < 3; ++i)
        B.bm[i] = reinterpret_cast<unsigned long long*>(base + 4 * kNoAllocQueueBytes + i * kNoAllocBitmapBytes);
    B.visited = reinterpret_cast<unsigned long long*>(base + 4 * kNoAllocQueueBytes + 3 * kNoAllocBitmapBytes);
    CtlStorage* ctl = reinterpret_cast<CtlStorage*>(base + 4 * kNoAllocQueueBytes + 4 * kNoAllocBitmapBytes);
    B.ctl = reinterpret_cast<BfsCtl*>(ctl);
    B.ctr = reinterpret_cast<BfsCounters*>(ctl + 1);
    B.sources = nullptr;  // one source: not used
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
