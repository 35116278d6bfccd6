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
//
// Near-queue entries are (vertex, slice) pairs: a vertex enters with one entry per
// kSlice of its out-edges, so a hub's relaxation spreads over as many workgroups as it
// has slices.  A workgroup relaxes up to 256 entries at a time (fewer when the queue is short,
// so that every workgroup gets some) through a block-wide prefix scan of their degrees, and
// stages its queue appends in LDS (one atomic per flush).
//
// While the near queue is short (a thin frontier), a workgroup keeps the vertices it lowered
// below the threshold in LDS and relaxes them itself in further rounds of the same launch,
// so the frontier advances several hops per superstep.  This stays exact: a vertex kept
// here is relaxed at the distance this workgroup wrote into it (if its atomicMin found a
// lower one, that value's writer relaxes the vertex), and any later, lower offer from
// another workgroup puts it in the near queue by stamp as usual.

#include <algorithm>
#include <chrono>

#include "engine.h"
#include "sched.h"

namespace medusa {

struct SsspCounters {
    MDS_COUNTER(uint32_t, n_near);              // entries in the next near queue
    MDS_COUNTER(uint32_t, n_far);               // entries in the far queue
    MDS_COUNTER(uint32_t, far_min);             // smallest distance among far entries kept by a split
    MDS_COUNTER(unsigned long long, near_deg);  // out-edges of the next near queue
    MDS_COUNTER(unsigned long long, relax);     // edges relaxed this superstep
};

constexpr uint32_t kSlice = 128;       // out-edges per near-queue entry
constexpr uint32_t kStage = 2048;      // LDS staging entries per queue per workgroup
constexpr uint32_t kStageSlices = 4;   // a vertex with more slices goes straight to HBM

__host__ __device__ inline uint32_t slices_of(uint32_t deg) { return deg <= kSlice ? 1u : (deg + kSlice - 1) / kSlice; }
// Entries one slot may append to a near queue.  A vertex enters once per kSlice of its
// out-edges and at most twice per slot: a split slot stamps the vertices it moves with the
// slot's epoch and the vertices its local rounds lower with the next one, so a vertex moved
// by the split and lowered again by a local round is appended under both.
__host__ inline size_t near_capacity(uint32_t n, uint64_t m) { return 2 * ((size_t)n + m / kSlice) + 64; }

__global__ void k_sssp_init(uint32_t n, uint32_t* __restrict__ dist, uint32_t* __restrict__ stamp, uint32_t src,
                            uint32_t src_slices, unsigned long long* __restrict__ near_q) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        dist[v] = v == src ? 0u : 0xFFFFFFFFu;
        stamp[v] = 0;
    }
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < src_slices; k += gridDim.x * blockDim.x)
        near_q[k] = ((unsigned long long)k << 32) | src;
}

__global__ void k_sssp_zero(SsspCounters* c, uint32_t keep_far) {
    c->n_near = 0;
    if (!keep_far) c->n_far = 0;
    c->far_min = 0xFFFFFFFFu;
    c->near_deg = 0;
    c->relax = 0;
}

// ---- workgroup-staged queue appends ----------------------------------------------------
// Appends from a workgroup collect in LDS and go to HBM in bulk, one atomic per queue per
// flush, instead of one atomic per wave per 64 edges on a single counter (which serialises
// on that address).

struct Stage {
    unsigned long long near[kStage];
    uint32_t far[kStage];
    uint32_t n_near, n_far, base_near, base_far, flush;
};

// Where a kernel's appends go.
struct Sinks {
    unsigned long long* near;
    uint32_t* near_cnt;
    uint32_t near_cap;
    uint32_t* far;
    uint32_t* far_cnt;
    uint32_t far_cap;
    const uint32_t* row_ptr;
};

__device__ __forceinline__ void stage_init(Stage& st) {
    if (threadIdx.x == 0) {
        st.n_near = 0;
        st.n_far = 0;
    }
    __syncthreads();
}

// One thread's appends: v (out-degree deg) to the near queue (as its slices) and/or to the
// far queue.
__device__ __forceinline__ void stage_put(Stage& st, const Sinks& k, bool to_near, bool to_far, uint32_t v,
                                          uint32_t deg, unsigned long long& ndeg) {
    if (to_near) {
        ndeg += deg;
        const uint32_t cnt = slices_of(deg);
        if (cnt > kStageSlices) {
            const uint32_t base = atomicAdd(k.near_cnt, cnt);
            for (uint32_t j = 0; j < cnt; ++j)
                if (base + j < k.near_cap) k.near[base + j] = ((unsigned long long)j << 32) | v;
        } else {
            const uint32_t pos = atomicAdd(&st.n_near, cnt);
            for (uint32_t j = 0; j < cnt; ++j) st.near[pos + j] = ((unsigned long long)j << 32) | v;
        }
    }
    if (to_far) st.far[atomicAdd(&st.n_far, 1u)] = v;
}

// Called by every thread of the block after each round of stage_put (at most one call per
// thread per round): copies the staged entries to HBM when another round could overflow
// the stage, or whenever anything is staged if `last`.
__device__ __forceinline__ void stage_flush(Stage& st, const Sinks& k, bool last) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t nn = st.n_near, nf = st.n_far;
        const bool go = last ? (nn | nf) != 0 : (nn > kStage - 256 * kStageSlices || nf > kStage - 256);
        st.flush = go ? 1u : 0u;
        if (go) {
            st.base_near = nn ? atomicAdd(k.near_cnt, nn) : 0u;
            st.base_far = nf ? atomicAdd(k.far_cnt, nf) : 0u;
        }
    }
    __syncthreads();
    if (!st.flush) return;  // uniform: written before the barrier above, rewritten only after the next one
    const uint32_t nn = st.n_near, nf = st.n_far;
    for (uint32_t i = threadIdx.x; i < nn; i += blockDim.x)
        if (st.base_near + i < k.near_cap) k.near[st.base_near + i] = st.near[i];
    for (uint32_t i = threadIdx.x; i < nf; i += blockDim.x)
        if (st.base_far + i < k.far_cap) k.far[st.base_far + i] = st.far[i];
    __syncthreads();
    if (threadIdx.x == 0) {
        st.n_near = 0;
        st.n_far = 0;
    }
    __syncthreads();
}

// Exclusive prefix sum over the 256 threads of the block; `total` = the block's sum.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t x, uint32_t* s_wave, uint32_t& total) {
    const uint32_t lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const uint32_t incl = wave_incl_scan(x);
    if (lane == kWave - 1) s_wave[wv] = incl;
    __syncthreads();
    uint32_t off = 0;
    for (uint32_t w = 0; w < wv; ++w) off += s_wave[w];
    total = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
    return off + incl - x;
}

// Vertices a workgroup lowered below the threshold and relaxes again itself, in the same
// launch, instead of handing them to the next superstep (two lists, one per round).  An
// entry carries what the next round needs (edge range and the distance this workgroup
// wrote), so a round starts without a global read.
constexpr uint32_t kLocal = 512;
constexpr uint32_t kLocalQueue = 32768;  // near entries up to which the local rounds run
constexpr uint32_t kLocalRounds = 32;     // default rounds on graphs of mean out-degree <= kLocalMeanDeg
constexpr uint32_t kLocalMeanDeg = 4;

struct LocalList {
    uint32_t beg[2][kLocal];  // row_ptr[v]
    uint32_t du[2][kLocal];   // the distance this workgroup offered v
    uint8_t deg[2][kLocal];   // out-degree, at most kSlice
    uint32_t n[2];
};

// Relaxes one tile of at most 256 entries with the whole block: thread t brings the edge
// range [beg, beg + deg) of its entry and the entry's distance du.  The block shares the
// out-edges through a prefix scan of the degrees and every thread binary-searches the
// owner of its edge.  A lowered vertex below the threshold goes to side `side` of the local
// list when one is given (side < 2, one slice, room left), else to the near queue once per
// superstep (by stamp); one at or above the threshold goes to the far queue.  Returns (in
// every thread) the edges relaxed and lowers `fmin` to the smallest distance it appended to
// the far queue; ends with a barrier.  kSpec: read the target's row
// offsets together with its distance whenever the offer is below the threshold, so a
// lowered vertex costs one round trip fewer (worth it on sparse graphs, where rounds are
// latency-bound; on dense ones most offers lose and the extra reads cost bandwidth).
template <bool kSpec>
__device__ __forceinline__ uint32_t relax_tile(uint32_t beg, uint32_t deg, uint32_t du,
                                               const uint32_t* __restrict__ col, const uint32_t* __restrict__ w,
                                               uint32_t* dist, uint32_t* __restrict__ stamp, uint32_t epoch,
                                               uint32_t thr, Stage& st, const Sinks& k, unsigned long long& ndeg,
                                               LocalList* loc, uint32_t side, uint32_t& fmin) {
    __shared__ uint32_t s_beg[256], s_off[256], s_du[256], s_wave[4];
    const uint32_t t = threadIdx.x;
    uint32_t total;
    const uint32_t excl = block_excl_scan(deg, s_wave, total);
    s_beg[t] = beg;
    s_off[t] = excl;
    s_du[t] = du;
    __syncthreads();
    for (uint32_t e0 = 0; e0 < total; e0 += 256) {  // uniform per block
        const uint32_t e = e0 + t;
        bool to_near = false, to_far = false;
        uint32_t v = 0, vdeg = 0;
        if (e < total) {
            uint32_t lo = 0;
#pragma unroll
            for (uint32_t step = 128; step; step >>= 1)
                if (s_off[lo + step] <= e) lo += step;
            const uint32_t x = s_beg[lo] + (e - s_off[lo]);
            v = col[x];
            const unsigned long long nd64 = (unsigned long long)s_du[lo] + w[x];
            uint32_t vb = 0, ve = 0;
            if constexpr (kSpec) {
                if (nd64 < thr) {
                    vb = k.row_ptr[v];
                    ve = k.row_ptr[v + 1];
                }
            }
            if (nd64 < dist[v]) {
                // dist[v] falls this superstep (to nd or below), so v needs relaxing at its
                // new distance whoever wins: the min needs no return value, and the stamp
                // keeps one near entry per vertex
                const uint32_t nd = (uint32_t)nd64;
                atomicMin(&dist[v], nd);
                if (nd < thr) {
                    if constexpr (!kSpec) {
                        vb = k.row_ptr[v];
                        ve = k.row_ptr[v + 1];
                    }
                    vdeg = ve - vb;
                    // kept here: this block relaxes v at nd, the value it wrote; if the
                    // atomicMin found a lower one, that one's writer relaxes v, and a later,
                    // lower offer from another block stamps v itself
                    bool kept = false;
                    if (loc && side < 2 && vdeg <= kSlice) {
                        if (vdeg == 0) {
                            kept = true;
                        } else {
                            const uint32_t pos = atomicAdd(&loc->n[side], 1u);
                            if (pos < kLocal) {
                                loc->beg[side][pos] = vb;
                                loc->du[side][pos] = nd;
                                loc->deg[side][pos] = (uint8_t)vdeg;
                                kept = true;
                            }
                        }
                    }
                    if (!kept) to_near = atomicMax(&stamp[v], epoch) < epoch;
                } else {
                    to_far = true;
                    fmin = min(fmin, nd);
                }
            }
        }
        stage_put(st, k, to_near, to_far, v, vdeg, ndeg);
        stage_flush(st, k, false);
    }
    __syncthreads();  // s_beg / s_off / s_du / s_wave and the local lists are reused by the next tile
    return total;
}

// Up to `rounds` tiles over the vertices the block kept in side 0 of its local list (the
// last round hands what it lowers to the near queue).  Returns (in every thread) the edges
// relaxed.  The list's side 0 must be complete: the caller's last append precedes a barrier.
template <bool kSpec>
__device__ __forceinline__ unsigned long long local_rounds(LocalList* loc, uint32_t rounds,
                                                           const uint32_t* __restrict__ col,
                                                           const uint32_t* __restrict__ w, uint32_t* dist,
                                                           uint32_t* __restrict__ stamp, uint32_t epoch, uint32_t thr,
                                                           Stage& st, const Sinks& k, unsigned long long& ndeg,
                                                           uint32_t& fmin) {
    const uint32_t t = threadIdx.x;
    unsigned long long relax = 0;
    for (uint32_t r = 0; r < rounds; ++r) {
        // list r & 1 is complete: the last tile ended with a barrier
        const uint32_t cur = r & 1, nxt = cur ^ 1;
        const uint32_t cnt = min(loc->n[cur], kLocal);
        if (cnt == 0) break;  // uniform: every thread reads the same LDS word
        if (t == 0) loc->n[nxt] = 0;
        __syncthreads();
        const bool more = r + 1 < rounds;
        for (uint32_t i0 = 0; i0 < cnt; i0 += 256) {  // uniform per block
            const uint32_t i = i0 + t;
            uint32_t beg = 0, deg = 0, du = 0;
            if (i < cnt) {
                beg = loc->beg[cur][i];
                deg = loc->deg[cur][i];
                du = loc->du[cur][i];
            }
            relax += relax_tile<kSpec>(beg, deg, du, col, w, dist, stamp, epoch, thr, st, k, ndeg, loc,
                                       more ? nxt : 2u, fmin);
        }
    }
    return relax;
}

// Entries per workgroup tile: 256, or fewer when the queue is short, so that every workgroup
// of the grid gets a tile (a queue of 1000 entries, hub slices of 128 out-edges among them,
// would otherwise give its edges to 4 workgroups).  With local rounds each workgroup then
// continues from the vertices its own tile lowered, so a short queue also spreads those
// rounds over the whole grid.
__device__ __forceinline__ uint32_t relax_tile_entries(uint32_t qn) {
    const uint32_t per = (uint32_t)(((uint64_t)qn + gridDim.x - 1) / gridDim.x);
    return per < 1u ? 1u : (per > 256u ? 256u : per);
}

// Relaxes near entries q[0, qn) with the whole block: workgroup b takes the T entries of
// chunks b, b + grid, ... (T = relax_tile_entries).  With a local list and rounds > 0, the
// block then runs up to `rounds` further tiles over the vertices it lowered itself.  Returns
// (in every thread) the edges relaxed by this block.  kSpec as for relax_tile.
template <bool kSpec>
__device__ __forceinline__ unsigned long long relax_block(const unsigned long long* __restrict__ q, uint32_t qn,
                                                          const uint32_t* __restrict__ col,
                                                          const uint32_t* __restrict__ w, uint32_t* dist,
                                                          uint32_t* __restrict__ stamp, uint32_t epoch, uint32_t thr,
                                                          Stage& st, const Sinks& k, unsigned long long& ndeg,
                                                          LocalList* loc = nullptr, uint32_t rounds = 0) {
    const uint32_t t = threadIdx.x;
    if (!loc) rounds = 0;
    unsigned long long relax = 0;
    uint32_t fmin = 0xFFFFFFFFu;  // not needed after a relaxation superstep
    const uint32_t T = relax_tile_entries(qn);
    for (uint32_t chunk = blockIdx.x; (uint64_t)chunk * T < qn; chunk += gridDim.x) {  // uniform per block
        const uint64_t i = (uint64_t)chunk * T + t;
        uint32_t beg = 0, deg = 0, du = 0;
        if (t < T && i < qn) {
            const unsigned long long ent = q[i];
            const uint32_t u = (uint32_t)ent, slice = (uint32_t)(ent >> 32);
            const uint32_t b = k.row_ptr[u], e = k.row_ptr[u + 1];
            beg = b + min(e - b, slice * kSlice);
            deg = min(e - beg, kSlice);
            du = dist[u];
        }
        relax += relax_tile<kSpec>(beg, deg, du, col, w, dist, stamp, epoch, thr, st, k, ndeg, loc,
                                   rounds ? 0u : 2u, fmin);
    }
    if (rounds) relax += local_rounds<kSpec>(loc, rounds, col, w, dist, stamp, epoch, thr, st, k, ndeg, fmin);
    return relax;
}

// Splits far entries (or, when `rescan`, every vertex) at a new threshold: distances in
// [old_thr, thr) move to the near queue (once, by stamp), distances >= thr stay far
// (deduplicated by stamp); below old_thr they were settled through the near queue.
// Returns the smallest kept far distance of this thread.  With a local list, a vertex moved
// to the near side (one slice) goes to side 0 of the list instead, for the block to relax in
// the same launch; the stamp still marks it moved, so duplicates in the far queue move once.
__device__ __forceinline__ uint32_t split_block(const uint32_t* __restrict__ far_in, uint32_t nin, bool rescan,
                                                const uint32_t* __restrict__ dist, uint32_t* __restrict__ stamp,
                                                uint32_t epoch, uint32_t old_thr, uint32_t thr, Stage& st,
                                                const Sinks& k, unsigned long long& ndeg, LocalList* loc = nullptr) {
    uint32_t fmin = 0xFFFFFFFFu;
    for (uint32_t i0 = blockIdx.x * 256; i0 < nin; i0 += gridDim.x * 256) {  // uniform per block
        const uint32_t i = i0 + threadIdx.x;
        bool to_near = false, keep = false;
        uint32_t v = 0, vdeg = 0;
        if (i < nin) {
            v = rescan ? i : far_in[i];
            const uint32_t d = dist[v];
            if (d != 0xFFFFFFFFu && d >= old_thr) {
                if (d < thr) {
                    const uint32_t vb = k.row_ptr[v];
                    vdeg = k.row_ptr[v + 1] - vb;
                    to_near = atomicMax(&stamp[v], epoch) < epoch;
                    if (to_near && loc && vdeg <= kSlice) {
                        if (vdeg == 0) {
                            to_near = false;
                        } else {
                            const uint32_t pos = atomicAdd(&loc->n[0], 1u);
                            if (pos < kLocal) {
                                loc->beg[0][pos] = vb;
                                loc->du[0][pos] = d;
                                loc->deg[0][pos] = (uint8_t)vdeg;
                                to_near = false;
                            }
                        }
                    }
                } else {
                    keep = atomicMax(&stamp[v], epoch) < epoch;
                    if (keep) fmin = min(fmin, d);
                }
            }
        }
        stage_put(st, k, to_near, keep, v, vdeg, ndeg);
        stage_flush(st, k, false);
    }
    return fmin;
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t x) {
#pragma unroll
    for (int off = kWave / 2; off >= 1; off >>= 1) x = min(x, (uint32_t)__shfl_xor(x, off, kWave));
    return x;
}

// One relaxation superstep over the near queue (host-driven loop); kLoc: with the local list
// (its LDS is allocated only in that variant).
template <bool kLoc>
__global__ __launch_bounds__(256) void k_sssp_relax(const unsigned long long* __restrict__ q, uint32_t qn,
                                                    const uint32_t* __restrict__ row_ptr,
                                                    const uint32_t* __restrict__ col, const uint32_t* __restrict__ w,
                                                    uint32_t* dist, uint32_t* __restrict__ stamp, uint32_t epoch,
                                                    uint32_t thr, unsigned long long* __restrict__ near_next,
                                                    uint32_t near_cap, uint32_t* __restrict__ far_q, uint32_t far_cap,
                                                    SsspCounters* __restrict__ ctr, uint32_t rounds) {
    __shared__ Stage st;
    LocalList* loc = nullptr;
    if constexpr (kLoc) {
        __shared__ LocalList s_loc;
        if (threadIdx.x == 0) s_loc.n[0] = 0;
        loc = &s_loc;
    }
    stage_init(st);  // its barrier publishes s_loc.n[0]
    const Sinks k{near_next, &ctr->n_near, near_cap, far_q, &ctr->n_far, far_cap, row_ptr};
    unsigned long long ndeg = 0;
    const unsigned long long relax =
        relax_block<kLoc>(q, qn, col, w, dist, stamp, epoch, thr, st, k, ndeg, loc, kLoc ? rounds : 0u);
    stage_flush(st, k, true);
    ndeg = wave_sum_u64(ndeg);
    if ((threadIdx.x & (kWave - 1)) == 0 && ndeg) atomicAdd(&ctr->near_deg, ndeg);
    if (threadIdx.x == 0 && relax) atomicAdd(&ctr->relax, relax);
}

// Split of the far queue at a new threshold (host-driven loop).
__global__ __launch_bounds__(256) void k_sssp_split(const uint32_t* __restrict__ far_in, uint32_t nfar,
                                                    uint32_t rescan, const uint32_t* __restrict__ dist,
                                                    uint32_t* __restrict__ stamp,
                                                    uint32_t epoch, uint32_t old_thr, uint32_t thr,
                                                    const uint32_t* __restrict__ row_ptr,
                                                    unsigned long long* __restrict__ near_q, uint32_t near_cap,
                                                    uint32_t* __restrict__ far_out, uint32_t far_cap,
                                                    SsspCounters* __restrict__ ctr) {
    __shared__ Stage st;
    stage_init(st);
    const Sinks k{near_q, &ctr->n_near, near_cap, far_out, &ctr->n_far, far_cap, row_ptr};
    unsigned long long ndeg = 0;
    uint32_t fmin = split_block(far_in, nfar, rescan != 0, dist, stamp, epoch, old_thr, thr, st, k, ndeg);
    stage_flush(st, k, true);
    ndeg = wave_sum_u64(ndeg);
    fmin = wave_min_u32(fmin);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        if (ndeg) atomicAdd(&ctr->near_deg, ndeg);
        if (fmin != 0xFFFFFFFFu) atomicMin(&ctr->far_min, fmin);
    }
}

// ---- device-driven loop ----------------------------------------------------------------
// Each slot of the captured loop is one work kernel:
//   work:   if a split is pending, move far entries below the new threshold to the near queue
//           (a full rescan of the distances if the far queue overflowed); otherwise relax the
//           current near queue, and, while that queue is short (at most kLocalQueue entries),
//           let each workgroup relax up to `local_cfg` more rounds of the vertices it lowered
//           itself, so a thin frontier (a road-like graph) advances several hops per superstep;
//   decide: the last workgroup to finish (by ticket) folds the counters into the control
//           block; after a relaxation that left the near queue empty the phase ends: it
//           raises the threshold by the run's bucket width (default, fixed, or chosen by the
//           learned policy before the run) and marks a split pending; after a split, an empty
//           bucket raises the threshold past the smallest far distance.
// kSsspBatch slots are captured once as a hipGraph and replayed until done.

constexpr uint32_t kSsspBatch = 32;
// Workgroups of a work launch: about the resident count (5 per CU at this LDS size) without local
// rounds, so every tile is in flight at once (RMAT-20 at width 250: 1.51 ms against 1.79 with
// 2048; ER 2^20/2^24: 1.15 against 1.36); twice that with local rounds, whose short queues spread
// better over more workgroups (2.4 M-vertex Chung-Lu: 0.74 against 0.77 ms with 1024).
constexpr uint32_t kSsspGrid = 1024;
constexpr uint32_t kSsspGridLocal = 2048;

struct SsspCtl {
    uint32_t done, epoch, thr, old_thr, delta, n_near, n_far, split_pending;
    uint32_t overflow, phase, supersteps, local_rounds, width, near_sel;
    uint32_t far_sel, local_cfg, trace, lost;  // lost: near entries beyond the queue
    unsigned long long n, m, relaxed, kernel_bytes;
};

// Phase marks of the device-driven loop when tracing: mark i (3 words) holds the features at
// the advance that starts phase i (phase index, far-queue length, threshold before the
// advance), the width, and the wall clock at the advance.  The mark
// after the last phase holds only the end time.  The host turns consecutive marks into
// phase durations, so the delta policy learns from the loop it steers.
constexpr uint32_t kMaxPhaseMarks = 1u << 16;

__device__ __forceinline__ void phase_mark(unsigned long long* marks, uint32_t i, uint32_t phase, uint32_t n_far,
                                           uint32_t thr, uint32_t delta) {
    if (i >= kMaxPhaseMarks) return;
    marks[3 * i] = ((unsigned long long)n_far << 32) | phase;
    marks[3 * i + 1] = ((unsigned long long)delta << 32) | thr;
    marks[3 * i + 2] = (unsigned long long)wall_clock64();
}

// Tickets of the last-workgroup election: workgroup b takes a ticket on line b mod kSsspSlots,
// and the last of a line's workgroups one on `top`, so no address sees more than
// grid / kSsspSlots atomics (same-address device atomics serialise).
constexpr uint32_t kSsspSlots = 32;
struct SsspTicket {
    alignas(128) uint32_t v;
};
struct SsspDevCounters {
    MDS_COUNTER(uint32_t, n_near);   // near entries written by this slot's work kernel
    MDS_COUNTER(uint32_t, n_far);    // far queue length (relax appends after the current entries)
    MDS_COUNTER(uint32_t, far_out);  // far entries kept by a split
    MDS_COUNTER(uint32_t, far_min);  // smallest distance kept by a split
    MDS_COUNTER(unsigned long long, relax);
    SsspTicket sub[kSsspSlots];
    MDS_COUNTER(uint32_t, top);
};
// storage units of the device-driven loop: the control block, then the counters
constexpr size_t kSsspCtlUnits = 1 + (sizeof(SsspDevCounters) + sizeof(CtlStorage) - 1) / sizeof(CtlStorage);

__device__ __forceinline__ uint32_t sat_add(uint32_t a, uint32_t b) {
    const unsigned long long s = (unsigned long long)a + b;
    return s >= 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)s;
}

// The slot's decision, by one thread of the work kernel's last workgroup (every other
// workgroup's counter atomics were performed before its ticket): folds the counters into the
// control block and clears them and the tickets for the next slot.  The next work kernel reads
// both after the kernel boundary.
__device__ void sssp_decide(SsspCtl* ctl, SsspDevCounters* ctr, uint32_t far_cap, uint32_t near_cap,
                            unsigned long long* __restrict__ marks) {
    SsspDevCounters k;
    k.n_near = load_now(&ctr->n_near);
    k.n_far = load_now(&ctr->n_far);
    k.far_out = load_now(&ctr->far_out);
    k.far_min = load_now(&ctr->far_min);
    k.relax = load_now(&ctr->relax);
    SsspCtl c = load_now_block(ctl);
    if (k.n_near > near_cap) {
        // entries past the queue were dropped by the writers (near_capacity rules this out):
        // stop, and let the host report it, rather than read past the queue
        k.n_near = near_cap;
        c.lost = 1;
        c.done = 1;
    }
    uint32_t n_far;
    // phase mark to write (trace on): index, then its fields; ~0u = none
    uint32_t mark = 0xFFFFFFFFu, mark_phase = 0, mark_far = 0, mark_thr = 0, mark_delta = 0;
    if (c.split_pending) {
        // after a split (which may also have relaxed the vertices it moved, under epoch + 1)
        c.far_sel ^= 1u;
        n_far = min(k.far_out, far_cap);
        c.n_far = n_far;
        c.overflow = k.far_out > far_cap ? 1u : 0u;  // kept entries beyond the cap: rescan next time
        c.epoch += 2;
        if (k.relax) {
            c.supersteps += 1;
            c.relaxed += k.relax;
            c.kernel_bytes += 16ull * k.relax + 8ull * k.n_near;
        }
        c.n_near = k.n_near;
        if (k.n_near > 0) {
            c.split_pending = 0;
        } else if (n_far == 0 && !c.overflow) {
            c.done = 1;
            mark = c.phase + 1;  // end time
        } else {
            // the bucket is finished: either this slot relaxed the vertices it moved and
            // lowered nothing below the threshold that is left for the near queue (a phase
            // ends), or it moved nothing (an empty bucket).  Either way the threshold restarts
            // from the smallest far distance.
            const uint32_t d = c.width;
            if (k.relax) {
                mark = c.phase + 1;
                mark_phase = c.phase;
                mark_far = n_far;
                mark_thr = c.thr;
                mark_delta = d;
                c.phase += 1;
            }
            c.old_thr = c.thr;
            c.delta = d;
            c.thr = sat_add(k.far_min > c.thr ? k.far_min : c.thr, d);
        }
    } else {
        // after a relaxation (of c.n_near entries)
        const uint32_t N = c.n_near;
        n_far = k.n_far;
        if (N > 0) {
            c.supersteps += 1;
            c.near_sel ^= 1u;
            c.relaxed += k.relax;
            c.kernel_bytes += 16ull * N + 16ull * k.relax + 8ull * k.n_near;
        }
        if (n_far > far_cap) {
            c.overflow = 1;
            n_far = far_cap;
        }
        c.n_near = k.n_near;
        c.n_far = n_far;
        c.epoch += 1;
        if (k.n_near == 0) {
            if (n_far == 0 && !c.overflow) {
                c.done = 1;
                mark = c.phase + 1;  // end time
            } else {
                const uint32_t d = c.width;
                mark = c.phase + 1;
                mark_phase = c.phase;
                mark_far = n_far;
                mark_thr = c.thr;
                mark_delta = d;
                c.old_thr = c.thr;
                c.delta = d;
                c.thr = sat_add(c.thr, d);
                c.split_pending = 1;
                c.phase += 1;
            }
        }
    }
    c.local_rounds = c.n_near <= kLocalQueue ? c.local_cfg : 0u;
    if (c.trace && mark != 0xFFFFFFFFu) phase_mark(marks, mark, mark_phase, mark_far, mark_thr, mark_delta);
    *ctl = c;
    ctr->n_near = 0;
    ctr->n_far = n_far;
    ctr->far_out = 0;
    ctr->far_min = 0xFFFFFFFFu;
    ctr->relax = 0;
    for (uint32_t i = 0; i < kSsspSlots; ++i) ctr->sub[i].v = 0;
    ctr->top = 0;
}

// End of a work kernel's workgroup: thread 0 takes the tickets after every counter atomic of
// the block has returned (the callers wait for their results before the barrier here, and
// `dep` carries thread 0's last one), so the workgroup that takes the last ticket sees every
// other workgroup's counts.  Returns true in every thread of that workgroup.
__device__ __forceinline__ bool sssp_last_block(SsspDevCounters* __restrict__ ctr, unsigned long long dep) {
    __shared__ uint32_t s_last;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t inc = 1;
        asm volatile("" : "+v"(inc) : "v"(dep));
        const uint32_t sl = blockIdx.x % kSsspSlots;
        const uint32_t members = (gridDim.x - sl + kSsspSlots - 1) / kSsspSlots;  // workgroups of line sl
        const uint32_t nslots = gridDim.x < kSsspSlots ? gridDim.x : kSsspSlots;
        bool last = false;
        if (atomicAdd(&ctr->sub[sl].v, inc) == members - 1) {
            uint32_t inc2 = 1;
            asm volatile("" : "+v"(inc2) : "v"(members));
            last = atomicAdd(&ctr->top, inc2) == nslots - 1;
        }
        s_last = last ? 1u : 0u;
    }
    __syncthreads();
    return s_last != 0;
}

template <bool kLoc>
__global__ __launch_bounds__(256) void k_sssp_work(SsspCtl* ctl,
                                                   unsigned long long* __restrict__ near0,
                                                   unsigned long long* __restrict__ near1, uint32_t near_cap,
                                                   uint32_t* __restrict__ far0, uint32_t* __restrict__ far1,
                                                   uint32_t far_cap, const uint32_t* __restrict__ row_ptr,
                                                   const uint32_t* __restrict__ col, const uint32_t* __restrict__ w,
                                                   uint32_t* dist, uint32_t* __restrict__ stamp,
                                                   SsspDevCounters* __restrict__ ctr,
                                                   unsigned long long* __restrict__ marks) {
    __shared__ Stage st;
    LocalList* loc = nullptr;
    if constexpr (kLoc) {
        __shared__ LocalList s_loc;
        loc = &s_loc;
    }
    const SsspCtl c = *ctl;
    // every field this kernel reads, loaded in one round trip
    keep_here(c.done);
    keep_here(c.local_rounds);
    keep_here(c.split_pending);
    keep_here(c.n_near);
    keep_here(c.epoch);
    keep_here(c.thr);
    keep_here(c.old_thr);
    keep_here(c.near_sel);
    keep_here(c.far_sel);
    keep_here(c.overflow);
    keep_here(c.n_far);
    keep_here((uint32_t)c.n);
    if (c.done) return;  // uniform: no workgroup takes a ticket
    const uint32_t epoch = c.epoch, thr = c.thr;
    unsigned long long ndeg = 0;
    unsigned long long dep = 0;  // thread 0's last counter atomic result
    if (c.split_pending) {
        const bool rescan = c.overflow != 0;
        const uint32_t nin = rescan ? (uint32_t)c.n : c.n_far;
        const Sinks k{c.near_sel ? near1 : near0, &ctr->n_near, near_cap, c.far_sel ? far0 : far1,
                      &ctr->far_out, far_cap, row_ptr};
        LocalList* sl = kLoc && c.local_rounds ? loc : nullptr;
        if (sl && threadIdx.x == 0) sl->n[0] = 0;
        stage_init(st);  // its barrier publishes sl->n[0]
        uint32_t fmin = split_block(c.far_sel ? far1 : far0, nin, rescan, dist, stamp, epoch, c.old_thr, thr,
                                    st, k, ndeg, sl);
        if (sl) {
            // the moved vertices this block kept are relaxed at once; their stamps use the next
            // epoch, so a vertex this split kept far still enters the near queue if it is lowered
            // below the threshold
            const unsigned long long relax =
                local_rounds<kLoc>(sl, c.local_rounds, col, w, dist, stamp, epoch + 1, thr, st, k, ndeg, fmin);
            if (threadIdx.x == 0 && relax) dep += atomicAdd(&ctr->relax, relax);
        }
        stage_flush(st, k, true);  // thread 0 stored its atomics' results in LDS
        fmin = wave_min_u32(fmin);
        if ((threadIdx.x & (kWave - 1)) == 0 && fmin != 0xFFFFFFFFu) {
            const uint32_t r = atomicMin(&ctr->far_min, fmin);
            asm volatile("" ::"v"(r));  // wait for it before the barrier in sssp_last_block
        }
        if (sssp_last_block(ctr, dep) && threadIdx.x == 0) sssp_decide(ctl, ctr, far_cap, near_cap, marks);
        return;
    }
    const uint32_t N = c.n_near;
    if (N == 0) {  // not reached: a decision leaves either a split pending or entries to relax
        if (sssp_last_block(ctr, dep) && threadIdx.x == 0) sssp_decide(ctl, ctr, far_cap, near_cap, marks);
        return;
    }
    const Sinks k{c.near_sel ? near0 : near1, &ctr->n_near, near_cap, c.far_sel ? far1 : far0, &ctr->n_far,
                  far_cap, row_ptr};
    if (loc && threadIdx.x == 0) loc->n[0] = 0;
    stage_init(st);  // its barrier publishes loc->n[0]
    const unsigned long long relax = relax_block<kLoc>(c.near_sel ? near1 : near0, N, col, w, dist, stamp, epoch,
                                                       thr, st, k, ndeg, loc, kLoc ? c.local_rounds : 0u);
    stage_flush(st, k, true);  // thread 0 stored its atomics' results in LDS
    if (threadIdx.x == 0 && relax) dep += atomicAdd(&ctr->relax, relax);
    if (sssp_last_block(ctr, dep) && threadIdx.x == 0) sssp_decide(ctl, ctr, far_cap, near_cap, marks);
}

__global__ void k_sssp_seed_ctl(SsspCtl* ctl, SsspDevCounters* c, uint32_t n, unsigned long long m, uint32_t thr,
                                uint32_t width, uint32_t src_slices, uint32_t local_cfg, uint32_t trace,
                                unsigned long long* marks) {
    SsspCtl z{};
    z.trace = trace;
    // phase 0: one source vertex, threshold 0 before the first advance
    if (trace) phase_mark(marks, 0, 0, 1, 0, width);
    z.local_cfg = local_cfg;
    z.local_rounds = src_slices <= kLocalQueue ? local_cfg : 0u;
    z.epoch = 1;
    z.thr = thr;
    z.delta = thr;
    z.n_near = src_slices;
    z.width = width;
    z.n = n;
    z.m = m;
    *ctl = z;
    c->n_near = 0;
    c->n_far = 0;
    c->far_out = 0;
    c->far_min = 0xFFFFFFFFu;
    c->relax = 0;
    for (uint32_t i = 0; i < kSsspSlots; ++i) c->sub[i].v = 0;
    c->top = 0;
}

// Local rounds per superstep (while the near queue is short): the caller's count, or by
// default kLocalRounds on sparse graphs (mean out-degree <= kLocalMeanDeg, road-like: thin
// frontiers, thousands of supersteps) and none on denser ones, where the extra rounds cost
// more than the supersteps they save.
static uint32_t sssp_local_cfg(const medusa_graph* g, const medusa_opts& o) {
    if (o.local_rounds == 0xFFFFFFFFu) return 0;
    if (o.local_rounds) return o.local_rounds;
    return g->m <= (uint64_t)kLocalMeanDeg * g->n ? kLocalRounds : 0u;
}

// Far-queue capacity: m + n entries, or the caller's smaller limit (a full queue is rebuilt
// from the distances, so any limit >= 1 gives the same result).
static size_t sssp_far_alloc(const medusa_graph* g) { return (size_t)std::max<uint64_t>(g->m, g->n) + g->n; }
static size_t sssp_far_cap(const medusa_graph* g, const medusa_opts& o) {
    const size_t full = sssp_far_alloc(g);
    return o.far_limit ? std::min<size_t>(o.far_limit, full) : full;
}

// Mean weight (sampled) and the default bucket width, a multiple of it: 8x with local rounds,
// which absorb the extra hops of a wider bucket (1400x1400 lattice, 32 rounds: 8.0 ms at 8x,
// 8.6 ms at 6x, 11.5 ms at 3x), else 2x.
static void sssp_widths(const medusa_graph* g, const medusa_opts& o, uint32_t& wmean, uint32_t& dflt) {
    uint64_t wsum = 0, wcnt = 0;
    const uint64_t m = g->m;
    const uint64_t wstride = std::max<uint64_t>(1, m / 4096);
    for (uint64_t e = 0; e < m; e += wstride, ++wcnt) wsum += g->h_w[e];
    wmean = (uint32_t)std::max<uint64_t>(1, wcnt ? wsum / wcnt : 1);
    const uint64_t mult = sssp_local_cfg(g, o) ? 8 : 2;
    dflt = (uint32_t)std::min<uint64_t>(0xFFFFFFFEull, mult * wmean);
}

// Field 4 of every phase record: the vertices the phase settled, i.e. those whose final
// distance lies between the phase's starting threshold and the next phase's (the last phase
// takes the rest).  Time per settled vertex is what a bucket width should minimise: a narrow
// width pays the per-phase cost over few vertices, a wide one relaxes vertices many times.
static void sssp_trace_settled(medusa_graph* g) {
    if (g->trace.empty()) return;
    std::vector<uint32_t> d(g->n);
    MDS_HIP(hipMemcpyAsync(d.data(), g->dist.p, (size_t)g->n * 4, hipMemcpyDeviceToHost, g->stream));
    MDS_HIP(hipStreamSynchronize(g->stream));
    std::vector<double> lo(g->trace.size());
    for (size_t i = 0; i < lo.size(); ++i) lo[i] = g->trace[i].f[3];
    std::vector<uint64_t> cnt(lo.size(), 0);
    for (uint32_t x : d) {
        if (x == 0xFFFFFFFFu) continue;
        const size_t k = (size_t)(std::upper_bound(lo.begin(), lo.end(), (double)x) - lo.begin());
        if (k) ++cnt[k - 1];
    }
    for (size_t i = 0; i < lo.size(); ++i) g->trace[i].f[4] = (double)cnt[i];
}

// Trace records of a device-driven run from its phase marks: one per phase, with the
// features and width of the mark and the time to the next mark.
static void sssp_device_trace(medusa_graph* g, uint32_t phases, uint32_t wmean) {
    const uint32_t nmarks = std::min<uint32_t>(phases + 2, kMaxPhaseMarks);
    std::vector<unsigned long long> h(3ull * nmarks);
    MDS_HIP(hipMemcpyAsync(h.data(), g->sssp_phases.p, h.size() * 8, hipMemcpyDeviceToHost, g->stream));
    MDS_HIP(hipStreamSynchronize(g->stream));
    int khz = 0;
    MDS_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, g->device));
    MDS_CHECK(khz > 0, MEDUSA_ERR_HIP, "sssp: no wall clock rate");
    g->trace.clear();
    for (uint32_t i = 0; i + 1 < nmarks; ++i) {
        Frontier f{};
        f.step = (uint32_t)h[3 * i];
        f.n_f = (uint32_t)(h[3 * i] >> 32);
        f.n = g->n;
        f.m = g->m;
        const uint32_t thr = (uint32_t)h[3 * i + 1], delta = (uint32_t)(h[3 * i + 1] >> 32);
        const double ms = (double)(h[3 * (i + 1) + 2] - h[3 * i + 2]) / (double)khz;
        TraceRec r;
        Sched::to_sssp_record(f, wmean, thr, (double)delta, ms, r.f);
        g->trace.push_back(r);
    }
}

static void sssp_run_device(medusa_graph* g, uint32_t src, uint32_t* out, const medusa_opts& o, medusa_stats& st) {
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    const size_t far_cap = sssp_far_cap(g, o);
    hipStream_t s = g->stream;
    uint32_t wmean, dflt;
    sssp_widths(g, o, wmean, dflt);
    Sched sched(o, Sched::kSssp);
    Frontier f{};
    f.n = n;
    f.m = m;
    f.n_f = 1;
    // one width for the whole run: the caller's, the default, or the learned policy's choice
    // from the features at the first phase
    const uint32_t d0 = sched.sssp_delta(f, o.delta ? o.delta : dflt, wmean, 0);
    const uint32_t thr0 = d0 >= 0xFFFFFFFFu - 1 ? 0xFFFFFFFFu : d0;
    // control block in unit 0 of the storage, counters and tickets after it
    g->sssp_ctl.ensure(kSsspCtlUnits);
    SsspCtl* ctl = reinterpret_cast<SsspCtl*>(g->sssp_ctl.p);
    static_assert(sizeof(SsspCtl) <= sizeof(CtlStorage) - 16, "control block storage (its last 16 bytes: reached counts)");
    SsspDevCounters* ctr = reinterpret_cast<SsspDevCounters*>(g->sssp_ctl.p + 1);
    uint32_t* far0 = g->q_aux.p;
    uint32_t* far1 = g->q_aux.p + sssp_far_alloc(g);
    const size_t near_cap = near_capacity(n, m);
    unsigned long long* near0 = g->near_q.p;
    unsigned long long* near1 = g->near_q.p + near_cap;
    const uint32_t src_slices = slices_of(g->h_row_ptr[src + 1] - g->h_row_ptr[src]);

    GraphCache& gc = g->sssp_graph;
    unsigned long long* marks = g->sssp_phases.p;
    const uint32_t local_cfg = sssp_local_cfg(g, o);
    // the marks buffer is 256-byte aligned, so its low bit can carry the kernel variant
    const uintptr_t key[6] = {(uintptr_t)g->dist.p, (uintptr_t)g->near_q.p, (uintptr_t)g->q_aux.p,
                              (uintptr_t)g->w.p, (uintptr_t)marks | (local_cfg ? 1u : 0u), (uintptr_t)ctl};
    if (!gc.exec || !std::equal(key, key + 6, gc.key)) {
        gc.reset();
        MDS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (uint32_t k = 0; k < kSsspBatch; ++k) {
            hipLaunchKernelGGL(local_cfg ? k_sssp_work<true> : k_sssp_work<false>,
                               dim3(local_cfg ? kSsspGridLocal : kSsspGrid), dim3(256), 0, s,
                               ctl, near0, near1, (uint32_t)near_cap, far0, far1, (uint32_t)far_cap,
                               g->row_ptr.p, g->col.p, g->w.p, g->dist.p, g->stamp.p, ctr, marks);
        }
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(s, &graph);
        if (e == hipSuccess) e = hipGraphInstantiate(&gc.exec, graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        MDS_HIP(e);
        std::copy(key, key + 6, gc.key);
    }

    MDS_HIP(hipEventRecord(g->ev_begin, s));
    hipLaunchKernelGGL(k_sssp_init, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, n, g->dist.p, g->stamp.p, src,
                       src_slices, near0);
    MDS_LAUNCHED();
    hipLaunchKernelGGL(k_sssp_seed_ctl, dim3(1), dim3(1), 0, s, ctl, ctr, n, (unsigned long long)m, thr0, d0,
                       src_slices, local_cfg, o.trace ? 1u : 0u, marks);
    MDS_LAUNCHED();
    SsspCtl h{};
    // every replay either relaxes a non-empty queue or raises the threshold
    const uint64_t max_batches = 4ull * n / kSsspBatch + 1024;
    for (uint64_t batch = 0;; ++batch) {
        MDS_HIP(hipGraphLaunch(gc.exec, s));
        MDS_HIP(hipMemcpyAsync(g->h_ctr_raw.p, ctl, sizeof(SsspCtl), hipMemcpyDeviceToHost, s));
        MDS_HIP(hipStreamSynchronize(s));
        std::memcpy(&h, g->h_ctr_raw.p, sizeof(SsspCtl));
        MDS_CHECK(!h.lost, MEDUSA_ERR_INTERNAL, "sssp: near queue overflow");
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
    if (o.trace) sssp_device_trace(g, h.phase, wmean);
    copy_out(g, out, g->dist.p, (size_t)n * 4);
}

static void sssp_run_host(medusa_graph* g, uint32_t src, uint32_t* out, const medusa_opts& o, medusa_stats& st) {
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    const size_t far_cap = sssp_far_cap(g, o);
    SsspCounters* C = dev_counters<SsspCounters>(g);
    hipStream_t s = g->stream;
    g->trace.clear();
    uint32_t wmean, delta_default;
    sssp_widths(g, o, wmean, delta_default);
    Sched sched(o, Sched::kSssp);

    const size_t near_cap = near_capacity(n, m);
    unsigned long long* near_cur = g->near_q.p;
    unsigned long long* near_nxt = g->near_q.p + near_cap;
    uint32_t* far_a = g->q_aux.p;
    uint32_t* far_b = g->q_aux.p + sssp_far_alloc(g);
    const uint32_t src_slices = slices_of(g->h_row_ptr[src + 1] - g->h_row_ptr[src]);

    MDS_HIP(hipEventRecord(g->ev_begin, s));
    hipLaunchKernelGGL(k_sssp_init, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, n, g->dist.p, g->stamp.p, src,
                       src_slices, near_cur);
    MDS_LAUNCHED();
    hipLaunchKernelGGL(k_sssp_zero, dim3(1), dim3(1), 0, s, C, 0u);
    MDS_LAUNCHED();
    uint32_t n_near = src_slices, n_far = 0, epoch = 1;
    bool far_incomplete = false;  // the far queue hit its cap: the next split rescans all distances
    uint64_t near_deg = g->h_row_ptr[src + 1] - g->h_row_ptr[src];
    uint64_t relaxed = 0;
    const bool timing = o.timing != 0;
    if (timing) ensure_events(g, 2);
    const uint32_t local_cfg = sssp_local_cfg(g, o);
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
    // one width for the whole run (as in the device-driven loop)
    const uint32_t delta = sched.sssp_delta(f, o.delta ? o.delta : delta_default, wmean, 0);
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
            if (n_far == 0 && !far_incomplete) break;
            close_phase();
            // advance the threshold and split the far queue; skip buckets that come out empty
            uint32_t old_thr = thr;
            f.step = phase;
            f.n_f = n_far;
            f.m_f = 0;
            phase_f = f;
            phase_thr = thr;
            phase_t0 = std::chrono::steady_clock::now();
            for (;;) {
                const uint64_t nt = (uint64_t)thr + delta;
                thr = nt >= 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)nt;
                ++epoch;
                hipLaunchKernelGGL(k_sssp_zero, dim3(1), dim3(1), 0, s, C, 0u);
                MDS_LAUNCHED();
                const uint32_t nin = far_incomplete ? n : n_far;
                hipLaunchKernelGGL(k_sssp_split, dim3(grid_for(nin, 256, 4096)), dim3(256), 0, s, far_a, nin,
                                   far_incomplete ? 1u : 0u,
                                   g->dist.p, g->stamp.p, epoch, old_thr, thr, g->row_ptr.p, near_cur,
                                   (uint32_t)near_cap, far_b, (uint32_t)far_cap, C);
                MDS_LAUNCHED();
                const SsspCounters sc = read_counters<SsspCounters>(g);
                std::swap(far_a, far_b);
                n_near = sc.n_near;
                MDS_CHECK(n_near <= near_cap, MEDUSA_ERR_INTERNAL, "sssp: near queue overflow");
                far_incomplete = sc.n_far > far_cap;
                n_far = std::min<uint32_t>(sc.n_far, (uint32_t)far_cap);
                near_deg = sc.near_deg;
                if (n_near || (n_far == 0 && !far_incomplete)) break;
                old_thr = thr;
                if (sc.far_min != 0xFFFFFFFFu && sc.far_min > thr) thr = sc.far_min;
            }
            if (n_near == 0) break;
        }
        f.step = step;
        f.n_f = n_near;
        f.m_f = near_deg;
        ++epoch;
        hipLaunchKernelGGL(k_sssp_zero, dim3(1), dim3(1), 0, s, C, 1u);
        MDS_LAUNCHED();
        if (timing) MDS_HIP(hipEventRecord(g->events[0], s));  // brackets the relaxation launch alone
        // 256 entries per workgroup chunk; the far queue keeps growing at far_a[n_far..)
        const uint32_t rounds = n_near <= kLocalQueue ? local_cfg : 0u;
        hipLaunchKernelGGL(rounds ? k_sssp_relax<true> : k_sssp_relax<false>,
                           dim3(grid_for(n_near, 1, local_cfg ? kSsspGridLocal : kSsspGrid)),
                           dim3(256), 0, s, near_cur, n_near, g->row_ptr.p, g->col.p, g->w.p, g->dist.p, g->stamp.p,
                           epoch, thr, near_nxt, (uint32_t)near_cap, far_a, (uint32_t)far_cap, C, rounds);
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
        // byte model: 16 per frontier entry (entry, offsets, dist), 16 per relaxed edge (col,
        // w, dist probe, atomic), 8 per enqueued entry
        st.kernel_bytes += 16ull * n_near + 16ull * c.relax + 8ull * (c.n_near);
        ++st.supersteps;
        ++st.push_steps;
        std::swap(near_cur, near_nxt);
        n_near = c.n_near;
        MDS_CHECK(n_near <= near_cap, MEDUSA_ERR_INTERNAL, "sssp: near queue overflow");
        near_deg = c.near_deg;
        if (c.n_far > far_cap) {  // the far queue is full: keep what fits, rescan at the next split
            far_incomplete = true;
            n_far = (uint32_t)far_cap;
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
        const auto lock = enter(g);
        MDS_CHECK(dist, MEDUSA_ERR_ARG, "sssp: null output");
        MDS_CHECK(source < g->n, MEDUSA_ERR_ARG, "sssp: source %u out of range (n=%u)", source, g->n);
        MDS_CHECK(g->weighted, MEDUSA_ERR_STATE, "sssp: graph has no weights");
        medusa_opts o{};
        if (opts) o = *opts;
        MDS_CHECK(o.local_rounds <= 64 || o.local_rounds == 0xFFFFFFFFu, MEDUSA_ERR_ARG,
                  "sssp: local_rounds %u above 64", o.local_rounds);
        const uint32_t n = g->n;
        const size_t far_cap = sssp_far_alloc(g);
        g->dist.ensure(n);
        g->stamp.ensure(n);
        g->near_q.ensure(2 * near_capacity(n, g->m));  // near queues A and B ((vertex, slice) entries)
        g->sssp_phases.ensure(3ull * kMaxPhaseMarks);     // phase marks of a traced device-driven run
        g->q_aux.ensure(2 * far_cap);                   // far queues A and B
        medusa_stats st{};
        // timing: host-driven, every relaxation launch timed; otherwise device-driven (traced
        // by phase marks when trace is on)
        if (o.timing) sssp_run_host(g, source, dist, o, st);
        else sssp_run_device(g, source, dist, o, st);
        if (o.trace) sssp_trace_settled(g);
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
