// Device-side graph object and helpers shared by the engine's HIP sources.
//
// One medusa_graph owns every device allocation for one graph on one GPU:
// CSR (out-edges) for push supersteps, CSC (in-edges, sources ascending) for
// pull supersteps and combines, plus lazily built per-algorithm layouts and
// work buffers.  Nothing lives in global symbols (the reference bound its
// arrays to __constant__ symbols, Algorithm/DeviceDataStructure.h:7-24), so any
// number of graphs can coexist in one process.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <vector>

#include "common.h"
#include "pr_bins.h"

namespace medusa {

#define MDS_HIP(expr)                                                                                \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess)                                                                        \
            ::medusa::fail(MEDUSA_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                           __LINE__);                                                                \
    } while (0)

#define MDS_LAUNCHED() MDS_HIP(hipGetLastError())

constexpr int kWave = 64;  // CDNA wavefront width

// RAII device array.
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count) {
            hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
            if (e != hipSuccess) {
                p = nullptr;
                fail(e == hipErrorOutOfMemory ? MEDUSA_ERR_NOMEM : MEDUSA_ERR_HIP, "hipMalloc(%zu bytes): %s",
                     count * sizeof(T), hipGetErrorString(e));
            }
            n = count;
        }
    }
    void ensure(size_t count) {
        if (n < count) alloc(count);
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void upload(const T* h, size_t count, hipStream_t s) {
        alloc(count);
        if (count) MDS_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
};

// RAII pinned host array (hipHostMalloc; no registration of existing memory).
template <class T>
struct HBuf {
    T* p = nullptr;
    size_t n = 0;
    HBuf() = default;
    HBuf(const HBuf&) = delete;
    HBuf& operator=(const HBuf&) = delete;
    ~HBuf() {
        if (p) (void)hipHostFree(p);
    }
    void ensure(size_t count) {
        if (n >= count) return;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
        MDS_HIP(hipHostMalloc((void**)&p, count * sizeof(T), hipHostMallocDefault));
        n = count;
    }
};

// A counter that workgroups update with atomics sits on its own 128-byte line: atomics on
// counters that share a line serialise on that line.
#define MDS_COUNTER(type, name) alignas(128) type name

#if defined(__HIPCC__)
// A load issued where it stands: relaxed atomic loads are not sunk into the branches that
// use them, so a one-workgroup decide kernel pays one memory round trip for all its inputs.
template <class T>
__device__ __forceinline__ T load_now(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Pins a uniform value in a scalar register at this point of the program, so the load that
// produced it is issued here with the loads around it rather than sunk into a later branch.
__device__ __forceinline__ void keep_here(uint32_t x) { asm volatile("" ::"s"(x)); }
// A whole control block, 8 bytes at a time, the same way.
template <class T>
__device__ __forceinline__ T load_now_block(const T* p) {
    static_assert(sizeof(T) % 8 == 0 && alignof(T) >= 8, "control blocks are 8-byte aligned");
    T out;
    auto* o = reinterpret_cast<unsigned long long*>(&out);
    const auto* q = reinterpret_cast<const unsigned long long*>(p);
#pragma unroll
    for (size_t i = 0; i < sizeof(T) / 8; ++i) o[i] = load_now(q + i);
    return out;
}
#endif

// Frontier counters written by the compaction kernels, read back each superstep.
struct FrontierCounters {
    MDS_COUNTER(uint32_t, n_small);              // frontier vertices with 1 <= outdeg <= large (wave queue)
    MDS_COUNTER(uint32_t, n_large);              // frontier vertices with outdeg > large (workgroup queue)
    MDS_COUNTER(uint32_t, n_new);                // vertices that joined the frontier (any outdeg)
    MDS_COUNTER(unsigned long long, deg_sum);    // sum of their out-degrees
    MDS_COUNTER(unsigned long long, probes);     // in-edges examined by a pull superstep
    MDS_COUNTER(unsigned long long, deg_large);  // out-degree sum of the workgroup (hub) queue
};

// One contiguous run of PageRank vertices handled with the same lane mapping.
struct PrRun {
    uint32_t begin;       // first internal vertex id
    uint32_t count;       // vertices in the run
    uint32_t kind;        // kPrLane / kPrWave / kPrZero
    uint32_t wg_begin;    // first workgroup of this run in the launch
    uint32_t deg;         // kPrLane: the in-degree shared by every vertex of the run (1..64)
    uint32_t ell_stride;  // kPrLane: column stride of the run's ELL slab (count rounded up to 64)
    uint32_t part;        // packed graph the run belongs to (0 for a single graph)
    uint32_t pad;
    uint64_t ell_base;    // kPrLane: offset of the slab in the ELL array
};
// A canonical block of the dangling-mass sum: up to 4096 consecutive internal ids.
struct PrBlock {
    uint32_t begin, len;
};
// Per packed graph: its dangling blocks and vertex count.
struct PrPart {
    uint32_t first_block, nblocks, n, pad;
};
struct PrBins {
    uint32_t nbins = 0, nhubs = 0, ntiles = 0;
    uint64_t nmsg = 0;     // messages per superstep (in-edges of non-hub destinations)
    uint64_t nentry = 0;   // message-array entries (runs padded for aligned access)
    DBuf<PrItem> items;    // combine items, bin by bin; then one kPrWave item per hub
    DBuf<PrBin> bins;
    DBuf<PrTile> tiles;
    DBuf<uint16_t> srcloc; // message order: source offset of every entry in its chunk
    DBuf<uint16_t> slot;   // bin order: LDS slot of every entry
    DBuf<uint32_t> qmsg;   // bin order: message-array position of every group of 4 entries, / 4
    DBuf<float> msg;       // the message array
};

// In-degree-sorted pull layout for PageRank (built on first use).
// Internal ids keep every packed graph's id range (a single graph is one range);
// inside a range: non-dangling vertices first, then dangling (outdeg 0); within
// each, in-degree descending, ties by original id.  Every run of equal in-degree
// (<= 64) is one contiguous id range inside one graph whose in-edge sources are
// stored as an ELL slab, column-major and 64-aligned, so the k-th source of 64
// consecutive vertices is one 256-byte load; writes of rank/contrib are
// contiguous, and each graph's dangling vertices form the tail of its range.
struct PrLayout {
    uint32_t n_dangling = 0;  // over all packed graphs
    std::vector<PrRun> runs;
    std::vector<PrBlock> blocks;
    std::vector<PrPart> parts;
    DBuf<PrBlock> d_blocks;
    DBuf<PrPart> d_parts;
    DBuf<uint32_t> d_part_begin;  // G+1 range starts (internal ids = original ids at range level)
    uint32_t total_wgs = 0;
    DBuf<PrRun> d_runs;
    DBuf<uint32_t> d_wg_run;  // run of every workgroup of the superstep launch (total_wgs)
    DBuf<uint32_t> ell;     // ELL slabs of the kPrLane runs
    DBuf<uint32_t> perm;    // internal -> original
    DBuf<uint32_t> iperm;   // original -> internal
    DBuf<uint32_t> cptr;    // in-edge offsets in internal order (n+1)
    DBuf<uint32_t> crow;    // in-edge sources, internal ids, original-source-ascending per vertex (m)
    DBuf<uint32_t> outdeg;  // out-degree per internal id (n)
    DBuf<uint32_t> optr;    // out-edge offsets in internal order (n+1), push variant
    DBuf<uint32_t> oslot;   // in-edge slot of every out-edge (m), push variant
    // host copies of cptr / crow / outdeg, from which the binned layout is built on first use
    std::vector<uint32_t> h_cptr, h_crow, h_od;
    std::unique_ptr<PrBins> bins;
    bool bins_refused = false;  // the binned plan failed for this layout
    // work
    DBuf<float> rank, rank_prev, contrib0, contrib1, msg, base, unperm;
    DBuf<float> partials;   // (s, c) pairs of the dangling-sum blocks
    DBuf<double> resid;
    // captured superstep sequence, reused while (steps, variant, damping) stay the same
    hipGraphExec_t exec = nullptr;
    uint32_t exec_steps = 0;
    uint64_t exec_sig = 0;  // signature of the per-superstep kernel variants
    float exec_damping = -1.f;
    ~PrLayout() {
        if (exec) (void)hipGraphExecDestroy(exec);
    }
};

// Layout for pull supersteps over 64-bit masks (MBFS): vertices with in-edges in in-degree-
// descending order, cut into runs as in PrLayout (equal in-degree 1..64: a lane per vertex
// over an ELL slab of ORIGINAL source ids, so the k-th sources of 64 vertices are one coalesced
// load; in-degree > 64: a wave per vertex over the CSC segment).  Vertices without in-edges
// never join through a pull and are left out.
struct PullLayout {
    std::vector<PrRun> runs;
    uint32_t total_wgs = 0;  // work items of 256 lane-vertices or 4 wave-vertices
    DBuf<PrRun> d_runs;
    DBuf<uint32_t> wg_run;   // run of every work item
    DBuf<uint32_t> perm;     // sorted position -> original vertex id
    DBuf<uint32_t> ell;      // ELL slabs (original source ids)
};

struct TraceRec {
    double f[MEDUSA_TRACE_FIELDS];
};

// An instantiated hipGraph of captured supersteps and the buffers it was captured with.
struct GraphCache {
    hipGraphExec_t exec = nullptr;
    uintptr_t key[6] = {0, 0, 0, 0, 0, 0};
    GraphCache() = default;
    GraphCache(const GraphCache&) = delete;
    GraphCache& operator=(const GraphCache&) = delete;
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
    }
    ~GraphCache() { reset(); }
};

// Slots in the first replay of a device-driven loop: the previous run's superstep count on
// the same handle (capped), or the regular batch when there is none.  Slots after the
// loop's end return at once, so the choice changes only the time, never the result.
constexpr uint32_t kMaxHeadSlots = 64;
inline uint32_t head_slots(uint32_t last_steps, uint32_t batch) {
    return last_steps ? (last_steps < kMaxHeadSlots ? last_steps : kMaxHeadSlots) : batch;
}

// 256 bytes of device memory for a device-driven loop's control block.
struct CtlStorage {
    unsigned long long w[32];
};
using BfsCtlStorage = CtlStorage;

}  // namespace medusa

struct medusa_graph {
    // serialises calls on this handle (recursive: the autotuner calls the BFS entry points)
    mutable std::recursive_mutex mu;
    uint32_t n = 0;
    uint64_t m = 0;
    bool weighted = false;
    int device = 0;
    hipStream_t stream = nullptr;
    // host copies (layouts are derived from them)
    std::vector<uint32_t> h_row_ptr, h_col, h_w, h_cptr, h_csrc, h_cw, h_slot;
    // device CSR / CSC
    medusa::DBuf<uint32_t> row_ptr, col, w, cptr, csrc, cw, slot;
    // packed batch of independent graphs: graph k owns vertices [parts[k], parts[k+1])
    // and no edge leaves its range; a plain graph is the single range {0, n}
    std::vector<uint32_t> parts;
    // launch parameters chosen by medusa_autotune (0 = engine default)
    struct {
        uint32_t bfs_large = 0, bfs_grid = 0, mbfs_large = 0, mbfs_grid = 0;
    } tuned;
    // traversal work buffers
    medusa::DBuf<int32_t> level;
    medusa::DBuf<uint32_t> dist, q_small, q_large, q_aux, stamp;
    medusa::DBuf<unsigned long long> bitmap;
    medusa::DBuf<unsigned long long> near_q;  // SSSP near queues of (slice << 32 | vertex) entries
    medusa::DBuf<unsigned long long> sssp_phases;  // SSSP device-loop phase marks (trace on)
    // per-superstep counters (device) and their pinned host mirror; each engine
    // views the 256 bytes as its own counter struct
    medusa::DBuf<unsigned long long> ctr_raw;
    medusa::HBuf<unsigned long long> h_ctr_raw;
    // device-driven superstep loops: control blocks and captured superstep batches
    medusa::DBuf<medusa::CtlStorage> bfs_ctl, mbfs_ctl, sssp_ctl;
    medusa::GraphCache bfs_graph, mbfs_graph, sssp_graph;
    // first replay of the BFS / MBFS loops: as many slots as the previous run on this handle
    // took supersteps, so a repeated traversal ends after one replay with no idle slots
    medusa::GraphCache bfs_head, mbfs_head;
    uint32_t bfs_last_steps = 0, mbfs_last_steps = 0;
    std::unique_ptr<medusa::PrLayout> pr;
    std::unique_ptr<medusa::PullLayout> pull;  // built on the first MBFS run
    std::vector<medusa::TraceRec> trace;
    std::vector<hipEvent_t> events;  // timing pool
    hipEvent_t ev_begin = nullptr, ev_end = nullptr;
};

namespace medusa {

// Checks g, takes its lock for the rest of the call and makes its device current.
[[nodiscard]] std::unique_lock<std::recursive_mutex> enter(const medusa_graph* g);
// Grows the timing-event pool to at least `count` events.
void ensure_events(medusa_graph* g, size_t count);
// Copies `bytes` from device pointer `src` to `dst`, which may be host or device memory.
void copy_out(medusa_graph* g, void* dst, const void* src, size_t bytes);
// Sums pairwise (start, stop) event intervals [0, pairs) in milliseconds.
double sum_event_pairs(medusa_graph* g, size_t pairs);
// Number of workgroups for a grid-stride launch over `items` with `per_wg` items each.
inline uint32_t grid_for(uint64_t items, uint32_t per_wg, uint32_t cap = 8192) {
    uint64_t g = (items + per_wg - 1) / per_wg;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (uint32_t)g;
}
void build_pr_layout(medusa_graph* g);
// Plans and uploads the binned PageRank layout (once per layout).  Returns false when the graph
// cannot be planned; with `required` that is a MEDUSA_ERR_STATE error instead.
bool build_pr_bins(medusa_graph* g, bool required);
// Send-tile workgroups of the binned PageRank that run at once on g's device, after the ones
// that sum `nblk` dangling blocks in the same launch.
uint32_t pr_send_tile_slots(const medusa_graph* g, uint32_t nblk);
// Binned-combine workgroups that run at once on g's device.
uint32_t pr_combine_slots(const medusa_graph* g);
void build_pull_layout(medusa_graph* g);

constexpr size_t kCounterWords = 256;  // 2 KiB
template <class T>
T* dev_counters(medusa_graph* g) {
    static_assert(sizeof(T) <= kCounterWords * 8, "counter struct too large");
    return reinterpret_cast<T*>(g->ctr_raw.p);
}
template <class T>
T* host_counters(medusa_graph* g) {
    return reinterpret_cast<T*>(g->h_ctr_raw.p);
}
// Copies the device counters to the pinned mirror and waits for the stream.
template <class T>
T read_counters(medusa_graph* g) {
    MDS_HIP(hipMemcpyAsync(g->h_ctr_raw.p, g->ctr_raw.p, sizeof(T), hipMemcpyDeviceToHost, g->stream));
    MDS_HIP(hipStreamSynchronize(g->stream));
    return *host_counters<T>(g);
}

}  // namespace medusa

// ---- device-side primitives ---------------------------------------------
#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)
namespace medusa {

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
// Number of set bits of `mask` in lanes below this one.
__device__ __forceinline__ uint32_t lanes_below(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, kWave);
        if (lane >= (uint32_t)off) x += y;
    }
    return x;
}
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long x) {
#pragma unroll
    for (int off = kWave / 2; off >= 1; off >>= 1) x += __shfl_xor(x, off, kWave);
    return x;
}

// Compensated summation used by every FP32 combine.  A pair (s, c) stands
// for the value s - c.  kahan_add is Kahan's update; kahan_combine merges two
// pairs with an error-free TwoSum of the leading parts.  Both are exact
// identities against the pair (0, 0), so zero-padded lanes do not change a sum.
__device__ __forceinline__ void kahan_add(float& s, float& c, float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}
__device__ __forceinline__ void kahan_combine(float& s, float& c, float s2, float c2) {
    const float t = s + s2;
    const float bp = t - s;
    const float err = (s - (t - bp)) + (s2 - bp);
    c = (c + c2) - err;
    s = t;
}

}  // namespace medusa
#endif
