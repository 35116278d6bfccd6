// Multi-source BFS: up to 64 sources advance together, one bit per source in a
// 64-bit mask per vertex (frontier, visited, next).
//
// No reference counterpart: "MBFS" in the reference is multi-GPU BFS
// (HowToBuild:13); the semantics here are those of k independent BFS runs
// (Algorithm/UserOperators.h:17-36 per source), checked against exactly that.
//
// Superstep:
//   push: active vertices (frontier mask != 0) OR their mask into next[v] of every
//         out-neighbour that still lacks some of those bits (64-bit atomicOr);
//   pull: every vertex not yet visited by all sources ORs the frontier masks of its
//         in-neighbours, stopping early once every missing bit has been found;
//   update: new = next & ~visited; visited |= new; frontier = new; level[s][v] =
//         step+1 for each new bit s; the active queue and counters are rebuilt with
//         ballot + in-workgroup prefix scan.

#include <algorithm>

#include "engine.h"
#include "sched.h"

namespace medusa {

struct MbfsCounters {
    uint32_t n_active;
    uint32_t saturated;  // a level >= 255 was clamped to 0xFF
    unsigned long long active_deg;
    unsigned long long new_bits;
    unsigned long long reached_edges;  // sum over new (source, vertex) pairs of outdeg(vertex)
    unsigned long long probes;
};

__global__ void k_mbfs_zero(MbfsCounters* c) {
    c->n_active = 0;
    c->saturated = 0;
    c->active_deg = 0;
    c->new_bits = 0;
    c->reached_edges = 0;
    c->probes = 0;
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

// Seeds the k sources (one thread: k <= 64) and the first active queue.
__global__ void k_mbfs_seed(uint32_t n, const uint32_t* __restrict__ srcs, uint32_t k, uint8_t* __restrict__ level,
                            unsigned long long* __restrict__ visited, unsigned long long* __restrict__ frontier,
                            const uint32_t* __restrict__ row_ptr, uint32_t* __restrict__ active,
                            MbfsCounters* __restrict__ ctr) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    uint32_t na = 0;
    unsigned long long deg = 0, re = 0;
    for (uint32_t s = 0; s < k; ++s) {
        const uint32_t v = srcs[s];
        if (frontier[v] == 0) {
            active[na++] = v;
            deg += row_ptr[v + 1] - row_ptr[v];
        }
        frontier[v] |= 1ull << s;
        visited[v] |= 1ull << s;
        level[(size_t)s * n + v] = 0;
        re += row_ptr[v + 1] - row_ptr[v];
    }
    ctr->n_active = na;
    ctr->active_deg = deg;
    ctr->new_bits = k;
    ctr->reached_edges = re;
}

__global__ __launch_bounds__(256) void k_mbfs_push(const uint32_t* __restrict__ q, uint32_t qn,
                                                   const uint32_t* __restrict__ row_ptr,
                                                   const uint32_t* __restrict__ col,
                                                   const unsigned long long* __restrict__ frontier,
                                                   const unsigned long long* __restrict__ visited,
                                                   unsigned long long* __restrict__ next) {
    __shared__ uint32_t s_beg[4][kWave], s_off[4][kWave];
    __shared__ unsigned long long s_mask[4][kWave];
    const uint32_t lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const uint32_t total_waves = gridDim.x * 4;
    for (uint32_t base = (blockIdx.x * 4 + wv) * kWave; base < qn; base += total_waves * kWave) {
        const uint32_t i = base + lane;
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

__global__ __launch_bounds__(256) void k_mbfs_pull(uint32_t n, unsigned long long full,
                                                   const uint32_t* __restrict__ cptr,
                                                   const uint32_t* __restrict__ csrc,
                                                   const unsigned long long* __restrict__ frontier,
                                                   const unsigned long long* __restrict__ visited,
                                                   unsigned long long* __restrict__ next,
                                                   MbfsCounters* __restrict__ ctr) {
    unsigned long long probes = 0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        const unsigned long long vis = visited[v];
        if (vis == full) continue;
        unsigned long long acc = 0;
        const uint32_t b = cptr[v], e = cptr[v + 1];
        for (uint32_t k = b; k < e; ++k) {
            acc |= frontier[csrc[k]];
            ++probes;
            if ((acc | vis) == full) break;
        }
        acc &= ~vis;
        if (acc) next[v] = acc;
    }
    probes = wave_sum_u64(probes);
    if ((threadIdx.x & (kWave - 1)) == 0 && probes) atomicAdd(&ctr->probes, probes);
}

constexpr uint32_t kMbfsTile = 4 * kWave * 4;  // vertices per workgroup tile (4 chunks per wave)

__global__ __launch_bounds__(256) void k_mbfs_update(uint32_t n, uint32_t k, int32_t level_val,
                                                     const uint32_t* __restrict__ row_ptr,
                                                     unsigned long long* __restrict__ visited,
                                                     unsigned long long* __restrict__ frontier,
                                                     unsigned long long* __restrict__ next,
                                                     uint8_t* __restrict__ level, uint32_t* __restrict__ active,
                                                     MbfsCounters* __restrict__ ctr) {
    __shared__ uint32_t s_cnt[4];
    __shared__ uint32_t s_base;
    const uint32_t lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const uint8_t lv = level_val >= 255 ? (uint8_t)0xFF : (uint8_t)level_val;
    for (uint32_t tile = blockIdx.x; (uint64_t)tile * kMbfsTile < n; tile += gridDim.x) {
        unsigned long long amask[4];
        uint32_t cnt = 0;
        unsigned long long adeg = 0, bits = 0, redges = 0;
#pragma unroll
        for (uint32_t c = 0; c < 4; ++c) {
            const uint32_t v = (tile * 16 + c * 4 + wv) * kWave + lane;
            unsigned long long nx = 0;
            if (v < n) {
                const unsigned long long raw = next[v];
                const unsigned long long fr = frontier[v];
                if (raw) {
                    next[v] = 0;
                    nx = raw & ~visited[v];
                }
                if (nx) visited[v] |= nx;
                if (fr != nx) frontier[v] = nx;
            }
            amask[c] = __ballot(nx != 0);
            if (amask[c]) {
                const uint32_t deg = nx ? row_ptr[v + 1] - row_ptr[v] : 0u;
                adeg += deg;
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
                cnt += __popcll(amask[c]);
            }
        }
        if (lane == 0) s_cnt[wv] = cnt;
        adeg = wave_sum_u64(adeg);
        bits = wave_sum_u64(bits);
        redges = wave_sum_u64(redges);
        if (lane == 0 && (adeg | bits)) {
            atomicAdd(&ctr->active_deg, adeg);
            atomicAdd(&ctr->new_bits, bits);
            atomicAdd(&ctr->reached_edges, redges);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t t = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
            s_base = t ? atomicAdd(&ctr->n_active, t) : 0u;
            if (level_val >= 255 && t) ctr->saturated = 1;
        }
        __syncthreads();
        uint32_t o = s_base;
        for (uint32_t w = 0; w < wv; ++w) o += s_cnt[w];
#pragma unroll
        for (uint32_t c = 0; c < 4; ++c) {
            const uint32_t v = (tile * 16 + c * 4 + wv) * kWave + lane;
            if ((amask[c] >> lane) & 1ull) active[o + lanes_below(amask[c])] = v;
            o += __popcll(amask[c]);
        }
        __syncthreads();
    }
}

static void mbfs_run(medusa_graph* g, const uint32_t* srcs, uint32_t k, uint8_t* out, const medusa_opts& o,
                     medusa_stats& st) {
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    const unsigned long long full = k == 64 ? ~0ull : ((1ull << k) - 1);
    g->q_small.ensure(n);
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
    hipStream_t s = g->stream;
    MDS_HIP(hipMemcpyAsync(g->q_aux.p, srcs, (size_t)k * 4, hipMemcpyHostToDevice, s));
    g->trace.clear();
    Sched sched(o, Sched::kMbfs);
    const bool timing = o.timing != 0;
    if (timing) ensure_events(g, 2);

    MDS_HIP(hipEventRecord(g->ev_begin, s));
    hipLaunchKernelGGL(k_mbfs_fill, dim3(grid_for(std::max<uint64_t>(n, (level_bytes + 3) / 4), 256, 8192)),
                       dim3(256), 0, s, n, (uint64_t)((level_bytes + 3) / 4), lvbuf.p, visited, frontier, next);
    MDS_LAUNCHED();
    hipLaunchKernelGGL(k_mbfs_zero, dim3(1), dim3(1), 0, s, C);
    MDS_LAUNCHED();
    hipLaunchKernelGGL(k_mbfs_seed, dim3(1), dim3(1), 0, s, n, (const uint32_t*)g->q_aux.p, k, level, visited,
                       frontier, g->row_ptr.p, g->q_small.p, C);
    MDS_LAUNCHED();
    MbfsCounters c = read_counters<MbfsCounters>(g);
    uint64_t reached_edges = c.reached_edges, reached = c.new_bits;
    Frontier f{};
    f.n = n;
    f.m = m;
    f.n_u = (uint64_t)n * k - c.new_bits;
    f.m_u = (uint64_t)m * k - c.reached_edges;
    bool pull = false;
    uint32_t saturated = 0;
    for (uint32_t t = 0;; ++t) {
        f.step = t;
        f.n_f = c.n_active;
        f.m_f = c.active_deg;
        if (f.n_f == 0) break;
        if (o.direction == MEDUSA_DIR_PUSH) pull = false;
        else if (o.direction == MEDUSA_DIR_PULL) pull = true;
        else pull = sched.bfs_pull(f, pull);
        if (timing) MDS_HIP(hipEventRecord(g->events[0], s));
        hipLaunchKernelGGL(k_mbfs_zero, dim3(1), dim3(1), 0, s, C);
        MDS_LAUNCHED();
        if (!pull) {
            const uint32_t waves = (c.n_active + kWave - 1) / kWave;
            hipLaunchKernelGGL(k_mbfs_push, dim3(grid_for(waves, 4, 16384)), dim3(256), 0, s,
                               (const uint32_t*)g->q_small.p, c.n_active, g->row_ptr.p, g->col.p,
                               (const unsigned long long*)frontier, (const unsigned long long*)visited, next);
        } else {
            hipLaunchKernelGGL(k_mbfs_pull, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, n, full, g->cptr.p,
                               g->csrc.p, (const unsigned long long*)frontier, (const unsigned long long*)visited,
                               next, C);
        }
        MDS_LAUNCHED();
        if (timing) MDS_HIP(hipEventRecord(g->events[1], s));
        hipLaunchKernelGGL(k_mbfs_update, dim3(grid_for(n, kMbfsTile, 4096)), dim3(256), 0, s, n, k, (int32_t)(t + 1),
                           g->row_ptr.p, visited, frontier, next, level, g->q_small.p, C);
        MDS_LAUNCHED();
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
        // byte model: push = queue+offsets+mask per active vertex, col+visited+next probe per
        // edge; pull = offsets+visited per vertex, source+frontier mask per probe; update = 24n
        const uint64_t kb = pull ? 4ull * (n + 1) + 8ull * n + 12ull * c.probes
                                 : 20ull * prev.n_active + 20ull * prev.active_deg;
        st.kernel_bytes += kb;
        st.alg_bytes += kb + 24ull * n + c.new_bits;
        if (pull) ++st.pull_steps; else ++st.push_steps;
        sched.record(g, f, pull ? 1.0 : 0.0, kms);
        ++st.supersteps;
        saturated |= c.saturated;
        reached += c.new_bits;
        reached_edges += c.reached_edges;
        f.n_u -= c.new_bits;
        f.m_u -= c.reached_edges;
    }
    MDS_HIP(hipEventRecord(g->ev_end, s));
    MDS_HIP(hipStreamSynchronize(s));
    float ms = 0;
    MDS_HIP(hipEventElapsedTime(&ms, g->ev_begin, g->ev_end));
    st.device_ms = ms;
    st.edges_traversed = reached_edges;
    st.reached = (uint32_t)std::min<uint64_t>(reached, 0xFFFFFFFFull);
    st.alg_bytes += level_bytes;
    st.reserved[0] = saturated;
    copy_out(g, out, level, level_bytes);
}

}  // namespace medusa

using namespace medusa;

extern "C" int medusa_mbfs(medusa_graph* g, const uint32_t* sources, uint32_t k, uint8_t* levels,
                           const medusa_opts* opts, medusa_stats* stats) {
    return guard([&] {
        enter(g);
        MDS_CHECK(levels && sources, MEDUSA_ERR_ARG, "mbfs: null argument");
        MDS_CHECK(k >= 1 && k <= 64, MEDUSA_ERR_ARG, "mbfs: k must be in [1, 64]");
        for (uint32_t s = 0; s < k; ++s)
            MDS_CHECK(sources[s] < g->n, MEDUSA_ERR_ARG, "mbfs: source %u out of range", sources[s]);
        medusa_opts o{};
        if (opts) o = *opts;
        MDS_CHECK(o.direction >= 0 && o.direction <= 2, MEDUSA_ERR_ARG, "mbfs: bad direction %d", o.direction);
        medusa_stats st{};
        mbfs_run(g, sources, k, levels, o, st);
        if (stats) *stats = st;
    });
}
