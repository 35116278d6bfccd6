// CPU oracle for the BSP superstep engine — TEST INFRASTRUCTURE ONLY.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// this library, and only as the checker (or the timed CPU baseline).  The
// product path never calls it.
//
// Two families:
//  (1) restatements of the reference CPU path (the CPU baseline):
//      BFS edge sweep        Algorithm/CPUFunctorHoster.cpp:16-65
//      SSSP message array    SSSP/CPUFunctorHoster.cpp:15-99
//      PageRank vertex push  MPG/CPUFunctorHoster.cpp:17-169 (reference formula)
//      CSR/CSC/msgDstID      Algorithm/VertexDataType.cu:60-80, SSSP/EdgeDataType.cu:138-186
//      segmented scan        CUDPP cudppSegmentedScan as called at SSSP/MsgCombiner.h:128,
//                            MPG/MsgCombiner.h:103 (forward, inclusive, flag-headed segments)
//  (2) textbook oracles: queue BFS (the never-called CPUBFS, Algorithm/main.cu:95-137),
//      binary-heap Dijkstra (in place of the O(V^2) CPUSSSP, SSSP/main.cu:172-208),
//      Bellman-Ford brute force, FP64 PageRank power iteration, k independent BFS for
//      multi-source BFS, and an FP32 restatement of the engine's canonical
//      compensated-combine order for PageRank.
//
// The reference's termination vote `bool cont[24]` indexed by omp_get_thread_num()
// (Algorithm/CPUFunctorHoster.cpp:3,27) is out of bounds above 24 threads; the
// restatements use an OpenMP reduction(||) instead.

#include <omp.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <queue>
#include <vector>

namespace {

constexpr uint32_t kUnreached = 0xFFFFFFFFu;

void set_threads(int threads) {
    if (threads > 0) omp_set_num_threads(threads);
}

// CSC with sources ascending in every destination segment; slot[e] = CSC position
// of CSR edge e (the reference's msgDstID).
void build_csc(uint32_t n, uint64_t m, const uint32_t* row_ptr, const uint32_t* col, std::vector<uint32_t>& cptr,
               std::vector<uint32_t>& csrc, std::vector<uint32_t>& slot) {
    cptr.assign((size_t)n + 1, 0);
    for (uint64_t e = 0; e < m; ++e) ++cptr[col[e] + 1];
    for (uint32_t v = 0; v < n; ++v) cptr[v + 1] += cptr[v];
    csrc.resize(m);
    slot.resize(m);
    std::vector<uint32_t> fill(cptr.begin(), cptr.end() - 1);
    for (uint32_t u = 0; u < n; ++u)
        for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; ++e) {
            const uint32_t pos = fill[col[e]]++;
            csrc[pos] = u;
            slot[e] = pos;
        }
}

// ---- canonical compensated summation (FP32) -------------------------------
// Definition followed by every GPU kernel variant: blocks of 4096; lane j of a
// block Kahan-adds elements j, j+64, ... from (0,0); lanes merge down a tree with
// offsets 32..1 (lane i absorbs lane i+off); blocks merge left to right; value s-c.
inline void kahan_add(float& s, float& c, float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}
inline void kahan_combine(float& s, float& c, float s2, float c2) {
    const float t = s + s2;
    const float bp = t - s;
    const float err = (s - (t - bp)) + (s2 - bp);
    c = (c + c2) - err;
    s = t;
}
template <class Get>
float canonical_sum(uint64_t len, Get get) {
    float as = 0.f, ac = 0.f;
    for (uint64_t b0 = 0; b0 < len; b0 += 4096) {
        const uint64_t lim = std::min<uint64_t>(len - b0, 4096);
        float s[64], c[64];
        for (int j = 0; j < 64; ++j) {
            s[j] = 0.f;
            c[j] = 0.f;
            for (uint64_t k = j; k < lim; k += 64) kahan_add(s[j], c[j], get(b0 + k));
        }
        // lanes at or beyond `active` hold (0,0): merging with them is an exact identity
        uint32_t active = (uint32_t)std::min<uint64_t>(lim, 64);
        for (uint32_t off = 32; off >= 1; off >>= 1) {
            for (uint32_t i = 0; i < off && i + off < active; ++i) kahan_combine(s[i], c[i], s[i + off], c[i + off]);
            active = std::min(active, off);
        }
        if (b0 == 0) {
            as = s[0];
            ac = c[0];
        } else {
            kahan_combine(as, ac, s[0], c[0]);
        }
    }
    return as - ac;
}

}  // namespace

extern "C" {

int orc_version(void) { return 1; }

// CSR -> CSC (sources ascending per destination) and the CSR-edge -> CSC-slot map.
// Restates VertexDataType::build's msg_index prefix sum and EdgeArray::buildAA's
// msgDstID (SSSP/VertexDataType.cu:64-85, SSSP/EdgeDataType.cu:138-186).
int orc_csc(uint32_t n, uint64_t m, const uint32_t* row_ptr, const uint32_t* col, uint32_t* cptr_out,
            uint32_t* csrc_out, uint32_t* slot_out) {
    std::vector<uint32_t> cptr, csrc, slot;
    build_csc(n, m, row_ptr, col, cptr, csrc, slot);
    std::memcpy(cptr_out, cptr.data(), ((size_t)n + 1) * 4);
    if (m) {
        std::memcpy(csrc_out, csrc.data(), m * 4);
        std::memcpy(slot_out, slot.data(), m * 4);
    }
    return 0;
}

// Forward inclusive segmented scan (op 0 = ADD, 1 = MIN) with segment heads where
// flags[i] != 0: the combiner arithmetic the reference delegates to CUDPP
// (MedusaRT/Combiner.cu:34-52 plan; reduced value = last element of a segment,
// SSSP/DeviceDataStructure.h:221-235).  Sequential left-to-right order.
int orc_segmented_scan(uint64_t len, const float* in, const uint32_t* flags, int op, float* out) {
    for (uint64_t i = 0; i < len; ++i) {
        if (i == 0 || flags[i]) {
            out[i] = in[i];
        } else {
            out[i] = op == 0 ? out[i - 1] + in[i] : std::min(out[i - 1], in[i]);
        }
    }
    return 0;
}

// Textbook queue BFS (Algorithm/main.cu:95-137).  level: -1 unreached.
int orc_bfs_queue(uint32_t n, const uint32_t* row_ptr, const uint32_t* col, uint32_t src, int32_t* level) {
    for (uint32_t v = 0; v < n; ++v) level[v] = -1;
    std::vector<uint32_t> q;
    q.reserve(n);
    level[src] = 0;
    q.push_back(src);
    for (size_t h = 0; h < q.size(); ++h) {
        const uint32_t u = q[h];
        for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; ++e) {
            const uint32_t v = col[e];
            if (level[v] == -1) {
                level[v] = level[u] + 1;
                q.push_back(v);
            }
        }
    }
    return 0;
}

// Reference CPU path for BFS: every superstep sweeps all edges; an edge whose
// source is at level `step` claims an unvisited destination (Algorithm/CPUFunctorHoster.cpp:16-65).
// Returns the number of supersteps (last one finds nothing).
int orc_bfs_sweep(uint32_t n, uint64_t m, const uint32_t* row_ptr, const uint32_t* col, uint32_t src, int32_t* level,
                  int threads) {
    set_threads(threads);
    std::vector<uint32_t> esrc(m);
    for (uint32_t u = 0; u < n; ++u)
        for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; ++e) esrc[e] = u;
    for (uint32_t v = 0; v < n; ++v) level[v] = -1;
    level[src] = 0;
    int step = 0;
    for (;;) {
        bool cont = false;
#pragma omp parallel for reduction(|| : cont) schedule(static)
        for (int64_t e = 0; e < (int64_t)m; ++e) {
            if (level[esrc[e]] == step) {
                const uint32_t d = col[e];
                if (level[d] == -1) {
                    level[d] = step + 1;
                    cont = true;
                }
            }
        }
        ++step;
        if (!cont) break;
    }
    return step;
}

// k independent BFS runs; levels[s*n + v] uint8, 0xFF unreached or >= 255.
int orc_mbfs(uint32_t n, const uint32_t* row_ptr, const uint32_t* col, const uint32_t* srcs, uint32_t k,
             uint8_t* levels, int threads) {
    set_threads(threads);
#pragma omp parallel
    {
        std::vector<int32_t> lv(n);
#pragma omp for schedule(dynamic, 1)
        for (int64_t s = 0; s < (int64_t)k; ++s) {
            orc_bfs_queue(n, row_ptr, col, srcs[s], lv.data());
            uint8_t* out = levels + (size_t)s * n;
            for (uint32_t v = 0; v < n; ++v) out[v] = (lv[v] < 0 || lv[v] >= 255) ? 0xFF : (uint8_t)lv[v];
        }
    }
    return 0;
}

// Binary-heap Dijkstra with 64-bit tentative distances.  dist: uint32, 0xFFFFFFFF
// unreached.  Returns -1 if a shortest distance does not fit in 32 bits.
int orc_sssp_dijkstra(uint32_t n, const uint32_t* row_ptr, const uint32_t* col, const uint32_t* w, uint32_t src,
                      uint32_t* dist) {
    std::vector<uint64_t> d(n, UINT64_MAX);
    using Item = std::pair<uint64_t, uint32_t>;
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
    d[src] = 0;
    pq.push({0, src});
    while (!pq.empty()) {
        const auto [du, u] = pq.top();
        pq.pop();
        if (du != d[u]) continue;
        for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; ++e) {
            const uint64_t nd = du + w[e];
            if (nd < d[col[e]]) {
                d[col[e]] = nd;
                pq.push({nd, col[e]});
            }
        }
    }
    int rc = 0;
    for (uint32_t v = 0; v < n; ++v) {
        if (d[v] == UINT64_MAX) {
            dist[v] = kUnreached;
        } else if (d[v] >= kUnreached) {
            dist[v] = kUnreached;
            rc = -1;
        } else {
            dist[v] = (uint32_t)d[v];
        }
    }
    return rc;
}

// Brute force: relax every edge until nothing changes (tiny graphs only).
int orc_sssp_bellman_ford(uint32_t n, const uint32_t* row_ptr, const uint32_t* col, const uint32_t* w, uint32_t src,
                          uint32_t* dist) {
    std::vector<uint64_t> d(n, UINT64_MAX);
    d[src] = 0;
    for (bool changed = true; changed;) {
        changed = false;
        for (uint32_t u = 0; u < n; ++u) {
            if (d[u] == UINT64_MAX) continue;
            for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; ++e)
                if (d[u] + w[e] < d[col[e]]) {
                    d[col[e]] = d[u] + w[e];
                    changed = true;
                }
        }
    }
    for (uint32_t v = 0; v < n; ++v) dist[v] = d[v] >= kUnreached ? kUnreached : (uint32_t)d[v];
    return 0;
}

// Reference CPU path for SSSP: every superstep fills all message slots with the
// init value, every edge sends dist[src] + w into its slot msgDstID, and every
// vertex takes the minimum over its slot segment (SSSP/CPUFunctorHoster.cpp:15-99).
// uint32 contract: an unreached source sends the init value (no wrap-around).
// Returns the number of supersteps.
int orc_sssp_bsp(uint32_t n, uint64_t m, const uint32_t* row_ptr, const uint32_t* col, const uint32_t* w,
                 uint32_t src, uint32_t* dist, int threads, int max_steps) {
    set_threads(threads);
    std::vector<uint32_t> cptr, csrc, slot, esrc(m), msg(m);
    build_csc(n, m, row_ptr, col, cptr, csrc, slot);
    for (uint32_t u = 0; u < n; ++u)
        for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; ++e) esrc[e] = u;
    for (uint32_t v = 0; v < n; ++v) dist[v] = kUnreached;
    dist[src] = 0;
    int step = 0;
    for (;;) {
        bool cont = false;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)m; ++i) msg[i] = kUnreached;
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < (int64_t)m; ++e) {
            const uint32_t du = dist[esrc[e]];
            const uint64_t nd = du == kUnreached ? kUnreached : (uint64_t)du + w[e];
            msg[slot[e]] = nd >= kUnreached ? kUnreached : (uint32_t)nd;
        }
#pragma omp parallel for reduction(|| : cont) schedule(static)
        for (int64_t v = 0; v < (int64_t)n; ++v) {
            uint32_t mn = kUnreached;
            for (uint32_t i = cptr[v]; i < cptr[v + 1]; ++i) mn = std::min(mn, msg[i]);
            if (mn < dist[v]) {
                dist[v] = mn;
                cont = true;
            }
        }
        ++step;
        if (!cont || (max_steps > 0 && step >= max_steps)) break;
    }
    return step;
}

// Supersteps the reference CPU path needs for SSSP, from final distances: the
// synchronous relaxation settles a vertex after as many supersteps as the fewest arcs on
// any of its shortest paths, plus one final superstep that changes nothing.
int orc_sssp_rounds(uint32_t n, const uint32_t* row_ptr, const uint32_t* col, const uint32_t* w, uint32_t src,
                    const uint32_t* dist) {
    std::vector<uint32_t> order;
    for (uint32_t v = 0; v < n; ++v)
        if (dist[v] != kUnreached) order.push_back(v);
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return dist[a] < dist[b]; });
    std::vector<uint32_t> hops(n, 0xFFFFFFFFu);
    hops[src] = 0;
    uint32_t worst = 0;
    // zero-weight arcs are impossible (weights >= 1), so distance order is a topological order
    for (uint32_t u : order) {
        if (hops[u] == 0xFFFFFFFFu) continue;
        worst = std::max(worst, hops[u]);
        for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; ++e) {
            const uint32_t v = col[e];
            if ((uint64_t)dist[u] + w[e] == dist[v]) hops[v] = std::min(hops[v], hops[u] + 1);
        }
    }
    return (int)worst + 1;
}

// FP64 PageRank power iteration with the engine's formula:
//   r_0 = 1/n, r_{t+1}(v) = (1-d)/n + d*(sum_{u->v} r_t(u)/outdeg(u) + D_t/n).
int orc_pagerank_f64(uint32_t n, uint64_t m, const uint32_t* row_ptr, const uint32_t* col, double d, uint32_t steps,
                     double* rank, int threads) {
    set_threads(threads);
    std::vector<uint32_t> cptr, csrc, slot;
    build_csc(n, m, row_ptr, col, cptr, csrc, slot);
    std::vector<double> r(n, 1.0 / n), nr(n), contrib(n);
    for (uint32_t t = 0; t < steps; ++t) {
        double D = 0;
#pragma omp parallel for reduction(+ : D) schedule(static)
        for (int64_t u = 0; u < (int64_t)n; ++u) {
            const uint32_t od = row_ptr[u + 1] - row_ptr[u];
            contrib[u] = od ? r[u] / od : 0.0;
            if (!od) D += r[u];
        }
        const double base = (1.0 - d) / n + d * D / n;
#pragma omp parallel for schedule(static)
        for (int64_t v = 0; v < (int64_t)n; ++v) {
            double s = 0;
            for (uint32_t i = cptr[v]; i < cptr[v + 1]; ++i) s += contrib[csrc[i]];
            nr[v] = base + d * s;
        }
        r.swap(nr);
    }
    std::memcpy(rank, r.data(), (size_t)n * 8);
    return 0;
}

// FP32 PageRank in the engine's canonical order (see canonical_sum).  Scalars:
// d_f=(float)d, tele=(float)((1-d)/n), r0=(float)(1/n), base=fmaf(d_f, D/(float)n, tele),
// r=fmaf(d_f, S, base), contrib=r/(float)outdeg.  Dangling vertices are summed in the
// order (in-degree descending, id ascending).
int orc_pagerank_f32(uint32_t n, uint64_t m, const uint32_t* row_ptr, const uint32_t* col, float d, uint32_t steps,
                     float* rank, int threads) {
    set_threads(threads);
    std::vector<uint32_t> cptr, csrc, slot;
    build_csc(n, m, row_ptr, col, cptr, csrc, slot);
    std::vector<uint32_t> dang;
    for (uint32_t v = 0; v < n; ++v)
        if (row_ptr[v + 1] == row_ptr[v]) dang.push_back(v);
    std::stable_sort(dang.begin(), dang.end(), [&](uint32_t a, uint32_t b) {
        return (cptr[a + 1] - cptr[a]) > (cptr[b + 1] - cptr[b]);
    });
    const double dd = (double)d;
    const float nf = (float)n;
    const float tele = (float)((1.0 - dd) / (double)n);
    const float r0 = (float)(1.0 / (double)n);
    std::vector<float> r(n, r0), nr(n), contrib(n, 0.f);
    auto dangling_base = [&](const std::vector<float>& rv) {
        const float D = canonical_sum(dang.size(), [&](uint64_t i) { return rv[dang[i]]; });
        return std::fmaf(d, D / nf, tele);
    };
    for (uint32_t v = 0; v < n; ++v) {
        const uint32_t od = row_ptr[v + 1] - row_ptr[v];
        if (od) contrib[v] = r0 / (float)od;
    }
    float base = dangling_base(r);
    for (uint32_t t = 0; t < steps; ++t) {
#pragma omp parallel for schedule(dynamic, 4096)
        for (int64_t v = 0; v < (int64_t)n; ++v) {
            const uint32_t b = cptr[v];
            const float S = canonical_sum(cptr[v + 1] - b, [&](uint64_t k) { return contrib[csrc[b + k]]; });
            nr[v] = std::fmaf(d, S, base);
        }
        r.swap(nr);
        for (uint32_t v = 0; v < n; ++v) {
            const uint32_t od = row_ptr[v + 1] - row_ptr[v];
            if (od) contrib[v] = r[v] / (float)od;
        }
        base = dangling_base(r);
    }
    std::memcpy(rank, r.data(), (size_t)n * 4);
    return 0;
}

// Reference CPU path for PageRank, reference formula: init 1.0; each superstep
// zero the message slots, every vertex pushes level/edge_count into the slots of
// its out-edges, every vertex sums its slots in order and sets
// level = 0.15 + 0.85*sum with the reference's float/double promotions
// (MPG/CPUFunctorHoster.cpp:17-31,143-169, MPG/main.cu:72).
int orc_pagerank_ref(uint32_t n, uint64_t m, const uint32_t* row_ptr, const uint32_t* col, uint32_t steps, float* rank,
                     int threads) {
    set_threads(threads);
    std::vector<uint32_t> cptr, csrc, slot;
    build_csc(n, m, row_ptr, col, cptr, csrc, slot);
    std::vector<float> level(n, 1.0f), msg(m);
    for (uint32_t t = 0; t < steps; ++t) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)m; ++i) msg[i] = 0.0f;
#pragma omp parallel for schedule(static)
        for (int64_t u = 0; u < (int64_t)n; ++u) {
            const uint32_t b = row_ptr[u], e = row_ptr[u + 1];
            const int edge_count = (int)(e - b);
            for (uint32_t k = b; k < e; ++k) msg[slot[k]] = level[u] / edge_count;
        }
#pragma omp parallel for schedule(static)
        for (int64_t v = 0; v < (int64_t)n; ++v) {
            float sum = 0;
            for (uint32_t i = cptr[v]; i < cptr[v + 1]; ++i) sum += msg[i];
            sum *= 0.85;
            level[v] = 0.15 + sum;
        }
    }
    std::memcpy(rank, level.data(), (size_t)n * 4);
    return 0;
}

}  // extern "C"

// ---- scheduler MLP in FP64 ---------------------------------------------------------
// Same network as the engine's: x[16] -> 32 -> 32 -> 1, leaky ReLU (slope 1/64), parameters
// W1[32][16], b1[32], W2[32][32], b2[32], W3[32], b3; loss = mean squared error.  Written
// from the chain rule in double precision; the GPU's FP32 gradients are checked against it.
extern "C" int orc_mlp_loss_grad(const float* params, const float* X, const float* y, uint32_t n, double* grad,
                                 double* loss) {
    constexpr int I = 16, H = 32;
    const float* W1 = params;
    const float* b1 = W1 + H * I;
    const float* W2 = b1 + H;
    const float* b2 = W2 + H * H;
    const float* W3 = b2 + H;
    const double b3 = W3[H];
    const double slope = 1.0 / 64.0;
    const int P = H * I + H + H * H + H + H + 1;
    for (int p = 0; p < P; ++p) grad[p] = 0.0;
    double L = 0.0;
    for (uint32_t s = 0; s < n; ++s) {
        const float* x = X + (size_t)s * I;
        double z1[H], a1[H], z2[H], a2[H];
        for (int i = 0; i < H; ++i) {
            double z = b1[i];
            for (int k = 0; k < I; ++k) z += (double)W1[i * I + k] * x[k];
            z1[i] = z;
            a1[i] = z > 0 ? z : z * slope;
        }
        double o = b3;
        for (int i = 0; i < H; ++i) {
            double z = b2[i];
            for (int k = 0; k < H; ++k) z += (double)W2[i * H + k] * a1[k];
            z2[i] = z;
            a2[i] = z > 0 ? z : z * slope;
            o += (double)W3[i] * a2[i];
        }
        const double e = o - (double)y[s];
        L += e * e;
        const double g = 2.0 * e / n;
        double d2[H];
        for (int i = 0; i < H; ++i) {
            grad[H * I + H + H * H + H + i] += g * a2[i];
            d2[i] = g * W3[i] * (z2[i] > 0 ? 1.0 : slope);
            grad[H * I + H + H * H + i] += d2[i];
            for (int k = 0; k < H; ++k) grad[H * I + H + i * H + k] += d2[i] * a1[k];
        }
        grad[P - 1] += g;
        for (int k = 0; k < H; ++k) {
            double acc = 0;
            for (int i = 0; i < H; ++i) acc += d2[i] * W2[i * H + k];
            const double d1 = acc * (z1[k] > 0 ? 1.0 : slope);
            grad[H * I + k] += d1;
            for (int j = 0; j < I; ++j) grad[k * I + j] += d1 * x[j];
        }
    }
    *loss = L / n;
    return 0;
}

extern "C" int orc_mlp_forward(const float* params, const float* X, uint32_t n, double* out) {
    constexpr int I = 16, H = 32;
    const float* W1 = params;
    const float* b1 = W1 + H * I;
    const float* W2 = b1 + H;
    const float* b2 = W2 + H * H;
    const float* W3 = b2 + H;
    for (uint32_t s = 0; s < n; ++s) {
        const float* x = X + (size_t)s * I;
        double a1[H];
        for (int i = 0; i < H; ++i) {
            double z = b1[i];
            for (int k = 0; k < I; ++k) z += (double)W1[i * I + k] * x[k];
            a1[i] = z > 0 ? z : z / 64.0;
        }
        double o = W3[H];
        for (int i = 0; i < H; ++i) {
            double z = b2[i];
            for (int k = 0; k < H; ++k) z += (double)W2[i * H + k] * a1[k];
            o += (double)W3[i] * (z > 0 ? z : z / 64.0);
        }
        out[s] = o;
    }
    return 0;
}
