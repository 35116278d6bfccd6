// Graph storage: CSR/CSC upload, the C-ABI library entry points, and the
// PageRank pull layout.
//
// Reference counterparts: VertexArray::build (CSR offsets, Algorithm/VertexDataType.cu:60-80),
// EdgeArray::buildAA (CSR order, CSC offsets msg_index, msgDstID slot map, segment-head
// flags; SSSP/EdgeDataType.cu:138-186), D_*::Fill uploads and InitDeviceDS
// (MultipleGPU/MultiPublicAPI.h:144-308), sortByDegree relabelling for MEG/HY
// (MedusaRT/GraphConverter.cu:136-165, MultiPublicAPI.h:52-53).

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <numeric>

#include "engine.h"

namespace medusa {

static thread_local std::string t_last_error;

void set_last_error(const char* msg) { t_last_error = msg; }
void clear_last_error() { t_last_error.clear(); }

void fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw Error(code, buf);
}

std::unique_lock<std::recursive_mutex> enter(const medusa_graph* g) {
    MDS_CHECK(g != nullptr, MEDUSA_ERR_ARG, "null graph handle");
    std::unique_lock<std::recursive_mutex> lock(g->mu);
    MDS_HIP(hipSetDevice(g->device));
    return lock;
}

void ensure_events(medusa_graph* g, size_t count) {
    while (g->events.size() < count) {
        hipEvent_t e;
        MDS_HIP(hipEventCreate(&e));
        g->events.push_back(e);
    }
}

double sum_event_pairs(medusa_graph* g, size_t pairs) {
    double ms = 0;
    for (size_t i = 0; i < pairs; ++i) {
        float t = 0;
        MDS_HIP(hipEventElapsedTime(&t, g->events[2 * i], g->events[2 * i + 1]));
        ms += t;
    }
    return ms;
}

void copy_out(medusa_graph* g, void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    hipPointerAttribute_t attr{};
    hipError_t e = hipPointerGetAttributes(&attr, dst);
    (void)hipGetLastError();  // host pointers unknown to HIP report an error; that is the "host" answer
    const bool dev = (e == hipSuccess && attr.type == hipMemoryTypeDevice);
    MDS_HIP(hipMemcpyAsync(dst, src, bytes, dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, g->stream));
    MDS_HIP(hipStreamSynchronize(g->stream));
}

// Internal id order for PageRank: inside every packed graph's id range,
// (dangling last, in-degree descending, original id).
void build_pull_layout(medusa_graph* g) {
    if (g->pull) return;
    const uint32_t n = g->n;
    auto L = std::make_unique<PullLayout>();
    std::vector<uint32_t> order;
    order.reserve(n);
    for (uint32_t v = 0; v < n; ++v)
        if (g->h_cptr[v + 1] > g->h_cptr[v]) order.push_back(v);
    auto indeg = [&](uint32_t v) { return g->h_cptr[v + 1] - g->h_cptr[v]; };
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return indeg(a) > indeg(b); });
    const uint32_t cnt = (uint32_t)order.size();
    uint32_t wg = 0;
    uint64_t ell_size = 0;
    for (uint32_t p = 0; p < cnt;) {
        const uint32_t deg = indeg(order[p]);
        const bool lane = deg <= kPrLaneMax;
        uint32_t q = p + 1;
        while (q < cnt && (lane ? indeg(order[q]) == deg : indeg(order[q]) > kPrLaneMax)) ++q;
        PrRun r{};
        r.begin = p;
        r.count = q - p;
        r.kind = lane ? kPrLane : kPrWave;
        r.wg_begin = wg;
        if (lane) {
            r.deg = deg;
            r.ell_stride = (r.count + 63u) & ~63u;
            r.ell_base = ell_size;
            ell_size += (uint64_t)deg * r.ell_stride;
        }
        wg += (r.count + (lane ? 255u : 3u)) / (lane ? 256u : 4u);
        L->runs.push_back(r);
        p = q;
    }
    L->total_wgs = wg;
    std::vector<uint32_t> ell(std::max<uint64_t>(ell_size, 1), 0u);
    for (const PrRun& r : L->runs) {
        if (r.kind != kPrLane) continue;
        for (uint32_t i = 0; i < r.count; ++i) {
            const uint32_t b = g->h_cptr[order[r.begin + i]];
            for (uint32_t k = 0; k < r.deg; ++k) ell[r.ell_base + (uint64_t)k * r.ell_stride + i] = g->h_csrc[b + k];
        }
    }
    std::vector<uint32_t> wg_run(std::max<uint32_t>(wg, 1), 0u);
    for (uint32_t r = 0; r < (uint32_t)L->runs.size(); ++r) {
        const uint32_t end = r + 1 < L->runs.size() ? L->runs[r + 1].wg_begin : wg;
        for (uint32_t w = L->runs[r].wg_begin; w < end; ++w) wg_run[w] = r;
    }
    hipStream_t s = g->stream;
    if (!L->runs.empty()) L->d_runs.upload(L->runs.data(), L->runs.size(), s);
    L->wg_run.upload(wg_run.data(), wg_run.size(), s);
    if (cnt) L->perm.upload(order.data(), cnt, s);
    L->ell.upload(ell.data(), ell.size(), s);
    MDS_HIP(hipStreamSynchronize(s));
    g->pull = std::move(L);
}

void build_pr_layout(medusa_graph* g) {
    if (g->pr) return;
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    const std::vector<uint32_t>& pb = g->parts;
    const uint32_t G = (uint32_t)pb.size() - 1;
    auto L = std::make_unique<PrLayout>();
    std::vector<uint32_t> perm(n), iperm(n), indeg(n), outdeg(n);
    for (uint32_t v = 0; v < n; ++v) {
        indeg[v] = g->h_cptr[v + 1] - g->h_cptr[v];
        outdeg[v] = g->h_row_ptr[v + 1] - g->h_row_ptr[v];
    }
    std::iota(perm.begin(), perm.end(), 0u);
    for (uint32_t k = 0; k < G; ++k)
        std::sort(perm.begin() + pb[k], perm.begin() + pb[k + 1], [&](uint32_t a, uint32_t b) {
            const bool da = outdeg[a] == 0, db = outdeg[b] == 0;
            if (da != db) return db;  // non-dangling first
            if (indeg[a] != indeg[b]) return indeg[a] > indeg[b];
            return a < b;
        });
    for (uint32_t p = 0; p < n; ++p) iperm[perm[p]] = p;
    std::vector<uint32_t> cptr(n + 1, 0), crow(m), od(n), optr(n + 1, 0), oslot(m);
    for (uint32_t p = 0; p < n; ++p) {
        cptr[p + 1] = cptr[p] + indeg[perm[p]];
        od[p] = outdeg[perm[p]];
        optr[p + 1] = optr[p] + od[p];
    }
    parallel_blocks(n, [&](uint64_t b, uint64_t e) {
        for (uint64_t p = b; p < e; ++p) {
            const uint32_t v = perm[p];
            const uint32_t s0 = g->h_cptr[v], s1 = g->h_cptr[v + 1];
            uint32_t o = cptr[p];
            for (uint32_t s = s0; s < s1; ++s) crow[o++] = iperm[g->h_csrc[s]];
        }
    });
    // push variant: for every out-edge of internal vertex p (in CSR order), its in-edge slot
    parallel_blocks(n, [&](uint64_t b, uint64_t e) {
        for (uint64_t p = b; p < e; ++p) {
            const uint32_t v = perm[p];
            uint32_t o = optr[p];
            for (uint32_t k = g->h_row_ptr[v]; k < g->h_row_ptr[v + 1]; ++k) {
                const uint32_t dst = g->h_col[k];
                // original CSC slot of edge k, re-based to the destination's internal segment
                oslot[o++] = cptr[iperm[dst]] + (g->h_slot[k] - g->h_cptr[dst]);
            }
        }
    });
    // dangling tail of every graph, cut into canonical blocks of 4096
    L->n_dangling = 0;
    for (uint32_t k = 0; k < G; ++k) {
        uint32_t db = pb[k];
        while (db < pb[k + 1] && od[db] != 0) ++db;
        PrPart P{(uint32_t)L->blocks.size(), 0, pb[k + 1] - pb[k], 0};
        for (uint32_t b = db; b < pb[k + 1]; b += 4096) {
            L->blocks.push_back({b, std::min(4096u, pb[k + 1] - b)});
            ++P.nblocks;
        }
        L->n_dangling += pb[k + 1] - db;
        L->parts.push_back(P);
    }
    // runs: equal in-degree for 1..64 (lane per vertex), >64 (wave per vertex), 0
    auto kind_of = [](uint32_t deg) -> uint32_t {
        if (deg == 0) return kPrZero;
        if (deg > kPrLaneMax) return kPrWave;
        return kPrLane;
    };
    auto per_wg = [](uint32_t kind) -> uint32_t { return kind == kPrWave ? 4u : 256u; };
    L->runs.clear();
    uint32_t wg = 0;
    uint64_t ell_size = 0;
    uint32_t part = 0;
    for (uint32_t p = 0; p < n;) {
        while (p >= pb[part + 1]) ++part;
        const uint32_t deg = cptr[p + 1] - cptr[p];
        const uint32_t kind = kind_of(deg);
        const bool dang = od[p] == 0;
        uint32_t q = p + 1;
        while (q < pb[part + 1] && kind_of(cptr[q + 1] - cptr[q]) == kind && (od[q] == 0) == dang &&
               (kind != kPrLane || cptr[q + 1] - cptr[q] == deg))
            ++q;
        PrRun r{};
        r.begin = p;
        r.count = q - p;
        r.kind = kind;
        r.wg_begin = wg;
        r.part = part;
        if (kind == kPrLane) {
            r.deg = deg;
            r.ell_stride = (r.count + 63u) & ~63u;
            r.ell_base = ell_size;
            ell_size += (uint64_t)deg * r.ell_stride;
        }
        wg += (r.count + per_wg(kind) - 1) / per_wg(kind);
        L->runs.push_back(r);
        p = q;
    }
    L->total_wgs = wg;
    // ELL slabs: element k of vertex i of a run at ell_base + k*stride + i (padding = 0)
    std::vector<uint32_t> ell(std::max<uint64_t>(ell_size, 1), 0u);
    for (const PrRun& r : L->runs) {
        if (r.kind != kPrLane) continue;
        for (uint32_t i = 0; i < r.count; ++i) {
            const uint32_t b = cptr[r.begin + i];
            for (uint32_t k = 0; k < r.deg; ++k) ell[r.ell_base + (uint64_t)k * r.ell_stride + i] = crow[b + k];
        }
    }
    hipStream_t s = g->stream;
    L->ell.upload(ell.data(), ell.size(), s);
    L->d_runs.upload(L->runs.data(), L->runs.size(), s);
    std::vector<uint32_t> wg_run(std::max<uint32_t>(L->total_wgs, 1), 0u);
    for (uint32_t r = 0; r < (uint32_t)L->runs.size(); ++r) {
        const uint32_t end = r + 1 < L->runs.size() ? L->runs[r + 1].wg_begin : L->total_wgs;
        for (uint32_t w = L->runs[r].wg_begin; w < end; ++w) wg_run[w] = r;
    }
    L->d_wg_run.upload(wg_run.data(), wg_run.size(), s);
    if (!L->blocks.empty()) L->d_blocks.upload(L->blocks.data(), L->blocks.size(), s);
    L->d_parts.upload(L->parts.data(), L->parts.size(), s);
    L->d_part_begin.upload(pb.data(), pb.size(), s);
    L->perm.upload(perm.data(), n, s);
    L->iperm.upload(iperm.data(), n, s);
    L->cptr.upload(cptr.data(), (size_t)n + 1, s);
    L->crow.upload(crow.data(), m, s);
    L->outdeg.upload(od.data(), n, s);
    L->optr.upload(optr.data(), (size_t)n + 1, s);
    L->oslot.upload(oslot.data(), m, s);
    MDS_HIP(hipStreamSynchronize(s));
    L->h_cptr = std::move(cptr);
    L->h_crow = std::move(crow);
    L->h_od = std::move(od);
    g->pr = std::move(L);
}

// Binned message exchange of the PageRank layout: planned on the host from the internal CSC,
// every index checked against the buffers it addresses, then uploaded.
bool build_pr_bins(medusa_graph* g, bool required) {
    build_pr_layout(g);
    PrLayout& L = *g->pr;
    if (L.bins) return true;
    if (L.bins_refused && !required) return false;
    PrBinsPlan P;
    const std::string err = plan_pr_bins(g->n, L.h_cptr.data(), L.h_crow.data(), g->parts.data(),
                                         (uint32_t)g->parts.size() - 1, P,
                                         pr_send_tile_slots(g, (uint32_t)L.blocks.size()), pr_combine_slots(g));
    if (!err.empty()) {
        L.bins_refused = true;
        MDS_CHECK(!required, MEDUSA_ERR_STATE, "%s", err.c_str());
        return false;
    }
    const std::string bad = check_pr_bins(g->n, L.h_cptr.data(), P);
    MDS_CHECK(bad.empty(), MEDUSA_ERR_INTERNAL, "binned PageRank plan: %s", bad.c_str());
    auto B = std::make_unique<PrBins>();
    hipStream_t s = g->stream;
    B->nbins = (uint32_t)P.bins.size();
    B->nhubs = P.nhubs;
    B->ntiles = (uint32_t)P.tiles.size();
    B->nmsg = P.nmsg;
    B->nentry = P.nentry;
    B->items.upload(P.items.data(), P.items.size(), s);
    B->bins.upload(P.bins.data(), P.bins.size(), s);
    B->tiles.upload(P.tiles.data(), P.tiles.size(), s);
    if (P.nentry) {
        B->srcloc.upload(P.srcloc.data(), P.nentry, s);
        B->slot.upload(P.slot.data(), P.nentry, s);
        B->qmsg.upload(P.qmsg.data(), P.qmsg.size(), s);
    } else {  // no message: every in-edge goes to a hub (or there are none)
        B->srcloc.alloc(1);
        B->slot.alloc(1);
        B->qmsg.alloc(1);
    }
    B->msg.alloc(std::max<uint64_t>(P.nentry, 1));
    MDS_HIP(hipStreamSynchronize(s));
    L.bins = std::move(B);
    return true;
}

}  // namespace medusa

using namespace medusa;

extern "C" const char* medusa_last_error(void) { return t_last_error.c_str(); }

extern "C" const char* medusa_version(void) { return "medusa-mi355x 0.1 (gfx950)"; }

extern "C" int medusa_device_count(int* count) {
    return guard([&] {
        MDS_CHECK(count, MEDUSA_ERR_ARG, "null count");
        int c = 0;
        if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
        (void)hipGetLastError();
        *count = c;
    });
}

extern "C" int medusa_graph_create(uint32_t n, uint64_t m, const uint32_t* row_ptr, const uint32_t* col,
                                   const uint32_t* weight, medusa_graph** out) {
    return guard([&] {
        MDS_CHECK(out, MEDUSA_ERR_ARG, "graph_create: null out");
        *out = nullptr;
        MDS_CHECK(n > 0, MEDUSA_ERR_ARG, "graph_create: n must be > 0");
        MDS_CHECK(m < (1ull << 32), MEDUSA_ERR_ARG, "graph_create: m must be < 2^32");
        MDS_CHECK(row_ptr && (m == 0 || col), MEDUSA_ERR_ARG, "graph_create: null CSR array");
        MDS_CHECK(row_ptr[0] == 0 && row_ptr[n] == m, MEDUSA_ERR_ARG,
                  "graph_create: row_ptr[0] must be 0 and row_ptr[n] must equal m");
        for (uint32_t v = 0; v < n; ++v)
            MDS_CHECK(row_ptr[v] <= row_ptr[v + 1], MEDUSA_ERR_ARG, "graph_create: row_ptr decreases at %u", v);
        for (uint64_t e = 0; e < m; ++e)
            MDS_CHECK(col[e] < n, MEDUSA_ERR_ARG, "graph_create: col[%llu]=%u out of range", (unsigned long long)e,
                      col[e]);
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess) ndev = 0;
        (void)hipGetLastError();
        MDS_CHECK(ndev > 0, MEDUSA_ERR_HIP, "graph_create: no GPU visible");

        auto g = std::make_unique<medusa_graph>();
        g->n = n;
        g->m = m;
        g->weighted = weight != nullptr;
        MDS_HIP(hipGetDevice(&g->device));
        MDS_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
        MDS_HIP(hipEventCreate(&g->ev_begin));
        MDS_HIP(hipEventCreate(&g->ev_end));
        g->parts = {0u, n};
        g->h_row_ptr.assign(row_ptr, row_ptr + n + 1);
        g->h_col.assign(col, col + m);
        if (weight) g->h_w.assign(weight, weight + m);
        // CSC: sources ascending within each destination (a stable pass over CSR);
        // slot[e] = CSC position of CSR edge e (the reference's msgDstID).
        g->h_cptr.assign((size_t)n + 1, 0);
        for (uint64_t e = 0; e < m; ++e) ++g->h_cptr[col[e] + 1];
        for (uint32_t v = 0; v < n; ++v) g->h_cptr[v + 1] += g->h_cptr[v];
        g->h_csrc.resize(m);
        g->h_slot.resize(m);
        if (weight) g->h_cw.resize(m);
        std::vector<uint32_t> fill(g->h_cptr.begin(), g->h_cptr.end() - 1);
        for (uint32_t u = 0; u < n; ++u)
            for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; ++e) {
                const uint32_t pos = fill[col[e]]++;
                g->h_csrc[pos] = u;
                g->h_slot[e] = pos;
                if (weight) g->h_cw[pos] = weight[e];
            }
        hipStream_t s = g->stream;
        g->row_ptr.upload(g->h_row_ptr.data(), (size_t)n + 1, s);
        g->col.upload(g->h_col.data(), m, s);
        g->cptr.upload(g->h_cptr.data(), (size_t)n + 1, s);
        g->csrc.upload(g->h_csrc.data(), m, s);
        g->slot.upload(g->h_slot.data(), m, s);
        if (weight) {
            g->w.upload(g->h_w.data(), m, s);
            g->cw.upload(g->h_cw.data(), m, s);
        }
        g->ctr_raw.alloc(kCounterWords);
        g->h_ctr_raw.ensure(kCounterWords);
        MDS_HIP(hipStreamSynchronize(s));
        *out = g.release();
    });
}

extern "C" int medusa_graph_set_batch(medusa_graph* g, uint32_t count, const uint32_t* graph_begin) {
    return guard([&] {
        const auto lock = enter(g);
        MDS_CHECK(count >= 1 && graph_begin, MEDUSA_ERR_ARG, "set_batch: need count >= 1 and graph_begin");
        MDS_CHECK(graph_begin[0] == 0 && graph_begin[count] == g->n, MEDUSA_ERR_ARG,
                  "set_batch: graph_begin must start at 0 and end at n=%u", g->n);
        for (uint32_t k = 0; k < count; ++k) {
            const uint32_t b = graph_begin[k], e = graph_begin[k + 1];
            MDS_CHECK(b < e, MEDUSA_ERR_ARG, "set_batch: graph %u is empty or out of order", k);
            for (uint64_t x = g->h_row_ptr[b]; x < g->h_row_ptr[e]; ++x)
                MDS_CHECK(g->h_col[x] >= b && g->h_col[x] < e, MEDUSA_ERR_ARG,
                          "set_batch: edge %llu leaves graph %u", (unsigned long long)x, k);
        }
        g->parts.assign(graph_begin, graph_begin + count + 1);
        g->pr.reset();  // the PageRank layout depends on the graph ranges
    });
}

extern "C" int medusa_graph_batch(const medusa_graph* g, uint32_t* count, uint32_t* graph_begin,
                                  uint32_t capacity) {
    return guard([&] {
        MDS_CHECK(g && count, MEDUSA_ERR_ARG, "null argument");
        const std::lock_guard<std::recursive_mutex> lock(g->mu);
        *count = (uint32_t)g->parts.size() - 1;
        MDS_CHECK(graph_begin || capacity == 0, MEDUSA_ERR_ARG, "null graph_begin");
        const size_t k = std::min<size_t>(capacity, g->parts.size());
        std::copy(g->parts.begin(), g->parts.begin() + k, graph_begin);
    });
}

extern "C" int medusa_graph_destroy(medusa_graph* g) {
    return guard([&] {
        if (!g) return;
        (void)hipSetDevice(g->device);
        if (g->stream) (void)hipStreamSynchronize(g->stream);
        for (hipEvent_t e : g->events) (void)hipEventDestroy(e);
        if (g->ev_begin) (void)hipEventDestroy(g->ev_begin);
        if (g->ev_end) (void)hipEventDestroy(g->ev_end);
        hipStream_t s = g->stream;
        delete g;  // DBuf destructors free device memory
        if (s) (void)hipStreamDestroy(s);
    });
}

extern "C" int medusa_graph_info(const medusa_graph* g, uint32_t* n, uint64_t* m, uint32_t* weighted) {
    return guard([&] {
        MDS_CHECK(g, MEDUSA_ERR_ARG, "null graph");
        if (n) *n = g->n;
        if (m) *m = g->m;
        if (weighted) *weighted = g->weighted ? 1u : 0u;
    });
}

extern "C" int medusa_graph_device_arrays(const medusa_graph* g, const uint32_t** row_ptr, const uint32_t** col,
                                          const uint32_t** weight, const uint32_t** csc_ptr,
                                          const uint32_t** csc_src, const uint32_t** msg_slot) {
    return guard([&] {
        MDS_CHECK(g, MEDUSA_ERR_ARG, "null graph");
        if (row_ptr) *row_ptr = g->row_ptr.p;
        if (col) *col = g->col.p;
        if (weight) *weight = g->w.p;
        if (csc_ptr) *csc_ptr = g->cptr.p;
        if (csc_src) *csc_src = g->csrc.p;
        if (msg_slot) *msg_slot = g->slot.p;
    });
}

extern "C" int medusa_trace_count(const medusa_graph* g, uint32_t* count) {
    return guard([&] {
        MDS_CHECK(g && count, MEDUSA_ERR_ARG, "null argument");
        const std::lock_guard<std::recursive_mutex> lock(g->mu);
        *count = (uint32_t)g->trace.size();
    });
}

extern "C" int medusa_trace_read(const medusa_graph* g, double* out, uint32_t capacity) {
    return guard([&] {
        MDS_CHECK(g && (out || capacity == 0), MEDUSA_ERR_ARG, "null argument");
        const std::lock_guard<std::recursive_mutex> lock(g->mu);
        const size_t k = std::min<size_t>(capacity, g->trace.size());
        for (size_t i = 0; i < k; ++i)
            for (int f = 0; f < MEDUSA_TRACE_FIELDS; ++f) out[i * MEDUSA_TRACE_FIELDS + f] = g->trace[i].f[f];
    });
}
