// Per-superstep scheduling: frontier features, the plain engine's heuristics,
// the learned policy hook, and the per-superstep trace that trains it.
#pragma once

#include "engine.h"

namespace medusa {

// Frontier features at the start of a superstep.
struct Frontier {
    uint32_t step = 0;
    uint32_t n = 0;
    uint64_t m = 0;
    uint64_t n_f = 0;  // frontier vertices
    uint64_t m_f = 0;  // their out-edges
    uint64_t n_u = 0;  // unvisited vertices
    uint64_t m_u = 0;  // their out-edges
};

struct Sched {
    enum Algo { kBfs, kSssp, kMbfs };
    // Direction-optimising thresholds (Beamer, Asanovic, Patterson, SC'12).
    static constexpr double kAlpha = 14.0, kBeta = 24.0;
    const medusa_opts& o;
    Algo algo;
    Sched(const medusa_opts& opts, Algo a) : o(opts), algo(a) {}

    // Plain engine: switch to pull when the frontier's out-edges exceed the
    // unvisited out-edges / alpha; back to push when the frontier shrinks below n / beta.
    bool heuristic_pull(const Frontier& f, bool cur_pull) const {
        if (!cur_pull) return (double)f.m_f > (double)f.m_u / kAlpha;
        return !((double)f.n_f < (double)f.n / kBeta);
    }
    bool bfs_pull(const Frontier& f, bool cur_pull) const { return heuristic_pull(f, cur_pull); }
    // SSSP bucket width for the next threshold advance.
    uint32_t sssp_delta(const Frontier& f, uint32_t dflt) const {
        (void)f;
        return dflt;
    }

    // Appends one trace record when tracing is on.
    void record(medusa_graph* g, const Frontier& f, double action, double kernel_ms) const {
        if (!o.trace) return;
        TraceRec r;
        r.f[0] = f.step;
        r.f[1] = (double)f.n_f;
        r.f[2] = (double)f.m_f;
        r.f[3] = (double)f.n_u;
        r.f[4] = (double)f.m_u;
        r.f[5] = f.n;
        r.f[6] = (double)f.m;
        r.f[7] = action;
        r.f[8] = kernel_ms;
        g->trace.push_back(r);
    }
};

}  // namespace medusa
