// Autotuner for the frontier engines' launch shape.
//
// The reference fixes its launch shape (grid 6*SM x 256 grid-stride, MultipleGPU/
// MultiAPIHoster.h:649-754, MedusaRT/APIKernelWrappers.h:8-218) and its hybrid ELL
// threshold (16, MultipleGPU/MultiPublicAPI.h:82).  Here the two parameters that decide
// how a frontier is spread over the GPU, the out-degree above which a vertex goes to the
// workgroup-per-hub queue and the number of workgroups of the expansion launch, are
// chosen per graph by timing the candidates on the caller's sources.  The choice never
// changes results (every candidate runs the same supersteps), only their duration.

#include <algorithm>
#include <vector>

#include "engine.h"

namespace medusa {

static const uint32_t kTuneLarge[] = {128, 256, 512, 1024, 2048, 4096, 16384};
static const uint32_t kTuneGrid[] = {512, 1024, 2048, 4096, 8192};

}  // namespace medusa

using namespace medusa;

extern "C" int medusa_bfs_multi(medusa_graph* g, const uint32_t* sources, uint32_t count, int32_t* levels,
                                const medusa_opts* opts, medusa_stats* stats);
extern "C" int medusa_mbfs(medusa_graph* g, const uint32_t* sources, uint32_t k, uint8_t* levels,
                           const medusa_opts* opts, medusa_stats* stats);

extern "C" int medusa_autotune(medusa_graph* g, uint32_t algo, const uint32_t* sources, uint32_t count,
                               uint32_t reps, medusa_tuning* result) {
    return guard([&] {
        const auto lock = enter(g);
        MDS_CHECK(algo == MEDUSA_TUNE_BFS || algo == MEDUSA_TUNE_MBFS, MEDUSA_ERR_ARG, "autotune: unknown algo %u",
                  algo);
        if (reps == 0) {  // forget the tuned shape
            if (algo == MEDUSA_TUNE_BFS) g->tuned.bfs_large = g->tuned.bfs_grid = 0;
            else g->tuned.mbfs_large = g->tuned.mbfs_grid = 0;
            if (result) *result = medusa_tuning{};
            return;
        }
        MDS_CHECK(sources && count >= 1, MEDUSA_ERR_ARG, "autotune: need sources");
        MDS_CHECK(algo != MEDUSA_TUNE_MBFS || count <= 64, MEDUSA_ERR_ARG, "autotune: MBFS takes at most 64 sources");
        const size_t out_bytes = algo == MEDUSA_TUNE_BFS ? (size_t)g->n * 4 : (size_t)g->n * count;
        DBuf<uint32_t> out;
        out.alloc((out_bytes + 3) / 4);
        auto run = [&](uint32_t large, uint32_t grid) -> double {
            medusa_opts o{};
            o.large_degree = large;
            o.grid = grid;
            double best = 1e300;
            for (uint32_t r = 0; r < reps + 1; ++r) {  // the first run captures the superstep graph
                medusa_stats st{};
                const int rc = algo == MEDUSA_TUNE_BFS
                                   ? medusa_bfs_multi(g, sources, count, reinterpret_cast<int32_t*>(out.p), &o, &st)
                                   : medusa_mbfs(g, sources, count, reinterpret_cast<uint8_t*>(out.p), &o, &st);
                if (rc != 0) throw Error(rc, medusa_last_error());
                if (r > 0) best = std::min(best, st.device_ms);
            }
            return best;
        };
        const uint32_t dl = algo == MEDUSA_TUNE_BFS ? 2048u : 1024u, dg = 2048u;  // engine defaults
        medusa_tuning t{};
        t.default_ms = run(dl, dg);
        t.best_ms = t.default_ms;
        t.large_degree = dl;
        t.grid = dg;
        for (uint32_t large : kTuneLarge)
            for (uint32_t grid : kTuneGrid) {
                if (large == dl && grid == dg) continue;
                const double ms = run(large, grid);
                ++t.candidates;
                if (ms < t.best_ms) {
                    t.best_ms = ms;
                    t.large_degree = large;
                    t.grid = grid;
                }
            }
        ++t.candidates;
        if (algo == MEDUSA_TUNE_BFS) {
            g->tuned.bfs_large = t.large_degree;
            g->tuned.bfs_grid = t.grid;
        } else {
            g->tuned.mbfs_large = t.large_degree;
            g->tuned.mbfs_grid = t.grid;
        }
        if (result) *result = t;
    });
}
