// Graph input and result output for the example applications.
//
// Graph file: either a GTgraph text file ("p sp N M" / "a u v w", 1-based, read by
// GTGraph::get_parition) or a binary CSR file starting with the magic "MDSCSR01":
//   u32 n, u64 m, u32 has_weight, u32 row_ptr[n+1], u32 col[m], u32 weight[m] (if has_weight).
// Binary arcs are added in CSR order to the GraphIR.
#pragma once

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

inline void LoadGraph(const char* path, GraphPartition* gp) {
    std::ifstream f(path, std::ios::binary);
    if (!f.good()) throw std::runtime_error(std::string("cannot open ") + path);
    char magic[8] = {0};
    f.read(magic, 8);
    if (f.gcount() == 8 && std::memcmp(magic, "MDSCSR01", 8) == 0) {
        uint32_t n = 0, has_w = 0;
        uint64_t m = 0;
        f.read(reinterpret_cast<char*>(&n), 4);
        f.read(reinterpret_cast<char*>(&m), 8);
        f.read(reinterpret_cast<char*>(&has_w), 4);
        std::vector<uint32_t> rp(n + 1), col(m), w(has_w ? m : 0);
        f.read(reinterpret_cast<char*>(rp.data()), (n + 1) * 4ull);
        f.read(reinterpret_cast<char*>(col.data()), m * 4ull);
        if (has_w) f.read(reinterpret_cast<char*>(w.data()), m * 4ull);
        if (!f) throw std::runtime_error("truncated CSR file");
        gp[0].graph.resize((int)n);
        gp[0].l2p.resize(n);
        for (uint32_t v = 0; v < n; ++v) gp[0].l2p[v] = (int)v;
        // AddEdge prepends per source, so add each source's arcs in reverse to keep CSR order
        for (uint32_t u = 0; u < n; ++u)
            for (uint32_t k = rp[u + 1]; k-- > rp[u];) {
                Edge e;
                e.srcVertexID = (int)u;
                e.dstVertexID = (int)col[k];
                e.weight = has_w ? (double)w[k] : 0.0;
                gp[0].graph.AddEdge(e);
            }
        return;
    }
    GTGraph gt;
    gt.get_parition(gp, path, nullptr);
}

template <class T>
inline void WriteArray(const char* path, const T* data, size_t count) {
    FILE* f = std::fopen(path, "wb");
    if (!f) throw std::runtime_error(std::string("cannot write ") + path);
    std::fwrite(data, sizeof(T), count, f);
    std::fclose(f);
}

// The reference's BFS/SSSP root: the first vertex of maximum out-degree (Algorithm/main.cu:209-212).
inline int MaxOutDegreeVertex(const VertexArray& V) {
    int best = 0;
    for (int v = 1; v < V.size; ++v)
        if (V.edge_count[v] > V.edge_count[best]) best = v;
    return best;
}

// The application's BSP loop over `exec` (its Medusa_Exec), from superstep 0 and the initial
// data InitDeviceDS uploaded: on the device with Medusa_Run, or, with MEDUSA_HOST_LOOP=1 in
// the environment, the reference's host loop (ResetToExecute / exec / RetriveToExecute per
// superstep).  `vote`: the functors' vote ends the run; otherwise exactly `max_supersteps`
// run.  With MEDUSA_APP_REPEAT=r the run is repeated r times, the data re-uploaded before
// each; returns the median milliseconds of the runs after the first (which also captures the
// device loop), or of the one run (the GPU idle before and after each).
inline double RunLoop(void (*exec)(), int max_supersteps, bool vote) {
    const char* hl = std::getenv("MEDUSA_HOST_LOOP");
    const char* rp = std::getenv("MEDUSA_APP_REPEAT");
    const bool host = hl && std::atoi(hl) == 1;
    const int reps = rp && std::atoi(rp) > 1 ? std::atoi(rp) : 1;
    std::vector<double> times;
    for (int r = 0; r < reps; ++r) {
        if (r) ResetGraphData();
        MGLOBAL::super_step = 0;
        MEDUSA_HIP_OK(hipDeviceSynchronize());
        const auto t0 = std::chrono::steady_clock::now();
        if (host) {
            bool more = true;
            while (more && MGLOBAL::super_step < max_supersteps) {
                if (vote) ResetToExecute();
                exec();
                more = vote ? RetriveToExecute() : true;
                ++MGLOBAL::super_step;
            }
        } else {
            Medusa_Run(exec, max_supersteps, vote);
        }
        MEDUSA_HIP_OK(hipDeviceSynchronize());
        times.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
    if (times.size() > 1) times.erase(times.begin());
    std::sort(times.begin(), times.end());
    return times[times.size() / 2];
}
