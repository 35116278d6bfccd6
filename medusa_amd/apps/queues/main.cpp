// The queue operations of Medusa's user API (QAPIType, MultipleGPU/MultiAPIHoster.h:18-23;
// EdgeQueue2EdgeQueue, MedusaRT/APIKernelWrappers.h:53-140).
//   BFS:       seed an edge queue with the source's out-edges (VQ2EQ), then EQ2EQ with a
//              functor that claims the destination's level, until the queue is empty;
//   then every other operation once on the result (weights below `light` are "light"):
//   VA2AQ reached vertices; VQ2VQ those at even levels; VQConsume marks them; VA2EQ out-edges
//   of level-1 vertices; EA2AQ destinations of light edges; EA2EQ out-edges of those
//   destinations; EQ2VQ destinations of the odd-weight edges among them; EQConsume counts
//   them per source in `hits`.
// Usage: queues_app <graph with weights> <out> [source] [light]
// Output (int32): sections [count, items...] in the order above: levels[n], reached,
// even, mark[n], va2eq, ea2aq, ea2eq, eq2vq, hits[n], then [1, supersteps].
#include "Message.h"
#include "Vertex.h"
#include "Edge.h"
#include <medusa/medusa.h>
#include "../common/app_io.h"

struct Always {
    __device__ bool operator()(D_Vertex, int) const { return true; }
};
struct ClaimDst {
    __device__ bool operator()(D_Edge e, int step) const {
        return atomicCAS(D_Vertex(e.get_dstVertexID()).ptr_level(), -1, step + 1) == -1;
    }
};
struct Reached {
    __device__ bool operator()(D_Vertex v, int) const { return v.get_level() >= 0; }
};
struct EvenLevel {
    __device__ bool operator()(D_Vertex v, int) const { return (v.get_level() & 1) == 0; }
};
struct Mark {
    __device__ void operator()(D_Vertex v, int) const { v.set_mark(1); }
};
struct LevelOne {
    __device__ bool operator()(D_Vertex v, int) const { return v.get_level() == 1; }
};
struct Light {
    int below;
    __device__ bool operator()(D_Edge e, int) const { return e.get_weight() < below; }
};
struct OddWeight {
    __device__ bool operator()(D_Edge e, int) const { return (e.get_weight() & 1) != 0; }
};
struct CountHits {
    __device__ void operator()(D_Edge e, int) const { atomicAdd(D_Vertex(e.get_srcVertexID()).ptr_hits(), 1); }
};

static void put(std::vector<int>& out, const std::vector<int>& items) {
    out.push_back((int)items.size());
    out.insert(out.end(), items.begin(), items.end());
}

int main(int argc, char** argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s <graph> <out> [source] [light]\n", argv[0]);
        return 2;
    }
    try {
        InitConfig("1", "0", NO_MESSAGE, false);
        GraphPartition gp[1];
        LoadGraph(argv[1], gp);
        InitHostDS(gp);
        VertexArray& V = MGLOBAL::gpu_def[0].vertexArray;
        const int src = argc > 3 ? atoi(argv[3]) : MaxOutDegreeVertex(V);
        const int light = argc > 4 ? atoi(argv[4]) : 3;
        if (src < 0 || src >= V.size) throw std::runtime_error("source out of range");
        for (int v = 0; v < V.size; ++v) {
            V.level[v] = v == src ? 0 : -1;
            V.mark[v] = 0;
            V.hits[v] = 0;
        }
        InitDeviceDS();
        MedusaQueue vq = CreateVertexQueue(), vq2 = CreateVertexQueue();
        // EA2EQ below queues the out-edges of a destination once per light edge into it
        const long long ecap = 4LL * MGLOBAL::gpu_def[0].edgeArray.size + 64;
        MedusaQueue eq = CreateEdgeQueue(ecap), eq2 = CreateEdgeQueue(ecap);
        QueueAssign(vq, &src, 1);
        QFunctorHolder<VQ2EQ>::Run(Always(), vq, eq);
        int steps = 0;
        for (MGLOBAL::super_step = 0; QueueSize(eq) > 0; ++MGLOBAL::super_step, ++steps) {
            QFunctorHolder<EQ2EQ>::Run(ClaimDst(), eq, eq2);
            std::swap(eq, eq2);
        }
        std::vector<int> out;
        QFunctorHolder<VA2AQ>::Run(Reached(), vq);
        const std::vector<int> reached = QueueItems(vq);
        QFunctorHolder<VQ2VQ>::Run(EvenLevel(), vq, vq2);
        const std::vector<int> even = QueueItems(vq2);
        QFunctorHolder<VQConsume>::Run(Mark(), vq2);
        QFunctorHolder<VA2EQ>::Run(LevelOne(), eq);
        const std::vector<int> va2eq = QueueItems(eq);
        QFunctorHolder<EA2AQ>::Run(Light{light}, vq);
        const std::vector<int> ea2aq = QueueItems(vq);
        QFunctorHolder<EA2EQ>::Run(Light{light}, eq);
        const std::vector<int> ea2eq = QueueItems(eq);
        QFunctorHolder<EQ2VQ>::Run(OddWeight(), eq, vq2);
        const std::vector<int> eq2vq = QueueItems(vq2);
        QFunctorHolder<EQConsume>::Run(CountHits(), eq);
        DumpVertexData();
        put(out, std::vector<int>(V.level, V.level + V.size));
        put(out, reached);
        put(out, even);
        put(out, std::vector<int>(V.mark, V.mark + V.size));
        put(out, va2eq);
        put(out, ea2aq);
        put(out, ea2eq);
        put(out, eq2vq);
        put(out, std::vector<int>(V.hits, V.hits + V.size));
        put(out, {steps});
        WriteArray(argv[2], out.data(), out.size());
        printf("queues n=%d source=%d supersteps=%d reached=%zu\n", V.size, src, steps, reached.size());
        FreeQueue(vq);
        FreeQueue(vq2);
        FreeQueue(eq);
        FreeQueue(eq2);
        ReleaseDeviceDS();
    } catch (const std::exception& e) {
        fprintf(stderr, "queues_app: %s\n", e.what());
        return 1;
    }
    return 0;
}
