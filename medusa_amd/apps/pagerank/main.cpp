// PageRank through Medusa's user API.  Usage: pagerank_app <graph> <rank.out> [supersteps=100]
#include "Message.h"
#include "Vertex.h"
#include "Edge.h"
#include <medusa/medusa.h>
#include "../common/app_io.h"
#include "UserOperators.h"

int main(int argc, char** argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s <graph> <rank.out> [supersteps]\n", argv[0]);
        return 2;
    }
    try {
        const int steps = argc > 3 ? atoi(argv[3]) : 100;
        InitConfig("1", "0", EDGE_MESSAGE, true);
        MGLOBAL::com.init(MEDUSA_COMBINE_FLOAT, MEDUSA_COMBINE_ADD);
        GraphPartition gp[1];
        LoadGraph(argv[1], gp);
        InitHostDS(gp);
        VertexArray& V = MGLOBAL::gpu_def[0].vertexArray;
        for (int v = 0; v < V.size; ++v) V.level[v] = 1.0f;
        InitDeviceDS();
        const double ms = RunLoop(Medusa_Exec, steps, false);
        DumpVertexData();
        WriteArray(argv[2], V.level, V.size);
        printf("pagerank supersteps=%d loop_ms=%.4f\n", MGLOBAL::super_step, ms);
        ReleaseDeviceDS();
    } catch (const std::exception& e) {
        fprintf(stderr, "pagerank_app: %s\n", e.what());
        return 1;
    }
    return 0;
}
