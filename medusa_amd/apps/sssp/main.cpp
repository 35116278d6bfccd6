// SSSP through Medusa's user API.  Usage: sssp_app <graph> <dist.out> [root]
// Output: float distances, FLT_MAX for unreached vertices.
#include "Message.h"
#include "Vertex.h"
#include "Edge.h"
#include <medusa/medusa.h>
#include "../common/app_io.h"
#include "UserOperators.h"

int main(int argc, char** argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s <graph> <dist.out> [root]\n", argv[0]);
        return 2;
    }
    try {
        InitConfig("1", "0", EDGE_MESSAGE, false);
        MGLOBAL::com.init(MEDUSA_COMBINE_FLOAT, MEDUSA_COMBINE_MIN);
        GraphPartition gp[1];
        LoadGraph(argv[1], gp);
        InitHostDS(gp);
        VertexArray& V = MGLOBAL::gpu_def[0].vertexArray;
        const int root = argc > 3 ? atoi(argv[3]) : MaxOutDegreeVertex(V);
        for (int v = 0; v < V.size; ++v) {
            V.distance[v] = FLT_MAX;
            V.updated[v] = false;
        }
        V.distance[root] = 0;
        V.updated[root] = true;
        InitDeviceDS();
        const double ms = RunLoop(Medusa_Exec, INT_MAX, true);
        DumpVertexData();
        WriteArray(argv[2], V.distance, V.size);
        printf("sssp root=%d supersteps=%d loop_ms=%.4f\n", root, MGLOBAL::super_step, ms);
        ReleaseDeviceDS();
    } catch (const std::exception& e) {
        fprintf(stderr, "sssp_app: %s\n", e.what());
        return 1;
    }
    return 0;
}
