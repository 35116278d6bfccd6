// BFS through Medusa's user API.  Usage: bfs_app <graph> <levels.out> [root]
// Without a root, the first vertex of maximum out-degree is used.
#include "Message.h"
#include "Vertex.h"
#include "Edge.h"
#include <medusa/medusa.h>
#include "../common/app_io.h"
#include "UserOperators.h"

int main(int argc, char** argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s <graph> <levels.out> [root]\n", argv[0]);
        return 2;
    }
    try {
        InitConfig("1", "0", NO_MESSAGE, false);
        GraphPartition gp[1];
        LoadGraph(argv[1], gp);
        InitHostDS(gp);
        VertexArray& V = MGLOBAL::gpu_def[0].vertexArray;
        const int root = argc > 3 ? atoi(argv[3]) : MaxOutDegreeVertex(V);
        for (int v = 0; v < V.size; ++v) V.level[v] = MVT_Init_Value;
        V.level[root] = 0;
        InitDeviceDS();
        const double ms = RunLoop(Medusa_Exec, INT_MAX, true);
        DumpVertexData();
        WriteArray(argv[2], V.level, V.size);
        printf("bfs root=%d supersteps=%d loop_ms=%.4f\n", root, MGLOBAL::super_step, ms);
        ReleaseDeviceDS();
    } catch (const std::exception& e) {
        fprintf(stderr, "bfs_app: %s\n", e.what());
        return 1;
    }
    return 0;
}
