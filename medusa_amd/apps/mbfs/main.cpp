// Multi-source BFS through Medusa's user API.
// Usage: mbfs_app <graph> <levels.out> <src0> [src1 ... src63]; output uint8 [k][n].
#include "Message.h"
#include "Vertex.h"
#include "Edge.h"
#include <medusa/medusa.h>
#include "../common/app_io.h"
#include "UserOperators.h"

int main(int argc, char** argv) {
    if (argc < 4 || argc > 3 + 64) {
        fprintf(stderr, "usage: %s <graph> <levels.out> <src0> [... up to 64 sources]\n", argv[0]);
        return 2;
    }
    try {
        InitConfig("1", "0", NO_MESSAGE, false);
        GraphPartition gp[1];
        LoadGraph(argv[1], gp);
        InitHostDS(gp);
        VertexArray& V = MGLOBAL::gpu_def[0].vertexArray;
        const int n = V.size, k = argc - 3;
        std::vector<unsigned char> levels((size_t)k * n, 0xFF);
        for (int s = 0; s < k; ++s) {
            const int v = atoi(argv[3 + s]);
            if (v < 0 || v >= n) throw std::runtime_error("source out of range");
            V.frontier[v] |= 1ull << s;
            V.visited[v] |= 1ull << s;
            levels[(size_t)s * n + v] = 0;
        }
        unsigned char* d_levels = medusa_api::dev_alloc_copy(levels.data(), levels.size());
        MEDUSA_HIP_OK(hipMemcpyToSymbol(HIP_SYMBOL(g_levels), &d_levels, sizeof(d_levels)));
        MEDUSA_HIP_OK(hipMemcpyToSymbol(HIP_SYMBOL(g_n), &n, sizeof(int)));
        InitDeviceDS();
        const double ms = RunLoop(Medusa_Exec, INT_MAX, true);
        MEDUSA_HIP_OK(hipMemcpy(levels.data(), d_levels, levels.size(), hipMemcpyDeviceToHost));
        WriteArray(argv[2], levels.data(), levels.size());
        printf("mbfs sources=%d supersteps=%d loop_ms=%.4f\n", k, MGLOBAL::super_step, ms);
        (void)hipFree(d_levels);
        ReleaseDeviceDS();
    } catch (const std::exception& e) {
        fprintf(stderr, "mbfs_app: %s\n", e.what());
        return 1;
    }
    return 0;
}
