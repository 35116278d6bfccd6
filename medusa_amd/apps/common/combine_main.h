// Message combiners through Medusa's user API: in one superstep every edge sends its weight
// to its destination, MGLOBAL::com combines each vertex's messages with the operator given
// on the command line, and a vertex functor stores the result.  Exercises the combiner
// (MedusaRT/Combiner.cu:21-142; CUDPP ADD/MULTIPLY/MIN/MAX in the reference) for the
// message type the including app defines as MVT.
// Usage: <app> <graph with weights> <out> <add|min|max|mul>
#include <cstdlib>
#include <cstring>

struct SendWeight {
    __device__ void operator()(D_Edge edge, int) { edge.sendMsg(edge.get_weight()); }
};

struct StoreCombined {
    __device__ void operator()(D_Vertex v, int) { v.set_result(v.get_combined_msg()); }
};

inline MedusaCombinerOp ParseOp(const char* s) {
    if (!std::strcmp(s, "add")) return MEDUSA_COMBINE_ADD;
    if (!std::strcmp(s, "min")) return MEDUSA_COMBINE_MIN;
    if (!std::strcmp(s, "max")) return MEDUSA_COMBINE_MAX;
    if (!std::strcmp(s, "mul")) return MEDUSA_COMBINE_MUL;
    throw std::runtime_error(std::string("unknown combiner ") + s);
}

inline int CombineMain(int argc, char** argv, MedusaCombinerType type, const char* name) {
    if (argc < 4) {
        fprintf(stderr, "usage: %s <graph> <out> <add|min|max|mul>\n", argv[0]);
        return 2;
    }
    try {
        InitConfig("1", "0", EDGE_MESSAGE, false);
        MGLOBAL::com.init(type, ParseOp(argv[3]));
        GraphPartition gp[1];
        LoadGraph(argv[1], gp);
        InitHostDS(gp);
        VertexArray& V = MGLOBAL::gpu_def[0].vertexArray;
        for (int v = 0; v < V.size; ++v) V.result[v] = 0;
        InitDeviceDS();
        ResetToExecute();
        Message none;
        none.val = (MVT)-1;  // vertices without in-edges read this
        InitMessageBuffer(none);
        FunctorHolder<EDGE>::Run(SendWeight());
        MGLOBAL::com.combineAllDevice();
        FunctorHolder<VERTEX>::Run(StoreCombined());
        (void)RetriveToExecute();
        DumpVertexData();
        WriteArray(argv[2], V.result, V.size);
        printf("%s %s n=%d\n", name, argv[3], V.size);
        ReleaseDeviceDS();
    } catch (const std::exception& e) {
        fprintf(stderr, "%s: %s\n", name, e.what());
        return 1;
    }
    return 0;
}
