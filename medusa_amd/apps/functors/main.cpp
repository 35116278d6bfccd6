// The EDGELIST and MESSAGE functor paths of Medusa's user API
// (FunctorHolder<EDGELIST/MESSAGE>::Run, MultipleGPU/MultiAPIHoster.h:706-735).  One superstep:
//   EDGELIST: every vertex walks its out-edges, sums their weights into `wsum` and sends
//             2*weight along each edge;
//   MESSAGE:  every message slot adds 1 to its value;
//   combine (ADD) and VERTEX: each vertex stores the sum of its incoming messages in `result`
//             (the init value -1 without in-edges).
// Usage: functors_app <graph with weights> <out>; writes wsum[n] then result[n] (int32).
#include "Message.h"
#include "Vertex.h"
#include "Edge.h"
#include <medusa/medusa.h>
#include "../common/app_io.h"

struct WalkOutEdges {
    __device__ void operator()(D_Vertex v, EdgeList& el, int) {
        MVT s = 0;
        const int deg = el.getEdgeCount();
        for (int i = 0; i < deg; ++i) {
            const D_Edge e = el.getNextEdge();
            s += e.get_weight();
            e.sendMsg(2 * e.get_weight());
        }
        v.set_wsum(s);
    }
};

struct BumpMessage {
    __device__ void operator()(D_Message msg, int) { msg.set_val(msg.get_val() + 1); }
};

struct StoreCombined {
    __device__ void operator()(D_Vertex v, int) { v.set_result(v.get_combined_msg()); }
};

int main(int argc, char** argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s <graph> <out>\n", argv[0]);
        return 2;
    }
    try {
        InitConfig("1", "0", EDGE_MESSAGE, false);
        MGLOBAL::com.init(MEDUSA_COMBINE_INT, MEDUSA_COMBINE_ADD);
        GraphPartition gp[1];
        LoadGraph(argv[1], gp);
        InitHostDS(gp);
        VertexArray& V = MGLOBAL::gpu_def[0].vertexArray;
        for (int v = 0; v < V.size; ++v) {
            V.wsum[v] = 0;
            V.result[v] = 0;
        }
        InitDeviceDS();
        ResetToExecute();
        Message none;
        none.val = -1;
        InitMessageBuffer(none);
        FunctorHolder<EDGELIST>::Run(WalkOutEdges());
        FunctorHolder<MESSAGE>::Run(BumpMessage());
        MGLOBAL::com.combineAllDevice();
        FunctorHolder<VERTEX>::Run(StoreCombined());
        (void)RetriveToExecute();
        DumpVertexData();
        std::vector<MVT> both(V.wsum, V.wsum + V.size);
        both.insert(both.end(), V.result, V.result + V.size);
        WriteArray(argv[2], both.data(), both.size());
        printf("functors n=%d\n", V.size);
        ReleaseDeviceDS();
    } catch (const std::exception& e) {
        fprintf(stderr, "functors_app: %s\n", e.what());
        return 1;
    }
    return 0;
}
