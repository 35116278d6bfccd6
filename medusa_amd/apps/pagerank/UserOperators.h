// PageRank with the reference application's formula: every vertex sends rank/outdeg
// along its out-edges, the ADD combiner sums each vertex's slots, and the vertex
// functor sets rank = 0.15 + 0.85 * sum (ranks start at 1.0; dangling mass is not
// redistributed).
#pragma once

struct ShareRank {
    __device__ void operator()(D_Edge edge, int) {
        D_Vertex from(edge.get_srcVertexID());
        edge.sendMsg(from.get_level() / from.get_edge_count());
    }
};

struct GatherRank {
    __device__ void operator()(D_Vertex v, int) {
        MVT sum = v.get_combined_msg();
        sum *= 0.85;
        v.set_level(0.15 + sum);
    }
};

inline void Medusa_Exec() {
    Message zero;
    zero.val = 0.0f;
    InitMessageBuffer(zero);
    FunctorHolder<EDGE>::Run(ShareRank());
    MGLOBAL::com.combineAllDevice();
    FunctorHolder<VERTEX>::Run(GatherRank());
}
