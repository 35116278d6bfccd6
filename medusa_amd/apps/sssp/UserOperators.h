// Bellman-Ford in BSP form: vertices improved last superstep send dist+w along their
// out-edges, the MIN combiner keeps each vertex's best offer, and a vertex functor
// accepts it if it is shorter.
#pragma once
#include <cfloat>

struct OfferDistance {
    __device__ void operator()(D_Edge edge, int) {
        D_Vertex from(edge.get_srcVertexID());
        if (from.get_updated()) edge.sendMsg(from.get_distance() + edge.get_weight());
    }
};

struct AcceptDistance {
    __device__ void operator()(D_Vertex v, int) {
        const MVT best = v.get_combined_msg();
        const bool better = best < v.get_distance();
        if (better) {
            v.set_distance(best);
            Medusa_Continue();
        }
        v.set_updated(better);
    }
};

inline void Medusa_Exec() {
    Message none;
    none.val = FLT_MAX;
    InitMessageBuffer(none);
    FunctorHolder<EDGE>::Run(OfferDistance());
    MGLOBAL::com.combineAllDevice();
    FunctorHolder<VERTEX>::Run(AcceptDistance());
}
