// BFS as an edge functor: an edge whose source sits on the current level claims an
// unreached destination for the next level and votes to continue.
#pragma once

struct ExpandLevel {
    __device__ void operator()(D_Edge edge, int step) {
        D_Vertex from(edge.get_srcVertexID());
        if (from.get_level() != step) return;
        D_Vertex to(edge.get_dstVertexID());
        if (to.get_level() == MVT_Init_Value) {
            to.set_level(step + 1);
            Medusa_Continue();
        }
    }
};

inline void Medusa_Exec() { FunctorHolder<EDGE>::Run(ExpandLevel()); }
