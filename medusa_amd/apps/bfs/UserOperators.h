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

// The reference's dispatch (Algorithm/UserOperators.h:44-53): multi-hop and multi-GPU holders
// when more than one GPU is used, the single-GPU holder otherwise.
inline void Medusa_Exec() {
    ExpandLevel ul;
    if (MGLOBAL::num_gpu_to_use > 1 && MGLOBAL::max_hop >= 1)
        MGFunctorHolder<MGMH_EDGE>::Run(ul);
    else if (MGLOBAL::num_gpu_to_use > 1)
        MGFunctorHolder<MG_EDGE>::Run(ul);
    else
        FunctorHolder<EDGE>::Run(ul);
}
