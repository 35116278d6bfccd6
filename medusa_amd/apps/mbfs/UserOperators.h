// Up to 64 BFS runs at once.  An edge ORs its source's frontier bits, minus those the
// destination already has, into the destination's `next`; the vertex functor turns new
// bits into the next frontier and records their level.
#pragma once

static __device__ unsigned char* g_levels;  // [k][n] bytes, 0xFF = unreached
static __device__ int g_n;

struct SpreadMask {
    __device__ void operator()(D_Edge edge, int) {
        D_Vertex from(edge.get_srcVertexID());
        const MVT bits = from.get_frontier();
        if (!bits) return;
        D_Vertex to(edge.get_dstVertexID());
        const MVT add = bits & ~to.get_visited();
        if (add && (to.get_next() & add) != add) atomicOr(to.ptr_next(), add);
    }
};

struct AdvanceMask {
    __device__ void operator()(D_Vertex v, int step) {
        const MVT fresh = v.get_next() & ~v.get_visited();
        v.set_next(0);
        v.set_frontier(fresh);
        if (!fresh) return;
        v.set_visited(v.get_visited() | fresh);
        const unsigned char lv = step + 1 >= 255 ? 0xFF : (unsigned char)(step + 1);
        for (MVT b = fresh; b; b &= b - 1) g_levels[(size_t)__builtin_ctzll(b) * g_n + v.index] = lv;
        Medusa_Continue();
    }
};

inline void Medusa_Exec() {
    FunctorHolder<EDGE>::Run(SpreadMask());
    FunctorHolder<VERTEX>::Run(AdvanceMask());
}
