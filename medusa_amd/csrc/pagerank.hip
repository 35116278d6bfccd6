// PageRank superstep engine (FP32, compensated combine in a fixed order).
//
// Reference: MPG/UserOperators.h:175-270 runs per superstep InitMessageBuffer(0.0)
// (MPG/InitMessage.h), SendMsg* scatter of rank/outdeg into per-in-edge message
// slots (MPG/UserOperators.h:29-163), a CUDPP segmented ADD scan (MPG/MsgCombiner.h:68-117)
// and UpdateVertex rank = 0.15 + 0.85*sum (MPG/UserOperators.h:16-26); fixed 100
// supersteps (MPG/main.cu:166).  Here the formula is the normalised one with
// dangling mass redistributed:
//     r_0 = 1/n,  r_{t+1}(v) = (1-d)/n + d*(S_t(v) + D_t/n)
//     S_t(v) = sum over in-edges u->v of c_t(u),  c_t(u) = r_t(u)/outdeg(u)
//     D_t = sum of r_t(u) over outdeg(u) == 0.
//
// A packed batch of disjoint graphs runs in the same launches: graph g uses its own
// vertex count n_g for r_0, (1-d)/n_g and D_t/n_g, and D_t sums its own dangling vertices,
// so every graph's ranks are bit-identical to a run on that graph alone.
//
// Canonical combine order (identical for every kernel variant and lane mapping,
// so push/pull and every binning give bit-identical ranks):
//   a segment of values x_0..x_{L-1} (in-edges in ascending original source id,
//   or dangling vertices in internal order) is cut into blocks of 4096; in a
//   block, lane j (0..63) Kahan-adds x_{j}, x_{j+64}, ... in order starting from
//   (0,0); lanes are then merged by kahan_combine down a tree with offsets
//   32,16,8,4,2,1 (lane i absorbs lane i+off); block results are merged left to
//   right with kahan_combine; the value is s - c.
// Scalars: d_f = (float)d, tele = (float)((1-d)/n), base_t = fmaf(d_f, D_t/(float)n, tele),
// r_0 = (float)(1/n),
// r_{t+1}(v) = fmaf(d_f, S_t(v), base_t), c = r/(float)outdeg (IEEE division).

#include <cmath>

#include "engine.h"
#include "sched.h"

namespace medusa {

constexpr uint32_t kBlock = 4096;  // elements per canonical block (64 lanes x 64 rounds)

// Tree-merge of the (s, c) pairs held by groups of g lanes (g a power of two,
// groups aligned within the wave); lane 0 of each group ends with the result.
__device__ __forceinline__ void group_tree(float& s, float& c, uint32_t g) {
    for (uint32_t off = g >> 1; off; off >>= 1) {
        const float s2 = __shfl_down(s, off, kWave);
        const float c2 = __shfl_down(c, off, kWave);
        kahan_combine(s, c, s2, c2);
    }
}

// Canonical sum of `len` values x(k), computed by one whole wave; every lane
// returns the same (s - c).  Each lane keeps 16 loads in flight and adds them
// in ascending order.
template <class Load>
__device__ __forceinline__ float wave_segment_sum(uint32_t len, Load load) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    float as = 0.f, ac = 0.f;
    for (uint32_t b0 = 0; b0 < len; b0 += kBlock) {
        const uint32_t lim = min(len - b0, kBlock);
        float s = 0.f, c = 0.f;
        for (uint32_t r0 = 0; r0 * kWave < lim; r0 += 16) {
            float x[16];
#pragma unroll
            for (uint32_t j = 0; j < 16; ++j) {
                const uint32_t k = lane + (r0 + j) * kWave;
                x[j] = k < lim ? load(b0 + k) : 0.f;
            }
#pragma unroll
            for (uint32_t j = 0; j < 16; ++j)
                if (lane + (r0 + j) * kWave < lim) kahan_add(s, c, x[j]);
        }
        group_tree(s, c, kWave);
        s = __shfl(s, 0, kWave);
        c = __shfl(c, 0, kWave);
        if (b0 == 0) {
            as = s;
            ac = c;
        } else {
            kahan_combine(as, ac, s, c);
        }
    }
    return as - ac;
}

// 16-leaf canonical subtree in registers: slot j absorbs slot j+off for
// off = 8, 4, 2, 1.  Absent leaves are (0,0), an exact identity.
__device__ __forceinline__ void tree16(float (&s)[16], float (&c)[16]) {
#pragma unroll
    for (int off = 8; off >= 1; off >>= 1)
#pragma unroll
        for (int j = 0; j < off; ++j) kahan_combine(s[j], c[j], s[j + off], c[j + off]);
}

// Leaves r, r+st, ..., r+15*st of a vertex of in-degree `deg` (one lane), reduced
// by tree16.  Element i of the vertex sits at col[i*stride] (ELL slab, pull) or at
// vals[seg + i] (message slots, push).
template <bool kFromMsg>
__device__ __forceinline__ void lane_subtree(float& S, float& C, uint32_t r, uint32_t st, uint32_t deg,
                                             const uint32_t* __restrict__ col, uint32_t stride,
                                             const float* __restrict__ vals, uint32_t seg) {
    uint32_t idx[16];
    float s[16], c[16];
#pragma unroll
    for (uint32_t j = 0; j < 16; ++j) {
        const uint32_t i = r + st * j;
        idx[j] = (i < deg) ? (kFromMsg ? seg + i : col[(size_t)i * stride]) : 0u;
    }
#pragma unroll
    for (uint32_t j = 0; j < 16; ++j) {
        s[j] = (r + st * j < deg) ? vals[idx[j]] : 0.f;
        c[j] = 0.f;
    }
    tree16(s, c);
    S = s[0];
    C = c[0];
}

// Canonical sum of one vertex of in-degree deg <= 64 held by a single lane.
// The 64-slot tree splits into 1, 2 or 4 interleaved 16-leaf subtrees.
template <bool kFromMsg>
__device__ __forceinline__ float lane_vertex_sum(uint32_t deg, const uint32_t* __restrict__ col, uint32_t stride,
                                                 const float* __restrict__ vals, uint32_t seg) {
    float s0, c0, s1, c1;
    if (deg <= 16) {
        lane_subtree<kFromMsg>(s0, c0, 0, 1, deg, col, stride, vals, seg);
    } else if (deg <= 32) {
        lane_subtree<kFromMsg>(s0, c0, 0, 2, deg, col, stride, vals, seg);
        lane_subtree<kFromMsg>(s1, c1, 1, 2, deg, col, stride, vals, seg);
        kahan_combine(s0, c0, s1, c1);
    } else {
        float s2, c2, s3, c3;
        lane_subtree<kFromMsg>(s0, c0, 0, 4, deg, col, stride, vals, seg);
        lane_subtree<kFromMsg>(s2, c2, 2, 4, deg, col, stride, vals, seg);
        kahan_combine(s0, c0, s2, c2);
        lane_subtree<kFromMsg>(s1, c1, 1, 4, deg, col, stride, vals, seg);
        lane_subtree<kFromMsg>(s3, c3, 3, 4, deg, col, stride, vals, seg);
        kahan_combine(s1, c1, s3, c3);
        kahan_combine(s0, c0, s1, c1);
    }
    return s0 - c0;
}

// The same canonical sum for a lane whose elements come from a functor (LDS slots of the
// binned variant).
template <class Load>
__device__ __forceinline__ void lane_subtree_at(float& S, float& C, uint32_t r, uint32_t st, uint32_t deg, Load load) {
    float s[16], c[16];
#pragma unroll
    for (uint32_t j = 0; j < 16; ++j) {
        s[j] = (r + st * j < deg) ? load(r + st * j) : 0.f;
        c[j] = 0.f;
    }
    tree16(s, c);
    S = s[0];
    C = c[0];
}
// The canonical 64-slot tree over a lane's first L slots (L a power of two <= 16, deg <= L):
// slot j absorbs slot j+off for off = L/2, ..., 1.  Every slot from deg on is (0,0), so the
// 64-slot tree's levels above L/2 only merge (0,0) pairs, an exact identity, and this is the
// whole sum.
template <uint32_t L, class Load>
__device__ __forceinline__ float lane_tree_small(uint32_t deg, Load load) {
    float s[L], c[L];
#pragma unroll
    for (uint32_t j = 0; j < L; ++j) {
        s[j] = j < deg ? load(j) : 0.f;
        c[j] = 0.f;
    }
#pragma unroll
    for (uint32_t off = L / 2; off >= 1; off >>= 1)
#pragma unroll
        for (uint32_t j = 0; j < off; ++j) kahan_combine(s[j], c[j], s[j + off], c[j + off]);
    return s[0] - c[0];
}

template <class Load>
__device__ __forceinline__ float lane_vertex_sum_at(uint32_t deg, Load load) {
    if (deg <= 1) return lane_tree_small<1>(deg, load);
    if (deg <= 2) return lane_tree_small<2>(deg, load);
    if (deg <= 4) return lane_tree_small<4>(deg, load);
    if (deg <= 8) return lane_tree_small<8>(deg, load);
    // the 1, 2 or 4 interleaved subtrees one at a time (a loop, so only one subtree's leaves
    // are live), then merged as the 64-slot tree merges them
    const uint32_t st = deg <= 16 ? 1u : deg <= 32 ? 2u : 4u;
    float ss[4], cc[4];
#pragma unroll 1
    for (uint32_t r = 0; r < st; ++r) {
        float S, C;
        lane_subtree_at(S, C, r, st, deg, load);
        ss[r] = S;
        cc[r] = C;
    }
    if (st == 2) {
        kahan_combine(ss[0], cc[0], ss[1], cc[1]);
    } else if (st == 4) {
        kahan_combine(ss[0], cc[0], ss[2], cc[2]);
        kahan_combine(ss[1], cc[1], ss[3], cc[3]);
        kahan_combine(ss[0], cc[0], ss[1], cc[1]);
    }
    return ss[0] - cc[0];
}

// One superstep for every vertex.  kFromMsg: values come from the message
// array in slot order (push variant) instead of contrib[src] (pull variant).
template <bool kFromMsg>
__global__ __launch_bounds__(256) void k_pr_main(const PrRun* __restrict__ runs, const uint32_t* __restrict__ wg_run,
                                                 const uint32_t* __restrict__ ell, const uint32_t* __restrict__ cptr,
                                                 const uint32_t* __restrict__ crow,
                                                 const float* __restrict__ vals,
                                                 const uint32_t* __restrict__ outdeg,
                                                 const float* __restrict__ base_p, float d,
                                                 float* __restrict__ rank_out, float* __restrict__ contrib_out) {
    const uint32_t wg = blockIdx.x;
    const PrRun R = runs[wg_run[wg]];  // a table lookup, not a search at every workgroup's start
    const uint32_t local = wg - R.wg_begin;
    const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const float base = base_p[R.part];

    auto finish = [&](uint32_t p, float S) {
        const float r = fmaf(d, S, base);
        rank_out[p] = r;
        const uint32_t od = outdeg[p];
        if (od) contrib_out[p] = r / (float)od;
    };

    if (R.kind == kPrWave) {
        const uint32_t p = R.begin + local * 4 + wave;
        if (p >= R.begin + R.count) return;  // uniform across the wave
        const uint32_t b = cptr[p], len = cptr[p + 1] - b;
        const float S = wave_segment_sum(len, [&](uint32_t k) {
            return kFromMsg ? vals[b + k] : vals[crow[b + k]];
        });
        if (lane == 0) finish(p, S);
        return;
    }
    const uint32_t i = local * 256 + threadIdx.x;
    if (i >= R.count) return;
    const uint32_t p = R.begin + i;
    if (R.kind == kPrZero) {
        finish(p, 0.f);
        return;
    }
    const uint32_t seg = kFromMsg ? cptr[p] : 0u;
    finish(p, lane_vertex_sum<kFromMsg>(R.deg, ell + R.ell_base + i, R.ell_stride, vals, seg));
}

// Canonical (s, c) pair of one dangling block, computed by one wave (lane 0 holds it).
__device__ __forceinline__ void dangle_block_sum(const float* __restrict__ rank, const PrBlock B, float& s, float& c) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const float* x = rank + B.begin;
    s = 0.f;
    c = 0.f;
    for (uint32_t r0 = 0; r0 * kWave < B.len; r0 += 16) {
        float v[16];
#pragma unroll
        for (uint32_t j = 0; j < 16; ++j) {
            const uint32_t k = lane + (r0 + j) * kWave;
            v[j] = k < B.len ? x[k] : 0.f;
        }
#pragma unroll
        for (uint32_t j = 0; j < 16; ++j)
            if (lane + (r0 + j) * kWave < B.len) kahan_add(s, c, v[j]);
    }
    group_tree(s, c, kWave);
}

// Dangling mass, step 1: one wave per canonical block of a graph's dangling tail
// writes the block's (s, c) pair.
__global__ __launch_bounds__(256) void k_pr_dangle(const float* __restrict__ rank, const PrBlock* __restrict__ blocks,
                                                   uint32_t nblk, float* __restrict__ partials) {
    const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const uint32_t b = blockIdx.x * 4 + wave;
    if (b >= nblk) return;  // uniform across the wave
    float s, c;
    dangle_block_sum(rank, blocks[b], s, c);
    if (lane == 0) {
        partials[2 * b] = s;
        partials[2 * b + 1] = c;
    }
}

// base = fmaf(d, D / (float)n_g, (float)((1 - d) / n_g)) for dangling mass D.
__device__ __forceinline__ float pr_base_value(float D, float d, uint32_t n_g) {
    const float tele = (float)((1.0 - (double)d) / (double)n_g);
    return fmaf(d, D / (float)n_g, tele);
}

// base = fmaf(d, D / (float)n_g, (float)((1 - d) / n_g)) of a graph whose `nblocks` block pairs
// start at partials[2 * first]: the pairs merge left to right.  Called by a whole wave; every
// lane returns the value.
__device__ __forceinline__ float pr_base_fold(const float* __restrict__ partials, uint32_t first, uint32_t nblocks,
                                              uint32_t n_g, float d) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    float as = 0.f, ac = 0.f;
    for (uint32_t c0 = 0; c0 < nblocks; c0 += kWave) {
        const uint32_t i = c0 + lane;
        const float ps = i < nblocks ? partials[2 * (first + i)] : 0.f;
        const float pc = i < nblocks ? partials[2 * (first + i) + 1] : 0.f;
        const uint32_t lim = min(nblocks - c0, (uint32_t)kWave);
        for (uint32_t j = 0; j < lim; ++j) {
            const float s2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, ps), j));
            const float c2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, pc), j));
            if (c0 + j == 0) {
                as = s2;
                ac = c2;
            } else {
                kahan_combine(as, ac, s2, c2);
            }
        }
    }
    return pr_base_value(as - ac, d, n_g);
}

// Dangling mass, step 2 (after the block kernel, so its partials are visible): one wave
// per graph merges the graph's block pairs and writes its base.
__global__ __launch_bounds__(256) void k_pr_base(const float* __restrict__ partials, const PrPart* __restrict__ parts,
                                                 uint32_t nparts, float d, float* __restrict__ base_out) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t w = blockIdx.x * 4 + threadIdx.x / kWave;
    if (w >= nparts) return;  // uniform across the wave
    const PrPart P = parts[w];
    const float base = pr_base_fold(partials, P.first_block, P.nblocks, P.n, d);
    if (lane == 0) base_out[w] = base;
}

constexpr uint32_t kSmallBlocks = 64;  // dangling blocks up to which one workgroup does both steps

// Both dangling-mass steps in one workgroup when there are at most kSmallBlocks blocks:
// wave w sums blocks w, w+4, ... into LDS, then wave w merges graphs w, w+4, ... in block
// order.  Same operations and order as k_pr_dangle followed by k_pr_base.
__global__ __launch_bounds__(256) void k_pr_dangle_base(const float* __restrict__ rank,
                                                        const PrBlock* __restrict__ blocks, uint32_t nblk,
                                                        const PrPart* __restrict__ parts, uint32_t nparts, float d,
                                                        float* __restrict__ base_out) {
    __shared__ float s_s[kSmallBlocks], s_c[kSmallBlocks];
    const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    for (uint32_t b = wave; b < nblk; b += 4) {
        const PrBlock B = blocks[b];
        const float* x = rank + B.begin;
        float s = 0.f, c = 0.f;
        for (uint32_t r0 = 0; r0 * kWave < B.len; r0 += 16) {
            float v[16];
#pragma unroll
            for (uint32_t j = 0; j < 16; ++j) {
                const uint32_t k = lane + (r0 + j) * kWave;
                v[j] = k < B.len ? x[k] : 0.f;
            }
#pragma unroll
            for (uint32_t j = 0; j < 16; ++j)
                if (lane + (r0 + j) * kWave < B.len) kahan_add(s, c, v[j]);
        }
        group_tree(s, c, kWave);
        if (lane == 0) {
            s_s[b] = s;
            s_c[b] = c;
        }
    }
    __syncthreads();
    if (lane != 0) return;
    for (uint32_t g = wave; g < nparts; g += 4) {
        const PrPart P = parts[g];
        float as = 0.f, ac = 0.f;
        for (uint32_t i = 0; i < P.nblocks; ++i) {
            if (i == 0) {
                as = s_s[P.first_block];
                ac = s_c[P.first_block];
            } else {
                kahan_combine(as, ac, s_s[P.first_block + i], s_c[P.first_block + i]);
            }
        }
        const float D = as - ac;
        const float tele = (float)((1.0 - (double)d) / (double)P.n);
        base_out[g] = fmaf(d, D / (float)P.n, tele);
    }
}

// Workgroups of a launch are dispatched round-robin over the 8 XCDs (each with its own L2):
// item b of a launch of `count` maps to XCD (b mod 8)'s contiguous share of the items, so
// consecutive items run on one XCD.
__device__ __forceinline__ uint32_t xcd_major(uint32_t b, uint32_t count) {
    const uint32_t q = count / 8, r = count % 8, x = b % 8, j = b / 8;
    return x * q + min(x, r) + j;
}

// Workgroup b of a launch of `count` (dispatched to XCD b mod 8) takes item
// ((j / kXcdGroup) 8 + x) kXcdGroup + j mod kXcdGroup, b = 8 j + x: blocks of kXcdGroup
// consecutive items run on one XCD at once, and the blocks go round-robin over the XCDs, so
// consecutive items share an L2 while every XCD gets a mix of the items; the tail past the
// last whole round of blocks maps to itself.  A bijection on [0, count).
constexpr uint32_t kXcdGroup = 8;
__device__ __forceinline__ uint32_t xcd_grouped(uint32_t b, uint32_t count) {
    const uint32_t full = count / (8 * kXcdGroup) * (8 * kXcdGroup);
    if (b >= full) return b;
    const uint32_t x = b % 8, j = b / 8;
    return ((j / kXcdGroup) * 8 + x) * kXcdGroup + j % kXcdGroup;
}

constexpr uint32_t kSendThreads = 512;
constexpr uint32_t kSendWaves = kSendThreads / kWave;

// Binned send (one launch per superstep).  Workgroups [0, ntiles) each compute their chunk's
// contributions c = r / outdeg into LDS and write one tile of the message array (contiguous;
// runs are padded to 4 entries, so every access is 16-byte aligned).  Tiles of one chunk run on
// one XCD, so the chunk's ranks are read from HBM once per XCD.
// Workgroups from ntiles on write the (s, c) pairs of the dangling blocks of `rank` (one block
// per wave), which the combine of the same superstep merges into its base.
__global__ __launch_bounds__(kSendThreads) void k_pr_send(const PrTile* __restrict__ tiles, uint32_t ntiles,
                                                 const uint16_t* __restrict__ srcloc, const float* __restrict__ rank,
                                                 const uint32_t* __restrict__ outdeg, float* __restrict__ msg,
                                                 const PrBlock* __restrict__ blocks,
                                                 uint32_t nblk, float* __restrict__ partials) {
    __shared__ float s_c[kPrChunkSources];
    using u16x4 = unsigned short __attribute__((ext_vector_type(4)));
    using f32x4 = float __attribute__((ext_vector_type(4)));
    using u32x4 = uint32_t __attribute__((ext_vector_type(4)));
    if (blockIdx.x >= ntiles) {
        const uint32_t lane = threadIdx.x & (kWave - 1);
        const uint32_t b = (blockIdx.x - ntiles) * kSendWaves + threadIdx.x / kWave;
        if (b >= nblk) return;  // uniform across the wave
        float bs, bc;
        dangle_block_sum(rank, blocks[b], bs, bc);
        if (lane == 0) {
            partials[2 * b] = bs;
            partials[2 * b + 1] = bc;
        }
        return;
    }
    const PrTile T = tiles[xcd_major(blockIdx.x, ntiles)];
    const u16x4* __restrict__ sl = reinterpret_cast<const u16x4*>(srcloc + T.msg_begin);
    f32x4* __restrict__ out = reinterpret_cast<f32x4*>(msg + T.msg_begin);
    const uint32_t nq = (T.msg_end - T.msg_begin) / 4;
    auto fetch = [&](uint32_t q0, u16x4(&a)[4]) {
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
            const uint32_t q = q0 + u * kSendThreads + threadIdx.x;
            a[u] = q < nq ? sl[q] : u16x4{0, 0, 0, 0};
        }
    };
    u16x4 a[4];
    fetch(0, a);  // in flight while the chunk's contributions are computed
    {  // chunks start on multiples of kPrChunkSources, so the slice is 16-byte aligned
        auto contrib = [](float r, uint32_t od) { return od ? r / (float)od : 0.f; };
        const f32x4* r4 = reinterpret_cast<const f32x4*>(rank + T.src_begin);
        const u32x4* o4 = reinterpret_cast<const u32x4*>(outdeg + T.src_begin);
        f32x4* s4 = reinterpret_cast<f32x4*>(s_c);
        const uint32_t n4 = T.src_count / 4;
        for (uint32_t i = threadIdx.x; i < n4; i += kSendThreads) {
            const f32x4 r = r4[i];
            const u32x4 o = o4[i];
            s4[i] = f32x4{contrib(r.x, o.x), contrib(r.y, o.y), contrib(r.z, o.z), contrib(r.w, o.w)};
        }
        for (uint32_t i = 4 * n4 + threadIdx.x; i < T.src_count; i += kSendThreads)
            s_c[i] = contrib(rank[T.src_begin + i], outdeg[T.src_begin + i]);
    }
    __syncthreads();
    for (uint32_t q0 = 0; q0 < nq; q0 += 4 * kSendThreads) {
        u16x4 nxt[4];
        fetch(q0 + 4 * kSendThreads, nxt);  // the next round's offsets fly while this round writes
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
            const uint32_t q = q0 + u * kSendThreads + threadIdx.x;
            if (q < nq) out[q] = f32x4{s_c[a[u].x], s_c[a[u].y], s_c[a[u].z], s_c[a[u].w]};
        }
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) a[u] = nxt[u];
    }
}

uint32_t pr_send_tile_slots(const medusa_graph* g, uint32_t nblk) {
    int cus = 0, per_cu = 0;
    MDS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device));
    MDS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k_pr_send),
                                                         (int)kSendThreads, 0));
    const uint64_t slots = (uint64_t)std::max(cus, 0) * (uint64_t)std::max(per_cu, 0);
    const uint64_t dwg = (nblk + kSendWaves - 1) / kSendWaves;
    return slots > dwg ? (uint32_t)std::min<uint64_t>(slots - dwg, 1u << 20) : 0u;
}

constexpr uint32_t kCombineThreads = 512;
constexpr uint32_t kCombineWaves = kCombineThreads / kWave;
constexpr uint32_t kCombineUnroll = 4;  // groups of four entries per thread in flight

// Binned combine (one launch per superstep).  Workgroup b < nbins takes bin xcd_grouped(b):
//   1. zeroes its LDS slots while every wave loads the records of its items (items w, w+8, ...
//      for wave w, one per lane) into registers;
//   2. reads its entries four at a time: a group's message-array position (qmsg, planned on
//      the host; a group never straddles two runs) and its four slots (8 bytes) load together,
//      then its four messages (16 bytes); each message goes to its LDS slot (padding entries to
//      a dummy slot);
//   3. after the barrier, wave w sums its items from LDS in the canonical order.  A lane
//      vertex sums as many slots as the largest in-degree of its group: the slots past its own
//      in-degree hold +0.0, which merges as the exact identity the canonical tree uses for
//      absent elements, so no per-vertex degree is read.
// Workgroups from nbins on sum the hub destinations (one wave each) by gathering r / outdeg of
// their sources over the CSC, with the same division the send does.  kFold: a single graph with at most 64
// dangling blocks, whose base every wave merges here from the send's block pairs; otherwise
// base[part] was written by k_pr_base.  kContrib: also write r / outdeg (the next superstep
// runs another variant, which reads it).
template <bool kFold, bool kContrib>
__global__ __launch_bounds__(kCombineThreads) void k_pr_combine(
    const PrBin* __restrict__ bins, uint32_t nbins, const PrItem* __restrict__ items,
    const PrItem* __restrict__ hubs, uint32_t nhubs, const uint32_t* __restrict__ qmsg,
    const float* __restrict__ msg, const uint16_t* __restrict__ slot, const uint32_t* __restrict__ cptr,
    const uint32_t* __restrict__ crow, const float* __restrict__ rank_in, const uint32_t* __restrict__ outdeg,
    const float* __restrict__ base_p, const float* __restrict__ partials, uint32_t nblk, uint32_t n, float d,
    float* __restrict__ rank_out, float* __restrict__ contrib_out) {
    __shared__ float lds[kPrBinValues + 4];  // + the dummy slot of padding entries
    using u16x4 = unsigned short __attribute__((ext_vector_type(4)));
    using f32x4 = float __attribute__((ext_vector_type(4)));
    const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    auto finish = [&](uint32_t p, float S, float base) {
        const float r = fmaf(d, S, base);
        rank_out[p] = r;
        if (kContrib) {
            const uint32_t od = outdeg[p];
            if (od) contrib_out[p] = r / (float)od;
        }
    };
    if (blockIdx.x >= nbins) {
        const uint32_t h = (blockIdx.x - nbins) * kCombineWaves + wave;
        if (h >= nhubs) return;  // uniform across the wave
        const PrItem I = hubs[h];
        const float base = kFold ? pr_base_fold(partials, 0, nblk, n, d) : base_p[I.part];
        const uint32_t b = cptr[I.v], len = cptr[I.v + 1] - b;
        const float S = wave_segment_sum(len, [&](uint32_t k) {
            const uint32_t u = crow[b + k];
            return rank_in[u] / (float)outdeg[u];  // u has an out-edge
        });
        if (lane == 0) finish(I.v, S, base);
        return;
    }
    // neighbouring bins' runs share their end lines in the message array: grouped on one XCD,
    // the second fetch of such a line hits its L2 (RMAT-20: 29 MB less fetched per superstep)
    const PrBin B = bins[xcd_grouped(blockIdx.x, nbins)];
    // 1. this wave's items, one record per lane (a bin has at most 64 items per wave)
    const uint32_t my_item = B.item_begin + wave + kCombineWaves * lane;
    uint32_t iv = 0, icount = 0, ikind = 0, ilds = 0, ipart = 0, ideg = 0;
    if (my_item < B.item_end) {
        const PrItem I = items[my_item];
        iv = I.v;
        icount = I.count;
        ikind = I.kind;
        ilds = I.lds;
        ipart = I.part;
        ideg = I.deg;
    }
    float ps = 0.f, pc = 0.f;  // kFold: this lane's dangling block pair
    if (kFold && lane < nblk) {
        ps = partials[2 * lane];
        pc = partials[2 * lane + 1];
    }
    // 2. entries into their slots: the first round's loads are issued before the slots are
    // zeroed, so they fly during the zeroing and its barrier
    const u16x4* __restrict__ s4 = reinterpret_cast<const u16x4*>(slot + B.slot_begin);
    const uint32_t* __restrict__ g4 = qmsg + B.slot_begin / 4;
    const f32x4* __restrict__ m4 = reinterpret_cast<const f32x4*>(msg);
    const uint32_t nq = B.entries / 4;
    f32x4 x[kCombineUnroll];
    u16x4 t[kCombineUnroll];
    auto fetch = [&](uint32_t q0) {
        uint32_t at[kCombineUnroll];
#pragma unroll
        for (uint32_t u = 0; u < kCombineUnroll; ++u) {
            const uint32_t q = q0 + u * kCombineThreads + threadIdx.x;
            if (q >= nq) continue;
            at[u] = g4[q];
            t[u] = s4[q];
        }
#pragma unroll
        for (uint32_t u = 0; u < kCombineUnroll; ++u)
            if (q0 + u * kCombineThreads + threadIdx.x < nq) x[u] = m4[at[u]];
    };
    auto drop = [&](uint32_t q0) {
#pragma unroll
        for (uint32_t u = 0; u < kCombineUnroll; ++u) {
            if (q0 + u * kCombineThreads + threadIdx.x >= nq) continue;
            lds[t[u].x] = x[u].x;
            lds[t[u].y] = x[u].y;
            lds[t[u].z] = x[u].z;
            lds[t[u].w] = x[u].w;
        }
    };
    fetch(0);
    {
        f32x4* z = reinterpret_cast<f32x4*>(lds);
        for (uint32_t i = threadIdx.x; i < kPrBinValues / 4; i += kCombineThreads) z[i] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    __syncthreads();
    drop(0);
    for (uint32_t q0 = kCombineThreads * kCombineUnroll; q0 < nq; q0 += kCombineThreads * kCombineUnroll) {
        fetch(q0);
        drop(q0);
    }
    __syncthreads();
    // 3. sums
    float base_all = 0.f;
    if (kFold) {
        float as = 0.f, ac = 0.f;
        for (uint32_t j = 0; j < nblk; ++j) {
            const float s2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, ps), j));
            const float c2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, pc), j));
            if (j == 0) {
                as = s2;
                ac = c2;
            } else {
                kahan_combine(as, ac, s2, c2);
            }
        }
        base_all = pr_base_value(as - ac, d, n);
    }
    const uint32_t mine = B.item_end > B.item_begin + wave
                              ? (B.item_end - B.item_begin - wave + kCombineWaves - 1) / kCombineWaves
                              : 0u;
    for (uint32_t j = 0; j < mine; ++j) {
        const uint32_t v = __builtin_amdgcn_readlane(iv, j), cnt = __builtin_amdgcn_readlane(icount, j);
        const uint32_t kind = __builtin_amdgcn_readlane(ikind, j), at = __builtin_amdgcn_readlane(ilds, j);
        const uint32_t deg = __builtin_amdgcn_readlane(ideg, j);
        const float base = kFold ? base_all : base_p[__builtin_amdgcn_readlane(ipart, j)];
        if (kind == kPrWave) {
            const float S = wave_segment_sum(deg, [&](uint32_t k) { return lds[at + k]; });
            if (lane == 0) finish(v, S, base);
        } else if (lane < cnt) {
            if (kind == kPrZero) {
                finish(v + lane, 0.f, base);
            } else {
                const float* x = lds + at + lane;
                finish(v + lane, lane_vertex_sum_at(deg, [&](uint32_t k) { return x[k * kWave]; }), base);
            }
        }
    }
}

uint32_t pr_combine_slots(const medusa_graph* g) {
    int cus = 0, per_cu = 0;
    MDS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, g->device));
    MDS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(k_pr_combine<true, false>), (int)kCombineThreads, 0));
    return (uint32_t)std::min<uint64_t>((uint64_t)std::max(cus, 0) * (uint64_t)std::max(per_cu, 0), 1u << 20);
}

// r_0 = (float)(1 / n_g) for every vertex of graph g (graph found by binary search
// over the range starts, which internal ids keep).
__global__ void k_pr_init(uint32_t n, const uint32_t* __restrict__ part_begin, uint32_t nparts,
                          const uint32_t* __restrict__ outdeg, float* __restrict__ rank, float* __restrict__ contrib) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = nparts;  // last graph with begin <= p
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (part_begin[mid] <= p) lo = mid;
            else hi = mid;
        }
        const float r0 = (float)(1.0 / (double)(part_begin[lo + 1] - part_begin[lo]));
        rank[p] = r0;
        const uint32_t od = outdeg[p];
        if (od) contrib[p] = r0 / (float)od;
    }
}

// Push variant: every out-edge (internal CSR order) writes its source's
// contribution into its destination's in-edge slot.
__global__ void k_pr_scatter(uint32_t n, const uint32_t* __restrict__ optr, const uint32_t* __restrict__ oslot,
                             const float* __restrict__ contrib, float* __restrict__ msg) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t waves = gridDim.x * (blockDim.x / kWave);
    for (uint32_t p = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; p < n; p += waves) {
        const uint32_t b = optr[p], e = optr[p + 1];
        if (b == e) continue;
        const float x = contrib[p];
        for (uint32_t k = b + lane; k < e; k += kWave) msg[oslot[k]] = x;
    }
}

__global__ void k_pr_unpermute(uint32_t n, const uint32_t* __restrict__ iperm, const float* __restrict__ rank,
                               float* __restrict__ out) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        out[v] = rank[iperm[v]];
}

__global__ void k_pr_resid(uint32_t n, const float* __restrict__ a, const float* __restrict__ b,
                           double* __restrict__ out) {
    __shared__ double s_part[4];
    double acc = 0.0;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x)
        acc += fabs((double)a[p] - (double)b[p]);
#pragma unroll
    for (int off = kWave / 2; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) s_part[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(out, s_part[0] + s_part[1] + s_part[2] + s_part[3]);
}

__global__ void k_zero_f64(double* p) { *p = 0.0; }

__global__ void k_set_f32(float* p, float v) { *p = v; }

// A single graph with at most this many dangling blocks merges its base inside the binned
// combine (every wave merges the block pairs itself) instead of in a launch of its own.
constexpr uint32_t kFoldBlocks = 64;

static void pr_run(medusa_graph* g, float damping, uint32_t steps, float* out, const medusa_opts& o,
                   medusa_stats& st) {
    build_pr_layout(g);
    PrLayout& L = *g->pr;
    const uint32_t n = g->n;
    const uint64_t m = g->m;
    const uint32_t nd = L.n_dangling;
    const uint32_t nblk = (uint32_t)L.blocks.size();
    const uint32_t nparts = (uint32_t)L.parts.size();
    MDS_CHECK(o.variant == MEDUSA_PR_DEFAULT || o.variant == MEDUSA_PR_PULL || o.variant == MEDUSA_PR_PUSH ||
                  o.variant == MEDUSA_PR_BINNED,
              MEDUSA_ERR_ARG, "pagerank: unknown variant %d", o.variant);
    // the engine's default: the binned exchange wherever its layout can be planned
    const int dflt = o.variant != MEDUSA_PR_DEFAULT       ? o.variant
                     : build_pr_bins(g, /*required=*/false) ? MEDUSA_PR_BINNED
                                                            : MEDUSA_PR_PULL;
    // kernel variant of every superstep: the caller's, or the learned policy's per superstep
    // (it keeps the caller's unless it predicts a clear gain); the ranks are the same bits
    uint64_t wave_vertices = 0, lane_vertices = 0;
    for (const PrRun& r : L.runs) {
        if (r.kind == kPrWave) wave_vertices += r.count;
        if (r.kind == kPrLane) lane_vertices += r.count;
    }
    Sched sched(o, Sched::kBfs);
    const bool binned_ok = dflt == MEDUSA_PR_BINNED || (o.policy && build_pr_bins(g, /*required=*/false));
    std::vector<uint8_t> var(steps);
    uint64_t sig = 1469598103934665603ull;  // FNV-1a over the sequence, keys the captured graph
    for (uint32_t t = 0; t < steps; ++t) {
        const double rec[MEDUSA_TRACE_FIELDS] = {(double)t, (double)nd, (double)wave_vertices, (double)lane_vertices,
                                                 0.0, (double)n, (double)m, 0.0, 0.0};
        var[t] = (uint8_t)sched.pr_variant(rec, dflt, binned_ok);
        sig = (sig ^ var[t]) * 1099511628211ull;
    }
    const bool push = std::find(var.begin(), var.end(), (uint8_t)MEDUSA_PR_PUSH) != var.end();
    const bool binned = std::find(var.begin(), var.end(), (uint8_t)MEDUSA_PR_BINNED) != var.end();
    if (binned) build_pr_bins(g, /*required=*/true);
    PrBins* Bn = L.bins.get();
    const bool fold = nparts == 1 && nblk <= kFoldBlocks;
    L.rank.ensure(n);
    L.contrib0.ensure(n);
    L.contrib1.ensure(n);
    L.base.ensure(nparts);
    L.partials.ensure(2 * (size_t)std::max(nblk, 1u));
    L.resid.ensure(1);
    L.rank_prev.ensure(n);
    L.unperm.ensure(n);
    if (push) L.msg.ensure(m);
    hipStream_t s = g->stream;
    MDS_HIP(hipMemsetAsync(L.contrib0.p, 0, (size_t)n * 4, s));
    MDS_HIP(hipMemsetAsync(L.contrib1.p, 0, (size_t)n * 4, s));

    const float d_f = damping;
    const uint32_t dangle_wgs = (nblk + 3) / 4, base_wgs = (nparts + 3) / 4;
    // base of the next superstep from the ranks in `r`
    auto enqueue_base = [&](const float* r) {
        if (nblk <= kSmallBlocks) {
            hipLaunchKernelGGL(k_pr_dangle_base, dim3(1), dim3(256), 0, s, r, (const PrBlock*)L.d_blocks.p, nblk,
                               (const PrPart*)L.d_parts.p, nparts, d_f, L.base.p);
            MDS_LAUNCHED();
            return;
        }
        if (nblk) {
            hipLaunchKernelGGL(k_pr_dangle, dim3(dangle_wgs), dim3(256), 0, s, r, (const PrBlock*)L.d_blocks.p, nblk,
                               L.partials.p);
            MDS_LAUNCHED();
        }
        hipLaunchKernelGGL(k_pr_base, dim3(base_wgs), dim3(256), 0, s, (const float*)L.partials.p,
                           (const PrPart*)L.d_parts.p, nparts, d_f, L.base.p);
        MDS_LAUNCHED();
    };

    hipLaunchKernelGGL(k_pr_init, dim3(grid_for(n, 256)), dim3(256), 0, s, n, (const uint32_t*)L.d_part_begin.p,
                       nparts, L.outdeg.p, L.rank.p, L.contrib0.p);
    MDS_LAUNCHED();

    const bool timing = o.timing != 0 || o.trace != 0;
    // per superstep: [3t] superstep start, [3t+1] main kernel start, [3t+2] main kernel end
    if (timing) ensure_events(g, 3 * (size_t)steps);
    g->trace.clear();
    float* rk[2] = {L.rank.p, L.rank_prev.p};
    float* ct[2] = {L.contrib0.p, L.contrib1.p};

    auto enqueue_steps = [&](bool with_events) {
        for (uint32_t t = 0; t < steps; ++t) {
            const float* r_in = rk[t & 1];
            float* r_out = rk[(t + 1) & 1];
            const float* c_in = ct[t & 1];
            float* c_out = ct[(t + 1) & 1];
            const bool step_push = var[t] == MEDUSA_PR_PUSH;
            if (with_events) MDS_HIP(hipEventRecord(g->events[3 * t], s));
            if (var[t] == MEDUSA_PR_BINNED) {
                // send: the contributions of r_t as messages, and the dangling block pairs of r_t;
                // combine: r_{t+1} (and c_{t+1} when the next superstep runs another variant)
                const uint32_t dwg = (nblk + kSendWaves - 1) / kSendWaves;
                if (with_events) MDS_HIP(hipEventRecord(g->events[3 * t + 1], s));
                if (dwg + Bn->ntiles) {  // none when nothing dangles and every in-edge goes to a hub
                    hipLaunchKernelGGL(k_pr_send, dim3(Bn->ntiles + dwg), dim3(kSendThreads), 0, s,
                                       (const PrTile*)Bn->tiles.p, Bn->ntiles, (const uint16_t*)Bn->srcloc.p, r_in,
                                       (const uint32_t*)L.outdeg.p, Bn->msg.p, (const PrBlock*)L.d_blocks.p, nblk,
                                       L.partials.p);
                    MDS_LAUNCHED();
                }
                if (!fold) {
                    hipLaunchKernelGGL(k_pr_base, dim3(base_wgs), dim3(256), 0, s, (const float*)L.partials.p,
                                       (const PrPart*)L.d_parts.p, nparts, d_f, L.base.p);
                    MDS_LAUNCHED();
                }
                const bool keep = t + 1 < steps && var[t + 1] != MEDUSA_PR_BINNED;
                const uint32_t cwg = Bn->nbins + (Bn->nhubs + kCombineWaves - 1) / kCombineWaves;
                const PrItem* hubs = Bn->items.p + (Bn->items.n - Bn->nhubs);
                auto kern = fold ? (keep ? k_pr_combine<true, true> : k_pr_combine<true, false>)
                                 : (keep ? k_pr_combine<false, true> : k_pr_combine<false, false>);
                hipLaunchKernelGGL(kern, dim3(cwg), dim3(kCombineThreads), 0, s, (const PrBin*)Bn->bins.p,
                                   Bn->nbins, (const PrItem*)Bn->items.p, hubs, Bn->nhubs,
                                   (const uint32_t*)Bn->qmsg.p, (const float*)Bn->msg.p,
                                   (const uint16_t*)Bn->slot.p, L.cptr.p, L.crow.p, r_in, L.outdeg.p,
                                   (const float*)L.base.p, (const float*)L.partials.p, nblk, n, d_f, r_out, c_out);
                MDS_LAUNCHED();
                if (with_events) MDS_HIP(hipEventRecord(g->events[3 * t + 2], s));
                continue;
            }
            enqueue_base(r_in);
            if (step_push) {
                hipLaunchKernelGGL(k_pr_scatter, dim3(grid_for(n, 4, 16384)), dim3(256), 0, s, n, L.optr.p,
                                   L.oslot.p, c_in, L.msg.p);
                MDS_LAUNCHED();
            }
            if (with_events) MDS_HIP(hipEventRecord(g->events[3 * t + 1], s));
            if (step_push)
                hipLaunchKernelGGL(k_pr_main<true>, dim3(L.total_wgs), dim3(256), 0, s, (const PrRun*)L.d_runs.p,
                                   (const uint32_t*)L.d_wg_run.p, (const uint32_t*)L.ell.p, L.cptr.p, L.crow.p,
                                   (const float*)L.msg.p, L.outdeg.p, (const float*)L.base.p, d_f, r_out, c_out);
            else
                hipLaunchKernelGGL(k_pr_main<false>, dim3(L.total_wgs), dim3(256), 0, s, (const PrRun*)L.d_runs.p,
                                   (const uint32_t*)L.d_wg_run.p, (const uint32_t*)L.ell.p, L.cptr.p, L.crow.p, c_in,
                                   L.outdeg.p, (const float*)L.base.p, d_f, r_out, c_out);
            MDS_LAUNCHED();
            if (with_events) MDS_HIP(hipEventRecord(g->events[3 * t + 2], s));
        }
    };

    MDS_HIP(hipEventRecord(g->ev_begin, s));
    if (o.use_graph && !timing && steps > 0) {
        // the captured sequence only depends on (steps, variants, damping): buffers are fixed per layout
        if (!L.exec || L.exec_steps != steps || L.exec_sig != sig || L.exec_damping != damping) {
            if (L.exec) (void)hipGraphExecDestroy(L.exec);
            L.exec = nullptr;
            L.exec_sig = 0;
            // capture must not include the event record issued above
            MDS_HIP(hipStreamSynchronize(s));
            hipGraph_t graph = nullptr;
            MDS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try {
                enqueue_steps(false);
            } catch (...) {
                hipGraph_t dummy = nullptr;
                (void)hipStreamEndCapture(s, &dummy);
                if (dummy) (void)hipGraphDestroy(dummy);
                throw;
            }
            MDS_HIP(hipStreamEndCapture(s, &graph));
            hipError_t e = hipGraphInstantiate(&L.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            MDS_HIP(e);
            L.exec_steps = steps;
            L.exec_sig = sig;
            L.exec_damping = damping;
        }
        MDS_HIP(hipGraphLaunch(L.exec, s));
    } else {
        enqueue_steps(timing);
    }
    MDS_HIP(hipEventRecord(g->ev_end, s));
    float* final_rank = rk[steps & 1];
    const float* prev_rank = rk[(steps + 1) & 1];
    if (steps > 0) {
        hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(1), 0, s, L.resid.p);
        MDS_LAUNCHED();
        hipLaunchKernelGGL(k_pr_resid, dim3(grid_for(n, 256, 512)), dim3(256), 0, s, n, (const float*)final_rank,
                           prev_rank, L.resid.p);
        MDS_LAUNCHED();
    }
    hipLaunchKernelGGL(k_pr_unpermute, dim3(grid_for(n, 256)), dim3(256), 0, s, n, L.iperm.p,
                       (const float*)final_rank, L.unperm.p);
    MDS_LAUNCHED();
    MDS_HIP(hipStreamSynchronize(s));
    float ms = 0;
    MDS_HIP(hipEventElapsedTime(&ms, g->ev_begin, g->ev_end));
    st.device_ms = ms;
    // byte model per superstep (pull over CSC): 4(n+1) offsets + 8m (source id and gathered
    // contribution per in-edge) + 16n (rank write, contribution write, out-degree read,
    // dangling read); the push variant reads message slots instead of gathering
    // contributions and adds its scatter
    const uint64_t pull_bytes = 4ull * (n + 1) + 8ull * m + 16ull * n;
    const uint64_t push_bytes = 4ull * m + 8ull * n + 8ull * n + 4ull * (n - nd);
    const uint64_t scatter_bytes = 4ull * (n + 1) + 8ull * m + 4ull * n;
    for (uint32_t t = 0; t < steps; ++t) {
        // the binned variant computes the same superstep as the pull and is costed by the same model
        const bool p = var[t] == MEDUSA_PR_PUSH;
        st.kernel_bytes += p ? push_bytes : pull_bytes;
        st.alg_bytes += (p ? push_bytes + scatter_bytes : pull_bytes) + 4ull * nd;
        if (p) ++st.push_steps;
        else if (var[t] == MEDUSA_PR_PULL) ++st.pull_steps;
        if (timing) {
            float k_ms = 0, s_ms = 0;
            MDS_HIP(hipEventElapsedTime(&k_ms, g->events[3 * t + 1], g->events[3 * t + 2]));
            MDS_HIP(hipEventElapsedTime(&s_ms, g->events[3 * t], g->events[3 * t + 2]));
            st.kernel_ms += k_ms;
            if (o.trace) {
                TraceRec r;
                const double rec[MEDUSA_TRACE_FIELDS] = {(double)t, (double)nd, (double)wave_vertices,
                                                         (double)lane_vertices, 0.0, (double)n, (double)m,
                                                         (double)var[t], (double)s_ms};
                std::copy(rec, rec + MEDUSA_TRACE_FIELDS, r.f);
                g->trace.push_back(r);
            }
        }
    }
    st.kernel_launches = steps;
    st.edges_traversed = (uint64_t)steps * m;
    st.supersteps = steps;
    st.reached = n;
    if (steps > 0) MDS_HIP(hipMemcpy(&st.residual_l1, L.resid.p, sizeof(double), hipMemcpyDeviceToHost));
    copy_out(g, out, L.unperm.p, (size_t)n * 4);
}

}  // namespace medusa

using namespace medusa;

extern "C" int medusa_pagerank(medusa_graph* g, float damping, uint32_t steps, float* rank, const medusa_opts* opts,
                               medusa_stats* stats) {
    return guard([&] {
        const auto lock = enter(g);
        MDS_CHECK(rank, MEDUSA_ERR_ARG, "pagerank: null output");
        MDS_CHECK(damping >= 0.f && damping < 1.f, MEDUSA_ERR_ARG, "pagerank: damping must be in [0,1)");
        medusa_opts o{};
        if (opts) o = *opts;
        medusa_stats st{};
        pr_run(g, damping, steps, rank, o, st);
        if (stats) *stats = st;
    });
}
