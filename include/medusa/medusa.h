// medusa/medusa.h — Medusa's user API on the MI355X engine (header-only, compile with hipcc).
//
// An application keeps the reference's shape (Vertex/Edge/Message declarations,
// functors in UserOperators.h, a BSP loop in main) and includes this header
// after declaring, in its own headers:
//
//   #define MVT float                          // message value type   (reference Message.h)
//   #define MVT_Init_Value -1                  //                      (Algorithm/Message.h:9-10)
//   #define MEDUSA_VERTEX_FIELDS(F) F(MVT, level)   // user vertex fields (reference Vertex.h)
//   #define MEDUSA_EDGE_FIELDS(F)   F(MVT, weight)  // user edge fields   (reference Edge.h)
//
// and gets, with the reference's names:
//   VertexArray / D_VertexArray / EdgeArray / D_EdgeArray / MessageArray / D_MessageArray
//       (Algorithm/VertexDataType.h:16-55, EdgeDataType.h:16-63, MessageDataType.h:26-47;
//        the reference notes they "should be automatically generated" — here they are)
//   D_Vertex / D_Edge / D_Message / EdgeList with get_<field>/set_<field> accessors
//       (Algorithm/DeviceDataStructure.h:33-233)
//   Medusa_Continue / Medusa_Break (MedusaRT/SystemLibGPU.h:14-22)
//   FunctorHolder<EDGE|VERTEX|EDGELIST|MESSAGE|EDGE_HT>::Run(op) (MultipleGPU/MultiAPIHoster.h:649-754),
//   MGFunctorHolder<MG_EDGE|MGMH_EDGE|MG_VERTEX|MGMH_VERTEX>::Run(op) on one GPU (MultiAPIHoster.h:28-640)
//   InitConfig, InitHostDS, CheckOutgoingEdge, InitDeviceDS, ResetGraphData, InitMessageBuffer,
//   NULL_Combine, ResetToExecute, RetriveToExecute (MultipleGPU/MultiPublicAPI.h:16-401,
//   MultiUtilities.h:48-69), MGLOBAL::com.combineAllDevice() (MedusaRT/Combiner.cu:82-142)
//   GraphIR, GraphPartition, GTGraph::get_parition (MedusaRT/GraphConverter.h:37-89,
//       MultipleGPU/PartitionManager.cu:116-534, one partition).
//   QFunctorHolder<EQ2EQ|EQ2VQ|EQConsume|VQ2EQ|VQ2VQ|VQConsume|VA2EQ|VA2AQ|EA2EQ|EA2AQ>::Run
//       over MedusaQueue (the reference's planned queue API, MultipleGPU/MultiAPIHoster.h:18-23,
//       MedusaRT/APIKernelWrappers.h:53-140).
//
// Addition: Medusa_Run(Medusa_Exec[, max_supersteps, vote]) runs the BSP loop on the device
// (captured batches of supersteps, the vote read once per batch).
//
// Storage comes from libmedusa_amd.so (medusa_graph_create: CSR + CSC + slot map in HBM);
// user fields are device arrays allocated here.  Every call runs on the current GPU; the
// multi-GPU arguments of the reference API are accepted only with one GPU.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../medusa_c.h"

#ifndef MVT
#error "define MVT (the message value type) before including <medusa/medusa.h>"
#endif
#ifndef MVT_Init_Value
#error "define MVT_Init_Value before including <medusa/medusa.h>"
#endif
#ifndef MEDUSA_VERTEX_FIELDS
#define MEDUSA_VERTEX_FIELDS(F)
#endif
#ifndef MEDUSA_EDGE_FIELDS
#define MEDUSA_EDGE_FIELDS(F)
#endif

// ---- configuration enums (Algorithm/Configuration.h:50-51) ---------------------
enum MessageMode { EDGE_MESSAGE, VERTEX_MESSAGE, NO_MESSAGE };
enum MedusaDS { DS_MEG, DS_AA, DS_HY, DS_ELL };
// combiner element type / operator (the reference passes CUDPP enums, MPG/main.cu:42-43)
enum MedusaCombinerType { MEDUSA_COMBINE_FLOAT, MEDUSA_COMBINE_INT, MEDUSA_COMBINE_UINT };
enum MedusaCombinerOp { MEDUSA_COMBINE_ADD, MEDUSA_COMBINE_MIN, MEDUSA_COMBINE_MAX, MEDUSA_COMBINE_MUL };
// functor kinds (MultipleGPU/MultiAPIHoster.h:18)
// functor kinds, with the reference's names and order (MultipleGPU/MultiAPIHoster.h:18); the MG_
// and MGMH_ kinds are the multi-GPU holders, which on one GPU do what EDGE / VERTEX do
enum EMVAPIType { EDGE, EDGELIST, MESSAGE, MESSAGELIST, VERTEX, MG_VERTEX, EDGE_HT, MG_EDGE, MGMH_EDGE, MGMH_VERTEX };

#define MEDUSA_HIP_OK(expr)                                                                                \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ---- host graph IR (MedusaRT/GraphConverter.h:37-89) ----------------------------
struct Edge {
    int srcVertexID = 0;
    int dstVertexID = 0;
    int msgDstID = 0;
    double weight = 0;
};

struct GraphIR {
    int vertexNum = 0;
    int totalEdgeCount = 0;
    std::vector<std::vector<Edge>> out;  // per source, insertion order
    std::vector<int> incoming_edge_count;
    GraphIR() = default;
    explicit GraphIR(int num) { resize(num); }
    void resize(int num) {
        vertexNum = num;
        out.assign(num, {});
        incoming_edge_count.assign(num, 0);
        totalEdgeCount = 0;
    }
    // Reference semantics: AddEdge prepends (GraphConverter.cu:87-103), so a vertex's
    // out-edges end up in reverse insertion order in the edge array.
    void AddEdge(Edge e) {
        if (e.srcVertexID < 0 || e.srcVertexID >= vertexNum || e.dstVertexID < 0 || e.dstVertexID >= vertexNum)
            throw std::runtime_error("GraphIR::AddEdge: vertex id out of range");
        out[e.srcVertexID].push_back(e);
        ++incoming_edge_count[e.dstVertexID];
        ++totalEdgeCount;
    }
    // CSR in the reference's edge order; weights as given.
    void toCSR(std::vector<uint32_t>& row_ptr, std::vector<uint32_t>& col, std::vector<double>& w) const {
        row_ptr.assign((size_t)vertexNum + 1, 0);
        col.clear();
        w.clear();
        col.reserve(totalEdgeCount);
        w.reserve(totalEdgeCount);
        for (int v = 0; v < vertexNum; ++v) {
            for (auto it = out[v].rbegin(); it != out[v].rend(); ++it) {
                col.push_back((uint32_t)it->dstVertexID);
                w.push_back(it->weight);
            }
            row_ptr[v + 1] = (uint32_t)col.size();
        }
    }
};

struct GraphPartition {
    GraphIR graph;
    std::vector<int> l2p;  // logical (file) id -> physical id; identity on one GPU
};

// GTgraph / DIMACS-style reader: "c ..." comments, "p sp N M", "a u v w" with 1-based ids
// (MultipleGPU/PartitionManager.cu:156-213).  One GPU: every vertex in partition 0 and the
// partition file is not needed.
struct GTGraph {
    void get_parition(GraphPartition* gp, const char* gt_file, const char* partition_file = nullptr) {
        (void)partition_file;
        std::ifstream f(gt_file);
        if (!f.good()) throw std::runtime_error(std::string("GTGraph: cannot open ") + gt_file);
        std::string line;
        int n = -1;
        long long declared = -1;
        while (std::getline(f, line)) {
            if (line.empty() || line[0] == 'c') continue;
            std::istringstream is(line);
            char tag;
            is >> tag;
            if (tag == 'p') {
                std::string kind;
                is >> kind >> n >> declared;
                gp[0].graph.resize(n);
                gp[0].l2p.resize(n);
                for (int i = 0; i < n; ++i) gp[0].l2p[i] = i;
            } else if (tag == 'a') {
                if (n < 0) throw std::runtime_error("GTGraph: 'a' line before 'p' line");
                long long u, v;
                double w = 0;
                is >> u >> v;
                if (!(is >> w)) w = 0;
                Edge e;
                e.srcVertexID = (int)(u - 1);
                e.dstVertexID = (int)(v - 1);
                e.weight = w;
                gp[0].graph.AddEdge(e);
            }
        }
        if (n < 0) throw std::runtime_error("GTGraph: no 'p' line");
        if (declared >= 0 && declared != gp[0].graph.totalEdgeCount)
            fprintf(stderr, "GTGraph: header declares %lld arcs, read %d\n", declared, gp[0].graph.totalEdgeCount);
    }
};

// ---- SoA arrays generated from the field lists -------------------------------------
#define MDS_HOST_FIELD(T, N) T* N = nullptr;
#define MDS_DEV_FIELD(T, N) T* d_##N = nullptr;

struct Vertex {
    int edge_count = 0;
};

struct VertexArray {
    int size = 0;
    int* edge_count = nullptr;
    int* edge_index = nullptr;  // CSR offsets (n+1)
    int* msg_index = nullptr;   // CSC offsets (n+1)
    MEDUSA_VERTEX_FIELDS(MDS_HOST_FIELD)
};
struct D_VertexArray {
    int size = 0;
    const int* d_edge_count = nullptr;
    const int* d_edge_index = nullptr;
    const int* d_msg_index = nullptr;
    MEDUSA_VERTEX_FIELDS(MDS_DEV_FIELD)
};
struct EdgeArray {
    int size = 0;
    int* srcVertexID = nullptr;
    int* dstVertexID = nullptr;
    int* msgDstID = nullptr;
    MEDUSA_EDGE_FIELDS(MDS_HOST_FIELD)
};
struct D_EdgeArray {
    int size = 0;
    const int* d_srcVertexID = nullptr;
    const int* d_dstVertexID = nullptr;
    const int* d_msgDstID = nullptr;
    MEDUSA_EDGE_FIELDS(MDS_DEV_FIELD)
};
struct Message {
    MVT val;
};
struct MessageArray {
    MVT* val = nullptr;
    int size = 0;
};
struct D_MessageArray {
    MVT* d_val = nullptr;
    int size = 0;
};

// ---- device-side global state (Algorithm/DeviceDataStructure.h:7-24) ---------------
namespace GRAPH_STORAGE_GPU {
static __constant__ D_VertexArray d_vertexArray;
static __constant__ D_EdgeArray d_edgeArray;
static __constant__ D_MessageArray d_messageArray;
static __constant__ MVT* d_messageArrayBuf;  // combined message per vertex
static __device__ bool d_toExecute;           // BSP vote
static __device__ MVT d_init_msg_val;
// Medusa_Run's loop on the device: running, the superstep being run, its last superstep
// (exclusive) and whether the vote ends it
static __device__ int d_active;
static __device__ int d_super_step;
static __device__ int d_max_step;
static __device__ int d_vote;
}  // namespace GRAPH_STORAGE_GPU

// ---- system calls for functors (MedusaRT/SystemLibGPU.h:14-22) ----------------------
__device__ inline void Medusa_Break() { GRAPH_STORAGE_GPU::d_toExecute = false; }
__device__ inline void Medusa_Continue() { GRAPH_STORAGE_GPU::d_toExecute = true; }

// ---- accessors ----------------------------------------------------------------------
#define MDS_VERTEX_ACCESS(T, N)                                                                      \
    __device__ T get_##N() const { return GRAPH_STORAGE_GPU::d_vertexArray.d_##N[index]; }          \
    __device__ void set_##N(T v) const { GRAPH_STORAGE_GPU::d_vertexArray.d_##N[index] = v; }       \
    __device__ T* ptr_##N() const { return &GRAPH_STORAGE_GPU::d_vertexArray.d_##N[index]; }
#define MDS_EDGE_ACCESS(T, N)                                                                        \
    __device__ T get_##N() const { return GRAPH_STORAGE_GPU::d_edgeArray.d_##N[edge_index]; }       \
    __device__ void set_##N(T v) const { GRAPH_STORAGE_GPU::d_edgeArray.d_##N[edge_index] = v; }    \
    __device__ T* ptr_##N() const { return &GRAPH_STORAGE_GPU::d_edgeArray.d_##N[edge_index]; }

struct D_Message {
    int index;
    __device__ explicit D_Message(int i) : index(i) {}
    __device__ MVT get_val() const { return GRAPH_STORAGE_GPU::d_messageArray.d_val[index]; }
    __device__ void set_val(MVT v) const { GRAPH_STORAGE_GPU::d_messageArray.d_val[index] = v; }
};

struct D_Edge {
    int edge_index;
    __device__ explicit D_Edge(int i) : edge_index(i) {}
    __device__ int get_srcVertexID() const { return GRAPH_STORAGE_GPU::d_edgeArray.d_srcVertexID[edge_index]; }
    __device__ int get_dstVertexID() const { return GRAPH_STORAGE_GPU::d_edgeArray.d_dstVertexID[edge_index]; }
    __device__ int get_msgDstID() const { return GRAPH_STORAGE_GPU::d_edgeArray.d_msgDstID[edge_index]; }
    // Sends a message along this edge: it lands in the destination's slot (msgDstID).
    __device__ void sendMsg(MVT v) const {
        GRAPH_STORAGE_GPU::d_messageArray.d_val[GRAPH_STORAGE_GPU::d_edgeArray.d_msgDstID[edge_index]] = v;
    }
    MEDUSA_EDGE_FIELDS(MDS_EDGE_ACCESS)
};

struct D_Vertex {
    int index;
    __device__ explicit D_Vertex(int i) : index(i) {}
    __device__ int get_edge_count() const { return GRAPH_STORAGE_GPU::d_vertexArray.d_edge_count[index]; }
    __device__ int get_msg_index() const { return GRAPH_STORAGE_GPU::d_vertexArray.d_msg_index[index]; }
    // Combined message of this vertex's in-edge slots, or the init value without in-edges
    // (SSSP/DeviceDataStructure.h:221-235).
    __device__ MVT get_combined_msg() const {
        const int a = GRAPH_STORAGE_GPU::d_vertexArray.d_msg_index[index];
        const int b = GRAPH_STORAGE_GPU::d_vertexArray.d_msg_index[index + 1];
        return a == b ? GRAPH_STORAGE_GPU::d_init_msg_val : GRAPH_STORAGE_GPU::d_messageArrayBuf[index];
    }
    MEDUSA_VERTEX_FIELDS(MDS_VERTEX_ACCESS)
};

// Out-edges of a vertex, walked in CSR order (the reference offered this on MEG only,
// Algorithm/DeviceDataStructure.h:118-142).
struct EdgeList {
    int vertex_index;
    int current_index;
    int edge_number;
    __device__ explicit EdgeList(int v)
        : vertex_index(v), current_index(GRAPH_STORAGE_GPU::d_vertexArray.d_edge_index[v] - 1), edge_number(0) {}
    __device__ D_Edge getNextEdge() {
        ++current_index;
        ++edge_number;
        return D_Edge(current_index);
    }
    __device__ D_Edge getPreviousEdge() {
        --current_index;
        --edge_number;
        return D_Edge(current_index);
    }
    __device__ int getEdgeCount() const { return GRAPH_STORAGE_GPU::d_vertexArray.d_edge_count[vertex_index]; }
};

// ---- global host state (MultipleGPU/GPUDef.h, MultiGraphStorage.h) -----------------------
struct Medusa_Combiner {
    void init(MedusaCombinerType dt, MedusaCombinerOp op, int gpu_id = 0);
    void combineAllDevice();
};

struct GPUDef {
    VertexArray vertexArray;
    EdgeArray edgeArray;
    D_MessageArray d_messageArray;
    D_MessageArray d_messageArrayBuf;
    D_EdgeArray d_edgeArray;
    D_VertexArray d_vertexArray;
    hipDeviceProp_t device_prop{};
    medusa_graph* graph = nullptr;
    std::vector<uint32_t> esrc;
};

namespace MGLOBAL {
inline GPUDef* gpu_def = nullptr;
inline int super_step = 0;
inline bool toExecute = false;
inline Medusa_Combiner com;
inline MedusaCombinerType combiner_datatype = MEDUSA_COMBINE_FLOAT;
inline MedusaCombinerOp combiner_operator = MEDUSA_COMBINE_ADD;
inline MessageMode message_mode = NO_MESSAGE;
inline MedusaDS graph_ds_type = DS_AA;
inline int num_gpu_to_use = 1;
inline hipStream_t stream = nullptr;  // every launch of the API; blocking, so it orders with hipMemcpy
inline bool device_loop = false;      // Medusa_Run is capturing: wrappers read the loop state
inline const int* d_vertex_order = nullptr;  // vertices by out-degree class, heaviest first (EDGELIST)
inline unsigned run_generation = 1;   // bumped when captured launches may hold stale pointers
inline hipGraphExec_t run_exec = nullptr;
inline unsigned run_exec_generation = 0;
inline const void* run_exec_fn = nullptr;
inline int max_hop = 0;
inline int* total_edge_count = nullptr;
}  // namespace MGLOBAL

namespace medusa_api {
inline hipStream_t stream() { return MGLOBAL::stream; }
inline int grid_size() { return MGLOBAL::gpu_def ? 8 * MGLOBAL::gpu_def[0].device_prop.multiProcessorCount : 2048; }
inline void check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
template <class T>
T* dev_alloc_copy(const T* h, size_t n) {
    T* d = nullptr;
    if (n == 0) return nullptr;
    MEDUSA_HIP_OK(hipMalloc((void**)&d, n * sizeof(T)));
    if (h) MEDUSA_HIP_OK(hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice));
    return d;
}
inline void upload_symbols() {
    GPUDef& G = MGLOBAL::gpu_def[0];
    MEDUSA_HIP_OK(hipMemcpyToSymbol(HIP_SYMBOL(GRAPH_STORAGE_GPU::d_vertexArray), &G.d_vertexArray, sizeof(D_VertexArray)));
    MEDUSA_HIP_OK(hipMemcpyToSymbol(HIP_SYMBOL(GRAPH_STORAGE_GPU::d_edgeArray), &G.d_edgeArray, sizeof(D_EdgeArray)));
    MEDUSA_HIP_OK(hipMemcpyToSymbol(HIP_SYMBOL(GRAPH_STORAGE_GPU::d_messageArray), &G.d_messageArray, sizeof(D_MessageArray)));
    MEDUSA_HIP_OK(hipMemcpyToSymbol(HIP_SYMBOL(GRAPH_STORAGE_GPU::d_messageArrayBuf), &G.d_messageArrayBuf.d_val, sizeof(MVT*)));
}

// ---- functor wrappers (MedusaRT/APIKernelWrappers.h:8-218), grid-stride ------------------
// kLoop: launched inside Medusa_Run's captured loop; the superstep number comes from the device
// and a superstep after the loop ended returns at once.
template <bool kLoop>
__device__ inline bool bsp_live(int& super_step) {
    if constexpr (kLoop) {
        if (!GRAPH_STORAGE_GPU::d_active) return false;
        super_step = GRAPH_STORAGE_GPU::d_super_step;
    }
    return true;
}
template <bool kLoop, class O>
__global__ void vertexProcWrapper(int super_step, O op, int n) {
    if (!bsp_live<kLoop>(super_step)) return;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) op(D_Vertex(v), super_step);
}
template <bool kLoop, class O>
__global__ void edgeProcWrapper(int super_step, O op, int m) {
    if (!bsp_live<kLoop>(super_step)) return;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) op(D_Edge(e), super_step);
}
// EDGELIST: one thread per vertex with its whole out-edge list (the functor's contract), in
// the order of `order`: vertices grouped by out-degree class, heaviest class first, so the
// lanes of a wave walk lists of similar length and the long lists start first.
template <bool kLoop, class O>
__global__ void edgeListProcWrapper(int super_step, O op, int n, const int* __restrict__ order) {
    if (!bsp_live<kLoop>(super_step)) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int v = order[i];
        EdgeList el(v);
        op(D_Vertex(v), el, super_step);
    }
}
template <bool kLoop, class O>
__global__ void messageProcWrapper(int super_step, O op, int slots) {
    if (!bsp_live<kLoop>(super_step)) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < slots; i += gridDim.x * blockDim.x) op(D_Message(i), super_step);
}
// BSP vote and loop control, each one thread on the API's stream.
__global__ inline void bsp_reset_vote(bool v) { GRAPH_STORAGE_GPU::d_toExecute = v; }
__global__ inline void bsp_set_init_msg(MVT v) { GRAPH_STORAGE_GPU::d_init_msg_val = v; }
__global__ inline void bsp_start(int step, int max_step, int vote) {
    GRAPH_STORAGE_GPU::d_active = step < max_step ? 1 : 0;
    GRAPH_STORAGE_GPU::d_super_step = step;
    GRAPH_STORAGE_GPU::d_max_step = max_step;
    GRAPH_STORAGE_GPU::d_vote = vote;
}
__global__ inline void bsp_step_begin() {
    if (GRAPH_STORAGE_GPU::d_active) GRAPH_STORAGE_GPU::d_toExecute = false;
}
__global__ inline void bsp_step_end() {
    if (!GRAPH_STORAGE_GPU::d_active) return;
    const int s = ++GRAPH_STORAGE_GPU::d_super_step;
    if ((GRAPH_STORAGE_GPU::d_vote && !GRAPH_STORAGE_GPU::d_toExecute) || s >= GRAPH_STORAGE_GPU::d_max_step)
        GRAPH_STORAGE_GPU::d_active = 0;
}
template <class T>
__global__ void fill_kernel(T* p, T v, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// Segmented combine of every vertex's in-edge slots: one wave per vertex.  Float ADD
// uses the engine's canonical compensated order (lanes j, j+64, ... Kahan; tree 32..1;
// blocks of 4096 merged left to right), so it is bit-stable; MIN/MAX and integer ADD
// are exact in any order.
__device__ inline void kahan_add_(float& s, float& c, float x) {
    const float y = x - c, t = s + y;
    c = (t - s) - y;
    s = t;
}
__device__ inline void kahan_combine_(float& s, float& c, float s2, float c2) {
    const float t = s + s2, bp = t - s, err = (s - (t - bp)) + (s2 - bp);
    c = (c + c2) - err;
    s = t;
}
template <class T>
__device__ inline T combine_op(T a, T b, int op) {
    switch (op) {
        case MEDUSA_COMBINE_ADD: return a + b;
        case MEDUSA_COMBINE_MUL: return a * b;
        case MEDUSA_COMBINE_MIN: return b < a ? b : a;
        default: return b > a ? b : a;
    }
}
template <class T>
__global__ void combine_kernel(int n, const int* __restrict__ msg_index, const T* __restrict__ msg, T* __restrict__ out,
                               int op) {
    const int lane = threadIdx.x & 63;
    const int waves = gridDim.x * (blockDim.x / 64);
    for (int v = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; v < n; v += waves) {
        const int a = msg_index[v], len = msg_index[v + 1] - a;
        if (len == 0) continue;
        if constexpr (sizeof(T) == 4 && (T)0.5 == 0.5) {
            if (op == MEDUSA_COMBINE_ADD) {
                float as = 0, ac = 0;
                for (int b0 = 0; b0 < len; b0 += 4096) {
                    const int lim = min(len - b0, 4096);
                    float s = 0, c = 0;
                    for (int k = lane; k < lim; k += 64) kahan_add_(s, c, (float)msg[a + b0 + k]);
                    for (int off = 32; off; off >>= 1) {
                        const float s2 = __shfl_down(s, off, 64), c2 = __shfl_down(c, off, 64);
                        kahan_combine_(s, c, s2, c2);
                    }
                    s = __shfl(s, 0, 64);
                    c = __shfl(c, 0, 64);
                    if (b0 == 0) {
                        as = s;
                        ac = c;
                    } else {
                        kahan_combine_(as, ac, s, c);
                    }
                }
                if (lane == 0) out[v] = (T)(as - ac);
                continue;
            }
        }
        bool have = false;
        T acc{};
        for (int k = lane; k < len; k += 64) {
            const T x = msg[a + k];
            acc = have ? combine_op(acc, x, op) : x;
            have = true;
        }
        for (int off = 32; off; off >>= 1) {
            const T o = __shfl_down(acc, off, 64);
            const bool oh = __shfl_down((int)have, off, 64) != 0;
            if (lane + off < 64 && oh) {
                acc = have ? combine_op(acc, o, op) : o;
                have = true;
            }
        }
        if (lane == 0) out[v] = acc;
    }
}
}  // namespace medusa_api

template <EMVAPIType f>
struct FunctorHolder {};

// Launches a wrapper on the API's stream: the loop form inside Medusa_Run's capture, else the
// plain one.
#define MDS_LAUNCH_FUNCTOR(WRAPPER, OP, ...)                                                             \
    do {                                                                                                 \
        if (MGLOBAL::device_loop)                                                                        \
            hipLaunchKernelGGL((medusa_api::WRAPPER<true, OP>), dim3(medusa_api::grid_size()), dim3(256), 0, \
                               medusa_api::stream(), MGLOBAL::super_step, op, __VA_ARGS__);             \
        else                                                                                             \
            hipLaunchKernelGGL((medusa_api::WRAPPER<false, OP>), dim3(medusa_api::grid_size()), dim3(256), 0, \
                               medusa_api::stream(), MGLOBAL::super_step, op, __VA_ARGS__);             \
        medusa_api::check_launch(#WRAPPER);                                                              \
    } while (0)

template <>
struct FunctorHolder<EDGE> {
    template <class OP>
    static void Run(OP op) {
        MDS_LAUNCH_FUNCTOR(edgeProcWrapper, OP, MGLOBAL::gpu_def[0].edgeArray.size);
    }
};
template <>
struct FunctorHolder<EDGE_HT> : FunctorHolder<EDGE> {};
template <>
struct FunctorHolder<VERTEX> {
    template <class OP>
    static void Run(OP op) {
        MDS_LAUNCH_FUNCTOR(vertexProcWrapper, OP, MGLOBAL::gpu_def[0].vertexArray.size);
    }
};
template <>
struct FunctorHolder<EDGELIST> {
    template <class OP>
    static void Run(OP op) {
        MDS_LAUNCH_FUNCTOR(edgeListProcWrapper, OP, MGLOBAL::gpu_def[0].vertexArray.size, MGLOBAL::d_vertex_order);
    }
};
template <>
struct FunctorHolder<MESSAGE> {
    template <class OP>
    static void Run(OP op) {
        MDS_LAUNCH_FUNCTOR(messageProcWrapper, OP, MGLOBAL::gpu_def[0].d_messageArray.size);
    }
};

// Multi-GPU holders (MultipleGPU/MultiAPIHoster.h:28-640).  One GPU has no replicas to gather
// or scatter, so each does what its single-GPU kind does; more GPUs are refused.
template <EMVAPIType f>
struct MGFunctorHolder {};
namespace medusa_api {
inline void require_one_gpu(const char* what) {
    if (MGLOBAL::num_gpu_to_use != 1)
        throw std::runtime_error(std::string(what) + ": this engine runs on one GPU (num_gpu_to_use must be 1)");
}
}  // namespace medusa_api
template <>
struct MGFunctorHolder<MG_EDGE> {
    template <class OP>
    static void Run(OP op) {
        medusa_api::require_one_gpu("MGFunctorHolder<MG_EDGE>");
        FunctorHolder<EDGE>::Run(op);
    }
};
template <>
struct MGFunctorHolder<MGMH_EDGE> {
    template <class OP>
    static void Run(OP op) {
        medusa_api::require_one_gpu("MGFunctorHolder<MGMH_EDGE>");
        FunctorHolder<EDGE>::Run(op);
    }
};
template <>
struct MGFunctorHolder<MG_VERTEX> {
    template <class OP>
    static void Run(OP op) {
        medusa_api::require_one_gpu("MGFunctorHolder<MG_VERTEX>");
        FunctorHolder<VERTEX>::Run(op);
    }
};
template <>
struct MGFunctorHolder<MGMH_VERTEX> {
    template <class OP>
    static void Run(OP op) {
        medusa_api::require_one_gpu("MGFunctorHolder<MGMH_VERTEX>");
        FunctorHolder<VERTEX>::Run(op);
    }
};

// ---- queue API (MultipleGPU/MultiAPIHoster.h:18-23 names the operations; the reference
// gives a body only to EdgeQueue2EdgeQueue, MedusaRT/APIKernelWrappers.h:53-140, compiled
// for MBFS only) ----------------------------------------------------------------------------
// X2Y runs the functor once per item of X and writes Y:
//   X: VQ = a vertex queue, EQ = an edge queue, VA = every vertex, EA = every edge;
//   Y: VQ / AQ = a vertex queue, EQ = an edge queue, Consume = nothing.
// The functor returns whether the item activates its target vertex (a vertex item: itself;
// an edge item: its destination, as SetActive(dstVertexID) in EdgeQueue2EdgeQueue).  An
// activated vertex goes to a vertex queue as itself and to an edge queue as its out-edges
// (CSR edge indices, as EdgeQueue2EdgeQueue enqueues the activated vertex's adjacency).
// Functors: bool operator()(D_Vertex v, int super_step) or (D_Edge e, int super_step); a
// Consume functor may return void.  A vertex activated twice is queued twice: functors that
// must not do so claim the vertex with an atomic (as the reference's ATOMIC variant does).
// The order of a queue's items is unspecified.
enum QAPIType { EQ2EQ, EQ2VQ, EQConsume, VQ2EQ, VQ2VQ, VQConsume, VA2EQ, VA2AQ, EA2EQ, EA2AQ };

struct MedusaQueue {
    int* d_items = nullptr;
    unsigned long long* d_count = nullptr;  // may exceed capacity: then the queue overflowed
    long long capacity = 0;
    bool of_edges = false;
};

namespace medusa_api {
template <class O, class It>
__device__ inline bool activates(O& op, It item, int step) {
    if constexpr (std::is_same<decltype(op(item, step)), void>::value) {
        op(item, step);
        return false;
    } else {
        return static_cast<bool>(op(item, step));
    }
}

// Item i of the input (a queue entry or the index itself) through the functor; returns the
// activated target vertex or -1.
template <bool kEdges, bool kFromQueue, class O>
__device__ inline int queue_visit(O& op, const int* __restrict__ in, unsigned long long i, int step) {
    const int item = kFromQueue ? in[i] : (int)i;
    if constexpr (kEdges) {
        return activates(op, D_Edge(item), step) ? GRAPH_STORAGE_GPU::d_edgeArray.d_dstVertexID[item] : -1;
    } else {
        return activates(op, D_Vertex(item), step) ? item : -1;
    }
}

__device__ inline unsigned long long queue_len(const unsigned long long* cnt, long long cap) {
    const unsigned long long c = *cnt;
    return c < (unsigned long long)cap ? c : (unsigned long long)cap;
}

// Exclusive prefix sum over a 256-thread block (4 waves); every thread gets the total.
__device__ inline unsigned long long block_scan_u64(unsigned long long x, unsigned long long* s_wave,
                                                    unsigned long long& total) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    unsigned long long incl = x;
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned long long y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    if (lane == 63) s_wave[wv] = incl;
    __syncthreads();
    unsigned long long before = 0;
    for (int w = 0; w < wv; ++w) before += s_wave[w];
    total = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
    __syncthreads();  // s_wave is reused by the next call
    return before + incl - x;
}

// X2VQ / X2AQ / XConsume: one slot per activated item, reserved with one atomic per block.
template <bool kEdges, bool kFromQueue, class O>
__global__ __launch_bounds__(256) void queueFilterWrapper(int step, O op, const int* __restrict__ in,
                                                          const unsigned long long* __restrict__ in_count,
                                                          long long in_n, int* __restrict__ out,
                                                          unsigned long long* __restrict__ out_count,
                                                          long long out_cap) {
    __shared__ unsigned long long s_wave[4];
    __shared__ unsigned long long s_base;
    const unsigned long long cnt = kFromQueue ? queue_len(in_count, in_n) : (unsigned long long)in_n;
    for (unsigned long long b = (unsigned long long)blockIdx.x * 256; b < cnt; b += (unsigned long long)gridDim.x * 256) {
        const unsigned long long i = b + threadIdx.x;
        const int t = i < cnt ? queue_visit<kEdges, kFromQueue>(op, in, i, step) : -1;
        if (!out) continue;  // Consume
        unsigned long long total;
        const unsigned long long off = block_scan_u64(t >= 0 ? 1ull : 0ull, s_wave, total);
        if (threadIdx.x == 0) s_base = total ? atomicAdd(out_count, total) : 0ull;
        __syncthreads();
        if (t >= 0 && s_base + off < (unsigned long long)out_cap) out[s_base + off] = t;
        __syncthreads();  // s_base is rewritten in the next round
    }
}

// X2EQ: the out-edges of every activated vertex.  The block reserves its whole output with
// one atomic and writes it cooperatively: output j belongs to the thread whose exclusive
// offset is the last one <= j.
template <bool kEdges, bool kFromQueue, class O>
__global__ __launch_bounds__(256) void queueExpandWrapper(int step, O op, const int* __restrict__ in,
                                                          const unsigned long long* __restrict__ in_count,
                                                          long long in_n, int* __restrict__ out,
                                                          unsigned long long* __restrict__ out_count,
                                                          long long out_cap) {
    __shared__ unsigned long long s_wave[4];
    __shared__ unsigned long long s_off[256];
    __shared__ int s_first[256];
    __shared__ unsigned long long s_base;
    const int* __restrict__ row = GRAPH_STORAGE_GPU::d_vertexArray.d_edge_index;
    const unsigned long long cnt = kFromQueue ? queue_len(in_count, in_n) : (unsigned long long)in_n;
    for (unsigned long long b = (unsigned long long)blockIdx.x * 256; b < cnt; b += (unsigned long long)gridDim.x * 256) {
        const unsigned long long i = b + threadIdx.x;
        const int t = i < cnt ? queue_visit<kEdges, kFromQueue>(op, in, i, step) : -1;
        const int first = t >= 0 ? row[t] : 0;
        const unsigned long long deg = t >= 0 ? (unsigned long long)(row[t + 1] - first) : 0ull;
        unsigned long long total;
        const unsigned long long off = block_scan_u64(deg, s_wave, total);
        s_off[threadIdx.x] = off;
        s_first[threadIdx.x] = first;
        if (threadIdx.x == 0) s_base = total ? atomicAdd(out_count, total) : 0ull;
        __syncthreads();
        for (unsigned long long j = threadIdx.x; j < total; j += 256) {
            int lo = 0, hi = 256;  // last k with s_off[k] <= j; that thread has deg > 0
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_off[mid] <= j) lo = mid;
                else hi = mid;
            }
            const unsigned long long pos = s_base + j;
            if (pos < (unsigned long long)out_cap) out[pos] = s_first[lo] + (int)(j - s_off[lo]);
        }
        __syncthreads();  // LDS is rewritten in the next round
    }
}

inline void queue_set_count(unsigned long long* c, unsigned long long v) {
    hipLaunchKernelGGL(fill_kernel<unsigned long long>, dim3(1), dim3(64), 0, stream(), c, v, 1);
    check_launch("queue count");
}

inline MedusaQueue make_queue(bool edges, long long capacity) {
    if (!MGLOBAL::gpu_def) throw std::runtime_error("queues need InitConfig and InitHostDS first");
    const GPUDef& G = MGLOBAL::gpu_def[0];
    MedusaQueue q;
    q.of_edges = edges;
    q.capacity = capacity >= 0 ? capacity
                               : edges ? G.edgeArray.size : std::max<long long>(G.vertexArray.size, G.edgeArray.size);
    q.d_items = dev_alloc_copy<int>(nullptr, q.capacity > 0 ? (size_t)q.capacity : 1);
    MEDUSA_HIP_OK(hipMalloc((void**)&q.d_count, 128));
    queue_set_count(q.d_count, 0ull);
    return q;
}

template <bool kExpand, bool kEdges, bool kFromQueue, class O>
inline void run_queue_op(O op, const MedusaQueue* in, MedusaQueue* out, const char* what) {
    if (!MGLOBAL::gpu_def) throw std::runtime_error(std::string(what) + ": InitConfig/InitDeviceDS first");
    if (kFromQueue && (!in || !in->d_items || in->of_edges != kEdges))
        throw std::runtime_error(std::string(what) + ": input must be a " + (kEdges ? "edge" : "vertex") + " queue");
    if (out && (!out->d_items || out->of_edges != kExpand))
        throw std::runtime_error(std::string(what) + ": output must be " + (kExpand ? "an edge" : "a vertex") + " queue");
    if (in && out && in->d_items == out->d_items)
        throw std::runtime_error(std::string(what) + ": input and output must be different queues");
    const GPUDef& G = MGLOBAL::gpu_def[0];
    const long long in_n = kFromQueue ? in->capacity : (kEdges ? G.edgeArray.size : G.vertexArray.size);
    int* out_items = out ? out->d_items : nullptr;
    unsigned long long* out_count = out ? out->d_count : nullptr;
    const long long out_cap = out ? out->capacity : 0;
    if (out) queue_set_count(out->d_count, 0ull);
    const int* in_items = kFromQueue ? in->d_items : nullptr;
    const unsigned long long* in_count = kFromQueue ? in->d_count : nullptr;
    if constexpr (kExpand)
        hipLaunchKernelGGL((queueExpandWrapper<kEdges, kFromQueue, O>), dim3(grid_size()), dim3(256), 0, stream(),
                           MGLOBAL::super_step, op, in_items, in_count, in_n, out_items, out_count, out_cap);
    else
        hipLaunchKernelGGL((queueFilterWrapper<kEdges, kFromQueue, O>), dim3(grid_size()), dim3(256), 0, stream(),
                           MGLOBAL::super_step, op, in_items, in_count, in_n, out_items, out_count, out_cap);
    check_launch(what);
}
}  // namespace medusa_api

// Queue capacity in items: by default m for an edge queue and max(n, m) for a vertex queue
// (one item per edge fits).  An edge queue written by X2EQ holds the activated vertices'
// out-edges, which exceed m when vertices are activated more than once: give such a queue
// a larger capacity.  InitHostDS first.
inline MedusaQueue CreateVertexQueue(long long capacity = -1) { return medusa_api::make_queue(false, capacity); }
inline MedusaQueue CreateEdgeQueue(long long capacity = -1) { return medusa_api::make_queue(true, capacity); }
inline void FreeQueue(MedusaQueue& q) {
    MEDUSA_HIP_OK(hipDeviceSynchronize());
    if (q.d_items) (void)hipFree(q.d_items);
    if (q.d_count) (void)hipFree(q.d_count);
    q = MedusaQueue{};
}
// Number of items; throws if the last operation that wrote the queue overflowed it.
inline long long QueueSize(const MedusaQueue& q) {
    unsigned long long c = 0;
    MEDUSA_HIP_OK(hipMemcpy(&c, q.d_count, sizeof(c), hipMemcpyDeviceToHost));
    if (c > (unsigned long long)q.capacity)
        throw std::runtime_error("queue overflow: " + std::to_string(c) + " items for capacity " +
                                 std::to_string(q.capacity));
    return (long long)c;
}
// Replaces the queue's contents with ids[0, count) (vertex ids or CSR edge indices).
inline void QueueAssign(MedusaQueue& q, const int* ids, long long count) {
    if (count < 0 || count > q.capacity) throw std::runtime_error("QueueAssign: count exceeds the queue's capacity");
    const int limit = q.of_edges ? MGLOBAL::gpu_def[0].edgeArray.size : MGLOBAL::gpu_def[0].vertexArray.size;
    for (long long i = 0; i < count; ++i)
        if (ids[i] < 0 || ids[i] >= limit) throw std::runtime_error("QueueAssign: id out of range");
    if (count) MEDUSA_HIP_OK(hipMemcpy(q.d_items, ids, (size_t)count * sizeof(int), hipMemcpyHostToDevice));
    medusa_api::queue_set_count(q.d_count, (unsigned long long)count);
}
inline std::vector<int> QueueItems(const MedusaQueue& q) {
    std::vector<int> v((size_t)QueueSize(q));
    if (!v.empty()) MEDUSA_HIP_OK(hipMemcpy(v.data(), q.d_items, v.size() * sizeof(int), hipMemcpyDeviceToHost));
    return v;
}

template <QAPIType q>
struct QFunctorHolder {};
// Run(op, in, out) for queue inputs, Run(op, out) for the whole vertex or edge array.
#define MDS_QUEUE_OP(Q, EXPAND, EDGES)                                                                    \
    template <>                                                                                           \
    struct QFunctorHolder<Q> {                                                                            \
        template <class OP>                                                                               \
        static void Run(OP op, const MedusaQueue& in, MedusaQueue& out) {                                 \
            medusa_api::run_queue_op<EXPAND, EDGES, true>(op, &in, &out, #Q);                              \
        }                                                                                                 \
    };
#define MDS_ARRAY_OP(Q, EXPAND, EDGES)                                                                    \
    template <>                                                                                           \
    struct QFunctorHolder<Q> {                                                                            \
        template <class OP>                                                                               \
        static void Run(OP op, MedusaQueue& out) {                                                        \
            medusa_api::run_queue_op<EXPAND, EDGES, false>(op, nullptr, &out, #Q);                         \
        }                                                                                                 \
    };
MDS_QUEUE_OP(EQ2EQ, true, true)
MDS_QUEUE_OP(EQ2VQ, false, true)
MDS_QUEUE_OP(VQ2EQ, true, false)
MDS_QUEUE_OP(VQ2VQ, false, false)
MDS_ARRAY_OP(VA2EQ, true, false)
MDS_ARRAY_OP(VA2AQ, false, false)
MDS_ARRAY_OP(EA2EQ, true, true)
MDS_ARRAY_OP(EA2AQ, false, true)
#undef MDS_QUEUE_OP
#undef MDS_ARRAY_OP
template <>
struct QFunctorHolder<EQConsume> {
    template <class OP>
    static void Run(OP op, const MedusaQueue& in) {
        medusa_api::run_queue_op<false, true, true>(op, &in, nullptr, "EQConsume");
    }
};
template <>
struct QFunctorHolder<VQConsume> {
    template <class OP>
    static void Run(OP op, const MedusaQueue& in) {
        medusa_api::run_queue_op<false, false, true>(op, &in, nullptr, "VQConsume");
    }
};

inline void Medusa_Combiner::init(MedusaCombinerType dt, MedusaCombinerOp op, int) {
    MGLOBAL::combiner_datatype = dt;
    MGLOBAL::combiner_operator = op;
}
inline void Medusa_Combiner::combineAllDevice() {
    GPUDef& G = MGLOBAL::gpu_def[0];
    if (MGLOBAL::message_mode != EDGE_MESSAGE) throw std::runtime_error("combineAllDevice needs EDGE_MESSAGE mode");
    hipLaunchKernelGGL(medusa_api::combine_kernel<MVT>, dim3(medusa_api::grid_size()), dim3(256), 0, medusa_api::stream(),
                       G.vertexArray.size, G.d_vertexArray.d_msg_index, (const MVT*)G.d_messageArray.d_val,
                       G.d_messageArrayBuf.d_val, (int)MGLOBAL::combiner_operator);
    medusa_api::check_launch("combine_kernel");
}

// ---- PublicAPI (MultipleGPU/MultiPublicAPI.h) ------------------------------------------------
// The graph layout is chosen at compile time as in the reference (define one of MEG, AA, HY,
// ELL; AA when none is).  AA, MEG and HY all run on the CSR/CSC built here: MEG and HY sort
// vertices by degree in the reference, and here every EDGELIST functor visits vertices grouped
// by out-degree class whatever the layout, with vertex ids unchanged, so results are the same.
// The reference's InitHostDS has no ELL builder (it exits), so ELL is refused here too.
inline MedusaDS compiled_ds() {
#if defined(ELL)
    return DS_ELL;
#elif defined(HY)
    return DS_HY;
#elif defined(MEG)
    return DS_MEG;
#else
    return DS_AA;
#endif
}

inline void InitConfig(const char* gpus, const char* hops, MessageMode mm, bool exe) {
    MGLOBAL::num_gpu_to_use = gpus ? atoi(gpus) : 1;
    MGLOBAL::max_hop = hops ? atoi(hops) : 0;
    if (MGLOBAL::num_gpu_to_use != 1)
        throw std::runtime_error("InitConfig: this engine runs on one GPU (multi-GPU partitioning is not built)");
    MGLOBAL::graph_ds_type = compiled_ds();
    if (MGLOBAL::graph_ds_type == DS_ELL)
        throw std::runtime_error("InitConfig: the ELL layout has no builder (as in the reference); use AA, MEG or HY");
    MGLOBAL::message_mode = mm;
    MGLOBAL::toExecute = exe;
    MGLOBAL::super_step = 0;
    delete[] MGLOBAL::gpu_def;
    MGLOBAL::gpu_def = new GPUDef[1];
    int dev = 0;
    MEDUSA_HIP_OK(hipGetDevice(&dev));
    MEDUSA_HIP_OK(hipGetDeviceProperties(&MGLOBAL::gpu_def[0].device_prop, dev));
    if (!MGLOBAL::stream) MEDUSA_HIP_OK(hipStreamCreate(&MGLOBAL::stream));
}

// Builds host SoA arrays and the HBM graph (CSR in the reference's edge order, CSC, slot map).
inline void InitHostDS(GraphPartition* gp) {
    GPUDef& G = MGLOBAL::gpu_def[0];
    GraphIR& ir = gp[0].graph;
    std::vector<uint32_t> row_ptr, col;
    std::vector<double> w;
    ir.toCSR(row_ptr, col, w);
    const int n = ir.vertexNum, m = ir.totalEdgeCount;
    if (G.graph) medusa_graph_destroy(G.graph);
    if (medusa_graph_create((uint32_t)n, (uint64_t)m, row_ptr.data(), col.data(), nullptr, &G.graph) != MEDUSA_OK)
        throw std::runtime_error(std::string("InitHostDS: ") + medusa_last_error());
    delete[] MGLOBAL::total_edge_count;
    MGLOBAL::total_edge_count = new int[1]{m};
    VertexArray& V = G.vertexArray;
    V.size = n;
    V.edge_count = new int[n];
    V.edge_index = new int[n + 1];
    V.msg_index = new int[n + 1];
    for (int v = 0; v <= n; ++v) V.edge_index[v] = (int)row_ptr[v];
    for (int v = 0; v < n; ++v) V.edge_count[v] = (int)(row_ptr[v + 1] - row_ptr[v]);
#define MDS_HOST_ALLOC_V(T, N) V.N = new T[n > 0 ? n : 1]();
    MEDUSA_VERTEX_FIELDS(MDS_HOST_ALLOC_V)
#undef MDS_HOST_ALLOC_V
    const uint32_t *d_row, *d_col, *d_cptr, *d_slot;
    medusa_graph_device_arrays(G.graph, &d_row, &d_col, nullptr, &d_cptr, nullptr, &d_slot);
    std::vector<uint32_t> cptr(n + 1), slot(m);
    MEDUSA_HIP_OK(hipMemcpy(cptr.data(), d_cptr, (n + 1) * 4ull, hipMemcpyDeviceToHost));
    if (m) MEDUSA_HIP_OK(hipMemcpy(slot.data(), d_slot, m * 4ull, hipMemcpyDeviceToHost));
    for (int v = 0; v <= n; ++v) V.msg_index[v] = (int)cptr[v];
    EdgeArray& E = G.edgeArray;
    E.size = m;
    E.srcVertexID = new int[m > 0 ? m : 1];
    E.dstVertexID = new int[m > 0 ? m : 1];
    E.msgDstID = new int[m > 0 ? m : 1];
    G.esrc.resize(m);
    for (int u = 0; u < n; ++u)
        for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; ++e) {
            E.srcVertexID[e] = u;
            G.esrc[e] = (uint32_t)u;
            E.dstVertexID[e] = (int)col[e];
            E.msgDstID[e] = (int)slot[e];
        }
    // user edge fields: a field named `weight` starts as the graph weight, others as zero
#define MDS_HOST_ALLOC_E(T, N)                                                               \
    E.N = new T[m > 0 ? m : 1]();                                                            \
    if (std::string(#N) == "weight")                                                         \
        for (int e = 0; e < m; ++e) E.N[e] = (T)w[e];
    MEDUSA_EDGE_FIELDS(MDS_HOST_ALLOC_E)
#undef MDS_HOST_ALLOC_E
}

// Counts vertices without out-edges (MultiPublicAPI.h:103-133).
inline int CheckOutgoingEdge(GraphPartition*) {
    const VertexArray& V = MGLOBAL::gpu_def[0].vertexArray;
    int c = 0;
    for (int v = 0; v < V.size; ++v) c += V.edge_count[v] == 0;
    return c;
}

inline void D_VertexArray_Fill(D_VertexArray& D, const VertexArray& V) {
    GPUDef& G = MGLOBAL::gpu_def[0];
    const uint32_t *d_row, *d_cptr;
    medusa_graph_device_arrays(G.graph, &d_row, nullptr, nullptr, &d_cptr, nullptr, nullptr);
    D.size = V.size;
    D.d_edge_index = reinterpret_cast<const int*>(d_row);
    D.d_msg_index = reinterpret_cast<const int*>(d_cptr);
    if (!D.d_edge_count) D.d_edge_count = medusa_api::dev_alloc_copy(V.edge_count, V.size);
#define MDS_FILL_V(T, N)                                                                             \
    if (!D.d_##N) D.d_##N = medusa_api::dev_alloc_copy<T>(nullptr, V.size > 0 ? V.size : 1);         \
    if (V.size) MEDUSA_HIP_OK(hipMemcpy(D.d_##N, V.N, V.size * sizeof(T), hipMemcpyHostToDevice));
    MEDUSA_VERTEX_FIELDS(MDS_FILL_V)
#undef MDS_FILL_V
}
inline void D_EdgeArray_Fill(D_EdgeArray& D, const EdgeArray& E) {
    GPUDef& G = MGLOBAL::gpu_def[0];
    const uint32_t *d_col, *d_slot;
    medusa_graph_device_arrays(G.graph, nullptr, &d_col, nullptr, nullptr, nullptr, &d_slot);
    D.size = E.size;
    D.d_dstVertexID = reinterpret_cast<const int*>(d_col);
    D.d_msgDstID = reinterpret_cast<const int*>(d_slot);
    if (!D.d_srcVertexID) D.d_srcVertexID = reinterpret_cast<const int*>(medusa_api::dev_alloc_copy(G.esrc.data(), E.size));
#define MDS_FILL_E(T, N)                                                                             \
    if (!D.d_##N) D.d_##N = medusa_api::dev_alloc_copy<T>(nullptr, E.size > 0 ? E.size : 1);         \
    if (E.size) MEDUSA_HIP_OK(hipMemcpy(D.d_##N, E.N, E.size * sizeof(T), hipMemcpyHostToDevice));
    MEDUSA_EDGE_FIELDS(MDS_FILL_E)
#undef MDS_FILL_E
}

// Uploads vertex/edge fields, allocates the message buffers and binds the device symbols
// (MultiPublicAPI.h:144-308).
// Vertices grouped by out-degree class (floor(log2(degree)) + 1, 0 for none), heaviest class
// first, ids ascending inside a class: the visiting order of EDGELIST functors.
inline std::vector<int> degree_order(const VertexArray& V) {
    constexpr int kClasses = 33;
    std::vector<int> cls(V.size), count(kClasses + 1, 0);
    for (int v = 0; v < V.size; ++v) {
        const unsigned d = (unsigned)V.edge_count[v];
        cls[v] = d ? 32 - __builtin_clz(d) : 0;
        ++count[kClasses - 1 - cls[v] + 1];
    }
    for (int c = 1; c <= kClasses; ++c) count[c] += count[c - 1];
    std::vector<int> order(V.size);
    for (int v = 0; v < V.size; ++v) order[count[kClasses - 1 - cls[v]]++] = v;
    return order;
}

inline void InitDeviceDS() {
    GPUDef& G = MGLOBAL::gpu_def[0];
    D_VertexArray_Fill(G.d_vertexArray, G.vertexArray);
    D_EdgeArray_Fill(G.d_edgeArray, G.edgeArray);
    if (MGLOBAL::d_vertex_order) MEDUSA_HIP_OK(hipFree((void*)MGLOBAL::d_vertex_order));
    const std::vector<int> order = degree_order(G.vertexArray);
    MGLOBAL::d_vertex_order = medusa_api::dev_alloc_copy(order.data(), order.size());
    ++MGLOBAL::run_generation;
    int slots = 0, buf = 0;
    if (MGLOBAL::message_mode == EDGE_MESSAGE) {
        slots = G.edgeArray.size;
        buf = G.vertexArray.size;
    } else if (MGLOBAL::message_mode == VERTEX_MESSAGE) {
        slots = buf = G.vertexArray.size;
    }
    if (slots && !G.d_messageArray.d_val) {
        G.d_messageArray.d_val = medusa_api::dev_alloc_copy<MVT>(nullptr, slots);
        G.d_messageArray.size = slots;
    }
    if (buf && !G.d_messageArrayBuf.d_val) {
        G.d_messageArrayBuf.d_val = medusa_api::dev_alloc_copy<MVT>(nullptr, buf);
        G.d_messageArrayBuf.size = buf;
    }
    MGLOBAL::super_step = 0;
    medusa_api::upload_symbols();
    MEDUSA_HIP_OK(hipMemcpyToSymbol(HIP_SYMBOL(GRAPH_STORAGE_GPU::d_toExecute), &MGLOBAL::toExecute, sizeof(bool)));
}

// Re-uploads host vertex/edge fields for another run (MultiPublicAPI.h:310-326).
inline void ResetGraphData() {
    GPUDef& G = MGLOBAL::gpu_def[0];
    ++MGLOBAL::run_generation;
    D_VertexArray_Fill(G.d_vertexArray, G.vertexArray);
    D_EdgeArray_Fill(G.d_edgeArray, G.edgeArray);
    medusa_api::upload_symbols();
    MGLOBAL::super_step = 0;
}

// Downloads the vertex fields into the host arrays (D_VertexArray::Dump, VertexDataType.cu:115-131).
inline void DumpVertexData() {
    GPUDef& G = MGLOBAL::gpu_def[0];
    MEDUSA_HIP_OK(hipDeviceSynchronize());
#define MDS_DUMP_V(T, N) \
    if (G.vertexArray.size) MEDUSA_HIP_OK(hipMemcpy(G.vertexArray.N, G.d_vertexArray.d_##N, G.vertexArray.size * sizeof(T), hipMemcpyDeviceToHost));
    MEDUSA_VERTEX_FIELDS(MDS_DUMP_V)
#undef MDS_DUMP_V
}

// Fills every message slot with init_val and records it as the value of vertices without
// messages (MultiPublicAPI.h:344-376, Algorithm/InitMessage.h:4-9).
inline void InitMessageBuffer(Message init_val) {
    GPUDef& G = MGLOBAL::gpu_def[0];
    if (G.d_messageArray.size) {
        hipLaunchKernelGGL(medusa_api::fill_kernel<MVT>, dim3(medusa_api::grid_size()), dim3(256), 0, medusa_api::stream(),
                           G.d_messageArray.d_val, init_val.val, G.d_messageArray.size);
        medusa_api::check_launch("InitMessageBuffer");
    }
    // a kernel argument rather than a host copy, so Medusa_Run can capture it
    hipLaunchKernelGGL(medusa_api::bsp_set_init_msg, dim3(1), dim3(1), 0, medusa_api::stream(), init_val.val);
    medusa_api::check_launch("InitMessageBuffer");
}

// Swaps the message array and its buffer, so messages become readable without combining
// (MultiPublicAPI.h:386-401); both must have the same size (VERTEX_MESSAGE mode).
inline void NULL_Combine() {
    GPUDef& G = MGLOBAL::gpu_def[0];
    if (G.d_messageArray.size != G.d_messageArrayBuf.size)
        throw std::runtime_error("NULL_Combine: message array and buffer differ in size (use VERTEX_MESSAGE)");
    std::swap(G.d_messageArray, G.d_messageArrayBuf);
    medusa_api::upload_symbols();
    ++MGLOBAL::run_generation;
}

// BSP vote (MultipleGPU/MultiUtilities.h:48-69): the reset is a one-thread kernel on the API's
// stream and the read one copy, so a host-driven superstep waits for the GPU once.
inline void ResetToExecute() {
    hipLaunchKernelGGL(medusa_api::bsp_reset_vote, dim3(1), dim3(1), 0, medusa_api::stream(), MGLOBAL::toExecute);
    medusa_api::check_launch("ResetToExecute");
}
inline bool RetriveToExecute() {
    bool v = false;
    MEDUSA_HIP_OK(hipMemcpyFromSymbolAsync(&v, HIP_SYMBOL(GRAPH_STORAGE_GPU::d_toExecute), sizeof(bool), 0,
                                           hipMemcpyDeviceToHost, medusa_api::stream()));
    MEDUSA_HIP_OK(hipStreamSynchronize(medusa_api::stream()));
    return v;
}

// The BSP loop on the device (no reference counterpart; the reference's host loop is
// ResetToExecute / Medusa_Exec / RetriveToExecute per superstep, which still works).  Runs
// supersteps of `exec` (the application's Medusa_Exec) until the vote ends the run (with
// `vote`; functors call Medusa_Continue to go on) or `max_supersteps` have run.  Batches of
// kRunBatch supersteps are captured once as a hipGraph, each superstep bracketed by one-thread
// kernels that reset the vote and advance the superstep on the device; supersteps after the
// end return at once, and the host reads the loop state once per batch.  `exec` must enqueue
// the same work every superstep through this API (FunctorHolder, MGFunctorHolder,
// InitMessageBuffer, combineAllDevice); work that synchronises with the host cannot be
// captured and makes Medusa_Run throw (use the host loop for it).  Without `vote` there is
// nothing to wait for: the supersteps are enqueued back to back and waited for once.  Returns
// the supersteps run; MGLOBAL::super_step is advanced by as many.
constexpr int kRunBatch = 8;
inline int Medusa_Run(void (*exec)(), int max_supersteps = INT_MAX, bool vote = true) {
    hipStream_t s = medusa_api::stream();
    const int start = MGLOBAL::super_step;
    const long long last = (long long)start + (max_supersteps > 0 ? max_supersteps : 0);
    const int max_step = last > INT_MAX ? INT_MAX : (int)last;
    if (!vote) {  // a fixed number of supersteps: nothing to wait for, so they go back to back
        for (; MGLOBAL::super_step < max_step; ++MGLOBAL::super_step) exec();
        MEDUSA_HIP_OK(hipStreamSynchronize(s));
        return MGLOBAL::super_step - start;
    }
    if (!MGLOBAL::run_exec || MGLOBAL::run_exec_generation != MGLOBAL::run_generation ||
        MGLOBAL::run_exec_fn != (const void*)exec) {
        if (MGLOBAL::run_exec) (void)hipGraphExecDestroy(MGLOBAL::run_exec);
        MGLOBAL::run_exec = nullptr;
        MEDUSA_HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        std::string err;
        MGLOBAL::device_loop = true;
        try {
            for (int k = 0; k < kRunBatch; ++k) {
                hipLaunchKernelGGL(medusa_api::bsp_step_begin, dim3(1), dim3(1), 0, s);
                exec();
                hipLaunchKernelGGL(medusa_api::bsp_step_end, dim3(1), dim3(1), 0, s);
            }
        } catch (const std::exception& e) {
            err = e.what();
        }
        MGLOBAL::device_loop = false;
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(s, &graph);
        if (e == hipSuccess && err.empty()) e = hipGraphInstantiate(&MGLOBAL::run_exec, graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        if (!err.empty() || e != hipSuccess)
            throw std::runtime_error("Medusa_Run: the superstep could not be captured (" +
                                     (err.empty() ? std::string(hipGetErrorString(e)) : err) +
                                     "); Medusa_Exec must only enqueue work through the API");
        MGLOBAL::run_exec_generation = MGLOBAL::run_generation;
        MGLOBAL::run_exec_fn = (const void*)exec;
    }
    hipLaunchKernelGGL(medusa_api::bsp_start, dim3(1), dim3(1), 0, s, start, max_step, vote ? 1 : 0);
    medusa_api::check_launch("Medusa_Run");
    int state[2] = {1, start};  // active, superstep
    while (state[0]) {
        MEDUSA_HIP_OK(hipGraphLaunch(MGLOBAL::run_exec, s));
        MEDUSA_HIP_OK(hipMemcpyFromSymbolAsync(&state[0], HIP_SYMBOL(GRAPH_STORAGE_GPU::d_active), sizeof(int), 0,
                                               hipMemcpyDeviceToHost, s));
        MEDUSA_HIP_OK(hipMemcpyFromSymbolAsync(&state[1], HIP_SYMBOL(GRAPH_STORAGE_GPU::d_super_step), sizeof(int), 0,
                                               hipMemcpyDeviceToHost, s));
        MEDUSA_HIP_OK(hipStreamSynchronize(s));
    }
    MGLOBAL::super_step = state[1];
    return state[1] - start;
}

// Frees the device state of the run.
inline void ReleaseDeviceDS() {
    if (!MGLOBAL::gpu_def) return;
    GPUDef& G = MGLOBAL::gpu_def[0];
    MEDUSA_HIP_OK(hipDeviceSynchronize());
    if (G.d_vertexArray.d_edge_count) (void)hipFree((void*)G.d_vertexArray.d_edge_count);
    if (G.d_edgeArray.d_srcVertexID) (void)hipFree((void*)G.d_edgeArray.d_srcVertexID);
#define MDS_FREE_V(T, N) if (G.d_vertexArray.d_##N) (void)hipFree(G.d_vertexArray.d_##N);
    MEDUSA_VERTEX_FIELDS(MDS_FREE_V)
#undef MDS_FREE_V
#define MDS_FREE_E(T, N) if (G.d_edgeArray.d_##N) (void)hipFree(G.d_edgeArray.d_##N);
    MEDUSA_EDGE_FIELDS(MDS_FREE_E)
#undef MDS_FREE_E
    if (G.d_messageArray.d_val) (void)hipFree(G.d_messageArray.d_val);
    if (G.d_messageArrayBuf.d_val) (void)hipFree(G.d_messageArrayBuf.d_val);
    if (G.graph) medusa_graph_destroy(G.graph);
    G.graph = nullptr;
    if (MGLOBAL::d_vertex_order) (void)hipFree((void*)MGLOBAL::d_vertex_order);
    MGLOBAL::d_vertex_order = nullptr;
    if (MGLOBAL::run_exec) (void)hipGraphExecDestroy(MGLOBAL::run_exec);
    MGLOBAL::run_exec = nullptr;
    ++MGLOBAL::run_generation;
    G.d_vertexArray = D_VertexArray{};
    G.d_edgeArray = D_EdgeArray{};
    G.d_messageArray = D_MessageArray{};
    G.d_messageArrayBuf = D_MessageArray{};
}
