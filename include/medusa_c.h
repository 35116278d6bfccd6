/*
 * medusa_c.h — C ABI of the MI355X (gfx950) BSP superstep engine.
 *
 * Plain pointers and sizes only.  Every entry point returns an int status
 * (MEDUSA_OK == 0, negative on error) and never exits the process; the text
 * of the last error on the calling thread is returned by medusa_last_error().
 *
 * Threads: calls on one graph handle are serialised by a lock inside the handle, so
 * several host threads may share a handle; different handles run independently.  A
 * handle must not be destroyed while another thread uses it, and a policy must not be
 * trained while a run uses it.
 *
 * Output pointers (levels, dist, rank) may point to host memory or to device
 * memory of the graph's GPU; the library detects which and copies the final
 * result there.  Graph inputs are host arrays that are uploaded once by
 * medusa_graph_create(); every algorithm then runs on HBM-resident data.
 *
 * The reference has no library boundary: the app, runtime templates and
 * globals are compiled into one executable (HowToBuild:18-21).  Each entry
 * point below names the reference interface whose job it takes over.
 */
#ifndef MEDUSA_C_H
#define MEDUSA_C_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
/* Reference: printf + exit(-1) everywhere (e.g. MedusaRT/GraphConverter.cu:19-23,
 * MultipleGPU/MultiGraphStorage.cu:31-47); here errors are returned instead. */
#define MEDUSA_OK            0
#define MEDUSA_ERR_ARG      -1  /* invalid argument (bad id, size, null) */
#define MEDUSA_ERR_HIP      -2  /* HIP runtime error (no device, launch failure) */
#define MEDUSA_ERR_NOMEM    -3  /* host or device allocation failed */
#define MEDUSA_ERR_STATE    -4  /* call not valid for this graph (e.g. no weights) */
#define MEDUSA_ERR_INTERNAL -5  /* invariant violated inside the library */

/* BFS level of a vertex never reached; reference MVT_Init_Value (Algorithm/Message.h:10). */
#define MEDUSA_LEVEL_UNREACHED (-1)
/* SSSP distance of a vertex never reached (reference used MEDUSA_MAX/FLT_MAX floats,
 * SSSP/Configuration.h:6, SSSP/UserOperators.h:217). */
#define MEDUSA_DIST_UNREACHED 0xFFFFFFFFu
/* Multi-source BFS level of a (source, vertex) pair never reached. */
#define MEDUSA_MBFS_UNREACHED 0xFFu

/* BFS / MBFS / SSSP direction (opts.direction). */
#define MEDUSA_DIR_AUTO 0  /* heuristic (or the learned policy when opts.policy is set) */
#define MEDUSA_DIR_PUSH 1  /* top-down over CSR every superstep */
#define MEDUSA_DIR_PULL 2  /* bottom-up over CSC every superstep */

/* PageRank kernel variant (opts.variant); all variants give bit-identical ranks. */
#define MEDUSA_PR_DEFAULT 0 /* the engine's choice: MEDUSA_PR_BINNED, or MEDUSA_PR_PULL on a graph whose
                               binned layout cannot be planned */
#define MEDUSA_PR_PUSH 1   /* scatter into per-in-edge message slots, then combine (reference style) */
#define MEDUSA_PR_BINNED 2 /* message array blocked by source chunk and destination bin: streaming
                              writes of one message per out-edge, then a combine from LDS per bin */
#define MEDUSA_PR_PULL 3   /* gather contributions over the in-edge (CSC) layout */

typedef struct medusa_graph medusa_graph;   /* opaque; owns all device memory of one graph */
typedef struct medusa_policy medusa_policy; /* opaque; learned per-superstep scheduler */

/* Per-run options.  Zero-initialised means "defaults". */
typedef struct medusa_opts {
    uint32_t timing;       /* nonzero: record hipEvents around every launch of the dominant kernel */
    int32_t  direction;    /* MEDUSA_DIR_* for BFS/MBFS/SSSP */
    int32_t  variant;      /* MEDUSA_PR_* for PageRank */
    uint32_t delta;        /* SSSP bucket width; 0 = engine default; 0xFFFFFFFF = no buckets (Bellman-Ford) */
    uint32_t large_degree; /* push frontier: vertices above this out-degree get a workgroup; 0 = default */
    uint32_t trace;        /* nonzero: keep a per-superstep trace, read back with medusa_trace_* */
    const medusa_policy* policy; /* learned scheduler; NULL = plain engine */
    uint32_t use_graph;    /* nonzero: run the fixed superstep sequence as one captured hipGraph (PageRank) */
    uint32_t grid;         /* workgroups of the BFS/MBFS frontier-expansion launch; 0 = default (or tuned) */
    uint32_t far_limit;    /* SSSP: far-queue entries kept before it is rebuilt from the distances; 0 = m+n */
    uint32_t local_rounds; /* SSSP: while the near queue is short, each workgroup relaxes
                              up to this many more rounds of the vertices it lowered itself in the same
                              superstep; 0 = default (32 when the mean out-degree is at most 4,
                              else none), 0xFFFFFFFF = none, at most 64 */
    uint32_t reserved[4];
} medusa_opts;

/* Per-run statistics.  Reference: printf of wall time / steps (Algorithm/main.cu:316,355,
 * SSSP/main.cu:391-392, MPG/main.cu:173). */
typedef struct medusa_stats {
    double   device_ms;        /* hipEvent time of the superstep loop on HBM-resident data */
    double   kernel_ms;        /* summed duration of the dominant kernel's launches (opts.timing) */
    uint64_t kernel_launches;  /* launches of the dominant kernel */
    uint64_t kernel_bytes;     /* algorithmic bytes moved by those launches (byte model) */
    uint64_t alg_bytes;        /* algorithmic bytes of the whole run (byte model) */
    uint64_t edges_traversed;  /* GTEPS numerator: arcs whose source is reached (BFS/SSSP/MBFS: summed
                                  over sources), steps*m for PageRank */
    uint32_t supersteps;
    uint32_t push_steps;
    uint32_t pull_steps;
    uint32_t reached;          /* vertices reached (BFS/SSSP), source-vertex pairs (MBFS), n (PR) */
    double   residual_l1;      /* PageRank: FP64 L1 distance between the last two iterates */
    uint32_t reserved[8];
} medusa_stats;

/* ---- library ----------------------------------------------------------- */
const char* medusa_last_error(void);
const char* medusa_version(void);
/* Number of visible GPUs (0 on a machine without one). */
int medusa_device_count(int* count);

/* ---- graph storage ------------------------------------------------------
 * Replaces InitHostDS + InitDeviceDS (MultipleGPU/MultiPublicAPI.h:43-96, :144-308),
 * VertexArray::build (Algorithm/VertexDataType.cu:60-80: edge_index = CSR row offsets)
 * and EdgeArray::buildAA (SSSP/EdgeDataType.cu:138-186: CSR order, msg_index = CSC
 * offsets, msgDstID = CSR-edge -> CSC-slot map, slots of a destination ordered by
 * source).  Input is CSR: row_ptr[n+1], col[m], weight[m] (weight may be NULL).
 * In-edges of each vertex are kept in ascending source order, duplicates adjacent. */
int medusa_graph_create(uint32_t n, uint64_t m, const uint32_t* row_ptr, const uint32_t* col,
                        const uint32_t* weight, medusa_graph** out);
/* Replaces ResetGraphData's re-upload (MultiPublicAPI.h:310-326): frees everything. */
int medusa_graph_destroy(medusa_graph* g);
int medusa_graph_info(const medusa_graph* g, uint32_t* n, uint64_t* m, uint32_t* weighted);
/* Declares the handle a packed batch of `count` independent graphs: graph k owns the
 * vertex range [graph_begin[k], graph_begin[k+1]) (graph_begin[0] = 0, graph_begin[count]
 * = n, every range non-empty) and no edge may leave its range.  PageRank then normalises
 * each graph by its own vertex count and dangling mass (ranks bit-identical to a run on
 * each graph alone); BFS from one source per graph (medusa_bfs_multi) gives each graph's
 * own levels; all graphs share every launch.  No reference counterpart (the reference runs
 * one graph per process; its GraphPartition ranges, MultipleGPU/PartitionManager.cu:116-534,
 * split one graph across GPUs instead).  count = 1 with {0, n} restores a plain graph. */
int medusa_graph_set_batch(medusa_graph* g, uint32_t count, const uint32_t* graph_begin);
/* Number of packed graphs and up to `capacity` entries of graph_begin (count+1 in all). */
int medusa_graph_batch(const medusa_graph* g, uint32_t* count, uint32_t* graph_begin, uint32_t capacity);
/* Device addresses of the CSR/CSC arrays (for header-only functor kernels compiled by the
 * caller).  Replaces the __constant__ D_VertexArray/D_EdgeArray symbols
 * (Algorithm/DeviceDataStructure.h:7-24).  Any pointer argument may be NULL. */
int medusa_graph_device_arrays(const medusa_graph* g, const uint32_t** row_ptr, const uint32_t** col,
                               const uint32_t** weight, const uint32_t** csc_ptr,
                               const uint32_t** csc_src, const uint32_t** msg_slot);

/* ---- algorithms (one BSP run each) ---------------------------------------
 * BFS: replaces the BSP loop of Algorithm/main.cu:283-302 driving
 * Medusa_Exec -> FunctorHolder<EDGE>::Run(UpdateLevel) (Algorithm/UserOperators.h:17-58,
 * MultipleGPU/MultiAPIHoster.h:657-678) with ResetToExecute/RetriveToExecute
 * (MultipleGPU/MultiUtilities.h:48-69).  levels[n]: int32, -1 unreached. */
int medusa_bfs(medusa_graph* g, uint32_t source, int32_t* levels, const medusa_opts* opts,
               medusa_stats* stats);
/* BFS from `count` sources at once (duplicates allowed): every source starts at level 0
 * and levels[v] is the distance from the nearest one.  Same loop as medusa_bfs (which is
 * the count = 1 case); on a packed batch with one source per graph it runs every graph's
 * BFS in the same supersteps. */
int medusa_bfs_multi(medusa_graph* g, const uint32_t* sources, uint32_t count, int32_t* levels,
                     const medusa_opts* opts, medusa_stats* stats);

/* BFS from one source with no device allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  levels[n] must be device memory of the graph's
 * GPU; the run writes it in place (the same levels as medusa_bfs, heuristic direction).  The
 * frontier queues, bitmap and control block live in one statically sized device array per
 * GPU, held until the process ends, so n may not exceed MEDUSA_NOALLOC_MAX_VERTICES
 * (MEDUSA_ERR_ARG otherwise).  Every call on a GPU shares that array: calls that overlap on one
 * GPU (from several threads or handles) race and must be serialised by the caller.  The graph
 * handle itself is allocated by medusa_graph_create beforehand. */
#define MEDUSA_NOALLOC_MAX_VERTICES (1u << 21)
int medusa_bfs_noalloc(medusa_graph* g, uint32_t source, int32_t* levels, medusa_stats* stats);

/* SSSP: replaces SSSP/UserOperators.h:200-264 (InitMessageBuffer(FLT_MAX), ssspEdgeProcAA
 * scatter, combineAllDevice = CUDPP segmented MIN scan, UpdateVertex) and its loop
 * SSSP/main.cu:369-382.  uint32 weights and distances; MEDUSA_DIST_UNREACHED unreached.
 * Requires a weighted graph.  Distances are exact shortest-path lengths (sum < 2^32). */
int medusa_sssp(medusa_graph* g, uint32_t source, uint32_t* dist, const medusa_opts* opts,
                medusa_stats* stats);

/* PageRank: replaces MPG/UserOperators.h:175-270 (SendMsg scatter, CUDPP segmented ADD scan,
 * UpdateVertex) and the fixed-step loop MPG/main.cu:144-169.  Formula (FP32, Kahan-
 * compensated combine in a fixed order):
 *   r_0 = 1/n;  r_{t+1}(v) = (1-d)/n + d*( sum_{u->v} r_t(u)/outdeg(u) + D_t/n ),
 *   D_t = sum of r_t over vertices with outdeg 0.   rank[n] float.
 * On a packed batch (medusa_graph_set_batch) n and D_t are those of each vertex's graph. */
int medusa_pagerank(medusa_graph* g, float damping, uint32_t steps, float* rank,
                    const medusa_opts* opts, medusa_stats* stats);

/* Multi-source BFS: k <= 64 sources as one 64-bit frontier mask per vertex.
 * levels[k*n] uint8, row s = levels of source s; 0xFF unreached (levels >= 255 saturate
 * to 0xFF and set stats->reserved[0]).  No reference counterpart ("MBFS" in the
 * reference is multi-GPU BFS, HowToBuild:13); equals k independent medusa_bfs runs. */
int medusa_mbfs(medusa_graph* g, const uint32_t* sources, uint32_t k, uint8_t* levels,
                const medusa_opts* opts, medusa_stats* stats);

/* ---- autotuner ------------------------------------------------------------
 * Chooses, per graph, the launch shape of the frontier engines: the out-degree above
 * which a frontier vertex is split over whole workgroups (opts.large_degree) and the
 * workgroups of the expansion launch (opts.grid).  Every candidate is run `reps` times
 * from `sources` (BFS: any count, as medusa_bfs_multi; MBFS: at most 64) and timed with
 * hipEvents; the fastest is kept in the handle and used by every later run whose opts
 * leave large_degree / grid at 0.  reps = 0 forgets the tuned shape.  Results never
 * depend on the shape, only run time.  Replaces the reference's fixed launch shape
 * (MultipleGPU/MultiAPIHoster.h:649-754) and hybrid threshold (MultiPublicAPI.h:82). */
#define MEDUSA_TUNE_BFS  1
#define MEDUSA_TUNE_MBFS 2
typedef struct medusa_tuning {
    uint32_t large_degree;  /* chosen hub threshold */
    uint32_t grid;          /* chosen expansion workgroups */
    uint32_t candidates;    /* shapes timed */
    uint32_t reserved0;
    double   best_ms;       /* best run time of the chosen shape */
    double   default_ms;    /* best run time of the engine's default shape */
    uint32_t reserved[4];
} medusa_tuning;
int medusa_autotune(medusa_graph* g, uint32_t algo, const uint32_t* sources, uint32_t count, uint32_t reps,
                    medusa_tuning* result);

/* ---- per-superstep trace (feeds the learned scheduler) --------------------
 * Reference analogue: frontier counters of Algorithm/CPUFunctorHoster.cpp:13-15,21-30
 * and per-operator timers SSSP/UserOperators.h:198-262.  Each record has
 * MEDUSA_TRACE_FIELDS doubles: step, frontier vertices, frontier out-edge sum,
 * unvisited vertices, unvisited out-edge sum, n, m, action, kernel ms (BFS/MBFS);
 * phase, far entries, mean weight, threshold, vertices settled in the phase (final
 * distance from this phase's threshold up to the next one's), n, m, width, phase ms (SSSP; without
 * timing the device-driven loop marks each phase on the GPU clock, with timing the
 * host-driven loop times it on the host);
 * step, dangling vertices, vertices of in-degree > 64, vertices of in-degree 1..64, 0, n, m,
 * variant, superstep ms (PageRank). */
#define MEDUSA_TRACE_FIELDS 9
int medusa_trace_count(const medusa_graph* g, uint32_t* count);
int medusa_trace_read(const medusa_graph* g, double* out, uint32_t capacity);

/* ---- learned scheduler ------------------------------------------------------
 * A small MLP cost model (16 inputs -> 32 -> 32 -> 1, leaky ReLU) trained on the GPU by the
 * library's own forward/backward kernels.  Input row layout (MEDUSA_POLICY_INPUTS floats) is
 * produced by medusa_policy_features(); the engine evaluates every candidate action at each
 * superstep (the SSSP width: once per run, at its start) and takes the lowest predicted cost.
 * Results never depend on the choice.
 * No reference counterpart. */
#define MEDUSA_POLICY_INPUTS 16
#define MEDUSA_POLICY_PARAMS 1633
#define MEDUSA_POLICY_DIRECTION 1 /* BFS / MBFS: push or pull per superstep */
#define MEDUSA_POLICY_DELTA 2     /* SSSP: bucket width of a run, chosen at its start */
#define MEDUSA_POLICY_VARIANT 3   /* PageRank: pull or push kernel variant per superstep */
int medusa_policy_create(uint32_t kind, uint64_t seed, medusa_policy** out);
int medusa_policy_destroy(medusa_policy* p);
/* Reads (get != NULL) and/or replaces (set != NULL) the MEDUSA_POLICY_PARAMS parameters. */
int medusa_policy_params(medusa_policy* p, float* get, const float* set, uint32_t count);
/* Full-batch Adam on mean squared error over n rows X[n][16] -> y[n]. */
int medusa_policy_train(medusa_policy* p, const float* X, const float* y, uint32_t n, uint32_t iters, float lr,
                        float* final_loss);
/* GPU forward pass over n rows. */
int medusa_policy_predict(medusa_policy* p, const float* X, uint32_t n, float* out);
/* GPU gradient of the mean squared error at the current parameters (and the loss). */
int medusa_policy_gradients(medusa_policy* p, const float* X, const float* y, uint32_t n, float* grad, float* loss);
/* Feature row for one trace record (MEDUSA_TRACE_FIELDS doubles) and a candidate action:
 * DIRECTION kind: action 0 = push, 1 = pull; DELTA kind: action = bucket width. */
int medusa_policy_features(uint32_t kind, const double* trace_record, double action, float* row);

/* ---- synthetic graph generators (host) -----------------------------------
 * Deterministic from the seed (xoshiro256** streams seeded by splitmix64), independent of
 * thread count.  Arcs are written as (src[i], dst[i], weight[i]); weight may be NULL.
 * capacity = length of the caller's arrays; *m_out = arcs written.  The reference's
 * generators (MedusaRT/GraphGenerator.cu:21-209) are not reproduced. */
/* Erdos-Renyi G(n,m): m draws of (u,v) uniform, self-loops dropped, duplicates kept. */
int medusa_gen_erdos_renyi(uint32_t n, uint64_t m, uint64_t seed, uint64_t capacity, uint32_t* src,
                           uint32_t* dst, uint64_t* m_out);
/* R-MAT: 2^scale vertices, m arcs, quadrant probabilities a,b,c (d = 1-a-b-c) each scaled per
 * level by (1 + noise*U[-1,1]) and renormalised; self-loops and duplicates kept. */
int medusa_gen_rmat(uint32_t scale, uint64_t m, double a, double b, double c, double noise,
                    uint64_t seed, uint64_t capacity, uint32_t* src, uint32_t* dst, uint64_t* m_out);
/* rows x cols lattice, id r*cols+c; right/down edges kept with p_keep; each cell with r<rows-1
 * gets with p_diag one diagonal to (r+1,c+-1) (sign uniform; none if off the grid); both
 * directions; uint32 weight uniform in [wmin,wmax] per undirected edge. */
int medusa_gen_lattice(uint32_t rows, uint32_t cols, double p_keep, double p_diag, uint32_t wmin,
                       uint32_t wmax, uint64_t seed, uint64_t capacity, uint32_t* src, uint32_t* dst,
                       uint32_t* weight, uint64_t* m_out);
/* Chung-Lu: out-weight 0 with probability p_zero, else Zipf(exponent) on [1, max_degree];
 * each of m arcs picks src with probability proportional to weight, dst uniform. */
int medusa_gen_chung_lu(uint32_t n, uint64_t m, double exponent, uint32_t max_degree, double p_zero,
                        uint64_t seed, uint64_t capacity, uint32_t* src, uint32_t* dst, uint64_t* m_out);
/* Uniform integer weights in [wmin,wmax] for m arcs (stream `seed`). */
int medusa_gen_weights(uint64_t m, uint32_t wmin, uint32_t wmax, uint64_t seed, uint32_t* weight);

/* Stable counting sort of an arc list into CSR (arcs of a source keep input order).
 * Replaces GraphIR::AddEdge + buildAA (MedusaRT/GraphConverter.cu:87-103,
 * SSSP/EdgeDataType.cu:138-186), except that the reference prepends (reverse input order). */
int medusa_csr_from_arcs(uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                         const uint32_t* weight, uint32_t* row_ptr, uint32_t* col, uint32_t* weight_out);

#ifdef __cplusplus
}
#endif

#endif /* MEDUSA_C_H */
