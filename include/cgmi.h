/*
 * cgmi.h -- C ABI of the MI355X frontier-traversal engine (libcgmi.so).
 *
 * The reference (PengBo410/CGgraph) exposes no plugin or FFI surface; its
 * seams are C++ classes and templates.  Each entry point below replaces one of
 * them (paths relative to CGgraph/):
 *
 *   cg_graph_from_csr      CSR_Result_type           Basic/Graph/basic_struct.hpp:7-16
 *   cg_graph_load_files    load_binFile<T>           Basic/Other/IO.hpp:44-92
 *                          + main_partition CSR load my_project/CGgraph/main_partition.hpp:60-76
 *   cg_graph_generate_*    (synthetic workloads; no reference counterpart)
 *   cg_partition           TwoLevelPartition         my_project/CGgraph/Two-level_partition.cuh:106-748
 *   cg_engine_create       CGgraphEngine ctor        my_project/CGgraph/CGgraph_engine.cuh:154-253
 *                          (engine_type and use_device_num as main_partition.hpp:20-31,
 *                          162-557: CG_SINGLE_GPU runs every iteration on the GPU;
 *                          use_device_num 0, CG_MULTI_CORE and CG_SINGLE_CORE run every
 *                          iteration on the host team, Graph_Numa GraphNuma.cuh:166-314)
 *   cg_engine_run          CGgraphEngine::CG_co_execution(root) -> ms, result in
 *                          vertexValue               CGgraph_engine.cuh:265-974, :69
 *   cg_push_step_device    BFS_SPACE::bfs_device / SSSP_SPACE::sssp_device
 *                                                    my_project/CGgraph/device_algorithm.cuh:61-72, 264-278
 *   cg_batch_run           (batched queries; no reference counterpart)
 *   cg_check_result_file   checkBinResult<T>          Basic/Graph/util.cuh:60-118
 *
 * Conventions (differences from the reference are deliberate):
 *   - every function returns int status, CG_OK == 0; cg_last_error() returns a
 *     per-thread message.  Nothing calls exit() (the reference's CUDA_CHECK and
 *     assert_msg do, Basic/GPU/cuda_check.cuh:48-56).
 *   - the caller owns every input and output buffer; the library copies the
 *     graph into HBM and never frees caller memory.
 *   - one engine is used by one thread at a time; distinct engines may run
 *     concurrently.
 *   - vertex ids, values and weights are uint32; edge offsets are uint64
 *     internally and may be supplied as uint32 (the reference's countl_type).
 *   - unreached BFS/SSSP vertices hold 999 999 999 (VertexValue_MAX,
 *     Basic/Graph/basic_def.hpp:18).  Weights must stay below 2^32 - 10^9 so
 *     that value + weight never wraps.
 */
#ifndef CGMI_H
#define CGMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CG_UNREACHED 999999999u

/* ---- status codes ---- */
enum {
  CG_OK = 0,
  CG_ERR_INVALID = 1,      /* bad argument */
  CG_ERR_NO_DEVICE = 2,    /* a GPU engine was requested but no GPU is usable (never a silent
                              switch to the host team: ask for use_device_num 0 for that) */
  CG_ERR_HIP = 3,          /* a HIP runtime call failed */
  CG_ERR_OOM = 4,          /* host or device allocation failed */
  CG_ERR_IO = 5,           /* file missing or of the wrong size */
  CG_ERR_UNSUPPORTED = 6,  /* algorithm / engine combination not provided */
  CG_ERR_INTERNAL = 7,     /* invariant broken; the run's output is not written */
  CG_ERR_MISMATCH = 8      /* cg_check_result_file found a differing value */
};

/* Algorithm_type (Basic/Graph/basic_struct.hpp:144-151); CC is WCC. */
enum { CG_BFS = 0, CG_SSSP = 1, CG_CC = 2, CG_PR = 3 };

/* Engine_type (Basic/Graph/basic_struct.hpp:165-178), the values this library runs.
 * COORDINATION with use_device_num 1: GPU and host team per the scheduler; with
 * use_device_num 0: every iteration on the host team (the reference's LIGHT path,
 * CGgraph_engine.cuh:907-912).  SINGLE_CORE / MULTI_CORE: the host team with one /
 * cg_options.cpu_threads threads, whatever use_device_num says.  SINGLE_GPU: every
 * iteration on the GPU.  PageRank needs a GPU.  COOPERATION and the tool engines
 * (BLOCK_2D, COMPUTE_INDEGREE, REORDER, ACTIVEBLOCK) return CG_ERR_UNSUPPORTED. */
enum { CG_COORDINATION = 0, CG_SINGLE_CORE = 2, CG_MULTI_CORE = 3, CG_SINGLE_GPU = 4 };

/* Per-iteration scheduler of the COORDINATION engine. */
enum {
  CG_SCHED_HEURISTIC = 0, /* the reference's LIGHT/HEAVY rule (CGgraph_engine.cuh:312-348) */
  CG_SCHED_GPU = 1,       /* every iteration on the GPU, push with direction switching for BFS */
  CG_SCHED_LEARNED = 2,   /* MLP over frontier statistics, heuristic as fallback */
  CG_SCHED_FORCED = 3,    /* cg_options.forced_choice every iteration (tests) */
  CG_SCHED_PROBE = 4      /* trace collection: every iteration times all 16 choices from the
                             same state (restored in between), writes them to trace_path,
                             then continues with the CG_SCHED_GPU rule */
};

#define CG_NUM_CHOICES 16
#define CG_NUM_FEATURES 12

typedef struct cg_graph cg_graph;
typedef struct cg_engine cg_engine;

typedef struct cg_options {
  uint32_t segment_size;          /* SEGMENT_SIZE 1024 (Two-level_partition.cuh:7-8) */
  uint32_t heavy_active_vertices; /* HEAVY_ACTIVE_VERTICES 16384 (CGgraph_engine.cuh:21) */
  uint64_t heavy_active_edges;    /* HEAVY_ACTIVE_EDGES 25 000 000 (CGgraph_engine.cuh:22) */
  uint32_t bin_wave;              /* degree >= bin_wave: one wavefront per vertex */
  uint32_t bin_chunk;             /* degree >= bin_chunk: split into edge chunks */
  uint32_t chunk_edges;           /* edges per chunk of a split vertex */
  int sched_mode;                 /* CG_SCHED_* */
  int forced_choice;              /* 0..CG_NUM_CHOICES-1 for CG_SCHED_FORCED */
  int cpu_threads;                /* host team size; 0 = all cores */
  int pr_iterations;              /* 20 */
  float pr_damping;               /* 0.85 */
  int profile;                    /* 1: time every expansion kernel with HIP events */
  const char* trace_path;         /* per-iteration JSONL trace, NULL = off */
  const char* mlp_path;           /* learned-scheduler weights, NULL = built-in untrained */
  uint32_t fault_flags;           /* CG_FAULT_* (tests only) */
  uint32_t batch_block;           /* cg_batch_run: threads per workgroup (256 / 512 / 1024) */
  uint32_t batch_grid;            /* cg_batch_run: workgroups of the expansion kernel */
  uint32_t batch_levels;          /* cg_batch_run: levels captured per HIP-graph replay */
  uint32_t sssp_delta;            /* SSSP bucket width (delta-stepping): frontier vertices whose
                                     distance is >= the current threshold wait; 0 = relax the
                                     whole frontier every iteration */
  uint32_t pr_segment_log2;       /* PageRank source-window size 2^k vertices (cache blocking,
                                     pr_layout 3); see pr_layout for 0 and 31 */
  uint32_t atomic_return;         /* 1: activate a destination only when its atomic min returned a
                                     larger value (exact first-visit counts); 0: fire-and-forget
                                     atomics, activate on the strict-less pre-check (a superset,
                                     same fixed point) */
  uint32_t pr_layout;             /* PageRank in-edge layout: 0 automatic, 1 plain pull over the
                                     transposed graph, 2 pull over a copy renumbered by descending
                                     out-degree (gather locality on skewed graphs), 3 source
                                     windows of 2^pr_segment_log2 vertices, 4 propagation blocking
                                     (out-edges binned by 8192-vertex destination blocks, sums in
                                     64-bit fixed point; at most 2^25 vertices), 5 propagation
                                     blocking over (source, block) pairs of a copy renumbered by
                                     descending out-degree (one contribution per pair, 2 bytes
                                     per edge; at most 2^25 vertices, fewer than 2^32 edges).
                                     Automatic: a window size in [1, 30] selects 3, 31 selects 1;
                                     0 selects 5 from 2^24 to 2^25 vertices and 4 from 2^22 (with
                                     pr_kernel 0), else 2 above 65 536 vertices */
  uint32_t relabel;               /* BFS / SSSP: run on an engine-owned copy renumbered by
                                     descending out-degree (results come back in the caller's
                                     ids): 0 automatic (at least 2^21 vertices for SSSP, 2^23
                                     for BFS, and the top 1/64 of vertices send at least half of
                                     the edges), 1 never, 2 always */
  uint32_t pr_kernel;             /* PageRank sweep over layouts 1 and 2: 0 automatic (row tiers,
                                     the fastest measured), 1 row tiers (a lane /
                                     wavefront / workgroup per row by in-degree), 2 edge tiles
                                     (2048 consecutive in-edges per workgroup, rows combined by a
                                     segmented scan), 3 two source bands (renumbered layout: the
                                     sources below 2^pr_segment_log2, default 2^19, first, whose
                                     contributions stay in L2, then the rest) */
  uint32_t min_pull;              /* SSSP / label-propagation CC under the GPU and learned
                                     schedules: an iteration whose frontier edges reach
                                     min_pull / 1000 of E runs bottom-up (every vertex takes the
                                     minimum over its in-edges from the frontier; no per-edge
                                     atomics); 0 = always push */
  uint32_t expand_grid;           /* push expansion: most workgroups per launch, whose wavefronts
                                     stride over the work items; 0 = 2048 (256 CUs x 32
                                     wavefronts) */
  uint32_t batch_group;           /* cg_batch_run: frontier items per wavefront (1-64, 0 = 32);
                                     their edges are spread over the 64 lanes */
  uint32_t hub_degree;            /* SSSP under the GPU and learned schedules: a frontier vertex
                                     of out-degree >= hub_degree waits (stays in the frontier)
                                     while its distance is >= the frontier's smallest distance +
                                     hub_delta, so a hub is expanded once its distance has
                                     (nearly) settled instead of at every improvement;
                                     lower-degree vertices never wait.  0 = off; not combined
                                     with sssp_delta */
  uint32_t hub_delta;             /* window of hub_degree above the frontier's smallest distance;
                                     0 = 5 */
  uint32_t probe_stride;          /* CG_SCHED_PROBE: time every choice only at iterations k with
                                     k mod probe_stride in {0, 1} (pairs of consecutive
                                     iterations, what the lagged scheduler trains on); the others
                                     follow the GPU rule.  0 or 1 = every iteration */
  uint64_t hbm_budget;            /* bytes of HBM the partition may charge (TwoLevelPartition's
                                     budget, Two-level_partition.cuh:392-445); 0 = the device's
                                     memory.  A BFS / SSSP / CC graph over it runs partly
                                     resident: the 1024-vertex segments in ascending average-degree
                                     order keep their edges in HBM while the budget lasts, and the
                                     host team relaxes the other segments' frontier vertices every
                                     iteration (GPU push for the resident part, no bottom-up steps,
                                     no renumbering).  PageRank needs the whole graph in HBM */
  uint32_t batch_engine_v;        /* cg_batch_run (batched mode): the queries on a graph of at least
                                     this many vertices whose largest out-degree is at least 16
                                     times the average (a small-world graph: few, wide levels)
                                     run one by one on an engine of their own graph (GPU
                                     schedule; SSSP defers hubs of out-degree >= 128 within
                                     max weight / 50 of the frontier's smallest), on a host thread
                                     of their own while the other queries advance level by
                                     level; 0 = 2^18, 0xFFFFFFFF = every query batched.  The
                                     engines belong to the graph, so batches running at once
                                     from several threads take turns for their engine queries */
} cg_options;

/* Fault injection, forced from the host so every kernel still exits normally. */
enum {
  CG_FAULT_HIP_ERROR = 1u,     /* fail the run as if a HIP call had failed mid-run */
  CG_FAULT_MLP_NAN = 2u,       /* feed NaN weights to the learned scheduler */
  CG_FAULT_SLOW_GPU = 4u       /* report the GPU share as slow to the split model */
};

typedef struct cg_run_stats {
  uint32_t iterations;
  uint32_t heavy_iterations;      /* iterations with GPU work */
  uint32_t light_iterations;      /* iterations run by the host team only */
  uint32_t pull_iterations;
  double ms;                      /* wall time of the run, as CG_co_execution returns */
  uint64_t edges_traversed;       /* TEPS numerator: sum of out-degrees of reached vertices */
  uint64_t edges_relaxed;         /* edges actually processed over all iterations */
  uint64_t bytes_h2d, bytes_d2h;  /* PCIe traffic inside the run */
  uint64_t cpu_vertices, gpu_vertices; /* frontier vertices handled by each side */
  uint32_t expand_launches;       /* expansion kernels launched */
  double expand_kernel_ms;        /* summed HIP-event time of those launches (profile=1) */
  double expand_alg_bytes;        /* algorithmic bytes those launches moved */
  uint32_t learned_fallbacks;     /* iterations that took the GPU rule under the learned scheduler: the
                                     first, those after a frontier of fewer than E / 256 edges on
                                     the queued-ahead SSSP / CC pipeline (the network does not run
                                     there), and those whose network output was not confident */
  uint64_t resident_edges;        /* edges held in HBM (E unless the graph ran partly resident) */
  uint32_t pr_layout;             /* PageRank: the layout the engine chose (pr_layout values) */
  double pr_stream_bytes;         /* PageRank, layouts 4 and 5: bytes streamed per iteration by
                                     design (0 for the pull layouts, whose reads are gathers) */
} cg_run_stats;

typedef struct cg_partition_info {
  uint64_t segment_num;
  uint64_t bucket_start[5];       /* bucket boundaries in degree-sorted segment order */
  int can_hold_all;               /* 1: every segment is resident in HBM */
  uint64_t device_bytes;          /* HBM charged by the partition */
} cg_partition_info;

/* ---- options ---- */
int cg_options_default(cg_options* opt);

/* ---- graphs ---- */
/* offsets: V+1 entries of uint64 (offsets_u64 != 0) or uint32; dst: E entries;
 * w: E entries or NULL.  zero_outdeg_tail: vertices [V - tail, V) are declared
 * to have out-degree 0 (RCM_LAST_ZERO_OUTDEGREE ordering); 0 under NATIVE. */
int cg_graph_from_csr(uint32_t V, uint64_t E, const void* offsets, int offsets_u64,
                      const uint32_t* dst, const uint32_t* w, uint32_t zero_outdeg_tail,
                      cg_graph** g);
/* Reference raw little-endian array files (csrOffsetFile / csrDestFile /
 * csrWeightFile); w_file may be NULL.  Sizes are checked as load_binFile does. */
int cg_graph_load_files(const char* offset_file, int offsets_u64, const char* dst_file,
                        const char* w_file, uint32_t V, uint64_t E, cg_graph** g);
/* Edge list -> CSR on the GPU: self-loops dropped, parallel edges merged
 * keeping the smallest weight, rows ascending; symmetrize != 0 adds every
 * reverse edge.  w may be NULL.  (GraphBinReader::group_load_directedGraph +
 * sort_nbr + build_csr, Basic/Graph/graphBinReader.hpp:89-326.) */
int cg_graph_from_edges(uint32_t V, uint64_t m, const uint32_t* src, const uint32_t* dst, const uint32_t* w,
                        int symmetrize, cg_graph** g);
/* The reference's binary edge file: packed little-endian records
 * {u32 src, u32 dst, u32 weight} (12 bytes, graphBinReader.hpp:35-39) when
 * has_weights != 0, {u32 src, u32 dst} otherwise. */
int cg_graph_load_edge_file(const char* path, uint32_t V, int has_weights, int symmetrize, cg_graph** g);
/* R-MAT on the GPU: 2^scale vertices, edge_factor * 2^scale generated edges,
 * self-loops and duplicates removed, neighbours ascending; weights uniform in
 * [wlo, whi] (whi == 0: no weights); symmetrize != 0 adds every reverse edge. */
int cg_graph_generate_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                           uint64_t seed, uint32_t wlo, uint32_t whi, int symmetrize,
                           cg_graph** g);
/* n x n 4-neighbour lattice, each undirected edge dropped with probability drop. */
int cg_graph_generate_grid(uint32_t n, double drop, uint64_t seed, uint32_t wlo, uint32_t whi,
                           cg_graph** g);
int cg_graph_info(const cg_graph* g, uint32_t* V, uint64_t* E, int* has_weights,
                  int* symmetric, uint32_t* max_outdeg_vertex);
/* Copies the CSR out (any pointer may be NULL). */
int cg_graph_export(const cg_graph* g, uint64_t* offsets, uint32_t* dst, uint32_t* w);
/* Builds the transposed graph in HBM (needed by pull steps and PageRank);
 * engines call it on demand. */
int cg_graph_build_transpose(cg_graph* g);
void cg_graph_destroy(cg_graph* g);

/* Segment / bucket partition of the reference (1024-vertex segments, 4
 * degree buckets cut at 17 / 33 / 257), charged against this GPU's HBM. */
int cg_partition(const cg_graph* g, int algorithm, cg_partition_info* info);

/* ---- engines ---- */
int cg_engine_create(const cg_graph* g, int algorithm, int engine_type, int use_device_num,
                     const cg_options* opt, cg_engine** e);
/* Runs one query.  out: uint32[V] (BFS/SSSP/CC) or float[V] (PR) in host
 * memory, or NULL to leave the result in HBM (read it with cg_engine_result).
 * stats may be NULL.  With out == NULL the call may return while iterations
 * that find nothing to do are still queued on the engine's stream (the
 * values are final); the next run and both result calls wait for them. */
int cg_engine_run(cg_engine* e, uint32_t root, void* out, cg_run_stats* stats);
int cg_engine_result(cg_engine* e, void* out);
/* Device address of the engine's value array (uint32[V] or float[V]), once
 * every kernel of the last run has finished. */
int cg_engine_result_device(cg_engine* e, void** dev_ptr);
void cg_engine_destroy(cg_engine* e);

/* One push relaxation step over device buffers: every vertex whose bit is set
 * in front_in relaxes its out-edges into val with atomic min; improved
 * destinations get their bit set in front_out (which the caller zeroes).
 * stream is a hipStream_t (NULL = the null stream).  The graph must be on the
 * device (cg_graph_device_csr). */
int cg_push_step_device(int algorithm, const cg_graph* g, uint32_t* val,
                        const uint64_t* front_in, uint64_t* front_out, void* stream);
int cg_graph_device_csr(const cg_graph* g, const uint64_t** offsets, const uint32_t** dst,
                        const uint32_t** w);

/* ---- batched queries ---- */
/* nq queries (algorithm[i] BFS or SSSP, root[i]) over graph[i]; outs[i] is a
 * host uint32[V_i] buffer or NULL.  mode 0 (batched): the queries on large
 * skewed graphs (opt->batch_engine_v) run one by one on an engine of their
 * graph from a host thread of their own, while all the others advance level
 * by level together, two launches per level, the per-level loop captured in
 * a HIP graph; stats: ms over both, iterations = levels, expand_launches =
 * graph replays, heavy_iterations = queries run on engines, edges_traversed =
 * the TEPS numerator over all queries.  mode 1: one query per engine launch
 * (the comparison path). */
int cg_batch_run(cg_graph* const* graphs, const int* algorithms, const uint32_t* roots, int nq,
                 uint32_t* const* outs, int mode, const cg_options* opt, cg_run_stats* stats);
/* The same batched run over engine handles: query i is engine i's algorithm
 * (BFS or SSSP) on engine i's graph from roots[i]; outs[i] is a host
 * uint32[V_i] buffer or NULL.  The batch takes engines[0]'s options.  Each
 * engine's own state is not touched, so engines may repeat in the list.
 * (Many CG_co_execution calls, CGgraph_engine.cuh:265, issued as one.) */
int cg_engine_run_batch(cg_engine* const* engines, const uint32_t* roots, int nq, void* const* outs,
                        cg_run_stats* stats);

/* ---- allocation-free run ---- */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc: the GPU schedule of cg_engine_run (statistics, degree-tier
 * compaction with the default tiers, push expansion) for BFS, SSSP or CC
 * (label propagation) over a CSR the caller holds in device memory (d_offsets
 * uint64[V+1], d_dst / d_w uint32[E]; d_w may be NULL except for SSSP) from
 * `root`, into d_val (device uint32[V]).  Frontiers, queues and counters live
 * in one static device array of CG_NOALLOC_SCRATCH_BYTES per GPU, held until
 * the process ends; a graph whose buffers do not fit (about 8.3 bytes per
 * vertex plus 3 bytes per 512 edges) returns CG_ERR_INVALID.  Nothing is
 * allocated: no hipMalloc, hipFree or pinned host memory.  Every call on one
 * GPU uses that GPU's array, so calls that overlap on one GPU race.  stream is
 * a hipStream_t (NULL = the null stream); the call returns when the run has
 * finished on it.  *iterations (may be NULL) receives the iteration count. */
#define CG_NOALLOC_SCRATCH_BYTES (64u << 20)
int cg_traverse_device_noalloc(int algorithm, const uint64_t* d_offsets, const uint32_t* d_dst,
                               const uint32_t* d_w, uint32_t V, uint64_t E, uint32_t root, uint32_t* d_val,
                               void* stream, uint32_t* iterations);

/* ---- learned scheduler ---- */
/* Trains the MLP (12 -> 64 -> 64 -> 16, ReLU, softmax) on the GPU from a
 * trace file written with cg_options.trace_path; writes weights to out_path. */
int cg_mlp_train(const char* trace_path, const char* out_path, int epochs, float lr,
                 uint64_t seed, float* final_loss);
/* Forward / backward of the MLP on the GPU for a batch (tests compare these with
 * the CPU oracle's gradients).  params: 6032 floats. */
int cg_mlp_forward(const float* params, const float* x, int n, float* probs);
int cg_mlp_loss_grad(const float* params, const float* x, const int* labels, int n,
                     float* loss, float* grad);

/* ---- result files ---- */
/* Compares values with a reference result file (<graph>_<ALG>_<root>.bin,
 * uint32 raw array), through old2new when it is not NULL. */
int cg_check_result_file(const char* path, const uint32_t* values, uint32_t V,
                         const uint32_t* old2new, uint64_t* first_bad);
int cg_save_bin_u32(const char* path, const uint32_t* values, uint64_t n);

const char* cg_last_error(void);
const char* cg_version(void);
/* sizeof of the ABI structs as this library was compiled (binding checks). */
int cg_abi_sizes(size_t* options, size_t* run_stats, size_t* partition_info);

#ifdef __cplusplus
}
#endif
#endif
