/*
 * CPU oracle for the frontier-traversal path (BFS / SSSP / CC / PR).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this code; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it, and only as the checker (or, for or_multicore_run, as the timed
 * CPU baseline that restates the reference's MULTI_CORE engine).
 *
 * Semantics follow the reference (paths relative to CGgraph/):
 *   - unreached sentinel 999 999 999          Basic/Graph/basic_def.hpp:18
 *   - u32 vertex ids / values, u32 weights    Basic/Type/data_type.hpp:6-15
 *   - BFS level = value[src] + 1, SSSP = value[src] + w, CC pushes value[src],
 *     each applied only when strictly smaller  my_project/CGgraph/host_algorithm.hpp:164-184, 277-300, 412-433
 *   - init: BFS/SSSP value = MAX, value[root] = 0; CC value[i] = i
 *                                               my_project/CGgraph/GraphNuma.cuh:521-563
 *   - PR: (1-d)/N + d * sum, d = 0.85, dangling mass dropped
 *                                               my_project/CGgraph/device_algorithm.cuh:543
 * The results of BFS/SSSP/CC are the unique least fixed points of those
 * relaxations, so textbook algorithms (queue BFS, Dijkstra, union-find /
 * reachability) reproduce them exactly.
 *
 * Pinning (what checks each function):
 *   or_bfs, or_dijkstra      toy_0/toy_1 golden vectors (computed independently
 *                            with scipy.sparse.csgraph), scipy on random graphs,
 *                            the shortest-path certificate, brute force.
 *   or_cc_undirected         scipy connected_components + per-component min.
 *   or_cc_directed           brute-force reachability on tiny graphs.
 *   or_pagerank64            closed forms (directed cycle => uniform 1/N; star),
 *                            dense-matrix power iteration in numpy.
 *   or_rmat_edges / weights  an independent numpy restatement of the same
 *                            counter-based generator (bit-exact edge lists).
 *   or_csr_build             numpy lexsort/unique on the same edges.
 *   or_check_sssp_cert       planted mistakes (one value off by one) must fail.
 *   or_multicore_run         compared with or_bfs / or_dijkstra / or_cc_directed.
 */
#ifndef CG_ORACLE_H
#define CG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_UNREACHED 999999999u

/* ---- deterministic generators (counter based; identical on every platform) ---- */
uint64_t or_mix64(uint64_t z);
/* Edges [first, first+count) of the R-MAT stream with the given scale. */
void or_rmat_edges(uint32_t scale, uint64_t first, uint64_t count, double a, double b,
                   double c, uint64_t seed, uint32_t* src, uint32_t* dst);
/* Weight of edge (u,v), uniform in [lo,hi]; symmetric != 0 hashes (min,max). */
uint32_t or_edge_weight(uint64_t seed, uint32_t u, uint32_t v, uint32_t lo, uint32_t hi,
                        int symmetric);
/* 4-neighbour n x n lattice, each undirected edge dropped with probability
 * `drop`; writes both directions.  Returns the number of directed edges
 * written (caller provides capacity 4*n*(n-1)). */
uint64_t or_grid_edges(uint32_t n, double drop, uint64_t seed, uint32_t* src, uint32_t* dst);

/* Edge list -> CSR: optional symmetrization, self-loops dropped, duplicates
 * dropped, neighbours strictly ascending.  off has V+1 entries; out_dst has
 * capacity `cap`.  Returns E, or -1 when cap is too small. */
int64_t or_csr_build(uint32_t V, uint64_t m, const uint32_t* src, const uint32_t* dst,
                     int symmetrize, uint64_t* off, uint32_t* out_dst, uint64_t cap);
void or_fill_weights(uint32_t V, const uint64_t* off, const uint32_t* dst, uint64_t seed,
                     uint32_t lo, uint32_t hi, int symmetric, uint32_t* w);
/* Transpose (CSR -> CSC); in-lists come out sorted by source id. */
void or_transpose(uint32_t V, const uint64_t* off, const uint32_t* dst, const uint32_t* w,
                  uint64_t* in_off, uint32_t* in_src, uint32_t* in_w);

/* ---- algorithms ---- */
void or_bfs(uint32_t V, const uint64_t* off, const uint32_t* dst, uint32_t root, uint32_t* out);
void or_dijkstra(uint32_t V, const uint64_t* off, const uint32_t* dst, const uint32_t* w,
                 uint32_t root, uint32_t* out);
/* Labels = minimum vertex id of each weakly connected component. */
void or_cc_undirected(uint32_t V, const uint64_t* off, const uint32_t* dst, uint32_t* out);
/* Labels = minimum id over all vertices that reach v (v included): the fixed
 * point of the reference's CC push on a directed graph. */
void or_cc_directed(uint32_t V, const uint64_t* off, const uint32_t* dst, uint32_t* out);
void or_pagerank64(uint32_t V, const uint64_t* off, const uint32_t* dst, int iters,
                   double damping, double* out);
/* One PageRank step in float64 from given (float32) ranks: the size-independent
 * check of a full-size run (PR_k must equal step(PR_{k-1}) up to FP32 rounding). */
void or_pagerank_step(uint32_t V, const uint64_t* off, const uint32_t* dst, const float* in,
                      double damping, double* out);

/* ---- checks ---- */
/* Shortest-path certificate (w == NULL means unit weights, i.e. BFS):
 * dist[root]==0; every edge u->v with dist[u] reached satisfies
 * dist[v] <= dist[u]+w or dist[u]+w >= UNREACHED; every reached v != root has
 * a tight in-edge.  Returns 0 when valid, else 1 + index of the first bad vertex
 * (or 1 + V for an edge violation). */
int64_t or_check_sssp_cert(uint32_t V, const uint64_t* off, const uint32_t* dst,
                           const uint32_t* w, uint32_t root, const uint32_t* dist);
uint64_t or_reached_edges(uint32_t V, const uint64_t* off, const uint32_t* val);
uint32_t or_max_outdeg_vertex(uint32_t V, const uint64_t* off);

/* ---- restated reference CPU engine (MULTI_CORE) ----
 * alg: 0 BFS, 1 SSSP, 2 CC.  Returns the wall time of the iterations in ms
 * (load and setup excluded), writes the values to out, iterations to *iters. */
int or_omp_max_threads(void);
double or_multicore_run(int alg, uint32_t V, const uint64_t* off, const uint32_t* dst,
                        const uint32_t* w, uint32_t root, int threads, uint32_t* out,
                        uint32_t* iters);

#ifdef __cplusplus
}
#endif
#endif
