/* ThreadSanitizer driver for the restated MULTI_CORE engine (oracle/multicore.c).
 *
 * Test infrastructure, like the rest of oracle/: tests/test_oracle_tsan.py
 * builds it with clang's ThreadSanitizer and libomp and runs it under libomp's
 * Archer tool (which tells TSan about OpenMP's barriers and joins).  CPU code
 * only; no GPU run loads it.
 * The reference relies on benign races -- plain reads of vertexValue beside
 * CAS/atomicMin (P/host_algorithm.hpp:176-181), a volatile CAS loop in write_min
 * (:23-39), busy-wait flags (P/CGgraph_engine.cuh:866-874) -- and the
 * restatement expresses each with __atomic operations instead, so a TSan run
 * over BFS, SSSP and CC on a skewed graph (one hub whose edges are donated in
 * chunks, the third stage of the stealer) must report nothing.
 *
 *   tsan_check            run the three algorithms, compare each with the
 *                         sequential oracle, print "mismatches N" per algorithm
 *   tsan_check plant      additionally run a deliberate plain-int race, so the
 *                         test can show that the sanitizer, set up this way,
 *                         does report races
 */
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static int planted_race(void) {
  int shared = 0;
#pragma omp parallel num_threads(2)
  {
    for (int i = 0; i < 1000; ++i) shared += omp_get_thread_num() + 1;  // unsynchronised on purpose
  }
  return shared;
}

int main(int argc, char** argv) {
  const int plant = argc > 1 && strcmp(argv[1], "plant") == 0;
  const uint32_t scale = 14, V = 1u << scale;
  const uint64_t m = 200000, hub = 40000, cap = m + hub;
  uint32_t* src = malloc(cap * sizeof(uint32_t));
  uint32_t* dst = malloc(cap * sizeof(uint32_t));
  uint64_t* off = malloc(((uint64_t)V + 1) * sizeof(uint64_t));
  uint32_t* adj = malloc(cap * sizeof(uint32_t));
  uint32_t* w = malloc(cap * sizeof(uint32_t));
  uint32_t* got = malloc((uint64_t)V * sizeof(uint32_t));
  uint32_t* ref = malloc((uint64_t)V * sizeof(uint32_t));
  if (!src || !dst || !off || !adj || !w || !got || !ref) return 2;
  or_rmat_edges(scale, 0, m, 0.57, 0.19, 0.19, 5, src, dst);
  for (uint64_t i = 0; i < hub; ++i) {  // vertex 0: a hub far above the donation threshold
    src[m + i] = 0;
    dst[m + i] = (uint32_t)((i * 7919u) % V);
  }
  const int64_t E = or_csr_build(V, cap, src, dst, 0, off, adj, cap);
  if (E < 0) return 2;
  or_fill_weights(V, off, adj, 5, 1, 100, 0, w);
  int bad_total = 0;
  for (int alg = 0; alg < 3; ++alg) {
    uint32_t iters = 0;
    or_multicore_run(alg, V, off, adj, w, 0, 8, got, &iters);
    if (alg == 0) or_bfs(V, off, adj, 0, ref);
    else if (alg == 1) or_dijkstra(V, off, adj, w, 0, ref);
    else or_cc_directed(V, off, adj, ref);
    int bad = 0;
    for (uint32_t v = 0; v < V; ++v) bad += got[v] != ref[v];
    printf("alg %d E %lld iterations %u mismatches %d\n", alg, (long long)E, iters, bad);
    bad_total += bad;
  }
  if (plant) printf("planted %d\n", planted_race());
  free(src);
  free(dst);
  free(off);
  free(adj);
  free(w);
  free(got);
  free(ref);
  return bad_total ? 1 : 0;
}
