// Host-team relaxation (see host_team.hpp).  The per-edge operation is the
// reference's bfs_numa / sssp_numa / cc_numa (my_project/CGgraph/
// host_algorithm.hpp:164-184, 277-300, 412-433) with write_min as a CAS loop
// (host_algorithm.hpp:23-39), applied to the shadow array.  Work is split in
// 64-vertex chunks over the OpenMP team, as VERTEXSTEAL_CHUNK does
// (Basic/Graph/basic_def.hpp:14, my_project/CGgraph/taskSteal.hpp:120).
#include "host_team.hpp"

#include <omp.h>

#include <algorithm>
#include <cstring>

#include "common.hpp"

namespace cg {

namespace {

inline bool write_min(uint32_t* p, uint32_t v) {
  uint32_t cur = __atomic_load_n(p, __ATOMIC_RELAXED);
  while (cur > v) {
    if (__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED))
      return true;
  }
  return false;
}

}  // namespace

void HostTeam::init(uint32_t V_, int threads_) {
  V = V_;
  threads = threads_ > 0 ? threads_ : omp_get_max_threads();
  shadow.assign(V, kUnreached);
  cid.assign(threads, {});
  cval.assign(threads, {});
}

void HostTeam::reset_values(int alg, uint32_t root) {
  if (alg == CG_CC) {
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t v = 0; v < (int64_t)V; ++v) shadow[v] = (uint32_t)v;
  } else {
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t v = 0; v < (int64_t)V; ++v) shadow[v] = kUnreached;
    if (root < V) shadow[root] = 0;
  }
}

uint64_t HostTeam::relax(int alg, const cg_graph* g, const uint32_t* ids, const uint32_t* vals,
                         const uint64_t* lo, const uint64_t* hi, uint64_t n,
                         std::vector<uint32_t>& out_id, std::vector<uint32_t>& out_val,
                         uint64_t* edges_done) {
  const uint64_t* off = g->h_off.data();
  const uint32_t* dst = g->h_dst.data();
  const uint32_t* w = g->has_w ? g->h_w.data() : nullptr;
  uint32_t* sh = shadow.data();
  uint64_t edges = 0;
#pragma omp parallel num_threads(threads) reduction(+ : edges)
  {
    const int t = omp_get_thread_num();
    std::vector<uint32_t>& my_id = cid[t];
    std::vector<uint32_t>& my_val = cval[t];
    my_id.clear();
    my_val.clear();
#pragma omp for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
      const uint32_t v = ids[i];
      const uint32_t sv = vals[i];
      // the gathered value is exact at gather time: tighten the shadow
      write_min(&sh[v], sv);
      const uint64_t s = lo ? lo[i] : off[v];
      const uint64_t e = hi ? hi[i] : off[v + 1];
      edges += e - s;
      for (uint64_t k = s; k < e; ++k) {
        const uint32_t d = dst[k];
        const uint32_t cand = alg == CG_BFS ? sv + 1 : (alg == CG_SSSP ? sv + w[k] : sv);
        if (cand < __atomic_load_n(&sh[d], __ATOMIC_RELAXED) && write_min(&sh[d], cand)) {
          my_id.push_back(d);
          my_val.push_back(cand);
        }
      }
    }
  }
  uint64_t total = 0;
  std::vector<uint64_t> base(threads + 1, 0);
  for (int t = 0; t < threads; ++t) base[t + 1] = base[t] + cid[t].size();
  total = base[threads];
  if (out_id.size() < total) { out_id.resize(total); out_val.resize(total); }
#pragma omp parallel for num_threads(threads) schedule(static, 1)
  for (int t = 0; t < threads; ++t) {
    if (!cid[t].empty()) {
      std::memcpy(out_id.data() + base[t], cid[t].data(), cid[t].size() * sizeof(uint32_t));
      std::memcpy(out_val.data() + base[t], cval[t].data(), cval[t].size() * sizeof(uint32_t));
    }
  }
  if (edges_done) *edges_done = edges;
  return total;
}

}  // namespace cg
