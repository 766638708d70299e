// ThreadSanitizer driver for the engine's host team (cggraph_amd/csrc/host_team.cpp).
//
// Test infrastructure: tests/test_oracle_tsan.py builds it with clang's
// ThreadSanitizer and libomp (host code only; no HIP call is made: the graph is
// host-only, so no device buffer is ever allocated or freed) and runs it under
// libomp's Archer tool.  It drives the team the way the COORDINATION engine's
// LIGHT iterations do (P/CGgraph_engine.cuh:1441-1469): the frontier's
// (vertex, value) pairs go in, candidates come out and are min-applied, and
// vertices above the donation threshold are handed over in edge chunks
// (P/taskSteal.hpp:89-274).  The result must equal the sequential oracle's.
#include <cstdio>
#include <vector>

#include "common.hpp"
#include "host_team.hpp"
#include "oracle.h"

int main() {
  const uint32_t scale = 13, V = 1u << scale;
  const uint64_t m = 120000, hub = 30000, cap = m + hub;
  std::vector<uint32_t> src(cap), dst(cap);
  or_rmat_edges(scale, 0, m, 0.57, 0.19, 0.19, 7, src.data(), dst.data());
  for (uint64_t i = 0; i < hub; ++i) {  // vertex 0: a hub split into chunks
    src[m + i] = 0;
    dst[m + i] = (uint32_t)((i * 7919u) % V);
  }
  cg_graph g;
  g.V = V;
  g.h_off.resize((uint64_t)V + 1);
  g.h_dst.resize(cap);
  const int64_t E = or_csr_build(V, cap, src.data(), dst.data(), 0, g.h_off.data(), g.h_dst.data(), cap);
  if (E < 0) return 2;
  g.E = (uint64_t)E;
  g.h_dst.resize(g.E);
  g.h_w.resize(g.E);
  or_fill_weights(V, g.h_off.data(), g.h_dst.data(), 7, 1, 100, 0, g.h_w.data());
  g.has_w = true;
  g.host_valid = true;

  cg::HostTeam team;
  team.init(V, 8);
  const uint64_t kChunk = 6400;  // edges per donated chunk
  int bad_total = 0;
  for (int alg = CG_BFS; alg <= CG_CC; ++alg) {
    const uint32_t root = 0;
    team.reset_values(alg, root);
    std::vector<uint32_t> val(V, cg::kUnreached), ref(V);
    std::vector<uint8_t> in_next(V, 0);
    std::vector<uint32_t> frontier;
    if (alg == CG_CC) {
      for (uint32_t v = 0; v < V; ++v) {
        val[v] = v;
        frontier.push_back(v);
      }
    } else {
      val[root] = 0;
      frontier.push_back(root);
    }
    std::vector<uint32_t> ids, vals, out_id, out_val;
    std::vector<uint64_t> lo, hi;
    uint32_t iters = 0;
    while (!frontier.empty()) {
      ++iters;
      ids.clear();
      vals.clear();
      lo.clear();
      hi.clear();
      for (uint32_t v : frontier) {
        const uint64_t s = g.h_off[v], e = g.h_off[v + 1];
        for (uint64_t a = s; a < e || a == s; a += kChunk) {
          ids.push_back(v);
          vals.push_back(val[v]);
          lo.push_back(a);
          hi.push_back(e - a > kChunk ? a + kChunk : e);
          if (e - a <= kChunk) break;
        }
      }
      uint64_t edges = 0;
      const uint64_t n = team.relax(alg, &g, ids.data(), vals.data(), lo.data(), hi.data(), ids.size(), out_id,
                                    out_val, &edges);
      frontier.clear();
      for (uint64_t i = 0; i < n; ++i) {
        const uint32_t d = out_id[i], c = out_val[i];
        if (c < val[d]) {
          val[d] = c;
          if (!in_next[d]) {
            in_next[d] = 1;
            frontier.push_back(d);
          }
        }
      }
      for (uint32_t d : frontier) in_next[d] = 0;
    }
    if (alg == CG_BFS) or_bfs(V, g.h_off.data(), g.h_dst.data(), root, ref.data());
    else if (alg == CG_SSSP) or_dijkstra(V, g.h_off.data(), g.h_dst.data(), g.h_w.data(), root, ref.data());
    else or_cc_directed(V, g.h_off.data(), g.h_dst.data(), ref.data());
    int bad = 0;
    for (uint32_t v = 0; v < V; ++v) bad += val[v] != ref[v];
    std::printf("alg %d E %lld iterations %u mismatches %d\n", alg, (long long)E, iters, bad);
    bad_total += bad;
  }
  return bad_total ? 1 : 0;
}
