// ThreadSanitizer driver for the engine's host team (cggraph_amd/csrc/host_team.cpp).
//
// Test infrastructure: tests/test_oracle_tsan.py builds it with clang's
// ThreadSanitizer and libomp (host code only; no HIP call is made: the graph is
// host-only, so no device buffer is ever allocated or freed) and runs it under
// libomp's Archer tool.  It drives the team in its three modes:
//   list    the COORDINATION engine's LIGHT iterations (P/CGgraph_engine.cuh:
//           1441-1469): (vertex, value) pairs in, candidates out, min-applied;
//           three-stage stealing with the hub donated in 6 400-edge pieces
//           (P/taskSteal.hpp:89-274);
//   bitmap  the engine without a GPU: the team relaxes a frontier bitmap on
//           its own value array;
//   tail    a co-executed iteration: an agent thread claims pieces from the
//           head of the frontier list (as the GPU agent does) while the team
//           claims from the tail through the same QueueEnds word
//           (P/CGgraph_engine.cuh:692-855, 1548-1624).
// Every mode's result must equal the sequential oracle's.
#include <cstdio>
#include <thread>
#include <vector>

#include "common.hpp"
#include "host_team.hpp"
#include "oracle.h"

namespace {

uint32_t cand_of(int alg, uint32_t sv, const cg_graph& g, uint64_t k) {
  return alg == CG_BFS ? sv + 1 : (alg == CG_SSSP ? sv + g.h_w[k] : sv);
}

}  // namespace

int main() {
  const uint32_t scale = 14, V = 1u << scale;
  // 16 000 distinct destinations for vertex 0: above the 12 800-edge threshold
  const uint64_t m = 160000, hub = 16000, cap = m + hub;
  std::vector<uint32_t> src(cap), dst(cap);
  or_rmat_edges(scale, 0, m, 0.57, 0.19, 0.19, 7, src.data(), dst.data());
  for (uint64_t i = 0; i < hub; ++i) {  // 7919 is odd: i * 7919 mod 2^14 are distinct
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
  int bad_total = 0;
  const char* modes[3] = {"list", "bitmap", "tail"};
  for (int alg = CG_BFS; alg <= CG_CC; ++alg) {
    const uint32_t root = 0;
    std::vector<uint32_t> ref(V);
    if (alg == CG_BFS) or_bfs(V, g.h_off.data(), g.h_dst.data(), root, ref.data());
    else if (alg == CG_SSSP) or_dijkstra(V, g.h_off.data(), g.h_dst.data(), g.h_w.data(), root, ref.data());
    else or_cc_directed(V, g.h_off.data(), g.h_dst.data(), ref.data());
    for (int mode = 0; mode < 3; ++mode) {
      team.reset_values(alg, root);
      std::vector<uint32_t> val(V, cg::kUnreached);
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
      uint32_t iters = 0;
      uint64_t donated = 0, stolen = 0, agent_entries = 0, team_entries = 0;
      if (mode == 1) {
        // the team's own values are the result
        const uint64_t words = (V + 63) / 64;
        std::vector<uint64_t> in(words + 1, 0), out(words + 1, 0);
        for (uint32_t v : frontier) in[v >> 6] |= 1ull << (v & 63);
        for (;;) {
          ++iters;
          std::fill(out.begin(), out.end(), 0ull);
          uint64_t edges = 0;
          const uint64_t upd = team.step_bitmap(alg, &g, in.data(), out.data(), words, &edges);
          donated += team.last_donated_vertices;
          stolen += team.last_stolen_units;
          if (!upd) break;
          in.swap(out);
        }
        val = team.shadow;
      } else {
        std::vector<uint32_t> ids, vals, out_id, out_val;
        while (!frontier.empty()) {
          ++iters;
          ids.assign(frontier.begin(), frontier.end());
          vals.clear();
          for (uint32_t v : frontier) vals.push_back(val[v]);
          uint64_t n = 0;
          std::vector<uint32_t> agent_id, agent_val;
          if (mode == 0) {
            uint64_t edges = 0;
            n = team.relax(alg, &g, ids.data(), vals.data(), ids.size(), out_id, out_val, &edges);
            donated += team.last_donated_vertices;
            stolen += team.last_stolen_units;
          } else {
            // the agent takes pieces from the head and relaxes them on its own
            // copy of the values; the team takes 64-entry pieces from the tail
            cg::QueueEnds q;
            const uint32_t total = (uint32_t)ids.size();
            q.reset(0, total, total / 3);  // the team holds the last two thirds only
            std::vector<uint32_t> agent_val_arr = val;
            std::thread agent([&] {
              uint32_t lo, hi, piece = 16;
              while (q.claim_head(piece, &lo, &hi)) {
                agent_entries += hi - lo;
                for (uint32_t i = lo; i < hi; ++i) {
                  const uint32_t v = ids[i], sv = vals[i];
                  for (uint64_t k = g.h_off[v]; k < g.h_off[v + 1]; ++k) {
                    const uint32_t d = g.h_dst[k], c = cand_of(alg, sv, g, k);
                    if (c < agent_val_arr[d]) {
                      agent_val_arr[d] = c;
                      agent_id.push_back(d);
                      agent_val.push_back(c);
                    }
                  }
                }
              }
            });
            uint64_t edges = 0, taken = 0;
            n = team.relax_tail(alg, &g, ids.data() + total / 3, vals.data() + total / 3, &q, out_id, out_val,
                                &edges, &taken);
            agent.join();
            team_entries += taken;
          }
          frontier.clear();
          auto apply = [&](uint32_t d, uint32_t c) {
            if (c < val[d]) {
              val[d] = c;
              if (!in_next[d]) {
                in_next[d] = 1;
                frontier.push_back(d);
              }
            }
          };
          for (uint64_t i = 0; i < n; ++i) apply(out_id[i], out_val[i]);
          for (size_t i = 0; i < agent_id.size(); ++i) apply(agent_id[i], agent_val[i]);
          for (uint32_t d : frontier) in_next[d] = 0;
        }
      }
      int bad = 0;
      for (uint32_t v = 0; v < V; ++v) bad += val[v] != ref[v];
      std::printf("alg %d mode %s E %lld iterations %u donated %llu stolen %llu agent %llu team %llu mismatches %d\n",
                  alg, modes[mode], (long long)E, iters, (unsigned long long)donated, (unsigned long long)stolen,
                  (unsigned long long)agent_entries, (unsigned long long)team_entries, bad);
      bad_total += bad;
    }
  }
  return bad_total ? 1 : 0;
}
