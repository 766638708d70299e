// Per-iteration schedule of the COORDINATION engine.
//
// A schedule choice (0..15) fixes, for one iteration:
//   bit 0     direction: 0 push, 1 pull (bottom-up; BFS only, push otherwise)
//   bits 1-2  degree-tier thresholds (kTierSets below)
//   bit 3     side: 0 the GPU expands the frontier, 1 the host team does
// The fixed heuristic is the reference's LIGHT/HEAVY rule
// (my_project/CGgraph/CGgraph_engine.cuh:312-348): LIGHT iterations run on the
// host team, HEAVY ones on the GPU with the host team taking a tail share sized
// by measured throughput (the analytical split of the reference's COOPERATION
// engine, GraphCooperation.cuh:1036-1086).
#pragma once

#include <cmath>
#include <cstdint>

#include "kernels.hpp"

namespace cg {

struct TierSet {
  uint32_t bin_wave, bin_chunk;
};
// bin_wave: degree from which a vertex gets a whole wavefront;
// bin_chunk: degree from which a vertex is cut into chunk_edges pieces.
constexpr TierSet kTierSets[4] = {{32, 4096}, {16, 1024}, {64, 16384}, {8, 256}};
// The network's choice is taken when its top probability reaches this.
constexpr float kLearnedMinProb = 0.35f;
// SSSP hub deferral: window above the frontier's smallest distance
constexpr uint32_t kHubDeltaDefault = 5;

struct Choice {
  int pull;
  int tier;
  int host;
};

inline Choice decode_choice(int c) { return Choice{c & 1, (c >> 1) & 3, (c >> 3) & 1}; }
inline int encode_choice(const Choice& ch) { return (ch.pull & 1) | ((ch.tier & 3) << 1) | ((ch.host & 1) << 3); }

// The 12 features of one iteration, all scaled to roughly [0, 1]:
//   0 log2(1+frontier vertices)/log2(1+V)   1 log2(1+frontier edges)/log2(1+E)
//   2 unvisited fraction                    3..10 log-degree histogram / frontier
//   11 min(iteration, 64)/64
// The same function runs on the host (traces) and in the stats kernel's
// epilogue (inference).
__host__ __device__ inline void make_features(const DevStats& s, uint32_t V, uint64_t E,
                                              uint64_t visited, uint32_t iter, float* f) {
  const double lv = log2(1.0 + (double)V), le = log2(1.0 + (double)E);
  f[0] = (float)(log2(1.0 + (double)s.n_active) / (lv > 0 ? lv : 1.0));
  f[1] = (float)(log2(1.0 + (double)s.m_active) / (le > 0 ? le : 1.0));
  f[2] = V ? (float)(1.0 - (double)visited / (double)V) : 0.f;
  const double na = s.n_active ? (double)s.n_active : 1.0;
  for (int k = 0; k < 8; ++k) f[3 + k] = (float)((double)s.hist[k] / na);
  f[11] = (float)((double)(iter < 64 ? iter : 64) / 64.0);
}

}  // namespace cg
