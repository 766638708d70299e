// The host team of the COORDINATION engine: OpenMP threads that relax the
// part of a frontier the scheduler gives to the CPU.
//
// The reference's CPU side relaxes edges directly on the host value array and
// the host later min-merges the GPU replica (CGgraph_engine.cuh:692-855,
// 1657-1727).  Here the authoritative values stay in HBM.  The host team
// receives (vertex, value) for its share of the frontier, relaxes the edges
// against a host "shadow" array -- an upper bound of every value, since values
// only decrease -- and returns (destination, candidate) pairs that a GPU kernel
// applies with atomic min.  Filtering with an upper bound never drops a
// candidate that could improve a value, so the fixed point reached is the same
// whichever side relaxes which vertex.
#pragma once

#include <cstdint>
#include <vector>

struct cg_graph;

namespace cg {

struct HostTeam {
  int threads = 1;
  uint32_t V = 0;
  std::vector<uint32_t> shadow;  // upper bounds of the device values
  // per-thread candidate buffers
  std::vector<std::vector<uint32_t>> cid, cval;

  void init(uint32_t V_, int threads_);
  // BFS/SSSP: all MAX, root 0; CC: identity.
  void reset_values(int alg, uint32_t root);
  // Relaxes the out-edges of `n` frontier vertices (ids[i] currently holds
  // vals[i]); edges of vertex i restricted to [lo[i], hi[i]) when lo/hi are
  // given (chunks of split vertices).  Returns the number of candidates,
  // gathered into out_id / out_val (capacity grows as needed).
  uint64_t relax(int alg, const cg_graph* g, const uint32_t* ids, const uint32_t* vals,
                 const uint64_t* lo, const uint64_t* hi, uint64_t n,
                 std::vector<uint32_t>& out_id, std::vector<uint32_t>& out_val,
                 uint64_t* edges_done);
};

}  // namespace cg
