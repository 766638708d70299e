// The host team of the COORDINATION engine: OpenMP threads that relax the
// part of a frontier the scheduler gives to the CPU, and the whole frontier
// when the engine runs without a GPU.
//
// The reference's CPU side relaxes edges directly on the host value array and
// the host later min-merges the GPU replica (CGgraph_engine.cuh:692-855,
// 1657-1727).  Here the authoritative values stay in HBM while a GPU takes
// part.  The host team then receives (vertex, value) for its share of the
// frontier, relaxes the edges against a host "shadow" array -- an upper bound
// of every value, since values only decrease -- and returns (destination,
// candidate) pairs that a GPU kernel applies with atomic min.  Filtering with
// an upper bound never drops a candidate that could improve a value, so the
// fixed point reached is the same whichever side relaxes which vertex.  In a
// run without a GPU the shadow array is the value array itself.
//
// Work distribution follows TaskSteal::threeStage_taskSteal
// (my_project/CGgraph/taskSteal.hpp:89-274): each thread owns a contiguous
// range of 64-slot units (VERTEXSTEAL_CHUNK), takes its own units, then steals
// units from the other threads' ranges, and a vertex of at least 12 800 edges
// (EDGESTEAL_THRESHOLD) is not relaxed by one thread: its owner publishes it
// and every thread claims 6 400-edge pieces (EDGESTEAL_CHUNK) of it until none
// is left (Basic/Graph/basic_def.hpp:14-17).
#pragma once

#include <atomic>
#include <cstdint>
#include <memory>
#include <vector>

struct cg_graph;

namespace cg {

constexpr uint64_t kVertexStealChunk = 64;
constexpr uint64_t kEdgeStealThreshold = 12800;
constexpr uint64_t kEdgeStealChunk = 6400;

// Head / tail of a queue region that the GPU agent consumes from the head and
// the host team from the tail until they meet (CGgraph_engine.cuh:692-855,
// 1548-1624).  Both ends live in one 64-bit word (head low, tail high) and
// every claim is one compare-and-swap on it, so no two claims overlap and no
// entry is left unclaimed.
struct QueueEnds {
  std::atomic<uint64_t> ends{0};
  uint32_t host_floor = 0;  // the host holds entries [host_floor, initial tail) only

  void reset(uint32_t head, uint32_t tail, uint32_t floor) {
    host_floor = floor;
    ends.store((uint64_t)tail << 32 | head, std::memory_order_release);
  }
  uint32_t head() const { return (uint32_t)ends.load(std::memory_order_acquire); }
  // GPU agent: up to n entries from the head, [*lo, *hi); false when none is left.
  bool claim_head(uint32_t n, uint32_t* lo, uint32_t* hi) {
    uint64_t cur = ends.load(std::memory_order_acquire);
    for (;;) {
      const uint32_t h = (uint32_t)cur, t = (uint32_t)(cur >> 32);
      if (h >= t) return false;
      const uint32_t nh = t - h > n ? h + n : t;
      if (ends.compare_exchange_weak(cur, (uint64_t)t << 32 | nh, std::memory_order_acq_rel)) {
        *lo = h;
        *hi = nh;
        return true;
      }
    }
  }
  // Host thread: up to n entries from the tail, never below host_floor.
  bool claim_tail(uint32_t n, uint32_t* lo, uint32_t* hi) {
    uint64_t cur = ends.load(std::memory_order_acquire);
    for (;;) {
      const uint32_t h = (uint32_t)cur, t = (uint32_t)(cur >> 32);
      const uint32_t bottom = h > host_floor ? h : host_floor;
      if (t <= bottom) return false;
      const uint32_t nt = t - bottom > n ? t - n : bottom;
      if (ends.compare_exchange_weak(cur, (uint64_t)nt << 32 | h, std::memory_order_acq_rel)) {
        *lo = nt;
        *hi = t;
        return true;
      }
    }
  }
};

struct HostTeam {
  int threads = 1;
  uint32_t V = 0;
  std::vector<uint32_t> shadow;  // upper bounds of the device values (the values, without a GPU)
  // per-thread candidate buffers
  std::vector<std::vector<uint32_t>> cid, cval;

  // Three-stage scheduler state, one cache line per thread.
  enum : int { kVertexWorking = 0, kEdgeDonate = 1, kVertexStealing = 2 };
  static constexpr uint64_t kNoDonation = ~0ull;
  struct alignas(64) Slot {
    std::atomic<uint64_t> cur{0};   // next unit of this thread's range
    uint64_t end = 0;               // end of the range (units)
    std::atomic<int> status{kVertexWorking};
    std::atomic<uint64_t> donate{kNoDonation};  // (entry << 32) | next 6 400-edge piece
  };
  std::unique_ptr<Slot[]> slots;
  alignas(64) std::atomic<int> working{0};

  // Counters of the last call (tests and traces)
  uint64_t last_donated_vertices = 0;
  uint64_t last_stolen_units = 0;

  void init(uint32_t V_, int threads_);
  // BFS/SSSP: all MAX, root 0; CC: identity.
  void reset_values(int alg, uint32_t root);

  // List mode.  Entry i is vertex ids[i], whose value was vals[i] when the
  // list was gathered; its out-edges are relaxed against the shadow array and
  // every candidate that lowered a shadow value is returned in out_id /
  // out_val.  Three-stage stealing over 64-entry units.
  uint64_t relax(int alg, const cg_graph* g, const uint32_t* ids, const uint32_t* vals, uint64_t n,
                 std::vector<uint32_t>& out_id, std::vector<uint32_t>& out_val, uint64_t* edges_done);

  // Tail mode (iterations the GPU shares).  ids / vals hold the queue entries
  // [q->host_floor, tail at the start); the threads claim 64-entry pieces from
  // the tail of *q until they meet the GPU agent's head.  Returns the number of
  // candidates (in out_id / out_val); *entries_done counts the entries taken.
  uint64_t relax_tail(int alg, const cg_graph* g, const uint32_t* ids, const uint32_t* vals, QueueEnds* q,
                      std::vector<uint32_t>& out_id, std::vector<uint32_t>& out_val, uint64_t* edges_done,
                      uint64_t* entries_done);
  // The same tail loop run by one thread of a parallel region the caller
  // opened (thread t of the team); candidates stay in cid[t] / cval[t] until
  // gather() collects them.
  void tail_worker(int t, int alg, const cg_graph* g, const uint32_t* ids, const uint32_t* vals, QueueEnds* q,
                   uint64_t* edges_done, uint64_t* entries_done);
  void clear_candidates();
  uint64_t gather(std::vector<uint32_t>& out_id, std::vector<uint32_t>& out_val);

  // Bitmap mode (an engine without a GPU): the frontier is the bitmap `in`
  // over `words` 64-bit words, the values are `shadow`; every improved
  // destination's bit is set in `out`.  Returns the number of successful
  // updates (the reference's activeNum).
  uint64_t step_bitmap(int alg, const cg_graph* g, const uint64_t* in, uint64_t* out, uint64_t words,
                       uint64_t* edges_done);

 private:
  template <class Src, class Sink>
  uint64_t three_stage(uint64_t units, Src& src, Sink& sink, const cg_graph* g, int alg);
};

}  // namespace cg
