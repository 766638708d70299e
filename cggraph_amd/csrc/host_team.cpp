// Host-team relaxation (see host_team.hpp).  The per-edge operation is the
// reference's bfs_numa / sssp_numa / cc_numa (my_project/CGgraph/
// host_algorithm.hpp:164-184, 277-300, 412-433) with write_min as a CAS loop
// (host_algorithm.hpp:23-39).  The scheduler is the three stages of
// TaskSteal::threeStage_taskSteal (my_project/CGgraph/taskSteal.hpp:89-274)
// on one socket; the tail loop is the CPU half of the co-executed iteration
// (my_project/CGgraph/CGgraph_engine.cuh:692-855).
#include "host_team.hpp"

#include <omp.h>

#include <algorithm>
#include <cstring>
#include <thread>

#include "common.hpp"

namespace cg {

namespace {

inline uint32_t ld(const uint32_t* p) { return __atomic_load_n(p, __ATOMIC_RELAXED); }

inline bool write_min(uint32_t* p, uint32_t v) {
  uint32_t cur = ld(p);
  while (cur > v) {
    if (__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) return true;
  }
  return false;
}

inline uint32_t edge_candidate(int alg, uint32_t sv, const uint32_t* w, uint64_t k) {
  return alg == CG_BFS ? sv + 1 : (alg == CG_SSSP ? sv + w[k] : sv);
}

// Candidates into per-thread lists (list and tail modes).
struct ListSink {
  HostTeam* team;
  void emit(int t, uint32_t d, uint32_t c) {
    team->cid[t].push_back(d);
    team->cval[t].push_back(c);
  }
};

// Improved destinations into the next frontier bitmap (bitmap mode).
struct BitmapSink {
  uint64_t* out;
  std::vector<uint64_t>* counts;  // one per thread, 8 words apart
  void emit(int t, uint32_t d, uint32_t) {
    __atomic_fetch_or(&out[d >> 6], 1ull << (d & 63), __ATOMIC_RELAXED);
    ++(*counts)[(size_t)t * 8];
  }
};

// Entries of a gathered (vertex, value) list.
struct ListSrc {
  const uint32_t* ids;
  const uint32_t* vals;
  uint64_t n;
  uint32_t* sh;
  template <class F>
  void for_each(uint64_t unit, F&& f) const {
    const uint64_t a = unit * kVertexStealChunk, b = std::min(n, a + kVertexStealChunk);
    for (uint64_t i = a; i < b; ++i) {
      // the gathered value is exact at gather time: tighten the shadow
      write_min(&sh[ids[i]], vals[i]);
      f(i, ids[i], vals[i]);
    }
  }
  void get(uint64_t ent, uint32_t* v, uint32_t* sv) const {
    *v = ids[ent];
    *sv = vals[ent];
  }
};

// Set bits of a frontier bitmap; the source value is read when it is used.
struct BitmapSrc {
  const uint64_t* in;
  const uint32_t* val;
  template <class F>
  void for_each(uint64_t unit, F&& f) const {
    for (uint64_t word = in[unit]; word; word &= word - 1) {
      const uint32_t v = (uint32_t)(unit * 64 + (uint64_t)__builtin_ctzll(word));
      f(v, v, ld(&val[v]));
    }
  }
  void get(uint64_t ent, uint32_t* v, uint32_t* sv) const {
    *v = (uint32_t)ent;
    *sv = ld(&val[ent]);
  }
};

inline uint32_t pack_lo(uint64_t x) { return (uint32_t)x; }
inline uint32_t pack_hi(uint64_t x) { return (uint32_t)(x >> 32); }

}  // namespace

void HostTeam::init(uint32_t V_, int threads_) {
  V = V_;
  threads = threads_ > 0 ? threads_ : omp_get_max_threads();
  if (threads < 1) threads = 1;
  shadow.assign(V, kUnreached);
  cid.assign(threads, {});
  cval.assign(threads, {});
  slots.reset(new Slot[threads]);
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

// The three stages over `units` 64-slot units; src yields (entry, vertex,
// value) per slot, sink takes each candidate that lowered a value.
template <class Src, class Sink>
uint64_t HostTeam::three_stage(uint64_t units, Src& src, Sink& sink, const cg_graph* g, int alg) {
  const int T = threads;
  const uint64_t per = units / (uint64_t)T;
  for (int t = 0; t < T; ++t) {
    slots[t].cur.store(per * (uint64_t)t, std::memory_order_relaxed);
    slots[t].end = t == T - 1 ? units : per * (uint64_t)(t + 1);
    slots[t].status.store(kVertexWorking, std::memory_order_relaxed);
    slots[t].donate.store(kNoDonation, std::memory_order_relaxed);
  }
  working.store(T, std::memory_order_release);
  const uint64_t* off = g->h_off.data();
  const uint32_t* dst = g->h_dst.data();
  const uint32_t* w = g->has_w ? g->h_w.data() : nullptr;
  uint32_t* sh = shadow.data();
  uint64_t edges = 0, stolen = 0, donated = 0;
#pragma omp parallel num_threads(T) reduction(+ : edges, stolen, donated)
  {
    const int tid = omp_get_thread_num();
    Slot& me = slots[tid];
    auto relax_edges = [&](uint32_t sv, uint64_t s, uint64_t e) {
      for (uint64_t k = s; k < e; ++k) {
        const uint32_t d = dst[k];
        const uint32_t cand = edge_candidate(alg, sv, w, k);
        if (cand < ld(&sh[d]) && write_min(&sh[d], cand)) sink.emit(tid, d, cand);
      }
      edges += e - s;
    };
    // Claims 6 400-edge pieces of the vertex donated in `d` until none is
    // left.  The (entry, piece) word is claimed by CAS, so a piece of a vertex
    // the donor has already left can never be taken.
    auto drain = [&](Slot& d) {
      for (;;) {
        uint64_t wv = d.donate.load(std::memory_order_acquire);
        if (wv == kNoDonation) return;
        uint32_t v, sv;
        src.get(pack_hi(wv), &v, &sv);
        const uint64_t s = off[v] + (uint64_t)pack_lo(wv) * kEdgeStealChunk, end = off[v + 1];
        if (s >= end) return;
        if (!d.donate.compare_exchange_weak(wv, wv + 1, std::memory_order_acq_rel, std::memory_order_acquire))
          continue;
        relax_edges(sv, s, std::min(s + kEdgeStealChunk, end));
      }
    };
    auto do_unit = [&](uint64_t u, bool stealing) {
      src.for_each(u, [&](uint64_t ent, uint32_t v, uint32_t sv) {
        const uint64_t s = off[v], e = off[v + 1], deg = e - s;
        // a working thread donates at >= the threshold, a stealing one above
        // it (taskSteal.hpp:136 vs :187)
        if (stealing ? deg > kEdgeStealThreshold : deg >= kEdgeStealThreshold) {
          me.donate.store(ent << 32, std::memory_order_release);
          me.status.store(kEdgeDonate, std::memory_order_release);
          drain(me);
          me.donate.store(kNoDonation, std::memory_order_release);
          me.status.store(stealing ? kVertexStealing : kVertexWorking, std::memory_order_release);
          ++donated;
        } else if (deg) {
          relax_edges(sv, s, e);
        }
      });
    };
    // 1.1 VERTEX_WORKING: the thread's own range
    for (;;) {
      const uint64_t u = me.cur.fetch_add(1, std::memory_order_relaxed);
      if (u >= me.end) break;
      do_unit(u, false);
    }
    // 1.2 VERTEX_STEALING: the other threads' ranges, nearest first
    me.status.store(kVertexStealing, std::memory_order_release);
    for (int o = 1; o < T; ++o) {
      Slot& h = slots[(tid + o) % T];
      while (h.status.load(std::memory_order_acquire) != kVertexStealing) {
        const uint64_t u = h.cur.fetch_add(1, std::memory_order_relaxed);
        if (u >= h.end) break;
        do_unit(u, true);
        ++stolen;
      }
    }
    // 1.3 EDGE_WORKING: help donors until every thread has got here
    working.fetch_sub(1, std::memory_order_acq_rel);
    // (a waiting thread yields now and then: when the process has more
    // runnable threads than cores, the thread everyone waits for may be the
    // one without a core)
    for (uint32_t spin = 1; working.load(std::memory_order_acquire) > 0; ++spin) {
      for (int o = 1; o < T; ++o) {
        Slot& h = slots[(tid + o) % T];
        if (h.status.load(std::memory_order_acquire) == kEdgeDonate) drain(h);
      }
      if (spin % 64 == 0) std::this_thread::yield();
      else __builtin_ia32_pause();
    }
  }
  last_stolen_units = stolen;
  last_donated_vertices = donated;
  return edges;
}

namespace {

// Concatenates the per-thread candidate lists in thread order.
uint64_t gather_candidates(HostTeam* t, std::vector<uint32_t>& out_id, std::vector<uint32_t>& out_val) {
  const int T = t->threads;
  std::vector<uint64_t> base(T + 1, 0);
  for (int k = 0; k < T; ++k) base[k + 1] = base[k] + t->cid[k].size();
  const uint64_t total = base[T];
  if (out_id.size() < total) {
    out_id.resize(total);
    out_val.resize(total);
  }
#pragma omp parallel for num_threads(T) schedule(static, 1)
  for (int k = 0; k < T; ++k) {
    if (!t->cid[k].empty()) {
      std::memcpy(out_id.data() + base[k], t->cid[k].data(), t->cid[k].size() * sizeof(uint32_t));
      std::memcpy(out_val.data() + base[k], t->cval[k].data(), t->cval[k].size() * sizeof(uint32_t));
    }
  }
  return total;
}

}  // namespace

uint64_t HostTeam::relax(int alg, const cg_graph* g, const uint32_t* ids, const uint32_t* vals, uint64_t n,
                         std::vector<uint32_t>& out_id, std::vector<uint32_t>& out_val, uint64_t* edges_done) {
  for (int t = 0; t < threads; ++t) {
    cid[t].clear();
    cval[t].clear();
  }
  ListSrc src{ids, vals, n, shadow.data()};
  ListSink sink{this};
  const uint64_t units = (n + kVertexStealChunk - 1) / kVertexStealChunk;
  const uint64_t edges = three_stage(units, src, sink, g, alg);
  if (edges_done) *edges_done = edges;
  return gather_candidates(this, out_id, out_val);
}

void HostTeam::clear_candidates() {
  for (int t = 0; t < threads; ++t) {
    cid[t].clear();
    cval[t].clear();
  }
}

uint64_t HostTeam::gather(std::vector<uint32_t>& out_id, std::vector<uint32_t>& out_val) {
  return gather_candidates(this, out_id, out_val);
}

void HostTeam::tail_worker(int t, int alg, const cg_graph* g, const uint32_t* ids, const uint32_t* vals,
                           QueueEnds* q, uint64_t* edges_done, uint64_t* entries_done) {
  const uint64_t* off = g->h_off.data();
  const uint32_t* dst = g->h_dst.data();
  const uint32_t* w = g->has_w ? g->h_w.data() : nullptr;
  uint32_t* sh = shadow.data();
  uint64_t edges = 0, taken = 0;
  for (;;) {
    uint32_t lo = 0, hi = 0;
    if (!q->claim_tail((uint32_t)kVertexStealChunk, &lo, &hi)) break;
    taken += hi - lo;
    for (uint32_t i = lo; i < hi; ++i) {
      const uint32_t v = ids[i - q->host_floor], sv = vals[i - q->host_floor];
      write_min(&sh[v], sv);
      const uint64_t s = off[v], e = off[v + 1];
      edges += e - s;
      for (uint64_t k = s; k < e; ++k) {
        const uint32_t d = dst[k];
        const uint32_t cand = edge_candidate(alg, sv, w, k);
        if (cand < ld(&sh[d]) && write_min(&sh[d], cand)) {
          cid[t].push_back(d);
          cval[t].push_back(cand);
        }
      }
    }
  }
  *edges_done = edges;
  *entries_done = taken;
}

uint64_t HostTeam::relax_tail(int alg, const cg_graph* g, const uint32_t* ids, const uint32_t* vals, QueueEnds* q,
                              std::vector<uint32_t>& out_id, std::vector<uint32_t>& out_val, uint64_t* edges_done,
                              uint64_t* entries_done) {
  clear_candidates();
  uint64_t edges = 0, taken = 0;
#pragma omp parallel num_threads(threads) reduction(+ : edges, taken)
  {
    uint64_t e = 0, n = 0;
    tail_worker(omp_get_thread_num(), alg, g, ids, vals, q, &e, &n);
    edges += e;
    taken += n;
  }
  if (edges_done) *edges_done = edges;
  if (entries_done) *entries_done = taken;
  return gather_candidates(this, out_id, out_val);
}

uint64_t HostTeam::step_bitmap(int alg, const cg_graph* g, const uint64_t* in, uint64_t* out, uint64_t words,
                               uint64_t* edges_done) {
  std::vector<uint64_t> counts((size_t)threads * 8, 0);
  BitmapSrc src{in, shadow.data()};
  BitmapSink sink{out, &counts};
  const uint64_t edges = three_stage(words, src, sink, g, alg);
  if (edges_done) *edges_done = edges;
  uint64_t updates = 0;
  for (int t = 0; t < threads; ++t) updates += counts[(size_t)t * 8];
  return updates;
}

}  // namespace cg
