// Frontier kernels of the traversal engine, written for gfx950 (wave64).
//
// One push iteration is four launches over the frontier bitmap:
//   k_stats    one lane per 64-bit frontier word: popcount, out-degree sum and
//              log-degree histogram, reduced per workgroup into a partials
//              array.  It also zeroes the next bitmap.
//   k_stats_reduce  one workgroup reduces the partials in a fixed order,
//              publishes the record to mapped host memory and -- in learned
//              mode -- runs the scheduler network on the iteration's features
//              (inference fused into the statistics pass).
//   k_compact  splits active vertices by out-degree into three work queues:
//              small (edges spread evenly over a wavefront's lanes), medium
//              (one wavefront per vertex) and chunked (a vertex of degree >=
//              bin_chunk becomes ceil(deg / chunk_edges) wavefront items).
//              Queue slots come from wave-level scans and one atomic per queue
//              per workgroup.
//   k_expand   one wavefront per work item; every edge relaxes value[dst] with
//              a strict-less pre-check and atomic min; an improvement sets the
//              destination's bit in the next frontier.
// The relaxation is the reference's edge operation (paths relative to
// CGgraph/my_project/CGgraph/): device_algorithm.cuh:51-54 (BFS: value+1),
// :253-257 (SSSP: value+w), host_algorithm.hpp:424-429 (CC: value).  The
// degree tiers replace WarpDonate::schedule (device_warpDonate.cuh:43-388),
// whose CTA / warp / half-warp tiers assume 32-lane warps.
#include <chrono>
#include <cstddef>

#include "kernels.hpp"
#include "mlp_dev.hpp"
#include "sched.hpp"

namespace cg {

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kFinishBlock = 256;
constexpr unsigned long long kNetEdgeShare = 256;  // the fused finish's network: frontiers of >= E / 256 edges  // the fused finish (kFinishParts compaction partials per thread)
constexpr int kFinishParts = 4;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint32_t wave_id() { return threadIdx.x / kWave; }

// Exclusive prefix sum over the 64 lanes; every lane must be active.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t x, uint32_t* total) {
  const uint32_t lane = lane_id();
  uint32_t v = x;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    uint32_t t = __shfl_up(v, d, kWave);
    if (lane >= (uint32_t)d) v += t;
  }
  *total = __shfl(v, kWave - 1, kWave);
  return v - x;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x) {
#pragma unroll
  for (int d = kWave / 2; d >= 1; d >>= 1) x += __shfl_xor(x, d, kWave);
  return x;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t x) {
#pragma unroll
  for (int d = kWave / 2; d >= 1; d >>= 1) x += __shfl_xor(x, d, kWave);
  return x;
}

__device__ __forceinline__ uint32_t hist_bin(uint64_t deg) {
  uint32_t l = 63 - __clzll(deg + 1);  // floor(log2(deg + 1))
  l >>= 1;
  return l > 7 ? 7 : l;
}

// A lane's 8-bin degree histogram in two registers of four 16-bit counts
// (an array indexed by the bin would live in scratch memory).  A lane counts
// at most 16 vertices per slice and a grid-stride pass gives it fewer than
// 2^12 slices, so no count reaches 2^16.
struct Hist8 {
  uint64_t lo = 0, hi = 0;
  __device__ __forceinline__ void add(uint32_t b) {
    const uint64_t inc = 1ull << (16 * (b & 3u));
    if (b & 4u) hi += inc;
    else lo += inc;
  }
  __device__ __forceinline__ uint32_t get(int k) const {
    return (uint32_t)(((k & 4) ? hi : lo) >> (16 * (k & 3))) & 0xFFFFu;
  }
};

__device__ __forceinline__ void stripe_add(uint32_t* stripes, uint32_t v) {
  if (v) atomicAdd(stripes + (blockIdx.x % kStripes) * kStripeStride, v);
}

// The second set of stripes holds minima (the next frontier's smallest value
// under hub deferral); k_stats_reduce reads and resets them.
__device__ __forceinline__ void stripe_min(uint32_t* stripes, uint32_t v) {
  if (v != 0xFFFFFFFFu && lane_id() == 0)
    atomicMin(stripes + kStripeMinBase + (blockIdx.x % kStripes) * kStripeStride, v);
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t x) {
#pragma unroll
  for (int d = kWave / 2; d >= 1; d >>= 1) {
    const uint32_t o = __shfl_xor(x, d, kWave);
    x = o < x ? o : x;
  }
  return x;
}

enum { ALG_BFS = CG_BFS, ALG_SSSP = CG_SSSP, ALG_CC = CG_CC };

// Edge operation shared by every push kernel.  Returns 1 when this update
// reached the vertex for the first time.  EXACT: activate only when the
// atomic min returned a larger value.  Otherwise the atomics return nothing
// (the wavefront does not wait for them) and the destination is activated on
// the strict-less pre-check: a pre-check that passes means the value is
// lowered in this iteration by this or another update (or the read was
// stale), so the active set is a superset and the fixed point is the same.
template <int ALG, class WT>
__device__ __forceinline__ uint32_t candidate(uint32_t sv, uint64_t e, const WT* __restrict__ w) {
  if constexpr (ALG == ALG_BFS) return sv + 1;
  else if constexpr (ALG == ALG_SSSP) return sv + w[e];
  else return sv;
}

// What the activation of a destination is folded into.  MinTrack (with
// TRACK): the smallest activating candidate, so the next frontier's smallest
// value is known without reading it back (every activated vertex ends at or
// below a tracked candidate, and the smallest tracked candidate is some
// vertex's final value).  FirstVisits: the vertices whose bit an activation
// set first in the next bitmap (the atomicOr returns the old word), with their
// out-degrees -- the next frontier's size and edges, counted where it is made.
struct MinTrack {
  static constexpr bool counts() { return false; }
  uint32_t mn = 0xFFFFFFFFu;
  __device__ __forceinline__ void track(uint32_t c) { mn = c < mn ? c : mn; }
  __device__ __forceinline__ void first(uint32_t) {}
};
struct FirstVisits {
  static constexpr bool counts() { return true; }
  const uint64_t* __restrict__ off;
  uint32_t n = 0, z = 0;
  unsigned long long m = 0;
  __device__ __forceinline__ void track(uint32_t) {}
  __device__ __forceinline__ void first(uint32_t d) {
    const uint64_t deg = off[d + 1] - off[d];
    ++n;
    m += deg;
    z += deg == 0 ? 1u : 0u;
  }
};

// The update of destination d with candidate cand, given a value `seen` read
// earlier (values only decrease, so a stale read is never below the current
// value and can only let an update through that the atomic then rejects).
template <int ALG, bool EXACT, bool TRACK = false, class A = MinTrack>
__device__ __forceinline__ uint32_t commit(uint32_t cand, uint32_t seen, uint32_t d, uint32_t* __restrict__ val,
                                           uint64_t* __restrict__ out, A& acc) {
  if (cand < seen) {
    bool act;
    uint32_t first;
    if constexpr (EXACT) {
      const uint32_t old = atomicMin(val + d, cand);
      act = cand < old;
      first = old == kUnreached ? 1u : 0u;
    } else {
      (void)atomicMin(val + d, cand);
      act = true;
      first = seen == kUnreached ? 1u : 0u;
    }
    if (act) {
      unsigned long long* const wp = reinterpret_cast<unsigned long long*>(out) + (d >> 6);
      const unsigned long long bit = 1ull << (d & 63);
      if constexpr (A::counts()) {
        if (!(atomicOr(wp, bit) & bit)) acc.first(d);
      } else {
        (void)atomicOr(wp, bit);
      }
      if constexpr (TRACK) acc.track(cand);
      return first;
    }
  }
  return 0u;
}

template <int ALG, bool EXACT, bool TRACK = false, class A = MinTrack, class WT = uint32_t>
__device__ __forceinline__ uint32_t relax(uint32_t sv, uint64_t e, const uint32_t* __restrict__ dst,
                                          const WT* __restrict__ w, uint32_t* __restrict__ val,
                                          uint64_t* __restrict__ out, A& acc) {
  const uint32_t d = dst[e];
  return commit<ALG, EXACT, TRACK>(candidate<ALG>(sv, e, w), val[d], d, val, out, acc);
}

// Edges [s, e) of one source with value sv, lane-strided over a wavefront.
// Four edges per lane are loaded (destination, weight, destination value)
// before any of them is committed, so four gathers are in flight per lane.
template <int ALG, bool EXACT, bool TRACK, class A, class WT>
__device__ __forceinline__ uint32_t relax_strided(uint32_t sv, uint64_t s, uint64_t e, uint32_t lane,
                                                  const uint32_t* __restrict__ dst, const WT* __restrict__ w,
                                                  uint32_t* __restrict__ val, uint64_t* __restrict__ out, A& acc) {
  uint32_t n = 0;
  uint64_t k = s + lane;
  for (; k + 3 * kWave < e; k += 4 * kWave) {
    const uint32_t d0 = dst[k], d1 = dst[k + kWave], d2 = dst[k + 2 * kWave], d3 = dst[k + 3 * kWave];
    const uint32_t c0 = candidate<ALG>(sv, k, w), c1 = candidate<ALG>(sv, k + kWave, w);
    const uint32_t c2 = candidate<ALG>(sv, k + 2 * kWave, w), c3 = candidate<ALG>(sv, k + 3 * kWave, w);
    const uint32_t s0 = val[d0], s1 = val[d1], s2 = val[d2], s3 = val[d3];
    n += commit<ALG, EXACT, TRACK>(c0, s0, d0, val, out, acc);
    n += commit<ALG, EXACT, TRACK>(c1, s1, d1, val, out, acc);
    n += commit<ALG, EXACT, TRACK>(c2, s2, d2, val, out, acc);
    n += commit<ALG, EXACT, TRACK>(c3, s3, d3, val, out, acc);
  }
  for (; k < e; k += kWave) n += relax<ALG, EXACT, TRACK>(sv, k, dst, w, val, out, acc);
  return n;
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_u64(uint64_t* __restrict__ p, uint64_t v, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_iota_u32(uint32_t* __restrict__ p, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = (uint32_t)i;
}

__global__ void k_bitmap_prefix(uint64_t* __restrict__ bm, uint64_t words, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < words;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t lo = i * 64;
    uint64_t word = 0;
    if (lo + 64 <= n) word = ~0ull;
    else if (lo < n) word = (1ull << (n - lo)) - 1;
    bm[i] = word;
  }
}

__global__ void k_set_root(uint32_t* __restrict__ val, uint64_t* __restrict__ bm, uint32_t root) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    val[root] = 0;
    bm[root >> 6] |= 1ull << (root & 63);
  }
}

__global__ void k_init_run_state(RunState* rs, unsigned long long visited) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    rs->visited = visited;
    rs->iter = 0;
    rs->pull = 0;
    rs->m_explored = 0;
    rs->prev_active = 0;
    rs->done = 0;
    rs->epoch = 0;
    rs->pad = 0;
  }
}

// One launch for everything a traversal run starts from (the reference's
// initAlgorithm, CGgraph_engine.cuh:1259-1316, plus the engine's own state):
// values and the first frontier, the other frontier bitmap, the striped
// counters, both statistics slots, both run states and, for BFS, the root's
// counts.  One kernel instead of a chain of small launches and copies.
__global__ __launch_bounds__(kBlock) void k_init_traverse(TraverseInit a) {
  const uint64_t t0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const uint64_t nt = (uint64_t)gridDim.x * blockDim.x;
  const bool cc = a.alg == CG_CC;
  const bool pull1 = cc && a.in_off;  // CC: the first round of labels pulled below
  // values (and records) four at a time with 16-byte stores, then the rest
  const uint64_t n4 = a.V / 4;
  for (uint64_t i = t0; i < n4; i += nt) {
    const uint32_t b = (uint32_t)(4 * i);
    uint4 x;
    if (cc) {
      x = make_uint4(b, b + 1, b + 2, b + 3);
    } else {
      const bool here = (a.root >> 2) == i;
      const uint32_t r = a.root & 3u;
      x = make_uint4(here && r == 0 ? 0u : kUnreached, here && r == 1 ? 0u : kUnreached,
                     here && r == 2 ? 0u : kUnreached, here && r == 3 ? 0u : kUnreached);
    }
    if (!pull1) reinterpret_cast<uint4*>(a.val)[i] = x;
    if (a.xval) reinterpret_cast<uint4*>(a.xval)[i] = make_uint4(0u, 0u, 0u, 0u);
  }
  for (uint64_t i = 4 * n4 + t0; i < a.V; i += nt) {
    if (!pull1) a.val[i] = cc ? (uint32_t)i : (i == a.root ? 0u : kUnreached);
    if (a.xval) a.xval[i] = 0;
  }
  for (uint64_t i = t0; i < a.bm_words; i += nt) {
    uint64_t w = 0;
    if (cc) {
      const uint64_t lo = i * 64;
      if (lo + 64 <= a.cc_active) w = ~0ull;
      else if (lo < a.cc_active) w = (1ull << (a.cc_active - lo)) - 1;
    } else if (i == (a.root >> 6)) {
      w = 1ull << (a.root & 63);
    }
    if (!pull1) a.bm0[i] = w;
    a.bm1[i] = 0;
  }
  if (pull1) {
    // Label propagation's first round as a pull: every vertex takes the
    // smallest of its id and its in-neighbours' ids, and the in-edges are
    // sorted by source, so that is its first in-edge's source.  The
    // frontier of the next round is the vertices whose label fell (a vertex
    // that kept its id has reached its out-neighbours through their pulls).
    // A wavefront owns a bitmap word: lane l is vertex 64 w + l.
    const uint64_t lane = threadIdx.x & (kWave - 1);
    const uint64_t w0 = t0 / kWave, nw = nt / kWave;
    for (uint64_t w = w0; w < a.bm_words; w += nw) {
      const uint64_t v = w * kWave + lane;
      uint32_t label = (uint32_t)v;
      if (v < a.V) {
        const uint64_t lo = a.in_off[v], hi = a.in_off[v + 1];
        if (hi > lo) {
          const uint32_t s0 = a.in_src[lo];
          label = s0 < label ? s0 : label;
        }
        a.val[v] = label;
      }
      const uint64_t fell = __ballot(v < a.V && label < (uint32_t)v);
      if (lane == 0) a.bm0[w] = fell;
    }
  }
  for (uint64_t i = t0; i < (uint64_t)kStripeWords; i += nt)
    a.stripes[i] = (a.minima_empty && i >= (uint64_t)kStripeMinBase) ? 0xFFFFFFFFu : 0u;
  if (a.acc)
    for (uint64_t i = t0; i < (uint64_t)kAccSlots; i += nt) a.acc[i] = 0;
  // BFS: iteration one's direction and queues for frontier {root}, as
  // k_bfs_front decides and compacts them (nothing explored yet, so the push
  // rule reads 15 * deg(root) > E; the chunk size shrinks as for any frontier)
  uint64_t f_deg = 0;
  uint32_t f_pull = 0, f_ce = 0, f_small = 0, f_med = 0, f_chunk = 0;
  if (a.bfs_first) {
    f_deg = a.off[a.root + 1] - a.off[a.root];
    f_pull = 15ull * f_deg > a.E ? 1u : 0u;
    f_ce = a.bins.chunk_edges;
    while (f_ce > (uint32_t)kWave && f_deg < (uint64_t)f_ce * kChunkSpread) f_ce >>= 1;
    if (!f_pull && f_deg) {
      if (f_deg < a.bins.bin_wave) f_small = 1;
      else if (f_deg < a.bins.bin_chunk) f_med = 1;
      else f_chunk = (uint32_t)((f_deg + f_ce - 1) / f_ce);
    }
    for (uint64_t i = t0; i < f_chunk; i += nt) a.q_chunk[i] = ((uint64_t)a.root << 32) | i;
    if (t0 == 0 && f_small) a.q_small[0] = a.root;
    if (t0 == 0 && f_med) a.q_med[0] = a.root;
  }
  // both statistics slots cleared; slot 1 (iteration one's) may carry the
  // learned schedule's "no choice" or the BFS decision and queue counts
  constexpr uint64_t kStU32 = sizeof(DevStats) / 4;
  for (uint64_t i = t0; i < 2 * kStU32; i += nt) {
    const uint64_t j = i - kStU32;  // field word of slot 1 (wraps for slot 0)
    uint32_t x = 0;
    if (i >= kStU32) {
      if (a.choice1_none && j == offsetof(DevStats, choice) / 4) x = 0xFFFFFFFFu;
      if (a.bfs_first) {
        if (j == offsetof(DevStats, n_small) / 4) x = f_small;
        else if (j == offsetof(DevStats, n_med) / 4) x = f_med;
        else if (j == offsetof(DevStats, n_chunk) / 4) x = f_chunk;
        else if (j == offsetof(DevStats, choice) / 4) x = f_pull;
        else if (j == offsetof(DevStats, chunk_edges) / 4) x = f_ce;
      }
    }
    reinterpret_cast<uint32_t*>(a.st)[i] = x;
  }
  if (a.bfs_cnt) {
    const uint64_t deg = a.off[a.root + 1] - a.off[a.root];
    for (uint64_t i = t0; i < 2ull * kBfsCntWords; i += nt)
      a.bfs_cnt[i] = i == 0 ? 1ull : (i == 1 ? deg : (i == 2 ? (deg == 0 ? 1ull : 0ull) : 0ull));
  }
  if (t0 == 0) {
    for (int k = 0; k < 2; ++k) {
      a.rs[k].visited = a.visited0;
      a.rs[k].iter = 0;
      a.rs[k].pull = 0;
      a.rs[k].m_explored = 0;
      a.rs[k].prev_active = 0;
      a.rs[k].pad = 0;
      a.rs[k].done = 0;
      a.rs[k].epoch = a.epoch;
    }
    if (a.bfs_first) {
      // run state after iteration one's decision, and its record
      a.rs[1].visited = a.visited0 + 1;
      a.rs[1].iter = 1;
      a.rs[1].pull = f_pull;
      a.rs[1].m_explored = f_deg;
      a.rs[1].prev_active = 1;
      DevStats d{};
      d.n_active = 1;
      d.m_active = f_deg;
      d.n_zero_deg = f_deg == 0 ? 1u : 0u;
      d.visited = a.visited0 + 1;
      d.choice = (int32_t)f_pull;
      d.iter = 0xFFFFFFFFu;
      *a.bfs_host_out = d;
      // the release store orders the record before `iter` at system scope
      __hip_atomic_store(&a.bfs_host_out->iter, record_tag(a.epoch, 1u), __ATOMIC_RELEASE,
                         __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

// The statistics and the compaction give each lane one 16-vertex slice of
// the frontier bitmap (four lanes per 64-bit word, all in one wavefront), so
// a small graph still spreads over many wavefronts and no lane walks more
// than 16 vertices.
constexpr uint64_t kSliceBits = 16;
constexpr uint64_t kSlicesPerWord = 64 / kSliceBits;

// Calls f(v, out_degree) for the set bits of the slice of vertices [base,
// base + 16), in vertex order.  A dense slice loads its 17 offsets at once
// (all in flight together) instead of one dependent pair per vertex; a slice
// that would read past off[V] takes the per-vertex path.
template <class F>
__device__ __forceinline__ void for_each_in_slice(uint32_t bits, uint64_t base, const uint64_t* __restrict__ off,
                                                  uint32_t V, F&& f) {
  if (!bits) return;
  if (__popc(bits) >= 4 && base + 16 <= V) {
    uint64_t o[17];
#pragma unroll
    for (int j = 0; j < 17; ++j) o[j] = off[base + j];
#pragma unroll
    for (int j = 0; j < 16; ++j)
      if ((bits >> j) & 1u) f((uint32_t)(base + j), o[j + 1] - o[j]);
  } else {
    for (; bits; bits &= bits - 1) {
      const uint32_t v = (uint32_t)(base + (uint32_t)__builtin_ctz(bits));
      f(v, off[v + 1] - off[v]);
    }
  }
}

// Dense frontiers: a wavefront's 64 slices cover 1024 consecutive vertices,
// whose 1025 offsets it loads coalesced (lane l takes entries l, l + 64, ...)
// into LDS, one pad entry after every 16 so that lane l's 17 offsets sit at
// 17l .. 17l + 15 and 17l + 17 (two-way bank conflicts at most).  Per-lane
// loads of those offsets touch a different line per lane and are bound by
// the L1's line rate instead of by HBM.
constexpr int kWaveVerts = kWave * (int)kSliceBits;                 // 1024
constexpr int kStageEntries = kWaveVerts + kWaveVerts / 16 + 2;      // 1090
__device__ __forceinline__ uint32_t stage_index(uint32_t i) { return i + (i >> 4); }

__device__ __forceinline__ void stage_offsets(const uint64_t* __restrict__ off, uint32_t V, uint64_t vb,
                                              uint64_t* __restrict__ lds) {
  for (uint32_t i = lane_id(); i <= (uint32_t)kWaveVerts; i += kWave) {
    const uint64_t v = vb + i;
    lds[stage_index(i)] = v <= V ? off[v] : 0ull;  // entries past V are never used (no bits there)
  }
}

// for_each_in_slice over offsets staged by stage_offsets; `row` is the
// wavefront's staging area, the lane's slice is its lane_id()-th.
template <class F>
__device__ __forceinline__ void for_each_in_slice_staged(uint32_t bits, uint64_t base,
                                                         const uint64_t* __restrict__ row, F&& f) {
  if (!bits) return;
  const uint32_t l0 = lane_id() * (uint32_t)kSliceBits;
  for (; bits; bits &= bits - 1) {
    const uint32_t j = (uint32_t)__builtin_ctz(bits);
    const uint64_t a = row[stage_index(l0 + j)], b = row[stage_index(l0 + j + 1)];
    f((uint32_t)(base + j), b - a);
  }
}

// Slice si of a bitmap: its 16 bits, and its first vertex.
__device__ __forceinline__ uint32_t slice_bits(uint64_t word, uint64_t si) {
  return (uint32_t)(word >> (kSliceBits * (si % kSlicesPerWord))) & 0xFFFFu;
}

// The four lanes of a word combine their 16-bit slices into the word (lane
// order within the wavefront puts them next to each other); every lane of
// the wavefront must call this.
__device__ __forceinline__ uint64_t join_slices(uint32_t bits, uint64_t si) {
  uint64_t w = (uint64_t)bits << (kSliceBits * (si % kSlicesPerWord));
  w |= __shfl_xor(w, 1, kWave);
  w |= __shfl_xor(w, 2, kWave);
  return w;
}

// k_stats grid cap: each workgroup leaves one partial for the reduction
constexpr uint64_t kStatsMaxBlocks = 2048;

// Sums the wavefronts' partials of a workgroup and stores the workgroup's
// partial.  No fence or ticket: the reduction runs in the next kernel, whose
// launch orders it after every partial.  (An agent-scope fence per workgroup
// writes back and invalidates the XCD's L2 -- tens of microseconds per
// statistics pass, and a cold L2 for the expansion behind it.)
__device__ __forceinline__ void stats_block_part(const StatsPart& p, StatsPart* __restrict__ parts) {
  __shared__ StatsPart wpart[kWavesPerBlock];
  if (lane_id() == 0) wpart[wave_id()] = p;
  __syncthreads();
  if (threadIdx.x == 0) {
    StatsPart b = wpart[0];
    for (int w = 1; w < kWavesPerBlock; ++w) {
      b.n_active += wpart[w].n_active;
      b.n_zero += wpart[w].n_zero;
      b.m_active += wpart[w].m_active;
      for (int k = 0; k < 8; ++k) b.hist[k] += wpart[w].hist[k];
      b.n_near += wpart[w].n_near;
      b.min_val = wpart[w].min_val < b.min_val ? wpart[w].min_val : b.min_val;
      b.n_def += wpart[w].n_def;
      b.m_def += wpart[w].m_def;
    }
    parts[blockIdx.x] = b;
  }
}

// One workgroup: fixed-order reduction of the partials, first-visit stripes,
// run state, features and (with weights) the scheduler network, then the
// counters of the next compaction and the publication to mapped host memory
// (`iter` last, so the host can poll it).
__device__ void stats_finish(const StatsPart* __restrict__ parts, uint32_t nparts, uint32_t* __restrict__ stripes,
                             RunState* __restrict__ rs, uint32_t V, uint64_t E, const float* __restrict__ mlp,
                             DevStats* __restrict__ st, DevStats* __restrict__ host_out, int direction,
                             const DevStats* __restrict__ prev, int min_from_stripes) {
  __shared__ unsigned long long s_m[kBlock];
  __shared__ uint32_t s_c[kBlock][11];
  __shared__ uint32_t s_min[kBlock];
  __shared__ float w_mlp[kMlpLds];
  MlpStage<kBlock> wst;  // weights loaded now, stored to LDS after the reduction
  if (mlp) wst.load(mlp);
  unsigned long long m = 0;
  uint32_t c[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  uint32_t mn = 0xFFFFFFFFu;
  for (uint32_t b = threadIdx.x; b < nparts; b += kBlock) {
    const StatsPart q = parts[b];
    m += q.m_active;
    c[0] += q.n_active;
    c[1] += q.n_zero;
    for (int k = 0; k < 8; ++k) c[2 + k] += q.hist[k];
    c[10] += q.n_near;
    mn = q.min_val < mn ? q.min_val : mn;
  }
  s_m[threadIdx.x] = m;
  for (int k = 0; k < 11; ++k) s_c[threadIdx.x][k] = c[k];
  s_min[threadIdx.x] = mn;
  __syncthreads();
  for (int w = kBlock / 2; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s_m[threadIdx.x] += s_m[threadIdx.x + w];
      for (int k = 0; k < 11; ++k) s_c[threadIdx.x][k] += s_c[threadIdx.x + w][k];
      const uint32_t o = s_min[threadIdx.x + w];
      if (o < s_min[threadIdx.x]) s_min[threadIdx.x] = o;
    }
    __syncthreads();
  }
  __shared__ float feat[CG_NUM_FEATURES];
  __shared__ float mlp_sh[2 * kMlpH + kMlpOut];
  __shared__ float probs[kMlpOut];
  __shared__ int best;
  // thread 0 builds the record in registers and stores it once: no read-back
  // of global memory on the serial path
  DevStats d{};
  if (threadIdx.x == 0) {
    uint32_t nn = 0;
#pragma unroll
    for (int k = 0; k < kStripes; ++k) nn += stripes[k * kStripeStride];
#pragma unroll
    for (int k = 0; k < kStripes; ++k) stripes[k * kStripeStride] = 0;
    const unsigned long long visited = rs->visited + nn;
    const uint32_t iter = rs->iter + 1;
    rs->visited = visited;
    rs->iter = iter;
    d.n_active = s_c[0][0];
    d.n_zero_deg = s_c[0][1];
    d.m_active = s_m[0];
    for (int k = 0; k < 8; ++k) d.hist[k] = s_c[0][2 + k];
    d.n_new = nn;
    d.n_near = s_c[0][10];
    d.min_val = s_min[0];
    d.visited = visited;
    d.iter = iter;
    d.choice = -1;
    if (prev && iter > 1) {
      // written by the previous iteration's compaction, which has finished
      d.n_def_prev = prev->n_def;
      d.m_def_prev = prev->m_def;
    }
    if (min_from_stripes) {
      // the frontier's smallest value, left by the previous expansion and
      // compaction (the run starts with 0: the root)
      uint32_t mv = 0xFFFFFFFFu;
#pragma unroll
      for (int k = 0; k < kStripes; ++k) {
        const uint32_t x = stripes[kStripeMinBase + k * kStripeStride];
        mv = x < mv ? x : mv;
        stripes[kStripeMinBase + k * kStripeStride] = 0xFFFFFFFFu;
      }
      d.min_val = mv;
    }
    if (direction) {
      // Beamer, Asanovic, Patterson (SC'12): bottom-up when the frontier's
      // edges exceed the unexplored edges / 15, back to push when the frontier
      // falls below V / 18 and shrinks
      const unsigned long long m_unexplored = E > rs->m_explored ? E - rs->m_explored : 0ull;
      bool pull;
      if (!rs->pull) pull = 15ull * d.m_active > m_unexplored;
      else pull = !(18ull * d.n_active < (unsigned long long)V && d.n_active < rs->prev_active);
      rs->pull = pull ? 1u : 0u;
      rs->m_explored += d.m_active;
      rs->prev_active = d.n_active;
      d.choice = pull ? 1 : 0;
    }
    // this iteration's compaction counts from zero (n_small / n_med / n_chunk
    // stay 0)
    *st = d;
    make_features(d, V, E, visited, iter, feat);
  }
  if (mlp) wst.store(w_mlp);
  __syncthreads();
  if (mlp) {
    mlp_forward_lds(w_mlp, feat, mlp_sh, probs, &best);
    if (threadIdx.x < kMlpOut) st->probs[threadIdx.x] = probs[threadIdx.x];
    if (threadIdx.x == 0) {
      // the network defers (choice -1) unless its output is a finite
      // distribution whose top choice has probability >= kLearnedMinProb
      float sum = 0.f;
      bool finite = true;
      for (int k = 0; k < kMlpOut; ++k) {
        finite = finite && isfinite(probs[k]);
        sum += probs[k];
      }
      const bool usable = finite && fabsf(sum - 1.f) < 1e-3f && probs[best] >= kLearnedMinProb;
      best = usable ? best : -1;
      st->choice = best;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && host_out) {
    if (mlp) {
      for (int k = 0; k < kMlpOut; ++k) d.probs[k] = probs[k];
      d.choice = best;
    }
    const uint32_t itv = record_tag(rs->epoch, d.iter);
    d.iter = 0xFFFFFFFFu;
    *host_out = d;
    // the release store orders the record before `iter` at system scope
    __hip_atomic_store(&host_out->iter, itv, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

__global__ __launch_bounds__(kBlock) void k_stats(const uint64_t* __restrict__ in,
                                                  uint64_t* __restrict__ zero_me, uint64_t words,
                                                  DevGraph g, StatsPart* __restrict__ parts,
                                                  const uint32_t* __restrict__ val, uint32_t threshold) {
  StatsPart p{};
  p.min_val = 0xFFFFFFFFu;
  uint32_t c_act = 0, c_zero = 0, c_near = 0, mn = 0xFFFFFFFFu;
  uint64_t m = 0;
  Hist8 h;
  // a capped grid strides over the slices (fewer workgroup partials)
  const uint64_t slices = words * kSlicesPerWord;
  for (uint64_t base = blockIdx.x * (uint64_t)kBlock; base < slices; base += (uint64_t)gridDim.x * kBlock) {
    const uint64_t si = base + threadIdx.x;
    uint32_t bits = 0;
    if (si < slices) {
      const uint64_t wi = si / kSlicesPerWord;
      bits = slice_bits(in[wi], si);
      if (zero_me && si % kSlicesPerWord == 0) zero_me[wi] = 0;
    }
    c_act += (uint32_t)__popc(bits);
    for_each_in_slice(bits, si * kSliceBits, g.off, g.V, [&](uint32_t v, uint64_t deg) {
      m += deg;
      h.add(hist_bin(deg));
      c_zero += deg == 0;
      if (val) {
        const uint32_t x = val[v];
        mn = x < mn ? x : mn;
        c_near += x < threshold;
      }
    });
  }
  p.n_active = wave_sum_u32(c_act);
  p.m_active = wave_sum_u64(m);
  p.n_zero = wave_sum_u32(c_zero);
#pragma unroll
  for (int k = 0; k < 8; ++k) p.hist[k] = wave_sum_u32(h.get(k));
  if (val) {
    p.n_near = wave_sum_u32(c_near);
#pragma unroll
    for (int d = kWave / 2; d >= 1; d >>= 1) {
      const uint32_t o = __shfl_xor(mn, d, kWave);
      mn = o < mn ? o : mn;
    }
    p.min_val = mn;
  }
  stats_block_part(p, parts);
}

__global__ __launch_bounds__(kBlock) void k_stats_reduce(const StatsPart* __restrict__ parts, uint32_t nparts,
                                                         uint32_t* __restrict__ stripes, RunState* __restrict__ rs,
                                                         uint32_t V, uint64_t E, const float* __restrict__ mlp,
                                                         DevStats* __restrict__ st,
                                                         DevStats* __restrict__ host_out, int direction,
                                                         const DevStats* __restrict__ prev, int min_from_stripes) {
  stats_finish(parts, nparts, stripes, rs, V, E, mlp, st, host_out, direction, prev, min_from_stripes);
}

// Chunk-tier items of a frontier vertex: nch entries (v, c).  A lane writes
// up to kLaneChunks of them itself; a longer run (a hub, or a small frontier
// cut into small chunks) is listed for its wavefront and written by all 64
// lanes once the pass is over, so one lane does not store thousands of
// entries while the others wait.  A full list falls back to the lane.
constexpr uint32_t kLaneChunks = 8;
constexpr uint32_t kHubSlots = 32;
struct WaveHubs {
  uint32_t n;
  uint32_t v[kHubSlots], nch[kHubSlots], base[kHubSlots];
};

__device__ __forceinline__ void put_chunks(uint64_t* __restrict__ q_chunk, uint32_t& b, uint32_t v, uint32_t nch,
                                           WaveHubs& h) {
  if (nch > kLaneChunks) {
    const uint32_t k = atomicAdd(&h.n, 1u);
    if (k < kHubSlots) {
      h.v[k] = v;
      h.nch[k] = nch;
      h.base[k] = b;
      b += nch;
      return;
    }
  }
  for (uint32_t c = 0; c < nch; ++c) q_chunk[b++] = ((uint64_t)v << 32) | c;
}

// The whole wavefront, converged, after its lanes' put_chunks calls.
__device__ __forceinline__ void write_hub_chunks(uint64_t* __restrict__ q_chunk, const WaveHubs& h) {
  const uint32_t n = h.n < kHubSlots ? h.n : kHubSlots;
  for (uint32_t k = 0; k < n; ++k) {
    const uint64_t v = (uint64_t)h.v[k] << 32;
    const uint32_t nch = h.nch[k], base = h.base[k];
    for (uint32_t c = lane_id(); c < nch; c += kWave) q_chunk[base + c] = v | c;
  }
}

__global__ __launch_bounds__(kBlock) void k_compact(const uint64_t* __restrict__ in, uint64_t words,
                                                    DevGraph g, Bins bins,
                                                    uint32_t* __restrict__ q_small,
                                                    uint32_t* __restrict__ q_med,
                                                    uint64_t* __restrict__ q_chunk,
                                                    DevStats* __restrict__ st,
                                                    const uint32_t* __restrict__ val, uint32_t threshold,
                                                    uint64_t* __restrict__ carry, int use_choice,
                                                    TierTable tiers, uint32_t skip_bits, HubDefer hub) {
  __shared__ uint32_t s_tot[kWavesPerBlock][3];
  __shared__ uint32_t s_base[kWavesPerBlock][3];
  __shared__ uint32_t s_ndef[kWavesPerBlock];
  __shared__ unsigned long long s_mdef[kWavesPerBlock];
  __shared__ WaveHubs s_hubs[kWavesPerBlock];
  if (use_choice) {
    // the network's choice, left by k_stats (-1: it deferred to the GPU rule)
    const int c = st->choice;
    if (c >= 0 && ((uint32_t)c & skip_bits)) return;  // host team or bottom-up step (uniform)
    const int tier = c > 0 ? (c >> 1) & 3 : 0;
    bins.bin_wave = tiers.wave[tier];
    bins.bin_chunk = tiers.chunk[tier];
  }
  const uint64_t si = blockIdx.x * (uint64_t)kBlock + threadIdx.x;
  const uint64_t wi = si / kSlicesPerWord;
  const uint64_t vbase = si * kSliceBits;
  uint32_t word = si < words * kSlicesPerWord ? slice_bits(in[wi], si) : 0u;
  uint32_t n_def = 0;
  unsigned long long m_def = 0;
  if (val && !hub.degree) {
    // delta-stepping: vertices at or above the threshold wait in the next
    // frontier; the four lanes of word wi own it in both bitmaps during the
    // compaction
    uint32_t far = 0;
    for (uint32_t t = word; t; t &= t - 1) {
      const int b = __builtin_ctz(t);
      if (val[vbase + b] >= threshold) far |= 1u << b;
    }
    const uint64_t far_word = join_slices(far, si);
    if (far_word && si % kSlicesPerWord == 0) carry[wi] = far_word;
    word &= ~far;
  }
  const bool any = __ballot(word != 0) != 0;  // wave-uniform
  uint32_t c_small = 0, c_med = 0, c_chunk = 0;
  uint32_t o_small = 0, o_med = 0, o_chunk = 0;
  uint32_t t_small = 0, t_med = 0, t_chunk = 0;
  if (any) {
    // Hub deferral, decided in the counting pass (the degree comes with the
    // offsets it loads): a vertex of out-degree >= hub.degree whose value is
    // >= the frontier's smallest + hub.delta waits in the next frontier
    uint32_t thr = 0xFFFFFFFFu;
    if (hub.degree) {
      const uint64_t t = (uint64_t)st->min_val + hub.delta;
      thr = t > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)t;
    }
    uint32_t wait = 0, wait_min = 0xFFFFFFFFu;
    for_each_in_slice(word, vbase, g.off, g.V, [&](uint32_t v, uint64_t deg) {
      if (deg == 0) return;
      if (deg >= hub.degree && hub.degree) {
        const uint32_t x = val[v];
        if (x >= thr) {
          wait |= 1u << (v - vbase);
          wait_min = x < wait_min ? x : wait_min;
          ++n_def;
          m_def += deg;
          return;
        }
      }
      if (deg < bins.bin_wave) c_small++;
      else if (deg < bins.bin_chunk) c_med++;
      else c_chunk += (uint32_t)((deg + bins.chunk_edges - 1) / bins.chunk_edges);
    });
    if (hub.degree) {
      const uint64_t wait_word = join_slices(wait, si);
      if (wait_word && si % kSlicesPerWord == 0) carry[wi] = wait_word;
      word &= ~wait;
      // the waiting vertices are part of the next frontier's smallest value
      if (hub.track) stripe_min(hub.track, wave_min_u32(wait_min));
    }
    o_small = wave_excl_scan(c_small, &t_small);
    o_med = wave_excl_scan(c_med, &t_med);
    o_chunk = wave_excl_scan(c_chunk, &t_chunk);
  }
  if (hub.degree) {
    // one atomic per workgroup that kept vertices waiting
    n_def = wave_sum_u32(n_def);
    m_def = wave_sum_u64(m_def);
    if (lane_id() == 0) {
      s_ndef[wave_id()] = n_def;
      s_mdef[wave_id()] = m_def;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t n = 0;
      unsigned long long m = 0;
      for (int w = 0; w < kWavesPerBlock; ++w) {
        n += s_ndef[w];
        m += s_mdef[w];
      }
      if (n) {
        atomicAdd(&st->n_def, n);
        atomicAdd(&st->m_def, m);
      }
    }
  }
  if (lane_id() == 0) {
    s_tot[wave_id()][0] = t_small;
    s_tot[wave_id()][1] = t_med;
    s_tot[wave_id()][2] = t_chunk;
    s_hubs[wave_id()].n = 0;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int q = threadIdx.x;
    uint32_t tot = 0;
    for (int w = 0; w < kWavesPerBlock; ++w) { s_base[w][q] = tot; tot += s_tot[w][q]; }
    uint32_t base = 0;
    if (tot) base = atomicAdd(q == 0 ? &st->n_small : (q == 1 ? &st->n_med : &st->n_chunk), tot);
    for (int w = 0; w < kWavesPerBlock; ++w) s_base[w][q] += base;
  }
  __syncthreads();
  WaveHubs& hubs = s_hubs[wave_id()];
  if (any) {
    uint32_t b_small = s_base[wave_id()][0] + o_small;
    uint32_t b_med = s_base[wave_id()][1] + o_med;
    uint32_t b_chunk = s_base[wave_id()][2] + o_chunk;
    for_each_in_slice(word, vbase, g.off, g.V, [&](uint32_t v, uint64_t deg) {
      if (deg == 0) return;
      if (deg < bins.bin_wave) q_small[b_small++] = v;
      else if (deg < bins.bin_chunk) q_med[b_med++] = v;
      else put_chunks(q_chunk, b_chunk, v, (uint32_t)((deg + bins.chunk_edges - 1) / bins.chunk_edges), hubs);
    });
  }
  __syncthreads();
  write_hub_chunks(q_chunk, hubs);
}

constexpr uint32_t kDenseBlock = 256;  // frontier vertices per 4096 that pay for staging offsets
constexpr uint64_t kFusedMaxBlocks = 1024;     // LDS staging: four workgroups per CU
// degree array, two passes: fewer workgroups take fewer queue atomics
// (SSSP s24 per query: 512 / 1024 / 2048 workgroups 4.28-4.30 / 4.19-4.24 /
// 4.35 ms; the one-pass form with 2048: 4.33-4.35 ms)
constexpr uint64_t kFusedMaxBlocksDeg = 1024;
// each workgroup of the fused compaction writes one statistics partial
static_assert(kFusedMaxBlocks <= kStatsMaxBlocks && kFusedMaxBlocksDeg <= kStatsMaxBlocks,
              "the partials buffer (stats_parts_needed) holds kStatsMaxBlocks + 1 entries");

// Statistics the compaction counts while it reads the frontier's offsets.
struct CompactAcc {
  uint32_t c_act = 0, c_zero = 0, n_def = 0;
  uint64_t m = 0, m_def = 0;
  Hist8 h;
};

// One compaction pass over frontier `in`, by every thread of the workgroup:
// k_compact's two passes over each 16-vertex slice (count, then place), plus
// the counts k_stats would take from the same offsets.  A grid capped at
// kFusedMaxBlocks strides over the slices (block-uniform loop; one
// queue-base atomic per queue per workgroup and stride step); every word of
// `out` is written -- zero, or the hubs kept waiting -- so no other kernel
// clears it.  A workgroup whose 4096 vertices hold at least kDenseBlock
// frontier vertices stages its offsets in LDS.  queue = false: nothing is
// queued and no hub waits (statistics only).
//
// With the graph's out-degree array (g.deg; the GPU and learned schedules)
// the workgroup makes two passes over its slices instead: it counts (the
// statistics, the hub decisions, its queue sizes), takes its three queue
// ranges with one atomic each, then places; the queue counters see one atomic
// per workgroup instead of one per workgroup and stride step, and neither
// pass has a barrier.  A lane places the items of all its slices
// contiguously.  The second pass decides the hubs again (no value changes
// during the compaction, so the decisions agree).  A lane with more than
// kFewBits frontier vertices in a slice loads the slice's 16 degrees with
// four 16-byte loads into its own LDS row; one with fewer loads each
// vertex's degree.
//
// With g.xval (SSSP / CC on the fused pipeline) the counting pass also drops
// every frontier vertex whose value is one its out-edges were all relaxed
// with (the expansion that read the value has finished, so each neighbour
// has seen that candidate already): it is not counted, queued or kept
// waiting, and the placing pass skips it (its bit is cleared in the cached
// word, or in `in` when the workgroup has more strides than it caches).  A
// vertex queued in chunks has its record cleared here; its chunks raise it
// to the largest value they read.
constexpr int kFewBits = 4;
constexpr int kPre = 4;
constexpr int kDegRow = (int)kSliceBits + 1;  // odd stride: lanes' rows on distinct banks
__device__ __forceinline__ void compact_frontier_deg(uint64_t* in, uint64_t* __restrict__ out,
                                                     uint64_t words, const DevGraph& g, const Bins& bins,
                                                     uint32_t* __restrict__ q_small, uint32_t* __restrict__ q_med,
                                                     uint64_t* __restrict__ q_chunk, DevStats* __restrict__ st,
                                                     const uint32_t* __restrict__ val, const HubDefer& hub,
                                                     bool queue, CompactAcc& a) {
  __shared__ uint32_t s_tot[kWavesPerBlock][3];
  __shared__ uint32_t s_base[kWavesPerBlock][3];
  __shared__ uint32_t s_deg[kWavesPerBlock][kWave * kDegRow];
  __shared__ WaveHubs s_hubs[kWavesPerBlock];
  WaveHubs& hubs = s_hubs[wave_id()];
  uint32_t* const drow = &s_deg[wave_id()][lane_id() * kDegRow];
  uint32_t thr = 0xFFFFFFFFu;
  if (hub.degree) {
    const uint64_t t = (uint64_t)st->min_val + hub.delta;
    thr = t > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)t;
  }
  const uint64_t slices = words * kSlicesPerWord;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  // calls f(v, deg) for the vertices of the lane's slice si in `bits`
  auto each = [&](uint64_t si, uint32_t bits, auto&& f) {
    const uint64_t vbase = si * kSliceBits;
    if (__popc(bits) > kFewBits) {
      const uint4* p = reinterpret_cast<const uint4*>(g.deg + vbase);
      const uint4 d0 = p[0], d1 = p[1], d2 = p[2], d3 = p[3];
      const uint32_t d[16] = {d0.x, d0.y, d0.z, d0.w, d1.x, d1.y, d1.z, d1.w,
                              d2.x, d2.y, d2.z, d2.w, d3.x, d3.y, d3.z, d3.w};
#pragma unroll
      for (int k = 0; k < 16; ++k) drow[k] = d[k];
      for (; bits; bits &= bits - 1) {
        const uint32_t j = (uint32_t)__builtin_ctz(bits);
        f((uint32_t)(vbase + j), (uint64_t)drow[j]);
      }
    } else {
      for (; bits; bits &= bits - 1) {
        const uint32_t v = (uint32_t)(vbase + (uint32_t)__builtin_ctz(bits));
        f(v, (uint64_t)g.deg[v]);
      }
    }
  };
  auto waits = [&](uint32_t v, uint64_t deg) -> bool {
    return hub.degree && deg >= hub.degree && val[v] >= thr;
  };
  const bool drops = queue && val && g.xval;  // kernel-uniform
  // The workgroup's slices, kPre strides at a time: their frontier words are
  // loaded together (kPre loads in flight per lane instead of one), and when
  // the workgroup has at most kPre strides the placing pass takes the words
  // from registers instead of reading them again.
  const uint64_t base0 = blockIdx.x * (uint64_t)kBlock;
  const uint64_t nstr = base0 < slices ? (slices - base0 + stride - 1) / stride : 0;  // block-uniform
  const bool cached = nstr <= (uint64_t)kPre;
  auto load_words = [&](uint64_t s0, uint64_t (&w)[kPre]) {
#pragma unroll
    for (int j = 0; j < kPre; ++j) {
      const uint64_t si = base0 + (s0 + j) * stride + threadIdx.x;
      w[j] = s0 + j < nstr && si < slices ? in[si / kSlicesPerWord] : 0ull;
    }
  };
  uint64_t keep[kPre] = {};  // the cached words, dropped vertices removed
  // pass 1: statistics, hubs kept waiting (written to the next bitmap, whose
  // every word this pass writes), queue sizes
  uint32_t c_small = 0, c_med = 0, c_chunk = 0, wait_min = 0xFFFFFFFFu;
  for (uint64_t s0 = 0; s0 < nstr; s0 += kPre) {
    uint64_t fulls[kPre];
    load_words(s0, fulls);
#pragma unroll
    for (int j = 0; j < kPre; ++j) {
      if (s0 + j >= nstr) break;  // block-uniform
      const uint64_t si = base0 + (s0 + j) * stride + threadIdx.x;
      const uint64_t wi = si / kSlicesPerWord;
      const uint32_t word = slice_bits(fulls[j], si);
      uint32_t wait = 0, drop = 0;
      // a dense slice wholly below V takes the drop test of its 16 vertices
      // from four 16-byte loads of the values and four of the propagated
      // values instead of two scalar loads per active vertex
      const bool vec_drop = drops && __popc(word) > kFewBits && (si + 1) * kSliceBits <= (uint64_t)g.V;
      uint32_t dmask = 0;
      if (vec_drop) {
        const uint4* pv = reinterpret_cast<const uint4*>(val + si * kSliceBits);
        const uint4* px = reinterpret_cast<const uint4*>(g.xval + si * kSliceBits);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint4 x = pv[q], r = px[q];
          dmask |= ((r.x != 0u && r.x - 1u == x.x) ? 1u : 0u) << (4 * q);
          dmask |= ((r.y != 0u && r.y - 1u == x.y) ? 2u : 0u) << (4 * q);
          dmask |= ((r.z != 0u && r.z - 1u == x.z) ? 4u : 0u) << (4 * q);
          dmask |= ((r.w != 0u && r.w - 1u == x.w) ? 8u : 0u) << (4 * q);
        }
      }
      each(si, word, [&](uint32_t v, uint64_t deg) {
        uint32_t x = 0;
        bool have_x = false;
        if (drops && deg != 0) {
          const uint32_t bit = 1u << (v - si * kSliceBits);
          if (vec_drop) {
            if (dmask & bit) {
              drop |= bit;
              return;
            }
          } else {
            x = val[v];
            have_x = true;
            const uint32_t xr = g.xval[v];
            if (xr != 0u && xr - 1u == x) {
              drop |= bit;
              return;
            }
          }
        }
        a.m += deg;
        a.h.add(hist_bin(deg));
        if (deg == 0) {
          ++a.c_zero;
          return;
        }
        if (hub.degree && deg >= hub.degree && (have_x ? x : (x = val[v])) >= thr) {
          wait |= 1u << (v - si * kSliceBits);
          wait_min = x < wait_min ? x : wait_min;
          ++a.n_def;
          a.m_def += deg;
          return;
        }
        if (deg < bins.bin_wave) c_small++;
        else if (deg < bins.bin_chunk) c_med++;
        else c_chunk += (uint32_t)((deg + bins.chunk_edges - 1) / bins.chunk_edges);
      });
      a.c_act += (uint32_t)__popc(word & ~drop);
      const uint64_t wait_word = join_slices(wait, si);
      if (si < slices && si % kSlicesPerWord == 0) out[wi] = wait_word;
      uint64_t kept = fulls[j];
      if (drops) {
        kept &= ~join_slices(drop, si);
        // read back by the placing pass after the workgroup's barriers
        if (!cached && kept != fulls[j] && si % kSlicesPerWord == 0) in[wi] = kept;
      }
      keep[j] = kept;
    }
  }
  // the waiting vertices are part of the next frontier's smallest value
  if (hub.track) stripe_min(hub.track, wave_min_u32(wait_min));
  if (!queue) c_small = c_med = c_chunk = 0;
  uint32_t t_small, t_med, t_chunk;
  const uint32_t o_small = wave_excl_scan(c_small, &t_small);
  const uint32_t o_med = wave_excl_scan(c_med, &t_med);
  const uint32_t o_chunk = wave_excl_scan(c_chunk, &t_chunk);
  if (lane_id() == 0) {
    s_tot[wave_id()][0] = t_small;
    s_tot[wave_id()][1] = t_med;
    s_tot[wave_id()][2] = t_chunk;
    hubs.n = 0;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int q = threadIdx.x;
    uint32_t tot = 0;
    for (int w = 0; w < kWavesPerBlock; ++w) { s_base[w][q] = tot; tot += s_tot[w][q]; }
    uint32_t b = 0;
    if (tot) b = atomicAdd(q == 0 ? &st->n_small : (q == 1 ? &st->n_med : &st->n_chunk), tot);
    for (int w = 0; w < kWavesPerBlock; ++w) s_base[w][q] += b;
  }
  __syncthreads();
  uint32_t any = 0;
#pragma unroll
  for (int w = 0; w < kWavesPerBlock; ++w) any |= s_tot[w][0] | s_tot[w][1] | s_tot[w][2];
  if (!any) return;  // uniform: nothing to place
  // pass 2: the same slices, placed
  uint32_t b_small = s_base[wave_id()][0] + o_small;
  uint32_t b_med = s_base[wave_id()][1] + o_med;
  uint32_t b_chunk = s_base[wave_id()][2] + o_chunk;
  for (uint64_t s0 = 0; s0 < nstr; s0 += kPre) {
    uint64_t fulls[kPre];
    if (cached) {
#pragma unroll
      for (int j = 0; j < kPre; ++j) fulls[j] = keep[j];
    } else {
      load_words(s0, fulls);
    }
#pragma unroll
    for (int j = 0; j < kPre; ++j) {
      if (s0 + j >= nstr) break;  // block-uniform
      const uint64_t si = base0 + (s0 + j) * stride + threadIdx.x;
      each(si, slice_bits(fulls[j], si), [&](uint32_t v, uint64_t deg) {
        if (deg == 0 || waits(v, deg)) return;
        if (deg < bins.bin_wave) q_small[b_small++] = v;
        else if (deg < bins.bin_chunk) q_med[b_med++] = v;
        else {
          if (drops) g.xval[v] = 0u;
          put_chunks(q_chunk, b_chunk, v, (uint32_t)((deg + bins.chunk_edges - 1) / bins.chunk_edges), hubs);
        }
      });
    }
  }
  write_hub_chunks(q_chunk, hubs);  // the wavefront's own list
}

template <bool DEG>
__device__ __forceinline__ void compact_frontier(uint64_t* in, uint64_t* __restrict__ out,
                                                 uint64_t words, const DevGraph& g, const Bins& bins,
                                                 uint32_t* __restrict__ q_small, uint32_t* __restrict__ q_med,
                                                 uint64_t* __restrict__ q_chunk, DevStats* __restrict__ st,
                                                 const uint32_t* __restrict__ val, const HubDefer& hub, bool queue,
                                                 CompactAcc& a) {
  if constexpr (DEG) {
    compact_frontier_deg(in, out, words, g, bins, q_small, q_med, q_chunk, st, val, hub, queue, a);
    return;
  }
  __shared__ uint32_t s_tot[kWavesPerBlock][3];
  __shared__ uint32_t s_base[kWavesPerBlock][3];
  __shared__ uint32_t s_cnt[kWavesPerBlock];
  __shared__ uint64_t s_off[DEG ? 1 : kWavesPerBlock][DEG ? 1 : kStageEntries];
  __shared__ WaveHubs s_hubs[kWavesPerBlock];
  WaveHubs& hubs = s_hubs[wave_id()];
  uint64_t* const row = s_off[DEG ? 0 : wave_id()];
  // hub threshold: the frontier's smallest value (left in st by the previous
  // iteration's finish) + delta
  uint32_t thr = 0xFFFFFFFFu;
  if (hub.degree) {
    const uint64_t t = (uint64_t)st->min_val + hub.delta;
    thr = t > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)t;
  }
  const uint64_t slices = words * kSlicesPerWord;
  // block-uniform loop: every thread runs the same number of steps
  for (uint64_t base = blockIdx.x * (uint64_t)kBlock; base < slices; base += (uint64_t)gridDim.x * kBlock) {
    const uint64_t si = base + threadIdx.x;
    const uint64_t wi = si / kSlicesPerWord;
    const uint64_t vbase = si * kSliceBits;
    uint32_t word = si < slices ? slice_bits(in[wi], si) : 0u;
    a.c_act += (uint32_t)__popc(word);
    const bool any = __ballot(word != 0) != 0;  // wave-uniform
    bool dense = false;
    {
      // block-uniform choice of the offsets' source
      const uint32_t wcnt = wave_sum_u32((uint32_t)__popc(word));
      if (lane_id() == 0) s_cnt[wave_id()] = wcnt;
      __syncthreads();
      uint32_t bcnt = 0;
#pragma unroll
      for (int w = 0; w < kWavesPerBlock; ++w) bcnt += s_cnt[w];
      dense = bcnt >= kDenseBlock;
      if (dense) {
        stage_offsets(g.off, g.V, (base + (uint64_t)wave_id() * kWave) * kSliceBits, row);
        __syncthreads();
      }
    }
    // sparse workgroups: one offset pair per frontier vertex
    auto slice_each = [&](uint32_t bits, auto&& f) {
      if (dense) {
        for_each_in_slice_staged(bits, vbase, row, f);
      } else {
        for (; bits; bits &= bits - 1) {
          const uint32_t v = (uint32_t)(vbase + (uint32_t)__builtin_ctz(bits));
          f(v, g.off[v + 1] - g.off[v]);
        }
      }
    };
    uint32_t o_small = 0, o_med = 0, o_chunk = 0;
    uint32_t t_small = 0, t_med = 0, t_chunk = 0;
    uint64_t wait_word = 0;
    if (any) {
      uint32_t c_small = 0, c_med = 0, c_chunk = 0;
      uint32_t wait = 0, wait_min = 0xFFFFFFFFu;
      slice_each(word, [&](uint32_t v, uint64_t deg) {
        a.m += deg;
        a.h.add(hist_bin(deg));
        if (deg == 0) {
          ++a.c_zero;
          return;
        }
        if (hub.degree && deg >= hub.degree) {
          const uint32_t x = val[v];
          if (x >= thr) {
            wait |= 1u << (v - vbase);
            wait_min = x < wait_min ? x : wait_min;
            ++a.n_def;
            a.m_def += deg;
            return;
          }
        }
        if (deg < bins.bin_wave) c_small++;
        else if (deg < bins.bin_chunk) c_med++;
        else c_chunk += (uint32_t)((deg + bins.chunk_edges - 1) / bins.chunk_edges);
      });
      if (hub.degree) {
        wait_word = join_slices(wait, si);
        word &= ~wait;
        // the waiting vertices are part of the next frontier's smallest value
        if (hub.track) stripe_min(hub.track, wave_min_u32(wait_min));
      }
      if (!queue) {
        word = 0;
        c_small = c_med = c_chunk = 0;
      }
      o_small = wave_excl_scan(c_small, &t_small);
      o_med = wave_excl_scan(c_med, &t_med);
      o_chunk = wave_excl_scan(c_chunk, &t_chunk);
    }
    if (si < slices && si % kSlicesPerWord == 0) out[wi] = wait_word;
    if (lane_id() == 0) {
      s_tot[wave_id()][0] = t_small;
      s_tot[wave_id()][1] = t_med;
      s_tot[wave_id()][2] = t_chunk;
      hubs.n = 0;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
      const int q = threadIdx.x;
      uint32_t tot = 0;
      for (int w = 0; w < kWavesPerBlock; ++w) { s_base[w][q] = tot; tot += s_tot[w][q]; }
      uint32_t b = 0;
      if (tot) b = atomicAdd(q == 0 ? &st->n_small : (q == 1 ? &st->n_med : &st->n_chunk), tot);
      for (int w = 0; w < kWavesPerBlock; ++w) s_base[w][q] += b;
    }
    __syncthreads();
    if (any) {
      uint32_t b_small = s_base[wave_id()][0] + o_small;
      uint32_t b_med = s_base[wave_id()][1] + o_med;
      uint32_t b_chunk = s_base[wave_id()][2] + o_chunk;
      slice_each(word, [&](uint32_t v, uint64_t deg) {
        if (deg == 0) return;
        if (deg < bins.bin_wave) q_small[b_small++] = v;
        else if (deg < bins.bin_chunk) q_med[b_med++] = v;
        else put_chunks(q_chunk, b_chunk, v, (uint32_t)((deg + bins.chunk_edges - 1) / bins.chunk_edges), hubs);
      });
    }
    __syncthreads();  // s_tot / s_base / s_cnt / s_off are reused by the next step
    write_hub_chunks(q_chunk, hubs);  // the list is the wavefront's own until its next reset
  }
}

// Compaction with the frontier statistics folded in (the GPU and learned
// schedules' SSSP / CC iterations), so the frontier and its offsets are read
// once per iteration instead of twice.
template <bool DEG>
__global__ __launch_bounds__(kBlock) void k_compact_fused(uint64_t* __restrict__ in,
                                                          uint64_t* __restrict__ out, uint64_t words, DevGraph g,
                                                          Bins bins, uint32_t* __restrict__ q_small,
                                                          uint32_t* __restrict__ q_med,
                                                          uint64_t* __restrict__ q_chunk,
                                                          DevStats* __restrict__ st,
                                                          const uint32_t* __restrict__ val, HubDefer hub,
                                                          StatsPart* __restrict__ parts, int use_choice,
                                                          TierTable tiers, const RunState* __restrict__ rs) {
  // an empty frontier was counted: this iteration is queued behind the
  // traversal's end and has nothing to do
  if (rs->done) return;
  // learned schedule: the choice the previous iteration's finish left in st
  // (-1: the network deferred; the tier table's entry 0 applies).  A
  // host-team choice queues nothing and keeps no hub waiting: the statistics
  // are still counted and the next bitmap cleared, and the host relaxes the
  // frontier after reading the record.
  bool queue = true;
  if (use_choice) {
    const int c = st->choice;
    const int tier = c > 0 ? (c >> 1) & 3 : 0;
    bins.bin_wave = tiers.wave[tier];
    bins.bin_chunk = tiers.chunk[tier];
    if (c >= 0 && (c & 8)) {
      queue = false;
      hub.degree = 0;
      hub.track = nullptr;
    }
  }
  CompactAcc a;
  compact_frontier<DEG>(in, out, words, g, bins, q_small, q_med, q_chunk, st, val, hub, queue, a);
  StatsPart p{};
  p.n_active = wave_sum_u32(a.c_act);
  p.m_active = wave_sum_u64(a.m);
  p.n_zero = wave_sum_u32(a.c_zero);
#pragma unroll
  for (int k = 0; k < 8; ++k) p.hist[k] = wave_sum_u32(a.h.get(k));
  p.min_val = 0xFFFFFFFFu;
  p.n_def = wave_sum_u32(a.n_def);
  p.m_def = wave_sum_u64(a.m_def);
  stats_block_part(p, parts);
}

// One workgroup behind the expansion of a fused iteration: fixed-order
// reduction of the compaction's partials, the record of the iteration
// (published to mapped host memory with `iter` last), the first-visit
// stripes of the expansion folded into the run state, and the next slot
// prepared for the next compaction.
__global__ __launch_bounds__(kFinishBlock) void k_fused_finish(const StatsPart* __restrict__ parts,
                                                               uint32_t nparts, uint32_t* __restrict__ stripes,
                                                               RunState* __restrict__ rs, DevStats* __restrict__ st,
                                                               DevStats* __restrict__ next,
                                                               DevStats* __restrict__ host_out, int track,
                                                               const float* __restrict__ mlp, uint32_t V, uint64_t E) {
  constexpr int kC = 11;  // n_active, n_zero, hist[8], n_def
  constexpr int kFW = kFinishBlock / kWave;
  __shared__ unsigned long long s_m[kFW], s_md[kFW];
  __shared__ uint32_t s_c[kFW][kC];
  __shared__ float w_mlp[kMlpLds];
  // what the record needs besides the partials -- the run state, this
  // slot's queue counts and minimum, both stripe sets -- is loaded by lanes
  // of wavefront 1 together with the partials, so the kernel waits on one
  // round of loads instead of a chain of them
  __shared__ RunState s_rs;
  __shared__ uint32_t s_cs[kStripes], s_ms[kStripes], s_q[4];
  {
    const uint32_t t = threadIdx.x;
    if (t >= 64 && t < 64 + kStripes) {
      s_cs[t - 64] = stripes[(t - 64) * kStripeStride];
    } else if (t >= 96 && t < 96 + kStripes) {
      s_ms[t - 96] = track ? stripes[kStripeMinBase + (t - 96) * kStripeStride] : 0xFFFFFFFFu;
    } else if (t == 128) {
      s_rs = *rs;
    } else if (t == 129) {
      s_q[0] = st->n_small;
      s_q[1] = st->n_med;
      s_q[2] = st->n_chunk;
      s_q[3] = st->min_val;
    }
  }
  // the network's weights are loaded now and stored to LDS after the
  // reduction, whose latency hides theirs
  MlpStage<kFinishBlock> wst;
  if (mlp) wst.load(mlp);
  // kFinishParts partials per thread, loaded together (a further round only
  // past kFinishBlock * kFinishParts partials); wavefront sums, then the
  // wavefronts' sums in order
  unsigned long long m = 0, md = 0;
  uint32_t c[kC] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (uint32_t b0 = 0; b0 < nparts; b0 += kFinishBlock * kFinishParts) {
    StatsPart q[kFinishParts];
#pragma unroll
    for (int j = 0; j < kFinishParts; ++j) {
      const uint32_t b = b0 + j * kFinishBlock + threadIdx.x;
      if (b < nparts) q[j] = parts[b];
      else q[j] = StatsPart{};
    }
#pragma unroll
    for (int j = 0; j < kFinishParts; ++j) {
      m += q[j].m_active;
      md += q[j].m_def;
      c[0] += q[j].n_active;
      c[1] += q[j].n_zero;
      for (int k = 0; k < 8; ++k) c[2 + k] += q[j].hist[k];
      c[10] += q[j].n_def;
    }
  }
  m = wave_sum_u64(m);
  md = wave_sum_u64(md);
#pragma unroll
  for (int k = 0; k < kC; ++k) c[k] = wave_sum_u32(c[k]);
  if (lane_id() == 0) {
    s_m[wave_id()] = m;
    s_md[wave_id()] = md;
#pragma unroll
    for (int k = 0; k < kC; ++k) s_c[wave_id()][k] = c[k];
  }
  __syncthreads();
  if (s_rs.done) return;  // queued behind the traversal's end (uniform)
  if (threadIdx.x < kWave) {
    // wavefront 0 sums the wavefronts' sums, one per lane
    const uint32_t l = threadIdx.x;
    m = wave_sum_u64(l < kFW ? s_m[l] : 0ull);
    md = wave_sum_u64(l < kFW ? s_md[l] : 0ull);
#pragma unroll
    for (int k = 0; k < kC; ++k) c[k] = wave_sum_u32(l < kFW ? s_c[l][k] : 0u);
  }
  __shared__ float feat[CG_NUM_FEATURES];
  __shared__ float mlp_sh[2 * kMlpH + kMlpOut];
  __shared__ float probs[kMlpOut];
  __shared__ int best;
  __shared__ int s_net;
  DevStats d{};
  if (threadIdx.x == 0) {
  uint32_t nn = 0;
#pragma unroll
  for (int k = 0; k < kStripes; ++k) nn += s_cs[k];
#pragma unroll
  for (int k = 0; k < kStripes; ++k) stripes[k * kStripeStride] = 0;
  d.n_small = s_q[0];
  d.n_med = s_q[1];
  d.n_chunk = s_q[2];
  d.min_val = s_q[3];
  d.n_active = c[0];
  d.n_zero_deg = c[1];
  d.m_active = m;
  for (int k = 0; k < 8; ++k) d.hist[k] = c[2 + k];
  d.n_def = c[10];
  d.m_def = md;
  // the record describes the frontier this iteration expanded: vertices
  // reached before it, and the first visits of the previous expansion
  d.visited = s_rs.visited;
  d.n_new = s_rs.pad;
  d.iter = s_rs.iter + 1;
  d.choice = -1;
  rs->visited = d.visited + nn;
  rs->pad = nn;
  rs->iter = d.iter;
  rs->done = d.n_active == 0 ? 1u : 0u;
  uint32_t mv = 0;
  if (track) {
    // the next frontier's smallest value: smallest activating candidate of
    // this expansion and smallest waiting hub of this compaction
    mv = 0xFFFFFFFFu;
#pragma unroll
    for (int k = 0; k < kStripes; ++k) {
      const uint32_t x = s_ms[k];
      mv = x < mv ? x : mv;
      stripes[kStripeMinBase + k * kStripeStride] = 0xFFFFFFFFu;
    }
  }
  next->n_small = 0;
  next->n_med = 0;
  next->n_chunk = 0;
  next->min_val = mv;
  make_features(d, V, E, d.visited, d.iter, feat);
  // the network runs on frontiers holding at least 1/kNetEdgeShare of the
  // edges; after a smaller one the next iteration takes the GPU rule (its
  // choices there are within noise of the tuned tiers, and the forward pass
  // would cost more than it could save)
  s_net = mlp && d.m_active * kNetEdgeShare >= E ? 1 : 0;
  }
  __syncthreads();
  const bool net = s_net != 0;
  if (net) wst.store(w_mlp);
  __syncthreads();
  if (net) {
    // the learned schedule of the next iteration, from this frontier's
    // features; the network defers (-1) unless its output is a finite
    // distribution whose top choice has probability >= kLearnedMinProb
    mlp_forward_lds(w_mlp, feat, mlp_sh, probs, &best);
  }
  if (threadIdx.x != 0) return;
  int choice = -1;
  if (net) {
    float sum = 0.f;
    bool finite = true;
    for (int k = 0; k < kMlpOut; ++k) {
      finite = finite && isfinite(probs[k]);
      sum += probs[k];
      d.probs[k] = probs[k];
    }
    if (finite && fabsf(sum - 1.f) < 1e-3f && probs[best] >= kLearnedMinProb) choice = best;
  }
  d.choice = choice;
  next->choice = choice;
  *st = d;
  if (host_out) {
    const uint32_t itv = record_tag(s_rs.epoch, d.iter);
    d.iter = 0xFFFFFFFFu;
    *host_out = d;
    // the release store orders the record before `iter` at system scope
    __hip_atomic_store(&host_out->iter, itv, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// The push expansion's work items (the queues of a compaction), strided over
// the wavefronts of the grid: chunks first (the heaviest), then medium
// vertices, then groups of 64 small ones.  `range` is the part of the small
// tier these items cover (the rest is the host team's, or another launch's)
// and whether the chunked and medium tiers come with it.  Returns the
// wavefront-lane's first-visit count; activations are folded into acc.
// The weight array of the width the expansion was instantiated for: the
// graph's own (uint32), or its narrow copy when every weight fits 8 or 16
// bits (the same values; fewer bytes streamed per edge).
template <class WT>
__device__ __forceinline__ const WT* graph_weights(const DevGraph& g) {
  if constexpr (sizeof(WT) == 1) return g.w8;
  else if constexpr (sizeof(WT) == 2) return g.w16;
  else return g.w;
}

template <int ALG, bool EXACT, bool TRACK, class A, class WT = uint32_t>
__device__ __forceinline__ uint32_t expand_items(const DevGraph& g, uint32_t* __restrict__ val,
                                                 uint64_t* __restrict__ out, const uint32_t* __restrict__ q_small,
                                                 const uint32_t* __restrict__ q_med,
                                                 const uint64_t* __restrict__ q_chunk, SmallRange range,
                                                 uint32_t chunk_edges, const DevStats* __restrict__ st, A& acc) {
  // The queue lengths come from the compaction; the wavefronts of a grid
  // capped near one chip's worth of resident workgroups stride over the items,
  // so no wavefront is launched only to find nothing to do.
  const uint32_t n_chunk = range.big ? st->n_chunk : 0u, n_med = range.big ? st->n_med : 0u;
  const WT* const w = graph_weights<WT>(g);
  const uint32_t s_end = st->n_small < range.end ? st->n_small : range.end;
  const uint32_t s_beg = range.begin < s_end ? range.begin : s_end;
  const uint32_t n_small = s_end - s_beg;
  const uint64_t n_items = (uint64_t)n_chunk + n_med + (n_small + kWave - 1) / kWave;
  const uint32_t lane = lane_id();
  const uint64_t stride = (uint64_t)gridDim.x * kWavesPerBlock;
  // the value each source's edges are relaxed with is recorded (plus one) for
  // the next compaction; a vertex in chunks keeps the largest its chunks read
  uint32_t* const xval = ALG == ALG_BFS ? nullptr : g.xval;
  uint32_t n_new = 0;
  for (uint64_t item = (uint64_t)blockIdx.x * kWavesPerBlock + wave_id(); item < n_items; item += stride) {
    if (item < n_chunk) {
      // a slice of a high-degree vertex: lanes stride its edges
      const uint64_t pk = q_chunk[item];
      const uint32_t v = (uint32_t)(pk >> 32);
      const uint64_t s = g.off[v] + (uint64_t)(uint32_t)pk * chunk_edges;
      const uint64_t end = g.off[v + 1];
      const uint64_t e = s + chunk_edges < end ? s + chunk_edges : end;
      const uint32_t sv = val[v];
      if (xval && lane == 0) atomicMax(xval + v, sv + 1u);
      n_new += relax_strided<ALG, EXACT, TRACK>(sv, s, e, lane, g.dst, w, val, out, acc);
    } else if (item < (uint64_t)n_chunk + n_med) {
      const uint32_t v = q_med[item - n_chunk];
      const uint32_t sv = val[v];
      if (xval && lane == 0) xval[v] = sv + 1u;
      n_new += relax_strided<ALG, EXACT, TRACK>(sv, g.off[v], g.off[v + 1], lane, g.dst, w, val, out, acc);
    } else {
      // 64 low-degree vertices per wavefront, their edges spread evenly over
      // the lanes: lane j of round r takes edge r*64+j of the concatenated
      // adjacency lists and finds its owner by binary search over the
      // exclusive degree prefix held one entry per lane.
      const uint64_t base = s_beg + (item - n_chunk - n_med) * kWave;
      const uint32_t cnt = (s_end - base) < (uint64_t)kWave ? (uint32_t)(s_end - base) : kWave;
      uint32_t deg = 0, sv = 0;
      uint64_t s = 0;
      if (lane < cnt) {
        const uint32_t v = q_small[base + lane];
        s = g.off[v];
        deg = (uint32_t)(g.off[v + 1] - s);
        sv = val[v];
        if (xval) xval[v] = sv + 1u;
      }
      uint32_t total;
      const uint32_t excl = wave_excl_scan(deg, &total);
      // edge j's owner, found by binary search over the exclusive prefix
      auto owner = [&](uint32_t j, uint64_t& e, uint32_t& osv) {
        uint32_t lo = 0;
#pragma unroll
        for (uint32_t step = kWave / 2; step >= 1; step >>= 1) {
          const uint32_t pm = __shfl(excl, (int)(lo + step), kWave);
          if (pm <= j) lo += step;
        }
        const uint64_t os = __shfl(s, (int)lo, kWave);
        osv = __shfl(sv, (int)lo, kWave);
        const uint32_t oex = __shfl(excl, (int)lo, kWave);
        e = os + (j - oex);
      };
      // SSSP takes two rounds of 64 edges at a time: both edges' destination,
      // weight and destination value are loaded before either is committed
      // (measured neutral to slower for the two-load relaxations)
      uint32_t j0 = 0;
      if constexpr (ALG == ALG_SSSP) {
        for (; j0 + kWave < total; j0 += 2 * kWave) {
          const uint32_t jb = j0 + kWave + lane;  // j0 + lane < total here
          uint64_t ea, eb;
          uint32_t sva, svb;
          owner(j0 + lane, ea, sva);
          owner(jb, eb, svb);
          const bool okb = jb < total;
          const uint32_t da = g.dst[ea];
          const uint32_t db = okb ? g.dst[eb] : 0u;
          const uint32_t ca = candidate<ALG>(sva, ea, w);
          const uint32_t cb = okb ? candidate<ALG>(svb, eb, w) : 0u;
          const uint32_t xa = val[da];
          const uint32_t xb = okb ? val[db] : 0u;
          n_new += commit<ALG, EXACT, TRACK>(ca, xa, da, val, out, acc);
          if (okb) n_new += commit<ALG, EXACT, TRACK>(cb, xb, db, val, out, acc);
        }
      }
      for (; j0 < total; j0 += kWave) {
        const uint32_t j = j0 + lane;
        uint64_t e;
        uint32_t osv;
        owner(j, e, osv);
        if (j < total) n_new += relax<ALG, EXACT, TRACK>(osv, e, g.dst, w, val, out, acc);
      }
    }
  }
  return n_new;
}

template <int ALG, bool EXACT, bool TRACK, class WT = uint32_t>
__global__ __launch_bounds__(kBlock) void k_expand(DevGraph g, uint32_t* __restrict__ val,
                                                   uint64_t* __restrict__ out,
                                                   const uint32_t* __restrict__ q_small,
                                                   const uint32_t* __restrict__ q_med,
                                                   const uint64_t* __restrict__ q_chunk,
                                                   SmallRange range, uint32_t chunk_edges,
                                                   const DevStats* __restrict__ st,
                                                   uint32_t* __restrict__ stripes) {
  MinTrack mt;  // smallest activating candidate (TRACK)
  const uint32_t n_new =
      expand_items<ALG, EXACT, TRACK, MinTrack, WT>(g, val, out, q_small, q_med, q_chunk, range, chunk_edges, st, mt);
  if constexpr (ALG != ALG_CC) {
    const uint32_t tot = wave_sum_u32(n_new);
    if (lane_id() == 0) stripe_add(stripes, tot);
  }
  if constexpr (TRACK) stripe_min(stripes, wave_min_u32(mt.mn));
}

// Bottom-up BFS (Beamer, Asanovic, Patterson, SC'12): every unreached vertex
// scans its in-edges for a parent in the current frontier.  One lane per
// vertex, so a wavefront owns one 64-bit word of the next frontier and writes
// it whole (no atomics).
__global__ __launch_bounds__(kBlock) void k_pull_bfs(DevGraph g, uint32_t* __restrict__ val,
                                                     const uint64_t* __restrict__ in,
                                                     uint64_t* __restrict__ out,
                                                     uint32_t* __restrict__ stripes,
                                                     const DevStats* __restrict__ cond) {
  if (cond) {
    const int c = cond->choice;
    if (!(c >= 0 && (c & 1))) return;  // this iteration pushes (uniform)
  }
  const uint64_t words = ((uint64_t)g.V + 63) / 64;
  const uint64_t nwaves = (uint64_t)gridDim.x * kWavesPerBlock;
  uint32_t n_new = 0;
  for (uint64_t w = (uint64_t)blockIdx.x * kWavesPerBlock + wave_id(); w < words; w += nwaves) {
    const uint64_t v = w * 64 + lane_id();
    bool found = false;
    if (v < g.V && val[v] == kUnreached) {
      const uint64_t s = g.in_off[v], e = g.in_off[v + 1];
      for (uint64_t k = s; k < e; ++k) {
        const uint32_t u = g.in_src[k];
        if ((in[u >> 6] >> (u & 63)) & 1ull) {
          val[v] = val[u] + 1;
          found = true;
          break;
        }
      }
    }
    const uint64_t mask = __ballot(found);
    if (lane_id() == 0) out[w] = mask;
    n_new += (uint32_t)__popcll(mask);
  }
  if (lane_id() == 0) stripe_add(stripes, n_new);
}

// BFS under the GPU schedule in two kernels per iteration.  The direction of
// iteration k (Beamer's rule) needs the size and out-edges of frontier k; the
// step kernel of iteration k - 1 counts exactly those where it makes the
// frontier (first visits: the bits it set first in the next bitmap, with
// their out-degrees), so no statistics pass over the frontier is needed.
//
// k_bfs_front: every workgroup takes the decision from the previous step's
// counts and run state (the same inputs give the same decision); workgroup 0
// stores it, the next run state and the iteration's record (published to
// mapped host memory, `iter` last) and clears the counters this step fills
// and the next slot's queue counters.  A push iteration then compacts the
// frontier into the degree-tier queues, writing every word of the next bitmap
// (zero); a bottom-up iteration leaves the next bitmap to the step, which
// writes it whole.
// k_bfs_step: push: the tiers of k_expand with first visits counted;
// bottom-up: lane per vertex, an unreached vertex looks for a parent in the
// frontier among its in-edges, a wavefront writes its word of the next
// bitmap and counts the vertices found and their out-degrees.
// Iteration one's front work (frontier {root}) is done by k_init_traverse.
template <bool DEG>
__global__ __launch_bounds__(kBlock) void k_bfs_front(const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                                                      uint64_t words, DevGraph g, uint64_t E, Bins bins,
                                                      uint32_t* __restrict__ q_small, uint32_t* __restrict__ q_med,
                                                      uint64_t* __restrict__ q_chunk, DevStats* __restrict__ st,
                                                      DevStats* __restrict__ st_next,
                                                      const unsigned long long* __restrict__ cnt_prev,
                                                      unsigned long long* __restrict__ cnt_cur,
                                                      const RunState* __restrict__ rs_prev,
                                                      RunState* __restrict__ rs_cur, DevStats* __restrict__ host_out) {
  unsigned long long n = 0, m = 0, z = 0;
#pragma unroll
  for (int k = 0; k < kStripes; ++k) {
    n += cnt_prev[k * kBfsCntStride];
    m += cnt_prev[k * kBfsCntStride + 1];
    z += cnt_prev[k * kBfsCntStride + 2];
  }
  const RunState r = *rs_prev;
  // Beamer, Asanovic, Patterson (SC'12): bottom-up when the frontier's edges
  // exceed the unexplored edges / 15, back to push when the frontier falls
  // below V / 18 and shrinks
  const unsigned long long m_unexplored = E > r.m_explored ? E - r.m_explored : 0ull;
  bool pull;
  if (!r.pull) pull = 15ull * m > m_unexplored;
  else pull = !(18ull * n < (unsigned long long)g.V && n < (unsigned long long)r.prev_active);
  // chunk size of this iteration's push queues (the step kernel reads it back)
  uint32_t ce = bins.chunk_edges;
  while (ce > (uint32_t)kWave && m < (unsigned long long)ce * kChunkSpread) ce >>= 1;
  bins.chunk_edges = ce;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    RunState w = r;
    w.pull = pull ? 1u : 0u;
    w.m_explored = r.m_explored + m;
    w.prev_active = (uint32_t)n;
    w.iter = r.iter + 1;
    w.visited = r.visited + n;
    *rs_cur = w;
    for (int k = 0; k < kStripes; ++k)
      for (int j = 0; j < 3; ++j) cnt_cur[k * kBfsCntStride + j] = 0;
    st->choice = pull ? 1 : 0;
    st->chunk_edges = ce;
    st_next->n_small = 0;
    st_next->n_med = 0;
    st_next->n_chunk = 0;
    if (host_out) {
      DevStats d{};
      d.n_active = (uint32_t)n;
      d.m_active = m;
      d.n_zero_deg = (uint32_t)z;
      d.visited = w.visited;
      d.choice = pull ? 1 : 0;
      d.iter = 0xFFFFFFFFu;
      *host_out = d;
      // the release store orders the record before `iter` at system scope
      __hip_atomic_store(&host_out->iter, record_tag(w.epoch, w.iter), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
  // Both the same in every workgroup.  An empty frontier ends the traversal:
  // the iterations queued behind this one find empty queues and zero counts
  // (the next run's initialisation clears both bitmaps).
  if (pull || n == 0) return;
  CompactAcc a;
  compact_frontier<DEG>(const_cast<uint64_t*>(in), out, words, g, bins, q_small, q_med, q_chunk, st, nullptr,
                        HubDefer{0, 1}, true, a);  // nothing dropped without values: `in` is only read
}

template <bool EXACT>
__global__ __launch_bounds__(kBlock) void k_bfs_step(DevGraph g, uint32_t* __restrict__ val,
                                                     const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                                                     const uint32_t* __restrict__ q_small,
                                                     const uint32_t* __restrict__ q_med,
                                                     const uint64_t* __restrict__ q_chunk, uint32_t chunk_edges,
                                                     const DevStats* __restrict__ st,
                                                     unsigned long long* __restrict__ cnt) {
  FirstVisits fv{g.off};
  if (st->choice == 1) {
    const uint64_t words = ((uint64_t)g.V + 63) / 64;
    const uint64_t nwaves = (uint64_t)gridDim.x * kWavesPerBlock;
    for (uint64_t w = (uint64_t)blockIdx.x * kWavesPerBlock + wave_id(); w < words; w += nwaves) {
      const uint64_t v = w * 64 + lane_id();
      bool found = false;
      if (v < g.V && val[v] == kUnreached) {
        const uint64_t s = g.in_off[v], e = g.in_off[v + 1];
        for (uint64_t k = s; k < e; ++k) {
          const uint32_t u = g.in_src[k];
          if ((in[u >> 6] >> (u & 63)) & 1ull) {
            val[v] = val[u] + 1;
            found = true;
            break;
          }
        }
      }
      const uint64_t mask = __ballot(found);
      if (lane_id() == 0) out[w] = mask;
      if (found) fv.first((uint32_t)v);
    }
  } else {
    const uint32_t ce = st->chunk_edges ? st->chunk_edges : chunk_edges;
    (void)expand_items<ALG_BFS, EXACT, false>(g, val, out, q_small, q_med, q_chunk, SmallRange{}, ce, st, fv);
  }
  const uint32_t n = wave_sum_u32(fv.n), z = wave_sum_u32(fv.z);
  const unsigned long long m = wave_sum_u64(fv.m);
  if (lane_id() == 0 && n) {
    unsigned long long* c = cnt + (blockIdx.x % kStripes) * kBfsCntStride;
    atomicAdd(c, (unsigned long long)n);
    atomicAdd(c + 1, m);
    if (z) atomicAdd(c + 2, (unsigned long long)z);
  }
}

// Bottom-up SSSP and label propagation: thread j of a workgroup takes kPullK
// consecutive in-edges of the transposed graph (whatever rows they belong
// to), keeps the candidates of those whose source is in the frontier, and
// folds them into one atomic min per row segment.  The minimum is exact and
// order-free, so no cross-thread combination is needed; the fixed point is
// the push engine's.
constexpr int kPullK = 8;
constexpr uint64_t kPullTile = (uint64_t)kBlock * kPullK;
typedef uint32_t pull_u32x4 __attribute__((ext_vector_type(4)));

__global__ void k_pull_tile_rows(const uint64_t* __restrict__ in_off, uint32_t V, uint64_t ntiles,
                                 uint32_t* __restrict__ tile_row) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t <= ntiles;
       t += (uint64_t)gridDim.x * blockDim.x) {
    if (t == ntiles) {
      tile_row[t] = V;
      continue;
    }
    const uint64_t e = t * kPullTile;
    uint32_t lo = 0, hi = V;  // in_off[V] = E > e
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo + 1) / 2;
      if (in_off[mid] <= e) lo = mid;
      else hi = mid - 1;
    }
    tile_row[t] = lo;
  }
}

__device__ __forceinline__ uint32_t pull_commit(uint32_t v, uint32_t m, uint32_t* __restrict__ val,
                                                uint64_t* __restrict__ out) {
  if (m < val[v]) {
    const uint32_t old = atomicMin(val + v, m);
    if (m < old) {
      atomicOr(reinterpret_cast<unsigned long long*>(out) + (v >> 6), 1ull << (v & 63));
      return old == kUnreached ? 1u : 0u;
    }
  }
  return 0u;
}

template <int ALG>
__global__ __launch_bounds__(kBlock) void k_pull_min(DevGraph g, uint64_t E, const uint32_t* __restrict__ tile_row,
                                                     uint32_t* __restrict__ val, const uint64_t* __restrict__ in,
                                                     uint64_t* __restrict__ out, uint32_t* __restrict__ stripes) {
  const uint64_t t = blockIdx.x;
  const uint64_t e0 = t * kPullTile + (uint64_t)threadIdx.x * kPullK;
  uint32_t n_new = 0;
  if (e0 < E) {
    const uint32_t nk = (E - e0) < (uint64_t)kPullK ? (uint32_t)(E - e0) : (uint32_t)kPullK;
    uint32_t lo = tile_row[t], hi = tile_row[t + 1];
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo + 1) / 2;
      if (g.in_off[mid] <= e0) lo = mid;
      else hi = mid - 1;
    }
    uint32_t u[kPullK], w[kPullK];
    if (nk == (uint32_t)kPullK) {
      const pull_u32x4* ps = reinterpret_cast<const pull_u32x4*>(g.in_src + e0);
      const pull_u32x4 a = ps[0], b = ps[1];
      u[0] = a.x; u[1] = a.y; u[2] = a.z; u[3] = a.w;
      u[4] = b.x; u[5] = b.y; u[6] = b.z; u[7] = b.w;
      if constexpr (ALG == ALG_SSSP) {
        const pull_u32x4* pw = reinterpret_cast<const pull_u32x4*>(g.in_w + e0);
        const pull_u32x4 c = pw[0], d = pw[1];
        w[0] = c.x; w[1] = c.y; w[2] = c.z; w[3] = c.w;
        w[4] = d.x; w[5] = d.y; w[6] = d.z; w[7] = d.w;
      }
    } else {
#pragma unroll
      for (int i = 0; i < kPullK; ++i) {
        u[i] = (uint32_t)i < nk ? g.in_src[e0 + i] : 0u;
        if constexpr (ALG == ALG_SSSP) w[i] = (uint32_t)i < nk ? g.in_w[e0 + i] : 0u;
      }
    }
    uint32_t cand[kPullK];
#pragma unroll
    for (int i = 0; i < kPullK; ++i) {
      cand[i] = 0xFFFFFFFFu;
      if ((uint32_t)i < nk && ((in[u[i] >> 6] >> (u[i] & 63)) & 1ull)) {
        if constexpr (ALG == ALG_SSSP) cand[i] = val[u[i]] + w[i];
        else cand[i] = val[u[i]];
      }
    }
    uint32_t r = lo;
    uint64_t next_end = g.in_off[r + 1];
    uint32_t m = 0xFFFFFFFFu;
#pragma unroll
    for (int i = 0; i < kPullK; ++i) {
      if ((uint32_t)i < nk) {
        m = cand[i] < m ? cand[i] : m;
        const uint64_t e = e0 + i + 1;
        if (next_end == e) {  // row r ends after this edge (and empty rows after it)
          if (m != 0xFFFFFFFFu) n_new += pull_commit(r, m, val, out);
          m = 0xFFFFFFFFu;
          while (next_end == e) {
            ++r;
            next_end = r < g.V ? g.in_off[r + 1] : ~0ull;
          }
        }
      }
    }
    if (m != 0xFFFFFFFFu) n_new += pull_commit(r, m, val, out);  // the row goes on past this thread
  }
  if constexpr (ALG != ALG_CC) {
    const uint32_t tot = wave_sum_u32(n_new);
    if (lane_id() == 0) stripe_add(stripes, tot);
  }
}

__global__ void k_apply_candidates(const uint32_t* __restrict__ ids, const uint32_t* __restrict__ vals,
                                   uint32_t n, uint32_t* __restrict__ val, uint64_t* __restrict__ out,
                                   uint32_t* __restrict__ stripes) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t fresh = 0;
  if (i < n) {
    const uint32_t d = ids[i], c = vals[i];
    if (c < val[d]) {
      const uint32_t old = atomicMin(val + d, c);
      if (c < old) {
        atomicOr(reinterpret_cast<unsigned long long*>(out) + (d >> 6), 1ull << (d & 63));
        fresh = old == kUnreached;
      }
    }
  }
  const uint32_t tot = wave_sum_u32(fresh);
  if (lane_id() == 0) stripe_add(stripes, tot);
}

// (vertex, value) of every vertex set in both `in` and `mask`, appended at
// *count (one atomic per wavefront that found any).  Lane per word.
__global__ __launch_bounds__(kBlock) void k_gather_masked(const uint64_t* __restrict__ in,
                                                          const uint64_t* __restrict__ mask, uint64_t words,
                                                          const uint32_t* __restrict__ val,
                                                          uint32_t* __restrict__ ids, uint32_t* __restrict__ vals,
                                                          uint32_t* __restrict__ count) {
  for (uint64_t base = blockIdx.x * (uint64_t)kBlock; base < words; base += (uint64_t)gridDim.x * kBlock) {
    const uint64_t w = base + threadIdx.x;
    const uint64_t bits = w < words ? in[w] & mask[w] : 0ull;
    const uint32_t c = (uint32_t)__popcll(bits);
    uint32_t total;
    const uint32_t off = wave_excl_scan(c, &total);
    uint32_t pos = 0;
    if (lane_id() == 0 && total) pos = atomicAdd(count, total);
    pos = __shfl(pos, 0, kWave) + off;
    for (uint64_t b = bits; b; b &= b - 1) {
      const uint32_t v = (uint32_t)(w * 64 + (uint64_t)__builtin_ctzll(b));
      ids[pos] = v;
      vals[pos] = val[v];
      ++pos;
    }
  }
}

__global__ void k_gather_values(const uint32_t* __restrict__ ids, uint32_t n,
                                const uint32_t* __restrict__ val, uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = val[ids[i]];
}

__global__ __launch_bounds__(kBlock) void k_finish_query(DevGraph g, const uint32_t* __restrict__ val,
                                                         unsigned long long* __restrict__ acc,
                                                         const uint32_t* __restrict__ new_of,
                                                         uint32_t* __restrict__ out) {
  __shared__ unsigned long long s[kWavesPerBlock];
  uint64_t sum = 0;
  const uint64_t t0 = blockIdx.x * (uint64_t)kBlock + threadIdx.x, nt = (uint64_t)gridDim.x * kBlock;
  uint64_t v0 = 0;  // vertices below v0 are summed four at a time
  if (acc && g.deg) {
    // 16-byte loads of four values and four out-degrees
    const uint64_t n4 = g.V / 4;
    const uint4* v4 = reinterpret_cast<const uint4*>(val);
    const uint4* d4 = reinterpret_cast<const uint4*>(g.deg);
    for (uint64_t i = t0; i < n4; i += nt) {
      const uint4 x = v4[i], d = d4[i];
      sum += (uint64_t)(x.x != kUnreached ? d.x : 0u) + (x.y != kUnreached ? d.y : 0u) +
             (x.z != kUnreached ? d.z : 0u) + (x.w != kUnreached ? d.w : 0u);
    }
    v0 = n4 * 4;
  }
  if (acc)
    for (uint64_t v = v0 + t0; v < g.V; v += nt)
      if (val[v] != kUnreached) sum += g.off[v + 1] - g.off[v];
  if (new_of) {
    // values back in the caller's ids: 16-byte loads of four ids and stores
    // of four values, so four gathers are in flight per lane (both buffers
    // are whole allocations, 16-byte aligned)
    const uint64_t n4 = g.V / 4;
    const uint4* o4 = reinterpret_cast<const uint4*>(new_of);
    uint4* w4 = reinterpret_cast<uint4*>(out);
    for (uint64_t i = t0; i < n4; i += nt) {
      const uint4 o = o4[i];
      w4[i] = make_uint4(val[o.x], val[o.y], val[o.z], val[o.w]);
    }
    for (uint64_t v = n4 * 4 + t0; v < g.V; v += nt) out[v] = val[new_of[v]];
  }
  if (!acc) return;  // uniform
  sum = wave_sum_u64(sum);
  if (lane_id() == 0) s[wave_id()] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < kWavesPerBlock; ++w) t += s[w];
    if (t) atomicAdd(acc + (blockIdx.x % kAccLines) * kAccStride, t);
  }
}

// Allocation-free push step straight from a bitmap: one lane per vertex, so a
// wavefront covers one frontier word; its active vertices' edges are spread
// over the lanes as in the small-degree tier of k_expand.
template <int ALG>
__global__ __launch_bounds__(kBlock) void k_push_bitmap(DevGraph g, uint32_t* __restrict__ val,
                                                        const uint64_t* __restrict__ in,
                                                        uint64_t* __restrict__ out) {
  const uint64_t v = blockIdx.x * (uint64_t)kBlock + threadIdx.x;
  const uint32_t lane = lane_id();
  const uint64_t words = ((uint64_t)g.V + 63) / 64;
  const uint64_t word = (v >> 6) < words ? in[v >> 6] : 0;
  if (word == 0) return;  // word is wave-uniform
  uint32_t deg = 0, sv = 0;
  uint64_t s = 0;
  if (v < g.V && ((word >> lane) & 1ull)) {
    s = g.off[v];
    deg = (uint32_t)(g.off[v + 1] - s);
    sv = val[v];
  }
  uint32_t total;
  const uint32_t excl = wave_excl_scan(deg, &total);
  for (uint32_t j0 = 0; j0 < total; j0 += kWave) {
    const uint32_t j = j0 + lane;
    uint32_t lo = 0;
#pragma unroll
    for (uint32_t step = kWave / 2; step >= 1; step >>= 1) {
      const uint32_t pm = __shfl(excl, (int)(lo + step), kWave);
      if (pm <= j) lo += step;
    }
    const uint64_t os = __shfl(s, (int)lo, kWave);
    const uint32_t osv = __shfl(sv, (int)lo, kWave);
    const uint32_t oex = __shfl(excl, (int)lo, kWave);
    MinTrack mt;  // not tracked here
    if (j < total) (void)relax<ALG, true>(osv, os + (j - oex), g.dst, g.w, val, out, mt);
  }
}

inline unsigned grid_for(uint64_t n, unsigned block = kBlock, unsigned cap = 1u << 20) {
  uint64_t b = (n + block - 1) / block;
  if (b == 0) b = 1;
  return (unsigned)(b < cap ? b : cap);
}

}  // namespace

#define CG_LAUNCH_CHECK()                                                               \
  do {                                                                                  \
    hipError_t e_ = hipGetLastError();                                                  \
    if (e_ != hipSuccess)                                                               \
      return set_error(CG_ERR_HIP, "kernel launch failed: %s (%s:%d)",                  \
                       hipGetErrorString(e_), __FILE__, __LINE__);                      \
  } while (0)

int launch_fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t s) {
  if (!n) return CG_OK;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, kBlock, 65536)), dim3(kBlock), 0, s, p, value, n);
  CG_LAUNCH_CHECK();
  return CG_OK;
}

int launch_fill_u64(uint64_t* p, uint64_t value, uint64_t n, hipStream_t s) {
  if (!n) return CG_OK;
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, kBlock, 65536)), dim3(kBlock), 0, s, p, value, n);
  CG_LAUNCH_CHECK();
  return CG_OK;
}

int launch_iota_u32(uint32_t* p, uint64_t n, hipStream_t s) {
  if (!n) return CG_OK;
  hipLaunchKernelGGL(k_iota_u32, dim3(grid_for(n, kBlock, 65536)), dim3(kBlock), 0, s, p, n);
  CG_LAUNCH_CHECK();
  return CG_OK;
}

int launch_bitmap_prefix(uint64_t* bm, uint64_t words, uint64_t n, hipStream_t s) {
  if (!words) return CG_OK;
  hipLaunchKernelGGL(k_bitmap_prefix, dim3(grid_for(words, kBlock, 65536)), dim3(kBlock), 0, s, bm,
                     words, n);
  CG_LAUNCH_CHECK();
  return CG_OK;
}

int launch_set_root(uint32_t* val, uint64_t* bm, uint32_t root, hipStream_t s) {
  hipLaunchKernelGGL(k_set_root, dim3(1), dim3(64), 0, s, val, bm, root);
  CG_LAUNCH_CHECK();
  return CG_OK;
}

int launch_init_run_state(RunState* rs, uint64_t visited, hipStream_t s) {
  hipLaunchKernelGGL(k_init_run_state, dim3(1), dim3(64), 0, s, rs, (unsigned long long)visited);
  CG_LAUNCH_CHECK();
  return CG_OK;
}

uint64_t stats_parts_needed(uint64_t words) {
  const uint64_t blocks = (words * kSlicesPerWord + kBlock - 1) / kBlock;
  return (blocks < kStatsMaxBlocks ? blocks : kStatsMaxBlocks) + 1;
}

int launch_stats(const uint64_t* in, uint64_t* zero_me, uint64_t words, const DevGraph& g, uint64_t E,
                 StatsPart* parts, uint32_t* stripes, RunState* rs, const float* mlp,
                 const uint32_t* val, uint32_t threshold, DevStats* st, DevStats* host_out,
                 hipStream_t s, int direction, const DevStats* prev, bool min_from_stripes) {
  const uint64_t slices = words * kSlicesPerWord;
  uint64_t blocks = slices ? (slices + kBlock - 1) / kBlock : 1;
  if (blocks > kStatsMaxBlocks) blocks = kStatsMaxBlocks;
  hipLaunchKernelGGL(k_stats, dim3((unsigned)blocks), dim3(kBlock), 0, s, in, zero_me, words, g, parts, val,
                     threshold);
  CG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_stats_reduce, dim3(1), dim3(kBlock), 0, s, parts, (uint32_t)blocks, stripes, rs, g.V, E, mlp,
                     st, host_out, direction, prev, min_from_stripes ? 1 : 0);
  CG_LAUNCH_CHECK();
  return CG_OK;
}

int launch_compact(const uint64_t* in, uint64_t words, const DevGraph& g, const Bins& bins,
                   uint32_t* q_small, uint32_t* q_med, uint64_t* q_chunk, DevStats* st,
                   const uint32_t* val, uint32_t threshold, uint64_t* carry, hipStream_t s,
                   const TierTable* choice_tiers, uint32_t skip_bits, HubDefer hub) {
  if (!words) return CG_OK;
  if (hub.degree && (!val || !hub.delta)) return set_error(CG_ERR_INVALID, "hub deferral needs values and delta >= 1");
  const uint64_t blocks = (words * kSlicesPerWord + kBlock - 1) / kBlock;
  const TierTable tiers = choice_tiers ? *choice_tiers : TierTable{};
  hipLaunchKernelGGL(k_compact, dim3((unsigned)blocks), dim3(kBlock), 0, s, in, words, g, bins,
                     q_small, q_med, q_chunk, st, val, threshold, carry, choice_tiers ? 1 : 0, tiers, skip_bits,
                     hub);
  CG_LAUNCH_CHECK();
  return CG_OK;
}

int launch_compact_fused(uint64_t* in, uint64_t* out, uint64_t words, const DevGraph& g, const Bins& bins,
                         uint32_t* q_small, uint32_t* q_med, uint64_t* q_chunk, DevStats* st,
                         const uint32_t* val, HubDefer hub, StatsPart* parts, uint32_t* nparts,
                         const RunState* rs, hipStream_t s, const TierTable* choice_tiers) {
  if (hub.degree && (!val || !hub.delta)) return set_error(CG_ERR_INVALID, "hub deferral needs values and delta >= 1");
  const uint64_t slices = words * kSlicesPerWord;
  uint64_t blocks = slices ? (slices + kBlock - 1) / kBlock : 1;
  const uint64_t cap = g.deg ? kFusedMaxBlocksDeg : kFusedMaxBlocks;
  if (blocks > cap) blocks = cap;
  const TierTable tiers = choice_tiers ? *choice_tiers : TierTable{};
  if (g.deg)
    hipLaunchKernelGGL(k_compact_fused<true>, dim3((unsigned)blocks), dim3(kBlock), 0, s, in, out, words, g, bins,
                       q_small, q_med, q_chunk, st, val, hub, parts, choice_tiers ? 1 : 0, tiers, rs);
  else
    hipLaunchKernelGGL(k_compact_fused<false>, dim3((unsigned)blocks), dim3(kBlock), 0, s, in, out, words, g, bins,
                       q_small, q_med, q_chunk, st, val, hub, parts, choice_tiers ? 1 : 0, tiers, rs);
  CG_LAUNCH_CHECK();
  *nparts = (uint32_t)blocks;
  return CG_OK;
}

int launch_fused_finish(const StatsPart* parts, uint32_t nparts, uint32_t* stripes, RunState* rs, DevStats* st,
                        DevStats* next, DevStats* host_out, bool track, const float* mlp, uint32_t V, uint64_t E,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_fused_finish, dim3(1), dim3(kFinishBlock), 0, s, parts, nparts, stripes, rs, st, next, host_out,
                     track ? 1 : 0, mlp, V, E);
  CG_LAUNCH_CHECK();
  return CG_OK;
}

uint64_t expand_item_bound(uint64_t n_active, uint64_t m_active, const Bins& bins) {
  // chunks <= vertices + edges / chunk_edges; medium <= vertices; small groups
  return n_active + m_active / bins.chunk_edges + n_active + n_active / kWave + 1;
}

int launch_expand_push(int alg, const DevGraph& g, uint32_t* val, uint64_t* out,
                       const uint32_t* q_small, const uint32_t* q_med, const uint64_t* q_chunk,
                       uint64_t item_bound, SmallRange range, const Bins& bins, DevStats* st,
                       uint32_t* stripes, bool exact, hipStream_t s, bool track_min) {
  if (!item_bound) return CG_OK;
  if (track_min && alg != CG_SSSP) return set_error(CG_ERR_INTERNAL, "minimum tracking is for SSSP");
  uint64_t blocks = (item_bound + kWavesPerBlock - 1) / kWavesPerBlock;
  uint64_t cap = bins.grid_cap ? bins.grid_cap : 2048;  // 256 CUs x 32 wavefronts
  if (cap > (1u << 20)) cap = 1u << 20;                 // the wavefronts stride; more adds nothing
  if (blocks > cap) blocks = cap;
  dim3 grid((unsigned)blocks), block(kBlock);
#define CG_EXPAND(A, X)                                                                          \
  hipLaunchKernelGGL((k_expand<A, X, false>), grid, block, 0, s, g, val, out, q_small, q_med, q_chunk, range, \
                     bins.chunk_edges, st, stripes)
  // SSSP reads the narrowest weight copy the graph has
#define CG_EXPAND_SSSP(X, T)                                                                              \
  do {                                                                                                    \
    if (g.w8)                                                                                             \
      hipLaunchKernelGGL((k_expand<ALG_SSSP, X, T, uint8_t>), grid, block, 0, s, g, val, out, q_small, q_med, \
                         q_chunk, range, bins.chunk_edges, st, stripes);                                  \
    else if (g.w16)                                                                                       \
      hipLaunchKernelGGL((k_expand<ALG_SSSP, X, T, uint16_t>), grid, block, 0, s, g, val, out, q_small,      \
                         q_med, q_chunk, range, bins.chunk_edges, st, stripes);                           \
    else                                                                                                  \
      hipLaunchKernelGGL((k_expand<ALG_SSSP, X, T, uint32_t>), grid, block, 0, s, g, val, out, q_small,      \
                         q_med, q_chunk, range, bins.chunk_edges, st, stripes);                           \
  } while (0)
  if (track_min) {
    if (exact) CG_EXPAND_SSSP(true, true);
    else CG_EXPAND_SSSP(false, true);
    CG_LAUNCH_CHECK();
    return CG_OK;
  }
  switch (alg * 2 + (exact ? 1 : 0)) {
    case CG_BFS * 2: CG_EXPAND(ALG_BFS, false); break;
    case CG_BFS * 2 + 1: CG_EXPAND(ALG_BFS, true); break;
    case CG_SSSP * 2: CG_EXPAND_SSSP(false, false); break;
    case CG_SSSP * 2 + 1: CG_EXPAND_SSSP(true, false); break;
    case CG_CC * 2: CG_EXPAND(ALG_CC, false); break;
    case CG_CC * 2 + 1: CG_EXPAND(ALG_CC, true); break;
    default:
      return set_error(CG_ERR_UNSUPPORTED, "push expansion: algorithm %d", alg);
  }
#undef CG_EXPAND
#undef CG_EXPAND_SSSP
  CG_LAUNCH_CHECK();
  return CG_OK;
}

int launch_pull_bfs(const DevGraph& g, uint32_t* val, const uint64_t* in, uint64_t* out,
                    uint32_t* stripes, hipStream_t s, const DevStats* cond) {
  if (!g.in_off || !g.in_src) return set_error(CG_ERR_INTERNAL, "pull step without in-edges");
  const uint64_t words = ((uint64_t)g.V + 63) / 64;
  uint64_t blocks = (words + kWavesPerBlock - 1) / kWavesPerBlock;
  if (blocks > 4096) blocks = 4096;  // wavefronts stride over the words
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_pull_bfs, dim3((unsigned)blocks), dim3(kBlock), 0, s, g, val, in, out, stripes, cond);
  CG_LAUNCH_CHECK();
  return CG_OK;
}

int launch_init_traverse(const TraverseInit& a, hipStream_t s) {
  uint64_t n = a.V / 4 > a.bm_words ? a.V / 4 : a.bm_words;
  if (n < 2ull * kBfsCntWords) n = 2ull * kBfsCntWords;
  hipLaunchKernelGGL(k_init_traverse, dim3(grid_for(n, kBlock, 65536)), dim3(kBlock), 0, s, a);
  CG_LAUNCH_CHECK();
  return CG_OK;
}

int launch_bfs_front(const uint64_t* in, uint64_t* out, uint64_t words, const DevGraph& g, uint64_t E,
                     const Bins& bins, uint32_t* q_small, uint32_t* q_med, uint64_t* q_chunk, DevStats* st,
                     DevStats* st_next, const unsigned long long* cnt_prev, unsigned long long* cnt_cur,
                     const RunState* rs_prev, RunState* rs_cur, DevStats* host_out, hipStream_t s) {
  const uint64_t slices = words * kSlicesPerWord;
  uint64_t fblocks = slices ? (slices + kBlock - 1) / kBlock : 1;
  const uint64_t cap = g.deg ? kFusedMaxBlocksDeg : kFusedMaxBlocks;
  if (fblocks > cap) fblocks = cap;
  if (g.deg)
    hipLaunchKernelGGL(k_bfs_front<true>, dim3((unsigned)fblocks), dim3(kBlock), 0, s, in, out, words, g, E, bins,
                       q_small, q_med, q_chunk, st, st_next, cnt_prev, cnt_cur, rs_prev, rs_cur, host_out);
  else
    hipLaunchKernelGGL(k_bfs_front<false>, dim3((unsigned)fblocks), dim3(kBlock), 0, s, in, out, words, g, E, bins,
                       q_small, q_med, q_chunk, st, st_next, cnt_prev, cnt_cur, rs_prev, rs_cur, host_out);
  CG_LAUNCH_CHECK();
  return CG_OK;
}

int launch_bfs_step(const uint64_t* in, uint64_t* out, uint64_t words, const DevGraph& g, const Bins& bins,
                    uint32_t* val, const uint32_t* q_small, const uint32_t* q_med, const uint64_t* q_chunk,
                    const DevStats* st, unsigned long long* cnt_cur, bool exact, hipStream_t s) {
  if (!g.in_off || !g.in_src) return set_error(CG_ERR_INTERNAL, "bottom-up steps without in-edges");
  // one grid for both directions: the push tiers' cap, at least the
  // bottom-up words' wavefronts up to 4096 workgroups
  uint64_t blocks = bins.grid_cap ? bins.grid_cap : 2048;
  uint64_t pblocks = (words + kWavesPerBlock - 1) / kWavesPerBlock;
  if (pblocks > 4096) pblocks = 4096;
  if (blocks < pblocks) blocks = pblocks;
  if (blocks < 1) blocks = 1;
  if (exact)
    hipLaunchKernelGGL((k_bfs_step<true>), dim3((unsigned)blocks), dim3(kBlock), 0, s, g, val, in, out, q_small, q_med,
                       q_chunk, bins.chunk_edges, st, cnt_cur);
  else
    hipLaunchKernelGGL((k_bfs_step<false>), dim3((unsigned)blocks), dim3(kBlock), 0, s, g, val, in, out, q_small,
                       q_med, q_chunk, bins.chunk_edges, st, cnt_cur);
  CG_LAUNCH_CHECK();
  return CG_OK;
}

uint64_t pull_tile_edges() { return kPullTile; }

uint64_t pull_tile_count(uint64_t E) { return (E + kPullTile - 1) / kPullTile; }

int build_pull_tiles(const uint64_t* in_off, uint32_t V, uint64_t E, uint32_t* tile_row, hipStream_t s) {
  const uint64_t nt = pull_tile_count(E);
  const uint64_t blocks = (nt + 1 + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_pull_tile_rows, dim3((unsigned)(blocks < 65536 ? blocks : 65536)), dim3(kBlock), 0, s, in_off,
                     V, nt, tile_row);
  CG_LAUNCH_CHECK();
  return CG_OK;
}

int launch_pull_min(int alg, const DevGraph& g, uint64_t E, const uint32_t* tile_row, uint32_t* val,
                    const uint64_t* in, uint64_t* out, uint32_t* stripes, hipStream_t s) {
  if (!g.in_off || !g.in_src || !tile_row) return set_error(CG_ERR_INTERNAL, "pull step without in-edges");
  if (alg == CG_SSSP && !g.in_w) return set_error(CG_ERR_INTERNAL, "pull SSSP step without in-edge weights");
  const uint64_t nt = pull_tile_count(E);
  if (!nt) return CG_OK;
  if (alg == CG_SSSP)
    hipLaunchKernelGGL(k_pull_min<ALG_SSSP>, dim3((unsigned)nt), dim3(kBlock), 0, s, g, E, tile_row, val, in, out,
                       stripes);
  else if (alg == CG_CC)
    hipLaunchKernelGGL(k_pull_min<ALG_CC>, dim3((unsigned)nt), dim3(kBlock), 0, s, g, E, tile_row, val, in, out,
                       stripes);
  else
    return set_error(CG_ERR_INTERNAL, "min-pull step for algorithm %d", alg);
  CG_LAUNCH_CHECK();
  return CG_OK;
}

int launch_apply_candidates(const uint32_t* ids, const uint32_t* vals, uint32_t n, uint32_t* val,
                            uint64_t* out, uint32_t* stripes, hipStream_t s) {
  if (!n) return CG_OK;
  hipLaunchKernelGGL(k_apply_candidates, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, ids,
                     vals, n, val, out, stripes);
  CG_LAUNCH_CHECK();
  return CG_OK;
}

int launch_gather_masked(const uint64_t* in, const uint64_t* mask, uint64_t words, const uint32_t* val, uint32_t* ids,
                         uint32_t* vals, uint32_t* count, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, count, 0u, (uint64_t)1);
  CG_LAUNCH_CHECK();
  if (!words) return CG_OK;
  hipLaunchKernelGGL(k_gather_masked, dim3(grid_for(words, kBlock, 4096)), dim3(kBlock), 0, s, in, mask, words, val,
                     ids, vals, count);
  CG_LAUNCH_CHECK();
  return CG_OK;
}

int launch_gather_values(const uint32_t* ids, uint32_t n, const uint32_t* val, uint32_t* out,
                         hipStream_t s) {
  if (!n) return CG_OK;
  hipLaunchKernelGGL(k_gather_values, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, ids, n,
                     val, out);
  CG_LAUNCH_CHECK();
  return CG_OK;
}

__global__ void k_degrees(const uint64_t* __restrict__ off, uint32_t V, uint64_t n, uint32_t* __restrict__ deg) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    deg[i] = i < V ? (uint32_t)(off[i + 1] - off[i]) : 0u;
}

__global__ void k_max_u32(const uint32_t* __restrict__ x, uint64_t n, uint32_t* __restrict__ mx) {
  uint32_t m = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    m = x[i] > m ? x[i] : m;
#pragma unroll
  for (int d = kWave / 2; d >= 1; d >>= 1) {
    const uint32_t o = __shfl_xor(m, d, kWave);
    m = o > m ? o : m;
  }
  if (lane_id() == 0 && m) atomicMax(mx, m);
}

template <class T>
__global__ void k_narrow(const uint32_t* __restrict__ w, uint64_t n, T* __restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = (T)w[i];
}

int ensure_narrow_weights(cg_graph* g, hipStream_t s) {
  if (g->dev_wn_valid || !g->has_w || !g->E) return CG_OK;
  CG_TRY(g->ensure_device());
  DevBuf<uint32_t> mx;
  CG_TRY(mx.ensure(1));
  CG_HIP(hipMemsetAsync(mx.p, 0, 4, s));
  const unsigned grid = grid_for(g->E, kBlock, 8192);
  hipLaunchKernelGGL(k_max_u32, dim3(grid), dim3(kBlock), 0, s, g->d_w.p, g->E, mx.p);
  CG_LAUNCH_CHECK();
  uint32_t m = 0;
  CG_HIP(hipMemcpyAsync(&m, mx.p, 4, hipMemcpyDeviceToHost, s));
  CG_HIP(hipStreamSynchronize(s));
  if (m < 256u) {
    CG_TRY(g->d_w8.ensure(g->E));
    hipLaunchKernelGGL(k_narrow<uint8_t>, dim3(grid), dim3(kBlock), 0, s, g->d_w.p, g->E, g->d_w8.p);
  } else if (m < 65536u) {
    CG_TRY(g->d_w16.ensure(g->E));
    hipLaunchKernelGGL(k_narrow<uint16_t>, dim3(grid), dim3(kBlock), 0, s, g->d_w.p, g->E, g->d_w16.p);
  }
  CG_LAUNCH_CHECK();
  CG_HIP(hipStreamSynchronize(s));
  g->max_w = m;
  g->dev_wn_valid = true;
  return CG_OK;
}

int ensure_degrees(cg_graph* g, hipStream_t s) {
  if (g->dev_deg_valid) return CG_OK;
  CG_TRY(g->ensure_device());
  const uint64_t n = ((uint64_t)g->V + 63) / 64 * 64 + 64;  // whole slices of every bitmap word
  CG_TRY(g->d_deg.ensure(n));
  hipLaunchKernelGGL(k_degrees, dim3(grid_for(n, kBlock, 8192)), dim3(kBlock), 0, s, g->d_off.p, g->V, n,
                     g->d_deg.p);
  CG_LAUNCH_CHECK();
  CG_HIP(hipStreamSynchronize(s));
  g->dev_deg_valid = true;
  return CG_OK;
}

int launch_finish_query(const DevGraph& g, const uint32_t* val, unsigned long long* acc, const uint32_t* new_of,
                        uint32_t* out, hipStream_t s) {
  if ((!acc && !new_of) || !g.V) return CG_OK;
  hipLaunchKernelGGL(k_finish_query, dim3(grid_for(g.V, kBlock, 8192)), dim3(kBlock), 0, s, g, val, acc, new_of,
                     out);
  CG_LAUNCH_CHECK();
  return CG_OK;
}

__global__ __launch_bounds__(kAccLines) void k_query_done(const unsigned long long* __restrict__ acc,
                                                           DevStats* __restrict__ host_out, uint32_t tag) {
  // one line's slot per thread, wavefront sums, the wavefronts' sums in order
  __shared__ unsigned long long s[kAccLines / kWave];
  const uint64_t x = wave_sum_u64(acc ? acc[threadIdx.x * kAccStride] : 0ull);
  if (lane_id() == 0) s[wave_id()] = x;
  __syncthreads();
  if (threadIdx.x != 0) return;
  DevStats d{};
  for (int w = 0; w < kAccLines / kWave; ++w) d.m_active += s[w];
  d.iter = 0xFFFFFFFFu;
  *host_out = d;
  // the release store orders the record before `iter` at system scope
  __hip_atomic_store(&host_out->iter, tag, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

int launch_query_done(const unsigned long long* acc, DevStats* host_out, uint32_t tag, hipStream_t s) {
  hipLaunchKernelGGL(k_query_done, dim3(1), dim3(kAccLines), 0, s, acc, host_out, tag);
  CG_LAUNCH_CHECK();
  return CG_OK;
}

hipError_t stream_wait(hipStream_t s) {
  const auto t0 = std::chrono::steady_clock::now();
  for (uint32_t spin = 1;; ++spin) {
    const hipError_t r = hipStreamQuery(s);
    if (r != hipErrorNotReady) return r;
    if ((spin & 255) == 0 &&
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() > 2.0)
      return hipStreamSynchronize(s);
    __builtin_ia32_pause();
  }
}

int push_step_from_bitmap(int alg, const DevGraph& g, uint32_t* val, const uint64_t* in,
                          uint64_t* out, hipStream_t s) {
  const uint64_t threads = ((uint64_t)g.V + 63) / 64 * 64;
  if (!threads) return CG_OK;
  dim3 grid((unsigned)((threads + kBlock - 1) / kBlock)), block(kBlock);
  switch (alg) {
    case CG_BFS: hipLaunchKernelGGL(k_push_bitmap<ALG_BFS>, grid, block, 0, s, g, val, in, out); break;
    case CG_SSSP: hipLaunchKernelGGL(k_push_bitmap<ALG_SSSP>, grid, block, 0, s, g, val, in, out); break;
    case CG_CC: hipLaunchKernelGGL(k_push_bitmap<ALG_CC>, grid, block, 0, s, g, val, in, out); break;
    default: return set_error(CG_ERR_UNSUPPORTED, "push step: algorithm %d", alg);
  }
  CG_LAUNCH_CHECK();
  return CG_OK;
}

}  // namespace cg
