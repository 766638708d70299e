// Batched BFS / SSSP queries over graphs of uneven size.
//
// All queries advance level by level together.  A level is two launches:
//   k_batch_plan    one workgroup: exclusive scan of the per-query frontier
//                   sizes into a global item space, clear of the next
//                   level's counters, per-query delta-stepping threshold
//                   update, done flag when nothing is left;
//   k_batch_expand  a fixed grid of workgroups whose wavefronts stride over
//                   groups of 64 items; an item is one frontier vertex of one
//                   query, and the 64 items' edges are spread evenly over the
//                   lanes (owner found by binary search over the wave's degree
//                   prefix) whatever query or graph each item belongs to.  An
//                   improved vertex joins its query's next frontier once,
//                   guarded by a per-query level stamp.  For SSSP a vertex
//                   whose distance is not below the query's threshold is
//                   carried to the next frontier unexpanded (delta-stepping,
//                   Meyer and Sanders 2003); the threshold moves past the
//                   smallest waiting distance when nothing below it is left.
// Up to kLevelMaxQ queries a level is one launch instead, k_batch_level:
// every workgroup scans the frontier sizes itself into LDS and takes the
// thresholds there, with counters and minima in rotating sets.
// The launch arguments never change from replay to replay (the level number,
// parity, thresholds and item count live in device memory), so
// `batch_levels` levels are captured once into a HIP graph and the graph is
// replayed until the done flag is set -- one host synchronisation per replay
// instead of one per level per query.  Queries on large small-world graphs
// run beside the levels on engines of their graphs (run_batched).  mode 1
// of cg_batch_run runs the same queries one per engine launch for
// comparison.
//
// The relaxation is the reference's (my_project/CGgraph/device_algorithm.cuh:
// 51-54 BFS, 253-257 SSSP), so each query ends at its unique fixed point,
// bit-identical to a single-query run.
#include <algorithm>
#include <chrono>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "engine.hpp"

namespace cg {

namespace {

struct QueryDesc {
  const uint64_t* off;
  const uint32_t* dst;
  const uint32_t* w;
  uint32_t* val;
  uint32_t* stamp;
  uint32_t* front[2];
  uint32_t V;
  int alg;
  uint32_t root;
  uint32_t delta;  // 0: no deferral
  // SSSP hub deferral: a frontier vertex of out-degree >= hub_degree waits
  // while its distance is >= the query's frontier minimum + hub_delta (0: off)
  uint32_t hub_degree, hub_delta;
};

struct BatchState {
  uint32_t next_level;   // level the next plan prepares
  uint32_t cur_level;    // level the running expansion processes
  uint32_t done;
  uint32_t total;        // items of the current level
  uint32_t levels_run;   // levels with work
  uint32_t base;         // one kernel per level: the level of the replay's first node
  uint32_t pad[2];
};

constexpr int kPlanBlock = 1024;
constexpr uint32_t kNone = 0xFFFFFFFFu;

__global__ void k_batch_init(const QueryDesc* __restrict__ q, int nq, const uint64_t* __restrict__ vpre,
                             uint64_t total) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    int lo = 0, hi = nq - 1;  // largest query with vpre[q] <= i
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (vpre[mid] <= i) lo = mid;
      else hi = mid - 1;
    }
    const uint32_t v = (uint32_t)(i - vpre[lo]);
    q[lo].val[v] = v == q[lo].root ? 0u : kUnreached;
    q[lo].stamp[v] = 0;
    if (v == q[lo].root) q[lo].front[0][0] = v;
  }
}

// cnt[2][nq] frontier sizes, qs = thr[nq] | minv[2][nq] (smallest value
// appended to each frontier).
__global__ void k_batch_reset(const QueryDesc* __restrict__ qd, uint32_t* __restrict__ cnt,
                              uint32_t* __restrict__ qs, int nq, BatchState* __restrict__ bs) {
  for (int i = threadIdx.x; i < nq; i += blockDim.x) {
    cnt[i] = 1;  // level 0: the root
    cnt[nq + i] = 0;
    qs[i] = qd[i].delta ? qd[i].delta : kNone;
    qs[nq + i] = 0;
    qs[2 * nq + i] = kNone;
  }
  if (threadIdx.x == 0) {
    bs->next_level = 0;
    bs->cur_level = 0;
    bs->done = 0;
    bs->total = 0;
    bs->levels_run = 0;
  }
}

__global__ __launch_bounds__(kPlanBlock) void k_batch_plan(const QueryDesc* __restrict__ qd,
                                                           uint32_t* __restrict__ cnt, uint32_t* __restrict__ qs,
                                                           int nq, uint32_t* __restrict__ item_off,
                                                           BatchState* __restrict__ bs) {
  __shared__ uint32_t wsum[kPlanBlock / 64];
  __shared__ uint32_t carry;
  const uint32_t L = bs->next_level;
  const uint32_t* cur = cnt + (L & 1) * nq;
  uint32_t* nxt = cnt + ((L + 1) & 1) * nq;
  uint32_t* thr = qs;
  const uint32_t* minv_cur = qs + nq + (L & 1) * nq;
  uint32_t* minv_nxt = qs + nq + ((L + 1) & 1) * nq;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nq; base += kPlanBlock) {
    const int i = base + threadIdx.x;
    const uint32_t x = i < nq ? cur[i] : 0u;
    // block exclusive scan: wave scan + wave totals
    uint32_t v = x;
    const uint32_t lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t t = __shfl_up(v, d, 64);
      if (lane >= (uint32_t)d) v += t;
    }
    if (lane == 63) wsum[threadIdx.x / 64] = v;
    __syncthreads();
    uint32_t wpre = 0;
    for (int w = 0; w < (int)(threadIdx.x / 64); ++w) wpre += wsum[w];
    if (i < nq) item_off[i] = carry + wpre + v - x;
    __syncthreads();
    if (threadIdx.x == kPlanBlock - 1) carry += wpre + v;
    __syncthreads();
  }
  for (int i = threadIdx.x; i < nq; i += kPlanBlock) {
    nxt[i] = 0;
    minv_nxt[i] = kNone;
    const uint32_t d = qd[i].delta;
    if (d && cur[i] && minv_cur[i] >= thr[i]) {
      // nothing of this frontier lies below the threshold: move it past the
      // smallest waiting distance
      const uint64_t t = (uint64_t)minv_cur[i] + d;
      thr[i] = t > 0xFFFFFFFEull ? 0xFFFFFFFEu : (uint32_t)t;
    }
  }
  if (threadIdx.x == 0) {
    item_off[nq] = carry;
    bs->total = carry;
    bs->cur_level = L;
    bs->next_level = L + 1;
    if (carry == 0) bs->done = 1;
    else bs->levels_run += 1;
  }
}

// Appends the flagged lanes' vertices to their queries' next frontiers with
// one atomic per distinct query in the wavefront, and folds their values into
// the queries' frontier minima.
__device__ __forceinline__ void wave_append(bool push, int q, uint32_t v, uint32_t value,
                                            const QueryDesc* __restrict__ qd, uint32_t* __restrict__ nxt_cnt,
                                            uint32_t* __restrict__ minv_nxt, int np, uint32_t lane) {
  uint64_t pending = __ballot(push);
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const int lq = __shfl(q, leader, 64);
    const bool mine = push && q == lq;
    const uint64_t same = __ballot(mine);
    uint32_t mn = mine ? value : kNone;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
      const uint32_t o = __shfl_xor(mn, d, 64);
      mn = o < mn ? o : mn;
    }
    uint32_t pos = 0;
    if ((int)lane == leader) {
      pos = atomicAdd(nxt_cnt + lq, (uint32_t)__popcll(same));
      if (qd[lq].delta || qd[lq].hub_degree) atomicMin(minv_nxt + lq, mn);
    }
    pos = __shfl(pos, leader, 64);
    if (mine) qd[lq].front[np][pos + (uint32_t)__popcll(same & ((1ull << lane) - 1))] = v;
    pending &= ~same;
  }
}

// One level's expansion: the wavefronts of the grid stride over groups of
// gsz items of the level's frontiers (item_off: exclusive prefix of their
// sizes over the queries, total items; thr / minv_cur: each query's
// threshold and frontier minimum for this level; nxt_cnt / minv_nxt: the
// next frontiers' counters and minima; np: the parity of the next fronts).
__device__ __forceinline__ void expand_level(const QueryDesc* __restrict__ qd, int nq, const uint32_t* item_off,
                                             uint32_t total, uint32_t L, const uint32_t* thr,
                                             const uint32_t* minv_cur, uint32_t* __restrict__ nxt_cnt,
                                             uint32_t* __restrict__ minv_nxt, int np, uint32_t gsz) {
  const int cp = np ^ 1;
  const uint32_t lane = __lane_id();
  const uint32_t waves = gridDim.x * (blockDim.x / 64);
  // wavefronts stride over groups of gsz items (no shared cursor to contend
  // on); smaller groups put more wavefronts on a level whose frontier is
  // small, shortening the level's latency chain
  for (uint64_t grp = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; grp * gsz < total; grp += waves) {
    const uint32_t i = (uint32_t)(grp * gsz) + lane;
    int q = 0;
    uint32_t deg = 0, sv = 0, v = 0;
    uint64_t s = 0;
    bool carry = false;
    if (lane < gsz && i < total) {
      int lo = 0, hi = nq - 1;  // largest q with item_off[q] <= i
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (item_off[mid] <= i) lo = mid;
        else hi = mid - 1;
      }
      q = lo;
      const QueryDesc& Q = qd[q];
      v = Q.front[cp][i - item_off[q]];
      sv = Q.val[v];
      s = Q.off[v];
      deg = (uint32_t)(Q.off[v + 1] - s);
      // a vertex waits for a later bucket (delta-stepping) or, as a hub, for
      // its distance to come within hub_delta of the frontier's smallest:
      // it stays in the frontier and expands nothing now
      bool wait = sv >= thr[q];
      if (!wait && Q.hub_degree && deg >= Q.hub_degree) {
        const uint64_t lim = (uint64_t)minv_cur[q] + Q.hub_delta;
        wait = (uint64_t)sv >= lim;
      }
      if (wait) {
        carry = atomicMax(Q.stamp + v, L + 1) < L + 1;
        deg = 0;
      }
    }
    wave_append(carry, q, v, sv, qd, nxt_cnt, minv_nxt, np, lane);
    // exclusive scan of degrees over the wave
    uint32_t excl = deg;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t t = __shfl_up(excl, d, 64);
      if (lane >= (uint32_t)d) excl += t;
    }
    const uint32_t tot = __shfl(excl, 63, 64);
    excl -= deg;
    for (uint32_t j0 = 0; j0 < tot; j0 += 64) {
      const uint32_t j = j0 + lane;
      uint32_t lo = 0;
#pragma unroll
      for (uint32_t step = 32; step >= 1; step >>= 1) {
        const uint32_t pm = __shfl(excl, (int)(lo + step), 64);
        if (pm <= j) lo += step;
      }
      const int oq = __shfl(q, (int)lo, 64);
      const uint64_t os = __shfl(s, (int)lo, 64);
      const uint32_t osv = __shfl(sv, (int)lo, 64);
      const uint32_t oex = __shfl(excl, (int)lo, 64);
      bool push = false;
      uint32_t d = 0, cand = 0;
      if (j < tot) {
        const QueryDesc& Q = qd[oq];
        const uint64_t e = os + (j - oex);
        d = Q.dst[e];
        cand = Q.alg == CG_BFS ? osv + 1 : osv + Q.w[e];
        if (cand < Q.val[d]) {
          const uint32_t old = atomicMin(Q.val + d, cand);
          if (cand < old) push = atomicMax(Q.stamp + d, L + 1) < L + 1;
        }
      }
      wave_append(push, oq, d, cand, qd, nxt_cnt, minv_nxt, np, lane);
    }
  }
}

__global__ __launch_bounds__(1024) void k_batch_expand(const QueryDesc* __restrict__ qd, int nq,
                                                       const uint32_t* __restrict__ item_off,
                                                       uint32_t* __restrict__ cnt, const uint32_t* __restrict__ qs,
                                                       uint32_t* __restrict__ qs_mut, BatchState* __restrict__ bs,
                                                       uint32_t gsz) {
  const uint32_t total = bs->total;
  if (total == 0) return;
  const uint32_t L = bs->cur_level;
  const int cp = L & 1, np = cp ^ 1;
  expand_level(qd, nq, item_off, total, L, qs, qs + nq + cp * nq, cnt + np * nq, qs_mut + nq + np * nq, np, gsz);
}

// One kernel per level (at most kLevelMaxQ queries): every workgroup scans
// the level's frontier sizes itself into LDS and takes each query's
// threshold for the level (k_batch_plan's rule), so no planning launch sits
// between two levels.  Counters and minima are in three rotating sets
// (level L reads set L % 3, appends to set (L + 1) % 3, and workgroup 0
// clears set (L + 2) % 3, which level L - 1 read and level L + 1 appends
// to), thresholds in two (workgroup 0 stores level L's for level L + 1).
// The level is bs->base + k, k the node's place in the captured replay,
// whose last node advances bs->base.
constexpr int kLevelMaxQ = 4096;

__global__ __launch_bounds__(1024) void k_batch_level(const QueryDesc* __restrict__ qd, int nq,
                                                      uint32_t* __restrict__ cnt3, uint32_t* __restrict__ thr2,
                                                      uint32_t* __restrict__ minv3, BatchState* __restrict__ bs,
                                                      uint32_t gsz, uint32_t k) {
  extern __shared__ uint32_t sh[];
  uint32_t* s_off = sh;            // nq + 1
  uint32_t* s_thr = sh + nq + 1;   // nq
  uint32_t* s_minv = s_thr + nq;   // nq
  __shared__ uint32_t wsum[16];
  __shared__ uint32_t carry;
  const uint32_t L = bs->base + k;
  const uint32_t* cur = cnt3 + (size_t)(L % 3) * nq;
  uint32_t* nxt = cnt3 + (size_t)((L + 1) % 3) * nq;
  uint32_t* clr = cnt3 + (size_t)((L + 2) % 3) * nq;
  const uint32_t* minv_c = minv3 + (size_t)(L % 3) * nq;
  uint32_t* minv_n = minv3 + (size_t)((L + 1) % 3) * nq;
  uint32_t* minv_clr = minv3 + (size_t)((L + 2) % 3) * nq;
  const uint32_t* thr_prev = thr2 + (size_t)((L + 1) & 1) * nq;
  uint32_t* thr_cur = thr2 + (size_t)(L & 1) * nq;
  const uint32_t lane = __lane_id(), wave = threadIdx.x / 64, nw = blockDim.x / 64;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  // exclusive scan of the frontier sizes over the queries
  for (int base = 0; base < nq; base += (int)blockDim.x) {
    const int i = base + (int)threadIdx.x;
    const uint32_t x = i < nq ? cur[i] : 0u;
    uint32_t v = x;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t t = __shfl_up(v, d, 64);
      if (lane >= (uint32_t)d) v += t;
    }
    if (lane == 63) wsum[wave] = v;
    __syncthreads();
    uint32_t wpre = 0;
    for (uint32_t w = 0; w < wave; ++w) wpre += wsum[w];
    if (i < nq) s_off[i] = carry + wpre + v - x;
    __syncthreads();
    if (threadIdx.x == blockDim.x - 1) carry += wpre + v;
    __syncthreads();
  }
  // the level's thresholds: a query with nothing below its threshold moves
  // it past its smallest waiting distance
  for (int i = threadIdx.x; i < nq; i += blockDim.x) {
    uint32_t t = thr_prev[i];
    const uint32_t mc = minv_c[i];
    const uint32_t d = qd[i].delta;
    if (d && cur[i] && mc >= t) {
      const uint64_t u = (uint64_t)mc + d;
      t = u > 0xFFFFFFFEull ? 0xFFFFFFFEu : (uint32_t)u;
    }
    s_thr[i] = t;
    s_minv[i] = mc;
    if (blockIdx.x == 0) {
      thr_cur[i] = t;
      clr[i] = 0;
      minv_clr[i] = kNone;
    }
  }
  if (threadIdx.x == 0) s_off[nq] = carry;
  __syncthreads();
  const uint32_t total = s_off[nq];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (total == 0) bs->done = 1;
    else bs->levels_run = L + 1;
  }
  if (total == 0) return;  // uniform
  (void)nw;
  expand_level(qd, nq, s_off, total, L, s_thr, s_minv, nxt, minv_n, (int)((L + 1) & 1), gsz);
}

// Counters for the one-kernel-per-level form: level 0's set holds each
// query's root (minimum 0), the others are empty; thresholds of "level -1".
__global__ void k_batch_level_reset(const QueryDesc* __restrict__ qd, uint32_t* __restrict__ cnt3,
                                    uint32_t* __restrict__ thr2, uint32_t* __restrict__ minv3, int nq,
                                    BatchState* __restrict__ bs) {
  for (int i = threadIdx.x; i < nq; i += blockDim.x) {
    cnt3[i] = 1;
    cnt3[nq + i] = 0;
    cnt3[2 * nq + i] = 0;
    minv3[i] = 0;
    minv3[nq + i] = kNone;
    minv3[2 * nq + i] = kNone;
    thr2[nq + i] = qd[i].delta ? qd[i].delta : kNone;
  }
  if (threadIdx.x == 0) {
    bs->done = 0;
    bs->total = 0;
    bs->levels_run = 0;
    bs->base = 0;
  }
}

__global__ void k_batch_advance(BatchState* __restrict__ bs, uint32_t levels) {
  if (threadIdx.x == 0 && blockIdx.x == 0) bs->base += levels;
}

__global__ void k_batch_reached(const QueryDesc* __restrict__ qd, int q, unsigned long long* __restrict__ acc) {
  const QueryDesc& Q = qd[q];
  unsigned long long sum = 0;
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < Q.V;
       v += (uint64_t)gridDim.x * blockDim.x)
    if (Q.val[v] != kUnreached) sum += Q.off[v + 1] - Q.off[v];
  for (int d = 32; d >= 1; d >>= 1) sum += __shfl_xor(sum, d, 64);
  if (__lane_id() == 0 && sum) atomicAdd(acc + (q % 64) * 16, sum);
}

#define BATCH_LAUNCH_OK()                                                                 \
  do {                                                                                    \
    hipError_t e_ = hipGetLastError();                                                    \
    if (e_ != hipSuccess)                                                                 \
      return set_error(CG_ERR_HIP, "batch kernel launch: %s (%s:%d)", hipGetErrorString(e_), \
                       __FILE__, __LINE__);                                               \
  } while (0)

struct GraphExec {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~GraphExec() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

// Batched queries on a graph of at least 2^20 vertices whose top 1/64 of
// vertices send at least half of the edges run on a copy renumbered by
// descending out-degree (the rule an engine applies to its own graph): the
// values those edges touch then share cache lines.  Built once per graph, on
// its first batched use; the copy (CSR, weights and the id map, about the
// size of the graph again) stays in HBM until the graph is destroyed.  A
// failed build is reported by that batch and tried again by the next one.
constexpr uint32_t kBatchRelabelMinV = 1u << 20;
std::mutex batch_relabel_mu;  // first uses of one graph from several threads

static int build_batch_relabel(cg_graph* g) {
  std::lock_guard<std::mutex> lock(batch_relabel_mu);
  if (g->batch_relabel_done) return CG_OK;
  DevBuf<uint32_t> perm, new_of;
  double share = 0.0;
  CG_TRY(degree_order(g, perm, new_of, &share));
  if (share < 0.5) {
    g->batch_relabel_done = true;  // not skewed: the batch runs on the graph itself
    return CG_OK;
  }
  std::unique_ptr<cg_graph> rg(new (std::nothrow) cg_graph());
  if (!rg) return set_error(CG_ERR_OOM, "renumbered graph object");
  CG_TRY(relabel_graph(g, new_of.p, rg.get()));
  std::vector<uint32_t> h(g->V);
  CG_HIP(hipMemcpy(h.data(), new_of.p, (size_t)g->V * 4, hipMemcpyDeviceToHost));
  CG_HIP(hipDeviceSynchronize());
  g->batch_new_of_h = std::move(h);
  g->batch_new_of = std::move(new_of);
  g->batch_rg = std::move(rg);
  g->batch_relabel_done = true;
  return CG_OK;
}

// Queries on a small-world graph of at least batch_engine_v vertices run one
// by one on an engine of their own graph -- renumbering, hub deferral, degree
// tiers, direction-optimising BFS: what one query of a large skewed graph
// runs best with -- from a host thread of their own, while the other queries
// (lattices and small graphs, whose many shallow levels are what batching
// amortises) advance level by level on the batch's stream.  The values of
// each engine query are copied on the engine's stream into a slot of the
// batch before the engine takes its next root.
constexpr uint32_t kBatchEngineMinV = 1u << 18;

// Whether the batch hands g's queries to an engine: at least min_v vertices
// and a largest out-degree of at least 16 times the average, the mark of a
// small-world graph (few, wide levels; a lattice or a road network has many
// narrow ones, which is what the level-synchronous batch amortises).
static bool batch_on_engine(const cg_graph* g, uint32_t min_v) {
  if (min_v == 0xFFFFFFFFu || g->V < min_v || !g->V) return false;
  const uint64_t avg = (g->E + g->V - 1) / g->V;
  return g->max_outdeg >= 16 * (avg ? avg : 1);
}

// The graph's engine for alg, created on first use and destroyed with the
// graph: GPU schedule, default tiers; SSSP defers hubs of out-degree >= 128
// while their distance is >= the frontier's smallest + max weight / 50 (the
// window the s24 tuning found for weights up to 100 scaled to the graph's
// weights: with weights up to 1000 a window of 2 took 3-4x more iterations).
static int batch_engine(cg_graph* g, int alg, cg_engine** out) {
  std::lock_guard<std::mutex> lock(batch_relabel_mu);
  auto& slot = g->batch_eng[alg == CG_SSSP ? 1 : 0];
  if (!slot) {
    cg_options eo;
    cg_options_default(&eo);
    eo.sched_mode = CG_SCHED_GPU;
    if (alg == CG_SSSP) {
      CG_TRY(ensure_narrow_weights(g, nullptr));
      eo.hub_degree = 128;
      eo.hub_delta = g->max_w / 50 > 1 ? g->max_w / 50 : 1;
    }
    cg_engine* e = nullptr;
    const int rc = cg_engine_create(g, alg, CG_COORDINATION, 1, &eo, &e);
    if (rc != CG_OK) return rc;
    slot = std::unique_ptr<cg_engine, void (*)(cg_engine*)>(e, cg_engine_destroy);
  }
  *out = slot.get();
  return CG_OK;
}

struct EngineQuery {
  cg_engine* e;
  uint32_t root;
  uint32_t* slot;  // the query's V values, in the caller's ids
};

// A graph's batch engines are shared by every batch over the graph and run
// one query at a time: batches running at once take turns for their engine
// queries (their level loops still overlap).
std::mutex batch_engine_run_mu;

// The engine queries in order, on the calling host thread; returns once every
// copy into a slot has finished.  Errors come back in rc / msg (the error
// text is per thread).
static void run_engine_queries(int dev, const std::vector<EngineQuery>* jobs, int* rc, std::string* msg,
                               unsigned long long* edges) {
  std::lock_guard<std::mutex> turn(batch_engine_run_mu);
  auto fail = [&](int code, const char* what) {
    if (*rc == CG_OK) {
      *rc = code;
      *msg = what;
    }
  };
  *rc = CG_OK;
  *edges = 0;
  if (hipSetDevice(dev) != hipSuccess) return fail(CG_ERR_HIP, "hipSetDevice on the engine thread failed");
  std::vector<cg_engine*> used;
  for (const EngineQuery& j : *jobs) {
    cg_run_stats st{};
    const int r = cg_engine_run(j.e, j.root, nullptr, &st);
    if (r != CG_OK) {
      fail(r, cg_last_error());
      break;
    }
    *edges += st.edges_traversed;
    const uint32_t* src = j.e->renumbered ? j.e->val_out.p : j.e->val.p;
    if (hipMemcpyAsync(j.slot, src, (size_t)j.e->caller_g->V * 4, hipMemcpyDeviceToDevice, j.e->stream) !=
        hipSuccess) {
      fail(CG_ERR_HIP, "copy of an engine query's values failed");
      break;
    }
    if (std::find(used.begin(), used.end(), j.e) == used.end()) used.push_back(j.e);
  }
  for (cg_engine* e : used)
    if (hipStreamSynchronize(e->stream) != hipSuccess) fail(CG_ERR_HIP, "engine stream failed");
}

int run_batched(cg_graph* const* graphs, const int* algs, const uint32_t* roots, int nq, uint32_t* const* outs,
                const cg_options& opt, cg_run_stats* rs) {
  CG_TRY(require_device());
  const uint32_t block = opt.batch_block ? opt.batch_block : 256;
  const uint32_t grid = opt.batch_grid ? opt.batch_grid : 2048;
  const uint32_t levels = opt.batch_levels ? opt.batch_levels : 16;
  const uint32_t gsz = opt.batch_group ? opt.batch_group : 32;
  const uint32_t min_v = opt.batch_engine_v ? opt.batch_engine_v : kBatchEngineMinV;
  if (gsz > 64) return set_error(CG_ERR_INVALID, "batch_group must be in [1, 64]");
  if (block % 64 || block > 1024) return set_error(CG_ERR_INVALID, "batch_block must be a multiple of 64, <= 1024");
  for (int i = 0; i < nq; ++i) {
    if (!graphs[i]) return set_error(CG_ERR_INVALID, "query %d has no graph", i);
    if (algs[i] != CG_BFS && algs[i] != CG_SSSP)
      return set_error(CG_ERR_UNSUPPORTED, "query %d: batched queries are BFS or SSSP", i);
    if (algs[i] == CG_SSSP && !graphs[i]->has_w) return set_error(CG_ERR_INVALID, "query %d: SSSP needs weights", i);
    if (roots[i] >= graphs[i]->V) return set_error(CG_ERR_INVALID, "query %d: root %u >= V", i, roots[i]);
    CG_TRY(graphs[i]->ensure_device());
  }
  // the queries engines run (jq, in order) and those the batch runs (bq)
  std::vector<int> bq, jq;
  std::vector<EngineQuery> jobs;
  for (int i = 0; i < nq; ++i) {
    if (batch_on_engine(graphs[i], min_v)) {
      cg_engine* e = nullptr;
      CG_TRY(batch_engine(graphs[i], algs[i], &e));
      jobs.push_back(EngineQuery{e, roots[i], nullptr});
      jq.push_back(i);
    } else {
      bq.push_back(i);
    }
  }
  uint64_t slot_v = 0;
  for (int i : jq) slot_v += graphs[i]->V;
  DevBuf<uint32_t> slots;
  if (slot_v) CG_TRY(slots.ensure(slot_v));
  std::vector<uint32_t*> jslot(jq.size());
  {
    uint64_t o = 0;
    for (size_t k = 0; k < jq.size(); ++k) {
      jslot[k] = jobs[k].slot = slots.p + o;
      o += graphs[jq[k]]->V;
    }
  }
  const int nb = (int)bq.size();
  // the graph each batched query runs on, its root there, and the map back
  std::vector<cg_graph*> qgraph(nb);
  std::vector<uint32_t> qroot(nb);
  std::vector<const uint32_t*> back(nb, nullptr);
  uint32_t back_v = 0;
  for (int k = 0; k < nb; ++k) {
    cg_graph* g = graphs[bq[k]];
    qgraph[k] = g;
    qroot[k] = roots[bq[k]];
    if (g->V >= kBatchRelabelMinV) CG_TRY(build_batch_relabel(g));
    if (g->batch_rg) {
      qgraph[k] = g->batch_rg.get();
      qroot[k] = g->batch_new_of_h[roots[bq[k]]];
      back[k] = g->batch_new_of.p;
      back_v = std::max(back_v, g->V);
    }
  }
  hipStream_t s = nullptr;
  CG_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  std::unique_ptr<std::remove_pointer<hipStream_t>::type, decltype(&hipStreamDestroy)> sguard(s, hipStreamDestroy);
  // per-query device arrays: values, stamps and two frontier queues
  std::vector<uint64_t> vpre(nb + 1, 0);
  for (int k = 0; k < nb; ++k) vpre[k + 1] = vpre[k] + qgraph[k]->V;
  const uint64_t totalV = vpre[nb];
  // the batch indexes its items (vertex slots of all queries) with 32 bits
  if (totalV > 0xFFFFFFFFull)
    return set_error(CG_ERR_INVALID, "the batch holds %llu vertex slots over its queries; at most 2^32 - 1",
                     (unsigned long long)totalV);
  DevBuf<uint32_t> pool;
  std::vector<QueryDesc> hq(nb);
  const bool one_kernel = nb <= kLevelMaxQ;
  const size_t level_lds = (3 * (size_t)nb + 1) * sizeof(uint32_t);
  DevBuf<QueryDesc> dq;
  DevBuf<uint64_t> dvpre;
  DevBuf<uint32_t> cnt, item_off, qs;
  DevBuf<BatchState> bs;
  GraphExec ge;
  if (nb) {
    CG_TRY(pool.ensure(4 * totalV));
    for (int k = 0; k < nb; ++k) {
      cg_graph* g = qgraph[k];
      const int alg = algs[bq[k]];
      uint32_t* b = pool.p + 4 * vpre[k];
      const uint32_t delta = alg == CG_SSSP ? opt.sssp_delta : 0;
      const uint32_t hub_deg = alg == CG_SSSP ? opt.hub_degree : 0;
      const uint32_t hub_delta = opt.hub_delta ? opt.hub_delta : kHubDeltaDefault;
      hq[k] = QueryDesc{g->d_off.p, g->d_dst.p, g->has_w ? g->d_w.p : nullptr, b, b + g->V,
                        {b + 2 * (uint64_t)g->V, b + 3 * (uint64_t)g->V}, g->V, alg, qroot[k], delta, hub_deg,
                        hub_delta};
    }
    CG_TRY(dq.ensure(nb));
    CG_TRY(dvpre.ensure(nb + 1));
    CG_TRY(cnt.ensure(3 * (uint64_t)nb));
    CG_TRY(qs.ensure(3 * (uint64_t)nb));
    CG_TRY(item_off.ensure(2 * (uint64_t)nb + 1));  // one kernel per level: the two threshold sets
    CG_TRY(bs.ensure(1));
    CG_HIP(hipMemcpyAsync(dq.p, hq.data(), nb * sizeof(QueryDesc), hipMemcpyHostToDevice, s));
    CG_HIP(hipMemcpyAsync(dvpre.p, vpre.data(), (nb + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    CG_HIP(hipStreamSynchronize(s));

    // capture `levels` levels once: one kernel per level (counters, minima
    // and thresholds in cnt / qs / item_off as three, three and two sets)
    // up to kLevelMaxQ queries, a planning and an expansion kernel beyond
    CG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (uint32_t l = 0; l < levels; ++l) {
      if (one_kernel) {
        hipLaunchKernelGGL(k_batch_level, dim3(grid), dim3(block), level_lds, s, dq.p, nb, cnt.p, item_off.p, qs.p,
                           bs.p, gsz, l);
      } else {
        hipLaunchKernelGGL(k_batch_plan, dim3(1), dim3(kPlanBlock), 0, s, dq.p, cnt.p, qs.p, nb, item_off.p, bs.p);
        hipLaunchKernelGGL(k_batch_expand, dim3(grid), dim3(block), 0, s, dq.p, nb, item_off.p, cnt.p, qs.p, qs.p,
                           bs.p, gsz);
      }
    }
    if (one_kernel) hipLaunchKernelGGL(k_batch_advance, dim3(1), dim3(64), 0, s, bs.p, levels);
    hipError_t ce = hipStreamEndCapture(s, &ge.graph);
    if (ce != hipSuccess) return set_error(CG_ERR_HIP, "graph capture: %s", hipGetErrorString(ce));
    CG_HIP(hipGraphInstantiate(&ge.exec, ge.graph, nullptr, nullptr, 0));
  }
  int dev = 0;
  CG_HIP(hipGetDevice(&dev));

  // ---- timed: the engine thread and the level-synchronous batch together ----
  BatchState hb{};
  uint32_t replays = 0;
  auto batch_loop = [&]() -> int {
    const unsigned ig = (unsigned)std::min<uint64_t>((totalV + 255) / 256, 65536);
    hipLaunchKernelGGL(k_batch_init, dim3(ig ? ig : 1), dim3(256), 0, s, dq.p, nb, dvpre.p, totalV);
    if (one_kernel)
      hipLaunchKernelGGL(k_batch_level_reset, dim3(1), dim3(1024), 0, s, dq.p, cnt.p, item_off.p, qs.p, nb, bs.p);
    else
      hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(1024), 0, s, dq.p, cnt.p, qs.p, nb, bs.p);
    BATCH_LAUNCH_OK();
    for (;;) {
      CG_HIP(hipGraphLaunch(ge.exec, s));
      ++replays;
      CG_HIP(hipMemcpyAsync(&hb, bs.p, sizeof(hb), hipMemcpyDeviceToHost, s));
      CG_HIP(hipStreamSynchronize(s));
      if (hb.done) return CG_OK;
      if (replays > (1u << 22)) return set_error(CG_ERR_INTERNAL, "batch did not converge");
    }
  };
  const auto t0 = std::chrono::steady_clock::now();
  int wrc = CG_OK;
  std::string wmsg;
  unsigned long long wedges = 0;
  std::thread worker;
  if (!jobs.empty()) worker = std::thread(run_engine_queries, dev, &jobs, &wrc, &wmsg, &wedges);
  const int brc = nb ? batch_loop() : CG_OK;
  if (worker.joinable()) worker.join();
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (brc != CG_OK) return brc;
  if (wrc != CG_OK) return set_error(wrc, "batched query on an engine: %s", wmsg.c_str());

  // TEPS numerator per query and the outputs, outside the timed region
  DevBuf<unsigned long long> acc;
  std::vector<unsigned long long> hacc(64 * 16, 0ull);
  if (nb) {
    CG_TRY(acc.ensure(64 * 16));
    CG_HIP(hipMemsetAsync(acc.p, 0, 64 * 16 * sizeof(unsigned long long), s));
    for (int k = 0; k < nb; ++k) {
      const unsigned bg = (unsigned)std::min<uint64_t>(((uint64_t)hq[k].V + 255) / 256, 2048);
      hipLaunchKernelGGL(k_batch_reached, dim3(bg ? bg : 1), dim3(256), 0, s, dq.p, k, acc.p);
    }
    BATCH_LAUNCH_OK();
    CG_HIP(hipMemcpyAsync(hacc.data(), acc.p, hacc.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  }
  DevBuf<uint32_t> unmapped;  // values of a renumbered query in the caller's ids
  if (outs && back_v) CG_TRY(unmapped.ensure(back_v));
  for (int k = 0; k < nb && outs; ++k) {
    if (!outs[bq[k]]) continue;
    const uint32_t* src = hq[k].val;
    if (back[k]) {
      // stream order: the previous query's copy out of `unmapped` is done
      CG_TRY(launch_renumber_back(back[k], hq[k].val, hq[k].V, unmapped.p, s));
      src = unmapped.p;
    }
    CG_HIP(hipMemcpyAsync(outs[bq[k]], src, (size_t)hq[k].V * 4, hipMemcpyDeviceToHost, s));
  }
  for (size_t k = 0; k < jq.size() && outs; ++k)
    if (outs[jq[k]])
      CG_HIP(hipMemcpyAsync(outs[jq[k]], jslot[k], (size_t)graphs[jq[k]]->V * 4, hipMemcpyDeviceToHost, s));
  CG_HIP(hipStreamSynchronize(s));
  if (rs) {
    std::memset(rs, 0, sizeof(*rs));
    rs->ms = ms;
    rs->iterations = hb.levels_run;
    rs->expand_launches = replays;
    unsigned long long m = wedges;
    for (auto x : hacc) m += x;
    rs->edges_traversed = m;
    rs->heavy_iterations = (uint32_t)jq.size();  // queries run on engines
  }
  return CG_OK;
}

int run_one_per_launch(cg_graph* const* graphs, const int* algs, const uint32_t* roots, int nq,
                       uint32_t* const* outs, const cg_options& opt, cg_run_stats* rs) {
  cg_options o = opt;
  o.sched_mode = CG_SCHED_GPU;
  o.trace_path = nullptr;
  double ms = 0;
  uint64_t m = 0;
  uint32_t iters = 0;
  for (int i = 0; i < nq; ++i) {
    cg_engine* e = nullptr;
    CG_TRY(cg_engine_create(graphs[i], algs[i], CG_COORDINATION, 1, &o, &e));
    std::unique_ptr<cg_engine, decltype(&cg_engine_destroy)> eg(e, cg_engine_destroy);
    cg_run_stats st{};
    CG_TRY(cg_engine_run(e, roots[i], nullptr, &st));
    ms += st.ms;
    m += st.edges_traversed;
    iters += st.iterations;
    if (outs && outs[i]) CG_TRY(cg_engine_result(e, outs[i]));
  }
  if (rs) {
    std::memset(rs, 0, sizeof(*rs));
    rs->ms = ms;
    rs->edges_traversed = m;
    rs->iterations = iters;
    rs->expand_launches = (uint32_t)nq;
  }
  return CG_OK;
}

}  // namespace

}  // namespace cg

extern "C" int cg_batch_run(cg_graph* const* graphs, const int* algorithms, const uint32_t* roots, int nq,
                            uint32_t* const* outs, int mode, const cg_options* opt, cg_run_stats* stats) {
  cg::clear_error();
  if (!graphs || !algorithms || !roots || nq <= 0) return cg::set_error(CG_ERR_INVALID, "bad batch arguments");
  cg_options o;
  if (opt) o = *opt;
  else cg_options_default(&o);
  if (mode == 0) return cg::run_batched(graphs, algorithms, roots, nq, outs, o, stats);
  if (mode == 1) return cg::run_one_per_launch(graphs, algorithms, roots, nq, outs, o, stats);
  return cg::set_error(CG_ERR_INVALID, "batch mode %d", mode);
}

// Engine-handle form: query i runs engine i's algorithm on engine i's graph;
// the batch takes engines[0]'s options (block, grid, levels, SSSP delta).
extern "C" int cg_engine_run_batch(cg_engine* const* engines, const uint32_t* roots, int nq, void* const* outs,
                                   cg_run_stats* stats) {
  cg::clear_error();
  if (!engines || !roots || nq <= 0) return cg::set_error(CG_ERR_INVALID, "bad batch arguments");
  std::vector<cg_graph*> graphs(nq);
  std::vector<int> algs(nq);
  std::vector<uint32_t*> o(nq, nullptr);
  for (int i = 0; i < nq; ++i) {
    if (!engines[i]) return cg::set_error(CG_ERR_INVALID, "engine %d is NULL", i);
    if (engines[i]->host_only)
      return cg::set_error(CG_ERR_UNSUPPORTED, "engine %d runs without a GPU; batched queries run on the GPU", i);
    if (engines[i]->alg != CG_BFS && engines[i]->alg != CG_SSSP)
      return cg::set_error(CG_ERR_UNSUPPORTED, "engine %d: batched queries are BFS or SSSP", i);
    graphs[i] = engines[i]->caller_g;
    algs[i] = engines[i]->alg;
    if (outs) o[i] = static_cast<uint32_t*>(outs[i]);
  }
  return cg::run_batched(graphs.data(), algs.data(), roots, nq, o.data(), engines[0]->opt, stats);
}
