// COORDINATION engine: the per-iteration loop of CGgraphEngine::CG_co_execution
// (my_project/CGgraph/CGgraph_engine.cuh:265-974) redesigned for one MI355X.
//
// Reference flow per iteration: classify LIGHT/HEAVY (:312-348); LIGHT runs
// on the CPU team; HEAVY copies the whole value array to the GPU (:467-473),
// lets a GPU agent consume a segment queue from the head while the CPU team
// steals from the tail (:692-855), copies the GPU replica back (:1636) and
// min-merges it on the host (:1657-1727).
//
// Here the values and both frontier bitmaps live in HBM for the whole run.
// Per iteration:
//   1. k_stats reads the frontier bitmap (popcount, edge sum, degree
//      histogram) and zeroes the next bitmap; the stats come back in one
//      small copy, which is the only synchronisation of the iteration.
//   2. the scheduler picks push / pull, degree tiers and the side;
//   3. GPU side: k_compact + k_expand (push) or k_pull_bfs (pull);
//      host side: the frontier's (vertex, value) pairs go to the host team,
//      whose candidates come back and are applied with atomic min; in a
//      co-executed iteration the host takes the tail of the small-degree queue
//      while the GPU expands the rest.
// Because every path applies the same monotone relaxation until no value
// changes, every schedule ends at the same unique fixed point.
#include "engine.hpp"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace cg {

namespace {

using clk = std::chrono::steady_clock;

double ms_since(clk::time_point t0) {
  return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
}

int ensure_team(cg_engine* e) {
  if (e->team_ready) return CG_OK;
  CG_TRY(e->g->ensure_host());
  e->team.init(e->g->V, e->opt.cpu_threads);
  e->team_ready = true;
  return CG_OK;
}

// Reference heuristic (CGgraph_engine.cuh:312-348).
bool heuristic_heavy(const cg_engine* e, const DevStats& s) {
  const uint64_t Vp = (uint64_t)e->g->V - e->g->zero_tail;
  if (s.n_active == 1) return false;
  if (s.n_active == Vp && e->g->E >= e->opt.heavy_active_edges) return true;
  return s.n_active >= e->opt.heavy_active_vertices && s.m_active >= e->opt.heavy_active_edges;
}

// Direction-optimising rule for BFS (Beamer, Asanovic, Patterson, SC'12):
// go bottom-up when the frontier's edges exceed the unexplored edges / 15,
// come back when the frontier falls below V / 18 and is shrinking.
bool beamer_pull(bool was_pull, const DevStats& s, uint64_t m_unexplored, uint32_t V,
                 uint32_t prev_active) {
  if (!was_pull) return (double)s.m_active > (double)m_unexplored / 15.0;
  return !((double)s.n_active < (double)V / 18.0 && s.n_active < prev_active);
}

struct IterRecord {
  uint32_t iter;
  int choice;
  int host_share;  // 0 none, 1 host only, 2 co-executed
  double us;
  DevStats s;
  float feat[CG_NUM_FEATURES];
};

void write_trace(cg_engine* e, const IterRecord& r, uint32_t root) {
  if (!e->trace) return;
  std::fprintf(e->trace,
               "{\"alg\":%d,\"V\":%u,\"E\":%llu,\"root\":%u,\"iter\":%u,\"choice\":%d,\"host\":%d,"
               "\"us\":%.3f,\"n_active\":%u,\"m_active\":%llu,\"features\":[",
               e->alg, e->g->V, (unsigned long long)e->g->E, root, r.iter, r.choice, r.host_share,
               r.us, r.s.n_active, (unsigned long long)r.s.m_active);
  for (int k = 0; k < CG_NUM_FEATURES; ++k)
    std::fprintf(e->trace, "%s%.6g", k ? "," : "", (double)r.feat[k]);
  std::fprintf(e->trace, "]}\n");
}

// Host side of one iteration: the frontier (or its small-tier tail) goes to
// the host team; candidates come back and are applied on the GPU.
int host_iteration(cg_engine* e, const DevGraph& dg, uint64_t* in, uint64_t* out, DevStats* dst_st,
                   uint32_t take_from, uint32_t n_take, cg_run_stats* rs) {
  hipStream_t s = e->stream;
  if (n_take) {
    CG_TRY(e->d_share_vals.ensure(n_take));
    CG_TRY(launch_gather_values(e->q_small.p + take_from, n_take, e->val.p, e->d_share_vals.p, s));
    CG_TRY(e->p_ids.ensure(n_take));
    CG_TRY(e->p_vals.ensure(n_take));
    CG_HIP(hipMemcpyAsync(e->p_ids.p, e->q_small.p + take_from, n_take * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, s));
    CG_HIP(hipMemcpyAsync(e->p_vals.p, e->d_share_vals.p, n_take * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, s));
    CG_HIP(hipEventRecord(e->ev_copy, s));
  }
  (void)in;
  (void)out;
  (void)dg;
  (void)dst_st;
  (void)rs;
  return CG_OK;
}

}  // namespace

int engine_run(cg_engine* e, uint32_t root, cg_run_stats* rs) {
  cg_graph& g = *e->g;
  const uint32_t V = g.V;
  const uint64_t words = ((uint64_t)V + 63) / 64;
  hipStream_t s = e->stream;
  const int alg = e->alg;
  const int mode = e->opt.sched_mode;
  const bool need_pull = alg == CG_BFS && (mode == CG_SCHED_GPU || mode == CG_SCHED_LEARNED ||
                                           mode == CG_SCHED_FORCED);
  if (need_pull) CG_TRY(build_transpose(&g));
  const bool host_possible = mode == CG_SCHED_HEURISTIC || mode == CG_SCHED_LEARNED ||
                             (mode == CG_SCHED_FORCED && decode_choice(e->opt.forced_choice).host);
  if (host_possible) CG_TRY(ensure_team(e));

  DevGraph dg{g.d_off.p, g.d_dst.p, g.has_w ? g.d_w.p : nullptr,
              g.dev_t_valid ? g.d_in_off.p : nullptr, g.dev_t_valid ? g.d_in_src.p : nullptr, V};
  if (alg == CG_SSSP && !g.has_w) return set_error(CG_ERR_INVALID, "SSSP needs edge weights");

  cg_run_stats local{};
  const auto t0 = clk::now();
  // ---- initAlgorithm (CGgraph_engine.cuh:1259-1316) ----
  CG_HIP(hipMemsetAsync(e->bm[0].p, 0, words * sizeof(uint64_t), s));
  CG_HIP(hipMemsetAsync(e->bm[1].p, 0, words * sizeof(uint64_t), s));
  uint64_t visited;
  if (alg == CG_CC) {
    CG_TRY(launch_iota_u32(e->val.p, V, s));
    CG_TRY(launch_bitmap_prefix(e->bm[0].p, words, (uint64_t)V - g.zero_tail, s));
    visited = V;
  } else {
    CG_TRY(launch_fill_u32(e->val.p, kUnreached, V, s));
    CG_TRY(launch_set_root(e->val.p, e->bm[0].p, root, s));
    visited = 1;
  }
  if (host_possible) e->team.reset_values(alg, root);

  int cur = 0;
  uint32_t it = 0, prev_active = 0;
  bool pull = false;
  uint64_t m_explored = 0;
  const Bins base_bins{e->opt.bin_wave, e->opt.bin_chunk, e->opt.chunk_edges};
  bool timing_pending = false;
  IterRecord rec{};
  auto tit = clk::now();
  for (;;) {
    ++it;
    const int slot = it & 1;
    DevStats* dst = e->st.p + slot;
    CG_HIP(hipMemsetAsync(dst, 0, sizeof(DevStats), s));
    CG_TRY(launch_stats(e->bm[cur].p, e->bm[cur ^ 1].p, words, dg, dst, s));
    CG_HIP(hipMemcpyAsync(e->hst.p, e->st.p, 2 * sizeof(DevStats), hipMemcpyDeviceToHost, s));
    CG_HIP(hipStreamSynchronize(s));
    if (e->opt.fault_flags & CG_FAULT_HIP_ERROR)
      return set_error(CG_ERR_HIP, "injected HIP failure at iteration %u", it);
    const DevStats sst = e->hst.p[slot];
    const DevStats prev = e->hst.p[slot ^ 1];
    if (it > 1) {
      visited += prev.n_new;
      if (timing_pending && e->opt.profile) {
        float ms = 0.f;
        CG_HIP(hipEventElapsedTime(&ms, e->ev0, e->ev1));
        local.expand_kernel_ms += ms;
        timing_pending = false;
      }
      rec.us = std::chrono::duration<double, std::micro>(clk::now() - tit).count();
      write_trace(e, rec, root);
    }
    tit = clk::now();
    if (sst.n_active == 0) break;

    // ---- schedule ----
    const uint64_t m_unexplored = g.E > m_explored ? g.E - m_explored : 0;
    Choice ch{0, 0, 0};
    int host_share = 0;  // 0 none, 1 host only, 2 co-executed
    if (mode == CG_SCHED_HEURISTIC) {
      const bool heavy = heuristic_heavy(e, sst);
      host_share = heavy ? 2 : 1;
    } else if (mode == CG_SCHED_GPU) {
      if (alg == CG_BFS) pull = beamer_pull(pull, sst, m_unexplored, V, prev_active);
      ch.pull = pull;
    } else if (mode == CG_SCHED_FORCED) {
      ch = decode_choice(e->opt.forced_choice);
      host_share = ch.host ? 1 : 0;
    } else {  // learned: decided further below once the MLP is wired in
      if (alg == CG_BFS) pull = beamer_pull(pull, sst, m_unexplored, V, prev_active);
      ch.pull = pull;
    }
    if (alg != CG_BFS) ch.pull = 0;
    if (ch.pull && host_share == 1) host_share = 0;  // pull steps run on the GPU
    pull = ch.pull;
    Bins bins = base_bins;
    if (mode == CG_SCHED_FORCED || mode == CG_SCHED_LEARNED) {
      bins.bin_wave = kTierSets[ch.tier].bin_wave;
      bins.bin_chunk = kTierSets[ch.tier].bin_chunk;
    }

    rec.iter = it;
    rec.choice = encode_choice(ch);
    rec.host_share = host_share;
    rec.s = sst;
    make_features(sst, V, g.E, visited, it, rec.feat);
    m_explored += sst.m_active;
    prev_active = sst.n_active;

    uint64_t* in = e->bm[cur].p;
    uint64_t* out = e->bm[cur ^ 1].p;
    if (ch.pull) {
      ++local.pull_iterations;
      ++local.heavy_iterations;
      local.gpu_vertices += sst.n_active;
      if (e->opt.profile) CG_HIP(hipEventRecord(e->ev0, s));
      CG_TRY(launch_pull_bfs(dg, e->val.p, in, out, dst, s));
      if (e->opt.profile) { CG_HIP(hipEventRecord(e->ev1, s)); timing_pending = true; }
      local.edges_relaxed += sst.m_active;
    } else if (host_share == 1) {
      // ---- LIGHT: the host team relaxes the whole frontier ----
      ++local.light_iterations;
      const Bins all_small{0xFFFFFFFFu, 0xFFFFFFFFu, e->opt.chunk_edges};
      CG_TRY(launch_compact(in, words, dg, all_small, e->q_small.p, e->q_med.p, e->q_chunk.p, dst, s));
      const uint32_t n = sst.n_active - sst.n_zero_deg;
      CG_TRY(e->d_share_vals.ensure(n ? n : 1));
      CG_TRY(launch_gather_values(e->q_small.p, n, e->val.p, e->d_share_vals.p, s));
      CG_TRY(e->p_ids.ensure(n ? n : 1));
      CG_TRY(e->p_vals.ensure(n ? n : 1));
      if (n) {
        CG_HIP(hipMemcpyAsync(e->p_ids.p, e->q_small.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        CG_HIP(hipMemcpyAsync(e->p_vals.p, e->d_share_vals.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
      }
      CG_HIP(hipStreamSynchronize(s));
      local.bytes_d2h += 8ull * n;
      uint64_t edges = 0;
      const auto th = clk::now();
      const uint64_t nc = e->team.relax(alg, &g, e->p_ids.p, e->p_vals.p, nullptr, nullptr, n,
                                        e->cand_id, e->cand_val, &edges);
      const double sec = ms_since(th) * 1e-3;
      if (edges && sec > 0) e->rate_cpu = 0.7 * e->rate_cpu + 0.3 * (edges / sec);
      local.edges_relaxed += edges;
      local.cpu_vertices += n;
      if (nc) {
        CG_TRY(e->d_cand_id.ensure(nc));
        CG_TRY(e->d_cand_val.ensure(nc));
        CG_HIP(hipMemcpyAsync(e->d_cand_id.p, e->cand_id.data(), nc * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        CG_HIP(hipMemcpyAsync(e->d_cand_val.p, e->cand_val.data(), nc * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        local.bytes_h2d += 8ull * nc;
        CG_TRY(launch_apply_candidates(e->d_cand_id.p, e->d_cand_val.p, (uint32_t)nc, e->val.p, out, dst, s));
        // the values just shipped and synchronised make the candidate buffers reusable
        CG_HIP(hipStreamSynchronize(s));
      }
    } else {
      // ---- GPU push, optionally co-executed with the host team ----
      ++local.heavy_iterations;
      CG_TRY(launch_compact(in, words, dg, bins, e->q_small.p, e->q_med.p, e->q_chunk.p, dst, s));
      uint32_t small_limit = 0xFFFFFFFFu, n_host = 0, n_small = 0;
      if (host_share == 2) {
        // Split the small tier by measured throughput (analytical split in the
        // spirit of GraphCooperation.cuh:1036-1086): host share of edges
        // = R_cpu / (R_cpu + R_gpu), taken from the queue tail.
        DevStats cnt;
        CG_HIP(hipMemcpyAsync(e->hst.p + 2, dst, sizeof(DevStats), hipMemcpyDeviceToHost, s));
        CG_HIP(hipStreamSynchronize(s));
        cnt = e->hst.p[2];
        n_small = cnt.n_small;
        double rg = e->rate_gpu;
        if (e->opt.fault_flags & CG_FAULT_SLOW_GPU) rg = e->rate_cpu * 0.25;
        const double frac = e->rate_cpu / (e->rate_cpu + rg);
        n_host = (uint32_t)std::min<double>(n_small, std::floor(frac * n_small));
        small_limit = n_small - n_host;
        if (n_host) CG_TRY(host_iteration(e, dg, in, out, dst, small_limit, n_host, &local));
      }
      const uint64_t bound = expand_item_bound(sst.n_active, sst.m_active, bins);
      if (e->opt.profile) CG_HIP(hipEventRecord(e->ev0, s));
      CG_TRY(launch_expand_push(alg, dg, e->val.p, out, e->q_small.p, e->q_med.p, e->q_chunk.p, bound,
                                small_limit, bins, dst, s));
      if (e->opt.profile) { CG_HIP(hipEventRecord(e->ev1, s)); timing_pending = true; }
      ++local.expand_launches;
      local.expand_alg_bytes += 20.0 * (sst.n_active - sst.n_zero_deg) +
                                (alg == CG_SSSP ? 12.0 : 8.0) * (double)sst.m_active;
      local.gpu_vertices += sst.n_active - n_host;
      local.edges_relaxed += sst.m_active;
      if (n_host) {
        // host share: its (id, value) copies were queued before the expansion
        CG_HIP(hipEventSynchronize(e->ev_copy));
        uint64_t edges = 0;
        const auto th = clk::now();
        const uint64_t nc = e->team.relax(alg, &g, e->p_ids.p, e->p_vals.p, nullptr, nullptr, n_host,
                                          e->cand_id, e->cand_val, &edges);
        const double sec = ms_since(th) * 1e-3;
        if (edges && sec > 0) e->rate_cpu = 0.7 * e->rate_cpu + 0.3 * (edges / sec);
        local.cpu_vertices += n_host;
        local.bytes_d2h += 8ull * n_host;
        if (nc) {
          CG_TRY(e->d_cand_id.ensure(nc));
          CG_TRY(e->d_cand_val.ensure(nc));
          CG_HIP(hipMemcpyAsync(e->d_cand_id.p, e->cand_id.data(), nc * sizeof(uint32_t), hipMemcpyHostToDevice, s));
          CG_HIP(hipMemcpyAsync(e->d_cand_val.p, e->cand_val.data(), nc * sizeof(uint32_t), hipMemcpyHostToDevice, s));
          local.bytes_h2d += 8ull * nc;
          CG_TRY(launch_apply_candidates(e->d_cand_id.p, e->d_cand_val.p, (uint32_t)nc, e->val.p, out, dst, s));
          CG_HIP(hipStreamSynchronize(s));
        }
      }
    }
    cur ^= 1;
  }
  CG_HIP(hipStreamSynchronize(s));
  local.iterations = it - 1;
  local.ms = ms_since(t0);
  e->cur = cur;
  // TEPS numerator, outside the timed region
  if (alg != CG_CC) {
    CG_HIP(hipMemsetAsync(e->acc.p, 0, sizeof(unsigned long long), s));
    CG_TRY(launch_reached_edges(dg, e->val.p, e->acc.p, s));
    unsigned long long m = 0;
    CG_HIP(hipMemcpyAsync(&m, e->acc.p, sizeof(m), hipMemcpyDeviceToHost, s));
    CG_HIP(hipStreamSynchronize(s));
    local.edges_traversed = m;
  } else {
    local.edges_traversed = g.E;
  }
  if (rs) *rs = local;
  return CG_OK;
}

}  // namespace cg

cg_engine::~cg_engine() {
  if (plan) cg::pr_plan_destroy(plan);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (ev_copy) (void)hipEventDestroy(ev_copy);
  if (stream) (void)hipStreamDestroy(stream);
  if (trace) std::fclose(trace);
}
