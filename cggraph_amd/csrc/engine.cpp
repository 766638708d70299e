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
//      histogram; in learned mode it also runs the scheduler network) and
//      zeroes the next bitmap; the stats come back in one small copy, which is
//      the only synchronisation of a GPU iteration.
//   2. the scheduler picks push / pull, degree tiers and the side;
//   3. GPU side: k_compact + k_expand (push) or k_pull_bfs (pull);
//      host side: the frontier's (vertex, value) pairs go to the host team,
//      whose candidates come back and are applied with atomic min; in a
//      co-executed iteration the host takes the tail of the small-degree queue
//      while the GPU expands the rest.
// Every path applies the same monotone relaxation until no value changes, so
// every schedule ends at the same unique fixed point.
#include "engine.hpp"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <string>

namespace cg {

namespace {

using clk = std::chrono::steady_clock;

// BFS on the GPU schedule: a frontier with fewer out-edges than this, not
// larger than the one before, ends the queueing ahead (see engine_run)
constexpr unsigned long long kBfsTailEdges = 256;
// fused SSSP / CC: a frontier below kTailEdges edges that did not grow has the
// expansions queued after it launched with at most kTailGrid workgroups
constexpr unsigned long long kTailEdges = 1ull << 21;
constexpr uint32_t kTailGrid = 2048;

double ms_since(clk::time_point t0) {
  return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
}

int ensure_team(cg_engine* e) {
  if (e->team_ready) return CG_OK;
  CG_TRY(e->host_g()->ensure_host());
  e->team.init(e->host_g()->V, e->opt.cpu_threads);
  e->team_ready = true;
  return CG_OK;
}

// Reference heuristic (CGgraph_engine.cuh:312-348).
// V' is the caller's: a renumbered copy puts every zero-out-degree vertex at
// the end, which can be more than the caller's declared tail.
bool heuristic_heavy(const cg_engine* e, const DevStats& s) {
  const uint64_t Vp = (uint64_t)e->caller_g->V - e->caller_g->zero_tail;
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
  int learned;     // 1 when the network's choice was taken
  double us;
  DevStats s;
  float feat[CG_NUM_FEATURES];
};

void write_trace(cg_engine* e, const IterRecord& r, uint32_t root) {
  if (!e->trace) return;
  std::fprintf(e->trace,
               "{\"alg\":%d,\"V\":%u,\"E\":%llu,\"root\":%u,\"iter\":%u,\"choice\":%d,\"host\":%d,"
               "\"learned\":%d,\"us\":%.3f,\"n_active\":%u,\"m_active\":%llu,\"features\":[",
               e->alg, e->g->V, (unsigned long long)e->g->E, root, r.iter, r.choice, r.host_share,
               r.learned, r.us, r.s.n_active, (unsigned long long)r.s.m_active);
  for (int k = 0; k < CG_NUM_FEATURES; ++k)
    std::fprintf(e->trace, "%s%.6g", k ? "," : "", (double)r.feat[k]);
  std::fprintf(e->trace, "]}\n");
}

void write_probe(cg_engine* e, const IterRecord& r, uint32_t root, const double* us) {
  if (!e->trace) return;
  std::fprintf(e->trace,
               "{\"probe\":1,\"alg\":%d,\"V\":%u,\"E\":%llu,\"root\":%u,\"iter\":%u,\"n_active\":%u,"
               "\"m_active\":%llu,\"features\":[",
               e->alg, e->g->V, (unsigned long long)e->g->E, root, r.iter, r.s.n_active,
               (unsigned long long)r.s.m_active);
  for (int k = 0; k < CG_NUM_FEATURES; ++k)
    std::fprintf(e->trace, "%s%.6g", k ? "," : "", (double)r.feat[k]);
  std::fprintf(e->trace, "],\"us_all\":[");
  for (int k = 0; k < CG_NUM_CHOICES; ++k) std::fprintf(e->trace, "%s%.3f", k ? "," : "", us[k]);
  std::fprintf(e->trace, "]}\n");
}

int apply_host_candidates(cg_engine* e, uint64_t nc, uint64_t* out, cg_run_stats* rs);

void shadow_ready(cg_engine* e) {
  if (!e->shadow_stale) return;
  e->team.reset_values(e->alg, e->shadow_root);
  e->shadow_stale = false;
}

// Relaxes the host team's share (ids/vals already in pinned memory) and
// applies its candidates on the GPU.
int host_relax_and_apply(cg_engine* e, uint32_t n, uint64_t* out, cg_run_stats* rs) {
  uint64_t edges = 0;
  shadow_ready(e);
  const auto th = clk::now();
  const uint64_t nc = e->team.relax(e->alg, e->host_g(), e->p_ids.p, e->p_vals.p, n, e->cand_id, e->cand_val, &edges);
  const double sec = ms_since(th) * 1e-3;
  if (edges && sec > 0) e->rate_cpu = 0.7 * e->rate_cpu + 0.3 * ((double)edges / sec);
  rs->edges_relaxed += edges;
  rs->cpu_vertices += n;
  rs->bytes_d2h += 8ull * n;
  return apply_host_candidates(e, nc, out, rs);
}

// Copies the host team's candidates (e->cand_id / cand_val) to the GPU and
// min-applies them.
int apply_host_candidates(cg_engine* e, uint64_t nc, uint64_t* out, cg_run_stats* rs) {
  hipStream_t s = e->stream;
  if (!nc) return CG_OK;
  if (nc > 0xFFFFFFFFull) return set_error(CG_ERR_INTERNAL, "host candidate list too long");
  CG_TRY(e->d_cand_id.ensure(nc));
  CG_TRY(e->d_cand_val.ensure(nc));
  CG_HIP(hipMemcpyAsync(e->d_cand_id.p, e->cand_id.data(), nc * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  CG_HIP(hipMemcpyAsync(e->d_cand_val.p, e->cand_val.data(), nc * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  rs->bytes_h2d += 8ull * nc;
  CG_TRY(launch_apply_candidates(e->d_cand_id.p, e->d_cand_val.p, (uint32_t)nc, e->val.p, out,
                                 e->stripes.p, s));
  // the candidate vectors are pageable host memory: finish the copies before reuse
  CG_HIP(hipStreamSynchronize(s));
  return CG_OK;
}

// Queues D2H copies of q_small[from, from + n) and their values; ev_copy
// marks their completion.
int queue_share_copy(cg_engine* e, uint32_t from, uint32_t n) {
  hipStream_t s = e->stream;
  CG_TRY(e->d_share_vals.ensure(n ? n : 1));
  CG_TRY(e->p_ids.ensure(n ? n : 1));
  CG_TRY(e->p_vals.ensure(n ? n : 1));
  if (n) {
    CG_TRY(launch_gather_values(e->q_small.p + from, n, e->val.p, e->d_share_vals.p, s));
    CG_HIP(hipMemcpyAsync(e->p_ids.p, e->q_small.p + from, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    CG_HIP(hipMemcpyAsync(e->p_vals.p, e->d_share_vals.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  }
  CG_HIP(hipEventRecord(e->ev_copy, s));
  return CG_OK;
}

bool probs_usable(const DevStats& s) {
  if (s.choice < 0 || s.choice >= CG_NUM_CHOICES) return false;
  float mx = 0.f, sum = 0.f;
  for (int k = 0; k < CG_NUM_CHOICES; ++k) {
    if (!std::isfinite(s.probs[k])) return false;
    mx = std::max(mx, s.probs[k]);
    sum += s.probs[k];
  }
  // the network defers to the heuristic when it is not confident
  return std::fabs(sum - 1.f) < 1e-3f && mx >= kLearnedMinProb;
}

// Waits for k_stats of iteration `it` to publish into mapped host memory:
// a short spin on the `iter` word (written last), then a stream
// synchronisation if the kernel takes longer.  Copies the stats to hst_last.
int wait_stats(cg_engine* e, int slot, uint32_t it) {
  DevStats* h = e->hst_map.p + slot;
  const uint32_t tag = record_tag(e->epoch, it);
  const auto t0 = clk::now();
  for (uint32_t spin = 1;; ++spin) {
    if (__atomic_load_n(&h->iter, __ATOMIC_ACQUIRE) == tag) {
      std::memcpy(&e->hst_last, h, sizeof(DevStats));
      e->hst_last.iter = it;
      return CG_OK;
    }
    __builtin_ia32_pause();
    if ((spin & 1023) == 0 && ms_since(t0) > 50.0) break;  // long iteration: block instead
  }
  CG_HIP(hipStreamSynchronize(e->stream));
  if (__atomic_load_n(&h->iter, __ATOMIC_ACQUIRE) != tag)
    return set_error(CG_ERR_INTERNAL, "statistics of iteration %u were not published", it);
  std::memcpy(&e->hst_last, h, sizeof(DevStats));
  e->hst_last.iter = it;
  return CG_OK;
}

struct Step {
  Choice ch;
  int host_share;      // 0 GPU only, 1 host team only, 2 co-executed
  uint32_t threshold;  // delta-stepping: GPU pushes expand only values below it
};

// A HEAVY iteration shared with the host team (CGgraph_engine.cuh:368-855,
// 1499-1650).  The compaction has written the degree-tier queues.  The host
// receives the tail of the small tier -- twice the share that the
// COOPERATION engine's split predicts, so a slow GPU can be relieved --
// and this thread becomes the GPU agent: it claims pieces of the small tier
// from the head, half of what is left first and shrinking to a floor
// (GPU_WORK_GRANULARITY / 2^k, :1550), and launches one expansion per piece
// with at most two in flight, while the other threads of the team claim
// 64-entry pieces from the tail (the CPU steal loop, :692-855).  Neither
// share is fixed before the launch: the two ends meet wherever the faster
// side got to.  The host's candidates are applied on the GPU afterwards.
int co_execute(cg_engine* e, const DevGraph& dg, const Bins& bins, const DevStats& sst, DevStats* dst,
               uint64_t* out, cg_run_stats* local) {
  hipStream_t s = e->stream;
  const int alg = e->alg;
  CG_HIP(hipMemcpyAsync(e->hst.p + 2, dst, sizeof(DevStats), hipMemcpyDeviceToHost, s));
  CG_HIP(hipStreamSynchronize(s));
  const DevStats q = e->hst.p[2];
  const uint32_t n_small = q.n_small;
  // GPU share of the frontier's edges W (GraphCooperation.cuh:1076-1086):
  // alpha = (W - t_pcie * R_cpu) / (W * (1 + R_cpu / R_gpu))
  double rg = e->rate_gpu;
  if (e->opt.fault_flags & CG_FAULT_SLOW_GPU) rg = e->rate_cpu * 0.25;
  const double W = (double)sst.m_active;
  double alpha = W > 0 ? (W - e->pcie_ms * 1e-3 * e->rate_cpu) / (W * (1.0 + e->rate_cpu / rg)) : 1.0;
  alpha = std::min(1.0, std::max(0.0, alpha));
  const int T = e->team.threads;
  uint32_t cap = (uint32_t)std::min<double>(n_small, std::ceil(2.0 * (1.0 - alpha) * n_small));
  cap = std::min<uint32_t>(n_small, std::max<uint32_t>(cap, (uint32_t)(kVertexStealChunk * T)));
  const uint32_t floor_ = n_small - cap;
  CG_HIP(hipEventRecord(e->ev_pcie[0], s));
  CG_TRY(queue_share_copy(e, floor_, cap));
  CG_HIP(hipEventRecord(e->ev_pcie[1], s));
  QueueEnds ends;
  ends.reset(0, n_small, floor_);
  shadow_ready(e);
  e->team.clear_candidates();
  const uint32_t min_piece = std::max<uint32_t>(4096u, n_small / 64);
  uint64_t host_entries = 0, host_edges = 0;
  int agent_status = CG_OK;
  std::string agent_error;
  uint32_t pieces = 0;
  double gpu_sec = 0, host_sec = 0;
  // one team of T threads: thread 0 is the agent, threads 1 .. T - 1 the
  // host's workers (a team larger than the host team would leave a thread
  // more than the cores spinning in the OpenMP pool afterwards)
#pragma omp parallel num_threads(T)
  {
    const int tid = omp_get_thread_num();
    if (tid == 0) {
      // the GPU agent
      const auto t0 = clk::now();
      uint32_t want = std::max<uint32_t>(n_small / 2, min_piece);
      uint32_t lo = 0, hi = 0;
      bool first = true;
      for (;;) {
        const bool got = ends.claim_head(want, &lo, &hi);
        if (!got && !first) break;
        if (!got) lo = hi = 0;  // the big tiers still run
        if (pieces >= 2 && hipEventSynchronize(e->ev_piece[pieces & 1]) != hipSuccess) {
          agent_status = set_error(CG_ERR_HIP, "expansion piece failed");
          break;
        }
        const uint64_t items = (first ? (uint64_t)q.n_chunk + q.n_med : 0) + (hi - lo + kWave - 1) / kWave + 1;
        int st = launch_expand_push(alg, dg, e->val.p, out, e->q_small.p, e->q_med.p, e->q_chunk.p, items,
                                    SmallRange{lo, hi, first ? 1u : 0u}, bins, dst, e->stripes.p,
                                    e->opt.atomic_return != 0, s);
        if (st == CG_OK && hipEventRecord(e->ev_piece[pieces & 1], s) != hipSuccess)
          st = set_error(CG_ERR_HIP, "hipEventRecord failed");
        if (st != CG_OK) {
          agent_status = st;
          agent_error = last_error();
          break;
        }
        ++pieces;
        first = false;
        if (!got) break;
        want = std::max<uint32_t>(want / 2, min_piece);
      }
      if (agent_status == CG_OK && hipStreamSynchronize(s) != hipSuccess)
        agent_status = set_error(CG_ERR_HIP, "expansion pieces failed");
      gpu_sec = ms_since(t0) * 1e-3;
    } else {
      const auto t0 = clk::now();
      if (hipEventSynchronize(e->ev_copy) == hipSuccess) {
        uint64_t ed = 0, en = 0;
        e->team.tail_worker(tid - 1, alg, e->host_g(), e->p_ids.p, e->p_vals.p, &ends, &ed, &en);
#pragma omp atomic
        host_entries += en;
#pragma omp atomic
        host_edges += ed;
      }
      const double sec = ms_since(t0) * 1e-3;
#pragma omp critical
      host_sec = std::max(host_sec, sec);
    }
  }
  if (agent_status != CG_OK) return set_error(agent_status, "%s", agent_error.empty() ? last_error() : agent_error.c_str());
  local->expand_launches += pieces;
  local->gpu_vertices += sst.n_active - host_entries;
  local->cpu_vertices += host_entries;
  local->edges_relaxed += sst.m_active;
  local->bytes_d2h += 8ull * cap;
  float pms = 0.f;
  if (hipEventElapsedTime(&pms, e->ev_pcie[0], e->ev_pcie[1]) == hipSuccess) e->pcie_ms = pms;
  else (void)hipGetLastError();
  // measured rates (edges/s) of both sides feed the next split
  const double gpu_edges = W - (double)host_edges;
  if (gpu_sec > 0 && gpu_edges > 0) e->rate_gpu = 0.7 * e->rate_gpu + 0.3 * (gpu_edges / gpu_sec);
  if (host_sec > 0 && host_edges > 0) e->rate_cpu = 0.7 * e->rate_cpu + 0.3 * ((double)host_edges / host_sec);
  const uint64_t nc = e->team.gather(e->cand_id, e->cand_val);
  return apply_host_candidates(e, nc, out, local);
}

// Executes one iteration's expansion from frontier `in` into `out`.
int exec_step(cg_engine* e, const DevGraph& dg, const Step& step, const DevStats& sst, DevStats* dst,
              uint64_t* in, uint64_t* out, bool* timing_pending, cg_run_stats* local) {
  const uint32_t* tval = step.threshold != 0xFFFFFFFFu ? e->val.p : nullptr;
  hipStream_t s = e->stream;
  const uint64_t words = ((uint64_t)dg.V + 63) / 64;
  const int alg = e->alg;
  // queue lengths restart at zero for this compaction
  // k_stats leaves them zero; probes run several compactions per iteration
  if (e->opt.sched_mode == CG_SCHED_PROBE) CG_HIP(hipMemsetAsync(dst, 0, 3 * sizeof(uint32_t), s));
  if (step.ch.pull) {
    ++local->pull_iterations;
    ++local->heavy_iterations;
    local->gpu_vertices += sst.n_active;
    if (e->opt.profile) CG_HIP(hipEventRecord(e->ev0, s));
    const double V1 = (double)dg.V;
    if (alg == CG_BFS) {
      CG_TRY(launch_pull_bfs(dg, e->val.p, in, out, e->stripes.p, s));
      // lower bound: offsets, value check, both bitmaps (the scanned in-edges
      // depend on where each search stops and are not counted)
      local->expand_alg_bytes += 8.0 * (V1 + 1) + 4.0 * V1 + 16.0 * (double)words;
    } else {
      CG_TRY(launch_pull_min(alg, dg, e->g->E, e->pull_tiles.p, e->val.p, in, out, e->stripes.p, s));
      // offsets, in-edge sources (+ weights), the frontier sources' values
      // (one per frontier out-edge), the row values, both bitmaps
      local->expand_alg_bytes += 8.0 * (V1 + 1) + (alg == CG_SSSP ? 8.0 : 4.0) * (double)e->g->E +
                                 4.0 * (double)sst.m_active + 4.0 * V1 + 16.0 * (double)words;
    }
    if (e->opt.profile) { CG_HIP(hipEventRecord(e->ev1, s)); *timing_pending = true; }
    ++local->expand_launches;
    local->edges_relaxed += sst.m_active;
    return CG_OK;
  }
  if (step.host_share == 1) {
    // LIGHT: the host team relaxes the whole frontier
    ++local->light_iterations;
    const Bins all_small{0xFFFFFFFFu, 0xFFFFFFFFu, e->opt.chunk_edges};
    CG_TRY(launch_compact(in, words, dg, all_small, e->q_small.p, e->q_med.p, e->q_chunk.p, dst, nullptr,
                          0xFFFFFFFFu, nullptr, s));
    const uint32_t n = sst.n_active - sst.n_zero_deg;
    CG_TRY(queue_share_copy(e, 0, n));
    CG_HIP(hipEventSynchronize(e->ev_copy));
    return host_relax_and_apply(e, n, out, local);
  }
  // GPU push, optionally co-executed with the host team
  ++local->heavy_iterations;
  Bins bins{e->opt.bin_wave, e->opt.bin_chunk, e->opt.chunk_edges, e->opt.expand_grid};
  if (step.ch.tier) {
    bins.bin_wave = kTierSets[step.ch.tier].bin_wave;
    bins.bin_chunk = kTierSets[step.ch.tier].bin_chunk;
  }
  CG_TRY(launch_compact(in, words, dg, bins, e->q_small.p, e->q_med.p, e->q_chunk.p, dst, tval,
                        step.threshold, out, s));
  local->expand_alg_bytes += 20.0 * (sst.n_active - sst.n_zero_deg) +
                             (alg == CG_SSSP ? 12.0 : 8.0) * (double)sst.m_active;
  if (step.host_share == 2) return co_execute(e, dg, bins, sst, dst, out, local);
  const uint64_t bound = expand_item_bound(sst.n_active, sst.m_active, bins);
  // partly resident graph: the frontier's non-resident vertices (no edges on
  // the device) are gathered with their values before the expansion and
  // relaxed by the host team while it runs
  uint32_t n_nr = 0;
  if (e->partial) {
    CG_TRY(launch_gather_masked(in, e->nonres_bm.p, words, e->val.p, e->d_nr_ids.p, e->d_nr_vals.p,
                                e->d_nr_cnt.p, s));
    CG_HIP(hipMemcpyAsync(e->p_nr_cnt.p, e->d_nr_cnt.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    CG_HIP(hipStreamSynchronize(s));
    n_nr = e->p_nr_cnt.p[0];
    if (n_nr > e->n_nonres) return set_error(CG_ERR_INTERNAL, "non-resident frontier larger than the non-resident set");
    if (n_nr) {
      CG_TRY(e->p_ids.ensure(n_nr));
      CG_TRY(e->p_vals.ensure(n_nr));
      CG_HIP(hipMemcpyAsync(e->p_ids.p, e->d_nr_ids.p, (size_t)n_nr * 4, hipMemcpyDeviceToHost, s));
      CG_HIP(hipMemcpyAsync(e->p_vals.p, e->d_nr_vals.p, (size_t)n_nr * 4, hipMemcpyDeviceToHost, s));
      CG_HIP(hipEventRecord(e->ev_copy, s));
    }
  }
  if (e->opt.profile) CG_HIP(hipEventRecord(e->ev0, s));
  CG_TRY(launch_expand_push(alg, dg, e->val.p, out, e->q_small.p, e->q_med.p, e->q_chunk.p, bound,
                            SmallRange{}, bins, dst, e->stripes.p, e->opt.atomic_return != 0, s));
  if (e->opt.profile) { CG_HIP(hipEventRecord(e->ev1, s)); *timing_pending = true; }
  ++local->expand_launches;
  local->gpu_vertices += sst.n_active - n_nr;
  local->edges_relaxed += sst.m_active;
  if (n_nr) {
    CG_HIP(hipEventSynchronize(e->ev_copy));
    CG_TRY(host_relax_and_apply(e, n_nr, out, local));
  }
  return CG_OK;
}

// CG_SCHED_PROBE: time every choice from the current state, restoring the
// values, the frontier, the new-vertex counters and the host shadow after each.
int probe_all(cg_engine* e, const DevGraph& dg, const DevStats& sst, DevStats* dst, uint64_t* in,
              uint64_t* out, uint32_t threshold, double* us) {
  hipStream_t s = e->stream;
  const uint32_t V = dg.V;
  const uint64_t words = ((uint64_t)V + 63) / 64;
  CG_TRY(e->val_snap.ensure(V ? V : 1));
  CG_TRY(e->bm_snap.ensure(words + 1));
  CG_HIP(hipMemcpyAsync(e->val_snap.p, e->val.p, (size_t)V * 4, hipMemcpyDeviceToDevice, s));
  CG_HIP(hipMemcpyAsync(e->bm_snap.p, in, words * 8, hipMemcpyDeviceToDevice, s));
  shadow_ready(e);
  std::vector<uint32_t> shadow_snap = e->team.shadow;
  CG_HIP(hipStreamSynchronize(s));
  // Choices that execute identically share one timing: a bottom-up step
  // ignores the tiers and the side, a host step ignores the tiers.
  double by_key[6] = {-1, -1, -1, -1, -1, -1};
  for (int c = 0; c < CG_NUM_CHOICES; ++c) {
    Step st{decode_choice(c), 0, threshold};
    if (e->alg != CG_BFS && !e->pull_tiles_ready) st.ch.pull = 0;
    st.host_share = (st.ch.host && !st.ch.pull) ? 1 : 0;
    const int key = st.ch.pull ? 4 : (st.host_share ? 5 : st.ch.tier);
    if (by_key[key] >= 0) {
      us[c] = by_key[key];
      continue;
    }
    cg_run_stats scratch{};
    bool tp = false;
    const auto t0 = clk::now();
    CG_TRY(exec_step(e, dg, st, sst, dst, in, out, &tp, &scratch));
    CG_HIP(hipStreamSynchronize(s));
    us[c] = std::chrono::duration<double, std::micro>(clk::now() - t0).count();
    by_key[key] = us[c];
    CG_HIP(hipMemcpyAsync(e->val.p, e->val_snap.p, (size_t)V * 4, hipMemcpyDeviceToDevice, s));
    CG_HIP(hipMemcpyAsync(in, e->bm_snap.p, words * 8, hipMemcpyDeviceToDevice, s));
    CG_HIP(hipMemsetAsync(out, 0, words * 8, s));
    CG_HIP(hipMemsetAsync(e->stripes.p, 0, kStripeWords * sizeof(uint32_t), s));
    if (st.host_share) e->team.shadow = shadow_snap;
  }
  CG_HIP(hipStreamSynchronize(s));
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
  if (alg == CG_SSSP && !g.has_w) return set_error(CG_ERR_INVALID, "SSSP needs edge weights");
  // A run that failed part-way can leave queued-ahead kernels on the stream;
  // drain them before the host resets the statistics records they publish
  // into (their error, if any, was reported by that run).
  if (!e->clean_end && stream_wait(s) != hipSuccess) (void)hipGetLastError();
  e->clean_end = false;
  e->epoch = (e->epoch + 1) & 0xFFu;
  if (!e->epoch) e->epoch = 1;
  // a partly resident graph has only its resident edges on the device: no
  // transposed graph (bottom-up steps would miss the others)
  const bool partial = e->partial;
  if (alg == CG_BFS && mode != CG_SCHED_HEURISTIC && !partial) CG_TRY(build_transpose(&g));
  // bottom-up min steps (SSSP, label propagation): transposed graph + edge tiles
  const bool forced_pull = mode == CG_SCHED_FORCED && decode_choice(e->opt.forced_choice).pull;
  const bool min_pull = (alg == CG_SSSP || alg == CG_CC) && g.E > 0 && !partial &&
                        ((e->opt.min_pull && (mode == CG_SCHED_GPU || mode == CG_SCHED_LEARNED)) || forced_pull);
  if (min_pull) {
    CG_TRY(build_transpose(&g));
    if (!e->pull_tiles_ready) {
      CG_TRY(e->pull_tiles.ensure(pull_tile_count(g.E) + 1));
      CG_TRY(build_pull_tiles(g.d_in_off.p, V, g.E, e->pull_tiles.p, s));
      CG_HIP(hipStreamSynchronize(s));
      e->pull_tiles_ready = true;
    }
  }
  const bool learned = mode == CG_SCHED_LEARNED;
  const bool probe = mode == CG_SCHED_PROBE;
  if (learned) CG_TRY(mlp_load(e));
  const bool host_possible = partial || mode == CG_SCHED_HEURISTIC || learned || probe ||
                             (mode == CG_SCHED_FORCED && decode_choice(e->opt.forced_choice).host);
  if (host_possible) CG_TRY(ensure_team(e));
  const float* mlp_dev = (learned && e->mlp_valid) ? e->d_mlp.p : nullptr;

  DevGraph dg{g.d_off.p, g.d_dst.p, g.has_w ? g.d_w.p : nullptr,
              g.dev_t_valid ? g.d_in_off.p : nullptr, g.dev_t_valid ? g.d_in_src.p : nullptr, V,
              g.dev_t_valid && g.has_w ? g.d_in_w.p : nullptr, g.dev_deg_valid && !partial ? g.d_deg.p : nullptr,
              g.dev_wn_valid && !partial && g.d_w8.p ? g.d_w8.p : nullptr,
              g.dev_wn_valid && !partial && g.d_w16.p ? g.d_w16.p : nullptr};

  cg_run_stats local{};
  const auto t0 = clk::now();
  // ---- initAlgorithm (CGgraph_engine.cuh:1259-1316): one launch, below ----
  for (int k = 0; k < 3; ++k) __atomic_store_n(&e->hst_map.p[k].iter, 0u, __ATOMIC_RELEASE);
  e->shadow_stale = host_possible;
  e->shadow_root = root;

  int cur = 0;
  uint32_t it = 0, prev_active = 0;
  bool pull = false;
  uint64_t m_explored = 0;
  bool timing_pending = false;
  IterRecord rec{};
  // delta-stepping threshold: frontier vertices at or above it wait
  const uint32_t delta = alg == CG_SSSP ? e->opt.sssp_delta : 0;
  uint32_t threshold = delta ? delta : 0xFFFFFFFFu;
  const uint32_t* tval = delta ? e->val.p : nullptr;
  // GPU schedule for SSSP / CC without bottom-up steps: the tiers are fixed
  // and nothing waits for the statistics, so compaction and expansion are
  // queued right behind them and the host reads the statistics while the GPU
  // goes on (BFS decides push or pull from the statistics first).
  // The learned schedule joins when its network runs on the device: the
  // network runs behind each expansion on the features of the frontier just
  // expanded and picks the next iteration's schedule; a GPU push choice is
  // applied there (tiers from the choice), a host choice leaves the queues
  // empty and the host team runs the iteration after reading its record.
  const bool early_learned = learned && mlp_dev && alg != CG_BFS && !min_pull && !delta && !partial;
  // BFS under the GPU schedule: the direction rule runs on the device too, and
  // both the push pair and the bottom-up kernel are queued; each runs only
  // when the rule picked it
  const bool early_bfs = mode == CG_SCHED_GPU && !host_possible && alg == CG_BFS && g.dev_t_valid && !partial;
  const bool early = (mode == CG_SCHED_GPU && !host_possible && alg != CG_BFS && !min_pull) || early_learned ||
                     early_bfs;
  // SSSP hub deferral (GPU-queued iterations only): the statistics read the
  // frontier's values for their smallest one, the compaction keeps hubs above
  // it + delta waiting
  const bool hub_on = alg == CG_SSSP && e->opt.hub_degree != 0 && early && !delta;
  const HubDefer hub{hub_on ? e->opt.hub_degree : 0u, e->opt.hub_delta ? e->opt.hub_delta : kHubDeltaDefault,
                     hub_on ? e->stripes.p : nullptr};
  const uint32_t* cval = hub_on ? e->val.p : tval;  // values the compaction reads
  TierTable choice_tiers{};
  for (int k = 0; k < 4; ++k) {
    choice_tiers.wave[k] = k ? kTierSets[k].bin_wave : e->opt.bin_wave;
    choice_tiers.chunk[k] = k ? kTierSets[k].bin_chunk : e->opt.bin_chunk;
  }
  const Bins fbins{e->opt.bin_wave, e->opt.bin_chunk, e->opt.chunk_edges, e->opt.expand_grid};
  // Pure GPU schedule: iteration it + 1 is queued before the host reads the
  // statistics of iteration it, so the host's launch work overlaps the GPU.
  // The iteration after the last one finds an empty frontier and does nothing;
  // a delta-stepping threshold moves one iteration later.  Under the learned
  // schedule a host-team choice is known one record ahead: the iteration it
  // names is already queued (and queues nothing on the GPU), the one after it
  // is queued once the host team has relaxed the frontier.
  const bool run_ahead = early;
  // SSSP / CC under the GPU and learned schedules without delta-stepping: the
  // statistics are counted by the compaction itself and reduced behind the
  // expansion (three kernels per iteration, one pass over the frontier)
  const bool fused = run_ahead && alg != CG_BFS && !delta;
  // the fused compaction drops frontier vertices already expanded with their
  // present value (it needs the degree array)
  if (fused && dg.deg && e->xval.p) dg.xval = e->xval.p;
  // label propagation on the fused pipeline pulls its first round from the
  // in-edges (built once per graph, kept with it)
  const bool cc_pull1 = fused && alg == CG_CC && !partial && g.E > 0;
  if (cc_pull1) CG_TRY(build_transpose(&g));
  {
    TraverseInit ti{};
    ti.val = e->val.p;
    ti.V = V;
    ti.alg = alg;
    ti.root = root;
    ti.cc_active = (uint64_t)V - g.zero_tail;  // CC starts from V' (CGgraph_engine.cuh:1303-1311)
    ti.bm0 = e->bm[0].p;
    ti.bm1 = e->bm[1].p;
    ti.bm_words = words;
    ti.stripes = e->stripes.p;
    // fused SSSP with hub deferral: the minima stripes start empty, and slot
    // 1 (the first iteration's) holds the root's value 0 as the frontier's
    // smallest (the cleared slot's min_val)
    ti.minima_empty = fused && hub_on;
    ti.st = e->st.p;
    // learned: the first iteration has no network choice (the GPU rule, -1)
    ti.choice1_none = early_learned;
    ti.rs = e->run_state.p;
    // BFS on the GPU schedule counts the root into its first frontier;
    // the other paths start with the root (or, for CC, every vertex) visited
    ti.visited0 = alg == CG_CC ? V : (early_bfs ? 0 : 1);
    // BFS: the direction of iteration 1 is taken from the root's counts,
    // and the initialisation does iteration 1's front work
    ti.bfs_cnt = early_bfs ? e->bfs_cnt.p : nullptr;
    ti.off = dg.off;
    ti.bfs_first = early_bfs ? 1 : 0;
    ti.E = g.E;
    ti.bins = fbins;
    ti.q_small = e->q_small.p;
    ti.q_med = e->q_med.p;
    ti.q_chunk = e->q_chunk.p;
    ti.bfs_host_out = e->hst_map_dev + 1;
    ti.acc = e->acc.p;
    ti.epoch = e->epoch;
    ti.xval = dg.xval;
    if (cc_pull1) {
      ti.in_off = g.d_in_off.p;
      ti.in_src = g.d_in_src.p;
    }
    CG_TRY(launch_init_traverse(ti, s));
  }
  bool early_timed[4] = {false, false, false, false};
  auto take_timing = [&](int k) -> int {
    if (!early_timed[k]) return CG_OK;
    float ms = 0.f;
    CG_HIP(hipEventSynchronize(e->ev_pair[k][1]));
    CG_HIP(hipEventElapsedTime(&ms, e->ev_pair[k][0], e->ev_pair[k][1]));
    local.expand_kernel_ms += ms;
    early_timed[k] = false;
    return CG_OK;
  };
  // Fused pipeline: once the frontiers shrink below kTailEdges edges, the
  // expansions queued next launch a grid of kTailGrid workgroups instead of
  // the tuned cap (the grid strides over whatever the queues hold, so any
  // grid gives the same result; an idle workgroup still costs dispatch time)
  uint64_t last_m = ~0ull, prev_m = 0;  // frontier edges of the last two records (none yet)
  // iteration k reads bm[(k - 1) & 1], writes bm[k & 1], uses stats slot k & 1
  auto enqueue = [&](uint32_t k) -> int {
    const int sl = k & 1, c = (k - 1) & 1;
    DevStats* dk = e->st.p + sl;
    if (fused) {
      Bins xb = fbins;
      if (last_m < kTailEdges && last_m <= prev_m && (xb.grid_cap == 0 || xb.grid_cap > kTailGrid))
        xb.grid_cap = kTailGrid;
      uint32_t nparts = 0;
      CG_TRY(launch_compact_fused(e->bm[c].p, e->bm[c ^ 1].p, words, dg, fbins, e->q_small.p, e->q_med.p,
                                  e->q_chunk.p, dk, e->val.p, hub, e->parts.p, &nparts, e->run_state.p, s,
                                  early_learned ? &choice_tiers : nullptr));
      const int tp = k & 3;
      if (e->opt.profile) CG_HIP(hipEventRecord(e->ev_pair[tp][0], s));
      CG_TRY(launch_expand_push(alg, dg, e->val.p, e->bm[c ^ 1].p, e->q_small.p, e->q_med.p, e->q_chunk.p,
                                ~0ull >> 8, SmallRange{}, xb, dk, e->stripes.p, e->opt.atomic_return != 0, s,
                                hub_on));
      if (e->opt.profile) {
        CG_HIP(hipEventRecord(e->ev_pair[tp][1], s));
        early_timed[tp] = true;
      }
      ++local.expand_launches;
      return launch_fused_finish(e->parts.p, nparts, e->stripes.p, e->run_state.p, dk, e->st.p + (sl ^ 1),
                                 e->hst_map_dev + sl, hub_on, early_learned ? mlp_dev : nullptr, V, g.E, s);
    }
    CG_TRY(launch_stats(e->bm[c].p, e->bm[c ^ 1].p, words, dg, g.E, e->parts.p, e->stripes.p, e->run_state.p,
                        mlp_dev, tval, threshold, dk, e->hst_map_dev + sl, s,
                        0, hub_on ? e->st.p + (sl ^ 1) : nullptr, hub_on));
    if (!early) return CG_OK;
    CG_TRY(launch_compact(e->bm[c].p, words, dg, fbins, e->q_small.p, e->q_med.p, e->q_chunk.p, dk, cval,
                          threshold, e->bm[c ^ 1].p, s, nullptr, 8u, hub));
    const int tp = k & 3;
    if (e->opt.profile) CG_HIP(hipEventRecord(e->ev_pair[tp][0], s));
    CG_TRY(launch_expand_push(alg, dg, e->val.p, e->bm[c ^ 1].p, e->q_small.p, e->q_med.p, e->q_chunk.p,
                              ~0ull >> 8, SmallRange{}, fbins, dk, e->stripes.p, e->opt.atomic_return != 0, s,
                              hub_on));
    if (e->opt.profile) {
      CG_HIP(hipEventRecord(e->ev_pair[tp][1], s));
      early_timed[tp] = true;
    }
    ++local.expand_launches;
    return CG_OK;
  };
  // BFS on the GPU schedule: the front kernel of iteration k (direction,
  // record, compaction) and its step kernel are queued separately; the
  // initialisation did front 1.  Queued ahead, iteration it + 1 is in the
  // stream while the host waits for record it.  A frontier of few edges that
  // did not grow is probably among the last ones: the next loop then queues
  // step it + 1 and front it + 2 only once record it + 1 has arrived (a front
  // kernel publishes its record before its work), so the run ends with a
  // single front kernel queued behind it, which finds the frontier empty.
  uint32_t bfs_front_q = 1, bfs_step_q = 0;
  bool bfs_ahead = true;
  uint32_t bfs_prev_n = 0;
  auto bfs_front = [&](uint32_t k) -> int {
    const int sl = k & 1, c = (k - 1) & 1;
    CG_TRY(launch_bfs_front(e->bm[c].p, e->bm[c ^ 1].p, words, dg, g.E, fbins, e->q_small.p, e->q_med.p,
                            e->q_chunk.p, e->st.p + sl, e->st.p + (sl ^ 1), e->bfs_cnt.p + (size_t)c * kBfsCntWords,
                            e->bfs_cnt.p + (size_t)sl * kBfsCntWords, e->run_state.p + c, e->run_state.p + sl,
                            e->hst_map_dev + sl, s));
    bfs_front_q = k;
    return CG_OK;
  };
  auto bfs_step = [&](uint32_t k) -> int {
    const int sl = k & 1, c = (k - 1) & 1, tp = k & 3;
    if (e->opt.profile) CG_HIP(hipEventRecord(e->ev_pair[tp][0], s));
    CG_TRY(launch_bfs_step(e->bm[c].p, e->bm[c ^ 1].p, words, dg, fbins, e->val.p, e->q_small.p, e->q_med.p,
                           e->q_chunk.p, e->st.p + sl, e->bfs_cnt.p + (size_t)sl * kBfsCntWords,
                           e->opt.atomic_return != 0, s));
    if (e->opt.profile) {
      CG_HIP(hipEventRecord(e->ev_pair[tp][1], s));
      early_timed[tp] = true;
    }
    ++local.expand_launches;
    bfs_step_q = k;
    return CG_OK;
  };
  // the fronts and steps of the iterations up to k, in stream order
  auto bfs_queue_upto = [&](uint32_t k) -> int {
    for (uint32_t j = bfs_step_q + 1; j <= k; ++j) {
      if (bfs_front_q < j) CG_TRY(bfs_front(j));
      CG_TRY(bfs_step(j));
    }
    return CG_OK;
  };
  // learned schedule: the choice the previous record made for this
  // iteration (-1: the GPU rule)
  int pending = -1;
  if (run_ahead && !early_bfs) CG_TRY(enqueue(1));
  auto tit = clk::now();
  for (;;) {
    ++it;
    const int slot = it & 1;
    DevStats* dst = e->st.p + slot;
    const bool host_now = early_learned && pending >= 0 && decode_choice(pending).host;
    if (!run_ahead) CG_TRY(enqueue(it));
    else if (early_bfs) {
      if (bfs_ahead) CG_TRY(bfs_queue_upto(it + 1));
    } else if (!host_now) {
      CG_TRY(enqueue(it + 1));
    }
    CG_TRY(wait_stats(e, slot, it));
    if (e->opt.fault_flags & CG_FAULT_HIP_ERROR)
      return set_error(CG_ERR_HIP, "injected HIP failure at iteration %u", it);
    const DevStats sst = e->hst_last;
    if (early) CG_TRY(take_timing((it - 1) & 3));  // the previous expansion has finished
    if (it > 1) {
      if (timing_pending) {
        float ms = 0.f;
        CG_HIP(hipEventSynchronize(e->ev1));
        CG_HIP(hipEventElapsedTime(&ms, e->ev0, e->ev1));
        local.expand_kernel_ms += ms;
        timing_pending = false;
      }
      rec.us = std::chrono::duration<double, std::micro>(clk::now() - tit).count();
      if (!probe) write_trace(e, rec, root);
    }
    tit = clk::now();
    if (hub_on) {
      // waiting hubs are counted as expanded below (the fused record carries
      // its own iteration's, the others the previous iteration's)
      const uint32_t n_def = fused ? sst.n_def : sst.n_def_prev;
      const unsigned long long m_def = fused ? sst.m_def : sst.m_def_prev;
      local.edges_relaxed -= m_def;
      local.gpu_vertices -= n_def;
      local.expand_alg_bytes -= 20.0 * n_def + 12.0 * (double)m_def;
    }
    if (sst.n_active == 0) break;
    prev_m = it > 1 ? last_m : 0;
    last_m = sst.m_active;
    if (early_bfs) {
      if (!bfs_ahead) {
        CG_TRY(bfs_queue_upto(it));
        if (bfs_front_q < it + 1) CG_TRY(bfs_front(it + 1));
      }
      bfs_ahead = !(sst.m_active < kBfsTailEdges && sst.n_active <= bfs_prev_n);
      bfs_prev_n = sst.n_active;
    }
    if (delta && sst.n_near == 0) {
      // nothing below the threshold: move it past the smallest waiting value
      const uint64_t t = (uint64_t)sst.min_val + delta;
      threshold = t > 0xFFFFFFFEull ? 0xFFFFFFFEu : (uint32_t)t;
    }

    // ---- schedule ----
    const uint64_t m_unexplored = g.E > m_explored ? g.E - m_explored : 0;
    Step step{Choice{0, 0, 0}, 0, threshold};
    int took_learned = 0;
    if (mode == CG_SCHED_HEURISTIC) {
      step.host_share = heuristic_heavy(e, sst) ? 2 : 1;
    } else if (mode == CG_SCHED_FORCED) {
      step.ch = decode_choice(e->opt.forced_choice);
      step.host_share = step.ch.host ? 1 : 0;
    } else {
      if (early_bfs) pull = sst.choice == 1;  // decided on the device
      else if (alg == CG_BFS) pull = beamer_pull(pull, sst, m_unexplored, V, prev_active);
      step.ch.pull = pull;
      if (learned) {
        if (pending >= 0) {
          step.ch = decode_choice(pending);
          step.host_share = step.ch.host ? 1 : 0;
          took_learned = 1;
        } else {
          ++local.learned_fallbacks;
        }
      }
      // bottom-up min step when the frontier's edges reach min_pull / 1000 of E
      if (min_pull) step.ch.pull = sst.m_active * 1000ull >= (uint64_t)e->opt.min_pull * g.E;
    }
    if (alg != CG_BFS && !(min_pull && (mode != CG_SCHED_HEURISTIC))) step.ch.pull = 0;
    if (step.ch.pull) step.host_share = 0;  // bottom-up steps run on the GPU
    if (partial) {
      // partly resident: the GPU pushes the resident vertices, the host team
      // the others (exec_step), every iteration
      step.ch = Choice{0, 0, 0};
      step.host_share = 0;
    }
    pull = step.ch.pull;

    rec.iter = it;
    rec.choice = encode_choice(Choice{step.ch.pull, step.ch.tier, step.host_share == 1});
    rec.host_share = step.host_share;
    rec.learned = took_learned;
    rec.s = sst;
    make_features(sst, V, g.E, sst.visited, sst.iter, rec.feat);
    m_explored += sst.m_active;
    prev_active = sst.n_active;

    uint64_t* in = e->bm[cur].p;
    uint64_t* out = e->bm[cur ^ 1].p;
    const uint32_t pstride = e->opt.probe_stride > 1 ? e->opt.probe_stride : 1;
    if (probe && (pstride == 1 || it % pstride <= 1)) {
      double us[CG_NUM_CHOICES];
      CG_TRY(probe_all(e, dg, sst, dst, in, out, threshold, us));
      write_probe(e, rec, root, us);
    }
    if (early && step.host_share) {
      // the network gave this iteration to the host team; the queued GPU
      // expansion found empty queues.  The next iteration is queued behind
      // the host team's candidates.
      CG_TRY(exec_step(e, dg, step, sst, dst, in, out, &timing_pending, &local));
      if (host_now) CG_TRY(enqueue(it + 1));
    } else if (early && step.ch.pull) {
      // the queued bottom-up kernel ran (same accounting as exec_step's)
      ++local.pull_iterations;
      ++local.heavy_iterations;
      local.gpu_vertices += sst.n_active;
      local.edges_relaxed += sst.m_active;
      local.expand_alg_bytes += 8.0 * ((double)V + 1) + 4.0 * (double)V + 16.0 * (double)words;
    } else if (early) {
      // already queued behind the statistics
      ++local.heavy_iterations;
      local.gpu_vertices += sst.n_active;
      local.edges_relaxed += sst.m_active;
      local.expand_alg_bytes += 20.0 * (sst.n_active - sst.n_zero_deg) +
                                (alg == CG_SSSP ? 12.0 : 8.0) * (double)sst.m_active;
    } else {
      CG_TRY(exec_step(e, dg, step, sst, dst, in, out, &timing_pending, &local));
    }
    // the network's choice for the next iteration, made from this frontier
    pending = (learned && mlp_dev && probs_usable(sst)) ? sst.choice : -1;
    cur ^= 1;
  }
  // The query's last pass, queued behind the iterations (those queued past
  // the empty frontier find nothing to do): the TEPS numerator -- BFS on the
  // GPU schedule has it already, every reached vertex being in exactly one
  // frontier whose out-edges the records count -- and, on a renumbered copy,
  // the values in the caller's ids.  The host then waits for the query-done
  // record that a one-lane kernel publishes behind it, with the numerator.
  // Where the record of the empty frontier came after the last expansion
  // (BFS and the fused SSSP / CC pipeline) and nothing else is to be done,
  // the values are final already: the kernels still queued are left to drain
  // (every read of the values waits for the stream).
  const bool count_edges = !early_bfs && alg != CG_CC;
  if (count_edges || e->renumbered) {
    DevGraph tg = dg;
    if (partial) tg.off = e->full_off.p;  // every edge of a reached vertex, resident or not
    CG_TRY(launch_finish_query(tg, e->val.p, count_edges ? e->acc.p : nullptr,
                               e->renumbered ? e->new_of_d.p : nullptr, e->val_out.p, s));
  }
  unsigned long long reached = 0;
  const bool wait_done = count_edges || e->renumbered || !(early_bfs || fused);
  if (wait_done) {
    CG_TRY(launch_query_done(count_edges ? e->acc.p : nullptr, e->hst_map_dev + 2, record_tag(e->epoch, 0), s));
    CG_TRY(wait_stats(e, 2, 0));
    reached = e->hst_last.m_active;
  }
  if (early)
    for (int k = 0; k < 4; ++k) CG_TRY(take_timing(k));
  local.iterations = it - 1;
  local.ms = ms_since(t0);
  e->cur = cur;
  if (e->trace) std::fflush(e->trace);
  if (early_bfs) local.edges_traversed = m_explored;
  else if (count_edges) local.edges_traversed = reached;
  else local.edges_traversed = e->caller_g->E;
  e->clean_end = true;
  if (rs) *rs = local;
  return CG_OK;
}

int engine_run_host(cg_engine* e, uint32_t root, cg_run_stats* rs) {
  cg_graph& g = *e->g;
  const uint32_t V = g.V;
  const int alg = e->alg;
  if (alg == CG_SSSP && !g.has_w) return set_error(CG_ERR_INVALID, "SSSP needs edge weights");
  CG_TRY(ensure_team(e));
  const uint64_t words = ((uint64_t)V + 63) / 64;
  std::vector<uint64_t> in, out;
  try {
    in.assign(words + 1, 0);
    out.assign(words + 1, 0);
  } catch (const std::bad_alloc&) {
    return set_error(CG_ERR_OOM, "host frontier bitmaps");
  }
  cg_run_stats local{};
  const auto t0 = clk::now();
  // ---- initAlgorithm (CGgraph_engine.cuh:1259-1316) ----
  e->team.reset_values(alg, root);
  if (alg == CG_CC) {
    const uint64_t Vp = (uint64_t)V - g.zero_tail;
    for (uint64_t i = 0; i < Vp / 64; ++i) in[i] = ~0ull;
    if (Vp % 64) in[Vp / 64] = (1ull << (Vp % 64)) - 1;
  } else {
    in[root >> 6] |= 1ull << (root & 63);
  }
  uint32_t it = 0;
  for (;;) {
    ++it;
    uint64_t frontier = 0;
    for (uint64_t i = 0; i < words; ++i) frontier += (uint64_t)__builtin_popcountll(in[i]);
    std::fill(out.begin(), out.end(), 0ull);
    uint64_t edges = 0;
    const uint64_t updates = e->team.step_bitmap(alg, &g, in.data(), out.data(), words, &edges);
    ++local.light_iterations;
    local.cpu_vertices += frontier;
    local.edges_relaxed += edges;
    if (e->trace) {
      std::fprintf(e->trace,
                   "{\"alg\":%d,\"V\":%u,\"E\":%llu,\"root\":%u,\"iter\":%u,\"host\":1,\"n_active\":%llu,"
                   "\"edges\":%llu,\"updates\":%llu,\"stolen_units\":%llu,\"donated_vertices\":%llu}\n",
                   alg, V, (unsigned long long)g.E, root, it, (unsigned long long)frontier,
                   (unsigned long long)edges, (unsigned long long)updates,
                   (unsigned long long)e->team.last_stolen_units, (unsigned long long)e->team.last_donated_vertices);
    }
    if (updates == 0) break;
    in.swap(out);
  }
  local.iterations = it;
  local.ms = ms_since(t0);
  if (e->trace) std::fflush(e->trace);
  // TEPS numerator, outside the timed region
  if (alg != CG_CC) {
    uint64_t m = 0;
    const uint32_t* val = e->team.shadow.data();
#pragma omp parallel for num_threads(e->team.threads) reduction(+ : m) schedule(static)
    for (int64_t v = 0; v < (int64_t)V; ++v)
      if (val[v] != kUnreached) m += g.h_off[v + 1] - g.h_off[v];
    local.edges_traversed = m;
  } else {
    local.edges_traversed = g.E;
  }
  if (rs) *rs = local;
  return CG_OK;
}

}  // namespace cg

cg_engine::~cg_engine() {
  // speculative iterations of the last run may still be queued: the buffers
  // they use are freed after this body
  if (stream) (void)hipStreamSynchronize(stream);
  if (plan) cg::pr_plan_destroy(plan);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (ev_copy) (void)hipEventDestroy(ev_copy);
  for (auto& pr : ev_pair)
    for (auto ev : pr)
      if (ev) (void)hipEventDestroy(ev);
  for (auto ev : ev_piece)
    if (ev) (void)hipEventDestroy(ev);
  for (auto ev : ev_pcie)
    if (ev) (void)hipEventDestroy(ev);
  if (stream) (void)hipStreamDestroy(stream);
  if (trace) std::fclose(trace);
}
