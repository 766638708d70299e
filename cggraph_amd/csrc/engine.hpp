// The engine object behind cg_engine*.
#pragma once

#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "host_team.hpp"
#include "kernels.hpp"
#include "sched.hpp"

struct cg_engine {
  cg_graph* g = nullptr;         // the graph the kernels run on
  cg_graph* caller_g = nullptr;  // the caller's graph (differs from g when renumbered)
  int alg = CG_BFS;
  int engine_type = CG_COORDINATION;
  int use_device = 1;
  cg_options opt{};
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_copy = nullptr;
  hipEvent_t ev_pair[4][2] = {};  // expansions queued ahead, by iteration mod 4
  hipEvent_t ev_piece[2] = {};    // co-executed iterations: the agent's last two pieces
  hipEvent_t ev_pcie[2] = {};     // co-executed iterations: around the host share's copy
  // no GPU: every iteration runs on the host team (use_device_num 0, MULTI_CORE, SINGLE_CORE)
  bool host_only = false;

  // values (uint32 for BFS/SSSP/CC, float for PR) and frontier state
  cg::DevBuf<uint32_t> val;
  cg::DevBuf<uint32_t> xval;  // SSSP / CC fused pipeline: DevGraph::xval
  cg::DevBuf<uint64_t> bm[2];
  cg::DevBuf<uint32_t> q_small, q_med;
  cg::DevBuf<uint64_t> q_chunk;
  cg::DevBuf<cg::DevStats> st;       // two slots, alternating per iteration
  cg::PinnedBuf<cg::DevStats> hst;
  cg::PinnedBuf<cg::DevStats> hst_map;  // mapped: k_stats publishes here
  cg::DevStats* hst_map_dev = nullptr;  // its device address
  cg::DevStats hst_last{};
  cg::DevBuf<unsigned long long> acc;  // kAccSlots accumulators (kAccLines used, a line apart)
  // run number tagged into the records (record_tag), and whether the last
  // run returned normally: only no-op iterations can then be left on the
  // stream, and the next run need not drain it (an error return may leave
  // queued work and a HIP error, which the next run waits for and clears)
  uint32_t epoch = 0;
  bool clean_end = true;
  cg::DevBuf<cg::StatsPart> parts;     // k_stats workgroup partials
  cg::DevBuf<uint32_t> stripes;        // striped new-vertex counters
  cg::DevBuf<cg::RunState> run_state;  // two: BFS's front kernel reads one, writes the other
  cg::DevBuf<unsigned long long> bfs_cnt;  // BFS first-visit counters, two parities
  cg::DevBuf<uint32_t> val_snap;       // probe mode: state restored between choices
  cg::DevBuf<uint64_t> bm_snap;

  // host-team traffic
  cg::DevBuf<uint32_t> d_share_vals, d_cand_id, d_cand_val;
  cg::PinnedBuf<uint32_t> p_ids, p_vals, p_cand_id, p_cand_val;
  cg::PinnedBuf<uint64_t> p_chunks;
  std::vector<uint32_t> cand_id, cand_val;
  std::vector<uint64_t> chunk_lo, chunk_hi;
  cg::HostTeam team;
  bool team_ready = false;
  // the team's shadow values (an upper bound of the run's values that
  // filters its candidates) are reset on the run's first host-team work, so
  // runs that never hand work to the team skip the pass over V
  bool shadow_stale = false;
  uint32_t shadow_root = 0;

  // renumbering by out-degree (BFS / SSSP): g is an engine-owned copy whose
  // vertex new_of[v] is the caller's v; results are gathered back through it
  bool renumbered = false;
  std::unique_ptr<cg_graph> own_g;
  cg::DevBuf<uint32_t> new_of_d, val_out;
  std::vector<uint32_t> new_of;

  // bottom-up min steps: row of the first in-edge of each edge tile
  cg::DevBuf<uint32_t> pull_tiles;
  bool pull_tiles_ready = false;

  // partial residency (TwoLevelPartition under hbm_budget): g is a device
  // view holding the resident segments' edges only; the host team relaxes
  // the frontier's vertices in nonres_bm from caller_g's host CSR
  bool partial = false;
  cg::DevBuf<uint64_t> nonres_bm;   // non-resident vertices with edges
  cg::DevBuf<uint64_t> full_off;    // the graph's own offsets (degrees, TEPS numerator)
  uint64_t n_nonres = 0, segments = 0, resident_segments = 0;
  cg::DevBuf<uint32_t> d_nr_cnt;    // non-resident frontier vertices gathered this iteration
  cg::DevBuf<uint32_t> d_nr_ids, d_nr_vals;  // and their (vertex, value), n_nonres at most
  cg::PinnedBuf<uint32_t> p_nr_cnt;
  // the CSR the host team reads
  cg_graph* host_g() const { return partial ? caller_g : g; }

  // PageRank
  cg::DevBuf<float> rank, rank2, contrib, contrib2;
  cg::PrPlan* plan = nullptr;

  // learned scheduler
  std::vector<float> mlp;   // 6032 parameters
  bool mlp_valid = false;
  cg::DevBuf<float> d_mlp;

  // measured throughputs (edges/s) and PCIe time (ms) of the last shared
  // iteration, for the host/GPU split
  double rate_gpu = 2.0e10, rate_cpu = 2.0e8, pcie_ms = 0.0;

  std::string trace_path, mlp_path;  // owned copies of the option strings
  FILE* trace = nullptr;
  bool result_valid = false;
  uint32_t last_root = 0;
  int cur = 0;  // bitmap holding the result's last frontier

  ~cg_engine();
};

namespace cg {
int engine_run(cg_engine* e, uint32_t root, cg_run_stats* rs);
// The engine without a GPU: every iteration is relaxed by the host team over
// a host frontier bitmap, on host values (the reference's LIGHT path,
// CGgraph_engine.cuh:907-912, and the MULTI_CORE engine, GraphNuma.cuh:166-314).
int engine_run_host(cg_engine* e, uint32_t root, cg_run_stats* rs);
int mlp_load(cg_engine* e);
// Learned-scheduler inference on the device, fused into the stats epilogue.
int launch_stats_mlp(const float* params, const DevStats* st, uint32_t V, uint64_t E,
                     uint64_t visited, uint32_t iter, float* probs, hipStream_t s);
}  // namespace cg
