// Host-side launchers of the HIP kernels.  Kernels live in the .hip files;
// engine.cpp and capi.cpp only see these declarations.
#pragma once

#include "common.hpp"

namespace cg {

// Counters the frontier kernels accumulate on the device; the host reads the
// whole struct back once per iteration (one small D2H copy).
struct DevStats {
  uint32_t n_small;            // frontier vertices with 0 < degree < bin_wave
  uint32_t n_med;              // bin_wave <= degree < bin_chunk (one wave each)
  uint32_t n_chunk;            // edge chunks of vertices with degree >= bin_chunk
  uint32_t n_active;           // frontier vertices (popcount of the bitmap)
  unsigned long long m_active; // sum of out-degrees over the frontier
  uint32_t hist[8];            // frontier vertices by floor(log2(degree + 1) / 2), last bin open
  uint32_t n_new;              // vertices first reached by the previous expansion
  uint32_t n_zero_deg;         // frontier vertices with out-degree 0
  unsigned long long visited;  // vertices reached so far (BFS/SSSP), V for CC
  uint32_t n_cand_applied;     // host candidates that improved a value
  uint32_t chunk_edges;         // BFS direction kernel: edges per chunk-tier item this iteration (0: the engine's)
  int32_t choice;              // learned-scheduler argmax, -1 when it did not run
  uint32_t iter;               // iteration index (1-based) seen by k_stats
  float probs[16];             // learned-scheduler softmax
  uint32_t min_val;            // smallest value on the frontier (threshold runs only)
  uint32_t n_near;             // frontier vertices below the threshold
  // hub deferral: frontier vertices the compaction kept waiting (added by
  // k_compact after the record was published), and the previous iteration's
  // counts, published with this record so the host can correct its totals
  uint32_t n_def, n_def_prev;
  unsigned long long m_def, m_def_prev;
};
static_assert(sizeof(DevStats) % 8 == 0, "DevStats layout");

// A BFS push frontier whose edges would give fewer than kChunkSpread chunk
// items at the engine's chunk size is cut into smaller chunks (halved down to
// one wavefront's width), so that a lone hub spreads over the machine; at
// most 2 * kChunkSpread extra items come of it.
constexpr uint32_t kChunkSpread = 2048;

// Counters written by many wavefronts are striped over kStripes cache lines
// so that their atomics do not serialise on one address.
constexpr int kStripes = 16;
constexpr int kStripeStride = 32;  // uint32 per stripe = 128 bytes
// a second set of kStripes stripes (minima) follows the counters
constexpr int kStripeMinBase = kStripes * kStripeStride;
constexpr int kStripeWords = 2 * kStripes * kStripeStride;

// Per-run device state the stats kernel carries from iteration to iteration.
struct RunState {
  unsigned long long visited;
  uint32_t iter;
  uint32_t pull;                   // device-side direction rule: last step was bottom-up
  unsigned long long m_explored;   // frontier edges of the iterations so far
  uint32_t prev_active;            // frontier size of the previous iteration
  uint32_t pad;                    // SSSP / CC fused pipeline: first visits of the last expansion
  uint32_t done;                   // SSSP / CC fused pipeline: an empty frontier was counted
  uint32_t epoch;                  // the run's number, tagged into the records it publishes
};

// A record published to mapped host memory carries, as its last store, the
// tag of the run (low 8 bits, never 0) and the iteration (low 24 bits) it
// belongs to, so records that iterations still queued from an earlier run
// publish late never match what the host waits for, and the host does not
// have to drain the stream before a run.
constexpr uint32_t record_tag(uint32_t epoch, uint32_t iter) { return (epoch << 24) | (iter & 0xFFFFFFu); }

// Per-workgroup partial of k_stats (and of the fused compaction, which also
// counts the hubs it keeps waiting), reduced by one workgroup.
struct StatsPart {
  uint32_t n_active, n_zero;
  unsigned long long m_active;
  uint32_t hist[8];
  uint32_t min_val, n_near;
  uint32_t n_def, pad;
  unsigned long long m_def;
};

struct DevGraph {
  const uint64_t* off;
  const uint32_t* dst;
  const uint32_t* w;
  const uint64_t* in_off;
  const uint32_t* in_src;
  uint32_t V;
  const uint32_t* in_w = nullptr;  // weights of the in-edges (bottom-up SSSP)
  // out-degrees, zero-padded to whole 64-vertex words (ensure_degrees); the
  // frontier compaction reads them instead of the offsets when present
  const uint32_t* deg = nullptr;
  // narrow copies of w (ensure_narrow_weights): at most one is set, when
  // every weight fits it; the push expansion of SSSP reads it instead of w
  const uint8_t* w8 = nullptr;
  const uint16_t* w16 = nullptr;
  // SSSP / CC push on the fused pipeline: per vertex, one plus a value with
  // which every out-edge has been relaxed (0: none yet).  The expansion
  // writes it; the compaction drops a frontier vertex whose value equals it,
  // since expanding it again would offer every neighbour the same candidates.
  uint32_t* xval = nullptr;
};

struct Bins {
  uint32_t bin_wave;
  uint32_t bin_chunk;
  uint32_t chunk_edges;
  uint32_t grid_cap = 0;  // push expansion: most workgroups launched (0: 2048)
};

// ---- traversal (traverse.hip) ----
int launch_fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t s);
int launch_iota_u32(uint32_t* p, uint64_t n, hipStream_t s);
int launch_fill_u64(uint64_t* p, uint64_t value, uint64_t n, hipStream_t s);
// Sets bits [0, n) of a bitmap of `words` words (the rest cleared).
int launch_bitmap_prefix(uint64_t* bm, uint64_t words, uint64_t n, hipStream_t s);
int launch_set_root(uint32_t* val, uint64_t* bm, uint32_t root, hipStream_t s);
int launch_init_run_state(RunState* rs, uint64_t visited, hipStream_t s);
// Frontier statistics of `in` (n_active, m_active, hist, n_zero_deg), two
// launches: workgroup partials go to `parts` (one StatsPart per workgroup,
// stats_parts_needed(words)) while `zero_me` (the bitmap the coming expansion
// writes) is cleared; then one workgroup reduces them in a fixed order, folds
// the striped new-vertex counters of the previous expansion (`stripes`,
// zeroed again) into rs->visited, and -- when `mlp` is not NULL -- runs the
// learned scheduler on the iteration's features and writes choice/probs into
// st.  With `val` not NULL it also reports the frontier's smallest value and
// how many frontier vertices lie below `threshold` (delta-stepping runs).
// The reduction zeroes st's queue counters and publishes the stats to
// `host_out` (mapped host memory, may be NULL) with `iter` written last.
// direction != 0: the reduction also takes the direction-optimising BFS
// decision on the device (the rule of the host's beamer_pull, state in rs) and
// leaves it in st->choice (1 bottom-up, 0 push), for the kernels queued behind.
uint64_t stats_parts_needed(uint64_t words);
// prev not NULL: the other statistics slot, whose deferral counts (left by the
// previous iteration's compaction) are published as n_def_prev / m_def_prev.
int launch_stats(const uint64_t* in, uint64_t* zero_me, uint64_t words, const DevGraph& g,
                 uint64_t E, StatsPart* parts, uint32_t* stripes, RunState* rs, const float* mlp,
                 const uint32_t* val, uint32_t threshold, DevStats* st,
                 DevStats* host_out, hipStream_t s, int direction = 0, const DevStats* prev = nullptr,
                 bool min_from_stripes = false);
// Writes the three work queues; lengths accumulate in st->n_small/n_med/n_chunk.
// Frontier vertices whose value is >= threshold are not queued but carried
// into `carry` (the next frontier) unchanged.
// choice_tiers not NULL: the tiers come from the choice k_stats left in
// st->choice (tier set k of the table; the table's entry 0 is also used when
// the network deferred, choice < 0), and a choice with any of `skip_bits` set
// (8: the host team takes the iteration; 1: it goes bottom-up) compacts
// nothing.
// hub.degree != 0 (SSSP): instead of `threshold`, a frontier vertex of
// out-degree >= hub.degree waits when its value is >= st->min_val + hub.delta
// (the frontier's smallest value: from k_stats with `val`, or from the minima
// stripes the expansion and this kernel keep, hub.track); lower-degree
// vertices never wait.  The waiting vertices are counted into st->n_def /
// st->m_def.
struct TierTable {
  uint32_t wave[4], chunk[4];
};
struct HubDefer {
  uint32_t degree, delta;
  uint32_t* track = nullptr;  // stripes: the waiting vertices' smallest value goes to their minima
};
int launch_compact(const uint64_t* in, uint64_t words, const DevGraph& g, const Bins& bins,
                   uint32_t* q_small, uint32_t* q_med, uint64_t* q_chunk, DevStats* st,
                   const uint32_t* val, uint32_t threshold, uint64_t* carry, hipStream_t s,
                   const TierTable* choice_tiers = nullptr, uint32_t skip_bits = 8,
                   HubDefer hub = HubDefer{0, 0});
// Fused compaction and statistics (GPU schedule, SSSP / CC without
// delta-stepping): one pass over frontier `in` writes the three work queues
// (lengths into st, whose counters start at zero), every word of `out` (zero,
// or the hubs kept waiting under `hub`) and one statistics partial per
// workgroup into `parts` (*nparts set by the launcher, at most
// stats_parts_needed(words) - 1).  The hub threshold is st->min_val +
// hub.delta.  With g.deg and g.xval, a frontier vertex whose value is the
// one its out-edges were last relaxed with is dropped: its bit is cleared in
// `in`, and it is neither queued nor counted.  launch_fused_finish reduces
// the partials after the expansion
// has run: it publishes the iteration's record (as k_stats_reduce does, plus
// n_def / m_def of this iteration), folds the expansion's first-visit stripes
// into the run state, and prepares `next` (the following iteration's slot):
// queue counters zeroed and, with `track`, min_val taken from the minima
// stripes, which it resets.
// Learned schedule (choice_tiers / mlp not NULL): the finish runs the
// network on the features of the frontier it reduced and leaves its choice
// for the next iteration in the record and in next->choice; the compaction
// takes the tiers of the choice in st->choice (entry 0 of the table when it
// is -1), and for a host-team choice queues nothing and keeps no hub waiting.
// rs: the run state k_fused_finish keeps; once it has counted an empty
// frontier, the iterations queued behind return at once.
int launch_compact_fused(uint64_t* in, uint64_t* out, uint64_t words, const DevGraph& g, const Bins& bins,
                         uint32_t* q_small, uint32_t* q_med, uint64_t* q_chunk, DevStats* st,
                         const uint32_t* val, HubDefer hub, StatsPart* parts, uint32_t* nparts,
                         const RunState* rs, hipStream_t s, const TierTable* choice_tiers = nullptr);
int launch_fused_finish(const StatsPart* parts, uint32_t nparts, uint32_t* stripes, RunState* rs, DevStats* st,
                        DevStats* next, DevStats* host_out, bool track, const float* mlp, uint32_t V, uint64_t E,
                        hipStream_t s);
// Everything a traversal run starts from, in one launch: val (BFS/SSSP:
// unreached, root 0; CC: identity), bm0 = the first frontier (the root, or
// vertices [0, cc_active) for CC), bm1 cleared (bm_words words each), the
// striped counters cleared (the minima set empty when minima_empty), both
// statistics slots cleared (slot 1's choice -1 when choice1_none), both run
// states with `visited0`, and -- bfs_cnt not NULL -- the BFS counters with
// the root's counts in parity 0 (off: the graph's offsets).
struct TraverseInit {
  uint32_t* val;
  uint64_t V;
  int alg;
  uint32_t root;
  uint64_t cc_active;
  uint64_t* bm0;
  uint64_t* bm1;
  uint64_t bm_words;
  uint32_t* stripes;
  int minima_empty;
  DevStats* st;
  int choice1_none;
  RunState* rs;
  unsigned long long visited0;
  unsigned long long* bfs_cnt;
  const uint64_t* off;
  // BFS on the GPU schedule: the direction kernel's work for frontier one
  // ({root}) is done here as well -- the decision, the root's queue entries
  // in slot 1 (iteration one's), run state 1 and the record of iteration one
  // (bfs_host_out) -- so iteration one is its step kernel alone
  int bfs_first;
  uint64_t E;
  Bins bins;
  uint32_t* q_small;
  uint32_t* q_med;
  uint64_t* q_chunk;
  DevStats* bfs_host_out;
  unsigned long long* acc;  // the query's striped accumulators (kAccSlots), cleared
  uint32_t epoch;           // the run's tag (record_tag)
  uint32_t* xval = nullptr; // DevGraph::xval of the run, cleared (V entries)
  // CC with the in-edges (sorted by source): the first round of labels is
  // pulled here, and bm0 holds the vertices whose label fell
  const uint64_t* in_off = nullptr;
  const uint32_t* in_src = nullptr;
};
int launch_init_traverse(const TraverseInit& a, hipStream_t s);
// BFS under the GPU schedule, two kernels per iteration (traverse.hip): the
// front kernel takes the direction decision from the counts the previous
// step left in cnt_prev (first visits, their out-edges, those without) and
// rs_prev, stores it with the next run state rs_cur and publishes the record
// (host_out), clears cnt_cur and st_next's queue counters, and for a push
// step compacts `in` into the queues (st's counters, zero on entry) and
// clears `out`; the step kernel expands (push) or searches parents bottom-up
// and counts its first visits into cnt_cur.  Counters: kStripes stripes of
// kBfsCntStride uint64 per parity; the traversal's initialisation writes the
// root's counts (frontier one) into parity 0 and clears parity 1.
constexpr int kBfsCntStride = 16;                    // 128 bytes per stripe
constexpr int kBfsCntWords = kStripes * kBfsCntStride;  // per parity
int launch_bfs_front(const uint64_t* in, uint64_t* out, uint64_t words, const DevGraph& g, uint64_t E,
                     const Bins& bins, uint32_t* q_small, uint32_t* q_med, uint64_t* q_chunk, DevStats* st,
                     DevStats* st_next, const unsigned long long* cnt_prev, unsigned long long* cnt_cur,
                     const RunState* rs_prev, RunState* rs_cur, DevStats* host_out, hipStream_t s);
int launch_bfs_step(const uint64_t* in, uint64_t* out, uint64_t words, const DevGraph& g, const Bins& bins,
                    uint32_t* val, const uint32_t* q_small, const uint32_t* q_med, const uint64_t* q_chunk,
                    const DevStats* st, unsigned long long* cnt_cur, bool exact, hipStream_t s);
// Upper bound on the expansion's work items for a frontier of n_active
// vertices and m_active edges.
uint64_t expand_item_bound(uint64_t n_active, uint64_t m_active, const Bins& bins);
// The part of the small tier one expansion launch takes: entries [begin,
// min(end, n_small)) of q_small; big: the launch also takes the medium and
// chunked tiers.
struct SmallRange {
  uint32_t begin = 0;
  uint32_t end = 0xFFFFFFFFu;
  uint32_t big = 1;
};
// Push expansion over the queues of launch_compact (lengths read on the
// device), restricted to `range`.
int launch_expand_push(int alg, const DevGraph& g, uint32_t* val, uint64_t* out,
                       const uint32_t* q_small, const uint32_t* q_med, const uint64_t* q_chunk,
                       uint64_t item_bound, SmallRange range, const Bins& bins, DevStats* st,
                       uint32_t* stripes, bool exact, hipStream_t s, bool track_min = false);
// Bottom-up BFS step: unreached vertices look for a parent in `in`.
// cond not NULL: runs only when cond->choice (left by k_stats) asks for a
// bottom-up step; the grid is capped and wavefronts stride over the words.
int launch_pull_bfs(const DevGraph& g, uint32_t* val, const uint64_t* in, uint64_t* out,
                    uint32_t* stripes, hipStream_t s, const DevStats* cond = nullptr);
// Bottom-up SSSP / label propagation: every vertex takes the minimum over its
// in-edges from frontier vertices.  The in-edges are cut into tiles of
// pull_tile_edges(); tile_row (pull_tile_count() + 1 entries, from
// build_pull_tiles) holds the row of each tile's first edge.
uint64_t pull_tile_edges();
uint64_t pull_tile_count(uint64_t E);
int build_pull_tiles(const uint64_t* in_off, uint32_t V, uint64_t E, uint32_t* tile_row, hipStream_t s);
int launch_pull_min(int alg, const DevGraph& g, uint64_t E, const uint32_t* tile_row, uint32_t* val,
                    const uint64_t* in, uint64_t* out, uint32_t* stripes, hipStream_t s);
// Applies host-produced (vertex, value) candidates with atomic min and marks
// improved vertices in `out`.
int launch_apply_candidates(const uint32_t* ids, const uint32_t* vals, uint32_t n, uint32_t* val,
                            uint64_t* out, uint32_t* stripes, hipStream_t s);
// (vertex, value) of the vertices set in both `in` and `mask` (words words
// each) into ids / vals, their number into *count (cleared first).
int launch_gather_masked(const uint64_t* in, const uint64_t* mask, uint64_t words, const uint32_t* val, uint32_t* ids,
                         uint32_t* vals, uint32_t* count, hipStream_t s);
// Gathers (id, value) of the frontier vertices listed in `ids`.
int launch_gather_values(const uint32_t* ids, uint32_t n, const uint32_t* val, uint32_t* out,
                         hipStream_t s);
// The query's last pass over its vertices, one kernel: with acc, the sum of
// out-degrees of the vertices whose value != CG_UNREACHED (the TEPS
// numerator), added into kAccLines slots one 128-byte line apart
// (acc[k * kAccStride], zero on entry; one atomic per workgroup, the
// workgroups spread over the lines),
// zero on entry); with new_of, the values in the caller's vertex ids
// (out[i] = val[new_of[i]]) for an engine that runs on a renumbered copy.
constexpr int kAccLines = 256, kAccStride = 16;
constexpr int kAccSlots = kAccLines * kAccStride;
// Builds the graph's device out-degree array once (DevGraph::deg).
int ensure_degrees(cg_graph* g, hipStream_t s);
// Builds, once, a copy of the weights in 8 or 16 bits when every weight fits
// (DevGraph::w8 / w16); the same values, fewer bytes per relaxed edge.
int ensure_narrow_weights(cg_graph* g, hipStream_t s);
int launch_finish_query(const DevGraph& g, const uint32_t* val, unsigned long long* acc, const uint32_t* new_of,
                        uint32_t* out, hipStream_t s);

// The query's end, published like an iteration's record (m_active: the sum
// of the striped accumulators when acc is given, iter: tag) once every kernel
// queued before it has finished.
int launch_query_done(const unsigned long long* acc, DevStats* host_out, uint32_t tag, hipStream_t s);

// Waits for a stream by polling it: a blocking wait can sleep through the end
// of a short kernel and take tens of microseconds to wake.  After 2 ms of
// polling it blocks.
hipError_t stream_wait(hipStream_t s);
// One push step straight from a bitmap (cg_push_step_device): compaction and
// expansion with scratch queues owned by the caller.
int push_step_from_bitmap(int alg, const DevGraph& g, uint32_t* val, const uint64_t* in,
                          uint64_t* out, hipStream_t s);

// ---- WCC by hooking (wcc.hip) ----
int wcc_hooking(const DevGraph& g, uint32_t* parent, uint32_t sample_rounds, hipStream_t s,
                uint64_t* edges_relaxed);

// ---- PageRank (pagerank.hip) ----
struct PrPlan;  // per-graph layout for the pull kernels (tiers, relabelled CSC or source windows)
// layout: 0 automatic, 1 plain pull, 2 pull over a CSC renumbered by out-degree,
// 3 source windows of 2^seg_log2 vertices.  Automatic: seg_log2 in [1, 30]
// asks for windows, 31 for plain pull, 0 picks relabelled pull above 2^16
// vertices.
// kernel (plain / relabelled layouts): 0 automatic (row tiers), 1 row tiers, 2 edge tiles,
// 3 two source bands (relabelled layout; boundary 2^seg_log2, default 2^19).
int pr_plan_create(cg_graph* g, uint32_t layout, uint32_t seg_log2, uint32_t kernel, PrPlan** plan,
                   hipStream_t s);
void pr_plan_destroy(PrPlan* plan);
// The layout a plan uses and the bytes it streams per iteration by design.
void pr_plan_info(const PrPlan* plan, uint32_t* layout, double* stream_bytes);
int pr_run(const DevGraph& g, PrPlan* plan, int iters, float damping, float* rank, float* tmp_rank,
           float* contrib, float* contrib2, hipStream_t s, int profile, double* kernel_ms,
           uint32_t* launches);

// ---- generators and CSR construction on the GPU (gen.hip) ----
int gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed,
             uint32_t wlo, uint32_t whi, int symmetrize, cg_graph* g);
int gen_grid(uint32_t n, double drop, uint64_t seed, uint32_t wlo, uint32_t whi, cg_graph* g);
int build_transpose(cg_graph* g);
int build_from_edges(const uint32_t* src, const uint32_t* dst, const uint32_t* w, uint64_t m, uint32_t V,
                     int symmetrize, cg_graph* g);
int device_max_outdeg(cg_graph* g);
// Source vertex of every edge of a CSR (u32[E]).
int edge_sources(const uint64_t* off, uint32_t V, uint64_t E, DevBuf<uint32_t>& src, hipStream_t s);
// Renumbering by descending out-degree (ties by ascending id): perm[new] = old,
// new_of[old] = new; top_share = share of the edges leaving the top 1/64 of
// vertices (NULL: not computed).
int degree_order(cg_graph* g, DevBuf<uint32_t>& perm, DevBuf<uint32_t>& new_of, double* top_share);
// rg = g with every vertex v renamed new_of[v] (device CSR, rows sorted).
int relabel_graph(cg_graph* g, const uint32_t* new_of, cg_graph* rg);
// out[v] = in[new_of[v]]: values of the renumbered graph back in the caller's ids
int launch_renumber_back(const uint32_t* new_of, const uint32_t* in, uint32_t V, uint32_t* out, hipStream_t s);

}  // namespace cg
