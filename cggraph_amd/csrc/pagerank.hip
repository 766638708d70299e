// PageRank by pulling over the transposed graph, FP32 with a fixed summation
// order.
//
// The reference's PR is not functional (it accumulates into the source on
// uint32, my_project/CGgraph/device_algorithm.cuh:505-545); only its update
// rule is recoverable: PR(v) = (1 - d)/N + d * sum_{u -> v} PR(u)/outdeg(u),
// d = 0.85 (:543), dangling mass dropped.  contrib(u) = PR(u)/outdeg(u) is
// written by the kernel that produces PR(u), so each iteration reads it once.
//
// Four layouts, chosen once per graph (the first three pull over the
// in-edges, the fourth pushes through destination bins):
//  plain       one sweep over the transposed graph per iteration, rows tiered
//              by in-degree: one lane / one wavefront per row, and rows of
//              8192+ in-edges cut into 8192-edge chunks of one workgroup each.
//  relabelled  the same sweep over a copy of the CSC whose vertex ids are
//              renumbered by descending out-degree (ties by id).  On skewed
//              graphs a small prefix of the ids then carries most edges'
//              sources (R-MAT: the top 1/16 of vertices are the source of
//              ~85 % of edges), so the contrib values those gathers read sit
//              in a few MB of contiguous lines that stay in L2 / Infinity
//              Cache instead of being spread over the whole array.  The ranks
//              live in the new numbering while iterating and are scattered
//              back to the caller's ids once at the end.
//  segmented   cache blocking of the source side: sources are cut into
//              windows of 2^k vertices; edges are ordered by (window,
//              destination, source), so window s is a list of runs "edges of
//              destination d from window s"; a run's partial sum (sequential,
//              or a fixed wavefront tree for long runs) is added to sum[d]
//              window after window, and a final sweep turns sum into PR.
//  blocked     propagation blocking (further below): contributions are
//              streamed into destination bins and summed in 64-bit fixed
//              point in LDS -- sequential traffic instead of random gathers.
// Every layout fixes each row's summation order from the graph alone (the
// blocked one sums in integers, where the order does not matter), so results
// are bit-identical from run to run.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <vector>

#include "kernels.hpp"

namespace cg {

enum {
  kLayoutAuto = 0,
  kLayoutPlain = 1,
  kLayoutRelabel = 2,
  kLayoutSegmented = 3,
  kLayoutBlocked = 4,
  kLayoutPairs = 5
};

struct PrPlan {
  int layout = kLayoutPlain;
  double stream_bytes = 0;  // per iteration, by design (edge and pair terms + 16 B per vertex)
  // plain and relabelled pull: tiers over the rows of the CSC in use (one
  // set per pass; the two-band sweep has two passes)
  struct Tiers {
    uint32_t n_med = 0, n_high = 0;
    DevBuf<uint32_t> med, high;
    // high rows cut into chunks of kHigh edges, one workgroup each; a row's
    // chunk sums are added in chunk order
    uint32_t n_chunks = 0;
    DevBuf<uint64_t> chunk_lo, chunk_hi;
    DevBuf<uint32_t> row_chunk;  // first chunk of high row i (n_high + 1 entries)
    DevBuf<float> partial;
  } tiers[2];
  bool banded = false;
  uint32_t band = 0;
  DevBuf<uint32_t> hot;
  DevBuf<float> part;
  DevBuf<uint32_t> deg;  // out-degree of each vertex (in the numbering in use)
  // edge tiles over the CSC in use (instead of the row tiers)
  bool tiled = false;
  uint64_t edges = 0;
  uint64_t ntiles = 0;
  uint32_t lead_rows = 0;  // rows before the first edge (no in-edges)
  DevBuf<uint32_t> tile_row;
  DevBuf<float> carry, head_part;
  // relabelled pull
  DevBuf<uint32_t> perm;  // new id -> caller's id
  DevBuf<uint64_t> r_off;
  DevBuf<uint32_t> r_src;
  // segmented pull
  uint32_t shift = 0, nseg = 0;
  uint64_t R = 0;
  DevBuf<uint32_t> src, run_dst, run_off, long_runs;
  std::vector<uint32_t> seg_run, seg_long;
  DevBuf<float> sum;
  // propagation blocking: out-edges binned by destination block
  uint32_t pb_nb = 0;                   // destination bins (= source chunks) of 2^13 vertices
  uint32_t pb_t1 = 0, pb_t2 = 0, pb_multi = 0;
  bool pb_narrow = false;               // every edge index fits 31 bits: 32-bit deltas in LDS
  DevBuf<uint32_t> pb_ent;              // chunk-major: source within chunk | bin << 13
  DevBuf<uint16_t> pb_dloc;             // bin-major: destination within bin
  DevBuf<int64_t> pb_delta;             // [chunk][bin]: bin-major minus chunk-major index of a run
  DevBuf<float> pb_val;                 // bin-major contributions, rewritten every iteration
  DevBuf<uint32_t> pb_t1_chunk, pb_t2_bin, pb_multi_bins;
  DevBuf<uint64_t> pb_t1_lo, pb_t1_hi, pb_t2_lo, pb_t2_hi;
  DevBuf<uint32_t> pb_t2_multi;         // tile's bin is cut into several tiles
  DevBuf<unsigned long long> pb_sum;    // fixed-point sums of the bins cut into several tiles
  // propagation blocking over (source, bin) pairs (layout 5): pb_ent and
  // pb_val hold one entry per pair, pb_dloc one per edge (destination within
  // its bin | 0x8000 on a pair's first edge), pb_gbase the pair of every 64th
  // edge; the graph is renumbered by descending out-degree (perm)
  DevBuf<uint32_t> pb_gbase;
  uint64_t pb_pairs = 0;
  // bins cut into several gather tiles: each tile writes its 8192 sums to a
  // row of its own (pb_t2_multi = row + 1), bin i's rows are
  // [pb_mrow[i], pb_mrow[i + 1]); one pass adds them up
  DevBuf<uint32_t> pb_mrow;
  DevBuf<unsigned long long> pb_part;
};

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kMed = 32, kHigh = 8192;
constexpr uint32_t kLongRun = 64;
constexpr int kMedUnroll = 8;  // gathers in flight per lane in the wavefront-per-row tier

// The in-edges the pull kernels read: row v's sources are src[off[v]..off[v+1]).
struct Csc {
  const uint64_t* off;
  const uint32_t* src;
  const uint32_t* deg;  // out-degree, for contrib = rank / deg
  uint32_t V;
  const uint32_t* hot;  // two-band sweep: sources of row v below the band boundary (a prefix)
};

unsigned grid_of(uint64_t n) {
  uint64_t b = (n + kBlock - 1) / kBlock;
  if (b < 1) b = 1;
  return (unsigned)(b < 65536 ? b : 65536);
}

uint32_t bits_for(uint64_t V) {
  uint32_t b = 1;
  while ((1ull << b) < V) ++b;
  return b;
}

// In-edge source ids, read once per sweep.  Ordinary (cached) loads: the
// non-temporal form measured 25 % slower per iteration at 2^19 vertices and
// 2-3 % at 2^25 (R-MAT, renumbered pull).
__device__ __forceinline__ uint32_t ld_src(const uint32_t* p) { return *p; }


__global__ void k_out_degree(const uint64_t* __restrict__ off, uint32_t V, uint32_t* __restrict__ deg) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x)
    deg[v] = (uint32_t)(off[v + 1] - off[v]);
}

__device__ __forceinline__ void finish(uint32_t v, float sum, float base, float damping,
                                       const uint32_t* __restrict__ deg, float* __restrict__ rank,
                                       float* __restrict__ contrib) {
  const float r = base + damping * sum;
  rank[v] = r;
  const uint32_t d = deg[v];
  contrib[v] = d ? r / (float)d : 0.0f;
}

// Row v's in-edges swept by pass P: 0 the whole row; in the two-band sweep,
// 1 the sources below the band boundary (a prefix, rows being sorted) and 2
// the rest.
template <int P>
__device__ __forceinline__ void row_range(const Csc& g, uint32_t v, uint64_t& s, uint64_t& e) {
  s = g.off[v];
  e = g.off[v + 1];
  if constexpr (P == 1) e = s + g.hot[v];
  else if constexpr (P == 2) s += g.hot[v];
}

// Pass 1 keeps the partial sum; pass 2 adds the rest to it, so every row is
// summed prefix first, in a fixed order.
template <int P>
__device__ __forceinline__ void emit(const Csc& g, uint32_t v, float sum, float base, float damping,
                                     float* __restrict__ part, float* __restrict__ rank,
                                     float* __restrict__ contrib) {
  if constexpr (P == 1) part[v] = sum;
  else if constexpr (P == 2) finish(v, part[v] + sum, base, damping, g.deg, rank, contrib);
  else finish(v, sum, base, damping, g.deg, rank, contrib);
}

// Rows with fewer than kMed in-edges in this pass: one lane per row, sources
// in order.
template <int P>
__global__ __launch_bounds__(kBlock) void k_pr_low(Csc g, const float* __restrict__ contrib_in, float base,
                                                   float damping, float* __restrict__ part, float* __restrict__ rank,
                                                   float* __restrict__ contrib_out) {
  const uint64_t v = blockIdx.x * (uint64_t)kBlock + threadIdx.x;
  if (v >= g.V) return;
  uint64_t s, e;
  row_range<P>(g, (uint32_t)v, s, e);
  if (e - s >= kMed) return;
  float sum = 0.0f;
  for (uint64_t k = s; k < e; ++k) sum += contrib_in[ld_src(g.src + k)];
  emit<P>(g, (uint32_t)v, sum, base, damping, part, rank, contrib_out);
}

// Lane-strided partial sum over [s, e): the lane adds sources s+lane,
// s+lane+stride, ... in that order; four sources are fetched ahead so four
// gathers are in flight per lane.
template <int STRIDE, int U = 4>
__device__ __forceinline__ float strided_sum(const uint32_t* __restrict__ src, const float* __restrict__ contrib,
                                             uint64_t s, uint64_t e, uint32_t lane) {
  float sum = 0.0f;
  uint64_t k = s + lane;
  for (; k + (U - 1) * STRIDE < e; k += U * STRIDE) {
    uint32_t u[U];
    float c[U];
#pragma unroll
    for (int j = 0; j < U; ++j) u[j] = ld_src(src + k + j * STRIDE);
#pragma unroll
    for (int j = 0; j < U; ++j) c[j] = contrib[u[j]];
#pragma unroll
    for (int j = 0; j < U; ++j) sum += c[j];
  }
  for (; k < e; k += STRIDE) sum += contrib[ld_src(src + k)];
  return sum;
}

template <int P>
__global__ __launch_bounds__(kBlock) void k_pr_med(Csc g, const uint32_t* __restrict__ list, uint32_t n,
                                                   const float* __restrict__ contrib_in, float base,
                                                   float damping, float* __restrict__ part, float* __restrict__ rank,
                                                   float* __restrict__ contrib_out) {
  const uint64_t i = blockIdx.x * (uint64_t)(kBlock / 64) + threadIdx.x / 64;
  if (i >= n) return;  // wave-uniform
  const uint32_t lane = __lane_id();
  const uint32_t v = list[i];
  uint64_t s, e;
  row_range<P>(g, v, s, e);
  float sum = strided_sum<64, kMedUnroll>(g.src, contrib_in, s, e, lane);
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) sum += __shfl_xor(sum, d, 64);
  if (lane == 0) emit<P>(g, v, sum, base, damping, part, rank, contrib_out);
}

// Edge ranges of the high rows in pass P (for the chunk table).
template <int P>
__global__ void k_high_ranges(Csc g, const uint32_t* __restrict__ list, uint32_t n, uint64_t* __restrict__ lo,
                              uint64_t* __restrict__ hi) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t s, e;
    row_range<P>(g, list[i], s, e);
    lo[i] = s;
    hi[i] = e;
  }
}

// One workgroup per chunk of a high row: fixed lane-strided partial sums and
// a fixed tree.
__global__ __launch_bounds__(kBlock) void k_pr_chunk(const uint32_t* __restrict__ src,
                                                     const uint64_t* __restrict__ chunk_lo,
                                                     const uint64_t* __restrict__ chunk_hi,
                                                     const float* __restrict__ contrib_in,
                                                     float* __restrict__ partial) {
  __shared__ float acc[kBlock];
  const uint64_t c = blockIdx.x;
  acc[threadIdx.x] = strided_sum<kBlock>(src, contrib_in, chunk_lo[c], chunk_hi[c], threadIdx.x);
  __syncthreads();
  for (int w = kBlock / 2; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) acc[threadIdx.x] += acc[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[c] = acc[0];
}

// One lane per high row: its chunk sums in chunk order.
template <int P>
__global__ void k_pr_chunk_finish(Csc g, const uint32_t* __restrict__ list, uint32_t n,
                                  const uint32_t* __restrict__ row_chunk, const float* __restrict__ partial,
                                  float base, float damping, float* __restrict__ part, float* __restrict__ rank,
                                  float* __restrict__ contrib_out) {
  const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  float sum = 0.0f;
  for (uint32_t c = row_chunk[i]; c < row_chunk[i + 1]; ++c) sum += partial[c];
  emit<P>(g, list[i], sum, base, damping, part, rank, contrib_out);
}

// hot[v] = sources of row v below 2^band (rows sorted ascending)
__global__ void k_band_split(const uint64_t* __restrict__ off, const uint32_t* __restrict__ src, uint32_t V,
                             uint32_t band, uint32_t* __restrict__ hot) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t lo = off[v], hi = off[v + 1];
    const uint64_t s = lo;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if ((src[mid] >> band) == 0) lo = mid + 1;
      else hi = mid;
    }
    hot[v] = (uint32_t)(lo - s);
  }
}

// Row lists of the medium and high tiers of pass P; one atomic per list per
// wavefront.
template <int P>
__global__ void k_classify(Csc g, uint32_t* __restrict__ med, uint32_t* __restrict__ high,
                           uint32_t* __restrict__ cnt) {
  const uint32_t lane = __lane_id();
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t b = blockIdx.x * (uint64_t)blockDim.x + (threadIdx.x & ~63u); b < g.V; b += stride) {
    const uint64_t v = b + lane;
    uint64_t s = 0, e = 0;
    if (v < g.V) row_range<P>(g, (uint32_t)v, s, e);
    const uint64_t d = e - s;
    const bool hi = d >= kHigh, md = d >= kMed && d < kHigh;
    const uint64_t bh = __ballot(hi), bm = __ballot(md);
    uint32_t base_h = 0, base_m = 0;
    if (lane == 0) {
      if (bh) base_h = atomicAdd(cnt + 1, (uint32_t)__popcll(bh));
      if (bm) base_m = atomicAdd(cnt, (uint32_t)__popcll(bm));
    }
    base_h = __shfl(base_h, 0, 64);
    base_m = __shfl(base_m, 0, 64);
    const uint64_t below = (1ull << lane) - 1;
    if (hi) high[base_h + __popcll(bh & below)] = (uint32_t)v;
    if (md) med[base_m + __popcll(bm & below)] = (uint32_t)v;
  }
}

__global__ void k_pr_init(uint32_t V, const uint32_t* __restrict__ deg, float r0, float* __restrict__ rank,
                          float* __restrict__ contrib) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    rank[v] = r0;
    const uint32_t d = deg[v];
    contrib[v] = d ? r0 / (float)d : 0.0f;
  }
}

// ---- edge tiles (merge-based pull) ----
//
// A workgroup takes kTileEdges consecutive in-edges of the CSC, whatever rows
// they belong to; thread j takes kTileK consecutive edges, loads their sources
// with two 16-byte loads and has all kTileK gathers in flight at once, then
// walks its edges adding into the current row.  A row that ends inside the
// thread's edges is finished there; the part of a row that continues across
// threads is combined by a segmented scan over the workgroup (fixed shape),
// and a row that continues across workgroups is finished by k_pr_fixup, which
// adds the workgroups' carries in workgroup order.  Every summation order is a
// function of the CSC alone.  (Merge-based sparse matrix-vector product,
// Merrill and Garland, SC'16.)

constexpr int kTileThreads = 256;
constexpr int kTileK = 8;
constexpr uint64_t kTileEdges = (uint64_t)kTileThreads * kTileK;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// tile_row[t] = row holding edge t * kTileEdges (largest r with off[r] <= it),
// tile_row[ntiles] = V.
__global__ void k_tile_rows(const uint64_t* __restrict__ off, uint32_t V, uint64_t ntiles,
                            uint32_t* __restrict__ tile_row) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t <= ntiles;
       t += (uint64_t)gridDim.x * blockDim.x) {
    if (t == ntiles) {
      tile_row[t] = V;
      continue;
    }
    const uint64_t e = t * kTileEdges;
    uint32_t lo = 0, hi = V;  // off[V] = E > e
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo + 1) / 2;
      if (off[mid] <= e) lo = mid;
      else hi = mid - 1;
    }
    tile_row[t] = lo;
  }
}

__global__ __launch_bounds__(kTileThreads) void k_pr_tile(Csc g, uint64_t E, const uint32_t* __restrict__ tile_row,
                                                          const float* __restrict__ contrib_in, float base,
                                                          float damping, float* __restrict__ rank,
                                                          float* __restrict__ contrib_out, float* __restrict__ carry,
                                                          float* __restrict__ head_part) {
  __shared__ float s_val[kTileThreads];
  __shared__ uint32_t s_flag[kTileThreads / 64];
  __shared__ float s_wval[kTileThreads / 64];
  const uint64_t t = blockIdx.x;
  const uint64_t tile0 = t * kTileEdges;
  const uint64_t e0 = tile0 + (uint64_t)threadIdx.x * kTileK;
  const uint32_t lane = __lane_id(), wave = threadIdx.x / 64;
  const bool has = e0 < E;
  const uint32_t nk = has ? (uint32_t)((E - e0) < (uint64_t)kTileK ? (E - e0) : kTileK) : 0u;

  // row holding edge e0
  uint32_t head = g.V;
  if (has) {
    uint32_t lo = tile_row[t], hi = tile_row[t + 1];
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo + 1) / 2;
      if (g.off[mid] <= e0) lo = mid;
      else hi = mid - 1;
    }
    head = lo;
  }
  // sources and gathers, all in flight before the first add
  float c[kTileK];
  if (nk == (uint32_t)kTileK) {
    const u32x4* p = reinterpret_cast<const u32x4*>(g.src + e0);
    const u32x4 a = p[0], b = p[1];
    c[0] = contrib_in[a.x];
    c[1] = contrib_in[a.y];
    c[2] = contrib_in[a.z];
    c[3] = contrib_in[a.w];
    c[4] = contrib_in[b.x];
    c[5] = contrib_in[b.y];
    c[6] = contrib_in[b.z];
    c[7] = contrib_in[b.w];
  } else {
#pragma unroll
    for (int i = 0; i < kTileK; ++i) c[i] = (uint32_t)i < nk ? contrib_in[g.src[e0 + i]] : 0.0f;
  }
  const bool inherited = has && g.off[head] < e0;  // the head row began before this thread
  uint32_t r = head;
  uint64_t next_end = has ? g.off[r + 1] : ~0ull;
  float acc = 0.0f, head_acc = 0.0f;
  bool first = true, head_done = false;
#pragma unroll
  for (int i = 0; i < kTileK; ++i) {
    if ((uint32_t)i < nk) {
      acc += c[i];
      const uint64_t e = e0 + i + 1;
      while (next_end == e) {  // row r ends after this edge (and empty rows after it)
        if (first && inherited) {
          head_acc = acc;
          head_done = true;
        } else {
          finish(r, acc, base, damping, g.deg, rank, contrib_out);
        }
        first = false;
        acc = 0.0f;
        ++r;
        next_end = r < g.V ? g.off[r + 1] : ~0ull;
      }
    }
  }
  // segmented inclusive scan of the tail parts: a thread whose tail row began
  // inside its own edges starts a new segment
  bool f = !(has && r == head && inherited);
  float v = acc;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const float vu = __shfl_up(v, d, 64);
    const int fu = __shfl_up((int)f, d, 64);
    if (lane >= (uint32_t)d) {
      if (!f) v = vu + v;
      f = f || fu;
    }
  }
  if (lane == 63) {
    s_flag[wave] = f;
    s_wval[wave] = v;
  }
  __syncthreads();
  if (!f) {
    // no segment start between the workgroup's first thread and this one in
    // this wavefront: prepend the earlier wavefronts' open segment
    float pre = 0.0f;
    for (uint32_t w = 0; w < wave; ++w) pre = s_flag[w] ? s_wval[w] : pre + s_wval[w];
    v = pre + v;
  }
  s_val[threadIdx.x] = v;
  __syncthreads();
  const float carry_in = threadIdx.x ? s_val[threadIdx.x - 1] : 0.0f;
  if (head_done) {
    const float total = carry_in + head_acc;
    if (g.off[head] < tile0) head_part[t] = total;  // began in an earlier workgroup
    else finish(head, total, base, damping, g.deg, rank, contrib_out);
  }
  if (threadIdx.x == kTileThreads - 1) carry[t] = v;
}

// Rows that began in an earlier workgroup, finished in the workgroup where
// they end: carries of the workgroups they span, in order, then that
// workgroup's part.
__global__ void k_pr_fixup(Csc g, uint64_t E, uint64_t ntiles, const uint32_t* __restrict__ tile_row,
                           const float* __restrict__ carry, const float* __restrict__ head_part, float base,
                           float damping, float* __restrict__ rank, float* __restrict__ contrib_out) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < ntiles;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t r = tile_row[t];
    const uint64_t tile0 = t * kTileEdges;
    const uint64_t tile1 = tile0 + kTileEdges < E ? tile0 + kTileEdges : E;
    const uint64_t s = g.off[r], e = g.off[r + 1];
    if (s >= tile0 || e > tile1) continue;  // not inherited, or does not end here
    float sum = 0.0f;
    for (uint64_t u = s / kTileEdges; u < t; ++u) sum += carry[u];
    finish(r, sum + head_part[t], base, damping, g.deg, rank, contrib_out);
  }
}

// rows [0, n) have no in-edges and precede the first edge
__global__ void k_pr_lead(uint32_t n, const uint32_t* __restrict__ deg, float base, float* __restrict__ rank,
                          float* __restrict__ contrib_out) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n;
       v += (uint64_t)gridDim.x * blockDim.x)
    finish((uint32_t)v, 0.0f, base, 0.0f, deg, rank, contrib_out);
}

__global__ void k_unpermute(const uint32_t* __restrict__ perm, const float* __restrict__ in, uint32_t V,
                            float* __restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < V;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[perm[i]] = in[i];
}

// ---- segmented layout construction ----

// key = window(u) << 2b | v << b | u for edge u -> v
__global__ void k_seg_keys(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst, uint64_t E,
                           uint32_t b, uint32_t shift, uint64_t* __restrict__ keys) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < E;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t u = src[e];
    keys[e] = ((u >> shift) << (2 * b)) | ((uint64_t)dst[e] << b) | u;
  }
}

__global__ void k_seg_flags(const uint64_t* __restrict__ keys, uint64_t E, uint32_t b,
                            uint32_t* __restrict__ src_out, uint32_t* __restrict__ flags) {
  const uint64_t mask = (1ull << b) - 1;
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < E;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t k = keys[e];
    src_out[e] = (uint32_t)(k & mask);
    flags[e] = (e == 0 || (keys[e - 1] >> b) != (k >> b)) ? 1u : 0u;
  }
}

__global__ void k_seg_runs(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ flags,
                           const uint32_t* __restrict__ pos, uint64_t E, uint32_t b,
                           uint32_t* __restrict__ run_dst, uint32_t* __restrict__ run_off) {
  const uint64_t mask = (1ull << b) - 1;
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < E;
       e += (uint64_t)gridDim.x * blockDim.x) {
    if (flags[e]) {
      run_dst[pos[e]] = (uint32_t)((keys[e] >> b) & mask);
      run_off[pos[e]] = (uint32_t)e;
    }
  }
}

// seg_run[s] = index of the first run of window s (runs are window-major).
__global__ void k_seg_bounds(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ pos, uint64_t E,
                             uint64_t R, uint32_t b, uint32_t nseg, uint32_t* __restrict__ seg_run) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > nseg) return;
  const uint64_t target = (uint64_t)s << (2 * b);
  uint64_t lo = 0, hi = E;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (keys[mid] < target) lo = mid + 1;
    else hi = mid;
  }
  seg_run[s] = lo < E ? pos[lo] : (uint32_t)R;
}

__global__ void k_long_flags(const uint32_t* __restrict__ run_off, uint64_t R, uint32_t* __restrict__ flags) {
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < R;
       r += (uint64_t)gridDim.x * blockDim.x)
    flags[r] = (run_off[r + 1] - run_off[r]) >= kLongRun ? 1u : 0u;
}

__global__ void k_long_scatter(const uint32_t* __restrict__ flags, const uint32_t* __restrict__ pos, uint64_t R,
                               uint32_t* __restrict__ long_runs) {
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < R;
       r += (uint64_t)gridDim.x * blockDim.x)
    if (flags[r]) long_runs[pos[r]] = (uint32_t)r;
}

// ---- segmented iteration ----

// One lane per run of the window; runs of kLongRun edges or more are left to
// k_seg_long.
__global__ __launch_bounds__(kBlock) void k_seg_short(const uint32_t* __restrict__ src,
                                                      const uint32_t* __restrict__ run_dst,
                                                      const uint32_t* __restrict__ run_off, uint32_t r0, uint32_t r1,
                                                      const float* __restrict__ contrib, float* __restrict__ sum) {
  const uint32_t r = r0 + blockIdx.x * kBlock + threadIdx.x;
  if (r >= r1) return;
  const uint32_t s = __builtin_nontemporal_load(run_off + r), e = __builtin_nontemporal_load(run_off + r + 1);
  if (e - s >= kLongRun) return;
  float p = 0.0f;
  for (uint32_t k = s; k < e; ++k) p += contrib[__builtin_nontemporal_load(src + k)];
  const uint32_t d = __builtin_nontemporal_load(run_dst + r);
  sum[d] += p;
}

// One wavefront per long run: lane j sums edges j, j+64, ... in order, then a
// fixed xor butterfly.
__global__ __launch_bounds__(kBlock) void k_seg_long(const uint32_t* __restrict__ src,
                                                     const uint32_t* __restrict__ run_dst,
                                                     const uint32_t* __restrict__ run_off,
                                                     const uint32_t* __restrict__ long_runs, uint32_t l0, uint32_t l1,
                                                     const float* __restrict__ contrib, float* __restrict__ sum) {
  const uint32_t i = l0 + blockIdx.x * (kBlock / 64) + threadIdx.x / 64;
  if (i >= l1) return;  // wave-uniform
  const uint32_t lane = __lane_id();
  const uint32_t r = long_runs[i];
  const uint32_t s = run_off[r], e = run_off[r + 1];
  float p = 0.0f;
  for (uint32_t k = s + lane; k < e; k += 64) p += contrib[__builtin_nontemporal_load(src + k)];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) p += __shfl_xor(p, d, 64);
  if (lane == 0) sum[run_dst[r]] += p;
}

__global__ void k_seg_finish(uint32_t V, const uint32_t* __restrict__ deg, float base, float damping,
                             float* __restrict__ sum, float* __restrict__ rank, float* __restrict__ contrib) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    finish((uint32_t)v, sum[v], base, damping, deg, rank, contrib);
    sum[v] = 0.0f;
  }
}

#define PR_LAUNCH_OK()                                                                  \
  do {                                                                                  \
    hipError_t e_ = hipGetLastError();                                                  \
    if (e_ != hipSuccess)                                                               \
      return set_error(CG_ERR_HIP, "PageRank kernel launch: %s (%s:%d)", hipGetErrorString(e_), \
                       __FILE__, __LINE__);                                             \
  } while (0)

// med / high row lists of pass P over a CSC (any order: each row's sum is
// fixed by the row).
template <int P>
int build_tiers(const Csc& g, PrPlan::Tiers& t, hipStream_t s) {
  DevBuf<uint32_t> cnt;
  CG_TRY(cnt.ensure(2));
  CG_TRY(t.med.ensure(g.V ? g.V : 1));
  CG_TRY(t.high.ensure(g.V ? g.V : 1));
  CG_HIP(hipMemsetAsync(cnt.p, 0, 2 * sizeof(uint32_t), s));
  if (g.V) {
    hipLaunchKernelGGL(k_classify<P>, dim3(grid_of(g.V)), dim3(kBlock), 0, s, g, t.med.p, t.high.p, cnt.p);
    PR_LAUNCH_OK();
  }
  uint32_t h[2] = {0, 0};
  CG_HIP(hipMemcpyAsync(h, cnt.p, sizeof(h), hipMemcpyDeviceToHost, s));
  CG_HIP(hipStreamSynchronize(s));
  t.n_med = h[0];
  t.n_high = h[1];
  if (!t.n_high) return CG_OK;
  // chunk table of the high rows
  DevBuf<uint64_t> lo, hi;
  CG_TRY(lo.ensure(t.n_high));
  CG_TRY(hi.ensure(t.n_high));
  hipLaunchKernelGGL(k_high_ranges<P>, dim3(grid_of(t.n_high)), dim3(kBlock), 0, s, g, t.high.p, t.n_high, lo.p,
                     hi.p);
  PR_LAUNCH_OK();
  std::vector<uint64_t> hl(t.n_high), hh(t.n_high);
  CG_HIP(hipMemcpyAsync(hl.data(), lo.p, t.n_high * 8, hipMemcpyDeviceToHost, s));
  CG_HIP(hipMemcpyAsync(hh.data(), hi.p, t.n_high * 8, hipMemcpyDeviceToHost, s));
  CG_HIP(hipStreamSynchronize(s));
  std::vector<uint64_t> clo, chi;
  std::vector<uint32_t> first(t.n_high + 1);
  for (uint32_t i = 0; i < t.n_high; ++i) {
    first[i] = (uint32_t)clo.size();
    for (uint64_t a = hl[i]; a < hh[i]; a += kHigh) {
      clo.push_back(a);
      chi.push_back(a + kHigh < hh[i] ? a + kHigh : hh[i]);
    }
  }
  first[t.n_high] = (uint32_t)clo.size();
  t.n_chunks = (uint32_t)clo.size();
  CG_TRY(t.chunk_lo.ensure(clo.size() ? clo.size() : 1));
  CG_TRY(t.chunk_hi.ensure(chi.size() ? chi.size() : 1));
  CG_TRY(t.row_chunk.ensure(first.size()));
  CG_TRY(t.partial.ensure(clo.size() ? clo.size() : 1));
  if (!clo.empty()) {
    CG_HIP(hipMemcpyAsync(t.chunk_lo.p, clo.data(), clo.size() * 8, hipMemcpyHostToDevice, s));
    CG_HIP(hipMemcpyAsync(t.chunk_hi.p, chi.data(), chi.size() * 8, hipMemcpyHostToDevice, s));
  }
  CG_HIP(hipMemcpyAsync(t.row_chunk.p, first.data(), first.size() * 4, hipMemcpyHostToDevice, s));
  CG_HIP(hipStreamSynchronize(s));
  return CG_OK;
}

// Two-band sweep over the relabelled CSC: pass 1 reads only the sources
// below 2^band, whose contributions (4 * 2^band bytes) stay in L2 for the
// whole pass; pass 2 reads the rest.
int build_bands(uint32_t V, uint32_t band, PrPlan* p, hipStream_t s) {
  CG_TRY(p->hot.ensure(V ? V : 1));
  CG_TRY(p->part.ensure(V ? V : 1));
  if (V) {
    hipLaunchKernelGGL(k_band_split, dim3(grid_of(V)), dim3(kBlock), 0, s, p->r_off.p, p->r_src.p, V, band,
                       p->hot.p);
    PR_LAUNCH_OK();
  }
  const Csc g{p->r_off.p, p->r_src.p, p->deg.p, V, p->hot.p};
  CG_TRY(build_tiers<1>(g, p->tiers[0], s));
  CG_TRY(build_tiers<2>(g, p->tiers[1], s));
  p->banded = true;
  p->band = band;
  return CG_OK;
}

int build_degrees(const DevGraph& g, PrPlan* p, hipStream_t s) {
  CG_TRY(p->deg.ensure(g.V ? g.V : 1));
  if (g.V) {
    hipLaunchKernelGGL(k_out_degree, dim3(grid_of(g.V)), dim3(kBlock), 0, s, g.off, g.V, p->deg.p);
    PR_LAUNCH_OK();
  }
  return CG_OK;
}

int build_tiles(const uint64_t* off, uint32_t V, uint64_t E, PrPlan* p, hipStream_t s) {
  p->edges = E;
  p->ntiles = (E + kTileEdges - 1) / kTileEdges;
  CG_TRY(p->tile_row.ensure(p->ntiles + 1));
  CG_TRY(p->carry.ensure(p->ntiles ? p->ntiles : 1));
  CG_TRY(p->head_part.ensure(p->ntiles ? p->ntiles : 1));
  hipLaunchKernelGGL(k_tile_rows, dim3(grid_of(p->ntiles + 1)), dim3(kBlock), 0, s, off, V, p->ntiles,
                     p->tile_row.p);
  PR_LAUNCH_OK();
  uint32_t lead = 0;
  CG_HIP(hipMemcpyAsync(&lead, p->tile_row.p, 4, hipMemcpyDeviceToHost, s));
  CG_HIP(hipStreamSynchronize(s));
  p->lead_rows = lead;
  p->tiled = true;
  return CG_OK;
}

// The CSC of the graph renumbered by descending out-degree.
int build_relabelled(cg_graph* g, PrPlan* p, hipStream_t s) {
  const uint32_t V = g->V;
  cg_graph rg;
  {
    DevBuf<uint32_t> new_of;
    CG_TRY(degree_order(g, p->perm, new_of, nullptr));
    CG_TRY(relabel_graph(g, new_of.p, &rg));
  }
  CG_TRY(build_transpose(&rg));
  const DevGraph rdg{rg.d_off.p, rg.d_dst.p, nullptr, rg.d_in_off.p, rg.d_in_src.p, V};
  CG_TRY(build_degrees(rdg, p, s));
  CG_HIP(hipStreamSynchronize(s));
  p->r_off = std::move(rg.d_in_off);
  p->r_src = std::move(rg.d_in_src);
  p->layout = kLayoutRelabel;
  return CG_OK;
}

int build_segmented(const DevGraph& g, uint64_t E, uint32_t shift, PrPlan* p, hipStream_t s) {
  const uint32_t V = g.V;
  const uint32_t b = bits_for(V);
  const uint32_t nseg = (uint32_t)(((uint64_t)V + (1ull << shift) - 1) >> shift);
  const uint32_t segbits = bits_for(nseg) + 1;
  p->shift = shift;
  p->nseg = nseg;
  DevBuf<uint64_t> keys, alt;
  CG_TRY(keys.ensure(E));
  CG_TRY(alt.ensure(E));
  {
    DevBuf<uint32_t> srcv;
    CG_TRY(edge_sources(g.off, V, E, srcv, s));
    hipLaunchKernelGGL(k_seg_keys, dim3(grid_of(E)), dim3(kBlock), 0, s, srcv.p, g.dst, E, b, shift, keys.p);
    PR_LAUNCH_OK();
    CG_HIP(hipStreamSynchronize(s));
  }
  rocprim::double_buffer<uint64_t> db(keys.p, alt.p);
  {
    size_t tb = 0;
    CG_HIP(rocprim::radix_sort_keys(nullptr, tb, db, E, 0, 2 * b + segbits, s));
    DevBuf<uint8_t> tmp;
    CG_TRY(tmp.ensure(tb));
    CG_HIP(rocprim::radix_sort_keys(tmp.p, tb, db, E, 0, 2 * b + segbits, s));
    CG_HIP(hipStreamSynchronize(s));
  }
  uint64_t* sorted = db.current();
  DevBuf<uint32_t> flags, pos;
  CG_TRY(p->src.ensure(E));
  CG_TRY(flags.ensure(E));
  CG_TRY(pos.ensure(E));
  hipLaunchKernelGGL(k_seg_flags, dim3(grid_of(E)), dim3(kBlock), 0, s, sorted, E, b, p->src.p, flags.p);
  PR_LAUNCH_OK();
  size_t sb = 0;
  CG_HIP(rocprim::exclusive_scan(nullptr, sb, flags.p, pos.p, 0u, E, rocprim::plus<uint32_t>(), s));
  DevBuf<uint8_t> st;
  CG_TRY(st.ensure(sb));
  CG_HIP(rocprim::exclusive_scan(st.p, sb, flags.p, pos.p, 0u, E, rocprim::plus<uint32_t>(), s));
  uint32_t last_pos = 0, last_flag = 0;
  CG_HIP(hipMemcpyAsync(&last_pos, pos.p + E - 1, 4, hipMemcpyDeviceToHost, s));
  CG_HIP(hipMemcpyAsync(&last_flag, flags.p + E - 1, 4, hipMemcpyDeviceToHost, s));
  CG_HIP(hipStreamSynchronize(s));
  const uint64_t R = (uint64_t)last_pos + last_flag;
  p->R = R;
  CG_TRY(p->run_dst.ensure(R));
  CG_TRY(p->run_off.ensure(R + 1));
  hipLaunchKernelGGL(k_seg_runs, dim3(grid_of(E)), dim3(kBlock), 0, s, sorted, flags.p, pos.p, E, b, p->run_dst.p,
                     p->run_off.p);
  PR_LAUNCH_OK();
  const uint32_t Eu = (uint32_t)E;
  CG_HIP(hipMemcpyAsync(p->run_off.p + R, &Eu, 4, hipMemcpyHostToDevice, s));
  DevBuf<uint32_t> segb;
  CG_TRY(segb.ensure(nseg + 1));
  hipLaunchKernelGGL(k_seg_bounds, dim3((nseg + 1 + 255) / 256), dim3(256), 0, s, sorted, pos.p, E, R, b, nseg,
                     segb.p);
  PR_LAUNCH_OK();
  p->seg_run.resize(nseg + 1);
  CG_HIP(hipMemcpyAsync(p->seg_run.data(), segb.p, (nseg + 1) * 4, hipMemcpyDeviceToHost, s));
  CG_HIP(hipStreamSynchronize(s));
  // long runs, in run order
  hipLaunchKernelGGL(k_long_flags, dim3(grid_of(R)), dim3(kBlock), 0, s, p->run_off.p, R, flags.p);
  PR_LAUNCH_OK();
  size_t sb2 = 0;
  CG_HIP(rocprim::exclusive_scan(nullptr, sb2, flags.p, pos.p, 0u, R, rocprim::plus<uint32_t>(), s));
  CG_TRY(st.ensure(sb2));
  CG_HIP(rocprim::exclusive_scan(st.p, sb2, flags.p, pos.p, 0u, R, rocprim::plus<uint32_t>(), s));
  CG_HIP(hipMemcpyAsync(&last_pos, pos.p + R - 1, 4, hipMemcpyDeviceToHost, s));
  CG_HIP(hipMemcpyAsync(&last_flag, flags.p + R - 1, 4, hipMemcpyDeviceToHost, s));
  CG_HIP(hipStreamSynchronize(s));
  const uint64_t NL = (uint64_t)last_pos + last_flag;
  CG_TRY(p->long_runs.ensure(NL ? NL : 1));
  hipLaunchKernelGGL(k_long_scatter, dim3(grid_of(R)), dim3(kBlock), 0, s, flags.p, pos.p, R, p->long_runs.p);
  PR_LAUNCH_OK();
  // window boundaries in the long-run list: first long run with id >= seg_run[w]
  std::vector<uint32_t> hl(NL);
  if (NL) CG_HIP(hipMemcpyAsync(hl.data(), p->long_runs.p, NL * 4, hipMemcpyDeviceToHost, s));
  CG_HIP(hipStreamSynchronize(s));
  p->seg_long.resize(nseg + 1);
  for (uint32_t w = 0; w <= nseg; ++w)
    p->seg_long[w] = (uint32_t)(std::lower_bound(hl.begin(), hl.end(), p->seg_run[w]) - hl.begin());
  CG_TRY(p->sum.ensure(V));
  CG_HIP(hipMemsetAsync(p->sum.p, 0, (size_t)V * sizeof(float), s));
  CG_TRY(build_degrees(g, p, s));
  CG_HIP(hipStreamSynchronize(s));
  p->layout = kLayoutSegmented;
  return CG_OK;
}

// ---- propagation blocking (layout 4) ----
//
// Beamer, Asanovic, Patterson, "Reducing PageRank Communication via
// Propagation Blocking" (IPDPS 2017), with the accumulation in 64-bit fixed
// point so that it does not depend on the order of the additions.
// Sources are cut into chunks and destinations into bins of 2^13 vertices.
// The out-edges are stored twice, once per pass:
//   ent  (chunk-major: chunk, bin, then any fixed order)  source within its
//        chunk | bin << 13
//   dloc (bin-major: bin, chunk, same order)              destination within
//        its bin
// and delta[chunk][bin] maps the run of (chunk, bin) from one order to the
// other.  An iteration is
//   k_pb_scatter  streams ent and writes contrib[source] to the edge's
//                 bin-major slot; the chunk's contributions and its row of
//                 deltas sit in LDS (64 KB), so the only global traffic is
//                 the streamed entries and the writes, which go to
//                 contiguous runs (one per bin);
//   k_pb_gather   streams a bin's (value, dloc) pairs and adds value * 2^56,
//                 rounded, into 8192 u64 sums in LDS, then writes rank and
//                 contrib (bins cut into several tiles add their LDS sums
//                 into a global array first, k_pb_finish converts it).
// Integer addition is associative, so every run gives the same sums; each
// sum is within E * 2^-57 of the exact one (far below FP32's rounding of the
// rank), so the L1 distance to the FP64 oracle stays at FP32 level.
constexpr uint32_t kPbShift = 13;        // destinations per bin: 2^13 (the LDS sums)
constexpr uint32_t kPbSize = 1u << kPbShift;
constexpr uint32_t kPbChunkShift = 13;   // sources per chunk: 2^13 (their contributions fill 32 KB of LDS)
constexpr uint32_t kPbChunk = 1u << kPbChunkShift;
constexpr uint64_t kPbTile = 1ull << 15;     // scatter: entries per workgroup (a slice of one
                                             // chunk's bins; fewer pages written per workgroup)
constexpr uint64_t kPbBinTile = 1ull << 19;  // gather: entries per workgroup (2^17 / 2^18 / 2^19 / 2^20:
                                             // 4.55 / 4.58 / 4.40 / 4.41 ms per iteration at s25/ef44;
                                             // larger tiles leave fewer bins cut in several tiles)
constexpr uint32_t kPbMaxBins = 4096;     // per-tile LDS tables of 12 B per bin stay within 48 KB
constexpr float kFixScale = 72057594037927936.0f;  // 2^56
constexpr double kFixInv = 1.0 / 72057594037927936.0;

__device__ __forceinline__ unsigned long long to_fixed(float x) { return __float2ull_rn(x * kFixScale); }

__global__ void k_pb_bounds(const uint64_t* __restrict__ off, uint32_t V, uint32_t nc, uint64_t* __restrict__ out) {
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c <= nc; c += gridDim.x * blockDim.x) {
    const uint64_t v = (uint64_t)c << kPbChunkShift;
    out[c] = off[v < V ? v : V];
  }
}

// Per tile of chunk c: how many of its edges land in each bin.
__global__ __launch_bounds__(kBlock) void k_pb_count(const uint32_t* __restrict__ dst,
                                                     const uint32_t* __restrict__ t_chunk,
                                                     const uint64_t* __restrict__ t_lo,
                                                     const uint64_t* __restrict__ t_hi, uint32_t nb,
                                                     uint32_t* __restrict__ cnt) {
  extern __shared__ uint32_t lc[];
  const uint32_t t = blockIdx.x;
  for (uint32_t b = threadIdx.x; b < nb; b += kBlock) lc[b] = 0;
  __syncthreads();
  for (uint64_t e = t_lo[t] + threadIdx.x; e < t_hi[t]; e += kBlock) atomicAdd(&lc[dst[e] >> kPbShift], 1u);
  __syncthreads();
  const uint64_t row = (uint64_t)t_chunk[t] * nb;
  for (uint32_t b = threadIdx.x; b < nb; b += kBlock)
    if (lc[b]) atomicAdd(&cnt[row + b], lc[b]);
}

// Exclusive prefix over n values read with `get`, written with `put`, plus
// the total; one workgroup, kBlock consecutive groups of values.
template <class Get, class Put>
__device__ void block_excl_scan(uint32_t n, uint64_t start, Get get, Put put, uint64_t* total) {
  __shared__ unsigned long long part[kBlock];
  const uint32_t per = (n + kBlock - 1) / kBlock;
  const uint32_t lo = threadIdx.x * per, hi = lo + per < n ? lo + per : n;
  unsigned long long sum = 0;
  for (uint32_t i = lo; i < hi; ++i) sum += get(i);
  part[threadIdx.x] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long run = start;
    for (int k = 0; k < kBlock; ++k) {
      const unsigned long long x = part[k];
      part[k] = run;
      run += x;
    }
    if (total) *total = run;
  }
  __syncthreads();
  unsigned long long run = part[threadIdx.x];
  for (uint32_t i = lo; i < hi; ++i) {
    const unsigned long long x = get(i);
    put(i, run);
    run += x;
  }
}

// Chunk-major start of run (c, b): the chunk's first edge + its runs of
// smaller bins.
__global__ __launch_bounds__(kBlock) void k_pb_rows(const uint32_t* __restrict__ cnt,
                                                    const uint64_t* __restrict__ bounds, uint32_t nb,
                                                    uint64_t* __restrict__ in_base) {
  const uint64_t row = (uint64_t)blockIdx.x * nb;
  block_excl_scan(
      nb, bounds[blockIdx.x], [&](uint32_t b) { return (unsigned long long)cnt[row + b]; },
      [&](uint32_t b, unsigned long long v) { in_base[row + b] = v; }, nullptr);
}

// Within bin b: the runs of smaller chunks, and the bin's size.
__global__ __launch_bounds__(kBlock) void k_pb_cols(const uint32_t* __restrict__ cnt, uint32_t nc, uint32_t nb,
                                                    uint64_t* __restrict__ col_pre, uint64_t* __restrict__ bin_size) {
  const uint32_t b = blockIdx.x;
  block_excl_scan(
      nc, 0, [&](uint32_t c) { return (unsigned long long)cnt[(uint64_t)c * nb + b]; },
      [&](uint32_t c, unsigned long long v) { col_pre[(uint64_t)c * nb + b] = v; }, bin_size + b);
}

__global__ __launch_bounds__(kBlock) void k_pb_bins(const uint64_t* __restrict__ bin_size, uint32_t nb,
                                                    uint64_t* __restrict__ bin_off) {
  block_excl_scan(
      nb, 0, [&](uint32_t b) { return (unsigned long long)bin_size[b]; },
      [&](uint32_t b, unsigned long long v) { bin_off[b] = v; }, bin_off + nb);
}

__global__ void k_pb_delta(const uint64_t* __restrict__ in_base, const uint64_t* __restrict__ col_pre,
                           const uint64_t* __restrict__ bin_off, uint32_t nb, uint64_t n,
                           int64_t* __restrict__ delta) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    delta[i] = (int64_t)(bin_off[i % nb] + col_pre[i]) - (int64_t)in_base[i];
}

// Places the edges of one tile: a range of each of its runs is reserved with
// one atomic per (tile, bin), slots inside it come from LDS counters.  The
// order inside a run may differ between builds; both arrays use the same
// slot, and the fixed-point sums do not depend on it.
__global__ __launch_bounds__(kBlock) void k_pb_fill(const uint32_t* __restrict__ src,
                                                    const uint32_t* __restrict__ dst,
                                                    const uint32_t* __restrict__ t_chunk,
                                                    const uint64_t* __restrict__ t_lo,
                                                    const uint64_t* __restrict__ t_hi, uint32_t nb,
                                                    unsigned long long* __restrict__ cursor,
                                                    const int64_t* __restrict__ delta,
                                                    uint32_t* __restrict__ ent, uint16_t* __restrict__ dloc) {
  extern __shared__ unsigned long long lbase[];  // nb bases, then nb u32 counters
  uint32_t* lc = reinterpret_cast<uint32_t*>(lbase + nb);
  const uint32_t t = blockIdx.x;
  const uint64_t lo = t_lo[t], hi = t_hi[t];
  const uint64_t row = (uint64_t)t_chunk[t] * nb;
  for (uint32_t b = threadIdx.x; b < nb; b += kBlock) lc[b] = 0;
  __syncthreads();
  for (uint64_t e = lo + threadIdx.x; e < hi; e += kBlock) atomicAdd(&lc[dst[e] >> kPbShift], 1u);
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < nb; b += kBlock) {
    if (lc[b]) lbase[b] = atomicAdd(&cursor[row + b], (unsigned long long)lc[b]);
    lc[b] = 0;
  }
  __syncthreads();
  for (uint64_t e = lo + threadIdx.x; e < hi; e += kBlock) {
    const uint32_t d = dst[e], b = d >> kPbShift;
    const uint64_t i = lbase[b] + atomicAdd(&lc[b], 1u);
    ent[i] = (src[e] & (kPbChunk - 1)) | (b << kPbChunkShift);
    dloc[(uint64_t)((int64_t)i + delta[row + b])] = (uint16_t)(d & (kPbSize - 1));
  }
}

// D: the LDS type of the deltas (int32_t when every index fits 31 bits, so
// three workgroups fit a CU).  Loads in flight per thread: 32 entries in the
// scatter, 32 pairs in the gather (measured per iteration, alternating
// processes: 8 / 4 -> 16 / 8 about 10 % faster, 16 / 16 another 6 %, 32 / 32
// another 1-4 %; the gather's 64 KB of LDS already limits it to two
// workgroups per CU, so its 114 registers cost no occupancy).
constexpr int kPbScatterUnroll = 32;
constexpr int kPbGatherPairs = 32;

template <typename D>
__global__ __launch_bounds__(kBlock) void k_pb_scatter(const uint32_t* __restrict__ ent,
                                                       const uint32_t* __restrict__ t_chunk,
                                                       const uint64_t* __restrict__ t_lo,
                                                       const uint64_t* __restrict__ t_hi, uint32_t nb, uint32_t V,
                                                       const int64_t* __restrict__ delta,
                                                       const float* __restrict__ contrib, float* __restrict__ val) {
  extern __shared__ float cc[];  // kPbChunk contributions, then the deltas of the tile's bins
  D* dr = reinterpret_cast<D*>(cc + kPbChunk);
  const uint32_t t = blockIdx.x;
  const uint32_t c = t_chunk[t];
  const uint64_t lo = t_lo[t], hi = t_hi[t];
  const uint64_t v0 = (uint64_t)c << kPbChunkShift;
  // the tile is a slice of the chunk's bin-sorted entries: its bins are
  // [first, last]
  const uint32_t b0 = ent[lo] >> kPbChunkShift, b1 = ent[hi - 1] >> kPbChunkShift;
  for (uint32_t k = threadIdx.x; k <= b1 - b0; k += kBlock) dr[k] = (D)delta[(uint64_t)c * nb + b0 + k];
  for (uint32_t k = threadIdx.x; k < kPbChunk; k += kBlock) cc[k] = v0 + k < V ? contrib[v0 + k] : 0.0f;
  __syncthreads();
  uint64_t i = lo + threadIdx.x;
  // ordinary (cached) loads and stores: a run's first and last lines are
  // shared with the neighbouring runs, written by other workgroups, and the
  // caches merge those partial lines (non-temporal accesses measured 10-20 %
  // slower)
  for (; i + (kPbScatterUnroll - 1) * kBlock < hi; i += kPbScatterUnroll * kBlock) {
    uint32_t x[kPbScatterUnroll];
#pragma unroll
    for (int j = 0; j < kPbScatterUnroll; ++j) x[j] = ent[i + j * kBlock];
#pragma unroll
    for (int j = 0; j < kPbScatterUnroll; ++j) {
      const int64_t slot = (int64_t)(i + j * kBlock) + (int64_t)dr[(x[j] >> kPbChunkShift) - b0];
      val[(uint64_t)slot] = cc[x[j] & (kPbChunk - 1)];
    }
  }
  for (; i < hi; i += kBlock) {
    const uint32_t x = ent[i];
    val[(uint64_t)((int64_t)i + (int64_t)dr[(x >> kPbChunkShift) - b0])] = cc[x & (kPbChunk - 1)];
  }
}

__global__ __launch_bounds__(kBlock) void k_pb_gather(const float* __restrict__ val, const uint16_t* __restrict__ dloc,
                                                      const uint32_t* __restrict__ t_bin,
                                                      const uint64_t* __restrict__ t_lo,
                                                      const uint64_t* __restrict__ t_hi,
                                                      const uint32_t* __restrict__ t_multi, uint32_t V,
                                                      const uint32_t* __restrict__ deg, float base, float damping,
                                                      unsigned long long* __restrict__ gsum,
                                                      float* __restrict__ rank, float* __restrict__ contrib) {
  __shared__ unsigned long long acc[kPbSize];
  const uint32_t t = blockIdx.x;
  for (uint32_t k = threadIdx.x; k < kPbSize; k += kBlock) acc[k] = 0;
  __syncthreads();
  const uint64_t lo = t_lo[t], hi = t_hi[t];
  // pairs of entries from an even index: one 8-byte value load and one
  // 4-byte destination load per pair; entries outside [lo, hi) are skipped
  typedef float f32x2 __attribute__((ext_vector_type(2)));
  const uint64_t p0 = lo >> 1, p1 = (hi + 1) >> 1;
  const f32x2* val2 = reinterpret_cast<const f32x2*>(val);
  const uint32_t* dloc2 = reinterpret_cast<const uint32_t*>(dloc);
  uint64_t pj = p0 + threadIdx.x;
  for (; pj + (kPbGatherPairs - 1) * kBlock < p1; pj += (kPbGatherPairs) * kBlock) {
    f32x2 x[kPbGatherPairs];
    uint32_t d[kPbGatherPairs];
#pragma unroll
    for (int u = 0; u < kPbGatherPairs; ++u) {
      x[u] = val2[pj + u * kBlock];  // ordinary loads: measured as fast or faster than non-temporal
      d[u] = dloc2[pj + u * kBlock];
    }
#pragma unroll
    for (int u = 0; u < kPbGatherPairs; ++u) {
      const uint64_t e = 2 * (pj + u * kBlock);
      if (e >= lo) atomicAdd(&acc[d[u] & 0xFFFFu], to_fixed(x[u].x));
      if (e + 1 < hi) atomicAdd(&acc[d[u] >> 16], to_fixed(x[u].y));
    }
  }
  for (; pj < p1; pj += kBlock) {
    const uint64_t e = 2 * pj;
    if (e >= lo) atomicAdd(&acc[dloc[e]], to_fixed(val[e]));
    if (e + 1 < hi) atomicAdd(&acc[dloc[e + 1]], to_fixed(val[e + 1]));
  }
  __syncthreads();
  const uint64_t v0 = (uint64_t)t_bin[t] << kPbShift;
  if (t_multi[t]) {
    for (uint32_t k = threadIdx.x; k < kPbSize; k += kBlock)
      if (acc[k]) atomicAdd(&gsum[v0 + k], acc[k]);
  } else {
    for (uint32_t k = threadIdx.x; k < kPbSize && v0 + k < V; k += kBlock)
      finish((uint32_t)(v0 + k), (float)((double)acc[k] * kFixInv), base, damping, deg, rank, contrib);
  }
}

// Bins cut into several tiles: their summed fixed-point values become ranks,
// and the sums are cleared for the next iteration.
__global__ __launch_bounds__(kBlock) void k_pb_finish(const uint32_t* __restrict__ bins, uint32_t V,
                                                      const uint32_t* __restrict__ deg, float base, float damping,
                                                      unsigned long long* __restrict__ gsum,
                                                      float* __restrict__ rank, float* __restrict__ contrib) {
  const uint64_t v = ((uint64_t)bins[blockIdx.x / (kPbSize / kBlock)] << kPbShift) +
                     (blockIdx.x % (kPbSize / kBlock)) * kBlock + threadIdx.x;
  if (v >= V) return;
  const unsigned long long x = gsum[v];
  gsum[v] = 0;
  finish((uint32_t)v, (float)((double)x * kFixInv), base, damping, deg, rank, contrib);
}

int build_blocked(const DevGraph& g, uint64_t E, PrPlan* p, hipStream_t s) {
  const uint32_t V = g.V;
  const uint32_t nb = (uint32_t)(((uint64_t)V + kPbSize - 1) >> kPbShift);
  const uint32_t nc = (uint32_t)(((uint64_t)V + kPbChunk - 1) >> kPbChunkShift);
  if (nb > kPbMaxBins) return set_error(CG_ERR_INVALID, "blocked PageRank holds at most 2^25 vertices");
  p->layout = kLayoutBlocked;
  p->pb_nb = nb;
  p->pb_narrow = E < (1ull << 31);
  // entries 4 B read + contributions 4 B written, contributions 4 B +
  // destinations 2 B read; contributions, degrees, ranks per vertex
  p->stream_bytes = 14.0 * (double)E + 16.0 * (double)V;
  CG_TRY(build_degrees(g, p, s));
  // chunk bounds -> tiles of at most kPbTile edges inside one chunk
  DevBuf<uint64_t> bounds;
  CG_TRY(bounds.ensure((uint64_t)nc + 1));
  hipLaunchKernelGGL(k_pb_bounds, dim3(grid_of((uint64_t)nc + 1)), dim3(kBlock), 0, s, g.off, V, nc, bounds.p);
  std::vector<uint64_t> hb(nc + 1);
  CG_HIP(hipMemcpyAsync(hb.data(), bounds.p, hb.size() * 8, hipMemcpyDeviceToHost, s));
  CG_HIP(hipStreamSynchronize(s));
  std::vector<uint32_t> t1c;
  std::vector<uint64_t> t1l, t1h;
  for (uint32_t c = 0; c < nc; ++c)
    for (uint64_t lo = hb[c]; lo < hb[c + 1]; lo += kPbTile) {
      t1c.push_back(c);
      t1l.push_back(lo);
      t1h.push_back(std::min(hb[c + 1], lo + kPbTile));
    }
  p->pb_t1 = (uint32_t)t1c.size();
  CG_TRY(p->pb_t1_chunk.ensure(t1c.size()));
  CG_TRY(p->pb_t1_lo.ensure(t1c.size()));
  CG_TRY(p->pb_t1_hi.ensure(t1c.size()));
  if (!t1c.empty()) {
    CG_HIP(hipMemcpyAsync(p->pb_t1_chunk.p, t1c.data(), t1c.size() * 4, hipMemcpyHostToDevice, s));
    CG_HIP(hipMemcpyAsync(p->pb_t1_lo.p, t1l.data(), t1l.size() * 8, hipMemcpyHostToDevice, s));
    CG_HIP(hipMemcpyAsync(p->pb_t1_hi.p, t1h.data(), t1h.size() * 8, hipMemcpyHostToDevice, s));
  }
  // run sizes, then the two prefix orders and their difference
  const uint64_t cells = (uint64_t)nc * nb;
  DevBuf<uint32_t> cnt;
  DevBuf<uint64_t> in_base, col_pre, bin_size, bin_off;
  CG_TRY(cnt.ensure(cells));
  CG_TRY(in_base.ensure(cells));
  CG_TRY(col_pre.ensure(cells));
  CG_TRY(bin_size.ensure(nb));
  CG_TRY(bin_off.ensure((uint64_t)nb + 1));
  CG_TRY(p->pb_delta.ensure(cells));
  CG_HIP(hipMemsetAsync(cnt.p, 0, cells * 4, s));
  if (p->pb_t1)
    hipLaunchKernelGGL(k_pb_count, dim3(p->pb_t1), dim3(kBlock), nb * 4, s, g.dst, p->pb_t1_chunk.p, p->pb_t1_lo.p,
                       p->pb_t1_hi.p, nb, cnt.p);
  hipLaunchKernelGGL(k_pb_rows, dim3(nc), dim3(kBlock), 0, s, cnt.p, bounds.p, nb, in_base.p);
  hipLaunchKernelGGL(k_pb_cols, dim3(nb), dim3(kBlock), 0, s, cnt.p, nc, nb, col_pre.p, bin_size.p);
  hipLaunchKernelGGL(k_pb_bins, dim3(1), dim3(kBlock), 0, s, bin_size.p, nb, bin_off.p);
  hipLaunchKernelGGL(k_pb_delta, dim3(grid_of(cells)), dim3(kBlock), 0, s, in_base.p, col_pre.p, bin_off.p, nb,
                     cells, p->pb_delta.p);
  CG_HIP(hipGetLastError());
  cnt.reset();
  col_pre.reset();
  // place every edge in both orders
  DevBuf<uint32_t> src;
  CG_TRY(edge_sources(g.off, V, E, src, s));
  CG_TRY(p->pb_ent.ensure(E));
  // + 2: the gather reads entries in aligned pairs, one past an odd end
  CG_TRY(p->pb_dloc.ensure(E + 2));
  CG_TRY(p->pb_val.ensure(E + 2));
  if (p->pb_t1)
    hipLaunchKernelGGL(k_pb_fill, dim3(p->pb_t1), dim3(kBlock), nb * 12, s, src.p, g.dst, p->pb_t1_chunk.p,
                       p->pb_t1_lo.p, p->pb_t1_hi.p, nb, reinterpret_cast<unsigned long long*>(in_base.p),
                       p->pb_delta.p, p->pb_ent.p, p->pb_dloc.p);
  CG_HIP(hipGetLastError());
  // bins -> tiles of at most kPbBinTile entries; every bin has one (its ranks
  // are written by it even when it has no in-edges)
  std::vector<uint64_t> ho(nb + 1);
  CG_HIP(hipMemcpyAsync(ho.data(), bin_off.p, ho.size() * 8, hipMemcpyDeviceToHost, s));
  CG_HIP(hipStreamSynchronize(s));
  std::vector<uint32_t> t2b, t2m, multi;
  std::vector<uint64_t> t2l, t2h;
  for (uint32_t b = 0; b < nb; ++b) {
    const uint64_t n = ho[b + 1] - ho[b];
    const uint32_t k = n > kPbBinTile ? (uint32_t)((n + kPbBinTile - 1) / kPbBinTile) : 1u;
    if (k > 1) multi.push_back(b);
    for (uint32_t i = 0; i < k; ++i) {
      t2b.push_back(b);
      t2l.push_back(ho[b] + (uint64_t)i * kPbBinTile);
      t2h.push_back(std::min(ho[b + 1], ho[b] + (uint64_t)(i + 1) * kPbBinTile));
      t2m.push_back(k > 1 ? 1u : 0u);
    }
  }
  p->pb_t2 = (uint32_t)t2b.size();
  p->pb_multi = (uint32_t)multi.size();
  CG_TRY(p->pb_t2_bin.ensure(t2b.size()));
  CG_TRY(p->pb_t2_lo.ensure(t2b.size()));
  CG_TRY(p->pb_t2_hi.ensure(t2b.size()));
  CG_TRY(p->pb_t2_multi.ensure(t2b.size()));
  CG_TRY(p->pb_multi_bins.ensure(multi.size()));
  CG_HIP(hipMemcpyAsync(p->pb_t2_bin.p, t2b.data(), t2b.size() * 4, hipMemcpyHostToDevice, s));
  CG_HIP(hipMemcpyAsync(p->pb_t2_lo.p, t2l.data(), t2l.size() * 8, hipMemcpyHostToDevice, s));
  CG_HIP(hipMemcpyAsync(p->pb_t2_hi.p, t2h.data(), t2h.size() * 8, hipMemcpyHostToDevice, s));
  CG_HIP(hipMemcpyAsync(p->pb_t2_multi.p, t2m.data(), t2m.size() * 4, hipMemcpyHostToDevice, s));
  if (!multi.empty())
    CG_HIP(hipMemcpyAsync(p->pb_multi_bins.p, multi.data(), multi.size() * 4, hipMemcpyHostToDevice, s));
  CG_TRY(p->pb_sum.ensure((uint64_t)nb << kPbShift));
  CG_HIP(hipMemsetAsync(p->pb_sum.p, 0, ((uint64_t)nb << kPbShift) * 8, s));
  // the host vectors must outlive the copies
  CG_HIP(hipStreamSynchronize(s));
  return CG_OK;
}

// ---- propagation blocking over (source, bin) pairs (layout 5) ----
//
// An edge's contribution is its source's, so the edges of one source into one
// bin carry the same value.  Writing and reading it once per (source, bin)
// pair instead of once per edge leaves a pass of 2 bytes per edge (the
// destination) plus 12 bytes per pair.  On a graph renumbered by descending
// out-degree the hot destinations share few bins, so the pairs are few:
// R-MAT edge factor 44 has 0.08 / 0.13 / 0.18 pairs per edge at scale
// 18 / 20 / 22 (8192-vertex bins), against 1 per edge before.
//   scatter  k_pb_scatter over the pairs' chunk-major entries (source within
//            its chunk | bin << 13): one store per pair into its bin-major slot;
//   gather   k_pbp_gather: a wavefront takes 256 consecutive edges of a bin
//            (four per lane, one 8-byte load); the pair of each edge is the
//            group's first pair (pb_gbase) plus the pair starts (0x8000
//            flags) before it, counted from one ballot per edge slot; its value is
//            one load from the pair's slot (neighbouring edges share pairs,
//            so these loads hit the same lines), added in 64-bit fixed point
//            into the bin's LDS sums.
// Pairs inside one (chunk, bin) run are ordered by source in both orders
// (a radix sort on (chunk, bin, source)), so one delta per run maps a pair's
// chunk-major index to its bin-major one, as for the edges of layout 4.
constexpr int kPpEdges = 4;                       // edges per lane in a group (one 8-byte load)
constexpr int kPpGroupShift = 8;                  // edges per group: 64 lanes x kPpEdges = 2^8
constexpr uint64_t kPpGroup = 1ull << kPpGroupShift;
// Gather shape, measured at R-MAT s25 / ef44 (ms per iteration): workgroups
// of 256 / 512 / 1024 threads with 16 / 8 / 2 groups in flight per wavefront
// 4.85-4.90 / 3.60-3.63 / 2.89-2.91; at 1024 threads 1 / 3 / 4 groups
// 2.92-2.95 / 3.77-3.85 / 3.81-3.84 (three groups and more need more than the
// 64 registers that eight wavefronts per SIMD leave); tiles of 2^18 / 2^19 /
// 2^20 edges 3.05 / 2.89-2.91 / 2.79-2.91.  The LDS sums (64 KB) allow two
// workgroups per CU, so the workgroup size sets the wavefronts per SIMD.
constexpr int kPpGroups = 2;                      // groups in flight per wavefront
constexpr int kPpBlock = 1024;                    // gather workgroup: 16 wavefronts share the LDS sums
constexpr uint64_t kPpBinTile = 1ull << 20;       // gather: edges per workgroup
constexpr uint64_t kPpPad = 2 * kPpGroup;         // zero entries past the last edge

// set bits of m below this lane (v_mbcnt)
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__global__ void k_pp_flags(const uint64_t* __restrict__ off, const uint32_t* __restrict__ dst,
                           const uint32_t* __restrict__ src, uint64_t E, uint32_t* __restrict__ flags) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < E; e += (uint64_t)gridDim.x * blockDim.x)
    flags[e] = (e == off[src[e]] || (dst[e] >> kPbShift) != (dst[e - 1] >> kPbShift)) ? 1u : 0u;
}

// incl[e]: pair starts in [0, e], so edge e belongs to pair incl[e] - 1
// (pairs numbered in source-major order: source, then bin)
__global__ void k_pp_pairs(const uint32_t* __restrict__ flags, const uint32_t* __restrict__ incl,
                           const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst, uint64_t E,
                           uint32_t nb, uint64_t* __restrict__ pfirst, uint64_t* __restrict__ key,
                           uint32_t* __restrict__ id, uint32_t* __restrict__ cnt) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < E; e += (uint64_t)gridDim.x * blockDim.x) {
    if (!flags[e]) continue;
    const uint32_t j = incl[e] - 1, u = src[e], b = dst[e] >> kPbShift, c = u >> kPbChunkShift;
    pfirst[j] = e;
    key[j] = ((uint64_t)c << 25) | ((uint64_t)b << kPbChunkShift) | (u & (kPbChunk - 1));
    id[j] = j;
    atomicAdd(&cnt[(uint64_t)c * nb + b], 1u);
  }
}

// Chunk-major start of chunk c: the pair of its first edge (pairs of a
// source are contiguous, sources ascending), or the pair count.
__global__ void k_pp_bounds(const uint64_t* __restrict__ off, const uint32_t* __restrict__ incl, uint32_t V,
                            uint64_t E, uint64_t P, uint32_t nc, uint64_t* __restrict__ out) {
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c <= nc; c += gridDim.x * blockDim.x) {
    const uint64_t v = (uint64_t)c << kPbChunkShift;
    const uint64_t e = off[v < V ? v : V];
    out[c] = e < E ? (uint64_t)incl[e] - 1 : P;
  }
}

__global__ void k_pp_place(const uint64_t* __restrict__ key, const uint32_t* __restrict__ id, uint64_t P,
                           uint32_t nb, const int64_t* __restrict__ delta, const uint64_t* __restrict__ pfirst,
                           uint32_t* __restrict__ ent, uint32_t* __restrict__ bm_of, uint32_t* __restrict__ count) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < P; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t k = key[i];
    const uint32_t c = (uint32_t)(k >> 25), b = (uint32_t)(k >> kPbChunkShift) & (kPbMaxBins - 1);
    ent[i] = (uint32_t)(k & (kPbChunk - 1)) | (b << kPbChunkShift);
    const uint64_t j = (uint64_t)((int64_t)i + delta[(uint64_t)c * nb + b]);
    const uint32_t sm = id[i];
    bm_of[sm] = (uint32_t)j;
    count[j] = (uint32_t)(pfirst[sm + 1] - pfirst[sm]);
  }
}

__global__ void k_pp_edges(const uint32_t* __restrict__ dst, const uint32_t* __restrict__ incl, uint64_t E,
                           const uint64_t* __restrict__ pfirst, const uint32_t* __restrict__ bm_of,
                           const uint64_t* __restrict__ eoff, uint16_t* __restrict__ dloc,
                           uint32_t* __restrict__ gbase) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < E; e += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t sm = incl[e] - 1, j = bm_of[sm];
    const uint64_t k = e - pfirst[sm], pos = eoff[j] + k;
    dloc[pos] = (uint16_t)((dst[e] & (kPbSize - 1)) | (k == 0 ? 0x8000u : 0u));
    if ((pos & (kPpGroup - 1)) == 0) gbase[pos >> kPpGroupShift] = j;
  }
}

// first edge (bin-major) of every bin: the first edge of its first pair
__global__ void k_pp_bin_edges(const uint64_t* __restrict__ bin_off, const uint64_t* __restrict__ eoff,
                               uint32_t nb, uint64_t* __restrict__ out) {
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b <= nb; b += gridDim.x * blockDim.x)
    out[b] = eoff[bin_off[b]];
}

__global__ __launch_bounds__(kPpBlock) void k_pbp_gather(const float* __restrict__ val,
                                                         const uint16_t* __restrict__ dloc,
                                                         const uint32_t* __restrict__ gbase,
                                                         const uint32_t* __restrict__ t_bin,
                                                         const uint64_t* __restrict__ t_lo,
                                                         const uint64_t* __restrict__ t_hi,
                                                         const uint32_t* __restrict__ t_multi, uint32_t V,
                                                         const uint32_t* __restrict__ deg, float base,
                                                         float damping, unsigned long long* __restrict__ rows,
                                                         float* __restrict__ rank, float* __restrict__ contrib) {
  __shared__ unsigned long long acc[kPbSize];
  const uint32_t t = blockIdx.x;
  for (uint32_t k = threadIdx.x; k < kPbSize; k += kPpBlock) acc[k] = 0;
  __syncthreads();
  const uint64_t lo = t_lo[t], hi = t_hi[t];
  if (hi > lo) {
    // a group is 2^8 consecutive edges, four per lane (lane l: 4l .. 4l + 3);
    // an edge's pair is the group's first pair plus the pair starts (0x8000
    // flags) after the group's first edge up to it: per edge slot k, a
    // ballot of the flags; the lanes below count with v_mbcnt
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    constexpr uint32_t kWaves = kPpBlock / 64;
    const uint64_t* dl4 = reinterpret_cast<const uint64_t*>(dloc);
    const uint64_t g0 = lo >> kPpGroupShift, g1 = (hi - 1) >> kPpGroupShift;  // inclusive
    for (uint64_t g = g0 + wave; g <= g1; g += (uint64_t)kPpGroups * kWaves) {
      uint64_t dd[kPpGroups];
      uint32_t gb[kPpGroups];
#pragma unroll
      for (int u = 0; u < kPpGroups; ++u) {
        const uint64_t gg = g + (uint64_t)u * kWaves;
        const bool ok = gg <= g1;
        dd[u] = ok ? dl4[gg * 64 + lane] : 0ull;
        gb[u] = ok ? gbase[gg] : 0u;
      }
      float x[kPpGroups * kPpEdges];
#pragma unroll
      for (int u = 0; u < kPpGroups; ++u) {
        uint32_t pi = gb[u];
        uint32_t f[kPpEdges];
#pragma unroll
        for (int k = 0; k < kPpEdges; ++k) {
          f[k] = (uint32_t)(dd[u] >> (16 * k + 15)) & 1u;
          uint64_t b = __ballot(f[k] != 0);
          if (k == 0) b &= ~1ull;  // the group's first edge belongs to pair gb
          pi += lanes_below(b);
        }
        if (lane == 0) f[0] = 0;
        const uint64_t gg = g + (uint64_t)u * kWaves, e0 = (gg << kPpGroupShift) + 4 * lane;
        const bool whole = gg <= g1 && (gg << kPpGroupShift) >= lo && ((gg + 1) << kPpGroupShift) <= hi;  // uniform
#pragma unroll
        for (int k = 0; k < kPpEdges; ++k) {
          pi += f[k];
          const bool in = whole || (gg <= g1 && e0 + k >= lo && e0 + k < hi);
          x[u * kPpEdges + k] = in ? val[pi] : 0.0f;
        }
      }
#pragma unroll
      for (int u = 0; u < kPpGroups; ++u) {
        const uint64_t gg = g + (uint64_t)u * kWaves, e0 = (gg << kPpGroupShift) + 4 * lane;
        const bool whole = gg <= g1 && (gg << kPpGroupShift) >= lo && ((gg + 1) << kPpGroupShift) <= hi;
#pragma unroll
        for (int k = 0; k < kPpEdges; ++k) {
          const bool in = whole || (gg <= g1 && e0 + k >= lo && e0 + k < hi);
          if (in) atomicAdd(&acc[(uint32_t)(dd[u] >> (16 * k)) & (kPbSize - 1)], to_fixed(x[u * kPpEdges + k]));
        }
      }
    }
  }
  __syncthreads();
  const uint64_t v0 = (uint64_t)t_bin[t] << kPbShift;
  if (t_multi[t]) {
    unsigned long long* row = rows + (uint64_t)(t_multi[t] - 1) * kPbSize;
    for (uint32_t k = threadIdx.x; k < kPbSize; k += kPpBlock) row[k] = acc[k];
  } else {
    for (uint32_t k = threadIdx.x; k < kPbSize && v0 + k < V; k += kPpBlock)
      finish((uint32_t)(v0 + k), (float)((double)acc[k] * kFixInv), base, damping, deg, rank, contrib);
  }
}

// Bins cut into several gather tiles: the tiles' rows added up per vertex
// (in row order), then the ranks.
__global__ __launch_bounds__(kBlock) void k_pbp_finish(const uint32_t* __restrict__ bins,
                                                       const uint32_t* __restrict__ mrow, uint32_t V,
                                                       const uint32_t* __restrict__ deg, float base, float damping,
                                                       const unsigned long long* __restrict__ part,
                                                       float* __restrict__ rank, float* __restrict__ contrib) {
  const uint32_t i = blockIdx.x / (kPbSize / kBlock);
  const uint32_t k = (blockIdx.x % (kPbSize / kBlock)) * kBlock + threadIdx.x;
  const uint64_t v = ((uint64_t)bins[i] << kPbShift) + k;
  if (v >= V) return;
  unsigned long long x = 0;
  for (uint32_t r = mrow[i]; r < mrow[i + 1]; ++r) x += part[(uint64_t)r * kPbSize + k];
  finish((uint32_t)v, (float)((double)x * kFixInv), base, damping, deg, rank, contrib);
}

// Tiles of the two passes: chunk c's pairs [hb[c], hb[c + 1]) in pieces of
// kPbTile, bin b's edges [he[b], he[b + 1]) in pieces of bin_tile (every
// bin gets a tile, which writes its ranks even when it has no in-edges).
int pb_tiles(PrPlan* p, const std::vector<uint64_t>& hb, const std::vector<uint64_t>& he, uint64_t bin_tile,
             hipStream_t s) {
  const uint32_t nc = (uint32_t)hb.size() - 1, nb = (uint32_t)he.size() - 1;
  std::vector<uint32_t> t1c, t2b, t2m, multi, mrow;
  std::vector<uint64_t> t1l, t1h, t2l, t2h;
  uint32_t rows = 0;
  for (uint32_t c = 0; c < nc; ++c)
    for (uint64_t lo = hb[c]; lo < hb[c + 1]; lo += kPbTile) {
      t1c.push_back(c);
      t1l.push_back(lo);
      t1h.push_back(std::min(hb[c + 1], lo + kPbTile));
    }
  for (uint32_t b = 0; b < nb; ++b) {
    const uint64_t n = he[b + 1] - he[b];
    const uint32_t k = n > bin_tile ? (uint32_t)((n + bin_tile - 1) / bin_tile) : 1u;
    if (k > 1) {
      multi.push_back(b);
      mrow.push_back(rows);
    }
    for (uint32_t i = 0; i < k; ++i) {
      t2b.push_back(b);
      t2l.push_back(he[b] + (uint64_t)i * bin_tile);
      t2h.push_back(std::min(he[b + 1], he[b] + (uint64_t)(i + 1) * bin_tile));
      t2m.push_back(k > 1 ? ++rows : 0u);
    }
  }
  mrow.push_back(rows);
  p->pb_t1 = (uint32_t)t1c.size();
  p->pb_t2 = (uint32_t)t2b.size();
  p->pb_multi = (uint32_t)multi.size();
  CG_TRY(p->pb_t1_chunk.ensure(t1c.size() + 1));
  CG_TRY(p->pb_t1_lo.ensure(t1c.size() + 1));
  CG_TRY(p->pb_t1_hi.ensure(t1c.size() + 1));
  CG_TRY(p->pb_t2_bin.ensure(t2b.size()));
  CG_TRY(p->pb_t2_lo.ensure(t2b.size()));
  CG_TRY(p->pb_t2_hi.ensure(t2b.size()));
  CG_TRY(p->pb_t2_multi.ensure(t2b.size()));
  CG_TRY(p->pb_multi_bins.ensure(multi.size() + 1));
  if (!t1c.empty()) {
    CG_HIP(hipMemcpyAsync(p->pb_t1_chunk.p, t1c.data(), t1c.size() * 4, hipMemcpyHostToDevice, s));
    CG_HIP(hipMemcpyAsync(p->pb_t1_lo.p, t1l.data(), t1l.size() * 8, hipMemcpyHostToDevice, s));
    CG_HIP(hipMemcpyAsync(p->pb_t1_hi.p, t1h.data(), t1h.size() * 8, hipMemcpyHostToDevice, s));
  }
  CG_HIP(hipMemcpyAsync(p->pb_t2_bin.p, t2b.data(), t2b.size() * 4, hipMemcpyHostToDevice, s));
  CG_HIP(hipMemcpyAsync(p->pb_t2_lo.p, t2l.data(), t2l.size() * 8, hipMemcpyHostToDevice, s));
  CG_HIP(hipMemcpyAsync(p->pb_t2_hi.p, t2h.data(), t2h.size() * 8, hipMemcpyHostToDevice, s));
  CG_HIP(hipMemcpyAsync(p->pb_t2_multi.p, t2m.data(), t2m.size() * 4, hipMemcpyHostToDevice, s));
  if (!multi.empty())
    CG_HIP(hipMemcpyAsync(p->pb_multi_bins.p, multi.data(), multi.size() * 4, hipMemcpyHostToDevice, s));
  CG_TRY(p->pb_mrow.ensure(mrow.size()));
  CG_HIP(hipMemcpyAsync(p->pb_mrow.p, mrow.data(), mrow.size() * 4, hipMemcpyHostToDevice, s));
  CG_TRY(p->pb_part.ensure(((uint64_t)rows + 1) << kPbShift));
  CG_HIP(hipStreamSynchronize(s));  // the host vectors must outlive the copies
  return CG_OK;
}

int build_pairs(cg_graph* graph, PrPlan* p, hipStream_t s) {
  const uint32_t V = graph->V;
  const uint64_t E = graph->E;
  const uint32_t nb = (uint32_t)(((uint64_t)V + kPbSize - 1) >> kPbShift);
  const uint32_t nc = (uint32_t)(((uint64_t)V + kPbChunk - 1) >> kPbChunkShift);
  if (nb > kPbMaxBins) return set_error(CG_ERR_INVALID, "blocked PageRank holds at most 2^25 vertices");
  if (E == 0 || E >= 0xFFFFFFFFull) return set_error(CG_ERR_INVALID, "pair blocking needs 0 < E < 2^32");
  p->layout = kLayoutPairs;
  p->pb_nb = nb;
  // the graph renumbered by descending out-degree, rows sorted by destination
  cg_graph rg;
  {
    DevBuf<uint32_t> new_of;
    CG_TRY(degree_order(graph, p->perm, new_of, nullptr));
    CG_TRY(relabel_graph(graph, new_of.p, &rg));
  }
  const DevGraph g{rg.d_off.p, rg.d_dst.p, nullptr, nullptr, nullptr, V};
  CG_TRY(build_degrees(g, p, s));
  DevBuf<uint32_t> src, flags, incl;
  CG_TRY(edge_sources(g.off, V, E, src, s));
  CG_TRY(flags.ensure(E));
  CG_TRY(incl.ensure(E));
  hipLaunchKernelGGL(k_pp_flags, dim3(grid_of(E)), dim3(kBlock), 0, s, g.off, g.dst, src.p, E, flags.p);
  PR_LAUNCH_OK();
  {
    size_t tb = 0;
    CG_HIP(rocprim::inclusive_scan(nullptr, tb, flags.p, incl.p, E, rocprim::plus<uint32_t>(), s));
    DevBuf<uint8_t> tmp;
    CG_TRY(tmp.ensure(tb));
    CG_HIP(rocprim::inclusive_scan(tmp.p, tb, flags.p, incl.p, E, rocprim::plus<uint32_t>(), s));
  }
  uint32_t P32 = 0;
  CG_HIP(hipMemcpyAsync(&P32, incl.p + E - 1, 4, hipMemcpyDeviceToHost, s));
  CG_HIP(hipStreamSynchronize(s));
  const uint64_t P = P32;
  p->pb_pairs = P;
  // pairs: entry 4 B read, contribution 4 B written and 4 B read; edges:
  // 2 B read (+ one 4-byte group pair per 256); per vertex as layout 4
  p->stream_bytes = 12.0 * (double)P + (2.0 + 4.0 / (double)kPpGroup) * (double)E + 16.0 * (double)V;
  p->pb_narrow = P < (1ull << 31);
  // the pairs in source-major order, and the pairs of every (chunk, bin) run
  const uint64_t cells = (uint64_t)nc * nb;
  DevBuf<uint64_t> pfirst, key, key2;
  DevBuf<uint32_t> id, id2, cnt;
  CG_TRY(pfirst.ensure(P + 1));
  CG_TRY(key.ensure(P));
  CG_TRY(key2.ensure(P));
  CG_TRY(id.ensure(P));
  CG_TRY(id2.ensure(P));
  CG_TRY(cnt.ensure(cells));
  CG_HIP(hipMemsetAsync(cnt.p, 0, cells * 4, s));
  CG_HIP(hipMemcpyAsync(pfirst.p + P, &E, 8, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_pp_pairs, dim3(grid_of(E)), dim3(kBlock), 0, s, flags.p, incl.p, src.p, g.dst, E, nb,
                     pfirst.p, key.p, id.p, cnt.p);
  PR_LAUNCH_OK();
  flags.reset();
  // chunk-major order: (chunk, bin, source)
  {
    rocprim::double_buffer<uint64_t> kb(key.p, key2.p);
    rocprim::double_buffer<uint32_t> vb(id.p, id2.p);
    size_t tb = 0;
    CG_HIP(rocprim::radix_sort_pairs(nullptr, tb, kb, vb, P, 0, 37, s));
    DevBuf<uint8_t> tmp;
    CG_TRY(tmp.ensure(tb));
    CG_HIP(rocprim::radix_sort_pairs(tmp.p, tb, kb, vb, P, 0, 37, s));
    if (kb.current() != key.p) std::swap(key, key2);
    if (vb.current() != id.p) std::swap(id, id2);
  }
  key2.reset();
  id2.reset();
  // run starts in both orders and their differences (the kernels of layout 4)
  DevBuf<uint64_t> bounds, in_base, col_pre, bin_size, bin_off;
  CG_TRY(bounds.ensure((uint64_t)nc + 1));
  CG_TRY(in_base.ensure(cells));
  CG_TRY(col_pre.ensure(cells));
  CG_TRY(bin_size.ensure(nb));
  CG_TRY(bin_off.ensure((uint64_t)nb + 1));
  CG_TRY(p->pb_delta.ensure(cells));
  hipLaunchKernelGGL(k_pp_bounds, dim3(grid_of((uint64_t)nc + 1)), dim3(kBlock), 0, s, g.off, incl.p, V, E, P, nc,
                     bounds.p);
  hipLaunchKernelGGL(k_pb_rows, dim3(nc), dim3(kBlock), 0, s, cnt.p, bounds.p, nb, in_base.p);
  hipLaunchKernelGGL(k_pb_cols, dim3(nb), dim3(kBlock), 0, s, cnt.p, nc, nb, col_pre.p, bin_size.p);
  hipLaunchKernelGGL(k_pb_bins, dim3(1), dim3(kBlock), 0, s, bin_size.p, nb, bin_off.p);
  hipLaunchKernelGGL(k_pb_delta, dim3(grid_of(cells)), dim3(kBlock), 0, s, in_base.p, col_pre.p, bin_off.p, nb,
                     cells, p->pb_delta.p);
  PR_LAUNCH_OK();
  cnt.reset();
  col_pre.reset();
  in_base.reset();
  // entries, bin-major slots and counts of the pairs
  DevBuf<uint32_t> bm_of, count;
  CG_TRY(p->pb_ent.ensure(P));
  CG_TRY(p->pb_val.ensure(P));
  CG_TRY(bm_of.ensure(P));
  CG_TRY(count.ensure(P));
  hipLaunchKernelGGL(k_pp_place, dim3(grid_of(P)), dim3(kBlock), 0, s, key.p, id.p, P, nb, p->pb_delta.p, pfirst.p,
                     p->pb_ent.p, bm_of.p, count.p);
  PR_LAUNCH_OK();
  key.reset();
  id.reset();
  DevBuf<uint64_t> eoff;
  CG_TRY(eoff.ensure(P + 1));
  {
    size_t tb = 0;
    CG_HIP(rocprim::exclusive_scan(nullptr, tb, count.p, eoff.p, (uint64_t)0, P, rocprim::plus<uint64_t>(), s));
    DevBuf<uint8_t> tmp;
    CG_TRY(tmp.ensure(tb));
    CG_HIP(rocprim::exclusive_scan(tmp.p, tb, count.p, eoff.p, (uint64_t)0, P, rocprim::plus<uint64_t>(), s));
  }
  CG_HIP(hipMemcpyAsync(eoff.p + P, &E, 8, hipMemcpyHostToDevice, s));
  count.reset();
  // the edges: destination within its bin, flagged at each pair's first edge;
  // zero entries past the end keep a tile's last group in bounds
  CG_TRY(p->pb_dloc.ensure(E + kPpPad));
  CG_TRY(p->pb_gbase.ensure((E + kPpGroup - 1) / kPpGroup + 2));
  CG_HIP(hipMemsetAsync(p->pb_dloc.p + E, 0, kPpPad * sizeof(uint16_t), s));
  hipLaunchKernelGGL(k_pp_edges, dim3(grid_of(E)), dim3(kBlock), 0, s, g.dst, incl.p, E, pfirst.p, bm_of.p, eoff.p,
                     p->pb_dloc.p, p->pb_gbase.p);
  PR_LAUNCH_OK();
  DevBuf<uint64_t> bin_edges;
  CG_TRY(bin_edges.ensure((uint64_t)nb + 1));
  hipLaunchKernelGGL(k_pp_bin_edges, dim3(grid_of((uint64_t)nb + 1)), dim3(kBlock), 0, s, bin_off.p, eoff.p, nb,
                     bin_edges.p);
  PR_LAUNCH_OK();
  std::vector<uint64_t> hb(nc + 1), he(nb + 1);
  CG_HIP(hipMemcpyAsync(hb.data(), bounds.p, hb.size() * 8, hipMemcpyDeviceToHost, s));
  CG_HIP(hipMemcpyAsync(he.data(), bin_edges.p, he.size() * 8, hipMemcpyDeviceToHost, s));
  CG_HIP(hipStreamSynchronize(s));
  return pb_tiles(p, hb, he, kPpBinTile, s);
}

template <int P>
void launch_tiers(const Csc& g, const PrPlan::Tiers& t, unsigned vblocks, const float* contrib_in, float base,
                  float damping, float* part, float* rank, float* contrib_out, hipStream_t s) {
  hipLaunchKernelGGL(k_pr_low<P>, dim3(vblocks), dim3(kBlock), 0, s, g, contrib_in, base, damping, part, rank,
                     contrib_out);
  if (t.n_med)
    hipLaunchKernelGGL(k_pr_med<P>, dim3((t.n_med + 3) / 4), dim3(kBlock), 0, s, g, t.med.p, t.n_med, contrib_in,
                       base, damping, part, rank, contrib_out);
  if (t.n_high) {
    // the longest rows (up to ~1.5 M in-edges at s25) would keep one workgroup
    // busy long after the rest: their chunks spread over the chip
    if (t.n_chunks)
      hipLaunchKernelGGL(k_pr_chunk, dim3(t.n_chunks), dim3(kBlock), 0, s, g.src, t.chunk_lo.p, t.chunk_hi.p,
                         contrib_in, t.partial.p);
    hipLaunchKernelGGL(k_pr_chunk_finish<P>, dim3((t.n_high + kBlock - 1) / kBlock), dim3(kBlock), 0, s, g,
                       t.high.p, t.n_high, t.row_chunk.p, t.partial.p, base, damping, part, rank, contrib_out);
  }
}

}  // namespace

int pr_plan_create(cg_graph* graph, uint32_t layout, uint32_t seg_log2, uint32_t kernel, PrPlan** plan,
                   hipStream_t s) {
  if (kernel > 3) return set_error(CG_ERR_INVALID, "pr_kernel %u outside [0, 3]", kernel);
  if (layout > kLayoutPairs) return set_error(CG_ERR_INVALID, "pr_layout %u outside [0, 5]", layout);
  // automatic: propagation blocking from 2^22 vertices up to its 2^25 limit
  // (measured per iteration on R-MAT edge factor 16 against the renumbered
  // pull: 0.30-0.35 vs 0.39-0.43 ms at 2^22, 0.54 vs 0.93 at 2^23; at 2^21
  // the two are within box-to-box spread, 0.17-0.20 vs 0.17 ms, and below it
  // the pull wins: 0.10 vs 0.12-0.16 ms at 2^20)
  // and over (source, bin) pairs from 2^24 (per iteration, R-MAT: ef16 0.82 vs
  // 1.09 ms at 2^24, 1.26 vs 2.12 at 2^25, ef44 2.92 vs 4.61 at 2^25; at 2^23
  // 0.66 vs 0.57 ms and below it the 2^13-vertex bins are too few to fill the
  // chip)
  if (layout == kLayoutAuto && seg_log2 == 0 && kernel == 0 && graph->V >= (1u << 22) &&
      ((uint64_t)graph->V + kPbSize - 1) / kPbSize <= kPbMaxBins && graph->E > 0)
    layout = graph->V >= (1u << 24) && graph->E < 0xFFFFFFFFull ? kLayoutPairs : kLayoutBlocked;
  if (layout == kLayoutPairs) {
    PrPlan* p = new PrPlan();
    const int st = build_pairs(graph, p, s);
    if (st != CG_OK) {
      delete p;
      return st;
    }
    *plan = p;
    return CG_OK;
  }
  if (layout == kLayoutBlocked) {
    // out-edges binned by destination; no transposed graph needed
    const DevGraph g{graph->d_off.p, graph->d_dst.p, nullptr, nullptr, nullptr, graph->V};
    PrPlan* p = new PrPlan();
    const int st = build_blocked(g, graph->E, p, s);
    if (st != CG_OK) {
      delete p;
      return st;
    }
    *plan = p;
    return CG_OK;
  }
  CG_TRY(build_transpose(graph));
  const DevGraph g{graph->d_off.p, graph->d_dst.p, nullptr, graph->d_in_off.p, graph->d_in_src.p, graph->V};
  const uint64_t E = graph->E;
  const uint32_t b = bits_for(g.V);
  if (layout == kLayoutAuto) {
    // an explicit window size asks for the segmented layout, 31 for plain pull;
    // otherwise relabelled pull above 2^16 vertices
    if (seg_log2 >= 1 && seg_log2 <= 30) layout = kLayoutSegmented;
    else if (seg_log2 == 0 && g.V > (1u << 16) && E > 0 && 2 * b <= 64) layout = kLayoutRelabel;
    else layout = kLayoutPlain;
  }
  uint32_t shift = seg_log2 >= 1 && seg_log2 <= 30 ? seg_log2 : 20;
  if (layout == kLayoutSegmented) {
    const uint32_t nseg = (uint32_t)(((uint64_t)g.V + (1ull << shift) - 1) >> shift);
    // segmented needs 32-bit run offsets and keys within 64 bits; plain otherwise
    if (E == 0 || E >= 0xFFFFFFFFull || 2 * b + bits_for(nseg) + 1 > 64 || (uint64_t)g.V <= (1ull << shift))
      layout = kLayoutPlain;
  }
  if (layout == kLayoutRelabel && (E == 0 || 2 * b > 64)) layout = kLayoutPlain;
  PrPlan* p = new PrPlan();
  int st = CG_OK;
  if (layout == kLayoutSegmented) {
    st = build_segmented(g, E, shift, p, s);
  } else if (layout == kLayoutRelabel) {
    st = build_relabelled(graph, p, s);
  } else {
    st = build_degrees(g, p, s);
    if (st == CG_OK) st = build_tiers<0>(Csc{g.in_off, g.in_src, p->deg.p, g.V, nullptr}, p->tiers[0], s);
    p->layout = kLayoutPlain;
  }
  if (st == CG_OK && p->layout == kLayoutRelabel) {
    if (kernel == 3) {
      st = build_bands(g.V, seg_log2 >= 1 && seg_log2 <= 30 ? seg_log2 : 19, p, s);
    } else if (kernel == 2) {
      st = build_tiles(p->r_off.p, g.V, E, p, s);
    } else {
      st = build_tiers<0>(Csc{p->r_off.p, p->r_src.p, p->deg.p, g.V, nullptr}, p->tiers[0], s);
    }
  } else if (st == CG_OK && p->layout == kLayoutPlain && kernel == 2) {
    st = build_tiles(g.in_off, g.V, E, p, s);
  }
  if (st != CG_OK) {
    delete p;
    return st;
  }
  *plan = p;
  return CG_OK;
}

void pr_plan_destroy(PrPlan* plan) { delete plan; }

void pr_plan_info(const PrPlan* plan, uint32_t* layout, double* stream_bytes) {
  *layout = (uint32_t)plan->layout;
  *stream_bytes = plan->stream_bytes;
}

int pr_run(const DevGraph& g, PrPlan* plan, int iters, float damping, float* rank, float* tmp_rank,
           float* contrib, float* contrib2, hipStream_t s, int profile, double* kernel_ms,
           uint32_t* launches) {
  const uint32_t V = g.V;
  if (!V) return CG_OK;
  const float base = (1.0f - damping) / (float)V;
  const unsigned vblocks = (unsigned)(((uint64_t)V + kBlock - 1) / kBlock);
  const bool relabel = plan->layout == kLayoutRelabel || plan->layout == kLayoutPairs;
  const Csc csc = relabel ? Csc{plan->r_off.p, plan->r_src.p, plan->deg.p, V, plan->hot.p}
                          : Csc{g.in_off, g.in_src, plan->deg.p, V, nullptr};
  const uint64_t plan_edges = plan->edges;
  hipLaunchKernelGGL(k_pr_init, dim3(std::min<unsigned>(vblocks, 65536)), dim3(kBlock), 0, s, V, plan->deg.p,
                     1.0f / (float)V, rank, contrib);
  float* r_cur = rank;
  float* r_nxt = tmp_rank;
  float* c_cur = contrib;
  float* c_nxt = contrib2;
  hipEvent_t a = nullptr, b = nullptr;
  if (profile) {
    CG_HIP(hipEventCreate(&a));
    CG_HIP(hipEventCreate(&b));
  }
  double total = 0.0;
  for (int it = 0; it < iters; ++it) {
    if (profile) CG_HIP(hipEventRecord(a, s));
    if (plan->layout == kLayoutBlocked || plan->layout == kLayoutPairs) {
      if (plan->pb_t1) {
        if (plan->pb_narrow)
          hipLaunchKernelGGL(k_pb_scatter<int32_t>, dim3(plan->pb_t1), dim3(kBlock),
                             plan->pb_nb * sizeof(int32_t) + kPbChunk * sizeof(float), s, plan->pb_ent.p,
                             plan->pb_t1_chunk.p, plan->pb_t1_lo.p, plan->pb_t1_hi.p, plan->pb_nb, V,
                             plan->pb_delta.p, c_cur, plan->pb_val.p);
        else
          hipLaunchKernelGGL(k_pb_scatter<int64_t>, dim3(plan->pb_t1), dim3(kBlock),
                             plan->pb_nb * sizeof(int64_t) + kPbChunk * sizeof(float), s, plan->pb_ent.p,
                             plan->pb_t1_chunk.p, plan->pb_t1_lo.p, plan->pb_t1_hi.p, plan->pb_nb, V,
                             plan->pb_delta.p, c_cur, plan->pb_val.p);
      }
      if (plan->layout == kLayoutPairs)
        hipLaunchKernelGGL(k_pbp_gather, dim3(plan->pb_t2), dim3(kPpBlock), 0, s, plan->pb_val.p, plan->pb_dloc.p,
                           plan->pb_gbase.p, plan->pb_t2_bin.p, plan->pb_t2_lo.p, plan->pb_t2_hi.p,
                           plan->pb_t2_multi.p, V, plan->deg.p, base, damping, plan->pb_part.p, r_nxt, c_nxt);
      else
        hipLaunchKernelGGL(k_pb_gather, dim3(plan->pb_t2), dim3(kBlock), 0, s, plan->pb_val.p, plan->pb_dloc.p,
                           plan->pb_t2_bin.p, plan->pb_t2_lo.p, plan->pb_t2_hi.p, plan->pb_t2_multi.p, V,
                           plan->deg.p, base, damping, plan->pb_sum.p, r_nxt, c_nxt);
      if (plan->pb_multi && plan->layout == kLayoutPairs)
        hipLaunchKernelGGL(k_pbp_finish, dim3(plan->pb_multi * (kPbSize / kBlock)), dim3(kBlock), 0, s,
                           plan->pb_multi_bins.p, plan->pb_mrow.p, V, plan->deg.p, base, damping, plan->pb_part.p,
                           r_nxt, c_nxt);
      else if (plan->pb_multi)
        hipLaunchKernelGGL(k_pb_finish, dim3(plan->pb_multi * (kPbSize / kBlock)), dim3(kBlock), 0, s,
                           plan->pb_multi_bins.p, V, plan->deg.p, base, damping, plan->pb_sum.p, r_nxt, c_nxt);
    } else if (plan->layout == kLayoutSegmented) {
      for (uint32_t w = 0; w < plan->nseg; ++w) {
        const uint32_t r0 = plan->seg_run[w], r1 = plan->seg_run[w + 1];
        if (r1 > r0)
          hipLaunchKernelGGL(k_seg_short, dim3((r1 - r0 + kBlock - 1) / kBlock), dim3(kBlock), 0, s, plan->src.p,
                             plan->run_dst.p, plan->run_off.p, r0, r1, c_cur, plan->sum.p);
        const uint32_t l0 = plan->seg_long[w], l1 = plan->seg_long[w + 1];
        if (l1 > l0)
          hipLaunchKernelGGL(k_seg_long, dim3((l1 - l0 + 3) / 4), dim3(kBlock), 0, s, plan->src.p, plan->run_dst.p,
                             plan->run_off.p, plan->long_runs.p, l0, l1, c_cur, plan->sum.p);
      }
      hipLaunchKernelGGL(k_seg_finish, dim3(std::min<unsigned>(vblocks, 65536)), dim3(kBlock), 0, s, V,
                         plan->deg.p, base, damping, plan->sum.p, r_nxt, c_nxt);
    } else if (plan->tiled) {
      if (plan->lead_rows)
        hipLaunchKernelGGL(k_pr_lead, dim3(grid_of(plan->lead_rows)), dim3(kBlock), 0, s, plan->lead_rows,
                           plan->deg.p, base, r_nxt, c_nxt);
      if (plan->ntiles) {
        hipLaunchKernelGGL(k_pr_tile, dim3((unsigned)plan->ntiles), dim3(kTileThreads), 0, s, csc, plan_edges,
                           plan->tile_row.p, c_cur, base, damping, r_nxt, c_nxt, plan->carry.p, plan->head_part.p);
        hipLaunchKernelGGL(k_pr_fixup, dim3(grid_of(plan->ntiles)), dim3(kBlock), 0, s, csc, plan_edges,
                           plan->ntiles, plan->tile_row.p, plan->carry.p, plan->head_part.p, base, damping, r_nxt,
                           c_nxt);
      }
    } else if (plan->banded) {
      launch_tiers<1>(csc, plan->tiers[0], vblocks, c_cur, base, damping, plan->part.p, r_nxt, c_nxt, s);
      launch_tiers<2>(csc, plan->tiers[1], vblocks, c_cur, base, damping, plan->part.p, r_nxt, c_nxt, s);
    } else {
      launch_tiers<0>(csc, plan->tiers[0], vblocks, c_cur, base, damping, nullptr, r_nxt, c_nxt, s);
    }
    if (profile) {
      CG_HIP(hipEventRecord(b, s));
      CG_HIP(hipEventSynchronize(b));
      float ms = 0.f;
      CG_HIP(hipEventElapsedTime(&ms, a, b));
      total += ms;
    }
    std::swap(r_cur, r_nxt);
    std::swap(c_cur, c_nxt);
  }
  if (relabel) {
    // back to the caller's ids: r_nxt is free after the last swap
    hipLaunchKernelGGL(k_unpermute, dim3(std::min<unsigned>(vblocks, 65536)), dim3(kBlock), 0, s, plan->perm.p,
                       r_cur, V, r_nxt);
    std::swap(r_cur, r_nxt);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(CG_ERR_HIP, "PageRank kernels: %s", hipGetErrorString(e));
  if (r_cur != rank) CG_HIP(hipMemcpyAsync(rank, r_cur, (uint64_t)V * sizeof(float), hipMemcpyDeviceToDevice, s));
  CG_HIP(hipStreamSynchronize(s));
  if (profile) {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  }
  if (kernel_ms) *kernel_ms = total;
  if (launches) *launches = (uint32_t)iters;
  return CG_OK;
}

}  // namespace cg
