// PageRank by pulling over the transposed graph, FP32 with a fixed summation
// order.
//
// The reference's PR is not functional (it accumulates into the source on
// uint32, my_project/CGgraph/device_algorithm.cuh:505-545); only its update
// rule is recoverable: PR(v) = (1 - d)/N + d * sum_{u -> v} PR(u)/outdeg(u),
// d = 0.85 (:543), dangling mass dropped.  contrib(u) = PR(u)/outdeg(u) is
// written by the kernel that produces PR(u), so each iteration reads it once.
//
// Three layouts of the in-edges (CSC), chosen once per graph:
//  plain       one sweep over the transposed graph per iteration, rows tiered
//              by in-degree: one lane / one wavefront / one workgroup per row.
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
// Every layout fixes each row's summation order from the graph alone, so
// results are bit-identical from run to run.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <vector>

#include "kernels.hpp"

namespace cg {

enum { kLayoutAuto = 0, kLayoutPlain = 1, kLayoutRelabel = 2, kLayoutSegmented = 3 };

struct PrPlan {
  int layout = kLayoutPlain;
  // plain and relabelled pull: tiers over the rows of the CSC in use
  uint32_t n_med = 0, n_high = 0;
  DevBuf<uint32_t> med, high;
  DevBuf<uint32_t> deg;  // out-degree of each vertex (in the numbering in use)
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
};

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kMed = 32, kHigh = 8192;
constexpr uint32_t kLongRun = 64;

// The in-edges the pull kernels read: row v's sources are src[off[v]..off[v+1]).
struct Csc {
  const uint64_t* off;
  const uint32_t* src;
  const uint32_t* deg;  // out-degree, for contrib = rank / deg
  uint32_t V;
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

__device__ __forceinline__ uint32_t ld_stream(const uint32_t* p) { return __builtin_nontemporal_load(p); }

__global__ void k_classify(const uint64_t* __restrict__ in_off, uint32_t V, uint32_t* __restrict__ med,
                           uint32_t* __restrict__ high, uint32_t* __restrict__ cnt) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t d = in_off[v + 1] - in_off[v];
    if (d >= kHigh) high[atomicAdd(cnt + 1, 1u)] = (uint32_t)v;
    else if (d >= kMed) med[atomicAdd(cnt, 1u)] = (uint32_t)v;
  }
}

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

// Rows with fewer than kMed in-edges: one lane per row, sources in order.
__global__ __launch_bounds__(kBlock) void k_pr_low(Csc g, const float* __restrict__ contrib_in, float base,
                                                   float damping, float* __restrict__ rank,
                                                   float* __restrict__ contrib_out) {
  const uint64_t v = blockIdx.x * (uint64_t)kBlock + threadIdx.x;
  if (v >= g.V) return;
  const uint64_t s = g.off[v], e = g.off[v + 1];
  if (e - s >= kMed) return;
  float sum = 0.0f;
  for (uint64_t k = s; k < e; ++k) sum += contrib_in[ld_stream(g.src + k)];
  finish((uint32_t)v, sum, base, damping, g.deg, rank, contrib_out);
}

// Lane-strided partial sum over [s, e): the lane adds sources s+lane,
// s+lane+stride, ... in that order; four sources are fetched ahead so four
// gathers are in flight per lane.
template <int STRIDE>
__device__ __forceinline__ float strided_sum(const uint32_t* __restrict__ src, const float* __restrict__ contrib,
                                             uint64_t s, uint64_t e, uint32_t lane) {
  float sum = 0.0f;
  uint64_t k = s + lane;
  for (; k + 3 * STRIDE < e; k += 4 * STRIDE) {
    const uint32_t u0 = ld_stream(src + k), u1 = ld_stream(src + k + STRIDE);
    const uint32_t u2 = ld_stream(src + k + 2 * STRIDE), u3 = ld_stream(src + k + 3 * STRIDE);
    const float c0 = contrib[u0], c1 = contrib[u1], c2 = contrib[u2], c3 = contrib[u3];
    sum += c0;
    sum += c1;
    sum += c2;
    sum += c3;
  }
  for (; k < e; k += STRIDE) sum += contrib[ld_stream(src + k)];
  return sum;
}

__global__ __launch_bounds__(kBlock) void k_pr_med(Csc g, const uint32_t* __restrict__ list, uint32_t n,
                                                   const float* __restrict__ contrib_in, float base,
                                                   float damping, float* __restrict__ rank,
                                                   float* __restrict__ contrib_out) {
  const uint64_t i = blockIdx.x * (uint64_t)(kBlock / 64) + threadIdx.x / 64;
  if (i >= n) return;  // wave-uniform
  const uint32_t lane = __lane_id();
  const uint32_t v = list[i];
  float sum = strided_sum<64>(g.src, contrib_in, g.off[v], g.off[v + 1], lane);
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) sum += __shfl_xor(sum, d, 64);
  if (lane == 0) finish(v, sum, base, damping, g.deg, rank, contrib_out);
}

__global__ __launch_bounds__(kBlock) void k_pr_high(Csc g, const uint32_t* __restrict__ list,
                                                    const float* __restrict__ contrib_in, float base,
                                                    float damping, float* __restrict__ rank,
                                                    float* __restrict__ contrib_out) {
  __shared__ float part[kBlock];
  const uint32_t v = list[blockIdx.x];
  part[threadIdx.x] = strided_sum<kBlock>(g.src, contrib_in, g.off[v], g.off[v + 1], threadIdx.x);
  __syncthreads();
  for (int w = kBlock / 2; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) finish(v, part[0], base, damping, g.deg, rank, contrib_out);
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

// med / high row lists of a CSC (any order: each row's sum is fixed by the row).
int build_tiers(const uint64_t* in_off, uint32_t V, PrPlan* p, hipStream_t s) {
  DevBuf<uint32_t> cnt;
  CG_TRY(cnt.ensure(2));
  CG_TRY(p->med.ensure(V ? V : 1));
  CG_TRY(p->high.ensure(V ? V : 1));
  CG_HIP(hipMemsetAsync(cnt.p, 0, 2 * sizeof(uint32_t), s));
  if (V) {
    hipLaunchKernelGGL(k_classify, dim3(grid_of(V)), dim3(kBlock), 0, s, in_off, V, p->med.p, p->high.p, cnt.p);
    PR_LAUNCH_OK();
  }
  uint32_t h[2] = {0, 0};
  CG_HIP(hipMemcpyAsync(h, cnt.p, sizeof(h), hipMemcpyDeviceToHost, s));
  CG_HIP(hipStreamSynchronize(s));
  p->n_med = h[0];
  p->n_high = h[1];
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
  CG_TRY(build_tiers(p->r_off.p, V, p, s));
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

}  // namespace

int pr_plan_create(cg_graph* graph, uint32_t layout, uint32_t seg_log2, PrPlan** plan, hipStream_t s) {
  CG_TRY(build_transpose(graph));
  const DevGraph g{graph->d_off.p, graph->d_dst.p, nullptr, graph->d_in_off.p, graph->d_in_src.p, graph->V};
  const uint64_t E = graph->E;
  if (layout > kLayoutSegmented) return set_error(CG_ERR_INVALID, "pr_layout %u outside [0, 3]", layout);
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
    if (st == CG_OK) st = build_tiers(g.in_off, g.V, p, s);
    p->layout = kLayoutPlain;
  }
  if (st != CG_OK) {
    delete p;
    return st;
  }
  *plan = p;
  return CG_OK;
}

void pr_plan_destroy(PrPlan* plan) { delete plan; }

int pr_run(const DevGraph& g, PrPlan* plan, int iters, float damping, float* rank, float* tmp_rank,
           float* contrib, float* contrib2, hipStream_t s, int profile, double* kernel_ms,
           uint32_t* launches) {
  const uint32_t V = g.V;
  if (!V) return CG_OK;
  const float base = (1.0f - damping) / (float)V;
  const unsigned vblocks = (unsigned)(((uint64_t)V + kBlock - 1) / kBlock);
  const bool relabel = plan->layout == kLayoutRelabel;
  const Csc csc = relabel ? Csc{plan->r_off.p, plan->r_src.p, plan->deg.p, V}
                          : Csc{g.in_off, g.in_src, plan->deg.p, V};
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
    if (plan->layout == kLayoutSegmented) {
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
    } else {
      hipLaunchKernelGGL(k_pr_low, dim3(vblocks), dim3(kBlock), 0, s, csc, c_cur, base, damping, r_nxt, c_nxt);
      if (plan->n_med)
        hipLaunchKernelGGL(k_pr_med, dim3((plan->n_med + 3) / 4), dim3(kBlock), 0, s, csc, plan->med.p,
                           plan->n_med, c_cur, base, damping, r_nxt, c_nxt);
      if (plan->n_high)
        hipLaunchKernelGGL(k_pr_high, dim3(plan->n_high), dim3(kBlock), 0, s, csc, plan->high.p, c_cur, base,
                           damping, r_nxt, c_nxt);
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
