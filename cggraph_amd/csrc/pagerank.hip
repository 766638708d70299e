// PageRank by pulling over the transposed graph, FP32 with a fixed summation
// order.
//
// The reference's PR is not functional (it accumulates into the source on
// uint32, my_project/CGgraph/device_algorithm.cuh:505-545); only its update
// rule is recoverable: PR(v) = (1 - d)/N + d * sum_{u -> v} PR(u)/outdeg(u),
// d = 0.85 (:543), dangling mass dropped.  One iteration here is one sweep per
// in-degree tier:
//   low    (in-degree < 32)    one lane per vertex, sequential sum in CSC order
//   medium (32 .. 8191)        one wavefront per vertex: lane j sums edges
//                              j, j+64, ... in order, then a fixed xor-butterfly
//   high   (>= 8192)           one workgroup per vertex: 256 strided partial
//                              sums, then a fixed LDS tree
// Tier membership depends only on the graph, so the result is bit-identical
// from run to run.  Each sweep also writes contrib(v) = PR(v)/outdeg(v) for the
// next iteration, so the division is fused into the producing kernel.
#include <algorithm>

#include "kernels.hpp"

namespace cg {

struct PrPlan {
  uint32_t n_med = 0, n_high = 0;
  DevBuf<uint32_t> med, high;
};

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kMed = 32, kHigh = 8192;

__global__ void k_classify(const uint64_t* __restrict__ in_off, uint32_t V, uint32_t* __restrict__ med,
                           uint32_t* __restrict__ high, uint32_t* __restrict__ cnt) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t d = in_off[v + 1] - in_off[v];
    if (d >= kHigh) high[atomicAdd(cnt + 1, 1u)] = (uint32_t)v;
    else if (d >= kMed) med[atomicAdd(cnt, 1u)] = (uint32_t)v;
  }
}

__device__ __forceinline__ void finish(uint32_t v, float sum, float base, float damping,
                                       const uint64_t* __restrict__ off, float* __restrict__ rank,
                                       float* __restrict__ contrib) {
  const float r = base + damping * sum;
  rank[v] = r;
  const uint64_t deg = off[v + 1] - off[v];
  contrib[v] = deg ? r / (float)deg : 0.0f;
}

__global__ __launch_bounds__(kBlock) void k_pr_low(DevGraph g, const float* __restrict__ contrib_in,
                                                   float base, float damping, float* __restrict__ rank,
                                                   float* __restrict__ contrib_out) {
  const uint64_t v = blockIdx.x * (uint64_t)kBlock + threadIdx.x;
  if (v >= g.V) return;
  const uint64_t s = g.in_off[v], e = g.in_off[v + 1];
  if (e - s >= kMed) return;
  float sum = 0.0f;
  for (uint64_t k = s; k < e; ++k) sum += contrib_in[g.in_src[k]];
  finish((uint32_t)v, sum, base, damping, g.off, rank, contrib_out);
}

__global__ __launch_bounds__(kBlock) void k_pr_med(DevGraph g, const uint32_t* __restrict__ list,
                                                   uint32_t n, const float* __restrict__ contrib_in,
                                                   float base, float damping, float* __restrict__ rank,
                                                   float* __restrict__ contrib_out) {
  const uint64_t i = blockIdx.x * (uint64_t)(kBlock / 64) + threadIdx.x / 64;
  if (i >= n) return;  // wave-uniform
  const uint32_t lane = __lane_id();
  const uint32_t v = list[i];
  const uint64_t s = g.in_off[v], e = g.in_off[v + 1];
  float sum = 0.0f;
  for (uint64_t k = s + lane; k < e; k += 64) sum += contrib_in[g.in_src[k]];
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) sum += __shfl_xor(sum, d, 64);
  if (lane == 0) finish(v, sum, base, damping, g.off, rank, contrib_out);
}

__global__ __launch_bounds__(kBlock) void k_pr_high(DevGraph g, const uint32_t* __restrict__ list,
                                                    const float* __restrict__ contrib_in, float base,
                                                    float damping, float* __restrict__ rank,
                                                    float* __restrict__ contrib_out) {
  __shared__ float part[kBlock];
  const uint32_t v = list[blockIdx.x];
  const uint64_t s = g.in_off[v], e = g.in_off[v + 1];
  float sum = 0.0f;
  for (uint64_t k = s + threadIdx.x; k < e; k += kBlock) sum += contrib_in[g.in_src[k]];
  part[threadIdx.x] = sum;
  __syncthreads();
  for (int w = kBlock / 2; w >= 1; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) finish(v, part[0], base, damping, g.off, rank, contrib_out);
}

__global__ void k_pr_init(DevGraph g, float r0, float* __restrict__ rank, float* __restrict__ contrib) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < g.V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    rank[v] = r0;
    const uint64_t deg = g.off[v + 1] - g.off[v];
    contrib[v] = deg ? r0 / (float)deg : 0.0f;
  }
}

}  // namespace

int pr_plan_create(const DevGraph& g, PrPlan** plan, hipStream_t s) {
  PrPlan* p = new PrPlan();
  DevBuf<uint32_t> cnt;
  int st = cnt.ensure(2);
  if (st == CG_OK) st = p->med.ensure(g.V ? g.V : 1);
  if (st == CG_OK) st = p->high.ensure(g.V ? g.V : 1);
  if (st != CG_OK) { delete p; return st; }
  hipError_t he = hipMemsetAsync(cnt.p, 0, 2 * sizeof(uint32_t), s);
  if (he == hipSuccess && g.V) {
    unsigned blocks = (unsigned)std::min<uint64_t>(((uint64_t)g.V + kBlock - 1) / kBlock, 65536);
    hipLaunchKernelGGL(k_classify, dim3(blocks), dim3(kBlock), 0, s, g.in_off, g.V, p->med.p, p->high.p,
                       cnt.p);
    he = hipGetLastError();
  }
  uint32_t h[2] = {0, 0};
  if (he == hipSuccess) he = hipMemcpyAsync(h, cnt.p, sizeof(h), hipMemcpyDeviceToHost, s);
  if (he == hipSuccess) he = hipStreamSynchronize(s);
  if (he != hipSuccess) {
    delete p;
    return set_error(CG_ERR_HIP, "PageRank plan: %s", hipGetErrorString(he));
  }
  p->n_med = h[0];
  p->n_high = h[1];
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
  hipLaunchKernelGGL(k_pr_init, dim3(std::min<unsigned>(vblocks, 65536)), dim3(kBlock), 0, s, g,
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
    hipLaunchKernelGGL(k_pr_low, dim3(vblocks), dim3(kBlock), 0, s, g, c_cur, base, damping, r_nxt, c_nxt);
    if (plan->n_med)
      hipLaunchKernelGGL(k_pr_med, dim3((plan->n_med + 3) / 4), dim3(kBlock), 0, s, g, plan->med.p,
                         plan->n_med, c_cur, base, damping, r_nxt, c_nxt);
    if (plan->n_high)
      hipLaunchKernelGGL(k_pr_high, dim3(plan->n_high), dim3(kBlock), 0, s, g, plan->high.p, c_cur, base,
                         damping, r_nxt, c_nxt);
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
