// WCC (weakly connected components) by min-label hooking on a symmetric graph.
//
// The reference computes CC by pushing labels: value[i] = i, then every
// active vertex lowers its neighbours' labels with write_min / atomicMin
// (my_project/CGgraph/host_algorithm.hpp:412-433, device_algorithm.cuh:430-465).
// On a symmetric graph the fixed point is the minimum vertex id of each
// connected component.  This file reaches the same labels by hooking:
//   parent[v] = v; for every edge (u, v) link the two trees by pointing the
//   larger root at the smaller one (CAS); finally compress every path.
// parent[x] <= x holds throughout, so each tree's root is its minimum id and
// the final labels equal the label-propagation fixed point bit for bit.
// The edge order follows Afforest (Sutton, Ben-Nun, Bar, IPDPS'18): link the
// first `sample_rounds` neighbours of every vertex, compress, find the most
// frequent root by sampling, then let only vertices outside that component
// link their remaining edges -- an edge between the large component and
// another vertex is seen from the other side, because the graph is symmetric.
#include <algorithm>
#include <vector>

#include "kernels.hpp"

namespace cg {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ uint32_t ld_relaxed(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Link the trees of u and v (Afforest's Link): climb towards the roots; when
// the larger candidate is a root, CAS it onto the smaller one.
__device__ void link(uint32_t u, uint32_t v, uint32_t* __restrict__ parent) {
  uint32_t p1 = ld_relaxed(parent + u);
  uint32_t p2 = ld_relaxed(parent + v);
  while (p1 != p2) {
    const uint32_t high = p1 > p2 ? p1 : p2;
    const uint32_t low = p1 + p2 - high;
    const uint32_t p_high = ld_relaxed(parent + high);
    if (p_high == low) break;
    if (p_high == high) {
      uint32_t expected = high;
      if (__hip_atomic_compare_exchange_strong(parent + high, &expected, low, __ATOMIC_RELAXED,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        break;
    }
    p1 = ld_relaxed(parent + ld_relaxed(parent + high));
    p2 = ld_relaxed(parent + low);
  }
}

__global__ void k_link_round(DevGraph g, uint32_t r, uint32_t* __restrict__ parent) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < g.V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t s = g.off[v], e = g.off[v + 1];
    if (s + r < e) link((uint32_t)v, g.dst[s + r], parent);
  }
}

__global__ void k_compress(uint32_t V, uint32_t* __restrict__ parent) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t p = ld_relaxed(parent + v);
    uint32_t gp = ld_relaxed(parent + p);
    while (p != gp) {
      __hip_atomic_store(parent + v, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      p = gp;
      gp = ld_relaxed(parent + p);
    }
  }
}

__global__ void k_sample_roots(const uint32_t* __restrict__ parent, uint32_t V, uint64_t seed,
                               uint32_t n, uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    uint64_t z = seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    out[i] = parent[(uint32_t)(z % V)];
  }
}

// The most frequent root among the samples (ties to the smaller id), on the
// device so the host never waits: one workgroup of kSamples threads, each
// counting its sample's occurrences in LDS.  Any choice gives the same
// labels; a good one lets most vertices skip their remaining edges.
constexpr uint32_t kSamples = 1024;

__global__ __launch_bounds__(kSamples) void k_sample_mode(const uint32_t* __restrict__ samp,
                                                          uint32_t* __restrict__ mode) {
  __shared__ uint32_t x[kSamples];
  __shared__ unsigned long long best[kSamples / 64];
  const uint32_t t = threadIdx.x;
  x[t] = samp[t];
  __syncthreads();
  uint32_t c = 0;
  for (uint32_t j = 0; j < kSamples; ++j) c += x[j] == x[t];
  // key: larger count first, then smaller id
  unsigned long long key = ((unsigned long long)c << 32) | (0xFFFFFFFFu - x[t]);
  for (int d = 32; d >= 1; d >>= 1) {
    const unsigned long long o = __shfl_xor(key, d, 64);
    key = o > key ? o : key;
  }
  if (__lane_id() == 0) best[t / 64] = key;
  __syncthreads();
  if (t == 0) {
    unsigned long long k = best[0];
    for (uint32_t w = 1; w < kSamples / 64; ++w) k = best[w] > k ? best[w] : k;
    *mode = 0xFFFFFFFFu - (uint32_t)(k & 0xFFFFFFFFu);
  }
}

// Remaining edges of vertices outside the large component.  A wavefront
// takes 64 consecutive vertices, tests them one per lane (parent outside the
// large component, edges left after the sampled ones), then links the rows
// that need it one at a time with the row spread over the 64 lanes -- most
// vertices are in the large component or have no edges left, so the test is
// the bulk of the work and runs 64 wide.
__global__ __launch_bounds__(kBlock) void k_link_rest(DevGraph g, uint32_t r, const uint32_t* __restrict__ mode,
                                                      uint32_t* __restrict__ parent) {
  const uint32_t skip = *mode;
  const uint32_t lane = __lane_id();
  const uint64_t waves = (uint64_t)gridDim.x * (kBlock / 64);
  for (uint64_t base = (blockIdx.x * (uint64_t)(kBlock / 64) + threadIdx.x / 64) * 64; base < g.V;
       base += waves * 64) {
    const uint64_t v = base + lane;
    uint64_t s = 0, e = 0;
    if (v < g.V && ld_relaxed(parent + v) != skip) {
      s = g.off[v] + r;
      e = g.off[v + 1];
    }
    for (uint64_t need = __ballot(s < e); need; need &= need - 1) {
      const int l = __builtin_ctzll(need);
      const uint64_t rs = __shfl(s, l, 64), re = __shfl(e, l, 64);
      for (uint64_t k = rs + lane; k < re; k += 64) link((uint32_t)(base + l), g.dst[k], parent);
    }
  }
}

unsigned grid_of(uint64_t n) {
  uint64_t b = (n + kBlock - 1) / kBlock;
  if (b < 1) b = 1;
  return (unsigned)(b < 65536 ? b : 65536);
}

}  // namespace

int wcc_hooking(const DevGraph& g, uint32_t* parent, uint32_t sample_rounds, hipStream_t s,
                uint64_t* edges_relaxed) {
  const uint32_t V = g.V;
  if (edges_relaxed) *edges_relaxed = 0;
  if (V == 0) return CG_OK;
  CG_TRY(launch_iota_u32(parent, V, s));
  for (uint32_t r = 0; r < sample_rounds; ++r) {
    hipLaunchKernelGGL(k_link_round, dim3(grid_of(V)), dim3(kBlock), 0, s, g, r, parent);
    hipLaunchKernelGGL(k_compress, dim3(grid_of(V)), dim3(kBlock), 0, s, V, parent);
  }
  // most frequent root among kSamples samples, found on the device
  DevBuf<uint32_t> samp;
  CG_TRY(samp.ensure(kSamples + 1));
  hipLaunchKernelGGL(k_sample_roots, dim3(kSamples / kBlock), dim3(kBlock), 0, s, parent, V,
                     (uint64_t)0x5EEDull, kSamples, samp.p);
  hipLaunchKernelGGL(k_sample_mode, dim3(1), dim3(kSamples), 0, s, samp.p, samp.p + kSamples);
  const uint64_t waves_grid = std::min<uint64_t>(((uint64_t)V + 255) / 256, 65536);
  hipLaunchKernelGGL(k_link_rest, dim3((unsigned)(waves_grid ? waves_grid : 1)), dim3(kBlock), 0, s, g,
                     sample_rounds, samp.p + kSamples, parent);
  hipLaunchKernelGGL(k_compress, dim3(grid_of(V)), dim3(kBlock), 0, s, V, parent);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(CG_ERR_HIP, "WCC kernels: %s", hipGetErrorString(e));
  // the sample buffer is freed on return: the kernels reading it end first
  CG_HIP(hipStreamSynchronize(s));
  if (edges_relaxed) *edges_relaxed = 0;
  return CG_OK;
}

}  // namespace cg
