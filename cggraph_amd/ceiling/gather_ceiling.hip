// gather_ceiling: what one MI355X sustains on the memory pattern of the
// traversal kernels, measured rather than taken from the data sheet.
//
// The push expansion and the PageRank pull each spend almost all their time
// on 4-byte gathers at data-dependent addresses (value[dst[e]] and
// contrib[src[e]]); every one moves a whole cache line, so their rate, not
// the HBM streaming peak, bounds the kernels.  This program times
//   * a streaming copy (the HBM reference),
//   * gathers  s += value[idx[i]]      (idx streamed, 8 loads in flight per lane),
//   * relaxations  if (c < value[d]) atomicMin(value + d, c)
// over value arrays of 2^24 and 2^25 entries, with the index stream drawn
//   * uniformly, or
//   * like R-MAT destinations (each of the scale bits set with probability
//     b + d = 0.24) and then renumbered by expected degree (ascending number
//     of set bits), which is where the engine's out-degree renumbering puts
//     an R-MAT graph's hot vertices.
// It prints one JSON object on stdout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <vector>

#define CK(x)                                                                   \
  do {                                                                          \
    hipError_t e_ = (x);                                                        \
    if (e_ != hipSuccess) {                                                     \
      fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      return 1;                                                                 \
    }                                                                           \
  } while (0)

namespace {

constexpr int kBlock = 256;
constexpr int kGrid = 8192;
constexpr uint64_t kN = 256ull << 20;  // indices per pass (1 GiB)

__device__ __forceinline__ uint64_t splitmix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// idx[i] in [0, 2^scale): uniform, or R-MAT-like bits mapped through rank[]
__global__ void k_indices(uint32_t* __restrict__ idx, uint64_t n, int scale, const uint32_t* __restrict__ rank) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t r = splitmix(i * 0x2545F4914F6CDD1Dull + 1);
    uint32_t v = 0;
    if (rank) {
      for (int b = 0; b < scale; ++b) {
        if (b && (b & 7) == 0) r = splitmix(r);
        if (((r >> ((b & 7) * 8)) & 0xFF) >= 194) v |= 1u << b;  // 62/256 = 0.242
      }
      v = rank[v];
    } else {
      v = (uint32_t)(r & ((1ull << scale) - 1));
    }
    idx[i] = v;
  }
}

__global__ __launch_bounds__(kBlock) void k_gather(const uint32_t* __restrict__ idx, uint64_t n,
                                                   const uint32_t* __restrict__ value, uint32_t* __restrict__ out) {
  constexpr int U = 8;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  uint32_t s = 0;
  for (; i + (U - 1) * stride < n; i += U * stride) {
    uint32_t d[U];
#pragma unroll
    for (int j = 0; j < U; ++j) d[j] = idx[i + j * stride];
#pragma unroll
    for (int j = 0; j < U; ++j) s += value[d[j]];
  }
  for (; i < n; i += stride) s += value[idx[i]];
  out[blockIdx.x * (uint64_t)blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void k_relax(const uint32_t* __restrict__ idx, uint64_t n,
                                                  uint32_t* __restrict__ value) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t d = idx[i], c = (uint32_t)(i & 127) + 1;
    if (c < value[d]) atomicMin(value + d, c);
  }
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// streaming copy, four 16-byte loads in flight per lane, non-temporal
__global__ __launch_bounds__(kBlock) void k_copy(const u32x4* __restrict__ a, u32x4* __restrict__ b, uint64_t n) {
  constexpr int U = 4;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  for (; i + (U - 1) * stride < n; i += U * stride) {
    u32x4 t[U];
#pragma unroll
    for (int j = 0; j < U; ++j) t[j] = __builtin_nontemporal_load(a + i + j * stride);
#pragma unroll
    for (int j = 0; j < U; ++j) __builtin_nontemporal_store(t[j], b + i + j * stride);
  }
  for (; i < n; i += stride) b[i] = a[i];
}

// rank[v] = position of v when [0, 2^scale) is ordered by (set bits, v)
std::vector<uint32_t> degree_rank(int scale) {
  const uint32_t V = 1u << scale;
  std::vector<uint64_t> start(scale + 2, 0);
  for (uint32_t v = 0; v < V; ++v) ++start[__builtin_popcount(v) + 1];
  for (int k = 1; k <= scale + 1; ++k) start[k] += start[k - 1];
  std::vector<uint32_t> rank(V);
  for (uint32_t v = 0; v < V; ++v) rank[v] = (uint32_t)start[__builtin_popcount(v)]++;
  return rank;
}

}  // namespace

int main() {
  uint32_t *idx = nullptr, *out = nullptr;
  u32x4 *a = nullptr, *b = nullptr;
  CK(hipMalloc(&idx, kN * 4));
  CK(hipMalloc(&out, (uint64_t)kGrid * kBlock * 4));
  const uint64_t copy_bytes = 1ull << 30;
  CK(hipMalloc(&a, copy_bytes));
  CK(hipMalloc(&b, copy_bytes));
  CK(hipMemset(a, 1, copy_bytes));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  auto best_of = [&](int reps, auto&& launch, float& best) -> hipError_t {
    best = 1e30f;
    for (int r = 0; r < reps; ++r) {
      hipError_t err = launch();
      if (err != hipSuccess) return err;
      if ((err = hipEventRecord(e0)) != hipSuccess) return err;
      launch.timed();
      if ((err = hipEventRecord(e1)) != hipSuccess) return err;
      if ((err = hipEventSynchronize(e1)) != hipSuccess) return err;
      float ms = 0.f;
      if ((err = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) return err;
      best = std::min(best, ms);
    }
    return hipGetLastError();
  };
  struct Copy {
    u32x4 *a, *b;
    uint64_t n;
    hipError_t operator()() const { return hipSuccess; }
    void timed() const { hipLaunchKernelGGL(k_copy, dim3(kGrid), dim3(kBlock), 0, 0, a, b, n); }
  } copy{a, b, copy_bytes / 16};
  float ms = 0.f;
  CK(best_of(4, copy, ms));
  printf("{\"copy_GBps\": %.1f", 2.0 * copy_bytes / ms / 1e6);

  for (int scale : {24, 25}) {
    const uint64_t V = 1ull << scale;
    uint32_t *value = nullptr, *rank_d = nullptr;
    CK(hipMalloc(&value, V * 4));
    CK(hipMalloc(&rank_d, V * 4));
    const std::vector<uint32_t> rank = degree_rank(scale);
    CK(hipMemcpy(rank_d, rank.data(), V * 4, hipMemcpyHostToDevice));
    for (int skewed : {0, 1}) {
      hipLaunchKernelGGL(k_indices, dim3(kGrid), dim3(kBlock), 0, 0, idx, kN, scale, skewed ? rank_d : nullptr);
      CK(hipGetLastError());
      CK(hipMemset(value, 0x7f, V * 4));
      struct Gather {
        const uint32_t* idx;
        const uint32_t* value;
        uint32_t* out;
        hipError_t operator()() const { return hipSuccess; }
        void timed() const { hipLaunchKernelGGL(k_gather, dim3(kGrid), dim3(kBlock), 0, 0, idx, kN, value, out); }
      } gather{idx, value, out};
      CK(best_of(3, gather, ms));
      const char* kind = skewed ? "skewed" : "uniform";
      printf(", \"gather_%s_v%d_G_per_s\": %.1f", kind, scale, kN / ms / 1e6);
      struct Relax {
        const uint32_t* idx;
        uint32_t* value;
        uint64_t V;
        hipError_t operator()() const { return hipMemset(value, 0x7f, V * 4); }
        void timed() const { hipLaunchKernelGGL(k_relax, dim3(kGrid), dim3(kBlock), 0, 0, idx, kN, value); }
      } relax{idx, value, V};
      CK(best_of(3, relax, ms));
      printf(", \"relax_%s_v%d_G_per_s\": %.1f", kind, scale, kN / ms / 1e6);
      fflush(stdout);
    }
    CK(hipFree(value));
    CK(hipFree(rank_d));
  }
  printf(", \"indices_per_pass\": %llu, \"workgroups\": %d}\n", (unsigned long long)kN, kGrid);
  CK(hipFree(idx));
  CK(hipFree(out));
  CK(hipFree(a));
  CK(hipFree(b));
  return 0;
}
