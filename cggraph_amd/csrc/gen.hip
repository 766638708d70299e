// Graph construction on the GPU: the synthetic workloads (R-MAT, lattice),
// edge-list -> CSR (sort, drop self-loops and duplicates), the transposed
// graph for pull steps, and the max-out-degree root.
//
// Keys are (src << bits) | dst with bits = ceil(log2 V), sorted with
// rocPRIM's radix sort over 2*bits bits only.  The generators are counter
// based: edge i of the R-MAT stream depends on (seed, i) alone, so the GPU and
// the CPU oracle produce identical edge lists.  The reference builds its CSR
// on the host from packed 12-byte edge records (Basic/Graph/graphBinReader.hpp:
// 89-326) and requires strictly ascending neighbours (my_project/CGgraph/
// sortNbr.hpp:84-91); the CSR built here has that property by construction.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "kernels.hpp"

namespace cg {

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;
constexpr int kBlock = 256;

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

uint32_t prob_threshold(double p) {
  double t = p * 4294967296.0;
  if (t <= 0.0) return 0;
  if (t >= 4294967295.0) return 0xFFFFFFFFu;
  return (uint32_t)t;
}

__device__ __forceinline__ uint32_t edge_weight(uint64_t wkey, uint32_t u, uint32_t v, uint32_t lo,
                                                uint64_t span, int symmetric) {
  if (symmetric && u > v) { uint32_t t = u; u = v; v = t; }
  const uint64_t h = mix64(wkey + ((((uint64_t)u) << 32) | v) * kGolden);
  return lo + (uint32_t)(((h >> 32) * span) >> 32);
}

__global__ void k_rmat_keys(uint64_t* __restrict__ keys, uint64_t m, uint32_t scale, uint32_t tA,
                            uint32_t tAB, uint32_t tABC, uint64_t key, int sym) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t s = 0, d = 0;
    uint64_t r = 0;
    for (uint32_t l = 0; l < scale; ++l) {
      if ((l & 1) == 0) r = mix64(key + (i * 16 + (l >> 1) + 1) * kGolden);
      const uint32_t x = (uint32_t)(r >> ((l & 1) * 32));
      const uint32_t q = x < tA ? 0u : (x < tAB ? 1u : (x < tABC ? 2u : 3u));
      s = (s << 1) | (q >> 1);
      d = (d << 1) | (q & 1);
    }
    if (sym) {
      keys[2 * i] = ((uint64_t)s << scale) | d;
      keys[2 * i + 1] = ((uint64_t)d << scale) | s;
    } else {
      keys[i] = ((uint64_t)s << scale) | d;
    }
  }
}

// Undirected lattice edge e (horizontal edges first); a dropped edge becomes
// the self-loop 0 -> 0, which the CSR build removes.
__global__ void k_grid_keys(uint64_t* __restrict__ keys, uint32_t n, uint64_t key, uint32_t tdrop,
                            uint32_t bits) {
  const uint64_t nh = (uint64_t)n * (n - 1);
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < 2 * nh;
       e += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t u, v;
    if (e < nh) {
      const uint32_t r = (uint32_t)(e / (n - 1)), c = (uint32_t)(e % (n - 1));
      u = r * n + c;
      v = u + 1;
    } else {
      const uint64_t f = e - nh;
      const uint32_t r = (uint32_t)(f / n), c = (uint32_t)(f % n);
      u = r * n + c;
      v = u + n;
    }
    const uint32_t x = (uint32_t)(mix64(key + (e + 1) * kGolden) >> 32);
    if (x < tdrop) { u = 0; v = 0; }
    keys[2 * e] = ((uint64_t)u << bits) | v;
    keys[2 * e + 1] = ((uint64_t)v << bits) | u;
  }
}

__global__ void k_mark(const uint64_t* __restrict__ keys, uint64_t n, uint32_t bits,
                       uint32_t* __restrict__ flags) {
  const uint64_t mask = (bits >= 32) ? 0xFFFFFFFFull : ((1ull << bits) - 1);
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t k = keys[i];
    const bool self = (k >> bits) == (k & mask);
    const bool dup = i > 0 && keys[i - 1] == k;
    flags[i] = (!self && !dup) ? 1u : 0u;
  }
}

__global__ void k_scatter(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ flags,
                          const uint64_t* __restrict__ pos, uint64_t n, uint32_t bits,
                          uint64_t* __restrict__ kept, uint32_t* __restrict__ dst) {
  const uint64_t mask = (bits >= 32) ? 0xFFFFFFFFull : ((1ull << bits) - 1);
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (flags[i]) {
      const uint64_t p = pos[i];
      kept[p] = keys[i];
      dst[p] = (uint32_t)(keys[i] & mask);
    }
  }
}

// off[v] = first position whose key has source >= v (lower bound), v in [0, V].
__global__ void k_offsets(const uint64_t* __restrict__ kept, uint64_t E, uint32_t V, uint32_t bits,
                          uint64_t* __restrict__ off) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v <= V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t target = v << bits;
    uint64_t lo = 0, hi = E;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (kept[mid] < target) lo = mid + 1;
      else hi = mid;
    }
    off[v] = lo;
  }
}

__global__ void k_weights(const uint64_t* __restrict__ kept, uint64_t E, uint32_t bits,
                          uint64_t wkey, uint32_t lo, uint64_t span, int symmetric,
                          uint32_t* __restrict__ w) {
  const uint64_t mask = (bits >= 32) ? 0xFFFFFFFFull : ((1ull << bits) - 1);
  for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < E;
       p += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t k = kept[p];
    w[p] = edge_weight(wkey, (uint32_t)(k >> bits), (uint32_t)(k & mask), lo, span, symmetric);
  }
}

// Source of every edge without a search: each non-empty vertex writes its id
// at its first edge, then an inclusive max-scan carries it across the row.
__global__ void k_row_heads(const uint64_t* __restrict__ off, uint32_t V, uint32_t* __restrict__ head) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x)
    if (off[v + 1] > off[v]) head[off[v]] = (uint32_t)v;
}

__global__ void k_transpose_keys(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
                                 uint64_t E, uint32_t bits, uint64_t* __restrict__ keys) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < E;
       e += (uint64_t)gridDim.x * blockDim.x)
    keys[e] = ((uint64_t)dst[e] << bits) | src[e];
}

__global__ void k_low_bits(const uint64_t* __restrict__ keys, uint64_t n, uint32_t bits,
                           uint32_t* __restrict__ out) {
  const uint64_t mask = (bits >= 32) ? 0xFFFFFFFFull : ((1ull << bits) - 1);
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = (uint32_t)(keys[i] & mask);
}

__global__ void k_max_deg(const uint64_t* __restrict__ off, uint32_t V,
                          unsigned long long* __restrict__ best) {
  unsigned long long b = 0;
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t deg = off[v + 1] - off[v];
    // larger degree wins; among equal degrees the smaller id wins
    const unsigned long long k = ((unsigned long long)deg << 32) | (0xFFFFFFFFull - v);
    if (k > b) b = k;
  }
  for (int d = 32; d >= 1; d >>= 1) {
    unsigned long long o = __shfl_xor(b, d, 64);
    if (o > b) b = o;
  }
  if (__lane_id() == 0 && b) atomicMax(best, b);
}

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

#define LAUNCH_OK()                                                                     \
  do {                                                                                  \
    hipError_t e_ = hipGetLastError();                                                  \
    if (e_ != hipSuccess)                                                               \
      return set_error(CG_ERR_HIP, "kernel launch failed: %s (%s:%d)",                  \
                       hipGetErrorString(e_), __FILE__, __LINE__);                      \
  } while (0)

// Sorts keys[0..n) (double buffer with alt), drops self-loops and
// duplicates, and fills the graph's device CSR.  On return keys/alt are free
// to reuse.
int csr_from_keys(DevBuf<uint64_t>& keys, DevBuf<uint64_t>& alt, uint64_t n, uint32_t V,
                  uint32_t bits, cg_graph* g, uint64_t wkey, uint32_t wlo, uint32_t whi,
                  int sym_weights, hipStream_t s) {
  rocprim::double_buffer<uint64_t> db(keys.p, alt.p);
  size_t tmp_bytes = 0;
  CG_HIP(rocprim::radix_sort_keys(nullptr, tmp_bytes, db, n, 0, 2 * bits, s));
  DevBuf<uint8_t> tmp;
  CG_TRY(tmp.ensure(tmp_bytes));
  CG_HIP(rocprim::radix_sort_keys(tmp.p, tmp_bytes, db, n, 0, 2 * bits, s));
  uint64_t* sorted = db.current();
  uint64_t* other = db.alternate();

  DevBuf<uint32_t> flags;
  DevBuf<uint64_t> pos;
  CG_TRY(flags.ensure(n));
  CG_TRY(pos.ensure(n));
  hipLaunchKernelGGL(k_mark, dim3(grid_of(n)), dim3(kBlock), 0, s, sorted, n, bits, flags.p);
  LAUNCH_OK();
  size_t scan_bytes = 0;
  CG_HIP(rocprim::exclusive_scan(nullptr, scan_bytes, flags.p, pos.p, (uint64_t)0, n,
                                 rocprim::plus<uint64_t>(), s));
  DevBuf<uint8_t> scan_tmp;
  CG_TRY(scan_tmp.ensure(scan_bytes));
  CG_HIP(rocprim::exclusive_scan(scan_tmp.p, scan_bytes, flags.p, pos.p, (uint64_t)0, n,
                                 rocprim::plus<uint64_t>(), s));
  uint64_t last_pos = 0;
  uint32_t last_flag = 0;
  if (n) {
    CG_HIP(hipMemcpyAsync(&last_pos, pos.p + n - 1, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    CG_HIP(hipMemcpyAsync(&last_flag, flags.p + n - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    CG_HIP(hipStreamSynchronize(s));
  }
  const uint64_t E = n ? last_pos + last_flag : 0;
  g->V = V;
  g->E = E;
  CG_TRY(g->d_off.ensure((uint64_t)V + 1));
  CG_TRY(g->d_dst.ensure(E));
  hipLaunchKernelGGL(k_scatter, dim3(grid_of(n)), dim3(kBlock), 0, s, sorted, flags.p, pos.p, n, bits,
                     other, g->d_dst.p);
  LAUNCH_OK();
  flags.reset();
  pos.reset();
  hipLaunchKernelGGL(k_offsets, dim3(grid_of((uint64_t)V + 1)), dim3(kBlock), 0, s, other, E, V, bits,
                     g->d_off.p);
  LAUNCH_OK();
  g->has_w = whi != 0;
  if (g->has_w) {
    CG_TRY(g->d_w.ensure(E));
    hipLaunchKernelGGL(k_weights, dim3(grid_of(E)), dim3(kBlock), 0, s, other, E, bits, wkey, wlo,
                       (uint64_t)(whi - wlo) + 1, sym_weights, g->d_w.p);
    LAUNCH_OK();
  }
  CG_HIP(hipStreamSynchronize(s));
  g->dev_valid = true;
  g->host_valid = false;
  return CG_OK;
}

}  // namespace

int gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed,
             uint32_t wlo, uint32_t whi, int symmetrize, cg_graph* g) {
  if (scale < 1 || scale > 31) return set_error(CG_ERR_INVALID, "R-MAT scale %u outside [1, 31]", scale);
  if (whi && wlo > whi) return set_error(CG_ERR_INVALID, "weight range [%u, %u] is empty", wlo, whi);
  if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
    return set_error(CG_ERR_INVALID, "R-MAT probabilities a=%g b=%g c=%g are not a distribution", a, b, c);
  CG_TRY(require_device());
  const uint64_t m = ((uint64_t)1 << scale) * edge_factor;
  const uint64_t n = symmetrize ? 2 * m : m;
  hipStream_t s = nullptr;
  DevBuf<uint64_t> keys, alt;
  CG_TRY(keys.ensure(n));
  CG_TRY(alt.ensure(n));
  const uint64_t key = mix64(seed * kGolden + 0x5851F42D4C957F2Dull);
  hipLaunchKernelGGL(k_rmat_keys, dim3(grid_of(m)), dim3(kBlock), 0, s, keys.p, m, scale,
                     prob_threshold(a), prob_threshold(a + b), prob_threshold(a + b + c), key,
                     symmetrize);
  LAUNCH_OK();
  const uint64_t wkey = mix64(seed * kGolden + 0x2545F4914F6CDD1Dull);
  CG_TRY(csr_from_keys(keys, alt, n, 1u << scale, scale, g, wkey, wlo, whi, 0, s));
  g->symmetric = symmetrize != 0;
  g->zero_tail = 0;
  return device_max_outdeg(g);
}

int gen_grid(uint32_t n, double drop, uint64_t seed, uint32_t wlo, uint32_t whi, cg_graph* g) {
  if (n < 2 || (uint64_t)n * n > 0xFFFFFFFFull)
    return set_error(CG_ERR_INVALID, "grid side %u outside [2, 65535]", n);
  if (whi && wlo > whi) return set_error(CG_ERR_INVALID, "weight range [%u, %u] is empty", wlo, whi);
  CG_TRY(require_device());
  const uint64_t V = (uint64_t)n * n;
  const uint32_t bits = bits_for(V);
  const uint64_t keys_n = 4ull * n * (n - 1);
  hipStream_t s = nullptr;
  DevBuf<uint64_t> keys, alt;
  CG_TRY(keys.ensure(keys_n));
  CG_TRY(alt.ensure(keys_n));
  const uint64_t key = mix64(seed * kGolden + 0x94D049BB133111EBull);
  hipLaunchKernelGGL(k_grid_keys, dim3(grid_of(keys_n / 2)), dim3(kBlock), 0, s, keys.p, n, key,
                     prob_threshold(drop), bits);
  LAUNCH_OK();
  const uint64_t wkey = mix64(seed * kGolden + 0x2545F4914F6CDD1Dull);
  CG_TRY(csr_from_keys(keys, alt, keys_n, (uint32_t)V, bits, g, wkey, wlo, whi, 1, s));
  g->symmetric = true;
  g->zero_tail = 0;
  return device_max_outdeg(g);
}

namespace {

__global__ void k_edge_keys(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
                            const uint32_t* __restrict__ w, uint64_t m, uint32_t bits, int sym,
                            uint64_t* __restrict__ keys, uint32_t* __restrict__ wk) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t s = src[i], d = dst[i];
    const uint32_t wi = w ? w[i] : 0u;
    if (sym) {
      keys[2 * i] = (s << bits) | d;
      keys[2 * i + 1] = (d << bits) | s;
      if (wk) { wk[2 * i] = wi; wk[2 * i + 1] = wi; }
    } else {
      keys[i] = (s << bits) | d;
      if (wk) wk[i] = wi;
    }
  }
}

// For every first occurrence of a key: the smallest weight among its copies
// (shortest paths only ever use the lightest of parallel edges).
__global__ void k_run_min(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ wk,
                          const uint32_t* __restrict__ flags, const uint64_t* __restrict__ pos, uint64_t n,
                          uint32_t* __restrict__ w_out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (!flags[i]) continue;
    uint32_t m = wk[i];
    for (uint64_t j = i + 1; j < n && keys[j] == keys[i]; ++j) m = wk[j] < m ? wk[j] : m;
    w_out[pos[i]] = m;
  }
}

}  // namespace

int build_from_edges(const uint32_t* h_src, const uint32_t* h_dst, const uint32_t* h_w, uint64_t m, uint32_t V,
                     int symmetrize, cg_graph* g) {
  if (V == 0) return set_error(CG_ERR_INVALID, "V must be positive");
  CG_TRY(require_device());
  for (uint64_t i = 0; i < m; ++i)
    if (h_src[i] >= V || h_dst[i] >= V)
      return set_error(CG_ERR_INVALID, "edge %llu (%u -> %u) has an endpoint >= V", (unsigned long long)i, h_src[i],
                       h_dst[i]);
  if (h_w)
    for (uint64_t i = 0; i < m; ++i)
      if (h_w[i] > 0xFFFFFFFFu - kUnreached)
        return set_error(CG_ERR_INVALID, "weight %u of edge %llu would let value + weight wrap", h_w[i],
                         (unsigned long long)i);
  hipStream_t s = nullptr;
  const uint32_t bits = bits_for(V);
  const uint64_t n = symmetrize ? 2 * m : m;
  DevBuf<uint32_t> ds, dd, dw;
  CG_TRY(ds.ensure(m));
  CG_TRY(dd.ensure(m));
  if (m) {
    CG_HIP(hipMemcpy(ds.p, h_src, m * 4, hipMemcpyHostToDevice));
    CG_HIP(hipMemcpy(dd.p, h_dst, m * 4, hipMemcpyHostToDevice));
  }
  if (h_w) {
    CG_TRY(dw.ensure(m));
    if (m) CG_HIP(hipMemcpy(dw.p, h_w, m * 4, hipMemcpyHostToDevice));
  }
  DevBuf<uint64_t> keys, kalt;
  DevBuf<uint32_t> wk, walt;
  CG_TRY(keys.ensure(n));
  CG_TRY(kalt.ensure(n));
  if (h_w) {
    CG_TRY(wk.ensure(n));
    CG_TRY(walt.ensure(n));
  }
  hipLaunchKernelGGL(k_edge_keys, dim3(grid_of(m)), dim3(kBlock), 0, s, ds.p, dd.p, h_w ? dw.p : nullptr, m, bits,
                     symmetrize, keys.p, h_w ? wk.p : nullptr);
  LAUNCH_OK();
  ds.reset();
  dd.reset();
  dw.reset();
  rocprim::double_buffer<uint64_t> kb(keys.p, kalt.p);
  size_t tb = 0;
  DevBuf<uint8_t> tmp;
  if (h_w) {
    rocprim::double_buffer<uint32_t> vb(wk.p, walt.p);
    CG_HIP(rocprim::radix_sort_pairs(nullptr, tb, kb, vb, n, 0, 2 * bits, s));
    CG_TRY(tmp.ensure(tb));
    CG_HIP(rocprim::radix_sort_pairs(tmp.p, tb, kb, vb, n, 0, 2 * bits, s));
    if (vb.current() != wk.p) std::swap(wk, walt);
  } else {
    CG_HIP(rocprim::radix_sort_keys(nullptr, tb, kb, n, 0, 2 * bits, s));
    CG_TRY(tmp.ensure(tb));
    CG_HIP(rocprim::radix_sort_keys(tmp.p, tb, kb, n, 0, 2 * bits, s));
  }
  if (kb.current() != keys.p) std::swap(keys, kalt);
  // from here: keys.p sorted, wk.p weights in the same order, kalt.p free
  DevBuf<uint32_t> flags;
  DevBuf<uint64_t> pos;
  CG_TRY(flags.ensure(n));
  CG_TRY(pos.ensure(n));
  hipLaunchKernelGGL(k_mark, dim3(grid_of(n)), dim3(kBlock), 0, s, keys.p, n, bits, flags.p);
  LAUNCH_OK();
  size_t sb = 0;
  CG_HIP(rocprim::exclusive_scan(nullptr, sb, flags.p, pos.p, (uint64_t)0, n, rocprim::plus<uint64_t>(), s));
  DevBuf<uint8_t> st;
  CG_TRY(st.ensure(sb));
  CG_HIP(rocprim::exclusive_scan(st.p, sb, flags.p, pos.p, (uint64_t)0, n, rocprim::plus<uint64_t>(), s));
  uint64_t last_pos = 0;
  uint32_t last_flag = 0;
  if (n) {
    CG_HIP(hipMemcpyAsync(&last_pos, pos.p + n - 1, 8, hipMemcpyDeviceToHost, s));
    CG_HIP(hipMemcpyAsync(&last_flag, flags.p + n - 1, 4, hipMemcpyDeviceToHost, s));
    CG_HIP(hipStreamSynchronize(s));
  }
  const uint64_t E = n ? last_pos + last_flag : 0;
  g->V = V;
  g->E = E;
  CG_TRY(g->d_off.ensure((uint64_t)V + 1));
  CG_TRY(g->d_dst.ensure(E));
  hipLaunchKernelGGL(k_scatter, dim3(grid_of(n)), dim3(kBlock), 0, s, keys.p, flags.p, pos.p, n, bits, kalt.p,
                     g->d_dst.p);
  LAUNCH_OK();
  g->has_w = h_w != nullptr;
  if (g->has_w) {
    CG_TRY(g->d_w.ensure(E));
    hipLaunchKernelGGL(k_run_min, dim3(grid_of(n)), dim3(kBlock), 0, s, keys.p, wk.p, flags.p, pos.p, n, g->d_w.p);
    LAUNCH_OK();
  }
  hipLaunchKernelGGL(k_offsets, dim3(grid_of((uint64_t)V + 1)), dim3(kBlock), 0, s, kalt.p, E, V, bits, g->d_off.p);
  LAUNCH_OK();
  CG_HIP(hipStreamSynchronize(s));
  g->dev_valid = true;
  g->host_valid = false;
  g->symmetric = symmetrize != 0;
  g->zero_tail = 0;
  return device_max_outdeg(g);
}

int build_transpose(cg_graph* g) {
  if (g->dev_t_valid) return CG_OK;
  CG_TRY(g->ensure_device());
  hipStream_t s = nullptr;
  const uint64_t E = g->E;
  const uint32_t V = g->V;
  const uint32_t bits = bits_for(V);
  CG_TRY(g->d_in_off.ensure((uint64_t)V + 1));
  CG_TRY(g->d_in_src.ensure(E));
  if (E == 0) {
    CG_TRY(launch_fill_u64(g->d_in_off.p, 0, (uint64_t)V + 1, s));
    CG_HIP(hipStreamSynchronize(s));
    g->dev_t_valid = true;
    return CG_OK;
  }
  DevBuf<uint64_t> keys, alt;
  CG_TRY(keys.ensure(E));
  CG_TRY(alt.ensure(E));
  {
    DevBuf<uint32_t> head, src;
    CG_TRY(head.ensure(E));
    CG_TRY(src.ensure(E));
    CG_HIP(hipMemsetAsync(head.p, 0, E * sizeof(uint32_t), s));
    hipLaunchKernelGGL(k_row_heads, dim3(grid_of(V)), dim3(kBlock), 0, s, g->d_off.p, V, head.p);
    LAUNCH_OK();
    size_t sb = 0;
    CG_HIP(rocprim::inclusive_scan(nullptr, sb, head.p, src.p, E, rocprim::maximum<uint32_t>(), s));
    DevBuf<uint8_t> st;
    CG_TRY(st.ensure(sb));
    CG_HIP(rocprim::inclusive_scan(st.p, sb, head.p, src.p, E, rocprim::maximum<uint32_t>(), s));
    hipLaunchKernelGGL(k_transpose_keys, dim3(grid_of(E)), dim3(kBlock), 0, s, src.p, g->d_dst.p, E,
                       bits, keys.p);
    LAUNCH_OK();
    CG_HIP(hipStreamSynchronize(s));
  }
  rocprim::double_buffer<uint64_t> db(keys.p, alt.p);
  size_t tmp_bytes = 0;
  DevBuf<uint8_t> tmp;
  if (g->has_w) {
    DevBuf<uint32_t> walt;
    CG_TRY(g->d_in_w.ensure(E));
    CG_TRY(walt.ensure(E));
    CG_HIP(hipMemcpyAsync(g->d_in_w.p, g->d_w.p, E * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    rocprim::double_buffer<uint32_t> vb(g->d_in_w.p, walt.p);
    CG_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, db, vb, E, 0, 2 * bits, s));
    CG_TRY(tmp.ensure(tmp_bytes));
    CG_HIP(rocprim::radix_sort_pairs(tmp.p, tmp_bytes, db, vb, E, 0, 2 * bits, s));
    if (vb.current() != g->d_in_w.p)
      CG_HIP(hipMemcpyAsync(g->d_in_w.p, vb.current(), E * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    CG_HIP(hipStreamSynchronize(s));
  } else {
    CG_HIP(rocprim::radix_sort_keys(nullptr, tmp_bytes, db, E, 0, 2 * bits, s));
    CG_TRY(tmp.ensure(tmp_bytes));
    CG_HIP(rocprim::radix_sort_keys(tmp.p, tmp_bytes, db, E, 0, 2 * bits, s));
  }
  uint64_t* sorted = db.current();
  hipLaunchKernelGGL(k_offsets, dim3(grid_of((uint64_t)V + 1)), dim3(kBlock), 0, s, sorted, E, V, bits,
                     g->d_in_off.p);
  LAUNCH_OK();
  hipLaunchKernelGGL(k_low_bits, dim3(grid_of(E)), dim3(kBlock), 0, s, sorted, E, bits, g->d_in_src.p);
  LAUNCH_OK();
  CG_HIP(hipStreamSynchronize(s));
  g->dev_t_valid = true;
  return CG_OK;
}

int device_max_outdeg(cg_graph* g) {
  DevBuf<unsigned long long> best;
  CG_TRY(best.ensure(1));
  CG_HIP(hipMemset(best.p, 0, sizeof(unsigned long long)));
  if (g->V) {
    hipLaunchKernelGGL(k_max_deg, dim3(grid_of(g->V)), dim3(kBlock), 0, 0, g->d_off.p, g->V, best.p);
    LAUNCH_OK();
  }
  unsigned long long b = 0;
  CG_HIP(hipMemcpy(&b, best.p, sizeof(b), hipMemcpyDeviceToHost));
  g->max_outdeg = b >> 32;
  g->max_outdeg_vertex = b ? (uint32_t)(0xFFFFFFFFull - (b & 0xFFFFFFFFull)) : 0;
  return CG_OK;
}

// ---- renumbering by descending out-degree ----

namespace {

// Sort key for descending out-degree; the radix sort is stable, so equal
// degrees keep ascending ids.
__global__ void k_degree_order_keys(const uint64_t* __restrict__ off, uint32_t V, uint32_t* __restrict__ key,
                                    uint32_t* __restrict__ id) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    key[v] = ~(uint32_t)(off[v + 1] - off[v]);
    id[v] = (uint32_t)v;
  }
}

__global__ void k_new_ids(const uint32_t* __restrict__ perm, uint32_t V, uint32_t* __restrict__ new_of) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < V;
       i += (uint64_t)gridDim.x * blockDim.x)
    new_of[perm[i]] = (uint32_t)i;
}

__global__ void k_relabel_keys(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst, uint64_t E,
                               const uint32_t* __restrict__ new_of, uint32_t bits, uint64_t* __restrict__ keys) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < E;
       e += (uint64_t)gridDim.x * blockDim.x)
    keys[e] = ((uint64_t)new_of[src[e]] << bits) | new_of[dst[e]];
}

// out[v] = in[new_of[v]]: coalesced writes, gathered reads
__global__ void k_gather_u32(const uint32_t* __restrict__ new_of, const uint32_t* __restrict__ in, uint32_t V,
                             uint32_t* __restrict__ out) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x)
    out[v] = in[new_of[v]];
}

}  // namespace

int edge_sources(const uint64_t* off, uint32_t V, uint64_t E, DevBuf<uint32_t>& src, hipStream_t s) {
  DevBuf<uint32_t> head;
  CG_TRY(head.ensure(E));
  CG_TRY(src.ensure(E));
  if (!E) return CG_OK;
  CG_HIP(hipMemsetAsync(head.p, 0, E * sizeof(uint32_t), s));
  hipLaunchKernelGGL(k_row_heads, dim3(grid_of(V)), dim3(kBlock), 0, s, off, V, head.p);
  LAUNCH_OK();
  size_t sb = 0;
  CG_HIP(rocprim::inclusive_scan(nullptr, sb, head.p, src.p, E, rocprim::maximum<uint32_t>(), s));
  DevBuf<uint8_t> st;
  CG_TRY(st.ensure(sb));
  CG_HIP(rocprim::inclusive_scan(st.p, sb, head.p, src.p, E, rocprim::maximum<uint32_t>(), s));
  CG_HIP(hipStreamSynchronize(s));
  return CG_OK;
}

int degree_order(cg_graph* g, DevBuf<uint32_t>& perm, DevBuf<uint32_t>& new_of, double* top_share) {
  CG_TRY(g->ensure_device());
  hipStream_t s = nullptr;
  const uint32_t V = g->V;
  CG_TRY(perm.ensure(V));
  CG_TRY(new_of.ensure(V));
  DevBuf<uint32_t> key, id, key_out;
  CG_TRY(key.ensure(V));
  CG_TRY(id.ensure(V));
  CG_TRY(key_out.ensure(V));
  hipLaunchKernelGGL(k_degree_order_keys, dim3(grid_of(V)), dim3(kBlock), 0, s, g->d_off.p, V, key.p, id.p);
  LAUNCH_OK();
  size_t tb = 0;
  CG_HIP(rocprim::radix_sort_pairs(nullptr, tb, key.p, key_out.p, id.p, perm.p, (size_t)V, 0, 32, s));
  DevBuf<uint8_t> tmp;
  CG_TRY(tmp.ensure(tb));
  CG_HIP(rocprim::radix_sort_pairs(tmp.p, tb, key.p, key_out.p, id.p, perm.p, (size_t)V, 0, 32, s));
  hipLaunchKernelGGL(k_new_ids, dim3(grid_of(V)), dim3(kBlock), 0, s, perm.p, V, new_of.p);
  LAUNCH_OK();
  if (top_share) {
    // share of the edges leaving the top 1/64 of vertices by out-degree
    const uint32_t k = V / 64 ? V / 64 : 1;
    std::vector<uint32_t> hk(k);
    CG_HIP(hipMemcpyAsync(hk.data(), key_out.p, (size_t)k * 4, hipMemcpyDeviceToHost, s));
    CG_HIP(hipStreamSynchronize(s));
    uint64_t sum = 0;
    for (uint32_t x : hk) sum += ~x;
    *top_share = g->E ? (double)sum / (double)g->E : 0.0;
  }
  CG_HIP(hipStreamSynchronize(s));
  return CG_OK;
}

int relabel_graph(cg_graph* g, const uint32_t* new_of, cg_graph* rg) {
  CG_TRY(g->ensure_device());
  hipStream_t s = nullptr;
  const uint32_t V = g->V;
  const uint64_t E = g->E;
  const uint32_t bits = bits_for(V);
  rg->V = V;
  rg->E = E;
  rg->has_w = g->has_w;
  rg->symmetric = g->symmetric;
  rg->zero_tail = 0;
  CG_TRY(rg->d_off.ensure((uint64_t)V + 1));
  CG_TRY(rg->d_dst.ensure(E));
  if (E == 0) {
    CG_TRY(launch_fill_u64(rg->d_off.p, 0, (uint64_t)V + 1, s));
  } else {
    DevBuf<uint64_t> keys, alt;
    CG_TRY(keys.ensure(E));
    CG_TRY(alt.ensure(E));
    {
      DevBuf<uint32_t> src;
      CG_TRY(edge_sources(g->d_off.p, V, E, src, s));
      hipLaunchKernelGGL(k_relabel_keys, dim3(grid_of(E)), dim3(kBlock), 0, s, src.p, g->d_dst.p, E, new_of, bits,
                         keys.p);
      LAUNCH_OK();
      CG_HIP(hipStreamSynchronize(s));
    }
    rocprim::double_buffer<uint64_t> db(keys.p, alt.p);
    size_t tb = 0;
    DevBuf<uint8_t> tmp;
    if (g->has_w) {
      DevBuf<uint32_t> walt;
      CG_TRY(rg->d_w.ensure(E));
      CG_TRY(walt.ensure(E));
      CG_HIP(hipMemcpyAsync(rg->d_w.p, g->d_w.p, E * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
      rocprim::double_buffer<uint32_t> vb(rg->d_w.p, walt.p);
      CG_HIP(rocprim::radix_sort_pairs(nullptr, tb, db, vb, E, 0, 2 * bits, s));
      CG_TRY(tmp.ensure(tb));
      CG_HIP(rocprim::radix_sort_pairs(tmp.p, tb, db, vb, E, 0, 2 * bits, s));
      if (vb.current() != rg->d_w.p)
        CG_HIP(hipMemcpyAsync(rg->d_w.p, vb.current(), E * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
      CG_HIP(hipStreamSynchronize(s));
    } else {
      CG_HIP(rocprim::radix_sort_keys(nullptr, tb, db, E, 0, 2 * bits, s));
      CG_TRY(tmp.ensure(tb));
      CG_HIP(rocprim::radix_sort_keys(tmp.p, tb, db, E, 0, 2 * bits, s));
    }
    uint64_t* sorted = db.current();
    hipLaunchKernelGGL(k_offsets, dim3(grid_of((uint64_t)V + 1)), dim3(kBlock), 0, s, sorted, E, V, bits,
                       rg->d_off.p);
    LAUNCH_OK();
    hipLaunchKernelGGL(k_low_bits, dim3(grid_of(E)), dim3(kBlock), 0, s, sorted, E, bits, rg->d_dst.p);
    LAUNCH_OK();
  }
  CG_HIP(hipStreamSynchronize(s));
  rg->dev_valid = true;
  rg->host_valid = false;
  rg->dev_t_valid = false;
  return device_max_outdeg(rg);
}

int launch_renumber_back(const uint32_t* new_of, const uint32_t* in, uint32_t V, uint32_t* out, hipStream_t s) {
  if (!V) return CG_OK;
  hipLaunchKernelGGL(k_gather_u32, dim3(grid_of(V)), dim3(kBlock), 0, s, new_of, in, V, out);
  LAUNCH_OK();
  return CG_OK;
}

}  // namespace cg
