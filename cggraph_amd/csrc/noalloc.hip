// cg_traverse_device_noalloc: cg_engine_run's GPU schedule (statistics,
// degree-tier compaction, push expansion; see traverse.hip) over a CSR and a
// value array that the caller holds in device memory, for callers that cannot
// allocate (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  Every buffer the engine would allocate -- the two
// frontier bitmaps, the three work queues, the statistics records and
// partials, the striped counters and the run state -- is carved out of one
// file-scope device array of kNoallocScratchBytes.  Nothing here calls
// hipMalloc, hipFree or hipHostMalloc.
//
// The array is one per GPU and lives until the process ends.  In host code its
// name is not a device address: the address each GPU gives it is taken once
// per device with hipGetSymbolAddress (with that device current) and kept in
// scratch_dev[]; only that pointer is handed to kernels.  Every call on one GPU
// uses the same array, so calls that overlap on one GPU race; each call clears
// what must start at zero with fill kernels (no hipMemsetAsync), and the array
// keeps the last call's contents afterwards.
#include <mutex>

#include "kernels.hpp"

namespace cg {

namespace {

constexpr size_t kNoallocScratchBytes = (size_t)CG_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

std::mutex scratch_mu;
void* scratch_dev[kMaxDevices] = {};

// The device address of `scratch` on the current device, taken on the first
// call on that device.
int scratch_address(void** out) {
  int dev = 0;
  CG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return set_error(CG_ERR_UNSUPPORTED, "device %d: at most %d devices", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return set_error(CG_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev, hipGetErrorString(e));
    }
    scratch_dev[dev] = p;
  }
  *out = scratch_dev[dev];
  return CG_OK;
}

// Carves 256-byte-aligned pieces out of the scratch array (base NULL: only
// counts the bytes).
struct Carver {
  unsigned char* base;
  size_t used = 0;
  template <class T>
  T* take(size_t count) {
    T* p = base ? reinterpret_cast<T*>(base + used) : nullptr;
    used += (count * sizeof(T) + 255) & ~(size_t)255;
    return p;
  }
};

struct Layout {
  uint64_t* bm[2];
  uint32_t* q_small;
  uint32_t* q_med;
  uint64_t* q_chunk;
  DevStats* st;
  StatsPart* parts;
  uint32_t* stripes;
  RunState* run_state;
  size_t bytes;
};

// Queue capacities as cg_engine_create sizes them: a chunked vertex has at
// least bin_chunk edges, so the chunk queue holds at most E / chunk_edges +
// E / bin_chunk items.
Layout layout(unsigned char* base, uint32_t V, uint64_t E, const Bins& bins) {
  const uint64_t words = ((uint64_t)V + 63) / 64 + 1;
  Carver c{base};
  Layout l{};
  l.bm[0] = c.take<uint64_t>(words);
  l.bm[1] = c.take<uint64_t>(words);
  l.q_small = c.take<uint32_t>(V ? V : 1);
  l.q_med = c.take<uint32_t>(V ? V : 1);
  l.q_chunk = c.take<uint64_t>(E / bins.chunk_edges + E / bins.bin_chunk + 1);
  l.st = c.take<DevStats>(2);
  l.parts = c.take<StatsPart>(stats_parts_needed(words));
  l.stripes = c.take<uint32_t>(kStripeWords);
  l.run_state = c.take<RunState>(1);
  l.bytes = c.used;
  return l;
}

}  // namespace

}  // namespace cg

using namespace cg;

extern "C" int cg_traverse_device_noalloc(int algorithm, const uint64_t* d_offsets, const uint32_t* d_dst,
                                          const uint32_t* d_w, uint32_t V, uint64_t E, uint32_t root,
                                          uint32_t* d_val, void* stream, uint32_t* iterations) {
  clear_error();
  if (iterations) *iterations = 0;
  if (algorithm != CG_BFS && algorithm != CG_SSSP && algorithm != CG_CC)
    return set_error(CG_ERR_UNSUPPORTED, "algorithm %d: the allocation-free run is BFS, SSSP or CC", algorithm);
  if (!d_offsets || !d_val || (E && !d_dst)) return set_error(CG_ERR_INVALID, "NULL argument");
  if (algorithm == CG_SSSP && E && !d_w) return set_error(CG_ERR_INVALID, "SSSP needs edge weights");
  if (algorithm != CG_CC && root >= V) return set_error(CG_ERR_INVALID, "root %u >= V = %u", root, V);
  if (V == 0) return CG_OK;
  const Bins bins{32, 4096, 2048, 0};  // the default tiers of cg_options_default
  {
    const Layout need = layout(nullptr, V, E, bins);
    if (need.bytes > kNoallocScratchBytes)
      return set_error(CG_ERR_INVALID,
                       "V = %u, E = %llu need %zu bytes of scratch; the allocation-free run has %zu "
                       "(about 8.3 bytes per vertex plus 3 bytes per 512 edges)",
                       V, (unsigned long long)E, need.bytes, kNoallocScratchBytes);
  }
  void* base = nullptr;
  CG_TRY(scratch_address(&base));
  const Layout l = layout(static_cast<unsigned char*>(base), V, E, bins);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint64_t words = ((uint64_t)V + 63) / 64;
  const DevGraph dg{d_offsets, d_dst, algorithm == CG_SSSP ? d_w : nullptr, nullptr, nullptr, V};
  // ---- initAlgorithm (CGgraph_engine.cuh:1259-1316), cleared with fill kernels ----
  CG_TRY(launch_fill_u64(l.bm[0], 0, words + 1, s));
  CG_TRY(launch_fill_u64(l.bm[1], 0, words + 1, s));
  CG_TRY(launch_fill_u32(l.stripes, 0, kStripeWords, s));
  CG_TRY(launch_fill_u32(reinterpret_cast<uint32_t*>(l.st), 0, 2 * sizeof(DevStats) / 4, s));
  if (algorithm == CG_CC) {
    CG_TRY(launch_iota_u32(d_val, V, s));
    CG_TRY(launch_bitmap_prefix(l.bm[0], words, V, s));
    CG_TRY(launch_init_run_state(l.run_state, V, s));
  } else {
    CG_TRY(launch_fill_u32(d_val, kUnreached, V, s));
    CG_TRY(launch_set_root(d_val, l.bm[0], root, s));
    CG_TRY(launch_init_run_state(l.run_state, 1, s));
  }
  // ---- the iteration loop: statistics, then compaction and expansion ----
  uint32_t it = 0;
  for (int cur = 0;; cur ^= 1) {
    DevStats* st = l.st + (it & 1);
    CG_TRY(launch_stats(l.bm[cur], l.bm[cur ^ 1], words, dg, E, l.parts, l.stripes, l.run_state, nullptr, nullptr,
                        0xFFFFFFFFu, st, nullptr, s));
    DevStats h;
    CG_HIP(hipMemcpyAsync(&h, st, sizeof(DevStats), hipMemcpyDeviceToHost, s));
    CG_HIP(hipStreamSynchronize(s));
    if (h.n_active == 0) break;
    ++it;
    CG_TRY(launch_compact(l.bm[cur], words, dg, bins, l.q_small, l.q_med, l.q_chunk, st, nullptr, 0xFFFFFFFFu,
                          l.bm[cur ^ 1], s));
    CG_TRY(launch_expand_push(algorithm, dg, d_val, l.bm[cur ^ 1], l.q_small, l.q_med, l.q_chunk,
                              expand_item_bound(h.n_active, h.m_active, bins), SmallRange{}, bins, st, l.stripes,
                              true, s));
  }
  if (iterations) *iterations = it;
  return CG_OK;
}
