// Shared definitions of libcgmi: status handling, device buffers and the
// graph object.  Every translation unit of the library includes this file.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/cgmi.h"

namespace cg {

// Per-thread last error (cg_last_error).
int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
const char* last_error();
void clear_error();

#define CG_HIP(call)                                                                      \
  do {                                                                                    \
    hipError_t cg_e_ = (call);                                                            \
    if (cg_e_ != hipSuccess)                                                              \
      return ::cg::set_error(CG_ERR_HIP, "%s failed: %s (%s:%d)", #call,                  \
                             hipGetErrorString(cg_e_), __FILE__, __LINE__);               \
  } while (0)

#define CG_TRY(expr)                 \
  do {                               \
    int cg_s_ = (expr);              \
    if (cg_s_ != CG_OK) return cg_s_; \
  } while (0)

constexpr uint32_t kUnreached = CG_UNREACHED;
constexpr int kWave = 64;  // gfx950 wavefront width

// Owning device allocation.  Freed with hipFree on destruction.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { reset(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  // (Re)allocates when the current allocation is smaller than `count`.
  int ensure(size_t count) {
    if (count <= n && p) return CG_OK;
    reset();
    size_t bytes = (count ? count : 1) * sizeof(T);
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), bytes);
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      return set_error(CG_ERR_OOM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    }
    n = count;
    return CG_OK;
  }
  size_t bytes() const { return n * sizeof(T); }
};

// Pinned host allocation (hipHostMalloc; never hipHostRegister).
template <class T>
struct PinnedBuf {
  T* p = nullptr;
  size_t n = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf() { reset(); }
  void reset() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
  // flags: hipHostMallocDefault, or hipHostMallocMapped | hipHostMallocCoherent
  // for memory a kernel writes directly (then see device_ptr()).
  int ensure(size_t count, unsigned flags = hipHostMallocDefault) {
    if (count <= n && p) return CG_OK;
    reset();
    size_t bytes = (count ? count : 1) * sizeof(T);
    hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&p), bytes, flags);
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      return set_error(CG_ERR_OOM, "hipHostMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    }
    n = count;
    return CG_OK;
  }
  // Device address of a mapped allocation.
  T* device_ptr() const {
    void* d = nullptr;
    if (!p || hipHostGetDevicePointer(&d, p, 0) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    return static_cast<T*>(d);
  }
};

// Returns CG_OK when a GPU is present and selects device 0.
int require_device();
bool device_available();

}  // namespace cg

// The graph object behind cg_graph*.  The host copy is filled lazily for
// graphs built on the GPU (the host team and cg_graph_export need it).
struct cg_graph {
  uint32_t V = 0;
  uint64_t E = 0;
  uint32_t zero_tail = 0;   // vertices [V - zero_tail, V) have out-degree 0
  bool symmetric = false;   // every edge u->v has v->u (set by generators / checked on load)
  bool has_w = false;
  uint32_t max_outdeg_vertex = 0;
  uint64_t max_outdeg = 0;
  // host CSR
  bool host_valid = false;
  std::vector<uint64_t> h_off;
  std::vector<uint32_t> h_dst, h_w;
  // device CSR (HBM)
  bool dev_valid = false;
  cg::DevBuf<uint64_t> d_off;
  cg::DevBuf<uint32_t> d_dst, d_w;
  // out-degrees u32, padded to whole 64-vertex words, built on demand
  bool dev_deg_valid = false;
  cg::DevBuf<uint32_t> d_deg;
  // the weights in 8 or 16 bits when all of them fit (at most one is built)
  bool dev_wn_valid = false;
  uint32_t max_w = 0;  // the largest weight (set with dev_wn_valid)
  cg::DevBuf<uint8_t> d_w8;
  cg::DevBuf<uint16_t> d_w16;
  // transposed graph (in-edges sorted by source), built on demand
  bool dev_t_valid = false;
  cg::DevBuf<uint64_t> d_in_off;
  cg::DevBuf<uint32_t> d_in_src, d_in_w;
  // batched queries: a copy renumbered by descending out-degree, built on
  // first use when the graph is large and skewed (vertex batch_new_of[v] of
  // the copy is the caller's v)
  bool batch_relabel_done = false;
  std::unique_ptr<cg_graph> batch_rg;
  cg::DevBuf<uint32_t> batch_new_of;
  std::vector<uint32_t> batch_new_of_h;
  // batched queries that run on an engine of their own (a large
  // small-world graph): the engines (BFS, SSSP), created on first use with
  // cg_engine_destroy as their deleter; declared last, so they go before
  // the buffers they refer to
  std::unique_ptr<cg_engine, void (*)(cg_engine*)> batch_eng[2] = {{nullptr, nullptr}, {nullptr, nullptr}};

  int ensure_host();    // D2H when only the device copy exists
  int ensure_device();  // H2D when only the host copy exists
};
