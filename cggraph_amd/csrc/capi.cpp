// extern "C" entry points of libcgmi (declared in include/cgmi.h), graph
// objects, file formats and the segment partition.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <memory>
#include <new>
#include <string>

#include "engine.hpp"

namespace cg {

namespace {
thread_local std::string g_err;
}

int set_error(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

const char* last_error() { return g_err.c_str(); }
void clear_error() { g_err.clear(); }

bool device_available() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return n > 0;
}

int require_device() {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return set_error(CG_ERR_NO_DEVICE, "no HIP device is available (hipGetDeviceCount: %s, %d devices)",
                     hipGetErrorString(e), n);
  }
  CG_HIP(hipSetDevice(0));
  return CG_OK;
}

}  // namespace cg

using namespace cg;

int cg_graph::ensure_host() {
  if (host_valid) return CG_OK;
  if (!dev_valid) return set_error(CG_ERR_INTERNAL, "graph has neither a host nor a device copy");
  try {
    h_off.resize((size_t)V + 1);
    h_dst.resize(E);
    if (has_w) h_w.resize(E);
  } catch (const std::bad_alloc&) {
    return set_error(CG_ERR_OOM, "host copy of the graph (%llu edges)", (unsigned long long)E);
  }
  CG_HIP(hipMemcpy(h_off.data(), d_off.p, ((size_t)V + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost));
  if (E) CG_HIP(hipMemcpy(h_dst.data(), d_dst.p, E * sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (has_w && E) CG_HIP(hipMemcpy(h_w.data(), d_w.p, E * sizeof(uint32_t), hipMemcpyDeviceToHost));
  host_valid = true;
  return CG_OK;
}

int cg_graph::ensure_device() {
  if (dev_valid) return CG_OK;
  if (!host_valid) return set_error(CG_ERR_INTERNAL, "graph has neither a host nor a device copy");
  CG_TRY(require_device());
  CG_TRY(d_off.ensure((size_t)V + 1));
  CG_TRY(d_dst.ensure(E));
  CG_HIP(hipMemcpy(d_off.p, h_off.data(), ((size_t)V + 1) * sizeof(uint64_t), hipMemcpyHostToDevice));
  if (E) CG_HIP(hipMemcpy(d_dst.p, h_dst.data(), E * sizeof(uint32_t), hipMemcpyHostToDevice));
  if (has_w) {
    CG_TRY(d_w.ensure(E));
    if (E) CG_HIP(hipMemcpy(d_w.p, h_w.data(), E * sizeof(uint32_t), hipMemcpyHostToDevice));
  }
  dev_valid = true;
  return CG_OK;
}

namespace {

// Host-side facts of a CSR: max-degree vertex, sortedness, symmetry.
void host_summary(cg_graph* g) {
  uint64_t best = 0;
  uint32_t bv = 0;
  for (uint32_t v = 0; v < g->V; ++v) {
    const uint64_t d = g->h_off[v + 1] - g->h_off[v];
    if (d > best) { best = d; bv = v; }
  }
  g->max_outdeg = best;
  g->max_outdeg_vertex = bv;
  // symmetric: every row sorted and every u->v has v->u (binary search)
  bool sym = true;
  for (uint32_t u = 0; u < g->V && sym; ++u)
    for (uint64_t e = g->h_off[u]; e + 1 < g->h_off[u + 1]; ++e)
      if (g->h_dst[e] >= g->h_dst[e + 1]) { sym = false; break; }
  for (uint32_t u = 0; u < g->V && sym; ++u)
    for (uint64_t e = g->h_off[u]; e < g->h_off[u + 1]; ++e) {
      const uint32_t v = g->h_dst[e];
      const uint32_t* b = g->h_dst.data() + g->h_off[v];
      const uint32_t* en = g->h_dst.data() + g->h_off[v + 1];
      if (!std::binary_search(b, en, u)) { sym = false; break; }
    }
  g->symmetric = sym;
}

int read_file(const char* path, void* dst, uint64_t bytes) {
  int fd = ::open(path, O_RDONLY);
  if (fd < 0) return set_error(CG_ERR_IO, "cannot open %s", path);
  struct stat stt;
  if (fstat(fd, &stt) != 0 || (uint64_t)stt.st_size != bytes) {
    ::close(fd);
    return set_error(CG_ERR_IO, "%s holds %lld bytes, expected %llu (load_binFile size check)", path,
                     (long long)stt.st_size, (unsigned long long)bytes);
  }
  uint64_t done = 0;
  char* p = static_cast<char*>(dst);
  while (done < bytes) {
    ssize_t r = ::read(fd, p + done, std::min<uint64_t>(bytes - done, 1ull << 30));
    if (r <= 0) {
      ::close(fd);
      return set_error(CG_ERR_IO, "read error on %s", path);
    }
    done += (uint64_t)r;
  }
  ::close(fd);
  return CG_OK;
}

int validate_and_adopt(cg_graph* g) {
  if (g->h_off[0] != 0) return set_error(CG_ERR_INVALID, "offsets[0] must be 0");
  for (uint32_t v = 0; v < g->V; ++v)
    if (g->h_off[v + 1] < g->h_off[v]) return set_error(CG_ERR_INVALID, "offsets decrease at vertex %u", v);
  if (g->h_off[g->V] != g->E)
    return set_error(CG_ERR_INVALID, "offsets[V] = %llu but E = %llu", (unsigned long long)g->h_off[g->V],
                     (unsigned long long)g->E);
  for (uint64_t e = 0; e < g->E; ++e)
    if (g->h_dst[e] >= g->V) return set_error(CG_ERR_INVALID, "destination %u at edge %llu is >= V", g->h_dst[e],
                                              (unsigned long long)e);
  if (g->has_w)
    for (uint64_t e = 0; e < g->E; ++e)
      if (g->h_w[e] > 0xFFFFFFFFu - CG_UNREACHED)
        return set_error(CG_ERR_INVALID, "weight %u at edge %llu would let value + weight wrap", g->h_w[e],
                         (unsigned long long)e);
  if (g->zero_tail > g->V) return set_error(CG_ERR_INVALID, "zero_outdeg_tail > V");
  for (uint32_t v = g->V - g->zero_tail; v < g->V; ++v)
    if (g->h_off[v + 1] != g->h_off[v])
      return set_error(CG_ERR_INVALID, "vertex %u lies in the declared zero-out-degree tail but has edges", v);
  g->host_valid = true;
  host_summary(g);
  return CG_OK;
}

}  // namespace

namespace cg {

int64_t device_memory_bytes() {
  if (!device_available()) return 0;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return (int64_t)tot;
}

// TwoLevelPartition (my_project/CGgraph/Two-level_partition.cuh:213-476) on
// one GPU: segments of 1024 vertices over V', sorted by average out-degree
// (ties by id) and cut into buckets at 17 / 33 / 257 by lower bound; the GPU
// takes segments in that order while the budget lasts, charging the fixed
// arrays (values, offsets, 1 GiB of working space) first and then 8 bytes per
// edge for SSSP (destination + weight), 4 otherwise.  resident (may be NULL)
// receives one flag per segment.
int plan_partition(const cg_graph* g, int algorithm, int64_t budget, cg_partition_info* info,
                   std::vector<uint8_t>* resident) {
  const uint64_t seg = 1024;
  const uint64_t Vp = (uint64_t)g->V - g->zero_tail;
  const uint64_t nseg = (Vp + seg - 1) / seg;
  struct S { double deg; uint32_t id; uint64_t edges; };
  std::vector<S> segs(nseg);
  for (uint64_t s = 0; s < nseg; ++s) {
    const uint64_t a = s * seg, b = std::min<uint64_t>(Vp, a + seg);
    const uint64_t edges = g->h_off[b] - g->h_off[a];
    segs[s] = S{(double)edges / (double)(b - a), (uint32_t)s, edges};
  }
  std::sort(segs.begin(), segs.end(), [](const S& x, const S& y) {
    return x.deg < y.deg || (x.deg == y.deg && x.id < y.id);
  });
  auto lb = [&](double t) {
    return (uint64_t)(std::lower_bound(segs.begin(), segs.end(), t,
                                       [](const S& x, double v) { return x.deg < v; }) - segs.begin());
  };
  info->segment_num = nseg;
  info->bucket_start[0] = 0;
  info->bucket_start[1] = lb(17.0);
  info->bucket_start[2] = lb(33.0);
  info->bucket_start[3] = lb(257.0);
  info->bucket_start[4] = nseg;
  const uint64_t fixed = 4ull * g->V + 8ull * (Vp + 1) + (1ull << 30);
  uint64_t used = fixed;
  bool all = budget > 0 && (int64_t)fixed <= budget;
  if (resident) resident->assign(nseg, 0);
  for (const S& s : segs) {
    const uint64_t b = s.edges * (algorithm == CG_SSSP ? 8 : 4);
    if (!all || (int64_t)(used + b) > budget) {
      all = false;
      break;
    }
    used += b;
    if (resident) (*resident)[s.id] = 1;
  }
  info->can_hold_all = all ? 1 : 0;
  info->device_bytes = used;
  return CG_OK;
}

namespace {

// The device view of a partly resident graph: the same vertices, the
// offsets and edges of the resident segments only (a non-resident vertex has
// no edges on the device), the bitmap of the non-resident vertices that have
// edges, and the full offsets for degree-based counts.
int build_resident_view(cg_engine* e, const std::vector<uint8_t>& resident) {
  const cg_graph& g = *e->caller_g;
  const uint64_t V = g.V, seg = 1024;
  std::vector<uint64_t> roff((size_t)V + 1, 0);
  std::vector<uint64_t> nonres(((size_t)V + 63) / 64 + 1, 0);
  uint64_t er = 0, n_nonres = 0;
  for (uint64_t v = 0; v < V; ++v) {
    roff[v] = er;
    const uint64_t deg = g.h_off[v + 1] - g.h_off[v];
    const uint64_t s = v / seg;
    const bool res = s >= resident.size() || resident[s];  // the zero-out-degree tail has no edges
    if (res) er += deg;
    else if (deg) {
      nonres[v >> 6] |= 1ull << (v & 63);
      ++n_nonres;
    }
  }
  roff[V] = er;
  std::vector<uint32_t> rdst, rw;
  try {
    rdst.resize(er);
    if (g.has_w) rw.resize(er);
  } catch (const std::bad_alloc&) {
    return set_error(CG_ERR_OOM, "resident edges");
  }
  for (uint64_t v = 0; v < V; ++v) {
    const uint64_t deg = roff[v + 1] - roff[v];
    if (!deg) continue;
    std::memcpy(rdst.data() + roff[v], g.h_dst.data() + g.h_off[v], deg * 4);
    if (g.has_w) std::memcpy(rw.data() + roff[v], g.h_w.data() + g.h_off[v], deg * 4);
  }
  std::unique_ptr<cg_graph> rg(new (std::nothrow) cg_graph());
  if (!rg) return set_error(CG_ERR_OOM, "resident graph object");
  rg->V = g.V;
  rg->E = er;
  rg->zero_tail = g.zero_tail;
  rg->has_w = g.has_w;
  rg->symmetric = false;
  rg->max_outdeg_vertex = g.max_outdeg_vertex;
  rg->max_outdeg = g.max_outdeg;
  CG_TRY(rg->d_off.ensure(V + 1));
  CG_TRY(rg->d_dst.ensure(er ? er : 1));
  if (g.has_w) CG_TRY(rg->d_w.ensure(er ? er : 1));
  CG_HIP(hipMemcpy(rg->d_off.p, roff.data(), (V + 1) * 8, hipMemcpyHostToDevice));
  if (er) CG_HIP(hipMemcpy(rg->d_dst.p, rdst.data(), er * 4, hipMemcpyHostToDevice));
  if (er && g.has_w) CG_HIP(hipMemcpy(rg->d_w.p, rw.data(), er * 4, hipMemcpyHostToDevice));
  rg->dev_valid = true;
  CG_TRY(e->nonres_bm.ensure(nonres.size()));
  CG_HIP(hipMemcpy(e->nonres_bm.p, nonres.data(), nonres.size() * 8, hipMemcpyHostToDevice));
  CG_TRY(e->full_off.ensure(V + 1));
  CG_HIP(hipMemcpy(e->full_off.p, g.h_off.data(), (V + 1) * 8, hipMemcpyHostToDevice));
  e->n_nonres = n_nonres;
  CG_TRY(e->d_nr_ids.ensure(n_nonres ? n_nonres : 1));
  CG_TRY(e->d_nr_vals.ensure(n_nonres ? n_nonres : 1));
  CG_TRY(e->d_nr_cnt.ensure(1));
  CG_TRY(e->p_nr_cnt.ensure(1));
  e->own_g = std::move(rg);
  e->g = e->own_g.get();
  e->partial = true;
  return CG_OK;
}

}  // namespace

}  // namespace cg

extern "C" {

int cg_options_default(cg_options* o) {
  if (!o) return set_error(CG_ERR_INVALID, "options pointer is NULL");
  std::memset(o, 0, sizeof(*o));
  o->segment_size = 1024;
  o->heavy_active_vertices = 16384;
  o->heavy_active_edges = 25000000ull;
  o->bin_wave = 32;
  o->bin_chunk = 4096;
  o->chunk_edges = 2048;
  o->sched_mode = CG_SCHED_GPU;
  o->forced_choice = 0;
  o->cpu_threads = 0;
  o->pr_iterations = 20;
  o->pr_damping = 0.85f;
  o->profile = 0;
  o->trace_path = nullptr;
  o->mlp_path = nullptr;
  o->fault_flags = 0;
  o->batch_block = 256;
  o->batch_grid = 2048;
  o->batch_levels = 16;
  o->sssp_delta = 0;
  o->pr_segment_log2 = 0;
  o->atomic_return = 1;
  o->pr_layout = 0;
  o->relabel = 0;
  o->pr_kernel = 0;
  o->min_pull = 0;
  o->expand_grid = 0;
  o->batch_group = 0;
  o->hub_degree = 0;
  o->hub_delta = 0;
  o->probe_stride = 0;
  o->batch_engine_v = 0;
  return CG_OK;
}

int cg_graph_from_csr(uint32_t V, uint64_t E, const void* offsets, int offsets_u64, const uint32_t* dst,
                      const uint32_t* w, uint32_t zero_outdeg_tail, cg_graph** out) {
  clear_error();
  if (!out || !offsets || (E && !dst)) return set_error(CG_ERR_INVALID, "NULL argument");
  if (!offsets_u64 && E > 0xFFFFFFFFull) return set_error(CG_ERR_INVALID, "uint32 offsets cannot hold E");
  std::unique_ptr<cg_graph> g(new (std::nothrow) cg_graph());
  if (!g) return set_error(CG_ERR_OOM, "graph object");
  g->V = V;
  g->E = E;
  g->zero_tail = zero_outdeg_tail;
  g->has_w = w != nullptr;
  try {
    g->h_off.resize((size_t)V + 1);
    if (offsets_u64) std::memcpy(g->h_off.data(), offsets, ((size_t)V + 1) * sizeof(uint64_t));
    else {
      const uint32_t* o32 = static_cast<const uint32_t*>(offsets);
      for (size_t i = 0; i <= V; ++i) g->h_off[i] = o32[i];
    }
    g->h_dst.assign(dst, dst + E);
    if (w) g->h_w.assign(w, w + E);
  } catch (const std::bad_alloc&) {
    return set_error(CG_ERR_OOM, "host copy of the graph");
  }
  CG_TRY(validate_and_adopt(g.get()));
  *out = g.release();
  return CG_OK;
}

int cg_graph_load_files(const char* offset_file, int offsets_u64, const char* dst_file, const char* w_file,
                        uint32_t V, uint64_t E, cg_graph** out) {
  clear_error();
  if (!out || !offset_file || !dst_file) return set_error(CG_ERR_INVALID, "NULL argument");
  std::unique_ptr<cg_graph> g(new (std::nothrow) cg_graph());
  if (!g) return set_error(CG_ERR_OOM, "graph object");
  g->V = V;
  g->E = E;
  g->has_w = w_file != nullptr;
  try {
    g->h_off.resize((size_t)V + 1);
    g->h_dst.resize(E);
    if (w_file) g->h_w.resize(E);
  } catch (const std::bad_alloc&) {
    return set_error(CG_ERR_OOM, "host copy of the graph");
  }
  if (offsets_u64) {
    CG_TRY(read_file(offset_file, g->h_off.data(), ((uint64_t)V + 1) * 8));
  } else {
    std::vector<uint32_t> o32((size_t)V + 1);
    CG_TRY(read_file(offset_file, o32.data(), ((uint64_t)V + 1) * 4));
    for (size_t i = 0; i <= V; ++i) g->h_off[i] = o32[i];
  }
  CG_TRY(read_file(dst_file, g->h_dst.data(), E * 4));
  if (w_file) CG_TRY(read_file(w_file, g->h_w.data(), E * 4));
  CG_TRY(validate_and_adopt(g.get()));
  *out = g.release();
  return CG_OK;
}

int cg_graph_from_edges(uint32_t V, uint64_t m, const uint32_t* src, const uint32_t* dst, const uint32_t* w,
                        int symmetrize, cg_graph** out) {
  clear_error();
  if (!out || (m && (!src || !dst))) return set_error(CG_ERR_INVALID, "NULL argument");
  std::unique_ptr<cg_graph> g(new (std::nothrow) cg_graph());
  if (!g) return set_error(CG_ERR_OOM, "graph object");
  CG_TRY(build_from_edges(src, dst, w, m, V, symmetrize, g.get()));
  *out = g.release();
  return CG_OK;
}

int cg_graph_load_edge_file(const char* path, uint32_t V, int has_weights, int symmetrize, cg_graph** out) {
  clear_error();
  if (!path || !out) return set_error(CG_ERR_INVALID, "NULL argument");
  struct stat stt;
  if (::stat(path, &stt) != 0) return set_error(CG_ERR_IO, "cannot open %s", path);
  const uint64_t rec = has_weights ? 12 : 8;
  if ((uint64_t)stt.st_size % rec)
    return set_error(CG_ERR_IO, "%s holds %lld bytes, not a whole number of %llu-byte records", path,
                     (long long)stt.st_size, (unsigned long long)rec);
  const uint64_t m = (uint64_t)stt.st_size / rec;
  std::vector<uint32_t> raw;
  try {
    raw.resize(m * (rec / 4));
  } catch (const std::bad_alloc&) {
    return set_error(CG_ERR_OOM, "edge file buffer");
  }
  CG_TRY(read_file(path, raw.data(), (uint64_t)stt.st_size));
  std::vector<uint32_t> s(m), d(m), w(has_weights ? m : 0);
  const uint64_t k = rec / 4;
  for (uint64_t i = 0; i < m; ++i) {
    s[i] = raw[i * k];
    d[i] = raw[i * k + 1];
    if (has_weights) w[i] = raw[i * k + 2];
  }
  return cg_graph_from_edges(V, m, s.data(), d.data(), has_weights ? w.data() : nullptr, symmetrize, out);
}

int cg_graph_generate_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed,
                           uint32_t wlo, uint32_t whi, int symmetrize, cg_graph** out) {
  clear_error();
  if (!out) return set_error(CG_ERR_INVALID, "NULL argument");
  std::unique_ptr<cg_graph> g(new (std::nothrow) cg_graph());
  if (!g) return set_error(CG_ERR_OOM, "graph object");
  CG_TRY(gen_rmat(scale, edge_factor, a, b, c, seed, wlo, whi, symmetrize, g.get()));
  *out = g.release();
  return CG_OK;
}

int cg_graph_generate_grid(uint32_t n, double drop, uint64_t seed, uint32_t wlo, uint32_t whi, cg_graph** out) {
  clear_error();
  if (!out) return set_error(CG_ERR_INVALID, "NULL argument");
  std::unique_ptr<cg_graph> g(new (std::nothrow) cg_graph());
  if (!g) return set_error(CG_ERR_OOM, "graph object");
  CG_TRY(gen_grid(n, drop, seed, wlo, whi, g.get()));
  *out = g.release();
  return CG_OK;
}

int cg_graph_info(const cg_graph* g, uint32_t* V, uint64_t* E, int* has_w, int* symmetric, uint32_t* mv) {
  if (!g) return set_error(CG_ERR_INVALID, "NULL graph");
  if (V) *V = g->V;
  if (E) *E = g->E;
  if (has_w) *has_w = g->has_w;
  if (symmetric) *symmetric = g->symmetric;
  if (mv) *mv = g->max_outdeg_vertex;
  return CG_OK;
}

int cg_graph_export(const cg_graph* cg, uint64_t* offsets, uint32_t* dst, uint32_t* w) {
  if (!cg) return set_error(CG_ERR_INVALID, "NULL graph");
  cg_graph* g = const_cast<cg_graph*>(cg);
  CG_TRY(g->ensure_host());
  if (offsets) std::memcpy(offsets, g->h_off.data(), ((size_t)g->V + 1) * sizeof(uint64_t));
  if (dst && g->E) std::memcpy(dst, g->h_dst.data(), g->E * sizeof(uint32_t));
  if (w && g->has_w && g->E) std::memcpy(w, g->h_w.data(), g->E * sizeof(uint32_t));
  return CG_OK;
}

int cg_graph_build_transpose(cg_graph* g) {
  clear_error();
  if (!g) return set_error(CG_ERR_INVALID, "NULL graph");
  CG_TRY(g->ensure_device());
  return build_transpose(g);
}

int cg_graph_device_csr(const cg_graph* cg, const uint64_t** offsets, const uint32_t** dst, const uint32_t** w) {
  if (!cg) return set_error(CG_ERR_INVALID, "NULL graph");
  cg_graph* g = const_cast<cg_graph*>(cg);
  CG_TRY(g->ensure_device());
  if (offsets) *offsets = g->d_off.p;
  if (dst) *dst = g->d_dst.p;
  if (w) *w = g->has_w ? g->d_w.p : nullptr;
  return CG_OK;
}

void cg_graph_destroy(cg_graph* g) { delete g; }

// TwoLevelPartition (my_project/CGgraph/Two-level_partition.cuh:213-476):
// 1024-vertex segments over V' = V - zero tail, sorted by average degree (then
// id), cut into 4 buckets at average degree 17 / 33 / 257 (lower_bound), and
// charged against this GPU's HBM in ascending-degree order.  The average is
// taken in floating point (the reference divides integers first, :253).
int cg_partition(const cg_graph* cg, int algorithm, cg_partition_info* info) {
  clear_error();
  if (!cg || !info) return set_error(CG_ERR_INVALID, "NULL argument");
  cg_graph* g = const_cast<cg_graph*>(cg);
  CG_TRY(g->ensure_host());
  return cg::plan_partition(g, algorithm, cg::device_memory_bytes(), info, nullptr);
}

int cg_engine_create(const cg_graph* cg, int algorithm, int engine_type, int use_device_num, const cg_options* opt,
                     cg_engine** out) {
  clear_error();
  if (!cg || !out) return set_error(CG_ERR_INVALID, "NULL argument");
  if (algorithm < CG_BFS || algorithm > CG_PR) return set_error(CG_ERR_INVALID, "unknown algorithm %d", algorithm);
  if (engine_type != CG_COORDINATION && engine_type != CG_SINGLE_GPU && engine_type != CG_MULTI_CORE &&
      engine_type != CG_SINGLE_CORE)
    return set_error(CG_ERR_UNSUPPORTED,
                     "engine type %d is not provided; this library runs COORDINATION (0), SINGLE_CORE (2), "
                     "MULTI_CORE (3) and SINGLE_GPU (4)",
                     engine_type);
  if (use_device_num < 0) return set_error(CG_ERR_INVALID, "use_device_num = %d", use_device_num);
  if (use_device_num > 1)
    return set_error(CG_ERR_UNSUPPORTED, "use_device_num = %d: libcgmi drives one GPU per engine", use_device_num);
  // Without a GPU (use_device_num 0, or a CPU engine type) every iteration
  // runs on the host team; the caller asked for it, so this is not a fallback:
  // a GPU engine on a machine without one still fails with CG_ERR_NO_DEVICE.
  const bool host_only = engine_type == CG_MULTI_CORE || engine_type == CG_SINGLE_CORE || use_device_num == 0;
  if (host_only) {
    if (engine_type == CG_SINGLE_GPU)
      return set_error(CG_ERR_INVALID, "SINGLE_GPU with use_device_num = 0");
    if (algorithm == CG_PR)
      return set_error(CG_ERR_UNSUPPORTED, "PageRank runs on the GPU engines only");
    cg_graph* g = const_cast<cg_graph*>(cg);
    if (algorithm == CG_SSSP && !g->has_w) return set_error(CG_ERR_INVALID, "SSSP needs edge weights");
    CG_TRY(g->ensure_host());
    std::unique_ptr<cg_engine> e(new (std::nothrow) cg_engine());
    if (!e) return set_error(CG_ERR_OOM, "engine object");
    e->g = g;
    e->caller_g = g;
    e->alg = algorithm;
    e->engine_type = engine_type;
    e->use_device = 0;
    e->host_only = true;
    if (opt) e->opt = *opt;
    else cg_options_default(&e->opt);
    if (e->opt.trace_path) { e->trace_path = e->opt.trace_path; e->opt.trace_path = e->trace_path.c_str(); }
    e->opt.mlp_path = nullptr;
    if (engine_type == CG_SINGLE_CORE) e->opt.cpu_threads = 1;
    if (e->opt.trace_path) {
      e->trace = std::fopen(e->opt.trace_path, "a");
      if (!e->trace) return set_error(CG_ERR_IO, "cannot open trace file %s", e->opt.trace_path);
    }
    *out = e.release();
    return CG_OK;
  }
  CG_TRY(require_device());
  cg_graph* g = const_cast<cg_graph*>(cg);
  if (algorithm == CG_SSSP && !g->has_w) return set_error(CG_ERR_INVALID, "SSSP needs edge weights");
  std::unique_ptr<cg_engine> e(new (std::nothrow) cg_engine());
  if (!e) return set_error(CG_ERR_OOM, "engine object");
  e->g = g;
  e->caller_g = g;
  e->alg = algorithm;
  e->engine_type = engine_type;
  e->use_device = use_device_num;
  if (opt) e->opt = *opt;
  else cg_options_default(&e->opt);
  // TwoLevelPartition's placement under the HBM budget: a graph that does
  // not fit runs partly resident (BFS / SSSP / CC)
  {
    const int64_t budget = e->opt.hbm_budget ? (int64_t)e->opt.hbm_budget : device_memory_bytes();
    if (!g->dev_valid || e->opt.hbm_budget) {
      CG_TRY(g->ensure_host());
      cg_partition_info pinfo{};
      std::vector<uint8_t> resident;
      CG_TRY(plan_partition(g, algorithm, budget, &pinfo, &resident));
      e->segments = pinfo.segment_num;
      if (!pinfo.can_hold_all) {
        if (algorithm == CG_PR)
          return set_error(CG_ERR_UNSUPPORTED, "PageRank needs the whole graph in HBM (%llu bytes over the budget)",
                           (unsigned long long)pinfo.device_bytes);
        if (e->opt.sched_mode == CG_SCHED_PROBE)
          return set_error(CG_ERR_UNSUPPORTED, "probe traces need the whole graph in HBM");
        CG_TRY(build_resident_view(e.get(), resident));
        for (uint8_t r : resident) e->resident_segments += r;
      }
    }
    if (!e->partial) CG_TRY(g->ensure_device());
  }
  // keep private copies of the caller's strings
  if (e->opt.trace_path) { e->trace_path = e->opt.trace_path; e->opt.trace_path = e->trace_path.c_str(); }
  if (e->opt.mlp_path) { e->mlp_path = e->opt.mlp_path; e->opt.mlp_path = e->mlp_path.c_str(); }
  if (engine_type == CG_SINGLE_GPU) e->opt.sched_mode = CG_SCHED_GPU;
  if (e->opt.chunk_edges == 0 || e->opt.bin_wave == 0 || e->opt.bin_chunk < e->opt.bin_wave)
    return set_error(CG_ERR_INVALID, "degree tiers need 0 < bin_wave <= bin_chunk and chunk_edges > 0");
  if (e->opt.sched_mode < CG_SCHED_HEURISTIC || e->opt.sched_mode > CG_SCHED_PROBE)
    return set_error(CG_ERR_INVALID, "unknown scheduler mode %d", e->opt.sched_mode);
  if (e->opt.sched_mode == CG_SCHED_PROBE && !e->opt.trace_path)
    return set_error(CG_ERR_INVALID, "probe mode needs a trace_path");
  if (e->opt.sched_mode == CG_SCHED_FORCED && (e->opt.forced_choice < 0 || e->opt.forced_choice >= CG_NUM_CHOICES))
    return set_error(CG_ERR_INVALID, "forced_choice %d outside [0, 16)", e->opt.forced_choice);
  CG_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  CG_HIP(hipEventCreate(&e->ev0));
  CG_HIP(hipEventCreate(&e->ev1));
  CG_HIP(hipEventCreateWithFlags(&e->ev_copy, hipEventDisableTiming));
  for (auto& pr : e->ev_pair)
    for (auto& ev : pr) CG_HIP(hipEventCreate(&ev));
  for (auto& ev : e->ev_piece) CG_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  for (auto& ev : e->ev_pcie) CG_HIP(hipEventCreate(&ev));
  if (e->opt.relabel > 2) return set_error(CG_ERR_INVALID, "relabel %u outside [0, 2]", e->opt.relabel);
  if (e->opt.hub_degree && e->opt.sssp_delta)
    return set_error(CG_ERR_INVALID, "hub_degree and sssp_delta are alternatives; set one");
  // Automatic renumbering pays once the values no longer fit the caches
  // (measured on R-MAT: SSSP from 2^21 vertices, BFS from 2^23; below that
  // the renumbered run is up to 15 % slower).
  const uint32_t relabel_min_v = algorithm == CG_BFS ? (1u << 23) : (1u << 21);
  if ((algorithm == CG_BFS || algorithm == CG_SSSP) && g->V > 0 && !e->partial &&
      (e->opt.relabel == 2 || (e->opt.relabel == 0 && g->V >= relabel_min_v))) {
    // Renumber by descending out-degree when asked, or (automatic) when the
    // top 1/64 of vertices send at least half of the edges: the values those
    // edges touch then share few cache lines.
    cg::DevBuf<uint32_t> perm;
    double share = 0.0;
    CG_TRY(degree_order(g, perm, e->new_of_d, &share));
    if (e->opt.relabel == 2 || share >= 0.5) {
      e->own_g.reset(new (std::nothrow) cg_graph());
      if (!e->own_g) return set_error(CG_ERR_OOM, "renumbered graph object");
      CG_TRY(relabel_graph(g, e->new_of_d.p, e->own_g.get()));
      e->new_of.resize(g->V);
      CG_HIP(hipMemcpy(e->new_of.data(), e->new_of_d.p, (size_t)g->V * 4, hipMemcpyDeviceToHost));
      CG_TRY(e->val_out.ensure(g->V));
      e->g = e->own_g.get();
      e->renumbered = true;
      g = e->g;
    } else {
      e->new_of_d.reset();
    }
  }
  const uint64_t V = g->V;
  const uint64_t words = (V + 63) / 64 + 1;
  if (algorithm == CG_PR) {
    CG_TRY(e->rank.ensure(V));
    CG_TRY(e->rank2.ensure(V));
    CG_TRY(e->contrib.ensure(V));
    CG_TRY(e->contrib2.ensure(V));
    CG_TRY(pr_plan_create(g, e->opt.pr_layout, e->opt.pr_segment_log2, e->opt.pr_kernel, &e->plan, e->stream));
  } else {
    // the frontier compaction reads out-degrees rather than offset pairs
    if (!e->partial && !e->host_only) CG_TRY(ensure_degrees(g, e->stream));
    if (algorithm == CG_SSSP && !e->partial && !e->host_only) CG_TRY(ensure_narrow_weights(g, e->stream));
    CG_TRY(e->val.ensure(V));
    if (algorithm != CG_BFS && !e->partial && !e->host_only) CG_TRY(e->xval.ensure(V));
    CG_TRY(e->bm[0].ensure(words));
    CG_TRY(e->bm[1].ensure(words));
    CG_TRY(e->q_small.ensure(V));
    CG_TRY(e->q_med.ensure(V));
    uint32_t min_chunk = e->opt.chunk_edges;
    CG_TRY(e->q_chunk.ensure(V + g->E / min_chunk + 2 * kChunkSpread + 1));
    CG_TRY(e->st.ensure(2));
    CG_TRY(e->hst.ensure(3));
    CG_TRY(e->hst_map.ensure(3, hipHostMallocMapped | hipHostMallocCoherent));  // two iterations, query done
    e->hst_map_dev = e->hst_map.device_ptr();
    if (!e->hst_map_dev) return set_error(CG_ERR_HIP, "mapped statistics buffer has no device address");
    CG_TRY(e->acc.ensure(kAccSlots));
    CG_TRY(e->parts.ensure(stats_parts_needed(words)));
    CG_TRY(e->stripes.ensure(kStripeWords));
    CG_TRY(e->run_state.ensure(2));
    if (algorithm == CG_BFS) CG_TRY(e->bfs_cnt.ensure(2 * kBfsCntWords));
  }
  if (e->opt.trace_path) {
    e->trace = std::fopen(e->opt.trace_path, "a");
    if (!e->trace) return set_error(CG_ERR_IO, "cannot open trace file %s", e->opt.trace_path);
  }
  *out = e.release();
  return CG_OK;
}

int cg_engine_run(cg_engine* e, uint32_t root, void* out, cg_run_stats* stats) {
  clear_error();
  if (!e) return set_error(CG_ERR_INVALID, "NULL engine");
  cg_graph& g = *e->g;
  if (e->alg != CG_CC && e->alg != CG_PR && root >= g.V)
    return set_error(CG_ERR_INVALID, "root %u >= V = %u", root, g.V);
  e->result_valid = false;
  cg_run_stats rs{};
  int st = CG_OK;
  if (e->host_only) {
    CG_TRY(engine_run_host(e, root, &rs));
    e->result_valid = true;
    e->last_root = root;
    if (stats) *stats = rs;
    if (out) return cg_engine_result(e, out);
    return CG_OK;
  }
  CG_HIP(hipSetDevice(0));
  if (e->alg == CG_PR) {
    DevGraph dg{g.d_off.p, g.d_dst.p, nullptr, g.d_in_off.p, g.d_in_src.p, g.V};
    const auto t0 = std::chrono::steady_clock::now();
    double kms = 0;
    uint32_t launches = 0;
    st = pr_run(dg, e->plan, e->opt.pr_iterations, e->opt.pr_damping, e->rank.p, e->rank2.p, e->contrib.p,
                e->contrib2.p, e->stream, e->opt.profile, &kms, &launches);
    rs.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    rs.iterations = e->opt.pr_iterations;
    rs.heavy_iterations = rs.iterations;
    rs.expand_launches = launches;
    rs.expand_kernel_ms = kms;
    rs.edges_relaxed = g.E * (uint64_t)e->opt.pr_iterations;
    rs.edges_traversed = g.E * (uint64_t)e->opt.pr_iterations;
    pr_plan_info(e->plan, &rs.pr_layout, &rs.pr_stream_bytes);
    rs.expand_alg_bytes = (double)e->opt.pr_iterations *
                          (8.0 * (double)g.E + 8.0 * ((double)g.V + 1) + 4.0 * (double)g.V);
  } else if (e->alg == CG_CC && g.symmetric &&
             (e->opt.sched_mode == CG_SCHED_GPU || e->opt.sched_mode == CG_SCHED_LEARNED)) {
    // Symmetric input: min-label hooking reaches the label-propagation fixed
    // point directly; the per-iteration schedulers apply to the frontier
    // engine (directed input, or the heuristic / forced / probe modes).
    DevGraph dg{g.d_off.p, g.d_dst.p, nullptr, nullptr, nullptr, g.V};
    const auto t0 = std::chrono::steady_clock::now();
    if (e->opt.profile) CG_HIP(hipEventRecord(e->ev0, e->stream));
    st = wcc_hooking(dg, e->val.p, 2, e->stream, nullptr);
    if (st == CG_OK && e->opt.profile) CG_HIP(hipEventRecord(e->ev1, e->stream));
    if (st == CG_OK) CG_HIP(hipStreamSynchronize(e->stream));
    rs.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (st == CG_OK && e->opt.profile) {
      float ms = 0;
      CG_HIP(hipEventElapsedTime(&ms, e->ev0, e->ev1));
      rs.expand_kernel_ms = ms;
    }
    rs.iterations = 1;
    rs.heavy_iterations = 1;
    rs.expand_launches = 1;
    rs.edges_traversed = g.E;
    rs.edges_relaxed = g.E;
    rs.expand_alg_bytes = 8.0 * ((double)g.V + 1) + 4.0 * (double)g.E + 8.0 * (double)g.V;
  } else if (e->renumbered) {
    // the query's last pass brings the values back to the caller's ids
    st = engine_run(e, e->new_of[root], &rs);
  } else {
    st = engine_run(e, root, &rs);
  }
  if (st != CG_OK) return st;
  e->result_valid = true;
  e->last_root = root;
  rs.resident_edges = e->g->E;
  if (stats) *stats = rs;
  if (out) return cg_engine_result(e, out);
  return CG_OK;
}

int cg_engine_result(cg_engine* e, void* out) {
  if (!e || !out) return set_error(CG_ERR_INVALID, "NULL argument");
  if (!e->result_valid) return set_error(CG_ERR_INVALID, "no completed run to read");
  if (e->host_only) {
    std::memcpy(out, e->team.shadow.data(), (size_t)e->g->V * 4);
    return CG_OK;
  }
  const void* src = e->alg == CG_PR ? (const void*)e->rank.p
                    : e->renumbered   ? (const void*)e->val_out.p
                                      : (const void*)e->val.p;
  // the engine's stream does not synchronise with the null stream, and a run
  // may return with no-op iterations still queued on it
  CG_HIP(stream_wait(e->stream));
  CG_HIP(hipMemcpy(out, src, (size_t)e->g->V * 4, hipMemcpyDeviceToHost));
  return CG_OK;
}

int cg_engine_result_device(cg_engine* e, void** dev_ptr) {
  if (!e || !dev_ptr) return set_error(CG_ERR_INVALID, "NULL argument");
  if (e->host_only) return set_error(CG_ERR_UNSUPPORTED, "an engine without a GPU keeps its values in host memory");
  CG_HIP(stream_wait(e->stream));  // the values are complete when the pointer is handed out
  *dev_ptr = e->alg == CG_PR ? (void*)e->rank.p : e->renumbered ? (void*)e->val_out.p : (void*)e->val.p;
  return CG_OK;
}

void cg_engine_destroy(cg_engine* e) { delete e; }

int cg_push_step_device(int algorithm, const cg_graph* cg, uint32_t* val, const uint64_t* front_in,
                        uint64_t* front_out, void* stream) {
  clear_error();
  if (!cg || !val || !front_in || !front_out) return set_error(CG_ERR_INVALID, "NULL argument");
  if (algorithm == CG_PR) return set_error(CG_ERR_UNSUPPORTED, "PageRank has no push step");
  cg_graph* g = const_cast<cg_graph*>(cg);
  if (algorithm == CG_SSSP && !g->has_w) return set_error(CG_ERR_INVALID, "SSSP needs edge weights");
  CG_TRY(g->ensure_device());
  DevGraph dg{g->d_off.p, g->d_dst.p, g->has_w ? g->d_w.p : nullptr, nullptr, nullptr, g->V};
  return push_step_from_bitmap(algorithm, dg, val, front_in, front_out, static_cast<hipStream_t>(stream));
}

int cg_check_result_file(const char* path, const uint32_t* values, uint32_t V, const uint32_t* old2new,
                         uint64_t* first_bad) {
  clear_error();
  if (!path || !values) return set_error(CG_ERR_INVALID, "NULL argument");
  std::vector<uint32_t> ref(V);
  CG_TRY(read_file(path, ref.data(), (uint64_t)V * 4));
  for (uint32_t i = 0; i < V; ++i) {
    const uint32_t mine = old2new ? values[old2new[i]] : values[i];
    if (mine != ref[i]) {
      if (first_bad) *first_bad = i;
      return set_error(CG_ERR_MISMATCH, "value of vertex %u is %u, the result file holds %u", i, mine, ref[i]);
    }
  }
  return CG_OK;
}

int cg_save_bin_u32(const char* path, const uint32_t* values, uint64_t n) {
  clear_error();
  if (!path || (n && !values)) return set_error(CG_ERR_INVALID, "NULL argument");
  FILE* f = std::fopen(path, "wb");
  if (!f) return set_error(CG_ERR_IO, "cannot create %s", path);
  const size_t wr = n ? std::fwrite(values, 4, n, f) : 0;
  std::fclose(f);
  if (wr != n) return set_error(CG_ERR_IO, "short write to %s", path);
  return CG_OK;
}

int cg_abi_sizes(size_t* options, size_t* run_stats, size_t* partition_info) {
  if (options) *options = sizeof(cg_options);
  if (run_stats) *run_stats = sizeof(cg_run_stats);
  if (partition_info) *partition_info = sizeof(cg_partition_info);
  return CG_OK;
}

const char* cg_last_error(void) { return last_error(); }
const char* cg_version(void) { return "cgmi 0.1 (gfx950)"; }

}  // extern "C"
