/*
 * Restatement of the reference's MULTI_CORE CPU engine -- the CPU baseline
 * that bench.py times beside the GPU path, and a second checker for tests.
 * TEST INFRASTRUCTURE ONLY (never linked into the product library).
 *
 * Follows (paths relative to CGgraph/my_project/CGgraph/):
 *   Graph_Numa::graphProcess            GraphNuma.cuh:166-314   iteration loop
 *   Graph_Numa::init_algorithm          GraphNuma.cuh:521-563   per-algorithm init
 *   TaskSteal::threeStage_taskSteal     taskSteal.hpp:89-274    three-stage scheduler
 *   bfs_numa / sssp_numa / cc_numa      host_algorithm.hpp:164-184, 277-300, 412-433
 *   Gemini_atomic::write_min            host_algorithm.hpp:23-39
 *   dense_bitset::set_bit               Basic/Bitmap/dense_bitset.hpp:202-208
 * Constants: VERTEXSTEAL_CHUNK 64, EDGESTEAL_THRESHOLD 12800, EDGESTEAL_CHUNK
 * 6400 (Basic/Graph/basic_def.hpp:14-17).
 *
 * Differences, on purpose:
 *   - one socket (no libnuma split); the per-socket loops collapse to one.
 *   - the edge-donation record is one 64-bit word (vertex << 32 | next chunk)
 *     claimed by CAS, so a helper can never claim a chunk of a vertex the donor
 *     has already moved past (the reference reads vertex/end/cur separately
 *     and CASes only cur, taskSteal.hpp:231-236).
 *   - plain reads racing with write_min are relaxed atomic loads.
 */
#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "oracle.h"

#define VSTEAL_CHUNK 64
#define ESTEAL_THRESHOLD 12800
#define ESTEAL_CHUNK 6400
#define NONE_VERTEX 0xFFFFFFFFull

enum { VERTEX_WORKING = 0, EDGE_DONATE = 1, VERTEX_STEALING = 2 };

typedef struct {
  uint64_t cur, end;   /* vertex range [cur, end), claimed 64 at a time */
  int status;
  uint64_t donate;     /* (vertex << 32) | next chunk, NONE_VERTEX << 32 when idle */
} __attribute__((aligned(64))) tstate;

typedef struct {
  int alg;
  const uint64_t* off;
  const uint32_t* dst;
  const uint32_t* w;
  uint32_t* val;
  uint64_t* in;
  uint64_t* out;
} eng;

static inline uint32_t ld(const uint32_t* p) { return __atomic_load_n(p, __ATOMIC_RELAXED); }

static inline int write_min(uint32_t* p, uint32_t v) {
  uint32_t cur = ld(p);
  while (cur > v) {
    if (__atomic_compare_exchange_n(p, &cur, v, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) return 1;
  }
  return 0;
}

static inline void set_bit(uint64_t* bm, uint32_t v) {
  __atomic_fetch_or(&bm[v >> 6], 1ull << (v & 63), __ATOMIC_RELAXED);
}

/* processEdge: relax edges [s, e) of vertex v; returns successful updates. */
static uint32_t process_edge(eng* g, uint32_t v, uint64_t s, uint64_t e) {
  uint32_t cnt = 0;
  uint32_t sv = ld(&g->val[v]);
  for (uint64_t k = s; k < e; ++k) {
    uint32_t d = g->dst[k];
    uint32_t cand = g->alg == 0 ? sv + 1 : (g->alg == 1 ? sv + g->w[k] : sv);
    if (cand < ld(&g->val[d]) && write_min(&g->val[d], cand)) {
      set_bit(g->out, d);
      ++cnt;
    }
  }
  return cnt;
}

/* Donate the edges of a high-degree vertex in ESTEAL_CHUNK pieces; the donor
 * and any helper claim pieces through the same CAS on the donation word. */
static uint32_t claim_donated(eng* g, uint64_t* word) {
  uint32_t cnt = 0;
  for (;;) {
    uint64_t wv = __atomic_load_n(word, __ATOMIC_ACQUIRE);
    uint64_t v = wv >> 32;
    if (v == NONE_VERTEX) return cnt;
    uint64_t ci = wv & 0xFFFFFFFFull;
    uint64_t s = g->off[v] + ci * ESTEAL_CHUNK, end = g->off[v + 1];
    if (s >= end) return cnt;
    if (!__atomic_compare_exchange_n(word, &wv, wv + 1, 0, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE))
      continue;
    uint64_t e = s + ESTEAL_CHUNK < end ? s + ESTEAL_CHUNK : end;
    cnt += process_edge(g, (uint32_t)v, s, e);
  }
}

static uint32_t do_chunk(eng* g, tstate* me, uint64_t vbase, int stealing) {
  uint32_t cnt = 0;
  uint64_t word = __atomic_load_n(&g->in[vbase >> 6], __ATOMIC_RELAXED);
  uint64_t v = vbase;
  while (word) {
    if (word & 1) {
      uint64_t s = g->off[v], e = g->off[v + 1];
      uint64_t deg = e - s;
      /* VERTEX_WORKING donates at >= threshold, VERTEX_STEALING at > threshold
       * (taskSteal.hpp:136 vs :187). */
      if (stealing ? deg > ESTEAL_THRESHOLD : deg >= ESTEAL_THRESHOLD) {
        __atomic_store_n(&me->donate, (v << 32), __ATOMIC_RELEASE);
        __atomic_store_n(&me->status, EDGE_DONATE, __ATOMIC_RELEASE);
        cnt += claim_donated(g, &me->donate);
        __atomic_store_n(&me->donate, NONE_VERTEX << 32, __ATOMIC_RELEASE);
        __atomic_store_n(&me->status, stealing ? VERTEX_STEALING : VERTEX_WORKING, __ATOMIC_RELEASE);
      } else if (deg > 0) {
        cnt += process_edge(g, (uint32_t)v, s, e);
      }
    }
    ++v;
    word >>= 1;
  }
  return cnt;
}

static uint64_t three_stage(eng* g, uint32_t V, tstate* ts, int T, int* working) {
  uint64_t per = ((uint64_t)V / T) / VSTEAL_CHUNK * VSTEAL_CHUNK;
  for (int t = 0; t < T; ++t) {
    ts[t].cur = per * t;
    ts[t].end = (t == T - 1) ? V : per * (t + 1);
    ts[t].status = VERTEX_WORKING;
    ts[t].donate = NONE_VERTEX << 32;
  }
  __atomic_store_n(working, T, __ATOMIC_RELEASE);
  uint64_t total = 0;
#pragma omp parallel num_threads(T) reduction(+ : total)
  {
    int tid = omp_get_thread_num();
    tstate* me = &ts[tid];
    uint64_t local = 0;
    /* 1.1 VERTEX_WORKING */
    for (;;) {
      uint64_t c = __atomic_fetch_add(&me->cur, VSTEAL_CHUNK, __ATOMIC_RELAXED);
      if (c >= me->end) break;
      local += do_chunk(g, me, c, 0);
    }
    /* 1.2 VERTEX_STEALING */
    __atomic_store_n(&me->status, VERTEX_STEALING, __ATOMIC_RELEASE);
    for (int o = 1; o < T; ++o) {
      tstate* h = &ts[(tid + o) % T];
      while (__atomic_load_n(&h->status, __ATOMIC_ACQUIRE) != VERTEX_STEALING) {
        uint64_t c = __atomic_fetch_add(&h->cur, VSTEAL_CHUNK, __ATOMIC_RELAXED);
        if (c >= h->end) break;
        local += do_chunk(g, me, c, 1);
      }
    }
    /* 1.3 EDGE_WORKING: help donors until every thread has reached here */
    __atomic_fetch_sub(working, 1, __ATOMIC_ACQ_REL);
    while (__atomic_load_n(working, __ATOMIC_ACQUIRE) > 0) {
      for (int o = 1; o < T; ++o) {
        tstate* h = &ts[(tid + o) % T];
        if (__atomic_load_n(&h->status, __ATOMIC_ACQUIRE) == EDGE_DONATE)
          local += claim_donated(g, &h->donate);
      }
    }
    total += local;
  }
  return total;
}

static double now_ms(void) {
  struct timeval tv;
  gettimeofday(&tv, NULL);
  return tv.tv_sec * 1e3 + tv.tv_usec * 1e-3;
}

double or_multicore_run(int alg, uint32_t V, const uint64_t* off, const uint32_t* dst,
                        const uint32_t* w, uint32_t root, int threads, uint32_t* out,
                        uint32_t* iters) {
  if (threads <= 0) threads = omp_get_max_threads();
  uint64_t words = ((uint64_t)V + 63) / 64 + 1;
  eng g;
  g.alg = alg; g.off = off; g.dst = dst; g.w = w; g.val = out;
  g.in = (uint64_t*)calloc(words, sizeof(uint64_t));
  g.out = (uint64_t*)calloc(words, sizeof(uint64_t));
  tstate* ts = (tstate*)aligned_alloc(64, sizeof(tstate) * (size_t)threads);
  int working = 0;
  /* init_algorithm (GraphNuma.cuh:521-563) */
  if (alg == 2) {
    for (uint32_t v = 0; v < V; ++v) out[v] = v;
    for (uint32_t v = 0; v < V; ++v) g.in[v >> 6] |= 1ull << (v & 63);
  } else {
    for (uint32_t v = 0; v < V; ++v) out[v] = OR_UNREACHED;
    if (root < V) { out[root] = 0; g.in[root >> 6] |= 1ull << (root & 63); }
  }
  uint32_t ite = 0;
  double t0 = now_ms();
  for (;;) {
    ++ite;
    memset(g.out, 0, words * sizeof(uint64_t));
    uint64_t active = three_stage(&g, V, ts, threads, &working);
    if (active == 0) break;
    uint64_t* t = g.in; g.in = g.out; g.out = t;
  }
  double ms = now_ms() - t0;
  if (iters) *iters = ite;
  free(g.in); free(g.out); free(ts);
  return ms;
}

/* omp_get_max_threads() of this library's OpenMP runtime (the thread count a
 * run with threads <= 0 uses), for the baseline's report. */
int or_omp_max_threads(void) { return omp_get_max_threads(); }
