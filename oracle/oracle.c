/*
 * CPU oracle: textbook restatements of the reference's traversal results.
 * TEST INFRASTRUCTURE ONLY -- see oracle.h for what may call it and how each
 * function is pinned.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

#define GOLDEN 0x9E3779B97F4A7C15ULL

/* SplitMix64 finaliser (Steele, Lea, Flood, "Fast splittable pseudorandom
 * number generators", OOPSLA 2014). */
uint64_t or_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static uint32_t prob_threshold(double p) {
  double t = p * 4294967296.0;
  if (t <= 0.0) return 0;
  if (t >= 4294967295.0) return 0xFFFFFFFFu;
  return (uint32_t)t;
}

/* R-MAT stream: edge i descends `scale` levels of the adjacency matrix.  At
 * level l it takes 32 random bits from draw(i, l/2) and picks quadrant
 * 0 (a) / 1 (b) / 2 (c) / 3 (d) by integer thresholds; the quadrant's high bit
 * extends the source id, its low bit the destination id. */
void or_rmat_edges(uint32_t scale, uint64_t first, uint64_t count, double a, double b,
                   double c, uint64_t seed, uint32_t* src, uint32_t* dst) {
  const uint32_t tA = prob_threshold(a), tAB = prob_threshold(a + b),
                 tABC = prob_threshold(a + b + c);
  const uint64_t key = or_mix64(seed * GOLDEN + 0x5851F42D4C957F2DULL);
  for (uint64_t k = 0; k < count; ++k) {
    const uint64_t i = first + k;
    uint32_t s = 0, d = 0;
    uint64_t r = 0;
    for (uint32_t l = 0; l < scale; ++l) {
      if ((l & 1) == 0) r = or_mix64(key + (i * 16 + (l >> 1) + 1) * GOLDEN);
      uint32_t x = (uint32_t)(r >> ((l & 1) * 32));
      uint32_t q = x < tA ? 0u : (x < tAB ? 1u : (x < tABC ? 2u : 3u));
      s = (s << 1) | (q >> 1);
      d = (d << 1) | (q & 1);
    }
    src[k] = s;
    dst[k] = d;
  }
}

uint32_t or_edge_weight(uint64_t seed, uint32_t u, uint32_t v, uint32_t lo, uint32_t hi,
                        int symmetric) {
  if (symmetric && u > v) { uint32_t t = u; u = v; v = t; }
  const uint64_t key = or_mix64(seed * GOLDEN + 0x2545F4914F6CDD1DULL);
  uint64_t h = or_mix64(key + ((((uint64_t)u) << 32) | v) * GOLDEN);
  uint64_t span = (uint64_t)(hi - lo) + 1;
  return lo + (uint32_t)(((h >> 32) * span) >> 32);
}

uint64_t or_grid_edges(uint32_t n, double drop, uint64_t seed, uint32_t* src, uint32_t* dst) {
  const uint64_t key = or_mix64(seed * GOLDEN + 0x94D049BB133111EBULL);
  const uint32_t tdrop = prob_threshold(drop);
  uint64_t m = 0;
  const uint64_t nh = (uint64_t)n * (n - 1); /* horizontal edges */
  for (uint64_t e = 0; e < 2 * nh; ++e) {
    uint32_t u, v;
    if (e < nh) {
      uint32_t r = (uint32_t)(e / (n - 1)), c = (uint32_t)(e % (n - 1));
      u = r * n + c; v = u + 1;
    } else {
      uint64_t f = e - nh;
      uint32_t r = (uint32_t)(f / n), c = (uint32_t)(f % n);
      u = r * n + c; v = u + n;
    }
    uint32_t x = (uint32_t)(or_mix64(key + (e + 1) * GOLDEN) >> 32);
    if (x < tdrop) continue;
    src[m] = u; dst[m] = v; ++m;
    src[m] = v; dst[m] = u; ++m;
  }
  return m;
}

/* LSD radix sort of 64-bit keys, 16-bit digits, skipping digits that are zero
 * in every key. */
static void radix_sort_u64(uint64_t* k, uint64_t n) {
  if (n < 2) return;
  uint64_t* tmp = (uint64_t*)malloc(n * sizeof(uint64_t));
  uint64_t all_or = 0;
  for (uint64_t i = 0; i < n; ++i) all_or |= k[i];
  uint64_t* cnt = (uint64_t*)malloc(65536 * sizeof(uint64_t));
  uint64_t *a = k, *b = tmp;
  for (int shift = 0; shift < 64; shift += 16) {
    if (((all_or >> shift) & 0xFFFF) == 0) continue;
    memset(cnt, 0, 65536 * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) cnt[(a[i] >> shift) & 0xFFFF]++;
    uint64_t s = 0;
    for (int d = 0; d < 65536; ++d) { uint64_t t = cnt[d]; cnt[d] = s; s += t; }
    for (uint64_t i = 0; i < n; ++i) b[cnt[(a[i] >> shift) & 0xFFFF]++] = a[i];
    uint64_t* t = a; a = b; b = t;
  }
  if (a != k) memcpy(k, a, n * sizeof(uint64_t));
  free(cnt);
  free(tmp);
}

int64_t or_csr_build(uint32_t V, uint64_t m, const uint32_t* src, const uint32_t* dst,
                     int symmetrize, uint64_t* off, uint32_t* out_dst, uint64_t cap) {
  uint64_t n = symmetrize ? 2 * m : m;
  uint64_t* keys = (uint64_t*)malloc((n ? n : 1) * sizeof(uint64_t));
  uint64_t q = 0;
  for (uint64_t i = 0; i < m; ++i) {
    if (src[i] == dst[i]) continue;
    keys[q++] = (((uint64_t)src[i]) << 32) | dst[i];
    if (symmetrize) keys[q++] = (((uint64_t)dst[i]) << 32) | src[i];
  }
  radix_sort_u64(keys, q);
  memset(off, 0, ((uint64_t)V + 1) * sizeof(uint64_t));
  uint64_t E = 0;
  for (uint64_t i = 0; i < q; ++i) {
    if (i > 0 && keys[i] == keys[i - 1]) continue;
    if (E >= cap) { free(keys); return -1; }
    out_dst[E++] = (uint32_t)keys[i];
    off[(keys[i] >> 32) + 1]++;
  }
  for (uint32_t v = 0; v < V; ++v) off[v + 1] += off[v];
  free(keys);
  return (int64_t)E;
}

void or_fill_weights(uint32_t V, const uint64_t* off, const uint32_t* dst, uint64_t seed,
                     uint32_t lo, uint32_t hi, int symmetric, uint32_t* w) {
  for (uint32_t u = 0; u < V; ++u)
    for (uint64_t e = off[u]; e < off[u + 1]; ++e)
      w[e] = or_edge_weight(seed, u, dst[e], lo, hi, symmetric);
}

void or_transpose(uint32_t V, const uint64_t* off, const uint32_t* dst, const uint32_t* w,
                  uint64_t* in_off, uint32_t* in_src, uint32_t* in_w) {
  memset(in_off, 0, ((uint64_t)V + 1) * sizeof(uint64_t));
  const uint64_t E = off[V];
  for (uint64_t e = 0; e < E; ++e) in_off[dst[e] + 1]++;
  for (uint32_t v = 0; v < V; ++v) in_off[v + 1] += in_off[v];
  uint64_t* pos = (uint64_t*)malloc(((uint64_t)V + 1) * sizeof(uint64_t));
  memcpy(pos, in_off, ((uint64_t)V + 1) * sizeof(uint64_t));
  /* sources visited in ascending order, so every in-list is sorted */
  for (uint32_t u = 0; u < V; ++u)
    for (uint64_t e = off[u]; e < off[u + 1]; ++e) {
      uint64_t p = pos[dst[e]]++;
      in_src[p] = u;
      if (w && in_w) in_w[p] = w[e];
    }
  free(pos);
}

/* ------------------------------------------------------------------------ */

void or_bfs(uint32_t V, const uint64_t* off, const uint32_t* dst, uint32_t root, uint32_t* out) {
  for (uint32_t v = 0; v < V; ++v) out[v] = OR_UNREACHED;
  if (root >= V) return;
  uint32_t* q = (uint32_t*)malloc((uint64_t)V * sizeof(uint32_t));
  uint64_t h = 0, t = 0;
  out[root] = 0;
  q[t++] = root;
  while (h < t) {
    uint32_t u = q[h++];
    uint32_t lv = out[u] + 1;
    if (lv >= OR_UNREACHED) continue;
    for (uint64_t e = off[u]; e < off[u + 1]; ++e) {
      uint32_t v = dst[e];
      if (out[v] == OR_UNREACHED) { out[v] = lv; q[t++] = v; }
    }
  }
  free(q);
}

/* Binary min-heap of (dist, vertex) with lazy deletion. */
typedef struct { uint64_t d; uint32_t v; } hent;
static void heap_push(hent* h, uint64_t* n, hent x) {
  uint64_t i = (*n)++;
  while (i > 0) {
    uint64_t p = (i - 1) / 2;
    if (h[p].d <= x.d) break;
    h[i] = h[p]; i = p;
  }
  h[i] = x;
}
static hent heap_pop(hent* h, uint64_t* n) {
  hent top = h[0], x = h[--(*n)];
  uint64_t i = 0;
  for (;;) {
    uint64_t l = 2 * i + 1, r = l + 1, m = i;
    uint64_t md = x.d;
    if (l < *n && h[l].d < md) { m = l; md = h[l].d; }
    if (r < *n && h[r].d < md) { m = r; }
    if (m == i) break;
    h[i] = h[m]; i = m;
  }
  h[i] = x;
  return top;
}

void or_dijkstra(uint32_t V, const uint64_t* off, const uint32_t* dst, const uint32_t* w,
                 uint32_t root, uint32_t* out) {
  for (uint32_t v = 0; v < V; ++v) out[v] = OR_UNREACHED;
  if (root >= V) return;
  uint64_t E = off[V];
  uint64_t* dist = (uint64_t*)malloc((uint64_t)V * sizeof(uint64_t));
  for (uint32_t v = 0; v < V; ++v) dist[v] = UINT64_MAX;
  hent* h = (hent*)malloc((E + 1) * sizeof(hent));
  uint64_t n = 0;
  dist[root] = 0;
  heap_push(h, &n, (hent){0, root});
  while (n) {
    hent x = heap_pop(h, &n);
    if (x.d != dist[x.v]) continue;
    for (uint64_t e = off[x.v]; e < off[x.v + 1]; ++e) {
      uint64_t nd = x.d + (w ? w[e] : 1);
      uint32_t v = dst[e];
      if (nd < dist[v]) { dist[v] = nd; heap_push(h, &n, (hent){nd, v}); }
    }
  }
  /* A true distance >= the sentinel is never written by the reference's
   * strict-less relaxation, so it reads as unreached. */
  for (uint32_t v = 0; v < V; ++v) out[v] = dist[v] < OR_UNREACHED ? (uint32_t)dist[v] : OR_UNREACHED;
  free(h);
  free(dist);
}

static uint32_t uf_find(uint32_t* p, uint32_t x) {
  while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
  return x;
}

void or_cc_undirected(uint32_t V, const uint64_t* off, const uint32_t* dst, uint32_t* out) {
  for (uint32_t v = 0; v < V; ++v) out[v] = v;
  for (uint32_t u = 0; u < V; ++u)
    for (uint64_t e = off[u]; e < off[u + 1]; ++e) {
      uint32_t a = uf_find(out, u), b = uf_find(out, dst[e]);
      if (a == b) continue;
      if (a < b) out[b] = a; else out[a] = b; /* the smaller id stays root */
    }
  for (uint32_t v = 0; v < V; ++v) out[v] = uf_find(out, v);
}

void or_cc_directed(uint32_t V, const uint64_t* off, const uint32_t* dst, uint32_t* out) {
  /* Sources in ascending id order; a vertex keeps the first (smallest) source
   * that reaches it.  Anything reachable from an already-labelled vertex was
   * reached by that smaller source too, so labelled vertices are not re-entered. */
  for (uint32_t v = 0; v < V; ++v) out[v] = OR_UNREACHED;
  uint32_t* q = (uint32_t*)malloc((uint64_t)(V ? V : 1) * sizeof(uint32_t));
  for (uint32_t s = 0; s < V; ++s) {
    if (out[s] != OR_UNREACHED) continue;
    uint64_t h = 0, t = 0;
    out[s] = s; q[t++] = s;
    while (h < t) {
      uint32_t u = q[h++];
      for (uint64_t e = off[u]; e < off[u + 1]; ++e) {
        uint32_t v = dst[e];
        if (out[v] == OR_UNREACHED) { out[v] = s; q[t++] = v; }
      }
    }
  }
  free(q);
}

void or_pagerank64(uint32_t V, const uint64_t* off, const uint32_t* dst, int iters,
                   double damping, double* out) {
  double* nxt = (double*)malloc((uint64_t)(V ? V : 1) * sizeof(double));
  for (uint32_t v = 0; v < V; ++v) out[v] = 1.0 / V;
  const double base = (1.0 - damping) / V;
  for (int it = 0; it < iters; ++it) {
    for (uint32_t v = 0; v < V; ++v) nxt[v] = 0.0;
    for (uint32_t u = 0; u < V; ++u) {
      uint64_t deg = off[u + 1] - off[u];
      if (!deg) continue; /* dangling mass is dropped */
      double c = out[u] / (double)deg;
      for (uint64_t e = off[u]; e < off[u + 1]; ++e) nxt[dst[e]] += c;
    }
    for (uint32_t v = 0; v < V; ++v) out[v] = base + damping * nxt[v];
  }
  free(nxt);
}

void or_pagerank_step(uint32_t V, const uint64_t* off, const uint32_t* dst, const float* in,
                      double damping, double* out) {
  const double base = (1.0 - damping) / V;
  for (uint32_t v = 0; v < V; ++v) out[v] = 0.0;
  for (uint32_t u = 0; u < V; ++u) {
    uint64_t deg = off[u + 1] - off[u];
    if (!deg) continue;
    double c = (double)in[u] / (double)deg;
    for (uint64_t e = off[u]; e < off[u + 1]; ++e) out[dst[e]] += c;
  }
  for (uint32_t v = 0; v < V; ++v) out[v] = base + damping * out[v];
}

int64_t or_check_sssp_cert(uint32_t V, const uint64_t* off, const uint32_t* dst,
                           const uint32_t* w, uint32_t root, const uint32_t* dist) {
  if (root >= V) return 1;
  if (dist[root] != 0) return 1 + (int64_t)root;
  uint8_t* tight = (uint8_t*)calloc(V ? V : 1, 1);
  tight[root] = 1;
  for (uint32_t u = 0; u < V; ++u) {
    if (dist[u] >= OR_UNREACHED) continue;
    for (uint64_t e = off[u]; e < off[u + 1]; ++e) {
      uint64_t nd = (uint64_t)dist[u] + (w ? w[e] : 1);
      uint32_t v = dst[e];
      if (nd >= OR_UNREACHED) continue;
      if (nd < dist[v]) { free(tight); return 1 + (int64_t)V; }
      if (nd == dist[v]) tight[v] = 1;
    }
  }
  for (uint32_t v = 0; v < V; ++v)
    if (dist[v] < OR_UNREACHED && !tight[v]) { free(tight); return 1 + (int64_t)v; }
  /* With positive weights a tight parent has a strictly smaller distance, so
   * tight edges form a forest rooted at `root`.  With zero-weight edges this
   * certificate is necessary but not sufficient; callers use it on w >= 1. */
  free(tight);
  return 0;
}

uint64_t or_reached_edges(uint32_t V, const uint64_t* off, const uint32_t* val) {
  uint64_t s = 0;
  for (uint32_t v = 0; v < V; ++v)
    if (val[v] != OR_UNREACHED) s += off[v + 1] - off[v];
  return s;
}

uint32_t or_max_outdeg_vertex(uint32_t V, const uint64_t* off) {
  uint32_t best = 0;
  uint64_t bd = 0;
  for (uint32_t v = 0; v < V; ++v) {
    uint64_t d = off[v + 1] - off[v];
    if (d > bd) { bd = d; best = v; }
  }
  return best;
}
