/*
 * gv_oracle.c -- host-sequential CPU restatement of GraphVine's intended
 * semantics (see gv_oracle.h for the contract).
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker, never the thing measured or
 * shipped.  Plain C99, single-threaded, FP64 analytics.
 *
 * Store: one sorted, duplicate-free uint32 array per vertex.  A batch is
 * turned into sorted unique (src,dst) keys and merged per source segment
 * (set union for insert, set difference for delete), which is the set algebra
 * the reference intends in GraphVine::batchInsert / batchDelete
 * (graphvine.cu:222-451) once its defects D3-D7 are taken out.
 */
#include "gv_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define GOLDEN 0x9E3779B97F4A7C15ULL

/* ---------------------------------------------------------------- RNG --- */

/* SplitMix64 finaliser (Steele, Lea, Flood 2014). */
uint64_t orc_fmix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

/* Key of an independent stream: mix(mix(seed) ^ stream * odd constant). */
uint64_t orc_stream_key(uint64_t seed, uint64_t stream) {
  return orc_fmix64(orc_fmix64(seed) ^ (stream * 0xD6E8FEB86659FD93ULL));
}

/* ctr-th output of a SplitMix64 generator whose state starts at key. */
uint64_t orc_rand64(uint64_t key, uint64_t ctr) {
  return orc_fmix64(key + (ctr + 1) * GOLDEN);
}

static inline uint32_t mulhi32(uint32_t r, uint32_t n) {
  return (uint32_t)(((uint64_t)r * (uint64_t)n) >> 32);
}

/* R-MAT: edge e draws one 64-bit value per level l (counter e*scale+l) and
 * compares its high 32 bits with the cumulative quadrant thresholds
 * ta=a, tab=a+b, tabc=a+b+c (each scaled by 2^32).  Level 0 sets the most
 * significant bit.  Quadrants: a=(0,0) b=(0,1) c=(1,0) d=(1,1).  No noise, no
 * vertex permutation: hubs sit at low IDs. */
void orc_gen_rmat(uint64_t key, int scale, uint64_t n, uint32_t ta, uint32_t tab,
                  uint32_t tabc, uint32_t *src, uint32_t *dst) {
  for (uint64_t e = 0; e < n; ++e) {
    uint32_t u = 0, v = 0;
    for (int l = 0; l < scale; ++l) {
      uint32_t r = (uint32_t)(orc_rand64(key, e * (uint64_t)scale + (uint64_t)l) >> 32);
      uint32_t bit = 1u << (scale - 1 - l);
      if (r < ta) {
      } else if (r < tab) {
        v |= bit;
      } else if (r < tabc) {
        u |= bit;
      } else {
        u |= bit;
        v |= bit;
      }
    }
    src[e] = u;
    dst[e] = v;
  }
}

/* Uniform pairs: u = hi32*V>>32, v = lo32*V>>32; a self-loop moves v to v+1
 * (mod V), so every pair has u != v when V > 1. */
void orc_gen_uniform(uint64_t key, uint32_t V, uint64_t n, uint32_t *src, uint32_t *dst) {
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t r = orc_rand64(key, i);
    uint32_t u = mulhi32((uint32_t)(r >> 32), V);
    uint32_t v = mulhi32((uint32_t)r, V);
    if (u == v && V > 1) v = (v + 1 == V) ? 0 : v + 1;
    src[i] = u;
    dst[i] = v;
  }
}

/* Sample n edges (with replacement) from a CSR: index = rand64 % E. */
void orc_gen_sample(uint64_t key, uint32_t V, const uint64_t *offsets, const uint32_t *cols,
                    uint64_t n, uint32_t *src, uint32_t *dst) {
  uint64_t E = offsets[V];
  for (uint64_t i = 0; i < n; ++i) {
    if (E == 0) {
      src[i] = dst[i] = 0;
      continue;
    }
    uint64_t idx = orc_rand64(key, i) % E;
    /* largest u with offsets[u] <= idx */
    uint32_t lo = 0, hi = V;
    while (hi - lo > 1) {
      uint32_t mid = lo + (hi - lo) / 2;
      if (offsets[mid] <= idx) lo = mid; else hi = mid;
    }
    src[i] = lo;
    dst[i] = cols[idx];
  }
}

/* Chung-Lu: weight w_i = (i+1)^(-1/(gamma-1)), scaled so the mean expected
 * degree is avg_deg, capped at max_deg.  Both endpoints are drawn with
 * probability proportional to w by inverse CDF over a 2^62-scaled integer
 * cumulative table (binary search), so host and device agree bit for bit
 * when they share the table construction. */
void orc_gen_chung_lu(uint64_t key, uint32_t V, double gamma, double avg_deg, double max_deg,
                      uint64_t n, uint32_t *src, uint32_t *dst) {
  double *w = (double *)malloc(sizeof(double) * V);
  uint64_t *cum = (uint64_t *)malloc(sizeof(uint64_t) * ((size_t)V + 1));
  double ex = -1.0 / (gamma - 1.0), s = 0.0;
  for (uint32_t i = 0; i < V; ++i) { w[i] = pow((double)i + 1.0, ex); s += w[i]; }
  double scale = avg_deg * (double)V / s, tot = 0.0;
  for (uint32_t i = 0; i < V; ++i) {
    w[i] *= scale;
    if (w[i] > max_deg) w[i] = max_deg;
    tot += w[i];
  }
  double acc = 0.0;
  cum[0] = 0;
  for (uint32_t i = 0; i < V; ++i) {
    acc += w[i];
    cum[i + 1] = (uint64_t)(acc / tot * 4611686018427387904.0); /* 2^62 */
  }
  cum[V] = 4611686018427387904ULL;
  for (uint64_t i = 0; i < 2 * n; ++i) {
    uint64_t r = orc_rand64(key, i) >> 2; /* uniform in [0, 2^62) */
    uint32_t lo = 0, hi = V; /* largest lo with cum[lo] <= r */
    while (hi - lo > 1) {
      uint32_t mid = lo + (hi - lo) / 2;
      if (cum[mid] <= r) lo = mid; else hi = mid;
    }
    if (i & 1) dst[i >> 1] = lo; else src[i >> 1] = lo;
  }
  free(w);
  free(cum);
}

/* -------------------------------------------------------------- store --- */

typedef struct {
  uint32_t *a;
  uint32_t n, cap;
} adjv;

struct orc_graph {
  uint32_t V;
  int directed;
  adjv *out;
};

orc_graph *orc_create(uint32_t V, int directed) {
  orc_graph *g = (orc_graph *)calloc(1, sizeof(orc_graph));
  g->V = V;
  g->directed = directed;
  g->out = (adjv *)calloc(V ? V : 1, sizeof(adjv));
  return g;
}

void orc_destroy(orc_graph *g) {
  if (!g) return;
  for (uint32_t i = 0; i < g->V; ++i) free(g->out[i].a);
  free(g->out);
  free(g);
}

uint32_t orc_num_vertices(const orc_graph *g) { return g->V; }

uint64_t orc_num_edges(const orc_graph *g) {
  uint64_t s = 0;
  for (uint32_t i = 0; i < g->V; ++i) s += g->out[i].n;
  return s;
}

void orc_degrees(const orc_graph *g, uint32_t *deg) {
  for (uint32_t i = 0; i < g->V; ++i) deg[i] = g->out[i].n;
}

void orc_export_csr(const orc_graph *g, uint64_t *offsets, uint32_t *cols) {
  uint64_t o = 0;
  for (uint32_t i = 0; i < g->V; ++i) {
    offsets[i] = o;
    if (cols && g->out[i].n) memcpy(cols + o, g->out[i].a, sizeof(uint32_t) * g->out[i].n);
    o += g->out[i].n;
  }
  offsets[g->V] = o;
}

static void radix_sort_u64(uint64_t *a, uint64_t n) {
  uint64_t *tmp = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
  for (int sh = 0; sh < 64; sh += 16) {
    uint64_t cnt[65537];
    memset(cnt, 0, sizeof(cnt));
    for (uint64_t i = 0; i < n; ++i) cnt[((a[i] >> sh) & 0xFFFF) + 1]++;
    for (int d = 0; d < 65536; ++d) cnt[d + 1] += cnt[d];
    for (uint64_t i = 0; i < n; ++i) tmp[cnt[(a[i] >> sh) & 0xFFFF]++] = a[i];
    memcpy(a, tmp, sizeof(uint64_t) * n);
  }
  free(tmp);
}

/* Turn a batch into sorted unique directed keys (src<<32|dst).  Returns -1 if
 * any ID is out of range (the batch is then rejected as a whole). Self-loops
 * are dropped (intent of insert_kernels.cu:171-177, without defect D4). */
static int64_t stage_batch(const orc_graph *g, const uint32_t *src, const uint32_t *dst,
                           uint64_t n, uint64_t **keys_out) {
  for (uint64_t i = 0; i < n; ++i)
    if (src[i] >= g->V || dst[i] >= g->V) return -1;
  uint64_t cap = g->directed ? n : 2 * n;
  uint64_t *k = (uint64_t *)malloc(sizeof(uint64_t) * (cap ? cap : 1));
  uint64_t m = 0;
  for (uint64_t i = 0; i < n; ++i) {
    if (src[i] == dst[i]) continue;
    k[m++] = ((uint64_t)src[i] << 32) | dst[i];
    if (!g->directed) k[m++] = ((uint64_t)dst[i] << 32) | src[i];
  }
  radix_sort_u64(k, m);
  uint64_t w = 0;
  for (uint64_t i = 0; i < m; ++i)
    if (i == 0 || k[i] != k[i - 1]) k[w++] = k[i];
  *keys_out = k;
  return (int64_t)w;
}

static void adj_reserve(adjv *a, uint32_t need) {
  if (need <= a->cap) return;
  uint32_t c = a->cap ? a->cap : 4;
  while (c < need) c *= 2;
  a->a = (uint32_t *)realloc(a->a, sizeof(uint32_t) * c);
  a->cap = c;
}

/* N(u) |= B_u (both sorted, unique); returns the number added. */
static uint64_t merge_union(adjv *a, const uint64_t *k, uint64_t m) {
  uint32_t *old = a->a;
  uint32_t n = a->n;
  uint32_t *out = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)n + m));
  uint64_t i = 0, j = 0, o = 0;
  while (i < n || j < m) {
    uint32_t bj = j < m ? (uint32_t)k[j] : 0;
    if (j >= m || (i < n && old[i] < bj)) out[o++] = old[i++];
    else if (i >= n || bj < old[i]) { out[o++] = bj; ++j; }
    else { out[o++] = old[i++]; ++j; }
  }
  uint64_t added = o - n;
  free(old);
  a->a = out;
  a->n = (uint32_t)o;
  a->cap = (uint32_t)((size_t)n + m);
  return added;
}

/* N(u) \= B_u; returns the number removed. */
static uint64_t merge_diff(adjv *a, const uint64_t *k, uint64_t m) {
  uint32_t n = a->n, w = 0;
  uint64_t j = 0;
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t x = a->a[i];
    while (j < m && (uint32_t)k[j] < x) ++j;
    if (j < m && (uint32_t)k[j] == x) continue;
    a->a[w++] = x;
  }
  uint64_t removed = n - w;
  a->n = w;
  return removed;
}

static int apply(orc_graph *g, const uint32_t *src, const uint32_t *dst, uint64_t n, int del,
                 uint64_t *count) {
  uint64_t *k = NULL;
  int64_t m = stage_batch(g, src, dst, n, &k);
  if (m < 0) return -1;
  uint64_t c = 0;
  for (int64_t s = 0; s < m;) {
    uint32_t u = (uint32_t)(k[s] >> 32);
    int64_t e = s;
    while (e < m && (uint32_t)(k[e] >> 32) == u) ++e;
    if (del) c += merge_diff(&g->out[u], k + s, (uint64_t)(e - s));
    else {
      adj_reserve(&g->out[u], g->out[u].n);
      c += merge_union(&g->out[u], k + s, (uint64_t)(e - s));
    }
    s = e;
  }
  free(k);
  if (count) *count = c;
  return 0;
}

/* GraphVine::bulkBuild (graphvine.cu:181-220): the graph starts empty. */
int orc_bulk_build(orc_graph *g, const uint32_t *src, const uint32_t *dst, uint64_t n,
                   uint64_t *inserted) {
  for (uint32_t i = 0; i < g->V; ++i) g->out[i].n = 0;
  return apply(g, src, dst, n, 0, inserted);
}

/* GraphVine::batchInsert (graphvine.cu:222-371): N(u) |= {v}. */
int orc_batch_insert(orc_graph *g, const uint32_t *src, const uint32_t *dst, uint64_t n,
                     uint64_t *inserted) {
  return apply(g, src, dst, n, 0, inserted);
}

/* GraphVine::batchDelete (graphvine.cu:373-451) + batch_delete_VC
 * (delete_kernels.cu:3-48): N(u) \= {v}; missing edges are no-ops. */
int orc_batch_delete(orc_graph *g, const uint32_t *src, const uint32_t *dst, uint64_t n,
                     uint64_t *deleted) {
  return apply(g, src, dst, n, 1, deleted);
}

/* ---------------------------------------------------------- analytics --- */

/* static_pagerank (pagerank.cu:17-59) with pageRank_kernel_HD and
 * pagerank_post_kernel (pr_kernel.cu:3-48), in FP64:
 *   x0 = 1/V; x'[v] = (1-a)/V + a * sum_{u: v in N(u)} x[u]/|N(u)|;
 *   r = sum |x'-x|; stop after the first update with r < eps or after
 *   max_iter updates.  Dangling mass is dropped (pr_kernel.cu:40).
 * x0 (may be NULL) replaces the 1/V start: the warm start of
 * dynamic_pagerank (pagerank.cu:70-132), which iterates the same update
 * from the previous vector (its affected-set heuristic, pr_kernel.cu:88, is
 * not reproduced).  fixed_iters > 0 runs exactly that many updates (replay
 * of a GPU run).  Returns the number of updates performed. */
int orc_pagerank(const orc_graph *g, double alpha, double eps, int max_iter, int fixed_iters,
                 const double *x0, double *x, double *residual) {
  uint32_t V = g->V;
  double *acc = (double *)malloc(sizeof(double) * (V ? V : 1));
  for (uint32_t i = 0; i < V; ++i) x[i] = x0 ? x0[i] : 1.0 / (double)V;
  double base = (1.0 - alpha) / (double)V, r = 0.0;
  int limit = fixed_iters > 0 ? fixed_iters : max_iter, it = 0;
  while (it < limit) {
    memset(acc, 0, sizeof(double) * V);
    for (uint32_t u = 0; u < V; ++u) {
      uint32_t d = g->out[u].n;
      if (!d) continue;
      double c = x[u] / (double)d;
      for (uint32_t j = 0; j < d; ++j) acc[g->out[u].a[j]] += c;
    }
    r = 0.0;
    for (uint32_t v = 0; v < V; ++v) {
      double xn = base + alpha * acc[v];
      r += fabs(xn - x[v]);
      x[v] = xn;
    }
    ++it;
    if (fixed_iters <= 0 && r < eps) break;
  }
  if (residual) *residual = r;
  free(acc);
  return it;
}

/* Personalized PageRank (static_pagerank's iteration with the teleport term
 * on one seed per column): x_c0 = e_{s_c};
 *   x_c'[v] = (1-a)[v == s_c] + a * sum_{u: v in N(u)} x_c[u]/|N(u)|,
 * dangling mass dropped; stop after the first update whose largest
 * per-column L1 residual is below eps (or max_iter; fixed_iters > 0 runs
 * exactly that many).  x is row-major V x ncols, FP64.  Returns updates. */
int orc_pagerank_personalized(const orc_graph *g, double alpha, double eps, int max_iter,
                              int fixed_iters, const uint32_t *seeds, uint32_t ncols, double *x,
                              double *max_residual) {
  uint32_t V = g->V;
  size_t n = (size_t)V * ncols;
  double *acc = (double *)malloc(sizeof(double) * (n ? n : 1));
  double *r = (double *)malloc(sizeof(double) * (ncols ? ncols : 1));
  /* x_c[u]/|N(u)| of the current source row, computed once per row (the
   * same quotients every edge of the row would compute) */
  double *q = (double *)malloc(sizeof(double) * (ncols ? ncols : 1));
  for (size_t i = 0; i < n; ++i) x[i] = 0.0;
  for (uint32_t c = 0; c < ncols; ++c) x[(size_t)seeds[c] * ncols + c] = 1.0;
  int limit = fixed_iters > 0 ? fixed_iters : max_iter, it = 0;
  double mr = 0.0;
  while (it < limit) {
    memset(acc, 0, sizeof(double) * n);
    for (uint32_t u = 0; u < V; ++u) {
      uint32_t d = g->out[u].n;
      if (!d) continue;
      for (uint32_t c = 0; c < ncols; ++c) q[c] = x[(size_t)u * ncols + c] / (double)d;
      for (uint32_t j = 0; j < d; ++j) {
        double *av = acc + (size_t)g->out[u].a[j] * ncols;
        for (uint32_t c = 0; c < ncols; ++c) av[c] += q[c];
      }
    }
    for (uint32_t c = 0; c < ncols; ++c) r[c] = 0.0;
    for (uint32_t v = 0; v < V; ++v)
      for (uint32_t c = 0; c < ncols; ++c) {
        size_t i = (size_t)v * ncols + c;
        double xn = alpha * acc[i] + (seeds[c] == v ? 1.0 - alpha : 0.0);
        r[c] += fabs(xn - x[i]);
        x[i] = xn;
      }
    mr = 0.0;
    for (uint32_t c = 0; c < ncols; ++c) mr = r[c] > mr ? r[c] : mr;
    ++it;
    if (fixed_iters <= 0 && mr < eps) break;
  }
  if (max_residual) *max_residual = mr;
  free(acc);
  free(r);
  free(q);
  return it;
}

static int has_edge(const adjv *a, uint32_t x) {
  uint32_t lo = 0, hi = a->n;
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2;
    if (a->a[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo < a->n && a->a[lo] == x;
}

/* tc_kernel_VC (tc_kernel.cu:27-73), literal definition with exact integer
 * counts (the reference's FP32 counts, defect D2, are not reproduced):
 *   t[u] = #{(v,w) : u<v<w, v in N(u), w in N(v), u in N(w)}. */
uint64_t orc_triangle_count(const orc_graph *g, uint64_t *per_vertex) {
  uint64_t total = 0;
  for (uint32_t u = 0; u < g->V; ++u) {
    uint64_t c = 0;
    const adjv *nu = &g->out[u];
    for (uint32_t i = 0; i < nu->n; ++i) {
      uint32_t v = nu->a[i];
      if (v <= u) continue;
      const adjv *nv = &g->out[v];
      for (uint32_t j = 0; j < nv->n; ++j) {
        uint32_t w = nv->a[j];
        if (w <= v) continue;
        if (has_edge(&g->out[w], u)) ++c;
      }
    }
    if (per_vertex) per_vertex[u] = c;
    total += c;
  }
  return total;
}

/* Brandes betweenness (intent of bc.cu:42-80 and bc_kernel.cu:19-119) over
 * the directed adjacency, for the given sources, FP64 sigma and delta:
 *   delta_s(v) = sum_{w in N(v), d(w)=d(v)+1} sigma_v/sigma_w (1+delta_s(w)),
 *   BC[v] = sum_{s, v != s} delta_s(v).  Unnormalised, not halved. */
void orc_bc(const orc_graph *g, const uint32_t *sources, uint32_t n_sources, double *bc) {
  uint32_t V = g->V;
  int64_t *dist = (int64_t *)malloc(sizeof(int64_t) * (V ? V : 1));
  double *sigma = (double *)malloc(sizeof(double) * (V ? V : 1));
  double *delta = (double *)malloc(sizeof(double) * (V ? V : 1));
  uint32_t *order = (uint32_t *)malloc(sizeof(uint32_t) * (V ? V : 1));
  for (uint32_t i = 0; i < V; ++i) bc[i] = 0.0;
  for (uint32_t si = 0; si < n_sources; ++si) {
    uint32_t s = sources[si];
    for (uint32_t i = 0; i < V; ++i) { dist[i] = -1; sigma[i] = 0.0; delta[i] = 0.0; }
    dist[s] = 0;
    sigma[s] = 1.0;
    uint32_t head = 0, tail = 0;
    order[tail++] = s;
    while (head < tail) {
      uint32_t u = order[head++];
      const adjv *nu = &g->out[u];
      for (uint32_t j = 0; j < nu->n; ++j) {
        uint32_t w = nu->a[j];
        if (dist[w] < 0) { dist[w] = dist[u] + 1; order[tail++] = w; }
        if (dist[w] == dist[u] + 1) sigma[w] += sigma[u];
      }
    }
    for (uint32_t k = tail; k-- > 0;) {
      uint32_t v = order[k];
      const adjv *nv = &g->out[v];
      double d = 0.0;
      for (uint32_t j = 0; j < nv->n; ++j) {
        uint32_t w = nv->a[j];
        if (dist[w] == dist[v] + 1) d += sigma[v] / sigma[w] * (1.0 + delta[w]);
      }
      delta[v] = d;
      if (v != s) bc[v] += d;
    }
  }
  free(dist);
  free(sigma);
  free(delta);
  free(order);
}
