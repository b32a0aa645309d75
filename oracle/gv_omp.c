/*
 * gv_omp.c -- OpenMP CSR baseline (CPU timing reference only).
 *
 * TEST/BENCH INFRASTRUCTURE ONLY: used by bench.py's cpu_baseline leg and by
 * tests that check it against the sequential oracle.  It computes the same
 * set semantics as gv_oracle.c on an immutable sorted CSR that is rebuilt per
 * update batch by a parallel sort-merge (the usual static-graph answer to a
 * dynamic workload), plus pull PageRank (symmetric and directed) and
 * Brandes BC parallel over sources, all in FP64: they are the FP64 references
 * of the precision checks and the CPU timings of the analytics.
 */
#include <omp.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static uint32_t id_bits(uint32_t V) {
  uint32_t b = 0;
  while ((1ull << b) < V) ++b;
  return b ? b : 1;
}

/* Parallel LSD radix sort (8-bit digits) of n keys on their low `bits` bits;
 * the result ends in a. */
static void psort(uint64_t *a, uint64_t *t, uint64_t n, uint32_t bits) {
  int nt = omp_get_max_threads();
  uint64_t *hist = (uint64_t *)malloc(sizeof(uint64_t) * 256 * (size_t)nt);
  uint64_t *src = a, *dst = t;
  for (uint32_t sh = 0; sh < bits; sh += 8) {
#pragma omp parallel num_threads(nt)
    {
      int tid = omp_get_thread_num();
      uint64_t lo = n * (uint64_t)tid / nt, hi = n * (uint64_t)(tid + 1) / nt;
      uint64_t *h = hist + 256 * (size_t)tid;
      memset(h, 0, 256 * sizeof(uint64_t));
      for (uint64_t i = lo; i < hi; ++i) h[(src[i] >> sh) & 255]++;
#pragma omp barrier
#pragma omp single
      {
        uint64_t s = 0;
        for (int d = 0; d < 256; ++d)
          for (int k = 0; k < nt; ++k) {
            uint64_t c = hist[256 * (size_t)k + d];
            hist[256 * (size_t)k + d] = s;
            s += c;
          }
      }
      for (uint64_t i = lo; i < hi; ++i) dst[h[(src[i] >> sh) & 255]++] = src[i];
    }
    uint64_t *x = src;
    src = dst;
    dst = x;
  }
  if (src != a) memcpy(a, src, n * sizeof(uint64_t));
  free(hist);
}

/* Stage (src,dst) pairs into sorted unique keys; mirror adds (dst,src). */
static uint64_t stage(uint32_t V, int directed, const uint32_t *src, const uint32_t *dst, uint64_t n,
                      uint64_t **out) {
  uint32_t bV = id_bits(V);
  uint64_t m = directed ? n : 2 * n;
  uint64_t *k = (uint64_t *)malloc(sizeof(uint64_t) * (m ? m : 1));
  uint64_t *t = (uint64_t *)malloc(sizeof(uint64_t) * (m ? m : 1));
  const uint64_t SENT = ~0ull;
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t u = src[i], v = dst[i];
    int drop = u == v || u >= V || v >= V;
    if (directed) k[i] = drop ? SENT : (((uint64_t)u << bV) | v);
    else {
      k[2 * i] = drop ? SENT : (((uint64_t)u << bV) | v);
      k[2 * i + 1] = drop ? SENT : (((uint64_t)v << bV) | u);
    }
  }
  psort(k, t, m, 2 * bV);
  free(t);
  /* unique + drop sentinels (sequential; batches are small next to E) */
  uint64_t w = 0;
  for (uint64_t i = 0; i < m; ++i)
    if (k[i] != SENT && (w == 0 || k[w - 1] != k[i])) k[w++] = k[i];
  *out = k;
  return w;
}

static uint64_t lower_src(const uint64_t *k, uint64_t m, uint64_t key) {
  uint64_t lo = 0, hi = m;
  while (lo < hi) {
    uint64_t mid = lo + (hi - lo) / 2;
    if (k[mid] < key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

/* New CSR = (old U ins) \ del per vertex, deletes applied after inserts
 * (callers pass one kind per call to keep batch order).  off_out must hold
 * V+1 entries, cols_out enough for the result (old E + inserts).
 * Returns the new edge count. */
uint64_t omp_apply_batch(uint32_t V, int directed, const uint64_t *off, const uint32_t *cols,
                         const uint32_t *src, const uint32_t *dst, uint64_t n, int del,
                         uint64_t *off_out, uint32_t *cols_out) {
  uint32_t bV = id_bits(V);
  uint64_t *k = NULL;
  uint64_t m = stage(V, directed, src, dst, n, &k);
  uint64_t *cnt = (uint64_t *)malloc(sizeof(uint64_t) * ((size_t)V + 1));
  uint64_t mask = (1ull << bV) - 1;
#pragma omp parallel for schedule(dynamic, 1024)
  for (uint32_t u = 0; u < V; ++u) {
    uint64_t b0 = lower_src(k, m, (uint64_t)u << bV), b1 = lower_src(k, m, (uint64_t)(u + 1) << bV);
    uint64_t i = off[u], ie = off[u + 1], j = b0, c = 0;
    while (i < ie || j < b1) {
      uint32_t bj = j < b1 ? (uint32_t)(k[j] & mask) : 0;
      if (j >= b1 || (i < ie && cols[i] < bj)) { ++c; ++i; }
      else if (i >= ie || bj < cols[i]) { if (!del) ++c; ++j; }
      else { if (!del) ++c; ++i; ++j; }
    }
    cnt[u] = c;
  }
  uint64_t s = 0;
  for (uint32_t u = 0; u < V; ++u) { off_out[u] = s; s += cnt[u]; }
  off_out[V] = s;
#pragma omp parallel for schedule(dynamic, 1024)
  for (uint32_t u = 0; u < V; ++u) {
    uint64_t b0 = lower_src(k, m, (uint64_t)u << bV), b1 = lower_src(k, m, (uint64_t)(u + 1) << bV);
    uint64_t i = off[u], ie = off[u + 1], j = b0, o = off_out[u];
    while (i < ie || j < b1) {
      uint32_t bj = j < b1 ? (uint32_t)(k[j] & mask) : 0;
      if (j >= b1 || (i < ie && cols[i] < bj)) cols_out[o++] = cols[i++];
      else if (i >= ie || bj < cols[i]) { if (!del) cols_out[o++] = bj; ++j; }
      else { if (!del) cols_out[o++] = cols[i]; ++i; ++j; }
    }
  }
  free(cnt);
  free(k);
  return s;
}

/* Pull PageRank over a symmetric CSR (in = out), same rule as the oracle
 * (orc_pagerank: pagerank.cu:17-59); FP64.  x0 (may be NULL = 1/V) is the
 * start vector; fixed_iters > 0 runs exactly that many updates.  Returns the
 * number of updates; *residual (may be NULL) gets the last L1 residual. */
int omp_pagerank_sym(uint32_t V, const uint64_t *off, const uint32_t *cols, double alpha, double eps,
                     int max_iter, int fixed_iters, const double *x0, double *x,
                     double *residual) {
  double *c = (double *)malloc(sizeof(double) * (V ? V : 1));
  double *xn = (double *)malloc(sizeof(double) * (V ? V : 1));
  for (uint32_t v = 0; v < V; ++v) x[v] = x0 ? x0[v] : 1.0 / V;
  double base = (1.0 - alpha) / V;
  const int limit = fixed_iters > 0 ? fixed_iters : max_iter;
  int it = 0;
  double rlast = 0.0;
  while (it < limit) {
#pragma omp parallel for schedule(static)
    for (uint32_t v = 0; v < V; ++v) {
      uint64_t d = off[v + 1] - off[v];
      c[v] = d ? x[v] / (double)d : 0.0;
    }
    double r = 0.0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : r)
    for (uint32_t v = 0; v < V; ++v) {
      double s = 0.0;
      for (uint64_t j = off[v]; j < off[v + 1]; ++j) s += c[cols[j]];
      xn[v] = base + alpha * s;
      r += fabs(xn[v] - x[v]);
    }
    memcpy(x, xn, sizeof(double) * V);
    ++it;
    rlast = r;
    if (fixed_iters <= 0 && r < eps) break;
  }
  if (residual) *residual = rlast;
  free(c);
  free(xn);
  return it;
}

/* Exact triangle count of an undirected graph given as a sorted symmetric
 * CSR, with the oracle's per-vertex definition (tc_kernel.cu:27-73):
 * t[u] = triangles whose minimum vertex ID is u.  Each triangle is found
 * once over degree-oriented lists (u -> v when (deg u, u) < (deg v, v)):
 * a thread marks N+(u) in its own byte map, then every w in N+(v) for
 * v in N+(u) that is marked closes {u, v, w}.  The count goes to
 * min(u, v, w); each thread keeps its own per-vertex array and the arrays
 * are added in thread order, so the result is exact and independent of
 * scheduling.  per_vertex may be NULL.  Returns the total. */
uint64_t omp_triangle_count(uint32_t V, const uint64_t *off, const uint32_t *cols,
                            uint64_t *per_vertex) {
  const int nt = omp_get_max_threads();
  const uint64_t E = off[V];
  uint64_t *poff = (uint64_t *)malloc(sizeof(uint64_t) * ((size_t)V + 1));
  uint32_t *pcols = (uint32_t *)malloc(sizeof(uint32_t) * (E ? E : 1));
#define GV_BEFORE(a, b)                                                                        \
  ((off[(a) + 1] - off[a]) < (off[(b) + 1] - off[b]) ||                                      \
   ((off[(a) + 1] - off[a]) == (off[(b) + 1] - off[b]) && (a) < (b)))
#pragma omp parallel for schedule(static)
  for (uint32_t u = 0; u < V; ++u) {
    uint64_t c = 0;
    for (uint64_t j = off[u]; j < off[u + 1]; ++j)
      if (GV_BEFORE(u, cols[j])) ++c;
    poff[u + 1] = c;
  }
  poff[0] = 0;
  for (uint32_t u = 0; u < V; ++u) poff[u + 1] += poff[u];
#pragma omp parallel for schedule(static)
  for (uint32_t u = 0; u < V; ++u) {
    uint64_t o = poff[u];
    for (uint64_t j = off[u]; j < off[u + 1]; ++j)
      if (GV_BEFORE(u, cols[j])) pcols[o++] = cols[j];
  }
#undef GV_BEFORE
  uint64_t *part = NULL;
  if (per_vertex) part = (uint64_t *)calloc((size_t)nt * (V ? V : 1), sizeof(uint64_t));
  uint64_t total = 0;
#pragma omp parallel num_threads(nt) reduction(+ : total)
  {
    uint64_t *t = per_vertex ? part + (size_t)omp_get_thread_num() * V : NULL;
    uint8_t *mark = (uint8_t *)calloc(V ? V : 1, 1);
#pragma omp for schedule(dynamic, 64)
    for (uint32_t u = 0; u < V; ++u) {
      for (uint64_t j = poff[u]; j < poff[u + 1]; ++j) mark[pcols[j]] = 1;
      for (uint64_t j = poff[u]; j < poff[u + 1]; ++j) {
        const uint32_t v = pcols[j];
        const uint32_t muv = u < v ? u : v;
        for (uint64_t k = poff[v]; k < poff[v + 1]; ++k) {
          const uint32_t w = pcols[k];
          if (mark[w]) {
            ++total;
            if (t) t[w < muv ? w : muv]++;
          }
        }
      }
      for (uint64_t j = poff[u]; j < poff[u + 1]; ++j) mark[pcols[j]] = 0;
    }
    free(mark);
  }
  if (per_vertex) {
#pragma omp parallel for schedule(static)
    for (uint32_t v = 0; v < V; ++v) {
      uint64_t s = 0;
      for (int k = 0; k < nt; ++k) s += part[(size_t)k * V + v];
      per_vertex[v] = s;
    }
    free(part);
  }
  free(poff);
  free(pcols);
  return total;
}

/* In-lists of a CSR of out-lists: in_off[V+1], in_cols[E], each in-list in
 * ascending source order (counting sort by destination). */
void omp_transpose(uint32_t V, const uint64_t *off, const uint32_t *cols, uint64_t *in_off,
                   uint32_t *in_cols) {
  const uint64_t E = off[V];
  memset(in_off, 0, sizeof(uint64_t) * ((size_t)V + 1));
  for (uint64_t j = 0; j < E; ++j) in_off[cols[j] + 1]++;
  for (uint32_t v = 0; v < V; ++v) in_off[v + 1] += in_off[v];
  uint64_t *pos = (uint64_t *)malloc(sizeof(uint64_t) * (V ? V : 1));
  memcpy(pos, in_off, sizeof(uint64_t) * V);
  for (uint32_t u = 0; u < V; ++u)
    for (uint64_t j = off[u]; j < off[u + 1]; ++j) in_cols[pos[cols[j]]++] = u;
  free(pos);
}

/* Pull PageRank on a directed graph (in-lists + out-degrees), FP64, the
 * oracle's rule: x0 = 1/V, dangling mass dropped, stop after the first update
 * whose L1 residual is below eps or after max_iter; fixed_iters > 0 runs
 * exactly that many updates.  Returns the number of updates. */
int omp_pagerank_dir(uint32_t V, const uint64_t *in_off, const uint32_t *in_cols,
                     const uint32_t *out_deg, double alpha, double eps, int max_iter,
                     int fixed_iters, double *x, double *residual) {
  double *c = (double *)malloc(sizeof(double) * (V ? V : 1));
  double *xn = (double *)malloc(sizeof(double) * (V ? V : 1));
  for (uint32_t v = 0; v < V; ++v) x[v] = 1.0 / V;
  const double base = (1.0 - alpha) / V;
  const int limit = fixed_iters > 0 ? fixed_iters : max_iter;
  int it = 0;
  double r = 0.0;
  while (it < limit) {
#pragma omp parallel for schedule(static)
    for (uint32_t v = 0; v < V; ++v) c[v] = out_deg[v] ? x[v] / (double)out_deg[v] : 0.0;
    r = 0.0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : r)
    for (uint32_t v = 0; v < V; ++v) {
      double s = 0.0;
      for (uint64_t j = in_off[v]; j < in_off[v + 1]; ++j) s += c[in_cols[j]];
      xn[v] = base + alpha * s;
      r += fabs(xn[v] - x[v]);
    }
    memcpy(x, xn, sizeof(double) * V);
    ++it;
    if (fixed_iters <= 0 && r < eps) break;
  }
  if (residual) *residual = r;
  free(c);
  free(xn);
  return it;
}

/* Brandes BC over a CSR of out-lists from the given sources (unnormalised,
 * not halved, BC[v] = sum over sources s != v of delta_s(v)), FP64 sigma and
 * delta.  Sources are split statically over the threads; each thread keeps
 * its own BC vector and the vectors are added in thread order. */
void omp_bc_dir(uint32_t V, const uint64_t *off, const uint32_t *cols, const uint32_t *src,
                uint32_t ns, double *bc) {
  const int nt = omp_get_max_threads();
  double *part = (double *)calloc((size_t)nt * (V ? V : 1), sizeof(double));
#pragma omp parallel num_threads(nt)
  {
    const int tid = omp_get_thread_num();
    double *b = part + (size_t)tid * (V ? V : 1);
    int *dist = (int *)malloc(sizeof(int) * (V ? V : 1));
    double *sigma = (double *)malloc(sizeof(double) * (V ? V : 1));
    double *delta = (double *)malloc(sizeof(double) * (V ? V : 1));
    uint32_t *q = (uint32_t *)malloc(sizeof(uint32_t) * (V ? V : 1));
    for (uint32_t k = (uint32_t)tid; k < ns; k += (uint32_t)nt) {
      const uint32_t s = src[k];
      for (uint32_t v = 0; v < V; ++v) {
        dist[v] = -1;
        sigma[v] = 0.0;
        delta[v] = 0.0;
      }
      uint64_t qh = 0, qt = 0;
      dist[s] = 0;
      sigma[s] = 1.0;
      q[qt++] = s;
      while (qh < qt) {
        const uint32_t u = q[qh++];
        for (uint64_t j = off[u]; j < off[u + 1]; ++j) {
          const uint32_t w = cols[j];
          if (dist[w] < 0) {
            dist[w] = dist[u] + 1;
            q[qt++] = w;
          }
          if (dist[w] == dist[u] + 1) sigma[w] += sigma[u];
        }
      }
      for (uint64_t i = qt; i-- > 0;) {
        const uint32_t v = q[i];
        double d = 0.0;
        for (uint64_t j = off[v]; j < off[v + 1]; ++j) {
          const uint32_t w = cols[j];
          if (dist[w] == dist[v] + 1) d += sigma[v] / sigma[w] * (1.0 + delta[w]);
        }
        delta[v] = d;
        if (v != s) b[v] += d;
      }
    }
    free(dist);
    free(sigma);
    free(delta);
    free(q);
  }
  for (uint32_t v = 0; v < V; ++v) {
    double t = 0.0;
    for (int k = 0; k < nt; ++k) t += part[(size_t)k * V + v];
    bc[v] = t;
  }
  free(part);
}

/* Personalized PageRank of the out-adjacency CSR, ncols columns (seed s_c
 * each), the oracle's rule and arithmetic (orc_pagerank_personalized,
 * gv_oracle.c): push x_c[u]/|N(u)| along every out-edge in (u, j) order,
 * then x_c' = a*acc + (1-a)[v == s_c], residual summed over v in order, stop
 * on the largest per-column residual below eps.  Columns are independent, so
 * threads take whole columns (held column-major, one V-array each) and every
 * column sees exactly the oracle's operations in the oracle's order: the
 * result equals the oracle's bit for bit.  x is row-major V x ncols. */
int omp_pagerank_personalized(uint32_t V, const uint64_t *off, const uint32_t *cols,
                              double alpha, double eps, int max_iter, int fixed_iters,
                              const uint32_t *seeds, uint32_t ncols, double *x,
                              double *max_residual) {
  size_t n = (size_t)V * ncols;
  double *xc = (double *)malloc(sizeof(double) * (n ? n : 1));
  double *acc = (double *)malloc(sizeof(double) * (n ? n : 1));
  double *r = (double *)malloc(sizeof(double) * (ncols ? ncols : 1));
  int limit = fixed_iters > 0 ? fixed_iters : max_iter, it = 0;
  double mr = 0.0;
#pragma omp parallel for schedule(static, 1)
  for (uint32_t c = 0; c < ncols; ++c) {
    double *xv = xc + (size_t)c * V;
    for (uint32_t v = 0; v < V; ++v) xv[v] = 0.0;
    xv[seeds[c]] = 1.0;
  }
  while (it < limit) {
#pragma omp parallel for schedule(static, 1)
    for (uint32_t c = 0; c < ncols; ++c) {
      double *xv = xc + (size_t)c * V, *av = acc + (size_t)c * V;
      memset(av, 0, sizeof(double) * V);
      for (uint32_t u = 0; u < V; ++u) {
        uint64_t b = off[u], e = off[u + 1];
        if (b == e) continue;
        double q = xv[u] / (double)(e - b);
        for (uint64_t j = b; j < e; ++j) av[cols[j]] += q;
      }
      double rc = 0.0;
      for (uint32_t v = 0; v < V; ++v) {
        double xn = alpha * av[v] + (seeds[c] == v ? 1.0 - alpha : 0.0);
        rc += fabs(xn - xv[v]);
        xv[v] = xn;
      }
      r[c] = rc;
    }
    mr = 0.0;
    for (uint32_t c = 0; c < ncols; ++c) mr = r[c] > mr ? r[c] : mr;
    ++it;
    if (fixed_iters <= 0 && mr < eps) break;
  }
#pragma omp parallel for schedule(static)
  for (uint32_t v = 0; v < V; ++v)
    for (uint32_t c = 0; c < ncols; ++c) x[(size_t)v * ncols + c] = xc[(size_t)c * V + v];
  if (max_residual) *max_residual = mr;
  free(xc);
  free(acc);
  free(r);
  return it;
}

int omp_threads(void) { return omp_get_max_threads(); }
