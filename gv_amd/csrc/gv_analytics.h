// gv_analytics.h -- host-side entry points of PageRank, triangle counting and
// betweenness centrality over the device-resident store (gv_analytics.hip).
#pragma once
#include "gv_internal.h"

#include <vector>

namespace gv {

// Device state of one PageRank run.
struct PrCtl {
  double residual;
  unsigned int iters;
  unsigned int done;
  unsigned int blocks_done;
  unsigned int n_bin[4];
  unsigned int cursor[4];
  unsigned int n_chunks;      // chunks of the chunked bin (counted)
  unsigned int chunk_cursor;  // chunk allocation cursor
  unsigned int pad;
};

struct PrWork {
  uint32_t V = 0;
  float *x[2] = {nullptr, nullptr};
  float *contrib[2] = {nullptr, nullptr};
  uint32_t *list = nullptr;      // vertices grouped by degree bin
  uint32_t *hub_base = nullptr;  // per chunked vertex: first chunk partial
  uint32_t *chunk_hub = nullptr; // per chunk: index into the chunked-vertex list
  float *chunk_part = nullptr;   // per chunk: (sum, compensation)
  double *partials = nullptr;    // per block residual partials, then per hub
  uint64_t chunk_cap = 0, part_cap = 0;
  PrCtl *ctl = nullptr;
  PrCtl *h_ctl = nullptr;        // pinned host mirror
  std::vector<hipEvent_t> ev;    // per-iteration start/stop events (timed runs)
  // CSR snapshot of the in-adjacency (live slots only), rebuilt per call.
  uint64_t *snap_off = nullptr;   // [V+1]
  uint32_t *snap_cols = nullptr;  // [E]
  uint64_t snap_cap = 0;
  uint32_t *snap_istart = nullptr;  // [V+1] first chunk item of each vertex
  uint32_t *snap_ivert = nullptr;   // per item: its vertex
  uint64_t *snap_ipos = nullptr;    // per item: live count, then output position
  uint64_t snap_icap = 0;
  void *snap_tmp = nullptr;
  size_t snap_tmp_bytes = 0;
  // Degree-ordered relabelling of the snapshot: new vertex r is old vertex
  // rl_order[r]; rl_newid is the inverse; rl_len / rl_dout are in-row length
  // and out-degree in the new order.
  uint32_t *rl_order = nullptr, *rl_newid = nullptr, *rl_len = nullptr, *rl_dout = nullptr;
  uint64_t *rl_keys[2] = {nullptr, nullptr};
  void *rl_tmp = nullptr;
  uint32_t rl_V = 0;
};

// Run parameters of one PageRank call.
struct PrOpts {
  float alpha = 0.85f;
  double eps = 1e-6;
  int max_iter = 100;
  const float *init = nullptr;  // device vector to start from (warm start), or null = 1/V
  int compensated = 1;          // per-vertex Kahan sums
};

struct PrBins {
  uint32_t thread_max, group_max, wave_max;
};

// Workspace of the XCD-segmented PageRank snapshot (gv_pagerank_seg.hip).
struct PsWork {
  uint32_t V = 0;
  uint64_t E_cap = 0, part_cap = 0;
  uint32_t *src = nullptr;   // sources (relabelled), grouped by (segment, vertex)
  uint32_t *off = nullptr;   // [8 * V + 1] offsets into src
  uint32_t *list = nullptr;  // relabelled vertices grouped by degree bin
  float *P = nullptr;        // [8 * V] per-(segment, vertex) partial sums
  float *xp[2] = {nullptr, nullptr};
  float *cp[2] = {nullptr, nullptr};
  uint32_t *din = nullptr, *dout = nullptr;
  uint32_t *heavy = nullptr;  // bit per relabelled vertex: handled by its own unit
  double *partials = nullptr;
  std::vector<hipEvent_t> ev;
};

// Bin helpers shared by both PageRank layouts.
void pr_ctl_reset(PrCtl *ctl, hipStream_t s);
void pr_bins_count(const uint32_t *deg, uint32_t V, const PrBins &b, PrCtl *ctl, hipStream_t s);
void pr_bins_place(const uint32_t *deg, uint32_t V, const PrBins &b, uint32_t *list, PrCtl *ctl,
                   hipStream_t s);
void ps_release(PsWork &w);
// PageRank on the XCD-segmented snapshot; same contract as pagerank_run.
gv_status pagerank_seg_run(const Store &in, const uint32_t *deg_out, const PrBins &bins,
                           const PrOpts &o, float *out_dev, int *iters, double *residual,
                           PrWork &pw, PsWork &w, hipStream_t s, double *iter_ms = nullptr);

gv_status pr_reserve(PrWork &w, uint32_t V);
void pr_release(PrWork &w);
// Pull PageRank over `in` (in-adjacency; the out-store itself for undirected
// graphs) with out-degrees deg_out.  On return out_dev[V] holds x.
// snapshot = 1: iterate over a CSR copy of the live in-adjacency built at the
// start of the call, vertices relabelled by descending out-degree; 2: the
// same copy in the original vertex order; 0: iterate over the block store.
gv_status pagerank_run(const Store &in, const uint32_t *deg_out, const PrBins &bins,
                       const PrOpts &o, float *out_dev, int *iters, double *residual, PrWork &w,
                       hipStream_t s, double *iter_ms = nullptr, int snapshot = 1);

// Triangle count on an undirected store.  per_vertex_dev may be null.
gv_status triangle_count_run(const Store &st, uint64_t *per_vertex_dev, uint64_t *total,
                             hipStream_t s);

// Brandes BC from the given device source list; out_dev[V] receives BC.
gv_status bc_run(const Store &st, const uint32_t *sources_host, uint32_t n_sources,
                 float *out_dev, hipStream_t s);

// Personalized PageRank for ncols <= 64 seeds (device array); out_dev is
// row-major V x ncols.  max_residual is the largest per-column L1 residual of
// the last update.
gv_status ppr_run(const Store &in, const uint32_t *deg_out, float alpha, double eps, int max_iter,
                  const uint32_t *seeds_dev, uint32_t ncols, float *out_dev, int *iters,
                  double *max_residual, double *iter_ms, hipStream_t s);
// Brandes BC from host source list, 32 sources per batch.
gv_status bc_multi_run(const Store &st, const uint32_t *sources, uint32_t n_sources,
                       float *out_dev, hipStream_t s);

}  // namespace gv
