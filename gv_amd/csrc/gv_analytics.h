// gv_analytics.h -- host-side entry points of PageRank, triangle counting and
// betweenness centrality over the device-resident store (gv_analytics.hip).
#pragma once
#include "gv_internal.h"

#include <vector>

namespace gv {

// Flat stream layout of the PageRank snapshot (gv_pagerank_flat.hip): the
// row-major in-adjacency is cut into tiles of PF_T entries; lane l of a
// wavefront owns entries [l*PF_K, (l+1)*PF_K) of a tile, and entry k of
// every lane is stored contiguously (PF_T/PF_K lanes apart in the stream
// order), so each of a lane's PF_K loads is one contiguous 256-byte read for
// the wavefront.  The first entry of every row carries PF_FLAG.
constexpr uint32_t PF_K = 16;
constexpr uint32_t PF_T = 64 * PF_K;
constexpr uint32_t PF_FLAG = 0x80000000u;
constexpr uint32_t PF_PAD = 0x7FFFFFFFu;  // padding entry of the last tile
__host__ __device__ inline uint64_t pf_phys(uint64_t p) {
  const uint64_t t = p / PF_T, r = p % PF_T;
  return t * PF_T + (r % PF_K) * 64 + r / PF_K;
}

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

// Sizes of the propagation-blocked layout of one snapshot, computed on the
// device (so that a captured PageRank needs no host round trip): stream
// lengths in whole tiles, tile and bin counts, the (source block, bin)
// table size, long rows, expansion slots and image groups.  fail != 0 (a
// snapshot larger than the buffers reserved for a captured call, or a failed
// relabelling sort) makes every later kernel of the call exit at once.
struct PbDims {
  unsigned long long E, hend, cend;
  unsigned int nht, nct, NB, table, n_long, P1, G, fail;
};

// Capacities (bytes) of the propagation-blocked layout's buffers.
struct PbCaps {
  size_t hitem = 0, hoff = 0, coff = 0, lidxr = 0, lng = 0, Sh = 0, Sc = 0, partials = 0;
  size_t hot = 0, cold = 0, csrc = 0, lidx = 0, vals = 0, htb = 0, ctb = 0, PQh = 0, PQc = 0;
  size_t cnt = 0, ngrp = 0, gaddr = 0, nit = 0, tmp = 0, items = 0;
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
  PrCtl *h_ctl = nullptr;        // pinned host mirror: [0] final, [1..2] per iteration group
  std::vector<hipEvent_t> ev;    // timed runs: start, then one event after each iteration
  hipEvent_t ev_grp[2] = {nullptr, nullptr};  // end of an iteration group (flat form)
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
  uint64_t snap_E = 0;  // live entries of the last snapshot
  // entries the snapshot buffers are sized for at least (a caller that must
  // not reallocate while other streams run presizes with headroom)
  uint64_t snap_min_cap = 0;
  uint32_t pf_grid = 0;  // persistent workgroups of the flat iteration (0 = one per CU)
  // flat iteration (per snapshot): per tile, the nonempty row in progress
  // before the tile and the partial sums of its first and last rows; per
  // nonempty row, its label, stream offset and in-tile sum; empty rows' labels
  uint32_t pf_V = 0;
  uint64_t pf_tiles_cap = 0;
  int32_t *pf_tbase = nullptr;
  float2 *pf_P = nullptr, *pf_Q = nullptr, *pf_S = nullptr;
  uint32_t *pf_nzlabel = nullptr, *pf_nzoff = nullptr, *pf_empty = nullptr, *pf_nzidx = nullptr;
  uint32_t *pf_lidx = nullptr, *pf_long = nullptr;  // rows spanning many tiles
  // sizes of the prepared flat layout (pagerank_flat_prepare)
  int pf_ready = 0;
  uint32_t pf_prepV = 0, pf_n_nz = 0, pf_n_long = 0;
  uint64_t pf_E = 0, pf_n_tiles = 0;
  double *pf_partials = nullptr;
  // propagation-blocked layout (gv_pagerank_pb.hip), built per snapshot:
  // hot and cold 2-byte streams, their row offsets (label order), tile bases
  // and row partials; the expansion stream of cold source offsets, the
  // values it fills, the (source block, bin) table, per bin the offset of its
  // image groups and per image group its place in the values, the expansion
  // work items
  uint16_t *pb_hot = nullptr, *pb_cold = nullptr, *pb_lidx = nullptr;
  uint32_t *pb_csrc = nullptr, *pb_hoff = nullptr, *pb_coff = nullptr;
  int32_t *pb_htb = nullptr, *pb_ctb = nullptr;
  float2 *pb_Sh = nullptr, *pb_Sc = nullptr, *pb_PQh = nullptr, *pb_PQc = nullptr;
  float *pb_vals = nullptr;
  uint32_t *pb_hitem = nullptr, *pb_cnt = nullptr, *pb_ngrp = nullptr, *pb_nit = nullptr;
  uint32_t *pb_gaddr = nullptr;
  uint4 *pb_items = nullptr;
  uint32_t *pb_lidxr = nullptr, *pb_long = nullptr;
  void *pb_tmp = nullptr;
  double *pb_partials = nullptr;
  PbCaps pb_caps;
  int pb_ready = 0;
  uint32_t pb_NS = 0, pb_NB = 0, pb_nht = 0, pb_nct = 0, pb_n_long = 0, pb_prepV = 0;
  uint32_t pb_bin_tiles = 0;
  uint64_t pb_E = 0;
  PbDims *pb_dims = nullptr;   // device sizes of the prepared layout
  PbDims *h_pb_dims = nullptr; // pinned host mirror
  // capture reservation (pagerank_pb_reserve): vertex count and live
  // in-edges every buffer is sized for; 0 = none
  uint32_t pb_res_V = 0;
  uint64_t pb_res_E = 0;
  uint64_t gen = 0;  // bumped whenever a buffer of this workspace moves
};

// Run parameters of one PageRank call.
struct PrOpts {
  float alpha = 0.85f;
  double eps = 1e-6;
  int max_iter = 100;
  const float *init = nullptr;  // device vector to start from (warm start), or null = 1/V
  int compensated = 1;          // per-vertex Kahan sums
  // Called before each iteration is issued (the iterations then run one at
  // a time): lets a caller hold further work back while other work waits.
  void (*yield)(void *) = nullptr;
  void *yield_arg = nullptr;
};

struct PrBins {
  uint32_t thread_max, group_max, wave_max;
};

void pr_ctl_reset(PrCtl *ctl, hipStream_t s);

gv_status pr_reserve(PrWork &w, uint32_t V);
// CSR snapshot of the live in-adjacency (relabelled by descending out-degree
// when relabel); flat != 0 writes the tiled stream layout above.  hitem (may
// be null, pr_snapshot_icap(in) entries): per chunk item of the snapshot, the
// entries it wrote whose (relabelled) source is below hot_below.
uint64_t pr_snapshot_icap(const Store &in);
// fail (capture form, may be null): no host round trip; every buffer must
// already hold the reserved size (GV_E_INVALID otherwise), and a snapshot
// larger than the reservation, or a failed relabelling sort, sets *fail
// instead of being written.
gv_status pr_snapshot(const Store &in, const uint32_t *deg_out, bool relabel, PrWork &w,
                      hipStream_t s, int flat, uint32_t *hitem = nullptr, uint32_t hot_below = 0,
                      uint32_t *fail = nullptr);
void pr_init_vectors(float *x, float *cb, const uint32_t *deg_out, uint32_t V, float x0,
                     const float *init, const uint32_t *order, hipStream_t s);
void pr_unpermute(const float *x, const uint32_t *newid, uint32_t V, float *out, hipStream_t s);
// Flat-stream PageRank over the relabelled snapshot; same contract as
// pagerank_run.  Returns GV_E_UNSUPPORTED when the graph exceeds the
// layout's 32-bit stream positions (the caller then uses the row form).
gv_status pagerank_flat_run(const Store &in, const uint32_t *deg_out, const PrOpts &o,
                            float *out_dev, int *iters, double *residual, PrWork &w,
                            hipStream_t s, double *iter_ms);
// The same in two steps: the snapshot and its layout, ordered on stream s
// after everything before it (the store is not read again afterwards), then
// the iterations over w alone, on any stream ordered after the first step.
gv_status pagerank_flat_prepare(const Store &in, const uint32_t *deg_out, PrWork &w,
                                hipStream_t s);
gv_status pagerank_flat_iterate(const PrOpts &o, PrWork &w, float *out_dev, int *iters,
                                double *residual, hipStream_t s, double *iter_ms);
void pr_release(PrWork &w);
// Propagation-blocked PageRank over the relabelled snapshot (same contract
// as the flat form): prepare builds the snapshot and both streams on s and
// returns GV_E_UNSUPPORTED when the graph exceeds the layout's limits (more
// than 16384 cold source blocks, 32-bit stream positions); iterate runs the
// iterations over w alone.
// capture: the form a stream capture can record -- no host round trip and
// no allocation (pagerank_pb_reserve must have sized every buffer for V and
// at least the snapshot's live entries; a larger snapshot sets
// w.pb_dims->fail on the device and the call computes nothing), launches
// sized by capacity, max_iter iteration launches (those after convergence
// exit at once); iters, residual and iter_ms must be null.
gv_status pagerank_pb_prepare(const Store &in, const uint32_t *deg_out, PrWork &w, hipStream_t s,
                              bool capture = false);
gv_status pagerank_pb_iterate(const PrOpts &o, PrWork &w, float *out_dev, int *iters,
                              double *residual, hipStream_t s, double *iter_ms,
                              bool capture = false);
// Sizes every buffer of the propagation-blocked layout (and its snapshot)
// for graphs of V vertices and up to max_edges live in-edges, by preparing
// the current graph once with that headroom.  GV_E_UNSUPPORTED when the
// layout does not apply to such graphs.
gv_status pagerank_pb_reserve(const Store &in, const uint32_t *deg_out, PrWork &w,
                              uint64_t max_edges, hipStream_t s);
// Iterations and residual of the last (possibly captured) call; GV_E_INVALID
// when its snapshot exceeded the reservation (nothing was computed).
gv_status pagerank_pb_last(PrWork &w, int *iters, double *residual, hipStream_t s);
void pb_release(PrWork &w);
// Snapshot + layout for PageRank by layout id (5: propagation-blocked, with
// the flat stream as fallback; otherwise the flat stream), then iterations
// over whichever was prepared.
gv_status pagerank_snap_prepare(const Store &in, const uint32_t *deg_out, PrWork &w,
                                hipStream_t s, int layout);
gv_status pagerank_snap_iterate(const PrOpts &o, PrWork &w, float *out_dev, int *iters,
                                double *residual, hipStream_t s, double *iter_ms);
// Pull PageRank over `in` (in-adjacency; the out-store itself for undirected
// graphs) with out-degrees deg_out.  On return out_dev[V] holds x.
// snapshot = 5: the propagation-blocked layout (falling back to 1);
// snapshot = 1: iterate over a copy of the live in-adjacency built at the
// start of the call, vertices relabelled by descending out-degree, as the
// flat tiled stream (falling back to the row form of 3 when the graph
// exceeds its 32-bit positions); 3: the same relabelled copy as CSR rows
// cut into degree bins (bins), one row or row chunk per work item.
gv_status pagerank_run(const Store &in, const uint32_t *deg_out, const PrBins &bins,
                       const PrOpts &o, float *out_dev, int *iters, double *residual, PrWork &w,
                       hipStream_t s, double *iter_ms = nullptr, int snapshot = 1);

// Triangle count on an undirected store.  per_vertex_dev may be null.
gv_status triangle_count_run(const Store &st, uint64_t *per_vertex_dev, uint64_t *total,
                             hipStream_t s);

// Grow-only workspace of tc_delta.
struct TcdWork {
  uint32_t *cnt = nullptr, *start = nullptr, *iseg = nullptr;
  uint64_t *hset = nullptr;           // changed-edge set
  void *tmp = nullptr;
  unsigned long long *dev = nullptr;  // [0] total change, [1] block-level items
  size_t cnt_bytes = 0, start_bytes = 0, iseg_bytes = 0, hset_bytes = 0, tmp_bytes = 0,
         dev_bytes = 0;
};
void tcd_release(TcdWork &w);
// Change of the triangle counts made by the staged batch in sc (keys: its m
// sorted keys).  Changed edges are the first copies of keys whose flag equals
// sflag: new edges after an insert (sflag 1, counted in the graph after it)
// or present edges before a delete (sflag 0 after a membership probe, counted
// before the delete, neg = 1).  per_vertex_dev (may be null) is updated in
// place; *delta receives the change of the total.  Skips everything when
// the store holds an abort code.
gv_status tc_delta(const Store &st, const Scratch &sc, const uint64_t *keys, uint64_t m,
                   uint32_t sflag, int neg, uint64_t *per_vertex_dev, TcdWork &w,
                   long long *delta, hipStream_t s);

// Brandes BC from the given host source list; out_dev[V] receives BC.
gv_status bc_run(const Store &st, const uint32_t *sources_host, uint32_t n_sources,
                 float *out_dev, hipStream_t s);
// Workspace of single-source BC, reusable across calls.
struct BcWork {
  uint32_t V = 0;
  int *dist = nullptr;
  double *sigma = nullptr;
  float *delta = nullptr, *bcc = nullptr;
  uint32_t *queue = nullptr;
  unsigned int *cursor = nullptr, *h_cursor = nullptr;
  // rows longer than BC_HUB slots of the level piece in flight, spread over
  // the whole grid; backward partials per (hub, block)
  uint32_t *hubs = nullptr;
  unsigned int *nhub = nullptr;
  float2 *hpart = nullptr;
  uint64_t hpart_cap = 0;  // hubs hpart holds
  uint32_t max_blocks = 0;  // grid cap of the level kernels (0 = 8192)
  // Called before each level is issued, as PrOpts::yield.
  void (*yield)(void *) = nullptr;
  void *yield_arg = nullptr;
};
gv_status bc_reserve(BcWork &w, uint32_t V);
void bc_release(BcWork &w);
// The same over a CSR snapshot of the out-adjacency (offsets[V+1], cols).
gv_status bc_run_csr(const uint64_t *off, const uint32_t *cols, uint32_t V,
                     const uint32_t *sources_host, uint32_t n_sources, float *out_dev,
                     hipStream_t s, BcWork &w);

// Personalized PageRank for ncols <= 64 seeds (device array); out_dev is
// row-major V x ncols.  max_residual is the largest per-column L1 residual of
// the last update.
gv_status ppr_run(const Store &in, const uint32_t *deg_out, float alpha, double eps, int max_iter,
                  const uint32_t *seeds_dev, uint32_t ncols, float *out_dev, int *iters,
                  double *max_residual, double *iter_ms, hipStream_t s);
// Grow-only workspace of multi-source BC, kept with the graph handle so
// repeated calls do not allocate (or free, which waits for the device).
// n_* are capacities in elements.
struct BcmWork {
  double *sigma = nullptr;
  float *coef = nullptr, *bcc = nullptr;
  void *sn = nullptr, *fl = nullptr;  // uint2 per vertex
  uint32_t *queue = nullptr, *qmask = nullptr, *qpart = nullptr, *longl = nullptr;
  uint32_t *dsrc = nullptr;
  void *items = nullptr, *cand = nullptr, *part = nullptr;  // uint2, uint2, float2
  void *lv = nullptr, *ctl = nullptr, *h_ctl = nullptr;
  unsigned long long *used_tot = nullptr;
  uint64_t n_sigma = 0, n_coef = 0, n_bcc = 0, n_sn = 0, n_fl = 0, n_queue = 0, n_qmask = 0,
           n_qpart = 0, n_items = 0, n_longl = 0, n_cand = 0, n_part = 0, n_dsrc = 0, n_ctl = 0,
           n_used_tot = 0, lvcap = 0;
};
void bcm_release(BcmWork &w);
// Brandes BC from a host source list in sweeps of 16 (at most 16 sources) or
// 32 sources.  out: adjacency; in: in-adjacency for bottom-up levels (the
// same store for undirected graphs; null disables bottom-up); bu_ratio: a
// level goes bottom-up when its candidates' in-slots are below bu_ratio
// times its frontier's out-slots (0: always top-down).
gv_status bc_multi_run(const Store &out, const Store *in, const uint32_t *sources,
                       uint32_t n_sources, float *out_dev, BcmWork &w, float bu_ratio,
                       hipStream_t s);

// Request-loop hooks over a graph handle (gv_api.hip): the relabelled flat
// PageRank snapshot of the handle's in-adjacency, and a CSR snapshot of its
// out-adjacency (w.snap_off / w.snap_cols) for BC, each built on the
// handle's stream.
gv_status graph_pr_prepare(gv_graph *g, PrWork &w);
gv_status graph_bc_prepare(gv_graph *g, PrWork &w);

}  // namespace gv
