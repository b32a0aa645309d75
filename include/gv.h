/*
 * gv.h -- C ABI of the MI355X-native GraphVine dynamic graph store.
 *
 * Every entry point returns a gv_status (0 = GV_OK, negative = GV_E_*) and
 * never exits or throws; gv_last_error() returns thread-local text for the
 * last failure.  All device work of a handle is ordered on that handle's own
 * HIP stream.  A handle is not thread-safe: use one handle per host thread, or
 * serialise through one queue.
 *
 * Pointers are plain host or device addresses; a `mem` argument
 * (GV_MEM_HOST / GV_MEM_DEVICE) says which.  Vertex IDs are 0-based uint32.
 *
 * Reference interfaces replaced (paths relative to the GraphVine sources):
 *   GraphVine::GraphVine(CSR*)        include/dynamicStructure/graphvine.cuh:31
 *   GraphVine::bulkBuild()            include/dynamicStructure/graphvine.cuh:37
 *   GraphVine::batchInsert(CSR*,kk)   include/dynamicStructure/graphvine.cuh:39
 *   GraphVine::batchDelete(CSR*,kk)   include/dynamicStructure/graphvine.cuh:41
 *   GraphVine::compaction()           include/dynamicStructure/graphvine.cuh:43
 *   GraphVine::getVertexDictionary()  include/dynamicStructure/graphvine.cuh:47
 *   GraphVine::getVertexSize()        include/dynamicStructure/graphvine.cuh:56
 *   static_pagerank(...)              include/algorithm/pagerank/pagerank.cuh:11
 *   static_tc(...)                    include/algorithm/tc/tc.cuh:10
 *   dynamic_tc(...)                   src/algorithm/tc/tc.cu:29-64
 *   bc(GraphVine*, int)               include/algorithm/bc/bc.cuh:8
 *   EdgePreallocatedQueue (pool)      include/dynamicStructure/graphStructures.cuh:103-108
 */
#ifndef GV_H
#define GV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int gv_status;

#define GV_OK 0
#define GV_E_INVALID -1     /* bad argument or handle */
#define GV_E_RANGE -2       /* a vertex ID >= num_vertices; the batch was not applied */
#define GV_E_NOMEM -3       /* device allocation failed */
#define GV_E_POOL -4        /* edge-block pool exhausted; the batch was not applied */
#define GV_E_HIP -5         /* HIP runtime error */
#define GV_E_UNSUPPORTED -6 /* operation not defined for this graph kind */

#define GV_MEM_HOST 0
#define GV_MEM_DEVICE 1

typedef struct gv_graph gv_graph;

/* Store configuration.  gv_config_default() fills every field. */
typedef struct gv_config {
  uint32_t block_slots;       /* slots per edge block: 32, 64 or 128 (4 B each) */
  uint64_t pool_blocks;       /* edge blocks in the pool; 0 = sized from free HBM; at most
                                 2^32 / block_slots - 1 per store (32-bit slot indices) */
  double pool_fraction;       /* share of free HBM used when pool_blocks == 0 */
  float compact_threshold;    /* default tombstone fraction for gv_compact(g, -1) */
  uint32_t pr_thread_max;     /* degree bins of the PageRank row form (layout 4): a row with
                                 <= this entries -> 1 lane per vertex */
  uint32_t pr_group_max;      /*   <= this -> 8 lanes per vertex */
  uint32_t pr_wave_max;       /*   <= this -> 1 wavefront; larger -> chunked */
  int maintain_in_edges;      /* directed graphs: keep the in-adjacency mirror */
  int device;                 /* HIP device ordinal */
  int pr_layout;              /* PageRank: 5 = snapshot of the live in-adjacency, rebuilt per
                                 call, vertices relabelled by descending out-degree, split into
                                 a hot stream (sources among the highest out-degrees, staged in
                                 LDS) and a propagation-blocked cold stream (default; graphs
                                 beyond its limits use 1); 1 = the same snapshot iterated as one
                                 flat stream of equal tiles (graphs beyond its 32-bit positions
                                 use 4); 4 = the relabelled snapshot row by row, rows cut into
                                 the degree bins below */
  float bc_bu_ratio;          /* multi-source BC: a BFS level is expanded bottom-up (unreached
                                 vertices pull from the level over their in-adjacency) when
                                 the candidates' in-slots are below this multiple of the
                                 level's out-slots, else top-down; 0 = always top-down.
                                 Directed graphs need maintain_in_edges for bottom-up. */
} gv_config;

/* Result of one update call. */
typedef struct gv_batch_stats {
  uint64_t requested;        /* (src,dst) pairs passed in */
  uint64_t directed_entries; /* directed entries after self-loop removal */
  uint64_t applied;          /* slots written (insert) or tombstoned (delete) */
  uint64_t touched;          /* distinct source vertices in the batch */
  uint64_t blocks_allocated; /* edge blocks taken from the pool */
  uint64_t live_edges;       /* directed live edges after the call */
} gv_batch_stats;

gv_status gv_config_default(gv_config *cfg);

/* GraphVine(CSR*) without the bulk build: an empty graph on num_vertices
 * vertices.  directed = 0 stores every edge in both directions. */
gv_status gv_create(const gv_config *cfg, uint32_t num_vertices, int directed, gv_graph **out);
gv_status gv_destroy(gv_graph *g);

/* GraphVine::bulkBuild: clear the graph, then insert the COO edge list. */
gv_status gv_bulk_build(gv_graph *g, const uint32_t *src, const uint32_t *dst, uint64_t n,
                        int mem, gv_batch_stats *stats);
/* GraphVine::batchInsert: N(u) |= {v}; duplicates and self-loops dropped. */
gv_status gv_batch_insert(gv_graph *g, const uint32_t *src, const uint32_t *dst, uint64_t n,
                          int mem, gv_batch_stats *stats);
/* GraphVine::batchDelete: N(u) \= {v} (tombstones); missing edges are no-ops. */
gv_status gv_batch_delete(gv_graph *g, const uint32_t *src, const uint32_t *dst, uint64_t n,
                          int mem, gv_batch_stats *stats);
/* Batched form: a group of update batches in one pipeline.  Batch b is
 * entries [offsets[b], offsets[b+1]) of src/dst (offsets: host array of
 * n_batches+1, offsets[0] = 0) with kind kinds[b] (GV_OP_INSERT or
 * GV_OP_DELETE, host array).  The resulting graph equals applying the batches
 * one after another (an edge's final membership is decided by its last
 * operation in the group); stats cover the whole group (applied = slots
 * written + slots tombstoned).  compact_fraction >= 0 then compacts, at that
 * tombstone fraction, the vertices the group touched (only they can have
 * gained tombstones from it; an empty group compacts every vertex, like
 * gv_compact).  No counterpart in the reference, which applies one CSR batch
 * per call (graphvine.cu:222-451). */
#define GV_OP_INSERT 0
#define GV_OP_DELETE 1
gv_status gv_batch_apply(gv_graph *g, const uint32_t *src, const uint32_t *dst,
                         const uint64_t *offsets, const int *kinds, uint32_t n_batches, int mem,
                         float compact_fraction, gv_batch_stats *stats);
/* The batched form over a device-resident queue: src, dst (max_pairs
 * entries), offsets (max_batches+1) and kinds (max_batches) are all device
 * pointers, read when the call runs on the handle's stream, so a producer can
 * refill them between launches of a capture that recorded this call.  The
 * queue holds offsets[max_batches] <= max_pairs pairs: unused trailing
 * batches are empty (equal offsets) and entries past offsets[max_batches] are
 * ignored.  Inside a capture the call needs gv_reserve(g, max_pairs,
 * max_batches) first and never allocates; an eager call grows the staging
 * like gv_batch_apply.  Nothing is read back: offsets that do not start at
 * 0, that decrease or that end past max_pairs, kinds other than GV_OP_INSERT
 * or GV_OP_DELETE and out-of-range vertex IDs leave the graph unchanged and
 * set the handle's sticky error, which skips later updates until gv_sync (or
 * gv_capture_launch with stats) reports it (GV_E_INVALID / GV_E_RANGE) and
 * clears it.  No reference counterpart: for callers that drain a device-side
 * update queue without a host round trip; the reference takes one host-built
 * CSR batch per call (graphvine.cu:222-451). */
gv_status gv_batch_apply_device(gv_graph *g, const uint32_t *src, const uint32_t *dst,
                                const uint64_t *offsets, const int *kinds, uint32_t max_batches,
                                uint64_t max_pairs, float compact_fraction);
/* GraphVine::compaction: pack every vertex whose tombstones exceed
 * tombstone_fraction of its slots (0 = any tombstone, < 0 = the configured
 * default) and return emptied blocks to the pool. */
gv_status gv_compact(gv_graph *g, float tombstone_fraction, uint64_t *vertices_compacted);

/* Stats reads (these synchronise the handle's stream). */
gv_status gv_num_vertices(gv_graph *g, uint32_t *v);
gv_status gv_num_edges(gv_graph *g, uint64_t *live_edges);
gv_status gv_degrees(gv_graph *g, uint32_t *deg, int mem);
gv_status gv_pool_stats(gv_graph *g, uint64_t *free_blocks, uint64_t *total_blocks);

/* Sorted CSR export: offsets[V+1], cols[offsets[V]] (ascending per vertex).
 * With cols, a graph of more than 2^32-1 live directed edges is refused
 * (GV_E_UNSUPPORTED): the export sorts every entry at once. */
gv_status gv_export_csr(gv_graph *g, uint64_t *offsets, uint32_t *cols, int mem);
/* 1-hop lookups: offs[n+1]; out[offs[n]] sorted per query (out may be NULL
 * to size it first; with out, more than 2^32-1 neighbours in one call is
 * GV_E_UNSUPPORTED). */
gv_status gv_neighbors(gv_graph *g, const uint32_t *vs, uint32_t n, uint64_t *offs, uint32_t *out,
                       int mem);

/* gv_neighbors without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * vs[n], offs[n+1] and out are device pointers (out may be NULL to size the
 * answer first); results equal gv_neighbors'.  Working memory is one
 * file-scope device array of the library (36 MiB, held on every GPU the
 * library runs on until the process ends), so limits apply: n <= 65536
 * queries and offs[n] <= 2097152 neighbours per call (GV_E_INVALID beyond,
 * nothing written to out).  Every call on one GPU shares that array: calls
 * that overlap on one GPU (other threads, other handles) race. */
gv_status gv_neighbors_noalloc(gv_graph *g, const uint32_t *vs, uint32_t n, uint64_t *offs,
                               uint32_t *out);

/* static_pagerank: FP32 values, per-vertex compensated pull sums, FP64 L1
 * residual; stops after the first update with residual < eps or after
 * max_iter updates.  out[V] receives x. */
gv_status gv_pagerank(gv_graph *g, float alpha, double eps, int max_iter, float *out, int mem,
                      int *iters, double *residual);
/* gv_pagerank plus measurement: kernel_ms is the device time of the counted
 * iterations (HIP events on the handle's stream; in the flat layout one span
 * from the first iteration's launch to the end of the last counted one, launch
 * gaps included); edges is the live in-adjacency size each iteration reads. */
typedef struct gv_pagerank_stats {
  int iters;
  double residual;
  double kernel_ms;
  uint64_t edges;
} gv_pagerank_stats;
gv_status gv_pagerank_ex(gv_graph *g, float alpha, double eps, int max_iter, float *out, int mem,
                         gv_pagerank_stats *stats);
/* Full-control PageRank: warm start from a previous vector (an incremental
 * refresh after updates converges to the same fixed point in fewer
 * iterations; the reference's dynamic_pagerank heuristic, pagerank.cu:70-132,
 * has no exact semantics) and plain vs compensated FP32 sums.  stats may be
 * NULL. */
typedef struct gv_pagerank_opts {
  float alpha;       /* damping, 0.85 */
  double eps;        /* L1 residual threshold */
  int max_iter;      /* update limit */
  const float *init; /* NULL: x0 = 1/V; else V start values */
  int init_mem;      /* GV_MEM_HOST or GV_MEM_DEVICE for init */
  int compensated;   /* 1: per-vertex Kahan sums (default); 0: plain FP32 sums */
} gv_pagerank_opts;
gv_status gv_pagerank_opts_default(gv_pagerank_opts *opts);
gv_status gv_pagerank_run(gv_graph *g, const gv_pagerank_opts *opts, float *out, int mem,
                          gv_pagerank_stats *stats);
/* Sizes every buffer of the default PageRank layout for this graph's vertex
 * count and up to max_edges live (in-)edges, so that gv_pagerank_run on
 * such a graph allocates nothing and needs no host round trip inside a
 * capture (below).  Needs pr_layout 5; GV_E_UNSUPPORTED when the layout's
 * 32-bit positions do not cover max_edges.  No reference counterpart (for
 * captured update -> compact -> PageRank pipelines). */
gv_status gv_pagerank_reserve(gv_graph *g, uint64_t max_edges);
/* Iterations and final residual of the last default-layout PageRank call on
 * the handle, captured calls included (it waits for the handle's stream);
 * GV_E_INVALID when that call's snapshot exceeded the reservation (nothing
 * was computed then).  edges: live in-adjacency size now. */
gv_status gv_pagerank_last(gv_graph *g, gv_pagerank_stats *stats);
/* Personalized PageRank for up to 64 seeds in one FP32 column block: column
 * c starts at e_{seeds[c]} and teleports there, x_c' = (1-alpha) e_{s_c} +
 * alpha * sum_{u in N_in(v)} x_c[u]/|N(u)| (dangling mass dropped, as in
 * static_pagerank); stops after the first update whose largest per-column L1
 * residual is below eps, or after max_iter.  seeds: host or device (mem);
 * out: row-major V x n_cols (mem).  No counterpart in the reference. */
gv_status gv_pagerank_personalized(gv_graph *g, float alpha, double eps, int max_iter,
                                   const uint32_t *seeds, uint32_t n_cols, float *out, int mem,
                                   int *iters, double *max_residual);
/* static_tc on an undirected graph: exact uint64 counts; per_vertex (may be
 * NULL) gets t[u] = triangles whose minimum vertex ID is u. */
gv_status gv_triangle_count(gv_graph *g, uint64_t *per_vertex, int mem, uint64_t *total);
/* Apply one batch of kind GV_OP_INSERT or GV_OP_DELETE to an undirected
 * graph, exactly as gv_batch_insert / gv_batch_delete do, and bring triangle
 * counts of the graph before it up to date: per_vertex (V entries in pv_mem,
 * may be NULL) and *total are read and rewritten so that they equal what
 * gv_triangle_count returns after the batch.  Only the triangles through
 * edges the batch really adds or removes are visited (duplicates, self-loops,
 * inserts of present and deletes of absent edges change nothing).  On an
 * error (a bad ID, an exhausted pool) the batch is not applied and the
 * counts are left as they were.  Replaces dynamic_tc (tc.cu:29-64, with the
 * affected set of find_affected_nodes, insert_kernels.cu:182-198), whose
 * recount of the updated edges' endpoints misses triangles whose minimum
 * vertex is the third vertex. */
gv_status gv_triangle_count_update(gv_graph *g, int kind, const uint32_t *src, const uint32_t *dst,
                                   uint64_t n, int mem, uint64_t *per_vertex, int pv_mem,
                                   uint64_t *total, gv_batch_stats *stats);
/* bc: Brandes betweenness from the given sources (unnormalised, not halved);
 * sigma in FP64, dependencies in FP32 with compensated sums.  Several sources
 * run 32 at a time in one level-synchronous sweep. */
gv_status gv_bc(gv_graph *g, const uint32_t *sources, uint32_t n_sources, float *out, int mem);

/* Device generators for update batches and synthetic graphs.  Each draws
 * from a counter-based SplitMix64 stream (output i = mix(key + (i+1)*golden))
 * so any range of a stream can be produced independently.
 *   rmat:    one draw per level, high 32 bits vs cumulative quadrant
 *            thresholds ta=a, tab=a+b, tabc=a+b+c scaled by 2^32;
 *   uniform: u = hi32*V>>32, v = lo32*V>>32, a self-loop moves v to v+1 mod V;
 *   sample:  entry (draw % offsets[V]) of a sorted CSR (with replacement).
 * Outputs are device arrays of n entries. */
gv_status gv_gen_rmat(gv_graph *g, uint64_t key, int scale, uint64_t n, uint32_t ta, uint32_t tab,
                      uint32_t tabc, uint32_t *src_dev, uint32_t *dst_dev);
gv_status gv_gen_uniform(gv_graph *g, uint64_t key, uint32_t V, uint64_t n, uint32_t *src_dev,
                         uint32_t *dst_dev);
gv_status gv_gen_sample(gv_graph *g, uint64_t key, uint32_t V, const uint64_t *offsets_dev,
                        const uint32_t *cols_dev, uint64_t n, uint32_t *src_dev,
                        uint32_t *dst_dev);

/* Chung-Lu endpoints: draw i (draw 2e is the source of edge e, draw 2e+1 its
 * destination) is the vertex lo with cum[lo] <= (r_i >> 2) < cum[lo+1],
 * r_i the SplitMix64 stream value, where cum is the FP64 running sum in
 * vertex order of w_i = (i+1)^(-1/(gamma-1)) scaled to average degree
 * avg_deg and capped at max_deg, normalised to 2^62.  The table is built on
 * the host once per parameter set and kept with the handle; the draws run on
 * the device.  Synthetic input of the power-law workloads; the reference has
 * no generator of this kind (its only generator is csr.cu:80-125). */
gv_status gv_gen_chung_lu(gv_graph *g, uint64_t key, uint32_t V, double gamma, double avg_deg,
                          double max_deg, uint64_t n, uint32_t *src_dev, uint32_t *dst_dev);

/* MatrixMarket coordinate file -> 0-based (src,dst) pairs (readFile,
 * src/reader/marketReader.cu:3-121).  Lines starting with '%' are skipped;
 * the first other line holds "rows cols nnz"; each entry line holds a 1-based
 * "row col [value ...]" (values are ignored).  num_vertices = max(rows, cols).
 * Two calls: with src = dst = NULL it returns num_vertices and *n = nnz; then
 * with buffers of *n entries it fills them.  *symmetric (may be NULL) is 1
 * when the banner declares a symmetric or Hermitian matrix, whose file holds
 * one triangle: load it into an undirected graph (or insert both ways).
 * Fails with GV_E_INVALID on an unreadable file, a malformed line, an entry
 * outside 1..rows x 1..cols or an entry count other than nnz. */
gv_status gv_read_matrix_market(const char *path, uint32_t *num_vertices, uint64_t *n,
                                uint32_t *src, uint32_t *dst, int *symmetric);

/* Tuning.  gv_set_tuning changes, on a live handle, the fields that do not
 * move memory: compact_threshold, pr_thread_max, pr_group_max, pr_wave_max,
 * pr_layout and bc_bu_ratio (the other fields of cfg are ignored).  Used by the
 * autotuner; no reference counterpart (the reference's block size and
 * launch shapes are compile-time constants, constants.cuh:4-17). */
gv_status gv_get_config(gv_graph *g, gv_config *cfg);
gv_status gv_set_tuning(gv_graph *g, const gv_config *cfg);

/* Captured pipelines.  Records update, compaction and PageRank calls issued
 * on the handle's stream into one HIP graph that gv_capture_launch replays
 * with a single launch.  No reference counterpart (the reference runs each
 * stage synchronously, graphvine.cu:222-464, with a host read per PageRank
 * iteration, pagerank.cu:46); for callers that replay a fixed update ->
 * compact -> PageRank sequence over device buffers they refill between
 * launches.
 *   - gv_reserve sizes the staging buffers for groups of up to max_pairs
 *     (src,dst) pairs and max_batches batches: a call inside a capture never
 *     allocates device memory and fails if the reservation is too small;
 *   - gv_pagerank_reserve sizes the PageRank buffers the same way: a
 *     captured gv_pagerank_run builds its snapshot and layout on the device
 *     and records max_iter iterations (those after convergence exit at
 *     once); a replay whose graph outgrew the reservation computes nothing
 *     and gv_pagerank_last then fails;
 *   - between gv_capture_begin and gv_capture_end only gv_batch_insert,
 *     gv_batch_delete and gv_batch_apply with GV_MEM_DEVICE inputs and NULL
 *     stats, gv_batch_apply_device, gv_compact with a NULL count, and
 *     gv_pagerank_run with device vectors and NULL stats are allowed; every other call on the handle
 *     fails and a failed call makes gv_capture_end fail;
 *   - a replay reads the device inputs at the addresses given at capture time
 *     (their contents may change between launches) and the batch offsets and
 *     kinds given at capture time (gv_batch_apply_device reads those from
 *     device memory too, so the batch count and sizes may change as well);
 *   - a later reallocation of a buffer the capture uses (a group larger than
 *     reserved, a plain PageRank call on a graph larger than reserved)
 *     invalidates it: gv_capture_launch then returns GV_E_INVALID;
 *   - gv_capture_launch with stats waits for the replay; the counters in
 *     stats describe its last update call, live_edges the whole graph.
 * A capture outliving its graph can only be destroyed. */
typedef struct gv_capture gv_capture;
gv_status gv_reserve(gv_graph *g, uint64_t max_pairs, uint32_t max_batches);
gv_status gv_capture_begin(gv_graph *g);
gv_status gv_capture_end(gv_graph *g, gv_capture **out);
gv_status gv_capture_launch(gv_capture *c, gv_batch_stats *stats);
gv_status gv_capture_destroy(gv_capture *c);

/* Test hook: bound on the polls of one look-back wait inside the staging
 * radix sort (default 2^26).  0 makes every sort tile after the first give
 * up at once, so a test can check that the failure is reported (GV_E_HIP,
 * nothing applied) while every kernel still exits normally.  Process-wide;
 * not for production use.  No reference counterpart. */
gv_status gv_test_set_sort_spin_limit(uint32_t limit);

/* Stream plumbing. */
gv_status gv_sync(gv_graph *g);
gv_status gv_get_stream(gv_graph *g, void **hip_stream);
const char *gv_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
