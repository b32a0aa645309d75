/*
 * gv_serve.h -- request loop over one device-resident graph.
 *
 * Replaces the reference's driver, main.cu:20-134: a one-shot CLI that bulk
 * builds, runs a fixed sequence of batches and then the analytics (its
 * benchmark switch, main.cu:76-128, is dead code).  Here a server takes
 * update batches and queries as they arrive and serves them in arrival order
 * on one worker thread:
 *   - consecutive update requests (insert / delete batches) form one step,
 *     applied by gv_batch_apply as one group -- the same graph as applying
 *     them one after another -- followed by compaction at compact_fraction;
 *   - admission by free pool memory: a step takes update requests while the
 *     sum of their worst-case block needs (one new edge block for every
 *     directed entry written, 2 per pair) fits in the free pool blocks minus
 *     headroom_blocks; the first request that does not fit starts the next
 *     step.  A request that does not fit on its own is applied alone under the
 *     device's exact all-or-nothing pool check, and refused with GV_E_POOL if
 *     the pool cannot hold it (the graph is then unchanged by it).  A group
 *     that the device rejects (GV_E_POOL, GV_E_RANGE) is re-applied request by
 *     request so that only the offending requests fail;
 *   - consecutive neighbour lookups share one gv_neighbors call;
 *   - consecutive PageRank refreshes share one run, warm-started from the
 *     server's previous PageRank vector (a cold start the first time);
 *   - each BC request (one source) runs on its own.
 * PageRank and BC do not hold up the requests behind them: at its place in
 * the arrival order such a step only takes a snapshot of the graph (the
 * relabelled in-adjacency, or the out-adjacency as CSR) on the graph's
 * stream and hands it to a second worker, which computes on its own stream
 * while later updates and lookups go on.  A snapshot occupies one of two
 * slots until its result is out; while both are held the next PageRank or
 * BC step waits for one (admission by snapshot memory).  The analytics
 * kernels are sized to occupy analytics_cu_share of the compute units, so
 * the many short kernels of an update step do not queue behind a
 * whole-chip analytics kernel.
 * A query sees every update that arrived before it and none after it.
 * Latency = done_us - arrival_us, where done_us is taken after the request's
 * device work has finished.
 *
 * While a server runs it owns its graph: no other thread may call into that
 * handle until gv_server_destroy returns.
 */
#ifndef GV_SERVE_H
#define GV_SERVE_H

#include "gv.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gv_server gv_server;

#define GV_REQ_INSERT 0
#define GV_REQ_DELETE 1
#define GV_REQ_NEIGHBORS 2
#define GV_REQ_PAGERANK 3
#define GV_REQ_BC 4

typedef struct gv_request {
  int kind;             /* GV_REQ_* */
  int mem;              /* updates: GV_MEM_HOST or GV_MEM_DEVICE for src/dst */
  const uint32_t *src;  /* updates: n (src,dst) pairs */
  const uint32_t *dst;
  uint64_t n;           /* updates: pairs; neighbours: query vertices */
  const uint32_t *vs;   /* neighbours: n host vertex IDs */
  uint64_t *offs;       /* neighbours: n+1 host offsets out, or NULL */
  uint32_t *out;        /* neighbours: host neighbour lists out, or NULL */
  uint64_t out_cap;     /* neighbours: entries available in out */
  uint32_t source;      /* bc: the source vertex */
  float *x;             /* pagerank / bc: V host floats out, or NULL */
} gv_request;

typedef struct gv_response {
  int status;           /* GV_OK or a GV_E_* code */
  int kind;
  double arrival_us;    /* microseconds since gv_server_create */
  double start_us;      /* the serving step started */
  double done_us;       /* the step's device work finished */
  uint64_t step;        /* index of the serving step */
  uint32_t step_size;   /* requests served by that step */
  int iters;            /* pagerank: iterations */
  double residual;      /* pagerank: L1 residual of the last update */
  uint64_t out_len;     /* neighbours: neighbours of all n vertices (GV_E_INVALID
                           with this set when out_cap is too small) */
} gv_response;

typedef struct gv_server_opts {
  uint64_t max_step_pairs;    /* update pairs staged per step (larger requests run alone) */
  uint32_t max_step_requests; /* requests per step */
  uint64_t headroom_blocks;   /* pool blocks admission keeps free */
  float compact_fraction;     /* compaction after each update step; < 0 = none */
  float pr_alpha;             /* PageRank refresh: damping */
  double pr_eps;              /*   L1 residual stop */
  int pr_max_iter;            /*   iteration cap */
  float analytics_cu_share;   /* share of the compute units PageRank / BC kernels occupy (the
                                 persistent PageRank grid, the BC level grids), so update
                                 kernels always find free CUs; 1 = all */
} gv_server_opts;

typedef struct gv_server_stats {
  uint64_t requests_done;
  uint64_t steps;
  uint64_t update_pairs_applied; /* pairs of update requests that succeeded */
  uint64_t refused;              /* requests that ended with an error status */
  uint64_t free_blocks;          /* pool state after the last update step */
  uint64_t total_blocks;
} gv_server_stats;

gv_status gv_server_opts_default(gv_server_opts *opts);
gv_status gv_server_create(gv_graph *g, const gv_server_opts *opts, gv_server **out);
/* Thread-safe enqueue; the request's arrays must stay valid until it is done.
 * *ticket names the request for gv_server_wait. */
gv_status gv_server_submit(gv_server *s, const gv_request *req, uint64_t *ticket);
gv_status gv_server_wait(gv_server *s, uint64_t ticket, gv_response *out);
/* Open-loop replay: submits reqs[i] at arrival_us[i] microseconds after the
 * call (non-decreasing), then waits for all of them; resp[i] answers reqs[i]. */
gv_status gv_server_replay(gv_server *s, const gv_request *reqs, const double *arrival_us,
                           uint64_t n, gv_response *resp);
gv_status gv_server_drain(gv_server *s);
gv_status gv_server_get_stats(gv_server *s, gv_server_stats *out);
/* Drains, stops the worker and releases the server (not the graph). */
gv_status gv_server_destroy(gv_server *s);

#ifdef __cplusplus
}
#endif
#endif
