// gv_api.hip -- the C ABI (include/gv.h) over the store and analytics
// kernels.  Host orchestration only: every device step is a kernel launch on
// the handle's own stream; the host waits only where a call must return a
// value (stats, sizes, analytic results).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "gv_analytics.h"
#include "gv_store.h"

namespace gv {
static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
gv_status hip_fail(hipError_t e, const char *what) {
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation ? GV_E_NOMEM : GV_E_HIP;
}
}  // namespace gv

using namespace gv;

struct gv_capture;

struct gv_graph {
  gv_config cfg;
  uint32_t V = 0;
  int directed = 0;
  int has_in = 0;
  hipStream_t stream = nullptr;
  Store st[2];
  Scratch sc[2];
  uint32_t *vlist = nullptr;
  PrWork pr;
  TcdWork tcd;
  BcmWork bcm;
  Ctl *h_ctl = nullptr;  // pinned, 2 entries
  uint32_t *in_src = nullptr, *in_dst = nullptr;
  uint64_t in_cap = 0;
  uint64_t *grp_off = nullptr;   // device copy of a group's batch offsets
  uint32_t *grp_kind = nullptr;  // device copy of a group's batch kinds
  uint32_t grp_cap = 0;
  // Stream capture (gv_capture_begin .. gv_capture_end).
  int capturing = 0;
  gv_status cap_err = GV_OK;        // first failure of a call made while capturing
  uint64_t cap_pairs = 0;           // update pairs recorded in the open capture
  std::vector<void *> cap_pinned;   // pinned host copies read by recorded copies
  uint64_t buf_gen = 0;             // bumped whenever a buffer a capture may use moves
  std::vector<gv_capture *> captures;  // live captures (detached by gv_destroy)
  // Neighbour-query buffers, grown on demand and kept (a free per call would
  // wait for all device work, including other streams').
  uint64_t *nb_cnt = nullptr, *nb_k0 = nullptr, *nb_k1 = nullptr;
  uint32_t *nb_dvs = nullptr, *nb_dout = nullptr;
  uint32_t *nb_ist = nullptr, *nb_cur = nullptr, *nb_iq = nullptr;  // key emission items
  void *nb_tmp = nullptr;
  size_t nb_cnt_cap = 0, nb_k_cap = 0, nb_dvs_cap = 0, nb_dout_cap = 0, nb_tmp_cap = 0;
  size_t nb_ist_cap = 0, nb_cur_cap = 0, nb_iq_cap = 0;
  // Chung-Lu cumulative table on the device, cached by its parameters.
  uint64_t *cl_cum = nullptr;
  uint32_t cl_V = 0;
  double cl_gamma = 0, cl_avg = 0, cl_max = 0;
};

struct gv_capture {
  gv_graph *g = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t gen = 0;
  uint64_t pr_gen = 0;  // the PageRank workspace's generation at capture
  uint64_t pairs = 0;
  std::vector<void *> pinned;
};

extern "C" {

const char *gv_last_error(void) { return g_err.c_str(); }

gv_status gv_test_set_sort_spin_limit(uint32_t limit) {
  sort_set_spin_limit(limit);
  return GV_OK;
}

gv_status gv_config_default(gv_config *cfg) {
  if (!cfg) return GV_E_INVALID;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->block_slots = 64;
  cfg->pool_blocks = 0;
  cfg->pool_fraction = 0.25;
  cfg->compact_threshold = 0.25f;
  cfg->pr_thread_max = 8;
  cfg->pr_group_max = 96;
  cfg->pr_wave_max = 4096;
  cfg->maintain_in_edges = 1;
  cfg->device = 0;
  cfg->pr_layout = 5;
  cfg->bc_bu_ratio = 3.0f;
  return GV_OK;
}

}  // extern "C"

// capturable: the call may be recorded into an open stream capture.
static gv_status check_graph(gv_graph *g, bool capturable = false) {
  if (!g) {
    set_error("null graph handle");
    return GV_E_INVALID;
  }
  if (g->capturing && !capturable) {
    set_error("call not allowed while the handle's stream is being captured");
    if (!g->cap_err) g->cap_err = GV_E_INVALID;
    return GV_E_INVALID;
  }
  hipError_t e = hipSetDevice(g->cfg.device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  return GV_OK;
}

static gv_status read_ctl(gv_graph *g) {
  int n = g->has_in ? 2 : 1;
  for (int k = 0; k < n; ++k)
    GV_HIP(hipMemcpyAsync(&g->h_ctl[k], g->st[k].ctl, sizeof(Ctl), hipMemcpyDeviceToHost,
                          g->stream));
  GV_HIP(hipStreamSynchronize(g->stream));
  return GV_OK;
}

__global__ void k_clear_abort(Ctl *c) { c->abort = 0; }

// Report and clear a sticky abort code left by the device.
static gv_status take_abort(gv_graph *g) {
  unsigned ab = g->h_ctl[0].abort;
  if (g->has_in && !ab) ab = g->h_ctl[1].abort;
  if (!ab) return GV_OK;
  for (int k = 0; k < (g->has_in ? 2 : 1); ++k)
    k_clear_abort<<<1, 1, 0, g->stream>>>(g->st[k].ctl);
  GV_HIP(hipStreamSynchronize(g->stream));
  if (ab == ABORT_RANGE) {
    set_error("vertex ID out of range; batch not applied");
    return GV_E_RANGE;
  }
  if (ab == ABORT_POOL) {
    set_error("edge-block pool exhausted; batch not applied");
    return GV_E_POOL;
  }
  if (ab == ABORT_GROUP) {
    set_error("group offsets or kinds invalid (offsets must start at 0, not decrease and end "
              "within max_pairs; kinds GV_OP_INSERT or GV_OP_DELETE); group not applied");
    return GV_E_INVALID;
  }
  if (ab == ABORT_SORT) {
    set_error("radix sort of the staged keys failed (a tile's look-back gave up); nothing applied");
    return GV_E_HIP;
  }
  set_error("device abort");
  return GV_E_HIP;
}

// Bring a host or device (src,dst) batch to device pointers.
static gv_status device_inputs(gv_graph *g, const uint32_t *src, const uint32_t *dst, uint64_t n,
                               int mem, const uint32_t **ds, const uint32_t **dd) {
  if (mem == GV_MEM_DEVICE) {
    *ds = src;
    *dd = dst;
    return GV_OK;
  }
  if (mem != GV_MEM_HOST) {
    set_error("mem must be GV_MEM_HOST or GV_MEM_DEVICE");
    return GV_E_INVALID;
  }
  if (g->capturing) {
    set_error("a captured update needs device inputs (GV_MEM_DEVICE)");
    return GV_E_INVALID;
  }
  if (n > g->in_cap) {
    hipFree(g->in_src);
    hipFree(g->in_dst);
    g->in_src = g->in_dst = nullptr;
    uint64_t cap = std::max<uint64_t>(n, 1024);
    GV_HIP(hipMalloc(&g->in_src, cap * sizeof(uint32_t)));
    GV_HIP(hipMalloc(&g->in_dst, cap * sizeof(uint32_t)));
    g->in_cap = cap;
  }
  GV_HIP(hipMemcpyAsync(g->in_src, src, n * sizeof(uint32_t), hipMemcpyHostToDevice, g->stream));
  GV_HIP(hipMemcpyAsync(g->in_dst, dst, n * sizeof(uint32_t), hipMemcpyHostToDevice, g->stream));
  *ds = g->in_src;
  *dd = g->in_dst;
  return GV_OK;
}

static void fill_stats(gv_graph *g, uint64_t n, gv_batch_stats *stats) {
  const Ctl &c = g->h_ctl[0];
  stats->requested = n;
  stats->directed_entries = c.n_valid;
  stats->applied = c.applied;
  stats->touched = c.n_segs;
  stats->blocks_allocated = c.need_blocks + (g->has_in ? g->h_ctl[1].need_blocks : 0);
  stats->live_edges = c.live_edges;
}

// Scratch for m staged keys in store k.  Inside a capture the buffers must
// already be large enough (gv_reserve), since recorded launches keep their
// addresses; outside, a reallocation bumps buf_gen so that older captures
// refuse to launch.
static gv_status reserve_keys(gv_graph *g, int k, uint64_t m) {
  if (scratch_fits(g->sc[k], g->st[k], m)) return GV_OK;
  if (g->capturing) {
    set_error("capture: staging buffers too small; call gv_reserve before gv_capture_begin");
    return GV_E_INVALID;
  }
  ++g->buf_gen;
  return scratch_reserve(g->sc[k], g->st[k], m);
}

static gv_status reserve_groups(gv_graph *g, uint32_t nb) {
  if (nb <= g->grp_cap) return GV_OK;
  if (g->capturing) {
    set_error("capture: group buffers too small; call gv_reserve before gv_capture_begin");
    return GV_E_INVALID;
  }
  ++g->buf_gen;
  hipFree(g->grp_off);
  hipFree(g->grp_kind);
  g->grp_off = nullptr;
  g->grp_kind = nullptr;
  g->grp_cap = 0;
  uint32_t cap = std::max<uint32_t>(nb, 64);
  GV_HIP(hipMalloc(&g->grp_off, ((size_t)cap + 1) * sizeof(uint64_t)));
  GV_HIP(hipMalloc(&g->grp_kind, (size_t)cap * sizeof(uint32_t)));
  g->grp_cap = cap;
  return GV_OK;
}

enum class Op { Insert, Delete };

static gv_status apply_batch(gv_graph *g, const uint32_t *src, const uint32_t *dst, uint64_t n,
                             int mem, gv_batch_stats *stats, Op op) {
  gv_status rc = check_graph(g, true);
  if (rc) return rc;
  if (n && (!src || !dst)) {
    set_error("null batch pointer");
    return GV_E_INVALID;
  }
  if (g->capturing && stats) {
    set_error("a captured update cannot return stats; pass NULL");
    return GV_E_INVALID;
  }
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (n == 0) return GV_OK;
  if (n > 0x7FFFFFFFull) {
    set_error("batch larger than 2^31 pairs; split it");
    return GV_E_INVALID;
  }
  const uint32_t *ds = nullptr, *dd = nullptr;
  rc = device_inputs(g, src, dst, n, mem, &ds, &dd);
  if (rc) return rc;
  hipStream_t s = g->stream;
  const int nst = g->has_in ? 2 : 1;
  uint64_t m[2] = {0, 0};
  uint64_t *keys[2] = {nullptr, nullptr};
  for (int k = 0; k < nst; ++k) {
    const int mirror = g->directed ? 0 : 1;
    rc = reserve_keys(g, k, mirror ? 2 * n : n);
    if (rc) return rc;
  }
  if (g->capturing) g->cap_pairs += n;
  for (int k = 0; k < nst; ++k) {
    const int mirror = g->directed ? 0 : 1;
    keys[k] = stage(g->st[k], g->sc[k], ds, dd, n, mirror, k == 1 ? 1 : 0, m[k], s);
  }
  if (nst > 1) abort_join(g->st[0], g->st[1], g->sc[0], g->sc[1], s);
  for (int k = 0; k < nst; ++k) {
    segment(g->st[k], g->sc[k], keys[k], m[k], s);
    if (op == Op::Insert) insert_plan(g->st[k], g->sc[k], keys[k], m[k], s);
    else delete_apply(g->st[k], g->sc[k], keys[k], m[k], s);
  }
  if (op == Op::Insert) {
    commit_pair(g->st[0], g->sc[0], nst > 1 ? &g->st[1] : nullptr, nst > 1 ? &g->sc[1] : nullptr, s);
    for (int k = 0; k < nst; ++k) insert_write(g->st[k], g->sc[k], keys[k], m[k], s);
  }
  GV_HIP(hipGetLastError());
  if (!stats) return GV_OK;
  rc = read_ctl(g);
  if (rc) return rc;
  fill_stats(g, n, stats);
  return take_abort(g);
}

static void group_pipeline(gv_graph *g, const uint32_t *ds, const uint32_t *dd, uint64_t n,
                           const uint64_t *boff, const uint32_t *kinds, uint32_t nb);
static void group_compact(gv_graph *g, uint64_t n, float compact_fraction);

// A group of batches in one pipeline (see gv_batch_apply in gv.h).
static gv_status apply_group(gv_graph *g, const uint32_t *src, const uint32_t *dst,
                             const uint64_t *offsets, const int *kinds, uint32_t nb, int mem,
                             float compact_fraction, gv_batch_stats *stats) {
  gv_status rc = check_graph(g, true);
  if (rc) return rc;
  if (g->capturing && stats) {
    set_error("a captured update cannot return stats; pass NULL");
    return GV_E_INVALID;
  }
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (nb == 0) return GV_OK;
  if (!offsets || !kinds) {
    set_error("null offsets or kinds");
    return GV_E_INVALID;
  }
  const uint64_t n = offsets[nb];
  if (offsets[0] != 0) {
    set_error("offsets[0] must be 0");
    return GV_E_INVALID;
  }
  for (uint32_t b = 0; b < nb; ++b) {
    if (offsets[b + 1] < offsets[b]) {
      set_error("batch offsets must be non-decreasing");
      return GV_E_INVALID;
    }
    if (kinds[b] != GV_OP_INSERT && kinds[b] != GV_OP_DELETE) {
      set_error("batch kind must be GV_OP_INSERT or GV_OP_DELETE");
      return GV_E_INVALID;
    }
  }
  if (2 * id_bits(g->V) + id_bits(nb) > 64) {
    set_error("too many batches in one group for this vertex count");
    return GV_E_INVALID;
  }
  if (n > 0x7FFFFFFFull) {
    set_error("group larger than 2^31 pairs; split it");
    return GV_E_INVALID;
  }
  if (n && (!src || !dst)) {
    set_error("null batch pointer");
    return GV_E_INVALID;
  }
  hipStream_t s = g->stream;
  if (n) {
    const uint32_t *ds = nullptr, *dd = nullptr;
    rc = device_inputs(g, src, dst, n, mem, &ds, &dd);
    if (rc) return rc;
    rc = reserve_groups(g, nb);
    if (rc) return rc;
    const int nst = g->has_in ? 2 : 1;
    for (int k = 0; k < nst; ++k) {
      rc = reserve_keys(g, k, g->directed ? n : 2 * n);
      if (rc) return rc;
    }
    // Offsets and kinds go through one host block: pageable for an eager call
    // (a pageable host-to-device copy has consumed its source when
    // hipMemcpyAsync returns), pinned and kept alive by the capture when the
    // copy is recorded, since every replay reads it again.
    const size_t ob = ((size_t)nb + 1) * sizeof(uint64_t), kb = (size_t)nb * sizeof(uint32_t);
    std::vector<uint8_t> eager_buf;
    uint8_t *hb = nullptr;
    if (g->capturing) {
      void *p = nullptr;
      GV_HIP(hipHostMalloc(&p, ob + kb, hipHostMallocDefault));
      g->cap_pinned.push_back(p);
      hb = (uint8_t *)p;
      g->cap_pairs += n;
    } else {
      eager_buf.resize(ob + kb);
      hb = eager_buf.data();
    }
    std::memcpy(hb, offsets, ob);
    uint32_t *hk = (uint32_t *)(hb + ob);
    for (uint32_t b = 0; b < nb; ++b) hk[b] = kinds[b] == GV_OP_DELETE ? OP_DELETE : OP_INSERT;
    GV_HIP(hipMemcpyAsync(g->grp_off, hb, ob, hipMemcpyHostToDevice, s));
    GV_HIP(hipMemcpyAsync(g->grp_kind, hk, kb, hipMemcpyHostToDevice, s));
    group_pipeline(g, ds, dd, n, g->grp_off, g->grp_kind, nb);
  }
  group_compact(g, n, compact_fraction);
  GV_HIP(hipGetLastError());
  if (!stats) return GV_OK;
  rc = read_ctl(g);
  if (rc) return rc;
  fill_stats(g, n, stats);
  return take_abort(g);
}

// The group pipeline over device pairs: staging (entries past boff[nb]
// dropped), one sort, selection and segmentation, the merged probe, the
// joint pool commit and the writes, on the handle's stream.
static void group_pipeline(gv_graph *g, const uint32_t *ds, const uint32_t *dd, uint64_t n,
                           const uint64_t *boff, const uint32_t *kinds, uint32_t nb) {
  hipStream_t s = g->stream;
  const int nst = g->has_in ? 2 : 1;
  uint64_t m[2] = {0, 0};
  uint64_t *keys[2] = {nullptr, nullptr};
  for (int k = 0; k < nst; ++k) {
    const int mirror = g->directed ? 0 : 1;
    keys[k] = stage_group(g->st[k], g->sc[k], ds, dd, n, boff, kinds, nb, mirror,
                          k == 1 ? 1 : 0, m[k], s);
  }
  if (nst > 1) abort_join(g->st[0], g->st[1], g->sc[0], g->sc[1], s);
  for (int k = 0; k < nst; ++k) {
    segment_items(g->st[k], g->sc[k], m[k], s);
    mixed_plan(g->st[k], g->sc[k], keys[k], m[k], s);
  }
  commit_pair(g->st[0], g->sc[0], nst > 1 ? &g->st[1] : nullptr, nst > 1 ? &g->sc[1] : nullptr,
              s);
  for (int k = 0; k < nst; ++k) insert_write(g->st[k], g->sc[k], keys[k], m[k], s);
}

// Compaction after a group of n pairs (see gv_batch_apply).
static void group_compact(gv_graph *g, uint64_t n, float compact_fraction) {
  hipStream_t s = g->stream;
  if (compact_fraction >= 0.f) {
    // with updates: only the vertices this group touched can have gained
    // tombstones, so only they are candidates; an empty group compacts all
    for (int k = 0; k < (g->has_in ? 2 : 1); ++k) {
      if (n) {
        compact_touched(g->st[k], g->sc[k], g->directed ? n : 2 * n, compact_fraction, g->vlist,
                        s);
      } else {
        compact(g->st[k], compact_fraction, false, s);
      }
    }
  }
}

// A group whose pairs, batch offsets and kinds are all in device memory:
// up to max_batches batches over up to max_pairs pairs, read when the call
// runs on the stream (a producer fills them between launches of a capture).
static gv_status apply_group_device(gv_graph *g, const uint32_t *src, const uint32_t *dst,
                                    const uint64_t *offsets, const int *kinds,
                                    uint32_t max_batches, uint64_t max_pairs,
                                    float compact_fraction) {
  gv_status rc = check_graph(g, true);
  if (rc) return rc;
  if (max_batches == 0 || max_pairs == 0) return GV_OK;
  if (!src || !dst || !offsets || !kinds) {
    set_error("null device pointer");
    return GV_E_INVALID;
  }
  if (2 * id_bits(g->V) + id_bits(max_batches) > 64) {
    set_error("too many batches in one group for this vertex count");
    return GV_E_INVALID;
  }
  if (max_pairs > 0x7FFFFFFFull) {
    set_error("group larger than 2^31 pairs; split it");
    return GV_E_INVALID;
  }
  // staging for the whole capacity (inside a capture it must already be
  // reserved: reserve_keys then refuses instead of allocating)
  const int nst = g->has_in ? 2 : 1;
  for (int k = 0; k < nst; ++k) {
    rc = reserve_keys(g, k, g->directed ? max_pairs : 2 * max_pairs);
    if (rc) return rc;
  }
  if (g->capturing) g->cap_pairs += max_pairs;
  group_pipeline(g, src, dst, max_pairs, reinterpret_cast<const uint64_t *>(offsets),
                 reinterpret_cast<const uint32_t *>(kinds), max_batches);
  group_compact(g, max_pairs, compact_fraction);
  GV_HIP(hipGetLastError());
  return GV_OK;
}

namespace gv {

// Snapshots for analytics that run off the graph's stream (the request loop):
// both are ordered on the graph's stream after every call issued before them
// and read the store only there.
gv_status graph_pr_prepare(gv_graph *g, PrWork &w) {
  gv_status rc = check_graph(g);
  if (rc) return rc;
  if (g->directed && !g->has_in) {
    set_error("PageRank on a directed graph needs maintain_in_edges");
    return GV_E_UNSUPPORTED;
  }
  const Store &in = g->has_in ? g->st[1] : g->st[0];
  return pagerank_snap_prepare(in, g->st[0].deg, w, g->stream, g->cfg.pr_layout);
}

gv_status graph_bc_prepare(gv_graph *g, PrWork &w) {
  gv_status rc = check_graph(g);
  if (rc) return rc;
  return pr_snapshot(g->st[0], g->st[0].deg, false, w, g->stream, 0);
}

}  // namespace gv

extern "C" {

// A failure while the stream is being captured poisons the capture.
static gv_status note_capture(gv_graph *g, gv_status rc) {
  if (rc && g && g->capturing && !g->cap_err) g->cap_err = rc;
  return rc;
}

gv_status gv_batch_apply(gv_graph *g, const uint32_t *src, const uint32_t *dst,
                         const uint64_t *offsets, const int *kinds, uint32_t n_batches, int mem,
                         float compact_fraction, gv_batch_stats *stats) {
  return note_capture(
      g, apply_group(g, src, dst, offsets, kinds, n_batches, mem, compact_fraction, stats));
}

gv_status gv_batch_apply_device(gv_graph *g, const uint32_t *src, const uint32_t *dst,
                                const uint64_t *offsets, const int *kinds, uint32_t max_batches,
                                uint64_t max_pairs, float compact_fraction) {
  return note_capture(g, apply_group_device(g, src, dst, offsets, kinds, max_batches, max_pairs,
                                            compact_fraction));
}

gv_status gv_create(const gv_config *cfg_in, uint32_t num_vertices, int directed,
                    gv_graph **out) {
  if (!out) {
    set_error("null output handle");
    return GV_E_INVALID;
  }
  *out = nullptr;
  gv_config cfg;
  if (cfg_in) cfg = *cfg_in; else gv_config_default(&cfg);
  if (cfg.block_slots != 32 && cfg.block_slots != 64 && cfg.block_slots != 128) {
    set_error("block_slots must be 32, 64 or 128");
    return GV_E_INVALID;
  }
  if (num_vertices >= 0xFFFFFFFFu) {
    set_error("num_vertices must be below 2^32-1 (0xFFFFFFFF marks a tombstone)");
    return GV_E_INVALID;
  }
  if (cfg.pr_thread_max > cfg.pr_group_max || cfg.pr_group_max > cfg.pr_wave_max ||
      cfg.pr_wave_max == 0) {
    set_error("PageRank bins must satisfy thread_max <= group_max <= wave_max, wave_max > 0");
    return GV_E_INVALID;
  }
  if (cfg.pr_layout != 1 && cfg.pr_layout != 4 && cfg.pr_layout != 5) {
    set_error("pr_layout must be 1, 4 or 5");
    return GV_E_INVALID;
  }
  if (!(cfg.bc_bu_ratio >= 0.f) || !std::isfinite(cfg.bc_bu_ratio)) {
    set_error("bc_bu_ratio must be finite and >= 0");
    return GV_E_INVALID;
  }
  hipError_t e = hipSetDevice(cfg.device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  gv_graph *g = new gv_graph();
  g->cfg = cfg;
  g->V = num_vertices;
  g->directed = directed ? 1 : 0;
  g->has_in = (directed && cfg.maintain_in_edges) ? 1 : 0;
  const int nst = g->has_in ? 2 : 1;
  uint64_t P = cfg.pool_blocks;
  // slots, free-stack entry, table entries, fences (one word per 16 slots)
  const uint64_t bytes_per_block = (uint64_t)cfg.block_slots * 4 + 4 + 16 + cfg.block_slots / 4;
  if (P == 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = (size_t)1 << 32;
    double frac = cfg.pool_fraction > 0 && cfg.pool_fraction < 0.9 ? cfg.pool_fraction : 0.25;
    P = (uint64_t)((double)fr * frac / (double)bytes_per_block) / nst;
  }
  P = std::max<uint64_t>(P, 64);
  // slot indices are 32-bit (block ID * B + offset) and 0xFFFFFFFF marks
  // "no slot", so every slot index must stay below it
  P = std::min<uint64_t>(P, (1ull << 32) / cfg.block_slots - 1);
  uint64_t A = std::min<uint64_t>(4 * P + (uint64_t)num_vertices + 1024, 0xFFFFFFF0ull);
  gv_status rc = GV_OK;
  do {
    if ((e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking)) != hipSuccess) {
      rc = hip_fail(e, "hipStreamCreate");
      break;
    }
    for (int k = 0; k < nst && !rc; ++k) rc = store_alloc(g->st[k], num_vertices, cfg.block_slots, P, A);
    if (rc) break;
    if ((e = hipMalloc(&g->vlist, (size_t)(num_vertices ? num_vertices : 1) * 4)) != hipSuccess) {
      rc = hip_fail(e, "hipMalloc");
      break;
    }
    if ((e = hipHostMalloc(&g->h_ctl, 2 * sizeof(Ctl), hipHostMallocDefault)) != hipSuccess) {
      rc = hip_fail(e, "hipHostMalloc");
      break;
    }
    std::memset(g->h_ctl, 0, 2 * sizeof(Ctl));
    for (int k = 0; k < nst; ++k) store_reset(g->st[k], g->stream);
    if ((e = hipStreamSynchronize(g->stream)) != hipSuccess) rc = hip_fail(e, "store init");
  } while (0);
  if (rc) {
    gv_destroy(g);
    return rc;
  }
  *out = g;
  return GV_OK;
}

gv_status gv_destroy(gv_graph *g) {
  if (!g) return GV_OK;
  hipSetDevice(g->cfg.device);
  if (g->capturing) {
    hipGraph_t open = nullptr;
    hipStreamEndCapture(g->stream, &open);
    if (open) hipGraphDestroy(open);
    g->capturing = 0;
  }
  for (void *p : g->cap_pinned) hipHostFree(p);
  for (gv_capture *c : g->captures) c->g = nullptr;
  hipFree(g->cl_cum);
  hipFree(g->nb_cnt);
  hipFree(g->nb_k0);
  hipFree(g->nb_k1);
  hipFree(g->nb_dvs);
  hipFree(g->nb_dout);
  hipFree(g->nb_ist);
  hipFree(g->nb_cur);
  hipFree(g->nb_iq);
  hipFree(g->nb_tmp);
  if (g->stream) hipStreamSynchronize(g->stream);
  for (int k = 0; k < 2; ++k) {
    store_release(g->st[k]);
    scratch_release(g->sc[k]);
  }
  pr_release(g->pr);
  tcd_release(g->tcd);
  bcm_release(g->bcm);
  hipFree(g->vlist);
  hipFree(g->in_src);
  hipFree(g->in_dst);
  hipFree(g->grp_off);
  hipFree(g->grp_kind);
  if (g->h_ctl) hipHostFree(g->h_ctl);
  if (g->stream) hipStreamDestroy(g->stream);
  delete g;
  return GV_OK;
}

gv_status gv_bulk_build(gv_graph *g, const uint32_t *src, const uint32_t *dst, uint64_t n, int mem,
                        gv_batch_stats *stats) {
  gv_status rc = check_graph(g);
  if (rc) return rc;
  for (int k = 0; k < (g->has_in ? 2 : 1); ++k) store_reset(g->st[k], g->stream);
  return apply_batch(g, src, dst, n, mem, stats, Op::Insert);
}

gv_status gv_batch_insert(gv_graph *g, const uint32_t *src, const uint32_t *dst, uint64_t n,
                          int mem, gv_batch_stats *stats) {
  return note_capture(g, apply_batch(g, src, dst, n, mem, stats, Op::Insert));
}

gv_status gv_batch_delete(gv_graph *g, const uint32_t *src, const uint32_t *dst, uint64_t n,
                          int mem, gv_batch_stats *stats) {
  return note_capture(g, apply_batch(g, src, dst, n, mem, stats, Op::Delete));
}

gv_status gv_compact(gv_graph *g, float frac, uint64_t *vertices_compacted) {
  gv_status rc = check_graph(g, true);
  if (rc) return rc;
  if (g->capturing && vertices_compacted) {
    set_error("a captured compaction cannot return a count; pass NULL");
    return g->cap_err = GV_E_INVALID;
  }
  if (frac < 0.f) frac = g->cfg.compact_threshold;
  for (int k = 0; k < (g->has_in ? 2 : 1); ++k)
    compact(g->st[k], frac, vertices_compacted != nullptr, g->stream);
  GV_HIP(hipGetLastError());
  if (!vertices_compacted) return GV_OK;
  rc = read_ctl(g);
  if (rc) return rc;
  *vertices_compacted = g->h_ctl[0].compacted;
  return GV_OK;
}

gv_status gv_sync(gv_graph *g) {
  gv_status rc = check_graph(g);
  if (rc) return rc;
  rc = read_ctl(g);
  if (rc) return rc;
  return take_abort(g);
}

gv_status gv_get_stream(gv_graph *g, void **hip_stream) {
  if (!g || !hip_stream) return GV_E_INVALID;
  *hip_stream = (void *)g->stream;
  return GV_OK;
}

gv_status gv_num_vertices(gv_graph *g, uint32_t *v) {
  if (!g || !v) return GV_E_INVALID;
  *v = g->V;
  return GV_OK;
}

gv_status gv_num_edges(gv_graph *g, uint64_t *live) {
  gv_status rc = check_graph(g);
  if (rc) return rc;
  if (!live) return GV_E_INVALID;
  rc = read_ctl(g);
  if (rc) return rc;
  *live = g->h_ctl[0].live_edges;
  return GV_OK;
}

gv_status gv_pool_stats(gv_graph *g, uint64_t *free_blocks, uint64_t *total_blocks) {
  gv_status rc = check_graph(g);
  if (rc) return rc;
  rc = read_ctl(g);
  if (rc) return rc;
  uint64_t f = 0, t = 0;
  for (int k = 0; k < (g->has_in ? 2 : 1); ++k) {
    f += g->h_ctl[k].free_top;
    t += g->st[k].P;
  }
  if (free_blocks) *free_blocks = f;
  if (total_blocks) *total_blocks = t;
  return GV_OK;
}

gv_status gv_degrees(gv_graph *g, uint32_t *deg, int mem) {
  gv_status rc = check_graph(g);
  if (rc) return rc;
  if (!deg) return GV_E_INVALID;
  GV_HIP(hipMemcpyAsync(deg, g->st[0].deg, (size_t)g->V * 4,
                        mem == GV_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                        g->stream));
  GV_HIP(hipStreamSynchronize(g->stream));
  return GV_OK;
}

gv_status gv_export_csr(gv_graph *g, uint64_t *offsets, uint32_t *cols, int mem) {
  gv_status rc = check_graph(g);
  if (rc) return rc;
  if (!offsets) return GV_E_INVALID;
  hipStream_t s = g->stream;
  const uint32_t V = g->V, bV = id_bits(V);
  uint64_t *off = nullptr, *k0 = nullptr, *k1 = nullptr;
  uint32_t *dcols = nullptr, *ist = nullptr, *cur = nullptr, *iq = nullptr;
  void *tmp = nullptr;
  uint64_t E = 0;
  uint32_t items = 0;
  do {
    hipError_t e;
    if ((e = hipMalloc(&off, ((size_t)V + 1) * 8)) != hipSuccess ||
        (e = hipMalloc(&tmp, scan_tmp_bytes((uint64_t)V + 1))) != hipSuccess ||
        (cols && (e = hipMalloc(&ist, ((size_t)V + 1) * 4)) != hipSuccess) ||
        (cols && (e = hipMalloc(&cur, ((size_t)V + 1) * 4)) != hipSuccess)) {
      rc = hip_fail(e, "export alloc");
      break;
    }
    degrees_u64(g->st[0], off, s);
    scan_u64(off, off, V, nullptr, tmp, s);
    if (cols) emit_items(g->st[0], nullptr, V, ist, cur, tmp, s);
    if ((e = hipMemcpyAsync(&E, off + V, 8, hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (cols && (e = hipMemcpyAsync(&items, ist + V, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) ||
        (e = hipStreamSynchronize(s)) != hipSuccess) {
      rc = hip_fail(e, "export");
      break;
    }
    if ((e = hipMemcpyAsync(offsets, off, ((size_t)V + 1) * 8,
                            mem == GV_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                            s)) != hipSuccess) {
      rc = hip_fail(e, "export");
      break;
    }
    if (cols && E > SORT_MAX_KEYS) {
      // the export sorts all E entries at once, and a sort's offsets are 32-bit
      set_error("gv_export_csr: more than 2^32-1 live edges; export a vertex range with "
                "gv_neighbors instead");
      rc = GV_E_UNSUPPORTED;
      break;
    }
    if (cols && E) {
      hipFree(tmp);
      tmp = nullptr;
      if ((e = hipMalloc(&k0, E * 8)) != hipSuccess || (e = hipMalloc(&k1, E * 8)) != hipSuccess ||
          (e = hipMalloc(&tmp, sort_tmp_bytes(E))) != hipSuccess ||
          (e = hipMalloc(&iq, ((size_t)items + 1) * 4)) != hipSuccess) {
        rc = hip_fail(e, "export alloc");
        break;
      }
      if (mem != GV_MEM_DEVICE && (e = hipMalloc(&dcols, E * 4)) != hipSuccess) {
        rc = hip_fail(e, "export alloc");
        break;
      }
      emit_keys(g->st[0], nullptr, V, ist, iq, items, off, cur, k0, s);
      uint64_t *sorted = sort_keys(k0, k1, E, 2 * bV, tmp, s);
      if ((rc = sort_check(tmp, E, 0, 2 * bV, s)) != GV_OK) break;
      uint32_t *cdst = mem == GV_MEM_DEVICE ? cols : dcols;
      keys_to_cols(sorted, E, bV, cdst, s);
      if (mem != GV_MEM_DEVICE &&
          (e = hipMemcpyAsync(cols, dcols, E * 4, hipMemcpyDeviceToHost, s)) != hipSuccess) {
        rc = hip_fail(e, "export");
        break;
      }
    }
    if ((e = hipStreamSynchronize(s)) != hipSuccess) rc = hip_fail(e, "export");
  } while (0);
  hipFree(off);
  hipFree(k0);
  hipFree(k1);
  hipFree(dcols);
  hipFree(tmp);
  hipFree(ist);
  hipFree(cur);
  hipFree(iq);
  return rc;
}

// ----------------------------------------------- allocation-free lookups --
// gv_neighbors_noalloc works in one file-scope device array instead of
// buffers it allocates: the key ping-pong pair of the sort and the sort's
// workspace.  Every GPU the library runs on holds its own copy until the
// process ends; host code reaches it only through the address
// hipGetSymbolAddress returns on that device (the array's name in host code
// is a host-side stand-in, not a device address).
constexpr uint32_t NA_MAX_QUERIES = 1u << 16;
constexpr uint64_t NA_MAX_NEIGHBORS = 1ull << 21;
constexpr size_t NA_KEY_BYTES = NA_MAX_NEIGHBORS * sizeof(uint64_t);
constexpr size_t NA_TMP_BYTES = 5u << 20;  // >= sort_tmp_bytes(NA_MAX_NEIGHBORS), scan of 2^16+1
constexpr size_t NA_SCRATCH_BYTES = 2 * NA_KEY_BYTES + NA_TMP_BYTES;

static __device__ __align__(256) uint8_t scratch[NA_SCRATCH_BYTES];

constexpr int NA_MAX_DEVICES = 64;
static std::mutex g_na_mu;
static uint8_t *g_na_ptr[NA_MAX_DEVICES];

// Device address of this device's scratch (the current device is set by
// check_graph); looked up once per device.
static gv_status noalloc_scratch(int dev, uint8_t **out) {
  if (dev < 0 || dev >= NA_MAX_DEVICES) {
    set_error("gv_neighbors_noalloc: device ordinal out of range");
    return GV_E_INVALID;
  }
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (!g_na_ptr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) {
      set_error("gv_neighbors_noalloc: no device address for the scratch array");
      return GV_E_HIP;
    }
    g_na_ptr[dev] = (uint8_t *)p;
  }
  *out = g_na_ptr[dev];
  return GV_OK;
}

// Grow-only device buffer: reallocated (with headroom) only when too small.
static hipError_t nb_grow(void **p, size_t &cap, size_t need) {
  if (*p && need <= cap) return hipSuccess;
  hipFree(*p);
  *p = nullptr;
  cap = 0;
  const size_t c = need + need / 2 + 256;
  hipError_t e = hipMalloc(p, c);
  if (e == hipSuccess) cap = c;
  return e;
}

// The two key buffers of the neighbour sort share one capacity.
static hipError_t nb_grow_keys(gv_graph *g, size_t need) {
  if (g->nb_k0 && g->nb_k1 && need <= g->nb_k_cap) return hipSuccess;
  hipFree(g->nb_k0);
  hipFree(g->nb_k1);
  g->nb_k0 = g->nb_k1 = nullptr;
  g->nb_k_cap = 0;
  const size_t c = need + need / 2 + 256;
  hipError_t e = hipMalloc(&g->nb_k0, c);
  if (e == hipSuccess) e = hipMalloc(&g->nb_k1, c);
  if (e == hipSuccess) g->nb_k_cap = c;
  return e;
}

gv_status gv_neighbors_noalloc(gv_graph *g, const uint32_t *vs, uint32_t n, uint64_t *offs,
                               uint32_t *out) {
  gv_status rc = check_graph(g);
  if (rc) return rc;
  if (!offs || (n && !vs)) {
    set_error("gv_neighbors_noalloc: null pointer");
    return GV_E_INVALID;
  }
  if (n > NA_MAX_QUERIES) {
    set_error("gv_neighbors_noalloc: more than 65536 queries in one call (the no-alloc maximum)");
    return GV_E_INVALID;
  }
  hipStream_t s = g->stream;
  const uint32_t V = g->V, bV = id_bits(V), qb = id_bits(n ? n : 1);
  if (qb + bV > 64) {
    set_error("too many queries for one call");
    return GV_E_INVALID;
  }
  if (sort_tmp_bytes(NA_MAX_NEIGHBORS) > NA_TMP_BYTES ||
      scan_tmp_bytes((uint64_t)NA_MAX_QUERIES + 1) > NA_TMP_BYTES) {
    set_error("gv_neighbors_noalloc: scratch layout too small for its limits");
    return GV_E_HIP;
  }
  uint8_t *base = nullptr;
  if ((rc = noalloc_scratch(g->cfg.device, &base)) != GV_OK) return rc;
  uint64_t *k0 = (uint64_t *)base;
  uint64_t *k1 = (uint64_t *)(base + NA_KEY_BYTES);
  void *tmp = base + 2 * NA_KEY_BYTES;
  // counts go straight into the caller's offsets, scanned in place
  neighbor_counts(g->st[0], vs, n, offs, s);
  scan_u64(offs, offs, n, nullptr, tmp, s);
  uint64_t T = 0;
  hipError_t e;
  if ((e = hipMemcpyAsync(&T, offs + n, 8, hipMemcpyDeviceToHost, s)) != hipSuccess ||
      (e = hipStreamSynchronize(s)) != hipSuccess)
    return hip_fail(e, "neighbors_noalloc");
  if (!out || !T) return GV_OK;
  if (T > NA_MAX_NEIGHBORS) {
    set_error("gv_neighbors_noalloc: " + std::to_string(T) +
              " neighbours exceed the no-alloc maximum of 2097152 per call; split the query");
    return GV_E_INVALID;
  }
  neighbor_keys(g->st[0], vs, n, offs, qb, k0, s);
  uint64_t *sorted = sort_keys(k0, k1, T, qb + bV, tmp, s);
  if ((rc = sort_check(tmp, T, 0, qb + bV, s)) != GV_OK) return rc;
  keys_to_cols(sorted, T, bV, out, s);
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(e, "neighbors_noalloc");
  return GV_OK;
}

gv_status gv_neighbors(gv_graph *g, const uint32_t *vs, uint32_t n, uint64_t *offs, uint32_t *out,
                       int mem) {
  gv_status rc = check_graph(g);
  if (rc) return rc;
  if (!offs || (n && !vs)) return GV_E_INVALID;
  hipStream_t s = g->stream;
  const uint32_t V = g->V, bV = id_bits(V), qb = id_bits(n ? n : 1);
  if (qb + bV > 64) {
    set_error("too many queries for one call");
    return GV_E_INVALID;
  }
  if (mem == GV_MEM_HOST)
    for (uint32_t i = 0; i < n; ++i)
      if (vs[i] >= V) {
        set_error("query vertex out of range");
        return GV_E_RANGE;
      }
  const hipMemcpyKind out_kind = mem == GV_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  uint64_t T = 0;
  hipError_t e;
  const uint64_t *cnt = nullptr;
  if ((e = nb_grow((void **)&g->nb_cnt, g->nb_cnt_cap, ((size_t)n + 1) * 8)) != hipSuccess ||
      (e = nb_grow(&g->nb_tmp, g->nb_tmp_cap, scan_tmp_bytes((uint64_t)n + 1))) != hipSuccess)
    return hip_fail(e, "neighbors alloc");
  cnt = g->nb_cnt;
  const uint32_t *qv = vs;
  if (mem != GV_MEM_DEVICE && n) {
    if ((e = nb_grow((void **)&g->nb_dvs, g->nb_dvs_cap, (size_t)n * 4)) != hipSuccess ||
        (e = hipMemcpyAsync(g->nb_dvs, vs, (size_t)n * 4, hipMemcpyHostToDevice, s)) !=
            hipSuccess)
      return hip_fail(e, "neighbors");
    qv = g->nb_dvs;
  }
  if (out && n &&
      ((e = nb_grow((void **)&g->nb_ist, g->nb_ist_cap, ((size_t)n + 1) * 4)) != hipSuccess ||
       (e = nb_grow((void **)&g->nb_cur, g->nb_cur_cap, ((size_t)n + 1) * 4)) != hipSuccess))
    return hip_fail(e, "neighbors alloc");
  neighbor_counts(g->st[0], qv, n, g->nb_cnt, s);
  scan_u64(g->nb_cnt, g->nb_cnt, n, nullptr, g->nb_tmp, s);
  uint32_t items = 0;
  if (out && n) emit_items(g->st[0], qv, n, g->nb_ist, g->nb_cur, g->nb_tmp, s);
  if ((e = hipMemcpyAsync(&T, cnt + n, 8, hipMemcpyDeviceToHost, s)) != hipSuccess ||
      (out && n &&
       (e = hipMemcpyAsync(&items, g->nb_ist + n, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) ||
      (e = hipMemcpyAsync(offs, cnt, ((size_t)n + 1) * 8, out_kind, s)) != hipSuccess ||
      (e = hipStreamSynchronize(s)) != hipSuccess)
    return hip_fail(e, "neighbors");
  if (out && T > SORT_MAX_KEYS) {
    set_error("gv_neighbors: more than 2^32-1 neighbours in one call; split the query");
    return GV_E_UNSUPPORTED;
  }
  if (out && T) {
    if ((e = nb_grow_keys(g, T * 8)) != hipSuccess ||
        (e = nb_grow(&g->nb_tmp, g->nb_tmp_cap, sort_tmp_bytes(T))) != hipSuccess ||
        (e = nb_grow((void **)&g->nb_iq, g->nb_iq_cap, ((size_t)items + 1) * 4)) != hipSuccess ||
        (mem != GV_MEM_DEVICE &&
         (e = nb_grow((void **)&g->nb_dout, g->nb_dout_cap, T * 4)) != hipSuccess))
      return hip_fail(e, "neighbors alloc");
    emit_keys(g->st[0], qv, n, g->nb_ist, g->nb_iq, items, g->nb_cnt, g->nb_cur, g->nb_k0, s);
    uint64_t *sorted = sort_keys(g->nb_k0, g->nb_k1, T, qb + bV, g->nb_tmp, s);
    if ((rc = sort_check(g->nb_tmp, T, 0, qb + bV, s)) != GV_OK) return rc;
    uint32_t *cdst = mem == GV_MEM_DEVICE ? out : g->nb_dout;
    keys_to_cols(sorted, T, bV, cdst, s);
    if (mem != GV_MEM_DEVICE &&
        (e = hipMemcpyAsync(out, g->nb_dout, T * 4, hipMemcpyDeviceToHost, s)) != hipSuccess)
      return hip_fail(e, "neighbors");
  }
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(e, "neighbors");
  return GV_OK;
}

// Inside a capture: the default layout's capture form (reserved buffers, no
// host round trip, no allocation); the result lands in the device vector out
// when the capture is replayed, gv_pagerank_last reports its iterations.
static gv_status pagerank_captured(gv_graph *g, const gv_pagerank_opts *po, float *out, int mem) {
  if (mem != GV_MEM_DEVICE || (po->init && po->init_mem != GV_MEM_DEVICE)) {
    set_error("a captured PageRank takes device vectors only (GV_MEM_DEVICE)");
    return GV_E_INVALID;
  }
  if (g->cfg.pr_layout != 5 || po->max_iter <= 0) {
    set_error("a captured PageRank needs the default layout (pr_layout 5) and max_iter > 0");
    return GV_E_UNSUPPORTED;
  }
  const Store &in = g->has_in ? g->st[1] : g->st[0];
  PrOpts o;
  o.alpha = po->alpha;
  o.eps = po->eps;
  o.max_iter = po->max_iter;
  o.compensated = po->compensated;
  o.init = po->init;
  gv_status rc = pagerank_pb_prepare(in, g->st[0].deg, g->pr, g->stream, true);
  if (rc) return rc;
  return pagerank_pb_iterate(o, g->pr, out, nullptr, nullptr, g->stream, nullptr, true);
}

static gv_status pagerank_common(gv_graph *g, const gv_pagerank_opts *po, float *out, int mem,
                                 int *iters, double *residual, double *iter_ms) {
  gv_status rc = check_graph(g, iters == nullptr && residual == nullptr && iter_ms == nullptr);
  if (rc) return rc;
  if (!out || !po) return GV_E_INVALID;
  if (g->directed && !g->has_in) {
    set_error("PageRank on a directed graph needs maintain_in_edges");
    return GV_E_UNSUPPORTED;
  }
  if (g->capturing) return pagerank_captured(g, po, out, mem);
  const Store &in = g->has_in ? g->st[1] : g->st[0];
  PrBins bins{g->cfg.pr_thread_max, g->cfg.pr_group_max, g->cfg.pr_wave_max};
  const size_t vb = (size_t)(g->V ? g->V : 1) * 4;
  float *dout = out, *dinit = nullptr;
  if (mem != GV_MEM_DEVICE) {
    hipError_t e = hipMalloc(&dout, vb);
    if (e != hipSuccess) return hip_fail(e, "pagerank alloc");
  }
  PrOpts o;
  o.alpha = po->alpha;
  o.eps = po->eps;
  o.max_iter = po->max_iter;
  o.compensated = po->compensated;
  if (po->init) {
    if (po->init_mem == GV_MEM_DEVICE) {
      o.init = po->init;
    } else {
      hipError_t e = hipMalloc(&dinit, vb);
      if (e == hipSuccess)
        e = hipMemcpyAsync(dinit, po->init, (size_t)g->V * 4, hipMemcpyHostToDevice, g->stream);
      if (e != hipSuccess) {
        if (mem != GV_MEM_DEVICE) hipFree(dout);
        hipFree(dinit);
        return hip_fail(e, "pagerank init copy");
      }
      o.init = dinit;
    }
  }
  rc = pagerank_run(in, g->st[0].deg, bins, o, dout, iters, residual, g->pr, g->stream, iter_ms,
                    g->cfg.pr_layout == 4 ? 3 : g->cfg.pr_layout);
  if (!rc && mem != GV_MEM_DEVICE) {
    hipError_t e = hipMemcpyAsync(out, dout, (size_t)g->V * 4, hipMemcpyDeviceToHost, g->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
    if (e != hipSuccess) rc = hip_fail(e, "pagerank copy");
  }
  if (dinit) {
    hipStreamSynchronize(g->stream);
    hipFree(dinit);
  }
  if (mem != GV_MEM_DEVICE) hipFree(dout);
  return rc;
}

gv_status gv_pagerank_opts_default(gv_pagerank_opts *o) {
  if (!o) return GV_E_INVALID;
  std::memset(o, 0, sizeof(*o));
  o->alpha = 0.85f;
  o->eps = 1e-6;
  o->max_iter = 100;
  o->init = nullptr;
  o->init_mem = GV_MEM_HOST;
  o->compensated = 1;
  return GV_OK;
}

gv_status gv_pagerank_run(gv_graph *g, const gv_pagerank_opts *opts, float *out, int mem,
                          gv_pagerank_stats *stats) {
  gv_pagerank_stats st;
  std::memset(&st, 0, sizeof(st));
  if (g && g->capturing) {
    if (stats) {
      set_error("a captured PageRank cannot return stats; pass NULL (gv_pagerank_last)");
      if (!g->cap_err) g->cap_err = GV_E_INVALID;
      return GV_E_INVALID;
    }
    return note_capture(g, pagerank_common(g, opts, out, mem, nullptr, nullptr, nullptr));
  }
  gv_status rc = pagerank_common(g, opts, out, mem, &st.iters, &st.residual,
                                 stats ? &st.kernel_ms : nullptr);
  if (rc) return rc;
  if (stats) {
    rc = read_ctl(g);
    if (rc) return rc;
    st.edges = g->h_ctl[g->has_in ? 1 : 0].live_edges;
    *stats = st;
  }
  return GV_OK;
}

gv_status gv_pagerank(gv_graph *g, float alpha, double eps, int max_iter, float *out, int mem,
                      int *iters, double *residual) {
  gv_pagerank_opts o;
  gv_pagerank_opts_default(&o);
  o.alpha = alpha;
  o.eps = eps;
  o.max_iter = max_iter;
  return pagerank_common(g, &o, out, mem, iters, residual, nullptr);
}

gv_status gv_pagerank_ex(gv_graph *g, float alpha, double eps, int max_iter, float *out, int mem,
                         gv_pagerank_stats *stats) {
  if (!stats) return GV_E_INVALID;
  gv_pagerank_opts o;
  gv_pagerank_opts_default(&o);
  o.alpha = alpha;
  o.eps = eps;
  o.max_iter = max_iter;
  return gv_pagerank_run(g, &o, out, mem, stats);
}

gv_status gv_triangle_count(gv_graph *g, uint64_t *per_vertex, int mem, uint64_t *total) {
  gv_status rc = check_graph(g);
  if (rc) return rc;
  if (!total) return GV_E_INVALID;
  if (g->directed) {
    set_error("triangle counting is defined on undirected graphs");
    return GV_E_UNSUPPORTED;
  }
  uint64_t *dpv = per_vertex;
  if (per_vertex && mem != GV_MEM_DEVICE) {
    hipError_t e = hipMalloc(&dpv, (size_t)(g->V ? g->V : 1) * 8);
    if (e != hipSuccess) return hip_fail(e, "tc alloc");
  }
  rc = triangle_count_run(g->st[0], dpv, total, g->stream);
  if (!rc && per_vertex && mem != GV_MEM_DEVICE) {
    hipError_t e = hipMemcpyAsync(per_vertex, dpv, (size_t)g->V * 8, hipMemcpyDeviceToHost, g->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
    if (e != hipSuccess) rc = hip_fail(e, "tc copy");
  }
  if (per_vertex && mem != GV_MEM_DEVICE) hipFree(dpv);
  return rc;
}

gv_status gv_triangle_count_update(gv_graph *g, int kind, const uint32_t *src, const uint32_t *dst,
                                   uint64_t n, int mem, uint64_t *per_vertex, int pv_mem,
                                   uint64_t *total, gv_batch_stats *stats) {
  gv_status rc = check_graph(g);
  if (rc) return rc;
  if (!total || (n && (!src || !dst))) {
    set_error("null pointer argument");
    return GV_E_INVALID;
  }
  if (kind != GV_OP_INSERT && kind != GV_OP_DELETE) {
    set_error("kind must be GV_OP_INSERT or GV_OP_DELETE");
    return GV_E_INVALID;
  }
  if (per_vertex && pv_mem != GV_MEM_HOST && pv_mem != GV_MEM_DEVICE) {
    set_error("pv_mem must be GV_MEM_HOST or GV_MEM_DEVICE");
    return GV_E_INVALID;
  }
  if (g->directed) {
    set_error("triangle counting is defined on undirected graphs");
    return GV_E_UNSUPPORTED;
  }
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (n == 0) return GV_OK;
  if (n > 0x7FFFFFFFull) {
    set_error("batch larger than 2^31 pairs; split it");
    return GV_E_INVALID;
  }
  const uint32_t *ds = nullptr, *dd = nullptr;
  rc = device_inputs(g, src, dst, n, mem, &ds, &dd);
  if (rc) return rc;
  rc = reserve_keys(g, 0, 2 * n);
  if (rc) return rc;
  hipStream_t s = g->stream;
  uint64_t *dpv = per_vertex;
  if (per_vertex && pv_mem == GV_MEM_HOST) {
    hipError_t e = hipMalloc(&dpv, (size_t)g->V * 8);
    if (e != hipSuccess) return hip_fail(e, "tc update alloc");
    e = hipMemcpyAsync(dpv, per_vertex, (size_t)g->V * 8, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) {
      hipFree(dpv);
      return hip_fail(e, "tc update copy");
    }
  }
  Store &st = g->st[0];
  Scratch &sc = g->sc[0];
  uint64_t m = 0;
  long long delta = 0;
  uint64_t *keys = stage(st, sc, ds, dd, n, 1, 0, m, s);
  segment(st, sc, keys, m, s);
  if (kind == GV_OP_INSERT) {
    // the new edges are the keys the membership probe left flagged
    insert_plan(st, sc, keys, m, s);
    commit_pair(st, sc, nullptr, nullptr, s);
    insert_write(st, sc, keys, m, s);
    rc = tc_delta(st, sc, keys, m, 1, 0, dpv, g->tcd, &delta, s);
  } else {
    // removed triangles are counted before the tombstones go in
    probe_present(st, sc, keys, m, s);
    rc = tc_delta(st, sc, keys, m, 0, 1, dpv, g->tcd, &delta, s);
    if (!rc) delete_apply(st, sc, keys, m, s);
  }
  if (!rc) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = hip_fail(e, "tc update");
  }
  if (!rc) rc = read_ctl(g);
  if (!rc) {
    if (stats) fill_stats(g, n, stats);
    rc = take_abort(g);
  }
  if (!rc) *total += (uint64_t)delta;
  if (per_vertex && pv_mem == GV_MEM_HOST) {
    if (!rc) {
      hipError_t e = hipMemcpyAsync(per_vertex, dpv, (size_t)g->V * 8, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (e != hipSuccess) rc = hip_fail(e, "tc update copy");
    } else {
      hipStreamSynchronize(s);
    }
    hipFree(dpv);
  }
  return rc;
}

gv_status gv_bc(gv_graph *g, const uint32_t *sources, uint32_t n_sources, float *out, int mem) {
  gv_status rc = check_graph(g);
  if (rc) return rc;
  if (!out || (n_sources && !sources)) return GV_E_INVALID;
  std::vector<uint32_t> hs(n_sources);
  if (n_sources) {
    if (mem == GV_MEM_DEVICE) {
      GV_HIP(hipMemcpy(hs.data(), sources, (size_t)n_sources * 4, hipMemcpyDeviceToHost));
    } else {
      std::memcpy(hs.data(), sources, (size_t)n_sources * 4);
    }
  }
  for (uint32_t x : hs)
    if (x >= g->V) {
      set_error("BC source out of range");
      return GV_E_RANGE;
    }
  float *dout = out;
  if (mem != GV_MEM_DEVICE) {
    hipError_t e = hipMalloc(&dout, (size_t)(g->V ? g->V : 1) * 4);
    if (e != hipSuccess) return hip_fail(e, "bc alloc");
  }
  // bottom-up levels pull over the in-adjacency: the store itself when
  // undirected, its mirror when directed (none without maintain_in_edges)
  const Store *in = !g->directed ? &g->st[0] : (g->has_in ? &g->st[1] : nullptr);
  rc = n_sources > 1 ? bc_multi_run(g->st[0], in, hs.data(), n_sources, dout, g->bcm,
                                    g->cfg.bc_bu_ratio, g->stream)
                     : bc_run(g->st[0], hs.data(), n_sources, dout, g->stream);
  if (!rc && mem != GV_MEM_DEVICE) {
    hipError_t e = hipMemcpyAsync(out, dout, (size_t)g->V * 4, hipMemcpyDeviceToHost, g->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
    if (e != hipSuccess) rc = hip_fail(e, "bc copy");
  }
  if (mem != GV_MEM_DEVICE) hipFree(dout);
  return rc;
}

gv_status gv_pagerank_personalized(gv_graph *g, float alpha, double eps, int max_iter,
                                   const uint32_t *seeds, uint32_t n_cols, float *out, int mem,
                                   int *iters, double *max_residual) {
  gv_status rc = check_graph(g);
  if (rc) return rc;
  if (!out || (n_cols && !seeds)) return GV_E_INVALID;
  if (n_cols > 64) {
    set_error("at most 64 personalized columns per call");
    return GV_E_INVALID;
  }
  if (g->directed && !g->has_in) {
    set_error("PageRank on a directed graph needs maintain_in_edges");
    return GV_E_UNSUPPORTED;
  }
  std::vector<uint32_t> hs(n_cols);
  if (n_cols) {
    if (mem == GV_MEM_DEVICE) {
      GV_HIP(hipMemcpy(hs.data(), seeds, (size_t)n_cols * 4, hipMemcpyDeviceToHost));
    } else {
      std::memcpy(hs.data(), seeds, (size_t)n_cols * 4);
    }
  }
  for (uint32_t x : hs)
    if (x >= g->V) {
      set_error("personalized PageRank seed out of range");
      return GV_E_RANGE;
    }
  const Store &in = g->has_in ? g->st[1] : g->st[0];
  uint32_t *dseeds = nullptr;
  float *dout = out;
  const size_t ob = (size_t)(g->V ? g->V : 1) * (n_cols ? n_cols : 1) * 4;
  do {
    hipError_t e;
    if ((e = hipMalloc(&dseeds, 64 * sizeof(uint32_t))) != hipSuccess ||
        (n_cols && (e = hipMemcpyAsync(dseeds, hs.data(), (size_t)n_cols * 4,
                                       hipMemcpyHostToDevice, g->stream)) != hipSuccess)) {
      rc = hip_fail(e, "personalized PageRank");
      break;
    }
    if (mem != GV_MEM_DEVICE && (e = hipMalloc(&dout, ob)) != hipSuccess) {
      dout = nullptr;
      rc = hip_fail(e, "personalized PageRank alloc");
      break;
    }
    rc = ppr_run(in, g->st[0].deg, alpha, eps, max_iter, dseeds, n_cols, dout, iters,
                 max_residual, nullptr, g->stream);
    if (!rc && mem != GV_MEM_DEVICE) {
      e = hipMemcpyAsync(out, dout, (size_t)g->V * n_cols * 4, hipMemcpyDeviceToHost, g->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
      if (e != hipSuccess) rc = hip_fail(e, "personalized PageRank copy");
    }
  } while (0);
  hipFree(dseeds);
  if (mem != GV_MEM_DEVICE) hipFree(dout);
  return rc;
}

gv_status gv_gen_rmat(gv_graph *g, uint64_t key, int scale, uint64_t n, uint32_t ta, uint32_t tab,
                      uint32_t tabc, uint32_t *src_dev, uint32_t *dst_dev) {
  gv_status rc = check_graph(g);
  if (rc) return rc;
  if (scale < 1 || scale > 31 || (n && (!src_dev || !dst_dev))) return GV_E_INVALID;
  gen_rmat(key, scale, n, ta, tab, tabc, src_dev, dst_dev, g->stream);
  GV_HIP(hipGetLastError());
  return GV_OK;
}

gv_status gv_gen_uniform(gv_graph *g, uint64_t key, uint32_t V, uint64_t n, uint32_t *src_dev,
                         uint32_t *dst_dev) {
  gv_status rc = check_graph(g);
  if (rc) return rc;
  if (n && (!src_dev || !dst_dev)) return GV_E_INVALID;
  gen_uniform(key, V, n, src_dev, dst_dev, g->stream);
  GV_HIP(hipGetLastError());
  return GV_OK;
}

gv_status gv_gen_sample(gv_graph *g, uint64_t key, uint32_t V, const uint64_t *offsets_dev,
                        const uint32_t *cols_dev, uint64_t n, uint32_t *src_dev,
                        uint32_t *dst_dev) {
  gv_status rc = check_graph(g);
  if (rc) return rc;
  if (n && (!src_dev || !dst_dev || !offsets_dev)) return GV_E_INVALID;
  gen_sample(key, V, offsets_dev, cols_dev, n, src_dev, dst_dev, g->stream);
  GV_HIP(hipGetLastError());
  return GV_OK;
}

// ------------------------------------------------------------ capture --
gv_status gv_reserve(gv_graph *g, uint64_t max_pairs, uint32_t max_batches) {
  gv_status rc = check_graph(g);
  if (rc) return rc;
  if (max_pairs > 0x7FFFFFFFull) {
    set_error("max_pairs above 2^31");
    return GV_E_INVALID;
  }
  for (int k = 0; k < (g->has_in ? 2 : 1); ++k) {
    rc = reserve_keys(g, k, g->directed ? max_pairs : 2 * max_pairs);
    if (rc) return rc;
  }
  return reserve_groups(g, max_batches);
}

gv_status gv_pagerank_reserve(gv_graph *g, uint64_t max_edges) {
  gv_status rc = check_graph(g);
  if (rc) return rc;
  if (g->directed && !g->has_in) {
    set_error("PageRank on a directed graph needs maintain_in_edges");
    return GV_E_UNSUPPORTED;
  }
  if (g->cfg.pr_layout != 5) {
    set_error("gv_pagerank_reserve sizes the default layout (pr_layout 5)");
    return GV_E_UNSUPPORTED;
  }
  const Store &in = g->has_in ? g->st[1] : g->st[0];
  return pagerank_pb_reserve(in, g->st[0].deg, g->pr, max_edges, g->stream);
}

gv_status gv_pagerank_last(gv_graph *g, gv_pagerank_stats *stats) {
  gv_status rc = check_graph(g);
  if (rc) return rc;
  if (!stats) return GV_E_INVALID;
  std::memset(stats, 0, sizeof(*stats));
  rc = pagerank_pb_last(g->pr, &stats->iters, &stats->residual, g->stream);
  if (rc) return rc;
  rc = read_ctl(g);
  if (rc) return rc;
  stats->edges = g->h_ctl[g->has_in ? 1 : 0].live_edges;
  return GV_OK;
}

gv_status gv_capture_begin(gv_graph *g) {
  gv_status rc = check_graph(g);
  if (rc) return rc;
  GV_HIP(hipStreamSynchronize(g->stream));
  g->cap_err = GV_OK;
  g->cap_pairs = 0;
  GV_HIP(hipStreamBeginCapture(g->stream, hipStreamCaptureModeRelaxed));
  g->capturing = 1;
  return GV_OK;
}

gv_status gv_capture_end(gv_graph *g, gv_capture **out) {
  if (!g || !out) {
    set_error("null handle");
    return GV_E_INVALID;
  }
  *out = nullptr;
  if (!g->capturing) {
    set_error("no capture is open on this handle");
    return GV_E_INVALID;
  }
  hipSetDevice(g->cfg.device);
  g->capturing = 0;
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(g->stream, &graph);
  std::vector<void *> pinned;
  pinned.swap(g->cap_pinned);
  gv_status rc = g->cap_err;
  if (!rc && e != hipSuccess) rc = hip_fail(e, "hipStreamEndCapture");
  hipGraphExec_t exec = nullptr;
  if (!rc && (e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0)) != hipSuccess)
    rc = hip_fail(e, "hipGraphInstantiate");
  if (rc) {
    if (graph) hipGraphDestroy(graph);
    for (void *p : pinned) hipHostFree(p);
    if (g->cap_err) set_error("a call inside the capture failed; nothing was captured");
    g->cap_err = GV_OK;
    return rc;
  }
  gv_capture *c = new gv_capture();
  c->g = g;
  c->graph = graph;
  c->exec = exec;
  c->gen = g->buf_gen;
  c->pr_gen = g->pr.gen;
  c->pairs = g->cap_pairs;
  c->pinned.swap(pinned);
  g->captures.push_back(c);
  *out = c;
  return GV_OK;
}

gv_status gv_capture_launch(gv_capture *c, gv_batch_stats *stats) {
  if (!c) {
    set_error("null capture");
    return GV_E_INVALID;
  }
  gv_graph *g = c->g;
  if (!g) {
    set_error("the graph of this capture was destroyed");
    return GV_E_INVALID;
  }
  gv_status rc = check_graph(g);
  if (rc) return rc;
  if (c->gen != g->buf_gen || c->pr_gen != g->pr.gen) {
    set_error("buffers used by this capture were reallocated after it was made; capture again");
    return GV_E_INVALID;
  }
  GV_HIP(hipGraphLaunch(c->exec, g->stream));
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (!stats) return GV_OK;
  rc = read_ctl(g);
  if (rc) return rc;
  fill_stats(g, c->pairs, stats);
  return take_abort(g);
}

gv_status gv_capture_destroy(gv_capture *c) {
  if (!c) return GV_OK;
  if (gv_graph *g = c->g) {
    hipSetDevice(g->cfg.device);
    hipStreamSynchronize(g->stream);
    g->captures.erase(std::remove(g->captures.begin(), g->captures.end(), c), g->captures.end());
  }
  if (c->exec) hipGraphExecDestroy(c->exec);
  if (c->graph) hipGraphDestroy(c->graph);
  for (void *p : c->pinned) hipHostFree(p);
  delete c;
  return GV_OK;
}

// ------------------------------------------------------------- tuning --
gv_status gv_get_config(gv_graph *g, gv_config *cfg) {
  if (!g || !cfg) return GV_E_INVALID;
  *cfg = g->cfg;
  return GV_OK;
}

gv_status gv_set_tuning(gv_graph *g, const gv_config *cfg) {
  gv_status rc = check_graph(g);
  if (rc) return rc;
  if (!cfg) return GV_E_INVALID;
  if (cfg->pr_thread_max > cfg->pr_group_max || cfg->pr_group_max > cfg->pr_wave_max ||
      cfg->pr_wave_max == 0) {
    set_error("PageRank bins must satisfy thread_max <= group_max <= wave_max, wave_max > 0");
    return GV_E_INVALID;
  }
  if (cfg->pr_layout != 1 && cfg->pr_layout != 4 && cfg->pr_layout != 5) {
    set_error("pr_layout must be 1, 4 or 5");
    return GV_E_INVALID;
  }
  if (!(cfg->bc_bu_ratio >= 0.f) || !std::isfinite(cfg->bc_bu_ratio)) {
    set_error("bc_bu_ratio must be finite and >= 0");
    return GV_E_INVALID;
  }
  g->cfg.compact_threshold = cfg->compact_threshold;
  g->cfg.pr_thread_max = cfg->pr_thread_max;
  g->cfg.pr_group_max = cfg->pr_group_max;
  g->cfg.pr_wave_max = cfg->pr_wave_max;
  g->cfg.pr_layout = cfg->pr_layout;
  g->cfg.bc_bu_ratio = cfg->bc_bu_ratio;
  return GV_OK;
}

}  // extern "C"

// Chung-Lu cumulative weight table: w_i = (i+1)^(-1/(gamma-1)) scaled to
// average degree avg_deg, capped at max_deg, accumulated in vertex order in
// FP64 and scaled to 2^62.  The operations are kept in this exact order
// (no contraction) so every build of the table gives the same integers.
static void chung_lu_table(uint32_t V, double gamma, double avg_deg, double max_deg,
                           std::vector<uint64_t> &cum) {
#pragma clang fp contract(off)
  std::vector<double> w(V);
  const double ex = -1.0 / (gamma - 1.0);
  double s = 0.0;
  for (uint32_t i = 0; i < V; ++i) {
    w[i] = std::pow((double)i + 1.0, ex);
    s += w[i];
  }
  const double scale = avg_deg * (double)V / s;
  double tot = 0.0;
  for (uint32_t i = 0; i < V; ++i) {
    w[i] *= scale;
    if (w[i] > max_deg) w[i] = max_deg;
    tot += w[i];
  }
  cum.assign((size_t)V + 1, 0);
  double acc = 0.0;
  for (uint32_t i = 0; i < V; ++i) {
    acc += w[i];
    cum[(size_t)i + 1] = (uint64_t)(acc / tot * 4611686018427387904.0);
  }
  cum[V] = 4611686018427387904ull;
}

extern "C" {

gv_status gv_gen_chung_lu(gv_graph *g, uint64_t key, uint32_t V, double gamma, double avg_deg,
                          double max_deg, uint64_t n, uint32_t *src_dev, uint32_t *dst_dev) {
  gv_status rc = check_graph(g);
  if (rc) return rc;
  if (V == 0 || !(gamma > 1.0) || !(avg_deg > 0.0) || !(max_deg > 0.0) ||
      (n && (!src_dev || !dst_dev))) {
    set_error("Chung-Lu needs V > 0, gamma > 1, avg_deg > 0, max_deg > 0");
    return GV_E_INVALID;
  }
  if (!g->cl_cum || g->cl_V != V || g->cl_gamma != gamma || g->cl_avg != avg_deg ||
      g->cl_max != max_deg) {
    std::vector<uint64_t> cum;
    chung_lu_table(V, gamma, avg_deg, max_deg, cum);
    GV_HIP(hipStreamSynchronize(g->stream));
    hipFree(g->cl_cum);
    g->cl_cum = nullptr;
    g->cl_V = 0;
    GV_HIP(hipMalloc(&g->cl_cum, cum.size() * sizeof(uint64_t)));
    GV_HIP(hipMemcpy(g->cl_cum, cum.data(), cum.size() * sizeof(uint64_t),
                     hipMemcpyHostToDevice));
    g->cl_V = V;
    g->cl_gamma = gamma;
    g->cl_avg = avg_deg;
    g->cl_max = max_deg;
  }
  gen_table(key, g->cl_cum, V, n, src_dev, dst_dev, g->stream);
  GV_HIP(hipGetLastError());
  return GV_OK;
}

}  // extern "C"
