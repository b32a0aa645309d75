// gv_serve.hip -- request loop over one graph handle (include/gv_serve.h).
//
// Host orchestration only.  The update worker drains an arrival-ordered queue
// in steps (a run of update requests, a run of neighbour lookups, a run of
// PageRank refreshes, or one BC request), issuing the store's own entry
// points on the graph's stream and stamping completion after the device work
// of the step has finished.  A PageRank or BC step only takes its snapshot
// there; the analytics worker runs it on a second stream, so it overlaps the
// updates and lookups that arrive after it.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "gv_analytics.h"
#include "gv_internal.h"
#include "gv_serve.h"
#include "gv_store.h"

using namespace gv;

namespace {

using Clock = std::chrono::steady_clock;

struct Item {
  gv_request req;
  uint64_t ticket;
  double arrival_us;
};

enum class Cls { Update, Neighbors, Pagerank, Bc };

// Analytics snapshot slots: a PageRank or BC request holds one from the
// moment its snapshot is taken on the graph's stream until its result is out.
constexpr int AN_SLOTS = 2;

struct AnSlot {
  PrWork w;
  hipEvent_t ready = nullptr;
  bool busy = false;
};

// A PageRank run (consecutive refreshes share it) or one BC, handed from the
// update worker to the analytics worker with its snapshot.
struct AnJob {
  Cls cls;
  std::vector<Item> items;
  int slot;
  double t_start;
  uint64_t step;
};

Cls cls_of(int kind) {
  switch (kind) {
    case GV_REQ_INSERT:
    case GV_REQ_DELETE: return Cls::Update;
    case GV_REQ_NEIGHBORS: return Cls::Neighbors;
    case GV_REQ_PAGERANK: return Cls::Pagerank;
    default: return Cls::Bc;
  }
}

}  // namespace

struct gv_server {
  gv_graph *g = nullptr;
  gv_server_opts o;
  uint32_t V = 0;
  hipStream_t stream = nullptr;
  Clock::time_point t0;

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Item> queue;
  std::vector<gv_response> resp;  // by ticket
  std::vector<char> finished;     // by ticket
  uint64_t next_ticket = 0, n_done = 0;
  bool stop = false;
  std::thread worker;

  // analytics worker: own stream and thread, jobs in capture order
  hipStream_t astream = nullptr;
  uint32_t an_cus = 0;      // compute units the analytics kernels may occupy
  std::thread aworker;
  std::deque<AnJob> ajobs;  // guarded by mu
  std::condition_variable cv_ajob, cv_slot, cv_idle;
  bool u_busy = false;      // the update worker is serving a step (guarded by mu)
  bool astop = false;       // guarded by mu
  AnSlot slots[AN_SLOTS];   // busy flags guarded by mu
  BcWork bcw;               // analytics worker only

  // update-worker state
  uint32_t *stage_src = nullptr, *stage_dst = nullptr;
  // PageRank vectors (analytics worker; the update worker only in the
  // fallback path, when no analytics job is outstanding)
  float *pr[2] = {nullptr, nullptr};
  int pr_cur = -1;  // index of the last PageRank vector, -1 = none yet
  float *bc_out = nullptr;
  uint32_t *bc_src = nullptr;
  std::vector<uint64_t> nb_offs;
  std::vector<uint32_t> nb_vs, nb_out;
  uint64_t w_free = 0, w_total = 0;  // pool state seen by the worker
  gv_server_stats stats{};           // guarded by mu

  double now_us() const {
    return std::chrono::duration<double, std::micro>(Clock::now() - t0).count();
  }
};

static gv_status stage_copy(gv_server *s, uint64_t at, const gv_request &r) {
  const hipMemcpyKind k = r.mem == GV_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  if (!r.n) return GV_OK;
  GV_HIP(hipMemcpyAsync(s->stage_src + at, r.src, r.n * 4, k, s->stream));
  GV_HIP(hipMemcpyAsync(s->stage_dst + at, r.dst, r.n * 4, k, s->stream));
  return GV_OK;
}

static int op_of(const gv_request &r) {
  return r.kind == GV_REQ_DELETE ? GV_OP_DELETE : GV_OP_INSERT;
}

static gv_status apply_one(gv_server *s, const gv_request &r, const uint32_t *src,
                           const uint32_t *dst, int mem) {
  const uint64_t off[2] = {0, r.n};
  const int kind = op_of(r);
  return gv_batch_apply(s->g, src, dst, off, &kind, 1, mem, s->o.compact_fraction, nullptr);
}

// Run of update requests: one group if admission allows, else one at a time.
// Returns the pairs of the requests that succeeded.
static uint64_t serve_updates(gv_server *s, std::vector<Item> &step,
                              std::vector<gv_response> &out) {
  const size_t nb = step.size();
  std::vector<uint64_t> offs(nb + 1, 0);
  std::vector<int> kinds(nb);
  for (size_t i = 0; i < nb; ++i) {
    offs[i + 1] = offs[i] + step[i].req.n;
    kinds[i] = op_of(step[i].req);
  }
  gv_status rc = GV_OK;
  bool staged = offs[nb] <= s->o.max_step_pairs;
  if (staged) {
    for (size_t i = 0; i < nb && !rc; ++i) rc = stage_copy(s, offs[i], step[i].req);
    if (!rc)
      rc = gv_batch_apply(s->g, s->stage_src, s->stage_dst, offs.data(), kinds.data(),
                          (uint32_t)nb, GV_MEM_DEVICE, s->o.compact_fraction, nullptr);
    if (!rc) rc = gv_sync(s->g);  // surfaces a device abort (nothing applied)
  }
  if (staged && rc == GV_OK) {
    for (size_t i = 0; i < nb; ++i) out[i].status = GV_OK;
  } else if (nb > 1 || !staged || rc == GV_E_POOL || rc == GV_E_RANGE) {
    // one at a time: only the offending requests fail
    for (size_t i = 0; i < nb; ++i) {
      const gv_request &r = step[i].req;
      gv_status ri;
      if (staged) {
        ri = apply_one(s, r, s->stage_src + offs[i], s->stage_dst + offs[i], GV_MEM_DEVICE);
      } else {
        ri = apply_one(s, r, r.src, r.dst, r.mem);
      }
      if (!ri) ri = gv_sync(s->g);
      out[i].status = ri;
    }
  } else {
    for (size_t i = 0; i < nb; ++i) out[i].status = rc;
  }
  uint64_t applied = 0;
  for (size_t i = 0; i < nb; ++i)
    if (out[i].status == GV_OK) applied += step[i].req.n;
  uint64_t fb = 0, tb = 0;
  if (gv_pool_stats(s->g, &fb, &tb) == GV_OK) {
    s->w_free = fb;
    s->w_total = tb;
  }
  return applied;
}

static void serve_neighbors(gv_server *s, std::vector<Item> &step,
                            std::vector<gv_response> &out) {
  s->nb_vs.clear();
  for (auto &it : step) {
    bool ok = it.req.n == 0 || it.req.vs != nullptr;
    for (uint64_t i = 0; ok && i < it.req.n; ++i) ok = it.req.vs[i] < s->V;
    if (ok) s->nb_vs.insert(s->nb_vs.end(), it.req.vs, it.req.vs + it.req.n);
  }
  const uint64_t nq = s->nb_vs.size();
  s->nb_offs.assign(nq + 1, 0);
  gv_status rc = gv_neighbors(s->g, s->nb_vs.data(), (uint32_t)nq, s->nb_offs.data(), nullptr,
                              GV_MEM_HOST);
  if (!rc) {
    s->nb_out.resize(std::max<uint64_t>(s->nb_offs[nq], 1));
    rc = gv_neighbors(s->g, s->nb_vs.data(), (uint32_t)nq, s->nb_offs.data(), s->nb_out.data(),
                      GV_MEM_HOST);
  }
  uint64_t q = 0;
  for (size_t i = 0; i < step.size(); ++i) {
    const gv_request &r = step[i].req;
    bool ok = r.n == 0 || r.vs != nullptr;
    for (uint64_t k = 0; ok && k < r.n; ++k) ok = r.vs[k] < s->V;
    if (!ok) {
      out[i].status = GV_E_RANGE;
      continue;
    }
    if (rc) {
      out[i].status = rc;
      q += r.n;
      continue;
    }
    const uint64_t b = s->nb_offs[q], e = s->nb_offs[q + r.n];
    out[i].out_len = e - b;
    if (r.offs)
      for (uint64_t k = 0; k <= r.n; ++k) r.offs[k] = s->nb_offs[q + k] - b;
    if (r.out && e - b > r.out_cap) {
      out[i].status = GV_E_INVALID;
    } else {
      if (r.out && e > b) std::memcpy(r.out, s->nb_out.data() + b, (e - b) * 4);
      out[i].status = GV_OK;
    }
    q += r.n;
  }
}

static void serve_pagerank(gv_server *s, std::vector<Item> &step, std::vector<gv_response> &out) {
  gv_pagerank_opts o;
  gv_pagerank_opts_default(&o);
  o.alpha = s->o.pr_alpha;
  o.eps = s->o.pr_eps;
  o.max_iter = s->o.pr_max_iter;
  const int nxt = s->pr_cur == 0 ? 1 : 0;
  if (s->pr_cur >= 0) {
    o.init = s->pr[s->pr_cur];
    o.init_mem = GV_MEM_DEVICE;
  }
  gv_pagerank_stats st;
  gv_status rc = gv_pagerank_run(s->g, &o, s->pr[nxt], GV_MEM_DEVICE, &st);
  if (!rc) s->pr_cur = nxt;
  for (size_t i = 0; i < step.size(); ++i) {
    out[i].status = rc;
    if (rc) continue;
    out[i].iters = st.iters;
    out[i].residual = st.residual;
    if (step[i].req.x) {
      hipError_t e = hipMemcpy(step[i].req.x, s->pr[nxt], (size_t)s->V * 4, hipMemcpyDeviceToHost);
      if (e != hipSuccess) out[i].status = hip_fail(e, "pagerank copy");
    }
  }
}

// Record the responses of one served step (worker threads).
static void finish(gv_server *s, const std::vector<Item> &items, std::vector<gv_response> &out,
                   double t_start, double t_done, uint64_t sid, uint64_t applied) {
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stats.update_pairs_applied += applied;
    s->stats.free_blocks = s->w_free;
    s->stats.total_blocks = s->w_total;
    for (size_t i = 0; i < items.size(); ++i) {
      gv_response &r = out[i];
      r.kind = items[i].req.kind;
      r.arrival_us = items[i].arrival_us;
      r.start_us = t_start;
      r.done_us = t_done;
      r.step = sid;
      r.step_size = (uint32_t)items.size();
      s->resp[items[i].ticket] = r;
      s->finished[items[i].ticket] = 1;
      ++s->n_done;
      ++s->stats.requests_done;
      if (r.status) ++s->stats.refused;
    }
  }
  s->cv_done.notify_all();
}

static void release_slot(gv_server *s, int k) {
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->slots[k].busy = false;
  }
  s->cv_slot.notify_all();
}

// Update worker, PageRank or BC step: take a snapshot slot (waiting while
// both are held: admission by snapshot memory), build the snapshot on the
// graph's stream -- after every update that arrived before the step -- and
// hand the step to the analytics worker.  Returns false when the step was
// answered here instead (a refused request, or a graph the flat snapshot
// cannot hold, served synchronously).
static bool capture_analytics(gv_server *s, Cls c, std::vector<Item> &step,
                              std::vector<gv_response> &out, double t_start, uint64_t sid) {
  if (c == Cls::Bc && step[0].req.source >= s->V) {
    out[0].status = GV_E_RANGE;
    return false;
  }
  int k = -1;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_slot.wait(lk, [&] {
      for (int i = 0; i < AN_SLOTS; ++i)
        if (!s->slots[i].busy) return true;
      return false;
    });
    for (int i = 0; i < AN_SLOTS; ++i)
      if (!s->slots[i].busy) {
        k = i;
        break;
      }
    s->slots[k].busy = true;
  }
  AnSlot &sl = s->slots[k];
  gv_status rc = c == Cls::Pagerank ? graph_pr_prepare(s->g, sl.w) : graph_bc_prepare(s->g, sl.w);
  if (rc == GV_OK && hipEventRecord(sl.ready, s->stream) != hipSuccess)
    rc = hip_fail(hipErrorUnknown, "serve event");
  if (rc) {
    release_slot(s, k);
    if (rc == GV_E_UNSUPPORTED && c == Cls::Pagerank) {
      // wait for the analytics worker to finish its PageRank jobs (the
      // vectors are shared), then run this one here
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_slot.wait(lk, [&] {
        for (int i = 0; i < AN_SLOTS; ++i)
          if (s->slots[i].busy) return false;
        return true;
      });
      lk.unlock();
      serve_pagerank(s, step, out);
    } else {
      for (auto &o : out) o.status = rc;
    }
    return false;
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->ajobs.push_back(AnJob{c, step, k, t_start, sid});
  }
  s->cv_ajob.notify_one();
  return true;
}

// Before each PageRank iteration or BC level the analytics worker steps
// aside while update-side requests wait or run (for at most 5 ms a time):
// updates and lookups then wait for one analytics kernel, not for a whole
// job, and the job takes the device time they leave.
static void analytics_yield(void *arg) {
  gv_server *s = static_cast<gv_server *>(arg);
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_idle.wait_for(lk, std::chrono::milliseconds(5),
                      [&] { return s->stop || (!s->u_busy && s->queue.empty()); });
}

// Analytics worker: each job waits for its snapshot event on the analytics
// stream and runs there, concurrently with later updates on the graph's
// stream.
static void analytics_loop(gv_server *s) {
  {
    gv_config cfg;
    if (gv_get_config(s->g, &cfg) == GV_OK) hipSetDevice(cfg.device);
  }
  std::vector<gv_response> out;
  for (;;) {
    AnJob job;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_ajob.wait(lk, [&] { return s->astop || !s->ajobs.empty(); });
      if (s->ajobs.empty()) break;
      job = std::move(s->ajobs.front());
      s->ajobs.pop_front();
    }
    AnSlot &sl = s->slots[job.slot];
    out.assign(job.items.size(), gv_response{});
    gv_status rc = GV_OK;
    if (hipStreamWaitEvent(s->astream, sl.ready, 0) != hipSuccess)
      rc = hip_fail(hipErrorUnknown, "serve wait");
    if (job.cls == Cls::Pagerank) {
      PrOpts o;
      o.alpha = s->o.pr_alpha;
      o.eps = s->o.pr_eps;
      o.max_iter = s->o.pr_max_iter;
      o.compensated = 1;
      o.yield = analytics_yield;
      o.yield_arg = s;
      const int nxt = s->pr_cur == 0 ? 1 : 0;
      if (s->pr_cur >= 0) o.init = s->pr[s->pr_cur];
      int iters = 0;
      double res = 0.0;
      if (!rc) {
        if (o.max_iter <= 0) {
          // zero iterations: the starting vector itself
          const float x0 = 1.0f / (float)s->V;
          if (o.init) {
            if (hipMemcpyAsync(s->pr[nxt], o.init, (size_t)s->V * 4, hipMemcpyDeviceToDevice,
                               s->astream) != hipSuccess)
              rc = hip_fail(hipErrorUnknown, "pagerank copy");
          } else {
            fill_f32(s->pr[nxt], x0, s->V, s->astream);
          }
        } else {
          rc = pagerank_snap_iterate(o, sl.w, s->pr[nxt], &iters, &res, s->astream, nullptr);
        }
      }
      if (!rc) s->pr_cur = nxt;
      for (size_t i = 0; i < job.items.size(); ++i) {
        out[i].status = rc;
        if (rc) continue;
        out[i].iters = iters;
        out[i].residual = res;
        if (job.items[i].req.x) {
          hipError_t e = hipMemcpyAsync(job.items[i].req.x, s->pr[nxt], (size_t)s->V * 4,
                                        hipMemcpyDeviceToHost, s->astream);
          if (e == hipSuccess) e = hipStreamSynchronize(s->astream);
          if (e != hipSuccess) out[i].status = hip_fail(e, "pagerank copy");
        }
      }
    } else {
      const gv_request &r = job.items[0].req;
      if (!rc)
        rc = bc_run_csr(sl.w.snap_off, sl.w.snap_cols, s->V, &r.source, 1, s->bc_out, s->astream,
                        s->bcw);
      if (!rc && r.x) {
        hipError_t e = hipMemcpyAsync(r.x, s->bc_out, (size_t)s->V * 4, hipMemcpyDeviceToHost,
                                      s->astream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->astream);
        if (e != hipSuccess) rc = hip_fail(e, "bc copy");
      }
      out[0].status = rc;
    }
    hipStreamSynchronize(s->astream);
    const double t_done = s->now_us();
    release_slot(s, job.slot);
    finish(s, job.items, out, job.t_start, t_done, job.step, 0);
  }
}

static void set_idle(gv_server *s) {
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->u_busy = false;
  }
  s->cv_idle.notify_all();
}

static void worker_loop(gv_server *s) {
  hipSetDevice(0);
  {
    gv_config cfg;
    if (gv_get_config(s->g, &cfg) == GV_OK) hipSetDevice(cfg.device);
  }
  std::vector<Item> step;
  std::vector<gv_response> out;
  for (;;) {
    step.clear();
    uint64_t sid = 0;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty(); });
      if (s->queue.empty()) break;  // stop requested and nothing left
      const Cls c = cls_of(s->queue.front().req.kind);
      if (c == Cls::Update) {
        // admission: worst case one new block per directed entry (2 per pair)
        uint64_t pairs = 0, need = 0;
        const uint64_t fb = s->w_free;
        const uint64_t room = fb > s->o.headroom_blocks ? fb - s->o.headroom_blocks : 0;
        while (!s->queue.empty() && cls_of(s->queue.front().req.kind) == Cls::Update &&
               step.size() < s->o.max_step_requests) {
          const gv_request &r = s->queue.front().req;
          const uint64_t w = 2 * r.n;
          if (!step.empty() && (pairs + r.n > s->o.max_step_pairs || need + w > room)) break;
          pairs += r.n;
          need += w;
          step.push_back(s->queue.front());
          s->queue.pop_front();
          if (need > room) break;  // a first request over the bound runs alone
        }
      } else if (c == Cls::Bc) {
        step.push_back(s->queue.front());
        s->queue.pop_front();
      } else {
        uint64_t q = 0;
        while (!s->queue.empty() && cls_of(s->queue.front().req.kind) == c &&
               step.size() < s->o.max_step_requests) {
          const uint64_t n = s->queue.front().req.n;
          if (c == Cls::Neighbors && !step.empty() && q + n > 0x7FFFFFFFull) break;
          q += n;
          step.push_back(s->queue.front());
          s->queue.pop_front();
        }
      }
      sid = s->stats.steps++;
      s->u_busy = true;
    }
    const double t_start = s->now_us();
    out.assign(step.size(), gv_response{});
    uint64_t applied = 0;
    const Cls c = cls_of(step[0].req.kind);
    switch (c) {
      case Cls::Update: applied = serve_updates(s, step, out); break;
      case Cls::Neighbors: serve_neighbors(s, step, out); break;
      case Cls::Pagerank:
      case Cls::Bc:
        if (capture_analytics(s, c, step, out, t_start, sid)) {  // answered later
          set_idle(s);
          continue;
        }
        break;
    }
    hipStreamSynchronize(s->stream);
    finish(s, step, out, t_start, s->now_us(), sid, applied);
    set_idle(s);
  }
  set_idle(s);
}

extern "C" {

gv_status gv_server_opts_default(gv_server_opts *o) {
  if (!o) return GV_E_INVALID;
  std::memset(o, 0, sizeof(*o));
  o->max_step_pairs = 1ull << 24;
  o->max_step_requests = 4096;
  o->headroom_blocks = 0;
  o->compact_fraction = 0.25f;
  o->pr_alpha = 0.85f;
  o->pr_eps = 1e-6;
  o->pr_max_iter = 100;
  o->analytics_cu_share = 0.75f;
  return GV_OK;
}

// The analytics stream, and the share of the compute units its kernels may
// occupy: the persistent PageRank grid gets analytics_cu_share of the CUs
// (one workgroup each, all of its LDS), the BC level kernels a grid of 8
// workgroups per such CU, so the rest of the device stays free for the
// short kernels of update steps.
static hipError_t create_analytics_stream(gv_server *s) {
  gv_config cfg;
  int dev = 0, ncu = 0;
  if (gv_get_config(s->g, &cfg) == GV_OK) dev = cfg.device;
  hipError_t e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  uint32_t n = (uint32_t)((float)ncu * s->o.analytics_cu_share + 0.5f);
  s->an_cus = n ? n : 1u;
  return hipStreamCreateWithFlags(&s->astream, hipStreamNonBlocking);
}

gv_status gv_server_create(gv_graph *g, const gv_server_opts *opts, gv_server **out) {
  if (!g || !out) {
    set_error("null graph or output");
    return GV_E_INVALID;
  }
  *out = nullptr;
  gv_server_opts o;
  if (opts) o = *opts; else gv_server_opts_default(&o);
  if (o.max_step_pairs == 0 || o.max_step_requests == 0 || o.pr_max_iter < 0) {
    set_error("max_step_pairs and max_step_requests must be positive");
    return GV_E_INVALID;
  }
  if (!(o.analytics_cu_share > 0.f && o.analytics_cu_share <= 1.f)) {
    set_error("analytics_cu_share must be in (0, 1]");
    return GV_E_INVALID;
  }
  gv_server *s = new gv_server();
  s->g = g;
  s->o = o;
  gv_status rc = gv_num_vertices(g, &s->V);
  void *st = nullptr;
  if (!rc) rc = gv_get_stream(g, &st);
  s->stream = (hipStream_t)st;
  uint64_t fb = 0, tb = 0;
  if (!rc) rc = gv_pool_stats(g, &fb, &tb);
  s->stats.free_blocks = s->w_free = fb;
  s->stats.total_blocks = s->w_total = tb;
  if (!rc) {
    const size_t vb = (size_t)(s->V ? s->V : 1) * 4;
    hipError_t e;
    if ((e = hipMalloc(&s->stage_src, o.max_step_pairs * 4)) != hipSuccess ||
        (e = hipMalloc(&s->stage_dst, o.max_step_pairs * 4)) != hipSuccess ||
        (e = hipMalloc(&s->pr[0], vb)) != hipSuccess ||
        (e = hipMalloc(&s->pr[1], vb)) != hipSuccess ||
        (e = hipMalloc(&s->bc_out, vb)) != hipSuccess ||
        (e = hipMalloc(&s->bc_src, 4)) != hipSuccess ||
        (e = create_analytics_stream(s)) != hipSuccess)
      rc = hip_fail(e, "server alloc");
    for (int k = 0; k < AN_SLOTS && !rc; ++k)
      if ((e = hipEventCreateWithFlags(&s->slots[k].ready, hipEventDisableTiming)) != hipSuccess)
        rc = hip_fail(e, "server event");
  }
  // Everything the request path would otherwise allocate is sized now: a
  // device free waits for the whole device, so a reallocation in the middle
  // of the loop would stall updates behind a running PageRank or BC.
  //  - update staging for the largest step;
  if (!rc) rc = gv_reserve(g, o.max_step_pairs, o.max_step_requests);
  //  - both analytics slots, for either snapshot kind, with room for the
  //    graph to grow by half; BC workspace.
  uint64_t E = 0;
  if (!rc) rc = gv_num_edges(g, &E);
  for (int k = 0; k < AN_SLOTS && !rc; ++k) {
    s->slots[k].w.snap_min_cap = E + E / 2 + (1u << 20);
    s->slots[k].w.pf_grid = s->an_cus;
    gv_status r1 = graph_pr_prepare(g, s->slots[k].w);
    if (r1 && r1 != GV_E_UNSUPPORTED) rc = r1;
    if (!rc) rc = graph_bc_prepare(g, s->slots[k].w);
  }
  if (!rc) rc = bc_reserve(s->bcw, s->V);
  s->bcw.max_blocks = 8 * s->an_cus;
  s->bcw.yield = analytics_yield;
  s->bcw.yield_arg = s;
  if (!rc && hipStreamSynchronize(s->stream) != hipSuccess) rc = hip_fail(hipErrorUnknown, "server");
  if (rc) {
    for (int k = 0; k < AN_SLOTS; ++k) {
      pr_release(s->slots[k].w);
      if (s->slots[k].ready) hipEventDestroy(s->slots[k].ready);
    }
    bc_release(s->bcw);
    if (s->astream) hipStreamDestroy(s->astream);
    hipFree(s->stage_src);
    hipFree(s->stage_dst);
    hipFree(s->pr[0]);
    hipFree(s->pr[1]);
    hipFree(s->bc_out);
    hipFree(s->bc_src);
    delete s;
    return rc;
  }
  s->t0 = Clock::now();
  s->worker = std::thread(worker_loop, s);
  s->aworker = std::thread(analytics_loop, s);
  *out = s;
  return GV_OK;
}

gv_status gv_server_submit(gv_server *s, const gv_request *req, uint64_t *ticket) {
  if (!s || !req) {
    set_error("null server or request");
    return GV_E_INVALID;
  }
  if (req->kind < GV_REQ_INSERT || req->kind > GV_REQ_BC) {
    set_error("unknown request kind");
    return GV_E_INVALID;
  }
  if ((req->kind == GV_REQ_INSERT || req->kind == GV_REQ_DELETE) && req->n &&
      (!req->src || !req->dst || (req->mem != GV_MEM_HOST && req->mem != GV_MEM_DEVICE))) {
    set_error("update request needs src/dst and a valid mem");
    return GV_E_INVALID;
  }
  if (req->kind == GV_REQ_NEIGHBORS && (req->n > 0x7FFFFFFFull || (req->n && !req->vs))) {
    set_error("neighbour request needs vs and at most 2^31-1 vertices");
    return GV_E_INVALID;
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) {
      set_error("server is stopping");
      return GV_E_INVALID;
    }
    Item it{*req, s->next_ticket++, s->now_us()};
    s->resp.emplace_back();
    s->finished.push_back(0);
    s->queue.push_back(it);
    if (ticket) *ticket = it.ticket;
  }
  s->cv_work.notify_one();
  return GV_OK;
}

gv_status gv_server_wait(gv_server *s, uint64_t ticket, gv_response *out) {
  if (!s) return GV_E_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  if (ticket >= s->next_ticket) {
    set_error("unknown ticket");
    return GV_E_INVALID;
  }
  s->cv_done.wait(lk, [&] { return s->finished[ticket] != 0; });
  if (out) *out = s->resp[ticket];
  return GV_OK;
}

gv_status gv_server_drain(gv_server *s) {
  if (!s) return GV_E_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->n_done == s->next_ticket; });
  return GV_OK;
}

gv_status gv_server_replay(gv_server *s, const gv_request *reqs, const double *arrival_us,
                           uint64_t n, gv_response *resp) {
  if (!s || (n && (!reqs || !arrival_us || !resp))) return GV_E_INVALID;
  for (uint64_t i = 1; i < n; ++i)
    if (arrival_us[i] < arrival_us[i - 1]) {
      set_error("arrival times must be non-decreasing");
      return GV_E_INVALID;
    }
  std::vector<uint64_t> tickets(n, ~0ull);
  const auto start = Clock::now();
  gv_status rc = GV_OK;
  for (uint64_t i = 0; i < n && !rc; ++i) {
    std::this_thread::sleep_until(start + std::chrono::duration_cast<Clock::duration>(
                                              std::chrono::duration<double, std::micro>(
                                                  arrival_us[i])));
    rc = gv_server_submit(s, &reqs[i], &tickets[i]);
  }
  for (uint64_t i = 0; i < n; ++i) {
    if (tickets[i] == ~0ull) {
      std::memset(&resp[i], 0, sizeof(resp[i]));
      resp[i].status = GV_E_INVALID;
      resp[i].kind = reqs[i].kind;
      continue;
    }
    gv_server_wait(s, tickets[i], &resp[i]);
  }
  return rc;
}

gv_status gv_server_get_stats(gv_server *s, gv_server_stats *out) {
  if (!s || !out) return GV_E_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->stats;
  return GV_OK;
}

gv_status gv_server_destroy(gv_server *s) {
  if (!s) return GV_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->astop = true;  // the update worker has handed over its last job
  }
  s->cv_ajob.notify_all();
  if (s->aworker.joinable()) s->aworker.join();
  hipStreamSynchronize(s->stream);
  hipStreamSynchronize(s->astream);
  for (int k = 0; k < AN_SLOTS; ++k) {
    pr_release(s->slots[k].w);
    hipEventDestroy(s->slots[k].ready);
  }
  bc_release(s->bcw);
  hipStreamDestroy(s->astream);
  hipFree(s->stage_src);
  hipFree(s->stage_dst);
  hipFree(s->pr[0]);
  hipFree(s->pr[1]);
  hipFree(s->bc_out);
  hipFree(s->bc_src);
  delete s;
  return GV_OK;
}

}  // extern "C"
