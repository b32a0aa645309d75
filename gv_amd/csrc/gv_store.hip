// gv_store.hip -- batch staging, insert, delete, compaction and read-back
// kernels of the MI355X GraphVine store.
//
// Update pipeline for one batch (all stream-ordered, no host round trip):
//   k_stage        (src,dst) -> packed keys (u<<bV | v); self-loops and bad IDs
//                  become the sentinel (sorts last); bad IDs set abort=RANGE.
//   sort_keys      LSD radix sort on 2*bV bits.
//   k_seg_mark     head-of-source-segment and intra-batch duplicate flags.
//   scan           head flags -> segment index per key.
//   k_seg_build    per segment: first key, source, slots in use, work items.
//   scan + expand  work items of PROBE_CHUNK slots (hub adjacencies are split
//                  across many wavefronts).
//   k_probe        one wavefront per work item reads a chunk of the source's
//                  slots (coalesced, one block per 64 lanes) and binary-searches
//                  each live slot in the source's sorted batch segment:
//                  insert -> mark the batch key as already present;
//                  delete -> tombstone the slot.
//   insert only:   scan of survivors -> k_ins_plan (new blocks and table
//                  capacity per segment, packed and scanned together) ->
//                  k_commit (all-or-nothing pool check, bulk pop of the free
//                  stack) -> k_ins_tables -> k_ins_write.
//   delete only:   k_del_finish (degrees, counters).
// The reference does the same work with O(V) host and device passes per batch
// (graphvine.cu:222-451); here every pass is O(batch + touched adjacency).
#include "gv_store.h"

namespace gv {

// ------------------------------------------------------------------ fills --
template <typename T>
__global__ void k_fill(T *p, T v, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
void fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t s) {
  if (n) k_fill<uint32_t><<<grid_for(n), 256, 0, s>>>(p, v, n);
}
void fill_u64(uint64_t *p, uint64_t v, uint64_t n, hipStream_t s) {
  if (n) k_fill<uint64_t><<<grid_for(n), 256, 0, s>>>(p, v, n);
}
void fill_f32(float *p, float v, uint64_t n, hipStream_t s) {
  if (n) k_fill<float><<<grid_for(n), 256, 0, s>>>(p, v, n);
}

// ------------------------------------------------------------ store init --
gv_status store_alloc(Store &st, uint32_t V, uint32_t B, uint64_t P, uint64_t A) {
  st.V = V;
  st.B = B;
  st.log2B = ceil_log2(B);
  st.P = P;
  st.A = A;
  if (A >= 0xFFFFFFFFull || P >= 0xFFFFFFFFull) {
    set_error("pool or table arena exceeds 2^32 entries");
    return GV_E_INVALID;
  }
  GV_HIP(hipMalloc(&st.slots, P * B * sizeof(uint32_t)));
  GV_HIP(hipMalloc(&st.free_ids, P * sizeof(uint32_t)));
  GV_HIP(hipMalloc(&st.deg, (size_t)(V ? V : 1) * sizeof(uint32_t)));
  GV_HIP(hipMalloc(&st.used, (size_t)(V ? V : 1) * sizeof(uint32_t)));
  GV_HIP(hipMalloc(&st.tbl_off, (size_t)(V ? V : 1) * sizeof(uint32_t)));
  GV_HIP(hipMalloc(&st.tbl_cap, (size_t)(V ? V : 1) * sizeof(uint32_t)));
  GV_HIP(hipMalloc(&st.sorted, (size_t)(V ? V : 1) * sizeof(uint32_t)));
  GV_HIP(hipMalloc(&st.fence, P * B / FENCE_SLOTS * sizeof(uint32_t)));
  GV_HIP(hipMalloc(&st.arena, A * sizeof(uint32_t)));
  GV_HIP(hipMalloc(&st.ctl, sizeof(Ctl)));
  return GV_OK;
}

void store_release(Store &st) {
  hipFree(st.slots);
  hipFree(st.free_ids);
  hipFree(st.deg);
  hipFree(st.used);
  hipFree(st.tbl_off);
  hipFree(st.tbl_cap);
  hipFree(st.sorted);
  hipFree(st.fence);
  hipFree(st.arena);
  hipFree(st.ctl);
  st = Store();
}

__global__ void k_store_reset(uint32_t *free_ids, uint64_t P, uint32_t *deg, uint32_t *used,
                              uint32_t *tbl_off, uint32_t *tbl_cap, uint32_t *sorted, uint32_t V,
                              Ctl *ctl) {
  uint64_t n = P > V ? P : V;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (i < P) free_ids[i] = (uint32_t)i;
    if (i < V) {
      deg[i] = 0;
      used[i] = 0;
      tbl_off[i] = 0;
      tbl_cap[i] = 0;
      sorted[i] = 0;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    Ctl c = {};
    c.free_top = P;
    *ctl = c;
  }
}

void store_reset(Store &st, hipStream_t s) {
  uint64_t n = st.P > st.V ? st.P : st.V;
  k_store_reset<<<grid_for(n), 256, 0, s>>>(st.free_ids, st.P, st.deg, st.used, st.tbl_off,
                                            st.tbl_cap, st.sorted, st.V, st.ctl);
}

bool scratch_fits(const Scratch &sc, const Store &st, uint64_t m) {
  uint64_t item_need = m + cdiv(st.P * (uint64_t)st.B, PROBE_CHUNK) + 1;
  return m <= sc.cap && item_need <= sc.item_cap;
}

gv_status scratch_reserve(Scratch &sc, const Store &st, uint64_t m) {
  if (scratch_fits(sc, st, m)) return GV_OK;
  scratch_release(sc);
  uint64_t cap = 1024;
  while (cap < m) cap <<= 1;
  sc.cap = cap;
  sc.item_cap = cap + cdiv(st.P * (uint64_t)st.B, PROBE_CHUNK) + 1;
  GV_HIP(hipMalloc(&sc.k0, cap * sizeof(uint64_t)));
  GV_HIP(hipMalloc(&sc.k1, cap * sizeof(uint64_t)));
  uint32_t **arrs[] = {&sc.flag, &sc.segof, &sc.rank, &sc.seg_start, &sc.seg_src, &sc.seg_used,
                       &sc.seg_cnt, &sc.item_start, &sc.seg_blk, &sc.op};
  for (auto a : arrs) GV_HIP(hipMalloc(a, (cap + 1) * sizeof(uint32_t)));
  GV_HIP(hipMalloc(&sc.plan, (cap + 1) * sizeof(uint64_t)));
  GV_HIP(hipMalloc(&sc.item_seg, sc.item_cap * sizeof(uint32_t)));
  size_t t = sort_tmp_bytes(cap);
  size_t t2 = scan_tmp_bytes(cap + 1);
  sc.tmp_bytes = t > t2 ? t : t2;
  GV_HIP(hipMalloc(&sc.tmp, sc.tmp_bytes));
  return GV_OK;
}

void scratch_release(Scratch &sc) {
  hipFree(sc.k0);
  hipFree(sc.k1);
  uint32_t *arrs[] = {sc.flag,    sc.segof,   sc.rank,       sc.seg_start, sc.seg_src,
                      sc.seg_used, sc.seg_cnt, sc.item_start, sc.seg_blk,   sc.item_seg,
                      sc.op};
  for (auto a : arrs) hipFree(a);
  hipFree(sc.plan);
  hipFree(sc.tmp);
  sc = Scratch();
}

__device__ __forceinline__ void op_reset(Ctl *ctl) {
  ctl->need_blocks = 0;
  ctl->need_arena = 0;
  ctl->blk_base = 0;
  ctl->arena_base = 0;
  ctl->applied = 0;
  ctl->blocks_freed = 0;
  ctl->compacted = 0;
  ctl->n_valid = 0;
  ctl->n_segs = 0;
  ctl->n_items = 0;
}
__global__ void k_op_begin(Ctl *ctl) { op_reset(ctl); }
void op_begin(Store &st, hipStream_t s) { k_op_begin<<<1, 1, 0, s>>>(st.ctl); }

// --------------------------------------------------------------- staging --
// Also starts the operation: the per-op counters are reset (op_begin) and
// the sort's workspace words clr[0, nclr) are zeroed, saving two launches.
__global__ void k_stage(const uint32_t *src, const uint32_t *dst, uint64_t n, uint32_t V,
                        uint32_t bV, int mirror, int swap, uint64_t *keys, Ctl *ctl,
                        uint32_t *clr, uint64_t nclr) {
  if (blockIdx.x == 0 && threadIdx.x == 0) op_reset(ctl);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nclr;
       i += (uint64_t)gridDim.x * blockDim.x)
    clr[i] = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t u = src[i], v = dst[i];
    if (swap) {
      uint32_t t = u;
      u = v;
      v = t;
    }
    bool bad = u >= V || v >= V;
    if (bad) atomicCAS(&ctl->abort, 0u, ABORT_RANGE);
    bool drop = bad || u == v;
    uint64_t k = drop ? KEY_SENT : (((uint64_t)u << bV) | v);
    if (mirror) {
      keys[2 * i] = k;
      keys[2 * i + 1] = drop ? KEY_SENT : (((uint64_t)v << bV) | u);
    } else {
      keys[i] = k;
    }
  }
}

uint64_t *stage(Store &st, Scratch &sc, const uint32_t *src, const uint32_t *dst, uint64_t n,
                int mirror, int swap, uint64_t &m, hipStream_t s) {
  m = mirror ? 2 * n : n;
  sc.sort_err = nullptr;
  if (!n) {
    op_begin(st, s);
    return sc.k0;
  }
  const uint32_t bV = id_bits(st.V);
  const SortClear c = sort_clear_span(sc.tmp, m, 0, 2 * bV);
  k_stage<<<grid_for(n), 256, 0, s>>>(src, dst, n, st.V, bV, mirror, swap, sc.k0, st.ctl, c.w,
                                      c.words);
  if (!c.err) return sort_keys(sc.k0, sc.k1, m, 2 * bV, sc.tmp, s, &st.ctl->abort);
  // the sort's failure word is folded into abort by segment (or abort_join)
  sc.sort_err = c.err;
  return sort_keys(sc.k0, sc.k1, m, 2 * bV, sc.tmp, s, nullptr, true);
}

// Group staging: key = ((u << bV | v) << sb) | batch, so after the sort the
// entries of one edge sit together in batch order and the last one holds the
// edge's final operation.
// Like k_stage, also starts the operation and clears the sort's workspace.
constexpr uint32_t STAGE_SEQ_LDS = 2048;
__global__ void k_stage_seq(const uint32_t *src, const uint32_t *dst, uint64_t n,
                            const uint64_t *boff, const uint32_t *kinds, uint32_t nb, uint32_t V,
                            uint32_t bV, uint32_t sb, int mirror, int swap, uint64_t *keys,
                            Ctl *ctl, uint32_t *clr, uint64_t nclr) {
  // the batch offsets in LDS when they fit (the search below runs per key)
  __shared__ uint64_t s_off[STAGE_SEQ_LDS];
  const bool in_lds = nb < STAGE_SEQ_LDS;
  if (in_lds) {
    for (uint32_t k = threadIdx.x; k <= nb; k += blockDim.x) s_off[k] = boff[k];
    __syncthreads();
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) op_reset(ctl);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nclr;
       i += (uint64_t)gridDim.x * blockDim.x)
    clr[i] = 0;
  // offsets start at 0, never decrease and end within the n entries; kinds
  // are OP_INSERT or OP_DELETE (a device-filled group can break either: the
  // group is then not applied)
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k <= nb;
       k += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t o = boff[k];
    const bool bad = (k == 0 ? o != 0 : boff[k - 1] > o) || (k == nb && o > n) ||
                     (k < nb && kinds[k] > OP_DELETE);
    // ABORT_GROUP is the largest code: it replaces a range error raised by
    // entries that bad offsets put inside the group
    if (bad) atomicMax(&ctl->abort, ABORT_GROUP);
  }
  // entries past the last batch (a device-filled group shorter than its
  // capacity n) are dropped like self-loops
  const uint64_t n_used = in_lds ? s_off[nb] : boff[nb];
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (i >= n_used) {
      if (mirror) {
        keys[2 * i] = KEY_SENT;
        keys[2 * i + 1] = KEY_SENT;
      } else {
        keys[i] = KEY_SENT;
      }
      continue;
    }
    uint32_t lo = 0, hi = nb;  // largest b with boff[b] <= i
    if (in_lds) {
      while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (s_off[mid] <= i) lo = mid; else hi = mid;
      }
    } else {
      while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (boff[mid] <= i) lo = mid; else hi = mid;
      }
    }
    uint32_t u = src[i], v = dst[i];
    if (swap) {
      uint32_t t = u;
      u = v;
      v = t;
    }
    bool bad = u >= V || v >= V;
    if (bad) atomicCAS(&ctl->abort, 0u, ABORT_RANGE);
    bool drop = bad || u == v;
    uint64_t k = drop ? KEY_SENT : (((((uint64_t)u << bV) | v) << sb) | lo);
    if (mirror) {
      keys[2 * i] = k;
      keys[2 * i + 1] = drop ? KEY_SENT : (((((uint64_t)v << bV) | u) << sb) | lo);
    } else {
      keys[i] = k;
    }
  }
}

// ------------------------------------------ group selection + segments --
// After the group sort, index i of the sorted keys is kept when it is the
// last entry of its edge (the edge's final op); a kept key heads a segment
// when no kept key of its source comes before it.  Two passes over tiles of
// GS_TILE sorted keys do the selection and the segmentation: k_gs_count
// counts each tile's kept keys and heads, one scan over the tiles gives
// their bases, and k_gs_write writes the kept keys (tag removed), their ops,
// insert flags and segment numbers, and each segment's description.  Row j
// of wave w of tile t holds indexes t * GS_TILE + j * 256 + w * 64 + lane,
// so a key's neighbours in index order are its neighbour lanes.
constexpr int GS_ROWS = 16;
constexpr uint64_t GS_TILE = 256 * GS_ROWS;
// k_gs_write scans the (row, wave) cells of a tile with one wavefront, a cell
// per lane
static_assert(GS_ROWS * 4 == WAVE, "one cell per lane");

// A tile's keys are loaded before any flag is evaluated; the neighbour keys
// that the shuffles of a row cannot supply (lane 63: key i + 1, lane 0: key
// i - 1) come from the first and last keys of every (row, wave) cell,
// exchanged through LDS, and the keys just outside the tile.
struct GsEdges {
  uint64_t first[GS_ROWS * 4], last[GS_ROWS * 4], before, after;
};

__device__ __forceinline__ void gs_load_tile(const uint64_t *keys, uint64_t m, uint64_t t0,
                                             uint32_t lane, uint64_t (&k)[GS_ROWS],
                                             GsEdges &ed) {
  const uint32_t w = threadIdx.x / WAVE;
#pragma unroll
  for (int j = 0; j < GS_ROWS; ++j) {
    const uint64_t i = t0 + (uint64_t)j * 256 + threadIdx.x;
    k[j] = i < m ? keys[i] : KEY_SENT;
  }
  if (threadIdx.x == 0) ed.before = t0 > 0 && t0 < m ? keys[t0 - 1] : KEY_SENT;
  if (threadIdx.x == 1) ed.after = t0 + GS_TILE < m ? keys[t0 + GS_TILE] : KEY_SENT;
#pragma unroll
  for (int j = 0; j < GS_ROWS; ++j) {
    if (lane == 0) ed.first[j * 4 + w] = k[j];
    if (lane == WAVE - 1) ed.last[j * 4 + w] = k[j];
  }
  __syncthreads();
}

// The edge-lane neighbour of row j: key i + 1 for lane 63, key i - 1 for lane 0.
__device__ __forceinline__ uint64_t gs_edge(const GsEdges &ed, int j, uint32_t lane) {
  const uint32_t c = (uint32_t)j * 4 + threadIdx.x / WAVE;
  if (lane == WAVE - 1) return c + 1 < GS_ROWS * 4 ? ed.first[c + 1] : ed.after;
  return c > 0 ? ed.last[c - 1] : ed.before;
}

// Flags of index i (every lane of the wavefront calls it): keep = last entry
// of its edge, head = kept and the entry before the edge's first one has
// another source.  After an abort (a failed sort leaves arbitrary keys)
// nothing is kept.
__device__ __forceinline__ void gs_flags(const uint64_t *keys, uint64_t m, uint64_t i,
                                         uint32_t sb, uint32_t bV, bool ab, uint32_t lane,
                                         uint64_t k, uint64_t x, bool &keep, bool &head) {
  uint64_t nxt = __shfl_down(k, 1, WAVE);
  if (lane == WAVE - 1) nxt = x;
  uint64_t prv = __shfl_up(k, 1, WAVE);
  if (lane == 0) prv = x;
  const uint64_t e = k >> sb;
  keep = !ab && i < m && k != KEY_SENT && (i + 1 == m || (nxt >> sb) != e);
  head = false;
  if (keep) {
    // walk back over earlier entries of the same edge (repeats across batches)
    uint64_t j = i, before = i > 0 ? prv : KEY_SENT;
    while (j > 0 && (before >> sb) == e) {
      --j;
      before = j > 0 ? keys[j - 1] : KEY_SENT;
    }
    head = j == 0 || (before >> (sb + bV)) != (e >> bV);
  }
}

// tcnt[t] = (kept keys of tile t) << 32 | (heads of tile t)
__global__ __launch_bounds__(256) void k_gs_count(const uint64_t *keys, uint64_t m, uint32_t sb,
                                                  uint32_t bV, const Ctl *ctl, uint64_t *tcnt) {
  __shared__ GsEdges ed;
  const bool ab = ctl->abort != 0;
  const uint32_t lane = lane_id();
  const uint64_t t0 = (uint64_t)blockIdx.x * GS_TILE, base = t0 + threadIdx.x;
  uint64_t k[GS_ROWS];
  gs_load_tile(keys, m, t0, lane, k, ed);
  uint32_t nk = 0, nh = 0;
#pragma unroll
  for (int j = 0; j < GS_ROWS; ++j) {
    bool keep, head;
    gs_flags(keys, m, base + (uint64_t)j * 256, sb, bV, ab, lane, k[j], gs_edge(ed, j, lane),
             keep, head);
    nk += keep ? 1u : 0u;
    nh += head ? 1u : 0u;
  }
  const unsigned long long t =
      block_sum<unsigned long long, 256>(((unsigned long long)nk << 32) | nh);
  if (threadIdx.x == 0) tcnt[blockIdx.x] = t;
}

struct SegArgs {
  uint32_t *segof, *seg_start, *seg_src, *seg_used, *seg_cnt, *seg_blk;
};

// toff: exclusive scan of tcnt over the nt tiles, toff[nt] = the totals.
// Writes, for the kept key of rank p: out[p] (tag removed), op[p], flag[p]
// (an insert: no duplicates are left) and segof[p] (its segment); per
// segment: first key, source, used slots, probe items; n_valid and n_segs.
// The rest of out is padded with the sentinel.
__global__ __launch_bounds__(256) void k_gs_write(const uint64_t *keys, uint64_t m, uint32_t sb,
                                                  uint32_t bV, const uint64_t *toff, uint64_t nt,
                                                  const uint32_t *kinds, uint64_t *out,
                                                  uint32_t *op, uint32_t *flag, SegArgs a,
                                                  const uint32_t *used, const uint32_t *sorted,
                                                  Ctl *ctl) {
  __shared__ uint32_t cell[GS_ROWS * 4];  // per (row, wave): kept << 16 | heads
  __shared__ GsEdges ed;
  const bool ab = ctl->abort != 0;
  const uint32_t lane = lane_id(), w = threadIdx.x / WAVE;
  const uint64_t t0 = (uint64_t)blockIdx.x * GS_TILE, base = t0 + threadIdx.x;
  const uint64_t smask = (1ull << sb) - 1;
  uint64_t key[GS_ROWS];
  gs_load_tile(keys, m, t0, lane, key, ed);
  uint32_t bits = 0;  // bit j: row j kept, bit GS_ROWS + j: row j heads a segment
#pragma unroll
  for (int j = 0; j < GS_ROWS; ++j) {
    bool keep, head;
    gs_flags(keys, m, base + (uint64_t)j * 256, sb, bV, ab, lane, key[j], gs_edge(ed, j, lane),
             keep, head);
    bits |= (keep ? 1u << j : 0u) | (head ? 1u << (GS_ROWS + j) : 0u);
    const unsigned long long mk = __ballot(keep), mh = __ballot(head);
    if (lane == 0) cell[j * 4 + w] = ((uint32_t)__popcll(mk) << 16) | (uint32_t)__popcll(mh);
  }
  __syncthreads();
  // exclusive prefix of the 64 cells in index order (row-major), lane c = cell c
  const uint32_t c = cell[lane];
  const uint32_t cex = wave_incl_scan<uint32_t>(c) - c;
  const uint64_t tb = toff[blockIdx.x], tot = toff[nt];
  const uint32_t K = (uint32_t)(tot >> 32), H = (uint32_t)tot;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->n_valid = K;
    ctl->n_segs = H;
    a.seg_start[H] = K;
  }
  // the keys are read again (cache hits) rather than held across the loop
#pragma unroll 4
  for (int j = 0; j < GS_ROWS; ++j) {
    const uint64_t i = base + (uint64_t)j * 256;
    const uint32_t ex = __shfl(cex, j * 4 + (int)w, WAVE);
    const bool keep = (bits >> j) & 1u, head = (bits >> (GS_ROWS + j)) & 1u;
    const unsigned long long mk = __ballot(keep), mh = __ballot(head);
    if (keep) {
      const uint32_t pk = (uint32_t)(tb >> 32) + (ex >> 16) + mask_lt_count(mk);
      const uint32_t ph = (uint32_t)tb + (ex & 0xFFFFu) + mask_lt_count(mh);  // heads before
      const uint64_t kk = keys[i];
      const uint64_t e = kk >> sb;
      const uint32_t o = kinds[kk & smask] == OP_INSERT ? OP_INSERT : OP_DELETE;
      out[pk] = e;
      op[pk] = o;
      flag[pk] = o == OP_INSERT ? 1u : 0u;
      a.segof[pk] = head ? ph : ph - 1u;
      if (head) {
        const uint32_t u = (uint32_t)(e >> bV), uo = used[u];
        a.seg_start[ph] = pk;
        a.seg_src[ph] = u;
        a.seg_used[ph] = uo;
        a.seg_cnt[ph] = probe_items(uo, sorted[u]);
        a.seg_blk[ph] = 0;
      }
    }
    if (i < m && i >= K) {
      out[i] = KEY_SENT;
      flag[i] = 0;
    }
  }
}

uint64_t *stage_group(Store &st, Scratch &sc, const uint32_t *src, const uint32_t *dst,
                      uint64_t n, const uint64_t *boff, const uint32_t *kinds, uint32_t nb,
                      int mirror, int swap, uint64_t &m, hipStream_t s) {
  m = mirror ? 2 * n : n;
  sc.sort_err = nullptr;
  if (!n) {
    op_begin(st, s);
    return sc.k0;
  }
  const uint32_t bV = id_bits(st.V), sb = id_bits(nb);
  const SortClear c = sort_clear_span(sc.tmp, m, sb, 2 * bV + sb);
  k_stage_seq<<<grid_for(n), 256, 0, s>>>(src, dst, n, boff, kinds, nb, st.V, bV, sb, mirror,
                                          swap, sc.k0, st.ctl, c.w, c.words);
  // the keys are staged in batch order, so a stable sort on the (u, v) bits
  // alone leaves equal edges in batch order: the same array as a sort on the
  // whole key, with one pass fewer when the tag crosses a digit boundary
  uint64_t *sorted = sort_keys_range(sc.k0, sc.k1, m, sb, 2 * bV + sb, sc.tmp, s,
                                     &st.ctl->abort, c.words != 0);
  uint64_t *other = sorted == sc.k0 ? sc.k1 : sc.k0;
  // tile counts in sc.plan (cap + 1 words; the insert plan is formed later)
  const uint64_t nt = cdiv(m, GS_TILE);
  k_gs_count<<<dim3((unsigned)nt), 256, 0, s>>>(sorted, m, sb, bV, st.ctl, sc.plan);
  scan_u64(sc.plan, sc.plan, nt, nullptr, sc.tmp, s);
  SegArgs a{sc.segof, sc.seg_start, sc.seg_src, sc.seg_used, sc.seg_cnt, sc.seg_blk};
  k_gs_write<<<dim3((unsigned)nt), 256, 0, s>>>(sorted, m, sb, bV, sc.plan, nt, kinds, other,
                                                sc.op, sc.flag, a, st.used, st.sorted, st.ctl);
  return other;
}

// Both stores of a directed graph stage the same batch; an abort in either
// (a failed sort) must stop both before anything is written.
__global__ void k_abort_join(Ctl *a, Ctl *b, const uint32_t *err_a, const uint32_t *err_b) {
  if (err_a && *err_a && !a->abort) a->abort = ABORT_SORT;
  if (err_b && *err_b && !b->abort) b->abort = ABORT_SORT;
  const unsigned ab = a->abort ? a->abort : b->abort;
  if (!a->abort) a->abort = ab;
  if (!b->abort) b->abort = ab;
}
void abort_join(Store &a, Store &b, const Scratch &sa, const Scratch &sb, hipStream_t s) {
  k_abort_join<<<1, 1, 0, s>>>(a.ctl, b.ctl, sa.sort_err, sb.sort_err);
}

// ---------------------------------------------------------- segmentation --
// op (may be null): per-key operation of a merged group, OP_DELETE keys are
// never insert candidates.  After an abort nothing is segmented: the keys of
// a failed sort are arbitrary, and no later step may index with them.
// serr (may be null): the staging sort's failure word, folded into abort here.
__global__ void k_seg_mark(const uint64_t *keys, uint64_t m, uint32_t bV, uint32_t *head,
                           uint32_t *flag, const uint32_t *op, Ctl *ctl, const uint32_t *serr) {
  const bool sfail = serr && *serr;
  if (sfail && blockIdx.x == 0 && threadIdx.x == 0) atomicCAS(&ctl->abort, 0u, ABORT_SORT);
  const bool ab = sfail || ctl->abort != 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t k = keys[i];
    bool valid = !ab && k != KEY_SENT;
    uint64_t prev = i ? keys[i - 1] : KEY_SENT;
    bool h = valid && (i == 0 || (prev >> bV) != (k >> bV));
    bool dup = valid && i > 0 && prev == k;
    head[i] = h ? 1u : 0u;
    flag[i] = (valid && !dup && (!op || op[i] == OP_INSERT)) ? 1u : 0u;
    if (valid) {
      uint64_t next = (i + 1 < m) ? keys[i + 1] : KEY_SENT;
      if (next == KEY_SENT) ctl->n_valid = (unsigned)(i + 1);
    }
  }
}

__global__ void k_seg_build(const uint64_t *keys, uint64_t m, uint32_t bV, SegArgs a,
                            const uint32_t *used, const uint32_t *sorted, Ctl *ctl) {
  const uint32_t nv = ctl->n_valid;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t k = keys[i];
    bool h = i == 0 || (keys[i - 1] >> bV) != (k >> bV);
    uint32_t sidx = a.segof[i] + (h ? 1u : 0u) - 1u;
    a.segof[i] = sidx;
    if (h) {
      uint32_t u = (uint32_t)(k >> bV);
      uint32_t uo = used[u];
      a.seg_start[sidx] = (uint32_t)i;
      a.seg_src[sidx] = u;
      a.seg_used[sidx] = uo;
      a.seg_cnt[sidx] = probe_items(uo, sorted[u]);
      a.seg_blk[sidx] = 0;
    }
    if (i == 0) {
      uint32_t T = a.segof[m];
      a.seg_start[T] = nv;
      ctl->n_segs = T;
    }
  }
}

__global__ void k_item_expand(const uint32_t *seg_cnt, const uint32_t *item_start,
                              uint32_t *item_seg, Ctl *ctl) {
  const uint32_t T = ctl->n_segs;
  for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < T;
       s += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t b = item_start[s], c = seg_cnt[s];
    for (uint32_t j = 0; j < c; ++j) item_seg[b + j] = (uint32_t)s;
    if (s == 0) ctl->n_items = item_start[T];
  }
}

void segment(Store &st, Scratch &sc, const uint64_t *keys, uint64_t m, hipStream_t s,
             const uint32_t *op) {
  if (!m) return;
  const uint32_t bV = id_bits(st.V);
  k_seg_mark<<<grid_for(m), 256, 0, s>>>(keys, m, bV, sc.segof, sc.flag, op, st.ctl,
                                         sc.sort_err);
  segment_marked(st, sc, keys, m, s);
}

void segment_marked(Store &st, Scratch &sc, const uint64_t *keys, uint64_t m, hipStream_t s) {
  if (!m) return;
  const uint32_t bV = id_bits(st.V);
  scan_u32(sc.segof, sc.segof, m, nullptr, sc.tmp, s);
  SegArgs a{sc.segof, sc.seg_start, sc.seg_src, sc.seg_used, sc.seg_cnt, sc.seg_blk};
  k_seg_build<<<grid_for(m), 256, 0, s>>>(keys, m, bV, a, st.used, st.sorted, st.ctl);
  segment_items(st, sc, m, s);
}

void segment_items(Store &st, Scratch &sc, uint64_t m, hipStream_t s) {
  if (!m) return;
  scan_u32(sc.seg_cnt, sc.item_start, m, &st.ctl->n_segs, sc.tmp, s);
  k_item_expand<<<grid_for(m), 256, 0, s>>>(sc.seg_cnt, sc.item_start, sc.item_seg, st.ctl);
}

// ----------------------------------------------------------------- probe --
struct ProbeArgs {
  const uint64_t *keys;
  const uint32_t *item_seg, *item_start, *seg_src, *seg_used, *seg_start;
  uint32_t *flag;       // insert: survivor flags
  uint32_t *seg_del;    // delete: per-segment tombstone counts
  const uint32_t *op;   // mixed: per-key operation
};

// Probe modes: every matching slot is an existing edge; what happens to it
// depends on the batch kind (or, in a merged group, on the key's last op).
enum : int { PROBE_INSERT = 0, PROBE_DELETE = 1, PROBE_MIXED = 2 };

// Act on a live slot that matches batch key p.
template <int MODE>
__device__ __forceinline__ void probe_hit(const Store &st, const ProbeArgs &a, uint32_t idx,
                                          uint32_t p, uint32_t &cnt) {
  bool del = MODE == PROBE_DELETE || (MODE == PROBE_MIXED && a.op[p] == OP_DELETE);
  if (del) {
    st.slots[idx] = TOMB;
    ++cnt;
  } else {
    a.flag[p] = 0;
  }
}

constexpr int PROBE_ROUNDS = PROBE_CHUNK / WAVE;
// A scan item of at most PROBE_LANE_SLOTS slots whose source has at most
// LaneProbe::keys batch keys is probed by one lane on its own, 16 slots per
// round (a multiple of 16, at most PROBE_CHUNK): 4 keys in a plain batch,
// 8 in a merged group, whose touched sources carry several batches' keys
// (about 5 at 10 R-MAT scale-22 batches, 10 at 20).
constexpr uint32_t PROBE_LANE_SLOTS = 128;
constexpr uint32_t PROBE_SMALL_KEYS = 4;
template <int MODE>
struct LaneProbe {
  static constexpr uint32_t slots = PROBE_LANE_SLOTS;
  static constexpr uint32_t keys = MODE == 2 ? 8 : PROBE_SMALL_KEYS;
  static constexpr uint32_t kbits = MODE == 2 ? 3 : 2;  // bits of a packed key index
};
// A wavefront-wide item whose source has at most this many batch keys
// compares each slot with every key.
constexpr uint32_t PROBE_FEW_KEYS = 8;
// A fence item looks its keys' chunks up when its source has more than 64
// batch keys (a scan would search memory for every slot) or at most one per
// PROBE_SPARSE slots of sorted prefix; otherwise it is scanned.
constexpr uint32_t PROBE_SPARSE = 32;

// First key of the source's sorted batch segment (keys[bks, bks + k), k > 64)
// that is >= value x, relative to bks: p = the number of the 64 key samples
// (lane l of the wavefront holding keys[bks + l*k/64]) below x brackets it,
// then memory is searched over ~k/64 keys.
__device__ __forceinline__ uint32_t probe_seg_lb(const uint64_t *keys, uint32_t bks, uint32_t k,
                                                 uint64_t ukey, uint32_t p, uint32_t x) {
  uint32_t l2 = p ? bks + (uint32_t)(((uint64_t)(p - 1) * k) >> 6) + 1 : bks;
  uint32_t h2 = p < (uint32_t)WAVE ? bks + (uint32_t)(((uint64_t)p * k) >> 6) : bks + k;
  const uint64_t key = ukey | x;
  while (l2 < h2) {
    const uint32_t mid = (l2 + h2) >> 1;
    if (keys[mid] < key) l2 = mid + 1; else h2 = mid;
  }
  return l2 - bks;
}

// Key kt against the 16 slots x of a chunk starting at slot index cb: the
// values in a store adjacency are distinct, so at most one slot matches.
template <int MODE>
__device__ __forceinline__ void probe_chunk_hit(const Store &st, const ProbeArgs &a,
                                                const uint32_t (&x)[FENCE_SLOTS], uint32_t kt,
                                                uint32_t cb, uint32_t p, uint32_t &cnt) {
  uint32_t hm = 0;
#pragma unroll
  for (uint32_t j = 0; j < FENCE_SLOTS; ++j) hm |= x[j] == kt ? 1u << j : 0u;
  if (hm) probe_hit<MODE>(st, a, cb + (uint32_t)__ffs(hm) - 1, p, cnt);
}

// A fence item: prefix slots [blo, bhi) of one source (multiples of
// FENCE_SLOTS, at most 64 chunks), lane l owning chunk blo/16 + l.  The lane
// reads its chunk's fence and the next one, which bound the chunk's values;
// only when a batch key of the source lies in that range does it read the
// chunk's 16 slots (four 16-byte loads) and compare them with those keys.  sv:
// the source's keys one per lane (k <= 64, padded with TOMB) or 64 key samples.
// The first of equal keys acts, as on the scan path.
template <int MODE>
__device__ __forceinline__ void probe_fences(const Store &st, const ProbeArgs &a, uint64_t mask,
                                             uint64_t ukey, uint32_t lane, uint32_t blo,
                                             uint32_t bhi, uint32_t bso, uint32_t btbl,
                                             uint32_t bks, uint32_t k, uint32_t sv,
                                             uint32_t &cnt) {
  const uint32_t nch = (bhi - blo) / FENCE_SLOTS;
  const bool on = lane < nch;
  uint32_t cb = 0, f = TOMB;
  if (on) {
    cb = slot_index(st.arena, btbl, blo + lane * FENCE_SLOTS, st.log2B, st.B);
    f = st.fence[cb / FENCE_SLOTS];
  }
  // exclusive upper bound of the chunk's values: the next chunk's fence, or
  // TOMB (above every ID) past the prefix
  uint32_t fn = __shfl(f, (int)((lane + 1) & (WAVE - 1)), WAVE);
  if (lane + 1 == nch) {
    fn = TOMB;
    if (bhi < bso) fn = st.fence[slot_index(st.arena, btbl, bhi, st.log2B, st.B) / FENCE_SLOTS];
  }
  const uint32_t last = __shfl(sv, WAVE - 1, WAVE);
  uint32_t p0 = wave_lower_bound(sv, f), p1 = wave_lower_bound(sv, fn);
  if (p0 == WAVE - 1 && last < f) p0 = WAVE;
  if (p1 == WAVE - 1 && last < fn) p1 = WAVE;
  // keys [i0, i1) of the segment (relative to bks) fall in the chunk's range
  uint32_t i0 = 0, i1 = 0;
  if (on) {
    if (k <= (uint32_t)WAVE) {
      i0 = p0;
      i1 = p1;
    } else {
      i0 = probe_seg_lb(a.keys, bks, k, ukey, p0, f);
      i1 = fn == TOMB ? k : probe_seg_lb(a.keys, bks, k, ukey, p1, fn);
    }
  }
  const uint32_t nk = i1 - i0;
  uint32_t mk = nk;
#pragma unroll
  for (int d = WAVE / 2; d >= 1; d >>= 1) {
    const uint32_t o = __shfl_xor(mk, d, WAVE);
    mk = o > mk ? o : mk;
  }
  if (!mk) return;
  uint32_t x[FENCE_SLOTS];
  const uint4 *p4 = reinterpret_cast<const uint4 *>(st.slots + cb);
#pragma unroll
  for (uint32_t r = 0; r < FENCE_SLOTS / 4; ++r) {
    const uint4 q = nk ? p4[r] : make_uint4(TOMB, TOMB, TOMB, TOMB);
    x[4 * r] = q.x;
    x[4 * r + 1] = q.y;
    x[4 * r + 2] = q.z;
    x[4 * r + 3] = q.w;
  }
  uint32_t kprev = TOMB;
  if (k <= (uint32_t)WAVE) {
    // keys from their lanes: a uniform loop, every lane shuffling
    for (uint32_t t = 0; t < mk; ++t) {
      const uint32_t kt = __shfl(sv, (int)((i0 + t) & (WAVE - 1)), WAVE);
      if (t < nk && kt != kprev) probe_chunk_hit<MODE>(st, a, x, kt, cb, bks + i0 + t, cnt);
      kprev = kt;
    }
  } else {
    for (uint32_t t = 0; t < nk; ++t) {
      const uint32_t kt = (uint32_t)(a.keys[bks + i0 + t] & mask);
      if (kt != kprev) probe_chunk_hit<MODE>(st, a, x, kt, cb, bks + i0 + t, cnt);
      kprev = kt;
    }
  }
}

// A scan item: slots [blo, bhi) of one source (blo a multiple of
// PROBE_CHUNK), read in 64-slot rounds, each live slot searched among the
// source's batch keys (sv as in probe_fences).
template <int MODE>
__device__ __forceinline__ void probe_scan(const Store &st, const ProbeArgs &a, uint32_t bV,
                                           uint32_t bu, uint32_t lane, uint32_t blo,
                                           uint32_t bhi, uint32_t btbl, uint32_t bks,
                                           uint32_t bke, uint32_t k, uint32_t sv,
                                           uint32_t &cnt) {
  uint32_t idx[PROBE_ROUNDS], wv[PROBE_ROUNDS];
  if (st.log2B >= 6) {
    // blo is a multiple of PROBE_CHUNK, so each 64-slot round lies in one
    // block: one uniform block-table load per round
#pragma unroll
    for (int q = 0; q < PROBE_ROUNDS; ++q) {
      const uint32_t base = blo + q * WAVE;
      const uint32_t blk = base < bhi ? st.arena[btbl + (base >> st.log2B)] : 0u;
      const uint32_t pos = base + lane;
      idx[q] = pos < bhi ? blk * st.B + (pos & (st.B - 1)) : 0xFFFFFFFFu;
    }
  } else {
#pragma unroll
    for (int q = 0; q < PROBE_ROUNDS; ++q) {
      const uint32_t pos = blo + q * WAVE + lane;
      idx[q] = pos < bhi ? slot_index(st.arena, btbl, pos, st.log2B, st.B) : 0xFFFFFFFFu;
    }
  }
#pragma unroll
  for (int q = 0; q < PROBE_ROUNDS; ++q)
    wv[q] = idx[q] != 0xFFFFFFFFu ? st.slots[idx[q]] : TOMB;
  if (k <= PROBE_FEW_KEYS) {
    // few keys (most touched vertices): wave-uniform copies, one compare
    // per key and slot instead of a search over the lanes
    uint32_t kv[PROBE_FEW_KEYS];
#pragma unroll
    for (uint32_t j = 0; j < PROBE_FEW_KEYS; ++j)
      kv[j] = __builtin_amdgcn_readlane(sv, (int)j);
#pragma unroll
    for (int q = 0; q < PROBE_ROUNDS; ++q) {
      uint32_t p = PROBE_FEW_KEYS;
#pragma unroll
      for (uint32_t j = PROBE_FEW_KEYS; j-- > 0;)
        if (kv[j] == wv[q]) p = j;
      if (wv[q] != TOMB && p < k) probe_hit<MODE>(st, a, idx[q], bks + p, cnt);
    }
  } else if (k <= (uint32_t)WAVE) {
#pragma unroll
    for (int q = 0; q < PROBE_ROUNDS; ++q) {
      const uint32_t p = wave_lower_bound(sv, wv[q]);
      const uint32_t hit = __shfl(sv, (int)p, WAVE);
      if (wv[q] != TOMB && hit == wv[q]) probe_hit<MODE>(st, a, idx[q], bks + p, cnt);
    }
  } else {
    // the key samples bracket the lower bound: keys[f(p-1)] < x <= keys[f(p)]
    // with f(l) = bks + l*k/64, so memory is searched over ~k/64 keys
    const uint64_t ukey = (uint64_t)bu << bV;
    const uint32_t last = __shfl(sv, WAVE - 1, WAVE);
#pragma unroll
    for (int q = 0; q < PROBE_ROUNDS; ++q) {
      uint32_t p = wave_lower_bound(sv, wv[q]);  // at most 63: all 64 samples < x gives 64
      if (p == WAVE - 1 && last < wv[q]) p = WAVE;
      if (wv[q] == TOMB) continue;
      const uint64_t key = ukey | wv[q];
      uint32_t l2 = p ? bks + (uint32_t)(((uint64_t)(p - 1) * k) >> 6) + 1 : bks;
      uint32_t h2 = p < (uint32_t)WAVE ? bks + (uint32_t)(((uint64_t)p * k) >> 6) : bke;
      while (l2 < h2) {
        const uint32_t mid = (l2 + h2) >> 1;
        if (a.keys[mid] < key) l2 = mid + 1; else h2 = mid;
      }
      if (l2 < bke && a.keys[l2] == key) probe_hit<MODE>(st, a, idx[q], l2, cnt);
    }
  }
}

// Work items (fence items over a source's sorted prefix, scan items of
// PROBE_CHUNK slots over the rest) are dealt to lanes: wave w of
// W, step r, lane l takes item (r * 64 + l) * W + w, so the items of one hub
// spread over many wavefronts.  A small item (<= PROBE_LANE_SLOTS slots,
// <= LaneProbe::keys batch keys -- most touched vertices of an R-MAT batch)
// is probed by its lane alone: 64 of them share one chain of dependent loads
// instead of paying it one after another.  The other items of the step are
// then probed by the whole wavefront, one at a time, their metadata broadcast
// from the owning lane (probe_fences, probe_scan); the source's sorted batch
// segment is held in registers (one key per lane) when it has at most 64
// keys, otherwise searched in memory.
template <int MODE>
__global__ __launch_bounds__(256, 8) void k_probe(Store st, ProbeArgs a, uint32_t bV, Ctl *ctl) {
  if (MODE != PROBE_INSERT && ctl->abort) return;
  const uint32_t nitems = ctl->n_items;
  const uint32_t lane = lane_id();
  const uint64_t mask = (1ull << bV) - 1;
  const uint64_t W = (uint64_t)gridDim.x * blockDim.x / WAVE;
  const uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
  for (uint64_t r = 0; r * WAVE * W + w < nitems; ++r) {
    const uint64_t it = (r * WAVE + lane) * W + w;
    const bool valid = it < nitems;
    uint32_t sidx = 0, lo = 0, hi = 0, u = 0, tbl = 0, ks = 0, ke = 0, uo = 0, so = 0;
    bool fitem = false;
    if (valid) {
      sidx = a.item_seg[it];
      const uint32_t c = (uint32_t)it - a.item_start[sidx];
      u = a.seg_src[sidx];
      uo = a.seg_used[sidx];
      tbl = st.tbl_off[u];
      so = st.sorted[u];
      ks = a.seg_start[sidx];
      ke = a.seg_start[sidx + 1];
      // the item list of probe_items: fence items over [0, pf), then scan
      // items over [pf, uo)
      const uint32_t pf = probe_pf(so);
      const uint32_t nf = (pf + FENCE_ITEM - 1) / FENCE_ITEM;
      if (c < nf) {
        fitem = true;
        lo = c * FENCE_ITEM;
        hi = lo + FENCE_ITEM < pf ? lo + FENCE_ITEM : pf;
      } else {
        lo = pf + (c - nf) * PROBE_CHUNK;
        hi = lo + PROBE_CHUNK < uo ? lo + PROBE_CHUNK : uo;
      }
    }
    using LP = LaneProbe<MODE>;
    const bool small = valid && !fitem && hi - lo <= LP::slots && ke - ks <= LP::keys;
    if (small) {
      // the keys' destinations, TOMB past the segment (never equal to a live slot)
      uint32_t kv[LP::keys];
#pragma unroll
      for (uint32_t j = 0; j < LP::keys; ++j)
        kv[j] = ks + j < ke ? (uint32_t)(a.keys[ks + j] & mask) : TOMB;
      uint32_t cnt = 0;
      const uint32_t n = hi - lo;
      // lo is a multiple of PROBE_CHUNK, so each round of 16 slots lies in
      // one block; the block-table entry is loaded as a round enters a new
      // block, then the round's slots as four 16-byte loads issued together
      uint32_t blk = 0;
#pragma unroll 1
      for (uint32_t r = 0; 16 * r < n; ++r) {
        const uint32_t pos = lo + 16 * r;
        if (r == 0 || (pos & (st.B - 1)) == 0) blk = st.arena[tbl + (pos >> st.log2B)];
        const uint32_t rb = blk * st.B + (pos & (st.B - 1));
        const uint32_t nr = n - 16 * r;
        const uint4 *p4 = reinterpret_cast<const uint4 *>(st.slots + rb);
        uint32_t x[16];
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) {
          const uint4 v = 4 * q < nr ? p4[q] : make_uint4(TOMB, TOMB, TOMB, TOMB);
          x[4 * q] = v.x;
          x[4 * q + 1] = v.y;
          x[4 * q + 2] = v.z;
          x[4 * q + 3] = v.w;
        }
        // matching slots (hm) and, LP::kbits bits each, their first matching key
        // (the lower bound, as on the wavefront path); hits are acted on after
        uint32_t hm = 0;
        uint64_t hp = 0;
#pragma unroll
        for (uint32_t j = 0; j < 16; ++j) {
          uint32_t p = LP::keys;
#pragma unroll
          for (uint32_t i = LP::keys; i-- > 0;)
            if (kv[i] == x[j]) p = i;
          const bool hit = j < nr && x[j] != TOMB && p < LP::keys;
          hm |= hit ? 1u << j : 0u;
          hp |= hit ? (uint64_t)p << (LP::kbits * j) : 0ull;
        }
        while (hm) {
          const uint32_t j = (uint32_t)__ffs(hm) - 1;
          hm &= hm - 1;
          probe_hit<MODE>(st, a, rb + j,
                          ks + (uint32_t)((hp >> (LP::kbits * j)) & ((1u << LP::kbits) - 1)), cnt);
        }
      }
      if (MODE != PROBE_INSERT && cnt) atomicAdd(&a.seg_del[sidx], cnt);
    }
    unsigned long long big = __ballot(valid && !small);
    while (big) {
      const int src = __ffsll(big) - 1;
      big &= big - 1;
      // the item's metadata is wave-uniform: scalar registers
      const uint32_t bsidx = __builtin_amdgcn_readfirstlane(__shfl(sidx, src, WAVE));
      const uint32_t blo = __builtin_amdgcn_readfirstlane(__shfl(lo, src, WAVE));
      const uint32_t bhi = __builtin_amdgcn_readfirstlane(__shfl(hi, src, WAVE));
      const uint32_t bu = __builtin_amdgcn_readfirstlane(__shfl(u, src, WAVE));
      const uint32_t btbl = __builtin_amdgcn_readfirstlane(__shfl(tbl, src, WAVE));
      const uint32_t bks = __builtin_amdgcn_readfirstlane(__shfl(ks, src, WAVE));
      const uint32_t bke = __builtin_amdgcn_readfirstlane(__shfl(ke, src, WAVE));
      const uint32_t bso = __builtin_amdgcn_readfirstlane(__shfl(so, src, WAVE));
      const uint32_t bfit =
          __builtin_amdgcn_readfirstlane(__shfl(fitem ? 1u : 0u, src, WAVE));
      const uint32_t k = bke - bks;
      // batch keys first, independent of the slot chain: the whole segment
      // (one key per lane) when it has <= 64 keys, else 64 fences, lane l
      // holding key bks + l*k/64
      uint32_t sv = 0xFFFFFFFFu;
      if (k <= (uint32_t)WAVE) {
        if (lane < k) sv = (uint32_t)(a.keys[bks + lane] & mask);
      } else {
        sv = (uint32_t)(a.keys[bks + (uint32_t)(((uint64_t)lane * k) >> 6)] & mask);
      }
      uint32_t cnt = 0;
      if (bfit && (k > (uint32_t)WAVE || (uint64_t)k * PROBE_SPARSE <= bso)) {
        probe_fences<MODE>(st, a, mask, (uint64_t)bu << bV, lane, blo, bhi, bso, btbl, bks, k,
                           sv, cnt);
      } else if (bfit) {
        // at most 64 keys (held in registers), dense in the prefix: reading
        // every chunk costs no more than looking each key's chunk up
        for (uint32_t q0 = blo; q0 < bhi; q0 += PROBE_CHUNK)
          probe_scan<MODE>(st, a, bV, bu, lane, q0, q0 + PROBE_CHUNK < bhi ? q0 + PROBE_CHUNK : bhi,
                           btbl, bks, bke, k, sv, cnt);
      } else {
        probe_scan<MODE>(st, a, bV, bu, lane, blo, bhi, btbl, bks, bke, k, sv, cnt);
      }
      if (MODE != PROBE_INSERT) {
        cnt = wave_sum(cnt);
        if (lane == 0 && cnt) atomicAdd(&a.seg_del[bsidx], cnt);
      }
    }
  }
}

static unsigned probe_grid(uint64_t items_bound) {
  uint64_t g = cdiv(items_bound ? items_bound : 1, 4 * 16);  // 4 waves per block, >= 16 items each
  return (unsigned)(g < 2048 ? g : 2048);
}

// ---------------------------------------------------------------- insert --
struct PlanArgs {
  const uint32_t *seg_src, *seg_used, *seg_start, *rank;
  uint64_t *plan;
};

__device__ __forceinline__ uint32_t pow2ceil(uint32_t x) {
  return x <= 1 ? 1u : 1u << (32 - __clz(x - 1));
}

// Per segment: new blocks (high 32 bits) and new table capacity (low 32 bits)
// packed in one word, so a single exclusive scan gives every segment its
// block base and table base with no atomics.
__global__ void k_ins_plan(Store st, PlanArgs a, Ctl *ctl) {
  const uint32_t T = ctl->n_segs;
  for (uint64_t sidx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; sidx < T;
       sidx += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t u = a.seg_src[sidx];
    uint32_t uo = a.seg_used[sidx];
    uint32_t surv = a.rank[a.seg_start[sidx + 1]] - a.rank[a.seg_start[sidx]];
    uint32_t nbo = (uo + st.B - 1) >> st.log2B;
    uint32_t nbn = (uo + surv + st.B - 1) >> st.log2B;
    uint32_t ncap = nbn > st.tbl_cap[u] ? pow2ceil(nbn) : 0u;
    a.plan[sidx] = ((uint64_t)(nbn - nbo) << 32) | ncap;
  }
}

__device__ bool commit_one(Ctl *c, uint64_t A) {
  return c->need_blocks <= c->free_top && c->arena_top + c->need_arena <= A;
}
__device__ void commit_take(Ctl *c) {
  c->blk_base = c->free_top - c->need_blocks;
  c->free_top = c->blk_base;
  c->arena_base = c->arena_top;
  c->arena_top += c->need_arena;
}

// All-or-nothing: either both stores get their blocks or neither does.
// rank_* hold each store's survivor prefix, plan_* the scanned plan words.
__global__ void k_commit(Ctl *a, uint64_t Aa, Ctl *b, uint64_t Ab, const uint32_t *rank_a,
                         const uint32_t *rank_b, const uint64_t *plan_a, const uint64_t *plan_b) {
  {
    uint64_t pa = plan_a[a->n_segs];
    a->need_blocks = pa >> 32;
    a->need_arena = pa & 0xFFFFFFFFull;
    if (b) {
      uint64_t pb = plan_b[b->n_segs];
      b->need_blocks = pb >> 32;
      b->need_arena = pb & 0xFFFFFFFFull;
    }
  }
  unsigned ab = a->abort | (b ? b->abort : 0u);
  if (ab) {
    a->abort = a->abort ? a->abort : ab;
    if (b) b->abort = b->abort ? b->abort : ab;
    return;
  }
  bool ok = commit_one(a, Aa) && (!b || commit_one(b, Ab));
  if (!ok) {
    a->abort = ABORT_POOL;
    if (b) b->abort = ABORT_POOL;
    return;
  }
  commit_take(a);
  unsigned long long ap = rank_a[a->n_valid];
  a->applied += ap;
  a->live_edges += ap;
  if (b) {
    commit_take(b);
    unsigned long long bp = rank_b[b->n_valid];
    b->applied += bp;
    b->live_edges += bp;
  }
}

// seg_tbl / seg_rb (written here, read by k_ins_write): per segment with
// survivors, its source's block table after growth and its survivors' rank
// base; they reuse the segment arrays of the delete counts and item counts,
// which nothing reads after the probe.
struct TblArgs {
  const uint32_t *seg_src, *seg_used, *seg_start, *rank;
  const uint64_t *plan;
  uint32_t *seg_tbl, *seg_rb;
};

__global__ void k_ins_tables(Store st, TblArgs a, Ctl *ctl) {
  if (ctl->abort) return;
  const uint32_t T = ctl->n_segs;
  const uint64_t bb0 = ctl->blk_base, ab0 = ctl->arena_base;
  for (uint64_t sidx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; sidx < T;
       sidx += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t u = a.seg_src[sidx];
    uint32_t uo = a.seg_used[sidx];
    const uint32_t rb = a.rank[a.seg_start[sidx]];
    uint32_t surv = a.rank[a.seg_start[sidx + 1]] - rb;
    if (!surv) continue;
    uint32_t nbo = (uo + st.B - 1) >> st.log2B;
    uint32_t nbn = (uo + surv + st.B - 1) >> st.log2B;
    uint32_t tbl = st.tbl_off[u];
    // the survivors of an empty vertex are written in key order: a sorted
    // prefix (a run into a non-empty vertex starts or extends its tail)
    if (uo == 0) st.sorted[u] = surv;
    const uint64_t pl = a.plan[sidx];
    if (nbn > st.tbl_cap[u]) {
      uint32_t noff = (uint32_t)(ab0 + (pl & 0xFFFFFFFFull));
      for (uint32_t j = 0; j < nbo; ++j) st.arena[noff + j] = st.arena[tbl + j];
      tbl = noff;
      st.tbl_off[u] = noff;
      st.tbl_cap[u] = pow2ceil(nbn);
    }
    uint64_t bb = bb0 + (pl >> 32);
    for (uint32_t j = nbo; j < nbn; ++j) st.arena[tbl + j] = st.free_ids[bb + (j - nbo)];
    st.used[u] = uo + surv;
    st.deg[u] += surv;
    a.seg_tbl[sidx] = tbl;
    a.seg_rb[sidx] = rb;
  }
}

struct WriteArgs {
  const uint64_t *keys;
  const uint32_t *flag, *segof, *rank, *seg_used, *seg_tbl, *seg_rb;
};

// One thread per staged key (the grid covers them): three dependent load
// levels -- the key's own words, its segment's words, the block table --
// then the slot.  The survivors of a source that was empty form its sorted
// prefix (k_ins_tables set it), so only they write fences.
__global__ void k_ins_write(Store st, WriteArgs a, uint32_t bV, Ctl *ctl) {
  if (ctl->abort) return;
  const uint32_t nv = ctl->n_valid;
  const uint64_t mask = (1ull << bV) - 1;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (!a.flag[i]) continue;
    const uint32_t sidx = a.segof[i], r = a.rank[i];
    const uint64_t k = a.keys[i];
    const uint32_t uo = a.seg_used[sidx];
    const uint32_t pos = uo + (r - a.seg_rb[sidx]);
    const uint32_t idx = slot_index(st.arena, a.seg_tbl[sidx], pos, st.log2B, st.B);
    const uint32_t v = (uint32_t)(k & mask);
    st.slots[idx] = v;
    if ((pos & (FENCE_SLOTS - 1)) == 0 && uo == 0) st.fence[idx / FENCE_SLOTS] = v;
  }
}

void probe_present(Store &st, Scratch &sc, const uint64_t *keys, uint64_t m, hipStream_t s) {
  if (!m) return;
  const uint32_t bV = id_bits(st.V);
  ProbeArgs pa{keys, sc.item_seg, sc.item_start, sc.seg_src, sc.seg_used, sc.seg_start, sc.flag,
               nullptr, nullptr};
  k_probe<PROBE_INSERT><<<probe_grid(m), 256, 0, s>>>(st, pa, bV, st.ctl);
}

void insert_plan(Store &st, Scratch &sc, const uint64_t *keys, uint64_t m, hipStream_t s) {
  if (!m) return;
  probe_present(st, sc, keys, m, s);
  insert_alloc(st, sc, m, s);
}

void insert_alloc(Store &st, Scratch &sc, uint64_t m, hipStream_t s) {
  if (!m) return;
  scan_u32(sc.flag, sc.rank, m, &st.ctl->n_valid, sc.tmp, s);
  PlanArgs pl{sc.seg_src, sc.seg_used, sc.seg_start, sc.rank, sc.plan};
  k_ins_plan<<<grid_for(m), 256, 0, s>>>(st, pl, st.ctl);
  scan_u64(sc.plan, sc.plan, m, &st.ctl->n_segs, sc.tmp, s);
}

void commit_pair(Store &a, Scratch &sa, Store *b, Scratch *sb, hipStream_t s) {
  k_commit<<<1, 1, 0, s>>>(a.ctl, a.A, b ? b->ctl : nullptr, b ? b->A : 0, sa.rank,
                           sb ? sb->rank : nullptr, sa.plan, sb ? sb->plan : nullptr);
}

void insert_write(Store &st, Scratch &sc, const uint64_t *keys, uint64_t m, hipStream_t s) {
  if (!m) return;
  const uint32_t bV = id_bits(st.V);
  TblArgs ta{sc.seg_src, sc.seg_used, sc.seg_start, sc.rank, sc.plan, sc.seg_blk, sc.seg_cnt};
  k_ins_tables<<<grid_for(m), 256, 0, s>>>(st, ta, st.ctl);
  WriteArgs wa{keys, sc.flag, sc.segof, sc.rank, sc.seg_used, sc.seg_blk, sc.seg_cnt};
  k_ins_write<<<grid_for(m, 256, 1u << 20), 256, 0, s>>>(st, wa, bV, st.ctl);
}

// ---------------------------------------------------------------- delete --
__global__ void k_del_finish(Store st, const uint32_t *seg_src, const uint32_t *seg_del,
                             Ctl *ctl) {
  if (ctl->abort) return;
  const uint32_t T = ctl->n_segs;
  unsigned long long mine = 0;
  // four segments per thread and step, their loads issued together
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t s0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s0 < T; s0 += 4 * stride) {
    uint32_t d[4], u[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint64_t sidx = s0 + j * stride;
      d[j] = sidx < T ? seg_del[sidx] : 0u;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) u[j] = d[j] ? seg_src[s0 + j * stride] : 0u;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (d[j]) st.deg[u[j]] -= d[j];
      mine += d[j];
    }
  }
  // one atomic pair per block (the grid is capped at 256 blocks: the pairs
  // all land on the same two words)
  unsigned long long tot = block_sum<unsigned long long, 256>(mine);
  if (threadIdx.x == 0 && tot) {
    atomicAdd(&ctl->applied, tot);
    atomicAdd(&ctl->live_edges, 0ull - tot);
  }
}

void delete_apply(Store &st, Scratch &sc, const uint64_t *keys, uint64_t m, hipStream_t s) {
  if (!m) return;
  const uint32_t bV = id_bits(st.V);
  ProbeArgs pa{keys, sc.item_seg, sc.item_start, sc.seg_src, sc.seg_used, sc.seg_start, nullptr,
               sc.seg_blk, nullptr};
  k_probe<PROBE_DELETE><<<probe_grid(m), 256, 0, s>>>(st, pa, bV, st.ctl);
  k_del_finish<<<grid_for(m, 256, 256), 256, 0, s>>>(st, sc.seg_src, sc.seg_blk, st.ctl);
}

// Merged group: one probe pass tombstones keys whose last op is a delete and
// marks present keys whose last op is an insert; then the insert machinery
// places the remaining inserts.
void mixed_plan(Store &st, Scratch &sc, const uint64_t *keys, uint64_t m, hipStream_t s) {
  if (!m) return;
  const uint32_t bV = id_bits(st.V);
  ProbeArgs pa{keys, sc.item_seg, sc.item_start, sc.seg_src, sc.seg_used, sc.seg_start, sc.flag,
               sc.seg_blk, sc.op};
  k_probe<PROBE_MIXED><<<probe_grid(m), 256, 0, s>>>(st, pa, bV, st.ctl);
  k_del_finish<<<grid_for(m, 256, 256), 256, 0, s>>>(st, sc.seg_src, sc.seg_blk, st.ctl);
  insert_alloc(st, sc, m, s);
}

// ------------------------------------------------------------ compaction --
__device__ __forceinline__ bool cmp_wanted(const Store &st, uint32_t v, float frac) {
  uint32_t us = st.used[v], d = st.deg[v];
  uint32_t tomb = us - d;
  return tomb > 0 && (frac <= 0.f || (float)tomb > frac * (float)us);
}

// One wavefront packs vertex u in place: round r reads 64 slots, keeps the
// live ones at the running write position (never ahead of the read
// position), then the emptied tail blocks go back on the free stack.
// Intent of compactionVertexCentric (delete_kernels.cu:50-261) without its
// defects D5/D6.
__device__ __forceinline__ void cmp_vertex(const Store &st, uint32_t u, uint32_t lane, Ctl *ctl) {
  const uint32_t uo = st.used[u];
  const uint32_t tbl = st.tbl_off[u];
  // the live slots of the sorted prefix stay first and in order: they are
  // the new prefix, and its chunks get fences from their new first slots
  const uint32_t so = st.sorted[u];
  uint32_t wpos = 0, nso = 0;
  for (uint32_t base = 0; base < uo; base += WAVE) {
    uint32_t pos = base + lane;
    uint32_t w = TOMB;
    if (pos < uo) w = st.slots[slot_index(st.arena, tbl, pos, st.log2B, st.B)];
    bool keep = w != TOMB;
    unsigned long long m = __ballot(keep);
    if (keep) {
      uint32_t dp = wpos + mask_lt_count(m);
      const uint32_t didx = slot_index(st.arena, tbl, dp, st.log2B, st.B);
      st.slots[didx] = w;
      if (pos < so && (dp & (FENCE_SLOTS - 1)) == 0) st.fence[didx / FENCE_SLOTS] = w;
    }
    const unsigned long long below =
        so <= base ? 0ull : (so - base >= (uint32_t)WAVE ? ~0ull : (1ull << (so - base)) - 1);
    nso += (uint32_t)__popcll(m & below);
    wpos += (uint32_t)__popcll(m);
  }
  const uint32_t nbo = (uo + st.B - 1) >> st.log2B;
  const uint32_t nbn = (wpos + st.B - 1) >> st.log2B;
  const uint32_t nf = nbo - nbn;
  unsigned long long fb = 0;
  if (lane == 0 && nf) fb = atomicAdd(&ctl->free_top, (unsigned long long)nf);
  fb = __shfl(fb, 0, WAVE);
  for (uint32_t j = lane; j < nf; j += WAVE) st.free_ids[fb + j] = st.arena[tbl + nbn + j];
  if (lane == 0) {
    st.used[u] = wpos;
    st.sorted[u] = nso;
  }
}

// Whole-graph compaction in one launch: lane l of a wavefront checks vertex
// base + l (coalesced reads of used and deg), then the wavefront packs the
// vertices it selected one after another.  COUNT: their number is also added
// to ctl->compacted (reset by op_begin), one atomic per block.
template <bool COUNT>
__global__ __launch_bounds__(256) void k_cmp_all(Store st, float frac, Ctl *ctl) {
  const uint32_t lane = lane_id();
  const uint64_t W = (uint64_t)gridDim.x * blockDim.x / WAVE;
  const uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
  uint32_t mine = 0;
  for (uint64_t base = w * WAVE; base < st.V; base += W * WAVE) {
    const uint64_t v = base + lane;
    const bool sel = v < st.V && cmp_wanted(st, (uint32_t)v, frac);
    unsigned long long m = __ballot(sel);
    mine += (uint32_t)__popcll(m);
    while (m) {
      const int j = __ffsll(m) - 1;
      m &= m - 1;
      cmp_vertex(st, (uint32_t)(base + j), lane, ctl);
    }
  }
  if (COUNT) {
    const uint32_t tot = block_sum<uint32_t, 256>(lane == 0 ? mine : 0u);
    if (threadIdx.x == 0 && tot) atomicAdd(&ctl->compacted, (unsigned long long)tot);
  }
}

// One wavefront per listed vertex (vlist[0, ctl->compacted)).
__global__ __launch_bounds__(256) void k_cmp_run(Store st, const uint32_t *vlist, Ctl *ctl) {
  const uint64_t n = ctl->compacted;
  const uint32_t lane = lane_id();
  const uint64_t nwaves = (uint64_t)gridDim.x * blockDim.x / WAVE;
  for (uint64_t i = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE; i < n; i += nwaves)
    cmp_vertex(st, vlist[i], lane, ctl);
}

// Same selection over the sources of the last staged batch only (segments
// [0, ctl->n_segs) of sc): a vertex the batch did not touch gained no
// tombstones from it.
__global__ __launch_bounds__(256) void k_cmp_select_touched(Store st, float frac,
                                                            const uint32_t *seg_src,
                                                            uint32_t *vlist, Ctl *ctl) {
  const uint64_t ns = ctl->n_segs;
  const uint64_t chunk = cdiv_dev(ns, gridDim.x);
  const uint64_t lo = blockIdx.x * chunk;
  const uint64_t hi = lo + chunk < ns ? lo + chunk : ns;
  uint32_t mine = 0;
  for (uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x)
    mine += cmp_wanted(st, seg_src[i], frac);
  uint32_t tot = block_sum<uint32_t, 256>(mine);
  __shared__ unsigned long long base;
  if (threadIdx.x == 0) base = tot ? atomicAdd(&ctl->compacted, (unsigned long long)tot) : 0ull;
  __syncthreads();
  if (!tot) return;
  unsigned long long b = base;
  for (uint64_t t0 = lo; t0 < hi; t0 += blockDim.x) {
    uint64_t i = t0 + threadIdx.x;
    uint32_t v = i < hi ? seg_src[i] : 0u;
    uint32_t sel = i < hi ? cmp_wanted(st, v, frac) : 0u;
    uint32_t ttot;
    uint32_t ex = block_excl_scan<uint32_t, 256>(sel, ttot);
    if (sel) vlist[b + ex] = v;
    b += ttot;
  }
}

void compact(Store &st, float frac, bool count, hipStream_t s) {
  const unsigned grid = grid_for(st.V, 256 * 8, 2048);
  if (count) {
    op_begin(st, s);
    k_cmp_all<true><<<grid, 256, 0, s>>>(st, frac, st.ctl);
  } else {
    k_cmp_all<false><<<grid, 256, 0, s>>>(st, frac, st.ctl);
  }
}

void compact_touched(Store &st, const Scratch &sc, uint64_t m, float frac, uint32_t *vlist,
                     hipStream_t s) {
  k_cmp_select_touched<<<grid_for(m, 256, 2048), 256, 0, s>>>(st, frac, sc.seg_src, vlist,
                                                              st.ctl);
  k_cmp_run<<<probe_grid(m), 256, 0, s>>>(st, vlist, st.ctl);
}

// ------------------------------------------------------------- read-back --
__global__ void k_deg64(const uint32_t *deg, uint32_t V, uint64_t *out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < V;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = deg[i];
}
void degrees_u64(const Store &st, uint64_t *out, hipStream_t s) {
  k_deg64<<<grid_for(st.V), 256, 0, s>>>(st.deg, st.V, out);
}

// Live slots of each vertex, as keys (tag << bV | dst), packed at offs[tag].
__global__ __launch_bounds__(256) void k_emit_keys(Store st, const uint32_t *vs, uint32_t n,
                                                   const uint64_t *offs, uint32_t bV,
                                                   uint64_t *keys) {
  const uint32_t lane = lane_id();
  const uint64_t nwaves = (uint64_t)gridDim.x * blockDim.x / WAVE;
  for (uint64_t i = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE; i < n; i += nwaves) {
    const uint32_t u = vs ? vs[i] : (uint32_t)i;
    if (u >= st.V) continue;
    const uint32_t uo = st.used[u];
    const uint32_t tbl = st.tbl_off[u];
    const uint64_t o = offs[i], cap = offs[i + 1] - offs[i];
    const uint64_t tag = (uint64_t)i << bV;
    uint32_t wpos = 0;
    for (uint32_t base = 0; base < uo; base += WAVE) {
      uint32_t pos = base + lane;
      uint32_t w = TOMB;
      if (pos < uo) w = st.slots[slot_index(st.arena, tbl, pos, st.log2B, st.B)];
      bool keep = w != TOMB;
      unsigned long long m = __ballot(keep);
      uint32_t r = wpos + mask_lt_count(m);
      if (keep && r < cap) keys[o + r] = tag | w;
      wpos += (uint32_t)__popcll(m);
    }
  }
}

// ------------------------------------------------ balanced key emission --
// The live slots of query i (vertex vs[i], or i) as keys (i << bV | dst) in
// [offs[i], offs[i + 1]).  Every query's slots are cut into items of EMIT_CH
// slots, dealt to lanes (wave w of W, step r, lane l takes item
// (r * 64 + l) * W + w), so a hub's items spread over many wavefronts; an
// item of at most EMIT_LANE slots is written by its lane, a longer one by
// the whole wavefront.  An item of a query without tombstones writes at
// fixed positions (output position = slot position); a query with
// tombstones takes each item's range from a per-query cursor after counting
// its live slots.  Callers sort the keys, so their order within a query's
// range does not matter.
constexpr uint32_t EMIT_CH = 1024;
constexpr uint32_t EMIT_LANE = 64;

__global__ void k_emit_items(Store st, const uint32_t *vs, uint32_t n, uint32_t *icnt,
                             uint32_t *cur) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t u = vs ? vs[i] : (uint32_t)i;
    const uint32_t us = u < st.V ? st.used[u] : 0u;
    icnt[i] = (us + EMIT_CH - 1) / EMIT_CH;
    cur[i] = 0;
  }
}

__global__ void k_emit_expand(const uint32_t *istart, uint32_t n, uint32_t *iq) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    for (uint32_t j = istart[i]; j < istart[i + 1]; ++j) iq[j] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void k_emit_chunks(Store st, const uint32_t *vs, uint32_t n,
                                                     const uint32_t *istart, const uint32_t *iq,
                                                     const uint64_t *offs, uint32_t bV,
                                                     uint32_t *cur, uint64_t *keys) {
  const uint32_t ni = istart[n];
  const uint32_t lane = lane_id();
  const uint64_t W = (uint64_t)gridDim.x * blockDim.x / WAVE;
  const uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
  for (uint64_t r = 0; r * WAVE * W + w < ni; ++r) {
    const uint64_t it = (r * WAVE + lane) * W + w;
    uint32_t lo = 0, hi = 0, tbl = 0, q = 0;
    uint64_t dst = 0;
    bool tomb = false;
    if (it < ni) {
      q = iq[it];
      const uint32_t u = vs ? vs[q] : q;
      const uint32_t us = st.used[u];
      tomb = us != st.deg[u];
      tbl = st.tbl_off[u];
      lo = ((uint32_t)it - istart[q]) * EMIT_CH;
      hi = lo + EMIT_CH < us ? lo + EMIT_CH : us;
      dst = offs[q] + lo;
    }
    if (it < ni && hi - lo <= EMIT_LANE) {
      const uint64_t tag = (uint64_t)q << bV;
      if (tomb) {
        uint32_t cnt = 0;
        for (uint32_t p = lo; p < hi; ++p)
          cnt += st.slots[slot_index(st.arena, tbl, p, st.log2B, st.B)] != TOMB ? 1u : 0u;
        dst = offs[q] + (cnt ? atomicAdd(&cur[q], cnt) : 0u);
      }
      for (uint32_t p = lo; p < hi; p += 4) {
        uint32_t x[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k)
          x[k] = p + k < hi ? st.slots[slot_index(st.arena, tbl, p + k, st.log2B, st.B)] : TOMB;
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
          if (x[k] == TOMB) continue;
          keys[dst++] = tag | x[k];
        }
      }
    }
    for (unsigned long long m = __ballot(it < ni && hi - lo > EMIT_LANE); m; m &= m - 1) {
      const int j = __ffsll(m) - 1;
      const uint32_t jlo = __builtin_amdgcn_readlane(lo, j), jhi = __builtin_amdgcn_readlane(hi, j);
      const uint32_t jtbl = __builtin_amdgcn_readlane(tbl, j);
      const uint32_t jq = __builtin_amdgcn_readlane(q, j);
      const bool jtomb = __builtin_amdgcn_readlane(tomb ? 1u : 0u, j) != 0u;
      uint64_t jdst = ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(dst >> 32), j) << 32) |
                      __builtin_amdgcn_readlane((uint32_t)dst, j);
      const uint64_t tag = (uint64_t)jq << bV;
      if (jtomb) {
        uint32_t cnt = 0;
        for (uint32_t p = jlo; p < jhi; p += WAVE) {
          const bool live = p + lane < jhi &&
                            st.slots[slot_index(st.arena, jtbl, p + lane, st.log2B, st.B)] != TOMB;
          cnt += (uint32_t)__popcll(__ballot(live));
        }
        uint32_t base = 0;
        if (lane == 0 && cnt) base = atomicAdd(&cur[jq], cnt);
        jdst = offs[jq] + __shfl(base, 0, WAVE);
      }
      for (uint32_t p = jlo; p < jhi; p += WAVE) {
        uint32_t x = TOMB;
        if (p + lane < jhi) x = st.slots[slot_index(st.arena, jtbl, p + lane, st.log2B, st.B)];
        const bool live = x != TOMB;
        const unsigned long long lm = __ballot(live);
        if (live) keys[jdst + mask_lt_count(lm)] = tag | x;
        jdst += (uint64_t)__popcll(lm);
      }
    }
  }
}

void emit_items(const Store &st, const uint32_t *vs, uint32_t n, uint32_t *istart,
                uint32_t *cur, void *tmp, hipStream_t s) {
  if (!n) return;
  k_emit_items<<<grid_for(n), 256, 0, s>>>(st, vs, n, istart, cur);
  scan_u32(istart, istart, n, nullptr, tmp, s);
}

void emit_keys(const Store &st, const uint32_t *vs, uint32_t n, const uint32_t *istart,
               uint32_t *iq, uint64_t items, const uint64_t *offs, uint32_t *cur, uint64_t *keys,
               hipStream_t s) {
  if (!n || !items) return;
  k_emit_expand<<<grid_for(n), 256, 0, s>>>(istart, n, iq);
  k_emit_chunks<<<grid_for(items * WAVE, 256, 2048), 256, 0, s>>>(st, vs, n, istart, iq, offs,
                                                                  id_bits(st.V), cur, keys);
}

__global__ void k_keys_to_cols(const uint64_t *keys, uint64_t n, uint64_t mask, uint32_t *cols) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    cols[i] = (uint32_t)(keys[i] & mask);
}
void keys_to_cols(const uint64_t *keys, uint64_t n, uint32_t bV, uint32_t *cols, hipStream_t s) {
  if (n) k_keys_to_cols<<<grid_for(n), 256, 0, s>>>(keys, n, (1ull << bV) - 1, cols);
}

__global__ void k_nbr_counts(const uint32_t *deg, uint32_t V, const uint32_t *vs, uint32_t n,
                             uint64_t *cnt) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t u = vs[i];
    cnt[i] = u < V ? deg[u] : 0;
  }
}
void neighbor_counts(const Store &st, const uint32_t *vs, uint32_t n, uint64_t *cnt,
                     hipStream_t s) {
  if (n) k_nbr_counts<<<grid_for(n), 256, 0, s>>>(st.deg, st.V, vs, n, cnt);
}
void neighbor_keys(const Store &st, const uint32_t *vs, uint32_t n, const uint64_t *offs,
                   uint32_t qbits, uint64_t *keys, hipStream_t s) {
  (void)qbits;
  if (n) k_emit_keys<<<probe_grid(n), 256, 0, s>>>(st, vs, n, offs, id_bits(st.V), keys);
}

// ------------------------------------------------------------ generators --
__device__ __forceinline__ uint64_t fmix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t rand64(uint64_t key, uint64_t ctr) {
  return fmix64(key + (ctr + 1) * 0x9E3779B97F4A7C15ull);
}
__device__ __forceinline__ uint32_t mulhi32(uint32_t r, uint32_t n) {
  return (uint32_t)(((uint64_t)r * n) >> 32);
}

__global__ void k_gen_rmat(uint64_t key, int scale, uint64_t n, uint32_t ta, uint32_t tab,
                           uint32_t tabc, uint32_t *src, uint32_t *dst) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t u = 0, v = 0;
    for (int l = 0; l < scale; ++l) {
      uint32_t r = (uint32_t)(rand64(key, e * (uint64_t)scale + (uint64_t)l) >> 32);
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

__global__ void k_gen_uniform(uint64_t key, uint32_t V, uint64_t n, uint32_t *src, uint32_t *dst) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t r = rand64(key, i);
    uint32_t u = mulhi32((uint32_t)(r >> 32), V);
    uint32_t v = mulhi32((uint32_t)r, V);
    if (u == v && V > 1) v = (v + 1 == V) ? 0 : v + 1;
    src[i] = u;
    dst[i] = v;
  }
}

__global__ void k_gen_sample(uint64_t key, uint32_t V, const uint64_t *offsets,
                             const uint32_t *cols, uint64_t n, uint32_t *src, uint32_t *dst) {
  const uint64_t E = offsets[V];
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (E == 0) {
      src[i] = dst[i] = 0;
      continue;
    }
    uint64_t idx = rand64(key, i) % E;
    uint32_t lo = 0, hi = V;
    while (hi - lo > 1) {
      uint32_t mid = lo + (hi - lo) / 2;
      if (offsets[mid] <= idx) lo = mid; else hi = mid;
    }
    src[i] = lo;
    dst[i] = cols[idx];
  }
}

// Endpoint draws against a cumulative weight table cum[0..V] (cum[0] = 0,
// cum[V] = 2^62): draw i picks the largest vertex lo with cum[lo] <= r_i >> 2.
// Edge e takes draw 2e as its source and draw 2e+1 as its destination.
__global__ void k_gen_table(uint64_t key, const uint64_t *cum, uint32_t V, uint64_t n,
                            uint32_t *src, uint32_t *dst) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t out[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const uint64_t r = rand64(key, 2 * e + k) >> 2;
      uint32_t lo = 0, hi = V;
      while (hi - lo > 1) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (cum[mid] <= r) lo = mid; else hi = mid;
      }
      out[k] = lo;
    }
    src[e] = out[0];
    dst[e] = out[1];
  }
}

void gen_table(uint64_t key, const uint64_t *cum, uint32_t V, uint64_t n, uint32_t *src,
               uint32_t *dst, hipStream_t s) {
  if (n) k_gen_table<<<grid_for(n), 256, 0, s>>>(key, cum, V, n, src, dst);
}

void gen_rmat(uint64_t key, int scale, uint64_t n, uint32_t ta, uint32_t tab, uint32_t tabc,
              uint32_t *src, uint32_t *dst, hipStream_t s) {
  if (n) k_gen_rmat<<<grid_for(n, 256, 16384), 256, 0, s>>>(key, scale, n, ta, tab, tabc, src, dst);
}
void gen_uniform(uint64_t key, uint32_t V, uint64_t n, uint32_t *src, uint32_t *dst,
                 hipStream_t s) {
  if (n) k_gen_uniform<<<grid_for(n), 256, 0, s>>>(key, V, n, src, dst);
}
void gen_sample(uint64_t key, uint32_t V, const uint64_t *offsets, const uint32_t *cols,
                uint64_t n, uint32_t *src, uint32_t *dst, hipStream_t s) {
  if (n) k_gen_sample<<<grid_for(n), 256, 0, s>>>(key, V, offsets, cols, n, src, dst);
}

}  // namespace gv
