// gv_internal.h -- device data model and shared primitives of the MI355X
// GraphVine store.  Not part of the public ABI (include/gv.h is).
//
// Layout in HBM (one Store per adjacency direction):
//   slots[P*B]      edge-block pool; block b owns slots [b*B, (b+1)*B), B in
//                   {32,64,128} uint32 destination IDs (128/256/512 B, i.e.
//                   whole 128 B lines; B=64 is one dword per lane of a wave).
//   free_ids[P]     free-block stack; entries [0, free_top) are free block IDs.
//   deg[V]          live out-degree |N(v)|.
//   used[V]         slots in use (live + tombstones); slots [0, used) valid.
//   tbl_off[V]      offset of v's block table in arena[].
//   tbl_cap[V]      capacity of that table (power of two, 0 = none).
//   sorted[V]       slots [0, sorted[v]) of v hold ascending values, tombstones
//                   aside (the sorted prefix: the first insert run into an
//                   empty vertex, e.g. a bulk build, less later tombstones;
//                   later runs are appended after it, compaction keeps its
//                   order).
//   fence[P*B/16]   fence[idx >> 4], for a slot idx at a position that is a
//                   multiple of 16 below sorted[v]: the value that slot held
//                   when the prefix was formed.  Prefix chunk c of 16 slots
//                   then holds only values in [fence of c, fence of c + 1),
//                   so a key can only be in one chunk of the prefix.
//   arena[A]        block tables: arena[tbl_off[v] + i] = i-th block of v in
//                   level order.  The reference's complete binary tree of
//                   blocks (children 2i+1, 2i+2; device_functions.cu:42-100)
//                   is kept as this level order, but block i is addressed in
//                   O(1) by index instead of a root-to-node pointer chase.
// Slot j of v lives at slots[arena[tbl_off[v] + (j >> log2B)] * B + (j & (B-1))].
// A tombstone is TOMB (0xFFFFFFFF); IDs must be < 0xFFFFFFFF.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/gv.h"

namespace gv {

constexpr uint32_t TOMB = 0xFFFFFFFFu;
constexpr uint64_t KEY_SENT = ~0ull;
constexpr int WAVE = 64;

// Device-resident control block of one Store.  Counters are 64-bit so the
// pool and arena can exceed 2^32 slots.
struct Ctl {
  unsigned long long free_top;    // number of free block IDs
  unsigned long long arena_top;   // bump pointer into arena
  unsigned long long live_edges;  // directed live edges
  // per-operation scratch counters (reset by k_op_begin)
  unsigned long long need_blocks;
  unsigned long long need_arena;
  unsigned long long blk_base;    // absolute index into free_ids of this op's blocks
  unsigned long long arena_base;  // absolute arena offset of this op's tables
  unsigned long long applied;
  unsigned long long blocks_freed;  // unused (kept for the layout)
  unsigned long long compacted;
  unsigned int n_valid;           // valid (non-sentinel) keys of the staged batch
  unsigned int n_segs;            // source segments of the staged batch
  unsigned int n_items;           // work items of the membership/tombstone pass
  unsigned int abort;             // sticky error code (positive GV error magnitude)
  unsigned int pad[4];
};

struct Store {
  uint32_t V = 0, B = 64, log2B = 6;
  uint64_t P = 0;
  uint64_t A = 0;  // arena entries
  uint32_t *slots = nullptr;
  uint32_t *free_ids = nullptr;
  uint32_t *deg = nullptr, *used = nullptr, *tbl_off = nullptr, *tbl_cap = nullptr;
  uint32_t *sorted = nullptr, *fence = nullptr;
  uint32_t *arena = nullptr;
  Ctl *ctl = nullptr;
};

// Batch staging scratch, sized for `cap` directed entries.
struct Scratch {
  uint64_t cap = 0;
  uint64_t item_cap = 0;
  uint64_t *k0 = nullptr, *k1 = nullptr;  // radix sort ping-pong
  uint32_t *flag = nullptr;               // per key: 1 = survives
  uint32_t *segof = nullptr;              // per key: inclusive head count (= seg + 1)
  uint32_t *rank = nullptr;               // per key: exclusive survivor prefix (cap+1)
  uint32_t *seg_start = nullptr;          // per segment: first key (cap+1)
  uint32_t *seg_src = nullptr;            // per segment: source vertex
  uint32_t *seg_used = nullptr;           // per segment: used slots before the op
  uint32_t *seg_cnt = nullptr;            // per segment: items, then blocks/deletes
  uint32_t *item_start = nullptr;         // per segment: first item (cap+1)
  uint32_t *item_seg = nullptr;           // per item: segment
  uint32_t *seg_blk = nullptr;            // per segment: tombstones written (delete)
  uint64_t *plan = nullptr;               // per segment: (new blocks << 32 | table cap), scanned
  uint32_t *op = nullptr;                 // per key: final op of a merged group
  const uint32_t *sort_err = nullptr;     // failure word of the last staging sort, if not folded
  void *tmp = nullptr;                    // scan/sort partials
  size_t tmp_bytes = 0;
};

// ------------------------------------------------------------ host errors --
void set_error(const std::string &msg);
gv_status hip_fail(hipError_t e, const char *what);

#define GV_HIP(call)                                              \
  do {                                                            \
    hipError_t _e = (call);                                       \
    if (_e != hipSuccess) return ::gv::hip_fail(_e, #call);       \
  } while (0)

inline uint32_t ceil_log2(uint64_t x) {
  uint32_t r = 0;
  while ((1ull << r) < x) ++r;
  return r;
}
inline uint32_t id_bits(uint32_t V) {
  uint32_t b = ceil_log2(V);
  return b ? b : 1;
}
inline uint64_t cdiv(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
__device__ __forceinline__ uint64_t cdiv_dev(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// ------------------------------------------------------- device helpers --
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// popcount of mask bits below this lane
__device__ __forceinline__ uint32_t mask_lt_count(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Inclusive prefix sum over all 64 lanes (every lane must be active).
template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  const uint32_t l = lane_id();
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    T o = __shfl_up(v, d, WAVE);
    if (l >= (uint32_t)d) v += o;
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int d = WAVE / 2; d >= 1; d >>= 1) v += __shfl_xor(v, d, WAVE);
  return v;
}

// Sum over the block (all NT threads must call it).
template <typename T, int NT>
__device__ __forceinline__ T block_sum(T v) {
  __shared__ T ws[NT / WAVE];
  v = wave_sum(v);
  if (lane_id() == 0) ws[threadIdx.x / WAVE] = v;
  __syncthreads();
  T t = 0;
#pragma unroll
  for (int i = 0; i < NT / WAVE; ++i) t += ws[i];
  __syncthreads();
  return t;
}

// Exclusive scan over the block; returns this thread's exclusive prefix and
// the block total (all NT threads must call it).
template <typename T, int NT>
__device__ __forceinline__ T block_excl_scan(T v, T &total) {
  __shared__ T wsum[NT / WAVE];
  T incl = wave_incl_scan<T>(v);
  const int w = threadIdx.x / WAVE;
  if (lane_id() == WAVE - 1) wsum[w] = incl;
  __syncthreads();
  T before = 0, all = 0;
#pragma unroll
  for (int i = 0; i < NT / WAVE; ++i) {
    T x = wsum[i];
    if (i < w) before += x;
    all += x;
  }
  __syncthreads();
  total = all;
  return before + incl - v;
}

// Lower bound of x in the ascending values held one per lane (sv, padded with
// 0xFFFFFFFF past the segment): the number of lanes whose sv is < x, capped
// at 63 (all 64 below x also gives 63).  Six uniform ds_bpermute steps, no
// divergence.
__device__ __forceinline__ uint32_t wave_lower_bound(uint32_t sv, uint32_t x) {
  uint32_t pos = 0;
#pragma unroll
  for (uint32_t step = WAVE / 2; step >= 1; step >>= 1) {
    uint32_t v = __shfl(sv, (int)(pos + step - 1), WAVE);
    if (v < x) pos += step;
  }
  return pos;
}

// Orders a wavefront's LDS writes before its later LDS reads (lanes of one
// wavefront exchanging data through LDS without a block barrier).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wave-aggregated atomic add: every lane of the wave calls it (inc may be 0);
// one atomic per wave; returns this lane's exclusive base.
__device__ __forceinline__ unsigned long long wave_atomic_add(unsigned long long *ctr,
                                                              uint32_t inc) {
  unsigned long long incl = wave_incl_scan<unsigned long long>(inc);
  unsigned long long total = __shfl(incl, WAVE - 1, WAVE);
  unsigned long long base = 0;
  if (lane_id() == WAVE - 1 && total) base = atomicAdd(ctr, total);
  base = __shfl(base, WAVE - 1, WAVE);
  return base + incl - inc;
}

__device__ __forceinline__ uint32_t slot_index(const uint32_t *arena, uint32_t tbl, uint32_t pos,
                                               uint32_t log2B, uint32_t B) {
  uint32_t blk = arena[tbl + (pos >> log2B)];
  return blk * B + (pos & (B - 1));
}

// TwoSum-compensated pair (value, compensation) add.
__device__ __forceinline__ void kahan_add(float &s, float &c, float x) {
  float y = x - c;
  float t = s + y;
  c = (t - s) - y;
  s = t;
}

// --------------------------------------------------------- primitives --
// (gv_prims.hip)
size_t scan_tmp_bytes(uint64_t n);
// Exclusive scan of in[0..n_eff) into out[0..n_eff], out[n_eff] = total, with
// n_eff = n_dev ? min(n, *n_dev) : n.  in may alias out.
void scan_u32(const uint32_t *in, uint32_t *out, uint64_t n, const uint32_t *n_dev, void *tmp,
              hipStream_t s);
void scan_u64(const uint64_t *in, uint64_t *out, uint64_t n, const uint32_t *n_dev, void *tmp,
              hipStream_t s);
// Workspace of any sort of at most n keys.
size_t sort_tmp_bytes(uint64_t n);
// Largest key count a sort takes (its histograms and offsets are 32-bit).
constexpr uint64_t SORT_MAX_KEYS = 0xFFFFFFFFull;
// Abort code a sort leaves in the caller's abort word when a tile of the
// single-sweep path could not find its offsets (the output is then invalid).
constexpr uint32_t ABORT_SORT = 5;
// LSD radix sort of n <= SORT_MAX_KEYS keys on their low `bits` bits;
// returns the buffer that holds the result (a or b).  abort (device word, may
// be null) receives ABORT_SORT if the sort failed.  cleared: the caller has
// already zeroed the words sort_clear_span names (the sort then skips that
// launch).
uint64_t *sort_keys(uint64_t *a, uint64_t *b, uint64_t n, uint32_t bits, void *tmp, hipStream_t s,
                    uint32_t *abort = nullptr, bool cleared = false);
// Same, ordering on bits [lo, hi) only (stable for keys equal there).
uint64_t *sort_keys_range(uint64_t *a, uint64_t *b, uint64_t n, uint32_t lo, uint32_t hi,
                          void *tmp, hipStream_t s, uint32_t *abort = nullptr,
                          bool cleared = false);
// The workspace words a sort of n keys on bits [lo, hi) needs zeroed before
// it starts (w[0, words)), and its failure word (nonzero after the sort if a
// tile gave up); both empty when the sort takes the path that needs neither.
struct SortClear {
  uint32_t *w = nullptr;
  uint64_t words = 0;
  const uint32_t *err = nullptr;
};
SortClear sort_clear_span(void *tmp, uint64_t n, uint32_t lo, uint32_t hi);
// After a sort of n keys on bits [lo, hi) without an abort word: waits for
// the stream and returns GV_E_HIP (error text set) if the sort failed.
gv_status sort_check(const void *tmp, uint64_t n, uint32_t lo, uint32_t hi, hipStream_t s);
// Bound on the look-back polls of the single-sweep sort (tests lower it to
// exercise the failure path; 0 makes every tile after the first give up).
void sort_set_spin_limit(uint32_t limit);

}  // namespace gv
