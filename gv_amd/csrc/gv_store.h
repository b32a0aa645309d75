// gv_store.h -- host-side entry points of the store kernels (gv_store.hip).
#pragma once
#include "gv_internal.h"

namespace gv {

constexpr uint32_t PROBE_CHUNK = 256;  // slots per membership/tombstone work item
constexpr uint32_t FENCE_SLOTS = 16;   // slots per fence chunk of a sorted prefix
constexpr uint32_t FENCE_ITEM = 64 * FENCE_SLOTS;  // prefix slots per fence item (a chunk per lane)

// Work items of a source with `used` slots and sorted prefix `sorted`: the
// prefix up to a multiple of PROBE_CHUNK (pf) in fence items of FENCE_ITEM
// slots, then [pf, used) in scan items of PROBE_CHUNK slots.  Never more
// than ceil(used / PROBE_CHUNK).
__host__ __device__ inline uint32_t probe_pf(uint32_t sorted) {
  return sorted / PROBE_CHUNK * PROBE_CHUNK;
}
__host__ __device__ inline uint32_t probe_items(uint32_t used, uint32_t sorted) {
  const uint32_t pf = probe_pf(sorted);
  return (pf + FENCE_ITEM - 1) / FENCE_ITEM + (used - pf + PROBE_CHUNK - 1) / PROBE_CHUNK;
}

// Operation of a key in a merged group of batches.
constexpr uint32_t OP_INSERT = 0;
constexpr uint32_t OP_DELETE = 1;

// Abort codes kept in Ctl::abort (magnitudes of the GV_E_* values).
constexpr uint32_t ABORT_RANGE = 2;
constexpr uint32_t ABORT_POOL = 4;
constexpr uint32_t ABORT_GROUP = 6;  // group offsets or kinds invalid (device-filled group)

gv_status store_alloc(Store &st, uint32_t V, uint32_t B, uint64_t P, uint64_t A);
void store_release(Store &st);
void store_reset(Store &st, hipStream_t s);
gv_status scratch_reserve(Scratch &sc, const Store &st, uint64_t m);
// True when scratch_reserve(sc, st, m) would not reallocate.
bool scratch_fits(const Scratch &sc, const Store &st, uint64_t m);
void scratch_release(Scratch &sc);

// Per-op counter reset (keeps the sticky abort code).
void op_begin(Store &st, hipStream_t s);
// Fold the staging sorts' failure words into the stores' abort codes, then
// copy an abort code set in either store into the other.
void abort_join(Store &a, Store &b, const Scratch &sa, const Scratch &sb, hipStream_t s);

// Start an operation (the per-op counters reset, as op_begin) and stage n
// (src,dst) pairs as sorted keys in sc; mirror adds (dst,src) too, swap
// stores (dst,src) instead.  Returns the sorted key buffer; m = keys.  A
// failure of the sort is left in sc.sort_err for segment (or abort_join) to
// fold into the abort code.
uint64_t *stage(Store &st, Scratch &sc, const uint32_t *src, const uint32_t *dst, uint64_t n,
                int mirror, int swap, uint64_t &m, hipStream_t s);
// Start an operation (as op_begin) and stage a group of nb batches (entries
// [boff[b], boff[b+1]) of batch b, op
// kinds[b]; both device arrays) as unique sorted keys with their final op in
// sc.op, insert flags in sc.flag and their source segments built (as
// segment_marked leaves them; segment_items then lists the probe items).
// Returns the key buffer (m keys, sentinel-padded).
uint64_t *stage_group(Store &st, Scratch &sc, const uint32_t *src, const uint32_t *dst,
                      uint64_t n, const uint64_t *boff, const uint32_t *kinds, uint32_t nb,
                      int mirror, int swap, uint64_t &m, hipStream_t s);
// Segment the staged keys and build the probe work list (op: per-key final
// operation of a merged group, or null).
void segment(Store &st, Scratch &sc, const uint64_t *keys, uint64_t m, hipStream_t s,
             const uint32_t *op = nullptr);
// The same after stage_group, which has already written the keys' segment
// marks (sc.segof heads, sc.flag insert candidates, n_valid).
void segment_marked(Store &st, Scratch &sc, const uint64_t *keys, uint64_t m, hipStream_t s);
// The probe work list of segments already built (n_segs, seg_cnt).
void segment_items(Store &st, Scratch &sc, uint64_t m, hipStream_t s);
// Merged group: probe (tombstone / mark present), fix degrees, plan inserts.
void mixed_plan(Store &st, Scratch &sc, const uint64_t *keys, uint64_t m, hipStream_t s);
// Survivor scan + per-segment block/table plan (after an insert probe).
void insert_alloc(Store &st, Scratch &sc, uint64_t m, hipStream_t s);
// Membership: clears sc.flag of the staged keys already in the store.
void probe_present(Store &st, Scratch &sc, const uint64_t *keys, uint64_t m, hipStream_t s);
// Insert: probe existing slots, plan block/table allocation.
void insert_plan(Store &st, Scratch &sc, const uint64_t *keys, uint64_t m, hipStream_t s);
// Joint all-or-nothing pool check over one or two stores (b may be null).
void commit_pair(Store &a, Scratch &sa, Store *b, Scratch *sb, hipStream_t s);
void insert_write(Store &st, Scratch &sc, const uint64_t *keys, uint64_t m, hipStream_t s);
// Delete: tombstone matches and fix degrees.
void delete_apply(Store &st, Scratch &sc, const uint64_t *keys, uint64_t m, hipStream_t s);
// Compaction of vertices whose tombstones exceed frac of their slots; count:
// also leave their number in ctl->compacted.
void compact(Store &st, float frac, bool count, hipStream_t s);
// Compaction restricted to the sources of the last staged batch of sc (m =
// staged key bound), for pipelines that compact what they just updated.
void compact_touched(Store &st, const Scratch &sc, uint64_t m, float frac, uint32_t *vlist,
                     hipStream_t s);

// Read-back helpers.
void degrees_u64(const Store &st, uint64_t *out, hipStream_t s);
// Balanced key emission for n queries (vertices vs[i], or i when vs is
// null): emit_items writes the per-query item counts into istart[0, n)
// (n + 1 words), scans them (istart[n] = items, read back by the caller to
// size iq) and clears cur[0, n); emit_keys then writes query i's live slots
// as keys (i << bV | dst) into [offs[i], offs[i + 1]), in no fixed order.
void emit_items(const Store &st, const uint32_t *vs, uint32_t n, uint32_t *istart,
                uint32_t *cur, void *tmp, hipStream_t s);
void emit_keys(const Store &st, const uint32_t *vs, uint32_t n, const uint32_t *istart,
               uint32_t *iq, uint64_t items, const uint64_t *offs, uint32_t *cur, uint64_t *keys,
               hipStream_t s);
void keys_to_cols(const uint64_t *keys, uint64_t n, uint32_t bV, uint32_t *cols, hipStream_t s);
void neighbor_counts(const Store &st, const uint32_t *vs, uint32_t n, uint64_t *cnt,
                     hipStream_t s);
void neighbor_keys(const Store &st, const uint32_t *vs, uint32_t n, const uint64_t *offs,
                   uint32_t qbits, uint64_t *keys, hipStream_t s);

// Generators.
void gen_rmat(uint64_t key, int scale, uint64_t n, uint32_t ta, uint32_t tab, uint32_t tabc,
              uint32_t *src, uint32_t *dst, hipStream_t s);
void gen_uniform(uint64_t key, uint32_t V, uint64_t n, uint32_t *src, uint32_t *dst,
                 hipStream_t s);
void gen_sample(uint64_t key, uint32_t V, const uint64_t *offsets, const uint32_t *cols,
                uint64_t n, uint32_t *src, uint32_t *dst, hipStream_t s);
// Endpoints drawn against a device cumulative table cum[0..V] scaled to 2^62.
void gen_table(uint64_t key, const uint64_t *cum, uint32_t V, uint64_t n, uint32_t *src,
               uint32_t *dst, hipStream_t s);

void fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t s);
void fill_u64(uint64_t *p, uint64_t v, uint64_t n, hipStream_t s);
void fill_f32(float *p, float v, uint64_t n, hipStream_t s);

inline unsigned grid_for(uint64_t n, unsigned threads = 256, unsigned cap = 4096) {
  uint64_t g = cdiv(n ? n : 1, threads);
  return (unsigned)(g < cap ? g : cap);
}

}  // namespace gv
