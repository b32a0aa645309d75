// gv.hpp -- C++ facade over the C ABI (gv.h) that restores the reference's
// names, so a caller of GraphVine's host API can switch by changing includes:
//   class GraphVine            include/dynamicStructure/graphvine.cuh:29-105
//   static_pagerank(...)       include/algorithm/pagerank/pagerank.cuh:11
//   static_tc(...)             include/algorithm/tc/tc.cuh:10
//   dynamic_tc(...)            src/algorithm/tc/tc.cu:29-64
//   bc(GraphVine*, ...)        include/algorithm/bc/bc.cuh:8
// Differences: vertex IDs are 0-based, batches are COO host vectors, results
// are returned in host vectors, and failures throw gv::Error (a
// std::runtime_error carrying the gv_status) instead of exiting.
#ifndef GV_HPP
#define GV_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "gv.h"

namespace gv {

class Error : public std::runtime_error {
 public:
  Error(gv_status s, const std::string &what) : std::runtime_error(what), status(s) {}
  gv_status status;
};

inline void check(gv_status s) {
  if (s != GV_OK) throw Error(s, gv_last_error());
}

class GraphVine {
 public:
  // GraphVine(CSR*): create and bulk-build from a COO edge list.
  GraphVine(uint32_t num_vertices, const std::vector<uint32_t> &src,
            const std::vector<uint32_t> &dst, bool directed = false,
            const gv_config *cfg = nullptr)
      : V_(num_vertices) {
    check(gv_create(cfg, num_vertices, directed ? 1 : 0, &g_));
    bulkBuild(src, dst);
  }
  ~GraphVine() { gv_destroy(g_); }
  GraphVine(const GraphVine &) = delete;
  GraphVine &operator=(const GraphVine &) = delete;

  gv_batch_stats bulkBuild(const std::vector<uint32_t> &src, const std::vector<uint32_t> &dst) {
    gv_batch_stats st;
    check(gv_bulk_build(g_, src.data(), dst.data(), src.size(), GV_MEM_HOST, &st));
    return st;
  }
  gv_batch_stats batchInsert(const std::vector<uint32_t> &src, const std::vector<uint32_t> &dst) {
    gv_batch_stats st;
    check(gv_batch_insert(g_, src.data(), dst.data(), src.size(), GV_MEM_HOST, &st));
    return st;
  }
  gv_batch_stats batchDelete(const std::vector<uint32_t> &src, const std::vector<uint32_t> &dst) {
    gv_batch_stats st;
    check(gv_batch_delete(g_, src.data(), dst.data(), src.size(), GV_MEM_HOST, &st));
    return st;
  }
  uint64_t compaction(float tombstone_fraction = 0.f) {
    uint64_t n = 0;
    check(gv_compact(g_, tombstone_fraction, &n));
    return n;
  }
  unsigned long getVertexSize() const { return V_; }
  void exportCSR(std::vector<uint64_t> &offsets, std::vector<uint32_t> &cols) {
    offsets.resize((size_t)V_ + 1);
    check(gv_export_csr(g_, offsets.data(), nullptr, GV_MEM_HOST));
    cols.resize(offsets.back());
    check(gv_export_csr(g_, offsets.data(), cols.data(), GV_MEM_HOST));
  }
  gv_graph *handle() { return g_; }

 private:
  gv_graph *g_ = nullptr;
  uint32_t V_;
};

// static_pagerank: result in x (FP32), returns the number of updates.
inline int static_pagerank(GraphVine *g, float alpha, float epsilon, int max_iter,
                           std::vector<float> &x) {
  x.resize(g->getVertexSize());
  int iters = 0;
  double res = 0;
  check(gv_pagerank(g->handle(), alpha, epsilon, max_iter, x.data(), GV_MEM_HOST, &iters, &res));
  return iters;
}

// static_tc: exact per-vertex counts (min-ID attribution); returns the total.
inline uint64_t static_tc(GraphVine *g, std::vector<uint64_t> &per_vertex) {
  per_vertex.resize(g->getVertexSize());
  uint64_t total = 0;
  check(gv_triangle_count(g->handle(), per_vertex.data(), GV_MEM_HOST, &total));
  return total;
}

// dynamic_tc (tc.cu:29-64): applies one insert or delete batch and updates
// counts taken by static_tc before it (per_vertex and total) to the counts
// after it.  The batch is explicit here; the reference reads the affected
// set its last batchInsert left behind.
inline void dynamic_tc(GraphVine *g, bool insert, const std::vector<uint32_t> &src,
                       const std::vector<uint32_t> &dst, std::vector<uint64_t> &per_vertex,
                       uint64_t &total) {
  if (src.size() != dst.size()) throw Error(GV_E_INVALID, "src and dst sizes differ");
  per_vertex.resize(g->getVertexSize());
  check(gv_triangle_count_update(g->handle(), insert ? GV_OP_INSERT : GV_OP_DELETE, src.data(),
                                 dst.data(), src.size(), GV_MEM_HOST, per_vertex.data(),
                                 GV_MEM_HOST, &total, nullptr));
}

// bc: Brandes from the given sources (the reference ignores its source
// argument and runs all vertices; pass every vertex ID to reproduce that).
inline std::vector<float> bc(GraphVine *g, const std::vector<uint32_t> &sources) {
  std::vector<float> out(g->getVertexSize());
  check(gv_bc(g->handle(), sources.data(), (uint32_t)sources.size(), out.data(), GV_MEM_HOST));
  return out;
}

}  // namespace gv

#endif
