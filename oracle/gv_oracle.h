/*
 * gv_oracle.h -- host-sequential CPU restatement of GraphVine's dynamic graph
 * store and analytics.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness checker for the
 * HIP implementation in gv_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product path never does.
 *
 * Semantics (restated from the reference, whose literal outputs are defective
 * in places; each function cites the reference lines it follows):
 *   - vertex IDs are 0-based uint32, the vertex set is fixed at creation;
 *   - N(u) is a SET of out-neighbours: duplicates collapse, self-loops drop
 *     (insert_kernels.cu:158-180, device_remove_batch_duplicates);
 *   - undirected graphs apply every update (u,v) as (u,v) and (v,u)
 *     (marketReader.cu:58-64 symmetrises on read);
 *   - PageRank drops dangling mass and stops after the first update whose L1
 *     residual is below eps (pagerank.cu:17-59, pr_kernel.cu:21-45);
 *   - triangle count t[u] = #{(v,w): u<v<w, v in N(u), w in N(v), u in N(w)}
 *     (tc_kernel.cu:27-73);
 *   - betweenness centrality is Brandes over the directed adjacency for a
 *     given source list, unnormalised and not halved (intent of
 *     bc.cu:42-80, bc_kernel.cu:19-119).
 */
#ifndef GV_ORACLE_H
#define GV_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_graph orc_graph;

/* counter-based generators (SplitMix64, identical to the device generators) */
uint64_t orc_fmix64(uint64_t z);
uint64_t orc_stream_key(uint64_t seed, uint64_t stream);
uint64_t orc_rand64(uint64_t key, uint64_t ctr);
void orc_gen_rmat(uint64_t key, int scale, uint64_t n, uint32_t ta, uint32_t tab,
                  uint32_t tabc, uint32_t *src, uint32_t *dst);
void orc_gen_uniform(uint64_t key, uint32_t V, uint64_t n, uint32_t *src, uint32_t *dst);
void orc_gen_sample(uint64_t key, uint32_t V, const uint64_t *offsets, const uint32_t *cols,
                    uint64_t n, uint32_t *src, uint32_t *dst);
void orc_gen_chung_lu(uint64_t key, uint32_t V, double gamma, double avg_deg, double max_deg,
                      uint64_t n, uint32_t *src, uint32_t *dst);

/* graph store */
orc_graph *orc_create(uint32_t V, int directed);
void orc_destroy(orc_graph *g);
int orc_bulk_build(orc_graph *g, const uint32_t *src, const uint32_t *dst, uint64_t n,
                   uint64_t *inserted);
int orc_batch_insert(orc_graph *g, const uint32_t *src, const uint32_t *dst, uint64_t n,
                     uint64_t *inserted);
int orc_batch_delete(orc_graph *g, const uint32_t *src, const uint32_t *dst, uint64_t n,
                     uint64_t *deleted);
uint64_t orc_num_edges(const orc_graph *g);
uint32_t orc_num_vertices(const orc_graph *g);
void orc_degrees(const orc_graph *g, uint32_t *deg);
void orc_export_csr(const orc_graph *g, uint64_t *offsets, uint32_t *cols);

/* analytics */
/* x0: start vector (NULL = 1/V everywhere) */
int orc_pagerank(const orc_graph *g, double alpha, double eps, int max_iter, int fixed_iters,
                 const double *x0, double *x, double *residual);
int orc_pagerank_personalized(const orc_graph *g, double alpha, double eps, int max_iter,
                              int fixed_iters, const uint32_t *seeds, uint32_t ncols, double *x,
                              double *max_residual);
uint64_t orc_triangle_count(const orc_graph *g, uint64_t *per_vertex);
void orc_bc(const orc_graph *g, const uint32_t *sources, uint32_t n_sources, double *bc);

#ifdef __cplusplus
}
#endif
#endif
