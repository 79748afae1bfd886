/* include/g2host.h -- host-side C ABI of the MI355X subgraph-search library.
 *
 * libg2host.so holds everything that runs on the CPU around the device path:
 * CSR graph files, the synthetic workload generators, the degeneracy recode
 * and view-bin maps, query patterns with their symmetry-breaking conditions,
 * and the query plan.  No HIP types cross this boundary; all sizes are u64,
 * vertex ids u32, row offsets u64 (the reference's uintV/uintE,
 * common/meta.h:15-16).
 *
 * Return convention: 0 = OK, < 0 = error; g2h_last_error() describes the last
 * error of the calling thread.
 */
#ifndef G2HOST_H
#define G2HOST_H

#include <stdint.h>
#include "g2plan.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- CSR graph handle (replaces common/graph.h:57-107 `Graph` data side) ---- */
typedef struct g2h_graph g2h_graph;

/* .bin reader / writer, byte layout of common/graph.h:360-398 and :539-562:
 * u64 sizeof(uintV)=4, u64 sizeof(uintE)=8, u64 V, u64 E, u64 row_ptrs[V+1],
 * u32 cols[E]; no trailer. */
int  g2h_graph_read_bin(const char* path, g2h_graph** out);
int  g2h_graph_write_bin(const g2h_graph* g, const char* path);
/* wrap a copy of caller arrays (rows must be sorted) */
int  g2h_graph_from_csr(const uint64_t* row_ptrs, uint64_t nv, const uint32_t* cols, uint64_t ne,
                        g2h_graph** out);
/* symmetrise + sort + drop self loops / parallel edges from an edge list */
int  g2h_graph_from_edges(const uint32_t* src, const uint32_t* dst, uint64_t m, uint64_t nv,
                          g2h_graph** out);
uint64_t        g2h_graph_nv(const g2h_graph* g);
uint64_t        g2h_graph_ne(const g2h_graph* g);
const uint64_t* g2h_graph_row_ptrs(const g2h_graph* g);
const uint32_t* g2h_graph_cols(const g2h_graph* g);
uint64_t        g2h_graph_max_degree(const g2h_graph* g);
void            g2h_graph_free(g2h_graph* g);

/* Graph::Preprocess (common/graph.h:720-811): drop self loops, parallel
 * edges and degree-0 vertices, relabel densely in id order.  In place. */
int  g2h_graph_preprocess(g2h_graph* g);

/* kcore recode (kmeans/kcore.cpp:39-191): level-synchronous peeling, each
 * k-shell sorted by degree, vertex ids recoded into that order.  In place.
 * order_out (nullable, nv entries) receives old id -> new id.
 * shell_offsets_out (nullable, nv+2 entries) receives the shell boundaries in
 * the new order; *num_shells their count. */
int  g2h_graph_kcore_recode(g2h_graph* g, uint32_t* order_out, uint64_t* shell_offsets_out,
                            uint64_t* num_shells, uint32_t* degeneracy);

/* view-bin map of kmeans/kcore.cpp:193-229 ("kcore1" blocks of
 * floor(w/nb)+1 per shell) for a graph already recoded; bins[v] in [0,nb) */
int  g2h_kcore_bins(const uint64_t* shell_offsets, uint64_t num_shells, uint32_t nb,
                    uint32_t* bins, uint64_t nv);
/* .vbmap text file (view/view_bin_manager.h:42-46): "K b0 b1 ... b(V-1)" */
int  g2h_write_vbmap(const char* path, uint32_t nbins, const uint32_t* bins, uint64_t nv);
int  g2h_read_vbmap(const char* path, uint32_t* nbins, uint32_t* bins, uint64_t nv);

/* ---- offline input tools (the reference kmeans directory) ---- */
/* SNAP text edge list (common/graph.h:268-358): lines starting with a digit
 * hold "a b"; ids shifted by the smallest; both directions stored, rows
 * sorted; parallel edges and self loops kept (g2h_graph_preprocess drops them) */
int  g2h_graph_read_snap(const char* path, g2h_graph** out);
/* k-means view binning (kmeans/kmeans.cpp:53-795): cosine similarity of
 * adjacency rows scaled by 1/sqrt(degree), dense centroids, seeded
 * farthest-first initialisation, `iters` rounds; k2 > 1 clusters each of
 * the k1 clusters again (the reference's "k1xk2").  cluster[v] in
 * [0, *nclusters), empty clusters removed. */
int  g2h_kmeans(const uint64_t* row_ptrs, uint64_t nv, const uint32_t* cols, uint32_t k1, uint32_t k2,
                uint32_t iters, uint64_t seed, uint32_t* cluster, uint32_t* nclusters);
/* view packing (kmeans/view_pack.cpp:123-328, kmeans/view_bin.h): each
 * cluster's vertices (id order, or by decreasing degree) go to the open bin
 * whose hop closure grows least within max_entries row entries, pool_size
 * bins open at a time; bins then merged next-fit by size.  bins[v] in
 * [0, *nbins). */
int  g2h_view_pack(const uint64_t* row_ptrs, uint64_t nv, const uint32_t* cols, const uint32_t* cluster,
                   uint32_t nclusters, int hop, uint64_t max_entries, uint32_t pool_size, int by_degree,
                   uint32_t* bins, uint32_t* nbins);
/* .vbmap with one id per line, as view_pack writes it (kmeans/view_pack.cpp:316-322);
 * g2h_read_vbmap reads both layouts */
int  g2h_write_vbmap_lines(const char* path, uint32_t nbins, const uint32_t* bins, uint64_t nv);

/* ---- synthetic workloads (seeded: splitmix64 -> xoshiro256**) ---- */
/* Chung-Lu: w_i = w_max*(1+i/i0)^(-1/(gamma-1)) with i0 solved so that
 * sum(w) = 2*target_edges, clipped below at w_min; round(sum/2) endpoint
 * pairs drawn proportional to w; self loops and duplicates removed. */
int  g2h_gen_chung_lu(uint64_t n, double gamma, double w_min, double w_max, uint64_t target_edges,
                      uint64_t seed, g2h_graph** out);
/* R-MAT: 2^scale vertices, edge_factor*2^scale draws with quadrant
 * probabilities (a,b,c,1-a-b-c), no noise, then a seeded vertex permutation;
 * symmetrised, deduplicated, no self loops (isolated vertices kept). */
int  g2h_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                  uint64_t seed, g2h_graph** out);
/* Chung-Lu backbone plus planted overlapping communities: ncomm communities
 * with sizes drawn from a discrete power law P(s) ~ s^-comm_alpha on
 * [smin,smax], members uniform over the vertices, intra-edge probability
 * p ~ U[pmin,pmax] resampled (at most 64 times, else pmin) until the expected
 * number of maximal cliques of G(s,p) is <= ecap. */
int  g2h_gen_social(uint64_t n, double gamma, double w_min, double w_max, uint64_t backbone_edges,
                    uint64_t ncomm, uint32_t smin, uint32_t smax, double comm_alpha,
                    double pmin, double pmax, double ecap, uint64_t seed, g2h_graph** out);

/* ---- query patterns and plans ---- */
/* preset pattern (common/graph.h:813-1104), remapped as Graph::ReMapVertexId
 * (common/graph.h:171-266); adjacency returned as bitmasks qadj[u] */
int  g2h_query_preset(int preset, uint32_t* n, uint32_t qadj[G2_MAXQ]);
/* remap an arbitrary connected pattern given by adjacency bitmasks */
int  g2h_query_remap(uint32_t n, const uint32_t qadj_in[G2_MAXQ], uint32_t qadj_out[G2_MAXQ]);
/* symmetry-breaking pairs (a,b) meaning M(a) < M(b): the reference's
 * Graph::GetConditions (common/graph.h:564-686) over the full automorphism
 * group, enumerated directly instead of through bliss. pairs[2*k], pairs[2*k+1] */
int  g2h_query_conditions(uint32_t n, const uint32_t qadj[G2_MAXQ], uint32_t* npairs,
                          uint32_t pairs[2 * G2_MAXQ * G2_MAXQ]);
/* |Aut| of the pattern; 0 when qadj is not a valid pattern (n in 1..G2_MAXQ,
 * symmetric masks, no self loops, no bits at or above n) */
uint64_t g2h_query_automorphisms(uint32_t n, const uint32_t qadj[G2_MAXQ]);
/* isomorphism-invariant code of a pattern (equal codes <=> isomorphic), 0
 * for an invalid pattern; no counterpart in the reference: lets a batch count
 * each class once */
uint64_t g2h_query_canonical(uint32_t n, const uint32_t qadj[G2_MAXQ]);
/* full plan of system/plan.h:51-372 for a remapped pattern */
int  g2h_plan_build(uint32_t n, const uint32_t qadj[G2_MAXQ], g2_plan* plan);

/* seeded random connected pattern: n vertices, density in [0,1]: a random
 * spanning tree plus random extra edges up to round(density*n(n-1)/2) */
int  g2h_query_random(uint32_t n, double density, uint64_t seed, uint32_t qadj[G2_MAXQ]);

/* ---- view bins (view/view_bin.h:41-159) ---- */
/* hop closure of the sources: vertices within hop-1 steps of a source whose
 * degree is >= root_degree keep their full rows; every other row is empty.
 * Output arrays: row_ptrs (nv+1) and cols (capacity cols_cap). */
int  g2h_view_materialize(const g2h_graph* g, const uint32_t* sources, uint64_t ns, int hop,
                          uint64_t root_degree, uint64_t* row_ptrs_out, uint32_t* cols_out,
                          uint64_t cols_cap, uint64_t* ne_out);

/* ---- AIMD root-chunk controller (system/subgraph_container.h:178-211) ----
 * chunk starts at 1e6 roots; an overflowing iteration divides it by 8 (not
 * below 1) and sets the additive step to max(5% of the new chunk, 1) and the
 * threshold to max(2000, old chunk / 2); a clean iteration below the
 * threshold adds `step` and doubles it, at or above it adds the fixed step. */
typedef struct g2h_aimd {
    uint32_t chunk, step, add_step, threshold;
    int32_t  recover;
} g2h_aimd;
void g2h_aimd_init(g2h_aimd* a);
/* returns the reference's trace prefix: "--- chunk: ", "+++ chunk: " or "... chunk: " */
const char* g2h_aimd_adjust(g2h_aimd* a, int overflow);

/* ---- learned root scheduler, host side (no counterpart in the reference;
 * the device side is g2_features / g2_mlp_* in g2hip.h) ---- */
#define G2H_NFEAT_INT 5   /* integer device features per root (g2_features) */
#define G2H_NFEAT     8   /* cost-model inputs per root */
/* core number of every vertex (bucket peel, any CSR with sorted rows) */
int  g2h_core_numbers(const uint64_t* row_ptrs, uint64_t nv, const uint32_t* cols, uint32_t* core);
/* the cost model's 8 inputs of ns roots, float32 row-major: log1p(degree),
 * log1p(core number), log1p(2-hop ball size), log1p(triangles), local clustering
 * 2t/(d(d-1)), id/nv (degeneracy rank of a kcore-recoded graph), shared-
 * neighbour fraction, log1p(later neighbours); feat = g2_features' ns x 5 */
int  g2h_root_features(const int64_t* feat, const uint32_t* sources, uint64_t ns, const uint32_t* core,
                       uint64_t nv, float* out);
/* per-feature standardisation fitted on training rows (population sd,
 * floored at 1e-6) */
typedef struct g2h_norm { float mu[G2H_NFEAT], sd[G2H_NFEAT]; } g2h_norm;
int  g2h_norm_fit(const float* x, uint64_t n, g2h_norm* out);
void g2h_norm_apply(const g2h_norm* norm, float* x, uint64_t n);  /* in place */
/* processing order (longest predicted cost first, stable) and per-root split
 * counts from predicted log(1 + subtree nodes): a root above the per-wave
 * share of the predicted total is dealt into ceil(cost / share) parts, at
 * most later[i] (its possible first-level children) and max_split (0 = 4096) */
int  g2h_learned_schedule(const float* pred_log_nodes, const int64_t* later, uint64_t ns, uint64_t waves,
                          uint32_t max_split, uint32_t* order, uint32_t* split);
/* trained model file (standardiser + MLP parameters), text, exact round trip */
int  g2h_model_save(const char* path, const g2h_norm* norm, const float* params, uint64_t nparams);
int  g2h_model_load(const char* path, g2h_norm* norm, float* params, uint64_t nparams);

const char* g2h_last_error(void);
/* progress lines on stderr from long host steps (generators, CSR build,
 * preprocess, kcore); off by default */
void g2h_set_progress(int on);

#ifdef __cplusplus
}
#endif
#endif
