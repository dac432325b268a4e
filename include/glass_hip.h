/* glass_hip.h — C ABI of the MI355X query-time graph search library
 * (libglass_hip.so).  Plain pointers and sizes only; every function returns
 * 0 on success or a negative status, with a message in glass_hip_last_error().
 *
 * The entry points are the ones the reference's Python binding and C++
 * searcher interface bind for this path:
 *   glass.Graph               python/bindings.cc:40-58, glass/graph.hpp:106-153
 *   glass.Searcher            python/bindings.cc:90-154,258-269
 *   GraphSearcherBase         glass/searcher/searcher_base.hpp:7-19
 *   create_searcher           glass/searcher/graph_searcher.hpp:256-391
 * Host pointers are expected unless a name ends in _dev (device pointers).
 */
#ifndef GLASS_HIP_H
#define GLASS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct glass_hip_graph glass_hip_graph;
typedef struct glass_hip_searcher glass_hip_searcher;

enum {
    GLASS_HIP_OK = 0,
    GLASS_HIP_EINVAL = -1,   /* bad argument (unknown metric/quantizer, shape, id range) */
    GLASS_HIP_EIO = -2,      /* graph file unreadable or malformed */
    GLASS_HIP_ENOMEM = -3,   /* device or host allocation failed */
    GLASS_HIP_EDEVICE = -4,  /* HIP runtime / kernel launch error, or no device */
    GLASS_HIP_ESTATE = -5,   /* call out of order (e.g. search before set_data) */
    GLASS_HIP_EINTERNAL = -6 /* unexpected C++ exception inside the library */
};

/* Message of the last failing call on this thread. */
const char *glass_hip_last_error(void);
/* Library version string. */
const char *glass_hip_version(void);
/* Number of visible HIP devices (0 when none). */
int glass_hip_device_count(void);

/* ---- Graph (glass::Graph<int32_t>) ------------------------------------ */

/* Graph(filename, format) / Graph::load: format "glass" (graph.hpp:136-153),
 * "nsg" (graph.hpp:220-245), "diskann" (graph.hpp:155-182) or "hnswlib"
 * (graph.hpp:184-218, level 0 of a graph with at most one upper level). */
int glass_hip_graph_load(const char *path, const char *format, glass_hip_graph **out);
/* Graph::save (graph.hpp:106-119), .glass layout incl. HNSW upper layers. */
int glass_hip_graph_save(const glass_hip_graph *g, const char *path);
/* Build from arrays: adj is N*K int32, -1 padded; eps the entry points. */
int glass_hip_graph_create(int32_t n, int32_t k, const int32_t *adj, int32_t nep, const int32_t *eps,
                           glass_hip_graph **out);
/* Attach HNSW upper layers (HNSWInitializer, HNSWInitializer.hpp:13-87):
 * levels[n]; lists concatenated per node, levels[i]*up_k ints each. */
int glass_hip_graph_set_upper(glass_hip_graph *g, int32_t up_k, int32_t ep, const int32_t *levels,
                              const int32_t *lists);
int glass_hip_graph_info(const glass_hip_graph *g, int32_t *n, int32_t *k, int32_t *nep, int32_t *has_upper);
/* Copy adjacency (n*k) / entry points out. */
int glass_hip_graph_copy(const glass_hip_graph *g, int32_t *adj, int32_t *eps);
void glass_hip_graph_free(glass_hip_graph *g);
/* Index(index_type, metric, R, L).build(data) / build_graph
 * (python/bindings.cc:60-88,231-237): CPU graph construction with fp32
 * distances.  "HNSW" restates glass/hnsw/hnsw.hpp:27-69 (M = R/2,
 * efConstruction = L); "NSG" / "SSG" restate glass/nsg/nsg.hpp:22-434 over
 * an NN-Descent kNN graph (L2 whatever `metric` says, as create_nsg).
 * num_threads <= 0 uses every core.  *seconds (nullable) receives the
 * construction time. */
int glass_hip_build_graph(const char *index_type, const float *data, int32_t n, int32_t dim, const char *metric,
                          int32_t R, int32_t L, int32_t num_threads, glass_hip_graph **out, double *seconds);

/* ---- Searcher (GraphSearcherBase) -------------------------------------- */

/* create_searcher(graph, metric, quantizer, refine_quant): metric "L2"/"IP";
 * quantizer "SQ8U", "SQ4U", "SQ2U", "E5M2" (alias "FP8_E5M2"), "SQ8", "SQ6",
 * "SQ4", "PQ8" (dim divisible by 4), "FP32", "FP16", "BF16"
 * (graph_searcher.hpp:263-381); refine_quant "", "FP32", "FP16", "SQ8U" or
 * "SQ8" (make_refiner, graph_searcher.hpp:240-254).  The graph's contents move into the searcher
 * and `g` is left empty (python/bindings.cc:94-98). */
int glass_hip_searcher_create(glass_hip_graph *g, const char *metric, const char *quantizer, const char *refine_quant,
                              glass_hip_searcher **out);
/* SetData: train the quantizer(s), encode every row on the device and keep
 * codes, adjacency and FP32 rows resident in HBM. */
int glass_hip_searcher_set_data(glass_hip_searcher *s, const float *data, int32_t n, int32_t dim);
int glass_hip_searcher_set_ef(glass_hip_searcher *s, int32_t ef);
int32_t glass_hip_searcher_get_ef(const glass_hip_searcher *s);
/* Search: one query, GraphSearcher::Search semantics (pool (max(k,ef),max(k,ef))),
 * one launch.  dst[k] ids, scores[k] distances (scores may be NULL). */
int glass_hip_searcher_search(glass_hip_searcher *s, const float *q, int32_t k, int32_t *dst, float *scores);
/* SearchBatch: nq queries, SearchBatch semantics (pool (ef, max(k,ef))),
 * one persistent launch; dst[nq*k], scores[nq*k] (may be NULL). */
int glass_hip_searcher_search_batch(glass_hip_searcher *s, const float *qs, int32_t nq, int32_t k, int32_t *dst,
                                    float *scores);
/* Optimize: autotune the launch configuration on sampled base rows. */
int glass_hip_searcher_optimize(glass_hip_searcher *s, int32_t num_threads);
/* GetLastSearchAvgDistCmps: mean distance evaluations of the last batch. */
double glass_hip_searcher_last_avg_dist_cmps(const glass_hip_searcher *s);
void glass_hip_searcher_free(glass_hip_searcher *s);

/* ---- Extensions beyond the reference interface ------------------------ */

/* Batch with per-query ef[nq] and k[nq] (SearchBatch semantics per query);
 * results of query i start at out_off[i] (exclusive prefix sum of k), so dst
 * and scores hold sum(k) entries.  counters (nullable) gets 3 ints per query:
 * distance evaluations, expansions, upper-layer evaluations. */
int glass_hip_searcher_search_batch_efk(glass_hip_searcher *s, const float *qs, int32_t nq, const int32_t *ef,
                                        const int32_t *k, int32_t *dst, float *scores, int32_t *counters);
/* Device-resident batch: qs_dev[nq*dim], dst_dev[nq*k], scores_dev (nullable),
 * counters_dev (nullable, 3 per query), all device pointers; enqueued on
 * `stream` (a hipStream_t, NULL = the searcher's stream) without host sync.
 * mode 0 = Search semantics, 1 = SearchBatch semantics. */
int glass_hip_searcher_search_batch_dev(glass_hip_searcher *s, const float *qs_dev, int32_t nq, int32_t k,
                                        int32_t mode, int32_t *dst_dev, float *scores_dev, int32_t *counters_dev,
                                        void *stream);
/* search_batch_dev that allocates nothing.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The work
 * counter and the per-wave overflow bitsets (and PQ8 lookup tables) live in a
 * fixed 128 MiB device array that every GPU the library runs on holds until
 * the process ends; the persistent grid is capped to the waves whose slot
 * areas fit in it (N/8 bytes each, 4 waves at least: N up to about 268M
 * rows), and larger requests fail with GLASS_HIP_EINVAL.  Requires set_data
 * first; launches eagerly (no graph replay).  Every call on a GPU shares that
 * GPU's array, so calls that overlap on one GPU race: order them (one stream,
 * or events). */
int glass_hip_searcher_search_batch_dev_noalloc(glass_hip_searcher *s, const float *qs_dev, int32_t nq, int32_t k,
                                                int32_t mode, int32_t *dst_dev, float *scores_dev,
                                                int32_t *counters_dev, void *stream);
/* Device-resident batch with per-query ef_dev[nq] and k_dev[nq] (SearchBatch
 * semantics per query); results of query i start at off_dev[i] in dst_dev /
 * scores_dev.  max_ef and max_k bound every entry of ef_dev and k_dev (the
 * launch is sized for them; a query beyond them gets ids -1).  Enqueued on
 * `stream` (NULL = the searcher's stream) without host sync. */
int glass_hip_searcher_search_batch_efk_dev(glass_hip_searcher *s, const float *qs_dev, int32_t nq,
                                            const int32_t *ef_dev, const int32_t *k_dev, const int64_t *off_dev,
                                            int32_t max_ef, int32_t max_k, int32_t *dst_dev, float *scores_dev,
                                            int32_t *counters_dev, void *stream);
/* search_batch_dev replays a captured hipGraph (counter-reset kernel +
 * search kernel) when every kernel argument and the stream repeat the
 * previous call's; the first call with new arguments runs eagerly and is
 * then captured.  on = 0 always launches eagerly.  Default on. */
int glass_hip_searcher_set_graph_capture(glass_hip_searcher *s, int32_t on);
/* Re-rank depth: reorder_k = int(k * factor).  0 restores the reference's
 * get_refine_factor(quantizer) (graph_searcher.hpp:227-238: 1.5 for
 * SQ8U/SQ8/SQ4U, 3.0 for SQ2U), which Refiner takes as its constructor
 * argument (refiner.hpp:22-24). */
int glass_hip_searcher_set_refine_factor(glass_hip_searcher *s, float factor);
/* Launch shape.  Optimize tunes, per ef bucket (<=16, 32, 64, 128, 256,
 * 512, larger), the dwords each lane gathers per candidate row (team: 4, 8
 * or 16), the waves per workgroup (1-4) and the LDS visited-table size.
 * set_launch forces team / waves_per_block for every bucket (0 = tuned) and
 * caps the persistent grid at `blocks` workgroups (0 = fill the device). */
int glass_hip_searcher_set_launch(glass_hip_searcher *s, int32_t team, int32_t waves_per_block, int32_t blocks);
/* Integer codes: 1 = gather every neighbour's code row as soon as the
 * adjacency row is known and run the visited test while the loads are in
 * flight (used when one lane group per neighbour fits a wave), 0 = gather
 * only the unvisited neighbours, -1 = tuned by optimize (default). */
int glass_hip_searcher_set_gather(glass_hip_searcher *s, int32_t mode);
/* Waves per query: 1 = one wave per query; 2 or 4 = the waves of one
 * workgroup share a query (the leading wave keeps the pool, the visited set
 * and the merges; every wave computes a share of each expansion's candidate
 * distances), which cuts a query's latency where code rows are long or the
 * batch is too small to fill the device (graphs of at most 64 neighbours per
 * row).  0 (default) = the choice optimize() tunes per ef bucket on a batch
 * of a quarter of the resident waves, used when a batch would leave most
 * wave slots idle. */
int glass_hip_searcher_set_query_waves(glass_hip_searcher *s, int32_t waves);
/* Waves per query of the last launch, and optimize()'s choice for small
 * batches at the current ef. */
int glass_hip_searcher_get_query_waves(const glass_hip_searcher *s, int32_t *last, int32_t *tuned);
/* LDS visited-table entries per query (power of two in [256, 16384]; 0 =
 * tuned / sized from ef).  A query that visits more than half as many nodes
 * continues on the global bitset of its slot. */
int glass_hip_searcher_set_visited_capacity(glass_hip_searcher *s, int32_t entries);
/* Shape of the last launch: team, waves per workgroup, workgroups,
 * visited-table entries (a batch with per-query ef runs with the tuning of
 * its largest ef).  Before the first launch: the shape a batch at the current
 * ef would use, with 0 workgroups. */
int glass_hip_searcher_get_launch(const glass_hip_searcher *s, int32_t *team, int32_t *waves_per_block,
                                  int32_t *blocks, int32_t *vis_entries);
/* Last launch: kernel VGPRs, LDS bytes per workgroup, resident workgroups
 * per CU the persistent grid was sized for. */
int glass_hip_searcher_kernel_info(const glass_hip_searcher *s, int32_t *vgprs, int32_t *lds_per_block,
                                   int32_t *blocks_per_cu);
/* Current device: "name (arch)", compute units, LDS bytes per CU as the
 * runtime reports it, HBM bytes, clock in kHz. */
int glass_hip_device_info(char *name, int32_t name_len, int32_t *cu_count, int64_t *lds_per_cu, int64_t *hbm_bytes,
                          int32_t *clock_khz);
/* Sizes for accounting: code bytes per row, FP32 re-rank bytes per row,
 * graph degree, number of rows, refine factor. */
int glass_hip_searcher_info(const glass_hip_searcher *s, int32_t *code_size, int32_t *rerank_row_bytes,
                            int32_t *degree, int32_t *n, float *refine_factor);
/* Copy the device-resident codes (n*code_size bytes) back to the host. */
int glass_hip_searcher_get_codes(const glass_hip_searcher *s, uint8_t *codes);
/* Calibration parameters: [min, dif] (affine), [max] (symmetric),
 * [mins[dim_align], difs[dim_align]] (SQ8 / SQ6 / SQ4 per-dim) or the PQ8
 * codebook [dim/4][256][4]; returns count via *n. */
int glass_hip_searcher_get_calibration(const glass_hip_searcher *s, float *params, int32_t *n);
/* Synchronize the searcher's stream. */
int glass_hip_searcher_sync(glass_hip_searcher *s);

/* ---- Served form: continuous batching ---------------------------------- */
/* The reference serves one call at a time (python/bindings.cc:100-154: the
 * GIL is released around Search/SearchBatch).  These entry points queue
 * requests from any thread; a dispatcher thread packs whatever is queued (up
 * to max_batch queries; a larger request is split over several batches) into
 * one launch as soon as one of the in-flight batch slots is free, so a lone
 * request never waits for a batch to fill.  Batches in flight are admitted
 * by device memory: up to max_inflight (at most 16), as many as half of the
 * free device memory holds (each needs its waves' visited bitsets and its
 * staging).  Each query gets SearchBatch semantics with its request's ef and
 * k; its result does not depend on what else is in the batch.  The searcher
 * must not be reconfigured (set_ef/optimize/set_launch) while a server on it
 * runs. */
typedef struct glass_hip_server glass_hip_server;

int glass_hip_server_create(glass_hip_searcher *s, int32_t max_batch, int32_t max_k, int32_t max_inflight,
                            glass_hip_server **out);
/* Batches in flight the device memory admitted, max_batch, max_k. */
int glass_hip_server_info(const glass_hip_server *sv, int32_t *max_inflight, int32_t *max_batch, int32_t *max_k);
/* Queue nq query rows (copied) with their ef and k; *ticket identifies the request. */
int glass_hip_server_submit(glass_hip_server *sv, const float *qs, int32_t nq, int32_t ef, int32_t k,
                            int64_t *ticket);
/* Non-blocking: *done = 1 and nq*k ids/scores written (ticket released) when
 * the request has completed, else *done = 0. */
int glass_hip_server_poll(glass_hip_server *sv, int64_t ticket, int32_t *dst, float *scores, int32_t *done);
/* Blocking variant of poll. */
int glass_hip_server_wait(glass_hip_server *sv, int64_t ticket, int32_t *dst, float *scores);
/* Drop a request that has not been dispatched yet (GLASS_HIP_ESTATE once it has). */
int glass_hip_server_cancel(glass_hip_server *sv, int64_t ticket);
/* Offered-load driver: the calling thread queues request i (nq[i] rows of qs,
 * in order, with ef[i], k[i]) arrival_ns[i] after the call starts (sorted),
 * waits for all, and returns each request's arrival-to-result latency; the
 * results of request i follow those of requests < i in dst / scores. */
int glass_hip_server_replay(glass_hip_server *sv, const float *qs, int64_t nreq, const int32_t *nq, const int32_t *ef,
                            const int32_t *k, const int64_t *arrival_ns, int64_t *latency_ns, int32_t *dst,
                            float *scores);
/* Batches launched, queries and requests completed so far. */
int glass_hip_server_stats(glass_hip_server *sv, int64_t *batches, int64_t *queries, int64_t *requests);
/* Stops the dispatcher after the batches in flight; queued requests fail. */
void glass_hip_server_free(glass_hip_server *sv);

/* ---- Quantizer trainers (python/bindings.cc:206-229) -------------------- */

/* quant_<name>(dim).train(data): calibrates `quantizer` ("SQ8U", "SQ8",
 * "FP32", "FP16", "BF16", "E5M2", and every other searcher quantizer) on n
 * rows of dim floats and returns its parameters in the layout of
 * glass_hip_searcher_get_calibration (none for the float formats).  With
 * params == NULL only *nparams is set.  The reference's trainers use the L2
 * metric; metric NULL means "L2". */
int glass_hip_quant_train(const char *quantizer, const char *metric, const float *data, int32_t n, int32_t dim,
                          float *params, int32_t *nparams);

/* ---- Clustering (glass/algorithms/clustering.hpp, python/bindings.cc:155-203) */

typedef struct glass_hip_clustering glass_hip_clustering;
/* Clustering(n_cluster, epochs, init): init "random" (the Python default) or
 * "kmeans++". */
int glass_hip_clustering_create(int32_t n_cluster, int32_t epochs, const char *init, glass_hip_clustering **out);
void glass_hip_clustering_free(glass_hip_clustering *c);
/* fit(data): k-means over min(n - 1, 100000) sampled rows; distance sweeps on
 * the GPU, seeding and centroid updates on the host in the reference's order. */
int glass_hip_clustering_fit(glass_hip_clustering *c, const float *data, int32_t n, int32_t dim);
/* predict(data): each row's nearest centroid (first on ties). */
int glass_hip_clustering_predict(const glass_hip_clustering *c, const float *data, int32_t n, int32_t dim,
                                 int32_t *ids);
/* transform(data, inplace=True): subtracts each row's nearest centroid in place. */
int glass_hip_clustering_transform(const glass_hip_clustering *c, float *data, int32_t n, int32_t dim);
/* The fitted centroids (n_cluster x dim floats); dim is 0 before fit. */
int glass_hip_clustering_centroids(const glass_hip_clustering *c, float *out, int32_t *n_cluster, int32_t *dim);

#ifdef __cplusplus
}
#endif
#endif
