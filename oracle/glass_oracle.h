/* Test infrastructure: the oracle.
 *
 * A plain-C restatement of pyglass's query-time graph search, used only by
 * tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as the
 * checker for the HIP path.  Nothing in pyglass_amd/ includes or links it.
 * Each function cites the reference file:line it restates.  Parity pinning of
 * each function is listed in glass_oracle.c's header comment.
 */
#ifndef GLASS_ORACLE_H
#define GLASS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* quantizer kinds (names as in glass/quant/quant.hpp:23-43) */
enum { ORQ_FP32 = 0, ORQ_FP16 = 1, ORQ_BF16 = 2, ORQ_E5M2 = 3, ORQ_SQ8U = 4, ORQ_SQ8 = 5, ORQ_SQ6 = 6, ORQ_SQ4U = 7, ORQ_SQ4 = 9, ORQ_SQ2U = 11, ORQ_PQ8 = 13 };
enum { ORQ_L2 = 0, ORQ_IP = 1 };

typedef struct {
    int32_t kind, metric, dim, dim_align, code_size, nbits, align;
    float min, dif, max; /* affine (min, dif) or symmetric (max) calibration */
    float *mins, *difs;  /* per-dim calibration, dim_align entries (SQ8) */
    float *codebook;     /* PQ8: nsq x 256 x 4 centroids (nsq = dim_align = dim / 4) */
} orq_quant;

typedef struct {
    int32_t N, K;
    int32_t *adj; /* N*K, -1 padded */
    int32_t nep;
    int32_t *eps;
    int32_t has_init; /* HNSW upper layers present */
    int32_t init_K, ep;
    int32_t *levels;  /* N */
    int64_t *offsets; /* N+1 offsets into lists */
    int32_t *lists;   /* concatenated upper-level lists */
} orq_graph;

typedef struct {
    int64_t dist_cmps;   /* every distance evaluation incl. initializer */
    int64_t expansions;  /* pool pops */
    int64_t upper_evals; /* distance evaluations in the upper-layer descent */
} orq_counters;

/* Quantizer set-up (train = calibrate).  Returns 0 or -1 on bad args. */
int orq_quant_init(orq_quant *q, int32_t kind, int32_t metric, int32_t dim);
void orq_quant_free(orq_quant *q);
int orq_quant_train(orq_quant *q, const float *data, int64_t n);
/* Encode one row (base or query) into code_size bytes (zero padded). */
void orq_encode(const orq_quant *q, const float *x, uint8_t *out);
void orq_encode_rows(const orq_quant *q, const float *data, int64_t n, uint8_t *codes);
/* helpa::bf16(float) with ROUND_MODE_TO_NEAREST (helpa/types.hpp:22-80) */
uint16_t orq_f32_to_bf16(float f);
/* Query-side representation: encoded code (integer quantizers) or the fp32
 * query zero-padded to dim_align (E5M2, SQ8, FP32).  Size in bytes: */
int64_t orq_query_size(const orq_quant *q);
void orq_prepare_query(const orq_quant *q, const float *x, uint8_t *qbuf);
/* Distance of prepared query to one code row, as the reference computer. */
double orq_distance(const orq_quant *q, const uint8_t *qbuf, const uint8_t *code);
/* ctx: where the reference's computer is inlined (its FMA contraction can
 * differ per context); ORQ_CTX_INIT is or-ed in for initialize_search. */
enum { ORQ_CTX_COMPUTER = 0, ORQ_CTX_SEARCH = 1, ORQ_CTX_BATCH = 2, ORQ_CTX_INIT = 4 };
double orq_distance_ctx(const orq_quant *q, const uint8_t *qbuf, const uint8_t *code, int32_t ctx);

/* .glass graph I/O (glass/graph.hpp:106-153) */
int orq_graph_load(orq_graph *g, const char *path);
int orq_graph_from_arrays(orq_graph *g, int32_t N, int32_t K, const int32_t *adj, int32_t nep, const int32_t *eps);
void orq_graph_free(orq_graph *g);

/* One query.  batch_mode=0 follows GraphSearcher::Search (pool (max(k,ef),
 * max(k,ef))), batch_mode=1 SearchBatch (pool (ef, max(k,ef))).  Writes k
 * ids/dists; entries past the pool size are -1 / +inf. */
void orq_search(const orq_graph *g, const orq_quant *q, const uint8_t *codes, const float *x, int32_t k, int32_t ef,
                int32_t batch_mode, int32_t *ids, float *dists, orq_counters *cnt);

/* Refiner::Search/SearchBatch (glass/searcher/refiner.hpp:45-110): inner search
 * with reorder_k = int(k*factor), then exact FP32 re-rank of those ids. */
void orq_search_refine(const orq_graph *g, const orq_quant *q, const uint8_t *codes, const orq_quant *rq,
                       const uint8_t *rcodes, float factor, const float *x, int32_t k, int32_t ef, int32_t batch_mode,
                       int32_t *ids, float *dists, orq_counters *cnt);

/* get_refine_factor (glass/searcher/graph_searcher.hpp:227-238) */
float orq_refine_factor(int32_t kind);

/* Batched driver over many queries with OpenMP (nthreads<=0: all). */
void orq_search_many(const orq_graph *g, const orq_quant *q, const uint8_t *codes, const orq_quant *rq,
                     const uint8_t *rcodes, float factor, const float *xs, int64_t nq, int32_t k, int32_t ef,
                     int32_t batch_mode, int32_t *ids, float *dists, int64_t *dist_cmps, int64_t *expansions,
                     int32_t nthreads);

/* LinearPool driver for model tests: ops[i] >= 0 inserts (ops[i], dists[i]),
 * ops[i] < 0 pops if has_next() (popped[i] = id, else -1; inserts record -2).
 * Writes the final pool (ids with checked bit 31), its size and cur_. */
int orq_pool_run(int32_t cap, int32_t ef, int32_t nops, const int32_t *ops, const double *dists, int32_t *popped,
                 int32_t *out_ids, double *out_dists, int32_t *out_size, int32_t *out_cur);

/* PQ8 pieces, exposed for tests: std::mt19937 (seed 5489) outputs, the
 * k-means fit of one 4-dim sub-space (glass/algorithms/clustering.hpp) and
 * the LUT sum (ProductQuant::dist_func). */
void orq_mt19937_outputs(int32_t count, uint32_t *out);
void orq_kmeans_fit4(const float *rows, int32_t n, int32_t k, int32_t epochs, float *medoids);
/* glass::Clustering (glass/algorithms/clustering.hpp) over rows of d floats */
void orq_clustering_fit(const float *rows, int32_t n, int32_t d, int32_t k, int32_t epochs, int32_t kmeanspp,
                        float *medoids);
void orq_clustering_predict(const float *medoids, int32_t k, int32_t d, const float *x, int32_t n, int32_t *ids);
void orq_clustering_transform(const float *medoids, int32_t k, int32_t d, float *x, int32_t n);
float orq_pq_lut_sum(const float *lut, const uint8_t *code, int32_t nsq);

/* fp32 -> fp16 bits, round to nearest even (F16C semantics). */
uint16_t orq_f32_to_f16(float f);
float orq_f16_to_f32(uint16_t h);

#ifdef __cplusplus
}
#endif
#endif
