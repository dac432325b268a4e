// Internal interface between the host runtime (glass_hip.cpp) and the HIP
// kernels (search_kernels.hip).  Not part of the public C ABI.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <random>

namespace glass_hip {

// Quantizer kinds; the numbering follows glass/quant/quant.hpp:23 (QuantizerType).
enum QuantKind : int32_t { QK_FP32 = 0, QK_FP16 = 1, QK_BF16 = 2, QK_E5M2 = 3, QK_SQ8U = 4, QK_SQ8 = 5,
                           QK_SQ6 = 6, QK_SQ4U = 7, QK_SQ4 = 9, QK_SQ2U = 11, QK_PQ8 = 13, QK_NONE = -1 };
enum MetricKind : int32_t { MK_L2 = 0, MK_IP = 1 };

// Search-mode semantics of one query (glass/searcher/graph_searcher.hpp):
// MODE_SEARCH   = GraphSearcher::Search      pool (max(k,ef), max(k,ef))  (:134)
// MODE_BATCH    = GraphSearcher::SearchBatch pool (ef, max(k,ef))         (:152)
enum SearchMode : int32_t { MODE_SEARCH = 0, MODE_BATCH = 1 };

struct DeviceIndex {
    // graph (glass/graph.hpp): N x K int32, -1 padded; rows deduplicated
    const int32_t *adj;
    int32_t N, K;
    // entry: HNSW upper layers (glass/hnsw/HNSWInitializer.hpp) or eps list
    int32_t has_init, up_K, ep, nep;
    const int32_t *levels;   // N
    const int64_t *up_off;   // N (offset into up_lists, valid when levels[u] > 0)
    const int32_t *up_lists; // concatenated lists, level l at (l-1)*up_K
    const int32_t *eps;      // nep
    // search quantizer
    int32_t quant, metric, dim, dim_align, code_size;
    const uint8_t *codes;    // N x code_size
    float qmin, qdif, qmax;  // affine / symmetric calibration
    const float *sq8_mins;   // dim_align (SQ8 / SQ6 / SQ4 per-dim)
    const float *sq8_difs;
    const float *pq_codebook; // PQ8: nsq x 256 x 4 centroids (nsq = code_size = dim / 4)
    // re-rank rows of the Refiner's quantizer (make_refiner,
    // graph_searcher.hpp:240-254), or null: N x r_code_size bytes of fp32
    // (rerank_kind QK_FP32) or fp16 (QK_FP16) values, SQ8U codes (QK_SQ8U,
    // integer distances) or per-dim SQ8 codes (QK_SQ8)
    const void *rerank;
    int32_t r_dim_align;
    int32_t rerank_kind;
    float refine_factor;     // get_refine_factor(quantizer)
    int32_t r_code_size;     // bytes per re-rank row
    float r_qmin, r_qdif, r_qmax;  // SQ8U re-rank calibration (affine for L2, symmetric for IP)
    const float *r_mins;     // SQ8 re-rank calibration, r_dim_align each
    const float *r_difs;
};

struct QueryBatch {
    const float *queries;    // nq x dim, fp32, device
    int32_t nq;
    int32_t dim;
    const int32_t *ef;       // per-query ef (device) or null -> ef_uniform
    const int32_t *k;        // per-query k (device) or null -> k_uniform
    int32_t ef_uniform, k_uniform;
    int32_t mode;            // SearchMode
    const int64_t *out_off;  // per-query output offset (device) or null -> q * k_uniform
    int32_t *out_ids;
    float *out_dists;        // may be null
    int32_t *out_counters;   // nq x 3: dist_cmps, expansions, upper evals; may be null
    uint64_t *out_pairs;     // if set: results as (id | dist bits << 32), one 64-bit store each, instead of
                             // out_ids/out_dists (lets a host poll host-mapped results without fences)
};

struct LaunchConfig {
    int32_t team;            // lanes per candidate (integer quantizers)
    int32_t waves_per_block;
    int32_t blocks;          // persistent grid size
    int32_t cap_max;         // LDS pool capacity per wave (>= every query's max(k',ef))
    int32_t vis_hash;        // LDS visited-table entries per wave (power of two >= 256)
    int32_t spec;            // integer codes: gather every neighbour's row before the visited test
    int32_t qwaves;          // waves per query: 1, or 2 / 4 (waves_per_block) waves of a workgroup on one query
};

struct SlotState {
    uint32_t *visited;       // nslots x vis_stride words, each slot: [LUT | dirty list | visited bitset]; the
                             // bitsets are all zero between queries
    int64_t vis_stride;      // words per slot (multiple of 4)
    int64_t lut_words;       // PQ8: the slot's first lut_words words hold the query's LUT (multiple of 4)
    int32_t dirty_words;     // then the bitset words a query set first (multiple of 4), then the bitset

    int32_t nslots;
    int32_t *work_counter;   // device counter for the persistent queue
};

// Launches the persistent search kernel for one batch on `stream`.
// Returns hipSuccess or the launch error.
hipError_t launch_search(const DeviceIndex &ix, const QueryBatch &qb, const LaunchConfig &lc, const SlotState &ss,
                         hipStream_t stream);

// The device array the no-alloc entry point runs in (no hipMalloc): its
// address on device `device` (looked up once per device with that device
// current) and its size.  0 or a GLASS_HIP_* status.
int noalloc_scratch(int device, void **ptr, size_t *bytes);

// Dynamic LDS bytes of one workgroup of a launch.
size_t lds_bytes_per_block(const DeviceIndex &ix, const LaunchConfig &lc);

// Per-wave dynamic LDS bytes for a configuration.
size_t lds_bytes_per_wave(const DeviceIndex &ix, int32_t cap_max, int32_t vis_hash);

// Resident blocks per CU for the kernel `launch_search` would pick (0 on error).
int max_blocks_per_cu(const DeviceIndex &ix, const LaunchConfig &lc);
// VGPRs of the kernel `launch_search` would pick.
int kernel_vgprs(const DeviceIndex &ix, const LaunchConfig &lc);

// Valid team widths for the index's code layout (powers of two dividing
// code_size/4 with at most 16 dwords per lane).
int32_t default_team(const DeviceIndex &ix);
bool team_valid(const DeviceIndex &ix, int32_t team);

// Sets the calling thread's last-error text (glass_hip_last_error) and returns code.
int set_error(int code, const char *msg);
// Status for the exception being handled: no C++ exception crosses the C ABI.
// Every entry point that returns a status catches (...) and returns this.
int exception_status() noexcept;

// GenRandom (glass/utils.hpp:13-27): size sorted, distinct, rotated draws in [0, N).
void gen_random(std::mt19937 &rng, int32_t *addr, int32_t size, int32_t N);

// k-means as glass::Clustering::fit runs it (glass/algorithms/clustering.hpp),
// over sample rows the backend evaluates distances for.
namespace km {

// helpa::l2_fp32_fp32 in the reference's compiled order, on the host.
float l2(const float *x, const float *y, int32_t d);

struct Backend {
    virtual ~Backend() = default;
    // out[i] = l2(row, sample i) for every sample (k-means++ seeding)
    virtual int dist_to(const float *row, float *out) = 0;
    // ids[i] = first nearest of the k centroids to sample i (a Lloyd step)
    virtual int assign(const float *cent, int32_t k, int32_t *ids) = 0;
};

// Clustering::fit on n rows of d floats: samples min(n - 1, 100000) rows with
// GenRandom, seeds k centroids (kmeans++ or random), runs `epochs` Lloyd
// steps.  `make` builds the backend over the sampled rows.  Returns 0 or an
// error code (the backend's).
int fit(const float *rows, int32_t n, int32_t d, int32_t k, int32_t epochs, bool kmeanspp, float *centroids,
        Backend *(*make)(const float *samples, int32_t m, int32_t d, void *ctx), void *ctx);

// A backend evaluating on the host (small d, e.g. PQ8's 4-dim sub-spaces).
Backend *host_backend(const float *samples, int32_t m, int32_t d, void *ctx);

}  // namespace km

hipError_t launch_cluster_assign(const float *x, int32_t n, int32_t d, const float *cent, int32_t k, int32_t *ids,
                                 hipStream_t stream);
hipError_t launch_cluster_dist(const float *x, int32_t n, int32_t d, const float *pivot, float *out,
                               hipStream_t stream);

}  // namespace glass_hip
