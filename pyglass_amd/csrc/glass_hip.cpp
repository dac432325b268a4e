// Host runtime of libglass_hip.so: the C ABI declared in include/glass_hip.h.
//
// It owns the graph on the host, trains the scalar calibrators
// (glass/quant/calibrator.hpp) on the CPU, uploads the base rows once, encodes
// them on the device, keeps adjacency + codes + FP32 rows resident in HBM and
// drives the HIP search kernels.  Searcher semantics follow
// glass/searcher/graph_searcher.hpp and glass/searcher/refiner.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <fstream>
#include <memory>
#include <mutex>
#include <new>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/glass_hip.h"
#include "glass_hip_internal.h"

namespace glass_hip {
hipError_t launch_encode(const float *data_dev, int64_t n, int32_t dim, int32_t code_size, int32_t quant,
                         int32_t metric, float mn, float dif, float mx, const float *mins, const float *difs,
                         uint8_t *codes_dev, hipStream_t stream);
hipError_t launch_pad_rows(const float *data_dev, int64_t n, int32_t dim, int32_t dim_align, float *out,
                           hipStream_t stream);
hipError_t launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t stream);
hipError_t launch_radix_hist(const float *data, int32_t nseg, int64_t seg_len, int64_t seg_stride, int32_t absval,
                             int32_t shift, const uint32_t *prefix, unsigned long long *hist, hipStream_t stream);
hipError_t launch_pq_encode(const float *data_dev, int64_t n, int32_t nsq, const float *codebook, uint8_t *codes_dev,
                            hipStream_t stream);
hipError_t launch_radix_pick(int32_t ntargets, int32_t shift, const unsigned long long *hist, uint32_t *prefix,
                             unsigned long long *rank, hipStream_t stream);
}  // namespace glass_hip

namespace glass_build {
int build_hnsw(const float *X, int32_t n, int32_t d, bool ip, int32_t R, int32_t L, int32_t num_threads,
               std::vector<int32_t> &adj, int32_t &ep, std::vector<int32_t> &levels, std::vector<int32_t> &lists,
               double *seconds);
int build_nsg(const float *X, int32_t n, int32_t d, int32_t R, int32_t L, bool ssg, int32_t num_threads,
              std::vector<int32_t> &adj, int32_t &ep, double *seconds);
}  // namespace glass_build

using namespace glass_hip;

namespace {

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess) {                                                                \
            return fail(_e == hipErrorOutOfMemory ? GLASS_HIP_ENOMEM : GLASS_HIP_EDEVICE,      \
                        "%s failed: %s", #expr, hipGetErrorString(_e));                        \
        }                                                                                      \
    } while (0)

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t count) {
        if (count <= n && p) return hipSuccess;
        release();
        hipError_t e = hipMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) n = count;
        else p = nullptr;
        return e;
    }
};

// quantizer_map (glass/quant/quant.hpp:25-43) restricted to the kinds with
// device kernels; "FP8_E5M2" is accepted as the README spelling of E5M2.
int32_t parse_quant(const std::string &s) {
    if (s == "SQ8U") return QK_SQ8U;
    if (s == "SQ4U") return QK_SQ4U;
    if (s == "SQ2U") return QK_SQ2U;
    if (s == "E5M2" || s == "FP8_E5M2") return QK_E5M2;
    if (s == "SQ8") return QK_SQ8;
    if (s == "SQ6") return QK_SQ6;
    if (s == "SQ4") return QK_SQ4;
    if (s == "PQ8") return QK_PQ8;
    if (s == "FP32") return QK_FP32;
    if (s == "FP16") return QK_FP16;
    if (s == "BF16") return QK_BF16;
    return QK_NONE;
}

// Quantizer<metric, AlignWidth, NBits> per kind (glass/quant/*_quant.hpp).
void quant_layout(int32_t kind, int32_t dim, int32_t *dim_align, int32_t *code_size) {
    int32_t align = 16, nbits = 8;
    switch (kind) {
        case QK_SQ8U: align = 64; nbits = 8; break;
        case QK_SQ4U: align = 128; nbits = 4; break;
        case QK_SQ2U: align = 256; nbits = 2; break;
        case QK_E5M2: align = 16; nbits = 8; break;
        case QK_SQ8: align = 16; nbits = 8; break;
        case QK_SQ6: align = 8; nbits = 6; break;  // sq6_quant.hpp:13-15
        case QK_SQ4: align = 16; nbits = 4; break;  // sq4_quant.hpp:13-15
        case QK_FP32: align = 16; nbits = 32; break;
        // fp16_quant.hpp:9-15 / bf16_quant.hpp:9-15 without the AVX512FP16 /
        // AVX512BF16 opt-in macros
        case QK_FP16: align = 16; nbits = 16; break;
        case QK_BF16: align = 16; nbits = 16; break;
    }
    if (kind == QK_PQ8) {  // ProductQuant(dim, sd = 4): Quantizer<metric, 1, 8> over dim / 4 sub-spaces
        *dim_align = *code_size = dim / 4;
        return;
    }
    *dim_align = (dim + align - 1) / align * align;  // Tensor ctor, glass/storage/tensor.hpp:30-35
    *code_size = *dim_align * nbits / 8;
}

// get_refine_factor (glass/searcher/graph_searcher.hpp:227-238)
float refine_factor_of(int32_t kind) {
    if (kind == QK_SQ8U || kind == QK_SQ8 || kind == QK_SQ4U || kind == QK_PQ8) return 1.5f;
    if (kind == QK_SQ2U) return 3.0f;
    return 1.0f;
}

}  // namespace

int glass_hip::set_error(int code, const char *msg) { return fail(code, "%s", msg); }

int glass_hip::exception_status() noexcept {
    try {
        throw;
    } catch (const std::bad_alloc &) {
        return fail(GLASS_HIP_ENOMEM, "host allocation failed");
    } catch (const std::length_error &e) {
        return fail(GLASS_HIP_EINVAL, "size too large: %s", e.what());
    } catch (const std::exception &e) {
        return fail(GLASS_HIP_EINTERNAL, "internal error: %s", e.what());
    } catch (...) {
        return fail(GLASS_HIP_EINTERNAL, "internal error: unknown exception");
    }
}

// GenRandom (glass/utils.hpp:13-27) with a default-seeded mt19937: the rows
// GraphSearcher::SetData samples for Optimize, and Clustering's samples.
void glass_hip::gen_random(std::mt19937 &rng, int32_t *addr, int32_t size, int32_t N) {
    for (int32_t i = 0; i < size; ++i) addr[i] = (int32_t)(rng() % (uint32_t)(N - size));
    std::sort(addr, addr + size);
    for (int32_t i = 1; i < size; ++i) {
        if (addr[i] <= addr[i - 1]) addr[i] = addr[i - 1] + 1;
    }
    int32_t off = (int32_t)(rng() % (uint32_t)N);
    for (int32_t i = 0; i < size; ++i) addr[i] = (addr[i] + off) % N;
}

struct glass_hip_graph {
    int32_t N = 0, K = 0;
    std::vector<int32_t> adj;
    std::vector<int32_t> eps;
    bool has_upper = false;
    int32_t up_K = 0, ep = 0;
    std::vector<int32_t> levels;
    std::vector<int64_t> up_off;
    std::vector<int32_t> up_lists;
};

// Everything one batch launch passes to its kernels, plus the stream: two
// launches with equal keys are identical, so a captured graph can replay one.
struct GraphKey {
    glass_hip::DeviceIndex ix;
    glass_hip::QueryBatch qb;
    glass_hip::LaunchConfig lc;
    glass_hip::SlotState ss;
    hipStream_t stream;
};

// Launch state of one stream: its work counter, slot areas and captured
// graph.  Launches on one stream share it (the stream orders them); launches
// on different streams use different states, so they run concurrently (one
// batch's last round overlaps the next batch's first).
struct StreamState {
    hipStream_t stream = nullptr;
    DevBuf<uint32_t> visited;
    int32_t nslots = 0;
    DevBuf<int32_t> counter;
    hipGraphExec_t gexec = nullptr;
    GraphKey gkey{};
    hipEvent_t done = nullptr;  // recorded after each launch that uses this state
    uint64_t last_use = 0;
    StreamState() = default;
    StreamState(const StreamState &) = delete;
    StreamState &operator=(const StreamState &) = delete;
    ~StreamState() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (done) (void)hipEventDestroy(done);
    }
};

struct glass_hip_searcher {
    glass_hip_graph graph;
    int32_t metric = MK_L2, quant = QK_SQ8U, refine = QK_NONE;
    int32_t dim = 0, dim_align = 0, code_size = 0, r_dim_align = 0;
    int32_t nb = 0;
    float qmin = 0, qdif = 0, qmax = 0;
    std::vector<float> sq8_mins, sq8_difs;
    // the re-rank quantizer's row size and calibration (SQ8U / SQ8 re-rank)
    int32_t r_code_size = 0;
    float r_qmin = 0, r_qdif = 0, r_qmax = 0;
    std::vector<float> r_mins, r_difs;
    std::vector<float> pq_codebook;  // PQ8: nsq x 256 x 4
    int32_t ef = 32;
    double last_avg_dist_cmps = 0.0;
    bool has_data = false;

    // launch configuration, tuned per ef bucket by Optimize
    struct Tune {
        int32_t team = 8, waves_per_block = 4, vis_hash = 0;  // vis_hash 0: size from ef
        int32_t spec = 1;  // integer codes: gather all neighbours' rows before the visited test
        int32_t qwaves_small = 1;  // waves per query for batches that leave wave slots idle
    };
    static constexpr int kBuckets = 7;  // ef <= 16, 32, 64, 128, 256, 512, larger
    Tune tune[kBuckets];
    int32_t blocks = 0;           // persistent grid cap (0 = fill the device)
    int32_t vis_override = 0;     // visited-table entries forced by the caller (0 = tuned/auto)
    int32_t force_team = 0, force_wpb = 0;
    int32_t force_spec = -1;      // -1 tuned, 0 off, 1 on
    int32_t force_qwaves = 0;     // waves per query forced by the caller (0 = tuned)
    int32_t cu_count = 0;
    int32_t device = 0;           // HIP device the stream and buffers live on
    LaunchConfig last_launch{};

    static int bucket(int32_t ef) {
        int b = 0;
        for (int32_t e = 16; b < kBuckets - 1 && ef > e; e <<= 1) ++b;
        return b;
    }
    // LDS visited table sized to stay under half full for a typical search
    // visiting ~16-24 nodes per unit of ef.
    static int32_t auto_vis(int32_t ef) {
        int32_t v = 1024;
        while (v < 48 * ef && v < 4096) v <<= 1;
        return v;
    }

    // device-resident index
    DevBuf<int32_t> d_adj, d_levels, d_up_lists, d_eps;
    DevBuf<int64_t> d_up_off;
    DevBuf<uint8_t> d_codes;
    DevBuf<uint8_t> d_rerank;  // re-rank rows, fp32 or fp16 (refine)
    DevBuf<float> d_mins, d_difs, d_codebook;
    DevBuf<float> d_rmins, d_rdifs;  // SQ8 re-rank calibration
    int64_t lut_words = 0;  // PQ8: per-slot LUT words ahead of the visited bitset
    // per-slot scratch of every stream state: words per slot
    int64_t vis_stride = 0;
    // launch states of the streams that launched last (least recently used
    // first to be reused)
    static constexpr int kStreamStates = 4;
    std::vector<std::unique_ptr<StreamState>> sstates;
    uint64_t use_clock = 0;
    // staging for host-pointer calls
    DevBuf<float> d_q;
    DevBuf<int32_t> d_ids, d_counters, d_ef, d_k;
    DevBuf<float> d_dists;
    DevBuf<int64_t> d_off;
    hipStream_t stream = nullptr;
    std::mutex mu;

    // Optimize sample (GraphSearcher::SetData keeps <=1000 base rows, graph_searcher.hpp:62-69)
    std::vector<float> optimize_queries;
    int32_t sample_points_num = 0;

    // device-pointer batches replay a captured graph when their arguments repeat
    bool graph_capture = true;

    ~glass_hip_searcher() {
        sstates.clear();
        if (stream) (void)hipStreamDestroy(stream);
    }

    // re-rank depth multiplier: get_refine_factor(quantizer) unless overridden
    float refine_override = 0.0f;
    float factor() const { return refine_override > 0.0f ? refine_override : refine_factor_of(quant); }

    DeviceIndex index() const {
        DeviceIndex ix{};
        ix.adj = d_adj.p;
        ix.N = graph.N;
        ix.K = graph.K;
        ix.has_init = graph.has_upper ? 1 : 0;
        ix.up_K = graph.up_K;
        ix.ep = graph.ep;
        ix.nep = (int32_t)graph.eps.size();
        ix.levels = d_levels.p;
        ix.up_off = d_up_off.p;
        ix.up_lists = d_up_lists.p;
        ix.eps = d_eps.p;
        ix.quant = quant;
        ix.metric = metric;
        ix.dim = dim;
        ix.dim_align = dim_align;
        ix.code_size = code_size;
        ix.codes = d_codes.p;
        ix.qmin = qmin;
        ix.qdif = qdif;
        ix.qmax = qmax;
        ix.sq8_mins = d_mins.p;
        ix.sq8_difs = d_difs.p;
        ix.pq_codebook = d_codebook.p;
        ix.rerank = refine != QK_NONE ? (const void *)d_rerank.p : nullptr;
        ix.r_dim_align = refine != QK_NONE ? r_dim_align : 0;
        ix.r_code_size = refine != QK_NONE ? r_code_size : 0;
        ix.r_qmin = r_qmin;
        ix.r_qdif = r_qdif;
        ix.r_qmax = r_qmax;
        ix.r_mins = d_rmins.p;
        ix.r_difs = d_rdifs.p;
        ix.rerank_kind = refine;
        ix.refine_factor = refine != QK_NONE ? factor() : 1.0f;
        return ix;
    }
};

namespace {

// Graph rows as the kernels expect them: ids in range, each row cut at its
// first -1 and later duplicates of an id removed.  A repeated id in a row is
// skipped by the visited test in SearchImpl1/2, so dropping it leaves the
// visit order unchanged.
int sanitize_graph(glass_hip_graph &g) {
    const int32_t N = g.N, K = g.K;
    std::atomic<int> bad{0};
    auto work = [&](int64_t lo, int64_t hi) {
        std::vector<int32_t> row(K), tmp(K);
        for (int64_t i = lo; i < hi; ++i) {
            int32_t *r = g.adj.data() + i * K;
            int n = 0;
            while (n < K && r[n] != -1) ++n;
            for (int j = 0; j < n; ++j) {
                if (r[j] < 0 || r[j] >= N) bad = 1;
            }
            std::copy(r, r + n, tmp.begin());
            std::sort(tmp.begin(), tmp.begin() + n);
            bool dup = std::adjacent_find(tmp.begin(), tmp.begin() + n) != tmp.begin() + n;
            if (dup) {
                int m = 0;
                for (int j = 0; j < n; ++j) {
                    bool seen = false;
                    for (int t = 0; t < m; ++t) seen |= row[t] == r[j];
                    if (!seen) row[m++] = r[j];
                }
                std::copy(row.begin(), row.begin() + m, r);
                n = m;
            }
            for (int j = n; j < K; ++j) r[j] = -1;
        }
    };
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    int64_t chunk = ((int64_t)N + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        int64_t lo = t * chunk, hi = std::min<int64_t>(N, lo + chunk);
        if (lo < hi) th.emplace_back(work, lo, hi);
    }
    for (auto &t : th) t.join();
    if (bad) return fail(GLASS_HIP_EINVAL, "graph has neighbour ids outside [0, %d)", N);
    for (int32_t e : g.eps) {
        if (e < 0 || e >= N) return fail(GLASS_HIP_EINVAL, "graph entry point %d outside [0, %d)", e, N);
    }
    if (g.has_upper) {
        if (g.ep < 0 || g.ep >= N) return fail(GLASS_HIP_EINVAL, "HNSW entry point out of range");
        if ((int32_t)g.levels.size() != N) return fail(GLASS_HIP_EINVAL, "HNSW levels size mismatch");
        // every node listed at level l must itself have a level-l list, so the
        // greedy descent never reads past a node's lists
        for (int32_t u = 0; u < N; ++u) {
            for (int32_t lv = 1; lv <= g.levels[u]; ++lv) {
                const int32_t *list = g.up_lists.data() + g.up_off[u] + (int64_t)(lv - 1) * g.up_K;
                for (int32_t i = 0; i < g.up_K && list[i] != -1; ++i) {
                    const int32_t v = list[i];
                    if (v < 0 || v >= N || g.levels[v] < lv)
                        return fail(GLASS_HIP_EINVAL, "HNSW upper-layer id %d invalid at level %d", v, lv);
                }
            }
        }
    }
    return GLASS_HIP_OK;
}

bool read_all(std::ifstream &f, void *p, size_t bytes) {
    f.read((char *)p, (std::streamsize)bytes);
    return (size_t)f.gcount() == bytes;
}

// Graph::load (glass/graph.hpp:136-153) + HNSWInitializer::load (HNSWInitializer.hpp:64-75),
// with 64-bit byte counts.
int load_glass(const char *path, glass_hip_graph &g) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return fail(GLASS_HIP_EIO, "cannot open %s", path);
    int32_t nep = 0;
    if (!read_all(f, &nep, 4) || nep < 0 || nep > (1 << 24)) return fail(GLASS_HIP_EIO, "%s: bad header", path);
    g.eps.resize(nep);
    if (nep && !read_all(f, g.eps.data(), 4 * (size_t)nep)) return fail(GLASS_HIP_EIO, "%s: truncated eps", path);
    if (!read_all(f, &g.N, 4) || !read_all(f, &g.K, 4) || g.N <= 0 || g.K <= 0)
        return fail(GLASS_HIP_EIO, "%s: bad N/K", path);
    g.adj.resize((size_t)g.N * g.K);
    if (!read_all(f, g.adj.data(), (size_t)g.N * g.K * 4)) return fail(GLASS_HIP_EIO, "%s: truncated adjacency", path);
    int32_t hdr[3];
    f.read((char *)hdr, 12);
    if (f.gcount() == 12) {
        if (hdr[0] != g.N || hdr[1] <= 0) return fail(GLASS_HIP_EIO, "%s: bad upper-layer header", path);
        g.has_upper = true;
        g.up_K = hdr[1];
        g.ep = hdr[2];
        g.levels.assign(g.N, 0);
        g.up_off.assign(g.N, 0);
        g.up_lists.clear();
        for (int32_t i = 0; i < g.N; ++i) {
            int32_t cur = 0;
            if (!read_all(f, &cur, 4) || cur < 0 || cur % g.up_K) return fail(GLASS_HIP_EIO, "%s: bad level", path);
            g.levels[i] = cur / g.up_K;
            g.up_off[i] = (int64_t)g.up_lists.size();
            if (cur) {
                size_t o = g.up_lists.size();
                g.up_lists.resize(o + cur);
                if (!read_all(f, g.up_lists.data() + o, 4 * (size_t)cur))
                    return fail(GLASS_HIP_EIO, "%s: truncated upper lists", path);
            }
        }
    }
    return GLASS_HIP_OK;
}

// Graph::load_nsg (graph.hpp:220-245): K, ep, then per node (cnt, ids[cnt]).
int load_nsg(const char *path, glass_hip_graph &g) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) return fail(GLASS_HIP_EIO, "cannot open %s", path);
    int64_t size = f.tellg();
    f.seekg(0);
    int32_t K = 0, ep = 0;
    if (!read_all(f, &K, 4) || !read_all(f, &ep, 4) || K <= 0) return fail(GLASS_HIP_EIO, "%s: bad nsg header", path);
    std::vector<std::vector<int32_t>> rows;
    while ((int64_t)f.tellg() < size) {
        int32_t c = 0;
        if (!read_all(f, &c, 4) || c < 0 || c > K) return fail(GLASS_HIP_EIO, "%s: bad nsg row", path);
        rows.emplace_back(c);
        if (c && !read_all(f, rows.back().data(), 4 * (size_t)c)) return fail(GLASS_HIP_EIO, "%s: truncated", path);
    }
    g.N = (int32_t)rows.size();
    g.K = K;
    g.eps = {ep};
    g.adj.assign((size_t)g.N * K, -1);
    for (int32_t i = 0; i < g.N; ++i) std::copy(rows[i].begin(), rows[i].end(), g.adj.begin() + (size_t)i * K);
    return GLASS_HIP_OK;
}

// Graph::load_hnswlib (graph.hpp:184-218): an hnswlib index file.  Header:
// u64 offset_level0, u64 max_elements (= N), u64 cur_count, u64
// size_per_element, u64 label_offset, u64 offset_data, i32 max_level, i32
// entry point, u64 maxM, u64 maxM0 (= K), u64 M, f64 mult, u64 ef_construction;
// then N elements of size_per_element bytes starting with the level-0 link
// list (count, maxM0 ids).  Only level 0 is read; files with more than one
// upper level are refused, as in the reference.  The count is the low 16
// bits of the list header (hnswlib keeps its delete mark in the third byte).
int load_hnswlib(const char *path, glass_hip_graph &g) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return fail(GLASS_HIP_EIO, "cannot open %s", path);
    uint64_t off0 = 0, max_el = 0, cur = 0, spe = 0, lab = 0, offd = 0, maxM = 0, maxM0 = 0, M = 0, efc = 0;
    int32_t max_level = 0, ep = 0;
    double mult = 0.0;
    if (!read_all(f, &off0, 8) || !read_all(f, &max_el, 8) || !read_all(f, &cur, 8) || !read_all(f, &spe, 8) ||
        !read_all(f, &lab, 8) || !read_all(f, &offd, 8) || !read_all(f, &max_level, 4) || !read_all(f, &ep, 4) ||
        !read_all(f, &maxM, 8) || !read_all(f, &maxM0, 8) || !read_all(f, &M, 8) || !read_all(f, &mult, 8) ||
        !read_all(f, &efc, 8))
        return fail(GLASS_HIP_EIO, "%s: bad hnswlib header", path);
    if (max_level > 1) return fail(GLASS_HIP_EINVAL, "%s: hnswlib graphs with %d levels are not supported", path, max_level);
    if (max_el == 0 || max_el >= (uint64_t)INT32_MAX || maxM0 == 0 || maxM0 > 4096 || spe < 4 * (1 + maxM0))
        return fail(GLASS_HIP_EIO, "%s: bad hnswlib sizes", path);
    g.N = (int32_t)max_el;
    g.K = (int32_t)maxM0;
    g.eps = {ep};
    g.adj.assign((size_t)g.N * g.K, -1);
    std::vector<uint8_t> buf(spe);
    for (int32_t i = 0; i < g.N; ++i) {
        if (!read_all(f, buf.data(), spe)) return fail(GLASS_HIP_EIO, "%s: truncated at element %d", path, i);
        uint32_t hdr;
        std::memcpy(&hdr, buf.data(), 4);
        const uint32_t c = hdr & 0xffffu;
        if (c > maxM0) return fail(GLASS_HIP_EIO, "%s: element %d has %u links > maxM0", path, i, c);
        std::memcpy(g.adj.data() + (size_t)i * g.K, buf.data() + 4, 4 * (size_t)c);
    }
    return GLASS_HIP_OK;
}

// Graph::load_diskann (graph.hpp:155-182): u64 size, K, ep, u64, rows (cnt, ids).
int load_diskann(const char *path, glass_hip_graph &g) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return fail(GLASS_HIP_EIO, "cannot open %s", path);
    uint64_t size = 0, x = 0;
    int32_t K = 0, ep = 0;
    if (!read_all(f, &size, 8) || !read_all(f, &K, 4) || !read_all(f, &ep, 4) || !read_all(f, &x, 8) || K <= 0)
        return fail(GLASS_HIP_EIO, "%s: bad diskann header", path);
    std::vector<std::vector<int32_t>> rows;
    while ((uint64_t)f.tellg() < size) {
        int32_t c = 0;
        if (!read_all(f, &c, 4) || c < 0 || c > K) return fail(GLASS_HIP_EIO, "%s: bad diskann row", path);
        rows.emplace_back(c);
        if (c && !read_all(f, rows.back().data(), 4 * (size_t)c)) return fail(GLASS_HIP_EIO, "%s: truncated", path);
    }
    g.N = (int32_t)rows.size();
    g.K = K;
    g.eps = {ep};
    g.adj.assign((size_t)g.N * K, -1);
    for (int32_t i = 0; i < g.N; ++i) std::copy(rows[i].begin(), rows[i].end(), g.adj.begin() + (size_t)i * K);
    return GLASS_HIP_OK;
}

// Inverse of the device's order-preserving float key.
float key_value(uint32_t k) {
    const uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// The top-th smallest and top-th largest value of each of nseg segments of
// seg_len device floats (segment j at data + j * seg_stride), or of their
// absolute values: the order statistic find_minmax / find_absmax keep in a
// MaxHeap (glass/quant/utils.hpp:31-50), i.e. the elements at positions
// top - 1 and seg_len - top of the sorted segment.  Found exactly on the
// device by a radix select over order-preserving keys, 8 bits per pass, both
// targets of every segment in one pass over the data.
int device_order_stats(hipStream_t stream, const float *data, int32_t nseg, int64_t seg_len, int64_t seg_stride,
                       bool absval, int64_t top, std::vector<float> &lo, std::vector<float> &hi) {
    if (seg_len <= 0) return fail(GLASS_HIP_EINVAL, "too few rows to calibrate");
    if (nseg < 1 || nseg > 65535) return fail(GLASS_HIP_EINVAL, "%d calibration segments (at most 65535)", nseg);
    top = std::max<int64_t>(1, std::min<int64_t>(top, seg_len));
    const int32_t nt = 2 * nseg;
    std::vector<unsigned long long> rank(nt);
    std::vector<uint32_t> prefix(nt, 0u);
    for (int32_t j = 0; j < nseg; ++j) {
        rank[2 * j] = (unsigned long long)(top - 1);
        rank[2 * j + 1] = (unsigned long long)(seg_len - top);
    }
    DevBuf<unsigned long long> d_rank, d_hist;
    DevBuf<uint32_t> d_prefix;
    HIP_TRY(d_rank.alloc(nt));
    HIP_TRY(d_prefix.alloc(nt));
    HIP_TRY(d_hist.alloc((size_t)nt * 256));
    HIP_TRY(hipMemcpyAsync(d_rank.p, rank.data(), sizeof(unsigned long long) * nt, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(d_prefix.p, prefix.data(), 4 * (size_t)nt, hipMemcpyHostToDevice, stream));
    for (int shift = 24; shift >= 0; shift -= 8) {
        HIP_TRY(launch_fill_i32((int32_t *)d_hist.p, (int64_t)nt * 256 * 2, 0, stream));
        HIP_TRY(launch_radix_hist(data, nseg, seg_len, seg_stride, absval ? 1 : 0, shift, d_prefix.p, d_hist.p,
                                  stream));
        HIP_TRY(launch_radix_pick(nt, shift, d_hist.p, d_prefix.p, d_rank.p, stream));
    }
    HIP_TRY(hipMemcpyAsync(prefix.data(), d_prefix.p, 4 * (size_t)nt, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    lo.resize(nseg);
    hi.resize(nseg);
    for (int32_t j = 0; j < nseg; ++j) {
        lo[j] = key_value(prefix[2 * j]);
        hi[j] = key_value(prefix[2 * j + 1]);
    }
    return GLASS_HIP_OK;
}

// PQ8 training: ProductQuant::train (glass/quant/product_quant.hpp:26-41), one
// glass::Clustering of 256 centroids (kmeans++ seeding, 10 Lloyd steps) per
// 4-dim sub-space on the host k-means backend, sub-spaces spread over host
// threads.
namespace pq {

void train(const float *data, int64_t n, int32_t dim, std::vector<float> &codebook) {
    const int32_t nsq = dim / 4;
    codebook.assign((size_t)nsq * 256 * 4, 0.0f);
    const int32_t nt = std::max(1, std::min<int32_t>(nsq, (int32_t)std::thread::hardware_concurrency()));
    std::atomic<int32_t> next{0};
    auto work = [&]() {
        std::vector<float> rows(4 * (size_t)n);
        for (int32_t sq; (sq = next.fetch_add(1)) < nsq;) {
            for (int64_t i = 0; i < n; ++i) std::memcpy(&rows[4 * (size_t)i], data + i * dim + 4 * sq, 16);
            km::fit(rows.data(), (int32_t)n, 4, 256, 10, true, &codebook[(size_t)sq * 256 * 4], km::host_backend,
                    nullptr);
        }
    };
    std::vector<std::thread> th;
    for (int32_t t = 1; t < nt; ++t) th.emplace_back(work);
    work();
    for (auto &t : th) t.join();
}

}  // namespace pq

// Calibration (train) of the search quantizer from the rows already in HBM.
int train_quant(glass_hip_searcher &s, const float *data_dev, int64_t n) {
    const int64_t nd = n * (int64_t)s.dim;
    std::vector<float> lo, hi;
    switch (s.quant) {
        case QK_SQ8U:
            if (s.metric == MK_L2) {  // AffineCalibrator<127>, drop 0 (sq8u_quant.hpp:19-28): min and max
                if (int r = device_order_stats(s.stream, data_dev, 1, nd, 0, false, 1, lo, hi)) return r;
                s.qmin = lo[0];
                s.qdif = hi[0] - lo[0];
            } else {  // SymCalibrator<127>: largest absolute value
                if (int r = device_order_stats(s.stream, data_dev, 1, nd, 0, true, 1, lo, hi)) return r;
                s.qmax = hi[0];
            }
            break;
        case QK_SQ4U:
        case QK_SQ2U: {
            // first 10% of the flattened items; drop 1% (SQ4U) / 5% (SQ2U) tails;
            // counts computed in float as in sq4u_quant.hpp:95-97 and utils.hpp:31-33
            const int64_t nitems = (int64_t)((float)nd * 0.1f);
            const float ratio = s.quant == QK_SQ4U ? 0.01f : 0.05f;
            if (nitems <= 0) return fail(GLASS_HIP_EINVAL, "too few rows to calibrate");
            const int64_t top = (int64_t)((float)nitems * ratio) + 1;
            if (int r = device_order_stats(s.stream, data_dev, 1, nitems, 0, false, top, lo, hi)) return r;
            s.qmin = lo[0];
            s.qdif = hi[0] - lo[0];
            break;
        }
        case QK_SQ8:
        case QK_SQ6:
        case QK_SQ4: {
            // AffinePerDimCalibrator<255> (SQ8, drop 0), <63> (SQ6) and <15>
            // (SQ4, both drop 0.01) through find_minmax_perdim
            // (glass/quant/utils.hpp:52-70, sq8_quant.hpp, sq6_quant.hpp:17-31,
            // sq4_quant.hpp:17-31), which bins flattened item i into
            // "dimension" i / n: segment j is items [j*n, (j+1)*n), not
            // column j.  Padding dims keep min = dif = 0.
            const int64_t top = s.quant == QK_SQ8 ? 1 : (int64_t)((float)n * 0.01f) + 1;
            if (int r = device_order_stats(s.stream, data_dev, s.dim, n, n, false, top, lo, hi)) return r;
            s.sq8_mins.assign(s.dim_align, 0.0f);
            s.sq8_difs.assign(s.dim_align, 0.0f);
            for (int32_t j = 0; j < s.dim; ++j) {
                s.sq8_mins[j] = lo[j];
                s.sq8_difs[j] = hi[j] - lo[j];
            }
            break;
        }
        default:
            break;
    }
    return GLASS_HIP_OK;
}

// The re-rank quantizer's calibration: Refiner::SetData trains its own
// quantizer on the same rows (refiner.hpp:26-37), so train_quant runs for
// that quantizer and its results move to the r_* fields.
int train_refine(glass_hip_searcher &s, const float *data_dev, int64_t n) {
    const int32_t q = s.quant, da = s.dim_align;
    const float qmin = s.qmin, qdif = s.qdif, qmax = s.qmax;
    std::vector<float> mins = std::move(s.sq8_mins), difs = std::move(s.sq8_difs);
    s.quant = s.refine;
    s.dim_align = s.r_dim_align;
    const int r = train_quant(s, data_dev, n);
    s.r_qmin = s.qmin;
    s.r_qdif = s.qdif;
    s.r_qmax = s.qmax;
    s.r_mins = std::move(s.sq8_mins);
    s.r_difs = std::move(s.sq8_difs);
    s.quant = q;
    s.dim_align = da;
    s.qmin = qmin;
    s.qdif = qdif;
    s.qmax = qmax;
    s.sq8_mins = std::move(mins);
    s.sq8_difs = std::move(difs);
    return r;
}

int ensure_stream(glass_hip_searcher &s) {
    if (!s.stream) {
        HIP_TRY(hipGetDevice(&s.device));
        HIP_TRY(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    }
    if (!s.cu_count) {
        int dev = 0;
        HIP_TRY(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        HIP_TRY(hipGetDeviceProperties(&prop, dev));
        s.cu_count = prop.multiProcessorCount;
    }
    return GLASS_HIP_OK;
}

// Persistent grid size: as many blocks as can be resident at once (kernel
// occupancy with this LDS footprint), never more slots than queries.
int32_t grid_blocks(const glass_hip_searcher &s, const LaunchConfig &lc, int32_t nq) {
    int32_t per_cu = max_blocks_per_cu(s.index(), lc);
    if (per_cu <= 0) per_cu = 1;
    int32_t full = per_cu * std::max(1, s.cu_count);
    if (s.blocks > 0) full = std::min(full, s.blocks);
    int32_t need = lc.qwaves > 1 ? nq : (nq + lc.waves_per_block - 1) / lc.waves_per_block;
    return std::max(1, std::min(full, need));
}

// Slot areas (visited bitset, PQ8 table) a workgroup uses: one per wave, or
// one when its waves share a query.
int32_t slots_per_block(const LaunchConfig &lc) { return lc.qwaves > 1 ? 1 : lc.waves_per_block; }

// Dirty-list words per slot: the bitset words a query may record as set
// first; a query that sets more clears its whole bitset at its end.
constexpr int32_t kDirtyWords = 4096;

// Words of slot scratch per resident wave: the PQ8 lookup table, the dirty
// list, then an N-bit visited bitset (each a multiple of 4 words).
int64_t slot_stride(const glass_hip_searcher &s) {
    return s.lut_words + kDirtyWords + ((int64_t)(s.graph.N + 31) / 32 + 3) / 4 * 4;
}

// Slot areas start zeroed (the visited bitsets must be all zero between
// queries); a fill kernel on the launch's stream, not hipMemsetAsync, so the
// launch that follows is ordered after it.
int ensure_slots(glass_hip_searcher &s, StreamState &st, int32_t slots, hipStream_t stream) {
    if (slots <= st.nslots && st.visited.p) return GLASS_HIP_OK;
    s.vis_stride = slot_stride(s);
    HIP_TRY(st.visited.alloc((size_t)slots * (size_t)s.vis_stride));
    HIP_TRY(launch_fill_i32((int32_t *)st.visited.p, (int64_t)slots * s.vis_stride, 0, stream));
    st.nslots = slots;
    return GLASS_HIP_OK;
}

// The launch state of `stream`: its own if it has one, else a new one, else
// the least recently used one, which `stream` first waits for (that state's
// last launch may still be running on its old stream).
int state_for(glass_hip_searcher &s, hipStream_t stream, StreamState **out) {
    StreamState *st = nullptr;
    for (auto &p : s.sstates)
        if (p->stream == stream) st = p.get();
    if (!st && (int)s.sstates.size() < glass_hip_searcher::kStreamStates) {
        auto n = std::make_unique<StreamState>();
        HIP_TRY(hipEventCreateWithFlags(&n->done, hipEventDisableTiming));
        n->stream = stream;
        st = n.get();
        s.sstates.push_back(std::move(n));
    }
    if (!st) {
        for (auto &p : s.sstates)
            if (!st || p->last_use < st->last_use) st = p.get();
        HIP_TRY(hipStreamWaitEvent(stream, st->done, 0));
        st->stream = stream;
    }
    st->last_use = ++s.use_clock;
    *out = st;
    return GLASS_HIP_OK;
}

// The all-neighbour gather needs one lane group per adjacency entry in a wave.
bool spec_fits(const glass_hip_searcher &s, int32_t team) {
    const int32_t code_dw = s.code_size / 4;
    const int32_t parts = (code_dw + team - 1) / team;
    int32_t slot = 1;
    while (slot < parts) slot <<= 1;
    return slot * s.graph.K <= 64;
}

// Pool capacity the launch sizes LDS for: the batch's largest pool rounded up
// to 16 entries (not to a power of two: LDS per wave decides how many waves a
// CU holds).
int32_t round_cap(int32_t cap) { return (std::max(cap, 1) + 15) / 16 * 16; }

// LDS of one CU (CDNA4: 160 KiB); a workgroup's per-wave areas must fit it.
constexpr size_t kLdsPerCu = 160 * 1024;

// The pool lives in LDS: a capacity is served when one wave's area fits a CU
// with the smallest visited table (about 19,000 entries at 128 dimensions).
int check_cap(const glass_hip_searcher &s, int32_t max_cap) {
    const size_t need = lds_bytes_per_wave(s.index(), round_cap(max_cap), 256);
    if (need > kLdsPerCu)
        return fail(GLASS_HIP_EINVAL, "max(k, ef) = %d needs %zu bytes of LDS per query (a CU has %zu)", max_cap, need,
                    kLdsPerCu);
    return GLASS_HIP_OK;
}

// Per-launch state for batches that run concurrently on several streams:
// their own work counter and visited bitsets (max_slots x vis_stride words),
// and optionally (id, dist) pair outputs.
struct BatchState {
    uint32_t *visited = nullptr;
    int32_t max_slots = 0;
    int32_t *counter = nullptr;
    uint64_t *out_pairs = nullptr;
};

// Enqueue one batch: queries/outputs are device pointers.
int run_batch(glass_hip_searcher &s, const float *qs_dev, int32_t nq, const int32_t *ef_dev, const int32_t *k_dev,
              int32_t ef_uniform, int32_t k_uniform, int32_t max_cap, int32_t max_ef, int32_t mode,
              const int64_t *off_dev, int32_t *ids_dev, float *dists_dev, int32_t *counters_dev, hipStream_t stream,
              const BatchState *bs = nullptr, bool graph = false) {
    if (!s.has_data) return fail(GLASS_HIP_ESTATE, "search before set_data");
    if (nq <= 0) return GLASS_HIP_OK;
    if (int r = ensure_stream(s)) return r;
    if (!stream) stream = s.stream;
    LaunchConfig lc{};
    const glass_hip_searcher::Tune &tn = s.tune[glass_hip_searcher::bucket(max_ef)];
    lc.team = s.force_team ? s.force_team : tn.team;
    if (!team_valid(s.index(), lc.team)) lc.team = default_team(s.index());
    lc.waves_per_block = s.force_wpb ? s.force_wpb : tn.waves_per_block;
    lc.cap_max = round_cap(max_cap);
    lc.vis_hash = s.vis_override ? s.vis_override : (tn.vis_hash ? tn.vis_hash : glass_hip_searcher::auto_vis(max_ef));
    lc.spec = s.force_spec >= 0 ? s.force_spec : tn.spec;
    lc.qwaves = 1;
    {
        // Several waves per query when the batch is small (tuned per ef
        // bucket), or when the caller forces it; rows of at most 64
        // neighbours.  A tuned choice is kept while the batch fits in one
        // round of the queries that kernel holds resident, or in two when
        // every query has its own ef: such a batch waits for its few
        // slowest queries, and more waves on each outweigh a partial second
        // round (1M x 960, ef 16-512 and k 1-100 per query, 1,000 queries:
        // 4 waves per query in 1.3 rounds, 4.30 ms, against 2 waves in one
        // round, 5.04 ms; with ef 256 for every query the 1.3 rounds lose).
        const int32_t want = s.force_qwaves ? s.force_qwaves : tn.qwaves_small;
        const int64_t rounds = ef_dev ? 2 : 1;
        if (want > 1 && s.graph.K <= 64) {
            int32_t qw = want;
            if (!s.force_qwaves) {
                auto resident_queries = [&](int32_t w) {
                    LaunchConfig mw = lc;
                    mw.qwaves = w;
                    mw.waves_per_block = w;
                    mw.spec = 0;
                    return (int64_t)std::max(1, max_blocks_per_cu(s.index(), mw)) * std::max(1, s.cu_count);
                };
                while (qw > 1 && (int64_t)nq > rounds * resident_queries(qw)) qw /= 2;
            }
            if (qw > 1) {
                lc.qwaves = qw;
                lc.waves_per_block = qw;
                lc.spec = 0;
            }
        }
    }
    // Large pools: fewer waves per workgroup, then a smaller visited table
    // (exact at any size: it moves to the global bitset when it fills), until
    // a workgroup's LDS fits the CU.
    auto lds_of = [&](const LaunchConfig &c) { return lds_bytes_per_block(s.index(), c); };
    while (lds_of(lc) > kLdsPerCu && lc.qwaves == 1 && lc.waves_per_block > 1) lc.waves_per_block -= 1;
    while (lds_of(lc) > kLdsPerCu && lc.vis_hash > 256) lc.vis_hash /= 2;
    const size_t lds = lds_of(lc);
    if (lds > kLdsPerCu) return fail(GLASS_HIP_EINVAL, "pool capacity %d needs %zu bytes of LDS", lc.cap_max, lds);
    int32_t blocks = grid_blocks(s, lc, nq);
    if (bs) blocks = std::max(1, std::min(blocks, bs->max_slots / slots_per_block(lc)));
    lc.blocks = blocks;
    StreamState *st = nullptr;
    if (!bs) {
        if (int r = state_for(s, stream, &st)) return r;
        if (int r = ensure_slots(s, *st, blocks * slots_per_block(lc), stream)) return r;
        if (!st->counter.p) HIP_TRY(st->counter.alloc(1));
    }
    int32_t *counter = bs ? bs->counter : st->counter.p;
    QueryBatch qb{};
    qb.queries = qs_dev;
    qb.nq = nq;
    qb.dim = s.dim;
    qb.ef = ef_dev;
    qb.k = k_dev;
    qb.ef_uniform = ef_uniform;
    qb.k_uniform = k_uniform;
    qb.mode = mode;
    qb.out_off = off_dev;
    qb.out_ids = ids_dev;
    qb.out_dists = dists_dev;
    qb.out_counters = counters_dev;
    qb.out_pairs = bs ? bs->out_pairs : nullptr;
    SlotState ss{};
    ss.visited = bs ? bs->visited : st->visited.p;
    ss.vis_stride = s.vis_stride;
    ss.lut_words = s.lut_words;
    ss.dirty_words = kDirtyWords;
    ss.nslots = bs ? bs->max_slots : st->nslots;
    ss.work_counter = counter;
    const DeviceIndex ix = s.index();
    // counter reset (a fill kernel, not hipMemsetAsync, so it can sit in a graph) + search
    auto enqueue = [&]() -> hipError_t {
        hipError_t e = launch_fill_i32(counter, 1, 0, stream);
        return e == hipSuccess ? launch_search(ix, qb, lc, ss, stream) : e;
    };
    s.last_launch = lc;
    auto mark = [&]() -> int {
        if (st) HIP_TRY(hipEventRecord(st->done, stream));
        return GLASS_HIP_OK;
    };
    if (!graph || !st) {
        HIP_TRY(enqueue());
        return mark();
    }
    // Graph replay: the key is every kernel argument plus the stream, so a
    // replay launches exactly what an eager call with these arguments would.
    GraphKey key{};
    key.ix = ix;
    key.qb = qb;
    key.lc = lc;
    key.ss = ss;
    key.stream = stream;
    if (st->gexec && std::memcmp(&key, &st->gkey, sizeof(key)) == 0) {
        HIP_TRY(hipGraphLaunch(st->gexec, stream));
        return mark();
    }
    // first call with these arguments: run eagerly (this also does any
    // one-time set-up), then capture the same two launches for next time
    HIP_TRY(enqueue());
    if (int r = mark()) return r;
    if (st->gexec) {
        (void)hipGraphExecDestroy(st->gexec);
        st->gexec = nullptr;
    }
    hipGraph_t gr = nullptr;
    HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    const hipError_t ce = enqueue();
    const hipError_t ee = hipStreamEndCapture(stream, &gr);
    if (ce != hipSuccess || ee != hipSuccess) {
        if (gr) (void)hipGraphDestroy(gr);
        return fail(GLASS_HIP_EDEVICE, "graph capture failed: %s", hipGetErrorString(ce != hipSuccess ? ce : ee));
    }
    const hipError_t ie = hipGraphInstantiate(&st->gexec, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (ie != hipSuccess) {
        st->gexec = nullptr;
        return fail(GLASS_HIP_EDEVICE, "graph instantiate failed: %s", hipGetErrorString(ie));
    }
    st->gkey = key;
    return GLASS_HIP_OK;
}

int32_t inner_k(const glass_hip_searcher &s, int32_t k) {
    if (s.refine != QK_NONE) {
        int32_t rk = (int32_t)((float)k * s.factor());
        return std::max(rk, k);
    }
    return k;
}

// Host-pointer batch used by search / search_batch / search_batch_efk.
int host_batch(glass_hip_searcher &s, const float *qs, int32_t nq, const int32_t *ef, const int32_t *k,
               int32_t k_uniform, int32_t mode, int32_t *dst, float *scores, int32_t *counters_out,
               bool record_avg) {
    if (!s.has_data) return fail(GLASS_HIP_ESTATE, "search before set_data");
    if (nq <= 0) return GLASS_HIP_OK;
    if (int r = ensure_stream(s)) return r;
    int64_t total = 0;
    int32_t max_cap = 1, max_ef = s.ef;
    std::vector<int64_t> off;
    if (k) {
        off.resize(nq);
        max_ef = 1;
        for (int32_t i = 0; i < nq; ++i) {
            if (k[i] <= 0 || ef[i] <= 0) return fail(GLASS_HIP_EINVAL, "k and ef must be positive");
            off[i] = total;
            total += k[i];
            max_cap = std::max(max_cap, std::max(inner_k(s, k[i]), ef[i]));
            max_ef = std::max(max_ef, ef[i]);
        }
    } else {
        if (k_uniform <= 0) return fail(GLASS_HIP_EINVAL, "k must be positive");
        total = (int64_t)nq * k_uniform;
        max_cap = std::max(inner_k(s, k_uniform), s.ef);
    }
    if (int r = check_cap(s, max_cap)) return r;
    HIP_TRY(s.d_q.alloc((size_t)nq * s.dim));
    HIP_TRY(s.d_ids.alloc((size_t)total));
    HIP_TRY(s.d_dists.alloc((size_t)total));
    HIP_TRY(s.d_counters.alloc((size_t)nq * 3));
    HIP_TRY(hipMemcpyAsync(s.d_q.p, qs, sizeof(float) * (size_t)nq * s.dim, hipMemcpyHostToDevice, s.stream));
    const int32_t *d_ef = nullptr, *d_k = nullptr;
    const int64_t *d_off = nullptr;
    if (k) {
        HIP_TRY(s.d_ef.alloc(nq));
        HIP_TRY(s.d_k.alloc(nq));
        HIP_TRY(s.d_off.alloc(nq));
        HIP_TRY(hipMemcpyAsync(s.d_ef.p, ef, 4 * (size_t)nq, hipMemcpyHostToDevice, s.stream));
        HIP_TRY(hipMemcpyAsync(s.d_k.p, k, 4 * (size_t)nq, hipMemcpyHostToDevice, s.stream));
        HIP_TRY(hipMemcpyAsync(s.d_off.p, off.data(), 8 * (size_t)nq, hipMemcpyHostToDevice, s.stream));
        d_ef = s.d_ef.p;
        d_k = s.d_k.p;
        d_off = s.d_off.p;
    }
    int r = run_batch(s, s.d_q.p, nq, d_ef, d_k, s.ef, k_uniform, max_cap, max_ef, mode, d_off, s.d_ids.p, s.d_dists.p,
                      s.d_counters.p, s.stream);
    if (r) return r;
    HIP_TRY(hipMemcpyAsync(dst, s.d_ids.p, 4 * (size_t)total, hipMemcpyDeviceToHost, s.stream));
    if (scores)
        HIP_TRY(hipMemcpyAsync(scores, s.d_dists.p, 4 * (size_t)total, hipMemcpyDeviceToHost, s.stream));
    std::vector<int32_t> cnt((size_t)nq * 3);
    HIP_TRY(hipMemcpyAsync(cnt.data(), s.d_counters.p, 4 * cnt.size(), hipMemcpyDeviceToHost, s.stream));
    HIP_TRY(hipStreamSynchronize(s.stream));
    if (counters_out) std::copy(cnt.begin(), cnt.end(), counters_out);
    if (record_avg) {
        int64_t sum = 0;
        for (int32_t i = 0; i < nq; ++i) sum += cnt[(size_t)i * 3];
        s.last_avg_dist_cmps = (double)sum / nq;
    }
    return GLASS_HIP_OK;
}

}  // namespace

extern "C" {

const char *glass_hip_last_error(void) { return g_last_error.c_str(); }

const char *glass_hip_version(void) { return "glass_hip 0.1.0 (gfx950)"; }

int glass_hip_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int glass_hip_graph_load(const char *path, const char *format, glass_hip_graph **out) {
    try {
        if (!path || !out) return fail(GLASS_HIP_EINVAL, "null argument");
        std::string fmt = format ? format : "glass";
        auto g = std::make_unique<glass_hip_graph>();
        int r;
        if (fmt == "glass") r = load_glass(path, *g);
        else if (fmt == "nsg") r = load_nsg(path, *g);
        else if (fmt == "diskann") r = load_diskann(path, *g);
        else if (fmt == "hnswlib") r = load_hnswlib(path, *g);
        else return fail(GLASS_HIP_EINVAL, "unknown graph format '%s'", fmt.c_str());
        if (r) return r;
        *out = g.release();
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_graph_save(const glass_hip_graph *g, const char *path) {
    try {
        if (!g || !path) return fail(GLASS_HIP_EINVAL, "null argument");
        std::ofstream f(path, std::ios::binary);
        if (!f) return fail(GLASS_HIP_EIO, "cannot write %s", path);
        int32_t nep = (int32_t)g->eps.size();
        f.write((const char *)&nep, 4);
        f.write((const char *)g->eps.data(), 4 * (size_t)nep);
        f.write((const char *)&g->N, 4);
        f.write((const char *)&g->K, 4);
        f.write((const char *)g->adj.data(), (std::streamsize)(4 * (size_t)g->N * g->K));
        if (g->has_upper) {  // HNSWInitializer::save (HNSWInitializer.hpp:77-86)
            f.write((const char *)&g->N, 4);
            f.write((const char *)&g->up_K, 4);
            f.write((const char *)&g->ep, 4);
            for (int32_t i = 0; i < g->N; ++i) {
                int32_t cur = g->levels[i] * g->up_K;
                f.write((const char *)&cur, 4);
                if (cur) f.write((const char *)(g->up_lists.data() + g->up_off[i]), 4 * (size_t)cur);
            }
        }
        if (!f) return fail(GLASS_HIP_EIO, "write to %s failed", path);
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_graph_create(int32_t n, int32_t k, const int32_t *adj, int32_t nep, const int32_t *eps,
                           glass_hip_graph **out) {
    try {
        if (n <= 0 || k <= 0 || !adj || nep < 0 || (nep && !eps) || !out) return fail(GLASS_HIP_EINVAL, "bad graph arrays");
        auto g = std::make_unique<glass_hip_graph>();
        g->N = n;
        g->K = k;
        g->adj.assign(adj, adj + (size_t)n * k);
        g->eps.assign(eps, eps + nep);
        *out = g.release();
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_graph_set_upper(glass_hip_graph *g, int32_t up_k, int32_t ep, const int32_t *levels,
                              const int32_t *lists) {
    try {
        if (!g || up_k <= 0 || !levels) return fail(GLASS_HIP_EINVAL, "bad upper-layer arrays");
        g->has_upper = true;
        g->up_K = up_k;
        g->ep = ep;
        g->levels.assign(levels, levels + g->N);
        g->up_off.assign(g->N, 0);
        int64_t total = 0;
        for (int32_t i = 0; i < g->N; ++i) {
            if (levels[i] < 0) return fail(GLASS_HIP_EINVAL, "negative level");
            g->up_off[i] = total;
            total += (int64_t)levels[i] * up_k;
        }
        if (total && !lists) return fail(GLASS_HIP_EINVAL, "missing upper lists");
        g->up_lists.assign(lists, lists + total);
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_graph_info(const glass_hip_graph *g, int32_t *n, int32_t *k, int32_t *nep, int32_t *has_upper) {
    try {
        if (!g) return fail(GLASS_HIP_EINVAL, "null graph");
        if (n) *n = g->N;
        if (k) *k = g->K;
        if (nep) *nep = (int32_t)g->eps.size();
        if (has_upper) *has_upper = g->has_upper ? 1 : 0;
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_graph_copy(const glass_hip_graph *g, int32_t *adj, int32_t *eps) {
    try {
        if (!g) return fail(GLASS_HIP_EINVAL, "null graph");
        if (adj) std::copy(g->adj.begin(), g->adj.end(), adj);
        if (eps) std::copy(g->eps.begin(), g->eps.end(), eps);
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

void glass_hip_graph_free(glass_hip_graph *g) { delete g; }

int glass_hip_build_graph(const char *index_type, const float *data, int32_t n, int32_t dim, const char *metric,
                          int32_t R, int32_t L, int32_t num_threads, glass_hip_graph **out, double *seconds) {
    try {
        if (!index_type || !data || n <= 0 || dim <= 0 || !metric || !out || R < 2 || R > 128)
            return fail(GLASS_HIP_EINVAL, "bad build arguments");
        std::string t(index_type), m(metric);
        if (m != "L2" && m != "IP") return fail(GLASS_HIP_EINVAL, "unknown metric '%s'", metric);
        if (t == "NSG" || t == "SSG") {
            // create_nsg (glass/nsg/nsg.hpp:411-434): the metric is ignored, graphs are L2
            auto g = std::make_unique<glass_hip_graph>();
            int32_t ep = 0;
            if (glass_build::build_nsg(data, n, dim, R, L, t == "SSG", num_threads, g->adj, ep, seconds))
                return fail(GLASS_HIP_EINVAL, "%s build failed", index_type);
            g->N = n;
            g->K = R;
            g->eps = {ep};
            *out = g.release();
            return GLASS_HIP_OK;
        }
        if (t != "HNSW") return fail(GLASS_HIP_EINVAL, "unsupported index type '%s'", index_type);
        auto g = std::make_unique<glass_hip_graph>();
        std::vector<int32_t> levels, lists;
        int32_t ep = 0;
        if (glass_build::build_hnsw(data, n, dim, m == "IP", R, L, num_threads, g->adj, ep, levels, lists, seconds))
            return fail(GLASS_HIP_EINVAL, "HNSW build failed");
        g->N = n;
        g->K = R;
        g->eps = {ep};
        if (int r = glass_hip_graph_set_upper(g.get(), R / 2, ep, levels.data(), lists.data())) return r;
        *out = g.release();
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_searcher_create(glass_hip_graph *g, const char *metric, const char *quantizer, const char *refine_quant,
                              glass_hip_searcher **out) {
    try {
        if (!g || !metric || !quantizer || !out) return fail(GLASS_HIP_EINVAL, "null argument");
        if (g->N <= 0) return fail(GLASS_HIP_EINVAL, "empty graph (was it moved into another searcher?)");
        std::string m(metric), q(quantizer), r(refine_quant ? refine_quant : "");
        int32_t mk;
        if (m == "L2") mk = MK_L2;
        else if (m == "IP") mk = MK_IP;
        else return fail(GLASS_HIP_EINVAL, "unknown metric '%s'", metric);
        int32_t qk = parse_quant(q);
        if (qk == QK_NONE) return fail(GLASS_HIP_EINVAL, "unsupported quantizer '%s'", quantizer);
        // make_refiner (graph_searcher.hpp:240-254): FP32, FP16, SQ8U and SQ8 re-rank rows
        int32_t rk = QK_NONE;
        if (!r.empty()) {
            if (r == "FP32") rk = QK_FP32;
            else if (r == "FP16") rk = QK_FP16;
            else if (r == "SQ8U") rk = QK_SQ8U;
            else if (r == "SQ8") rk = QK_SQ8;
            else return fail(GLASS_HIP_EINVAL, "unsupported refine quantizer '%s'", refine_quant);
        }
        if (g->K > 128) return fail(GLASS_HIP_EINVAL, "graph degree %d > 128 is not supported", g->K);
        if (g->has_upper && g->up_K > 64) return fail(GLASS_HIP_EINVAL, "upper-layer degree %d > 64", g->up_K);
        if (!g->has_upper && g->eps.empty()) return fail(GLASS_HIP_EINVAL, "graph has no entry point");
        auto s = std::make_unique<glass_hip_searcher>();
        s->graph = std::move(*g);
        *g = glass_hip_graph();  // the caller's graph is left empty (bindings.cc:94-98)
        s->metric = mk;
        // SQ2U with IP instantiates the L2 searcher (graph_searcher.hpp:357)
        s->quant = qk;
        s->refine = rk;
        if (int rr = sanitize_graph(s->graph)) return rr;
        *out = s.release();
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_searcher_set_data(glass_hip_searcher *s, const float *data, int32_t n, int32_t dim) {
    try {
        if (!s || !data || n <= 0 || dim <= 0) return fail(GLASS_HIP_EINVAL, "bad data");
        if (n != s->graph.N) return fail(GLASS_HIP_EINVAL, "data has %d rows but the graph has %d nodes", n, s->graph.N);
        std::lock_guard<std::mutex> lk(s->mu);
        if (int r = ensure_stream(*s)) return r;
        if (s->quant == QK_PQ8 && dim % 4 != 0)
            return fail(GLASS_HIP_EINVAL, "PQ8 needs a dimension divisible by 4 (got %d)", dim);
        s->nb = n;
        s->dim = dim;
        quant_layout(s->quant, dim, &s->dim_align, &s->code_size);
        s->lut_words = s->quant == QK_PQ8 ? (int64_t)s->code_size * 256 : 0;
        for (auto &st : s->sstates) st->nslots = 0;  // the slot stride depends on lut_words and N: re-sized at the next launch
        s->vis_stride = slot_stride(*s);
        int32_t r_row_bytes = 0;
        if (s->refine != QK_NONE) quant_layout(s->refine, dim, &s->r_dim_align, &r_row_bytes);
        s->r_code_size = r_row_bytes;
        // graph
        const glass_hip_graph &g = s->graph;
        HIP_TRY(s->d_adj.alloc(g.adj.size()));
        HIP_TRY(hipMemcpy(s->d_adj.p, g.adj.data(), 4 * g.adj.size(), hipMemcpyHostToDevice));
        HIP_TRY(s->d_eps.alloc(std::max<size_t>(1, g.eps.size())));
        if (!g.eps.empty()) HIP_TRY(hipMemcpy(s->d_eps.p, g.eps.data(), 4 * g.eps.size(), hipMemcpyHostToDevice));
        if (g.has_upper) {
            HIP_TRY(s->d_levels.alloc(g.levels.size()));
            HIP_TRY(hipMemcpy(s->d_levels.p, g.levels.data(), 4 * g.levels.size(), hipMemcpyHostToDevice));
            HIP_TRY(s->d_up_off.alloc(g.up_off.size()));
            HIP_TRY(hipMemcpy(s->d_up_off.p, g.up_off.data(), 8 * g.up_off.size(), hipMemcpyHostToDevice));
            HIP_TRY(s->d_up_lists.alloc(std::max<size_t>(1, g.up_lists.size())));
            if (!g.up_lists.empty())
                HIP_TRY(hipMemcpy(s->d_up_lists.p, g.up_lists.data(), 4 * g.up_lists.size(), hipMemcpyHostToDevice));
        }
        // base rows -> device, encode there
        DevBuf<float> raw;
        HIP_TRY(raw.alloc((size_t)n * dim));
        HIP_TRY(hipMemcpy(raw.p, data, sizeof(float) * (size_t)n * dim, hipMemcpyHostToDevice));
        if (int r = train_quant(*s, raw.p, n)) return r;
        if (s->quant == QK_SQ8 || s->quant == QK_SQ6 || s->quant == QK_SQ4) {
            HIP_TRY(s->d_mins.alloc(s->dim_align));
            HIP_TRY(s->d_difs.alloc(s->dim_align));
            HIP_TRY(hipMemcpy(s->d_mins.p, s->sq8_mins.data(), 4 * (size_t)s->dim_align, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(s->d_difs.p, s->sq8_difs.data(), 4 * (size_t)s->dim_align, hipMemcpyHostToDevice));
        }
        HIP_TRY(s->d_codes.alloc((size_t)n * s->code_size));
        if (s->quant == QK_FP32) {
            HIP_TRY(launch_pad_rows(raw.p, n, dim, s->dim_align, (float *)s->d_codes.p, s->stream));
        } else if (s->quant == QK_PQ8) {
            // k-means on the host, as ProductQuant::train; nearest-centroid codes on the device
            pq::train(data, n, dim, s->pq_codebook);
            HIP_TRY(s->d_codebook.alloc(s->pq_codebook.size()));
            HIP_TRY(hipMemcpy(s->d_codebook.p, s->pq_codebook.data(), 4 * s->pq_codebook.size(), hipMemcpyHostToDevice));
            HIP_TRY(launch_pq_encode(raw.p, n, s->code_size, s->d_codebook.p, s->d_codes.p, s->stream));
        } else {
            HIP_TRY(launch_encode(raw.p, n, dim, s->code_size, s->quant, s->metric, s->qmin, s->qdif, s->qmax,
                                  s->d_mins.p, s->d_difs.p, s->d_codes.p, s->stream));
        }
        if (s->refine == QK_FP32) {
            HIP_TRY(s->d_rerank.alloc((size_t)n * r_row_bytes));
            HIP_TRY(launch_pad_rows(raw.p, n, dim, s->r_dim_align, (float *)s->d_rerank.p, s->stream));
        } else if (s->refine == QK_FP16) {  // FP16Quantizer::add, as for the search codes
            HIP_TRY(s->d_rerank.alloc((size_t)n * r_row_bytes));
            HIP_TRY(launch_encode(raw.p, n, dim, r_row_bytes, QK_FP16, s->metric, 0.0f, 0.0f, 0.0f, nullptr, nullptr,
                                  s->d_rerank.p, s->stream));
        } else if (s->refine == QK_SQ8U || s->refine == QK_SQ8) {
            // SQ8QuantizerUniform / SQ8Quantizer trained and encoded like search codes
            if (int r = train_refine(*s, raw.p, n)) return r;
            if (s->refine == QK_SQ8) {
                HIP_TRY(s->d_rmins.alloc(s->r_dim_align));
                HIP_TRY(s->d_rdifs.alloc(s->r_dim_align));
                HIP_TRY(hipMemcpy(s->d_rmins.p, s->r_mins.data(), 4 * (size_t)s->r_dim_align, hipMemcpyHostToDevice));
                HIP_TRY(hipMemcpy(s->d_rdifs.p, s->r_difs.data(), 4 * (size_t)s->r_dim_align, hipMemcpyHostToDevice));
            }
            HIP_TRY(s->d_rerank.alloc((size_t)n * r_row_bytes));
            HIP_TRY(launch_encode(raw.p, n, dim, r_row_bytes, s->refine, s->metric, s->r_qmin, s->r_qdif, s->r_qmax,
                                  s->d_rmins.p, s->d_rdifs.p, s->d_rerank.p, s->stream));
        }
        HIP_TRY(hipStreamSynchronize(s->stream));
        raw.release();
        // Optimize sample (graph_searcher.hpp:62-69)
        s->sample_points_num = std::min(1000, n - 1);
        if (s->sample_points_num > 0) {
            std::vector<int32_t> pts(s->sample_points_num);
            std::mt19937 rng;
            gen_random(rng, pts.data(), s->sample_points_num, n);
            s->optimize_queries.resize((size_t)s->sample_points_num * dim);
            for (int32_t i = 0; i < s->sample_points_num; ++i)
                std::memcpy(s->optimize_queries.data() + (size_t)i * dim, data + (size_t)pts[i] * dim, 4 * (size_t)dim);
        }
        for (auto &t : s->tune) t.team = default_team(s->index());
        s->has_data = true;
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_searcher_set_ef(glass_hip_searcher *s, int32_t ef) {
    try {
        if (!s || ef <= 0) return fail(GLASS_HIP_EINVAL, "ef must be positive");
        s->ef = ef;
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

int32_t glass_hip_searcher_get_ef(const glass_hip_searcher *s) { return s ? s->ef : 0; }

int glass_hip_searcher_search(glass_hip_searcher *s, const float *q, int32_t k, int32_t *dst, float *scores) {
    try {
        if (!s || !q || !dst || k <= 0) return fail(GLASS_HIP_EINVAL, "bad search arguments");
        std::lock_guard<std::mutex> lk(s->mu);
        // one query per launch: a single wave
        const int32_t saved = s->blocks, saved_wpb = s->force_wpb;
        s->blocks = 1;
        s->force_wpb = 1;
        int r = host_batch(*s, q, 1, nullptr, nullptr, k, MODE_SEARCH, dst, scores, nullptr, false);
        s->blocks = saved;
        s->force_wpb = saved_wpb;
        return r;
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_searcher_search_batch(glass_hip_searcher *s, const float *qs, int32_t nq, int32_t k, int32_t *dst,
                                    float *scores) {
    try {
        if (!s || (!qs && nq) || (!dst && nq) || nq < 0 || k <= 0) return fail(GLASS_HIP_EINVAL, "bad batch arguments");
        std::lock_guard<std::mutex> lk(s->mu);
        return host_batch(*s, qs, nq, nullptr, nullptr, k, MODE_BATCH, dst, scores, nullptr, true);
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_searcher_search_batch_efk(glass_hip_searcher *s, const float *qs, int32_t nq, const int32_t *ef,
                                        const int32_t *k, int32_t *dst, float *scores, int32_t *counters) {
    try {
        if (!s || nq < 0 || (nq && (!qs || !ef || !k || !dst))) return fail(GLASS_HIP_EINVAL, "bad batch arguments");
        std::lock_guard<std::mutex> lk(s->mu);
        return host_batch(*s, qs, nq, ef, k, 0, MODE_BATCH, dst, scores, counters, true);
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_searcher_search_batch_dev(glass_hip_searcher *s, const float *qs_dev, int32_t nq, int32_t k,
                                        int32_t mode, int32_t *dst_dev, float *scores_dev, int32_t *counters_dev,
                                        void *stream) {
    try {
        if (!s || nq < 0 || k <= 0 || (nq && (!qs_dev || !dst_dev))) return fail(GLASS_HIP_EINVAL, "bad batch arguments");
        if (mode != MODE_SEARCH && mode != MODE_BATCH) return fail(GLASS_HIP_EINVAL, "mode must be 0 or 1");
        std::lock_guard<std::mutex> lk(s->mu);
        const int32_t max_cap = std::max(inner_k(*s, k), s->ef);
        if (int r = check_cap(*s, max_cap)) return r;
        return run_batch(*s, qs_dev, nq, nullptr, nullptr, s->ef, k, max_cap, s->ef, mode, nullptr, dst_dev, scores_dev,
                         counters_dev, (hipStream_t)stream, nullptr, s->graph_capture);
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_searcher_search_batch_dev_noalloc(glass_hip_searcher *s, const float *qs_dev, int32_t nq, int32_t k,
                                                int32_t mode, int32_t *dst_dev, float *scores_dev,
                                                int32_t *counters_dev, void *stream) {
    try {
        if (!s || nq < 0 || k <= 0 || (nq && (!qs_dev || !dst_dev))) return fail(GLASS_HIP_EINVAL, "bad batch arguments");
        if (mode != MODE_SEARCH && mode != MODE_BATCH) return fail(GLASS_HIP_EINVAL, "mode must be 0 or 1");
        if (!s->has_data || !s->stream) return fail(GLASS_HIP_ESTATE, "search before set_data");
        std::lock_guard<std::mutex> lk(s->mu);
        const int32_t max_cap = std::max(inner_k(*s, k), s->ef);
        if (int r = check_cap(*s, max_cap)) return r;
        // the device's scratch array: [work counter | 256 bytes][one slot area per resident wave]
        void *base = nullptr;
        size_t bytes = 0;
        if (int r = noalloc_scratch(s->device, &base, &bytes)) return r;
        const int64_t slot_bytes = 4 * slot_stride(*s);
        const int64_t slots = ((int64_t)bytes - 256) / slot_bytes;
        if (slots < 4)
            return fail(GLASS_HIP_EINVAL,
                        "%d rows need %lld bytes of scratch per wave: the %zu-byte no-alloc scratch holds fewer "
                        "than 4 waves",
                        s->graph.N, (long long)slot_bytes, bytes);
        // The bitsets must be all zero between queries; every query leaves
        // its own so, but a different slot layout (another searcher's N or
        // PQ8 table) reads other words as bitsets: zero the array then.
        {
            static std::mutex lay_mu;
            static int64_t lay_stride[64] = {}, lay_lut[64] = {};
            const int d = s->device;
            std::lock_guard<std::mutex> g(lay_mu);
            if (d < 0 || d >= 64) return fail(GLASS_HIP_EINVAL, "device index out of range");
            if (lay_stride[d] != slot_stride(*s) || lay_lut[d] != s->lut_words) {
                hipStream_t st = stream ? (hipStream_t)stream : s->stream;
                HIP_TRY(launch_fill_i32((int32_t *)base, (int64_t)(bytes / 4), 0, st));
                lay_stride[d] = slot_stride(*s);
                lay_lut[d] = s->lut_words;
            }
        }
        BatchState bs;
        bs.counter = (int32_t *)base;
        bs.visited = (uint32_t *)((uint8_t *)base + 256);
        bs.max_slots = (int32_t)std::min<int64_t>(slots, 1 << 20);
        return run_batch(*s, qs_dev, nq, nullptr, nullptr, s->ef, k, max_cap, s->ef, mode, nullptr, dst_dev,
                         scores_dev, counters_dev, (hipStream_t)stream, &bs, false);
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_searcher_search_batch_efk_dev(glass_hip_searcher *s, const float *qs_dev, int32_t nq,
                                            const int32_t *ef_dev, const int32_t *k_dev, const int64_t *off_dev,
                                            int32_t max_ef, int32_t max_k, int32_t *dst_dev, float *scores_dev,
                                            int32_t *counters_dev, void *stream) {
    try {
        if (!s || nq < 0 || max_ef <= 0 || max_k <= 0 ||
            (nq && (!qs_dev || !ef_dev || !k_dev || !off_dev || !dst_dev)))
            return fail(GLASS_HIP_EINVAL, "bad batch arguments");
        std::lock_guard<std::mutex> lk(s->mu);
        const int32_t max_cap = std::max(inner_k(*s, max_k), max_ef);
        if (int r = check_cap(*s, max_cap)) return r;
        return run_batch(*s, qs_dev, nq, ef_dev, k_dev, 0, 0, max_cap, max_ef, MODE_BATCH, off_dev, dst_dev,
                         scores_dev, counters_dev, (hipStream_t)stream, nullptr, s->graph_capture);
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_searcher_set_graph_capture(glass_hip_searcher *s, int32_t on) {
    try {
        if (!s) return fail(GLASS_HIP_EINVAL, "null searcher");
        std::lock_guard<std::mutex> lk(s->mu);
        s->graph_capture = on != 0;
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_searcher_optimize(glass_hip_searcher *s, int32_t num_threads) {
    try {
        (void)num_threads;  // unused by the reference too (graph_searcher.hpp:77-79)
        if (!s) return fail(GLASS_HIP_EINVAL, "null searcher");
        if (!s->has_data) return fail(GLASS_HIP_ESTATE, "optimize before set_data");
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->sample_points_num <= 0) return GLASS_HIP_OK;
        // GraphSearcher::Optimize (graph_searcher.hpp:76-129) times SearchBatch on
        // the sampled base rows at k=10 over a grid of prefetch knobs and keeps
        // the fastest.  Here the knobs are the launch shape (lanes per candidate,
        // waves per workgroup, LDS visited-table size), tuned per ef bucket on the
        // samples tiled to a device-filling batch.
        const int32_t kTryK = 10;
        const int32_t base = s->sample_points_num;
        const int32_t nq = std::max(base, std::min(8192, 8 * base));
        std::vector<float> qs((size_t)nq * s->dim);
        for (int32_t i = 0; i < nq; ++i)
            std::memcpy(qs.data() + (size_t)i * s->dim, s->optimize_queries.data() + (size_t)(i % base) * s->dim,
                        4 * (size_t)s->dim);
        std::vector<int32_t> teams;
        for (int32_t t : {4, 8, 16}) {
            if (team_valid(s->index(), t)) teams.push_back(t);
        }
        const int32_t wpbs[] = {1, 2, 4};
        // small tables hold more waves per CU where most queries overflow to the global bitset anyway
        const int32_t vhs[] = {256, 512, 1024, 2048, 4096};
        std::vector<int32_t> ids((size_t)nq * kTryK);
        const int32_t saved_ef = s->ef;
        const int32_t saved_vo = s->vis_override, saved_ft = s->force_team, saved_fw = s->force_wpb,
                      saved_fs = s->force_spec, saved_fq = s->force_qwaves;
        s->force_qwaves = 1;
        const bool int_codes = s->quant == QK_SQ8U || s->quant == QK_SQ4U || s->quant == QK_SQ2U;
        auto timeit_n = [&](int32_t n) -> double {
            auto t0 = std::chrono::high_resolution_clock::now();
            int r = host_batch(*s, qs.data(), n, nullptr, nullptr, kTryK, MODE_BATCH, ids.data(), nullptr, nullptr,
                               false);
            auto t1 = std::chrono::high_resolution_clock::now();
            return r ? 1e30 : std::chrono::duration<double>(t1 - t0).count();
        };
        auto timeit = [&]() -> double { return timeit_n(nq); };
        for (int b = 0; b < glass_hip_searcher::kBuckets - 1; ++b) {
            s->ef = 16 << b;  // upper edge of the bucket
            double best = 1e30;
            glass_hip_searcher::Tune bt = s->tune[b];
            for (int32_t t : teams) {
                for (int32_t w : wpbs) {
                    for (int32_t v : vhs) {
                        for (int32_t sp : {0, 1}) {
                            if (sp && !(int_codes && spec_fits(*s, t))) continue;
                            s->force_team = t;
                            s->force_wpb = w;
                            s->vis_override = v;
                            s->force_spec = sp;
                            const size_t lds = lds_bytes_per_wave(
                                                   s->index(), round_cap(std::max(inner_k(*s, kTryK), s->ef)), v) *
                                               (size_t)w;
                            if (lds > kLdsPerCu) continue;
                            timeit();
                            // A workgroup of w waves starts only when w wave slots
                            // of one CU are free at once, so launches that overlap
                            // (batches on two streams, concurrent served batches)
                            // refill a draining CU more coarsely than a launch
                            // timed alone shows: 2% per doubling of w.
                            double e = std::min(timeit(), timeit()) * (w == 1 ? 1.0 : (w == 2 ? 1.02 : 1.04));
                            if (e < best) {
                                best = e;
                                bt.team = t;
                                bt.waves_per_block = w;
                                bt.vis_hash = v;
                                bt.spec = sp;
                            }
                        }
                    }
                }
            }
            s->tune[b] = bt;
            // Waves per query for batches that leave wave slots idle: a batch
            // of a quarter of the resident waves at the tuned shape, timed at
            // 1, 2 and 4 waves per query.
            s->tune[b].qwaves_small = 1;
            if (s->graph.K <= 64) {
                s->force_team = bt.team;
                s->force_wpb = bt.waves_per_block;
                s->vis_override = bt.vis_hash;
                s->force_spec = bt.spec;
                LaunchConfig lc{};
                lc.team = bt.team;
                lc.waves_per_block = bt.waves_per_block;
                lc.cap_max = round_cap(std::max(inner_k(*s, kTryK), s->ef));
                lc.vis_hash = bt.vis_hash;
                lc.spec = bt.spec;
                lc.qwaves = 1;
                const int64_t resident = (int64_t)std::max(1, max_blocks_per_cu(s->index(), lc)) *
                                         bt.waves_per_block * std::max(1, s->cu_count);
                const int32_t nsmall = (int32_t)std::max<int64_t>(1, std::min<int64_t>(nq, resident / 4));
                double best_q = 1e30;
                for (int32_t qw : {1, 2, 4}) {
                    s->force_qwaves = qw;
                    timeit_n(nsmall);
                    const double e = std::min(timeit_n(nsmall), timeit_n(nsmall));
                    if (e < best_q) {
                        best_q = e;
                        s->tune[b].qwaves_small = qw;
                    }
                }
                s->force_qwaves = 1;
            }
        }
        s->tune[glass_hip_searcher::kBuckets - 1] = s->tune[glass_hip_searcher::kBuckets - 2];
        s->ef = saved_ef;
        s->vis_override = saved_vo;
        s->force_team = saved_ft;
        s->force_wpb = saved_fw;
        s->force_spec = saved_fs;
        s->force_qwaves = saved_fq;
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

double glass_hip_searcher_last_avg_dist_cmps(const glass_hip_searcher *s) { return s ? s->last_avg_dist_cmps : 0.0; }

void glass_hip_searcher_free(glass_hip_searcher *s) { delete s; }

int glass_hip_searcher_set_refine_factor(glass_hip_searcher *s, float factor) {
    try {
        if (!s || !(factor >= 0.0f) || factor > 64.0f) return fail(GLASS_HIP_EINVAL, "refine factor must be in [0, 64]");
        std::lock_guard<std::mutex> lk(s->mu);
        s->refine_override = factor;
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_searcher_set_launch(glass_hip_searcher *s, int32_t team, int32_t waves_per_block, int32_t blocks) {
    try {
        if (!s) return fail(GLASS_HIP_EINVAL, "null searcher");
        if (waves_per_block < 0 || waves_per_block > 4 || blocks < 0) return fail(GLASS_HIP_EINVAL, "bad launch config");
        if (team && s->has_data && !team_valid(s->index(), team))
            return fail(GLASS_HIP_EINVAL, "team %d invalid for this layout", team);
        std::lock_guard<std::mutex> lk(s->mu);
        s->force_team = team;
        s->force_wpb = waves_per_block;
        s->blocks = blocks;
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_searcher_set_query_waves(glass_hip_searcher *s, int32_t waves) {
    try {
        if (!s) return fail(GLASS_HIP_EINVAL, "null searcher");
        if (waves != 0 && waves != 1 && waves != 2 && waves != 4)
            return fail(GLASS_HIP_EINVAL, "waves per query must be 0 (tuned), 1, 2 or 4");
        std::lock_guard<std::mutex> lk(s->mu);
        s->force_qwaves = waves;
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_searcher_get_query_waves(const glass_hip_searcher *s, int32_t *last, int32_t *tuned) {
    if (!s) return fail(GLASS_HIP_EINVAL, "null searcher");
    if (last) *last = s->last_launch.qwaves > 1 ? s->last_launch.qwaves : 1;
    if (tuned) *tuned = s->tune[glass_hip_searcher::bucket(s->ef)].qwaves_small;
    return GLASS_HIP_OK;
}

int glass_hip_searcher_set_gather(glass_hip_searcher *s, int32_t mode) {
    try {
        if (!s || mode < -1 || mode > 1) return fail(GLASS_HIP_EINVAL, "gather mode must be -1, 0 or 1");
        std::lock_guard<std::mutex> lk(s->mu);
        s->force_spec = mode;
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_searcher_set_visited_capacity(glass_hip_searcher *s, int32_t entries) {
    try {
        if (!s) return fail(GLASS_HIP_EINVAL, "null searcher");
        if (entries != 0 && (entries < 256 || entries > 16384 || (entries & (entries - 1))))
            return fail(GLASS_HIP_EINVAL, "visited table size must be 0 or a power of two in [256, 16384]");
        std::lock_guard<std::mutex> lk(s->mu);
        s->vis_override = entries;
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_searcher_get_launch(const glass_hip_searcher *s, int32_t *team, int32_t *waves_per_block,
                                  int32_t *blocks, int32_t *slots) {
    try {
        if (!s) return fail(GLASS_HIP_EINVAL, "null searcher");
        const LaunchConfig &ll = s->last_launch;
        if (ll.team) {
            // the shape the last launch ran with: a batch with per-query ef
            // takes the tuning of its largest ef, not of the current one
            if (team) *team = ll.team;
            if (waves_per_block) *waves_per_block = ll.waves_per_block;
            if (blocks) *blocks = ll.blocks;
            if (slots) *slots = ll.vis_hash;
            return GLASS_HIP_OK;
        }
        // no launch yet: the shape a batch at the current ef would use (as run_batch picks it)
        const glass_hip_searcher::Tune &tn = s->tune[glass_hip_searcher::bucket(s->ef)];
        if (team) *team = s->force_team ? s->force_team : tn.team;
        if (waves_per_block) *waves_per_block = s->force_wpb ? s->force_wpb : tn.waves_per_block;
        if (blocks) *blocks = 0;
        if (slots) *slots = s->vis_override ? s->vis_override : (tn.vis_hash ? tn.vis_hash : glass_hip_searcher::auto_vis(s->ef));
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_searcher_kernel_info(const glass_hip_searcher *s, int32_t *vgprs, int32_t *lds_per_block,
                                   int32_t *blocks_per_cu) {
    try {
        if (!s) return fail(GLASS_HIP_EINVAL, "null searcher");
        if (!s->has_data || !s->last_launch.team) return fail(GLASS_HIP_ESTATE, "no launch yet");
        const LaunchConfig &lc = s->last_launch;
        if (vgprs) *vgprs = kernel_vgprs(s->index(), lc);
        if (lds_per_block) *lds_per_block = (int32_t)lds_bytes_per_block(s->index(), lc);
        if (blocks_per_cu) *blocks_per_cu = max_blocks_per_cu(s->index(), lc);
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_device_info(char *name, int32_t name_len, int32_t *cu_count, int64_t *lds_per_cu, int64_t *hbm_bytes,
                          int32_t *clock_khz) {
    try {
        int dev = 0;
        HIP_TRY(hipGetDevice(&dev));
        hipDeviceProp_t p;
        HIP_TRY(hipGetDeviceProperties(&p, dev));
        if (name && name_len > 0) {
            std::snprintf(name, (size_t)name_len, "%s (%s)", p.name, p.gcnArchName);
        }
        if (cu_count) *cu_count = p.multiProcessorCount;
        if (lds_per_cu) *lds_per_cu = (int64_t)p.maxSharedMemoryPerMultiProcessor;
        if (hbm_bytes) *hbm_bytes = (int64_t)p.totalGlobalMem;
        if (clock_khz) *clock_khz = p.clockRate;
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_searcher_info(const glass_hip_searcher *s, int32_t *code_size, int32_t *rerank_row_bytes,
                            int32_t *degree, int32_t *n, float *refine_factor) {
    try {
        if (!s) return fail(GLASS_HIP_EINVAL, "null searcher");
        if (code_size) *code_size = s->code_size;
        if (rerank_row_bytes)
            *rerank_row_bytes = s->refine != QK_NONE ? s->r_code_size : 0;
        if (degree) *degree = s->graph.K;
        if (n) *n = s->graph.N;
        if (refine_factor) *refine_factor = s->refine != QK_NONE ? s->factor() : 1.0f;
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_searcher_get_codes(const glass_hip_searcher *s, uint8_t *codes) {
    try {
        if (!s || !codes) return fail(GLASS_HIP_EINVAL, "null argument");
        if (!s->has_data) return fail(GLASS_HIP_ESTATE, "no data");
        HIP_TRY(hipMemcpy(codes, s->d_codes.p, (size_t)s->nb * s->code_size, hipMemcpyDeviceToHost));
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_searcher_get_calibration(const glass_hip_searcher *s, float *params, int32_t *n) {
    try {
        if (!s || !n) return fail(GLASS_HIP_EINVAL, "null argument");
        if (s->quant == QK_PQ8) {
            *n = (int32_t)s->pq_codebook.size();
            if (params) std::copy(s->pq_codebook.begin(), s->pq_codebook.end(), params);
        } else if (s->quant == QK_SQ8 || s->quant == QK_SQ6 || s->quant == QK_SQ4) {
            *n = 2 * s->dim_align;
            if (params) {
                std::copy(s->sq8_mins.begin(), s->sq8_mins.end(), params);
                std::copy(s->sq8_difs.begin(), s->sq8_difs.end(), params + s->dim_align);
            }
        } else if (s->quant == QK_SQ8U && s->metric == MK_IP) {
            *n = 1;
            if (params) params[0] = s->qmax;
        } else if (s->quant == QK_SQ8U || s->quant == QK_SQ4U || s->quant == QK_SQ2U) {
            *n = 2;
            if (params) {
                params[0] = s->qmin;
                params[1] = s->qdif;
            }
        } else {
            *n = 0;
        }
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_quant_train(const char *quantizer, const char *metric, const float *data, int32_t n, int32_t dim,
                          float *params, int32_t *nparams) {
    try {
        if (!quantizer || !nparams || dim <= 0) return fail(GLASS_HIP_EINVAL, "bad quantizer arguments");
        const int32_t q = parse_quant(quantizer);
        if (q == QK_NONE) return fail(GLASS_HIP_EINVAL, "unknown quantizer %s", quantizer);
        if (q == QK_PQ8 && dim % 4 != 0) return fail(GLASS_HIP_EINVAL, "PQ8 needs a dimension divisible by 4");
        glass_hip_searcher t;
        t.quant = q;
        t.metric = (metric && std::string(metric) == "IP") ? MK_IP : MK_L2;
        t.dim = dim;
        quant_layout(q, dim, &t.dim_align, &t.code_size);
        if (!params) {  // size query: the layout of glass_hip_searcher_get_calibration
            if (q == QK_PQ8) *nparams = (dim / 4) * 256 * 4;
            else if (q == QK_SQ8 || q == QK_SQ6 || q == QK_SQ4) *nparams = 2 * t.dim_align;
            else if (q == QK_SQ8U && t.metric == MK_IP) *nparams = 1;
            else if (q == QK_SQ8U || q == QK_SQ4U || q == QK_SQ2U) *nparams = 2;
            else *nparams = 0;
            return GLASS_HIP_OK;
        }
        if (!data || n <= 0) return fail(GLASS_HIP_EINVAL, "no training data");
        if (q == QK_PQ8) {
            pq::train(data, n, dim, t.pq_codebook);
        } else if (q != QK_FP32 && q != QK_FP16 && q != QK_BF16 && q != QK_E5M2) {
            if (int r = ensure_stream(t)) return r;
            DevBuf<float> raw;
            HIP_TRY(raw.alloc((size_t)n * dim));
            HIP_TRY(hipMemcpy(raw.p, data, sizeof(float) * (size_t)n * dim, hipMemcpyHostToDevice));
            if (int r = train_quant(t, raw.p, n)) return r;
        }
        return glass_hip_searcher_get_calibration(&t, params, nparams);
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_searcher_sync(glass_hip_searcher *s) {
    try {
        if (!s) return fail(GLASS_HIP_EINVAL, "null searcher");
        if (s->stream) HIP_TRY(hipStreamSynchronize(s->stream));
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Served form: continuous batching.
//
// Requests (one or more queries with their own ef and k) queue on the host.
// A dispatcher thread packs whatever is queued, up to max_batch queries, into
// a free staging slot, copies it to the device in one transfer and launches
// the persistent search kernel with per-query ef/k (SearchBatch semantics per
// query).  Each slot has its own stream, work counter and per-wave visited
// bitsets, so up to max_inflight batches run on the GPU at once and a new
// batch fills the compute units a finishing batch leaves.  The kernel writes
// every query's results straight into host-mapped memory as (id, distance)
// pairs, one 64-bit store each, over a sentinel the host wrote before the
// launch; a request completes as soon as none of its results holds the
// sentinel, not when the slowest query of its batch does, and no fence is
// needed on the device.  A batch is formed as soon as
// a slot is free and anything is queued.  Each query's result depends only on
// (query, ef, k, index), so a request gets the same ids and distances alone
// or in any mix.

namespace {

using Clock = std::chrono::steady_clock;

int64_t now_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now().time_since_epoch()).count();
}

enum ReqState : int32_t { RQ_QUEUED = 0, RQ_RUNNING = 1, RQ_DONE = 2, RQ_FAILED = 3 };

struct ServerReq {
    int64_t ticket = 0;
    int32_t nq = 0, ef = 0, k = 0;
    const float *q = nullptr;        // nq x dim rows (owned copy or caller's replay array)
    int32_t *ids = nullptr;          // nq x k results
    float *dists = nullptr;
    std::vector<float> own_q;        // submit(): copies owned by the request
    std::vector<int32_t> own_ids;
    std::vector<float> own_dists;
    int64_t arrival_ns = 0, done_ns = 0;
    // Written under the server lock, except that a request lying whole in one
    // batch is completed by that batch's worker without it (done_ns first,
    // then a release store of RQ_DONE; the worker touches the request no more).
    std::atomic<int32_t> state{RQ_QUEUED};
    int32_t err = GLASS_HIP_OK;      // RQ_FAILED by its own batch: the status and message
    std::string err_msg;
    int64_t *group_done = nullptr;   // replay(): completed-request count of its call (under the server lock)
    int32_t dispatched = 0;          // queries handed to batches so far (a request larger than a batch is split)
    int32_t parts_open = 0;          // dispatched parts not yet complete
};

struct ServerSlot {
    // pinned staging in: [ef | k | off | queries]; host-mapped coherent out:
    // (id, dist) pairs the kernel writes with one 64-bit store each
    uint8_t *h_in = nullptr;
    uint64_t *h_out = nullptr;
    uint64_t *d_out = nullptr;       // device address of h_out
    DevBuf<uint8_t> d_in;
    DevBuf<int32_t> counter;
    uint32_t *visited = nullptr;     // this slot's share of the server's visited bitsets
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    // A part is queries [q0, q0 + nq) of one request.  r is cleared once the
    // part completes: its request's owner may release the request from then on.
    struct Part {
        ServerReq *r;
        int32_t q0, nq;
        int32_t pos;                 // first query of the part in the batch
        int64_t res;                 // first result of the part in h_out
        int64_t checked;             // leading results of the part seen complete
        bool done;
        bool whole;                  // the part is its whole request
    };
    std::vector<Part> parts;
    std::vector<std::pair<int64_t *, int64_t>> groups;  // server_collect: completions per replay() call
    size_t first_open = 0;           // parts before it are complete
    int32_t pending = 0;             // parts of this batch not yet complete
    int32_t nq = 0;
    // The slot's worker packs, launches and collects its batches, so the
    // host work of batches in flight runs on as many threads.  phase (under
    // the server lock): 0 free, 1 filled by the dispatcher, -1 exit.
    std::thread worker;
    std::condition_variable cv;
    int32_t phase = 0;
};

}  // namespace

struct glass_hip_server {
    glass_hip_searcher *s = nullptr;
    int32_t max_batch = 0, max_k = 0, max_inflight = 0, dim = 0, slots_per_batch = 0;
    size_t in_meta = 0, in_bytes = 0, out_bytes = 0;
    std::vector<std::unique_ptr<ServerSlot>> slots;
    DevBuf<uint32_t> visited;
    std::thread worker;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ServerReq *> pending;
    std::atomic<int64_t> pending_n{0};  // pending.size(), readable without the lock
    std::unordered_map<int64_t, std::unique_ptr<ServerReq>> owned;  // submit() requests by ticket
    int64_t next_ticket = 1;
    bool stop = false;
    int err_code = GLASS_HIP_OK;
    std::string err_msg;
    int64_t n_batches = 0;                 // under the lock
    // Completed queries and requests: added to before any of them is
    // published, so a caller that has seen its request complete also sees it
    // counted by glass_hip_server_stats.
    std::atomic<int64_t> n_queries{0}, n_requests{0};

    ~glass_hip_server() {
        // workers still running (creation failed after they started): stop them
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
            for (auto &sl : slots) {
                sl->phase = -1;
                sl->cv.notify_all();
            }
        }
        cv_work.notify_all();
        if (worker.joinable()) worker.join();
        for (auto &sl : slots)
            if (sl->worker.joinable()) sl->worker.join();
        for (auto &sl : slots) {
            if (sl->h_in) (void)hipHostFree(sl->h_in);
            if (sl->h_out) (void)hipHostFree(sl->h_out);
            if (sl->done) (void)hipEventDestroy(sl->done);
            if (sl->stream) (void)hipStreamDestroy(sl->stream);
        }
    }
};

namespace {

// A device error ends the server: the queued requests fail here, and each
// slot's worker fails the requests of its own batch when it sees err_code.
void server_fail_all(glass_hip_server &sv, int code, const std::string &msg) {
    std::lock_guard<std::mutex> lk(sv.mu);
    if (!sv.err_code) {
        sv.err_code = code;
        sv.err_msg = msg;
    }
    for (ServerReq *r : sv.pending) r->state = RQ_FAILED;
    sv.pending.clear();
    sv.pending_n.store(0);
    sv.cv_done.notify_all();
    sv.cv_work.notify_all();
}

// The requests of a slot's batch after a device error (under the server
// lock).  Completed parts no longer name their request (server_collect
// clears them): its owner may already have released it.
void server_fail_slot_parts(glass_hip_server &sv, ServerSlot &sl) {
    for (auto &pt : sl.parts)
        if (pt.r && pt.r->state != RQ_DONE) pt.r->state = RQ_FAILED;
    sl.parts.clear();
    sv.cv_done.notify_all();
}

// Publishes a request whose every query has been dispatched and whose every
// part has completed or failed (under the server lock): from here on its
// owner may take or release it.
void server_finish(glass_hip_server &sv, ServerReq *r, int64_t t) {
    if (!r->err) sv.n_requests.fetch_add(1, std::memory_order_relaxed);
    if (r->group_done) *r->group_done += 1;
    r->done_ns = t;
    r->state = r->err ? RQ_FAILED : RQ_DONE;
}

// A batch that could not be launched fails its own requests only; the
// dispatcher and every other request carry on.  A split request fails as a
// whole: its queries not yet dispatched leave the queue, and it is published
// once its parts in other batches have completed (they still write into it).
void server_fail_batch(glass_hip_server &sv, ServerSlot &sl, int code, const std::string &msg) {
    (void)hipStreamSynchronize(sl.stream);
    std::lock_guard<std::mutex> lk(sv.mu);
    for (auto &pt : sl.parts) {
        ServerReq *r = pt.r;
        if (!r) continue;
        pt.r = nullptr;
        r->parts_open -= 1;
        if (!r->err) {
            r->err = code;
            r->err_msg = msg;
        }
        if (r->dispatched < r->nq) {
            auto it = std::find(sv.pending.begin(), sv.pending.end(), r);
            if (it != sv.pending.end()) sv.pending.erase(it);
            sv.pending_n.store((int64_t)sv.pending.size());
            r->dispatched = r->nq;
        }
        if (r->parts_open == 0 && r->state == RQ_RUNNING) server_finish(sv, r, now_ns());
    }
    sl.parts.clear();
    sv.cv_done.notify_all();
}

// Packs the requests of `sl` and enqueues copy-in and search on the slot's stream.
int server_launch(glass_hip_server &sv, ServerSlot &sl) {
    glass_hip_searcher &s = *sv.s;
    const int32_t mb = sv.max_batch;
    int32_t *h_ef = (int32_t *)sl.h_in;
    int32_t *h_k = h_ef + mb;
    int64_t *h_off = (int64_t *)(sl.h_in + 8 * (size_t)mb);
    float *h_q = (float *)(sl.h_in + sv.in_meta);
    int32_t pos = 0, max_cap = 1, max_ef = 1;
    int64_t res = 0;
    const float *run_src = nullptr;  // rows of consecutive parts that are contiguous are copied at once
    int32_t run_pos = 0, run_len = 0;
    auto flush = [&]() {
        if (run_len) std::memcpy(h_q + (size_t)run_pos * sv.dim, run_src, sizeof(float) * (size_t)run_len * sv.dim);
        run_len = 0;
    };
    for (auto &pt : sl.parts) {
        ServerReq *r = pt.r;
        pt.pos = pos;
        pt.res = res;
        pt.checked = 0;
        pt.done = false;
        const float *src = r->q + (size_t)pt.q0 * sv.dim;
        if (run_len && src == run_src + (size_t)run_len * sv.dim) {
            run_len += pt.nq;
        } else {
            flush();
            run_src = src;
            run_pos = pos;
            run_len = pt.nq;
        }
        for (int32_t j = 0; j < pt.nq; ++j) {
            h_ef[pos + j] = r->ef;
            h_k[pos + j] = r->k;
            h_off[pos + j] = res + (int64_t)j * r->k;
        }
        pos += pt.nq;
        res += (int64_t)pt.nq * r->k;
        max_cap = std::max(max_cap, std::max(inner_k(s, r->k), r->ef));
        max_ef = std::max(max_ef, r->ef);
    }
    flush();
    sl.pending = (int32_t)sl.parts.size();
    sl.first_open = 0;
    sl.nq = pos;
    // sentinel id 0xFEFEFEFE in every result this batch writes (ids are -1 or
    // in [0, N)); no kernel of this slot is running
    std::memset(sl.h_out, 0xFE, 8 * (size_t)res);
    std::lock_guard<std::mutex> lk(s.mu);
    const size_t in_used = sv.in_meta + sizeof(float) * (size_t)pos * sv.dim;
    HIP_TRY(hipMemcpyAsync(sl.d_in.p, sl.h_in, in_used, hipMemcpyHostToDevice, sl.stream));
    const int32_t *d_ef = (const int32_t *)sl.d_in.p;
    const int32_t *d_k = d_ef + mb;
    const int64_t *d_off = (const int64_t *)(sl.d_in.p + 8 * (size_t)mb);
    const float *d_q = (const float *)(sl.d_in.p + sv.in_meta);
    BatchState bs;
    bs.visited = sl.visited;
    bs.max_slots = sv.slots_per_batch;
    bs.counter = sl.counter.p;
    bs.out_pairs = sl.d_out;
    if (int r = run_batch(s, d_q, pos, d_ef, d_k, 0, 0, max_cap, max_ef, MODE_BATCH, d_off, nullptr, nullptr, nullptr,
                          sl.stream, &bs))
        return r;
    HIP_TRY(hipEventRecord(sl.done, sl.stream));
    return GLASS_HIP_OK;
}

constexpr uint32_t kSentinelId = 0xFEFEFEFEu;

// Completes the requests of a running slot whose results have all arrived
// (no pair still holds the sentinel id; every request when `all` is set: the
// batch has finished).  Returns true when every request of the slot is complete.
bool server_collect(glass_hip_server &sv, ServerSlot &sl, bool all) {
    // Copy the results of every part whose pairs have all arrived into its
    // request (no lock: a request is not published until server_finish),
    // then publish the parts under one lock.
    // Queries start in batch order, so completions cluster after the first
    // open part: a poll looks at no more than kScanWindow open parts past
    // it (the batch's final collect, `all`, takes every part).
    constexpr int32_t kScanWindow = 1024;
    int64_t t = 0;
    size_t first_done = sl.parts.size();
    while (sl.first_open < sl.parts.size() && sl.parts[sl.first_open].done) ++sl.first_open;
    int32_t open_seen = 0;
    size_t scan_end = sl.parts.size();
    for (size_t i = sl.first_open; i < sl.parts.size(); ++i) {
        auto &pt = sl.parts[i];
        if (pt.done) continue;
        if (!all && ++open_seen > kScanWindow) {
            scan_end = i;
            break;
        }
        ServerReq *r = pt.r;
        const int64_t n = (int64_t)pt.nq * r->k;
        const uint64_t *res = sl.h_out + pt.res;
        if (!all) {
            // results arrive roughly in order: resume where the last scan stopped
            while (pt.checked < n && (uint32_t)__atomic_load_n(&res[pt.checked], __ATOMIC_ACQUIRE) != kSentinelId)
                ++pt.checked;
            if (pt.checked < n) continue;
        }
        if (!t) t = now_ns();
        int32_t *ids = r->ids + (int64_t)pt.q0 * r->k;
        float *dists = r->dists + (int64_t)pt.q0 * r->k;
        for (int64_t j = 0; j < n; ++j) {
            const uint64_t pr = __atomic_load_n(&res[j], __ATOMIC_ACQUIRE);
            ids[j] = (int32_t)(uint32_t)pr;
            uint32_t db = (uint32_t)(pr >> 32);
            std::memcpy(&dists[j], &db, 4);
        }
        pt.done = true;
        pt.checked = -1;  // done and not yet published
        if (first_done == sl.parts.size()) first_done = i;
    }
    if (first_done == sl.parts.size()) return sl.pending == 0;
    // A part that is its whole request completes here, without the lock: its
    // owner reads the request's state (an atomic) and may take or release it
    // right after the store.  Parts of split requests, the counters and the
    // wake-up of waiters go under one lock.
    int64_t nq_done = 0, whole_done = 0;
    bool split = false;
    for (size_t i = first_done; i < scan_end; ++i) {
        const auto &pt = sl.parts[i];
        if (pt.checked != -1) continue;
        nq_done += pt.nq;
        if (pt.whole)
            whole_done += 1;
        else
            split = true;
    }
    // counted before any of them is published (the release stores below
    // order these adds before a caller's view of its completed request)
    sv.n_queries.fetch_add(nq_done, std::memory_order_relaxed);
    sv.n_requests.fetch_add(whole_done, std::memory_order_relaxed);
    sl.groups.clear();
    for (size_t i = first_done; i < scan_end; ++i) {
        auto &pt = sl.parts[i];
        if (pt.checked != -1) continue;
        sl.pending -= 1;
        if (!pt.whole) continue;
        ServerReq *r = pt.r;
        pt.checked = 0;
        pt.r = nullptr;
        if (int64_t *g = r->group_done) {
            if (sl.groups.empty() || sl.groups.back().first != g) sl.groups.emplace_back(g, 0);
            sl.groups.back().second += 1;
        }
        r->done_ns = t;
        r->state.store(RQ_DONE, std::memory_order_release);
    }
    {
        std::lock_guard<std::mutex> lk(sv.mu);
        for (auto &g : sl.groups) *g.first += g.second;
        for (size_t i = first_done; split && i < scan_end; ++i) {
            auto &pt = sl.parts[i];
            if (pt.checked != -1) continue;
            ServerReq *r = pt.r;
            pt.checked = 0;
            pt.r = nullptr;  // the request's owner may release it once it completes
            r->parts_open -= 1;
            if (r->parts_open == 0 && r->dispatched == r->nq && r->state == RQ_RUNNING) server_finish(sv, r, t);
        }
        sv.cv_done.notify_all();
    }
    return sl.pending == 0;
}

// One slot's worker: waits for the dispatcher to fill the slot, packs and
// launches the batch, then completes its requests as their results arrive.
void slot_loop(glass_hip_server *psv, ServerSlot *psl) {
    glass_hip_server &sv = *psv;
    ServerSlot &sl = *psl;
    const bool dev_ok = hipSetDevice(sv.s->device) == hipSuccess;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(sv.mu);
            sl.cv.wait(lk, [&] { return sl.phase != 0; });
            if (sl.phase < 0) return;
        }
        bool launched = false;
        if (!dev_ok) {
            server_fail_all(sv, GLASS_HIP_EDEVICE, "server thread: hipSetDevice failed");
        } else if (int rc = server_launch(sv, sl)) {
            server_fail_batch(sv, sl, rc, g_last_error);
        } else {
            launched = true;
            for (;;) {
                const hipError_t e = hipEventQuery(sl.done);
                if (e != hipSuccess && e != hipErrorNotReady) {
                    server_fail_all(sv, GLASS_HIP_EDEVICE, std::string("search batch failed: ") + hipGetErrorString(e));
                    (void)hipStreamSynchronize(sl.stream);
                    break;
                }
                const bool finished = e == hipSuccess;
                const bool progressed = sl.pending && server_collect(sv, sl, finished);
                if (finished) {
                    if (sl.pending) server_collect(sv, sl, true);
                    break;
                }
                if (!progressed) std::this_thread::yield();
            }
        }
        std::lock_guard<std::mutex> lk(sv.mu);
        if (sv.err_code) server_fail_slot_parts(sv, sl);
        if (launched) sv.n_batches += 1;
        sl.parts.clear();
        if (sl.phase > 0) sl.phase = 0;
        sv.cv_work.notify_all();
    }
}

// The dispatcher: fills a free slot with whatever is queued (up to
// max_batch queries; a larger request is split, its remaining queries
// staying at the queue's head) and hands it to the slot's worker.
void server_loop(glass_hip_server *psv) {
    glass_hip_server &sv = *psv;
    std::unique_lock<std::mutex> lk(sv.mu);
    for (;;) {
        ServerSlot *free_sl = nullptr;
        sv.cv_work.wait(lk, [&] {
            if (sv.err_code || (sv.stop && sv.pending.empty())) return true;
            if (sv.pending.empty()) return false;
            for (auto &p : sv.slots)
                if (p->phase == 0) return true;
            return false;
        });
        if (sv.err_code || (sv.stop && sv.pending.empty())) break;
        for (auto &p : sv.slots)
            if (p->phase == 0) {
                free_sl = p.get();
                break;
            }
        ServerSlot &sl = *free_sl;
        int32_t total = 0;
        while (!sv.pending.empty() && total < sv.max_batch) {
            ServerReq *r = sv.pending.front();
            const int32_t take = std::min(r->nq - r->dispatched, sv.max_batch - total);
            sl.parts.push_back(ServerSlot::Part{r, r->dispatched, take, 0, 0, 0, false, take == r->nq});
            r->dispatched += take;
            r->parts_open += 1;
            r->state.store(RQ_RUNNING, std::memory_order_relaxed);  // readers hold the lock
            total += take;
            if (r->dispatched == r->nq) sv.pending.pop_front();
        }
        sv.pending_n.store((int64_t)sv.pending.size());
        sl.phase = 1;
        sl.cv.notify_one();
    }
    // every worker finishes its batch, then exits
    for (auto &p : sv.slots) {
        sv.cv_work.wait(lk, [&] { return p->phase <= 0; });
        p->phase = -1;
        p->cv.notify_one();
    }
}

int server_enqueue(glass_hip_server &sv, ServerReq *r) {
    if (r->nq <= 0) return fail(GLASS_HIP_EINVAL, "request of %d queries", r->nq);
    if (r->k <= 0 || r->k > sv.max_k || r->ef <= 0)
        return fail(GLASS_HIP_EINVAL, "bad ef %d / k %d (max_k %d)", r->ef, r->k, sv.max_k);
    if (int rc = check_cap(*sv.s, std::max(inner_k(*sv.s, r->k), r->ef))) return rc;
    return GLASS_HIP_OK;
}

}  // namespace

extern "C" {

int glass_hip_server_create(glass_hip_searcher *s, int32_t max_batch, int32_t max_k, int32_t max_inflight,
                            glass_hip_server **out) {
    try {
        if (!s || !out || max_batch <= 0 || max_k <= 0 || max_inflight <= 0 || max_inflight > 16)
            return fail(GLASS_HIP_EINVAL, "bad server arguments");
        if (!s->has_data) return fail(GLASS_HIP_ESTATE, "server before set_data");
        auto sv = std::make_unique<glass_hip_server>();
        sv->s = s;
        sv->max_batch = max_batch;
        sv->max_k = max_k;
        sv->max_inflight = max_inflight;
        sv->dim = s->dim;
        sv->in_meta = 16 * (size_t)max_batch;
        sv->in_bytes = sv->in_meta + sizeof(float) * (size_t)max_batch * s->dim;
        sv->out_bytes = 8 * (size_t)max_batch * max_k;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (int r = ensure_stream(*s)) return r;
            // waves one batch can have resident: the most any launch shape allows
            LaunchConfig lc{};
            int32_t most = 0;
            for (int32_t team : {4, 8, 16}) {
                if (!team_valid(s->index(), team)) continue;
                lc.team = team;
                for (int32_t w = 1; w <= 4; ++w) {
                    // both kernel paths (gather on / off, small / large pools)
                    for (int32_t sp = 0; sp <= 1; ++sp) {
                        for (int32_t cap : {64, 129}) {
                            lc.waves_per_block = w;
                            lc.cap_max = cap;
                            lc.vis_hash = 256;
                            lc.spec = sp;
                            most = std::max(most,
                                            std::max(1, max_blocks_per_cu(s->index(), lc)) * std::max(1, s->cu_count) * w);
                        }
                    }
                }
            }
            if (most == 0) most = 32 * std::max(1, s->cu_count);  // 8 waves per SIMD, 4 SIMDs per CU
            sv->slots_per_batch = most;
            s->vis_stride = slot_stride(*s);
            // Batches in flight are admitted by device memory: each needs its
            // waves' slot areas (visited bitsets, PQ8 tables) and its staging;
            // at most half of the free memory goes to them.
            size_t free_b = 0, total_b = 0;
            HIP_TRY(hipMemGetInfo(&free_b, &total_b));
            const size_t per_batch = (size_t)most * (size_t)s->vis_stride * 4 + sv->in_bytes;
            const size_t fit = per_batch ? (free_b / 2) / per_batch : (size_t)max_inflight;
            if (fit < 1) return fail(GLASS_HIP_ENOMEM, "no device memory for one batch in flight (%zu bytes)", per_batch);
            max_inflight = (int32_t)std::min<size_t>((size_t)max_inflight, fit);
            sv->max_inflight = max_inflight;
            // the server's own slot areas, zeroed (bitsets are all zero between queries)
            const size_t words = (size_t)max_inflight * (size_t)most * (size_t)s->vis_stride;
            HIP_TRY(sv->visited.alloc(words));
            HIP_TRY(launch_fill_i32((int32_t *)sv->visited.p, (int64_t)words, 0, s->stream));
            HIP_TRY(hipStreamSynchronize(s->stream));
        }
        for (int32_t i = 0; i < max_inflight; ++i) {
            sv->slots.push_back(std::make_unique<ServerSlot>());
            ServerSlot &sl = *sv->slots.back();
            HIP_TRY(hipHostMalloc((void **)&sl.h_in, sv->in_bytes, hipHostMallocDefault));
            HIP_TRY(hipHostMalloc((void **)&sl.h_out, sv->out_bytes, hipHostMallocMapped | hipHostMallocCoherent));
            HIP_TRY(hipHostGetDevicePointer((void **)&sl.d_out, sl.h_out, 0));
            HIP_TRY(sl.d_in.alloc(sv->in_bytes));
            HIP_TRY(sl.counter.alloc(1));
            sl.visited = sv->visited.p + (size_t)i * (size_t)sv->slots_per_batch * (size_t)s->vis_stride;
            HIP_TRY(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
            HIP_TRY(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
        }
        glass_hip_server *p = sv.get();
        for (auto &sl : sv->slots) sl->worker = std::thread(slot_loop, p, sl.get());
        sv->worker = std::thread(server_loop, p);
        *out = sv.release();
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_server_submit(glass_hip_server *sv, const float *qs, int32_t nq, int32_t ef, int32_t k,
                            int64_t *ticket) {
    try {
        if (!sv || !qs || !ticket) return fail(GLASS_HIP_EINVAL, "null argument");
        auto r = std::make_unique<ServerReq>();
        r->nq = nq;
        r->ef = ef;
        r->k = k;
        if (int rc = server_enqueue(*sv, r.get())) return rc;
        r->own_q.assign(qs, qs + (size_t)nq * sv->dim);
        r->own_ids.assign((size_t)nq * k, -1);
        r->own_dists.assign((size_t)nq * k, 0.0f);
        r->q = r->own_q.data();
        r->ids = r->own_ids.data();
        r->dists = r->own_dists.data();
        r->arrival_ns = now_ns();
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            if (sv->err_code) return fail(sv->err_code, "%s", sv->err_msg.c_str());
            if (sv->stop) return fail(GLASS_HIP_ESTATE, "server is stopping");
            r->ticket = sv->next_ticket++;
            *ticket = r->ticket;
            sv->pending.push_back(r.get());
            sv->pending_n.store((int64_t)sv->pending.size());
            sv->owned.emplace(r->ticket, std::move(r));
        }
        sv->cv_work.notify_one();
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

static int server_take(glass_hip_server *sv, std::unique_lock<std::mutex> &lk, int64_t ticket, int32_t *dst,
                       float *scores, bool wait, int32_t *done) {
    auto it = sv->owned.find(ticket);
    if (it == sv->owned.end()) return fail(GLASS_HIP_EINVAL, "unknown ticket %lld", (long long)ticket);
    ServerReq *r = it->second.get();
    if (wait) sv->cv_done.wait(lk, [&] { return r->state == RQ_DONE || r->state == RQ_FAILED; });
    if (r->state == RQ_FAILED) {
        const int code = r->err ? r->err : (sv->err_code ? sv->err_code : GLASS_HIP_EDEVICE);
        const std::string msg = r->err ? r->err_msg : sv->err_msg;
        sv->owned.erase(it);
        return fail(code, "%s", msg.c_str());
    }
    if (r->state != RQ_DONE) {
        if (done) *done = 0;
        return GLASS_HIP_OK;
    }
    if (dst) std::copy(r->own_ids.begin(), r->own_ids.end(), dst);
    if (scores) std::copy(r->own_dists.begin(), r->own_dists.end(), scores);
    if (done) *done = 1;
    sv->owned.erase(it);
    return GLASS_HIP_OK;
}

int glass_hip_server_poll(glass_hip_server *sv, int64_t ticket, int32_t *dst, float *scores, int32_t *done) {
    try {
        if (!sv || !done) return fail(GLASS_HIP_EINVAL, "null argument");
        std::unique_lock<std::mutex> lk(sv->mu);
        return server_take(sv, lk, ticket, dst, scores, false, done);
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_server_wait(glass_hip_server *sv, int64_t ticket, int32_t *dst, float *scores) {
    try {
        if (!sv) return fail(GLASS_HIP_EINVAL, "null argument");
        std::unique_lock<std::mutex> lk(sv->mu);
        return server_take(sv, lk, ticket, dst, scores, true, nullptr);
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_server_cancel(glass_hip_server *sv, int64_t ticket) {
    try {
        if (!sv) return fail(GLASS_HIP_EINVAL, "null argument");
        std::lock_guard<std::mutex> lk(sv->mu);
        auto it = sv->owned.find(ticket);
        if (it == sv->owned.end()) return fail(GLASS_HIP_EINVAL, "unknown ticket %lld", (long long)ticket);
        ServerReq *r = it->second.get();
        if (r->state != RQ_QUEUED) return fail(GLASS_HIP_ESTATE, "request %lld already dispatched", (long long)ticket);
        sv->pending.erase(std::find(sv->pending.begin(), sv->pending.end(), r));
        sv->pending_n.store((int64_t)sv->pending.size());
        sv->owned.erase(it);
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_server_replay(glass_hip_server *sv, const float *qs, int64_t nreq, const int32_t *nq, const int32_t *ef,
                            const int32_t *k, const int64_t *arrival_ns, int64_t *latency_ns, int32_t *dst,
                            float *scores) {
    try {
        if (!sv || nreq < 0 || (nreq && (!qs || !nq || !ef || !k || !arrival_ns || !latency_ns || !dst || !scores)))
            return fail(GLASS_HIP_EINVAL, "bad replay arguments");
        std::vector<ServerReq> reqs((size_t)nreq);
        int64_t done_count = 0;  // guarded by sv->mu
        int64_t qoff = 0, roff = 0;
        for (int64_t i = 0; i < nreq; ++i) {
            ServerReq &r = reqs[(size_t)i];
            r.nq = nq[i];
            r.ef = ef[i];
            r.k = k[i];
            if (int rc = server_enqueue(*sv, &r)) return rc;
            if (i && arrival_ns[i] < arrival_ns[i - 1]) return fail(GLASS_HIP_EINVAL, "arrivals must be sorted");
            r.q = qs + qoff * sv->dim;
            r.group_done = &done_count;
            r.ids = dst + roff;
            r.dists = scores + roff;
            qoff += r.nq;
            roff += (int64_t)r.nq * r.k;
        }
        const int64_t t0 = now_ns();
        int64_t next = 0;
        std::vector<ServerReq *> burst;
        burst.reserve(1024);
        while (next < nreq) {
            const int64_t t = now_ns() - t0;
            if (arrival_ns[next] > t) {
                if (arrival_ns[next] - t > 200000) std::this_thread::sleep_for(std::chrono::microseconds(50));
                continue;
            }
            // at most 1,024 at a time, stamped outside the lock and spliced
            // into the queue in one insert: the lock is held about a
            // microsecond per burst, so a backlog being pushed does not keep
            // the dispatcher and the workers from it
            burst.clear();
            for (int c = 0; c < 1024 && next < nreq && arrival_ns[next] <= t; ++c) {
                ServerReq &r = reqs[(size_t)next];
                r.arrival_ns = t0 + arrival_ns[next];
                burst.push_back(&r);
                ++next;
            }
            {
                std::lock_guard<std::mutex> lk(sv->mu);
                if (sv->err_code) break;
                sv->pending.insert(sv->pending.end(), burst.begin(), burst.end());
                sv->pending_n.store((int64_t)sv->pending.size());
            }
            sv->cv_work.notify_one();
        }
        std::unique_lock<std::mutex> lk(sv->mu);
        sv->cv_done.wait(lk, [&] { return sv->err_code || done_count == nreq; });
        if (sv->err_code) {
            // requests of this call may still sit in the queue: drop them
            for (auto it = sv->pending.begin(); it != sv->pending.end();) {
                ServerReq *p = *it;
                it = (p >= reqs.data() && p < reqs.data() + reqs.size()) ? sv->pending.erase(it) : it + 1;
            }
            sv->pending_n.store((int64_t)sv->pending.size());
            return fail(sv->err_code, "%s", sv->err_msg.c_str());
        }
        for (const ServerReq &r : reqs)
            if (r.state == RQ_FAILED) return fail(r.err ? r.err : GLASS_HIP_EDEVICE, "%s", r.err_msg.c_str());
        for (int64_t i = 0; i < nreq; ++i) latency_ns[i] = reqs[(size_t)i].done_ns - reqs[(size_t)i].arrival_ns;
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_server_info(const glass_hip_server *sv, int32_t *max_inflight, int32_t *max_batch, int32_t *max_k) {
    if (!sv) return fail(GLASS_HIP_EINVAL, "null argument");
    if (max_inflight) *max_inflight = sv->max_inflight;
    if (max_batch) *max_batch = sv->max_batch;
    if (max_k) *max_k = sv->max_k;
    return GLASS_HIP_OK;
}

int glass_hip_server_stats(glass_hip_server *sv, int64_t *batches, int64_t *queries, int64_t *requests) {
    try {
        if (!sv) return fail(GLASS_HIP_EINVAL, "null argument");
        std::lock_guard<std::mutex> lk(sv->mu);
        if (batches) *batches = sv->n_batches;
        if (queries) *queries = sv->n_queries.load();
        if (requests) *requests = sv->n_requests.load();
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

void glass_hip_server_free(glass_hip_server *sv) {
    if (!sv) return;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
        for (ServerReq *r : sv->pending) r->state = RQ_FAILED;
        sv->pending.clear();
        sv->pending_n.store(0);
        sv->cv_done.notify_all();
    }
    sv->cv_work.notify_all();
    if (sv->worker.joinable()) sv->worker.join();  // it lets every slot's worker finish its batch and exit
    for (auto &sl : sv->slots)
        if (sl->worker.joinable()) sl->worker.join();
    for (auto &sl : sv->slots)
        if (sl->stream) (void)hipStreamSynchronize(sl->stream);
    delete sv;
}

}  // extern "C"
