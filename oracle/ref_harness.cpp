// Test infrastructure only: a thin extern "C" harness around the reference
// pyglass headers, compiled from /root/reference by oracle/Makefile into
// oracle/_ref/.  It is the checker's checker: golden vectors under
// tests/golden/ are written through it, the C restatement in oracle/ is
// validated against it, and bench.py may time it as the host-CPU baseline.
// Nothing in pyglass_amd/ links, loads or calls it.
//
// Every function here only forwards to the reference API:
//   glass::Graph::load / save            glass/graph.hpp:106-153
//   glass::create_searcher               glass/searcher/graph_searcher.hpp:256-391
//   GraphSearcherBase virtuals           glass/searcher/searcher_base.hpp:7-19
//   glass::create_hnsw / create_nsg      glass/hnsw/hnsw.hpp:74-133, glass/nsg/nsg.hpp:411-434
//   Quantizer train/add/get_computer     glass/quant/*.hpp
#include <omp.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <string>

#include "glass/builder.hpp"
#include "glass/graph.hpp"
#include "glass/hnsw/hnsw.hpp"
#include "glass/nsg/nsg.hpp"
#include "glass/quant/quant.hpp"
#include "glass/searcher/graph_searcher.hpp"

namespace {

struct RefSearcher {
    std::unique_ptr<glass::GraphSearcherBase> s;
};

// Copies codes and calibration parameters out of a trained quantizer and
// evaluates query->base distances through the quantizer's own computer.
template <typename Q>
int quant_probe(const float *data, int n, int d, const float *qs, int nq, const int *ids, int nids, char *codes_out,
                float *params_out, double *dist_out, int *code_size_out, int *dim_align_out) {
    Q quant(d);
    quant.train(data, n);
    quant.add(data, n);
    *code_size_out = quant.code_size();
    *dim_align_out = quant.dim_align();
    if (codes_out) {
        for (int i = 0; i < n; ++i) {
            std::memcpy(codes_out + (int64_t)i * quant.code_size(), quant.get_code(i), quant.code_size());
        }
    }
    if (params_out) {
        if constexpr (requires { quant.calibrator.min; }) {
            params_out[0] = quant.calibrator.min;
            params_out[1] = quant.calibrator.dif;
        } else if constexpr (requires { quant.calibrator.max; }) {
            params_out[0] = quant.calibrator.max;
        } else if constexpr (requires { quant.calibrator.mins; }) {
            for (int i = 0; i < quant.dim_align(); ++i) {
                params_out[i] = quant.calibrator.mins[i];
                params_out[quant.dim_align() + i] = quant.calibrator.difs[i];
            }
        }
    }
    if (dist_out) {
        for (int qi = 0; qi < nq; ++qi) {
            auto computer = quant.get_computer(qs + (int64_t)qi * d);
            for (int j = 0; j < nids; ++j) {
                dist_out[(int64_t)qi * nids + j] = (double)computer(ids[j]);
            }
        }
    }
    return 0;
}

template <glass::Metric M>
int quant_dispatch(const std::string &quant, const float *data, int n, int d, const float *qs, int nq, const int *ids,
                   int nids, char *codes_out, float *params_out, double *dist_out, int *cs, int *da) {
    using namespace glass;
    if (quant == "SQ8U") return quant_probe<SQ8QuantizerUniform<M>>(data, n, d, qs, nq, ids, nids, codes_out, params_out, dist_out, cs, da);
    if (quant == "SQ4U") return quant_probe<SQ4QuantizerUniform<M>>(data, n, d, qs, nq, ids, nids, codes_out, params_out, dist_out, cs, da);
    if (quant == "SQ2U") return quant_probe<SQ2QuantizerUniform<M>>(data, n, d, qs, nq, ids, nids, codes_out, params_out, dist_out, cs, da);
    if (quant == "E5M2") return quant_probe<E5M2Quantizer<M>>(data, n, d, qs, nq, ids, nids, codes_out, params_out, dist_out, cs, da);
    if (quant == "SQ8") return quant_probe<SQ8Quantizer<M>>(data, n, d, qs, nq, ids, nids, codes_out, params_out, dist_out, cs, da);
    if (quant == "FP32") return quant_probe<FP32Quantizer<M>>(data, n, d, qs, nq, ids, nids, codes_out, params_out, dist_out, cs, da);
    if (quant == "FP16") return quant_probe<FP16Quantizer<M>>(data, n, d, qs, nq, ids, nids, codes_out, params_out, dist_out, cs, da);
    if (quant == "BF16") return quant_probe<BF16Quantizer<M>>(data, n, d, qs, nq, ids, nids, codes_out, params_out, dist_out, cs, da);
    if (quant == "SQ6") return quant_probe<SQ6Quantizer<M>>(data, n, d, qs, nq, ids, nids, codes_out, params_out, dist_out, cs, da);
    if (quant == "SQ4") return quant_probe<SQ4Quantizer<M>>(data, n, d, qs, nq, ids, nids, codes_out, params_out, dist_out, cs, da);
    return -1;
}

// ProductQuant (glass/quant/product_quant.hpp): codebooks, codes, per-query
// lookup tables and computer distances.
template <glass::Metric M>
int pq_probe(const float *data, int n, int d, const float *qs, int nq, const int *ids, int nids, float *codebook_out,
             uint8_t *codes_out, float *lut_out, double *dist_out) {
    glass::ProductQuant<M> pq(d);
    pq.train(data, n);
    pq.add(data, n);
    const int nsq = pq.nsq, sd = pq.sd, K = glass::ProductQuant<M>::K;
    if (codebook_out) {
        for (int sq = 0; sq < nsq; ++sq)
            for (int k = 0; k < K; ++k)
                std::memcpy(codebook_out + ((int64_t)sq * K + k) * sd, pq.codebook[sq].get(k), sizeof(float) * sd);
    }
    if (codes_out) {
        for (int i = 0; i < n; ++i) std::memcpy(codes_out + (int64_t)i * nsq, pq.get_code(i), nsq);
    }
    for (int qi = 0; qi < nq; ++qi) {
        if (lut_out) pq.get_lut(qs + (int64_t)qi * d, lut_out + (int64_t)qi * nsq * K);
        if (dist_out) {
            auto computer = pq.get_computer(qs + (int64_t)qi * d);
            for (int j = 0; j < nids; ++j) dist_out[(int64_t)qi * nids + j] = (double)computer(ids[j]);
        }
    }
    return 0;
}

}  // namespace

extern "C" {

// glass::Clustering as python/bindings.cc:155-203 drives it: fit on data,
// centroids out, predict and in-place transform of `pred` (clustering.hpp).
int ref_clustering_probe(const float *data, int n, int d, int n_cluster, int epochs, const char *init,
                         float *centroids_out, float *pred, int npred, int *ids_out) {
    glass::Clustering c(n_cluster, epochs, init);
    c.fit(data, n, d);
    for (int i = 0; i < n_cluster; ++i) std::memcpy(centroids_out + (int64_t)i * d, c.medoids.get(i), sizeof(float) * d);
    if (npred > 0) {
        auto ids = c.predict(pred, npred, d);
        std::memcpy(ids_out, ids.data(), sizeof(int) * npred);
        c.transform(pred, npred, d);
    }
    return 0;
}

int ref_pq_probe(const char *metric, const float *data, int n, int d, const float *qs, int nq, const int *ids,
                 int nids, float *codebook_out, uint8_t *codes_out, float *lut_out, double *dist_out) {
    if (d % 4 != 0) return -1;
    if (std::string(metric) == "IP")
        return pq_probe<glass::Metric::IP>(data, n, d, qs, nq, ids, nids, codebook_out, codes_out, lut_out, dist_out);
    return pq_probe<glass::Metric::L2>(data, n, d, qs, nq, ids, nids, codebook_out, codes_out, lut_out, dist_out);
}

int ref_set_num_threads(int t) {
    if (t > 0) omp_set_num_threads(t);
    return omp_get_max_threads();
}

// glass.build_graph (python/bindings.cc:231-237) without the Python layer.
double ref_build_graph(const char *index_type, const float *data, int n, int d, const char *metric, const char *quant,
                       int R, int L, const char *path) {
    std::unique_ptr<glass::Builder> b;
    std::string t(index_type);
    if (t == "HNSW") {
        b = glass::create_hnsw(metric, quant, R, L);
    } else if (t == "NSG") {
        b = glass::create_nsg(quant, R, L, false);
    } else if (t == "SSG") {
        b = glass::create_nsg(quant, R, L, true);
    }
    if (!b) return -1.0;
    auto g = b->Build(data, n, d);
    g.save(path);
    return b->GetConstructionTime();
}

void *ref_searcher_new(const char *graph_path, const float *data, int n, int d, const char *metric, const char *quant,
                       const char *refine) {
    glass::Graph<int> g;
    g.load(graph_path);
    auto *h = new RefSearcher;
    h->s = glass::create_searcher(std::move(g), metric, quant, refine);
    if (!h->s) {
        delete h;
        return nullptr;
    }
    h->s->SetData(data, n, d);
    return h;
}

// Same searcher with the re-rank depth multiplier chosen by the caller: the
// reference's make_refiner (glass/searcher/graph_searcher.hpp:240-254) takes
// it as an argument; create_searcher passes get_refine_factor(quantizer).
void *ref_searcher_new_factor(const char *graph_path, const float *data, int n, int d, const char *metric,
                              const char *quant, const char *refine, float factor) {
    glass::Graph<int> g;
    g.load(graph_path);
    auto *h = new RefSearcher;
    h->s = glass::create_searcher(std::move(g), metric, quant, "");
    if (!h->s) {
        delete h;
        return nullptr;
    }
    if (refine && refine[0]) {
        if (std::string(metric) == "IP") {
            h->s = glass::make_refiner<glass::Metric::IP>(std::move(h->s), refine, factor);
        } else {
            h->s = glass::make_refiner<glass::Metric::L2>(std::move(h->s), refine, factor);
        }
    }
    h->s->SetData(data, n, d);
    return h;
}

// Graph::load(filename, format) then Graph::save (glass/graph.hpp:106-153):
// converts a foreign graph file (diskann, nsg, hnswlib) to the .glass layout.
int ref_graph_convert(const char *src, const char *format, const char *dst) {
    glass::Graph<int> g;
    g.load(src, format);
    g.save(dst);
    return g.N;
}

void ref_searcher_free(void *h) { delete (RefSearcher *)h; }

void ref_set_ef(void *h, int ef) { ((RefSearcher *)h)->s->SetEf(ef); }

void ref_optimize(void *h, int nt) { ((RefSearcher *)h)->s->Optimize(nt); }

void ref_search(void *h, const float *q, int k, int *ids, float *dis) { ((RefSearcher *)h)->s->Search(q, k, ids, dis); }

void ref_search_batch(void *h, const float *qs, int nq, int k, int *ids, float *dis) {
    ((RefSearcher *)h)->s->SearchBatch(qs, nq, k, ids, dis);
}

double ref_dist_cmps(void *h) { return ((RefSearcher *)h)->s->GetLastSearchAvgDistCmps(); }

// Trains `quant` on data, returns its codes / calibration and query->id
// distances as computed by the reference computer (glass/quant/computer.hpp:42-75).
int ref_quant_probe(const char *quant, const char *metric, const float *data, int n, int d, const float *qs, int nq,
                    const int *ids, int nids, char *codes_out, float *params_out, double *dist_out, int *code_size_out,
                    int *dim_align_out) {
    std::string m(metric);
    if (m == "IP") {
        return quant_dispatch<glass::Metric::IP>(quant, data, n, d, qs, nq, ids, nids, codes_out, params_out, dist_out,
                                                 code_size_out, dim_align_out);
    }
    return quant_dispatch<glass::Metric::L2>(quant, data, n, d, qs, nq, ids, nids, codes_out, params_out, dist_out,
                                             code_size_out, dim_align_out);
}

}  // extern "C"
