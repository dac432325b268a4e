// glass_hip_searcher.hpp -- the C++ drop-in over the C ABI: a
// glass::GraphSearcherBase (glass/searcher/searcher_base.hpp:7-19) whose
// methods call libglass_hip.so.  A maintainer copies this header into the
// reference tree as glass/searcher/hip_searcher.hpp; create_searcher
// (glass/searcher/graph_searcher.hpp:256-391) and the Searcher binding
// (python/bindings.cc:90-154) then use it unchanged.  It needs the
// reference's glass/graph.hpp and glass/searcher/searcher_base.hpp on the
// include path, and include/glass_hip.h from this repository.
#pragma once

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "glass/graph.hpp"
#include "glass/searcher/searcher_base.hpp"
#include "glass_hip.h"

namespace glass {

struct HipSearcher : GraphSearcherBase {
    glass_hip_searcher *h = nullptr;

    // The reference's searcher reports nothing; a failing call here throws,
    // which the pybind11 binding turns into a Python exception.
    static void check(int rc) {
        if (rc != GLASS_HIP_OK) throw std::runtime_error(glass_hip_last_error());
    }

    // create_searcher(graph, metric, quantizer, refine_quant): the graph's
    // adjacency, entry points and HNSW upper layers move into the library.
    HipSearcher(Graph<int32_t> &&graph, const std::string &metric, const std::string &quantizer,
                const std::string &refine_quant) {
        glass_hip_graph *g = nullptr;
        check(glass_hip_graph_create(graph.N, graph.K, graph.data, (int32_t)graph.eps.size(), graph.eps.data(), &g));
        if (graph.initializer) {
            const auto &ini = *graph.initializer;
            std::vector<int32_t> lists;
            for (int32_t i = 0; i < ini.N; ++i)
                if (ini.levels[i] > 0) lists.insert(lists.end(), ini.lists[i], ini.lists[i] + ini.levels[i] * ini.K);
            const int rc = glass_hip_graph_set_upper(g, ini.K, ini.ep, ini.levels.data(), lists.data());
            if (rc != GLASS_HIP_OK) glass_hip_graph_free(g);
            check(rc);
        }
        const int rc = glass_hip_searcher_create(g, metric.c_str(), quantizer.c_str(), refine_quant.c_str(), &h);
        glass_hip_graph_free(g);
        check(rc);
        Graph<int32_t> moved(std::move(graph));  // the caller's graph is left empty (bindings.cc:94-98)
    }
    HipSearcher(const HipSearcher &) = delete;
    HipSearcher &operator=(const HipSearcher &) = delete;
    ~HipSearcher() override { glass_hip_searcher_free(h); }

    void SetData(const float *data, int32_t n, int32_t dim) override {
        check(glass_hip_searcher_set_data(h, data, n, dim));
    }
    void SetEf(int32_t ef) override { check(glass_hip_searcher_set_ef(h, ef)); }
    int32_t GetEf() const override { return glass_hip_searcher_get_ef(h); }
    void Search(const float *q, int32_t k, int32_t *dst, float *scores = nullptr) const override {
        check(glass_hip_searcher_search(h, q, k, dst, scores));
    }
    void SearchBatch(const float *qs, int32_t nq, int32_t k, int32_t *dst, float *scores = nullptr) const override {
        check(glass_hip_searcher_search_batch(h, qs, nq, k, dst, scores));
    }
    void Optimize(int32_t num_threads = 0) override { check(glass_hip_searcher_optimize(h, num_threads)); }
    double GetLastSearchAvgDistCmps() const override { return glass_hip_searcher_last_avg_dist_cmps(h); }
};

}  // namespace glass
