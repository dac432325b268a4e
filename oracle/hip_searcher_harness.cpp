// Test infrastructure only: include/glass_hip_searcher.hpp (the C++ drop-in a
// maintainer adds to the reference as glass/searcher/hip_searcher.hpp)
// compiled against the reference's own headers where they lie
// (oracle/Makefile, target `ref`), and driven only through
// glass::GraphSearcherBase's virtuals (glass/searcher/searcher_base.hpp:7-19),
// as python/bindings.cc:90-154 drives a searcher.  The graph is read by the
// reference's Graph::load (glass/graph.hpp:136-153).  The library it calls
// is this repository's libglass_hip.so.
#include <cstdint>
#include <memory>
#include <string>

#include "glass/graph.hpp"
#include "glass_hip_searcher.hpp"

namespace {
glass::GraphSearcherBase *as_base(void *h) { return static_cast<glass::GraphSearcherBase *>(h); }
}  // namespace

extern "C" {

void *hs_new(const char *graph_path, const float *data, int n, int d, const char *metric, const char *quant,
             const char *refine) {
    try {
        glass::Graph<int> g;
        g.load(graph_path);
        std::unique_ptr<glass::GraphSearcherBase> s =
            std::make_unique<glass::HipSearcher>(std::move(g), metric, quant, refine ? refine : "");
        s->SetData(data, n, d);
        return s.release();
    } catch (...) {
        return nullptr;
    }
}

void hs_free(void *h) { delete as_base(h); }

int hs_set_ef(void *h, int ef) {
    try {
        as_base(h)->SetEf(ef);
        return 0;
    } catch (...) {
        return -1;
    }
}

int hs_get_ef(void *h) { return as_base(h)->GetEf(); }

int hs_optimize(void *h) {
    try {
        as_base(h)->Optimize(0);
        return 0;
    } catch (...) {
        return -1;
    }
}

int hs_search(void *h, const float *q, int k, int *ids, float *dis) {
    try {
        as_base(h)->Search(q, k, ids, dis);
        return 0;
    } catch (...) {
        return -1;
    }
}

int hs_search_batch(void *h, const float *qs, int nq, int k, int *ids, float *dis) {
    try {
        as_base(h)->SearchBatch(qs, nq, k, ids, dis);
        return 0;
    } catch (...) {
        return -1;
    }
}

double hs_dist_cmps(void *h) { return as_base(h)->GetLastSearchAvgDistCmps(); }

}  // extern "C"
