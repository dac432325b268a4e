// CPU graph builders (the producers of the search path's input; builders stay
// on the host).  HNSW restates the insertion algorithm the reference wraps in
// glass::HNSW::Build (glass/hnsw/hnsw.hpp:27-69, glass/hnswlib/hnswalg.h:53-349):
// M = R/2 links per upper-layer node, 2M at layer 0, heuristic-2 pruning,
// level ~ floor(-ln U * 1/ln M) from a default_random_engine seeded with 100,
// then the level-0 lists become the Graph and the upper layers the
// HNSWInitializer.  Distances are exact fp32 L2 (or negated inner product) on
// the raw rows, as the reference's default build quantizer "FP32".
//
// Two deliberate differences: levels are drawn up front in insertion order
// (the reference draws them inside concurrently running addPoint calls, so its
// level sequence depends on thread timing), and the visited set is a
// per-thread tagged array instead of a fresh N-bit bitset per search.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <queue>
#include <random>
#include <thread>
#include <vector>

namespace glass_build {

namespace {

using pair_t = std::pair<float, uint32_t>;

struct Dist {
    const float *X;
    int32_t d;
    bool ip;

    // 16 independent partial sums so the loop vectorises without reassociation.
    float operator()(uint32_t a, uint32_t b) const {
        const float *x = X + (int64_t)a * d, *y = X + (int64_t)b * d;
        float acc[16] = {0};
        int32_t j = 0;
        if (!ip) {
            for (; j + 16 <= d; j += 16)
                for (int l = 0; l < 16; ++l) {
                    float t = x[j + l] - y[j + l];
                    acc[l] += t * t;
                }
            for (; j < d; ++j) {
                float t = x[j] - y[j];
                acc[0] += t * t;
            }
        } else {
            for (; j + 16 <= d; j += 16)
                for (int l = 0; l < 16; ++l) acc[l] += x[j + l] * y[j + l];
            for (; j < d; ++j) acc[0] += x[j] * y[j];
        }
        float s = 0.0f;
        for (int l = 0; l < 16; ++l) s += acc[l];
        return ip ? -s : s;
    }
};

// Sorted candidate pool with LinearPool semantics (glass/neighbor.hpp:320-433).
struct Pool {
    std::vector<float> dist;
    std::vector<uint32_t> id;  // bit 31 = checked
    int32_t size = 0, cur = 0, cap = 0;

    void reset(int32_t c) {
        cap = c;
        size = cur = 0;
        dist.resize(c + 1);
        id.resize(c + 1);
    }
    void insert(uint32_t u, float d) {
        if (size == cap && d >= dist[size - 1]) return;
        int32_t lo = (int32_t)(std::upper_bound(dist.begin(), dist.begin() + size, d) - dist.begin());
        std::memmove(&dist[lo + 1], &dist[lo], sizeof(float) * (size_t)(size - lo));
        std::memmove(&id[lo + 1], &id[lo], sizeof(uint32_t) * (size_t)(size - lo));
        dist[lo] = d;
        id[lo] = u;
        if (size < cap) size++;
        if (lo < cur) cur = lo;
    }
    bool has_next() const { return cur < size; }
    uint32_t pop() {
        id[cur] |= 0x80000000u;
        int32_t pre = cur;
        while (cur < size && (id[cur] & 0x80000000u)) cur++;
        return id[pre] & 0x7fffffffu;
    }
};

struct Visited {
    std::vector<uint32_t> tag;
    uint32_t epoch = 0;
    void reset(size_t n) {
        if (tag.size() != n) {
            tag.assign(n, 0);
            epoch = 0;
        }
        if (++epoch == 0) {
            std::fill(tag.begin(), tag.end(), 0);
            epoch = 1;
        }
    }
    bool test_set(uint32_t v) {
        if (tag[v] == epoch) return true;
        tag[v] = epoch;
        return false;
    }
};

struct Hnsw {
    Dist dist;
    int32_t n, M, maxM, maxM0, efc;
    double mult;
    std::vector<int32_t> levels;
    std::vector<uint32_t> l0;          // n x (1 + maxM0): count, ids
    std::vector<std::vector<uint32_t>> up;  // per node: levels x (1 + maxM)
    std::vector<std::mutex> locks;
    std::mutex global;
    std::atomic<int32_t> ep{-1};
    int32_t maxlevel = -1;

    uint32_t *list(uint32_t u, int32_t level) {
        if (level == 0) return l0.data() + (size_t)u * (1 + maxM0);
        return up[u].data() + (size_t)(level - 1) * (1 + maxM);
    }

    // searchBaseLayer (hnswalg.h:100-144)
    void search_layer(uint32_t ep_id, uint32_t u, int32_t level, Pool &pool, Visited &vis,
                      std::vector<pair_t> &out) {
        vis.reset((size_t)n);
        pool.reset(efc);
        pool.insert(ep_id, dist(u, ep_id));
        vis.test_set(ep_id);
        std::vector<uint32_t> nb;
        while (pool.has_next()) {
            uint32_t x = pool.pop();
            {
                std::lock_guard<std::mutex> lk(locks[x]);
                uint32_t *l = list(x, level);
                nb.assign(l + 1, l + 1 + l[0]);
            }
            for (uint32_t y : nb) {
                if (vis.test_set(y)) continue;
                pool.insert(y, dist(u, y));
            }
        }
        out.clear();
        for (int32_t i = 0; i < pool.size; ++i) out.emplace_back(pool.dist[i], pool.id[i] & 0x7fffffffu);
    }

    // getNeighborsByHeuristic2 (hnswalg.h:146-184); cands is a max-heap.
    void heuristic(std::priority_queue<pair_t> &cands, size_t m) {
        if (cands.size() < m) return;
        std::priority_queue<pair_t> closest;
        std::vector<pair_t> ret;
        while (!cands.empty()) {
            auto t = cands.top();
            cands.pop();
            closest.emplace(-t.first, t.second);
        }
        while (!closest.empty() && ret.size() < m) {
            auto t = closest.top();
            closest.pop();
            float dq = -t.first;
            bool good = true;
            for (auto &r : ret) {
                if (dist(t.second, r.second) < dq) {
                    good = false;
                    break;
                }
            }
            if (good) ret.push_back(t);
        }
        for (auto &r : ret) cands.emplace(-r.first, r.second);
    }

    // mutuallyConnectNewElement (hnswalg.h:196-275)
    uint32_t connect(uint32_t u, std::priority_queue<pair_t> &top, int32_t level) {
        const size_t mcur = level ? (size_t)maxM : (size_t)maxM0;
        heuristic(top, (size_t)M);
        std::vector<uint32_t> sel;
        while (!top.empty()) {
            sel.push_back(top.top().second);
            top.pop();
        }
        uint32_t next = sel.back();
        {
            std::lock_guard<std::mutex> lk(locks[u]);
            uint32_t *l = list(u, level);
            l[0] = (uint32_t)sel.size();
            for (size_t i = 0; i < sel.size(); ++i) l[1 + i] = sel[i];
        }
        for (uint32_t s : sel) {
            std::lock_guard<std::mutex> lk(locks[s]);
            uint32_t *l = list(s, level);
            size_t sz = l[0];
            if (sz < mcur) {
                l[1 + sz] = u;
                l[0] = (uint32_t)(sz + 1);
            } else {
                std::priority_queue<pair_t> c;
                c.emplace(dist(u, s), u);
                for (size_t j = 0; j < sz; ++j) c.emplace(dist(l[1 + j], s), l[1 + j]);
                heuristic(c, mcur);
                size_t i = 0;
                while (!c.empty()) {
                    l[1 + i++] = c.top().second;
                    c.pop();
                }
                l[0] = (uint32_t)i;
            }
        }
        return next;
    }

    // addPoint (hnswalg.h:285-349)
    void add(uint32_t u, Pool &pool, Visited &vis) {
        std::unique_lock<std::mutex> lock_el(locks[u]);
        const int32_t curlevel = levels[u];
        std::unique_lock<std::mutex> templock(global);
        const int32_t maxlevelcopy = maxlevel;
        if (curlevel <= maxlevelcopy) templock.unlock();
        int32_t cur = ep.load();
        lock_el.unlock();
        if (cur != -1) {
            uint32_t curobj = (uint32_t)cur;
            if (curlevel < maxlevelcopy) {
                float curdist = dist(u, curobj);
                for (int32_t level = maxlevelcopy; level > curlevel; --level) {
                    bool changed = true;
                    while (changed) {
                        changed = false;
                        std::vector<uint32_t> nb;
                        {
                            std::lock_guard<std::mutex> lk(locks[curobj]);
                            uint32_t *l = list(curobj, level);
                            nb.assign(l + 1, l + 1 + l[0]);
                        }
                        for (uint32_t c : nb) {
                            float d = dist(u, c);
                            if (d < curdist) {
                                curdist = d;
                                curobj = c;
                                changed = true;
                            }
                        }
                    }
                }
            }
            std::vector<pair_t> res;
            for (int32_t level = std::min(curlevel, maxlevelcopy); level >= 0; --level) {
                search_layer(curobj, u, level, pool, vis, res);
                std::priority_queue<pair_t> top(res.begin(), res.end());
                curobj = connect(u, top, level);
            }
        } else {
            ep.store((int32_t)u);
            maxlevel = curlevel;
        }
        if (curlevel > maxlevelcopy) {
            ep.store((int32_t)u);
            maxlevel = curlevel;
        }
    }
};

}  // namespace

// Builds an HNSW graph; fills level-0 adjacency (n x 2M, -1 padded), the
// entry point, per-node levels and the concatenated upper-layer lists
// (levels[i] * M ints per node).  Returns 0 on success.
int build_hnsw(const float *X, int32_t n, int32_t d, bool ip, int32_t R, int32_t L, int32_t num_threads,
               std::vector<int32_t> &adj, int32_t &ep, std::vector<int32_t> &levels, std::vector<int32_t> &lists,
               double *seconds) {
    if (n <= 0 || d <= 0 || R < 2) return -1;
    auto t0 = std::chrono::high_resolution_clock::now();
    Hnsw h{Dist{X, d, ip}, n, R / 2, R / 2, R, std::max(L, R / 2), 1.0 / std::log(1.0 * (R / 2)), {}, {}, {}, std::vector<std::mutex>(n)};
    h.levels.resize(n);
    std::default_random_engine gen(100);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    for (int32_t i = 0; i < n; ++i) h.levels[i] = (int32_t)(-std::log(uni(gen)) * h.mult);
    h.l0.assign((size_t)n * (1 + h.maxM0), 0);
    h.up.resize(n);
    for (int32_t i = 0; i < n; ++i)
        if (h.levels[i] > 0) h.up[i].assign((size_t)h.levels[i] * (1 + h.maxM), 0);
    {
        Pool p;
        Visited v;
        h.add(0, p, v);
    }
    if (num_threads <= 0) num_threads = (int32_t)std::max(1u, std::thread::hardware_concurrency());
    std::atomic<int64_t> next{1};
    auto worker = [&]() {
        Pool p;
        Visited v;
        for (;;) {
            int64_t i = next.fetch_add(1);
            if (i >= n) break;
            h.add((uint32_t)i, p, v);
        }
    };
    std::vector<std::thread> th;
    for (int32_t t = 0; t < num_threads; ++t) th.emplace_back(worker);
    for (auto &t : th) t.join();
    // export (hnsw.hpp:42-65)
    adj.assign((size_t)n * R, -1);
    for (int32_t i = 0; i < n; ++i) {
        const uint32_t *l = h.l0.data() + (size_t)i * (1 + h.maxM0);
        for (uint32_t j = 0; j < l[0]; ++j) adj[(size_t)i * R + j] = (int32_t)l[1 + j];
    }
    ep = h.ep.load();
    levels = h.levels;
    lists.clear();
    for (int32_t i = 0; i < n; ++i) {
        for (int32_t lv = 1; lv <= h.levels[i]; ++lv) {
            const uint32_t *l = h.up[i].data() + (size_t)(lv - 1) * (1 + h.maxM);
            for (int32_t j = 0; j < h.maxM; ++j) lists.push_back(j < (int32_t)l[0] ? (int32_t)l[1 + j] : -1);
        }
    }
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
    return 0;
}

}  // namespace glass_build
