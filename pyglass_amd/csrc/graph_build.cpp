// CPU graph builders (the producers of the search path's input; builders stay
// on the host).  HNSW restates the insertion algorithm the reference wraps in
// glass::HNSW::Build (glass/hnsw/hnsw.hpp:27-69, glass/hnswlib/hnswalg.h:53-349):
// M = R/2 links per upper-layer node, 2M at layer 0, heuristic-2 pruning,
// level ~ floor(-ln U * 1/ln M) from a default_random_engine seeded with 100,
// then the level-0 lists become the Graph and the upper layers the
// HNSWInitializer.  Distances are exact fp32 L2 (or negated inner product) on
// the raw rows, as the reference's default build quantizer "FP32".
//
// NSG / SSG restate glass::NSG (glass/nsg/nsg.hpp:22-409) over an NN-Descent
// kNN graph (glass/nsg/nndescent.hpp, glass/nsg/nhood.hpp): K=64 neighbours,
// S=10, R=100, pool 114, 10 iterations; entry = the row nearest the centroid;
// every node searches the kNN graph with a pool of L, keeps up to R of the
// first C=R+100 candidates by RNG occlusion (NSG: d(j,k) < d(i,k); SSG: angle
// above 60 degrees), adds reverse links and finally attaches every node not
// reachable from the entry.  As in the reference the metric is ignored (L2).
// Differences: per-node random streams instead of one shared rand(), visited
// epochs instead of a fresh N-bit bitset per node, and a pruned reverse-link
// list leaves its unused tail empty rather than holding earlier entries.
//
// Two deliberate differences in HNSW: levels are drawn up front in insertion order
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


namespace {

// ---------------------------------------------------------------------------
// NN-Descent (glass/nsg/nndescent.hpp:20-226)

struct FlagNb {
    float d;
    uint32_t id;
    bool flag;
    bool operator<(const FlagNb &o) const { return d < o.d || (d == o.d && id < o.id); }
};

struct Nhood {
    std::vector<FlagNb> pool;  // max-heap by (d, id), at most L entries
    int32_t M = 0;
    std::mutex lock;
    std::vector<uint32_t> nn_new, nn_old, rnn_new, rnn_old;

    // Nhood::insert (nhood.hpp:41-55)
    void insert(uint32_t id, float d, size_t L) {
        std::lock_guard<std::mutex> g(lock);
        if (!pool.empty() && d > pool.front().d) return;
        for (const FlagNb &x : pool)
            if (x.id == id) return;
        if (pool.size() < L) {
            pool.push_back({d, id, true});
            std::push_heap(pool.begin(), pool.end());
        } else {
            std::pop_heap(pool.begin(), pool.end());
            pool.back() = {d, id, true};
            std::push_heap(pool.begin(), pool.end());
        }
    }
};

template <typename F>
void parallel_for(int64_t n, int32_t threads, F f) {
    std::atomic<int64_t> next{0};
    auto worker = [&](int32_t t) {
        for (;;) {
            int64_t b = next.fetch_add(64);
            if (b >= n) break;
            for (int64_t i = b; i < std::min<int64_t>(n, b + 64); ++i) f(i, t);
        }
    };
    std::vector<std::thread> th;
    for (int32_t t = 0; t < threads; ++t) th.emplace_back(worker, t);
    for (auto &x : th) x.join();
}

uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

struct NNDescent {
    const Dist &dist;
    int32_t n, K, S = 10, R = 100, L, iters = 10, threads;
    std::vector<Nhood> g;

    NNDescent(const Dist &d_, int32_t n_, int32_t K_, int32_t th) : dist(d_), n(n_), K(K_), L(K_ + 50), threads(th), g(n_) {}

    void init() {
        parallel_for(n, threads, [&](int64_t i, int32_t) {
            std::mt19937_64 rng(mix64(347u ^ ((uint64_t)i << 20)));
            Nhood &h = g[i];
            h.M = S;
            h.nn_new.resize(2 * S);
            for (auto &x : h.nn_new) x = (uint32_t)(rng() % (uint64_t)n);
            for (int32_t j = 0; j < S; ++j) {
                uint32_t id = (uint32_t)(rng() % (uint64_t)n);
                if (id == (uint32_t)i) continue;
                h.pool.push_back({dist((uint32_t)i, id), id, true});
            }
            std::make_heap(h.pool.begin(), h.pool.end());
            h.pool.reserve(L);
        });
    }

    // Nhood::join + NNDescent::Join (nhood.hpp:57-81, nndescent.hpp:86-98)
    void join() {
        parallel_for(n, threads, [&](int64_t u, int32_t) {
            const Nhood &h = g[u];
            auto cb = [&](uint32_t a, uint32_t b) {
                if (a == b) return;
                const float d = dist(a, b);
                g[a].insert(b, d, (size_t)L);
                g[b].insert(a, d, (size_t)L);
            };
            for (uint32_t a : h.nn_new) {
                for (uint32_t b : h.nn_new)
                    if (a < b) cb(a, b);
                for (uint32_t b : h.nn_old) cb(a, b);
            }
        });
    }

    // NNDescent::Update (nndescent.hpp:100-172)
    void update(int32_t iter) {
        parallel_for(n, threads, [&](int64_t i, int32_t) {
            std::vector<uint32_t>().swap(g[i].nn_new);
            std::vector<uint32_t>().swap(g[i].nn_old);
        });
        parallel_for(n, threads, [&](int64_t i, int32_t) {
            Nhood &h = g[i];
            std::sort(h.pool.begin(), h.pool.end());
            if ((int32_t)h.pool.size() > L) h.pool.resize(L);
            const int32_t maxl = std::min<int32_t>(h.M + S, (int32_t)h.pool.size());
            int32_t c = 0, l = 0;
            while (l < maxl && c < S) {
                if (h.pool[l].flag) ++c;
                ++l;
            }
            h.M = l;
        });
        parallel_for(n, threads, [&](int64_t i, int32_t) {
            Nhood &h = g[i];
            std::mt19937_64 rng(mix64(((uint64_t)iter << 40) ^ (uint64_t)i));
            for (int32_t l = 0; l < h.M; ++l) {
                FlagNb &nb = h.pool[l];
                Nhood &o = g[nb.id];
                std::vector<uint32_t> &mine = nb.flag ? h.nn_new : h.nn_old;
                mine.push_back(nb.id);
                // reverse candidate when this node lies beyond the other's pool
                bool far;
                {
                    std::lock_guard<std::mutex> lk(o.lock);
                    far = !o.pool.empty() && nb.d > o.pool.back().d;
                    if (far) {
                        std::vector<uint32_t> &r = nb.flag ? o.rnn_new : o.rnn_old;
                        if ((int32_t)r.size() < R) r.push_back((uint32_t)i);
                        else r[rng() % (uint64_t)R] = (uint32_t)i;
                    }
                }
                nb.flag = false;
            }
            std::make_heap(h.pool.begin(), h.pool.end());
        });
        parallel_for(n, threads, [&](int64_t i, int32_t) {
            Nhood &h = g[i];
            h.nn_new.insert(h.nn_new.end(), h.rnn_new.begin(), h.rnn_new.end());
            h.nn_old.insert(h.nn_old.end(), h.rnn_old.begin(), h.rnn_old.end());
            if ((int32_t)h.nn_old.size() > 2 * R) h.nn_old.resize(2 * R);
            std::vector<uint32_t>().swap(h.rnn_new);
            std::vector<uint32_t>().swap(h.rnn_old);
        });
    }

    std::vector<int32_t> build() {
        init();
        for (int32_t it = 1; it <= iters; ++it) {
            join();
            update(it);
        }
        std::vector<int32_t> knn((size_t)n * K, -1);
        parallel_for(n, threads, [&](int64_t i, int32_t) {
            std::sort(g[i].pool.begin(), g[i].pool.end());
            for (int32_t j = 0; j < K && j < (int32_t)g[i].pool.size(); ++j)
                knn[(size_t)i * K + j] = (int32_t)g[i].pool[j].id;
        });
        std::vector<Nhood>().swap(g);
        return knn;
    }
};

// ---------------------------------------------------------------------------
// NSG (glass/nsg/nsg.hpp)

struct NsgBuilder {
    const Dist &dist;
    int32_t n, R, L, C, threads;
    bool ssg;
    float threshold = (float)std::cos(60.0 / 180.0 * 3.14159265358979323846264);
    int32_t ep = 0;

    // search_on_graph (nsg.hpp:128-163): every evaluated (dist, id) goes to fullset
    void search(uint32_t p, const std::vector<int32_t> &adj, int32_t K, Pool &pool, Visited &vis,
                std::vector<pair_t> &fullset) const {
        pool.reset(L);
        vis.reset((size_t)n);
        const float de = dist(p, (uint32_t)ep);
        pool.insert((uint32_t)ep, de);
        vis.test_set((uint32_t)ep);
        fullset.push_back({de, (uint32_t)ep});
        while (pool.has_next()) {
            const uint32_t u = pool.pop();
            for (int32_t i = 0; i < K; ++i) {
                const int32_t v = adj[(size_t)u * K + i];
                if (v == -1) break;
                if (vis.test_set((uint32_t)v)) continue;
                const float dv = dist(p, (uint32_t)v);
                pool.insert((uint32_t)v, dv);
                fullset.push_back({dv, (uint32_t)v});
            }
        }
    }

    bool occluded(const pair_t &p, const std::vector<pair_t> &result, bool angle) const {
        for (const pair_t &r : result) {
            if (p.second == r.second) return true;
            const float djk = dist(r.second, p.second);
            if (angle) {
                const float cos_ij = (p.first + r.first - djk) / 2 / std::sqrt(p.first * r.first);
                if (cos_ij > threshold) return true;
            } else if (djk < p.first) {
                return true;
            }
        }
        return false;
    }

    // sync_prune (nsg.hpp:186-240)
    void prune(uint32_t q, std::vector<pair_t> &fullset, Visited &vis, const std::vector<int32_t> &knn, int32_t GK,
               std::vector<pair_t> &out) const {
        for (int32_t i = 0; i < GK; ++i) {
            const int32_t id = knn[(size_t)q * GK + i];
            if (id == -1) break;
            if (vis.test_set((uint32_t)id)) continue;
            fullset.push_back({dist(q, (uint32_t)id), (uint32_t)id});
        }
        std::sort(fullset.begin(), fullset.end());
        out.clear();
        size_t start = 0;
        if (!fullset.empty() && fullset[0].second == q) start++;
        if (start >= fullset.size()) return;
        out.push_back(fullset[start]);
        while ((int32_t)out.size() < R && ++start < fullset.size() && (int32_t)start < C) {
            if (!occluded(fullset[start], out, ssg)) out.push_back(fullset[start]);
        }
    }
};

}  // namespace

// Builds an NSG (ssg=false) or SSG (ssg=true) graph: adj is n x R, -1 padded;
// ep the single entry point.  Returns 0 on success.
int build_nsg(const float *X, int32_t n, int32_t d, int32_t R, int32_t L, bool ssg, int32_t num_threads,
              std::vector<int32_t> &adj, int32_t &ep, double *seconds) {
    if (n <= 1 || d <= 0 || R < 2) return -1;
    auto t0 = std::chrono::high_resolution_clock::now();
    if (num_threads <= 0) num_threads = (int32_t)std::max(1u, std::thread::hardware_concurrency());
    const Dist dist{X, d, false};
    const int32_t GK = std::min<int32_t>(64, n - 1);
    std::vector<int32_t> knn;
    {
        NNDescent nnd(dist, n, GK, num_threads);
        knn = nnd.build();
    }
    NsgBuilder b{dist, n, R, L > 0 ? L : R + 32, R + 100, num_threads, ssg};
    // Init (nsg.hpp:101-126): the row nearest the centroid
    {
        std::vector<double> c(d, 0.0);
        for (int64_t i = 0; i < n; ++i)
            for (int32_t j = 0; j < d; ++j) c[j] += X[i * d + j];
        std::vector<float> cf(d);
        for (int32_t j = 0; j < d; ++j) cf[j] = (float)(c[j] / n);
        float best = 0.0f;
        for (int64_t i = 0; i < n; ++i) {
            float s = 0.0f;
            for (int32_t j = 0; j < d; ++j) {
                const float t = X[i * d + j] - cf[j];
                s += t * t;
            }
            if (i == 0 || s < best) {
                best = s;
                b.ep = (int32_t)i;
            }
        }
    }
    // link (nsg.hpp:165-184): search + prune per node, then reverse links
    std::vector<std::vector<pair_t>> lists((size_t)n);
    {
        std::vector<Pool> pools(num_threads);
        std::vector<Visited> vis(num_threads);
        parallel_for(n, num_threads, [&](int64_t i, int32_t t) {
            std::vector<pair_t> fullset;
            b.search((uint32_t)i, knn, GK, pools[t], vis[t], fullset);
            b.prune((uint32_t)i, fullset, vis[t], knn, GK, lists[i]);
        });
    }
    // add_reverse_links (nsg.hpp:242-312)
    {
        std::vector<std::mutex> locks((size_t)n);
        std::vector<std::vector<pair_t>> fwd = lists;  // each node's own pruned list, read-only below
        parallel_for(n, num_threads, [&](int64_t q, int32_t) {
            for (const pair_t &e : fwd[q]) {
                const uint32_t des = e.second;
                const pair_t sn{e.first, (uint32_t)q};
                std::vector<pair_t> tmp;
                {
                    std::lock_guard<std::mutex> g(locks[des]);
                    bool dup = false;
                    for (const pair_t &x : lists[des]) {
                        if (x.second == (uint32_t)q) {
                            dup = true;
                            break;
                        }
                    }
                    if (dup) continue;
                    if ((int32_t)lists[des].size() < R) {
                        lists[des].push_back(sn);
                        continue;
                    }
                    tmp = lists[des];
                }
                tmp.push_back(sn);
                std::sort(tmp.begin(), tmp.end());
                std::vector<pair_t> res;
                res.push_back(tmp[0]);
                for (size_t st = 1; (int32_t)res.size() < R && st < tmp.size(); ++st) {
                    if (!b.occluded(tmp[st], res, false)) res.push_back(tmp[st]);
                }
                std::lock_guard<std::mutex> g(locks[des]);
                lists[des] = res;
            }
        });
    }
    adj.assign((size_t)n * R, -1);
    std::vector<int32_t> degree((size_t)n, 0);
    for (int64_t i = 0; i < n; ++i) {
        int32_t c = 0;
        for (const pair_t &e : lists[i]) {
            if (c >= R) break;
            adj[(size_t)i * R + c++] = (int32_t)e.second;
        }
        degree[i] = c;
    }
    std::vector<std::vector<pair_t>>().swap(lists);
    // tree_grow (nsg.hpp:314-409): attach every node the entry cannot reach
    {
        std::vector<uint8_t> seen((size_t)n, 0);
        int64_t cnt = 0;
        int32_t root = b.ep;
        Pool pool;
        Visited vis;
        std::mt19937_64 rng(347);
        for (;;) {
            // bfs (nsg.hpp:333-357)
            std::queue<int32_t> bq;
            bq.push(root);
            if (!seen[root]) {
                seen[root] = 1;
                cnt++;
            }
            while (!bq.empty()) {
                const int32_t u = bq.front();
                bq.pop();
                for (int32_t i = 0; i < R; ++i) {
                    const int32_t v = adj[(size_t)u * R + i];
                    if (v == -1) break;
                    if (seen[v]) continue;
                    seen[v] = 1;
                    cnt++;
                    bq.push(v);
                }
            }
            if (cnt >= n) break;
            // attach_unlinked (nsg.hpp:359-398)
            int32_t id = -1;
            for (int64_t i = 0; i < n; ++i) {
                if (!seen[i]) {
                    id = (int32_t)i;
                    break;
                }
            }
            std::vector<pair_t> fullset;
            b.search((uint32_t)id, adj, R, pool, vis, fullset);
            std::sort(fullset.begin(), fullset.end());
            int32_t node = -1;
            for (const pair_t &e : fullset) {
                if (degree[e.second] < R && (int32_t)e.second != id) {
                    node = (int32_t)e.second;
                    break;
                }
            }
            for (int64_t tries = 0; node < 0 && tries < 64 * (int64_t)n; ++tries) {
                const int32_t c = (int32_t)(rng() % (uint64_t)n);
                if (seen[c] && degree[c] < R && c != id) node = c;
            }
            for (int64_t i = 0; node < 0 && i < n; ++i) {
                if (seen[i] && degree[i] < R && (int32_t)i != id) node = (int32_t)i;
            }
            if (node < 0) return -3;  // every reachable node is full
            adj[(size_t)node * R + degree[node]] = id;
            degree[node] += 1;
            root = node;
        }
    }
    ep = b.ep;
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
    return 0;
}

}  // namespace glass_build
