// k-means (glass::Clustering, glass/algorithms/clustering.hpp): the engine the
// PQ8 trainer and the Clustering object of the C ABI share.  Seeding and the
// centroid update follow the reference step by step on the host (both are
// order-dependent: every k-means++ pick depends on the previous one, and a
// centroid is a float sum over its rows in row order); the distance sweeps
// run on a backend, the GPU for the Clustering object.
//
// Float steps are those of the reference as GCC -Ofast compiles it:
//   distances     helpa::l2_fp32_fp32 (l2 below, l2_helpa on the device)
//   seeding total in double: 8 floats per step into 4 lanes,
//                 a_l += d[8b+l] + d[8b+4+l], (a1 + a3) + (a0 + a2), one
//                 4-wide step (d1 + d3) + (d0 + d2), then one by one
//   centroid      sum * (1.0f / float(count))  (reciprocal math)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "../../include/glass_hip.h"
#include "glass_hip_internal.h"

namespace glass_hip {
namespace km {

float l2(const float *x, const float *y, int32_t d) {
    const int32_t da = d / 16 * 16;
    float A = 0.0f;
    if (da > 0) {
        float a[16] = {0};
        for (int32_t i = 0; i < da; i += 16)
            for (int l = 0; l < 16; ++l) {
                const float t = x[i + l] - y[i + l];
                a[l] = std::fma(t, t, a[l]);
            }
        float b[8], c[4];
        for (int l = 0; l < 8; ++l) b[l] = a[l + 8] + a[l];
        for (int l = 0; l < 4; ++l) c[l] = b[l + 4] + b[l];
        A = (c[0] + c[2]) + (c[1] + c[3]);
    }
    float T = 0.0f;
    int32_t i = da;
    if (d - i >= 8) {
        float u[4];
        for (int l = 0; l < 4; ++l) {
            const float t0 = x[i + l] - y[i + l], t1 = x[i + 4 + l] - y[i + 4 + l];
            u[l] = t1 * t1 + t0 * t0;
        }
        T = (u[1] + u[3]) + (u[0] + u[2]);
        i += 8;
    }
    if (d - i >= 4) {
        float q[4];
        for (int l = 0; l < 4; ++l) {
            const float t = x[i + l] - y[i + l];
            q[l] = t * t;
        }
        T = T + ((q[1] + q[3]) + (q[0] + q[2]));
        i += 4;
    }
    for (; i < d; ++i) {
        const float t = x[i] - y[i];
        T = std::fma(t, t, T);
    }
    return A + T;
}

namespace {

double seed_total(const float *d, int32_t n) {
    double S = 0.0;
    int32_t i = 0;
    if (n >= 8) {
        double a[4] = {0.0, 0.0, 0.0, 0.0};
        for (; i + 8 <= n; i += 8)
            for (int l = 0; l < 4; ++l) a[l] = a[l] + ((double)d[i + 4 + l] + (double)d[i + l]);
        S = (a[1] + a[3]) + (a[0] + a[2]);
    }
    if (n - i >= 4) {
        S = S + (((double)d[i + 1] + (double)d[i + 3]) + ((double)d[i] + (double)d[i + 2]));
        i += 4;
    }
    for (; i < n; ++i) S = S + (double)d[i];
    return S;
}

struct HostBackend : Backend {
    const float *x;
    int32_t m, d;
    int dist_to(const float *row, float *out) override {
        for (int32_t i = 0; i < m; ++i) out[i] = l2(row, x + (int64_t)i * d, d);
        return 0;
    }
    int assign(const float *cent, int32_t k, int32_t *ids) override {
        for (int32_t i = 0; i < m; ++i) {
            float best = HUGE_VALF;
            int32_t bid = -1;
            for (int32_t j = 0; j < k; ++j) {
                const float v = l2(x + (int64_t)i * d, cent + (int64_t)j * d, d);
                if (v < best) {
                    best = v;
                    bid = j;
                }
            }
            ids[i] = bid;
        }
        return 0;
    }
};

}  // namespace

Backend *host_backend(const float *samples, int32_t m, int32_t d, void *) {
    auto *b = new HostBackend;
    b->x = samples;
    b->m = m;
    b->d = d;
    return b;
}

int fit(const float *rows, int32_t n, int32_t d, int32_t k, int32_t epochs, bool kmeanspp, float *centroids,
        Backend *(*make)(const float *samples, int32_t m, int32_t d, void *ctx), void *ctx) {
    const int32_t m = std::min(n - 1, 100000);  // kSampleNum (clustering.hpp:24,75)
    std::vector<int32_t> samples(std::max(m, 1));
    std::mt19937 rng;
    gen_random(rng, samples.data(), m, n);
    std::vector<float> tr((size_t)d * std::max(m, 1));
    for (int32_t i = 0; i < m; ++i) std::memcpy(&tr[(size_t)d * i], rows + (int64_t)d * samples[i], 4 * (size_t)d);
    std::fill(centroids, centroids + (size_t)d * k, 0.0f);
    std::unique_ptr<Backend> be(make(tr.data(), m, d, ctx));
    if (!be) return GLASS_HIP_ENOMEM;
    if (kmeanspp) {  // init_kmeanspp (clustering.hpp:138-196)
        std::mt19937 r;
        std::uniform_real_distribution<> dart_dist(0, 1);
        std::uniform_int_distribution<int32_t> first(0, m - 1);
        std::vector<float> dist(std::max(m, 1)), fresh(std::max(m, 1));
        std::vector<int32_t> picked;
        const int32_t init_id = first(r);
        picked.push_back(init_id);
        std::memcpy(centroids, &tr[(size_t)d * init_id], 4 * (size_t)d);
        if (int e = be->dist_to(&tr[(size_t)d * init_id], dist.data())) return e;
        bool sum_flag = false;
        while ((int32_t)picked.size() < k) {
            double dart = dart_dist(r);
            const double sum = seed_total(dist.data(), m);
            if (sum < 1e-6) sum_flag = true;
            dart *= sum;
            double prefix = 0.0;
            int32_t pivot = 0;
            for (int32_t i = 0; i < m; ++i) {
                pivot = i;
                const double next = prefix + (double)dist[i];
                if (dart >= prefix && dart < next) break;
                prefix = next;
            }
            if (std::find(picked.begin(), picked.end(), pivot) != picked.end() && !sum_flag) continue;
            std::memcpy(centroids + (size_t)d * picked.size(), &tr[(size_t)d * pivot], 4 * (size_t)d);
            picked.push_back(pivot);
            if (int e = be->dist_to(&tr[(size_t)d * pivot], fresh.data())) return e;
            for (int32_t i = 0; i < m; ++i) dist[i] = std::min(dist[i], fresh[i]);
        }
    } else {  // init_random (clustering.hpp:127-136)
        std::mt19937 r;
        std::vector<int32_t> idx0(k);
        gen_random(r, idx0.data(), k, m);
        for (int32_t i = 0; i < k; ++i) std::memcpy(centroids + (size_t)d * i, &tr[(size_t)d * idx0[i]], 4 * (size_t)d);
    }
    // Lloyd steps: compute_dist + find_index, update_mediods (clustering.hpp:198-256)
    std::vector<int32_t> idx(std::max(m, 1)), cnt(k);
    std::vector<float> nm((size_t)d * k);
    for (int32_t e = 0; e < epochs; ++e) {
        if (int r = be->assign(centroids, k, idx.data())) return r;
        std::fill(nm.begin(), nm.end(), 0.0f);
        std::fill(cnt.begin(), cnt.end(), 0);
        for (int32_t i = 0; i < m; ++i) {
            if (idx[i] < 0) continue;  // no centroid compares less than +inf (NaN rows)
            cnt[idx[i]]++;
            float *to = &nm[(size_t)d * idx[i]];
            const float *from = &tr[(size_t)d * i];
            for (int32_t j = 0; j < d; ++j) to[j] += from[j];
        }
        for (int32_t c = 0; c < k; ++c) {
            if (cnt[c] == 0) continue;
            const float rcp = 1.0f / (float)cnt[c];
            for (int32_t j = 0; j < d; ++j) nm[(size_t)d * c + j] *= rcp;
        }
        std::copy(nm.begin(), nm.end(), centroids);
    }
    return 0;
}

}  // namespace km

namespace {

template <typename T>
struct DevArr {
    T *p = nullptr;
    DevArr() = default;
    DevArr(const DevArr &) = delete;
    DevArr &operator=(const DevArr &) = delete;
    ~DevArr() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t count) { return hipMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T)); }
};

int hip_fail(hipError_t e, const char *what) {
    std::string msg = std::string(what) + " failed: " + hipGetErrorString(e);
    return set_error(e == hipErrorOutOfMemory ? GLASS_HIP_ENOMEM : GLASS_HIP_EDEVICE, msg.c_str());
}

#define CL_TRY(expr)                                    \
    do {                                                \
        hipError_t _e = (expr);                         \
        if (_e != hipSuccess) return hip_fail(_e, #expr); \
    } while (0)

// Distance sweeps on the GPU: the sample rows stay resident, centroids and
// picked rows are copied in per call.
struct DeviceBackend : km::Backend {
    int32_t m = 0, d = 0, kmax = 0;
    DevArr<float> x, cent, row, out;
    DevArr<int32_t> ids;
    int dist_to(const float *r, float *o) override {
        CL_TRY(hipMemcpy(row.p, r, 4 * (size_t)d, hipMemcpyHostToDevice));
        CL_TRY(launch_cluster_dist(x.p, m, d, row.p, out.p, nullptr));
        CL_TRY(hipMemcpy(o, out.p, 4 * (size_t)m, hipMemcpyDeviceToHost));
        return 0;
    }
    int assign(const float *c, int32_t k, int32_t *o) override {
        CL_TRY(hipMemcpy(cent.p, c, 4 * (size_t)d * k, hipMemcpyHostToDevice));
        CL_TRY(launch_cluster_assign(x.p, m, d, cent.p, k, ids.p, nullptr));
        CL_TRY(hipMemcpy(o, ids.p, 4 * (size_t)m, hipMemcpyDeviceToHost));
        return 0;
    }
};

struct DeviceCtx {
    int32_t k;
    int error = 0;
};

km::Backend *device_backend(const float *samples, int32_t m, int32_t d, void *ctx) {
    auto *c = (DeviceCtx *)ctx;
    auto b = std::make_unique<DeviceBackend>();
    b->m = m;
    b->d = d;
    b->kmax = c->k;
    hipError_t e;
    if ((e = b->x.alloc((size_t)m * d)) != hipSuccess || (e = b->cent.alloc((size_t)c->k * d)) != hipSuccess ||
        (e = b->row.alloc(d)) != hipSuccess || (e = b->out.alloc(m)) != hipSuccess ||
        (e = b->ids.alloc(m)) != hipSuccess ||
        (e = hipMemcpy(b->x.p, samples, 4 * (size_t)m * d, hipMemcpyHostToDevice)) != hipSuccess) {
        c->error = hip_fail(e, "clustering device buffers");
        return nullptr;
    }
    return b.release();
}

// nearest centroid of every row, on the device
int predict_device(const float *cent, int32_t k, int32_t d, const float *x, int32_t n, int32_t *ids) {
    DevArr<float> dx, dc;
    DevArr<int32_t> di;
    CL_TRY(dx.alloc((size_t)n * d));
    CL_TRY(dc.alloc((size_t)k * d));
    CL_TRY(di.alloc(n));
    CL_TRY(hipMemcpy(dx.p, x, 4 * (size_t)n * d, hipMemcpyHostToDevice));
    CL_TRY(hipMemcpy(dc.p, cent, 4 * (size_t)k * d, hipMemcpyHostToDevice));
    CL_TRY(launch_cluster_assign(dx.p, n, d, dc.p, k, di.p, nullptr));
    CL_TRY(hipMemcpy(ids, di.p, 4 * (size_t)n, hipMemcpyDeviceToHost));
    return GLASS_HIP_OK;
}

}  // namespace
}  // namespace glass_hip

using namespace glass_hip;

struct glass_hip_clustering {
    int32_t k = 0, epochs = 10, dim = 0;
    bool kmeanspp = false;
    std::vector<float> centroids;
};

extern "C" {

int glass_hip_clustering_create(int32_t n_cluster, int32_t epochs, const char *init, glass_hip_clustering **out) {
    try {
        if (!out || n_cluster < 1 || epochs < 0) return set_error(GLASS_HIP_EINVAL, "bad clustering arguments");
        const std::string in = init ? init : "random";
        if (in != "random" && in != "kmeans++")
            return set_error(GLASS_HIP_EINVAL, "init must be \"random\" or \"kmeans++\"");
        auto *c = new glass_hip_clustering;
        c->k = n_cluster;
        c->epochs = epochs;
        c->kmeanspp = in == "kmeans++";
        *out = c;
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

void glass_hip_clustering_free(glass_hip_clustering *c) { delete c; }

int glass_hip_clustering_fit(glass_hip_clustering *c, const float *data, int32_t n, int32_t dim) {
    try {
        if (!c || !data || dim <= 0) return set_error(GLASS_HIP_EINVAL, "bad clustering input");
        // the fit samples min(n - 1, 100000) rows; random seeding draws k distinct ones of them
        if (n < 2 || (!c->kmeanspp && std::min(n - 1, 100000) < c->k))
            return set_error(GLASS_HIP_EINVAL, "too few sampled rows for the number of clusters");
        std::vector<float> cent((size_t)c->k * dim);
        DeviceCtx ctx{c->k};
        const int r = km::fit(data, n, dim, c->k, c->epochs, c->kmeanspp, cent.data(), device_backend, &ctx);
        if (r) return ctx.error ? ctx.error : r;
        c->centroids.swap(cent);
        c->dim = dim;
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_clustering_predict(const glass_hip_clustering *c, const float *data, int32_t n, int32_t dim,
                                 int32_t *ids) {
    try {
        if (!c || !data || !ids || n < 0) return set_error(GLASS_HIP_EINVAL, "bad clustering input");
        if (c->dim == 0) return set_error(GLASS_HIP_ESTATE, "clustering not fitted");
        if (dim != c->dim) {
            char msg[96];
            std::snprintf(msg, sizeof(msg), "data has dim %d, the centroids %d", dim, c->dim);
            return set_error(GLASS_HIP_EINVAL, msg);
        }
        if (n == 0) return GLASS_HIP_OK;
        return predict_device(c->centroids.data(), c->k, dim, data, n, ids);
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_clustering_transform(const glass_hip_clustering *c, float *data, int32_t n, int32_t dim) {
    try {
        std::vector<int32_t> ids(std::max(n, 1));
        if (int r = glass_hip_clustering_predict(c, data, n, dim, ids.data())) return r;
        for (int32_t i = 0; i < n; ++i) {
            if (ids[i] < 0) continue;  // a row with no nearest centroid is left as it is
            const float *ctr = &c->centroids[(size_t)dim * ids[i]];
            for (int32_t j = 0; j < dim; ++j) data[(int64_t)dim * i + j] -= ctr[j];
        }
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

int glass_hip_clustering_centroids(const glass_hip_clustering *c, float *out, int32_t *n_cluster, int32_t *dim) {
    try {
        if (!c) return set_error(GLASS_HIP_EINVAL, "null clustering");
        if (n_cluster) *n_cluster = c->k;
        if (dim) *dim = c->dim;
        if (out && c->dim) std::copy(c->centroids.begin(), c->centroids.end(), out);
        return GLASS_HIP_OK;
    } catch (...) {
        return exception_status();
    }
}

}  // extern "C"
