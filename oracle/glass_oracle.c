/* Test infrastructure: the oracle (CPU restatement of pyglass query-time
 * graph search).  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg load it, and only as the checker.
 *
 * Built with -fno-fast-math -ffp-contract=off: every rounding step of the
 * float paths is written out explicitly (fmaf where the compiled reference
 * fuses, a fixed reduction tree where it reduces).
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   calibration + encoders       codes equal the compiled reference's codes
 *                                (tests/golden/*.npz) and the textbook
 *                                quantile/rounding formula (numpy)
 *   integer distances            closed forms sum (x-y)^2 / -sum x*y in int64
 *   float distances              bit-equal to the compiled reference on golden
 *                                vectors; summation order shown to matter by a
 *                                planted-reorder test
 *   LinearPool / expansion       hypothesis model test against a sorted list,
 *                                brute force (ef >= N on a connected graph gives
 *                                the exact quantized kNN), and golden ids/dists
 *                                of the compiled reference Searcher
 *   HNSW initializer             hand-built 3-level toy graph + golden vectors
 *   refiner                      total order (dist, id) + golden vectors
 *   FP16 / BF16 / SQ6 / SQ4      golden codes, calibration and distances of the
 *                                compiled reference; numpy encoders; float64
 *                                closed forms of the decoded values
 *   PQ8 + k-means (Clustering)   golden codebooks, codes, LUTs, centroids; the
 *                                standard std::mt19937 known answer; float64
 *                                sub-vector distances and LUT sums
 */
#include "glass_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* fp16 / e5m2 (helpa/helpa/types.hpp:83-151 with F16C, glass/types.hpp:19-46) */

uint16_t orq_f32_to_f16(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) {
        if (ax > 0x7f800000u) return (uint16_t)(sign | 0x7e00u | ((ax >> 13) & 0x3ffu));
        return (uint16_t)(sign | 0x7c00u);
    }
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* >= 65520 rounds to inf */
    if (ax < 0x38800000u) {                                     /* below 2^-14: fp16 subnormal */
        uint32_t e = ax >> 23;
        if (e < 102) return (uint16_t)sign;
        uint32_t M = (ax & 0x7fffffu) | 0x800000u;
        int shift = 126 - (int)e;
        uint32_t q = M >> shift, rem = M & ((1u << shift) - 1u), half = 1u << (shift - 1);
        if (rem > half || (rem == half && (q & 1u))) q++;
        return (uint16_t)(sign | q);
    }
    uint32_t r = ax + 0x0fffu + ((ax >> 13) & 1u);
    r -= 0x38000000u;
    return (uint16_t)(sign | (r >> 13));
}

/* (bits + 0x8000) >> 16: round to nearest, ties away from zero in magnitude */
uint16_t orq_f32_to_bf16(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    return (uint16_t)((x + 0x8000u) >> 16);
}

static float bf16_to_f32(uint16_t h) {
    uint32_t x = (uint32_t)h << 16;
    float f;
    memcpy(&f, &x, 4);
    return f;
}

float orq_f16_to_f32(uint16_t h) {
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu, bits;
    if (e == 0) {
        if (m == 0) {
            bits = sign;
        } else { /* subnormal: normalise */
            int s = 0;
            while (!(m & 0x400u)) {
                m <<= 1;
                s++;
            }
            m &= 0x3ffu;
            bits = sign | ((uint32_t)(113 - s) << 23) | (m << 13);
        }
    } else if (e == 31) {
        bits = sign | 0x7f800000u | (m << 13);
    } else {
        bits = sign | ((e + 112u) << 23) | (m << 13);
    }
    float f;
    memcpy(&f, &bits, 4);
    return f;
}

/* ------------------------------------------------------------------ */
/* Calibration (glass/quant/calibrator.hpp:14-91, glass/quant/utils.hpp:31-70) */

static int cmp_float(const void *a, const void *b) {
    float x = *(const float *)a, y = *(const float *)b;
    return (x > y) - (x < y);
}

/* The reference keeps the `top` smallest values in a MaxHeap and returns its
 * root, i.e. the top-th smallest value (order statistic).  Restated with a
 * sort of a copy. */
static void order_stats(const float *data, int64_t nitems, int64_t top, float *kth_small, float *kth_large) {
    float *tmp = (float *)malloc(sizeof(float) * (size_t)nitems);
    memcpy(tmp, data, sizeof(float) * (size_t)nitems);
    qsort(tmp, (size_t)nitems, sizeof(float), cmp_float);
    if (top > nitems) top = nitems;
    *kth_small = tmp[top - 1];
    *kth_large = tmp[nitems - top];
    free(tmp);
}

/* find_minmax (glass/quant/utils.hpp:31-40): top = int64(float(nitems)*ratio)+1 */
static void find_minmax(const float *data, int64_t nitems, float ratio, float *mn, float *mx) {
    int64_t top = (int64_t)((float)nitems * ratio) + 1;
    order_stats(data, nitems, top, mn, mx);
}

/* find_absmax (glass/quant/utils.hpp:42-50) */
static float find_absmax(const float *data, int64_t nitems, float ratio) {
    int64_t top = (int64_t)((float)nitems * ratio) + 1;
    float *tmp = (float *)malloc(sizeof(float) * (size_t)nitems);
    for (int64_t i = 0; i < nitems; ++i) tmp[i] = fabsf(data[i]);
    float a, b;
    order_stats(tmp, nitems, top, &a, &b);
    free(tmp);
    return b;
}

int orq_quant_init(orq_quant *q, int32_t kind, int32_t metric, int32_t dim) {
    memset(q, 0, sizeof(*q));
    q->kind = kind;
    q->metric = metric;
    q->dim = dim;
    /* Quantizer<metric, AlignWidth, NBits> per quantizer
     * (sq8u_quant.hpp:14, sq4u_quant.hpp:13, sq2u_quant.hpp:11,
     *  e5m2_quant.hpp:11, sq8_quant.hpp:15-16, fp32_quant.hpp:8,
     *  fp16_quant.hpp:9-15 and bf16_quant.hpp:9-15 without the AVX512FP16 /
     *  AVX512BF16 opt-in macros) */
    switch (kind) {
        case ORQ_SQ8U: q->align = 64; q->nbits = 8; break;
        case ORQ_SQ4U: q->align = 128; q->nbits = 4; break;
        case ORQ_SQ2U: q->align = 256; q->nbits = 2; break;
        case ORQ_E5M2: q->align = 16; q->nbits = 8; break;
        case ORQ_SQ8: q->align = 16; q->nbits = 8; break;
        case ORQ_FP32: q->align = 16; q->nbits = 32; break;
        case ORQ_FP16: q->align = 16; q->nbits = 16; break;
        case ORQ_SQ6: q->align = 8; q->nbits = 6; break; /* sq6_quant.hpp:13-15 */
        case ORQ_SQ4: q->align = 16; q->nbits = 4; break; /* sq4_quant.hpp:13-15 */
        case ORQ_BF16: q->align = 16; q->nbits = 16; break;
        case ORQ_PQ8: /* ProductQuant(dim, sd = 4): Quantizer<metric, 1, 8> over nsq = dim / 4 (product_quant.hpp:10-24) */
            if (dim % 4 != 0) return -1;
            q->align = 1;
            q->nbits = 8;
            q->dim_align = q->code_size = dim / 4;
            return 0;
        default: return -1;
    }
    /* Tensor ctor (glass/storage/tensor.hpp:30-35) */
    q->dim_align = (dim + q->align - 1) / q->align * q->align;
    q->code_size = q->dim_align * q->nbits / 8;
    return 0;
}

void orq_quant_free(orq_quant *q) {
    free(q->mins);
    free(q->difs);
    free(q->codebook);
    q->mins = q->difs = NULL;
    q->codebook = NULL;
}

/* ------------------------------------------------------------------ */
/* PQ8: ProductQuant<metric> (glass/quant/product_quant.hpp) with K = 256
 * centroids per 4-dim sub-space, trained by glass::Clustering
 * (glass/algorithms/clustering.hpp: k-means++ seeding, 10 Lloyd epochs).
 *
 * Random numbers come from libstdc++ 11 (the library the reference is
 * compiled against): std::mt19937 default-seeded (5489), GenRandom
 * (glass/utils.hpp:13-27), std::uniform_int_distribution<int> (Lemire's
 * nearly divisionless method with 64-bit products, bits/uniform_int_dist.h
 * _S_nd) and std::generate_canonical<double, 53> (two draws,
 * bits/random.tcc).  Float steps follow the x86 code GCC -Ofast emits for the
 * reference (compiled here from its sources with its own flags):
 *   4-dim L2 / dot (helpa::l2_fp32_fp32 / dot_fp32_fp32 with d = 4, the
 *   scalar tail vectorised 4-wide): t_i = x_i - y_i, s_i = t_i * t_i,
 *   (s0 + s2) + (s1 + s3); dot: -((p0 + p2) + (p1 + p3))
 *   k-means++ total (double): 8 floats per step into 4 double lanes,
 *   a_l += d[8b+l] + d[8b+4+l], then (a1 + a3) + (a0 + a2), a 4-wide step
 *   (d1 + d3) + (d0 + d2), then the rest one by one
 *   new centroid = sum * (1.0f / float(count))   (reciprocal math)
 *   LUT sum: see orq_pq_lut_sum. */

typedef struct {
    uint32_t mt[624];
    int32_t i;
} mt19937_t;

static void mt_seed(mt19937_t *g, uint32_t seed) {
    g->mt[0] = seed;
    for (int i = 1; i < 624; ++i) g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->i = 624;
}

static uint32_t mt_next(mt19937_t *g) {
    if (g->i >= 624) {
        for (int k = 0; k < 624; ++k) {
            const uint32_t y = (g->mt[k] & 0x80000000u) | (g->mt[(k + 1) % 624] & 0x7fffffffu);
            g->mt[k] = g->mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        g->i = 0;
    }
    uint32_t y = g->mt[g->i++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

void orq_mt19937_outputs(int32_t count, uint32_t *out) {
    mt19937_t g;
    mt_seed(&g, 5489u);
    for (int32_t i = 0; i < count; ++i) out[i] = mt_next(&g);
}

/* uniform_int_distribution<int>(a, b) */
static int32_t uniform_int(mt19937_t *g, int32_t a, int32_t b) {
    const uint32_t range = (uint32_t)((int64_t)b - a) + 1u;
    uint64_t product = (uint64_t)mt_next(g) * range;
    uint32_t low = (uint32_t)product;
    if (low < range) {
        const uint32_t threshold = (0u - range) % range;
        while (low < threshold) {
            product = (uint64_t)mt_next(g) * range;
            low = (uint32_t)product;
        }
    }
    return a + (int32_t)(product >> 32);
}

/* uniform_real_distribution<double>(0, 1) = generate_canonical<double, 53> */
static double canonical(mt19937_t *g) {
    const double x1 = (double)mt_next(g);
    const double x2 = (double)mt_next(g);
    const double r = fma(x2, 4294967296.0, x1) * 0x1p-64;
    return r >= 1.0 ? nextafter(1.0, 0.0) : r;
}

/* GenRandom (glass/utils.hpp:13-27) */
static void gen_random(mt19937_t *g, int32_t *addr, int32_t size, int32_t N) {
    for (int32_t i = 0; i < size; ++i) addr[i] = (int32_t)(mt_next(g) % (uint32_t)(N - size));
    for (int32_t i = 1; i < size; ++i) { /* std::sort */
        int32_t v = addr[i], j = i - 1;
        while (j >= 0 && addr[j] > v) {
            addr[j + 1] = addr[j];
            --j;
        }
        addr[j + 1] = v;
    }
    for (int32_t i = 1; i < size; ++i)
        if (addr[i] <= addr[i - 1]) addr[i] = addr[i - 1] + 1;
    const int32_t off = (int32_t)(mt_next(g) % (uint32_t)N);
    for (int32_t i = 0; i < size; ++i) addr[i] = (addr[i] + off) % N;
}

static float l2_4(const float *x, const float *y) {
    float s[4];
    for (int i = 0; i < 4; ++i) {
        const float t = x[i] - y[i];
        s[i] = t * t;
    }
    return (s[0] + s[2]) + (s[1] + s[3]);
}

static float dot_4(const float *x, const float *y) {
    float p[4];
    for (int i = 0; i < 4; ++i) p[i] = x[i] * y[i];
    return -((p[0] + p[2]) + (p[1] + p[3]));
}

static double kpp_total(const float *d, int32_t n) {
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

static float reduce16_helpa(const float a[16]);

/* helpa::l2_fp32_fp32 for any d (x86/l2_impl.hpp:14-18,50-71 with
 * ref/l2_ref.hpp:7-15), as GCC -Ofast compiles it: the 16-lane AVX-512 body
 * with fma accumulators and helpa's (c0+c2)+(c1+c3) tree; the scalar tail
 * vectorised as one 8-wide step (squares rounded, u_i = s_(i+4) + s_i, then
 * (u1 + u3) + (u0 + u2)), one 4-wide step ((s1 + s3) + (s0 + s2), added),
 * then fused multiply-adds; body + tail. */
static float l2_helpa(const float *x, const float *y, int32_t d) {
    const int32_t da = d / 16 * 16;
    float A = 0.0f;
    if (da > 0) {
        float a[16] = {0};
        for (int32_t i = 0; i < da; i += 16)
            for (int l = 0; l < 16; ++l) {
                const float t = x[i + l] - y[i + l];
                a[l] = fmaf(t, t, a[l]);
            }
        A = reduce16_helpa(a);
    }
    float T = 0.0f;
    int32_t i = da;
    if (d - i >= 8) {
        float sq[8], u[4];
        for (int l = 0; l < 8; ++l) {
            const float t = x[i + l] - y[i + l];
            sq[l] = t * t;
        }
        for (int l = 0; l < 4; ++l) u[l] = sq[l + 4] + sq[l];
        T = (u[1] + u[3]) + (u[0] + u[2]);
        i += 8;
    }
    if (d - i >= 4) {
        float sq[4];
        for (int l = 0; l < 4; ++l) {
            const float t = x[i + l] - y[i + l];
            sq[l] = t * t;
        }
        T = T + ((sq[1] + sq[3]) + (sq[0] + sq[2]));
        i += 4;
    }
    for (; i < d; ++i) {
        const float t = x[i] - y[i];
        T = fmaf(t, t, T);
    }
    return A + T;
}

/* find_index over compute_dist (clustering.hpp:198-232): the first centroid
 * at least distance; ties keep the lower index. */
static int32_t nearest(const float *x, const float *medoids, int32_t k, int32_t d) {
    float best = HUGE_VALF;
    int32_t bid = -1;
    for (int32_t j = 0; j < k; ++j) {
        const float v = l2_helpa(x, medoids + (int64_t)j * d, d);
        if (v < best) {
            best = v;
            bid = j;
        }
    }
    return bid;
}

/* Clustering::fit(data, n, dim) (clustering.hpp:73-100) with k clusters,
 * kmeans++ (init_kmeanspp, clustering.hpp:138-196) or random
 * (init_random, clustering.hpp:127-136) seeding, then `epochs` Lloyd steps. */
void orq_clustering_fit(const float *rows, int32_t n, int32_t d, int32_t k, int32_t epochs, int32_t kmeanspp,
                        float *medoids) {
    const int32_t m = n - 1 < 100000 ? n - 1 : 100000; /* kSampleNum */
    int32_t *samples = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
    mt19937_t g;
    mt_seed(&g, 5489u);
    gen_random(&g, samples, m, n);
    float *tr = (float *)malloc(sizeof(float) * (size_t)d * (size_t)(m > 0 ? m : 1));
    for (int32_t i = 0; i < m; ++i) memcpy(tr + (int64_t)d * i, rows + (int64_t)d * samples[i], sizeof(float) * d);
    memset(medoids, 0, sizeof(float) * (size_t)d * (size_t)k);
    if (kmeanspp) {
        mt19937_t r;
        mt_seed(&r, 5489u);
        float *dist = (float *)malloc(sizeof(float) * (size_t)(m > 0 ? m : 1));
        int32_t *picked = (int32_t *)malloc(sizeof(int32_t) * (size_t)k);
        const int32_t init_id = uniform_int(&r, 0, m - 1);
        int32_t npicked = 1;
        picked[0] = init_id;
        memcpy(medoids, tr + (int64_t)d * init_id, sizeof(float) * d);
        for (int32_t i = 0; i < m; ++i) dist[i] = l2_helpa(tr + (int64_t)d * init_id, tr + (int64_t)d * i, d);
        int sum_flag = 0;
        while (npicked < k) {
            double dart = canonical(&r);
            const double sum = kpp_total(dist, m);
            if (sum < 1e-6) sum_flag = 1;
            dart *= sum;
            double prefix = 0.0;
            int32_t pivot = 0;
            for (int32_t i = 0; i < m; ++i) {
                pivot = i;
                const double next = prefix + (double)dist[i];
                if (dart >= prefix && dart < next) break;
                prefix = next;
            }
            int dup = 0;
            for (int32_t j = 0; j < npicked; ++j) dup |= picked[j] == pivot;
            if (dup && !sum_flag) continue;
            picked[npicked] = pivot;
            memcpy(medoids + (int64_t)d * npicked, tr + (int64_t)d * pivot, sizeof(float) * d);
            for (int32_t i = 0; i < m; ++i) {
                const float v = l2_helpa(tr + (int64_t)d * pivot, tr + (int64_t)d * i, d);
                dist[i] = v < dist[i] ? v : dist[i]; /* std::min */
            }
            npicked++;
        }
        free(dist);
        free(picked);
    } else {
        mt19937_t r;
        mt_seed(&r, 5489u);
        int32_t *idx0 = (int32_t *)malloc(sizeof(int32_t) * (size_t)k);
        gen_random(&r, idx0, k, m);
        for (int32_t i = 0; i < k; ++i) memcpy(medoids + (int64_t)d * i, tr + (int64_t)d * idx0[i], sizeof(float) * d);
        free(idx0);
    }
    int32_t *idx = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
    int32_t *cnt = (int32_t *)malloc(sizeof(int32_t) * (size_t)k);
    float *nm = (float *)malloc(sizeof(float) * (size_t)d * (size_t)k);
    for (int32_t e = 0; e < epochs; ++e) {
#pragma omp parallel for schedule(static)
        for (int32_t i = 0; i < m; ++i) idx[i] = nearest(tr + (int64_t)d * i, medoids, k, d);
        memset(nm, 0, sizeof(float) * (size_t)d * (size_t)k);
        memset(cnt, 0, sizeof(int32_t) * (size_t)k);
        for (int32_t i = 0; i < m; ++i) {
            cnt[idx[i]]++;
            for (int32_t j = 0; j < d; ++j) nm[(int64_t)d * idx[i] + j] += tr[(int64_t)d * i + j];
        }
        for (int32_t c = 0; c < k; ++c) {
            if (cnt[c] == 0) continue;
            const float rcp = 1.0f / (float)cnt[c];
            for (int32_t j = 0; j < d; ++j) nm[(int64_t)d * c + j] *= rcp;
        }
        memcpy(medoids, nm, sizeof(float) * (size_t)d * (size_t)k);
    }
    free(samples);
    free(tr);
    free(idx);
    free(cnt);
    free(nm);
}

/* Clustering::predict / transform (clustering.hpp:102-129) */
void orq_clustering_predict(const float *medoids, int32_t k, int32_t d, const float *x, int32_t n, int32_t *ids) {
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) ids[i] = nearest(x + (int64_t)d * i, medoids, k, d);
}

void orq_clustering_transform(const float *medoids, int32_t k, int32_t d, float *x, int32_t n) {
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) {
        const float *c = medoids + (int64_t)d * nearest(x + (int64_t)d * i, medoids, k, d);
        for (int32_t j = 0; j < d; ++j) x[(int64_t)d * i + j] -= c[j];
    }
}

/* the 4-dim fit ProductQuant::train runs per sub-space */
void orq_kmeans_fit4(const float *rows, int32_t n, int32_t k, int32_t epochs, float *medoids) {
    orq_clustering_fit(rows, n, 4, k, epochs, 1, medoids);
}

/* ProductQuant::dist_func, sum over nsq LUT entries as GCC -Ofast vectorises
 * it (AVX2 gathers): blocks of 32 into 8 lanes, acc_l += (G(l) + G(8+l)) +
 * (G(16+l) + G(24+l)), then h_i = acc_i + acc_(i+4), (h0 + h2) + (h1 + h3);
 * one block of 16 into 4 lanes, v_i = (G(i) + G(4+i)) + (G(8+i) + G(12+i)),
 * r += (v1 + v3) + (v0 + v2); then r += G(s) one by one. */
float orq_pq_lut_sum(const float *lut, const uint8_t *code, int32_t nsq) {
#define PQ_G(s) lut[(int64_t)(s) * 256 + code[(s)]]
    float r = 0.0f;
    int32_t c = 0;
    if (nsq >= 32) {
        float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (; c + 32 <= nsq; c += 32)
            for (int l = 0; l < 8; ++l)
                acc[l] = acc[l] + ((PQ_G(c + l) + PQ_G(c + 8 + l)) + (PQ_G(c + 16 + l) + PQ_G(c + 24 + l)));
        float h[4];
        for (int i = 0; i < 4; ++i) h[i] = acc[i] + acc[i + 4];
        r = (h[0] + h[2]) + (h[1] + h[3]);
    }
    if (nsq - c >= 16) {
        float v[4];
        for (int i = 0; i < 4; ++i) v[i] = (PQ_G(c + i) + PQ_G(c + 4 + i)) + (PQ_G(c + 8 + i) + PQ_G(c + 12 + i));
        r = r + ((v[1] + v[3]) + (v[0] + v[2]));
        c += 16;
    }
    for (; c < nsq; ++c) r = r + PQ_G(c);
    return r;
#undef PQ_G
}

/* ProductQuant::train: one Clustering per 4-dim sub-space (product_quant.hpp:26-41) */
static void pq_train(orq_quant *q, const float *data, int64_t n) {
    const int32_t nsq = q->dim / 4;
    q->codebook = (float *)calloc((size_t)nsq * 256 * 4, sizeof(float));
#pragma omp parallel for schedule(dynamic)
    for (int32_t sq = 0; sq < nsq; ++sq) {
        float *rows = (float *)malloc(sizeof(float) * 4 * (size_t)n);
        for (int64_t i = 0; i < n; ++i) memcpy(rows + 4 * i, data + i * q->dim + 4 * sq, 16);
        orq_kmeans_fit4(rows, (int32_t)n, 256, 10, q->codebook + (int64_t)sq * 256 * 4);
        free(rows);
    }
}

int orq_quant_train(orq_quant *q, const float *data, int64_t n) {
    int64_t nd = n * (int64_t)q->dim;
    if (q->kind == ORQ_PQ8) {
        pq_train(q, data, n);
        return 0;
    }
    float mn, mx;
    switch (q->kind) {
        case ORQ_SQ8U:
            if (q->metric == ORQ_L2) { /* AffineCalibrator<127>, drop 0 (sq8u_quant.hpp:19-28) */
                find_minmax(data, nd, 0.0f, &mn, &mx);
                q->min = mn;
                q->dif = mx - mn;
            } else { /* SymCalibrator<127> */
                q->max = find_absmax(data, nd, 0.0f);
            }
            break;
        case ORQ_SQ4U: { /* Affine<15>, first 10% of items, drop 1% (sq4u_quant.hpp:85-97) */
            int64_t nitems = (int64_t)((float)nd * 0.1f);
            find_minmax(data, nitems, 0.01f, &mn, &mx);
            q->min = mn;
            q->dif = mx - mn;
            break;
        }
        case ORQ_SQ2U: { /* Affine<3>, first 10% of items, drop 5% (sq2u_quant.hpp:164-176) */
            int64_t nitems = (int64_t)((float)nd * 0.1f);
            find_minmax(data, nitems, 0.05f, &mn, &mx);
            q->min = mn;
            q->dif = mx - mn;
            break;
        }
        case ORQ_SQ8: {
            /* AffinePerDimCalibrator<255> via find_minmax_perdim
             * (glass/quant/utils.hpp:52-70): "dimension" j is fed the
             * flattened items [j*n, (j+1)*n), not column j. */
            q->mins = (float *)calloc((size_t)q->dim_align, sizeof(float));
            q->difs = (float *)calloc((size_t)q->dim_align, sizeof(float));
            for (int32_t j = 0; j < q->dim; ++j) {
                const float *p = data + (int64_t)j * n;
                float lo = p[0], hi = p[0];
                for (int64_t i = 1; i < n; ++i) {
                    if (p[i] < lo) lo = p[i];
                    if (p[i] > hi) hi = p[i];
                }
                q->mins[j] = lo;
                q->difs[j] = hi - lo;
            }
            break;
        }
        case ORQ_SQ6:
        case ORQ_SQ4: {
            /* AffinePerDimCalibrator<63> (SQ6) / <15> (SQ4) with drop ratio
             * 0.01 through find_minmax_perdim (sq6_quant.hpp:17-31,
             * sq4_quant.hpp:17-31, utils.hpp:52-70):
             * chunk j = flattened items [j*n, (j+1)*n), its top-th smallest
             * and largest values, top = int64(n * 0.01f) + 1; padding dims
             * keep min = dif = 0 (mins/difs resized to dim_align). */
            q->mins = (float *)calloc((size_t)q->dim_align, sizeof(float));
            q->difs = (float *)calloc((size_t)q->dim_align, sizeof(float));
            for (int32_t j = 0; j < q->dim; ++j) {
                float lo, hi;
                find_minmax(data + (int64_t)j * n, n, 0.01f, &lo, &hi);
                q->mins[j] = lo;
                q->difs[j] = hi - lo;
            }
            break;
        }
        default:
            break;
    }
    return 0;
}

/* limit_range / limit_range_sym (glass/quant/utils.hpp:8-26) + std::round */
static int32_t affine_code(float x, float mn, float dif, int32_t mul) {
    x = (x - mn) / dif;
    if (x < 0.0f) x = 0.0f;
    if (x > 1.0f) x = 1.0f;
    return (int32_t)roundf(x * (float)mul);
}

static int32_t sym_code(float x, float mx, int32_t mul) {
    x = x / mx;
    if (x < -1.0f) x = -1.0f;
    if (x > 1.0f) x = 1.0f;
    return (int32_t)roundf(x * (float)mul);
}

void orq_encode(const orq_quant *q, const float *x, uint8_t *out) {
    memset(out, 0, (size_t)q->code_size);
    int32_t d = q->dim;
    switch (q->kind) {
        case ORQ_PQ8: /* ProductQuant::encode (product_quant.hpp:52-64): first nearest centroid */
            for (int32_t sq = 0; sq < q->code_size; ++sq) {
                float best = FLT_MAX;
                for (int32_t k = 0; k < 256; ++k) {
                    const float v = l2_4(x + 4 * sq, q->codebook + ((int64_t)sq * 256 + k) * 4);
                    if (v < best) {
                        best = v;
                        out[sq] = (uint8_t)k;
                    }
                }
            }
            break;
        case ORQ_SQ8U: /* sq8u_quant.hpp:38-42 */
            for (int32_t j = 0; j < d; ++j) {
                int32_t c = q->metric == ORQ_L2 ? affine_code(x[j], q->min, q->dif, 127) : sym_code(x[j], q->max, 127);
                out[j] = (uint8_t)(int8_t)c;
            }
            break;
        case ORQ_SQ4U: /* sq4u_quant.hpp:107-117: dim j -> nibble (j&1 ? hi : lo) of byte j/2 */
            for (int32_t j = 0; j < d; ++j) {
                uint8_t y = (uint8_t)affine_code(x[j], q->min, q->dif, 15);
                out[j / 2] |= (uint8_t)((j & 1) ? (y << 4) : y);
            }
            break;
        case ORQ_SQ2U: /* sq2u_quant.hpp:186-200: dim j -> bits 2*(j%4) of byte j/4 */
            for (int32_t j = 0; j < d; ++j) {
                uint8_t y = (uint8_t)affine_code(x[j], q->min, q->dif, 3);
                out[j / 4] |= (uint8_t)(y << (2 * (j % 4)));
            }
            break;
        case ORQ_E5M2: /* e5m2_quant.hpp:30-34: high byte of fp16(x) */
            for (int32_t j = 0; j < d; ++j) out[j] = (uint8_t)(orq_f32_to_f16(x[j]) >> 8);
            break;
        case ORQ_SQ8: /* sq8_quant.hpp:43-47 */
            for (int32_t j = 0; j < d; ++j) out[j] = (uint8_t)affine_code(x[j], q->mins[j], q->difs[j], 255);
            break;
        case ORQ_FP32:
            memcpy(out, x, sizeof(float) * (size_t)d);
            break;
        case ORQ_FP16: { /* FP16Quantizer::encode -> helpa::fp32_to_fp16 (RNE), padding as below */
            uint16_t *o = (uint16_t *)out;
            for (int32_t j = 0; j < d; ++j) o[j] = orq_f32_to_f16(x[j]);
            break;
        }
        case ORQ_SQ4: /* SQ4Quantizer::encode (sq4_quant.hpp:42-52): dim j -> nibble (j&1 ? hi : lo) of byte j/2 */
            for (int32_t j = 0; j < d; ++j) {
                const int32_t c = affine_code(x[j], q->mins[j], q->difs[j], 15);
                out[j >> 1] |= (uint8_t)((j & 1) ? c << 4 : c);
            }
            break;
        case ORQ_SQ6: /* SQ6Quantizer::encode_component (sq6_quant.hpp:41-62): 4 components per 3 bytes */
            for (int32_t j = 0; j < d; ++j) {
                const int32_t c = affine_code(x[j], q->mins[j], q->difs[j], 63);
                uint8_t *g = out + (j >> 2) * 3;
                switch (j & 3) {
                    case 0: g[0] |= (uint8_t)c; break;
                    case 1: g[0] |= (uint8_t)(c << 6); g[1] |= (uint8_t)(c >> 2); break;
                    case 2: g[1] |= (uint8_t)(c << 4); g[2] |= (uint8_t)(c >> 4); break;
                    default: g[2] |= (uint8_t)(c << 2); break;
                }
            }
            break;
        case ORQ_BF16: { /* BF16Quantizer::encode (bf16_quant.hpp:34-38), padding stays zero */
            uint16_t *o = (uint16_t *)out;
            for (int32_t j = 0; j < d; ++j) o[j] = orq_f32_to_bf16(x[j]);
            break;
        }
    }
}

void orq_encode_rows(const orq_quant *q, const float *data, int64_t n, uint8_t *codes) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        orq_encode(q, data + i * q->dim, codes + i * q->code_size);
        if (q->kind == ORQ_FP16) {
            /* helpa::fp32_to_fp16 (x86/data_convert.hpp:18-31) converts whole
             * 16-float blocks of the row, the last one included, so the
             * padding [dim, dim_align) holds fp16 of the floats that follow
             * the row in the input (the next row's first values).  Past the
             * end of the input the reference reads unowned memory; zeros here. */
            uint16_t *o = (uint16_t *)(codes + i * q->code_size);
            const int64_t total = n * (int64_t)q->dim;
            for (int32_t j = q->dim; j < q->dim_align; ++j) {
                const int64_t at = i * (int64_t)q->dim + j;
                o[j] = at < total ? orq_f32_to_f16(data[at]) : 0;
            }
        }
    }
}

/* ComputerImpl ctor (glass/quant/computer.hpp:54-63): integer quantizers
 * encode the query with the base calibrator; E5M2/SQ8/FP32 copy it (MemCpyTag)
 * into a zeroed dim_align float buffer. */
int64_t orq_query_size(const orq_quant *q) {
    if (q->kind == ORQ_SQ8U || q->kind == ORQ_SQ4U || q->kind == ORQ_SQ2U) return q->code_size;
    if (q->kind == ORQ_PQ8) return (int64_t)q->code_size * 256 * 4; /* the LUT */
    return (int64_t)q->dim_align * 4;
}

void orq_prepare_query(const orq_quant *q, const float *x, uint8_t *qbuf) {
    if (q->kind == ORQ_SQ8U || q->kind == ORQ_SQ4U || q->kind == ORQ_SQ2U) {
        orq_encode(q, x, qbuf);
    } else if (q->kind == ORQ_PQ8) { /* ProductQuant::get_lut (product_quant.hpp:72-82) */
        float *lut = (float *)qbuf;
        for (int32_t sq = 0; sq < q->code_size; ++sq)
            for (int32_t k = 0; k < 256; ++k) {
                const float *c = q->codebook + ((int64_t)sq * 256 + k) * 4;
                lut[(int64_t)sq * 256 + k] = q->metric == ORQ_L2 ? l2_4(x + 4 * sq, c) : dot_4(x + 4 * sq, c);
            }
    } else {
        memset(qbuf, 0, (size_t)q->dim_align * 4);
        memcpy(qbuf, x, sizeof(float) * (size_t)q->dim);
    }
}

/* ------------------------------------------------------------------ */
/* Distances */

/* glass::reduce_add_f32x16 (glass/simd/avx512_utils.hpp:12-22):
 * 16 -> 8 -> 4 by halves, then hadd: (c0+c1)+(c2+c3). */
static float reduce16_glass(const float a[16]) {
    float b[8], c[4];
    for (int i = 0; i < 8; ++i) b[i] = a[i] + a[i + 8];
    for (int i = 0; i < 4; ++i) c[i] = b[i] + b[i + 4];
    return (c[0] + c[1]) + (c[2] + c[3]);
}

/* helpa reduce_add_f32x16 = GCC _mm512_reduce_add_ps
 * (helpa/platform/x86/utils.hpp:61-64; avx512fintrin.h __MM512_REDUCE_OP):
 * 16 -> 8 -> 4 by halves, then (c0+c2)+(c1+c3). */
static float reduce16_helpa(const float a[16]) {
    float b[8], c[4];
    for (int i = 0; i < 8; ++i) b[i] = a[i + 8] + a[i];
    for (int i = 0; i < 4; ++i) c[i] = b[i + 4] + b[i];
    return (c[0] + c[2]) + (c[1] + c[3]);
}


/* ctx selects the FMA contraction GCC -Ofast chose where the reference's
 * SQ8 L2 kernel is inlined: ORQ_CTX_COMPUTER (a bare computer call) and
 * ORQ_CTX_SEARCH (SearchImpl1's expansions) fuse as fma(-c, dif, 256x -
 * 256min); ORQ_CTX_BATCH (SearchBatch/SearchImpl2) and ORQ_CTX_INIT
 * (initialize_search, graph.hpp:95-104, in Search as in SearchBatch) fuse
 * yy = fma(c, dif, 256min), t = 256x - yy.  Pinned by the golden file's
 * single-query distances: with the initializer's evaluations in the second
 * form, all of them are bit-equal; in the first form, 1 in 270 HNSW and NSG
 * results (the entry points and the upper-layer descent's last node) is one
 * ulp off.  SQ6 below splits the same way; every other kernel compiles the
 * same in all contexts. */
double orq_distance_ctx(const orq_quant *q, const uint8_t *qbuf, const uint8_t *code, int32_t ctx) {
    int32_t da = q->dim_align;
    switch (q->kind) {
        case ORQ_PQ8: /* ProductQuant::dist_func (product_quant.hpp:84-90) over the query's LUT */
            return (double)orq_pq_lut_sum((const float *)qbuf, code, q->code_size);
        case ORQ_SQ8U: {
            const int8_t *x = (const int8_t *)qbuf, *y = (const int8_t *)code;
            int64_t s = 0;
            if (q->metric == ORQ_L2) { /* helpa::l2_s7_s7 (x86/l2_impl.hpp:44-48,222-255) */
                for (int32_t j = 0; j < da; ++j) {
                    int32_t t = (int32_t)x[j] - (int32_t)y[j];
                    s += t * t;
                }
                return (double)(int32_t)s;
            }
            /* helpa::dot_s8_s8 (x86/dot_impl.hpp:46-50,210-229) */
            for (int32_t j = 0; j < da; ++j) s += (int32_t)x[j] * (int32_t)y[j];
            return (double)(int32_t)(-s);
        }
        case ORQ_SQ4U: { /* helpa::l2a_u4_u4 (x86/l2_impl.hpp:257-298) */
            int64_t s = 0;
            for (int32_t j = 0; j < da / 2; ++j) {
                int32_t a = qbuf[j] & 15, b = code[j] & 15, c = qbuf[j] >> 4, e = code[j] >> 4;
                s += (a - b) * (a - b) + (c - e) * (c - e);
            }
            return (double)(int32_t)s;
        }
        case ORQ_SQ2U: { /* helpa::l2a_u2_u2 (x86/l2_impl.hpp:300-367) */
            int64_t s = 0;
            for (int32_t j = 0; j < da / 4; ++j) {
                for (int sh = 0; sh < 8; sh += 2) {
                    int32_t a = (qbuf[j] >> sh) & 3, b = (code[j] >> sh) & 3;
                    s += (a - b) * (a - b);
                }
            }
            return (double)(int32_t)s;
        }
        case ORQ_E5M2: { /* glass::L2SqrE5M2 / IPE5M2 (glass/simd/avx512.hpp:15-38) */
            const float *x = (const float *)qbuf;
            float acc[16] = {0};
            for (int32_t i = 0; i < da; i += 16) {
                for (int l = 0; l < 16; ++l) {
                    float y = orq_f16_to_f32((uint16_t)(code[i + l] << 8));
                    if (q->metric == ORQ_L2) {
                        float t = x[i + l] - y;
                        acc[l] = fmaf(t, t, acc[l]);
                    } else {
                        acc[l] = fmaf(x[i + l], y, acc[l]);
                    }
                }
            }
            float r = reduce16_glass(acc);
            return q->metric == ORQ_L2 ? (double)r : (double)(-r);
        }
        case ORQ_SQ8: { /* glass::L2SqrSQ8_ext / IPSQ8_ext (glass/simd/avx512.hpp:40-77) */
            const float *x = (const float *)qbuf;
            float acc[16] = {0};
            for (int32_t i = 0; i < da; i += 16) {
                for (int l = 0; l < 16; ++l) {
                    float c = (float)code[i + l] + 0.5f;
                    if (q->metric == ORQ_L2) {
                        float t;
                        if ((ctx & 3) == ORQ_CTX_BATCH || (ctx & ORQ_CTX_INIT)) {
                            t = x[i + l] * 256.0f - fmaf(c, q->difs[i + l], q->mins[i + l] * 256.0f);
                        } else {
                            t = fmaf(-c, q->difs[i + l], x[i + l] * 256.0f - q->mins[i + l] * 256.0f);
                        }
                        acc[l] = fmaf(t, t, acc[l]);
                    } else {
                        float yy = fmaf(c, q->difs[i + l], q->mins[i + l] * 256.0f);
                        acc[l] = fmaf(x[i + l], yy, acc[l]);
                    }
                }
            }
            float r = reduce16_glass(acc);
            return q->metric == ORQ_L2 ? (double)r : (double)(-r);
        }
        case ORQ_SQ4: { /* glass::L2SqrSQ4_ext / IPSQ4_ext (glass/simd/avx512.hpp:106-144) */
            /* A scalar loop over element pairs inside FAST_BEGIN (associative
             * math).  The optimize pragma keeps GCC from inlining it, so every
             * context calls the same code, which GCC -Ofast compiles (x86
             * code of the reference built with its own flags, checked against
             * its golden distances) as one serial sum of per-pair terms, with
             * a = c + 0.5:
             *   L2  t = fma(-a, dif, (x - min) * 16),  p = fma(t0, t0, t1 * t1)
             *   IP  y = fma(a, dif * (1/16), min),      p = fma(x0, y0, x1 * y1)
             *   sum = ((p_0 + p_1) + p_2) + ...;  IP returns -sum. */
            const float *x = (const float *)qbuf;
            float sum = 0.0f;
            for (int32_t i = 0; i < da; i += 2) {
                const float a0 = (float)(code[i >> 1] & 15) + 0.5f, a1 = (float)(code[i >> 1] >> 4) + 0.5f;
                float p;
                if (q->metric == ORQ_L2) {
                    const float t0 = fmaf(-a0, q->difs[i], (x[i] - q->mins[i]) * 16.0f);
                    const float t1 = fmaf(-a1, q->difs[i + 1], (x[i + 1] - q->mins[i + 1]) * 16.0f);
                    p = fmaf(t0, t0, t1 * t1);
                } else {
                    const float y0 = fmaf(a0, q->difs[i] * 0.0625f, q->mins[i]);
                    const float y1 = fmaf(a1, q->difs[i + 1] * 0.0625f, q->mins[i + 1]);
                    p = fmaf(x[i], y0, x[i + 1] * y1);
                }
                sum = i == 0 ? p : sum + p;
            }
            return q->metric == ORQ_L2 ? (double)sum : (double)(-sum);
        }
        case ORQ_SQ6: { /* glass::L2SqrSQ6_ext / IPSQ6_ext (glass/simd/avx512.hpp:79-104), 8 accumulators */
            const float *x = (const float *)qbuf;
            float acc[8] = {0};
            for (int32_t i = 0; i < da; i += 8) {
                for (int l = 0; l < 8; ++l) {
                    const int32_t j = i + l;
                    const uint8_t *g = code + (j >> 2) * 3;
                    int32_t c;
                    switch (j & 3) { /* SQ6::load6 (avx2_utils.hpp:32-49) */
                        case 0: c = g[0] & 63; break;
                        case 1: c = (g[0] >> 6) | ((g[1] & 15) << 2); break;
                        case 2: c = (g[1] >> 4) | ((g[2] & 3) << 4); break;
                        default: c = g[2] >> 2; break;
                    }
                    /* GCC -Ofast reassociates the decode ((c + 0.5) / 63) * dif
                     * + mi per inlining context (x86 code GCC emits for
                     * glass/simd/avx512.hpp:79-104 with the reference's flags,
                     * checked against its golden distances).  With a = c + 0.5:
                     * a bare computer and Search's expansions scale dif first,
                     * kd = (1/63) * dif: L2 t = fma(-a, kd, x - mi), IP yy =
                     * fma(a, kd, mi); SearchBatch and initialize_search scale a
                     * first, ad = a * dif: L2 t = fma(-ad, 1/63, x - mi), IP
                     * yy = fma(ad, 1/63, mi). */
                    const float a = (float)c + 0.5f;
                    const int ad_form = (ctx & ORQ_CTX_INIT) || (ctx & 3) == ORQ_CTX_BATCH;
                    if (q->metric == ORQ_L2) {
                        const float t = ad_form ? fmaf(-(a * q->difs[j]), 1.0f / 63.0f, x[j] - q->mins[j])
                                                : fmaf(-a, (1.0f / 63.0f) * q->difs[j], x[j] - q->mins[j]);
                        acc[l] = fmaf(t, t, acc[l]);
                    } else {
                        const float yy = ad_form ? fmaf(a * q->difs[j], 1.0f / 63.0f, q->mins[j])
                                                 : fmaf(a, (1.0f / 63.0f) * q->difs[j], q->mins[j]);
                        acc[l] = fmaf(x[j], yy, acc[l]);
                    }
                }
            }
            /* glass::reduce_add_f32x8 (avx2_utils.hpp:11-16): h_i = a_i + a_(i+4), (h0 + h2) + (h1 + h3) */
            float h[4];
            for (int i = 0; i < 4; ++i) h[i] = acc[i] + acc[i + 4];
            const float r = (h[0] + h[2]) + (h[1] + h[3]);
            return q->metric == ORQ_L2 ? (double)r : (double)(-r);
        }
        case ORQ_FP32:   /* helpa::l2_fp32_fp32 / dot_fp32_fp32 (x86/l2_impl.hpp:14-18,50-71; x86/dot_impl.hpp:16-20,52-71) */
        case ORQ_FP16:   /* helpa::l2_fp32_fp16 / dot_fp32_fp16 (x86/l2_impl.hpp:20-24,74-84; x86/dot_impl.hpp:22-26,74-84) */
        case ORQ_BF16: { /* helpa::l2_fp32_bf16 / dot_fp32_bf16 (x86/l2_impl.hpp:32-36,145-156; x86/dot_impl.hpp:34-38,139-149) */
            const float *x = (const float *)qbuf;
            float acc[16] = {0};
            for (int32_t i = 0; i < da; i += 16) {
                for (int l = 0; l < 16; ++l) {
                    float y;
                    if (q->kind == ORQ_FP32) {
                        y = ((const float *)code)[i + l];
                    } else if (q->kind == ORQ_FP16) {
                        y = orq_f16_to_f32(((const uint16_t *)code)[i + l]);
                    } else {
                        y = bf16_to_f32(((const uint16_t *)code)[i + l]);
                    }
                    if (q->metric == ORQ_L2) {
                        float t = x[i + l] - y;
                        acc[l] = fmaf(t, t, acc[l]);
                    } else {
                        acc[l] = fmaf(x[i + l], y, acc[l]);
                    }
                }
            }
            float r = reduce16_helpa(acc);
            return q->metric == ORQ_L2 ? (double)r : (double)(-r);
        }
    }
    return 0.0;
}

double orq_distance(const orq_quant *q, const uint8_t *qbuf, const uint8_t *code) {
    return orq_distance_ctx(q, qbuf, code, ORQ_CTX_COMPUTER);
}

/* ------------------------------------------------------------------ */
/* Graph I/O (glass/graph.hpp:106-153, glass/hnsw/HNSWInitializer.hpp:64-86) */

void orq_graph_free(orq_graph *g) {
    free(g->adj);
    free(g->eps);
    free(g->levels);
    free(g->offsets);
    free(g->lists);
    memset(g, 0, sizeof(*g));
}

int orq_graph_from_arrays(orq_graph *g, int32_t N, int32_t K, const int32_t *adj, int32_t nep, const int32_t *eps) {
    memset(g, 0, sizeof(*g));
    g->N = N;
    g->K = K;
    g->adj = (int32_t *)malloc(sizeof(int32_t) * (size_t)N * (size_t)K);
    memcpy(g->adj, adj, sizeof(int32_t) * (size_t)N * (size_t)K);
    g->nep = nep;
    g->eps = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nep > 0 ? nep : 1));
    if (nep > 0) memcpy(g->eps, eps, sizeof(int32_t) * (size_t)nep);
    return 0;
}

int orq_graph_load(orq_graph *g, const char *path) {
    memset(g, 0, sizeof(*g));
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    int32_t nep;
    if (fread(&nep, 4, 1, f) != 1 || nep < 0) goto fail;
    g->nep = nep;
    g->eps = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nep > 0 ? nep : 1));
    if (nep && fread(g->eps, 4, (size_t)nep, f) != (size_t)nep) goto fail;
    if (fread(&g->N, 4, 1, f) != 1 || fread(&g->K, 4, 1, f) != 1) goto fail;
    g->adj = (int32_t *)malloc(sizeof(int32_t) * (size_t)g->N * (size_t)g->K);
    if (fread(g->adj, 4, (size_t)g->N * (size_t)g->K, f) != (size_t)g->N * (size_t)g->K) goto fail;
    {
        int32_t hdr[3];
        if (fread(hdr, 4, 3, f) == 3) {
            g->has_init = 1;
            g->init_K = hdr[1];
            g->ep = hdr[2];
            int32_t n = hdr[0];
            g->levels = (int32_t *)calloc((size_t)n, sizeof(int32_t));
            g->offsets = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
            size_t capn = 1024, used = 0;
            g->lists = (int32_t *)malloc(sizeof(int32_t) * capn);
            for (int32_t i = 0; i < n; ++i) {
                int32_t cur;
                if (fread(&cur, 4, 1, f) != 1) goto fail;
                g->levels[i] = cur / g->init_K;
                g->offsets[i] = (int64_t)used;
                if (used + (size_t)cur > capn) {
                    while (used + (size_t)cur > capn) capn *= 2;
                    g->lists = (int32_t *)realloc(g->lists, sizeof(int32_t) * capn);
                }
                if (cur && fread(g->lists + used, 4, (size_t)cur, f) != (size_t)cur) goto fail;
                used += (size_t)cur;
            }
            g->offsets[n] = (int64_t)used;
        }
    }
    fclose(f);
    return 0;
fail:
    fclose(f);
    orq_graph_free(g);
    return -2;
}

/* ------------------------------------------------------------------ */
/* LinearPool (glass/neighbor.hpp:320-433) */

typedef struct {
    int32_t size, cur, ef, cap;
    double *dist;
    int32_t *id; /* bit 31 = checked */
    uint8_t *vis;
} pool_t;

static void pool_insert(pool_t *p, int32_t u, double d) {
    /* insert (neighbor.hpp:382-396): reject when full and d >= last */
    if (p->size == p->cap && d >= p->dist[p->size - 1]) return;
    /* find_bsearch (neighbor.hpp:361-371): first index with dist > d */
    int32_t lo = 0, hi = p->size;
    while (lo < hi) {
        int32_t mid = (lo + hi) / 2;
        if (p->dist[mid] > d) hi = mid;
        else lo = mid + 1;
    }
    /* shift right; the slot past cap is scratch (data_ has cap+1 entries) */
    int32_t n = p->size - lo;
    memmove(&p->dist[lo + 1], &p->dist[lo], sizeof(double) * (size_t)n);
    memmove(&p->id[lo + 1], &p->id[lo], sizeof(int32_t) * (size_t)n);
    p->dist[lo] = d;
    p->id[lo] = u;
    if (p->size < p->cap) p->size++;
    if (lo < p->cur) p->cur = lo;
}

static int32_t pool_pop(pool_t *p) {
    /* pop (neighbor.hpp:398-405) */
    p->id[p->cur] |= (int32_t)0x80000000u;
    int32_t pre = p->cur;
    while (p->cur < p->size && (p->id[p->cur] & (int32_t)0x80000000u)) p->cur++;
    return p->id[pre] & 0x7fffffff;
}

static int pool_has_next(const pool_t *p) { return p->cur < p->size && p->cur < p->ef; }

int orq_pool_run(int32_t cap, int32_t ef, int32_t nops, const int32_t *ops, const double *dists, int32_t *popped,
                 int32_t *out_ids, double *out_dists, int32_t *out_size, int32_t *out_cur) {
    /* Drives one LinearPool through a sequence of operations, for model tests:
     * ops[i] >= 0 inserts (ops[i], dists[i]); ops[i] < 0 pops when has_next()
     * holds (popped[i] = id) and records -1 otherwise.  Final entries keep the
     * checked bit (bit 31). */
    if (cap < 1 || nops < 0) return -1;
    pool_t p;
    p.size = 0;
    p.cur = 0;
    p.ef = ef;
    p.cap = cap;
    p.dist = (double *)calloc((size_t)cap + 1, sizeof(double));
    p.id = (int32_t *)calloc((size_t)cap + 1, sizeof(int32_t));
    p.vis = NULL;
    for (int32_t i = 0; i < nops; ++i) {
        if (ops[i] >= 0) {
            pool_insert(&p, ops[i], dists[i]);
            popped[i] = -2;
        } else {
            popped[i] = pool_has_next(&p) ? pool_pop(&p) : -1;
        }
    }
    for (int32_t i = 0; i < p.size; ++i) {
        out_ids[i] = p.id[i];
        out_dists[i] = p.dist[i];
    }
    *out_size = p.size;
    *out_cur = p.cur;
    free(p.dist);
    free(p.id);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Search */

static double eval(const orq_quant *q, const uint8_t *qbuf, const uint8_t *codes, int32_t u, orq_counters *cnt,
                   int32_t ctx) {
    cnt->dist_cmps++;
    return orq_distance_ctx(q, qbuf, codes + (int64_t)u * q->code_size, ctx);
}

void orq_search(const orq_graph *g, const orq_quant *q, const uint8_t *codes, const float *x, int32_t k, int32_t ef,
                int32_t batch_mode, int32_t *ids, float *dists, orq_counters *cnt) {
    orq_counters local = {0, 0, 0};
    if (!cnt) cnt = &local;
    int32_t ctx = batch_mode ? ORQ_CTX_BATCH : ORQ_CTX_SEARCH;
    uint8_t *qbuf = (uint8_t *)calloc(1, (size_t)orq_query_size(q));
    orq_prepare_query(q, x, qbuf);
    /* Search: reset(nb, max(k,ef), max(k,ef)) (graph_searcher.hpp:134);
     * SearchBatch: reset(nb, ef, max(k,ef)) (graph_searcher.hpp:152) */
    int32_t cap = k > ef ? k : ef;
    pool_t p;
    p.size = p.cur = 0;
    p.cap = cap;
    p.ef = batch_mode ? ef : cap;
    p.dist = (double *)calloc((size_t)cap + 1, sizeof(double));
    p.id = (int32_t *)calloc((size_t)cap + 1, sizeof(int32_t));
    p.vis = (uint8_t *)calloc((size_t)g->N, 1);

    /* initialize_search (graph.hpp:95-104) */
    if (g->has_init) {
        /* HNSWInitializer::initialize (HNSWInitializer.hpp:41-62): the list
         * scanned in one pass is the one of the node current at pass start. */
        int32_t u = g->ep;
        double du = eval(q, qbuf, codes, u, cnt, ctx | ORQ_CTX_INIT);
        for (int32_t level = g->levels[u]; level > 0; --level) {
            int changed = 1;
            while (changed) {
                changed = 0;
                const int32_t *list = g->lists + g->offsets[u] + (int64_t)(level - 1) * g->init_K;
                for (int32_t i = 0; i < g->init_K && list[i] != -1; ++i) {
                    int32_t v = list[i];
                    double dv = eval(q, qbuf, codes, v, cnt, ctx | ORQ_CTX_INIT);
                    cnt->upper_evals++;
                    if (dv < du) {
                        du = dv;
                        u = v;
                        changed = 1;
                    }
                }
            }
        }
        pool_insert(&p, u, du);
        p.vis[u] = 1;
    } else {
        for (int32_t i = 0; i < g->nep; ++i) {
            pool_insert(&p, g->eps[i], eval(q, qbuf, codes, g->eps[i], cnt, ctx | ORQ_CTX_INIT));
            p.vis[g->eps[i]] = 1;
        }
    }

    /* SearchImpl2 (graph_searcher.hpp:196-224); SearchImpl1 inserts the same
     * sequence (161-194) */
    while (pool_has_next(&p)) {
        int32_t u = pool_pop(&p);
        cnt->expansions++;
        const int32_t *row = g->adj + (int64_t)u * g->K;
        for (int32_t i = 0; i < g->K; ++i) {
            int32_t v = row[i];
            if (v == -1) break;
            if (p.vis[v]) continue;
            p.vis[v] = 1;
            pool_insert(&p, v, eval(q, qbuf, codes, v, cnt, ctx));
        }
    }

    /* to_sorted (neighbor.hpp:421-428); entries past size are -1/+inf here */
    for (int32_t i = 0; i < k; ++i) {
        if (i < p.size) {
            ids[i] = p.id[i] & 0x7fffffff;
            if (dists) dists[i] = (float)p.dist[i];
        } else {
            ids[i] = -1;
            if (dists) dists[i] = INFINITY;
        }
    }
    free(p.dist);
    free(p.id);
    free(p.vis);
    free(qbuf);
}

float orq_refine_factor(int32_t kind) {
    if (kind == ORQ_SQ8U || kind == ORQ_SQ8 || kind == ORQ_SQ4U || kind == ORQ_PQ8) return 1.5f;
    if (kind == ORQ_SQ2U) return 3.0f;
    return 1.0f;
}

void orq_search_refine(const orq_graph *g, const orq_quant *q, const uint8_t *codes, const orq_quant *rq,
                       const uint8_t *rcodes, float factor, const float *x, int32_t k, int32_t ef, int32_t batch_mode,
                       int32_t *ids, float *dists, orq_counters *cnt) {
    /* reorder_k = (int32_t)(k * reorder_mul) (refiner.hpp:46,73) */
    int32_t rk = (int32_t)((float)k * factor);
    if (!rq || rk == k) {
        orq_search(g, q, codes, x, k, ef, batch_mode, ids, dists, cnt);
        return;
    }
    int32_t *cand = (int32_t *)malloc(sizeof(int32_t) * (size_t)rk);
    orq_search(g, q, codes, x, rk, ef, batch_mode, cand, NULL, cnt);
    uint8_t *qbuf = (uint8_t *)calloc(1, (size_t)orq_query_size(rq));
    orq_prepare_query(rq, x, qbuf);
    /* RefineImpl (refiner.hpp:92-110): MaxHeap(k) by (dist, id)
     * (neighbor.hpp:197-258) == the k smallest (dist, id) pairs, ascending. */
    float *cd = (float *)malloc(sizeof(float) * (size_t)rk);
    int32_t m = 0;
    for (int32_t j = 0; j < rk; ++j) {
        if (cand[j] < 0) continue;
        cand[m] = cand[j];
        cd[m] = (float)orq_distance(rq, qbuf, rcodes + (int64_t)cand[j] * rq->code_size);
        m++;
    }
    for (int32_t i = 1; i < m; ++i) { /* insertion sort by (dist, id) */
        float dv = cd[i];
        int32_t iv = cand[i];
        int32_t j = i - 1;
        while (j >= 0 && (cd[j] > dv || (cd[j] == dv && cand[j] > iv))) {
            cd[j + 1] = cd[j];
            cand[j + 1] = cand[j];
            j--;
        }
        cd[j + 1] = dv;
        cand[j + 1] = iv;
    }
    for (int32_t i = 0; i < k; ++i) {
        ids[i] = i < m ? cand[i] : -1;
        if (dists) dists[i] = i < m ? cd[i] : INFINITY;
    }
    free(cand);
    free(cd);
    free(qbuf);
}

void orq_search_many(const orq_graph *g, const orq_quant *q, const uint8_t *codes, const orq_quant *rq,
                     const uint8_t *rcodes, float factor, const float *xs, int64_t nq, int32_t k, int32_t ef,
                     int32_t batch_mode, int32_t *ids, float *dists, int64_t *dist_cmps, int64_t *expansions,
                     int32_t nthreads) {
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
#endif
    for (int64_t i = 0; i < nq; ++i) {
        orq_counters c = {0, 0, 0};
        orq_search_refine(g, q, codes, rq, rcodes, factor, xs + i * q->dim, k, ef, batch_mode, ids + i * k,
                          dists ? dists + i * k : NULL, &c);
        if (dist_cmps) dist_cmps[i] = c.dist_cmps;
        if (expansions) expansions[i] = c.expansions;
    }
}
