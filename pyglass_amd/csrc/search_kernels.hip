// HIP kernels for pyglass query-time graph search on MI355X (gfx950).
//
// One wavefront (64 lanes) owns one query at a time and reproduces the
// reference's best-first expansion exactly:
//   * the candidate pool is glass::LinearPool (glass/neighbor.hpp:320-433)
//     kept sorted in LDS as (key, id|checked) pairs;
//   * the R neighbours of a popped node are loaded by lanes 0..R-1 from the
//     adjacency row (one coalesced 4R-byte read), tested-and-set in an exact
//     per-slot visited bitset (global, cleared per query) with atomic OR, and
//     compacted in adjacency order (glass/searcher/graph_searcher.hpp:196-224);
//   * their distances are computed by "teams" of lanes from coalesced code
//     gathers (integer quantizers: v_dot4 on byte-expanded differences, exact;
//     float quantizers: 16-lane accumulators in the reference's order);
//   * the whole batch of new candidates is merged into the pool at once.  The
//     merge reproduces sequential LinearPool::insert exactly: the pool after
//     inserting c_1..c_m in order holds the `cap` smallest (dist, insertion
//     order) keys, and `cur_` is the first unchecked index.
// A persistent grid pulls queries from a device work counter, so one launch
// serves any number of queries with per-query ef / k.
#include "glass_hip_internal.h"

namespace glass_hip {

namespace {

constexpr uint32_t kChecked = 0x80000000u;
constexpr uint32_t kIdMask = 0x7fffffffu;
constexpr uint32_t kKeyMax = 0xffffffffu;

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Order-preserving 32-bit keys: integer distances compare as signed ints,
// float distances as IEEE floats (-0 folded onto +0, as operator< does).
__device__ __forceinline__ uint32_t key_of_int(int32_t d) { return (uint32_t)d ^ 0x80000000u; }
__device__ __forceinline__ int32_t int_of_key(uint32_t k) { return (int32_t)(k ^ 0x80000000u); }
__device__ __forceinline__ uint32_t key_of_float(float f) {
    uint32_t b = __float_as_uint(f);
    if ((b & 0x7fffffffu) == 0) b = 0;
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float float_of_key(uint32_t k) {
    uint32_t b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(b);
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        uint32_t o = (uint32_t)__shfl_xor((int)v, off);
        v = o < v ? o : v;
    }
    return v;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        uint64_t o = __shfl_xor(v, off);
        v = o < v ? o : v;
    }
    return v;
}

__device__ __forceinline__ int32_t sdot4(uint32_t a, uint32_t b, int32_t c) {
    return __builtin_amdgcn_sdot4((int)a, (int)b, c, false);
}

// Variants of the distance kernel (one per quantizer x metric).
enum Variant : int {
    V_SQ8U_L2 = 0,
    V_SQ8U_IP = 1,
    V_SQ4U = 2,  // helpa::l2a_u4_u4 for both metrics (sq4u_quant.hpp:131)
    V_SQ2U = 3,  // helpa::l2a_u2_u2 for both metrics (sq2u_quant.hpp:202, graph_searcher.hpp:357)
    V_E5M2_L2 = 10,
    V_E5M2_IP = 11,
    V_SQ8_L2 = 12,
    V_SQ8_IP = 13,
    V_FP32_L2 = 14,
    V_FP32_IP = 15,
};

constexpr bool is_int_variant(int v) { return v < 10; }

// Per-wave LDS layout.  All offsets are multiples of 16 bytes.
struct WaveLds {
    uint32_t *pk;   // pool keys [cap_max]
    uint32_t *pid;  // pool ids (bit 31 = checked) [cap_max]
    uint32_t *ck;   // candidate keys [64]
    uint32_t *ci;   // candidate ids [64]
    float *qf;      // query, zero padded [qf_len]
    float *qm;      // SQ8 helpers: 256x-256min [qf_len], 256min [qf_len], dif [qf_len]
    uint8_t *qcode; // encoded query [code_size rounded to 16]
};

__host__ __device__ inline int32_t round16(int64_t x) { return (int32_t)((x + 15) / 16 * 16); }

__host__ __device__ inline int32_t qf_len(const DeviceIndex &ix) {
    int32_t a = ix.dim_align > ix.r_dim_align ? ix.dim_align : ix.r_dim_align;
    if (a < ix.dim) a = ix.dim;
    return round16(a);
}

__host__ __device__ inline size_t wave_lds_bytes(const DeviceIndex &ix, int32_t cap_max) {
    size_t b = 0;
    b += (size_t)round16((int64_t)cap_max * 4) * 2;
    b += 64 * 4 * 2;
    b += (size_t)qf_len(ix) * 4;
    if (ix.quant == QK_SQ8) b += (size_t)qf_len(ix) * 4 * 3;
    b += (size_t)round16(ix.code_size);
    return b;
}

__device__ inline WaveLds carve(uint8_t *base, const DeviceIndex &ix, int32_t cap_max) {
    WaveLds w;
    size_t off = 0;
    int32_t capb = round16((int64_t)cap_max * 4);
    w.pk = (uint32_t *)(base + off);
    off += capb;
    w.pid = (uint32_t *)(base + off);
    off += capb;
    w.ck = (uint32_t *)(base + off);
    off += 256;
    w.ci = (uint32_t *)(base + off);
    off += 256;
    int32_t ql = qf_len(ix);
    w.qf = (float *)(base + off);
    off += (size_t)ql * 4;
    w.qm = (float *)(base + off);
    if (ix.quant == QK_SQ8) off += (size_t)ql * 4 * 3;
    w.qcode = base + off;
    return w;
}

// ---------------------------------------------------------------------------
// Query encoding (ComputerImpl ctor, glass/quant/computer.hpp:54-63) with the
// base calibrator: AffineCalibrator / SymCalibrator::transform
// (glass/quant/calibrator.hpp:26-55), limit_range (glass/quant/utils.hpp:8-26).

__device__ __forceinline__ uint32_t affine_code(float x, float mn, float dif, float mul) {
    x = (x - mn) / dif;
    x = x < 0.0f ? 0.0f : x;
    x = x > 1.0f ? 1.0f : x;
    return (uint32_t)(int32_t)roundf(x * mul);
}

__device__ __forceinline__ uint32_t sym_code(float x, float mx) {
    x = x / mx;
    x = x < -1.0f ? -1.0f : x;
    x = x > 1.0f ? 1.0f : x;
    return (uint32_t)(int32_t)roundf(x * 127.0f);
}

template <int V>
__device__ void encode_query(const DeviceIndex &ix, const WaveLds &w) {
    const int l = lane_id();
    for (int b = l; b < ix.code_size; b += 64) {
        uint32_t byte = 0;
        if (V == V_SQ8U_L2) {  // sq8u_quant.hpp:38-42, AffineCalibrator<127>
            if (b < ix.dim) byte = affine_code(w.qf[b], ix.qmin, ix.qdif, 127.0f) & 0xffu;
        } else if (V == V_SQ8U_IP) {  // SymCalibrator<127>
            if (b < ix.dim) byte = sym_code(w.qf[b], ix.qmax) & 0xffu;
        } else if (V == V_SQ4U) {  // sq4u_quant.hpp:107-117: dim j -> nibble j&1 of byte j/2
            int j = 2 * b;
            if (j < ix.dim) byte |= affine_code(w.qf[j], ix.qmin, ix.qdif, 15.0f) & 0xfu;
            if (j + 1 < ix.dim) byte |= (affine_code(w.qf[j + 1], ix.qmin, ix.qdif, 15.0f) & 0xfu) << 4;
        } else if (V == V_SQ2U) {  // sq2u_quant.hpp:186-200: dim j -> bits 2(j%4) of byte j/4
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                int j = 4 * b + s;
                if (j < ix.dim) byte |= (affine_code(w.qf[j], ix.qmin, ix.qdif, 3.0f) & 0x3u) << (2 * s);
            }
        }
        w.qcode[b] = (uint8_t)byte;
    }
}

// ---------------------------------------------------------------------------
// Integer distances.  Exact for every decomposition, so a candidate row is
// split over `slot` lanes, DPL dwords each, and the partial sums are added.
//   SQ8U L2  helpa::l2_s7_s7   sum (x-y)^2 over int8 codes in [0,127]
//   SQ8U IP  helpa::dot_s8_s8  -sum x*y over int8 codes in [-127,127]
//   SQ4U     helpa::l2a_u4_u4  sum (x-y)^2 over nibbles
//   SQ2U     helpa::l2a_u2_u2  sum (x-y)^2 over 2-bit fields
// Unsigned fields are turned into signed byte differences without borrows:
// ((q | 0x80) - c) ^ 0x80 per byte equals q - c as int8 for q, c <= 127.

template <int V>
__device__ __forceinline__ int32_t int_partial_dword(uint32_t q, uint32_t c, int32_t acc) {
    if (V == V_SQ8U_L2) {
        uint32_t t = ((q | 0x80808080u) - c) ^ 0x80808080u;
        return sdot4(t, t, acc);
    } else if (V == V_SQ8U_IP) {
        return sdot4(q, c, acc);
    } else if (V == V_SQ4U) {
        uint32_t ql = (q & 0x0f0f0f0fu) | 0x80808080u, qh = ((q >> 4) & 0x0f0f0f0fu) | 0x80808080u;
        uint32_t t1 = (ql - (c & 0x0f0f0f0fu)) ^ 0x80808080u;
        uint32_t t2 = (qh - ((c >> 4) & 0x0f0f0f0fu)) ^ 0x80808080u;
        acc = sdot4(t1, t1, acc);
        return sdot4(t2, t2, acc);
    } else {  // V_SQ2U
#pragma unroll
        for (int s = 0; s < 8; s += 2) {
            uint32_t qs = ((q >> s) & 0x03030303u) | 0x80808080u;
            uint32_t t = (qs - ((c >> s) & 0x03030303u)) ^ 0x80808080u;
            acc = sdot4(t, t, acc);
        }
        return acc;
    }
}

template <int V, int DPL>
struct IntQuery {
    uint32_t q[DPL];
    int32_t slot;    // lanes per candidate (power of two)
    int32_t nparts;  // active parts per candidate
    int32_t part;    // this lane's part

    __device__ void init(const DeviceIndex &ix, const WaveLds &w) {
        const int code_dw = ix.code_size / 4;
        nparts = (code_dw + DPL - 1) / DPL;
        slot = 1;
        while (slot < nparts) slot <<= 1;
        part = lane_id() & (slot - 1);
        const uint32_t *qd = (const uint32_t *)w.qcode;
#pragma unroll
        for (int i = 0; i < DPL; ++i) {
            int d = part * DPL + i;
            q[i] = (part < nparts && d < code_dw) ? qd[d] : 0u;
        }
    }

    // Distances of candidates ci[0..m) -> keys ck[0..m).
    __device__ void distances(const DeviceIndex &ix, const WaveLds &w, int m) const {
        const int l = lane_id();
        const int cpr = 64 / slot;
        const int cj = l / slot;
        const int code_q = ix.code_size / 16;  // uint4 per row
        for (int base = 0; base < m; base += cpr) {
            const int j = base + cj;
            int32_t acc = 0;
            if (j < m && part < nparts) {
                const uint32_t id = w.ci[j];
                const uint4 *src = (const uint4 *)(ix.codes + (int64_t)id * ix.code_size);
                uint4 c[DPL / 4];
#pragma unroll
                for (int i = 0; i < DPL / 4; ++i) {
                    int qi = part * (DPL / 4) + i;
                    c[i] = qi < code_q ? src[qi] : make_uint4(0, 0, 0, 0);
                }
#pragma unroll
                for (int i = 0; i < DPL / 4; ++i) {
                    acc = int_partial_dword<V>(q[4 * i + 0], c[i].x, acc);
                    acc = int_partial_dword<V>(q[4 * i + 1], c[i].y, acc);
                    acc = int_partial_dword<V>(q[4 * i + 2], c[i].z, acc);
                    acc = int_partial_dword<V>(q[4 * i + 3], c[i].w, acc);
                }
            }
            for (int off = slot >> 1; off >= 1; off >>= 1) acc += __shfl_xor(acc, off);
            if (j < m && part == 0) {
                int32_t d = (V == V_SQ8U_IP) ? -acc : acc;
                w.ck[j] = key_of_int(d);
            }
        }
    }
};

// ---------------------------------------------------------------------------
// Float distances: 16 lanes per candidate, lane l accumulates elements
// j = l, l+16, ... in increasing j (the AVX-512 lane order), then the 16
// partial sums are reduced in the reference's tree.
//   E5M2  glass::L2SqrE5M2 / IPE5M2   (glass/simd/avx512.hpp:15-38)
//   SQ8   glass::L2SqrSQ8_ext / IPSQ8_ext (glass/simd/avx512.hpp:40-77)
//   FP32  helpa::l2_fp32_fp32 / dot_fp32_fp32 (x86/l2_impl.hpp:50-71, x86/dot_impl.hpp:52-71)
// The fused multiply-adds are where the compiled reference fuses them.

// glass::reduce_add_f32x16 (glass/simd/avx512_utils.hpp:12-22): (c0+c1)+(c2+c3)
__device__ __forceinline__ float reduce16_glass(float a) {
    a = a + __shfl_xor(a, 8);
    a = a + __shfl_xor(a, 4);
    a = a + __shfl_xor(a, 1);
    a = a + __shfl_xor(a, 2);
    return a;
}

// GCC _mm512_reduce_add_ps as used by helpa (x86/utils.hpp:61-64): (c0+c2)+(c1+c3)
__device__ __forceinline__ float reduce16_helpa(float a) {
    a = a + __shfl_xor(a, 8);
    a = a + __shfl_xor(a, 4);
    a = a + __shfl_xor(a, 2);
    a = a + __shfl_xor(a, 1);
    return a;
}

__device__ __forceinline__ float e5m2_to_float(uint32_t byte) {
    return __half2float(__ushort_as_half((unsigned short)(byte << 8)));
}

// One float candidate row, lanes of a 16-lane segment; returns the reduced sum.
template <int V>
__device__ __forceinline__ float float_row(const DeviceIndex &ix, const WaveLds &w, const uint8_t *row, int l16,
                                           int32_t mode, int32_t da) {
    float acc = 0.0f;
    if (V == V_E5M2_L2 || V == V_E5M2_IP) {
        for (int j = l16; j < da; j += 16) {
            float y = e5m2_to_float(row[j]);
            float x = w.qf[j];
            if (V == V_E5M2_L2) {
                float t = x - y;
                acc = fmaf(t, t, acc);
            } else {
                acc = fmaf(x, y, acc);
            }
        }
        return reduce16_glass(acc);
    } else if (V == V_SQ8_L2 || V == V_SQ8_IP) {
        const int ql = qf_len(ix);
        const float *m256 = w.qm + ql, *dif = w.qm + 2 * ql;
        for (int j = l16; j < da; j += 16) {
            float c = (float)row[j] + 0.5f;
            if (V == V_SQ8_L2) {
                float t;
                if (mode == MODE_BATCH) {
                    // SearchImpl2 context: yy = fma(c, dif, 256min), t = 256x - yy
                    t = w.qf[j] * 256.0f - fmaf(c, dif[j], m256[j]);
                } else {
                    // Search / bare computer context: t = fma(-c, dif, 256x - 256min)
                    t = fmaf(-c, dif[j], w.qm[j]);
                }
                acc = fmaf(t, t, acc);
            } else {
                float yy = fmaf(c, dif[j], m256[j]);
                acc = fmaf(w.qf[j], yy, acc);
            }
        }
        return reduce16_glass(acc);
    } else {  // FP32
        const float *y = (const float *)row;
        for (int j = l16; j < da; j += 16) {
            float x = w.qf[j];
            if (V == V_FP32_L2) {
                float t = x - y[j];
                acc = fmaf(t, t, acc);
            } else {
                acc = fmaf(x, y[j], acc);
            }
        }
        return reduce16_helpa(acc);
    }
}

template <int V>
__device__ void float_distances(const DeviceIndex &ix, const WaveLds &w, int m, int32_t mode) {
    if (m <= 0) return;
    const int l = lane_id();
    const int l16 = l & 15, cj = l >> 4;
    for (int base = 0; base < m; base += 4) {
        const int j = base + cj;
        const int jj = j < m ? j : (m - 1);
        const uint32_t id = w.ci[jj];
        float r = float_row<V>(ix, w, ix.codes + (int64_t)id * ix.code_size, l16, mode, ix.dim_align);
        bool neg = (V == V_E5M2_IP || V == V_SQ8_IP || V == V_FP32_IP);
        if (j < m && l16 == 0) w.ck[j] = key_of_float(neg ? -r : r);
    }
}

// ---------------------------------------------------------------------------
// Pool operations.

struct PoolState {
    int32_t size, cur, cap, ef;
};

// LinearPool::pop (neighbor.hpp:398-405): mark data[cur] checked, advance
// cur to the next unchecked entry.
__device__ __forceinline__ uint32_t pool_pop(const WaveLds &w, PoolState &ps) {
    const int l = lane_id();
    const uint32_t uid = w.pid[ps.cur];
    wave_sync();
    if (l == 0) w.pid[ps.cur] = uid | kChecked;
    int c = ps.cur + 1;
    while (c < ps.size) {
        int p = c + l;
        uint32_t x = p < ps.size ? w.pid[p] : kChecked;
        uint64_t um = ballot(!(x & kChecked));
        if (um) {
            c += __builtin_ctzll(um);
            break;
        }
        c += 64;
    }
    ps.cur = c < ps.size ? c : ps.size;
    wave_sync();
    return uid & kIdMask;
}

__device__ __forceinline__ int upper_bound_lds(const uint32_t *pk, int n, uint32_t key) {
    int lo = 0, len = n;
    while (len > 0) {
        int half = len >> 1;
        int mid = lo + half;
        if (pk[mid] <= key) {
            lo = mid + 1;
            len = len - half - 1;
        } else {
            len = half;
        }
    }
    return lo;
}

// Merge candidates ck/ci[0..m) (in insertion order) into the pool.  Equivalent
// to m sequential LinearPool::insert calls (neighbor.hpp:382-396).
__device__ void pool_merge(const WaveLds &w, PoolState &ps, int m) {
    const int l = lane_id();
    uint32_t kj = l < m ? w.ck[l] : kKeyMax;
    uint32_t ij = l < m ? w.ci[l] : 0u;
    bool act = l < m;
    // insert() rejects d >= data[size-1] when full; such a candidate would
    // also land past cap in the merge and displaces nothing.
    if (ps.size == ps.cap) act = act && (kj < w.pk[ps.size - 1]);
    const uint64_t am = ballot(act);
    if (am == 0) return;
    // rank among the active candidates by (key, insertion order)
    int r = 0;
    uint64_t mm = am;
    while (mm) {
        const int i = __builtin_ctzll(mm);
        mm &= mm - 1;
        const uint32_t ki = (uint32_t)__builtin_amdgcn_readlane((int)kj, i);
        r += (ki < kj) | ((ki == kj) & (i < l));
    }
    // upper bound in the current pool: new keys go after equal old keys
    int ub = act ? upper_bound_lds(w.pk, ps.size, kj) : 0x7fffffff;
    const int np = r + ub;
    const bool ins = act && np < ps.cap;
    const int lo = (int)wave_min_u32((uint32_t)ub);
    const int ma = __builtin_popcountll(am);
    const int new_size = ps.size + ma < ps.cap ? ps.size + ma : ps.cap;
    int new_cur = 0x7fffffff;
    if (ps.cur < ps.size) {
        const uint32_t kc = w.pk[ps.cur];
        new_cur = ps.cur + __builtin_popcountll(ballot(act && kj < kc));
    }
    const int min_np = (int)wave_min_u32(ins ? (uint32_t)np : 0x7fffffffu);
    new_cur = new_cur < min_np ? new_cur : min_np;
    new_cur = new_cur < new_size ? new_cur : new_size;
    wave_sync();
    // shift pool entries [lo, size) right by the number of smaller candidates;
    // top chunk first so no entry is overwritten before it is read
    if (lo < ps.size) {
        const int nch = (ps.size - lo + 63) >> 6;
        for (int c = nch - 1; c >= 0; --c) {
            const int p = lo + 64 * c + l;
            const bool valid = p < ps.size;
            const uint32_t key = valid ? w.pk[p] : kKeyMax;
            const uint32_t id = valid ? w.pid[p] : 0u;
            int s = 0;
            uint64_t m2 = am;
            while (m2) {
                const int i = __builtin_ctzll(m2);
                m2 &= m2 - 1;
                const uint32_t ki = (uint32_t)__builtin_amdgcn_readlane((int)kj, i);
                s += ki < key;
            }
            wave_sync();
            const int dst = p + s;
            if (valid && dst < ps.cap) {
                w.pk[dst] = key;
                w.pid[dst] = id;
            }
            wave_sync();
        }
    }
    if (ins) {
        w.pk[np] = kj;
        w.pid[np] = ij;
    }
    wave_sync();
    ps.size = new_size;
    ps.cur = new_cur;
}

// ---------------------------------------------------------------------------
// The search kernel.

template <int V, int DPL>
__global__ __launch_bounds__(256) void search_kernel(DeviceIndex ix, QueryBatch qb, LaunchConfig lc, SlotState ss) {
    extern __shared__ __attribute__((aligned(16))) uint8_t lds_raw[];
    const int l = lane_id();
    const int wv = threadIdx.x >> 6;
    const size_t wbytes = wave_lds_bytes(ix, lc.cap_max);
    const WaveLds w = carve(lds_raw + (size_t)wv * wbytes, ix, lc.cap_max);
    const int slot = blockIdx.x * lc.waves_per_block + wv;
    uint32_t *vis = ss.visited + (int64_t)slot * ss.vis_stride;
    const int ql = qf_len(ix);

    for (;;) {
        int qi = 0;
        if (l == 0) qi = atomicAdd(ss.work_counter, 1);
        qi = __shfl(qi, 0);
        if (qi >= qb.nq) break;

        const int32_t ef = qb.ef ? qb.ef[qi] : qb.ef_uniform;
        const int32_t k = qb.k ? qb.k[qi] : qb.k_uniform;
        // Refiner::Search/SearchBatch: reorder_k = int(k * factor) (refiner.hpp:46,73)
        int32_t kk = k;
        bool refine = false;
        if (ix.rerank) {
            int32_t rk = (int32_t)((float)k * ix.refine_factor);
            if (rk != k) {
                kk = rk;
                refine = true;
            }
        }
        PoolState ps;
        ps.cap = kk > ef ? kk : ef;
        ps.ef = qb.mode == MODE_SEARCH ? ps.cap : ef;
        ps.size = 0;
        ps.cur = 0;
        int32_t cnt_d = 0, cnt_e = 0, cnt_u = 0;

        // stage the fp32 query (zero padded, as MemCpyTag copies into a zeroed buffer)
        const float *x = qb.queries + (int64_t)qi * qb.dim;
        for (int j = l; j < ql; j += 64) w.qf[j] = j < ix.dim ? x[j] : 0.0f;
        wave_sync();
        if (V == V_SQ8_L2 || V == V_SQ8_IP) {
            for (int j = l; j < ql; j += 64) {
                float mn = j < ix.dim_align ? ix.sq8_mins[j] : 0.0f;
                float df = j < ix.dim_align ? ix.sq8_difs[j] : 0.0f;
                w.qm[j] = w.qf[j] * 256.0f - mn * 256.0f;
                w.qm[ql + j] = mn * 256.0f;
                w.qm[2 * ql + j] = df;
            }
        }
        IntQuery<V, DPL> iq;
        if (is_int_variant(V)) {
            encode_query<V>(ix, w);
            wave_sync();
            iq.init(ix, w);
        }
        // clear this slot's visited bitset (Bitset::reset, neighbor.hpp:67-74)
        {
            uint4 *v4 = (uint4 *)vis;
            const int64_t n4 = ss.vis_stride / 4;
            for (int64_t i = l; i < n4; i += 64) v4[i] = make_uint4(0, 0, 0, 0);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        }
        wave_sync();

        auto dists = [&](int m) {
            if (is_int_variant(V)) {
                iq.distances(ix, w, m);
            } else {
                float_distances<V>(ix, w, m, qb.mode);
            }
            wave_sync();
        };

        // initialize_search (graph.hpp:95-104)
        if (ix.has_init) {
            // HNSWInitializer::initialize (HNSWInitializer.hpp:41-62)
            uint32_t u = (uint32_t)ix.ep;
            if (l == 0) w.ci[0] = u;
            wave_sync();
            dists(1);
            uint32_t du = w.ck[0];
            cnt_d += 1;
            for (int level = ix.levels[ix.ep]; level > 0; --level) {
                for (;;) {
                    const int32_t *list = ix.up_lists + ix.up_off[u] + (int64_t)(level - 1) * ix.up_K;
                    const int32_t v = l < ix.up_K ? list[l] : -1;
                    const uint64_t inv = ballot(l < ix.up_K && v == -1);
                    const int n = inv ? __builtin_ctzll(inv) : ix.up_K;
                    if (n == 0) break;
                    if (l < n) w.ci[l] = (uint32_t)v;
                    wave_sync();
                    dists(n);
                    cnt_d += n;
                    cnt_u += n;
                    const uint32_t mk = l < n ? w.ck[l] : kKeyMax;
                    const uint64_t best = wave_min_u64(((uint64_t)mk << 32) | (uint32_t)l);
                    const uint32_t bk = (uint32_t)(best >> 32);
                    const int bi = (int)(best & 0xffffffffu);
                    wave_sync();
                    if (bk < du) {  // strict <: the first minimum of the pass wins
                        du = bk;
                        u = (uint32_t)__builtin_amdgcn_readlane(v, bi);
                    } else {
                        break;
                    }
                }
            }
            if (l == 0) {
                w.pk[0] = du;
                w.pid[0] = u;
                __hip_atomic_fetch_or(&vis[u >> 5], 1u << (u & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            ps.size = 1;
            ps.cur = 0;
            wave_sync();
        } else {
            // insert every entry point in order, then mark it visited
            const int nep = ix.nep < 64 ? ix.nep : 64;
            if (l < nep) w.ci[l] = (uint32_t)ix.eps[l];
            wave_sync();
            dists(nep);
            cnt_d += nep;
            pool_merge(w, ps, nep);
            if (l < nep) {
                uint32_t e = (uint32_t)ix.eps[l];
                __hip_atomic_fetch_or(&vis[e >> 5], 1u << (e & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }

        // best-first expansion (SearchImpl2, graph_searcher.hpp:196-224)
        while (ps.cur < ps.size && ps.cur < ps.ef) {
            const uint32_t u = pool_pop(w, ps);
            cnt_e += 1;
            const int32_t v = l < ix.K ? ix.adj[(int64_t)u * ix.K + l] : -1;
            const uint64_t inv = ballot(l < ix.K && v == -1);
            const int n = inv ? __builtin_ctzll(inv) : ix.K;
            bool isnew = false;
            if (l < n) {
                const uint32_t bit = 1u << (v & 31);
                const uint32_t old = __hip_atomic_fetch_or(&vis[(uint32_t)v >> 5], bit, __ATOMIC_RELAXED,
                                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                isnew = !(old & bit);
            }
            const uint64_t nm = ballot(isnew);
            if (nm == 0) continue;
            const int m = __builtin_popcountll(nm);
            if (isnew) w.ci[mbcnt(nm)] = (uint32_t)v;
            wave_sync();
            dists(m);
            cnt_d += m;
            pool_merge(w, ps, m);
        }

        // outputs
        const int64_t ob = qb.out_off ? qb.out_off[qi] : (int64_t)qi * qb.k_uniform;
        if (!refine) {
            // LinearPool::to_sorted (neighbor.hpp:421-428)
            for (int i = l; i < k; i += 64) {
                int32_t id = -1;
                float d = __builtin_inff();
                if (i < ps.size) {
                    id = (int32_t)(w.pid[i] & kIdMask);
                    d = is_int_variant(V) ? (float)int_of_key(w.pk[i]) : float_of_key(w.pk[i]);
                }
                qb.out_ids[ob + i] = id;
                if (qb.out_dists) qb.out_dists[ob + i] = d;
            }
        } else {
            // Refiner::RefineImpl (refiner.hpp:92-110): exact FP32 distances of the
            // first reorder_k pool entries, the k smallest by (dist, id) ascending.
            const int m = kk < ps.size ? kk : ps.size;
            const int l16 = l & 15, cj = l >> 4;
            for (int base = 0; base < m; base += 4) {
                const int j = base + cj;
                const int jj = j < m ? j : m - 1;
                const uint32_t id = w.pid[jj] & kIdMask;
                float acc = 0.0f;
                const float *y = ix.rerank + (int64_t)id * ix.r_dim_align;
                if (ix.metric == MK_L2) {
                    for (int t = l16; t < ix.r_dim_align; t += 16) {
                        float dd = w.qf[t] - y[t];
                        acc = fmaf(dd, dd, acc);
                    }
                } else {
                    for (int t = l16; t < ix.r_dim_align; t += 16) acc = fmaf(w.qf[t], y[t], acc);
                }
                float rsum = reduce16_helpa(acc);
                wave_sync();
                if (j < m && l16 == 0) w.pk[j] = key_of_float(ix.metric == MK_L2 ? rsum : -rsum);
                wave_sync();
            }
            wave_sync();
            for (int j = l; j < m; j += 64) {
                const uint32_t kj = w.pk[j], ij = w.pid[j] & kIdMask;
                int rank = 0;
                for (int i = 0; i < m; ++i) {
                    const uint32_t ki = w.pk[i], ii = w.pid[i] & kIdMask;
                    rank += (ki < kj) | ((ki == kj) & (ii < ij));
                }
                if (rank < k) {
                    qb.out_ids[ob + rank] = (int32_t)ij;
                    if (qb.out_dists) qb.out_dists[ob + rank] = float_of_key(kj);
                }
            }
            for (int i = m + l; i < k; i += 64) {
                qb.out_ids[ob + i] = -1;
                if (qb.out_dists) qb.out_dists[ob + i] = __builtin_inff();
            }
        }
        if (qb.out_counters && l == 0) {
            qb.out_counters[(int64_t)qi * 3 + 0] = cnt_d;
            qb.out_counters[(int64_t)qi * 3 + 1] = cnt_e;
            qb.out_counters[(int64_t)qi * 3 + 2] = cnt_u;
        }
        wave_sync();
    }
}

template <int V, int DPL>
hipError_t launch_one(const DeviceIndex &ix, const QueryBatch &qb, const LaunchConfig &lc, const SlotState &ss,
                      hipStream_t stream) {
    const size_t lds = wave_lds_bytes(ix, lc.cap_max) * (size_t)lc.waves_per_block;
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void *)search_kernel<V, DPL>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            160 * 1024);
        attr_set = true;
    }
    hipLaunchKernelGGL((search_kernel<V, DPL>), dim3(lc.blocks), dim3(64 * lc.waves_per_block), lds, stream, ix, qb,
                       lc, ss);
    return hipGetLastError();
}

template <int V, int DPL>
int occupancy_one(const DeviceIndex &ix, const LaunchConfig &lc) {
    const size_t lds = wave_lds_bytes(ix, lc.cap_max) * (size_t)lc.waves_per_block;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void *)search_kernel<V, DPL>, 64 * lc.waves_per_block,
                                                     lds) != hipSuccess)
        return 0;
    return nb;
}

template <int V>
int occupancy_int(const DeviceIndex &ix, const LaunchConfig &lc) {
    switch (lc.team) {
        case 4: return occupancy_one<V, 4>(ix, lc);
        case 8: return occupancy_one<V, 8>(ix, lc);
        case 16: return occupancy_one<V, 16>(ix, lc);
        default: return 0;
    }
}

template <int V>
hipError_t launch_int(const DeviceIndex &ix, const QueryBatch &qb, const LaunchConfig &lc, const SlotState &ss,
                      hipStream_t stream) {
    switch (lc.team) {
        case 4: return launch_one<V, 4>(ix, qb, lc, ss, stream);
        case 8: return launch_one<V, 8>(ix, qb, lc, ss, stream);
        case 16: return launch_one<V, 16>(ix, qb, lc, ss, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

size_t lds_bytes_per_wave(const DeviceIndex &ix, int32_t cap_max) { return wave_lds_bytes(ix, cap_max); }

int max_blocks_per_cu(const DeviceIndex &ix, const LaunchConfig &lc) {
    const bool l2 = ix.metric == MK_L2;
    switch (ix.quant) {
        case QK_SQ8U: return l2 ? occupancy_int<V_SQ8U_L2>(ix, lc) : occupancy_int<V_SQ8U_IP>(ix, lc);
        case QK_SQ4U: return occupancy_int<V_SQ4U>(ix, lc);
        case QK_SQ2U: return occupancy_int<V_SQ2U>(ix, lc);
        case QK_E5M2: return l2 ? occupancy_one<V_E5M2_L2, 4>(ix, lc) : occupancy_one<V_E5M2_IP, 4>(ix, lc);
        case QK_SQ8: return l2 ? occupancy_one<V_SQ8_L2, 4>(ix, lc) : occupancy_one<V_SQ8_IP, 4>(ix, lc);
        case QK_FP32: return l2 ? occupancy_one<V_FP32_L2, 4>(ix, lc) : occupancy_one<V_FP32_IP, 4>(ix, lc);
        default: return 0;
    }
}

bool team_valid(const DeviceIndex &ix, int32_t team) {
    if (ix.quant != QK_SQ8U && ix.quant != QK_SQ4U && ix.quant != QK_SQ2U) return team == 16;
    if (team != 4 && team != 8 && team != 16) return false;
    const int code_dw = ix.code_size / 4;
    const int nparts = (code_dw + team - 1) / team;
    return nparts <= 64;
}

int32_t default_team(const DeviceIndex &ix) {
    if (ix.quant != QK_SQ8U && ix.quant != QK_SQ4U && ix.quant != QK_SQ2U) return 16;
    const int code_dw = ix.code_size / 4;
    // fewest lanes per candidate that still covers a 32-neighbour row in one pass
    if (code_dw <= 16) return 8;
    if (code_dw <= 64) return 16;
    return 16;
}

// Per-dword "DPL" is the team selector for integer quantizers; float
// quantizers always use 16-lane segments.
hipError_t launch_search(const DeviceIndex &ix, const QueryBatch &qb, const LaunchConfig &lc, const SlotState &ss,
                         hipStream_t stream) {
    const bool l2 = ix.metric == MK_L2;
    switch (ix.quant) {
        case QK_SQ8U: return l2 ? launch_int<V_SQ8U_L2>(ix, qb, lc, ss, stream) : launch_int<V_SQ8U_IP>(ix, qb, lc, ss, stream);
        case QK_SQ4U: return launch_int<V_SQ4U>(ix, qb, lc, ss, stream);
        case QK_SQ2U: return launch_int<V_SQ2U>(ix, qb, lc, ss, stream);
        case QK_E5M2: return l2 ? launch_one<V_E5M2_L2, 4>(ix, qb, lc, ss, stream) : launch_one<V_E5M2_IP, 4>(ix, qb, lc, ss, stream);
        case QK_SQ8: return l2 ? launch_one<V_SQ8_L2, 4>(ix, qb, lc, ss, stream) : launch_one<V_SQ8_IP, 4>(ix, qb, lc, ss, stream);
        case QK_FP32: return l2 ? launch_one<V_FP32_L2, 4>(ix, qb, lc, ss, stream) : launch_one<V_FP32_IP, 4>(ix, qb, lc, ss, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace glass_hip
