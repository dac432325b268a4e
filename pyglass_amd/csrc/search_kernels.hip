// HIP kernels for pyglass query-time graph search on MI355X (gfx950).
//
// One wavefront (64 lanes) owns one query at a time and reproduces the
// reference's best-first expansion exactly:
//   * the candidate pool is glass::LinearPool (glass/neighbor.hpp:320-433)
//     kept sorted in LDS as (key, id|checked) pairs;
//   * the R neighbours of a popped node are loaded from the adjacency row
//     (one coalesced 4R-byte read), tested-and-set in an exact visited set
//     (a per-wave LDS hash table that moves to the slot's global bitset if
//     it fills past half) and kept in adjacency order
//     (glass/searcher/graph_searcher.hpp:196-224);
//   * their distances are computed by "teams" of lanes from coalesced code
//     gathers (integer quantizers: |q|^2 + |c|^2 - 2 q.c with packed dot
//     products, exact; float quantizers: the reference's 16 accumulators and
//     reduction tree, bit-exact); integer codes may be gathered for every
//     neighbour before the visited test to overlap the two;
//   * the whole batch of new candidates is merged into the pool at once.  The
//     merge reproduces sequential LinearPool::insert exactly: the pool after
//     inserting c_1..c_m in order holds the `cap` smallest (dist, insertion
//     order) keys, and `cur_` is the first unchecked index.
// A persistent grid pulls queries from a device work counter, so one launch
// serves any number of queries with per-query ef / k.
#include <mutex>

#include "glass_hip.h"
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

// The wave64 ballot builtin: the lane mask of a compare is used as it is
// (__ballot goes through an int and compares again).
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Marks a wave-uniform value as such (moves it to a scalar register), so that
// branches on it are scalar branches and arithmetic on it runs on the SALU.
__device__ __forceinline__ int32_t uni(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }

// A wave-uniform value (already in a scalar register) that the compiler may
// not hoist tests on out of loops: a hoisted uniform test lives on as a
// 64-bit lane mask, and with many of them the kernel runs out of scalar
// registers and spills them to vector lanes.  Tested where used, the value
// costs one scalar register and an s_cmp.
__device__ __forceinline__ int32_t at_use(int32_t v) {
    v = uni(v);
    asm volatile("" : "+s"(v));
    return v;
}

// Cross-lane moves inside groups of lanes with DPP (no LDS round trip, unlike
// ds_bpermute).  dpp_xor1/dpp_xor2: partner lane l^1 / l^2 (quad_perm);
// dpp_down4/dpp_down8: lane l+4 / l+8 within its row of 16 (row_ror:12 /
// row_ror:8; row_ror:n moves data n lanes up), the partner of the group's
// first lane for groups of 8 / 16 lanes.
__device__ __forceinline__ int32_t dpp_xor1(int32_t v) { return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xf, 0xf, false); }
__device__ __forceinline__ int32_t dpp_xor2(int32_t v) { return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xf, 0xf, false); }
__device__ __forceinline__ int32_t dpp_down4(int32_t v) { return __builtin_amdgcn_mov_dpp(v, 0x12C, 0xf, 0xf, false); }
__device__ __forceinline__ int32_t dpp_down8(int32_t v) { return __builtin_amdgcn_mov_dpp(v, 0x128, 0xf, 0xf, false); }
__device__ __forceinline__ float dpp_xor1(float v) { return __int_as_float(dpp_xor1(__float_as_int(v))); }
__device__ __forceinline__ float dpp_xor2(float v) { return __int_as_float(dpp_xor2(__float_as_int(v))); }
// lane Q of each group of four, broadcast to the group (quad_perm [Q,Q,Q,Q])
template <int Q>
__device__ __forceinline__ float dpp_quad_bcast(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x55 * Q, 0xf, 0xf, false));
}

__device__ __forceinline__ int mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Minimum over the wave (all lanes active), returned to every lane: DPP
// within each row of 16 (quads, then the first lane of each 8- and 16-lane
// group), then the four row minima read with readlane.
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
    uint32_t o = (uint32_t)dpp_xor1((int)v);
    v = o < v ? o : v;
    o = (uint32_t)dpp_xor2((int)v);
    v = o < v ? o : v;
    o = (uint32_t)dpp_down4((int)v);
    v = o < v ? o : v;
    o = (uint32_t)dpp_down8((int)v);
    v = o < v ? o : v;
    const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)v, 0);
    const uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)v, 16);
    const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)v, 32);
    const uint32_t d = (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
    const uint32_t ab = a < b ? a : b, cd = c < d ? c : d;
    return ab < cd ? ab : cd;
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
    V_FP16_L2 = 16,  // helpa::l2_fp32_fp16 / dot_fp32_fp16 (x86/l2_impl.hpp:20-24,74-84, x86/dot_impl.hpp:22-26,74-84)
    V_FP16_IP = 17,
    V_BF16_L2 = 18,  // helpa::l2_fp32_bf16 / dot_fp32_bf16 (x86/l2_impl.hpp:32-36,145-156, x86/dot_impl.hpp:34-38,139-149)
    V_BF16_IP = 19,
    V_SQ6_L2 = 20,   // glass::L2SqrSQ6_ext / IPSQ6_ext (glass/simd/avx512.hpp:79-104)
    V_SQ6_IP = 21,
    V_SQ4_L2 = 22,   // glass::L2SqrSQ4_ext / IPSQ4_ext (glass/simd/avx512.hpp:106-144)
    V_SQ4_IP = 23,
    V_PQ8_L2 = 24,   // ProductQuant::dist_func over the query's LUT (glass/quant/product_quant.hpp:66-90)
    V_PQ8_IP = 25,
};

constexpr bool is_pq_variant(int v) { return v == V_PQ8_L2 || v == V_PQ8_IP; }

// Or-ed into the search mode for initialize_search's evaluations: the
// compiled reference's SQ6 kernel rounds differently there than in Search's
// expansions.
constexpr int32_t kInitCtx = 4;

// 16-bit codes: the fp32 kernel's arithmetic on widened values.
constexpr bool is_f16_variant(int v) { return v == V_FP16_L2 || v == V_FP16_IP; }
constexpr bool is_bf16_variant(int v) { return v == V_BF16_L2 || v == V_BF16_IP; }
constexpr bool is_l2_float(int v) { return v == V_FP32_L2 || v == V_FP16_L2 || v == V_BF16_L2; }

// Four 16-bit codes (one uint2) widened to fp32: fp16 exactly, bf16 as the
// high half of the float.
template <int V>
__device__ __forceinline__ float4 widen4(uint2 h) {
    if (is_bf16_variant(V)) {
        return make_float4(__uint_as_float(h.x << 16), __uint_as_float(h.x & 0xffff0000u), __uint_as_float(h.y << 16),
                           __uint_as_float(h.y & 0xffff0000u));
    }
    return make_float4(__half2float(__ushort_as_half((unsigned short)(h.x & 0xffffu))),
                       __half2float(__ushort_as_half((unsigned short)(h.x >> 16))),
                       __half2float(__ushort_as_half((unsigned short)(h.y & 0xffffu))),
                       __half2float(__ushort_as_half((unsigned short)(h.y >> 16))));
}

constexpr bool is_int_variant(int v) { return v < 10; }

// Per-wave LDS layout.  All offsets are multiples of 16 bytes.
struct WaveLds {
    uint32_t *vt;   // visited hash table, id+1 per slot, 0 = empty [vis_hash]
    uint32_t *pk;   // pool keys [cap_max]
    uint32_t *pid;  // pool ids (bit 31 = checked) [cap_max]
    uint32_t *ck;   // candidate keys [64]
    uint32_t *ci;   // candidate ids [64]
    float *qf;      // query, zero padded [qf_len]
    float *qm;      // SQ8 helpers: 256x-256min, 256min, dif; SQ6: x-min, min, dif; SQ4: 16(x-min), min, dif [3 x qf_len]
                    // (an SQ8 re-rank rewrites them with its own calibration after the search).  Only the L2
                    // kernels read the first, query-dependent row: for IP the area holds the last two rows and
                    // qm is based one row before it (qm[0, qf_len) is neither read nor written then)
    uint8_t *qcode; // encoded query [qcode_bytes]
};

__host__ __device__ inline int32_t round16(int64_t x) { return (int32_t)((x + 15) / 16 * 16); }

__host__ __device__ inline int32_t qf_len(const DeviceIndex &ix) {
    int32_t a = ix.dim_align > ix.r_dim_align ? ix.dim_align : ix.r_dim_align;
    if (a < ix.dim) a = ix.dim;
    return round16(a);
}

// The per-dim helper rows (qm): the search quantizer's, or an SQ8 re-rank's.
__host__ __device__ inline bool needs_qm(const DeviceIndex &ix) {
    return ix.quant == QK_SQ8 || ix.quant == QK_SQ6 || ix.quant == QK_SQ4 || (ix.rerank && ix.rerank_kind == QK_SQ8);
}

// Helper rows the kernels read: the x - min row only for L2.
__host__ __device__ inline int32_t qm_rows(const DeviceIndex &ix) { return ix.metric == MK_L2 ? 3 : 2; }

// Encoded query bytes: the integer search codes' (SQ8U, SQ4U, SQ2U), or an
// SQ8U re-rank's; the float kernels read the query from qf and need none.
__host__ __device__ inline int32_t qcode_bytes(const DeviceIndex &ix) {
    const bool int_codes = ix.quant == QK_SQ8U || ix.quant == QK_SQ4U || ix.quant == QK_SQ2U;
    const int32_t c = int_codes ? ix.code_size : 0;
    const int32_t r = ix.rerank && ix.rerank_kind == QK_SQ8U ? ix.r_code_size : 0;
    return round16(c > r ? c : r);
}

__host__ __device__ inline size_t wave_lds_bytes(const DeviceIndex &ix, int32_t cap_max, int32_t vis_hash) {
    size_t b = (size_t)vis_hash * 4;
    b += (size_t)round16((int64_t)cap_max * 4) * 2;
    b += 64 * 4 * 2;
    b += (size_t)qf_len(ix) * 4;
    if (needs_qm(ix)) b += (size_t)qf_len(ix) * 4 * qm_rows(ix);
    b += (size_t)qcode_bytes(ix);
    return b;
}

__device__ inline WaveLds carve(uint8_t *base, const DeviceIndex &ix, int32_t cap_max, int32_t vis_hash) {
    WaveLds w;
    w.vt = (uint32_t *)base;
    size_t off = (size_t)vis_hash * 4;
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
    w.qm = (float *)(base + off) - (needs_qm(ix) && qm_rows(ix) == 2 ? ql : 0);
    if (needs_qm(ix)) off += (size_t)ql * 4 * qm_rows(ix);
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
__device__ __forceinline__ void encode_query(const DeviceIndex &ix, const WaveLds &w) {
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
// The squared distances use sum (x-y)^2 = sum x^2 + sum y^2 - 2 sum x*y on
// unsigned fields: per code dword one dot instruction for sum x*y and one for
// sum y^2 (v_dot4_u32_u8 on bytes, v_dot8_u32_u4 on nibbles; 2-bit fields are
// split into two nibble dwords with a 0x33333333 mask), sum x^2 of the lane's
// slice of the query precomputed.  All terms are integers below 2^31, so the
// int32 result is the reference's value exactly.

__device__ __forceinline__ uint32_t udot4(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_udot4(a, b, c, false);
}
__device__ __forceinline__ uint32_t udot8(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_udot8(a, b, c, false);
}

template <int V, int DPL>
struct IntQuery {
    uint32_t q[DPL];
    int32_t qq;      // sum x^2 over this lane's slice (squared-L2 variants)
    int32_t slot;    // lanes per candidate (power of two)
    int32_t nparts;  // active parts per candidate
    int32_t part;    // this lane's part

    __device__ __forceinline__ void init(const DeviceIndex &ix, const WaveLds &w) {
        const int code_dw = ix.code_size / 4;
        nparts = (code_dw + DPL - 1) / DPL;
        slot = 1;
        while (slot < nparts) slot <<= 1;
        part = lane_id() & (slot - 1);
        const uint32_t *qd = (const uint32_t *)w.qcode;
        uint32_t acc = 0;
#pragma unroll
        for (int i = 0; i < DPL; ++i) {
            int d = part * DPL + i;
            q[i] = (part < nparts && d < code_dw) ? qd[d] : 0u;
            if (V == V_SQ8U_L2) acc = udot4(q[i], q[i], acc);
            if (V == V_SQ4U) acc = udot8(q[i], q[i], acc);
            if (V == V_SQ2U) {
                const uint32_t a0 = q[i] & 0x33333333u, a1 = (q[i] >> 2) & 0x33333333u;
                acc = udot8(a0, a0, acc);
                acc = udot8(a1, a1, acc);
            }
        }
        qq = (int32_t)acc;
    }

    // This lane's slice of code row `id` (zeros for slices past the row).
    // Unconditional loads (no exec-mask branch per load): a lane with no row
    // (on false) reads row 0, whose distance is never used; slices past the
    // row read as zero.
    __device__ __forceinline__ void load_row(const DeviceIndex &ix, uint32_t id, bool on, uint4 (&c4)[DPL / 4]) const {
        const int code_q = ix.code_size / 16;  // uint4 per row
        const uint4 *src = (const uint4 *)(ix.codes + (int64_t)(on ? id : 0u) * ix.code_size);
#pragma unroll
        for (int i = 0; i < DPL / 4; ++i) {
            const int qi = part * (DPL / 4) + i;
            const bool ok = part < nparts && qi < code_q;
            const uint4 x = src[ok ? qi : 0];
            c4[i] = ok ? x : make_uint4(0, 0, 0, 0);
        }
    }

    // Distance of a row whose slices c4 the lanes of one group hold; valid in
    // the group's first lane (part 0).
    __device__ __forceinline__ int32_t row_dist(const uint4 (&c4)[DPL / 4], bool on) const {
        // no branch: lanes past the row's parts hold zero query and code
        // dwords (partial 0), and a row that is off is loaded as zeros and
        // its result is never used
        (void)on;
        int32_t acc;
        {
            uint32_t qc = 0, cc = 0;
            int32_t ip = 0;
#pragma unroll
            for (int i = 0; i < DPL / 4; ++i) {
                const uint32_t cs[4] = {c4[i].x, c4[i].y, c4[i].z, c4[i].w};
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const uint32_t qv = q[4 * i + r], c = cs[r];
                    if (V == V_SQ8U_L2) {
                        qc = udot4(qv, c, qc);
                        cc = udot4(c, c, cc);
                    } else if (V == V_SQ8U_IP) {
                        ip = sdot4(qv, c, ip);
                    } else if (V == V_SQ4U) {
                        qc = udot8(qv, c, qc);
                        cc = udot8(c, c, cc);
                    } else {  // V_SQ2U
                        const uint32_t c0 = c & 0x33333333u, c1 = (c >> 2) & 0x33333333u;
                        qc = udot8(qv & 0x33333333u, c0, qc);
                        qc = udot8((qv >> 2) & 0x33333333u, c1, qc);
                        cc = udot8(c0, c0, cc);
                        cc = udot8(c1, c1, cc);
                    }
                }
            }
            acc = (V == V_SQ8U_IP) ? ip : (int32_t)((uint32_t)qq + cc - 2u * qc);
        }
        // nested, so that a group of two lanes leaves after two compares
        const int32_t s = at_use(slot);
        if (s >= 2) {
            acc += dpp_xor1(acc);
            if (s >= 4) {
                acc += dpp_xor2(acc);
                if (s >= 8) {
                    acc += dpp_down4(acc);
                    if (s >= 16) {
                        acc += dpp_down8(acc);
                        // across rows of 16: the first lane of each 16-lane group holds its partial sum
                        if (s >= 32) {
                            acc += __shfl_xor(acc, 16);
                            if (s >= 64) acc += __shfl_xor(acc, 32);
                        }
                    }
                }
            }
        }
        return (V == V_SQ8U_IP) ? -acc : acc;
    }

    // Distances of candidates ci[0..m) -> keys ck[0..m).  With several waves
    // on one query, wave `first` of `stride` takes passes first, first +
    // stride, ... of 64 / slot rows each.
    __device__ __forceinline__ void distances(const DeviceIndex &ix, const WaveLds &w, int m, int first = 0, int stride = 1) const {
        const int l = lane_id();
        const int cpr = 64 / slot;
        const int cj = l / slot;
        for (int base = first * cpr; base < m; base += stride * cpr) {
            const int j = base + cj;
            uint4 c4[DPL / 4];
            load_row(ix, j < m ? w.ci[j] : 0u, j < m, c4);
            const int32_t d = row_dist(c4, j < m);
            if (j < m && part == 0) w.ck[j] = key_of_int(d);
        }
    }
};

// ---------------------------------------------------------------------------
// Float distances in the reference's AVX-512 order.  The compiled reference
// keeps 16 accumulators, acc[l] summing elements j = l, l+16, ... in
// increasing j, then reduces the 16 partial sums in a fixed tree.  Here four
// lanes share one row: lane p owns accumulators 4p..4p+3 and reads elements
// 16t + 4p .. 16t + 4p + 3 (one dword of codes or one float4) per step t, so
// every accumulator sees exactly the reference's operand sequence, and 16
// rows are evaluated per pass.
//   E5M2  glass::L2SqrE5M2 / IPE5M2   (glass/simd/avx512.hpp:15-38)
//   SQ8   glass::L2SqrSQ8_ext / IPSQ8_ext (glass/simd/avx512.hpp:40-77)
//   FP32  helpa::l2_fp32_fp32 / dot_fp32_fp32 (x86/l2_impl.hpp:50-71, x86/dot_impl.hpp:52-71)
// The fused multiply-adds are where the compiled reference fuses them.

struct Acc4 {
    float a[4];
};

// Tree over the 16 partial sums held 4 per lane by a 4-lane group.
//   glass::reduce_add_f32x16 (glass/simd/avx512_utils.hpp:12-22):
//     b_i = a_i + a_{i+8}, c_i = b_i + b_{i+4}, result (c0+c1)+(c2+c3)
//   GCC _mm512_reduce_add_ps as helpa uses it (x86/utils.hpp:61-64):
//     same b and c, result (c0+c2)+(c1+c3)
// Lane p holds a_{4p+r}; a_{i+8} sits two lanes up, b_{i+4} one lane up.
template <bool kGlassTree>
__device__ __forceinline__ float reduce16_4lanes(Acc4 v) {
#pragma unroll
    for (int r = 0; r < 4; ++r) v.a[r] = v.a[r] + dpp_xor2(v.a[r]);
#pragma unroll
    for (int r = 0; r < 4; ++r) v.a[r] = v.a[r] + dpp_xor1(v.a[r]);
    return kGlassTree ? (v.a[0] + v.a[1]) + (v.a[2] + v.a[3]) : (v.a[0] + v.a[2]) + (v.a[1] + v.a[3]);
}

__device__ __forceinline__ float e5m2_to_float(uint32_t byte) {
    return __half2float(__ushort_as_half((unsigned short)(byte << 8)));
}

// One row shared by a 4-lane group (p = lane & 3); returns the reduced sum in
// every lane of the group.  The row is read in batches of KB steps of 16
// elements whose loads are all issued before any arithmetic, so a row of
// 16 * KB * n elements waits on memory n times (a 768-d SQ8 row at KB = 8:
// 6 times, against 12 at KB = 4); each accumulator still sums its elements
// in increasing order.  FP32 rows load a float4 per step and use half of KB.
template <int V, int KB = 8>
__device__ __forceinline__ float float_row4(const DeviceIndex &ix, const WaveLds &w, const uint8_t *row, int p,
                                            int32_t mode, int32_t da) {
    Acc4 acc = {{0.0f, 0.0f, 0.0f, 0.0f}};
    if (V == V_E5M2_L2 || V == V_E5M2_IP || V == V_SQ8_L2 || V == V_SQ8_IP) {
        constexpr int kB = KB;  // one code dword per step and lane
        const int ql = qf_len(ix);
        const float *m256 = w.qm + ql, *dif = w.qm + 2 * ql;
        for (int t0 = 0; t0 < da; t0 += 16 * kB) {
            uint32_t cw[kB];
#pragma unroll
            for (int b = 0; b < kB; ++b) {
                const int t = t0 + 16 * b;
                cw[b] = t < da ? *(const uint32_t *)(row + t + 4 * p) : 0u;
            }
#pragma unroll
            for (int b = 0; b < kB; ++b) {
                const int t = t0 + 16 * b;
                if (t >= da) break;
                const int j = t + 4 * p;
                if (V == V_E5M2_L2 || V == V_E5M2_IP) {
                    const float4 x = *(const float4 *)(w.qf + j);
                    const float xs[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const float y = e5m2_to_float((cw[b] >> (8 * r)) & 0xffu);
                        if (V == V_E5M2_L2) {
                            const float d = xs[r] - y;
                            acc.a[r] = fmaf(d, d, acc.a[r]);
                        } else {
                            acc.a[r] = fmaf(xs[r], y, acc.a[r]);
                        }
                    }
                } else {
                    // SearchImpl2 and initialize_search contexts (in Search too):
                    // yy = fma(c, dif, 256min), d = 256x - yy.  Search's expansions and a
                    // bare computer: d = fma(-c, dif, 256x - 256min).  One 16-byte LDS read
                    // per row and 4 dimensions (every LDS row starts on 16 bytes, j % 4 == 0).
                    const bool bform = (mode & 3) == MODE_BATCH || (mode & kInitCtx);
                    const float4 dv = *(const float4 *)(dif + j), mv = *(const float4 *)(m256 + j);
                    const float4 xv = *(const float4 *)((V == V_SQ8_L2 && !bform ? w.qm : w.qf) + j);
                    const float dif4[4] = {dv.x, dv.y, dv.z, dv.w}, m4[4] = {mv.x, mv.y, mv.z, mv.w};
                    const float x4[4] = {xv.x, xv.y, xv.z, xv.w};
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const float c = (float)((cw[b] >> (8 * r)) & 0xffu) + 0.5f;
                        if (V == V_SQ8_L2) {
                            const float d =
                                bform ? x4[r] * 256.0f - fmaf(c, dif4[r], m4[r]) : fmaf(-c, dif4[r], x4[r]);
                            acc.a[r] = fmaf(d, d, acc.a[r]);
                        } else {
                            const float yy = fmaf(c, dif4[r], m4[r]);
                            acc.a[r] = fmaf(x4[r], yy, acc.a[r]);
                        }
                    }
                }
            }
        }
        return reduce16_4lanes<true>(acc);
    } else {  // FP32, FP16, BF16
        constexpr bool kWide = V == V_FP32_L2 || V == V_FP32_IP;
        constexpr int kB = kWide ? KB / 2 : KB;  // 2 KB VGPRs of loads per batch
        for (int t0 = 0; t0 < da; t0 += 16 * kB) {
            float4 yv[kB];
#pragma unroll
            for (int b = 0; b < kB; ++b) {
                const int t = t0 + 16 * b;
                const int j = t + 4 * p;
                if (t >= da) {
                    yv[b] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                } else if (kWide) {
                    yv[b] = *(const float4 *)((const float *)row + j);
                } else {
                    yv[b] = widen4<V>(*(const uint2 *)((const uint16_t *)row + j));
                }
            }
#pragma unroll
            for (int b = 0; b < kB; ++b) {
                const int t = t0 + 16 * b;
                if (t >= da) break;
                const int j = t + 4 * p;
                const float4 x = *(const float4 *)(w.qf + j);
                const float xs[4] = {x.x, x.y, x.z, x.w}, ys[4] = {yv[b].x, yv[b].y, yv[b].z, yv[b].w};
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if (is_l2_float(V)) {
                        const float d = xs[r] - ys[r];
                        acc.a[r] = fmaf(d, d, acc.a[r]);
                    } else {
                        acc.a[r] = fmaf(xs[r], ys[r], acc.a[r]);
                    }
                }
            }
        }
        return reduce16_4lanes<false>(acc);
    }
}

// SQ6 codes (SQ6Quantizer::encode_component, glass/quant/sq6_quant.hpp:41-62):
// components 4g..4g+3 in the 24 bits of bytes 3g..3g+2, six bits each from
// the least significant end.  The reference keeps 8 accumulators, acc[l]
// summing elements j = l, l+8, ... (glass/simd/avx512.hpp:79-104); here lane p
// of a 4-lane group owns acc[2p] and acc[2p+1], i.e. bits [12p, 12p+12) of
// every 48-bit step of 8 components.  The decode ((c + 0.5) / 63) * dif + min
// is rounded as GCC -Ofast compiled it, which depends on the inlining
// context: Search's expansions scale dif first (kd = dif / 63, then
// fma(-a, kd, x - min) or fma(a, kd, min)); SearchBatch and
// initialize_search scale a first (ad = a * dif, then fma(-ad, 1/63, x - min)
// or fma(ad, 1/63, min)).
template <int V>
__device__ __forceinline__ float sq6_row4(const DeviceIndex &ix, const WaveLds &w, const uint8_t *row, int p,
                                          int32_t mode, int32_t da) {
    constexpr float kInv63 = 1.0f / 63.0f;
    const int ql = qf_len(ix);
    const float *xm = w.qm, *mi = w.qm + ql, *dif = w.qm + 2 * ql;
    const bool ad = mode != MODE_SEARCH;
    const uint32_t sh = 12u * (uint32_t)p;
    float acc0 = 0.0f, acc1 = 0.0f;
    auto step = [&](uint64_t bits48, int t) {
        const uint32_t b = (uint32_t)(bits48 >> sh);
        const int j = 8 * t + 2 * p;
        const float ca = (float)(b & 63u) + 0.5f, cb = (float)((b >> 6) & 63u) + 0.5f;
        const float2 dv = *(const float2 *)(dif + j);
        if (V == V_SQ6_L2) {
            const float2 xv = *(const float2 *)(xm + j);
            const float ta = ad ? fmaf(-(ca * dv.x), kInv63, xv.x) : fmaf(-ca, kInv63 * dv.x, xv.x);
            const float tb = ad ? fmaf(-(cb * dv.y), kInv63, xv.y) : fmaf(-cb, kInv63 * dv.y, xv.y);
            acc0 = fmaf(ta, ta, acc0);
            acc1 = fmaf(tb, tb, acc1);
        } else {
            const float2 xv = *(const float2 *)(w.qf + j);
            const float2 mv = *(const float2 *)(mi + j);
            const float ya = ad ? fmaf(ca * dv.x, kInv63, mv.x) : fmaf(ca, kInv63 * dv.x, mv.x);
            const float yb = ad ? fmaf(cb * dv.y, kInv63, mv.y) : fmaf(cb, kInv63 * dv.y, mv.y);
            acc0 = fmaf(xv.x, ya, acc0);
            acc1 = fmaf(xv.y, yb, acc1);
        }
    };
    const int steps = da >> 3;
    if ((ix.code_size & 3) == 0) {
        // rows are dword aligned and hold an even number of steps: 12 bytes per two steps
#pragma unroll 2
        for (int t = 0; t < steps; t += 2) {
            const uint32_t *r32 = (const uint32_t *)(row + 6 * t);
            const uint32_t d0 = r32[0], d1 = r32[1], d2 = r32[2];
            step(((uint64_t)d1 << 32) | d0, t);
            step(((uint64_t)d2 << 16) | (d1 >> 16), t + 1);
        }
    } else {
        // rows of 6 * odd bytes are only 2-byte aligned
#pragma unroll 2
        for (int t = 0; t < steps; ++t) {
            const uint16_t *r16 = (const uint16_t *)(row + 6 * t);
            step(((uint64_t)r16[2] << 32) | ((uint32_t)r16[1] << 16) | r16[0], t);
        }
    }
    // glass::reduce_add_f32x8 (glass/simd/avx2_utils.hpp:11-16): h_i = a_i +
    // a_(i+4), result (h0 + h2) + (h1 + h3).  a_(i+4) sits two lanes up, and
    // after that step lane p holds h_(2p), h_(2p+1) for p < 2: h_(i+2) is one
    // lane up.
    acc0 = acc0 + dpp_xor2(acc0);
    acc1 = acc1 + dpp_xor2(acc1);
    acc0 = acc0 + dpp_xor1(acc0);
    acc1 = acc1 + dpp_xor1(acc1);
    return acc0 + acc1;
}

// SQ4 codes (SQ4Quantizer::encode, glass/quant/sq4_quant.hpp:42-52): dim j in
// the low (even j) or high (odd j) nibble of byte j/2.  The reference's
// distance is a scalar loop over element pairs that GCC -Ofast (associative
// math, never inlined) compiles to one serial sum of per-pair terms, with
// a = c + 0.5:
//   L2  t = fma(-a, dif, 16(x - min)),  p = fma(t0, t0, t1 * t1)
//   IP  y = fma(a, dif / 16, min),       p = fma(x0, y0, x1 * y1), negated at the end
//   sum = ((p_0 + p_1) + p_2) + ...
// Lane q of a 4-lane group computes pair 4t + q (byte q of dword t of the
// row); the group then adds the four pairs of each dword in order, every
// lane holding the running sum.  Starting from -0 the first add returns p_0
// exactly, as the reference's first term does.
template <int V>
__device__ __forceinline__ float sq4_row4(const DeviceIndex &ix, const WaveLds &w, const uint8_t *row, int q,
                                          int32_t da) {
    const int ql = qf_len(ix);
    const float *xm = w.qm, *mi = w.qm + ql, *dif = w.qm + 2 * ql;
    float sum = -0.0f;
    // da is a multiple of 16, so a row is a whole number of 8-byte words
#pragma unroll 2
    for (int t = 0; t < (da >> 3); t += 2) {
        const uint2 c2 = *(const uint2 *)(row + 4 * t);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const uint32_t byte = ((h ? c2.y : c2.x) >> (8 * q)) & 0xffu;
            const int j = 8 * (t + h) + 2 * q;
            const float a0 = (float)(byte & 15u) + 0.5f, a1 = (float)(byte >> 4) + 0.5f;
            const float2 dv = *(const float2 *)(dif + j);
            float p;
            if (V == V_SQ4_L2) {
                const float2 xv = *(const float2 *)(xm + j);
                const float t0 = fmaf(-a0, dv.x, xv.x), t1 = fmaf(-a1, dv.y, xv.y);
                p = fmaf(t0, t0, t1 * t1);
            } else {
                const float2 xv = *(const float2 *)(w.qf + j), mv = *(const float2 *)(mi + j);
                const float y0 = fmaf(a0, dv.x * 0.0625f, mv.x), y1 = fmaf(a1, dv.y * 0.0625f, mv.y);
                p = fmaf(xv.x, y0, xv.y * y1);
            }
            sum = sum + dpp_quad_bcast<0>(p);
            sum = sum + dpp_quad_bcast<1>(p);
            sum = sum + dpp_quad_bcast<2>(p);
            sum = sum + dpp_quad_bcast<3>(p);
        }
    }
    return sum;
}

// Exact distances (Refiner, refiner.hpp:92-110) of pool entries [0, m),
// written over their keys in pk, from fp32 rows (VF = V_FP32_*) or fp16 rows
// (VF = V_FP16_*).  Rows of S * 16 values (S <= 8, a compile-time count, so
// the loop has no branches): each 4-lane group takes R rows per pass
// (entries base + 16r + g) and issues every load of the pass before any
// arithmetic, so a pass waits on memory once; each accumulator still sums
// its elements in increasing order, as float_row4 does.
template <int VF, int R, int S>
__device__ void rerank_rows(const DeviceIndex &ix, const WaveLds &w, int m) {
    constexpr bool kHalf = is_f16_variant(VF);
    const int l = lane_id();
    const int p4 = l & 3, g = l >> 2;
    for (int base = 0; base < m; base += 16 * R) {
        float4 y[R][S];
        uint2 h[R][S];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int j = base + 16 * r + g;
            const uint32_t id = w.pid[j < m ? j : m - 1] & kIdMask;
            const int64_t at = (int64_t)id * (16 * S) + 4 * p4;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                if (kHalf) {
                    h[r][s] = *(const uint2 *)((const uint16_t *)ix.rerank + at + 16 * s);
                } else {
                    y[r][s] = *(const float4 *)((const float *)ix.rerank + at + 16 * s);
                }
            }
        }
        if (kHalf) {
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int s = 0; s < S; ++s) y[r][s] = widen4<VF>(h[r][s]);
        }
        Acc4 acc[R];
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] = Acc4{{0.0f, 0.0f, 0.0f, 0.0f}};
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const float4 x = *(const float4 *)(w.qf + 16 * s + 4 * p4);
            const float xs[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float ys[4] = {y[r][s].x, y[r][s].y, y[r][s].z, y[r][s].w};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    if (is_l2_float(VF)) {
                        const float d = xs[e] - ys[e];
                        acc[r].a[e] = fmaf(d, d, acc[r].a[e]);
                    } else {
                        acc[r].a[e] = fmaf(xs[e], ys[e], acc[r].a[e]);
                    }
                }
            }
        }
        float rs[R];
#pragma unroll
        for (int r = 0; r < R; ++r) rs[r] = reduce16_4lanes<false>(acc[r]);
        wave_sync();
        if (p4 == 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int j = base + 16 * r + g;
                if (j < m) w.pk[j] = key_of_float(is_l2_float(VF) ? rs[r] : -rs[r]);
            }
        }
        wave_sync();
    }
}

// Re-rank with SQ8U codes (Refiner<SQ8QuantizerUniform>): the query is
// encoded with the re-rank calibrator (ComputerImpl, computer.hpp:54-63) and
// each row's integer distance (helpa::l2_s7_s7 / dot_s8_s8, exact in any
// order) is taken by 4 lanes, 16 rows per pass.  The reference's MaxHeap
// orders the integer distances and reports them as floats; both are exact
// below 2^24, so the float keys order them the same way.
// Compiled into the general path only (the common path's launches never
// carry an SQ8U / SQ8 re-rank), which keeps the common path's registers and
// leaves it without calls or a scratch segment.
template <bool kL2>
__device__ __forceinline__ void rerank_sq8u(const DeviceIndex &ix, const WaveLds &w, int m) {
    const int l = lane_id();
    for (int b = l; b < ix.r_code_size; b += 64) {
        uint32_t byte = 0;
        if (b < ix.dim) {
            byte = (kL2 ? affine_code(w.qf[b], ix.r_qmin, ix.r_qdif, 127.0f) : sym_code(w.qf[b], ix.r_qmax)) & 0xffu;
        }
        w.qcode[b] = (uint8_t)byte;
    }
    wave_sync();
    const int p4 = l & 3, g = l >> 2;
    const int dw = ix.r_code_size / 4;
    const uint32_t *qd = (const uint32_t *)w.qcode;
    for (int base = 0; base < m; base += 16) {
        const int j = base + g;
        const uint32_t id = w.pid[j < m ? j : m - 1] & kIdMask;
        const uint32_t *row = (const uint32_t *)((const uint8_t *)ix.rerank + (int64_t)id * ix.r_code_size);
        uint32_t qc = 0, cc = 0, qq = 0;
        int32_t ip = 0;
#pragma unroll 8
        for (int t = p4; t < dw; t += 4) {
            const uint32_t c = row[t], qv = qd[t];
            if (kL2) {
                qc = udot4(qv, c, qc);
                cc = udot4(c, c, cc);
                qq = udot4(qv, qv, qq);
            } else {
                ip = sdot4(qv, c, ip);
            }
        }
        int32_t acc = kL2 ? (int32_t)(qq + cc - 2u * qc) : ip;
        acc += dpp_xor1(acc);
        acc += dpp_xor2(acc);
        wave_sync();
        if (j < m && p4 == 0) w.pk[j] = key_of_float((float)(kL2 ? acc : -acc));
        wave_sync();
    }
}

// Re-rank with per-dim SQ8 codes (Refiner<SQ8Quantizer>): the helper rows
// are rebuilt from the re-rank calibration and each row goes through the SQ8
// kernel in the bare computer's rounding (glass::L2SqrSQ8_ext / IPSQ8_ext).
// General path only, as rerank_sq8u.
template <int VS, int KB>
__device__ __forceinline__ void rerank_sq8(const DeviceIndex &ix, const WaveLds &w, int m) {
    const int l = lane_id();
    const int ql = qf_len(ix);
    wave_sync();
    for (int j = l; j < ql; j += 64) {
        const float mn = j < ix.r_dim_align ? ix.r_mins[j] : 0.0f;
        const float df = j < ix.r_dim_align ? ix.r_difs[j] : 0.0f;
        if (VS == V_SQ8_L2) w.qm[j] = w.qf[j] * 256.0f - mn * 256.0f;
        w.qm[ql + j] = mn * 256.0f;
        w.qm[2 * ql + j] = df;
    }
    wave_sync();
    const int p4 = l & 3, g = l >> 2;
    for (int base = 0; base < m; base += 16) {
        const int j = base + g;
        const uint32_t id = w.pid[j < m ? j : m - 1] & kIdMask;
        const uint8_t *row = (const uint8_t *)ix.rerank + (int64_t)id * ix.r_code_size;
        const float r = float_row4<VS, KB>(ix, w, row, p4, MODE_SEARCH, ix.r_dim_align);
        wave_sync();
        if (j < m && p4 == 0) w.pk[j] = key_of_float(VS == V_SQ8_IP ? -r : r);
        wave_sync();
    }
}

// Exact distances of pool entries [0, m) into pk; rows longer than 128
// values go through float_row4, one row per 4-lane group per pass.
template <int VF, int KB>
__device__ __forceinline__ void rerank(const DeviceIndex &ix, const WaveLds &w, int m) {
    const int da = ix.r_dim_align;  // a multiple of 16
    switch (da) {
        case 16: return rerank_rows<VF, 1, 1>(ix, w, m);
        case 32: return rerank_rows<VF, 1, 2>(ix, w, m);
        case 48: return rerank_rows<VF, 1, 3>(ix, w, m);
        case 64: return rerank_rows<VF, 1, 4>(ix, w, m);
        case 80: return rerank_rows<VF, 1, 5>(ix, w, m);
        case 96: return rerank_rows<VF, 1, 6>(ix, w, m);
        case 112: return rerank_rows<VF, 1, 7>(ix, w, m);
        case 128: return rerank_rows<VF, 1, 8>(ix, w, m);
        default: break;
    }
    const int l = lane_id();
    const int p4 = l & 3, g = l >> 2;
    for (int base = 0; base < m; base += 16) {
        const int j = base + g;
        const uint32_t id = w.pid[j < m ? j : m - 1] & kIdMask;
        const uint8_t *row = (const uint8_t *)ix.rerank + (int64_t)id * da * (is_f16_variant(VF) ? 2 : 4);
        const float rsum = float_row4<VF, KB>(ix, w, row, p4, MODE_BATCH, da);
        wave_sync();
        if (j < m && p4 == 0) w.pk[j] = key_of_float(is_l2_float(VF) ? rsum : -rsum);
        wave_sync();
    }
}

// PQ8 lookup table (ProductQuant::get_lut, glass/quant/product_quant.hpp:66-76):
// entry (sq, k) is helpa::l2_fp32_fp32 (L2) or dot_fp32_fp32 (IP) of the
// query's sub-vector sq and centroid k with d = 4, which GCC -Ofast compiles
// to one 4-wide step: (t0^2 + t2^2) + (t1^2 + t3^2) with t_i = x_i - c_i, or
// -((p0 + p2) + (p1 + p3)) with p_i = x_i * c_i.  Written by the wave into
// its slot's scratch in HBM (nsq x 256 floats).
template <int V>
__device__ __forceinline__ void build_lut(const DeviceIndex &ix, const WaveLds &w, float *lut) {
    const int n = ix.code_size * 256;
    for (int e = lane_id(); e < n; e += 64) {
        const float4 c = *(const float4 *)(ix.pq_codebook + (int64_t)e * 4);
        const float4 x = *(const float4 *)(w.qf + 4 * (e >> 8));
        float v;
        if (V == V_PQ8_L2) {
            const float t0 = x.x - c.x, t1 = x.y - c.y, t2 = x.z - c.z, t3 = x.w - c.w;
            v = (t0 * t0 + t2 * t2) + (t1 * t1 + t3 * t3);
        } else {
            v = -((x.x * c.x + x.z * c.z) + (x.y * c.y + x.w * c.w));
        }
        lut[e] = v;
    }
}

// PQ8 distance: sum over the row's nsq codes of lut[sq][code], in the order
// GCC -Ofast vectorises ProductQuant::dist_func (AVX2 gathers): blocks of 32
// into 8 accumulators, acc_l += (G(l) + G(8+l)) + (G(16+l) + G(24+l)), reduced
// as h_i = acc_i + acc_(i+4), (h0 + h2) + (h1 + h3); one block of 16 into 4
// lanes, v_i = (G(i) + G(4+i)) + (G(8+i) + G(12+i)), added as
// (v1 + v3) + (v0 + v2); then the rest one by one.  Lane p of a 4-lane group
// owns acc_(2p) and acc_(2p+1) and v_p; the reductions are the DPP steps of
// the SQ6 kernel.
template <int V>
__device__ __forceinline__ float pq_row4(const DeviceIndex &ix, const float *lut, const uint8_t *row, int p) {
    const int nsq = ix.code_size;
    auto G = [&](int sq) { return lut[(sq << 8) + row[sq]]; };
    float r = 0.0f;
    int c = 0;
    if (nsq >= 32) {
        float a0 = 0.0f, a1 = 0.0f;
        for (; c + 32 <= nsq; c += 32) {
            const int l0 = c + 2 * p, l1 = l0 + 1;
            a0 = a0 + ((G(l0) + G(l0 + 8)) + (G(l0 + 16) + G(l0 + 24)));
            a1 = a1 + ((G(l1) + G(l1 + 8)) + (G(l1 + 16) + G(l1 + 24)));
        }
        a0 = a0 + dpp_xor2(a0);
        a1 = a1 + dpp_xor2(a1);
        a0 = a0 + dpp_xor1(a0);
        a1 = a1 + dpp_xor1(a1);
        r = a0 + a1;
    }
    if (nsq - c >= 16) {
        const float v = (G(c + p) + G(c + 4 + p)) + (G(c + 8 + p) + G(c + 12 + p));
        const float wv = v + dpp_xor2(v);  // lane 0: v0 + v2, lane 1: v1 + v3
        r = r + (dpp_quad_bcast<1>(wv) + dpp_quad_bcast<0>(wv));
        c += 16;
    }
    for (; c < nsq; ++c) r = r + G(c);
    return r;
}

template <int V, int KB = 8>
__device__ __forceinline__ void float_distances(const DeviceIndex &ix, const WaveLds &w, int m, int32_t mode,
                                                const float *lut, int first = 0, int stride = 1) {
    if (m <= 0) return;
    const int l = lane_id();
    const int p = l & 3, cj = l >> 2;
    for (int base = 16 * first; base < m; base += 16 * stride) {
        const int j = base + cj;
        const int jj = j < m ? j : (m - 1);
        const uint32_t id = w.ci[jj];
        const uint8_t *row = ix.codes + (int64_t)id * ix.code_size;
        float r;
        if (V == V_SQ6_L2 || V == V_SQ6_IP) {
            r = sq6_row4<V>(ix, w, row, p, mode, ix.dim_align);
        } else if (V == V_SQ4_L2 || V == V_SQ4_IP) {
            r = sq4_row4<V>(ix, w, row, p, ix.dim_align);
        } else if (is_pq_variant(V)) {
            r = pq_row4<V>(ix, lut, row, p);  // IP entries are already negated dot products
        } else {
            r = float_row4<V, KB>(ix, w, row, p, mode, ix.dim_align);
        }
        bool neg = (V == V_E5M2_IP || V == V_SQ8_IP || V == V_FP32_IP || V == V_FP16_IP || V == V_BF16_IP ||
                    V == V_SQ6_IP || V == V_SQ4_IP);
        if (j < m && p == 0) w.ck[j] = key_of_float(neg ? -r : r);
    }
}

// ---------------------------------------------------------------------------
// Exact visited set (Bitset, glass/neighbor.hpp:31-79) as an open-addressing
// hash table in LDS holding id+1.  Ids tested together come from one
// deduplicated adjacency row, so they are distinct: each lane either finds
// its id or claims an empty slot with a compare-and-swap.  The table is kept
// below half full; a query that would exceed that switches to the slot's
// global bitset for the ids it has not seen yet, and keeps consulting the
// table for the ids inserted before.  The bitset is all zero whenever no
// query is using it: a query records each word it is the first to set (the
// lane whose fetch-or returns 0) in the slot's dirty list, and clears those
// words when it ends (the whole bitset if the list overflowed).  A query
// thus writes about as many bitset words as it sets, instead of clearing
// N/8 bytes (1.25 MB at 10M rows) when it switches.
struct VisitedSet {
    uint32_t *tab;
    uint32_t mask, shift;
    int32_t count, thr;
    bool global;
    uint32_t *gbits;
    int64_t gwords;
    uint32_t *dirty;     // the slot's dirty list: bitset words this query set first
    int32_t dirty_cap;   // its capacity in words
    int32_t ndirty;      // words recorded (past dirty_cap: the whole bitset is cleared)

    __device__ __forceinline__ void reset(uint32_t *t, int32_t size, uint32_t *g, int64_t gw, uint32_t *dl,
                                          int32_t dcap) {
        tab = t;
        mask = (uint32_t)size - 1u;
        shift = 32u - (uint32_t)__builtin_ctz((uint32_t)size);
        count = 0;
        thr = size / 2;
        global = false;
        gbits = g;
        gwords = gw;
        dirty = dl;
        dirty_cap = dcap;
        ndirty = 0;
        uint4 *t4 = (uint4 *)t;
        for (int i = lane_id(); i < size / 4; i += 64) t4[i] = make_uint4(0, 0, 0, 0);
        wave_sync();
    }

    __device__ __forceinline__ uint32_t home(uint32_t v) const { return (v * 0x9E3779B1u) >> shift; }

    // true when v had not been visited; marks it visited.
    __device__ __forceinline__ bool test_set(uint32_t v, bool active) {
        bool isnew = false, fresh = false;
        if (active) {
            const uint32_t key = v + 1u;
            uint32_t h = home(v);
            if (!global) {
                // the first probe decides almost always: it stays out of the loop
                uint32_t old = atomicCAS(&tab[h], 0u, key);
                while (old != 0u && old != key) {
                    h = (h + 1u) & mask;
                    old = atomicCAS(&tab[h], 0u, key);
                }
                isnew = old == 0u;
            } else {
                bool found = false;
                for (;;) {
                    const uint32_t cur = tab[h];
                    if (cur == key) {
                        found = true;
                        break;
                    }
                    if (cur == 0u) break;
                    h = (h + 1u) & mask;
                }
                if (!found) {
                    const uint32_t bit = 1u << (v & 31u);
                    const uint32_t old =
                        __hip_atomic_fetch_or(&gbits[v >> 5], bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    isnew = !(old & bit);
                    fresh = old == 0u;  // lanes on one word are serialised: one sees 0
                }
            }
        }
        if (global) {
            const uint64_t fm = ballot(fresh);
            if (fm) {
                const int pos = ndirty + mbcnt(fm);
                if (fresh && pos < dirty_cap) dirty[pos] = v >> 5;
                ndirty += __builtin_popcountll(fm);
            }
        }
        return isnew;
    }

    // End of a query: the bitset words it set back to zero.
    __device__ __forceinline__ void cleanup() {
        if (!global) return;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        if (ndirty <= dirty_cap) {
            for (int i = lane_id(); i < ndirty; i += 64) gbits[dirty[i]] = 0u;
        } else {
            uint4 *g4 = (uint4 *)gbits;
            for (int64_t i = lane_id(); i < gwords / 4; i += 64) g4[i] = make_uint4(0, 0, 0, 0);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(0);
        global = false;
    }

    // Account for the ids a row inserted; switch to the global bitset (all
    // zero, see above) before the table could pass half full (a row inserts
    // at most 64 ids).
    __device__ __forceinline__ void account(int inserted) {
        if (global) return;
        count += inserted;
        if (count > thr - 64) global = true;
    }
};

// ---------------------------------------------------------------------------
// Pool operations.

struct PoolState {
    int32_t size, cur, cap, ef;
};

// LinearPool::pop (neighbor.hpp:398-405): mark data[cur] checked, advance
// cur to the next unchecked entry.  Here the pop only marks the entry; the
// scan for the next unchecked entry is left to whoever runs next: the
// register merge finds it from the pool chunks it loads anyway; other paths
// call pool_settle.  Until then ps.cur is a lower bound (every entry before
// it is checked).
__device__ __forceinline__ uint32_t pool_pop(const WaveLds &w, PoolState &ps, uint32_t known = kIdMask) {
    const uint32_t uid = known != kIdMask ? known : uni(w.pid[ps.cur]);
    wave_sync();
    w.pid[ps.cur] = uid | kChecked;  // every lane stores the same value: no exec mask
    wave_sync();
    ps.cur += 1;
    return uid & kIdMask;
}

// ps.cur <- the first unchecked entry at or after it (size if none).
__device__ __forceinline__ void pool_settle(const WaveLds &w, PoolState &ps) {
    const int l = lane_id();
    int c = ps.cur;
    while (c < ps.size) {
        int p = c + l;
        // unconditional read: p < size + 63 <= cap_max + 63, which stays
        // inside pid and the 256-byte ck array that follows it
        const uint32_t y = w.pid[p];
        const uint32_t x = p < ps.size ? y : kChecked;
        uint64_t um = ballot(!(x & kChecked));
        if (um) {
            c += __builtin_ctzll(um);
            break;
        }
        c += 64;
    }
    ps.cur = uni(c < ps.size ? c : ps.size);
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
// General form for any capacity: binary search and chunked shifts in LDS.
__device__ __forceinline__ void pool_merge_lds(const WaveLds &w, PoolState &ps, int m) {
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
    ps.size = uni(new_size);
    ps.cur = uni(new_cur);
}

// Pools of up to kMaxChunks*64 entries are merged from registers: lane l
// holds entries l, l+64, ... of the pool, NCH chunks of 64 (NCH = 2 when the
// launch's pools hold at most 128 entries, else 4).  Each candidate's insertion point
// (entries with key <= its key) is a sum of ballots, the first moved entry
// and the first inserted position come from the rank-0 candidate, and all
// writes happen after all reads.  The node the next pop will take is known
// at that point, so its adjacency row load is started here (pred_row) and
// overlaps the rest of the merge and the pop.
constexpr int kMaxChunks = 4;

// Candidates: lanes with `cand` set (cm = the mask of those lanes), keys kj /
// ids ij, insertion order = lane order.  nj is the adjacency-row entry this lane loads for the prefetch.
template <int NCH>
__device__ __forceinline__ uint32_t pool_merge_reg(const DeviceIndex &ix, const WaveLds &w, PoolState &ps, uint32_t kj,
                                   uint32_t ij, bool cand, uint64_t cm, int nj, int32_t &pred_row) {
    const int l = lane_id();
    uint32_t pkc[NCH], pic[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        const int p = 64 * c + l;
        const bool v = p < ps.size;
        // unconditional reads, then selects: p < 64 * NCH, and the arrays
        // that follow pk and pid (pid, ck, ci: at least 512 + round16(4 *
        // cap_max) bytes) cover 64 * NCH entries for both NCH = 2 and
        // NCH = 4 (used only when cap_max > 128).  Entries past the pool
        // read as checked, with the largest key.
        const uint32_t xk = w.pk[p], xi = w.pid[p];
        pkc[c] = v ? xk : kKeyMax;
        pic[c] = v ? xi : kChecked;
    }
    // the first unchecked entry (pool_pop leaves ps.cur a lower bound), its
    // key and its id, from the registers
    int cur0 = ps.size;
#pragma unroll
    for (int c = NCH - 1; c >= 0; --c) {
        const uint64_t um = ballot((int32_t)pic[c] >= 0);
        cur0 = um ? 64 * c + __builtin_ctzll(um) : cur0;
    }
    uint32_t kc = kKeyMax, idc = kIdMask;
    if (cur0 < ps.size) {
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            if ((cur0 >> 6) == c) {
                kc = (uint32_t)__builtin_amdgcn_readlane((int)pkc[c], cur0 & 63);
                idc = (uint32_t)__builtin_amdgcn_readlane((int)pic[c], cur0 & 63) & kIdMask;
            }
        }
    }
    // insert() rejects d >= data[size-1] when full; such a candidate would
    // also land past cap in the merge and displaces nothing.  An empty pool
    // reads entry 0, which it does not use (size < cap then).
    const uint32_t last = w.pk[ps.size > 0 ? ps.size - 1 : 0];
    const bool act = cand && (ps.size < ps.cap || kj < last);
    const uint64_t am = cm & (ps.size < ps.cap ? ~0ull : ballot(kj < last));  // the lanes' act bits
    int new_size = ps.size, new_cur = cur0, np0 = 0x7fffffff;
    uint32_t id0 = 0;
    if (am) {
        // the next node before the rank loop, so that its adjacency row is
        // requested as early as possible: the rank-0 candidate (least key,
        // then least lane) lands at the pool's upper bound of its key (new
        // keys go after equal old keys)
        const uint32_t key0 = uni(wave_min_u32(((am >> l) & 1) ? kj : kKeyMax));
        const int i0 = __builtin_ctzll(am & ballot(kj == key0));
        int ub0 = 0;
#pragma unroll
        for (int c = 0; c < NCH; ++c) ub0 += __builtin_popcountll(ballot(pkc[c] <= key0));
        id0 = (uint32_t)__builtin_amdgcn_readlane((int)ij, i0);
        np0 = ub0 < ps.cap ? ub0 : 0x7fffffff;
        const int ma = __builtin_popcountll(am);
        new_size = ps.size + ma < ps.cap ? ps.size + ma : ps.cap;
        new_cur = cur0 < ps.size ? cur0 + __builtin_popcountll(am & ballot(kj < kc)) : 0x7fffffff;
        new_cur = new_cur < np0 ? new_cur : np0;
        new_cur = new_cur < new_size ? new_cur : new_size;
    }
    // the next pop takes entry new_cur: the rank-0 candidate or the old first
    // unchecked entry
    uint32_t next = kIdMask;
    new_size = uni(new_size);
    new_cur = uni(new_cur);
    if (new_cur < new_size && new_cur < ps.ef) next = uni((new_cur == np0) ? id0 : idc);
    // an unconditional load (row 0 stands in when there is no next node, the
    // row's last entry for lanes past K; neither is used): a conditional one
    // would keep the old row live and make the loop copy the new one, which
    // waits for the load at the end of the expansion
    pred_row = ix.adj[(int64_t)(next != kIdMask ? next : 0u) * ix.K + (nj < ix.K ? nj : ix.K - 1)];
    int r = 0, ub = 0;
    // le[c]: active candidates whose key is >= pool entry 64c + l; the
    // entry's shift in the merge is the number of the others, ma - le[c]
    int le[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) le[c] = 0;
    if (am) {
        // rank among the active candidates by (key, insertion order) -- one
        // 64-bit compare of (key, lane) pairs -- and each candidate's upper
        // bound in the pool
        const uint64_t kl = ((uint64_t)kj << 32) | (uint32_t)l;
        uint64_t mm = am;
        while (mm) {
            const int i = __builtin_ctzll(mm);
            mm &= ~(1ull << i);
            const uint32_t ki = (uint32_t)__builtin_amdgcn_readlane((int)kj, i);
            r += (((uint64_t)ki << 32) | (uint32_t)i) < kl;
            // unused chunks hold kKeyMax, which no distance key reaches, so
            // their ballots are empty: no per-chunk branch in this loop
            int cnt = 0;
#pragma unroll
            for (int c = 0; c < NCH; ++c) {
                const bool below = pkc[c] <= ki;
                cnt += __builtin_popcountll(ballot(below));
                le[c] += below;
            }
            ub = (l == i) ? cnt : ub;
        }
    }
    if (am) {
        const int lo = np0 < ps.cap ? np0 : ps.cap;  // entries before the first insertion do not move
        // shift of each pool entry = number of candidates with a smaller key
        const int ma = __builtin_popcountll(am);
        int sh[NCH];
#pragma unroll
        for (int c = 0; c < NCH; ++c) sh[c] = ma - le[c];
        wave_sync();
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const int p = 64 * c + l;
            const int dst = p + sh[c];
            if (p >= lo && p < ps.size && dst < ps.cap) {
                w.pk[dst] = pkc[c];
                w.pid[dst] = pic[c];
            }
        }
        if (act && r + ub < ps.cap) {
            w.pk[r + ub] = kj;
            w.pid[r + ub] = ij;
        }
        wave_sync();
        ps.size = new_size;
    }
    ps.cur = new_cur;
    return next;
}

// Candidates 0..m-1 staged in LDS (ck / ci).
template <int NCH>
__device__ __forceinline__ uint32_t pool_merge_reg_lds(const DeviceIndex &ix, const WaveLds &w, PoolState &ps, int m,
                                                       int32_t &pred_row) {
    const int l = lane_id();
    const uint32_t kj = l < m ? w.ck[l] : kKeyMax;
    const uint32_t ij = l < m ? w.ci[l] : 0u;
    return pool_merge_reg<NCH>(ix, w, ps, kj, ij, l < m, m >= 64 ? ~0ull : (1ull << m) - 1ull, l, pred_row);
}

// One result: separate id / distance arrays, or one 64-bit (id, dist) pair
// store, which a host polling host-mapped memory sees whole.
__device__ __forceinline__ void put_result(const QueryBatch &qb, int64_t pos, int32_t id, float d) {
    if (qb.out_pairs) {
        qb.out_pairs[pos] = (uint64_t)(uint32_t)id | ((uint64_t)__float_as_uint(d) << 32);
    } else {
        qb.out_ids[pos] = id;
        if (qb.out_dists) qb.out_dists[pos] = d;
    }
}

// ---------------------------------------------------------------------------
// The search kernel.  P = 1 is the common case compiled on its own: pools of
// at most 128 entries (two-chunk register merge) and, for integer codes, the
// all-neighbour gather; the other paths are left out, which keeps the hot
// loop small and its scalar registers unspilled.  P = 0 handles everything
// with one wave per query; P = 2 is the same with the waves of a workgroup
// sharing each query (LaunchConfig::qwaves), compiled apart so that the
// other two carry none of its code.

// Register budgets: the common path keeps at least 4 waves per SIMD (at most
// 128 VGPRs; left to itself the compiler schedules the re-rank's loads into
// more registers than that).  The general path runs at 3 (168 VGPRs): its
// pools, long rows and visited bookkeeping spill at 128, and its launches are
// mostly held to fewer waves by LDS anyway.  The shared-query path runs at 3
// too: at 2, a batch of 1,000 queries at 4 waves each does not fit the device
// in one round (it spills at 128).  Its SQ2U instance with 16 dwords per lane
// spills at 168 and runs at 2.
template <int V, int DPL, int P>
constexpr int kWavesPerEu = P == 1 ? 4 : ((P == 2 && V == V_SQ2U && DPL == 16) ? 2 : 3);

template <int V, int DPL, int P>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kWavesPerEu<V, DPL, P>))) void search_kernel(DeviceIndex ix,
                                                                                            QueryBatch qb,
                                                                                            LaunchConfig lc,
                                                                                            SlotState ss) {
    constexpr bool kFast = P == 1;
    extern __shared__ __attribute__((aligned(16))) uint8_t lds_raw[];
    const int l = lane_id();
    const int wv = threadIdx.x >> 6;
    const size_t wbytes = wave_lds_bytes(ix, lc.cap_max, lc.vis_hash);
    // Waves per query: 1 (each wave its own query), or -- the P = 2 kernel --
    // every wave of the workgroup on one query: wave 0 leads (pool, visited
    // set, merges, initializer, outputs) and all waves share each expansion's
    // distance passes.  The query's LDS area is then the workgroup's, and two
    // control words follow it (the query index, the expansion's candidates).
    const int W = P == 2 ? at_use(lc.qwaves) : 1;
    const bool lead = W == 1 || wv == 0;
    const WaveLds w = carve(lds_raw + (W > 1 ? (size_t)0 : (size_t)wv * wbytes), ix, lc.cap_max, lc.vis_hash);
    int32_t *const ctl = (int32_t *)(lds_raw + wbytes);
    const int slot = W > 1 ? (int)blockIdx.x : (int)blockIdx.x * lc.waves_per_block + wv;
    // the slot's HBM scratch: PQ8 lookup table (lut_words), the visited
    // bitset's dirty list (dirty_words), then the bitset
    uint32_t *const slot_mem = ss.visited + (int64_t)slot * ss.vis_stride;
    float *const lut = (float *)slot_mem;
    uint32_t *const dlist = slot_mem + ss.lut_words;
    uint32_t *const vis = dlist + ss.dirty_words;
    const int ql = qf_len(ix);

    // The next query's index, and its row when ql <= 256 (xn, four floats
    // per lane), are requested before the current query's output phase, so
    // that their latency overlaps it.
    int qi_next = 0;
    if (W == 1) {
        if (l == 0) qi_next = atomicAdd(ss.work_counter, 1);
        qi_next = uni(qi_next);
    }
    bool have_row = false;
    float xn[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (;;) {
        if (W > 1) {
            if (wv == 0 && l == 0) ctl[0] = atomicAdd(ss.work_counter, 1);
            __syncthreads();
            qi_next = uni(ctl[0]);
        }
        const int qi = qi_next;
        if (qi >= qb.nq) break;

        const int32_t ef = uni(qb.ef ? qb.ef[qi] : qb.ef_uniform);
        const int32_t k = uni(qb.k ? qb.k[qi] : qb.k_uniform);
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
        // A pool larger than the launch sized LDS for (a caller's bound on ef
        // and k that did not hold) is never filled: the query is answered
        // with an empty pool (ids -1) instead.
        const bool fits = ps.cap <= lc.cap_max;

        // stage the fp32 query (zero padded, as MemCpyTag copies into a zeroed buffer)
        const float *x = qb.queries + (int64_t)qi * qb.dim;
        if (lead) {
        if (have_row) {
#pragma unroll
            for (int r = 0; r < 4; ++r)
                if (l + 64 * r < ql) w.qf[l + 64 * r] = xn[r];
        } else {
            for (int j = l; j < ql; j += 64) w.qf[j] = j < ix.dim ? x[j] : 0.0f;
        }
        wave_sync();
        if (V == V_SQ8_L2 || V == V_SQ8_IP) {
            for (int j = l; j < ql; j += 64) {
                float mn = j < ix.dim_align ? ix.sq8_mins[j] : 0.0f;
                float df = j < ix.dim_align ? ix.sq8_difs[j] : 0.0f;
                if (V == V_SQ8_L2) w.qm[j] = w.qf[j] * 256.0f - mn * 256.0f;
                w.qm[ql + j] = mn * 256.0f;
                w.qm[2 * ql + j] = df;
            }
        } else if (V == V_SQ6_L2 || V == V_SQ6_IP || V == V_SQ4_L2 || V == V_SQ4_IP) {
            constexpr float scale = (V == V_SQ4_L2 || V == V_SQ4_IP) ? 16.0f : 1.0f;
            for (int j = l; j < ql; j += 64) {
                float mn = j < ix.dim_align ? ix.sq8_mins[j] : 0.0f;
                float df = j < ix.dim_align ? ix.sq8_difs[j] : 0.0f;
                if (V == V_SQ6_L2 || V == V_SQ4_L2) w.qm[j] = (w.qf[j] - mn) * scale;
                w.qm[ql + j] = mn;
                w.qm[2 * ql + j] = df;
            }
        }
        if (is_pq_variant(V)) {
            build_lut<V>(ix, w, lut);
            __threadfence_block();  // the waves read the table back from the same CU's L1
            wave_sync();
        }
        if (is_int_variant(V)) encode_query<V>(ix, w);
        }  // lead
        if (W > 1) {
            __syncthreads();
        } else {
            wave_sync();
        }
        IntQuery<V, DPL> iq;
        if (is_int_variant(V)) iq.init(ix, w);
        // empty visited set (Bitset::reset, neighbor.hpp:67-74)
        VisitedSet vs;
        if (lead) vs.reset(w.vt, lc.vis_hash, vis, ss.vis_stride - ss.lut_words - ss.dirty_words, dlist, ss.dirty_words);

        auto dists = [&](int m, int32_t mode) __attribute__((always_inline)) {
            if (is_int_variant(V)) {
                iq.distances(ix, w, m);
            } else {
                float_distances<V, kFast ? 4 : 8>(ix, w, m, mode, lut);
            }
            wave_sync();
        };

        // initialize_search (graph.hpp:95-104)
        if (!fits || !lead) {
        } else if (ix.has_init) {
            // HNSWInitializer::initialize (HNSWInitializer.hpp:41-62)
            uint32_t u = (uint32_t)ix.ep;
            const int top = uni(ix.levels[ix.ep]);
            // offset of u's upper lists.  Each candidate's offset is read
            // together with its id, so a hop's list load waits on nothing but
            // the previous hop's distances.
            int64_t uo = 0;
            if (top > 0) {
                const int64_t e = ix.up_off[u];
                uo = ((int64_t)uni((int32_t)(e >> 32)) << 32) | (uint32_t)uni((int32_t)e);
            }
            if (l == 0) w.ci[0] = u;
            wave_sync();
            dists(1, qb.mode | kInitCtx);
            uint32_t du = uni(w.ck[0]);
            cnt_d += 1;
            for (int level = top; level > 0; --level) {
                for (;;) {
                    const int32_t *list = ix.up_lists + uo + (int64_t)(level - 1) * ix.up_K;
                    const int32_t v = l < ix.up_K ? list[l] : -1;
                    const uint64_t inv = ballot(l < ix.up_K && v == -1);
                    const int n = inv ? __builtin_ctzll(inv) : ix.up_K;
                    if (n == 0) break;
                    // a candidate of a level-l list has level >= l, so it has upper lists
                    const int64_t vo = l < n ? ix.up_off[v] : 0;
                    if (l < n) w.ci[l] = (uint32_t)v;
                    wave_sync();
                    dists(n, qb.mode | kInitCtx);
                    cnt_d += n;
                    cnt_u += n;
                    const uint32_t mk = l < n ? w.ck[l] : kKeyMax;
                    // smallest key; its first lane is the first minimum in list order
                    const uint32_t bk = uni(wave_min_u32(mk));
                    const int bi = __builtin_ctzll(ballot(mk == bk));
                    wave_sync();
                    if (bk < du) {  // strict <: the first minimum of the pass wins
                        du = bk;
                        u = (uint32_t)__builtin_amdgcn_readlane(v, bi);
                        uo = ((int64_t)__builtin_amdgcn_readlane((int32_t)(vo >> 32), bi) << 32) |
                             (uint32_t)__builtin_amdgcn_readlane((int32_t)vo, bi);
                    } else {
                        break;
                    }
                }
            }
            if (l == 0) {
                w.pk[0] = du;
                w.pid[0] = u;
            }
            vs.test_set(u, l == 0);
            vs.account(1);
            ps.size = 1;
            ps.cur = 0;
            wave_sync();
        } else {
            // insert every entry point in order, then mark it visited, 64 at
            // a time (insert never consults the visited set, so marking a
            // chunk after inserting it changes nothing).  A repeated entry
            // point is inserted again, as the reference's loop does.
            for (int e0 = 0; e0 < ix.nep; e0 += 64) {
                const int nep = ix.nep - e0 < 64 ? ix.nep - e0 : 64;
                const uint32_t ep = l < nep ? (uint32_t)ix.eps[e0 + l] : 0u;
                if (l < nep) w.ci[l] = ep;
                wave_sync();
                dists(nep, qb.mode | kInitCtx);
                cnt_d += nep;
                if (lc.cap_max <= 64 * kMaxChunks) {
                    int32_t unused = -1;
                    pool_merge_reg_lds<kMaxChunks>(ix, w, ps, nep, unused);
                } else {
                    pool_settle(w, ps);
                    pool_merge_lds(w, ps, nep);
                }
                vs.test_set(ep, l < nep);
                vs.account(nep);
            }
        }

        // best-first expansion (SearchImpl2, graph_searcher.hpp:196-224).
        // After each pop the adjacency row of the pool's next unchecked entry
        // is loaded speculatively, in flight together with the current
        // candidates' code gathers; the next pop uses it when it pops that
        // node (the merge did not insert anything closer), else reloads.
        //
        // Integer codes, when one lane group per neighbour fits the wave
        // (lc.spec): every group loads its neighbour's id straight from the
        // adjacency row and issues that neighbour's code gather at once; the
        // visited test runs while the gathers are in flight, and the rows
        // that turn out visited are discarded.  Candidates stay in their
        // lanes, so lane order is adjacency order for the merge.
        const int32_t spec = kFast ? is_int_variant(V) : (is_int_variant(V) && lc.spec && iq.slot * ix.K <= 64);
        if (W > 1) {
            // Several waves on this query (rows of at most 64 neighbours).
            // Per expansion: the leader pops nodes until one has unvisited
            // neighbours (or the pool is done) and posts their count; every
            // wave computes its share of their distances; the leader merges.
            uint32_t pred = kIdMask;
            int32_t pred_row = -1;
            for (;;) {
                if (lead) {
                    int m = -1;
                    while (fits && ps.cur < ps.size && ps.cur < ps.ef) {
                        const uint32_t u = pool_pop(w, ps, pred);
                        cnt_e += 1;
                        const int32_t v = (u == pred) ? pred_row : (l < ix.K ? ix.adj[(int64_t)u * ix.K + l] : -1);
                        const uint64_t inv = ballot(l < ix.K && v == -1);
                        const int n = inv ? __builtin_ctzll(inv) : ix.K;
                        const bool isnew = vs.test_set((uint32_t)v, l < n);
                        const uint64_t nm = ballot(isnew);
                        if (nm) {
                            m = __builtin_popcountll(nm);
                            vs.account(m);
                            if (isnew) w.ci[mbcnt(nm)] = (uint32_t)v;
                            break;
                        }
                        pred = kIdMask;
                        pool_settle(w, ps);
                        if (ps.cur < ps.size && ps.cur < ps.ef) {
                            pred = uni(w.pid[ps.cur]) & kIdMask;
                            pred_row = l < ix.K ? ix.adj[(int64_t)pred * ix.K + l] : -1;
                        }
                    }
                    if (l == 0) ctl[1] = m;
                }
                __syncthreads();
                const int m = uni(ctl[1]);
                if (m < 0) break;
                if (is_int_variant(V)) {
                    iq.distances(ix, w, m, wv, W);
                } else {
                    float_distances<V>(ix, w, m, qb.mode, lut, wv, W);
                }
                __syncthreads();
                if (lead) {
                    cnt_d += m;
                    if (at_use(lc.cap_max) <= 128) {
                        pred = pool_merge_reg_lds<2>(ix, w, ps, m, pred_row);
                    } else if (lc.cap_max <= 64 * kMaxChunks) {
                        pred = pool_merge_reg_lds<kMaxChunks>(ix, w, ps, m, pred_row);
                    } else {
                        pool_settle(w, ps);
                        pool_merge_lds(w, ps, m);
                        pred = kIdMask;
                    }
                }
            }
        } else {
        const int nj = spec ? l / iq.slot : l;  // adjacency entry this lane reads
        const uint64_t lead_k = ballot(iq.part == 0 && nj < ix.K);  // first lane of each group with an entry
        uint32_t pred = kIdMask;
        int32_t pred_row = -1;
        const int njc = nj < ix.K ? nj : ix.K - 1;  // lanes past K read the last entry and drop it
        while (fits && ps.cur < ps.size && ps.cur < ps.ef) {
            const uint32_t u = pool_pop(w, ps, pred);
            cnt_e += 1;
            const int32_t vr = (u == pred) ? pred_row : ix.adj[(int64_t)u * ix.K + njc];
            const int32_t v = nj < ix.K ? vr : -1;
            if (kFast ? is_int_variant(V) : at_use(spec) != 0) {
                const bool lead = iq.part == 0;
                const uint64_t inv = lead_k & ballot(v == -1);
                const int n = inv ? __builtin_ctzll(inv) / iq.slot : ix.K;
                const bool valid = nj < n;
                uint4 c4[DPL / 4];
                iq.load_row(ix, (uint32_t)v, valid, c4);
                const bool isnew = vs.test_set((uint32_t)v, lead && valid);
                const uint64_t nm = ballot(isnew);
                pred = kIdMask;
                if (nm == 0) {
                    pool_settle(w, ps);
                    if (ps.cur < ps.size && ps.cur < ps.ef) pred = uni(w.pid[ps.cur]) & kIdMask;
                    pred_row = ix.adj[(int64_t)(pred != kIdMask ? pred : 0u) * ix.K + njc];
                    continue;
                }
                const int m = __builtin_popcountll(nm);
                vs.account(m);
                const uint32_t kj = key_of_int(iq.row_dist(c4, valid));
                cnt_d += m;
                if (kFast || at_use(lc.cap_max) <= 128) {
                    pred = pool_merge_reg<2>(ix, w, ps, kj, (uint32_t)v, isnew, nm, nj, pred_row);
                } else if (lc.cap_max <= 64 * kMaxChunks) {
                    pred = pool_merge_reg<kMaxChunks>(ix, w, ps, kj, (uint32_t)v, isnew, nm, nj, pred_row);
                } else {
                    pred_row = -1;
                    if (isnew) {
                        w.ck[mbcnt(nm)] = kj;
                        w.ci[mbcnt(nm)] = (uint32_t)v;
                    }
                    wave_sync();
                    pool_settle(w, ps);
                    pool_merge_lds(w, ps, m);
                }
                continue;
            }
            // The row in chunks of 64 entries (rows of up to 128): each chunk's
            // new neighbours are merged before the next chunk is read, which
            // leaves the pool as the reference's one-at-a-time insertion in
            // adjacency order does.  The first chunk may come from the
            // prefetch; a merge's prefetch covers the first chunk of the next
            // node's row.
            pred = kIdMask;
            int32_t vc = v;
            for (int c0 = 0;; c0 += 64) {
                const int rem = ix.K - c0;  // row entries from c0 on
                const uint64_t inv = ballot(l < rem && vc == -1);
                const int n = inv ? __builtin_ctzll(inv) : (rem < 64 ? rem : 64);
                const bool isnew = vs.test_set((uint32_t)vc, l < n);
                const uint64_t nm = ballot(isnew);
                if (nm) {
                    const int m = __builtin_popcountll(nm);
                    vs.account(m);
                    if (isnew) w.ci[mbcnt(nm)] = (uint32_t)vc;
                    wave_sync();
                    dists(m, qb.mode);
                    cnt_d += m;
                    if (kFast || at_use(lc.cap_max) <= 128) {
                        pred = pool_merge_reg_lds<2>(ix, w, ps, m, pred_row);
                    } else if (lc.cap_max <= 64 * kMaxChunks) {
                        pred = pool_merge_reg_lds<kMaxChunks>(ix, w, ps, m, pred_row);
                    } else {
                        pool_settle(w, ps);
                        pool_merge_lds(w, ps, m);
                        pred = kIdMask;
                    }
                }
                if (inv || rem <= 64) break;
                vc = c0 + 64 + l < ix.K ? ix.adj[(int64_t)u * ix.K + c0 + 64 + l] : -1;
            }
            if (pred == kIdMask) {
                // the pool did not change since the last merge gave no next
                // node: the next pop takes the first unchecked entry
                pool_settle(w, ps);
                if (ps.cur < ps.size && ps.cur < ps.ef) pred = uni(w.pid[ps.cur]) & kIdMask;
                pred_row = ix.adj[(int64_t)(pred != kIdMask ? pred : 0u) * ix.K + njc];
            }
        }

        }  // one wave per query
        if (lead) vs.cleanup();
        // take the next query now (see above)
        if (W == 1) {
            int t = 0;
            if (l == 0) t = atomicAdd(ss.work_counter, 1);
            qi_next = uni(t);
            have_row = qi_next < qb.nq && ql <= 256;
            if (have_row) {
                const float *xq = qb.queries + (int64_t)qi_next * qb.dim;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int j = l + 64 * r;
                    xn[r] = j < ix.dim ? xq[j] : 0.0f;
                }
            }
        }

        // outputs
        const int64_t ob = qb.out_off ? qb.out_off[qi] : (int64_t)qi * qb.k_uniform;
        if (!lead) {
        } else if (!refine) {
            // LinearPool::to_sorted (neighbor.hpp:421-428)
            for (int i = l; i < k; i += 64) {
                int32_t id = -1;
                float d = __builtin_inff();
                if (i < ps.size) {
                    id = (int32_t)(w.pid[i] & kIdMask);
                    d = is_int_variant(V) ? (float)int_of_key(w.pk[i]) : float_of_key(w.pk[i]);
                }
                put_result(qb, ob + i, id, d);
            }
        } else {
            // Refiner::RefineImpl (refiner.hpp:92-110): exact FP32 distances of the
            // first reorder_k pool entries, the k smallest by (dist, id) ascending.
            const int m = kk < ps.size ? kk : ps.size;
            if (ix.rerank_kind == QK_FP16) {
                if (ix.metric == MK_L2) {
                    rerank<V_FP16_L2, kFast ? 4 : 8>(ix, w, m);
                } else {
                    rerank<V_FP16_IP, kFast ? 4 : 8>(ix, w, m);
                }
            } else if (!kFast && ix.rerank_kind == QK_SQ8U) {
                if (ix.metric == MK_L2) {
                    rerank_sq8u<true>(ix, w, m);
                } else {
                    rerank_sq8u<false>(ix, w, m);
                }
            } else if (!kFast && ix.rerank_kind == QK_SQ8) {
                if (ix.metric == MK_L2) {
                    rerank_sq8<V_SQ8_L2, kFast ? 4 : 8>(ix, w, m);
                } else {
                    rerank_sq8<V_SQ8_IP, kFast ? 4 : 8>(ix, w, m);
                }
            } else if (ix.metric == MK_L2) {
                rerank<V_FP32_L2, kFast ? 4 : 8>(ix, w, m);
            } else {
                rerank<V_FP32_IP, kFast ? 4 : 8>(ix, w, m);
            }
            wave_sync();
            // Ranks by (dist, id), with equal pairs (a repeated entry point
            // is in the pool twice, as in the reference's) ranked by pool
            // position, so the ranks are a permutation and each output
            // position is written once.  Equal pairs are indistinguishable,
            // so this is MaxHeap's output order.
            if (m <= 128) {
                // (key, id) pairs in registers, two per lane
                const int j1 = l + 64;
                const uint32_t ka = w.pk[l < m ? l : 0], ia = w.pid[l < m ? l : 0] & kIdMask;
                const uint32_t kb = w.pk[j1 < m ? j1 : 0], ib = w.pid[j1 < m ? j1 : 0] & kIdMask;
                const uint64_t ca = ((uint64_t)ka << 32) | ia, cb = ((uint64_t)kb << 32) | ib;
                int ra = 0, rb = 0;
                const int m0 = m < 64 ? m : 64;
                for (int i = 0; i < m0; ++i) {
                    const uint64_t ci = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)ka, i) << 32) |
                                        (uint32_t)__builtin_amdgcn_readlane((int)ia, i);
                    ra += (ci < ca) | ((ci == ca) & (i < l));
                    rb += ci <= cb;  // every i here is before position j1
                }
                for (int i = 64; i < m; ++i) {
                    const uint64_t ci = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)kb, i - 64) << 32) |
                                        (uint32_t)__builtin_amdgcn_readlane((int)ib, i - 64);
                    ra += ci < ca;  // every i here is after position l
                    rb += (ci < cb) | ((ci == cb) & (i < j1));
                }
                if (l < m && ra < k) put_result(qb, ob + ra, (int32_t)ia, float_of_key(ka));
                if (j1 < m && rb < k) put_result(qb, ob + rb, (int32_t)ib, float_of_key(kb));
            } else {
                for (int j = l; j < m; j += 64) {
                    const uint32_t kj = w.pk[j], ij = w.pid[j] & kIdMask;
                    int rank = 0;
                    for (int i = 0; i < m; ++i) {
                        const uint32_t ki = w.pk[i], ii = w.pid[i] & kIdMask;
                        rank += (ki < kj) | ((ki == kj) & (ii < ij)) | ((ki == kj) & (ii == ij) & (i < j));
                    }
                    if (rank < k) put_result(qb, ob + rank, (int32_t)ij, float_of_key(kj));
                }
            }
            for (int i = m + l; i < k; i += 64) put_result(qb, ob + i, -1, __builtin_inff());
        }
        if (qb.out_counters && lead && l == 0) {
            qb.out_counters[(int64_t)qi * 3 + 0] = cnt_d;
            qb.out_counters[(int64_t)qi * 3 + 1] = cnt_e;
            qb.out_counters[(int64_t)qi * 3 + 2] = cnt_u;
        }
        if (W > 1) {
            __syncthreads();
        } else {
            wave_sync();
        }
    }
}

// Whether a launch runs the P = 1 kernel: pools of at most 128 entries and,
// for integer codes, the all-neighbour gather with one lane group per
// neighbour (slot * K <= 64, slot as IntQuery::init computes it).
template <int V, int DPL>
bool fast_path(const DeviceIndex &ix, const LaunchConfig &lc) {
    if (lc.cap_max > 128) return false;
    if (ix.rerank && (ix.rerank_kind == QK_SQ8U || ix.rerank_kind == QK_SQ8)) return false;
    if (!is_int_variant(V)) return true;
    if (!lc.spec) return false;
    const int code_dw = ix.code_size / 4;
    const int nparts = (code_dw + DPL - 1) / DPL;
    int slot = 1;
    while (slot < nparts) slot <<= 1;
    return slot * ix.K <= 64;
}

template <int V, int DPL>
const void *kernel_fn(const DeviceIndex &ix, const LaunchConfig &lc) {
    if (lc.qwaves > 1) return (const void *)search_kernel<V, DPL, 2>;
    return fast_path<V, DPL>(ix, lc) ? (const void *)search_kernel<V, DPL, 1> : (const void *)search_kernel<V, DPL, 0>;
}

// A workgroup whose waves share one query: its query area plus the two
// control words (query index, candidates of the expansion), rounded to 16.
constexpr size_t kCtlBytes = 16;

__host__ __device__ inline size_t block_lds_bytes(const DeviceIndex &ix, const LaunchConfig &lc) {
    const size_t wb = wave_lds_bytes(ix, lc.cap_max, lc.vis_hash);
    return lc.qwaves > 1 ? wb + kCtlBytes : wb * (size_t)lc.waves_per_block;
}

template <int V, int DPL>
hipError_t launch_one(const DeviceIndex &ix, const QueryBatch &qb, const LaunchConfig &lc, const SlotState &ss,
                      hipStream_t stream) {
    const size_t lds = block_lds_bytes(ix, lc);
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void *)search_kernel<V, DPL, 0>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        (void)hipFuncSetAttribute((const void *)search_kernel<V, DPL, 1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        (void)hipFuncSetAttribute((const void *)search_kernel<V, DPL, 2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        attr_set = true;
    }
    if (lc.qwaves > 1) {
        hipLaunchKernelGGL((search_kernel<V, DPL, 2>), dim3(lc.blocks), dim3(64 * lc.waves_per_block), lds, stream,
                           ix, qb, lc, ss);
    } else if (fast_path<V, DPL>(ix, lc)) {
        hipLaunchKernelGGL((search_kernel<V, DPL, 1>), dim3(lc.blocks), dim3(64 * lc.waves_per_block), lds, stream,
                           ix, qb, lc, ss);
    } else {
        hipLaunchKernelGGL((search_kernel<V, DPL, 0>), dim3(lc.blocks), dim3(64 * lc.waves_per_block), lds, stream,
                           ix, qb, lc, ss);
    }
    return hipGetLastError();
}

// Resident workgroups per CU from the kernel's register count and the CDNA
// limits: 4 SIMDs x 512 VGPRs per lane (granule 8), at most 8 waves per SIMD,
// 160 KiB of LDS per CU.
template <int V, int DPL>
int occupancy_one(const DeviceIndex &ix, const LaunchConfig &lc) {
    const size_t lds = block_lds_bytes(ix, lc);
    hipFuncAttributes attr;
    if (hipFuncGetAttributes(&attr, kernel_fn<V, DPL>(ix, lc)) != hipSuccess) return 0;
    int vgpr = attr.numRegs > 0 ? attr.numRegs : 128;
    vgpr = (vgpr + 7) / 8 * 8;
    int waves_simd = 512 / vgpr;
    if (waves_simd > 8) waves_simd = 8;
    int by_regs = 4 * waves_simd / lc.waves_per_block;
    int by_lds = lds ? (int)((160 * 1024) / lds) : by_regs;
    int nb = by_regs < by_lds ? by_regs : by_lds;
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

// Scratch of glass_hip_searcher_search_batch_dev_noalloc: the persistent
// grid's work counter and one slot area (PQ8 lookup table, overflow visited
// bitset) per resident wave, for callers that cannot allocate device memory.
// Each device holds its own copy from module load until the process ends.
// Host code uses the address noalloc_scratch returns, never this name.
constexpr size_t kNoAllocScratchBytes = (size_t)128 << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

int noalloc_scratch(int device, void **ptr, size_t *bytes) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void *addr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) return set_error(GLASS_HIP_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[device]) {
        int cur = 0;
        if (hipGetDevice(&cur) != hipSuccess) return set_error(GLASS_HIP_EDEVICE, "hipGetDevice failed");
        if (cur != device && hipSetDevice(device) != hipSuccess) return set_error(GLASS_HIP_EDEVICE, "hipSetDevice failed");
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p) return set_error(GLASS_HIP_EDEVICE, "hipGetSymbolAddress(scratch) failed");
        addr[device] = p;
    }
    *ptr = addr[device];
    *bytes = kNoAllocScratchBytes;
    return 0;
}

size_t lds_bytes_per_wave(const DeviceIndex &ix, int32_t cap_max, int32_t vis_hash) {
    return wave_lds_bytes(ix, cap_max, vis_hash);
}

size_t lds_bytes_per_block(const DeviceIndex &ix, const LaunchConfig &lc) { return block_lds_bytes(ix, lc); }

template <int V, int DPL>
int vgprs_one(const DeviceIndex &ix, const LaunchConfig &lc) {
    hipFuncAttributes attr;
    if (hipFuncGetAttributes(&attr, kernel_fn<V, DPL>(ix, lc)) != hipSuccess) return 0;
    return attr.numRegs;
}

template <int V>
int vgprs_int(const DeviceIndex &ix, const LaunchConfig &lc) {
    switch (lc.team) {
        case 4: return vgprs_one<V, 4>(ix, lc);
        case 8: return vgprs_one<V, 8>(ix, lc);
        case 16: return vgprs_one<V, 16>(ix, lc);
        default: return 0;
    }
}

int kernel_vgprs(const DeviceIndex &ix, const LaunchConfig &lc) {
    const bool l2 = ix.metric == MK_L2;
    switch (ix.quant) {
        case QK_SQ8U: return l2 ? vgprs_int<V_SQ8U_L2>(ix, lc) : vgprs_int<V_SQ8U_IP>(ix, lc);
        case QK_SQ4U: return vgprs_int<V_SQ4U>(ix, lc);
        case QK_SQ2U: return vgprs_int<V_SQ2U>(ix, lc);
        case QK_E5M2: return l2 ? vgprs_one<V_E5M2_L2, 4>(ix, lc) : vgprs_one<V_E5M2_IP, 4>(ix, lc);
        case QK_SQ8: return l2 ? vgprs_one<V_SQ8_L2, 4>(ix, lc) : vgprs_one<V_SQ8_IP, 4>(ix, lc);
        case QK_FP32: return l2 ? vgprs_one<V_FP32_L2, 4>(ix, lc) : vgprs_one<V_FP32_IP, 4>(ix, lc);
        case QK_FP16: return l2 ? vgprs_one<V_FP16_L2, 4>(ix, lc) : vgprs_one<V_FP16_IP, 4>(ix, lc);
        case QK_BF16: return l2 ? vgprs_one<V_BF16_L2, 4>(ix, lc) : vgprs_one<V_BF16_IP, 4>(ix, lc);
        case QK_SQ6: return l2 ? vgprs_one<V_SQ6_L2, 4>(ix, lc) : vgprs_one<V_SQ6_IP, 4>(ix, lc);
        case QK_SQ4: return l2 ? vgprs_one<V_SQ4_L2, 4>(ix, lc) : vgprs_one<V_SQ4_IP, 4>(ix, lc);
        case QK_PQ8: return l2 ? vgprs_one<V_PQ8_L2, 4>(ix, lc) : vgprs_one<V_PQ8_IP, 4>(ix, lc);
        default: return 0;
    }
}

int max_blocks_per_cu(const DeviceIndex &ix, const LaunchConfig &lc) {
    const bool l2 = ix.metric == MK_L2;
    switch (ix.quant) {
        case QK_SQ8U: return l2 ? occupancy_int<V_SQ8U_L2>(ix, lc) : occupancy_int<V_SQ8U_IP>(ix, lc);
        case QK_SQ4U: return occupancy_int<V_SQ4U>(ix, lc);
        case QK_SQ2U: return occupancy_int<V_SQ2U>(ix, lc);
        case QK_E5M2: return l2 ? occupancy_one<V_E5M2_L2, 4>(ix, lc) : occupancy_one<V_E5M2_IP, 4>(ix, lc);
        case QK_SQ8: return l2 ? occupancy_one<V_SQ8_L2, 4>(ix, lc) : occupancy_one<V_SQ8_IP, 4>(ix, lc);
        case QK_FP32: return l2 ? occupancy_one<V_FP32_L2, 4>(ix, lc) : occupancy_one<V_FP32_IP, 4>(ix, lc);
        case QK_FP16: return l2 ? occupancy_one<V_FP16_L2, 4>(ix, lc) : occupancy_one<V_FP16_IP, 4>(ix, lc);
        case QK_BF16: return l2 ? occupancy_one<V_BF16_L2, 4>(ix, lc) : occupancy_one<V_BF16_IP, 4>(ix, lc);
        case QK_SQ6: return l2 ? occupancy_one<V_SQ6_L2, 4>(ix, lc) : occupancy_one<V_SQ6_IP, 4>(ix, lc);
        case QK_SQ4: return l2 ? occupancy_one<V_SQ4_L2, 4>(ix, lc) : occupancy_one<V_SQ4_IP, 4>(ix, lc);
        case QK_PQ8: return l2 ? occupancy_one<V_PQ8_L2, 4>(ix, lc) : occupancy_one<V_PQ8_IP, 4>(ix, lc);
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
        case QK_FP16: return l2 ? launch_one<V_FP16_L2, 4>(ix, qb, lc, ss, stream) : launch_one<V_FP16_IP, 4>(ix, qb, lc, ss, stream);
        case QK_BF16: return l2 ? launch_one<V_BF16_L2, 4>(ix, qb, lc, ss, stream) : launch_one<V_BF16_IP, 4>(ix, qb, lc, ss, stream);
        case QK_SQ6: return l2 ? launch_one<V_SQ6_L2, 4>(ix, qb, lc, ss, stream) : launch_one<V_SQ6_IP, 4>(ix, qb, lc, ss, stream);
        case QK_SQ4: return l2 ? launch_one<V_SQ4_L2, 4>(ix, qb, lc, ss, stream) : launch_one<V_SQ4_IP, 4>(ix, qb, lc, ss, stream);
        case QK_PQ8: return l2 ? launch_one<V_PQ8_L2, 4>(ix, qb, lc, ss, stream) : launch_one<V_PQ8_IP, 4>(ix, qb, lc, ss, stream);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace glass_hip
