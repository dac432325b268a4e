// Device-side SetData: the calibrators' order statistics (radix select below),
// then encode every base row with the trained calibrator
// (Quantizer::add, e.g. glass/quant/sq4u_quant.hpp:99-117), or convert it to
// fp16 / bf16 (fp16_quant.hpp:27-35, bf16_quant.hpp:26-38), and lay out the
// FP32 re-rank rows (FP32Quantizer::add, glass/quant/fp32_quant.hpp:25-31).
// One thread per output byte; the arithmetic is the reference's
// transform(): (x - min) / dif, clamp to [0,1], round half away from zero.
#include "glass_hip_internal.h"

namespace glass_hip {

namespace {

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

__global__ void encode_kernel(const float *data, int64_t n, int32_t dim, int32_t code_size, int32_t quant,
                              int32_t metric, float mn, float dif, float mx, const float *mins, const float *difs,
                              uint8_t *codes) {
    const int64_t total = n * (int64_t)code_size;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = idx / code_size;
        const int32_t b = (int32_t)(idx - row * code_size);
        const float *x = data + row * dim;
        uint32_t byte = 0;
        switch (quant) {
            case QK_SQ8U:  // sq8u_quant.hpp:38-42
                if (b < dim) byte = (metric == MK_L2 ? affine_code(x[b], mn, dif, 127.0f) : sym_code(x[b], mx)) & 0xffu;
                break;
            case QK_SQ4U: {  // sq4u_quant.hpp:107-117
                int j = 2 * b;
                if (j < dim) byte |= affine_code(x[j], mn, dif, 15.0f) & 0xfu;
                if (j + 1 < dim) byte |= (affine_code(x[j + 1], mn, dif, 15.0f) & 0xfu) << 4;
                break;
            }
            case QK_SQ2U: {  // sq2u_quant.hpp:186-200
                for (int s = 0; s < 4; ++s) {
                    int j = 4 * b + s;
                    if (j < dim) byte |= (affine_code(x[j], mn, dif, 3.0f) & 0x3u) << (2 * s);
                }
                break;
            }
            case QK_E5M2:  // e5m2_quant.hpp:30-34: high byte of the RNE fp16
                if (b < dim) byte = ((uint32_t)__half_as_ushort(__float2half_rn(x[b])) >> 8) & 0xffu;
                break;
            case QK_SQ8:  // sq8_quant.hpp:43-47, AffinePerDimCalibrator<255>
                if (b < dim) byte = affine_code(x[b], mins[b], difs[b], 255.0f) & 0xffu;
                break;
            case QK_SQ4: {  // sq4_quant.hpp:42-52, AffinePerDimCalibrator<15>: dim j -> nibble j&1 of byte j/2
                const int32_t j = 2 * b;
                if (j < dim) byte |= affine_code(x[j], mins[j], difs[j], 15.0f) & 0xfu;
                if (j + 1 < dim) byte |= (affine_code(x[j + 1], mins[j + 1], difs[j + 1], 15.0f) & 0xfu) << 4;
                break;
            }
            case QK_SQ6: {
                // SQ6Quantizer::encode_component (sq6_quant.hpp:41-62): byte r
                // of group g holds bits [8r, 8r+8) of components 4g..4g+3,
                // six bits each from the least significant end
                const int32_t g = b / 3, r = b - 3 * g;
                uint32_t bits = 0;
                for (int s = 0; s < 4; ++s) {
                    const int32_t j = 4 * g + s;
                    if (j < dim) bits |= (affine_code(x[j], mins[j], difs[j], 63.0f) & 63u) << (6 * s);
                }
                byte = (bits >> (8 * r)) & 0xffu;
                break;
            }
            case QK_FP16: {
                // FP16Quantizer::encode -> helpa::fp32_to_fp16 (RNE).  Its
                // block converter (x86/data_convert.hpp:18-31) also fills the
                // padding [dim, dim_align) from the floats that follow the row
                // in the input; past the end of the input the padding is 0.
                const int32_t j = b >> 1;
                const int64_t at = row * dim + j;
                float v = 0.0f;
                if (j < dim) v = x[j];
                else if (at < n * (int64_t)dim) v = data[at];
                const uint32_t h = __half_as_ushort(__float2half_rn(v));
                byte = (b & 1) ? (h >> 8) : (h & 0xffu);
                break;
            }
            case QK_BF16: {  // bf16_quant.hpp:34-38: helpa::bf16, (bits + 0x8000) >> 16; padding 0
                const int32_t j = b >> 1;
                const uint32_t h = j < dim ? ((__float_as_uint(x[j]) + 0x8000u) >> 16) : 0u;
                byte = (b & 1) ? (h >> 8) : (h & 0xffu);
                break;
            }
            default:
                break;
        }
        codes[idx] = (uint8_t)byte;
    }
}

// ProductQuant::encode (glass/quant/product_quant.hpp:52-64): for each 4-dim
// sub-space the first centroid of least L2 distance, the distance rounded as
// helpa::l2_fp32_fp32 with d = 4 (t_i = x_i - c_i, (t0^2 + t2^2) + (t1^2 + t3^2)).
// One thread per (row, sub-space).
__global__ void pq_encode_kernel(const float *data, int64_t n, int32_t nsq, const float *codebook, uint8_t *codes) {
    const int64_t total = n * (int64_t)nsq;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = idx / nsq;
        const int32_t sq = (int32_t)(idx - row * nsq);
        const float4 x = *(const float4 *)(data + row * (int64_t)nsq * 4 + 4 * sq);
        const float4 *cb = (const float4 *)(codebook + (int64_t)sq * 256 * 4);
        float best = 3.402823466e+38f;
        uint32_t code = 0;
        for (int k = 0; k < 256; ++k) {
            const float4 c = cb[k];
            const float t0 = x.x - c.x, t1 = x.y - c.y, t2 = x.z - c.z, t3 = x.w - c.w;
            const float v = (t0 * t0 + t2 * t2) + (t1 * t1 + t3 * t3);
            if (v < best) {
                best = v;
                code = (uint32_t)k;
            }
        }
        codes[idx] = (uint8_t)code;
    }
}

// helpa::l2_fp32_fp32 for any d (x86/l2_impl.hpp:14-18,50-71 with
// ref/l2_ref.hpp:7-15), rounded as GCC -Ofast compiles it for the reference:
// the 16-lane body accumulates fma(t, t, a_l) for elements l, l+16, ... and
// reduces b_i = a_(i+8) + a_i, c_i = b_(i+4) + b_i, (c0 + c2) + (c1 + c3);
// the tail of d mod 16 elements takes one 8-wide step (squares rounded,
// u_i = s_(i+4) + s_i, (u1 + u3) + (u0 + u2)), one 4-wide step
// ((s1 + s3) + (s0 + s2), added), then fused multiply-adds; body + tail.
__device__ float l2_helpa(const float *x, const float *y, int32_t d) {
    const int32_t da = d / 16 * 16;
    float A = 0.0f;
    if (da > 0) {
        float a[16];
#pragma unroll
        for (int l = 0; l < 16; ++l) a[l] = 0.0f;
        for (int32_t i = 0; i < da; i += 16) {
#pragma unroll
            for (int l = 0; l < 16; ++l) {
                const float t = x[i + l] - y[i + l];
                a[l] = fmaf(t, t, a[l]);
            }
        }
        float b[8], c[4];
#pragma unroll
        for (int l = 0; l < 8; ++l) b[l] = a[l + 8] + a[l];
#pragma unroll
        for (int l = 0; l < 4; ++l) c[l] = b[l + 4] + b[l];
        A = (c[0] + c[2]) + (c[1] + c[3]);
    }
    float T = 0.0f;
    int32_t i = da;
    if (d - i >= 8) {
        float u[4];
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            const float t0 = x[i + l] - y[i + l], t1 = x[i + 4 + l] - y[i + 4 + l];
            u[l] = t1 * t1 + t0 * t0;
        }
        T = (u[1] + u[3]) + (u[0] + u[2]);
        i += 8;
    }
    if (d - i >= 4) {
        float q[4];
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            const float t = x[i + l] - y[i + l];
            q[l] = t * t;
        }
        T = T + ((q[1] + q[3]) + (q[0] + q[2]));
        i += 4;
    }
    for (; i < d; ++i) {
        const float t = x[i] - y[i];
        T = fmaf(t, t, T);
    }
    return A + T;
}

// Clustering's compute_dist + find_index (glass/algorithms/clustering.hpp:198-232):
// each row's first nearest centroid.  One thread per row; the lanes of a wave
// read the same centroid at the same time.
__global__ void cluster_assign_kernel(const float *x, int32_t n, int32_t d, const float *cent, int32_t k,
                                      int32_t *ids) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *xi = x + i * d;
    float best = __int_as_float(0x7f800000);
    int32_t bid = -1;
    for (int32_t j = 0; j < k; ++j) {
        const float v = l2_helpa(xi, cent + (int64_t)j * d, d);
        if (v < best) {
            best = v;
            bid = j;
        }
    }
    ids[i] = bid;
}

// k-means++ seeding (clustering.hpp:138-196): distance of every row to one
// picked row, as the symmetric computer evaluates it (picked row first).
__global__ void cluster_dist_kernel(const float *x, int32_t n, int32_t d, const float *pivot, float *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = l2_helpa(pivot, x + i * d, d);
}

__global__ void pad_rows_kernel(const float *data, int64_t n, int32_t dim, int32_t dim_align, float *out) {
    const int64_t total = n * (int64_t)dim_align;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = idx / dim_align;
        const int32_t j = (int32_t)(idx - row * dim_align);
        out[idx] = j < dim ? data[row * dim + j] : 0.0f;
    }
}

__global__ void fill_i32_kernel(int32_t *p, int64_t n, int32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Order-preserving 32-bit key of a float (of |f| when absval): unsigned
// comparison of keys is the float order, and key_value inverts it.
__device__ __forceinline__ uint32_t order_key(float f, int32_t absval) {
    uint32_t u = __float_as_uint(f);
    if (absval) u &= 0x7fffffffu;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// One radix-select pass: per segment (blockIdx.y) and per target t in {0, 1},
// a 256-bin histogram of digit (key >> shift) & 255 over the items whose key
// agrees with prefix[t] above that digit.
__global__ void radix_hist_kernel(const float *data, int64_t seg_len, int64_t seg_stride, int32_t absval,
                                  int32_t shift, const uint32_t *prefix, unsigned long long *hist) {
    __shared__ unsigned long long h[512];
    for (int i = threadIdx.x; i < 512; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const int64_t seg = blockIdx.y;
    const float *p = data + seg * seg_stride;
    const uint32_t p0 = prefix[2 * seg], p1 = prefix[2 * seg + 1];
    const uint32_t hm = shift >= 24 ? 0u : (0xffffffffu << (shift + 8));
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < seg_len;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t k = order_key(p[i], absval);
        const uint32_t d = (k >> shift) & 255u;
        if (((k ^ p0) & hm) == 0) atomicAdd(&h[d], 1ull);
        if (((k ^ p1) & hm) == 0) atomicAdd(&h[256 + d], 1ull);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 512; i += blockDim.x) {
        const unsigned long long v = h[i];
        if (v) atomicAdd(&hist[seg * 512 + i], v);
    }
}

// Picks, for every (segment, target), the bin holding the target's remaining
// rank: appends the digit to the prefix and makes the rank relative to it.
__global__ void radix_pick_kernel(int32_t ntargets, int32_t shift, const unsigned long long *hist, uint32_t *prefix,
                                  unsigned long long *rank) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntargets) return;
    const unsigned long long *h = hist + (int64_t)t * 256;
    const unsigned long long r = rank[t];
    unsigned long long c = 0;
    int b = 0;
    for (; b < 255; ++b) {
        if (c + h[b] > r) break;
        c += h[b];
    }
    rank[t] = r - c;
    prefix[t] |= (uint32_t)b << shift;
}

int grid_for(int64_t total) {
    int64_t g = (total + 255) / 256;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (int)g;
}

}  // namespace

hipError_t launch_encode(const float *data_dev, int64_t n, int32_t dim, int32_t code_size, int32_t quant,
                         int32_t metric, float mn, float dif, float mx, const float *mins, const float *difs,
                         uint8_t *codes_dev, hipStream_t stream) {
    const int64_t total = n * (int64_t)code_size;
    hipLaunchKernelGGL(encode_kernel, dim3(grid_for(total)), dim3(256), 0, stream, data_dev, n, dim, code_size, quant,
                       metric, mn, dif, mx, mins, difs, codes_dev);
    return hipGetLastError();
}

hipError_t launch_pq_encode(const float *data_dev, int64_t n, int32_t nsq, const float *codebook, uint8_t *codes_dev,
                            hipStream_t stream) {
    const int64_t total = n * (int64_t)nsq;
    hipLaunchKernelGGL(pq_encode_kernel, dim3(grid_for(total)), dim3(256), 0, stream, data_dev, n, nsq, codebook,
                       codes_dev);
    return hipGetLastError();
}

hipError_t launch_cluster_assign(const float *x, int32_t n, int32_t d, const float *cent, int32_t k, int32_t *ids,
                                 hipStream_t stream) {
    hipLaunchKernelGGL(cluster_assign_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, x, n, d, cent, k, ids);
    return hipGetLastError();
}

hipError_t launch_cluster_dist(const float *x, int32_t n, int32_t d, const float *pivot, float *out,
                               hipStream_t stream) {
    hipLaunchKernelGGL(cluster_dist_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, x, n, d, pivot, out);
    return hipGetLastError();
}

hipError_t launch_pad_rows(const float *data_dev, int64_t n, int32_t dim, int32_t dim_align, float *out,
                           hipStream_t stream) {
    const int64_t total = n * (int64_t)dim_align;
    hipLaunchKernelGGL(pad_rows_kernel, dim3(grid_for(total)), dim3(256), 0, stream, data_dev, n, dim, dim_align,
                       out);
    return hipGetLastError();
}

hipError_t launch_radix_hist(const float *data, int32_t nseg, int64_t seg_len, int64_t seg_stride, int32_t absval,
                             int32_t shift, const uint32_t *prefix, unsigned long long *hist, hipStream_t stream) {
    int64_t bx = (seg_len + 256 * 16 - 1) / (256 * 16);
    if (bx > 2048) bx = 2048;
    if (bx < 1) bx = 1;
    hipLaunchKernelGGL(radix_hist_kernel, dim3((unsigned)bx, (unsigned)nseg), dim3(256), 0, stream, data, seg_len,
                       seg_stride, absval, shift, prefix, hist);
    return hipGetLastError();
}

hipError_t launch_radix_pick(int32_t ntargets, int32_t shift, const unsigned long long *hist, uint32_t *prefix,
                             unsigned long long *rank, hipStream_t stream) {
    hipLaunchKernelGGL(radix_pick_kernel, dim3((ntargets + 255) / 256), dim3(256), 0, stream, ntargets, shift, hist,
                       prefix, rank);
    return hipGetLastError();
}

// Graph-capture-safe clear of small device counters (no hipMemsetAsync).
hipError_t launch_fill_i32(int32_t *p, int64_t n, int32_t v, hipStream_t stream) {
    hipLaunchKernelGGL(fill_i32_kernel, dim3(grid_for(n)), dim3(256), 0, stream, p, n, v);
    return hipGetLastError();
}

}  // namespace glass_hip
