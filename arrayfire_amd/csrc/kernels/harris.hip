// Harris corner detector: af_harris (src/api/c/harris.cpp:52-95), whose
// backend (src/backend/cpu/harris.cpp:26-148, kernels
// src/backend/cpu/kernel/harris.hpp:18-118) is a caller of the separable
// convolution: it smooths the three gradient products with the window filter
// through convolve2<T, convAccT>(..., filter, filter, false).
//
//  * harris_deriv: iy = gradient along dim0, ix = gradient along dim1
//    (central differences x 0.5 inside, one-sided at the ends,
//    src/backend/cpu/kernel/gradient.hpp:18-80) and the products
//    ixx = ix*ix, ixy = ix*iy, iyy = iy*iy, written as the three slices of a
//    [W, H, 3] array so that one batched separable call smooths all three;
//  * harris_nms: the response det - k*tr^2 of a pixel and of its 8
//    neighbours (recomputed from the smoothed products, every one of them
//    inside the band where the reference computes responses), flagged when it
//    is above all 8 and at least the minimum response; the flagged pixels'
//    responses are kept as float, as the reference stores them;
//  * the flagged pixels are compacted in memory order, which is the
//    reference's scan order (dim1 outer, dim0 inner), by a stable device
//    select, and, when the strongest max_corners are asked for, ordered by a
//    stable descending radix sort (std::stable_sort in the reference's
//    sort_index);
//  * harris_out: the features' x (dim1 index), y (dim0 index), score,
//    orientation 0 and size 1.
// Every product and sum is rounded separately (no contraction), as the
// reference's CPU build evaluates them, so flags and scores are its own.
#include <hipcub/hipcub.hpp>

#include "kcommon.hpp"
#include "../../../include/afhip_kernels.h"

namespace afhip {

template <typename T>
__global__ __launch_bounds__(256) void harris_deriv(T *__restrict__ d, const T *__restrict__ in, int W, int H) {
#pragma clang fp contract(off)
    const long long n = (long long)W * H;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const int i0 = (int)(e % W), i1 = (int)(e / W);
        const T *p = in + e;
        T g0, g1;
        if (i0 == 0)
            g0 = p[1] - p[0];
        else if (i0 == W - 1)
            g0 = p[0] - p[-1];
        else
            g0 = T(0.5) * (p[1] - p[-1]);
        if (i1 == 0)
            g1 = p[W] - p[0];
        else if (i1 == H - 1)
            g1 = p[0] - p[-W];
        else
            g1 = T(0.5) * (p[W] - p[-W]);
        const T iy = g0, ix = g1;
        d[e] = ix * ix;
        d[n + e] = ix * iy;
        d[2 * n + e] = iy * iy;
    }
}

template <typename T>
__device__ inline T harris_response(const T *__restrict__ s, long long n, long long i, T k) {
#pragma clang fp contract(off)
    const T a = s[i], b = s[n + i], c = s[2 * n + i];  // ixx, ixy, iyy smoothed
    const T tr = a + c;
    const T det = a * c - b * b;
    return det - k * (tr * tr);
}

template <typename T>
__global__ __launch_bounds__(256) void harris_nms(unsigned char *__restrict__ flags, float *__restrict__ resp,
                                                  const T *__restrict__ s, int W, int H, float k_thr,
                                                  float min_r, int r1) {
    const long long n = (long long)W * H;
    const T k = (T)k_thr, mr = (T)min_r;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const int i0 = (int)(e % W), i1 = (int)(e / W);
        unsigned char f = 0;
        if (i1 >= r1 && i1 < H - r1 && i0 >= r1 && i0 < W - r1) {
            const T v = harris_response(s, n, e, k);
            T m = harris_response(s, n, e - W - 1, k);
            m = max(m, harris_response(s, n, e - 1, k));
            m = max(m, harris_response(s, n, e + W - 1, k));
            m = max(m, harris_response(s, n, e - W, k));
            m = max(m, harris_response(s, n, e + W, k));
            m = max(m, harris_response(s, n, e - W + 1, k));
            m = max(m, harris_response(s, n, e + 1, k));
            m = max(m, harris_response(s, n, e + W + 1, k));
            if (v > m && v >= mr) {
                f = 1;
                resp[e] = (float)v;
            }
        }
        flags[e] = f;
    }
}

__global__ __launch_bounds__(256) void harris_keys(float *__restrict__ keys, unsigned *__restrict__ vals,
                                                   const unsigned *__restrict__ sel, const float *__restrict__ resp,
                                                   long long m) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < m;
         i += (long long)gridDim.x * blockDim.x) {
        keys[i] = resp[sel[i]];
        vals[i] = (unsigned)i;
    }
}

__global__ __launch_bounds__(256) void harris_out(float *__restrict__ x, float *__restrict__ y,
                                                  float *__restrict__ score, float *__restrict__ orient,
                                                  float *__restrict__ size, const unsigned *__restrict__ sel,
                                                  const unsigned *__restrict__ order,
                                                  const float *__restrict__ sorted_keys,
                                                  const float *__restrict__ resp, int W, long long n) {
    for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < n;
         f += (long long)gridDim.x * blockDim.x) {
        const unsigned i = order ? order[f] : (unsigned)f;
        const unsigned idx = sel[i];
        x[f] = (float)(idx / (unsigned)W);
        y[f] = (float)(idx % (unsigned)W);
        score[f] = sorted_keys ? sorted_keys[f] : resp[idx];
        orient[f] = 0.f;
        size[f] = 1.f;
    }
}

static unsigned grid_of(long long n) {
    long long b = (n + 255) / 256;
    if (b > kMaxGridBlocks) b = kMaxGridBlocks;
    return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace afhip

using namespace afhip;

extern "C" int afhip_k_harris_deriv(int dtype, void *d3, const void *in, long long W, long long H, void *stream) {
    if (W < 2 || H < 2 || W * H >= (1LL << 31)) return K_ERR_SIZE;
    hipStream_t s = (hipStream_t)stream;
    const unsigned g = grid_of(W * H);
    if (dtype == DT_F32)
        harris_deriv<float><<<g, 256, 0, s>>>((float *)d3, (const float *)in, (int)W, (int)H);
    else if (dtype == DT_F64)
        harris_deriv<double><<<g, 256, 0, s>>>((double *)d3, (const double *)in, (int)W, (int)H);
    else
        return K_ERR_TYPE;
    return launch_status();
}

extern "C" int afhip_k_harris_nms(int dtype, unsigned char *flags, float *resp, const void *s3, long long W,
                                  long long H, float k_thr, float min_r, int border, void *stream) {
    if (W < 1 || H < 1 || W * H >= (1LL << 31)) return K_ERR_SIZE;
    hipStream_t s = (hipStream_t)stream;
    const unsigned g = grid_of(W * H);
    if (dtype == DT_F32)
        harris_nms<float><<<g, 256, 0, s>>>(flags, resp, (const float *)s3, (int)W, (int)H, k_thr, min_r, border);
    else if (dtype == DT_F64)
        harris_nms<double><<<g, 256, 0, s>>>(flags, resp, (const double *)s3, (int)W, (int)H, k_thr, min_r, border);
    else
        return K_ERR_TYPE;
    return launch_status();
}

extern "C" long long afhip_k_select_flagged_workspace(long long n) {
    size_t bytes = 0;
    if (hipcub::DeviceSelect::Flagged(nullptr, bytes, hipcub::CountingInputIterator<unsigned>(0),
                                      (const unsigned char *)nullptr, (unsigned *)nullptr, (unsigned *)nullptr,
                                      (int64_t)n, (hipStream_t)0) != hipSuccess)
        return -1;
    return (long long)bytes;
}

extern "C" int afhip_k_select_flagged(const unsigned char *flags, long long n, unsigned *out_idx, unsigned *d_num,
                                      void *temp, long long temp_bytes, void *stream) {
    size_t bytes = (size_t)temp_bytes;
    if (hipcub::DeviceSelect::Flagged(temp, bytes, hipcub::CountingInputIterator<unsigned>(0), flags, out_idx,
                                      d_num, (int64_t)n, (hipStream_t)stream) != hipSuccess)
        return K_ERR_RUNTIME;
    return launch_status();
}

extern "C" long long afhip_k_sort_desc_workspace(long long m) {
    size_t bytes = 0;
    if (hipcub::DeviceRadixSort::SortPairsDescending(nullptr, bytes, (const float *)nullptr, (float *)nullptr,
                                                     (const unsigned *)nullptr, (unsigned *)nullptr, (int)m, 0, 32,
                                                     (hipStream_t)0) != hipSuccess)
        return -1;
    return (long long)bytes;
}

extern "C" int afhip_k_harris_sort(const unsigned *sel, const float *resp, long long m, float *keys,
                                   unsigned *vals, float *keys_sorted, unsigned *vals_sorted, void *temp,
                                   long long temp_bytes, void *stream) {
    if (m < 1 || m >= (1LL << 31)) return K_ERR_SIZE;
    hipStream_t s = (hipStream_t)stream;
    harris_keys<<<grid_of(m), 256, 0, s>>>(keys, vals, sel, resp, m);
    size_t bytes = (size_t)temp_bytes;
    if (hipcub::DeviceRadixSort::SortPairsDescending(temp, bytes, keys, keys_sorted, vals, vals_sorted, (int)m, 0,
                                                     32, s) != hipSuccess)
        return K_ERR_RUNTIME;
    return launch_status();
}

extern "C" int afhip_k_harris_out(float *x, float *y, float *score, float *orient, float *size,
                                  const unsigned *sel, const unsigned *order, const float *sorted_keys,
                                  const float *resp, long long W, long long n, void *stream) {
    if (n <= 0) return K_OK;
    harris_out<<<grid_of(n), 256, 0, (hipStream_t)stream>>>(x, y, score, orient, size, sel, order, sorted_keys,
                                                            resp, (int)W, n);
    return launch_status();
}
