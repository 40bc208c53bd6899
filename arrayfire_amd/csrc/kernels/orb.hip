// ORB features: af_orb (src/api/c/orb.cpp:58-89) with its backend
// (src/backend/cpu/orb.cpp:44-283), a caller of the separable convolution:
// each pyramid level is smoothed by convolve2<T, convAccT>(img, gauss,
// gauss, false) with a 9-tap sigma-2 Gaussian before the descriptors are
// sampled. The kernels here are the rest of one level:
//
//  * orb_resize: the bilinear pyramid step (resize_op<T, AF_INTERP_BILINEAR>,
//    src/backend/cpu/kernel/resize.hpp:86-130): source coordinates and
//    weights in float, the four weighted taps summed in T;
//  * orb_fast: FAST corners (src/backend/cpu/kernel/fast.hpp:51-157), one
//    lane per interior pixel: the 16 circle tests as two bit masks, a corner
//    when either mask holds a cyclic run of arc_length ones (the reference's
//    sliding window sums reach +-arc_length exactly then), the score the
//    larger of the bright and dark sums taken in circle order. Flags and
//    scores are written at the pixel's scan-order rank (dim0 index outer,
//    dim1 index inner), so a stable device select lists the corners in the
//    reference's detection order;
//  * orb_fast_scatter + orb_fast_nonmax: the score map of the first max_feat
//    corners and the 8-neighbour suppression (non_maximal,
//    src/backend/cpu/kernel/fast.hpp:159-214, with its edge test that compares
//    y with dims[1] and x with dims[0]);
//  * orb_harris: the Harris response of each surviving corner over a 7x7
//    block (harris_response<T, false>, src/backend/cpu/kernel/orb.hpp:110-162);
//  * orb_angle: the kept features (after the stable descending sort of the
//    responses, done with the Harris sort) and their intensity-centroid
//    angle (centroid_angle, src/backend/cpu/kernel/orb.hpp:164-193), the two
//    moments accumulated in T in patch order by one lane each;
//  * orb_describe: the 256-bit rotated-BRIEF descriptor
//    (extract_orb/get_pixel, src/backend/cpu/kernel/orb.hpp:195-264), a lane
//    per point pair, the words assembled by wave ballots, and the positions
//    scaled to the level's scale.
//
// Every product and sum is rounded separately (no contraction), as the
// reference's CPU build evaluates them. Angles are atan2 of the moments and
// sin/cos of the angle evaluated in double and rounded to float.
#include "kcommon.hpp"
#include "../../../include/afhip_kernels.h"

namespace afhip {

// {x1, y1, x2, y2} per descriptor bit.
__constant__ signed char kOrbPairs[256][4] = {
#include "orb_pattern.inc"
};

static constexpr int kOrbGridMax = 65535;

static unsigned orb_grid(long long n) {
    long long b = (n + 255) / 256;
    if (b > kOrbGridMax) b = kOrbGridMax;
    return (unsigned)(b < 1 ? 1 : b);
}

template <typename T>
__global__ __launch_bounds__(256) void orb_resize(T *__restrict__ out, const T *__restrict__ in, int id0, int id1,
                                                  int od0, int od1) {
#pragma clang fp contract(off)
    const long long n = (long long)od0 * od1;
    const float sx = (float)od0 / (float)id0, sy = (float)od1 / (float)id1;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const int xo = (int)(e % od0), yo = (int)(e / od0);
        const float fx = (float)xo / sx, fy = (float)yo / sy;
        long long i1x = (long long)floorf(fx), i1y = (long long)floorf(fy);
        if (i1x >= id0) i1x = id0 - 1;
        if (i1y >= id1) i1y = id1 - 1;
        const float b = fx - (float)i1x, a = fy - (float)i1y;
        const long long i2x = i1x + 1 >= id0 ? id0 - 1 : i1x + 1;
        const long long i2y = i1y + 1 >= id1 ? id1 - 1 : i1y + 1;
        const T p1 = in[i1y * id0 + i1x], p2 = in[i2y * id0 + i1x];
        const T p3 = in[i1y * id0 + i2x], p4 = in[i2y * id0 + i2x];
        const T w1 = (T)((1.0f - a) * (1.0f - b)), w2 = (T)(a * (1.0f - b));
        const T w3 = (T)((1.0f - a) * b), w4 = (T)(a * b);
        out[e] = w1 * p1 + w2 * p2 + w3 * p3 + w4 * p4;
    }
}

// FAST's circle of 16 pixels (idx_y / idx_x, src/backend/cpu/kernel/fast.hpp:20-31)
__host__ __device__ constexpr int fast_clamp3(int v) { return v < -3 ? -3 : (v > 3 ? 3 : v); }
__host__ __device__ constexpr int fast_dy(int i) { return i >= 8 ? fast_clamp3(-(i - 8 - 4)) : fast_clamp3(i - 4); }
__host__ __device__ constexpr int fast_dx(int i) { return i < 12 ? fast_dy(i + 4) : fast_dy(i - 12); }

// True when the 16-bit cyclic mask m holds a run of `arc` consecutive ones.
__device__ inline bool cyclic_run(unsigned m, int arc) {
    const unsigned m2 = m | (m << 16);
    unsigned r = m2;
    for (int t = 1; t < arc; ++t) r &= m2 >> t;
    return (r & 0xffffu) != 0u;
}

template <typename T>
__global__ __launch_bounds__(256) void orb_fast(unsigned char *__restrict__ flags, float *__restrict__ score,
                                                const T *__restrict__ img, int d0, int edge, int ny, int nx,
                                                float thr, int arc) {
#pragma clang fp contract(off)
    const long long n = (long long)ny * nx;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const int yy = (int)(e % ny), xx = (int)(e / ny);
        const T *c = img + (long long)(edge + xx) * d0 + (edge + yy);
        const float p = (float)c[0];
        unsigned bright = 0u, dark = 0u;
        float sb = 0.f, sd = 0.f;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const float px = (float)c[(long long)fast_dx(i) * d0 + fast_dy(i)];
            const int gt = px > p + thr;
            const int lt = px < p - thr;
            bright |= (unsigned)gt << i;
            dark |= (unsigned)lt << i;
            sb = sb + (float)gt * (fabsf(px - p) - thr);
            sd = sd + (float)lt * (fabsf(p - px) - thr);
        }
        const long long k = (long long)yy * nx + xx;
        flags[k] = (cyclic_run(bright, arc) || cyclic_run(dark, arc)) ? 1 : 0;
        score[k] = sb < sd ? sd : sb;
    }
}

__global__ __launch_bounds__(256) void orb_zero(float *__restrict__ v, long long n) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x)
        v[e] = 0.f;
}

// The corner count in device memory, capped at n (the launch's bound).
__device__ inline long long capped(const unsigned *count, long long n) {
    return count ? min(n, (long long)*count) : n;
}

__global__ __launch_bounds__(256) void orb_fast_scatter(float *__restrict__ V, const unsigned *__restrict__ sel,
                                                        const float *__restrict__ score, long long n,
                                                        const unsigned *__restrict__ count, int d0, int edge,
                                                        int nx) {
    n = capped(count, n);
    for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < n;
         j += (long long)gridDim.x * blockDim.x) {
        const unsigned k = sel[j];
        const long long y = edge + (long long)(k / (unsigned)nx), x = edge + (long long)(k % (unsigned)nx);
        V[x * d0 + y] = score[k];
    }
}

__global__ __launch_bounds__(256) void orb_fast_nonmax(unsigned char *__restrict__ flags,
                                                       const float *__restrict__ V,
                                                       const unsigned *__restrict__ sel, long long n,
                                                       const unsigned *__restrict__ count, int d0, int d1, int edge,
                                                       int nx) {
    const long long ne = capped(count, n);
    for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < n;
         j += (long long)gridDim.x * blockDim.x) {
        if (j >= ne) {
            flags[j] = 0;
            continue;
        }
        const unsigned k = sel[j];
        const long long y = edge + (long long)(k / (unsigned)nx), x = edge + (long long)(k % (unsigned)nx);
        const float *c = V + x * d0 + y;
        const float v = c[0];
        float m = fmaxf(c[-d0 - 1], c[-1]);
        m = fmaxf(m, c[d0 - 1]);
        m = fmaxf(m, c[-d0]);
        m = fmaxf(m, c[d0]);
        m = fmaxf(m, c[-d0 + 1]);
        m = fmaxf(m, c[1]);
        m = fmaxf(m, c[d0 + 1]);
        // the reference's edge test: y against dims[1], x against dims[0]
        const bool skip = y >= (long long)d1 - edge - 1 || y <= edge + 1 || x >= (long long)d0 - edge - 1 ||
                          x <= edge + 1;
        flags[j] = (!skip && v > m) ? 1 : 0;
    }
}

// BLOCK is the Harris window (7 in ORB); the loop is unrolled so that a
// lane's loads are all in flight before the in-order sums.
template <typename T, int BLOCK>
__global__ __launch_bounds__(256) void orb_harris(unsigned char *__restrict__ flags, float *__restrict__ resp,
                                                  const T *__restrict__ img, const unsigned *__restrict__ sel1,
                                                  const unsigned *__restrict__ sel2, long long n,
                                                  const unsigned *__restrict__ count, int d0, int d1, int edge,
                                                  int nx, int patch_r, float k_thr) {
#pragma clang fp contract(off)
    constexpr int block = BLOCK;
    const long long ne = capped(count, n);
    for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < n;
         j += (long long)gridDim.x * blockDim.x) {
        if (j >= ne) {
            flags[j] = 0;
            resp[j] = 0.f;
            continue;
        }
        const unsigned k = sel1[sel2[j]];
        const long long y = edge + (long long)(k / (unsigned)nx), x = edge + (long long)(k % (unsigned)nx);
        if (x < patch_r || y < patch_r || x >= (long long)d1 - patch_r || y >= (long long)d0 - patch_r) {
            flags[j] = 0;
            resp[j] = 0.f;
            continue;
        }
        constexpr int r = block / 2;
        float ixx = 0.f, iyy = 0.f, ixy = 0.f;
#pragma unroll
        for (int q = 0; q < block * block; ++q) {
            const int i = q / block - r, jj = q % block - r;
            const T *c = img + (x + i) * d0 + (y + jj);
            const float ix = (float)(c[d0] - c[-d0]);
            const float iy = (float)(c[1] - c[-1]);
            ixx = ixx + ix * ix;
            iyy = iyy + iy * iy;
            ixy = ixy + ix * iy;
        }
        const float tr = ixx + iyy;
        const float det = ixx * iyy - ixy * ixy;
        const float rs0 = 0.001f;
        const float rs = rs0 * rs0 * rs0 * rs0;
        resp[j] = (det - k_thr * (tr * tr)) * rs;
        flags[j] = 1;
    }
}

// PATCH is ORB's 31; each patch row's PATCH values are loaded at once, then
// added in the reference's order (row by row, column by column).
template <typename T, int PATCH>
__global__ __launch_bounds__(256) void orb_angle(float *__restrict__ ox, float *__restrict__ oy,
                                                 float *__restrict__ oscore, float *__restrict__ oori,
                                                 const T *__restrict__ img, const unsigned *__restrict__ sel1,
                                                 const unsigned *__restrict__ sel2,
                                                 const unsigned *__restrict__ sel3,
                                                 const unsigned *__restrict__ order,
                                                 const float *__restrict__ keys, long long n, int d0, int d1,
                                                 int edge, int nx) {
#pragma clang fp contract(off)
    constexpr int patch = PATCH;
    for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < n;
         f += (long long)gridDim.x * blockDim.x) {
        const unsigned k = sel1[sel2[sel3[order[f]]]];
        const long long y = edge + (long long)(k / (unsigned)nx), x = edge + (long long)(k % (unsigned)nx);
        ox[f] = (float)x;
        oy[f] = (float)y;
        oscore[f] = keys[f];
        // The reference skips x > dims[1] - r; x == dims[1] - r would read one
        // column past the image, so it is skipped here too (features reach
        // this kernel only from the Harris step, which keeps them 22 pixels
        // inside, so neither test ever fires for them).
        constexpr int r = patch / 2;
        if (x < r || y < r || x + r >= (long long)d1 || y + r >= (long long)d0) {
            oori[f] = 0.f;
            continue;
        }
        T m01 = T(0), m10 = T(0);
        for (int ii = 0; ii < patch; ++ii) {
            const T *col = img + (x + ii - r) * d0 + (y - r);
            T v[patch];
#pragma unroll
            for (int jj = 0; jj < patch; ++jj) v[jj] = col[jj];
#pragma unroll
            for (int jj = 0; jj < patch; ++jj) {
                m01 = m01 + (T)(jj - r) * v[jj];
                m10 = m10 + (T)(ii - r) * v[jj];
            }
        }
        oori[f] = (float)atan2((double)m01, (double)m10);
    }
}

// One workgroup of 256 lanes per feature, one lane per point pair (bit b of
// the descriptor compares pair b); each wave's 64 comparisons become two
// 32-bit words by a ballot.
template <typename T>
__global__ __launch_bounds__(256) void orb_describe(unsigned *__restrict__ desc, float *__restrict__ ox,
                                                    float *__restrict__ oy, float *__restrict__ osize,
                                                    const float *__restrict__ fx, const float *__restrict__ fy,
                                                    const float *__restrict__ fori, const T *__restrict__ img,
                                                    long long n, int d0, int d1, int patch, int r, float scl) {
#pragma clang fp contract(off)
    const long long f = blockIdx.x;
    if (f >= n) return;  // uniform per workgroup
    const int b = threadIdx.x, lane = b & 63, wv = b >> 6;
    const long long x = (long long)roundf(fx[f]), y = (long long)roundf(fy[f]);
    if (x < r || y < r || x >= (long long)d1 - r || y >= (long long)d0 - r) {
        if (lane == 0) {
            desc[f * 8 + 2 * wv] = 0u;
            desc[f * 8 + 2 * wv + 1] = 0u;
        }
        if (b == 0) {
            ox[f] = fx[f];
            oy[f] = fy[f];
            osize[f] = 0.f;
        }
        return;
    }
    const float ori = fori[f];
    const float sn = (float)sin((double)ori), cs = (float)cos((double)ori);
    const float pscl = (float)patch / (float)patch;
    const signed char *pp = kOrbPairs[b];
    const float ax = (float)pp[0] * pscl, ay = (float)pp[1] * pscl;
    const float bx = (float)pp[2] * pscl, by = (float)pp[3] * pscl;
    const long long x1 = x + (long long)roundf(ax * cs - ay * sn);
    const long long y1 = y + (long long)roundf(ax * sn + ay * cs);
    const long long x2 = x + (long long)roundf(bx * cs - by * sn);
    const long long y2 = y + (long long)roundf(bx * sn + by * cs);
    const T p1 = img[x1 * d0 + y1], p2 = img[x2 * d0 + y2];
    const unsigned long long m = __ballot(p1 < p2);
    if (lane == 0) {
        desc[f * 8 + 2 * wv] = (unsigned)(m & 0xffffffffull);
        desc[f * 8 + 2 * wv + 1] = (unsigned)(m >> 32);
    }
    if (b == 0) {
        ox[f] = roundf((float)x * scl);
        oy[f] = roundf((float)y * scl);
        osize[f] = (float)patch * scl;
    }
}

}  // namespace afhip

using namespace afhip;

extern "C" int afhip_k_orb_resize(int dtype, void *out, const void *in, long long id0, long long id1, long long od0,
                                  long long od1, void *stream) {
    if (id0 < 1 || id1 < 1 || od0 < 1 || od1 < 1 || id0 * id1 >= (1LL << 31) || od0 * od1 >= (1LL << 31))
        return K_ERR_SIZE;
    hipStream_t s = (hipStream_t)stream;
    const unsigned g = orb_grid(od0 * od1);
    if (dtype == DT_F32)
        orb_resize<float><<<g, 256, 0, s>>>((float *)out, (const float *)in, (int)id0, (int)id1, (int)od0, (int)od1);
    else if (dtype == DT_F64)
        orb_resize<double><<<g, 256, 0, s>>>((double *)out, (const double *)in, (int)id0, (int)id1, (int)od0,
                                             (int)od1);
    else
        return K_ERR_TYPE;
    return launch_status();
}

extern "C" int afhip_k_orb_fast(int dtype, unsigned char *flags, float *score, const void *img, long long d0,
                                long long d1, int edge, float thr, int arc, void *stream) {
    if (arc < 9 || arc > 16 || edge < 3) return K_ERR_ARG;
    const long long ny = d0 - 2LL * edge, nx = d1 - 2LL * edge;
    if (ny < 1 || nx < 1 || d0 * d1 >= (1LL << 31)) return K_ERR_SIZE;
    hipStream_t s = (hipStream_t)stream;
    const unsigned g = orb_grid(ny * nx);
    if (dtype == DT_F32)
        orb_fast<float><<<g, 256, 0, s>>>(flags, score, (const float *)img, (int)d0, edge, (int)ny, (int)nx, thr,
                                          arc);
    else if (dtype == DT_F64)
        orb_fast<double><<<g, 256, 0, s>>>(flags, score, (const double *)img, (int)d0, edge, (int)ny, (int)nx, thr,
                                           arc);
    else
        return K_ERR_TYPE;
    return launch_status();
}

extern "C" int afhip_k_orb_fast_nonmax(unsigned char *flags, float *V, const unsigned *sel, const float *score,
                                       long long n, const unsigned *count, long long d0, long long d1, int edge,
                                       void *stream) {
    const long long nx = d1 - 2LL * edge;
    if (n < 1 || nx < 1 || edge < 1 || d0 * d1 >= (1LL << 31)) return K_ERR_SIZE;
    hipStream_t s = (hipStream_t)stream;
    orb_zero<<<orb_grid(d0 * d1), 256, 0, s>>>(V, d0 * d1);
    orb_fast_scatter<<<orb_grid(n), 256, 0, s>>>(V, sel, score, n, count, (int)d0, edge, (int)nx);
    orb_fast_nonmax<<<orb_grid(n), 256, 0, s>>>(flags, V, sel, n, count, (int)d0, (int)d1, edge, (int)nx);
    return launch_status();
}

extern "C" int afhip_k_orb_harris(int dtype, unsigned char *flags, float *resp, const void *img,
                                  const unsigned *sel1, const unsigned *sel2, long long n, const unsigned *count,
                                  long long d0, long long d1, int edge, int patch_r, int block, float k_thr,
                                  void *stream) {
    const long long nx = d1 - 2LL * edge;
    if (n < 1 || nx < 1 || d0 * d1 >= (1LL << 31)) return K_ERR_SIZE;
    if (block != 7 || patch_r < block / 2 + 1) return K_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    const unsigned g = orb_grid(n);
    if (dtype == DT_F32)
        orb_harris<float, 7><<<g, 256, 0, s>>>(flags, resp, (const float *)img, sel1, sel2, n, count, (int)d0,
                                               (int)d1, edge, (int)nx, patch_r, k_thr);
    else if (dtype == DT_F64)
        orb_harris<double, 7><<<g, 256, 0, s>>>(flags, resp, (const double *)img, sel1, sel2, n, count, (int)d0,
                                                (int)d1, edge, (int)nx, patch_r, k_thr);
    else
        return K_ERR_TYPE;
    return launch_status();
}

extern "C" int afhip_k_orb_angle(int dtype, float *x, float *y, float *score, float *ori, const void *img,
                                 const unsigned *sel1, const unsigned *sel2, const unsigned *sel3,
                                 const unsigned *order, const float *keys, long long n, long long d0, long long d1,
                                 int edge, int patch, void *stream) {
    const long long nx = d1 - 2LL * edge;
    if (n < 1 || nx < 1 || d0 * d1 >= (1LL << 31)) return K_ERR_SIZE;
    if (patch != 31 || patch / 2 > edge) return K_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    const unsigned g = orb_grid(n);
    if (dtype == DT_F32)
        orb_angle<float, 31><<<g, 256, 0, s>>>(x, y, score, ori, (const float *)img, sel1, sel2, sel3, order, keys,
                                               n, (int)d0, (int)d1, edge, (int)nx);
    else if (dtype == DT_F64)
        orb_angle<double, 31><<<g, 256, 0, s>>>(x, y, score, ori, (const double *)img, sel1, sel2, sel3, order,
                                                keys, n, (int)d0, (int)d1, edge, (int)nx);
    else
        return K_ERR_TYPE;
    return launch_status();
}

extern "C" int afhip_k_orb_describe(int dtype, unsigned *desc, float *ox, float *oy, float *osize, const float *x,
                                    const float *y, const float *ori, const void *img, long long n, long long d0,
                                    long long d1, int patch, int r, float scl, void *stream) {
    if (n < 1 || d0 * d1 >= (1LL << 31)) return K_ERR_SIZE;
    // every sample stays inside the image: |offset| <= 13*sqrt(2) < 19 < r
    if (patch != 31 || r < 19) return K_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    if (n >= (1LL << 31)) return K_ERR_SIZE;
    const unsigned g = (unsigned)n;  // a workgroup per feature, a lane per point pair
    if (dtype == DT_F32)
        orb_describe<float><<<g, 256, 0, s>>>(desc, ox, oy, osize, x, y, ori, (const float *)img, n, (int)d0,
                                              (int)d1, patch, r, scl);
    else if (dtype == DT_F64)
        orb_describe<double><<<g, 256, 0, s>>>(desc, ox, oy, osize, x, y, ori, (const double *)img, n, (int)d0,
                                               (int)d1, patch, r, scl);
    else
        return K_ERR_TYPE;
    return launch_status();
}
