// af_unwrap / af_wrap: im2col and col2im as public functions
// (src/api/c/unwrap.cpp:37-99, src/api/c/wrap.cpp:26-140;
// src/backend/cpu/kernel/unwrap.hpp:20-80, kernel/wrap.hpp:20-80).
//
// unwrap: out [wx*wy, nx*ny, C, N] (is_column) or its dim0/dim1 swap; patch
// col = winy * nx + winx starts at (winx*sx - px, winy*sy - py); element
// (x, y) of a patch is in[xpad = start_x + x*dx, ypad = start_y + y*dy], zero
// outside the image. Pure data movement, so each element type is moved as
// its byte width.
//
// wrap: the adjoint. The reference adds every patch element into out in
// patch order (col ascending; within a col each output pixel is hit once).
// Here each output pixel gathers its own contributions, walking the patches
// that cover it in ascending col order and adding in the element type, so
// the sum order, and with it every floating-point rounding, is the
// reference's. The gather needs no atomics and writes each output once.
// Integer types add modulo 2^bits (two's complement: signed and unsigned
// addition give the same bits), which is what C++'s `+=` on the reference's
// types stores.
#include "kcommon.hpp"
#include "../../../include/afhip_kernels.h"

namespace afhip {

struct b16 {
    uint64_t a, b;
};

struct WrapGeo {
    long long i[4];  // input dims
    long long o[4];  // output dims
    long long wx, wy, sx, sy, px, py, dx, dy, nx, ny;
    int is_column;
};

template <typename E>
__global__ __launch_bounds__(256) void unwrap_k(E *__restrict__ out, const E *__restrict__ in, WrapGeo g) {
    const long long total = g.o[0] * g.o[1] * g.o[2] * g.o[3];
    const long long plane_o = g.o[0] * g.o[1], plane_i = g.i[0] * g.i[1];
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long zw = e / plane_o;
        const long long r = e - zw * plane_o;
        const long long i0 = r % g.o[0], i1 = r / g.o[0];
        const long long pos = g.is_column ? i0 : i1;
        const long long col = g.is_column ? i1 : i0;
        const long long winy = col / g.nx, winx = col - winy * g.nx;
        const long long y = pos / g.wx, x = pos - y * g.wx;
        const long long xp = winx * g.sx - g.px + x * g.dx;
        const long long yp = winy * g.sy - g.py + y * g.dy;
        E v{};
        if (xp >= 0 && xp < g.i[0] && yp >= 0 && yp < g.i[1]) v = in[zw * plane_i + yp * g.i[0] + xp];
        out[e] = v;
    }
}

template <typename T>
struct WrapAdd {
    __device__ static T add(T a, T b) {
#pragma clang fp contract(off)
        return a + b;
    }
};
template <>
struct WrapAdd<cfloat> {
    __device__ static cfloat add(cfloat a, cfloat b) {
        a += b;
        return a;
    }
};
template <>
struct WrapAdd<cdouble> {
    __device__ static cdouble add(cdouble a, cdouble b) {
        a += b;
        return a;
    }
};

__device__ inline long long floor_div(long long a, long long b) {
    return a >= 0 ? a / b : -((-a + b - 1) / b);
}

// out pixel (xo, yo) <- (accumulate ? out : 0) + sum over covering patches in col order
template <typename T>
__global__ __launch_bounds__(256) void wrap_k(T *__restrict__ out, const T *__restrict__ in, WrapGeo g,
                                              int accumulate) {
    const long long total = g.o[0] * g.o[1] * g.o[2] * g.o[3];
    const long long plane_o = g.o[0] * g.o[1], plane_i = g.i[0] * g.i[1];
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long zw = e / plane_o;
        const long long r = e - zw * plane_o;
        const long long xo = r % g.o[0], yo = r / g.o[0];
        const long long ax = xo + g.px, ay = yo + g.py;  // position in the padded frame
        // patches with win*s <= a <= win*s + w - 1, clamped to the patch grid
        // (a can be negative when the padding is)
        long long wx0 = ax - g.wx + 1 > 0 ? (ax - g.wx + 1 + g.sx - 1) / g.sx : 0;
        long long wx1 = floor_div(ax, g.sx);
        if (wx1 > g.nx - 1) wx1 = g.nx - 1;
        long long wy0 = ay - g.wy + 1 > 0 ? (ay - g.wy + 1 + g.sy - 1) / g.sy : 0;
        long long wy1 = floor_div(ay, g.sy);
        if (wy1 > g.ny - 1) wy1 = g.ny - 1;
        T acc = accumulate ? out[e] : T{};
        const T *src = in + zw * plane_i;
        for (long long wy = wy0; wy <= wy1; ++wy) {
            const long long y = ay - wy * g.sy;
            for (long long wx = wx0; wx <= wx1; ++wx) {
                const long long x = ax - wx * g.sx;
                const long long col = wy * g.nx + wx, loc = y * g.wx + x;
                const long long idx = g.is_column ? loc + col * g.i[0] : col + loc * g.i[0];
                acc = WrapAdd<T>::add(acc, src[idx]);
            }
        }
        out[e] = acc;
    }
}

static unsigned grid_for(long long n) {
    long long b = (n + 255) / 256;
    if (b > kMaxGridBlocks) b = kMaxGridBlocks;
    return (unsigned)(b < 1 ? 1 : b);
}

static int esize_of(int dtype) {
    switch (dtype) {
        case DT_F32: case DT_S32: case DT_U32: return 4;
        case DT_F64: case DT_S64: case DT_U64: case DT_C32: return 8;
        case DT_C64: return 16;
        case DT_S16: case DT_U16: case DT_F16: case DT_BF16: return 2;
        case DT_U8: case DT_S8: case DT_B8: return 1;
        default: return 0;
    }
}

}  // namespace afhip

using namespace afhip;

extern "C" int afhip_k_unwrap(int dtype, void *out, const long long *odims, const void *in, const long long *idims,
                              long long wx, long long wy, long long sx, long long sy, long long px, long long py,
                              long long dx, long long dy, int is_column, void *stream) {
    if (wx <= 0 || wy <= 0 || sx <= 0 || sy <= 0 || dx <= 0 || dy <= 0 || px < 0 || py < 0) return K_ERR_ARG;
    WrapGeo g{};
    for (int k = 0; k < 4; ++k) {
        g.i[k] = idims[k];
        g.o[k] = odims[k];
    }
    g.wx = wx; g.wy = wy; g.sx = sx; g.sy = sy; g.px = px; g.py = py; g.dx = dx; g.dy = dy;
    g.nx = 1 + (idims[0] + 2 * px - ((wx - 1) * dx + 1)) / sx;
    g.ny = 1 + (idims[1] + 2 * py - ((wy - 1) * dy + 1)) / sy;
    g.is_column = is_column ? 1 : 0;
    const long long pd0 = is_column ? wx * wy : g.nx * g.ny, pd1 = is_column ? g.nx * g.ny : wx * wy;
    if (odims[0] != pd0 || odims[1] != pd1 || odims[2] != idims[2] || odims[3] != idims[3]) return K_ERR_SIZE;
    const long long n = odims[0] * odims[1] * odims[2] * odims[3];
    if (n == 0) return K_OK;
    hipStream_t s = (hipStream_t)stream;
    switch (esize_of(dtype)) {
        case 1: unwrap_k<uint8_t><<<grid_for(n), 256, 0, s>>>((uint8_t *)out, (const uint8_t *)in, g); break;
        case 2: unwrap_k<uint16_t><<<grid_for(n), 256, 0, s>>>((uint16_t *)out, (const uint16_t *)in, g); break;
        case 4: unwrap_k<uint32_t><<<grid_for(n), 256, 0, s>>>((uint32_t *)out, (const uint32_t *)in, g); break;
        case 8: unwrap_k<uint64_t><<<grid_for(n), 256, 0, s>>>((uint64_t *)out, (const uint64_t *)in, g); break;
        case 16: unwrap_k<b16><<<grid_for(n), 256, 0, s>>>((b16 *)out, (const b16 *)in, g); break;
        default: return K_ERR_TYPE;
    }
    return launch_status();
}

extern "C" int afhip_k_wrap(int dtype, void *out, const long long *odims, const void *in, const long long *idims,
                            long long wx, long long wy, long long sx, long long sy, long long px, long long py,
                            int is_column, int accumulate, void *stream) {
    if (wx <= 0 || wy <= 0 || sx <= 0 || sy <= 0) return K_ERR_ARG;
    WrapGeo g{};
    for (int k = 0; k < 4; ++k) {
        g.i[k] = idims[k];
        g.o[k] = odims[k];
    }
    g.wx = wx; g.wy = wy; g.sx = sx; g.sy = sy; g.px = px; g.py = py; g.dx = 1; g.dy = 1;
    g.nx = (odims[0] + 2 * px - wx) / sx + 1;
    g.ny = (odims[1] + 2 * py - wy) / sy + 1;
    g.is_column = is_column ? 1 : 0;
    const long long psz = is_column ? idims[0] : idims[1], np = is_column ? idims[1] : idims[0];
    if (psz != wx * wy || np != g.nx * g.ny || odims[2] != idims[2] || odims[3] != idims[3]) return K_ERR_SIZE;
    const long long n = odims[0] * odims[1] * odims[2] * odims[3];
    if (n == 0) return K_OK;
    hipStream_t s = (hipStream_t)stream;
    const unsigned gb = grid_for(n);
    switch (dtype) {
        case DT_F32: wrap_k<float><<<gb, 256, 0, s>>>((float *)out, (const float *)in, g, accumulate); break;
        case DT_F64: wrap_k<double><<<gb, 256, 0, s>>>((double *)out, (const double *)in, g, accumulate); break;
        case DT_C32: wrap_k<cfloat><<<gb, 256, 0, s>>>((cfloat *)out, (const cfloat *)in, g, accumulate); break;
        case DT_C64: wrap_k<cdouble><<<gb, 256, 0, s>>>((cdouble *)out, (const cdouble *)in, g, accumulate); break;
        case DT_S32: case DT_U32:
            wrap_k<uint32_t><<<gb, 256, 0, s>>>((uint32_t *)out, (const uint32_t *)in, g, accumulate); break;
        case DT_S64: case DT_U64:
            wrap_k<uint64_t><<<gb, 256, 0, s>>>((uint64_t *)out, (const uint64_t *)in, g, accumulate); break;
        case DT_S16: case DT_U16:
            wrap_k<uint16_t><<<gb, 256, 0, s>>>((uint16_t *)out, (const uint16_t *)in, g, accumulate); break;
        case DT_S8: case DT_U8: case DT_B8:
            wrap_k<uint8_t><<<gb, 256, 0, s>>>((uint8_t *)out, (const uint8_t *)in, g, accumulate); break;
        default: return K_ERR_TYPE;
    }
    return launch_status();
}
