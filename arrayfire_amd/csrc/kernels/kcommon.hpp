// Shared definitions for the gfx950 kernels: element types, conversions,
// launch helpers. Kernel sources include this file only.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace afhip {

typedef long long dim_t;

// Return codes of the kernel-level C ABI (values of af_err).
enum KStatus : int {
    K_OK = 0,
    K_ERR_RUNTIME = 103,
    K_ERR_ARG = 202,
    K_ERR_SIZE = 203,
    K_ERR_TYPE = 204,
    K_ERR_NOT_SUPPORTED = 301,
};

// af_dtype values (include/af/defines.h) plus the bfloat16 storage type.
enum DType : int {
    DT_F32 = 0, DT_C32 = 1, DT_F64 = 2, DT_C64 = 3, DT_B8 = 4, DT_S32 = 5,
    DT_U32 = 6, DT_U8 = 7, DT_S64 = 8, DT_U64 = 9, DT_S16 = 10, DT_U16 = 11,
    DT_F16 = 12, DT_S8 = 13, DT_BF16 = 64,
};

// Complex element types with the arithmetic the spatial kernels need.
template <typename R>
struct cplx {
    R re, im;
    __host__ __device__ cplx() : re(0), im(0) {}
    __host__ __device__ cplx(R r) : re(r), im(0) {}
    __host__ __device__ cplx(R r, R i) : re(r), im(i) {}
    __host__ __device__ cplx &operator+=(const cplx &o) {
#pragma clang fp contract(off)
        re += o.re;
        im += o.im;
        return *this;
    }
};
template <typename R>
__host__ __device__ inline cplx<R> operator*(const cplx<R> &a, const cplx<R> &b) {
    // Each product rounded before the sum, as the reference CPU evaluates it.
#pragma clang fp contract(off)
    return cplx<R>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
typedef cplx<float> cfloat;
typedef cplx<double> cdouble;

// bfloat16 / binary16 conversions (round to nearest even; NaN stays NaN).
// On the device this is gfx950's v_cvt_pk_bf16_f32 (same rounding); the host
// form is the reference for it.
__host__ __device__ inline uint16_t f32_to_bf16_bits(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_bit_cast(uint16_t, (__bf16)f);
#endif
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16_bits_to_f32(uint16_t h) {
    return __builtin_bit_cast(float, (uint32_t)h << 16);
}

inline int launch_status() {
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? K_OK : K_ERR_RUNTIME;
}

// Number of workgroups that fill the chip several times over; grid-stride
// kernels cap their grid here.
constexpr int kMaxGridBlocks = 256 * 64;

// Compute units of the current device (256 on MI355X), cached per device.
static inline int cu_count_current() {
    static int cus[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (cus[dev] <= 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cus[dev] = n;
    }
    return cus[dev];
}

// Remaps a linear workgroup id so that each of the 8 XCDs (which receive
// workgroups round-robin) walks a contiguous range of tiles; neighbouring
// tiles then share halo lines in that XCD's L2.
__device__ inline long long xcd_remap(long long bid, long long nblk) {
    const long long per = nblk / 8, rem = nblk % 8;
    const long long xcd = bid % 8, k = bid / 8;
    return xcd < rem ? xcd * (per + 1) + k : rem * (per + 1) + (xcd - rem) * per + k;
}

// 16-byte LDS accesses at addresses known to be 16-byte aligned. The load is
// volatile so that it stays whole: a load narrowed to the elements used is
// split into 4-byte reads, which conflict in the LDS banks when lanes are 16
// bytes apart. p must point into LDS.
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) f32x4_t lds_f32x4_t;
__device__ __attribute__((always_inline)) inline float4 lds_load16(const float *p) {
    const f32x4_t t = *(const volatile lds_f32x4_t *)p;
    return make_float4(t.x, t.y, t.z, t.w);
}
__device__ __attribute__((always_inline)) inline void lds_store16(float *p, float4 v) {
    *(lds_f32x4_t *)p = f32x4_t{v.x, v.y, v.z, v.w};
}

// Output rows per workgroup of a strip kernel (a multiple of 16, at most
// 256): long segments amortise the halo rows each segment loads and
// computes again, short ones keep at least about 2048 workgroups (8 per CU)
// when the strips alone are too few.
inline int strip_rows(long long rows, long long strips) {
    const long long steps = (rows + 15) / 16;
    long long per = steps * strips / 2048;
    if (per < 1) per = 1;
    if (per > 16) per = 16;
    return (int)(16 * per);
}

// Loads an LH x LW halo tile of an FP32 image into LDS (row pitch LS) with
// zeros outside [0,S0) x [0,S1). The tile origin (ix0, iy0) may be negative
// or unaligned: rows are read as aligned float4s (when S0 % 4 == 0) that
// cover the tile's columns, NB float4 slots per lane in flight before any
// LDS store, then scattered into place. NQ (float4 slots per row) must be at
// least (LW + 6) / 4; a compile-time NQ keeps the slot decomposition cheap.
// t: the calling thread's index among the NT threads that take part. Always
// inlined: out of line, the LDS tile is written through generic pointers from
// a called function with its own stack.
template <int NT, int NQ, int NB = 8>
__device__ __attribute__((always_inline)) inline void load_halo_tile_f32(float *tile, const float *__restrict__ img, long long S0,
                                          long long S1, long long ix0, long long iy0, int LH, int LW,
                                          int LS, int t = (int)threadIdx.x) {
    const long long a0 = ix0 - (((ix0 % 4) + 4) % 4);  // floor to a multiple of 4
    const int off = (int)(ix0 - a0);                  // 0..3
    const int slots = LH * NQ;
    const bool vec = (S0 & 3) == 0;
    for (int s0 = 0; s0 < slots; s0 += NB * NT) {
        float4 v[NB];
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const int slot = s0 + u * NT + t;
            v[u] = make_float4(0.f, 0.f, 0.f, 0.f);
            const int r = slot / NQ, q = slot - (slot / NQ) * NQ;
            if (slot < slots && 4 * q - off < LW) {
                const long long gy = iy0 + r;
                const long long gx = a0 + 4 * q;
                if (gy >= 0 && gy < S1) {
                    const float *row = img + gy * S0;
                    if (vec && gx >= 0 && gx + 3 < S0) {
                        v[u] = *reinterpret_cast<const float4 *>(row + gx);
                    } else {
                        float t[4];
#pragma unroll
                        for (int e = 0; e < 4; ++e) t[e] = (gx + e >= 0 && gx + e < S0) ? row[gx + e] : 0.f;
                        v[u] = make_float4(t[0], t[1], t[2], t[3]);
                    }
                }
            }
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const int slot = s0 + u * NT + t;
            const int r = slot / NQ, q = slot - (slot / NQ) * NQ;
            const int c0 = 4 * q - off;
            if (slot < slots && c0 < LW) {
                const float t[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (c0 + e >= 0 && c0 + e < LW) tile[r * LS + c0 + e] = t[e];
            }
        }
    }
}

}  // namespace afhip
