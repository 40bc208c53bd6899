// Elementwise helpers used around the convolution path: dtype conversion
// (castArray / af_cast), constant fill (af_constant, zeroing), and the
// scalar-filter scaling of the separable convolution.
#include "kcommon.hpp"
#include "../../../include/afhip_kernels.h"

namespace afhip {

struct bf16_t {
    uint16_t bits;
};

// Every storage type goes through these two hooks: to a wide value and back.
template <typename T>
struct Conv {
    typedef double wide;
    __device__ static double get(T v) { return (double)v; }
    __device__ static T put(double v) { return (T)v; }
};
template <>
struct Conv<float> {
    typedef float wide;
    __device__ static float get(float v) { return v; }
    __device__ static float put(double v) { return (float)v; }
};
template <>
struct Conv<_Float16> {
    __device__ static double get(_Float16 v) { return (double)(float)v; }
    __device__ static _Float16 put(double v) { return (_Float16)v; }
};
template <>
struct Conv<bf16_t> {
    __device__ static double get(bf16_t v) { return (double)bf16_bits_to_f32(v.bits); }
    // Rounds the double to float first, then to bfloat16: double rounding is
    // avoided for every value representable in float, which is every value
    // the library feeds here.
    __device__ static bf16_t put(double v) { return bf16_t{f32_to_bf16_bits((float)v)}; }
};
// Integer conversions go through long long / unsigned long long so that
// 64-bit values keep every bit.
template <>
struct Conv<long long> {
    __device__ static double get(long long v) { return (double)v; }
    __device__ static long long put(double v) { return (long long)v; }
};

template <typename O, typename I>
__device__ inline O convert(I v) {
    return Conv<O>::put(Conv<I>::get(v));
}
template <>
__device__ inline float convert<float, float>(float v) { return v; }
template <>
__device__ inline _Float16 convert<_Float16, float>(float v) { return (_Float16)v; }
template <>
__device__ inline bf16_t convert<bf16_t, float>(float v) { return bf16_t{f32_to_bf16_bits(v)}; }
template <>
__device__ inline float convert<float, bf16_t>(bf16_t v) { return bf16_bits_to_f32(v.bits); }
template <>
__device__ inline float convert<float, _Float16>(_Float16 v) { return (float)v; }
template <>
__device__ inline long long convert<long long, long long>(long long v) { return v; }
template <>
__device__ inline long long convert<long long, unsigned long long>(unsigned long long v) {
    return (long long)v;
}
template <>
__device__ inline unsigned long long convert<unsigned long long, long long>(long long v) {
    return (unsigned long long)v;
}

template <typename O, typename I>
__global__ __launch_bounds__(256) void cast_kernel(O *__restrict__ out,
                                                   const I *__restrict__ in,
                                                   dim_t n) {
    for (dim_t i = blockIdx.x * (dim_t)blockDim.x + threadIdx.x; i < n;
         i += (dim_t)gridDim.x * blockDim.x)
        out[i] = convert<O, I>(in[i]);
}

template <typename R, typename I>
__global__ __launch_bounds__(256) void cast_to_complex(cplx<R> *__restrict__ out,
                                                       const I *__restrict__ in,
                                                       dim_t n) {
    for (dim_t i = blockIdx.x * (dim_t)blockDim.x + threadIdx.x; i < n;
         i += (dim_t)gridDim.x * blockDim.x)
        out[i] = cplx<R>((R)Conv<I>::get(in[i]), (R)0);
}

template <typename R, typename S>
__global__ __launch_bounds__(256) void cast_complex(cplx<R> *__restrict__ out,
                                                    const cplx<S> *__restrict__ in,
                                                    dim_t n) {
    for (dim_t i = blockIdx.x * (dim_t)blockDim.x + threadIdx.x; i < n;
         i += (dim_t)gridDim.x * blockDim.x)
        out[i] = cplx<R>((R)in[i].re, (R)in[i].im);
}

static unsigned grid_for(dim_t n) {
    dim_t b = (n + 255) / 256;
    if (b > kMaxGridBlocks) b = kMaxGridBlocks;
    if (b < 1) b = 1;
    return (unsigned)b;
}

template <typename O>
static int cast_from(void *out, int in_type, const void *in, dim_t n, hipStream_t s) {
    const unsigned g = grid_for(n);
#define CASE(DT, IT)                                                                  \
    case DT:                                                                          \
        cast_kernel<O, IT><<<g, 256, 0, s>>>((O *)out, (const IT *)in, n);            \
        break;
    switch (in_type) {
        CASE(DT_F32, float)
        CASE(DT_F64, double)
        CASE(DT_F16, _Float16)
        CASE(DT_BF16, bf16_t)
        CASE(DT_S32, int32_t)
        CASE(DT_U32, uint32_t)
        CASE(DT_S16, int16_t)
        CASE(DT_U16, uint16_t)
        CASE(DT_S64, long long)
        CASE(DT_U64, unsigned long long)
        CASE(DT_U8, uint8_t)
        CASE(DT_S8, int8_t)
        CASE(DT_B8, int8_t)
        default: return K_ERR_TYPE;
    }
#undef CASE
    return launch_status();
}

template <typename R>
static int cast_to_cplx(void *out, int in_type, const void *in, dim_t n, hipStream_t s) {
    const unsigned g = grid_for(n);
    cplx<R> *o = (cplx<R> *)out;
#define CASE(DT, IT)                                                        \
    case DT:                                                                \
        cast_to_complex<R, IT><<<g, 256, 0, s>>>(o, (const IT *)in, n);     \
        break;
    switch (in_type) {
        case DT_C32: cast_complex<R, float><<<g, 256, 0, s>>>(o, (const cfloat *)in, n); break;
        case DT_C64: cast_complex<R, double><<<g, 256, 0, s>>>(o, (const cdouble *)in, n); break;
        CASE(DT_F32, float)
        CASE(DT_F64, double)
        CASE(DT_F16, _Float16)
        CASE(DT_BF16, bf16_t)
        CASE(DT_S32, int32_t)
        CASE(DT_U32, uint32_t)
        CASE(DT_S16, int16_t)
        CASE(DT_U16, uint16_t)
        CASE(DT_S64, long long)
        CASE(DT_U64, unsigned long long)
        CASE(DT_U8, uint8_t)
        CASE(DT_S8, int8_t)
        CASE(DT_B8, int8_t)
        default: return K_ERR_TYPE;
    }
#undef CASE
    return launch_status();
}

template <typename T>
__global__ __launch_bounds__(256) void fill_kernel(T *__restrict__ out, T v, dim_t n) {
    for (dim_t i = blockIdx.x * (dim_t)blockDim.x + threadIdx.x; i < n;
         i += (dim_t)gridDim.x * blockDim.x)
        out[i] = v;
}

template <typename T, typename AccT>
__global__ __launch_bounds__(256) void scale_kernel(T *__restrict__ out,
                                                    const T *__restrict__ in,
                                                    AccT scale, dim_t n) {
    for (dim_t i = blockIdx.x * (dim_t)blockDim.x + threadIdx.x; i < n;
         i += (dim_t)gridDim.x * blockDim.x)
        out[i] = (T)((AccT)in[i] * scale);
}

}  // namespace afhip

using namespace afhip;

extern "C" int afhip_k_cast(int out_type, void *out, int in_type, const void *in,
                            long long n, void *stream) {
    if (n <= 0) return K_OK;
    hipStream_t s = (hipStream_t)stream;
    switch (out_type) {
        case DT_F32: return cast_from<float>(out, in_type, in, n, s);
        case DT_F64: return cast_from<double>(out, in_type, in, n, s);
        case DT_F16: return cast_from<_Float16>(out, in_type, in, n, s);
        case DT_BF16: return cast_from<bf16_t>(out, in_type, in, n, s);
        case DT_S32: return cast_from<int32_t>(out, in_type, in, n, s);
        case DT_U32: return cast_from<uint32_t>(out, in_type, in, n, s);
        case DT_S16: return cast_from<int16_t>(out, in_type, in, n, s);
        case DT_U16: return cast_from<uint16_t>(out, in_type, in, n, s);
        case DT_S64: return cast_from<long long>(out, in_type, in, n, s);
        case DT_U64: return cast_from<unsigned long long>(out, in_type, in, n, s);
        case DT_U8: return cast_from<uint8_t>(out, in_type, in, n, s);
        case DT_S8: return cast_from<int8_t>(out, in_type, in, n, s);
        case DT_B8: return cast_from<int8_t>(out, in_type, in, n, s);
        case DT_C32: return cast_to_cplx<float>(out, in_type, in, n, s);
        case DT_C64: return cast_to_cplx<double>(out, in_type, in, n, s);
        default: return K_ERR_TYPE;
    }
}

extern "C" int afhip_k_fill(int dtype, void *out, double value, long long n, void *stream) {
    if (n <= 0) return K_OK;
    hipStream_t s = (hipStream_t)stream;
    const unsigned g = grid_for(n);
    switch (dtype) {
        case DT_F32: fill_kernel<float><<<g, 256, 0, s>>>((float *)out, (float)value, n); break;
        case DT_F64: fill_kernel<double><<<g, 256, 0, s>>>((double *)out, value, n); break;
        case DT_F16: fill_kernel<_Float16><<<g, 256, 0, s>>>((_Float16 *)out, (_Float16)(float)value, n); break;
        case DT_BF16:
            fill_kernel<uint16_t><<<g, 256, 0, s>>>((uint16_t *)out, f32_to_bf16_bits((float)value), n);
            break;
        case DT_C32:
            fill_kernel<cfloat><<<g, 256, 0, s>>>((cfloat *)out, cfloat((float)value), n);
            break;
        case DT_C64:
            fill_kernel<cdouble><<<g, 256, 0, s>>>((cdouble *)out, cdouble(value), n);
            break;
        case DT_S32: fill_kernel<int32_t><<<g, 256, 0, s>>>((int32_t *)out, (int32_t)value, n); break;
        case DT_U32: fill_kernel<uint32_t><<<g, 256, 0, s>>>((uint32_t *)out, (uint32_t)value, n); break;
        case DT_S16: fill_kernel<int16_t><<<g, 256, 0, s>>>((int16_t *)out, (int16_t)value, n); break;
        case DT_U16: fill_kernel<uint16_t><<<g, 256, 0, s>>>((uint16_t *)out, (uint16_t)value, n); break;
        case DT_S64: fill_kernel<long long><<<g, 256, 0, s>>>((long long *)out, (long long)value, n); break;
        case DT_U64:
            fill_kernel<unsigned long long><<<g, 256, 0, s>>>((unsigned long long *)out,
                                                              (unsigned long long)value, n);
            break;
        case DT_U8: fill_kernel<uint8_t><<<g, 256, 0, s>>>((uint8_t *)out, (uint8_t)value, n); break;
        case DT_S8:
        case DT_B8: fill_kernel<int8_t><<<g, 256, 0, s>>>((int8_t *)out, (int8_t)value, n); break;
        default: return K_ERR_TYPE;
    }
    return launch_status();
}

extern "C" int afhip_k_scale(int dtype, void *out, const void *in, double scale,
                             long long n, void *stream) {
    if (n <= 0) return K_OK;
    hipStream_t s = (hipStream_t)stream;
    const unsigned g = grid_for(n);
#define CASE(DT, T)                                                                 \
    case DT:                                                                        \
        scale_kernel<T, float><<<g, 256, 0, s>>>((T *)out, (const T *)in,           \
                                                 (float)scale, n);                  \
        break;
    switch (dtype) {
        case DT_F32:
            scale_kernel<float, float><<<g, 256, 0, s>>>((float *)out, (const float *)in,
                                                         (float)scale, n);
            break;
        case DT_F64:
            scale_kernel<double, double><<<g, 256, 0, s>>>((double *)out, (const double *)in,
                                                           scale, n);
            break;
        CASE(DT_S32, int32_t)
        CASE(DT_U32, uint32_t)
        CASE(DT_S16, int16_t)
        CASE(DT_U16, uint16_t)
        CASE(DT_S64, long long)
        CASE(DT_U64, unsigned long long)
        CASE(DT_U8, uint8_t)
        CASE(DT_S8, int8_t)
        CASE(DT_B8, int8_t)
        default: return K_ERR_TYPE;
    }
#undef CASE
    return launch_status();
}

namespace afhip {
template <typename R>
__global__ __launch_bounds__(256) void scale_complex_kernel(cplx<R> *__restrict__ out,
                                                            const cplx<R> *__restrict__ in,
                                                            cplx<R> scale, dim_t n) {
    for (dim_t i = blockIdx.x * (dim_t)blockDim.x + threadIdx.x; i < n;
         i += (dim_t)gridDim.x * blockDim.x)
        out[i] = in[i] * scale;
}
}  // namespace afhip

extern "C" int afhip_k_scale_complex(int dtype, void *out, const void *in, double re, double im,
                                     long long n, void *stream) {
    if (n <= 0) return K_OK;
    hipStream_t s = (hipStream_t)stream;
    const unsigned g = grid_for(n);
    switch (dtype) {
        case DT_C32:
            scale_complex_kernel<float><<<g, 256, 0, s>>>((cfloat *)out, (const cfloat *)in,
                                                          cfloat((float)re, (float)im), n);
            break;
        case DT_C64:
            scale_complex_kernel<double><<<g, 256, 0, s>>>((cdouble *)out, (const cdouble *)in,
                                                           cdouble(re, im), n);
            break;
        default: return K_ERR_TYPE;
    }
    return launch_status();
}
