// Kernels for the internal callers of the convolution path:
//  * region copy (af_fir's af_index of the first x0 rows of the expanded
//    convolution, src/api/c/iir.cpp:28-37);
//  * elementwise difference (af_dog's arithOp<af_sub_t>, src/api/c/dog.cpp:45-50);
//  * the IIR feedback recurrence over the first x0 rows of conv(x, b)
//    (src/backend/cpu/kernel/iir.hpp:20-55): per column,
//      y[i] = (c[i] + z[0]) / a[0];  z[k-1] = z[k] - a[k] * y[i], k = 1..na-1,
//    with z zero-initialised and z[na-1] never written. One lane per column;
//    the state z lives in registers for na <= 16 and in a workspace otherwise.
//    Products and differences are rounded separately (no contraction), as the
//    reference's x86 build evaluates them.
#include "kcommon.hpp"
#include "../../../include/afhip_kernels.h"

namespace afhip {

struct RegionGeo {
    long long od[4];   // region (= output) dims
    long long is[4];   // input strides (elements)
    long long off[4];  // region origin in the input
};

template <typename E>
__global__ __launch_bounds__(256) void region_copy(E *__restrict__ out, const E *__restrict__ in, RegionGeo g) {
    const long long total = g.od[0] * g.od[1] * g.od[2] * g.od[3];
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        long long r = e, src = 0;
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            const long long i = r % g.od[d];
            r /= g.od[d];
            src += (i + g.off[d]) * g.is[d];
        }
        out[e] = in[src];
    }
}

struct b16e {
    uint64_t a, b;
};

template <typename T>
__global__ __launch_bounds__(256) void sub_kernel(T *__restrict__ out, const T *__restrict__ a,
                                                  const T *__restrict__ b, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        out[i] = a[i] - b[i];
}

// complex arithmetic of the recurrence (std::complex semantics for + - *;
// the division is the textbook (c * conj(d)) / |d|^2 form)
// (each body switches contraction off: a pragma in the caller does not reach
// into these functions, and a product fused into the next subtraction would
// round differently from the reference)
template <typename T>
struct IirOps {
    __device__ static T add(T a, T b) {
#pragma clang fp contract(off)
        return a + b;
    }
    __device__ static T sub(T a, T b) {
#pragma clang fp contract(off)
        return a - b;
    }
    __device__ static T mul(T a, T b) {
#pragma clang fp contract(off)
        return a * b;
    }
    __device__ static T div(T a, T b) {
#pragma clang fp contract(off)
        return a / b;
    }
};
template <typename R>
struct IirOps<cplx<R>> {
    typedef cplx<R> C;
    __device__ static C add(C a, C b) {
#pragma clang fp contract(off)
        return C(a.re + b.re, a.im + b.im);
    }
    __device__ static C sub(C a, C b) {
#pragma clang fp contract(off)
        return C(a.re - b.re, a.im - b.im);
    }
    __device__ static C mul(C a, C b) { return a * b; }
    __device__ static C div(C a, C b) {
#pragma clang fp contract(off)
        const R den = b.re * b.re + b.im * b.im;
        return C((a.re * b.re + a.im * b.im) / den, (a.im * b.re - a.re * b.im) / den);
    }
};

struct IirGeo {
    long long yd[4];    // output dims (y[0] = x0)
    long long cs[4];    // strides of the convolution result (c[0] stride 1)
    long long na;       // feedback taps
    long long as1;      // a column stride (0: one a for every column)
    long long ncols;    // yd[1] * yd[2] * yd[3]
};

// NZ > 0: the state in NZ registers (na <= NZ); NZ == 0: in zw (na per column)
template <typename T, int NZ>
__global__ __launch_bounds__(64) void iir_kernel(T *__restrict__ y, const T *__restrict__ c,
                                                 const T *__restrict__ a, IirGeo g, T *__restrict__ zw) {
#pragma clang fp contract(off)
    using O = IirOps<T>;
    const long long col = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (col >= g.ncols) return;
    const long long j = col % g.yd[1], k = (col / g.yd[1]) % g.yd[2], l = col / (g.yd[1] * g.yd[2]);
    const T *cc = c + j * g.cs[1] + k * g.cs[2] + l * g.cs[3];
    T *yy = y + col * g.yd[0];
    const T *aa = a + col * g.as1;
    const int na = (int)g.na;
    if constexpr (NZ > 0) {
        T z[NZ];
#pragma unroll
        for (int q = 0; q < NZ; ++q) z[q] = T(0);
        T ac[NZ];
#pragma unroll
        for (int q = 0; q < NZ; ++q) ac[q] = q < na ? aa[q] : T(0);
        // the inputs arrive in chunks of CHK loaded a chunk ahead into the
        // other of two register buffers, so the recurrence never waits on a
        // single load; the arithmetic and its order are unchanged
        constexpr int CHK = sizeof(T) <= 4 ? 32 : (sizeof(T) <= 8 ? 16 : 8);
        const long long L = g.yd[0];
        T A[CHK], B[CHK];
        auto load = [&](long long i0, T (&dst)[CHK]) __attribute__((always_inline)) {
#pragma unroll
            for (int q = 0; q < CHK; ++q) dst[q] = i0 + q < L ? cc[i0 + q] : T(0);
        };
        auto run = [&](long long i0, const T (&src)[CHK]) __attribute__((always_inline)) {
#pragma unroll
            for (int q = 0; q < CHK; ++q) {
                if (i0 + q >= L) break;
                const T v = O::div(O::add(src[q], z[0]), ac[0]);
                yy[i0 + q] = v;
#pragma unroll
                for (int r = 1; r < NZ; ++r)
                    if (r < na) z[r - 1] = O::sub(z[r], O::mul(ac[r], v));
            }
        };
        load(0, A);
        for (long long i0 = 0; i0 < L; i0 += 2 * CHK) {
            if (i0 + CHK < L) load(i0 + CHK, B);
            run(i0, A);
            if (i0 + CHK >= L) break;
            if (i0 + 2 * CHK < L) load(i0 + 2 * CHK, A);
            run(i0 + CHK, B);
        }
    } else {
        T *z = zw + col * g.na;
        for (int q = 0; q < na; ++q) z[q] = T(0);
        for (long long i = 0; i < g.yd[0]; ++i) {
            const T v = O::div(O::add(cc[i], z[0]), aa[0]);
            yy[i] = v;
            for (int q = 1; q < na; ++q) z[q - 1] = O::sub(z[q], O::mul(aa[q], v));
        }
    }
}

// out[i, col] = b[i, col] / a[col or 0]: af_div(b, a, batch = true) for a
// with one element per column (src/api/c/iir.cpp:76-81)
template <typename T>
__global__ __launch_bounds__(256) void div_cols_kernel(T *__restrict__ out, const T *__restrict__ b,
                                                       const T *__restrict__ a, long long rows, long long n,
                                                       int per_col) {
#pragma clang fp contract(off)
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        out[i] = IirOps<T>::div(b[i], a[per_col ? i / rows : 0]);
}

static unsigned grid_of(long long n) {
    long long b = (n + 255) / 256;
    if (b > kMaxGridBlocks) b = kMaxGridBlocks;
    return (unsigned)(b < 1 ? 1 : b);
}

template <typename T>
static int launch_iir(void *y, const void *c, const void *a, const IirGeo &g, void *zw, hipStream_t s) {
    const unsigned blocks = (unsigned)((g.ncols + 63) / 64);
    if (g.na <= 4)
        iir_kernel<T, 4><<<blocks, 64, 0, s>>>((T *)y, (const T *)c, (const T *)a, g, nullptr);
    else if (g.na <= 16)
        iir_kernel<T, 16><<<blocks, 64, 0, s>>>((T *)y, (const T *)c, (const T *)a, g, nullptr);
    else
        iir_kernel<T, 0><<<blocks, 64, 0, s>>>((T *)y, (const T *)c, (const T *)a, g, (T *)zw);
    return launch_status();
}

}  // namespace afhip

using namespace afhip;

extern "C" int afhip_k_region_copy(int esize, void *out, const long long *odims, const void *in,
                                   const long long *idims, const long long *offsets, void *stream) {
    RegionGeo g{};
    long long st = 1;
    for (int d = 0; d < 4; ++d) {
        if (offsets[d] < 0 || odims[d] < 0 || offsets[d] + odims[d] > idims[d]) return K_ERR_SIZE;
        g.od[d] = odims[d];
        g.off[d] = offsets[d];
        g.is[d] = st;
        st *= idims[d];
    }
    const long long n = odims[0] * odims[1] * odims[2] * odims[3];
    if (n == 0) return K_OK;
    hipStream_t s = (hipStream_t)stream;
    const unsigned gb = grid_of(n);
    switch (esize) {
        case 1: region_copy<uint8_t><<<gb, 256, 0, s>>>((uint8_t *)out, (const uint8_t *)in, g); break;
        case 2: region_copy<uint16_t><<<gb, 256, 0, s>>>((uint16_t *)out, (const uint16_t *)in, g); break;
        case 4: region_copy<uint32_t><<<gb, 256, 0, s>>>((uint32_t *)out, (const uint32_t *)in, g); break;
        case 8: region_copy<uint64_t><<<gb, 256, 0, s>>>((uint64_t *)out, (const uint64_t *)in, g); break;
        case 16: region_copy<b16e><<<gb, 256, 0, s>>>((b16e *)out, (const b16e *)in, g); break;
        default: return K_ERR_TYPE;
    }
    return launch_status();
}

extern "C" int afhip_k_sub(int dtype, void *out, const void *a, const void *b, long long n, void *stream) {
    if (n == 0) return K_OK;
    hipStream_t s = (hipStream_t)stream;
    switch (dtype) {
        case DT_F32: sub_kernel<float><<<grid_of(n), 256, 0, s>>>((float *)out, (const float *)a, (const float *)b, n); break;
        case DT_F64: sub_kernel<double><<<grid_of(n), 256, 0, s>>>((double *)out, (const double *)a, (const double *)b, n); break;
        default: return K_ERR_TYPE;
    }
    return launch_status();
}

extern "C" int afhip_k_div_cols(int dtype, void *out, const void *b, const long long *bdims, const void *a,
                                int per_col, void *stream) {
    const long long n = bdims[0] * bdims[1] * bdims[2] * bdims[3];
    if (n == 0) return K_OK;
    hipStream_t s = (hipStream_t)stream;
    const unsigned g = grid_of(n);
    switch (dtype) {
        case DT_F32: div_cols_kernel<float><<<g, 256, 0, s>>>((float *)out, (const float *)b, (const float *)a, bdims[0], n, per_col); break;
        case DT_F64: div_cols_kernel<double><<<g, 256, 0, s>>>((double *)out, (const double *)b, (const double *)a, bdims[0], n, per_col); break;
        case DT_C32: div_cols_kernel<cfloat><<<g, 256, 0, s>>>((cfloat *)out, (const cfloat *)b, (const cfloat *)a, bdims[0], n, per_col); break;
        case DT_C64: div_cols_kernel<cdouble><<<g, 256, 0, s>>>((cdouble *)out, (const cdouble *)b, (const cdouble *)a, bdims[0], n, per_col); break;
        default: return K_ERR_TYPE;
    }
    return launch_status();
}

extern "C" long long afhip_k_iir_workspace(int dtype, long long na, const long long *ydims) {
    if (na <= 16) return 0;
    const long long es = dtype == DT_F32 ? 4 : (dtype == DT_F64 || dtype == DT_C32) ? 8 : dtype == DT_C64 ? 16 : 0;
    if (es == 0) return -1;
    return es * na * ydims[1] * ydims[2] * ydims[3];
}

extern "C" int afhip_k_iir(int dtype, void *y, const long long *ydims, const void *c, const long long *cdims,
                           const void *a, const long long *adims, void *workspace, long long workspace_bytes,
                           void *stream) {
    IirGeo g{};
    long long st = 1;
    for (int d = 0; d < 4; ++d) {
        g.yd[d] = ydims[d];
        g.cs[d] = st;
        st *= cdims[d];
    }
    if (cdims[0] < ydims[0] || cdims[1] != ydims[1] || cdims[2] != ydims[2] || cdims[3] != ydims[3]) return K_ERR_SIZE;
    g.na = adims[0];
    if (g.na < 1) return K_ERR_SIZE;
    const long long acols = adims[1] * adims[2] * adims[3];
    g.ncols = ydims[1] * ydims[2] * ydims[3];
    if (acols == 1) g.as1 = 0;
    else if (adims[1] == ydims[1] && adims[2] == ydims[2] && adims[3] == ydims[3]) g.as1 = g.na;
    else return K_ERR_SIZE;
    if (g.ncols == 0 || ydims[0] == 0) return K_OK;
    const long long need = afhip_k_iir_workspace(dtype, g.na, ydims);
    if (need < 0) return K_ERR_TYPE;
    if (need > 0 && (workspace == nullptr || workspace_bytes < need)) return K_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    switch (dtype) {
        case DT_F32: return launch_iir<float>(y, c, a, g, workspace, s);
        case DT_F64: return launch_iir<double>(y, c, a, g, workspace, s);
        case DT_C32: return launch_iir<cfloat>(y, c, a, g, workspace, s);
        case DT_C64: return launch_iir<cdouble>(y, c, a, g, workspace, s);
        default: return K_ERR_TYPE;
    }
}
