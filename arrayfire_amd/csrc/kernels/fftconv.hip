// Frequency-domain convolution: af_fft_convolve1/2/3 and the AUTO route of
// af_convolve1/2/3 (src/api/c/fftconvolve.cpp:123-264,
// src/backend/cpu/fftconvolve.cpp:42-203, the DIFF / complex fallback at
// src/api/c/fftconvolve.cpp:52-121).
//
// Semantics: every base dimension k < rank is zero-padded to
// L_k = nextpow2(s_k + f_k - 1); the signal and filter spectra are multiplied
// with the batch kind's pairing of slices (NONE, LHS, RHS, SAME, and DIFF's
// broadcast over dims where either operand has extent 1), transformed back,
// scaled by 1 / prod(L_k) and cropped at f_k / 2 (DEFAULT) or 0 (EXPAND).
// Arithmetic is in convT = float for f32, c32 and every integer type, double
// for f64 and c64. Integer outputs are rounded to nearest (roundf) on the
// batched path (kernel/fftconvolve.hpp:186-218) and truncated on the DIFF
// fallback (a plain cast, src/backend/cpu/cast.hpp).
//
// Layout in HBM: complex buffers [L0, L1, L2, batch] (dims past the rank are
// the batch, folded into one index), convT pairs.
//
// Kernels:
//  * fft_pad: signal or filter (any dtype) -> zero-padded complex buffer;
//  * fft_pass: a batched radix-2 FFT of power-of-two length L <= 4096 along
//    one axis (element stride `stride`), M transforms per 256-lane workgroup
//    staged in LDS (M * L <= 4096 complex values), loaded in bit-reversed
//    order and butterflied in place, one barrier per stage; columns of a
//    strided axis are taken M consecutive at a time so loads coalesce.
//    Axes longer than 4096 run as two passes of the four-step decomposition
//    L = L1 * L2, with the W_L^(n1 k2) twiddles fused into the first forward
//    pass (after the transform) or the last inverse pass (before it);
//  * fft_mul: spectrum product with the batch pairing (in place when the
//    output slice is the signal's or the filter's own);
//  * fft_crop: crop, scale and the output dtype's conversion rule.
// Twiddles W_4096^j (j < 2048) live in a static device table per precision,
// filled once per device by fft_twiddle_init (sincospi in double).
#include <mutex>

#include "kcommon.hpp"
#include "../../../include/afhip_kernels.h"

namespace afhip {

constexpr int FFT_LMAX = 4096;        // longest in-LDS transform
constexpr int FFT_LOGMAX = 12;
constexpr int FFT_THREADS = 256;

static __device__ float2 g_fft_tw32[FFT_LMAX / 2];
static __device__ double2 g_fft_tw64[FFT_LMAX / 2];

__global__ __launch_bounds__(256) void fft_twiddle_init() {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < FFT_LMAX / 2; j += gridDim.x * blockDim.x) {
        double s, c;
        sincospi(-2.0 * j / FFT_LMAX, &s, &c);
        g_fft_tw32[j] = make_float2((float)c, (float)s);
        g_fft_tw64[j] = make_double2(c, s);
    }
}

template <typename R>
struct FT;
template <>
struct FT<float> {
    typedef float2 C;
    __device__ static C tw(int j) { return g_fft_tw32[j]; }
    __device__ static C mk(float a, float b) { return make_float2(a, b); }
    // W_L^m for arbitrary L (four-step twiddles), evaluated in double
    __device__ static C wl(long long m, long long L, bool inv) {
        double s, c;
        sincospi((inv ? 2.0 : -2.0) * (double)m / (double)L, &s, &c);
        return make_float2((float)c, (float)s);
    }
};
template <>
struct FT<double> {
    typedef double2 C;
    __device__ static C tw(int j) { return g_fft_tw64[j]; }
    __device__ static C mk(double a, double b) { return make_double2(a, b); }
    __device__ static C wl(long long m, long long L, bool inv) {
        double s, c;
        sincospi((inv ? 2.0 : -2.0) * (double)m / (double)L, &s, &c);
        return make_double2(c, s);
    }
};

template <typename C>
__device__ inline C cmul2(C a, C b) {
    C r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}

struct FftPass {
    long long L;       // transform length (power of two, <= FFT_LMAX)
    long long stride;  // element stride along the transform
    long long ntrans;  // number of transforms
    int logL, M;       // log2(L), transforms per workgroup
    int inv;           // 0 forward (W = e^{-2 pi i / L}), 1 inverse (conjugate, unscaled)
    int tw_mode;       // 0 none; 1 post-multiply output k by W_TL^(n1 k); 2 pre-multiply input k by W_TL^-(n1 k)
    long long TL, tst; // four-step: full length and the axis stride (n1 = (i / tst) % (TL / L))
};

// Transform t covers elements base + n * stride, n < L, with
// base = (t % stride) + (t / stride) * L * stride.
template <typename R>
__global__ __launch_bounds__(FFT_THREADS) void fft_pass(typename FT<R>::C *__restrict__ data, FftPass p) {
    using C = typename FT<R>::C;
    __shared__ C buf[FFT_LMAX];
    const int L = (int)p.L, M = p.M, tid = threadIdx.x;
    const long long t0 = (long long)blockIdx.x * M;
    const int cnt = M * L;
    const int lshift = FFT_LOGMAX - p.logL;  // table index scale: W_L^k = W_4096^(k << lshift)
    auto base_of = [&](long long t) -> long long { return (t % p.stride) + (t / p.stride) * p.L * p.stride; };
    auto n1_of = [&](long long t) -> long long {
        const long long i = t % p.stride;
        return (i / p.tst) % (p.TL / p.L);
    };
    // load, bit-reversed, with the optional pre-twiddle
    for (int e = tid; e < cnt; e += FFT_THREADS) {
        int m, n;
        if (p.stride == 1) { n = e % L; m = e / L; } else { m = e % M; n = e / M; }
        const long long t = t0 + m;
        C v = FT<R>::mk(0, 0);
        if (t < p.ntrans) {
            v = data[base_of(t) + (long long)n * p.stride];
            if (p.tw_mode == 2) v = cmul2(v, FT<R>::wl((n1_of(t) * n) % p.TL, p.TL, true));
        }
        const int rn = (int)(__builtin_bitreverse32((unsigned)n) >> (32 - p.logL));
        buf[m * L + (p.logL ? rn : 0)] = v;
    }
    // radix-2 decimation-in-time stages
    const int half = cnt / 2;
    for (int s = 0; s < p.logL; ++s) {
        __syncthreads();
        const int h = 1 << s;
        for (int b = tid; b < half; b += FFT_THREADS) {
            const int m = b >> (p.logL - 1), j = b & (L / 2 - 1);
            const int k = j & (h - 1);
            const int i0 = m * L + ((j >> s) << (s + 1)) + k, i1 = i0 + h;
            C w = FT<R>::tw((k << (p.logL - 1 - s)) << lshift);
            if (p.inv) w.y = -w.y;
            const C a = buf[i0], bv = cmul2(buf[i1], w);
            C u, d;
            u.x = a.x + bv.x; u.y = a.y + bv.y;
            d.x = a.x - bv.x; d.y = a.y - bv.y;
            buf[i0] = u;
            buf[i1] = d;
        }
    }
    __syncthreads();
    for (int e = tid; e < cnt; e += FFT_THREADS) {
        int m, n;
        if (p.stride == 1) { n = e % L; m = e / L; } else { m = e % M; n = e / M; }
        const long long t = t0 + m;
        if (t >= p.ntrans) continue;
        C v = buf[m * L + n];
        if (p.tw_mode == 1) v = cmul2(v, FT<R>::wl((n1_of(t) * n) % p.TL, p.TL, false));
        data[base_of(t) + (long long)n * p.stride] = v;
    }
}

// element conversions into the complex compute type
template <typename T, typename R>
__device__ inline typename FT<R>::C to_c(const T &v) { return FT<R>::mk((R)v, (R)0); }
template <>
__device__ inline float2 to_c<cfloat, float>(const cfloat &v) { return make_float2(v.re, v.im); }
template <>
__device__ inline float2 to_c<cdouble, float>(const cdouble &v) { return make_float2((float)v.re, (float)v.im); }
template <>
__device__ inline double2 to_c<cfloat, double>(const cfloat &v) { return make_double2(v.re, v.im); }
template <>
__device__ inline double2 to_c<cdouble, double>(const cdouble &v) { return make_double2(v.re, v.im); }

struct FftGeo {
    long long L[3];       // padded base dims (1 past the rank)
    long long PS;         // padded elements per slice: L0 * L1 * L2
    long long d[4];       // source dims (pad) or output dims (crop)
    long long sb[4];      // pad: unused; mul: signal batch dims; crop: unused
    long long fb[4];      // mul: filter batch dims
    long long ob[4];      // mul: output batch dims
    long long off[3];     // crop: offset into the padded domain per base dim
    int rank;
};

// src [d0, d1, d2, d3] -> buffer [L0, L1, L2, batch]; base dims past d are zero.
template <typename T, typename R>
__global__ __launch_bounds__(256) void fft_pad(typename FT<R>::C *__restrict__ out, const T *__restrict__ src,
                                               FftGeo g, long long nbatch) {
    const long long total = g.PS * nbatch;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        long long r = e;
        const long long i0 = r % g.L[0];
        r /= g.L[0];
        const long long i1 = r % g.L[1];
        r /= g.L[1];
        const long long i2 = r % g.L[2];
        const long long b = r / g.L[2];
        typename FT<R>::C v = FT<R>::mk(0, 0);
        // batch b walks the source's dims past the rank in column-major order
        if (i0 < g.d[0] && (g.rank < 2 || i1 < g.d[1]) && (g.rank < 3 || i2 < g.d[2])) {
            long long off;
            if (g.rank == 1) off = i0 + b * g.d[0];
            else if (g.rank == 2) off = i0 + g.d[0] * (i1 + b * g.d[1]);
            else off = i0 + g.d[0] * (i1 + g.d[1] * (i2 + b * g.d[2]));
            v = to_c<T, R>(src[off]);
        }
        out[e] = v;
    }
}

// O[p, ob] = S[p, sidx(ob)] * F[p, fidx(ob)], batch multi-indices over dims
// rank..3 (an operand of extent 1 in a dim is broadcast along it).
template <typename R>
__global__ __launch_bounds__(256) void fft_mul(typename FT<R>::C *__restrict__ O,
                                               const typename FT<R>::C *__restrict__ S,
                                               const typename FT<R>::C *__restrict__ F, FftGeo g,
                                               long long nob) {
    const long long total = g.PS * nob;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long pidx = e % g.PS;
        long long ob = e / g.PS;
        long long si = 0, fi = 0, smul = 1, fmul = 1;
        for (int d = g.rank; d < 4; ++d) {
            const long long od = ob % g.ob[d];
            ob /= g.ob[d];
            si += (g.sb[d] == 1 ? 0 : od) * smul;
            fi += (g.fb[d] == 1 ? 0 : od) * fmul;
            smul *= g.sb[d];
            fmul *= g.fb[d];
        }
        O[e] = cmul2(S[pidx + si * g.PS], F[pidx + fi * g.PS]);
    }
}

// output conversion rules
template <typename T, typename R, int RULE>  // RULE 0: cast; 1: roundf (integers)
struct FftOut {
    __device__ static T put(typename FT<R>::C v, R scale) {
        const R x = v.x / scale;
        if constexpr (RULE == 1) return (T)roundf((float)x);
        else return (T)x;
    }
};
template <typename R, int RULE>
struct FftOut<cfloat, R, RULE> {
    __device__ static cfloat put(typename FT<R>::C v, R scale) { return cfloat((float)(v.x / scale), (float)(v.y / scale)); }
};
template <typename R, int RULE>
struct FftOut<cdouble, R, RULE> {
    __device__ static cdouble put(typename FT<R>::C v, R scale) { return cdouble(v.x / scale, v.y / scale); }
};

// out [d0..d3] (column-major) <- O[(o_k + off_k) per base dim, batch]
template <typename T, typename R, int RULE>
__global__ __launch_bounds__(256) void fft_crop(T *__restrict__ out, const typename FT<R>::C *__restrict__ O,
                                                FftGeo g, R scale) {
    const long long total = g.d[0] * g.d[1] * g.d[2] * g.d[3];
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        long long r = e, idx[4];
        for (int d = 0; d < 4; ++d) {
            idx[d] = r % g.d[d];
            r /= g.d[d];
        }
        long long b = 0, bm = 1;
        for (int d = g.rank; d < 4; ++d) {
            b += idx[d] * bm;
            bm *= g.d[d];
        }
        long long p = idx[0] + g.off[0];
        if (g.rank > 1) p += g.L[0] * (idx[1] + g.off[1]);
        if (g.rank > 2) p += g.L[0] * g.L[1] * (idx[2] + g.off[2]);
        out[e] = FftOut<T, R, RULE>::put(O[p + b * g.PS], scale);
    }
}

// ---------------------------------------------------------------- host side
static long long nextpow2ll(long long x) {
    long long p = 1;
    while (p < x) p <<= 1;
    return p;
}
static int log2ll(long long x) {
    int l = 0;
    while ((1LL << l) < x) ++l;
    return l;
}
static unsigned grid_of(long long n) {
    long long b = (n + 255) / 256;
    if (b > kMaxGridBlocks) b = kMaxGridBlocks;
    if (b < 1) b = 1;
    return (unsigned)b;
}

static std::mutex g_tw_mutex;
static bool g_tw_ready[64] = {};

static int ensure_twiddles(hipStream_t s) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return K_ERR_RUNTIME;
    std::lock_guard<std::mutex> lock(g_tw_mutex);
    if (!g_tw_ready[dev]) {
        // filled once per device; the wait makes the table visible to work
        // queued later on any stream of the device
        fft_twiddle_init<<<8, 256, 0, s>>>();
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) return K_ERR_RUNTIME;
        g_tw_ready[dev] = true;
    }
    return K_OK;
}

// One FFT along an axis of length L with element stride st, over a buffer of
// `total` complex elements (a four-step pair of passes when L > FFT_LMAX).
template <typename R>
static int fft_axis(typename FT<R>::C *buf, long long total, long long L, long long st, bool inv, hipStream_t s) {
    if (L <= 1) return K_OK;
    auto run = [&](long long len, long long stride, int tw_mode, long long TL, long long tst) {
        FftPass p;
        p.L = len;
        p.stride = stride;
        p.ntrans = total / len;
        p.logL = log2ll(len);
        int M = (int)(FFT_LMAX / len);
        if (stride > 1 && M > stride) M = (int)stride;
        if (stride == 1 && M > p.ntrans) M = (int)(p.ntrans > 0 ? p.ntrans : 1);
        p.M = M < 1 ? 1 : M;
        p.inv = inv ? 1 : 0;
        p.tw_mode = tw_mode;
        p.TL = TL;
        p.tst = tst;
        const long long blocks = (p.ntrans + p.M - 1) / p.M;
        fft_pass<R><<<dim3((unsigned)blocks), FFT_THREADS, 0, s>>>(buf, p);
    };
    if (L <= FFT_LMAX) {
        run(L, st, 0, 1, 1);
        return launch_status();
    }
    // four-step: n = n1 + L1 n2 (L1 fast), L1 = FFT_LMAX, L2 = L / L1
    const long long L1 = FFT_LMAX, L2 = L / L1;
    if (L2 > FFT_LMAX) return K_ERR_SIZE;
    if (!inv) {
        run(L2, L1 * st, 1, L, st);  // along n2 for each n1, then W_L^(n1 k2)
        run(L1, st, 0, 1, 1);        // along n1 for each k2
    } else {
        run(L1, st, 0, 1, 1);        // inverse along the k1 slot
        run(L2, L1 * st, 2, L, st);  // W_L^-(n1 k2), then inverse along the k2 slot
    }
    return launch_status();
}

struct FftPlan {
    long long L[3] = {1, 1, 1};
    long long PS = 1, nbs = 1, nbf = 1, nbo = 1;
    int esz = 8;  // bytes per complex element
    bool dbl = false;
};

static bool is_complex(int t) { return t == DT_C32 || t == DT_C64; }

static FftPlan fft_plan(int dtype, const long long *sd, const long long *fd, const long long *od, int rank) {
    FftPlan P;
    P.dbl = dtype == DT_F64 || dtype == DT_C64;
    P.esz = P.dbl ? 16 : 8;
    for (int k = 0; k < rank; ++k) {
        P.L[k] = nextpow2ll(sd[k] + fd[k] - 1);
        P.PS *= P.L[k];
    }
    for (int k = rank; k < 4; ++k) {
        P.nbs *= sd[k];
        P.nbf *= fd[k];
        P.nbo *= od[k];
    }
    return P;
}

template <typename T, typename R>
static void launch_pad(void *buf, const void *src, const long long *dims, const FftPlan &P, int rank, long long nb,
                       hipStream_t s) {
    FftGeo g{};
    for (int k = 0; k < 3; ++k) g.L[k] = P.L[k];
    g.PS = P.PS;
    for (int k = 0; k < 4; ++k) g.d[k] = dims[k];
    g.rank = rank;
    fft_pad<T, R><<<grid_of(P.PS * nb), 256, 0, s>>>((typename FT<R>::C *)buf, (const T *)src, g, nb);
}

template <typename R>
static int pad_any(int type, void *buf, const void *src, const long long *dims, const FftPlan &P, int rank,
                   long long nb, hipStream_t s) {
    switch (type) {
        case DT_F32: launch_pad<float, R>(buf, src, dims, P, rank, nb, s); break;
        case DT_F64: launch_pad<double, R>(buf, src, dims, P, rank, nb, s); break;
        case DT_C32: launch_pad<cfloat, R>(buf, src, dims, P, rank, nb, s); break;
        case DT_C64: launch_pad<cdouble, R>(buf, src, dims, P, rank, nb, s); break;
        case DT_S32: launch_pad<int32_t, R>(buf, src, dims, P, rank, nb, s); break;
        case DT_U32: launch_pad<uint32_t, R>(buf, src, dims, P, rank, nb, s); break;
        case DT_S16: launch_pad<int16_t, R>(buf, src, dims, P, rank, nb, s); break;
        case DT_U16: launch_pad<uint16_t, R>(buf, src, dims, P, rank, nb, s); break;
        case DT_S64: launch_pad<long long, R>(buf, src, dims, P, rank, nb, s); break;
        case DT_U64: launch_pad<unsigned long long, R>(buf, src, dims, P, rank, nb, s); break;
        case DT_U8: launch_pad<uint8_t, R>(buf, src, dims, P, rank, nb, s); break;
        case DT_S8:
        case DT_B8: launch_pad<int8_t, R>(buf, src, dims, P, rank, nb, s); break;
        default: return K_ERR_TYPE;
    }
    return launch_status();
}

template <typename T, typename R>
static void launch_crop(void *out, const void *O, const long long *od, const long long *fd, const FftPlan &P,
                        int rank, bool expand, int rule, hipStream_t s) {
    FftGeo g{};
    for (int k = 0; k < 3; ++k) {
        g.L[k] = P.L[k];
        g.off[k] = (k < rank && !expand) ? fd[k] / 2 : 0;
    }
    g.PS = P.PS;
    for (int k = 0; k < 4; ++k) g.d[k] = od[k];
    g.rank = rank;
    const R scale = (R)P.PS;
    const long long n = od[0] * od[1] * od[2] * od[3];
    if (rule == 1)
        fft_crop<T, R, 1><<<grid_of(n), 256, 0, s>>>((T *)out, (const typename FT<R>::C *)O, g, scale);
    else
        fft_crop<T, R, 0><<<grid_of(n), 256, 0, s>>>((T *)out, (const typename FT<R>::C *)O, g, scale);
}

template <typename R>
static int crop_any(int type, void *out, const void *O, const long long *od, const long long *fd, const FftPlan &P,
                    int rank, bool expand, bool diff, hipStream_t s) {
    const int rule = diff ? 0 : 1;  // integers: roundf on the batched path, truncation on the fallback
    switch (type) {
        case DT_F32: launch_crop<float, R>(out, O, od, fd, P, rank, expand, 0, s); break;
        case DT_F64: launch_crop<double, R>(out, O, od, fd, P, rank, expand, 0, s); break;
        case DT_C32: launch_crop<cfloat, R>(out, O, od, fd, P, rank, expand, 0, s); break;
        case DT_C64: launch_crop<cdouble, R>(out, O, od, fd, P, rank, expand, 0, s); break;
        case DT_S32: launch_crop<int32_t, R>(out, O, od, fd, P, rank, expand, rule, s); break;
        case DT_U32: launch_crop<uint32_t, R>(out, O, od, fd, P, rank, expand, rule, s); break;
        case DT_S16: launch_crop<int16_t, R>(out, O, od, fd, P, rank, expand, rule, s); break;
        case DT_U16: launch_crop<uint16_t, R>(out, O, od, fd, P, rank, expand, rule, s); break;
        case DT_S64: launch_crop<long long, R>(out, O, od, fd, P, rank, expand, rule, s); break;
        case DT_U64: launch_crop<unsigned long long, R>(out, O, od, fd, P, rank, expand, rule, s); break;
        case DT_U8: launch_crop<uint8_t, R>(out, O, od, fd, P, rank, expand, rule, s); break;
        case DT_S8:
        case DT_B8: launch_crop<int8_t, R>(out, O, od, fd, P, rank, expand, rule, s); break;
        default: return K_ERR_TYPE;
    }
    return launch_status();
}

template <typename R>
static int fft_convolve_run(int dtype, void *out, const long long *od, const void *in, const long long *sd,
                            int ftype, const void *filt, const long long *fd, int kind, int rank, bool expand,
                            char *ws, const FftPlan &P, hipStream_t s) {
    using C = typename FT<R>::C;
    C *S = (C *)ws;
    C *F = S + P.PS * P.nbs;
    const bool diff = kind == 5;
    C *O = diff ? F + P.PS * P.nbf : (kind == 3 ? F : S);  // RHS: in place in F; DIFF: its own buffer
    int rc = ensure_twiddles(s);
    if (rc) return rc;
    if ((rc = pad_any<R>(dtype, S, in, sd, P, rank, P.nbs, s))) return rc;
    if ((rc = pad_any<R>(ftype, F, filt, fd, P, rank, P.nbf, s))) return rc;
    long long st = 1;
    for (int k = 0; k < rank; ++k) {
        if ((rc = fft_axis<R>(S, P.PS * P.nbs, P.L[k], st, false, s))) return rc;
        if ((rc = fft_axis<R>(F, P.PS * P.nbf, P.L[k], st, false, s))) return rc;
        st *= P.L[k];
    }
    FftGeo g{};
    for (int k = 0; k < 3; ++k) g.L[k] = P.L[k];
    g.PS = P.PS;
    g.rank = rank;
    for (int k = 0; k < 4; ++k) {
        g.sb[k] = sd[k];
        g.fb[k] = fd[k];
        g.ob[k] = od[k];
    }
    fft_mul<R><<<grid_of(P.PS * P.nbo), 256, 0, s>>>(O, S, F, g, P.nbo);
    if ((rc = launch_status())) return rc;
    for (int k = rank - 1; k >= 0; --k) {
        st /= P.L[k];
        if ((rc = fft_axis<R>(O, P.PS * P.nbo, P.L[k], st, true, s))) return rc;
    }
    return crop_any<R>(dtype, out, O, od, fd, P, rank, expand, diff, s);
}

}  // namespace afhip

using namespace afhip;

extern "C" long long afhip_k_fft_convolve_workspace(int dtype, const long long *sdims, const long long *fdims,
                                                    const long long *odims, int kind, int rank) {
    if (rank < 1 || rank > 3) return -1;
    const FftPlan P = fft_plan(dtype, sdims, fdims, odims, rank);
    for (int k = 0; k < rank; ++k)
        if (P.L[k] > (long long)FFT_LMAX * FFT_LMAX) return -1;
    long long n = P.PS * (P.nbs + P.nbf + (kind == 5 ? P.nbo : 0));
    return n * P.esz;
}

extern "C" int afhip_k_fft_convolve(int dtype, void *out, const long long *odims, const void *in,
                                    const long long *sdims, int ftype, const void *filt, const long long *fdims,
                                    int kind, int rank, int expand, void *workspace, long long workspace_bytes,
                                    void *stream) {
    if (rank < 1 || rank > 3 || kind < 1 || kind > 5) return K_ERR_ARG;
    const long long need = afhip_k_fft_convolve_workspace(dtype, sdims, fdims, odims, kind, rank);
    if (need < 0) return K_ERR_SIZE;
    if (workspace == nullptr || workspace_bytes < need) return K_ERR_ARG;
    if (odims[0] * odims[1] * odims[2] * odims[3] == 0) return K_OK;
    const FftPlan P = fft_plan(dtype, sdims, fdims, odims, rank);
    hipStream_t s = (hipStream_t)stream;
    if (P.dbl)
        return fft_convolve_run<double>(dtype, out, odims, in, sdims, ftype, filt, fdims, kind, rank, expand != 0,
                                        (char *)workspace, P, s);
    return fft_convolve_run<float>(dtype, out, odims, in, sdims, ftype, filt, fdims, kind, rank, expand != 0,
                                   (char *)workspace, P, s);
}
