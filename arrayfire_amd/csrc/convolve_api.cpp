// The af_convolve* C API of libafhip and its backend host layer.
//
// Dispatch restates src/api/c/convolve.cpp (rank fallback, domain heuristic,
// batch-kind detection, dtype -> accumulator map, asserts and error codes).
// The backend layer ("detail", src/backend/cpu/convolve.cpp:38-243) computes
// output dims, allocates from the stream-ordered pool and queues the gfx950
// kernels through the kernel C ABI (include/afhip_kernels.h).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <memory>
#include <vector>
#include <string>
#include <functional>

#include "../../include/af/array.h"
#include "../../include/af/ml.h"
#include "../../include/af/signal.h"
#include "../../include/afhip_kernels.h"
#include "autotune.hpp"
#include "runtime.hpp"

namespace afhip {

// AF_BATCH_KIND, src/backend/common/internal_enums.hpp:15-22
enum BatchKind {
    BATCH_NONE = 1,
    BATCH_LHS = 2,
    BATCH_RHS = 3,
    BATCH_SAME = 4,
    BATCH_DIFF = 5,
    BATCH_UNSUPPORTED = 6
};

// src/api/c/convolve.cpp:77-97
static BatchKind identify_batch_kind(int rank, const dim4 &s, const dim4 &f) {
    const dim_t sn = s.ndims(), fn = f.ndims();
    if (sn == rank && fn == rank) return BATCH_NONE;
    if (sn == rank && (fn > rank && fn <= AF_MAX_DIMS)) return BATCH_RHS;
    if ((sn > rank && sn <= AF_MAX_DIMS) && fn == rank) return BATCH_LHS;
    if ((sn > rank && sn <= AF_MAX_DIMS) && (fn > rank && fn <= AF_MAX_DIMS)) {
        bool match = true, inter = true;
        for (int i = rank; i < AF_MAX_DIMS; ++i) {
            match &= s[i] == f[i];
            inter &= (s[i] == 1 || f[i] == 1 || s[i] == f[i]);
        }
        if (match) return BATCH_SAME;
        return inter ? BATCH_DIFF : BATCH_UNSUPPORTED;
    }
    return BATCH_UNSUPPORTED;
}

// src/api/c/convolve.cpp:99-130
static bool is_freq_domain(int rank, const dim4 &s, const dim4 &f, af_conv_domain domain) {
    if (domain == AF_CONV_FREQ) return true;
    if (domain != AF_CONV_AUTO) return false;
    if (identify_batch_kind(rank, s, f) == BATCH_DIFF) return true;
    dim_t kbatch = 1;
    for (int i = 3; i >= rank; --i) kbatch *= f[i];
    if (kbatch >= 10) return true;
    if (rank == 1 && f[0] > 128) return true;
    if (rank == 2) {
        if (f[0] > 17 || f[1] > 17) return true;
        if (f[0] != f[1] && f[0] > 5) return true;
    }
    if (rank == 3 && (f[0] > 5 || f[1] > 5 || f[2] > 5)) return true;
    return false;
}

// Accumulator dtype per signal dtype: src/api/c/convolve.cpp:155-209.
static af_dtype acc_type(af_dtype t) {
    switch ((int)t) {
        case c32: return c32;
        case c64: return c64;
        case f32: return f32;
        case f64: return f64;
        case u32:
        case s32:
        case u16:
        case s16:
        case u64:
        case s64:
        case u8:
        case s8:
        case b8: return f32;
        default: TYPE_ERROR(1, t);
    }
}

// castArray<accT>: a device copy converted to accT (shares storage when the
// type already matches).
static std::shared_ptr<Buffer> cast_buffer(const ArrayImpl &a, af_dtype to) {
    if (a.type == to) return a.buf;
    const dim_t n = a.elements();
    auto b = scratch((size_t)n * dtype_size(to));
    check_k(afhip_k_cast(to, b->ptr, a.type, a.data(), n, stream()), "cast");
    return b;
}

// ---------------------------------------------------------------- detail
namespace detail {

// Output dims: src/backend/cpu/convolve.cpp:44-58
static dim4 conv_out_dims(const dim4 &s, const dim4 &f, BatchKind kind, int rank, bool expand) {
    dim4 o(1, 1, 1, 1);
    if (expand) {
        for (int d = 0; d < AF_MAX_DIMS; ++d) {
            if (kind == BATCH_NONE || kind == BATCH_RHS)
                o[d] = s[d] + f[d] - 1;
            else
                o[d] = d < rank ? s[d] + f[d] - 1 : s[d];
        }
    } else {
        o = s;
        if (kind == BATCH_RHS)
            for (int i = rank; i < AF_MAX_DIMS; ++i) o[i] = f[i];
    }
    return o;
}

// cpu::convolve<T, accT>, src/backend/cpu/convolve.cpp:38-66
static af_array convolve(const ArrayImpl &sig, const ArrayImpl &filt, BatchKind kind, int rank,
                         bool expand) {
    const af_dtype at = acc_type(sig.type);
    auto fbuf = cast_buffer(filt, at);
    const dim4 od = conv_out_dims(sig.dims, filt.dims, kind, rank, expand);
    NewHandle out(make_handle(sig.type, od));
    if (od.elements() > 0) {
        check_k(afhip_k_convolve(sig.type, out.data(), od.d, sig.data(), sig.dims.d, fbuf->ptr,
                                 filt.dims.d, (int)kind, rank, expand ? 1 : 0, stream()),
                "convolve");
    }
    return out.release();
}

// cpu::convolve2<T, accT> (separable), src/backend/cpu/convolve.cpp:68-99
static af_array convolve2_sep(const ArrayImpl &sig, const std::shared_ptr<Buffer> &cf, dim_t cflen,
                              const std::shared_ptr<Buffer> &rf, dim_t rflen, bool expand) {
    dim4 od = sig.dims;
    dim4 td = sig.dims;
    if (expand) {
        od[0] += cflen - 1;
        od[1] += rflen - 1;
        td[0] += cflen - 1;
    }
    NewHandle out(make_handle(sig.type, od));
    if (od.elements() == 0) return out.release();
    // FP32 with filters of up to AFHIP_SEP_FUSED_MAX_TAPS runs fused (the
    // intermediate never leaves the chip); everything else runs the two
    // passes through the reference's T-typed temp.
    std::shared_ptr<Buffer> tmp;
    const bool fused = sig.type == f32 && cflen <= AFHIP_SEP_FUSED_MAX_TAPS &&
                       rflen <= AFHIP_SEP_FUSED_MAX_TAPS;
    if (!fused) tmp = scratch((size_t)td.elements() * dtype_size(sig.type));
    check_k(afhip_k_convolve2_sep(sig.type, out.data(), od.d, sig.data(), sig.dims.d, cf->ptr,
                                  cflen, rf->ptr, rflen, expand ? 1 : 0, tmp ? tmp->ptr : nullptr,
                                  stream()),
            "convolve2_sep");
    return out.release();
}

// convolve2_unwrap output size, src/backend/cpu/convolve.cpp:130-135
// (C++ integer division truncates toward zero).
static dim_t nn_out(dim_t s, dim_t f, dim_t st, dim_t pd, dim_t dl) {
    return 1 + (s + 2 * pd - (((f - 1) * dl) + 1)) / st;
}

struct NNGeom {
    dim4 x, w, y;
    long long st[2], pd[2], dl[2];
};

static NNGeom nn_geom(const dim4 &x, const dim4 &w, const dim4 &stride, const dim4 &padding,
                      const dim4 &dilation) {
    NNGeom g;
    g.x = x;
    g.w = w;
    for (int i = 0; i < 2; ++i) {
        g.st[i] = stride[i];
        g.pd[i] = padding[i];
        g.dl[i] = dilation[i];
    }
    const dim_t ow = nn_out(x[0], w[0], stride[0], padding[0], dilation[0]);
    const dim_t oh = nn_out(x[1], w[1], stride[1], padding[1], dilation[1]);
    g.y = dim4(ow, oh, w[3], x[3]);
    return g;
}

static void check_nn_geom(const NNGeom &g) {
    // The kernels index with 32-bit extents and need positive sizes.
    for (int i = 0; i < 4; ++i) {
        DIM_ASSERT(1, g.x[i] >= 0 && g.x[i] < (1LL << 31));
        DIM_ASSERT(2, g.w[i] >= 0 && g.w[i] < (1LL << 31));
    }
    DIM_ASSERT(1, g.y[0] >= 0 && g.y[1] >= 0);
    ARG_ASSERT(3, g.st[0] > 0 && g.st[1] > 0);
    ARG_ASSERT(7, g.dl[0] > 0 && g.dl[1] > 0);
}

// f64 inputs run the direct FP64 kernels (afhip_k_conv2_nn_f64); the MFMA
// implicit-GEMM kernels take FP32 and 16-bit operands only.

// The patch-staged kernels serve the forward pass and the stride-1 data
// gradient when the input patch fits in LDS, the row-staged kernel the filter
// gradient of 16-bit operands; AFHIP_NN_GATHER=1 forces the gather kernels
// (the tests use it to cover both paths).
static int device_cus() {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
        return 256;
    return n;
}

// ---- NN variant choice. Layers whose patch kernels would leave most CUs
// idle, and every layer the patch kernels do not serve, run the gather GEMM
// with a split reduction; the autotuner (autotune.cpp) times the candidate
// variants (the patch kernels where they serve; split counts 1, 2, 4, 8, 16)
// on the first call of a geometry and caches the fastest under a key of
// (pass, dtypes, precision, dims, stride, padding, dilation). Under stream
// capture only cached choices are used. AFHIP_NN_AUTOTUNE=0 (or an explicit
// AFHIP_NN_SPLIT) keeps the built-in rules.
static bool nn_autotune_on() {
    const char *a = std::getenv("AFHIP_NN_AUTOTUNE");
    const char *sp = std::getenv("AFHIP_NN_SPLIT");
    return !(a && a[0] == '0') && !(sp && sp[0] >= '1' && sp[0] <= '9');
}

static bool stream_capturing() {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(stream(), &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}

// The key names the arithmetic (the MMA operand mode), not the storage
// types, so that calls computing the same products (f32 arrays at BF16
// precision and bf16 arrays; f32 and 16-bit outputs) run the same variant and
// give the same sums.
static std::string nn_key(const char *tag, int pass, const ArrayImpl &in, af_dtype, int precision,
                          const NNGeom &g) {
    const int mode = precision != AFHIP_PRECISION_DEFAULT ? precision
                     : in.type == f32                     ? (int)AFHIP_PRECISION_F32
                     : in.type == afhip_bf16              ? (int)AFHIP_PRECISION_BF16
                                                          : (int)AFHIP_PRECISION_F16;
    // the storage type too: f32 arrays at a 16-bit precision run other
    // kernels (element gather) than 16-bit arrays (channel-last copies)
    std::string k = std::string(tag) + ":p" + std::to_string(pass) + ":m" + std::to_string(mode) + ":t" +
                    std::to_string((int)in.type);
    for (int i = 0; i < 4; ++i) k += (i ? "," : ":x") + std::to_string(g.x[i]);
    for (int i = 0; i < 4; ++i) k += (i ? "," : ":w") + std::to_string(g.w[i]);
    k += ":s" + std::to_string(g.st[0]) + "," + std::to_string(g.st[1]) + ":pd" + std::to_string(g.pd[0]) + "," +
         std::to_string(g.pd[1]) + ":d" + std::to_string(g.dl[0]) + "," + std::to_string(g.dl[1]);
    return k;
}

// Index of the variant to run: the autotuner's (timing them now if the key
// is new), or -1 for the built-in rule.
static int nn_pick(const std::string &key, const std::vector<std::function<void()>> &cands) {
    if (cands.size() < 2 || !nn_autotune_on()) return -1;
    if (stream_capturing()) {
        const int r = autotune_lookup(key);
        return r >= 0 && r < (int)cands.size() ? r : -1;
    }
    return autotune_pick(key, (int)cands.size(), [&](int c) { cands[c](); });
}

static const int kSplitCands[] = {1, 2, 4, 8, 16};

static void run_gemm_split(int pass, const ArrayImpl &in, const ArrayImpl &w, af_dtype out_type, int precision,
                           void *out, const NNGeom &g, int nsplit) {
    const long long need = afhip_k_conv2_nn_gemm_split_workspace(pass, in.type, precision, g.x.d, g.w.d, g.y.d,
                                                                 g.st, g.pd, g.dl, nsplit);
    std::shared_ptr<Buffer> ws;
    if (need > 0) ws = scratch((size_t)need);
    check_k(afhip_k_conv2_nn_gemm_split(pass, in.type, out_type, precision, out, in.data(), w.data(), g.x.d,
                                        g.w.d, g.y.d, g.st, g.pd, g.dl, nsplit, ws ? ws->ptr : nullptr, need,
                                        stream()),
            pass == 0 ? "convolve2_nn" : "convolve2_gradient_nn");
}

static void run_patch_kernels(int pass, const ArrayImpl &in, const ArrayImpl &w, af_dtype out_type, int precision,
                              void *out, const NNGeom &g, long long need) {
    auto ws = scratch((size_t)need);
    check_k(afhip_k_conv2_nn_patch(pass, in.type, out_type, precision, out, in.data(), w.data(), g.x.d, g.w.d,
                                   g.y.d, g.st, g.pd, g.dl, ws->ptr, need, stream()),
            pass == 0 ? "convolve2_nn" : "convolve2_gradient_nn");
}

// Candidate list: [patch kernels (when they serve)] + gather GEMM splits.
static std::vector<std::function<void()>> nn_candidates(int pass, const ArrayImpl &in, const ArrayImpl &w,
                                                        af_dtype out_type, int precision, void *out,
                                                        const NNGeom &g, long long patch_need) {
    std::vector<std::function<void()>> c;
    if (patch_need > 0)
        c.push_back([=, &in, &w, &g] { run_patch_kernels(pass, in, w, out_type, precision, out, g, patch_need); });
    for (int sp : kSplitCands)
        c.push_back([=, &in, &w, &g] { run_gemm_split(pass, in, w, out_type, precision, out, g, sp); });
    return c;
}

// The patch kernels for layers that fill the chip with them; otherwise the
// autotuned choice between them and the split gather GEMM (built-in rule:
// the GEMM when the patch kernels launch fewer workgroups than a quarter of
// the CUs). AFHIP_NN_SMALL=0 keeps every served layer on the patch kernels.
static bool run_patch(int pass, const ArrayImpl &in, const ArrayImpl &w, af_dtype out_type,
                      int precision, void *out, const NNGeom &g) {
    const char *env = std::getenv("AFHIP_NN_GATHER");
    if (env && env[0] == '1') return false;
    const long long need = afhip_k_conv2_nn_patch_workspace(pass, in.type, precision, g.x.d, g.w.d,
                                                            g.y.d, g.st, g.pd, g.dl);
    if (need <= 0) return false;
    const char *sm = std::getenv("AFHIP_NN_SMALL");
    const long long tiles = afhip_k_conv2_nn_patch_tiles(pass, in.type, precision, g.x.d, g.w.d, g.y.d, g.st,
                                                         g.pd, g.dl);
    const int cus = device_cus();
    if ((sm && sm[0] == '0') || tiles >= cus) {
        run_patch_kernels(pass, in, w, out_type, precision, out, g, need);
        return true;
    }
    const auto cands = nn_candidates(pass, in, w, out_type, precision, out, g, need);
    int c = nn_pick(nn_key("nn", pass, in, out_type, precision, g), cands);
    if (c < 0) c = (tiles > 0 && tiles * 4 >= cus) ? 0 : -1;  // the rule
    if (c >= 0)
        cands[c]();
    else
        run_gemm_split(pass, in, w, out_type, precision, out, g, 0);
    return true;
}

// The gather GEMM for every geometry, its split count autotuned (built-in
// rule: split when that fills the chip).
static void run_gemm(int pass, const ArrayImpl &in, const ArrayImpl &w, af_dtype out_type, int precision,
                     void *out, const NNGeom &g) {
    const auto cands = nn_candidates(pass, in, w, out_type, precision, out, g, 0);
    const int c = nn_pick(nn_key("nn_gemm", pass, in, out_type, precision, g), cands);
    if (c >= 0)
        cands[c]();
    else
        run_gemm_split(pass, in, w, out_type, precision, out, g, 0);
}

static af_dtype nn_out_type(af_dtype in, af_dtype requested) {
    if ((int)requested == -1) return in;
    return requested;
}

// cpu::convolve2<T> (NN forward), src/backend/cpu/convolve.cpp:160-167
static af_array conv2_nn(const ArrayImpl &x, const ArrayImpl &w, const NNGeom &g, int precision,
                         af_dtype out_type) {
    NewHandle out(make_handle(out_type, g.y));
    if (g.y.elements() == 0) return out.release();
    if (x.type == f64) {
        check_k(afhip_k_conv2_nn_f64(0, out.data(), x.data(), w.data(), g.x.d, g.w.d, g.y.d, g.st,
                                     g.pd, g.dl, stream()),
                "convolve2_nn");
        return out.release();
    }
    if (run_patch(0, x, w, out_type, precision, out.data(), g)) return out.release();
    run_gemm(0, x, w, out_type, precision, out.data(), g);
    return out.release();
}

// cpu::conv2DataGradient<T>, src/backend/cpu/convolve.cpp:179-211
static af_array conv2_data_grad(const ArrayImpl &dy, const ArrayImpl &x, const ArrayImpl &w,
                                const NNGeom &g, int precision, af_dtype out_type) {
    NewHandle out(make_handle(out_type, g.x));
    if (g.x.elements() == 0) return out.release();
    void *o = out.data();
    if (dy.elements() == 0 || g.w.elements() == 0) {
        check_k(afhip_k_fill(out_type, o, 0.0, g.x.elements(), stream()), "fill");
        return out.release();
    }
    if (x.type == f64) {
        check_k(afhip_k_conv2_nn_f64(1, o, dy.data(), w.data(), g.x.d, g.w.d, g.y.d, g.st, g.pd, g.dl,
                                     stream()),
                "convolve2_gradient_nn");
        return out.release();
    }
    if (run_patch(1, dy, w, out_type, precision, o, g)) return out.release();
    // strided layers: the polyphase split into stride-1 convolutions on the
    // patch kernels when its phases fill the chip, else the split gather
    // GEMM over phase-ordered pixels (AFHIP_NN_GATHER=1 keeps the gather
    // kernel)
    const char *gather = std::getenv("AFHIP_NN_GATHER");
    if (!(gather && gather[0] == '1')) {
        const long long pb = afhip_k_conv2_nn_dgrad_poly_workspace(x.type, out_type, precision, g.x.d, g.w.d,
                                                                  g.y.d, g.st, g.pd, g.dl);
        if (pb > 0) {
            auto ws = scratch((size_t)pb);
            check_k(afhip_k_conv2_nn_dgrad_poly(x.type, out_type, precision, o, dy.data(), w.data(), g.x.d, g.w.d,
                                                g.y.d, g.st, g.pd, g.dl, ws->ptr, pb, stream()),
                    "convolve2_gradient_nn");
            return out.release();
        }
    }
    run_gemm(1, dy, w, out_type, precision, o, g);
    return out.release();
}

// cpu::conv2FilterGradient<T>, src/backend/cpu/convolve.cpp:213-243
static af_array conv2_filter_grad(const ArrayImpl &dy, const ArrayImpl &x, const ArrayImpl &w,
                                  const NNGeom &g, int precision, af_dtype out_type) {
    NewHandle out(make_handle(out_type, g.w));
    if (g.w.elements() == 0) return out.release();
    void *o = out.data();
    if (dy.elements() == 0 || g.x.elements() == 0) {
        check_k(afhip_k_fill(out_type, o, 0.0, g.w.elements(), stream()), "fill");
        return out.release();
    }
    if (x.type == f64) {
        check_k(afhip_k_conv2_nn_f64(2, o, x.data(), dy.data(), g.x.d, g.w.d, g.y.d, g.st, g.pd, g.dl,
                                     stream()),
                "convolve2_gradient_nn");
        return out.release();
    }
    const char *env = std::getenv("AFHIP_NN_GATHER");
    if (!(env && env[0] == '1')) {
        const long long rb = afhip_k_conv2_nn_wgrad_rows_workspace(x.type, precision, g.x.d, g.w.d, g.y.d,
                                                                   g.st, g.pd, g.dl);
        if (rb > 0) {
            auto ws = scratch((size_t)rb);
            check_k(afhip_k_conv2_nn_wgrad_rows(x.type, out_type, precision, o, x.data(), dy.data(), g.x.d,
                                                g.w.d, g.y.d, g.st, g.pd, g.dl, ws->ptr, rb, stream()),
                    "convolve2_gradient_nn");
            return out.release();
        }
    }
    const long long wsb = afhip_k_conv2_nn_wgrad_workspace(g.x.d, g.w.d, g.y.d, g.st, g.pd, g.dl);
    std::shared_ptr<Buffer> ws;
    if (wsb > 0) ws = scratch((size_t)wsb);
    check_k(afhip_k_conv2_nn_wgrad(x.type, out_type, precision, o, x.data(), dy.data(), g.x.d, g.w.d,
                                   g.y.d, g.st, g.pd, g.dl, ws ? ws->ptr : nullptr, wsb, stream()),
            "convolve2_gradient_nn");
    return out.release();
}

// Frequency-domain convolution (src/backend/cpu/fftconvolve.cpp:42-203 and
// the fallback of src/api/c/fftconvolve.cpp:52-121). The batched path takes
// the filter converted to the signal's type (castArray<T>: fractional taps on
// an integer signal truncate); the fallback (DIFF batches, complex signals)
// converts it straight to the complex compute type.
static af_array fft_convolve(const ArrayImpl &sig, const ArrayImpl &filt, BatchKind kind, int rank,
                             bool expand) {
    const bool fallback = kind == BATCH_DIFF || sig.type == c32 || sig.type == c64;
    std::shared_ptr<Buffer> fb;
    af_dtype ft;
    if (fallback) {
        ft = (filt.type == f16 || filt.type == afhip_bf16) ? f32 : filt.type;
        fb = cast_buffer(filt, ft);
    } else {
        ft = sig.type;
        fb = cast_buffer(filt, ft);
    }
    dim4 od;
    if (kind == BATCH_DIFF) {
        // base dims as the batched kinds; dims past the rank: max(s, f)
        // (odims of src/api/c/fftconvolve.cpp:72-97)
        for (int d = 0; d < AF_MAX_DIMS; ++d)
            od[d] = d < rank ? (expand ? sig.dims[d] + filt.dims[d] - 1 : sig.dims[d])
                             : std::max(sig.dims[d], filt.dims[d]);
    } else {
        od = conv_out_dims(sig.dims, filt.dims, kind, rank, expand);
    }
    NewHandle out(make_handle(sig.type, od));
    if (od.elements() == 0) return out.release();
    const long long need =
        afhip_k_fft_convolve_workspace(sig.type, sig.dims.d, filt.dims.d, od.d, (int)kind, rank);
    if (need < 0) AFHIP_THROW(AF_ERR_SIZE, "fft_convolve: a padded axis exceeds 2^24 elements");
    auto ws = scratch((size_t)need);
    check_k(afhip_k_fft_convolve(sig.type, out.data(), od.d, sig.data(), sig.dims.d, ft, fb->ptr,
                                 filt.dims.d, (int)kind, rank, expand ? 1 : 0, ws->ptr, need, stream()),
            "fft_convolve");
    return out.release();
}

// Scalar/scalar separable filters on a complex signal (src/api/c/convolve.cpp:57-68
// with accT = cfloat / cdouble): y = x * (c * r), the filter product formed in
// accT with each real product rounded before the sum, as std::complex does.
template <typename R>
static void cmul(const R *a, const R *b, R *o) {
#pragma clang fp contract(off)
    o[0] = a[0] * b[0] - a[1] * b[1];
    o[1] = a[0] * b[1] + a[1] * b[0];
}

static af_array scale_complex(const ArrayImpl &sig, const std::shared_ptr<Buffer> &cb,
                              const std::shared_ptr<Buffer> &rb, af_dtype at) {
    double re, im;
    if (at == c32) {
        float c[2], r[2], p[2];
        check_hip(hipMemcpyAsync(c, cb->ptr, 8, hipMemcpyDeviceToHost, stream()), "copy");
        check_hip(hipMemcpyAsync(r, rb->ptr, 8, hipMemcpyDeviceToHost, stream()), "copy");
        sync_stream();
        cmul(c, r, p);
        re = p[0];
        im = p[1];
    } else {
        double c[2], r[2], p[2];
        check_hip(hipMemcpyAsync(c, cb->ptr, 16, hipMemcpyDeviceToHost, stream()), "copy");
        check_hip(hipMemcpyAsync(r, rb->ptr, 16, hipMemcpyDeviceToHost, stream()), "copy");
        sync_stream();
        cmul(c, r, p);
        re = p[0];
        im = p[1];
    }
    NewHandle o(make_handle(sig.type, sig.dims));
    check_k(afhip_k_scale_complex(sig.type, o.data(), sig.data(), re, im, sig.elements(), stream()),
            "convolve2_sep");
    return o.release();
}

}  // namespace detail

// ---------------------------------------------------------------- C API
// src/api/c/convolve.cpp:132-215
static af_err convolve_common(af_array *out, const af_array signal, const af_array filter,
                              const af_conv_mode mode, const int rank) {
    try {
        const ArrayImpl &s = get(signal);
        const ArrayImpl &f = get(filter);
        ARG_ASSERT(0, out != nullptr);
        if (f.dims.ndims() == 0 || s.dims.ndims() == 0) return af_retain_array(out, signal);
        const BatchKind kind = identify_batch_kind(rank, s.dims, f.dims);
        ARG_ASSERT(1, kind != BATCH_UNSUPPORTED && kind != BATCH_DIFF);
        const bool expand = mode == AF_CONV_EXPAND;
        (void)acc_type(s.type);  // TYPE_ERROR(1, stype) for f16 / bf16
        af_array o = detail::convolve(s, f, kind, rank, expand);
        *out = o;
    }
    CATCHALL;
    return AF_SUCCESS;
}

// src/api/c/fftconvolve.cpp:162-219: batch kind, dtype dispatch (every
// dtype but f16), then the batched path or, for DIFF batches and complex
// signals, the fallback (:52-121).
static af_err fft_convolve_common(af_array *out, const af_array signal, const af_array filter,
                                  const bool expand, const int rank) {
    try {
        const ArrayImpl &s = get(signal);
        const ArrayImpl &f = get(filter);
        ARG_ASSERT(0, out != nullptr);
        const BatchKind kind = identify_batch_kind(rank, s.dims, f.dims);
        ARG_ASSERT(1, kind != BATCH_UNSUPPORTED);
        (void)acc_type(s.type);  // TYPE_ERROR(1, stype) for f16 / bf16
        *out = detail::fft_convolve(s, f, kind, rank, expand);
    }
    CATCHALL;
    return AF_SUCCESS;
}

}  // namespace afhip

using namespace afhip;

extern "C" {

// src/api/c/fftconvolve.cpp:247-264
AFAPI af_err af_fft_convolve1(af_array *out, const af_array signal, const af_array filter,
                              const af_conv_mode mode) {
    return fft_convolve_common(out, signal, filter, mode == AF_CONV_EXPAND, 1);
}
AFAPI af_err af_fft_convolve2(af_array *out, const af_array signal, const af_array filter,
                              const af_conv_mode mode) {
    try {
        if (get(signal).dims.ndims() < 2 && get(filter).dims.ndims() < 2)
            return fft_convolve_common(out, signal, filter, mode == AF_CONV_EXPAND, 1);
        return fft_convolve_common(out, signal, filter, mode == AF_CONV_EXPAND, 2);
    }
    CATCHALL;
}
AFAPI af_err af_fft_convolve3(af_array *out, const af_array signal, const af_array filter,
                              const af_conv_mode mode) {
    try {
        if (get(signal).dims.ndims() < 3 && get(filter).dims.ndims() < 3)
            return fft_convolve_common(out, signal, filter, mode == AF_CONV_EXPAND, 2);
        return fft_convolve_common(out, signal, filter, mode == AF_CONV_EXPAND, 3);
    }
    CATCHALL;
}

// src/api/c/convolve.cpp:217-226
AFAPI af_err af_convolve1(af_array *out, const af_array signal, const af_array filter,
                          const af_conv_mode mode, af_conv_domain domain) {
    try {
        if (is_freq_domain(1, get(signal).dims, get(filter).dims, domain))
            return af_fft_convolve1(out, signal, filter, mode);
        return convolve_common(out, signal, filter, mode, 1);
    }
    CATCHALL;
}

// src/api/c/convolve.cpp:228-241
AFAPI af_err af_convolve2(af_array *out, const af_array signal, const af_array filter,
                          const af_conv_mode mode, af_conv_domain domain) {
    try {
        if (get(signal).dims.ndims() < 2 || get(filter).dims.ndims() < 2)
            return af_convolve1(out, signal, filter, mode, domain);
        if (is_freq_domain(2, get(signal).dims, get(filter).dims, domain))
            return af_fft_convolve2(out, signal, filter, mode);
        return convolve_common(out, signal, filter, mode, 2);
    }
    CATCHALL;
}

// src/api/c/convolve.cpp:243-256
AFAPI af_err af_convolve3(af_array *out, const af_array signal, const af_array filter,
                          const af_conv_mode mode, af_conv_domain domain) {
    try {
        if (get(signal).dims.ndims() < 3 || get(filter).dims.ndims() < 3)
            return af_convolve2(out, signal, filter, mode, domain);
        if (is_freq_domain(3, get(signal).dims, get(filter).dims, domain))
            return af_fft_convolve3(out, signal, filter, mode);
        return convolve_common(out, signal, filter, mode, 3);
    }
    CATCHALL;
}

// isVector, src/backend/common/ArrayInfo.cpp:122-130
static bool is_vector(const dim4 &d) {
    int singular = 0, nonsingular = 0;
    for (int i = 0; i < AF_MAX_DIMS; ++i) {
        nonsingular += (d[i] != 0 && d[i] != 1);
        singular += (d[i] == 1);
    }
    return singular == AF_MAX_DIMS - 1 && nonsingular == 1;
}

// src/api/c/convolve.cpp:258-334 and :50-75
AFAPI af_err af_convolve2_sep(af_array *out, const af_array col_filter, const af_array row_filter,
                              const af_array signal, const af_conv_mode mode) {
    try {
        const ArrayImpl &s = get(signal);
        ARG_ASSERT(1, s.dims.ndims() >= 2);
        ARG_ASSERT(0, out != nullptr);
        const af_dtype at = acc_type(s.type);
        const bool expand = mode == AF_CONV_EXPAND;
        const ArrayImpl &cf = get(col_filter);
        const ArrayImpl &rf = get(row_filter);
        auto cb = cast_buffer(cf, at);
        auto rb = cast_buffer(rf, at);
        if (cf.elements() == 1 && rf.elements() == 1) {
            // scalar/scalar: y = T(accT(x) * (c * r))
            if (at == c32 || at == c64) {
                *out = detail::scale_complex(s, cb, rb, at);
                return AF_SUCCESS;
            }
            if (at != f32 && at != f64) TYPE_ERROR(1, s.type);
            double scale;
            if (at == f32) {
                float c = 0, r = 0;
                check_hip(hipMemcpyAsync(&c, cb->ptr, 4, hipMemcpyDeviceToHost, stream()), "copy");
                check_hip(hipMemcpyAsync(&r, rb->ptr, 4, hipMemcpyDeviceToHost, stream()), "copy");
                sync_stream();
                scale = (double)(c * r);
            } else {
                double c = 0, r = 0;
                check_hip(hipMemcpyAsync(&c, cb->ptr, 8, hipMemcpyDeviceToHost, stream()), "copy");
                check_hip(hipMemcpyAsync(&r, rb->ptr, 8, hipMemcpyDeviceToHost, stream()), "copy");
                sync_stream();
                scale = c * r;
            }
            NewHandle o(make_handle(s.type, s.dims));
            check_k(afhip_k_scale(s.type, o.data(), s.data(), scale, s.elements(), stream()),
                    "convolve2_sep");
            *out = o.release();
            return AF_SUCCESS;
        }
        ARG_ASSERT(2, is_vector(cf.dims));
        ARG_ASSERT(3, is_vector(rf.dims));
        *out = detail::convolve2_sep(s, cb, cf.elements(), rb, rf.elements(), expand);
    }
    CATCHALL;
    return AF_SUCCESS;
}

static af_err nn_forward(af_array *out, const af_array signal, const af_array filter,
                         const unsigned stride_dims, const dim_t *strides,
                         const unsigned padding_dims, const dim_t *paddings,
                         const unsigned dilation_dims, const dim_t *dilations, int precision,
                         af_dtype out_type) {
    try {
        const ArrayImpl &s = get(signal);
        const ArrayImpl &f = get(filter);
        ARG_ASSERT(0, out != nullptr);
        dim4 stride(stride_dims, strides), padding(padding_dims, paddings),
            dilation(dilation_dims, dilations);
        const dim_t sn = stride.ndims(), pn = padding.ndims(), dn = dilation.ndims();
        ARG_ASSERT(3, sn > 0 && sn <= 2);
        ARG_ASSERT(5, pn >= 0 && pn <= 2);
        ARG_ASSERT(7, dn > 0 && dn <= 2);
        DIM_ASSERT(1, s.dims[2] == f.dims[2]);
        const af_dtype t = s.type;
        if (t != f32 && t != f64 && t != f16 && t != afhip_bf16) TYPE_ERROR(1, t);
        if (f.type != t) AFHIP_THROW(AF_ERR_INTERNAL, "Invalid type for input array.");
        const af_dtype ot = detail::nn_out_type(t, out_type);
        if (t == f64) {
            ARG_ASSERT(9, precision == AFHIP_PRECISION_DEFAULT);
            if (ot != f64) TYPE_ERROR(10, ot);
        }
        auto g = detail::nn_geom(s.dims, f.dims, stride, padding, dilation);
        detail::check_nn_geom(g);
        *out = detail::conv2_nn(s, f, g, precision, ot);
    }
    CATCHALL;
    return AF_SUCCESS;
}

// src/api/c/convolve.cpp:344-392
AFAPI af_err af_convolve2_nn(af_array *out, const af_array signal, const af_array filter,
                             const unsigned stride_dims, const dim_t *strides,
                             const unsigned padding_dims, const dim_t *paddings,
                             const unsigned dilation_dims, const dim_t *dilations) {
    return nn_forward(out, signal, filter, stride_dims, strides, padding_dims, paddings, dilation_dims,
                      dilations, AFHIP_PRECISION_DEFAULT, (af_dtype)-1);
}

AFAPI af_err afhip_convolve2_nn_ex(af_array *out, const af_array signal, const af_array filter,
                                   const unsigned stride_dims, const dim_t *strides,
                                   const unsigned padding_dims, const dim_t *paddings,
                                   const unsigned dilation_dims, const dim_t *dilations,
                                   afhip_precision precision, af_dtype out_type) {
    return nn_forward(out, signal, filter, stride_dims, strides, padding_dims, paddings, dilation_dims,
                      dilations, (int)precision, out_type);
}

static af_err nn_gradient(af_array *out, const af_array incoming_gradient,
                          const af_array original_signal, const af_array original_filter,
                          const af_array convolved_output, const unsigned stride_dims,
                          const dim_t *strides, const unsigned padding_dims, const dim_t *paddings,
                          const unsigned dilation_dims, const dim_t *dilations,
                          af_conv_gradient_type grad_type, int precision, af_dtype out_type,
                          bool strict_types) {
    try {
        const ArrayImpl &g = get(incoming_gradient);
        const ArrayImpl &s = get(original_signal);
        const ArrayImpl &f = get(original_filter);
        const ArrayImpl &o = get(convolved_output);
        ARG_ASSERT(0, out != nullptr);
        // src/api/c/convolve.cpp:434-437
        DIM_ASSERT(1, g.dims == o.dims);
        DIM_ASSERT(3, o.dims[2] == f.dims[3]);
        DIM_ASSERT(3, o.dims[3] == s.dims[3]);
        DIM_ASSERT(2, s.dims[2] == f.dims[2]);
        dim4 stride(stride_dims, strides), padding(padding_dims, paddings),
            dilation(dilation_dims, dilations);
        const dim_t sn = stride.ndims(), pn = padding.ndims(), dn = dilation.ndims();
        ARG_ASSERT(3, sn > 0 && sn <= 2);
        ARG_ASSERT(5, pn >= 0 && pn <= 2);
        ARG_ASSERT(7, dn > 0 && dn <= 2);
        // The reference takes T from convolved_output and requires all four
        // arrays to be T (src/api/c/convolve.cpp:452, handle.hpp:53-54). The
        // extension entry point reads only the dims of convolved_output (it
        // may hold the FP32 output of a BF16 forward), so T comes from the
        // incoming gradient there.
        const af_dtype t = strict_types ? o.type : g.type;
        if (t != f32 && t != f64 && t != f16 && t != afhip_bf16) TYPE_ERROR(1, t);
        if (g.type != t || s.type != t || f.type != t || (strict_types && o.type != t))
            AFHIP_THROW(AF_ERR_INTERNAL, "Invalid type for input array.");
        const af_dtype ot = detail::nn_out_type(t, out_type);
        if (t == f64) {
            ARG_ASSERT(12, precision == AFHIP_PRECISION_DEFAULT);
            if (ot != f64) TYPE_ERROR(13, ot);
        }
        auto geo = detail::nn_geom(s.dims, f.dims, stride, padding, dilation);
        detail::check_nn_geom(geo);
        // The incoming gradient must have the forward output's shape for the
        // kernels' indexing (the reference reads it through the same dims).
        DIM_ASSERT(1, g.dims == geo.y);
        if (grad_type == AF_CONV_GRADIENT_FILTER)
            *out = detail::conv2_filter_grad(g, s, f, geo, precision, ot);
        else
            *out = detail::conv2_data_grad(g, s, f, geo, precision, ot);
    }
    CATCHALL;
    return AF_SUCCESS;
}

// src/api/c/convolve.cpp:394-477
AFAPI af_err af_convolve2_gradient_nn(af_array *out, const af_array incoming_gradient,
                                      const af_array original_signal, const af_array original_filter,
                                      const af_array convolved_output, const unsigned stride_dims,
                                      const dim_t *strides, const unsigned padding_dims,
                                      const dim_t *paddings, const unsigned dilation_dims,
                                      const dim_t *dilations, af_conv_gradient_type grad_type) {
    return nn_gradient(out, incoming_gradient, original_signal, original_filter, convolved_output,
                       stride_dims, strides, padding_dims, paddings, dilation_dims, dilations,
                       grad_type, AFHIP_PRECISION_DEFAULT, (af_dtype)-1, true);
}

AFAPI af_err afhip_convolve2_gradient_nn_ex(
    af_array *out, const af_array incoming_gradient, const af_array original_signal,
    const af_array original_filter, const af_array convolved_output, const unsigned stride_dims,
    const dim_t *strides, const unsigned padding_dims, const dim_t *paddings,
    const unsigned dilation_dims, const dim_t *dilations, af_conv_gradient_type grad_type,
    afhip_precision precision, af_dtype out_type) {
    return nn_gradient(out, incoming_gradient, original_signal, original_filter, convolved_output,
                       stride_dims, strides, padding_dims, paddings, dilation_dims, dilations,
                       grad_type, (int)precision, out_type, false);
}

}  // extern "C"

struct afhip_nn_step_plan_t {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<void *> kept;                            // workspaces the graph uses
    std::vector<std::shared_ptr<afhip::Buffer>> arrays;  // x, w, dy, y, dx, dw
    ~afhip_nn_step_plan_t() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (void *p : kept) afhip::pool_free(p);
    }
};

extern "C" {

// The three passes of one step through the public entry points (checks,
// dispatch, allocation); the handles are released on failure.
static af_err nn_step_passes(af_array *y, af_array *dx, af_array *dw, const af_array x, const af_array w,
                             const af_array dy, const unsigned sdims, const dim_t *strides, const unsigned pdims,
                             const dim_t *paddings, const unsigned ddims, const dim_t *dilations, int precision,
                             af_dtype out_type) {
    *y = *dx = *dw = nullptr;
    af_err e = nn_forward(y, x, w, sdims, strides, pdims, paddings, ddims, dilations, precision, out_type);
    if (e == AF_SUCCESS)
        e = nn_gradient(dx, dy, x, w, *y, sdims, strides, pdims, paddings, ddims, dilations,
                        AF_CONV_GRADIENT_DATA, precision, out_type, false);
    if (e == AF_SUCCESS)
        e = nn_gradient(dw, dy, x, w, *y, sdims, strides, pdims, paddings, ddims, dilations,
                        AF_CONV_GRADIENT_FILTER, precision, out_type, false);
    if (e != AF_SUCCESS) {
        for (af_array *h : {y, dx, dw})
            if (*h) {
                destroy_handle(*h);
                *h = nullptr;
            }
    }
    return e;
}

AFAPI af_err afhip_nn_step_plan_create(afhip_nn_step_plan *plan, af_array *y, af_array *dx, af_array *dw,
                                       const af_array x, const af_array w, const af_array dy,
                                       const unsigned stride_dims, const dim_t *strides,
                                       const unsigned padding_dims, const dim_t *paddings,
                                       const unsigned dilation_dims, const dim_t *dilations,
                                       afhip_precision precision, af_dtype out_type) {
    try {
        ARG_ASSERT(0, plan != nullptr && y != nullptr && dx != nullptr && dw != nullptr);
        // eager run: argument checks, first-use kernel setup, pool warm-up
        af_array e[3];
        af_err rc = nn_step_passes(&e[0], &e[1], &e[2], x, w, dy, stride_dims, strides, padding_dims, paddings,
                                   dilation_dims, dilations, (int)precision, out_type);
        if (rc != AF_SUCCESS) return rc;
        for (af_array h : e) destroy_handle(h);
        sync_stream();

        auto p = std::make_unique<afhip_nn_step_plan_t>();
        af_array h[3] = {nullptr, nullptr, nullptr};
        hipStream_t s = stream();
        check_hip(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
        pool_keep(&p->kept);
        rc = nn_step_passes(&h[0], &h[1], &h[2], x, w, dy, stride_dims, strides, padding_dims, paddings,
                            dilation_dims, dilations, (int)precision, out_type);
        pool_keep(nullptr);
        const hipError_t ce = hipStreamEndCapture(s, &p->graph);
        if (rc != AF_SUCCESS || ce != hipSuccess) {
            for (af_array a : h)
                if (a) destroy_handle(a);
            if (rc != AF_SUCCESS) return rc;
            check_hip(ce, "hipStreamEndCapture");
        }
        const hipError_t ie = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
            for (af_array a : h) destroy_handle(a);
            check_hip(ie, "hipGraphInstantiate");
        }
        for (af_array a : {x, w, dy}) p->arrays.push_back(get(a).buf);
        for (af_array a : h) p->arrays.push_back(get(a).buf);
        *y = h[0];
        *dx = h[1];
        *dw = h[2];
        *plan = p.release();
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err afhip_nn_step_plan_run(afhip_nn_step_plan plan) {
    try {
        ARG_ASSERT(0, plan != nullptr);
        check_hip(hipGraphLaunch(plan->exec, stream()), "hipGraphLaunch");
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err afhip_nn_step_plan_release(afhip_nn_step_plan plan) {
    try {
        if (plan) {
            sync_stream();  // a replay in flight still uses the workspaces
            delete plan;
        }
    }
    CATCHALL;
    return AF_SUCCESS;
}

// include/af/data.h:515
AFAPI af_err af_constant(af_array *arr, const double val, const unsigned ndims,
                         const dim_t *const dims, const af_dtype type) {
    try {
        ARG_ASSERT(0, arr != nullptr);
        ARG_ASSERT(2, ndims >= 1 && ndims <= 4);
        if (!dtype_known(type)) TYPE_ERROR(4, type);
        dim4 d(ndims, dims);
        NewHandle o(make_handle(type, d));
        if (d.elements() > 0)
            check_k(afhip_k_fill(type, o.data(), val, d.elements(), stream()), "af_constant");
        *arr = o.release();
    }
    CATCHALL;
    return AF_SUCCESS;
}

// include/af/arith.h (af_cast)
AFAPI af_err af_cast(af_array *out, const af_array in, const af_dtype type) {
    try {
        ARG_ASSERT(0, out != nullptr);
        const ArrayImpl &a = get(in);
        if (!dtype_known(type)) TYPE_ERROR(2, type);
        if (a.type == type) return af_retain_array(out, in);
        NewHandle o(make_handle(type, a.dims));
        if (a.elements() > 0)
            check_k(afhip_k_cast(type, o.data(), a.type, a.data(), a.elements(), stream()),
                    "af_cast");
        *out = o.release();
    }
    CATCHALL;
    return AF_SUCCESS;
}

}  // extern "C"
