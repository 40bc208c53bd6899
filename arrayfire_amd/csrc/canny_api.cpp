// af_canny (src/api/c/canny.cpp:198-285): the 5-tap separable smoothing of
// the convolution path, then Sobel gradients, non-maximum suppression,
// thresholds (MANUAL or AUTO_OTSU) and edge tracking by hysteresis, all on
// the GPU (kernels/canny.hip). The host reads two scalars (the suppressed
// map's maximum, for the Otsu bin count) and one flag per batch of sweeps.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/af/array.h"
#include "../../include/af/image.h"
#include "../../include/afhip_kernels.h"
#include "runtime.hpp"

namespace afhip {

// cannyHelper's smoothing taps (src/api/c/canny.cpp:203-206)
static const float kCannySmooth[5] = {-0.11021f, -0.23691f, -0.30576f, -0.23691f, -0.11021f};

static af_array canny(const ArrayImpl &in, af_canny_threshold ct, float t1, float t2, bool is_fast) {
    const dim4 d = in.dims;
    const dim_t d0 = d[0], d1 = d[1], ns = d[2] * d[3], n = d.elements();
    hipStream_t s = stream();
    // cast<float, T>(in)
    auto xf = scratch(sizeof(float) * n);
    check_k(afhip_k_cast(f32, xf->ptr, in.type, in.data(), n, s), "canny");
    // convolve2<float, float>(x, cFilter, rFilter, false): the workspace
    // selects the two-pass kernel in the reference's order
    auto taps = scratch(sizeof(kCannySmooth));
    check_hip(hipMemcpyAsync(taps->ptr, kCannySmooth, sizeof(kCannySmooth), hipMemcpyHostToDevice, s), "canny");
    auto smt = scratch(sizeof(float) * n), tmp = scratch(sizeof(float) * n);
    check_k(afhip_k_convolve2_sep(f32, smt->ptr, d.d, xf->ptr, d.d, taps->ptr, 5, taps->ptr, 5, 0, tmp->ptr, s),
            "canny");
    auto gx = scratch(sizeof(float) * n), gy = scratch(sizeof(float) * n), mag = scratch(sizeof(float) * n);
    check_k(afhip_k_canny_grad((float *)gx->ptr, (float *)gy->ptr, (float *)mag->ptr, (const float *)smt->ptr, d0,
                               d1, ns, is_fast ? 1 : 0, s),
            "canny");
    auto sup = scratch(sizeof(float) * n);
    check_k(afhip_k_canny_nms((float *)sup->ptr, (const float *)mag->ptr, (const float *)gx->ptr,
                              (const float *)gy->ptr, d0, d1, ns, s),
            "canny");
    auto mm = scratch(2 * sizeof(float)), part = scratch(2048 * sizeof(float));
    check_k(afhip_k_canny_minmax((float *)mm->ptr, (const float *)sup->ptr, n, (float *)part->ptr, s), "canny");
    std::shared_ptr<Buffer> thr;
    int mode = 0;
    if (ct == AF_CANNY_THRESHOLD_AUTO_OTSU) {
        float h_mm[2];
        check_hip(hipMemcpyAsync(h_mm, mm->ptr, sizeof(h_mm), hipMemcpyDeviceToHost, s), "canny");
        sync_stream();
        const float maxv = h_mm[1];
        const unsigned nbins = static_cast<unsigned>(maxv);
        if (nbins < 1) AFHIP_THROW(AF_ERR_ARG, "canny: the gradient's maximum is below one Otsu bin");
        auto hist = scratch(sizeof(unsigned) * nbins * ns);
        check_k(afhip_k_fill(u32, hist->ptr, 0.0, (long long)nbins * ns, s), "canny");
        thr = scratch(sizeof(float) * ns);
        check_k(afhip_k_canny_otsu((float *)thr->ptr, (unsigned *)hist->ptr, (const float *)sup->ptr, d0 * d1, ns,
                                   (int)nbins, maxv, s),
                "canny");
        mode = 1;
    }
    auto strong = scratch((size_t)n), weak = scratch((size_t)n);
    check_k(afhip_k_canny_candidates((unsigned char *)strong->ptr, (unsigned char *)weak->ptr,
                                     (const float *)sup->ptr, (const float *)mm->ptr,
                                     thr ? (const float *)thr->ptr : nullptr, d0 * d1, ns, t1, t2, mode, s),
            "canny");
    NewHandle out(make_handle(b8, d));
    unsigned char *o = (unsigned char *)out.data();
    check_k(afhip_k_canny_seed(o, (const unsigned char *)strong->ptr, d0, d1, ns, s), "canny");
    // sweeps until one changes nothing: 16 per host check
    constexpr int kBatch = 16;
    auto flags = scratch(sizeof(unsigned) * kBatch);
    for (;;) {
        check_k(afhip_k_fill(u32, flags->ptr, 0.0, kBatch, s), "canny");
        for (int k = 0; k < kBatch; ++k)
            check_k(afhip_k_canny_sweep(o, (const unsigned char *)weak->ptr, d0, d1, ns, (unsigned *)flags->ptr + k,
                                        s),
                    "canny");
        unsigned last = 0;
        check_hip(hipMemcpyAsync(&last, (unsigned *)flags->ptr + kBatch - 1, sizeof(unsigned),
                                 hipMemcpyDeviceToHost, s),
                  "canny");
        sync_stream();
        if (!last) break;
    }
    return out.release();
}

}  // namespace afhip

using namespace afhip;

// src/api/c/canny.cpp:224-285
AFAPI af_err af_canny(af_array *out, const af_array in, const af_canny_threshold ct, const float t1, const float t2,
                      const unsigned sw, const bool isf) {
    try {
        ARG_ASSERT(0, out != nullptr);
        const ArrayImpl &a = get(in);
        const dim4 dims = a.dims;
        DIM_ASSERT(2, dims.ndims() >= 2);
        DIM_ASSERT(2, dims[0] >= 5 && dims[1] >= 5);
        ARG_ASSERT(5, sw == 3);
        switch (a.type) {
            case f32:
            case f64:
            case s32:
            case u32:
            case s16:
            case u16:
            case s8:
            case u8: break;
            default: TYPE_ERROR(1, a.type);
        }
        DIM_ASSERT(1, dims[0] * dims[1] < (1LL << 31));
        *out = canny(a, ct, t1, t2, isf);
    }
    CATCHALL;
    return AF_SUCCESS;
}
