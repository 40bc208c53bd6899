// af_features handles (src/api/c/features.cpp) and af_harris
// (src/api/c/harris.cpp:52-95 with src/backend/cpu/harris.cpp:26-148), a
// caller of the separable convolution: the three gradient products are
// smoothed by one batched af_convolve2_sep-style call with the window
// filter, then scored, suppressed, compacted and (for max_corners > 0)
// sorted on the GPU (kernels/harris.hip).
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/af/array.h"
#include "../../include/af/features.h"
#include "../../include/af/vision.h"
#include "../../include/afhip_kernels.h"
#include "features.hpp"
#include "runtime.hpp"

namespace afhip {

static Features &feat(const af_features f) {
    if (!f) AFHIP_THROW(AF_ERR_ARG, "Invalid features handle");
    return *static_cast<Features *>(f);
}

// The window of cpu::harris (src/backend/cpu/harris.cpp:34-44): a box
// 1/filter_len (in T) for sigma < 0.5, else gaussian1D
// (src/backend/cpu/utility.hpp:35-48): each tap evaluated in double and
// stored in T, the sum and the normalisation in T.
template <typename T>
static std::vector<T> harris_window(unsigned len, float sigma) {
    std::vector<T> h(len);
    if (sigma < 0.5f) {
        for (unsigned i = 0; i < len; ++i) h[i] = static_cast<T>(1) / static_cast<T>(len);
        return h;
    }
    double s = sigma;
    if (!(s > 0)) s = 0.25 * len;
    const double pi = 3.14159265358979323846;
    T sum = T(0);
    for (int i = 0; i < (int)len; ++i) {
        const int x = i - ((int)len - 1) / 2;
        const T el = (T)(1. / std::sqrt(2 * pi * s * s) * std::exp(-((x * x) / (2 * (s * s)))));
        h[i] = el;
        sum += el;
    }
    for (unsigned k = 0; k < len; ++k) h[k] /= sum;
    return h;
}

static af_array zeros_f32(dim_t n) { return make_handle(f32, dim4(n)); }

template <typename T>
static Features harris(const ArrayImpl &in, unsigned max_corners, float min_response, float sigma,
                       unsigned filter_len, float k_thr) {
    const af_dtype t = in.type;
    const dim_t W = in.dims[0], H = in.dims[1];
    DIM_ASSERT(1, W >= 2 && H >= 2);  // the gradient's one-sided ends need two samples
    DIM_ASSERT(1, W * H < (1LL << 31));
    const dim_t n = W * H;
    hipStream_t s = stream();

    std::vector<T> hw = harris_window<T>(filter_len, sigma);
    auto filt = scratch(sizeof(T) * filter_len);
    check_hip(hipMemcpyAsync(filt->ptr, hw.data(), sizeof(T) * filter_len, hipMemcpyHostToDevice, s), "harris");
    const unsigned border_len = filter_len / 2 + 1;

    auto d3 = scratch(sizeof(T) * 3 * n);
    auto s3 = scratch(sizeof(T) * 3 * n);
    auto tmp = scratch(sizeof(T) * 3 * n);
    check_k(afhip_k_harris_deriv(t, d3->ptr, in.data(), W, H, s), "harris");
    // convolve2<T, convAccT>(ixx | ixy | iyy, filter, filter, false): one
    // call over the three slices; the workspace selects the two-pass kernel
    // with the reference's T-typed temp and summation order
    const long long dims3[4] = {W, H, 3, 1};
    check_k(afhip_k_convolve2_sep(t, s3->ptr, dims3, d3->ptr, dims3, filt->ptr, filter_len, filt->ptr, filter_len,
                                  0, tmp->ptr, s),
            "harris");

    const unsigned corner_lim = (unsigned)((float)n * 0.2f);
    const unsigned min_r = (max_corners > 0) ? 0U : (unsigned)min_response;
    auto flags = scratch((size_t)n);
    auto resp = scratch(sizeof(float) * n);
    check_k(afhip_k_harris_nms(t, (unsigned char *)flags->ptr, (float *)resp->ptr, s3->ptr, W, H, k_thr,
                               (float)min_r, (int)border_len + 1, s),
            "harris");
    const long long sb = afhip_k_select_flagged_workspace(n);
    if (sb < 0) AFHIP_THROW(AF_ERR_INTERNAL, "harris: select workspace");
    auto sel = scratch(sizeof(unsigned) * n);
    auto num = scratch(sizeof(unsigned));
    auto stemp = scratch((size_t)(sb > 0 ? sb : 1));
    check_k(afhip_k_select_flagged((const unsigned char *)flags->ptr, n, (unsigned *)sel->ptr, (unsigned *)num->ptr,
                                   stemp->ptr, sb, s),
            "harris");
    unsigned found = 0;
    check_hip(hipMemcpyAsync(&found, num->ptr, sizeof(unsigned), hipMemcpyDeviceToHost, s), "harris");
    sync_stream();

    // non_maximal keeps the first corner_lim corners in scan order
    const unsigned stored = found < corner_lim ? found : corner_lim;
    const unsigned lim = max_corners > 0 ? max_corners : corner_lim;
    const unsigned corners_out = found < lim ? found : lim;
    Features f;
    f.n = corners_out;
    NewHandle hx(zeros_f32(corners_out)), hy(zeros_f32(corners_out)), hs(zeros_f32(corners_out)),
        ho(zeros_f32(corners_out)), hz(zeros_f32(corners_out));
    if (corners_out > 0) {
        const unsigned *order = nullptr;
        const float *keys_sorted = nullptr;
        std::shared_ptr<Buffer> keys, vals, ksort, vsort, ttemp;
        if (max_corners > 0 && found > corners_out) {
            const long long tb = afhip_k_sort_desc_workspace(stored);
            if (tb < 0) AFHIP_THROW(AF_ERR_INTERNAL, "harris: sort workspace");
            keys = scratch(sizeof(float) * stored);
            vals = scratch(sizeof(unsigned) * stored);
            ksort = scratch(sizeof(float) * stored);
            vsort = scratch(sizeof(unsigned) * stored);
            ttemp = scratch((size_t)(tb > 0 ? tb : 1));
            check_k(afhip_k_harris_sort((const unsigned *)sel->ptr, (const float *)resp->ptr, stored,
                                        (float *)keys->ptr, (unsigned *)vals->ptr, (float *)ksort->ptr,
                                        (unsigned *)vsort->ptr, ttemp->ptr, tb, s),
                    "harris");
            order = (const unsigned *)vsort->ptr;
            keys_sorted = (const float *)ksort->ptr;
        }
        check_k(afhip_k_harris_out((float *)hx.data(), (float *)hy.data(), (float *)hs.data(), (float *)ho.data(),
                                   (float *)hz.data(), (const unsigned *)sel->ptr, order, keys_sorted,
                                   (const float *)resp->ptr, W, corners_out, s),
                "harris");
        sync_stream();  // the sort's scratch is returned to the pool below
    }
    f.x = hx.release();
    f.y = hy.release();
    f.score = hs.release();
    f.orientation = ho.release();
    f.size = hz.release();
    return f;
}

}  // namespace afhip

using namespace afhip;

extern "C" {

// src/api/c/features.cpp (af_create_features): n f32 arrays of n elements
AFAPI af_err af_create_features(af_features *out, dim_t num) {
    try {
        ARG_ASSERT(0, out != nullptr);
        ARG_ASSERT(1, num >= 0);
        auto f = std::make_unique<Features>();
        f->n = num;
        NewHandle a(make_handle(f32, dim4(num))), b(make_handle(f32, dim4(num))), c(make_handle(f32, dim4(num))),
            d(make_handle(f32, dim4(num))), e(make_handle(f32, dim4(num)));
        f->x = a.release();
        f->y = b.release();
        f->score = c.release();
        f->orientation = d.release();
        f->size = e.release();
        *out = f.release();
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err af_retain_features(af_features *out, const af_features in) {
    try {
        ARG_ASSERT(0, out != nullptr);
        const Features &f = feat(in);
        auto g = std::make_unique<Features>();
        g->n = f.n;
        for (auto p : {std::make_pair(&g->x, f.x), std::make_pair(&g->y, f.y), std::make_pair(&g->score, f.score),
                       std::make_pair(&g->orientation, f.orientation), std::make_pair(&g->size, f.size)}) {
            af_err e = af_retain_array(p.first, p.second);
            if (e != AF_SUCCESS) return e;
        }
        *out = g.release();
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err af_get_features_num(dim_t *num, const af_features in) {
    try {
        ARG_ASSERT(0, num != nullptr);
        *num = feat(in).n;
    }
    CATCHALL;
    return AF_SUCCESS;
}

// The getters hand out the features' own handles (not retained), as the
// reference's do.
#define AFHIP_FEAT_GET(NAME, FIELD)                              \
    AFAPI af_err NAME(af_array *out, const af_features in) {     \
        try {                                                    \
            ARG_ASSERT(0, out != nullptr);                       \
            *out = feat(in).FIELD;                               \
        }                                                        \
        CATCHALL;                                                \
        return AF_SUCCESS;                                       \
    }
AFHIP_FEAT_GET(af_get_features_xpos, x)
AFHIP_FEAT_GET(af_get_features_ypos, y)
AFHIP_FEAT_GET(af_get_features_score, score)
AFHIP_FEAT_GET(af_get_features_orientation, orientation)
AFHIP_FEAT_GET(af_get_features_size, size)
#undef AFHIP_FEAT_GET

AFAPI af_err af_release_features(af_features in) {
    try {
        if (!in) return AF_SUCCESS;
        Features *f = static_cast<Features *>(in);
        for (af_array a : {f->x, f->y, f->score, f->orientation, f->size})
            if (a) destroy_handle(a);
        delete f;
    }
    CATCHALL;
    return AF_SUCCESS;
}

// src/api/c/harris.cpp:52-95
AFAPI af_err af_harris(af_features *out, const af_array in, const unsigned max_corners, const float min_response,
                       const float sigma, const unsigned block_size, const float k_thr) {
    try {
        ARG_ASSERT(0, out != nullptr);
        const ArrayImpl &a = get(in);
        const dim4 dims = a.dims;
        const dim_t in_ndims = dims.ndims();
        unsigned filter_len = (block_size == 0) ? static_cast<unsigned>(std::floor(6.f * sigma)) : block_size;
        if (block_size == 0 && filter_len % 2 == 0) filter_len--;
        const unsigned edge = (block_size > 0) ? block_size / 2 : filter_len / 2;
        DIM_ASSERT(1, in_ndims == 2);
        ARG_ASSERT(1, dims[0] >= (dim_t)(2 * edge + 1) || dims[1] >= (dim_t)(2 * edge + 1));
        ARG_ASSERT(3, max_corners > 0 || min_response > 0.0f);
        ARG_ASSERT(7, k_thr >= 0.01f);
        ARG_ASSERT(4, block_size > 2 || (sigma >= 0.5f && sigma <= 5.f));
        ARG_ASSERT(5, block_size <= 32);
        Features f;
        switch (a.type) {
            case f64: f = harris<double>(a, max_corners, min_response, sigma, filter_len, k_thr); break;
            case f32: f = harris<float>(a, max_corners, min_response, sigma, filter_len, k_thr); break;
            default: TYPE_ERROR(1, a.type);
        }
        *out = new Features(f);
    }
    CATCHALL;
    return AF_SUCCESS;
}

}  // extern "C"
