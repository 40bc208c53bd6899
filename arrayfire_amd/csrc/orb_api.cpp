// af_orb (src/api/c/orb.cpp:58-89 with src/backend/cpu/orb.cpp:44-283): an
// image pyramid, FAST corners, Harris ranking, centroid orientation and
// rotated-BRIEF descriptors, all computed on the GPU (kernels/orb.hip); each
// level's descriptor source is the level smoothed by the separable
// convolution path with a 9-tap sigma-2 Gaussian. The corner counts after
// detection and after suppression stay on the device (the kernels after
// them run over the level's cap and stop at the count); the host reads one
// count per level, the sort's size, and synchronises once at the end.
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

namespace {

constexpr unsigned kPatch = 31;      // REF_PAT_SIZE
constexpr int kArc = 9;              // FAST arc length
constexpr float kFeatRatio = 0.15f;  // FAST feature ratio
constexpr int kHarrisBlock = 7;
constexpr float kHarrisK = 0.04f;
constexpr int kGaussLen = 9;

// gaussian1D(out, 9, 2.0) (src/backend/cpu/utility.hpp:35-48): each tap in
// double, stored in T; the sum and the normalisation in T.
template <typename T>
std::vector<T> gauss_taps() {
    std::vector<T> h(kGaussLen);
    const double s = 2.0, pi = 3.14159265358979323846;
    T sum = T(0);
    for (int i = 0; i < kGaussLen; ++i) {
        const int x = i - (kGaussLen - 1) / 2;
        const T el = (T)(1. / std::sqrt(2 * pi * s * s) * std::exp(-((x * x) / (2 * (s * s)))));
        h[i] = el;
        sum += el;
    }
    for (int k = 0; k < kGaussLen; ++k) h[k] /= sum;
    return h;
}

// Buffers of one af_orb call, released together after its last sync.
typedef std::vector<std::shared_ptr<Buffer>> Keep;

std::shared_ptr<Buffer> kept(Keep &keep, size_t bytes) {
    keep.push_back(scratch(bytes));
    return keep.back();
}

// Stable compaction of the set flags: indices into sel; the count stays in
// device memory (the returned buffer).
std::shared_ptr<Buffer> select(Keep &keep, const std::shared_ptr<Buffer> &flags, long long n,
                               std::shared_ptr<Buffer> &sel) {
    const long long tb = afhip_k_select_flagged_workspace(n);
    if (tb < 0) AFHIP_THROW(AF_ERR_INTERNAL, "orb: select workspace");
    sel = kept(keep, sizeof(unsigned) * (size_t)n);
    auto num = kept(keep, sizeof(unsigned));
    auto temp = kept(keep, (size_t)(tb > 0 ? tb : 1));
    check_k(afhip_k_select_flagged((const unsigned char *)flags->ptr, n, (unsigned *)sel->ptr, (unsigned *)num->ptr,
                                   temp->ptr, tb, stream()),
            "orb");
    return num;
}

unsigned read_count(const std::shared_ptr<Buffer> &num) {
    unsigned v = 0;
    check_hip(hipMemcpyAsync(&v, num->ptr, sizeof(unsigned), hipMemcpyDeviceToHost, stream()), "orb");
    sync_stream();
    return v;
}

struct Level {
    unsigned n = 0;
    std::shared_ptr<Buffer> x, y, score, ori, size, desc;
};

template <typename T>
Level orb_level(Keep &keep, const ArrayImpl &img_arr, const void *img, dim_t d0, dim_t d1, float fast_thr,
                unsigned lvl_best, float lvl_scl, const std::vector<T> *gauss) {
    const int dt = img_arr.type;
    hipStream_t s = stream();
    Level L;
    // Round the patch to the nearest odd size; keep features whose widest
    // (45 degree) patch fits: edge = ceil(size * sqrt(2) / 2)
    const float size = 2.f * std::floor(static_cast<float>(kPatch) / 2.f) + 1.f;
    const int edge = (int)std::ceil(size * std::sqrt(2.f) / 2.f);
    const dim_t ny = d0 - 2 * (dim_t)edge, nx = d1 - 2 * (dim_t)edge;
    if (ny < 1 || nx < 1) return L;
    // fast(): at most ceil(elements * 0.15) corners are stored; the counts
    // of the first two compactions stay on the device, the kernels after
    // them run over this cap and stop at the count
    const unsigned max_fast = (unsigned)std::ceil(static_cast<float>(d0 * d1) * kFeatRatio);
    const long long ncap = std::min<long long>(max_fast, ny * nx);
    auto flags1 = kept(keep, (size_t)(ny * nx));
    auto score1 = kept(keep, sizeof(float) * (size_t)(ny * nx));
    check_k(afhip_k_orb_fast(dt, (unsigned char *)flags1->ptr, (float *)score1->ptr, img, d0, d1, edge, fast_thr,
                             kArc, s),
            "orb");
    std::shared_ptr<Buffer> sel1, sel2, sel3;
    auto num1 = select(keep, flags1, ny * nx, sel1);
    auto V = kept(keep, sizeof(float) * (size_t)(d0 * d1));
    auto flags2 = kept(keep, (size_t)ncap);
    check_k(afhip_k_orb_fast_nonmax((unsigned char *)flags2->ptr, (float *)V->ptr, (const unsigned *)sel1->ptr,
                                    (const float *)score1->ptr, ncap, (const unsigned *)num1->ptr, d0, d1, edge, s),
            "orb");
    auto num2 = select(keep, flags2, ncap, sel2);
    // harris_response<T, false> over a 7x7 block, patch radius as FAST's edge
    auto flags3 = kept(keep, (size_t)ncap);
    auto resp3 = kept(keep, sizeof(float) * (size_t)ncap);
    check_k(afhip_k_orb_harris(dt, (unsigned char *)flags3->ptr, (float *)resp3->ptr, img,
                               (const unsigned *)sel1->ptr, (const unsigned *)sel2->ptr, ncap,
                               (const unsigned *)num2->ptr, d0, d1, edge, edge, kHarrisBlock, kHarrisK, s),
            "orb");
    const unsigned n3 = read_count(select(keep, flags3, ncap, sel3));
    if (n3 == 0) return L;
    // sort_index(..., descending): stable descending radix sort of the
    // responses; the strongest lvl_best are kept
    const long long tb = afhip_k_sort_desc_workspace(n3);
    if (tb < 0) AFHIP_THROW(AF_ERR_INTERNAL, "orb: sort workspace");
    auto keys = kept(keep, sizeof(float) * n3), vals = kept(keep, sizeof(unsigned) * n3);
    auto ksort = kept(keep, sizeof(float) * n3), vsort = kept(keep, sizeof(unsigned) * n3);
    auto ttemp = kept(keep, (size_t)(tb > 0 ? tb : 1));
    check_k(afhip_k_harris_sort((const unsigned *)sel3->ptr, (const float *)resp3->ptr, n3, (float *)keys->ptr,
                                (unsigned *)vals->ptr, (float *)ksort->ptr, (unsigned *)vsort->ptr, ttemp->ptr, tb,
                                s),
            "orb");
    const unsigned usable = n3 < lvl_best ? n3 : lvl_best;
    if (usable == 0) return L;
    auto lx = kept(keep, sizeof(float) * usable), ly = kept(keep, sizeof(float) * usable);
    L.score = kept(keep, sizeof(float) * usable);
    L.ori = kept(keep, sizeof(float) * usable);
    check_k(afhip_k_orb_angle(dt, (float *)lx->ptr, (float *)ly->ptr, (float *)L.score->ptr, (float *)L.ori->ptr,
                              img, (const unsigned *)sel1->ptr, (const unsigned *)sel2->ptr,
                              (const unsigned *)sel3->ptr, (const unsigned *)vsort->ptr,
                              (const float *)ksort->ptr, usable, d0, d1, edge, (int)kPatch, s),
            "orb");
    // the descriptor source: the level smoothed by convolve2(img, g, g) in
    // the reference's two-pass order, or the level itself
    const void *src = img;
    if (gauss) {
        const size_t bytes = sizeof(T) * (size_t)(d0 * d1);
        auto filt = kept(keep, bytes), tmp = kept(keep, bytes), taps = kept(keep, sizeof(T) * kGaussLen);
        check_hip(hipMemcpyAsync(taps->ptr, gauss->data(), sizeof(T) * kGaussLen, hipMemcpyHostToDevice, s), "orb");
        const long long dd[4] = {d0, d1, 1, 1};
        check_k(afhip_k_convolve2_sep(dt, filt->ptr, dd, img, dd, taps->ptr, kGaussLen, taps->ptr, kGaussLen, 0,
                                      tmp->ptr, s),
                "orb");
        src = filt->ptr;
    }
    L.x = kept(keep, sizeof(float) * usable);
    L.y = kept(keep, sizeof(float) * usable);
    L.size = kept(keep, sizeof(float) * usable);
    L.desc = kept(keep, sizeof(unsigned) * 8 * (size_t)usable);
    const int r = (int)std::ceil(kPatch * std::sqrt(2.f) / 2.f);
    check_k(afhip_k_orb_describe(dt, (unsigned *)L.desc->ptr, (float *)L.x->ptr, (float *)L.y->ptr,
                                 (float *)L.size->ptr, (const float *)lx->ptr, (const float *)ly->ptr,
                                 (const float *)L.ori->ptr, src, usable, d0, d1, (int)kPatch, r, lvl_scl, s),
            "orb");
    L.n = usable;
    return L;
}

template <typename T>
void orb(Features &feat, af_array &desc_out, const ArrayImpl &in, float fast_thr, unsigned max_feat, float scl_fctr,
         unsigned levels, bool blur_img) {
    const dim_t id0 = in.dims[0], id1 = in.dims[1];
    // pyramid depth: levels whose smaller side stays at least the patch size
    float min_side = static_cast<float>(id0 < id1 ? id0 : id1);
    unsigned max_levels = 0;
    float scl_sum = 0.f;
    for (unsigned i = 0; i < levels; ++i) {
        min_side /= scl_fctr;
        if (min_side < static_cast<float>(kPatch) || max_levels == levels) break;
        max_levels++;
        scl_sum += 1.f / std::pow(scl_fctr, static_cast<float>(i));
    }
    std::vector<Level> lv;
    Keep keep;  // every level's buffers, released after the final sync
    // the blur taps: host memory that asynchronous copies read, alive until
    // the final sync
    const std::vector<T> g = blur_img ? gauss_taps<T>() : std::vector<T>();
    if (max_levels > 0) {
        // features per level: ceil(max_feat / scl_sum / scale), the rest to
        // the last level (unsigned arithmetic, as the reference's)
        std::vector<unsigned> lvl_best(max_levels);
        unsigned feat_sum = 0;
        for (unsigned i = 0; i + 1 < max_levels; ++i) {
            const float ls = std::pow(scl_fctr, static_cast<float>(i));
            lvl_best[i] = (unsigned)std::ceil((static_cast<float>(max_feat) / scl_sum) / ls);
            feat_sum += lvl_best[i];
        }
        lvl_best[max_levels - 1] = max_feat - feat_sum;
        std::shared_ptr<Buffer> prev;  // level i-1 (level 0 is the input)
        const void *prev_ptr = in.data();
        dim_t pd0 = id0, pd1 = id1;
        for (unsigned i = 0; i < max_levels; ++i) {
            const float ls = std::pow(scl_fctr, static_cast<float>(i));
            std::shared_ptr<Buffer> cur;
            const void *img = in.data();
            dim_t d0 = id0, d1 = id1;
            if (i > 0) {
                d0 = (dim_t)std::round(static_cast<float>(id0) / ls);
                d1 = (dim_t)std::round(static_cast<float>(id1) / ls);
                cur = kept(keep, sizeof(T) * (size_t)(d0 * d1));
                check_k(afhip_k_orb_resize(in.type, cur->ptr, prev_ptr, pd0, pd1, d0, d1, stream()), "orb");
                img = cur->ptr;
                prev = cur;
                prev_ptr = cur->ptr;
                pd0 = d0;
                pd1 = d1;
            }
            Level L = orb_level<T>(keep, in, img, d0, d1, fast_thr, lvl_best[i], ls, blur_img ? &g : nullptr);
            if (L.n > 0) lv.push_back(std::move(L));
        }
    }
    // Levels are concatenated in order. (The reference advances its copy
    // offset only past the level just before a non-empty one, so a level
    // with no features between two that have some makes it overwrite
    // earlier rows and leave the tail unset; here every level lands after
    // the previous non-empty one.)
    dim_t total = 0;
    for (const Level &L : lv) total += L.n;
    NewHandle hx(make_handle(f32, dim4(total))), hy(make_handle(f32, dim4(total))),
        hs(make_handle(f32, dim4(total))), ho(make_handle(f32, dim4(total))), hz(make_handle(f32, dim4(total)));
    NewHandle hd(make_handle(u32, total > 0 ? dim4(8, total) : dim4(0)));
    hipStream_t s = stream();
    dim_t off = 0;
    for (const Level &L : lv) {
        const size_t fb = sizeof(float) * L.n;
        check_hip(hipMemcpyAsync((float *)hx.data() + off, L.x->ptr, fb, hipMemcpyDeviceToDevice, s), "orb");
        check_hip(hipMemcpyAsync((float *)hy.data() + off, L.y->ptr, fb, hipMemcpyDeviceToDevice, s), "orb");
        check_hip(hipMemcpyAsync((float *)hs.data() + off, L.score->ptr, fb, hipMemcpyDeviceToDevice, s), "orb");
        check_hip(hipMemcpyAsync((float *)ho.data() + off, L.ori->ptr, fb, hipMemcpyDeviceToDevice, s), "orb");
        check_hip(hipMemcpyAsync((float *)hz.data() + off, L.size->ptr, fb, hipMemcpyDeviceToDevice, s), "orb");
        check_hip(hipMemcpyAsync((unsigned *)hd.data() + 8 * off, L.desc->ptr, 8 * sizeof(unsigned) * L.n,
                                 hipMemcpyDeviceToDevice, s),
                  "orb");
        off += L.n;
    }
    sync_stream();  // the levels' buffers (keep) go back to the pool on return
    feat.n = total;
    feat.x = hx.release();
    feat.y = hy.release();
    feat.score = hs.release();
    feat.orientation = ho.release();
    feat.size = hz.release();
    desc_out = hd.release();
}

}  // namespace

}  // namespace afhip

using namespace afhip;

// src/api/c/orb.cpp:58-89
AFAPI af_err af_orb(af_features *feat, af_array *desc, const af_array in, const float fast_thr,
                    const unsigned max_feat, const float scl_fctr, const unsigned levels, const bool blur_img) {
    try {
        ARG_ASSERT(0, feat != nullptr && desc != nullptr);
        const ArrayImpl &a = get(in);
        const dim4 dims = a.dims;
        ARG_ASSERT(2, (dims[0] >= 7 && dims[1] >= 7 && dims[2] == 1 && dims[3] == 1));
        ARG_ASSERT(3, fast_thr > 0.0f);
        ARG_ASSERT(4, max_feat > 0);
        ARG_ASSERT(5, scl_fctr > 1.0f);
        ARG_ASSERT(6, levels > 0);
        DIM_ASSERT(1, dims.ndims() == 2);
        DIM_ASSERT(1, dims[0] * dims[1] < (1LL << 31));
        Features f;
        af_array d = nullptr;
        switch (a.type) {
            case f32: orb<float>(f, d, a, fast_thr, max_feat, scl_fctr, levels, blur_img); break;
            case f64: orb<double>(f, d, a, fast_thr, max_feat, scl_fctr, levels, blur_img); break;
            default: TYPE_ERROR(1, a.type);
        }
        *feat = new Features(f);
        *desc = d;
    }
    CATCHALL;
    return AF_SUCCESS;
}
