// af_unwrap / af_wrap / af_wrap_v2: argument checks and output dims of
// src/api/c/unwrap.cpp:37-99 and src/api/c/wrap.cpp:26-115, queued on the
// gfx950 kernels of kernels/unwrap.hip.
#include "../../include/af/array.h"
#include "../../include/af/image.h"
#include "../../include/afhip_kernels.h"
#include "runtime.hpp"

namespace afhip {

// The dtypes the reference instantiates for both functions (no f16).
static bool wrap_type_ok(af_dtype t) {
    switch ((int)t) {
        case f32: case f64: case c32: case c64: case s32: case u32: case s64: case u64:
        case s16: case u16: case s8: case u8: case b8:
            return true;
        default:
            return false;
    }
}

}  // namespace afhip

using namespace afhip;

AFAPI af_err af_unwrap(af_array *out, const af_array in, const dim_t wx, const dim_t wy, const dim_t sx,
                       const dim_t sy, const dim_t px, const dim_t py, const bool is_column) {
    try {
        const ArrayImpl &a = get(in);
        const dim4 &id = a.dims;
        ARG_ASSERT(2, wx > 0 && wx <= id[0] + 2 * px);
        ARG_ASSERT(3, wy > 0 && wy <= id[1] + 2 * py);
        ARG_ASSERT(4, sx > 0);
        ARG_ASSERT(5, sy > 0);
        ARG_ASSERT(6, px >= 0 && px < wx);
        ARG_ASSERT(7, py >= 0 && py < wy);
        if (!wrap_type_ok(a.type)) TYPE_ERROR(1, a.type);
        // src/backend/cpu/unwrap.cpp:24-30 (dilation 1)
        const dim_t nx = 1 + (id[0] + 2 * px - wx) / sx;
        const dim_t ny = 1 + (id[1] + 2 * py - wy) / sy;
        dim4 od(wx * wy, nx * ny, id[2], id[3]);
        if (!is_column) std::swap(od[0], od[1]);
        NewHandle o(make_handle(a.type, od));
        if (od.elements() > 0)
            check_k(afhip_k_unwrap(a.type, o.data(), od.d, a.data(), id.d, wx, wy, sx, sy, px, py, 1, 1,
                                   is_column ? 1 : 0, stream()),
                    "unwrap");
        ARG_ASSERT(0, out != nullptr);
        *out = o.release();
    }
    CATCHALL;
    return AF_SUCCESS;
}

// src/api/c/wrap.cpp:26-85 (af_wrap_common)
static void wrap_common(af_array *out, const af_array in, const dim_t ox, const dim_t oy, const dim_t wx,
                        const dim_t wy, const dim_t sx, const dim_t sy, const dim_t px, const dim_t py,
                        const bool is_column, const bool allocate_out) {
    ARG_ASSERT(0, out != nullptr);
    ARG_ASSERT(1, in != nullptr);
    const ArrayImpl &a = get(in);
    const dim4 &id = a.dims;
    const dim4 od(ox, oy, id[2], id[3]);
    ARG_ASSERT(4, wx > 0);
    ARG_ASSERT(5, wy > 0);
    ARG_ASSERT(6, sx > 0);
    ARG_ASSERT(7, sy > 0);
    const dim_t nx = (ox + 2 * px - wx) / sx + 1;
    const dim_t ny = (oy + 2 * py - wy) / sy + 1;
    const dim_t patch_size = is_column ? id[0] : id[1];
    const dim_t num_patches = is_column ? id[1] : id[0];
    DIM_ASSERT(1, patch_size == wx * wy);
    DIM_ASSERT(1, num_patches == nx * ny);
    if (!wrap_type_ok(a.type)) TYPE_ERROR(1, a.type);
    if (allocate_out) {
        NewHandle o(make_handle(a.type, od));
        if (od.elements() > 0)
            check_k(afhip_k_wrap(a.type, o.data(), od.d, a.data(), id.d, wx, wy, sx, sy, px, py,
                                 is_column ? 1 : 0, 0, stream()),
                    "wrap");
        *out = o.release();
        return;
    }
    ArrayImpl &o = get(*out);
    DIM_ASSERT(0, o.dims[0] == od[0] && o.dims[1] == od[1] && o.dims[2] == od[2] && o.dims[3] == od[3]);
    if (o.type != a.type) TYPE_ERROR(0, o.type);
    if (od.elements() > 0)
        check_k(afhip_k_wrap(a.type, o.data(), od.d, a.data(), id.d, wx, wy, sx, sy, px, py, is_column ? 1 : 0, 1,
                             stream()),
                "wrap");
}

AFAPI af_err af_wrap(af_array *out, const af_array in, const dim_t ox, const dim_t oy, const dim_t wx,
                     const dim_t wy, const dim_t sx, const dim_t sy, const dim_t px, const dim_t py,
                     const bool is_column) {
    try {
        wrap_common(out, in, ox, oy, wx, wy, sx, sy, px, py, is_column, true);
    }
    CATCHALL;
    return AF_SUCCESS;
}

AFAPI af_err af_wrap_v2(af_array *out, const af_array in, const dim_t ox, const dim_t oy, const dim_t wx,
                        const dim_t wy, const dim_t sx, const dim_t sy, const dim_t px, const dim_t py,
                        const bool is_column) {
    try {
        ARG_ASSERT(0, out != nullptr);
        wrap_common(out, in, ox, oy, wx, wy, sx, sy, px, py, is_column, *out == nullptr);
    }
    CATCHALL;
    return AF_SUCCESS;
}
