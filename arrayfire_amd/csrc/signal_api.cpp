// Internal callers of the convolution path: af_fir, af_iir
// (src/api/c/iir.cpp:24-103, src/backend/cpu/iir.cpp:23-45), af_dog
// (src/api/c/dog.cpp:30-80) and af_gaussian_kernel
// (src/api/c/gaussian_kernel.cpp:36-103), which DoG builds its filters with.
// Each one reaches the same spatial (or, through af_convolve1's AUTO route,
// frequency-domain) convolution kernels as af_convolve*.
#include <cmath>
#include <vector>

#include "../../include/af/array.h"
#include "../../include/af/image.h"
#include "../../include/af/signal.h"
#include "../../include/afhip_kernels.h"
#include "runtime.hpp"

namespace afhip {

// gaussianKernel<float> (src/api/c/gaussian_kernel.cpp:36-91), evaluated
// in float in the reference's order: per element the squared normalised
// column offset plus the squared normalised row offset, times -1/2, exp;
// then divided by the column-major sum of the window.
static std::vector<float> gaussian_host(int rows, int cols, double sigma_r, double sigma_c) {
    std::vector<float> t((size_t)rows * cols, 0.0f);
    if (cols > 1) {
        const float c = (float)((double)(cols - 1) / 2.0);
        const float sig = (float)(sigma_c > 0 ? sigma_c : 0.25 * cols);
        for (int j = 0; j < cols; ++j)
            for (int i = 0; i < rows; ++i) {
                const float w = ((float)j - c) / sig;
                t[(size_t)j * rows + i] = w * w + t[(size_t)j * rows + i];
            }
    }
    if (rows > 1) {
        const float r = (float)((double)(rows - 1) / 2.0);
        const float sig = (float)(sigma_r > 0 ? sigma_r : 0.25 * rows);
        for (int j = 0; j < cols; ++j)
            for (int i = 0; i < rows; ++i) {
                const float w = ((float)i - r) / sig;
                t[(size_t)j * rows + i] = w * w + t[(size_t)j * rows + i];
            }
    }
    float sum = 0.0f;
    for (auto &v : t) {
        v = std::exp(0.0f - 0.5f * v);
        sum += v;
    }
    for (auto &v : t) v = v / sum;
    return t;
}

static af_array upload(af_dtype type, const dim4 &dims, const void *host, size_t bytes) {
    NewHandle h(make_handle(type, dims));
    if (bytes) {
        check_hip(hipMemcpyAsync(h.data(), host, bytes, hipMemcpyHostToDevice, stream()), "upload");
        sync_stream();  // the host vector is about to go away
    }
    return h.release();
}

// castArray<to>: a device copy converted to `to` (shares storage when equal)
static std::shared_ptr<Buffer> as_type(const ArrayImpl &a, af_dtype to) {
    if (a.type == to) return a.buf;
    auto b = scratch((size_t)a.elements() * dtype_size(to));
    check_k(afhip_k_cast(to, b->ptr, a.type, a.data(), a.elements(), stream()), "cast");
    return b;
}

}  // namespace afhip

using namespace afhip;

AFAPI af_err af_gaussian_kernel(af_array *out, const int rows, const int cols, const double sigma_r,
                                const double sigma_c) {
    try {
        ARG_ASSERT(0, out != nullptr);
        ARG_ASSERT(1, rows > 0);
        ARG_ASSERT(2, cols > 0);
        const std::vector<float> g = gaussian_host(rows, cols, sigma_r, sigma_c);
        *out = upload(f32, dim4(rows, cols), g.data(), g.size() * sizeof(float));
    }
    CATCHALL;
    return AF_SUCCESS;
}

// src/api/c/iir.cpp:24-41: the first x0 rows of af_convolve1(x, b, EXPAND, AUTO)
AFAPI af_err af_fir(af_array *y, const af_array b, const af_array x) {
    try {
        ARG_ASSERT(0, y != nullptr);
        af_array full = nullptr;
        const af_err e = af_convolve1(&full, x, b, AF_CONV_EXPAND, AF_CONV_AUTO);
        if (e != AF_SUCCESS) return e;
        NewHandle fh(full);
        const ArrayImpl &f = get(full);
        const dim4 xd = get(x).dims;
        dim4 od = f.dims;
        od[0] = std::min(xd[0], f.dims[0]);
        NewHandle o(make_handle(f.type, od));
        const long long off[4] = {0, 0, 0, 0};
        if (od.elements() > 0)
            check_k(afhip_k_region_copy((int)dtype_size(f.type), o.data(), od.d, f.data(), f.dims.d, off,
                                        stream()),
                    "fir");
        *y = o.release();
    }
    CATCHALL;
    return AF_SUCCESS;
}

// src/api/c/iir.cpp:49-103 and cpu::iir (src/backend/cpu/iir.cpp:23-45)
AFAPI af_err af_iir(af_array *y, const af_array b, const af_array a, const af_array x) {
    try {
        ARG_ASSERT(0, y != nullptr);
        const ArrayImpl &A = get(a);
        const ArrayImpl &B = get(b);
        const ArrayImpl &X = get(x);
        const af_dtype xt = X.type;
        ARG_ASSERT(1, A.type == xt);
        ARG_ASSERT(2, B.type == xt);
        ARG_ASSERT(1, B.dims.ndims() == A.dims.ndims());
        if (X.dims.ndims() == 0) return af_retain_array(y, x);
        if (X.dims.ndims() > 1 && B.dims.ndims() > 1)
            for (int i = 1; i < 3; ++i) ARG_ASSERT(1, B.dims[i] == X.dims[i]);
        if (xt != f32 && xt != f64 && xt != c32 && xt != c64) TYPE_ERROR(1, xt);
        if (A.dims[0] == 1) {
            // only a0: y = fir(b / a0, x), a0 broadcast along b's columns
            // (af_div(b, a, batch = true), src/api/c/iir.cpp:76-81)
            const bool per_col = A.elements() > 1;
            if (per_col)
                for (int i = 1; i < 4; ++i) DIM_ASSERT(1, A.dims[i] == B.dims[i]);
            NewHandle bnorm(make_handle(xt, B.dims));
            check_k(afhip_k_div_cols(xt, bnorm.data(), B.data(), B.dims.d, A.data(), per_col ? 1 : 0, stream()),
                    "iir");
            return af_fir(y, bnorm.get(), x);
        }
        // batch kind of the feedforward convolution (src/backend/cpu/iir.cpp:25-28)
        const dim_t xn = X.dims.ndims(), bn = B.dims.ndims();
        int kind = xn == 1 ? 1 : 4;  // NONE : SAME
        if (xn != bn) kind = xn < bn ? 3 : 2;  // RHS : LHS
        dim4 cd(1, 1, 1, 1);
        for (int d = 0; d < 4; ++d) {
            if (d == 0) cd[d] = X.dims[0] + B.dims[0] - 1;
            else if (kind == 1 || kind == 3) cd[d] = X.dims[d] + B.dims[d] - 1;
            else cd[d] = X.dims[d];
        }
        auto c = scratch((size_t)cd.elements() * dtype_size(xt));
        if (cd.elements() > 0)
            check_k(afhip_k_convolve(xt, c->ptr, cd.d, X.data(), X.dims.d, B.data(), B.dims.d, kind, 1, 1,
                                     stream()),
                    "iir");
        dim4 yd = cd;
        yd[0] = X.dims[0];
        NewHandle o(make_handle(xt, yd));
        const long long need = afhip_k_iir_workspace(xt, A.dims[0], yd.d);
        std::shared_ptr<Buffer> ws;
        if (need > 0) ws = scratch((size_t)need);
        if (yd.elements() > 0)
            check_k(afhip_k_iir(xt, o.data(), yd.d, c->ptr, cd.d, A.data(), A.dims.d, ws ? ws->ptr : nullptr,
                                need > 0 ? need : 0, stream()),
                    "iir");
        *y = o.release();
    }
    CATCHALL;
    return AF_SUCCESS;
}

// src/api/c/dog.cpp:30-80: two Gaussian smoothings (2r+1 square windows,
// sigma 0.25 * side) in accT and their difference.
AFAPI af_err af_dog(af_array *out, const af_array in, const int radius1, const int radius2) {
    try {
        ARG_ASSERT(0, out != nullptr);
        const ArrayImpl &X = get(in);
        const dim4 &id = X.dims;
        ARG_ASSERT(1, id.ndims() >= 2);
        ARG_ASSERT(1, id.ndims() <= 3);
        af_dtype at;
        switch ((int)X.type) {
            case f32: case b8: case s32: case u32: case s16: case u16: case s8: case u8: at = f32; break;
            case f64: at = f64; break;
            default: TYPE_ERROR(1, X.type);
        }
        ARG_ASSERT(2, radius1 >= 0);
        ARG_ASSERT(3, radius2 >= 0);
        auto xin = as_type(X, at);
        const int k1 = 2 * radius1 + 1, k2 = 2 * radius2 + 1;
        const std::vector<float> g1 = gaussian_host(k1, k1, 0.0, 0.0), g2 = gaussian_host(k2, k2, 0.0, 0.0);
        // the filters in accT, on the device
        auto fbuf = [&](const std::vector<float> &g) {
            auto b = scratch(g.size() * dtype_size(at));
            if (at == f32) {
                check_hip(hipMemcpyAsync(b->ptr, g.data(), g.size() * 4, hipMemcpyHostToDevice, stream()), "upload");
            } else {
                std::vector<double> d(g.begin(), g.end());
                check_hip(hipMemcpyAsync(b->ptr, d.data(), d.size() * 8, hipMemcpyHostToDevice, stream()), "upload");
                sync_stream();
            }
            return b;
        };
        auto f1 = fbuf(g1), f2 = fbuf(g2);
        sync_stream();  // host filter vectors end with this scope
        const int kind = id[2] > 1 ? 2 : 1;  // LHS : NONE
        const long long fd1[4] = {k1, k1, 1, 1}, fd2[4] = {k2, k2, 1, 1};
        auto s1 = scratch((size_t)id.elements() * dtype_size(at));
        auto s2 = scratch((size_t)id.elements() * dtype_size(at));
        NewHandle o(make_handle(at, id));
        if (id.elements() > 0) {
            check_k(afhip_k_convolve(at, s1->ptr, id.d, xin->ptr, id.d, f1->ptr, fd1, kind, 2, 0, stream()), "dog");
            check_k(afhip_k_convolve(at, s2->ptr, id.d, xin->ptr, id.d, f2->ptr, fd2, kind, 2, 0, stream()), "dog");
            check_k(afhip_k_sub(at, o.data(), s1->ptr, s2->ptr, id.elements(), stream()), "dog");
        }
        *out = o.release();
    }
    CATCHALL;
    return AF_SUCCESS;
}
