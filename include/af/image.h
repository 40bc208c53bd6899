/*
 * Image-processing entry points of the ArrayFire C API that sit on the
 * convolution path: the public im2col / col2im pair.
 */
#pragma once
#include <af/defines.h>

/* include/af/image.h:1483; src/api/c/unwrap.cpp:37-99. Patches of wx x wy
 * with strides (sx, sy) and zero padding (px, py): out is
 * [wx*wy, nx*ny, d2, d3] when is_column, else [nx*ny, wx*wy, d2, d3], with
 * nx = 1 + (d0 + 2*px - wx) / sx (and ny alike). */
AFAPI af_err af_unwrap(af_array *out, const af_array in, const dim_t wx,
                       const dim_t wy, const dim_t sx, const dim_t sy,
                       const dim_t px, const dim_t py, const bool is_column);
/* include/af/image.h:1524; src/api/c/wrap.cpp:87-99. The adjoint of
 * af_unwrap: patch elements are added into a zero [ox, oy, d2, d3] array. */
AFAPI af_err af_wrap(af_array *out, const af_array in, const dim_t ox,
                     const dim_t oy, const dim_t wx, const dim_t wy,
                     const dim_t sx, const dim_t sy, const dim_t px,
                     const dim_t py, const bool is_column);
/* include/af/image.h:1570; src/api/c/wrap.cpp:101-115. As af_wrap; when
 * *out is not NULL the patches are added into that existing array (whose
 * dims must be [ox, oy, d2, d3]) in place. */
AFAPI af_err af_wrap_v2(af_array *out, const af_array in, const dim_t ox,
                        const dim_t oy, const dim_t wx, const dim_t wy,
                        const dim_t sx, const dim_t sy, const dim_t px,
                        const dim_t py, const bool is_column);
/* include/af/image.h (af_gaussian_kernel); src/api/c/gaussian_kernel.cpp:
 * a rows x cols f32 Gaussian window (sigma 0.25 * side when not positive),
 * normalised to sum 1. */
AFAPI af_err af_gaussian_kernel(af_array *out, const int rows, const int cols,
                                const double sigma_r, const double sigma_c);
/* include/af/image.h (af_dog); src/api/c/dog.cpp:30-80: difference of two
 * Gaussian smoothings (windows 2*radius+1), computed in f32 (f64 for f64
 * input); 2-D or 3-D input (the third dim is a batch). */
AFAPI af_err af_dog(af_array *out, const af_array in, const int radius1,
                    const int radius2);
/* include/af/image.h:1683; src/api/c/canny.cpp:224-285. Binary (b8) edge
 * map of a 2-D image (or a batch of them) of f32, f64, s32, u32, s16, u16,
 * s8 or u8: 5-tap smoothing, 3x3 Sobel (sobel_window must be 3), L1
 * (is_fast) or L2 magnitude, non-maximum suppression, then hysteresis with
 * thresholds low/high of the normalised magnitude (MANUAL) or low * Otsu's
 * threshold / Otsu's threshold (AUTO_OTSU, high ignored). */
AFAPI af_err af_canny(af_array *out, const af_array in,
                      const af_canny_threshold threshold_type,
                      const float low_threshold_ratio,
                      const float high_threshold_ratio,
                      const unsigned sobel_window, const bool is_fast);
