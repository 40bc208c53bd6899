/*
 * Computer-vision entry points of the ArrayFire C API that call the
 * convolution path.
 */
#pragma once
#include <af/defines.h>
#include <af/features.h>

/* include/af/vision.h:402; src/api/c/harris.cpp:52-95. Harris corners of a
 * 2-D f32/f64 image: the window is block_size wide (a box filter), or, for
 * block_size 0, a Gaussian of sigma (0.5..5) and floor(6 sigma) taps (made
 * odd). max_corners > 0 keeps the strongest max_corners; max_corners 0
 * keeps every corner with a response of at least (unsigned)min_response.
 * x is the dim1 index, y the dim0 index. */
AFAPI af_err af_harris(af_features *out, const af_array in, const unsigned max_corners,
                       const float min_response, const float sigma,
                       const unsigned block_size, const float k_thr);

/* include/af/vision.h:429-431; src/api/c/orb.cpp:58-89. ORB features of a 2-D
 * f32/f64 image: FAST corners (threshold fast_thr) on up to `levels` pyramid
 * levels scaled by 1/scl_fctr each, the max_feat strongest by Harris
 * response spread over the levels, their centroid orientation and a 256-bit
 * rotated-BRIEF descriptor (8 u32 words per feature, desc dims [8, n]),
 * sampled from the level smoothed by a 9-tap sigma-2 Gaussian when blur_img.
 * x and y are in the input image's coordinates. */
AFAPI af_err af_orb(af_features *feat, af_array *desc, const af_array in, const float fast_thr,
                    const unsigned max_feat, const float scl_fctr, const unsigned levels,
                    const bool blur_img);
