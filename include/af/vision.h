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
