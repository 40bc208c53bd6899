/*
 * Neural-network convolution entry points of the ArrayFire C API.
 */
#pragma once
#include <af/defines.h>

/* include/af/signal.h:1440-1443; src/api/c/convolve.cpp:344-392.
 * signal [W,H,C,N], filter [fw,fh,C,K] -> [OW,OH,K,N]; f32, f64, f16. */
AFAPI af_err af_convolve2_nn(af_array *out, const af_array signal,
                             const af_array filter, const unsigned stride_dims,
                             const dim_t *strides, const unsigned padding_dims,
                             const dim_t *paddings, const unsigned dilation_dims,
                             const dim_t *dilations);

/* include/af/ml.h:84-92; src/api/c/convolve.cpp:394-477.
 * AF_CONV_GRADIENT_FILTER returns dW; every other grad_type returns dX. */
AFAPI af_err af_convolve2_gradient_nn(
    af_array *out, const af_array incoming_gradient,
    const af_array original_signal, const af_array original_filter,
    const af_array convolved_output, const unsigned stride_dims,
    const dim_t *strides, const unsigned padding_dims, const dim_t *paddings,
    const unsigned dilation_dims, const dim_t *dilations,
    af_conv_gradient_type grad_type);
