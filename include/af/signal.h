/*
 * Spatial convolution entry points of the ArrayFire C API.
 *
 * Drop-in replacements for the reference declarations cited per function;
 * dispatch semantics (rank fallback, domain heuristic, batch kinds, dtype to
 * accumulator mapping, error codes) restate src/api/c/convolve.cpp:77-334.
 */
#pragma once
#include <af/defines.h>

/* include/af/signal.h:1391; src/api/c/convolve.cpp:217-226 */
AFAPI af_err af_convolve1(af_array *out, const af_array signal,
                          const af_array filter, const af_conv_mode mode,
                          af_conv_domain domain);
/* include/af/signal.h:1408; src/api/c/convolve.cpp:228-241 */
AFAPI af_err af_convolve2(af_array *out, const af_array signal,
                          const af_array filter, const af_conv_mode mode,
                          af_conv_domain domain);
/* include/af/signal.h:1460; src/api/c/convolve.cpp:243-256 */
AFAPI af_err af_convolve3(af_array *out, const af_array signal,
                          const af_array filter, const af_conv_mode mode,
                          af_conv_domain domain);
/* include/af/signal.h:1478; src/api/c/convolve.cpp:258-334.
 * Note the argument order: filters first. */
AFAPI af_err af_convolve2_sep(af_array *out, const af_array col_filter,
                              const af_array row_filter, const af_array signal,
                              const af_conv_mode mode);
/* include/af/signal.h:1492-1520; src/api/c/fftconvolve.cpp:162-264.
 * Frequency-domain convolution (every batch kind including AF_BATCH_DIFF);
 * AUTO calls that the reference's domain heuristic routes to the frequency
 * domain land here too. */
AFAPI af_err af_fft_convolve1(af_array *out, const af_array signal,
                              const af_array filter, const af_conv_mode mode);
AFAPI af_err af_fft_convolve2(af_array *out, const af_array signal,
                              const af_array filter, const af_conv_mode mode);
AFAPI af_err af_fft_convolve3(af_array *out, const af_array signal,
                              const af_array filter, const af_conv_mode mode);

/* include/af/signal.h (af_fir); src/api/c/iir.cpp:24-41: the first x0 rows
 * of af_convolve1(x, b, AF_CONV_EXPAND, AF_CONV_AUTO). */
AFAPI af_err af_fir(af_array *y, const af_array b, const af_array x);
/* include/af/signal.h (af_iir); src/api/c/iir.cpp:49-103: feedforward
 * convolution with b, then the feedback recurrence with a (f32, f64, c32,
 * c64); a with a single tap normalises b and runs af_fir. */
AFAPI af_err af_iir(af_array *y, const af_array b, const af_array a,
                    const af_array x);
