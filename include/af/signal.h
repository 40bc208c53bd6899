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
/* include/af/signal.h:1492-1520. Frequency-domain convolution is not part of
 * this library: these return AF_ERR_NOT_SUPPORTED, and so do AUTO/FREQ calls
 * that the reference's domain heuristic routes to them. */
AFAPI af_err af_fft_convolve1(af_array *out, const af_array signal,
                              const af_array filter, const af_conv_mode mode);
AFAPI af_err af_fft_convolve2(af_array *out, const af_array signal,
                              const af_array filter, const af_conv_mode mode);
AFAPI af_err af_fft_convolve3(af_array *out, const af_array signal,
                              const af_array filter, const af_conv_mode mode);
