/*
 * Types and enums of the ArrayFire C API that the convolution path uses.
 *
 * Values are identical to the reference so that existing code compiled
 * against ArrayFire's headers binds to this library unchanged:
 *   af_err ................ include/af/defines.h:82-207
 *   af_dtype .............. include/af/defines.h:210-233
 *   AF_MAX_DIMS, af_array . include/af/defines.h:240-243
 *   af_conv_mode .......... include/af/defines.h:303-314
 *   af_conv_domain ........ include/af/defines.h:316-320
 *   af_conv_gradient_type . include/af/defines.h:545-550
 */
#pragma once

#include <stddef.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
#define AFAPI extern "C" __attribute__((visibility("default")))
#else
#define AFAPI __attribute__((visibility("default")))
#endif

typedef long long dim_t;
typedef void *af_array;

typedef enum {
    AF_SUCCESS = 0,
    AF_ERR_NO_MEM = 101,
    AF_ERR_DRIVER = 102,
    AF_ERR_RUNTIME = 103,
    AF_ERR_INVALID_ARRAY = 201,
    AF_ERR_ARG = 202,
    AF_ERR_SIZE = 203,
    AF_ERR_TYPE = 204,
    AF_ERR_DIFF_TYPE = 205,
    AF_ERR_BATCH = 207,
    AF_ERR_DEVICE = 208,
    AF_ERR_NOT_SUPPORTED = 301,
    AF_ERR_NOT_CONFIGURED = 302,
    AF_ERR_NONFREE = 303,
    AF_ERR_NO_DBL = 401,
    AF_ERR_NO_GFX = 402,
    AF_ERR_NO_HALF = 403,
    AF_ERR_LOAD_LIB = 501,
    AF_ERR_LOAD_SYM = 502,
    AF_ERR_ARR_BKND_MISMATCH = 503,
    AF_ERR_INTERNAL = 998,
    AF_ERR_UNKNOWN = 999
} af_err;

typedef enum {
    f32 = 0,
    c32 = 1,
    f64 = 2,
    c64 = 3,
    b8 = 4,
    s32 = 5,
    u32 = 6,
    u8 = 7,
    s64 = 8,
    u64 = 9,
    s16 = 10,
    u16 = 11,
    f16 = 12,
    s8 = 13,
    /* Not a reference dtype: widens the enum's range so that the bfloat16
     * storage type of include/af/hip.h (afhip_bf16 = 64) is a valid value. */
    afhip_dtype_ext_ = 64
} af_dtype;

#define AF_MAX_DIMS 4

typedef enum { AF_CONV_DEFAULT = 0, AF_CONV_EXPAND = 1 } af_conv_mode;

typedef enum {
    AF_CONV_AUTO = 0,
    AF_CONV_SPATIAL = 1,
    AF_CONV_FREQ = 2
} af_conv_domain;

/* include/af/defines.h:486-489 */
typedef enum {
    AF_CANNY_THRESHOLD_MANUAL = 0,
    AF_CANNY_THRESHOLD_AUTO_OTSU = 1
} af_canny_threshold;

typedef enum {
    AF_CONV_GRADIENT_DEFAULT = 0,
    AF_CONV_GRADIENT_FILTER = 1,
    AF_CONV_GRADIENT_DATA = 2,
    AF_CONV_GRADIENT_BIAS = 3
} af_conv_gradient_type;
