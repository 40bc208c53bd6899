/*
 * Feature lists of the ArrayFire C API (the output of af_harris): five f32
 * vectors of n elements (x, y, score, orientation, size).
 */
#pragma once
#include <af/defines.h>

/* include/af/features.h:13 */
typedef void *af_features;

/* include/af/features.h:92; src/api/c/features.cpp (af_create_features) */
AFAPI af_err af_create_features(af_features *feat, dim_t num);
/* include/af/features.h:102: a new handle retaining every array */
AFAPI af_err af_retain_features(af_features *out, const af_features feat);
/* include/af/features.h:109 */
AFAPI af_err af_get_features_num(dim_t *num, const af_features feat);
/* include/af/features.h:116-144: the features' own arrays (not retained) */
AFAPI af_err af_get_features_xpos(af_array *out, const af_features feat);
AFAPI af_err af_get_features_ypos(af_array *out, const af_features feat);
AFAPI af_err af_get_features_score(af_array *score, const af_features feat);
AFAPI af_err af_get_features_orientation(af_array *orientation, const af_features feat);
AFAPI af_err af_get_features_size(af_array *size, const af_features feat);
/* include/af/features.h:151 */
AFAPI af_err af_release_features(af_features feat);
