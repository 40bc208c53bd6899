/*
 * The af_array lifecycle subset needed to feed and read the convolution path.
 * Signatures follow the reference declarations cited on each line; each call
 * behaves as the reference's does (new handles are owned by the caller,
 * errors are returned as af_err and the text is kept for af_get_last_error).
 */
#pragma once
#include <af/defines.h>

/* include/af/array.h:1657 -- copies host data into a new device array */
AFAPI af_err af_create_array(af_array *arr, const void *const data,
                             const unsigned ndims, const dim_t *const dims,
                             const af_dtype type);
/* include/af/array.h:1689 -- copies device data to host (synchronises) */
AFAPI af_err af_get_data_ptr(void *data, const af_array arr);
/* include/af/array.h:1696 */
AFAPI af_err af_release_array(af_array arr);
/* include/af/array.h:1701 -- a new handle sharing the same data */
AFAPI af_err af_retain_array(af_array *out, const af_array in);
/* include/af/array.h:1713 -- arrays are always evaluated here; no-op */
AFAPI af_err af_eval(af_array in);
/* include/af/array.h:1744 */
AFAPI af_err af_get_elements(dim_t *elems, const af_array arr);
/* include/af/array.h:1754 */
AFAPI af_err af_get_type(af_dtype *type, const af_array arr);
/* include/af/array.h:1767 */
AFAPI af_err af_get_dims(dim_t *d0, dim_t *d1, dim_t *d2, dim_t *d3,
                         const af_array arr);
/* include/af/array.h:1778 */
AFAPI af_err af_get_numdims(unsigned *result, const af_array arr);
/* include/af/data.h:515 -- array filled with one value */
AFAPI af_err af_constant(af_array *arr, const double val, const unsigned ndims,
                         const dim_t *const dims, const af_dtype type);
/* include/af/exception.h:65 -- *msg is allocated with af_alloc_host */
AFAPI void af_get_last_error(char **msg, dim_t *len);
