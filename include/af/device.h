/*
 * Device-management subset of the ArrayFire C API (one GPU).
 */
#pragma once
#include <af/defines.h>

/* include/af/device.h:307 */
AFAPI af_err af_init(void);
/* include/af/device.h:329 */
AFAPI af_err af_get_device_count(int *num_of_devices);
/* include/af/device.h:344 -- only device 0 exists in this build */
AFAPI af_err af_set_device(const int device);
/* include/af/device.h:349 */
AFAPI af_err af_get_device(int *device);
/* include/af/device.h:354 -- waits for all work queued on the device */
AFAPI af_err af_sync(const int device);
/* include/af/device.h:453 -- wraps a device pointer; the array takes
 * ownership and frees it with hipFree when the last handle is released */
AFAPI af_err af_device_array(af_array *arr, void *data, const unsigned ndims,
                             const dim_t *const dims, const af_dtype type);
/* include/af/device.h:543 -- the array's buffer is no longer freed by the
 * library (also not when the last handle is released) */
AFAPI af_err af_lock_array(const af_array arr);
/* include/af/device.h:553 -- hands the buffer back to the library */
AFAPI af_err af_unlock_array(const af_array arr);
/* include/af/device.h:565 */
AFAPI af_err af_is_locked_array(bool *res, const af_array arr);
/* include/af/device.h:576 -- raw device pointer of an array */
AFAPI af_err af_get_device_ptr(void **ptr, const af_array arr);
/* include/af/device.h (af_alloc_host / af_free_host) -- host memory for
 * strings returned by the library */
AFAPI af_err af_alloc_host(void **ptr, const dim_t bytes);
AFAPI af_err af_free_host(void *ptr);
/* include/af/device.h (af_device_gc): return cached device buffers */
AFAPI af_err af_device_gc(void);
/* include/af/device.h (af_device_mem_info) */
AFAPI af_err af_device_mem_info(size_t *alloc_bytes, size_t *alloc_buffers,
                                size_t *lock_bytes, size_t *lock_buffers);
