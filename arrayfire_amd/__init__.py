"""arrayfire_amd: ArrayFire's convolution path on MI355X (gfx950).

The product is ``lib/libafhip.so``: the af_* C ABI (include/af/*.h) over
hand-written gfx950 kernels. This module is its Python host mirror, shaped
like ArrayFire's C++ API for the same path (src/api/cpp/convolve.cpp:20-90):

    convolve(signal, filter, mode, domain)      -> af::convolve (rank = min ndims)
    convolve1/2/3(signal, filter, mode, domain)  -> af::convolve1/2/3
    convolve_sep(col_filter, row_filter, signal, mode)
                                                 -> af::convolve(col, row, signal)
    convolve2NN(signal, filter, stride, padding, dilation)
    convolve2GradientNN(grad, signal, filter, output, stride, padding,
                        dilation, grad_type)

Arguments and results are ``Array`` handles (af_array owners). Errors raise
``AfError`` carrying the af_err code, as AF_THROW raises af::exception.
Loading fails loudly when the library has not been built: there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libafhip.so")

# af_dtype (include/af/defines.h)
f32, c32, f64, c64, b8, s32, u32, u8, s64, u64, s16, u16, f16, s8 = range(14)
bf16 = 64  # afhip_bf16 (include/af/hip.h)

AF_CONV_DEFAULT, AF_CONV_EXPAND = 0, 1
AF_CONV_AUTO, AF_CONV_SPATIAL, AF_CONV_FREQ = 0, 1, 2
AF_CONV_GRADIENT_DEFAULT, AF_CONV_GRADIENT_FILTER, AF_CONV_GRADIENT_DATA, \
    AF_CONV_GRADIENT_BIAS = 0, 1, 2, 3
PRECISION_DEFAULT, PRECISION_BF16, PRECISION_F16, PRECISION_SPLIT_BF16, \
    PRECISION_F32 = 0, 1, 2, 3, 4

AF_SUCCESS = 0
AF_ERR_ARG, AF_ERR_SIZE, AF_ERR_TYPE = 202, 203, 204
AF_ERR_NOT_SUPPORTED, AF_ERR_INTERNAL = 301, 998

_NP2AF = {np.dtype(np.float32): f32, np.dtype(np.complex64): c32,
          np.dtype(np.float64): f64, np.dtype(np.complex128): c64,
          np.dtype(np.bool_): b8, np.dtype(np.int32): s32,
          np.dtype(np.uint32): u32, np.dtype(np.uint8): u8,
          np.dtype(np.int64): s64, np.dtype(np.uint64): u64,
          np.dtype(np.int16): s16, np.dtype(np.uint16): u16,
          np.dtype(np.float16): f16, np.dtype(np.int8): s8}
_AF2NP = {v: k for k, v in _NP2AF.items()}
_AF2NP[bf16] = np.dtype(np.uint16)  # raw bfloat16 bits

EXPORTED = [
    # include/af/array.h
    "af_create_array", "af_get_data_ptr", "af_release_array", "af_retain_array",
    "af_eval", "af_get_elements", "af_get_type", "af_get_dims", "af_get_numdims",
    "af_constant", "af_get_last_error",
    # include/af/device.h
    "af_init", "af_get_device_count", "af_set_device", "af_get_device", "af_sync",
    "af_device_array", "af_get_device_ptr", "af_alloc_host", "af_free_host",
    "af_device_gc", "af_device_mem_info", "af_lock_array", "af_unlock_array",
    "af_is_locked_array",
    # include/af/signal.h
    "af_convolve1", "af_convolve2", "af_convolve3", "af_convolve2_sep",
    "af_fft_convolve1", "af_fft_convolve2", "af_fft_convolve3", "af_fir", "af_iir",
    # include/af/ml.h
    "af_convolve2_nn", "af_convolve2_gradient_nn",
    # include/af/image.h
    "af_unwrap", "af_wrap", "af_wrap_v2", "af_gaussian_kernel", "af_dog",
    # include/af/hip.h
    "afhip_get_stream", "af_cast", "afhip_convolve2_nn_ex",
    "afhip_convolve2_gradient_nn_ex",
    # include/afhip_kernels.h
    "afhip_k_convolve", "afhip_k_convolve2_sep", "afhip_k_cast", "afhip_k_fill",
    "afhip_k_scale", "afhip_k_scale_complex", "afhip_k_conv2_nn_fwd", "afhip_k_conv2_nn_dgrad",
    "afhip_k_conv2_nn_wgrad_workspace", "afhip_k_conv2_nn_wgrad",
    "afhip_k_conv2_nn_f64", "afhip_k_conv2_nn_patch_workspace", "afhip_k_conv2_nn_patch",
    "afhip_k_conv2_nn_wgrad_rows_workspace", "afhip_k_conv2_nn_wgrad_rows",
    "afhip_ragged_plan_create", "afhip_ragged_plan_run", "afhip_ragged_plan_variant",
    "afhip_ragged_plan_release", "afhip_convolve2_ragged", "afhip_k_ragged_tile",
    "afhip_k_convolve2_ragged", "afhip_k_conv2_nn_patch_pack", "afhip_k_conv2_nn_patch_run",
    "afhip_k_conv2_nn_wgrad_rows_partial", "afhip_k_conv2_nn_wgrad_rows_reduce",
    "afhip_k_fft_convolve_workspace", "afhip_k_fft_convolve",
    "afhip_k_unwrap", "afhip_k_wrap", "afhip_k_region_copy", "afhip_k_sub", "afhip_k_div_cols",
    "afhip_k_iir_workspace", "afhip_k_iir", "afhip_k_conv2_nn_patch_noalloc_limit",
    "afhip_k_conv2_nn_patch_noalloc", "afhip_k_conv2_nn_dgrad_poly_workspace",
    "afhip_k_conv2_nn_dgrad_poly", "afhip_k_conv2_nn_gemm_workspace", "afhip_k_conv2_nn_gemm",
    "afhip_k_conv2_nn_patch_tiles", "afhip_k_conv2_nn_gemm_split_workspace", "afhip_k_conv2_nn_gemm_split",
    "afhip_nn_step_plan_create", "afhip_nn_step_plan_run",
    "afhip_nn_step_plan_release",
    # include/af/features.h, include/af/vision.h
    "af_create_features", "af_retain_features", "af_get_features_num", "af_get_features_xpos",
    "af_get_features_ypos", "af_get_features_score", "af_get_features_orientation",
    "af_get_features_size", "af_release_features", "af_harris",
    "afhip_k_harris_deriv", "afhip_k_harris_nms", "afhip_k_select_flagged_workspace",
    "afhip_k_select_flagged", "afhip_k_sort_desc_workspace", "afhip_k_harris_sort",
    "afhip_k_harris_out", "af_canny", "afhip_k_canny_grad", "afhip_k_canny_nms",
    "afhip_k_canny_minmax", "afhip_k_canny_otsu", "afhip_k_canny_candidates", "afhip_k_canny_seed",
    "afhip_k_canny_sweep", "af_orb", "afhip_k_orb_resize", "afhip_k_orb_fast", "afhip_k_orb_fast_nonmax",
    "afhip_k_orb_harris", "afhip_k_orb_angle", "afhip_k_orb_describe",
]


class AfError(RuntimeError):
    def __init__(self, code: int, msg: str = ""):
        super().__init__(f"af_err {code}: {msg}")
        self.code = code


_lib = None


def lib() -> ctypes.CDLL:
    """The loaded libafhip.so. Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with "
                "`python -m arrayfire_amd.build` (there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        vp, i64p = ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong)
        u = ctypes.c_uint
        L.af_get_last_error.restype = None
        L.af_get_last_error.argtypes = [ctypes.POINTER(ctypes.c_char_p),
                                        ctypes.POINTER(ctypes.c_longlong)]
        L.af_free_host.argtypes = [vp]
        L.af_create_array.argtypes = [ctypes.POINTER(vp), vp, u, i64p, ctypes.c_int]
        L.af_constant.argtypes = [ctypes.POINTER(vp), ctypes.c_double, u, i64p, ctypes.c_int]
        L.af_get_data_ptr.argtypes = [vp, vp]
        L.af_release_array.argtypes = [vp]
        L.af_retain_array.argtypes = [ctypes.POINTER(vp), vp]
        L.af_get_type.argtypes = [ctypes.POINTER(ctypes.c_int), vp]
        L.af_get_dims.argtypes = [i64p, i64p, i64p, i64p, vp]
        L.af_get_numdims.argtypes = [ctypes.POINTER(u), vp]
        L.af_get_elements.argtypes = [i64p, vp]
        L.af_sync.argtypes = [ctypes.c_int]
        L.afhip_get_stream.argtypes = [ctypes.POINTER(vp), ctypes.c_int]
        L.af_get_device_ptr.argtypes = [ctypes.POINTER(vp), vp]
        L.af_device_array.argtypes = [ctypes.POINTER(vp), vp, u, i64p, ctypes.c_int]
        L.af_cast.argtypes = [ctypes.POINTER(vp), vp, ctypes.c_int]
        L.af_lock_array.argtypes = [vp]
        L.af_unlock_array.argtypes = [vp]
        L.af_is_locked_array.argtypes = [ctypes.POINTER(ctypes.c_bool), vp]
        for n in ("af_convolve1", "af_convolve2", "af_convolve3"):
            getattr(L, n).argtypes = [ctypes.POINTER(vp), vp, vp, ctypes.c_int, ctypes.c_int]
        for n in ("af_fft_convolve1", "af_fft_convolve2", "af_fft_convolve3"):
            getattr(L, n).argtypes = [ctypes.POINTER(vp), vp, vp, ctypes.c_int]
        L.af_convolve2_sep.argtypes = [ctypes.POINTER(vp), vp, vp, vp, ctypes.c_int]
        ll6 = [ctypes.c_longlong] * 6
        L.af_unwrap.argtypes = [ctypes.POINTER(vp), vp] + ll6 + [ctypes.c_bool]
        L.af_wrap.argtypes = [ctypes.POINTER(vp), vp] + [ctypes.c_longlong] * 8 + [ctypes.c_bool]
        L.af_wrap_v2.argtypes = [ctypes.POINTER(vp), vp] + [ctypes.c_longlong] * 8 + [ctypes.c_bool]
        L.af_fir.argtypes = [ctypes.POINTER(vp), vp, vp]
        L.af_iir.argtypes = [ctypes.POINTER(vp), vp, vp, vp]
        L.af_dog.argtypes = [ctypes.POINTER(vp), vp, ctypes.c_int, ctypes.c_int]
        L.af_gaussian_kernel.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int,
                                         ctypes.c_double, ctypes.c_double]
        L.af_convolve2_nn.argtypes = [ctypes.POINTER(vp), vp, vp, u, i64p, u, i64p, u, i64p]
        L.afhip_convolve2_nn_ex.argtypes = [ctypes.POINTER(vp), vp, vp, u, i64p, u, i64p, u,
                                            i64p, ctypes.c_int, ctypes.c_int]
        L.af_convolve2_gradient_nn.argtypes = [ctypes.POINTER(vp), vp, vp, vp, vp, u, i64p, u,
                                               i64p, u, i64p, ctypes.c_int]
        L.afhip_convolve2_gradient_nn_ex.argtypes = [ctypes.POINTER(vp), vp, vp, vp, vp, u, i64p,
                                                     u, i64p, u, i64p, ctypes.c_int, ctypes.c_int,
                                                     ctypes.c_int]
        L.afhip_nn_step_plan_create.argtypes = ([ctypes.POINTER(vp)] * 4 + [vp, vp, vp] + [u, i64p] * 3
                                                + [ctypes.c_int, ctypes.c_int])
        L.afhip_nn_step_plan_run.argtypes = [vp]
        L.af_harris.argtypes = [ctypes.POINTER(vp), vp, u, ctypes.c_float, ctypes.c_float, u, ctypes.c_float]
        L.af_create_features.argtypes = [ctypes.POINTER(vp), ctypes.c_longlong]
        L.af_retain_features.argtypes = [ctypes.POINTER(vp), vp]
        L.af_get_features_num.argtypes = [ctypes.POINTER(ctypes.c_longlong), vp]
        for n in ("xpos", "ypos", "score", "orientation", "size"):
            getattr(L, "af_get_features_" + n).argtypes = [ctypes.POINTER(vp), vp]
        L.af_release_features.argtypes = [vp]
        L.af_canny.argtypes = [ctypes.POINTER(vp), vp, ctypes.c_int, ctypes.c_float, ctypes.c_float, u,
                               ctypes.c_bool]
        L.af_orb.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(vp), vp, ctypes.c_float, u, ctypes.c_float, u,
                             ctypes.c_bool]
        L.afhip_nn_step_plan_release.argtypes = [vp]
        ll = ctypes.c_longlong
        L.afhip_k_convolve.argtypes = [ctypes.c_int, vp, i64p, vp, i64p, vp, i64p, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, vp]
        L.afhip_k_convolve2_sep.argtypes = [ctypes.c_int, vp, i64p, vp, i64p, vp, ll, vp, ll,
                                            ctypes.c_int, vp, vp]
        L.afhip_k_cast.argtypes = [ctypes.c_int, vp, ctypes.c_int, vp, ll, vp]
        L.afhip_k_fill.argtypes = [ctypes.c_int, vp, ctypes.c_double, ll, vp]
        for n in ("afhip_k_conv2_nn_fwd", "afhip_k_conv2_nn_dgrad"):
            getattr(L, n).argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp, i64p,
                                      i64p, i64p, i64p, i64p, i64p, vp]
        L.afhip_k_conv2_nn_wgrad_workspace.argtypes = [i64p] * 6
        L.afhip_k_conv2_nn_wgrad_workspace.restype = ll
        L.afhip_k_conv2_nn_wgrad.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp,
                                             i64p, i64p, i64p, i64p, i64p, i64p, vp, ll, vp]
        L.afhip_k_conv2_nn_patch_workspace.argtypes = [ctypes.c_int] * 3 + [i64p] * 6
        L.afhip_k_conv2_nn_patch_workspace.restype = ll
        L.afhip_k_conv2_nn_patch.argtypes = [ctypes.c_int] * 4 + [vp, vp, vp] + [i64p] * 6 + [vp, ll, vp]
        L.afhip_k_conv2_nn_wgrad_rows_workspace.argtypes = [ctypes.c_int] * 2 + [i64p] * 6
        L.afhip_k_conv2_nn_wgrad_rows_workspace.restype = ll
        L.afhip_k_conv2_nn_wgrad_rows.argtypes = [ctypes.c_int] * 3 + [vp, vp, vp] + [i64p] * 6 + [vp, ll, vp]
        L.afhip_k_conv2_nn_dgrad_poly_workspace.argtypes = [ctypes.c_int] * 3 + [i64p] * 6
        L.afhip_k_conv2_nn_dgrad_poly_workspace.restype = ll
        L.afhip_k_conv2_nn_dgrad_poly.argtypes = [ctypes.c_int] * 3 + [vp, vp, vp] + [i64p] * 6 + [vp, ll, vp]
        L.afhip_k_conv2_nn_gemm_workspace.argtypes = [ctypes.c_int] * 3 + [i64p] * 6
        L.afhip_k_conv2_nn_gemm_workspace.restype = ll
        L.afhip_k_conv2_nn_gemm.argtypes = [ctypes.c_int] * 4 + [vp, vp, vp] + [i64p] * 6 + [vp, ll, vp]
        L.afhip_k_conv2_nn_patch_tiles.argtypes = [ctypes.c_int] * 3 + [i64p] * 6
        L.afhip_k_conv2_nn_patch_tiles.restype = ll
        L.afhip_k_conv2_nn_patch_noalloc.argtypes = [ctypes.c_int] * 4 + [vp, vp, vp] + [i64p] * 6 + [vp]
        L.afhip_k_conv2_nn_patch_noalloc_limit.argtypes = []
        L.afhip_k_conv2_nn_patch_noalloc_limit.restype = ll
        L.afhip_k_conv2_nn_patch_pack.argtypes = [ctypes.c_int] * 3 + [vp] + [i64p] * 6 + [vp, ll, vp]
        L.afhip_k_conv2_nn_patch_run.argtypes = [ctypes.c_int] * 4 + [vp, vp] + [i64p] * 6 + [vp, ll, vp]
        L.afhip_k_conv2_nn_wgrad_rows_partial.argtypes = [ctypes.c_int] * 2 + [vp, vp] + [i64p] * 6 + [vp, ll, vp]
        L.afhip_k_conv2_nn_wgrad_rows_reduce.argtypes = [ctypes.c_int] * 3 + [vp] + [i64p] * 6 + [vp, ll, vp]
        L.afhip_k_ragged_tile.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.afhip_k_convolve2_ragged.argtypes = [ctypes.c_int, vp, vp, ll, ctypes.c_int, vp]
        L.afhip_k_conv2_nn_f64.argtypes = [ctypes.c_int, vp, vp, vp] + [i64p] * 6 + [vp]
        L.afhip_k_scale.argtypes = [ctypes.c_int, vp, vp, ctypes.c_double, ll, vp]
        L.afhip_ragged_plan_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(vp),
                                               ctypes.POINTER(vp), ctypes.POINTER(vp),
                                               ctypes.POINTER(ctypes.c_int), u]
        L.afhip_ragged_plan_run.argtypes = [vp]
        L.afhip_ragged_plan_variant.argtypes = [ctypes.POINTER(ctypes.c_int), vp]
        L.afhip_ragged_plan_release.argtypes = [vp]
        L.afhip_convolve2_ragged.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(vp),
                                             ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_int), u]
        _lib = L
    return _lib


def last_error() -> str:
    msg = ctypes.c_char_p()
    n = ctypes.c_longlong()
    lib().af_get_last_error(ctypes.byref(msg), ctypes.byref(n))
    if not msg.value:
        return ""
    s = msg.value.decode(errors="replace")
    lib().af_free_host(ctypes.cast(msg, ctypes.c_void_p))
    return s


def check(code: int) -> None:
    if code != AF_SUCCESS:
        raise AfError(code, last_error())


def dims_buf(d) -> ctypes.Array:
    d = list(d)
    return (ctypes.c_longlong * max(1, len(d)))(*d)


class Array:
    """Owner of one af_array handle (released on garbage collection)."""

    __slots__ = ("_h",)

    def __init__(self, data=None, dims=None, dtype=None, *, handle=None):
        self._h = ctypes.c_void_p(handle)
        if handle is not None:
            return
        a = np.asarray(data)
        if dtype == bf16:
            if a.dtype != np.uint16:
                raise TypeError("bf16 arrays are created from uint16 bit patterns")
            t = bf16
        else:
            if a.dtype not in _NP2AF:
                a = a.astype(np.float32)
            t = _NP2AF[a.dtype]
        if dims is None:
            dims = a.shape if a.ndim else (1,)
        dims = tuple(int(v) for v in dims)
        if len(dims) > 4:
            raise ValueError("at most 4 dims")
        flat = np.ascontiguousarray(a.reshape(-1, order="F")) if a.ndim else a.reshape(1)
        if flat.dtype == np.bool_:
            flat = flat.view(np.int8)
        h = ctypes.c_void_p()
        check(lib().af_create_array(ctypes.byref(h), flat.ctypes.data_as(ctypes.c_void_p),
                                    len(dims), dims_buf(dims), t))
        self._h = h

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    def __del__(self):
        if _lib is not None and self._h and self._h.value:
            _lib.af_release_array(self._h)
            self._h = ctypes.c_void_p()

    @property
    def dims(self) -> tuple:
        d = [ctypes.c_longlong() for _ in range(4)]
        check(lib().af_get_dims(*[ctypes.byref(x) for x in d], self._h))
        return tuple(x.value for x in d)

    @property
    def type(self) -> int:
        t = ctypes.c_int()
        check(lib().af_get_type(ctypes.byref(t), self._h))
        return t.value

    def numdims(self) -> int:
        n = ctypes.c_uint()
        check(lib().af_get_numdims(ctypes.byref(n), self._h))
        return n.value

    def elements(self) -> int:
        d = self.dims
        return d[0] * d[1] * d[2] * d[3]

    def device_ptr(self) -> int:
        p = ctypes.c_void_p()
        check(lib().af_get_device_ptr(ctypes.byref(p), self._h))
        return p.value or 0

    def to_numpy(self) -> np.ndarray:
        """Host copy with shape (d0, d1, d2, d3), element [i0,i1,i2,i3]."""
        d = self.dims
        dt = _AF2NP[self.type]
        out = np.empty(d[0] * d[1] * d[2] * d[3], dtype=np.int8 if dt == np.bool_ else dt)
        if out.size:
            check(lib().af_get_data_ptr(out.ctypes.data_as(ctypes.c_void_p), self._h))
        if dt == np.bool_:
            out = out.view(np.bool_)
        return out.reshape(d, order="F")

    def astype(self, t: int) -> "Array":
        h = ctypes.c_void_p()
        check(lib().af_cast(ctypes.byref(h), self._h, t))
        return Array(handle=h.value)


def constant(value: float, dims, dtype: int = f32) -> Array:
    h = ctypes.c_void_p()
    check(lib().af_constant(ctypes.byref(h), float(value), len(dims), dims_buf(dims), dtype))
    return Array(handle=h.value)


def sync() -> None:
    check(lib().af_sync(-1))


def stream() -> int:
    """The library's HIP stream (afhip_get_stream)."""
    s = ctypes.c_void_p()
    check(lib().afhip_get_stream(ctypes.byref(s), 0))
    return s.value or 0


def _out(fn, *args) -> Array:
    h = ctypes.c_void_p()
    check(fn(ctypes.byref(h), *args))
    return Array(handle=h.value)


def _arr(a) -> Array:
    return a if isinstance(a, Array) else Array(a)


# ------------------------------------------------ af:: convolution functions
# Inputs converted from numpy are bound to locals so that their handles stay
# alive for the duration of the call.
def convolve1(signal, filter, mode=AF_CONV_DEFAULT, domain=AF_CONV_AUTO) -> Array:
    s, f = _arr(signal), _arr(filter)
    return _out(lib().af_convolve1, s.handle, f.handle, mode, domain)


def convolve2(signal, filter, mode=AF_CONV_DEFAULT, domain=AF_CONV_AUTO) -> Array:
    s, f = _arr(signal), _arr(filter)
    return _out(lib().af_convolve2, s.handle, f.handle, mode, domain)


def convolve3(signal, filter, mode=AF_CONV_DEFAULT, domain=AF_CONV_AUTO) -> Array:
    s, f = _arr(signal), _arr(filter)
    return _out(lib().af_convolve3, s.handle, f.handle, mode, domain)


def convolve(signal, filter, mode=AF_CONV_DEFAULT, domain=AF_CONV_AUTO) -> Array:
    """af::convolve: rank = min(numdims) (src/api/cpp/convolve.cpp:20-32)."""
    s, f = _arr(signal), _arr(filter)
    r = min(s.numdims(), f.numdims())
    if r == 1:
        return convolve1(s, f, mode, domain)
    if r == 2:
        return convolve2(s, f, mode, domain)
    return convolve3(s, f, mode, domain)


def convolve_sep(col_filter, row_filter, signal, mode=AF_CONV_DEFAULT) -> Array:
    c, r, s = _arr(col_filter), _arr(row_filter), _arr(signal)
    return _out(lib().af_convolve2_sep, c.handle, r.handle, s.handle, mode)


def _two(v) -> ctypes.Array:
    v = tuple(v)
    return dims_buf(v)


def convolve2NN(signal, filter, stride=(1, 1), padding=(0, 0), dilation=(1, 1),
                precision: int | None = None, out_type: int | None = None) -> Array:
    """af::convolve2NN; like the C++ wrapper it passes 2 dims of each
    parameter (src/api/cpp/convolve.cpp:55-63). ``precision``/``out_type``
    select afhip_convolve2_nn_ex."""
    s, f = _arr(signal), _arr(filter)
    st, pd, dl = _two(stride), _two(padding), _two(dilation)
    if precision is None and out_type is None:
        return _out(lib().af_convolve2_nn, s.handle, f.handle, len(stride), st, len(padding), pd,
                    len(dilation), dl)
    return _out(lib().afhip_convolve2_nn_ex, s.handle, f.handle, len(stride), st, len(padding),
                pd, len(dilation), dl, precision or 0, -1 if out_type is None else out_type)


def convolve2GradientNN(incoming_gradient, original_signal, original_filter, convolved_output,
                        stride=(1, 1), padding=(0, 0), dilation=(1, 1),
                        grad_type=AF_CONV_GRADIENT_DEFAULT, precision: int | None = None,
                        out_type: int | None = None) -> Array:
    g, s, f, o = (_arr(incoming_gradient), _arr(original_signal), _arr(original_filter),
                  _arr(convolved_output))
    st, pd, dl = _two(stride), _two(padding), _two(dilation)
    if precision is None and out_type is None:
        return _out(lib().af_convolve2_gradient_nn, g.handle, s.handle, f.handle, o.handle,
                    len(stride), st, len(padding), pd, len(dilation), dl, grad_type)
    return _out(lib().afhip_convolve2_gradient_nn_ex, g.handle, s.handle, f.handle, o.handle,
                len(stride), st, len(padding), pd, len(dilation), dl, grad_type,
                precision or 0, -1 if out_type is None else out_type)


def unwrap(a, wx, wy, sx, sy, px=0, py=0, is_column=True) -> Array:
    """af::unwrap (src/api/cpp/unwrap.cpp): patches as columns (or rows)."""
    x = _arr(a)
    return _out(lib().af_unwrap, x.handle, wx, wy, sx, sy, px, py, bool(is_column))


def wrap(a, ox, oy, wx, wy, sx, sy, px=0, py=0, is_column=True, out: Array | None = None) -> Array:
    """af::wrap; with ``out`` the patches are added into that array in place
    (af_wrap_v2 with an existing output) and it is returned."""
    x = _arr(a)
    if out is None:
        return _out(lib().af_wrap, x.handle, ox, oy, wx, wy, sx, sy, px, py, bool(is_column))
    h = ctypes.c_void_p(out.handle.value)
    check(lib().af_wrap_v2(ctypes.byref(h), x.handle, ox, oy, wx, wy, sx, sy, px, py,
                           bool(is_column)))
    return out


def fir(b, x) -> Array:
    """af::fir (src/api/cpp/signal.cpp): y = the first x0 rows of
    convolve1(x, b, EXPAND)."""
    bb, xx = _arr(b), _arr(x)
    return _out(lib().af_fir, bb.handle, xx.handle)


def iir(b, a, x) -> Array:
    """af::iir: feedforward b, feedback a (a[0] normalises)."""
    bb, aa, xx = _arr(b), _arr(a), _arr(x)
    return _out(lib().af_iir, bb.handle, aa.handle, xx.handle)


def dog(image, radius1: int, radius2: int) -> Array:
    """af::dog: difference of Gaussian smoothings."""
    x = _arr(image)
    return _out(lib().af_dog, x.handle, int(radius1), int(radius2))


def gaussian_kernel(rows: int, cols: int, sigma_r: float = 0.0, sigma_c: float = 0.0) -> Array:
    """af::gaussianKernel."""
    h = ctypes.c_void_p()
    check(lib().af_gaussian_kernel(ctypes.byref(h), int(rows), int(cols), float(sigma_r),
                                   float(sigma_c)))
    return Array(handle=h.value)


class RaggedPlan:
    """afhip_ragged_plan: a captured persistent launch over many independent
    af_convolve2(signal, filter, mode, AF_CONV_SPATIAL) problems."""

    def __init__(self, signals, filters, modes):
        self._sig = [_arr(s) for s in signals]
        self._flt = [_arr(f) for f in filters]
        n = len(self._sig)
        if len(self._flt) != n or len(modes) != n:
            raise ValueError("signals, filters and modes differ in length")
        VP = ctypes.c_void_p * max(n, 1)
        outs = VP()
        sigs = VP(*[a.handle.value for a in self._sig])
        flts = VP(*[a.handle.value for a in self._flt])
        md = (ctypes.c_int * max(n, 1))(*modes)
        self._plan = ctypes.c_void_p()
        check(lib().afhip_ragged_plan_create(ctypes.byref(self._plan), outs, sigs, flts, md, n))
        self.outputs = [Array(handle=outs[i]) for i in range(n)]

    def run(self) -> None:
        check(lib().afhip_ragged_plan_run(self._plan))

    @property
    def variant(self) -> int:
        v = ctypes.c_int()
        check(lib().afhip_ragged_plan_variant(ctypes.byref(v), self._plan))
        return v.value

    def __del__(self):
        if _lib is not None and getattr(self, "_plan", None) and self._plan.value:
            _lib.afhip_ragged_plan_release(self._plan)
            self._plan = ctypes.c_void_p()


class NNStepPlan:
    """afhip_nn_step_plan: the convolve2_nn forward and both adjoints of one
    training step captured as one hipGraph. ``y``, ``dx`` and ``dw`` are
    rewritten by every run() from the current x, w and dy."""

    def __init__(self, x, w, dy, stride=(1, 1), padding=(0, 0), dilation=(1, 1),
                 precision: int | None = None, out_type: int | None = None):
        self._in = [_arr(x), _arr(w), _arr(dy)]
        st, pd, dl = _two(stride), _two(padding), _two(dilation)
        self._plan = ctypes.c_void_p()
        hs = [ctypes.c_void_p() for _ in range(3)]
        check(lib().afhip_nn_step_plan_create(
            ctypes.byref(self._plan), *[ctypes.byref(h) for h in hs],
            *[a.handle for a in self._in], len(stride), st, len(padding), pd, len(dilation), dl,
            precision or 0, -1 if out_type is None else out_type))
        self.y, self.dx, self.dw = [Array(handle=h.value) for h in hs]

    def run(self) -> None:
        check(lib().afhip_nn_step_plan_run(self._plan))

    def __del__(self):
        if _lib is not None and getattr(self, "_plan", None) and self._plan.value:
            _lib.afhip_nn_step_plan_release(self._plan)
            self._plan = ctypes.c_void_p()


class Features:
    """af_features: x (dim1 index), y (dim0 index), score, orientation and
    size vectors of n features. The arrays belong to the features handle."""

    def __init__(self, handle: int):
        self._h = ctypes.c_void_p(handle)

    @property
    def n(self) -> int:
        v = ctypes.c_longlong()
        check(lib().af_get_features_num(ctypes.byref(v), self._h))
        return v.value

    def _get(self, name: str) -> np.ndarray:
        h = ctypes.c_void_p()
        check(getattr(lib(), "af_get_features_" + name)(ctypes.byref(h), self._h))
        a = Array(handle=h.value)
        try:
            return a.to_numpy().reshape(-1)[: self.n]
        finally:
            a._h = ctypes.c_void_p()  # the features handle owns the array

    def arrays(self):
        """(x, y, score, orientation, size) as numpy vectors."""
        return tuple(self._get(k) for k in ("xpos", "ypos", "score", "orientation", "size"))

    def __del__(self):
        if _lib is not None and getattr(self, "_h", None) and self._h.value:
            _lib.af_release_features(self._h)
            self._h = ctypes.c_void_p()


def harris(image, max_corners: int = 500, min_response: float = 1e5, sigma: float = 1.0,
           block_size: int = 0, k_thr: float = 0.04) -> Features:
    """af::harris (include/af/vision.h:402)."""
    x = _arr(image)
    h = ctypes.c_void_p()
    check(lib().af_harris(ctypes.byref(h), x.handle, int(max_corners), float(min_response),
                          float(sigma), int(block_size), float(k_thr)))
    return Features(h.value)


def canny(image, threshold_type: int = 0, low: float = 0.1, high: float = 0.3, sobel_window: int = 3,
          is_fast: bool = False) -> Array:
    """af::canny (include/af/image.h:735): a b8 edge map."""
    x = _arr(image)
    return _out(lib().af_canny, x.handle, int(threshold_type), float(low), float(high), int(sobel_window),
                bool(is_fast))


def orb(image, fast_thr: float = 20.0, max_feat: int = 400, scl_fctr: float = 1.5, levels: int = 4,
        blur_img: bool = False):
    """af::orb (include/af/vision.h:429): (Features, descriptors), the
    descriptors a u32 Array of dims [8, n]."""
    x = _arr(image)
    h = ctypes.c_void_p()
    d = ctypes.c_void_p()
    check(lib().af_orb(ctypes.byref(h), ctypes.byref(d), x.handle, float(fast_thr), int(max_feat), float(scl_fctr),
                       int(levels), bool(blur_img)))
    return Features(h.value), Array(handle=d.value)


def convolve2_ragged(signals, filters, modes) -> list:
    """One-shot afhip_convolve2_ragged: returns the output Arrays."""
    sig = [_arr(s) for s in signals]
    flt = [_arr(f) for f in filters]
    n = len(sig)
    VP = ctypes.c_void_p * max(n, 1)
    outs = VP()
    check(lib().afhip_convolve2_ragged(outs, VP(*[a.handle.value for a in sig]),
                                       VP(*[a.handle.value for a in flt]),
                                       (ctypes.c_int * max(n, 1))(*modes), n))
    return [Array(handle=outs[i]) for i in range(n)]


def filter(image, kernel) -> Array:  # noqa: A001 - af::filter
    return convolve(image, kernel, AF_CONV_DEFAULT, AF_CONV_AUTO)
