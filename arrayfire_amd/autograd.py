"""torch.autograd.Function over the af C ABI: convolve2NN and its two adjoints.

This is a test harness for af_convolve2_nn / af_convolve2_gradient_nn
(src/api/c/convolve.cpp:344-477), not a second implementation: forward calls
af_convolve2_nn (or afhip_convolve2_nn_ex for a reduced precision) and
backward calls af_convolve2_gradient_nn twice (AF_CONV_GRADIENT_DATA and
AF_CONV_GRADIENT_FILTER), all through libafhip.so's exported symbols. The
arithmetic runs on the GPU in the library; torch only records the graph.

Tensors are host (CPU) tensors: each call copies its operands into new
af_arrays (af_create_array) and its results back (af_get_data_ptr), so torch
never initialises a device of its own next to the library's.

Layout: a contiguous row-major torch tensor (N, C, H, W) is the af array
[W, H, C, N] (column-major), and a filter (K, C, fh, fw) is [fw, fh, C, K];
the bytes are the same, only the dims are listed in reverse.

The geometry arguments follow torch's order, (height, width); they are passed
to the library in af's order, (dim0 = width, dim1 = height). af's
convolve2NN is a true convolution (the filter is flipped,
src/backend/cpu/convolve.cpp:141), so it equals
torch.nn.functional.conv2d(x, w.flip(-1, -2), ...).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

import arrayfire_amd as af

_T2AF = {torch.float32: af.f32, torch.float64: af.f64, torch.float16: af.f16,
         torch.bfloat16: af.bf16}
_AF2T = {v: k for k, v in _T2AF.items()}


def _to_af(t: torch.Tensor) -> af.Array:
    """A new af_array holding a copy of a host tensor, dims reversed."""
    if t.device.type != "cpu":
        raise ValueError("the harness takes host tensors; the library owns device memory")
    if t.dtype not in _T2AF:
        raise TypeError(f"unsupported dtype {t.dtype}")
    t = t.detach().contiguous()
    dims = tuple(reversed(t.shape)) or (1,)
    if len(dims) > 4:
        raise ValueError("at most 4 dims")
    raw = t.view(torch.int16) if t.dtype == torch.bfloat16 else t
    buf = np.ascontiguousarray(raw.numpy()).reshape(-1)
    h = ctypes.c_void_p()
    af.check(af.lib().af_create_array(ctypes.byref(h), buf.ctypes.data_as(ctypes.c_void_p), len(dims),
                                      af.dims_buf(dims), _T2AF[t.dtype]))
    return af.Array(handle=h.value)


def _to_torch(a: af.Array) -> torch.Tensor:
    """Host copy of an af array as a torch tensor in (N, C, H, W) order."""
    v = a.to_numpy()                                  # (d0, d1, d2, d3), column-major
    v = np.ascontiguousarray(v.transpose(3, 2, 1, 0))
    t = torch.from_numpy(v)
    return t.view(torch.bfloat16) if a.type == af.bf16 else t


def _af_geom(v):
    """(height, width) -> af's (dim0, dim1) = (width, height)."""
    v = tuple(int(x) for x in (v if isinstance(v, (tuple, list)) else (v, v)))
    return (v[1], v[0])


class Convolve2NN(torch.autograd.Function):
    """y = convolve2NN(x, w): x (N, C, H, W), w (K, C, fh, fw) -> (N, K, OH, OW).

    ``precision`` selects afhip_convolve2_nn_ex's mode for all three passes
    (None: the dtype's own af_convolve2_nn path)."""

    @staticmethod
    def forward(ctx, x, w, stride=(1, 1), padding=(0, 0), dilation=(1, 1), precision=None):
        st, pd, dl = _af_geom(stride), _af_geom(padding), _af_geom(dilation)
        X, W = _to_af(x), _to_af(w)
        Y = af.convolve2NN(X, W, st, pd, dl, precision=precision)
        y = _to_torch(Y)
        ctx.save_for_backward(x, w, y)
        ctx.geom = (st, pd, dl, precision)
        return y

    @staticmethod
    def backward(ctx, gy):
        x, w, y = ctx.saved_tensors
        st, pd, dl, precision = ctx.geom
        G, X, W, Y = _to_af(gy.to(y.dtype)), _to_af(x), _to_af(w), _to_af(y)
        dx = dw = None
        if ctx.needs_input_grad[0]:
            dx = _to_torch(af.convolve2GradientNN(G, X, W, Y, st, pd, dl, af.AF_CONV_GRADIENT_DATA,
                                                  precision=precision))
        if ctx.needs_input_grad[1]:
            dw = _to_torch(af.convolve2GradientNN(G, X, W, Y, st, pd, dl, af.AF_CONV_GRADIENT_FILTER,
                                                  precision=precision))
        return dx, dw, None, None, None, None


def convolve2_nn(x, w, stride=(1, 1), padding=(0, 0), dilation=(1, 1), precision=None):
    return Convolve2NN.apply(x, w, stride, padding, dilation, precision)
