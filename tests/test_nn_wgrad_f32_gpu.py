"""The FP32 filter gradient on the FP32 matrix cores (wgrad_f32 + the
fixed-order split sum) for the geometries it serves: 1 or 3 taps along each
axis, stride 1 or 2 along W, dilation 1 along W (the reference semantics are
src/backend/cpu/convolve.cpp:213-243).

Cases: the benchmarked layer on two images, the stride-2 gradient layer on
two images, widths that are not multiples of 4 (element-wise staging),
channel and filter counts off the 64-wide tile, 1x1, 3x1 and 1x3 filters,
padding 0 and 2, dilation 2 along H.

Bars: within 1e-5 relative L2 of the reference restatement (im2col + SGEMM)
and of the FP64 closed form; within 2e-6 of the gather kernel
(AFHIP_NN_WGRAD_F32=0)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

GEOMS = [
    ((56, 56, 64, 2), (3, 3, 64, 128), (1, 1), (1, 1), (1, 1)),
    ((28, 28, 128, 2), (3, 3, 128, 128), (2, 2), (1, 1), (1, 1)),
    ((30, 17, 70, 2), (3, 3, 70, 66), (1, 1), (1, 1), (1, 1)),
    ((27, 13, 24, 3), (3, 3, 24, 40), (2, 1), (0, 2), (1, 2)),
    ((16, 12, 33, 2), (1, 1, 33, 65), (1, 1), (0, 0), (1, 1)),
    ((20, 9, 16, 2), (3, 1, 16, 24), (1, 1), (1, 0), (1, 1)),
    ((24, 10, 8, 1), (1, 3, 8, 130), (2, 2), (0, 1), (1, 1)),
]


def _id(g):
    return "x{}_w{}_s{}_p{}_d{}".format(*["x".join(map(str, v)) for v in g])


@pytest.mark.parametrize("g", GEOMS, ids=_id)
def test_filter_gradient_f32(af, rng, g, monkeypatch):
    xd, wd, st, pd, dl = g
    x = rng.standard_normal(xd).astype(np.float32)
    w = (rng.standard_normal(wd) * np.sqrt(2.0 / (wd[0] * wd[1] * wd[2]))).astype(np.float32)
    y = oracle.conv2_nn(x, w, st, pd, dl)
    dy = rng.standard_normal(y.shape).astype(np.float32)
    dw = af.convolve2GradientNN(dy, x, w, y, st, pd, dl, af.AF_CONV_GRADIENT_FILTER)
    assert dw.dims == wd
    got = dw.to_numpy()
    assert oracle.rel_l2(got, oracle.conv2_filter_grad(dy, x, w, st, pd, dl)) <= 1e-5
    assert oracle.rel_l2(got, oracle.conv2_filter_grad_direct_f64(dy, x, w, st, pd, dl)) <= 1e-5
    monkeypatch.setenv("AFHIP_NN_WGRAD_F32", "0")
    ref = af.convolve2GradientNN(dy, x, w, y, st, pd, dl, af.AF_CONV_GRADIENT_FILTER).to_numpy()
    assert oracle.rel_l2(got, ref) <= 2e-6
