"""The pipelined convolve2_nn kernel (16-byte gather transposed in registers,
swizzled LDS, persistent tiles) on geometries it serves: input rows that are a
whole number of 16-byte segments (W % 8 == 0 for 16-bit operands, W % 4 == 0
for FP32). It covers both tile shapes (64 and 128 GEMM rows), one and three
taps per weight group, several channel chunks with a tail, more than one
filter tile, stride, dilation, a negative column pad in the data gradient
and more tiles than CUs (the persistent loop).

Bars (FP32 accumulation of operands stored in the array's type):
  * within 1e-5 relative L2 of the FP64 closed form on the same operands;
  * within 1e-5 of the conv_patch kernel (AFHIP_NN_PIPE=0) on the same input;
  * 16-bit results: within 4e-3 relative L2 (the final rounding).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _patch_kernels_for_small_layers(monkeypatch):
    """These small layers exercise the patch kernels, not the split GEMM that
    the dispatch would otherwise pick for them."""
    monkeypatch.setenv("AFHIP_NN_SMALL", "0")

# (x dims [W,H,C,N], w dims [fw,fh,C,K], stride, pad, dilation)
PIPE_GEOMS = [
    ((16, 12, 40, 3), (3, 3, 40, 24), (1, 1), (1, 1), (1, 1)),    # 64-row tile, 2 chunks with a tail
    ((24, 20, 64, 2), (3, 3, 64, 136), (1, 1), (1, 1), (1, 1)),   # 128-row tiles, K = 136: 2 filter tiles
    ((32, 9, 16, 2), (1, 1, 16, 70), (1, 1), (0, 0), (1, 1)),     # 1x1: one tap per group, one group per tile
    ((40, 24, 35, 2), (3, 3, 35, 50), (2, 2), (1, 1), (1, 1)),    # stride 2
    ((24, 16, 8, 2), (5, 5, 8, 20), (1, 1), (2, 2), (1, 1)),      # 5x5: 25 taps, one per group
    ((32, 16, 24, 1), (3, 3, 24, 32), (1, 1), (2, 2), (2, 2)),    # dilation 2
    ((14, 8, 12, 2), (1, 3, 12, 10), (1, 1), (1, 1), (1, 1)),     # data gradient: dy 16 wide, column pad -1
    ((32, 32, 32, 80), (3, 3, 32, 128), (1, 1), (1, 1), (1, 1)),  # 320 tiles of 256 pixels
    ((32, 32, 32, 160), (3, 3, 32, 64), (1, 1), (1, 1), (1, 1)),  # 320 tiles of 512 pixels
]


def _id(g):
    return "x{}_w{}_s{}_p{}_d{}".format(*["x".join(map(str, v)) for v in g])


def _rounded(a):
    v = a.to_numpy()
    if v.dtype == np.uint16:
        return (v.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    return v.astype(np.float64)


def _arrays(af, rng, g, dt):
    x = rng.standard_normal(g[0]).astype(np.float32)
    fan = g[1][0] * g[1][1] * g[1][2]
    w = (rng.standard_normal(g[1]) * np.sqrt(2.0 / fan)).astype(np.float32)
    t = {"bf16": af.bf16, "f16": af.f16, "f32": af.f32}[dt]
    X, W = af.Array(x), af.Array(w)
    if dt != "f32":
        X, W = X.astype(t), W.astype(t)
    return X, W


@pytest.mark.parametrize("dt", ["bf16", "f16", "f32"])
@pytest.mark.parametrize("g", PIPE_GEOMS, ids=_id)
def test_pipe_forward(af, rng, g, dt, monkeypatch):
    X, W = _arrays(af, rng, g, dt)
    y = af.convolve2NN(X, W, g[2], g[3], g[4], out_type=af.f32)
    ref = oracle.conv2_nn_direct_f64(_rounded(X), _rounded(W), *g[2:])
    assert y.dims == tuple(ref.shape)
    assert oracle.rel_l2(y.to_numpy(), ref) <= 1e-5
    monkeypatch.setenv("AFHIP_NN_PIPE", "0")
    y0 = af.convolve2NN(X, W, g[2], g[3], g[4], out_type=af.f32)
    assert oracle.rel_l2(y.to_numpy(), y0.to_numpy()) <= 1e-5
    if dt != "f32":  # results stored in the operand type
        monkeypatch.delenv("AFHIP_NN_PIPE")
        y16 = af.convolve2NN(X, W, g[2], g[3], g[4])
        assert oracle.rel_l2(_rounded(y16), ref) <= 4e-3


@pytest.mark.parametrize("dt", ["bf16", "f32"])
@pytest.mark.parametrize("g", PIPE_GEOMS, ids=_id)
def test_pipe_data_gradient(af, rng, g, dt, monkeypatch):
    monkeypatch.setenv("AFHIP_NN_WRES_DGRAD", "0")  # the pipelined kernel, not the weights-resident one
    X, W = _arrays(af, rng, g, dt)
    y = af.convolve2NN(X, W, g[2], g[3], g[4], out_type=af.f32)
    dy = rng.standard_normal(y.dims).astype(np.float32)
    DY = af.Array(dy) if dt == "f32" else af.Array(dy).astype(af.bf16)
    dx = af.convolve2GradientNN(DY, X, W, y, *g[2:], af.AF_CONV_GRADIENT_DATA, out_type=af.f32)
    ref = oracle.conv2_data_grad_direct_f64(_rounded(DY), _rounded(X), _rounded(W), *g[2:])
    assert dx.dims == X.dims
    assert oracle.rel_l2(dx.to_numpy(), ref) <= 1e-5
    monkeypatch.setenv("AFHIP_NN_PIPE", "0")
    dx0 = af.convolve2GradientNN(DY, X, W, y, *g[2:], af.AF_CONV_GRADIENT_DATA, out_type=af.f32)
    assert oracle.rel_l2(dx.to_numpy(), dx0.to_numpy()) <= 1e-5


def test_pipe_repeat_deterministic(af, rng):
    """The persistent kernel gives bit-identical results on repeated calls."""
    g = PIPE_GEOMS[7]
    X, W = _arrays(af, rng, g, "bf16")
    a = af.convolve2NN(X, W, g[2], g[3], g[4], out_type=af.f32).to_numpy()
    b = af.convolve2NN(X, W, g[2], g[3], g[4], out_type=af.f32).to_numpy()
    assert np.array_equal(a, b)
