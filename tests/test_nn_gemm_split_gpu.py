"""The split gather GEMM (afhip_k_conv2_nn_gemm) for small and strided
convolve2_nn layers: the reduction split into FP32 partial tiles summed in a
fixed order, and strided data gradients whose pixels are ordered by stride
phase so that only the taps of each phase are visited
(src/backend/cpu/convolve.cpp:123-211 for the semantics).

Cases: stride 2 and 3 (including phases that no tap reaches: stride larger
than the filter), unequal strides, dilation with stride, zero and large
padding, odd sizes, channel counts off the 32-deep slice, 1, 2, 3 and 7
splits (AFHIP_NN_SPLIT), the phased and unphased orders (AFHIP_NN_PHASED),
and the stride-2 gradient layer (N=16, C=K=128, 28x28, 3x3, s2 p1) at full
size through the af_* dispatch in FP32, BF16 and FP16.

Bars: FP32 within 1e-5 relative L2 of the reference restatement and of
FP64; 16-bit operands within 1e-2 (BF16) and 2e-3 (FP16) of FP64; every
split count within 2e-6 of the unsplit result (same products, another
summation order)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

# (x dims [W,H,C,N], w dims [fw,fh,C,K], stride, pad, dilation)
GEOMS = [
    ((28, 28, 40, 2), (3, 3, 40, 70), (2, 2), (1, 1), (1, 1)),
    ((15, 13, 33, 2), (3, 3, 33, 20), (2, 2), (1, 1), (1, 1)),
    ((17, 16, 8, 3), (3, 3, 8, 12), (3, 3), (1, 1), (1, 1)),
    ((19, 11, 16, 2), (1, 1, 16, 24), (3, 2), (0, 0), (1, 1)),     # s > f: phases without taps
    ((22, 18, 12, 2), (5, 3, 12, 40), (2, 3), (2, 1), (1, 1)),
    ((20, 20, 9, 2), (3, 3, 9, 17), (2, 2), (2, 2), (2, 2)),       # dilation with stride
    ((12, 14, 64, 1), (3, 5, 64, 64), (2, 1), (0, 2), (1, 1)),
    ((9, 9, 3, 4), (3, 3, 3, 5), (1, 1), (1, 1), (1, 1)),          # stride 1, split only
]


def _id(g):
    return "x{}_w{}_s{}_p{}_d{}".format(*["x".join(map(str, v)) for v in g])


def _data(rng, g):
    x = rng.standard_normal(g[0]).astype(np.float32)
    fan = g[1][0] * g[1][1] * g[1][2]
    w = (rng.standard_normal(g[1]) * np.sqrt(2.0 / fan)).astype(np.float32)
    return x, w


@pytest.mark.parametrize("g", GEOMS, ids=_id)
def test_forward_and_data_gradient_f32(af, rng, g):
    x, w = _data(rng, g)
    st, pd, dl = g[2], g[3], g[4]
    y = af.convolve2NN(x, w, st, pd, dl)
    ref = oracle.conv2_nn(x, w, st, pd, dl)
    assert y.dims == ref.shape
    assert oracle.rel_l2(y.to_numpy(), ref) <= 1e-5
    assert oracle.rel_l2(y.to_numpy(), oracle.conv2_nn_direct_f64(x, w, st, pd, dl)) <= 1e-5
    dy = rng.standard_normal(ref.shape).astype(np.float32)
    dx = af.convolve2GradientNN(dy, x, w, y, st, pd, dl, af.AF_CONV_GRADIENT_DATA)
    assert dx.dims == x.shape + (1,) * (4 - x.ndim)
    got = dx.to_numpy()
    assert oracle.rel_l2(got, oracle.conv2_data_grad(dy, x, w, st, pd, dl)) <= 1e-5
    assert oracle.rel_l2(got, oracle.conv2_data_grad_direct_f64(dy, x, w, st, pd, dl)) <= 1e-5


@pytest.mark.parametrize("split", ["1", "2", "3", "7"])
@pytest.mark.parametrize("g", GEOMS[:6], ids=_id)
def test_split_counts_agree(af, rng, g, split, monkeypatch):
    x, w = _data(rng, g)
    st, pd, dl = g[2], g[3], g[4]
    monkeypatch.setenv("AFHIP_NN_GATHER", "1")
    monkeypatch.setenv("AFHIP_NN_SPLIT", "1")
    y1 = af.convolve2NN(x, w, st, pd, dl).to_numpy()
    dy = rng.standard_normal(y1.shape).astype(np.float32)
    dx1 = af.convolve2GradientNN(dy, x, w, y1, st, pd, dl, af.AF_CONV_GRADIENT_DATA).to_numpy()
    monkeypatch.setenv("AFHIP_NN_SPLIT", split)
    y = af.convolve2NN(x, w, st, pd, dl).to_numpy()
    dx = af.convolve2GradientNN(dy, x, w, y1, st, pd, dl, af.AF_CONV_GRADIENT_DATA).to_numpy()
    assert oracle.rel_l2(y, y1) <= 2e-6
    assert oracle.rel_l2(dx, dx1) <= 2e-6
    assert oracle.rel_l2(dx, oracle.conv2_data_grad_direct_f64(dy, x, w, st, pd, dl)) <= 1e-5


@pytest.mark.parametrize("g", GEOMS[:7], ids=_id)
def test_phased_matches_unphased(af, rng, g, monkeypatch):
    x, w = _data(rng, g)
    st, pd, dl = g[2], g[3], g[4]
    y = oracle.conv2_nn(x, w, st, pd, dl)
    dy = rng.standard_normal(y.shape).astype(np.float32)
    monkeypatch.setenv("AFHIP_NN_GATHER", "1")
    dxp = af.convolve2GradientNN(dy, x, w, y, st, pd, dl, af.AF_CONV_GRADIENT_DATA).to_numpy()
    monkeypatch.setenv("AFHIP_NN_PHASED", "0")
    dxu = af.convolve2GradientNN(dy, x, w, y, st, pd, dl, af.AF_CONV_GRADIENT_DATA).to_numpy()
    assert oracle.rel_l2(dxp, dxu) <= 2e-6
    ref = oracle.conv2_data_grad_direct_f64(dy, x, w, st, pd, dl)
    assert oracle.rel_l2(dxp, ref) <= 1e-5
    # input pixels that no tap reaches (stride beyond the filter) are exact zeros
    assert np.array_equal(dxp[ref == 0], ref[ref == 0])


@pytest.mark.parametrize("prec,bound", [(None, 1e-5), (1, 1e-2), (2, 2e-3)], ids=["f32", "bf16", "f16"])
def test_stride2_layer_full_size(af, prec, bound):
    rng = np.random.default_rng(20261202 + 3)
    xd, wd = (28, 28, 128, 16), (3, 3, 128, 128)
    st, pd, dl = (2, 2), (1, 1), (1, 1)
    x = rng.standard_normal(xd).astype(np.float32)
    w = (rng.standard_normal(wd) * np.sqrt(2.0 / (128 * 9))).astype(np.float32)
    y = af.convolve2NN(x, w, st, pd, dl, precision=prec)
    assert y.dims == (14, 14, 128, 16)
    dy = rng.standard_normal(y.dims).astype(np.float32)
    dx = af.convolve2GradientNN(dy, x, w, y, st, pd, dl, af.AF_CONV_GRADIENT_DATA, precision=prec)
    dw = af.convolve2GradientNN(dy, x, w, y, st, pd, dl, af.AF_CONV_GRADIENT_FILTER, precision=prec)
    assert dx.dims == xd and dw.dims == wd
    assert oracle.rel_l2(y.to_numpy(), oracle.conv2_nn_direct_f64(x, w, st, pd, dl)) <= bound
    assert oracle.rel_l2(dx.to_numpy(), oracle.conv2_data_grad_direct_f64(dy, x, w, st, pd, dl)) <= bound
    assert oracle.rel_l2(dw.to_numpy(), oracle.conv2_filter_grad_direct_f64(dy, x, w, st, pd, dl)) <= bound


@pytest.mark.parametrize("dt", ["f32", "bf16", "f16"])
@pytest.mark.parametrize("g", GEOMS, ids=_id)
def test_channel_last_copies_match_direct_gather(af, rng, g, dt, monkeypatch):
    """The channel-last path (packed weights, NHWC activation) loads the
    same operand values into the same tiles as the element gather, so both
    give identical bits; channels past a 32-deep slice are zero-filled."""
    x, w = _data(rng, g)
    st, pd, dl = g[2], g[3], g[4]
    t = {"bf16": af.bf16, "f16": af.f16, "f32": af.f32}[dt]
    X, W = af.Array(x).astype(t), af.Array(w).astype(t)
    y0 = af.convolve2NN(X, W, st, pd, dl)
    DY = af.Array(rng.standard_normal(y0.dims).astype(np.float32)).astype(t)
    monkeypatch.setenv("AFHIP_NN_GATHER", "1")
    y = af.convolve2NN(X, W, st, pd, dl).to_numpy()
    dx = af.convolve2GradientNN(DY, X, W, y0, st, pd, dl, af.AF_CONV_GRADIENT_DATA).to_numpy()
    monkeypatch.setenv("AFHIP_NN_CL", "0")
    y1 = af.convolve2NN(X, W, st, pd, dl).to_numpy()
    dx1 = af.convolve2GradientNN(DY, X, W, y0, st, pd, dl, af.AF_CONV_GRADIENT_DATA).to_numpy()
    assert np.array_equal(y, y1)
    assert np.array_equal(dx, dx1)
