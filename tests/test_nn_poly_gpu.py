"""The polyphase strided data gradient (one stride-1 convolution of dy per
residue pair on the patch kernels, scattered into dx) against the FP64
adjoint and against the gather kernel (AFHIP_NN_GATHER=1).

Geometries: the stride-2 gradient layer's shape on fewer images, strides 2
and 3 (and mixed 2 x 3), paddings 0..2, filter sides 1..5 including filters
smaller than the stride (residues no tap reaches: zero gradient), input
sides that are not a multiple of the stride, and positions past the full
convolution's output. FP32 (1e-5 relative L2 vs FP64 and vs the gather
kernel), BF16 and FP16 operands with FP32 results (1e-5 vs FP64 of the
rounded operands)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

# (x dims [W,H,C,N], w dims [fw,fh,C,K], stride, pad)
POLY_GEOMS = [
    ((28, 28, 128, 2), (3, 3, 128, 128), (2, 2), (1, 1)),
    ((13, 11, 8, 2), (3, 3, 8, 12), (2, 3), (0, 1)),
    ((9, 10, 6, 1), (5, 4, 6, 10), (3, 2), (2, 1)),
    ((8, 8, 4, 2), (1, 1, 4, 6), (2, 2), (0, 0)),
    ((7, 9, 5, 2), (2, 3, 5, 7), (2, 2), (1, 0)),
    ((3, 4, 3, 1), (2, 2, 3, 4), (2, 2), (0, 0)),
    ((17, 15, 32, 3), (3, 3, 32, 40), (2, 2), (1, 1)),
    ((16, 16, 24, 2), (4, 4, 24, 16), (2, 2), (1, 1)),
]


def _id(g):
    return "x{}_w{}_s{}_p{}".format(*["x".join(map(str, v)) for v in g])


def _f64(a):
    v = a.to_numpy()
    if v.dtype == np.uint16:
        return (v.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    return v.astype(np.float64)


@pytest.fixture(autouse=True)
def _force_poly(monkeypatch):
    # these layers are small: without the switch the library would choose the
    # gather kernel for them (too few tiles per phase to fill the chip)
    monkeypatch.setenv("AFHIP_NN_POLY", "1")


@pytest.mark.parametrize("dt", ["f32", "bf16", "f16"])
@pytest.mark.parametrize("g", POLY_GEOMS, ids=_id)
def test_poly_data_gradient(af, g, dt, monkeypatch):
    rng = np.random.default_rng(abs(hash(g)) % 2**32)
    xd, wd, st, pd = g
    x = rng.standard_normal(xd).astype(np.float32)
    w = (rng.standard_normal(wd) * np.sqrt(2.0 / (wd[0] * wd[1] * wd[2]))).astype(np.float32)
    t = {"f32": af.f32, "bf16": af.bf16, "f16": af.f16}[dt]
    X, W = af.Array(x), af.Array(w)
    if dt != "f32":
        X, W = X.astype(t), W.astype(t)
    y = af.convolve2NN(X, W, st, pd, (1, 1), out_type=af.f32)
    dy = rng.standard_normal(y.dims).astype(np.float32)
    DY = af.Array(dy) if dt == "f32" else af.Array(dy).astype(t)
    dx = af.convolve2GradientNN(DY, X, W, y, st, pd, (1, 1), af.AF_CONV_GRADIENT_DATA, out_type=af.f32)
    assert dx.dims == xd
    ref = oracle.conv2_data_grad_direct_f64(_f64(DY), _f64(X), _f64(W), st, pd, (1, 1))
    got = dx.to_numpy()
    assert oracle.rel_l2(got, ref) <= 1e-5
    if dt == "f32":
        monkeypatch.delenv("AFHIP_NN_POLY")
        monkeypatch.setenv("AFHIP_NN_GATHER", "1")
        dx0 = af.convolve2GradientNN(DY, X, W, y, st, pd, (1, 1), af.AF_CONV_GRADIENT_DATA).to_numpy()
        assert oracle.rel_l2(got, dx0) <= 1e-5


def test_poly_bf16_output(af):
    """bfloat16 results (the af convention: output dtype = input dtype)."""
    rng = np.random.default_rng(3)
    xd, wd, st, pd = POLY_GEOMS[0]
    x = rng.standard_normal(xd).astype(np.float32)
    w = (rng.standard_normal(wd) * 0.04).astype(np.float32)
    X, W = af.Array(x).astype(af.bf16), af.Array(w).astype(af.bf16)
    y = af.convolve2NN(X, W, st, pd, (1, 1))
    DY = af.Array(rng.standard_normal(y.dims).astype(np.float32)).astype(af.bf16)
    dx = af.convolve2GradientNN(DY, X, W, y, st, pd, (1, 1), af.AF_CONV_GRADIENT_DATA)
    assert dx.type == af.bf16 and dx.dims == xd
    ref = oracle.conv2_data_grad_direct_f64(_f64(DY), _f64(X), _f64(W), st, pd, (1, 1))
    assert oracle.rel_l2(_f64(dx), ref) <= 4e-3
