"""The weights-resident convolve2_nn kernel (conv_wres: the weight fragments
of every k-step held in registers, the band's patch gathered once per unit)
on the geometries it serves: 16-bit operands stored as such, 3x3 filters
with dilation 1 along W, up to 128 reduction channels (32 filters per wave
up to 64 channels, 16 above); the forward for W % 8 == 0, the stride-1 data
gradient (dy's K channels reduced) for OW % 8 == 0.

Cases: the benchmarked layer's shape on fewer images, one and two channel
chunks with a tail (C = 16, 48), filters past one 64-filter tile and not a
multiple of it (K = 24, 70, 200), stride 2, padding 0 and 2, dilation 2 along
H, a wide row (200 pixels: 13 input segments), a single tall band, and more
units than CUs (the persistent loop reloading the weights when the filter
tile changes).

Bars: within 1e-5 relative L2 of the FP64 closed form on the same (rounded)
operands with FP32 output; within 1e-6 of the pipelined kernel
(AFHIP_NN_WRES=0) on the same input; 16-bit output within 4e-3."""
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
WRES_GEOMS = [
    ((56, 56, 64, 4), (3, 3, 64, 128), (1, 1), (1, 1), (1, 1)),
    ((56, 56, 16, 2), (3, 3, 16, 24), (1, 1), (1, 1), (1, 1)),
    ((24, 20, 48, 3), (3, 3, 48, 70), (1, 1), (1, 1), (1, 1)),
    ((8, 9, 32, 2), (3, 3, 32, 64), (1, 1), (1, 1), (1, 1)),
    ((40, 33, 64, 2), (3, 3, 64, 128), (2, 2), (1, 1), (1, 1)),
    ((16, 16, 8, 1), (3, 3, 8, 8), (1, 1), (0, 0), (1, 1)),
    ((32, 30, 64, 2), (3, 3, 64, 96), (1, 1), (2, 2), (1, 2)),
    ((48, 12, 64, 4), (3, 3, 64, 200), (1, 2), (1, 0), (1, 1)),
    ((200, 50, 64, 1), (3, 3, 64, 64), (1, 1), (1, 1), (1, 1)),
    ((16, 16, 32, 300), (3, 3, 32, 128), (1, 1), (1, 1), (1, 1)),
    # 16 filters per wave: 3 and 4 channel chunks
    ((32, 32, 128, 4), (3, 3, 128, 128), (2, 2), (1, 1), (1, 1)),
    ((56, 30, 96, 2), (3, 3, 96, 64), (1, 1), (1, 1), (1, 1)),
    ((24, 20, 100, 3), (3, 3, 100, 70), (1, 1), (0, 2), (1, 2)),
]


def _id(g):
    return "x{}_w{}_s{}_p{}_d{}".format(*["x".join(map(str, v)) for v in g])


def _rounded(a):
    v = a.to_numpy()
    if v.dtype == np.uint16:
        return (v.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    return v.astype(np.float64)


@pytest.mark.parametrize("dt", ["bf16", "f16"])
@pytest.mark.parametrize("g", WRES_GEOMS, ids=_id)
def test_wres_forward(af, g, dt, monkeypatch):
    rng = np.random.default_rng(abs(hash(g)) % 2**32)
    x = rng.standard_normal(g[0]).astype(np.float32)
    fan = g[1][0] * g[1][1] * g[1][2]
    w = (rng.standard_normal(g[1]) * np.sqrt(2.0 / fan)).astype(np.float32)
    t = af.bf16 if dt == "bf16" else af.f16
    X, W = af.Array(x).astype(t), af.Array(w).astype(t)
    y = af.convolve2NN(X, W, g[2], g[3], g[4], out_type=af.f32)
    ref = oracle.conv2_nn_direct_f64(_rounded(X), _rounded(W), *g[2:])
    assert y.dims == tuple(ref.shape)
    yn = y.to_numpy()
    assert oracle.rel_l2(yn, ref) <= 1e-5
    monkeypatch.setenv("AFHIP_NN_WRES", "0")
    y0 = af.convolve2NN(X, W, g[2], g[3], g[4], out_type=af.f32).to_numpy()
    e = oracle.rel_l2(yn, y0)
    print(f"{_id(g)} {dt}: vs pipelined {e:.2e} exact={np.array_equal(yn, y0)}", flush=True)
    assert e <= 1e-6
    monkeypatch.delenv("AFHIP_NN_WRES")
    y16 = af.convolve2NN(X, W, g[2], g[3], g[4])
    assert y16.type == t
    assert oracle.rel_l2(_rounded(y16), ref) <= 4e-3


# Data gradient: (x dims, w dims, stride 1, pad, dilation); dy has K channels
# (the reduction: 2, 3 or 4 chunks of 32) and OW % 8 == 0.
WRES_DGRAD_GEOMS = [
    ((56, 56, 64, 4), (3, 3, 64, 128), (1, 1), (1, 1), (1, 1)),
    ((56, 56, 128, 2), (3, 3, 128, 64), (1, 1), (1, 1), (1, 1)),
    ((24, 20, 70, 3), (3, 3, 70, 96), (1, 1), (1, 1), (1, 1)),
    ((16, 16, 24, 300), (3, 3, 24, 100), (1, 1), (1, 1), (1, 1)),
    ((18, 18, 32, 2), (3, 3, 32, 128), (1, 1), (0, 0), (1, 1)),
    ((32, 30, 64, 2), (3, 3, 64, 96), (1, 1), (1, 2), (1, 2)),
    ((200, 20, 96, 1), (3, 3, 96, 64), (1, 1), (1, 1), (1, 1)),
]


@pytest.mark.parametrize("dt", ["bf16", "f16"])
@pytest.mark.parametrize("g", WRES_DGRAD_GEOMS, ids=_id)
def test_wres_data_gradient(af, g, dt, monkeypatch):
    rng = np.random.default_rng(abs(hash(("dgrad",) + g)) % 2**32)
    x = rng.standard_normal(g[0]).astype(np.float32)
    fan = g[1][0] * g[1][1] * g[1][2]
    w = (rng.standard_normal(g[1]) * np.sqrt(2.0 / fan)).astype(np.float32)
    t = af.bf16 if dt == "bf16" else af.f16
    X, W = af.Array(x).astype(t), af.Array(w).astype(t)
    y = af.convolve2NN(X, W, g[2], g[3], g[4], out_type=af.f32)
    assert y.dims[0] % 8 == 0
    dy = rng.standard_normal(y.dims).astype(np.float32)
    DY = af.Array(dy).astype(t)
    dx = af.convolve2GradientNN(DY, X, W, y, *g[2:], af.AF_CONV_GRADIENT_DATA, out_type=af.f32)
    ref = oracle.conv2_data_grad_direct_f64(_rounded(DY), _rounded(X), _rounded(W), *g[2:])
    assert dx.dims == X.dims
    dxn = dx.to_numpy()
    assert oracle.rel_l2(dxn, ref) <= 1e-5
    monkeypatch.setenv("AFHIP_NN_WRES_DGRAD", "0")
    dx0 = af.convolve2GradientNN(DY, X, W, y, *g[2:], af.AF_CONV_GRADIENT_DATA, out_type=af.f32).to_numpy()
    e = oracle.rel_l2(dxn, dx0)
    print(f"dgrad {_id(g)} {dt}: vs pipelined {e:.2e} exact={np.array_equal(dxn, dx0)}", flush=True)
    assert e <= 1e-6
    monkeypatch.delenv("AFHIP_NN_WRES_DGRAD")
    y16 = af.convolve2NN(X, W, g[2], g[3], g[4])
    dx16 = af.convolve2GradientNN(DY, X, W, y16, *g[2:], af.AF_CONV_GRADIENT_DATA)
    assert dx16.type == t
    assert oracle.rel_l2(_rounded(dx16), ref) <= 4e-3
