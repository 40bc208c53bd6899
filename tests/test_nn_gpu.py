"""convolve2_nn forward and both adjoints on the GPU.

Bars (tolerances are written in each test):
  * FP32: output dims exact; within 1e-5 relative L2 of the reference-CPU
    restatement (im2col + SGEMM + col2im) and of the FP64 closed form.
  * BF16 / FP16 MFMA with FP32 accumulation: relative L2 vs FP64 below the
    bound stated per mode (BF16 1e-2, FP16 2e-3); split-BF16 below 1e-5.
  * f64 path: 1e-12 vs the FP64 closed form; FP64 central finite differences
    pin both adjoints on the reduced N=2, C=4, K=4, 8x8, stride 2 instance.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

# (x dims [W,H,C,N], w dims [fw,fh,C,K], stride, pad, dilation)
GEOMS = [
    ((10, 10, 1, 1), (3, 3, 1, 1), (1, 1), (1, 1), (1, 1)),    # sig101011_filt3311_s11_p11_d11
    ((8, 10, 3, 2), (3, 3, 3, 4), (1, 1), (1, 1), (1, 1)),     # sig81032_filt3334_s11_p11_d11
    ((10, 10, 1, 1), (3, 3, 1, 1), (3, 3), (1, 1), (1, 1)),    # s33
    ((10, 10, 1, 1), (5, 5, 1, 1), (5, 5), (1, 1), (1, 1)),    # s55
    ((10, 10, 1, 1), (3, 3, 1, 1), (1, 1), (1, 1), (2, 2)),    # d22
    ((10, 10, 1, 1), (3, 5, 1, 1), (3, 1), (1, 1), (2, 1)),    # filt35 s31 d21
    ((14, 12, 37, 3), (3, 3, 37, 45), (1, 1), (1, 1), (1, 1)),  # C, K not multiples of 32
    ((28, 28, 64, 2), (3, 3, 64, 70), (2, 2), (1, 1), (1, 1)),  # cfg-4-like stride 2
    ((17, 19, 33, 2), (5, 3, 33, 130), (2, 1), (2, 0), (1, 2)),
    ((9, 7, 5, 3), (1, 1, 5, 7), (1, 1), (0, 0), (1, 1)),      # 1x1
]


def _id(g):
    return "x{}_w{}_s{}_p{}_d{}".format(*["x".join(map(str, v)) for v in g])


def _data(rng, g, dt=np.float32):
    x = rng.standard_normal(g[0]).astype(dt)
    fan = g[1][0] * g[1][1] * g[1][2]
    w = (rng.standard_normal(g[1]) * np.sqrt(2.0 / fan)).astype(dt)
    return x, w


@pytest.mark.parametrize("g", GEOMS, ids=_id)
def test_forward_f32(af, rng, g):
    x, w = _data(rng, g)
    y = af.convolve2NN(x, w, g[2], g[3], g[4])
    ref = oracle.conv2_nn(x, w, g[2], g[3], g[4])
    assert y.dims == ref.shape
    got = y.to_numpy()
    assert oracle.rel_l2(got, ref) <= 1e-5
    assert oracle.rel_l2(got, oracle.conv2_nn_direct_f64(x, w, g[2], g[3], g[4])) <= 1e-5


@pytest.mark.parametrize("g", GEOMS, ids=_id)
def test_gradients_f32(af, rng, g):
    x, w = _data(rng, g)
    y = af.convolve2NN(x, w, g[2], g[3], g[4])
    dy = rng.standard_normal(y.dims).astype(np.float32)
    dx = af.convolve2GradientNN(dy, x, w, y, g[2], g[3], g[4], af.AF_CONV_GRADIENT_DATA)
    dw = af.convolve2GradientNN(dy, x, w, y, g[2], g[3], g[4], af.AF_CONV_GRADIENT_FILTER)
    assert dx.dims == tuple(g[0]) and dw.dims == tuple(g[1])
    rdx = oracle.conv2_data_grad(dy, x, w, g[2], g[3], g[4])
    rdw = oracle.conv2_filter_grad(dy, x, w, g[2], g[3], g[4])
    assert oracle.rel_l2(dx.to_numpy(), rdx) <= 1e-5
    assert oracle.rel_l2(dw.to_numpy(), rdw) <= 1e-5


def test_grad_type_default_and_bias_give_data_gradient(af, rng):
    # conv2GradCall: anything but FILTER returns the data gradient (convolve.cpp:401-411)
    g = GEOMS[1]
    x, w = _data(rng, g)
    y = af.convolve2NN(x, w, g[2], g[3], g[4])
    dy = rng.standard_normal(y.dims).astype(np.float32)
    ref = af.convolve2GradientNN(dy, x, w, y, g[2], g[3], g[4], af.AF_CONV_GRADIENT_DATA).to_numpy()
    for t in (af.AF_CONV_GRADIENT_DEFAULT, af.AF_CONV_GRADIENT_BIAS):
        got = af.convolve2GradientNN(dy, x, w, y, g[2], g[3], g[4], t).to_numpy()
        np.testing.assert_array_equal(got, ref)


# (mode, bound on rel-L2 vs FP64) -- the FP32 accumulation is common to all
PRECISIONS = [(1, 1e-2), (2, 2e-3), (3, 1e-5), (4, 1e-5)]


@pytest.mark.parametrize("prec,bound", PRECISIONS, ids=["bf16", "f16", "split", "f32"])
@pytest.mark.parametrize("g", [GEOMS[1], GEOMS[6], GEOMS[7], GEOMS[8]], ids=_id)
def test_reduced_precision_vs_fp64(af, rng, g, prec, bound):
    x, w = _data(rng, g)
    y = af.convolve2NN(x, w, g[2], g[3], g[4], precision=prec)
    y64 = oracle.conv2_nn_direct_f64(x, w, g[2], g[3], g[4])
    e = oracle.rel_l2(y.to_numpy(), y64)
    assert e <= bound, f"fwd rel-L2 {e:.3e} > {bound}"
    dy = rng.standard_normal(y.dims).astype(np.float32)
    dx = af.convolve2GradientNN(dy, x, w, y, g[2], g[3], g[4], af.AF_CONV_GRADIENT_DATA,
                                precision=prec)
    dw = af.convolve2GradientNN(dy, x, w, y, g[2], g[3], g[4], af.AF_CONV_GRADIENT_FILTER,
                                precision=prec)
    e1 = oracle.rel_l2(dx.to_numpy(), oracle.conv2_data_grad_direct_f64(dy, x, w, *g[2:]))
    e2 = oracle.rel_l2(dw.to_numpy(), oracle.conv2_filter_grad_direct_f64(dy, x, w, *g[2:]))
    assert e1 <= bound and e2 <= bound, (e1, e2)


def _bf16_bits(a):
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return r.astype(np.uint16)


def test_bf16_arrays(af, rng, monkeypatch):
    """bfloat16 arrays (afhip_bf16) run BF16 MFMA; output dtype = input dtype.
    With the split count fixed (the autotuner keys the two storage types
    apart and may pick different splits), the same operands give the same
    FP32 bits through either storage type."""
    monkeypatch.setenv("AFHIP_NN_SPLIT", "2")
    g = GEOMS[7]
    x, w = _data(rng, g)
    xb = af.Array(x).astype(af.bf16)
    wb = af.Array(w).astype(af.bf16)
    np.testing.assert_array_equal(xb.to_numpy().ravel(order="F"), _bf16_bits(x).ravel(order="F"))
    y = af.convolve2NN(xb, wb, g[2], g[3], g[4])
    assert y.type == af.bf16
    yf = af.convolve2NN(xb, wb, g[2], g[3], g[4], out_type=af.f32)
    assert yf.type == af.f32
    # same operands as the f32-input BF16 mode: identical FP32 results
    ym = af.convolve2NN(x, w, g[2], g[3], g[4], precision=af.PRECISION_BF16).to_numpy()
    np.testing.assert_array_equal(yf.to_numpy(), ym)
    # bf16 output is the RNE rounding of the FP32 result
    np.testing.assert_array_equal(y.to_numpy(), _bf16_bits(ym))


def test_f16_reference_dtype(af, rng):
    # af_convolve2_nn on f16: GEMM with FP32 accumulation rounded once to
    # half (src/backend/cpu/blas.cpp:334-343); reference tolerance 7e-2 abs
    g = GEOMS[1]
    x, w = _data(rng, g, np.float16)
    y = af.convolve2NN(x, w, g[2], g[3], g[4])
    assert y.type == af.f16
    ref = oracle.conv2_nn(x, w, g[2], g[3], g[4])
    assert np.max(np.abs(y.to_numpy().astype(np.float64) - ref.astype(np.float64))) <= 7e-2
    dy = rng.standard_normal(y.dims).astype(np.float16)
    dx = af.convolve2GradientNN(dy, x, w, y, g[2], g[3], g[4], af.AF_CONV_GRADIENT_DATA)
    rdx = oracle.conv2_data_grad_direct_f64(dy, x, w, *g[2:])
    assert np.max(np.abs(dx.to_numpy().astype(np.float64) - rdx)) <= 7e-2


@pytest.mark.parametrize("g", GEOMS[:6], ids=_id)
def test_f64_path(af, rng, g):
    x, w = _data(rng, g, np.float64)
    y = af.convolve2NN(x, w, g[2], g[3], g[4])
    assert y.type == af.f64
    assert oracle.rel_l2(y.to_numpy(), oracle.conv2_nn_direct_f64(x, w, *g[2:])) <= 1e-12
    dy = rng.standard_normal(y.dims)
    dx = af.convolve2GradientNN(dy, x, w, y, *g[2:], af.AF_CONV_GRADIENT_DATA).to_numpy()
    dw = af.convolve2GradientNN(dy, x, w, y, *g[2:], af.AF_CONV_GRADIENT_FILTER).to_numpy()
    assert oracle.rel_l2(dx, oracle.conv2_data_grad_direct_f64(dy, x, w, *g[2:])) <= 1e-12
    assert oracle.rel_l2(dw, oracle.conv2_filter_grad_direct_f64(dy, x, w, *g[2:])) <= 1e-12


def test_finite_differences_reduced_instance(af, rng):
    """FP64 central differences of L = <conv(x, w), g> on N=2, C=4, K=4,
    8x8, stride 2, pad 1: dL/dx and dL/dw must match the GPU adjoints."""
    st, pd, dl = (2, 2), (1, 1), (1, 1)
    x = rng.standard_normal((8, 8, 4, 2))
    w = rng.standard_normal((3, 3, 4, 4))
    y = af.convolve2NN(x, w, st, pd, dl)
    assert y.dims == (4, 4, 4, 2)
    g = rng.standard_normal(y.dims)
    dx = af.convolve2GradientNN(g, x, w, y, st, pd, dl, af.AF_CONV_GRADIENT_DATA).to_numpy()
    dw = af.convolve2GradientNN(g, x, w, y, st, pd, dl, af.AF_CONV_GRADIENT_FILTER).to_numpy()
    h = 1e-6

    def L(xx, ww):
        return float(np.sum(oracle.conv2_nn_direct_f64(xx, ww, st, pd, dl) * g))

    fdx = np.zeros_like(x)
    for idx in np.ndindex(*x.shape):
        xp, xm = x.copy(), x.copy()
        xp[idx] += h
        xm[idx] -= h
        fdx[idx] = (L(xp, w) - L(xm, w)) / (2 * h)
    fdw = np.zeros_like(w)
    for idx in np.ndindex(*w.shape):
        wp, wm = w.copy(), w.copy()
        wp[idx] += h
        wm[idx] -= h
        fdw[idx] = (L(x, wp) - L(x, wm)) / (2 * h)
    assert oracle.rel_l2(dx, fdx) <= 1e-7
    assert oracle.rel_l2(dw, fdw) <= 1e-7
    # FP32 MFMA adjoints against the same finite differences
    x32, w32, g32 = x.astype(np.float32), w.astype(np.float32), g.astype(np.float32)
    y32 = af.convolve2NN(x32, w32, st, pd, dl)
    dx32 = af.convolve2GradientNN(g32, x32, w32, y32, st, pd, dl, af.AF_CONV_GRADIENT_DATA)
    dw32 = af.convolve2GradientNN(g32, x32, w32, y32, st, pd, dl, af.AF_CONV_GRADIENT_FILTER)
    assert oracle.rel_l2(dx32.to_numpy(), fdx) <= 1e-5
    assert oracle.rel_l2(dw32.to_numpy(), fdw) <= 1e-5


CONSISTENCY = [((10, 10), (3, 3)), ((11, 11), (5, 5)), ((12, 12), (7, 7)), ((19, 19), (9, 9)),
               ((33, 33), (3, 3)), ((255, 255), (3, 3)), ((256, 256), (3, 3)),
               ((257, 257), (3, 3))]


@pytest.mark.parametrize("sd,fd", CONSISTENCY, ids=lambda v: "x".join(map(str, v)))
def test_conv2_consistency(af, rng, sd, fd):
    # test/convolve.cpp:828-894: NN(s1, p=f/2, d1) == spatial DEFAULT, 2e-5 abs
    x = rng.standard_normal(sd).astype(np.float32)
    w = rng.standard_normal(fd).astype(np.float32)
    native = af.convolve2(x, w, af.AF_CONV_DEFAULT, af.AF_CONV_AUTO).to_numpy()
    nn = af.convolve2NN(x, w, (1, 1), (fd[0] // 2, fd[1] // 2), (1, 1)).to_numpy()
    assert native.shape == nn.shape
    assert np.max(np.abs(native - nn)) <= 2e-5


def test_zero_padding_issue2817(af):
    # test/convolve.cpp:1170-1185
    signal = np.ones((5, 5), np.float32)
    filt = np.full((3, 3), 1 / 9, np.float32)
    conv = af.convolve2NN(signal, filt, (1, 1), (0, 0), (1, 1))
    assert conv.dims == (3, 3, 1, 1)
    assert np.sum(np.abs(conv.to_numpy()[:, :, 0, 0] - 1.0)) < 1e-5
    grad = np.full((3, 3), 1 / 9, np.float32)
    dw = af.convolve2GradientNN(grad, signal, filt, conv, (1, 1), (0, 0), (1, 1),
                                af.AF_CONV_GRADIENT_FILTER)
    assert np.sum(np.abs(conv.to_numpy() - dw.to_numpy())) < 1e-5


def test_padding_dims_zero_means_one(af, rng):
    # dim4(0, ptr) fills missing entries with 1 (src/backend/common/dim4.cpp:35-37)
    import ctypes
    x, w = _data(rng, GEOMS[1])
    X, W = af.Array(x), af.Array(w)
    h = ctypes.c_void_p()
    one = af.dims_buf([1, 1])
    af.check(af.lib().af_convolve2_nn(ctypes.byref(h), X.handle, W.handle, 2, one, 0, None, 2, one))
    y = af.Array(handle=h.value)
    ref = oracle.conv2_nn(x, w, (1, 1), (1, 1), (1, 1))
    assert y.dims == ref.shape
    assert oracle.rel_l2(y.to_numpy(), ref) <= 1e-5


def test_error_codes(af, rng):
    x, w = _data(rng, GEOMS[1])
    # channel mismatch -> AF_ERR_SIZE (convolve.cpp:370)
    with pytest.raises(af.AfError) as e:
        af.convolve2NN(x, w[:, :, :2, :].copy(), (1, 1), (1, 1), (1, 1))
    assert e.value.code == af.AF_ERR_SIZE
    # integer dtype -> AF_ERR_TYPE (convolve.cpp:386)
    with pytest.raises(af.AfError) as e:
        af.convolve2NN(x.astype(np.int32), w.astype(np.int32), (1, 1), (1, 1), (1, 1))
    assert e.value.code == af.AF_ERR_TYPE
    # filter dtype differs from signal -> getArray type check (handle.hpp:53-54)
    with pytest.raises(af.AfError) as e:
        af.convolve2NN(x, w.astype(np.float64), (1, 1), (1, 1), (1, 1))
    assert e.value.code == af.AF_ERR_INTERNAL
    # stride of 3 dims -> AF_ERR_ARG (convolve.cpp:365)
    with pytest.raises(af.AfError) as e:
        af.convolve2NN(x, w, (1, 1, 2), (1, 1), (1, 1))
    assert e.value.code == af.AF_ERR_ARG
    # gradient with mismatched incoming-gradient dims -> AF_ERR_SIZE (:434)
    y = af.convolve2NN(x, w, (1, 1), (1, 1), (1, 1))
    bad = np.zeros((3, 3, 4, 2), np.float32)
    with pytest.raises(af.AfError) as e:
        af.convolve2GradientNN(bad, x, w, y, (1, 1), (1, 1), (1, 1), af.AF_CONV_GRADIENT_DATA)
    assert e.value.code == af.AF_ERR_SIZE


def test_separable_arg_errors(af):
    # test/convolve.cpp:309-367: 1-D signal -> AF_ERR_ARG
    sig = np.ones(10, np.float32)
    filt = np.ones(4, np.int32)
    with pytest.raises(af.AfError) as e:
        af.convolve_sep(filt, filt, sig, af.AF_CONV_EXPAND)
    assert e.value.code == af.AF_ERR_ARG
    # non-vector filter -> AF_ERR_ARG
    with pytest.raises(af.AfError) as e:
        af.convolve_sep(np.ones((3, 3), np.float32), np.ones(3, np.float32),
                        np.ones((8, 8), np.float32))
    assert e.value.code == af.AF_ERR_ARG


@pytest.mark.parametrize("prec", [None, 1, 2, 3], ids=["f32", "bf16", "f16", "split"])
@pytest.mark.parametrize("g", GEOMS, ids=_id)
def test_patch_path_matches_gather_path(af, rng, g, prec, monkeypatch):
    """The LDS-patch kernels (forward, stride-1 data gradient) and the gather
    kernels compute the same sums in different orders."""
    monkeypatch.setenv("AFHIP_NN_SMALL", "0")
    x, w = _data(rng, g)
    y = af.convolve2NN(x, w, g[2], g[3], g[4], precision=prec)
    dy = rng.standard_normal(y.dims).astype(np.float32)
    dx = af.convolve2GradientNN(dy, x, w, y, *g[2:], af.AF_CONV_GRADIENT_DATA, precision=prec)
    monkeypatch.setenv("AFHIP_NN_GATHER", "1")
    y2 = af.convolve2NN(x, w, g[2], g[3], g[4], precision=prec)
    dx2 = af.convolve2GradientNN(dy, x, w, y, *g[2:], af.AF_CONV_GRADIENT_DATA, precision=prec)
    assert oracle.rel_l2(y.to_numpy(), y2.to_numpy()) <= 2e-6
    assert oracle.rel_l2(dx.to_numpy(), dx2.to_numpy()) <= 2e-6


WGRAD_GEOMS = GEOMS + [
    ((56, 56, 64, 3), (3, 3, 64, 128), (1, 1), (1, 1), (1, 1)),   # headline layer, N=3
    ((28, 28, 128, 2), (3, 3, 128, 128), (2, 2), (1, 1), (1, 1)),  # stride-2 layer, N=2
    ((30, 9, 70, 2), (3, 3, 70, 66), (2, 1), (0, 2), (1, 2)),
    ((13, 11, 5, 2), (1, 3, 5, 9), (2, 2), (0, 1), (1, 1)),
]


@pytest.mark.parametrize("dt", ["bf16", "f16"])
@pytest.mark.parametrize("g", WGRAD_GEOMS, ids=_id)
def test_filter_gradient_16bit_arrays(af, rng, g, dt, monkeypatch):
    """16-bit arrays take the row-staged filter-gradient kernel where it
    applies: against the FP64 closed form of the same rounded operands (the
    only error left is FP32 accumulation and the final rounding), and against
    the gather kernel."""
    x, w = _data(rng, g)
    t = af.bf16 if dt == "bf16" else af.f16
    X, Wt = af.Array(x).astype(t), af.Array(w).astype(t)
    y = af.convolve2NN(X, Wt, g[2], g[3], g[4], out_type=af.f32)
    dy = rng.standard_normal(y.dims).astype(np.float32)
    DY = af.Array(dy).astype(t)
    dw = af.convolve2GradientNN(DY, X, Wt, y, *g[2:], af.AF_CONV_GRADIENT_FILTER, out_type=af.f32)
    xr, dyr = _rounded(X), _rounded(DY)
    ref = oracle.conv2_filter_grad_direct_f64(dyr, xr, w, *g[2:])
    assert oracle.rel_l2(dw.to_numpy(), ref) <= 1e-5
    monkeypatch.setenv("AFHIP_NN_GATHER", "1")
    dw2 = af.convolve2GradientNN(DY, X, Wt, y, *g[2:], af.AF_CONV_GRADIENT_FILTER, out_type=af.f32)
    assert oracle.rel_l2(dw.to_numpy(), dw2.to_numpy()) <= 2e-6


@pytest.mark.parametrize("g", WGRAD_GEOMS, ids=_id)
def test_filter_gradient_single_buffer_kernel(af, rng, g, monkeypatch):
    """The single-buffer row-staged kernel (the fallback when two row blocks
    do not fit in LDS) on the same geometries, BF16, against the FP64 closed
    form of the rounded operands."""
    monkeypatch.setenv("AFHIP_NN_WGRAD_PC", "0")
    x, w = _data(rng, g)
    X, Wt = af.Array(x).astype(af.bf16), af.Array(w).astype(af.bf16)
    y = af.convolve2NN(X, Wt, g[2], g[3], g[4], out_type=af.f32)
    dy = rng.standard_normal(y.dims).astype(np.float32)
    DY = af.Array(dy).astype(af.bf16)
    dw = af.convolve2GradientNN(DY, X, Wt, y, *g[2:], af.AF_CONV_GRADIENT_FILTER, out_type=af.f32)
    ref = oracle.conv2_filter_grad_direct_f64(_rounded(DY), _rounded(X), w, *g[2:])
    assert oracle.rel_l2(dw.to_numpy(), ref) <= 1e-5


def _rounded(a):
    v = a.to_numpy()
    if v.dtype == np.uint16:
        return (v.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    return v.astype(np.float64)
