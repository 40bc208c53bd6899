"""The benchmark workloads at their full sizes, checked against the oracle.

Each workload runs at exactly the shape it is timed at, with the seeded
inputs the benchmark uses; where the oracle would take minutes on the whole
output, whole batch slices (or row slabs at segment edges and borders) are
checked instead:

  * 512 x 512, 5x5, DEFAULT, AUTO domain: the whole output;
  * 256 images of 1024 x 1024 with one 7x7 filter (AF_BATCH_LHS): whole
    slices 0 and 255, row slabs of two more;
  * the same images, 17-tap Gaussian column and row filters (sigma 3),
    af_convolve2_sep: whole slices 0, 128 and 255;
  * a 128^3 volume with a 3x3x3 filter: the whole output;
  * convolve2_nn, N=32, C=64, 56x56, K=128, 3x3, s1 p1: forward and data
    gradient in FP32, BF16, FP16 and split-BF16 against FP64 on two images;
    filter gradient over the whole batch; bfloat16 arrays as the benchmark
    runs them;
  * the stride-2 gradient layer, N=16, C=K=128, 28x28, 3x3, s2 p1: all three
    passes in FP32 and BF16 against FP64 over the whole batch.

Bars: FP32 1e-5 relative L2 (against the reference restatement and FP64),
BF16 1e-2, FP16 2e-3, split-BF16 1e-5 against FP64; output dims exact.
"""
import concurrent.futures as cf

import numpy as np
import pytest

import oracle
from slab import slab_check

pytestmark = pytest.mark.gpu

SEED = 20261202
DEFAULT, EXPAND = 0, 1
AUTO, SPATIAL = 0, 1


def _pool(fn, items, workers=16):
    with cf.ThreadPoolExecutor(workers) as ex:
        return list(ex.map(fn, items))


def test_single_image_512_5x5(af):
    rng = np.random.default_rng(SEED + 0)
    x = rng.random((512, 512), dtype=np.float32)
    w = rng.uniform(-1, 1, (5, 5)).astype(np.float32)
    assert not oracle.is_freq_domain(2, (512, 512, 1, 1), (5, 5, 1, 1), AUTO)
    got = af.convolve2(x, w, DEFAULT, AUTO).to_numpy()
    ref = oracle.convolve(x, w, 2, False)
    assert got.shape == ref.shape == (512, 512, 1, 1)
    assert oracle.rel_l2(got, ref) <= 1e-5


@pytest.fixture(scope="module")
def images_1024():
    rng = np.random.default_rng(SEED + 1)
    return rng.random((1024, 1024, 256), dtype=np.float32)


def test_batch_256x1024sq_7x7_lhs(af, images_1024):
    x = images_1024
    w = np.random.default_rng(SEED + 1).standard_normal((7, 7)).astype(np.float32)
    out = af.convolve2(x, w, DEFAULT, SPATIAL)
    assert out.dims == (1024, 1024, 256, 1)
    got = out.to_numpy()
    errs = _pool(lambda k: oracle.rel_l2(got[:, :, k, 0], oracle.convolve(
        np.ascontiguousarray(x[:, :, k]), w, 2, False)[:, :, 0, 0]), [0, 255])
    assert max(errs) <= 1e-5, errs
    for k in (1, 127):
        for y0 in (0, 496, 1008):
            slab_check(got, x, k, lambda xs: oracle.convolve(xs, w, 2, True), 7, 3, 3, y0)


def test_separable_gaussian17_256x1024sq(af, images_1024):
    x = images_1024
    g = oracle.gaussian_kernel(17, 1, 3.0).ravel()
    out = af.convolve_sep(g, g, x, DEFAULT)
    assert out.dims == (1024, 1024, 256, 1)
    got = out.to_numpy()
    errs = _pool(lambda k: oracle.rel_l2(got[:, :, k, 0], oracle.convolve2_sep(
        g, g, np.ascontiguousarray(x[:, :, k]), False)[:, :, 0, 0]), [0, 128, 255])
    assert max(errs) <= 1e-5, errs


def test_volume_128cubed_3x3x3(af):
    rng = np.random.default_rng(SEED + 3)
    x = rng.random((128, 128, 128), dtype=np.float32)
    w = rng.uniform(-1, 1, (3, 3, 3)).astype(np.float32)
    got = af.convolve3(x, w, DEFAULT, AUTO).to_numpy()
    ref = oracle.convolve(x, w, 3, False)
    assert got.shape == ref.shape == (128, 128, 128, 1)
    assert oracle.rel_l2(got, ref) <= 1e-5


# ------------------------------------------------------------ convolve2_nn
def _layer(seed, xd, wd):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(xd).astype(np.float32)
    w = (rng.standard_normal(wd) * np.sqrt(2.0 / (wd[0] * wd[1] * wd[2]))).astype(np.float32)
    return rng, x, w


HEAD = dict(xd=(56, 56, 64, 32), wd=(3, 3, 64, 128), st=(1, 1), pd=(1, 1), dl=(1, 1))
# (precision, bound vs FP64)
PRECS = [(4, 1e-5), (1, 1e-2), (2, 2e-3), (3, 1e-5)]


@pytest.mark.parametrize("prec,bound", PRECS, ids=["f32", "bf16", "f16", "split"])
def test_headline_layer_forward_and_data_gradient(af, prec, bound):
    st, pd, dl = HEAD["st"], HEAD["pd"], HEAD["dl"]
    rng, x, w = _layer(SEED + 2, HEAD["xd"], HEAD["wd"])
    y = af.convolve2NN(x, w, st, pd, dl, precision=prec)
    assert y.dims == (56, 56, 128, 32)
    dy = rng.standard_normal(y.dims).astype(np.float32)
    dx = af.convolve2GradientNN(dy, x, w, y, st, pd, dl, af.AF_CONV_GRADIENT_DATA, precision=prec)
    assert dx.dims == HEAD["xd"]
    yn, dxn = y.to_numpy(), dx.to_numpy()
    n = [0, 31]
    y64 = oracle.conv2_nn_direct_f64(x[..., n], w, st, pd, dl)
    dx64 = oracle.conv2_data_grad_direct_f64(dy[..., n], x[..., n], w, st, pd, dl)
    e0, e1 = oracle.rel_l2(yn[..., n], y64), oracle.rel_l2(dxn[..., n], dx64)
    print(f"headline layer prec {prec}: fwd {e0:.3e} dgrad {e1:.3e}", flush=True)
    assert e0 <= bound and e1 <= bound, (e0, e1)
    if prec == 4:   # against the im2col + SGEMM + col2im restatement too
        assert oracle.rel_l2(yn[..., n], oracle.conv2_nn(x[..., n], w, st, pd, dl)) <= 1e-5
        assert oracle.rel_l2(dxn[..., n], oracle.conv2_data_grad(dy[..., n], x[..., n], w,
                                                                 st, pd, dl)) <= 1e-5


@pytest.mark.parametrize("prec,bound", [(4, 1e-5), (1, 1e-2)], ids=["f32", "bf16"])
def test_headline_layer_filter_gradient(af, prec, bound):
    st, pd, dl = HEAD["st"], HEAD["pd"], HEAD["dl"]
    rng, x, w = _layer(SEED + 2, HEAD["xd"], HEAD["wd"])
    y = af.convolve2NN(x, w, st, pd, dl)
    dy = rng.standard_normal(y.dims).astype(np.float32)
    dw = af.convolve2GradientNN(dy, x, w, y, st, pd, dl, af.AF_CONV_GRADIENT_FILTER,
                                precision=prec).to_numpy()
    ref = oracle.conv2_filter_grad_direct_f64(dy, x, w, st, pd, dl)
    e = oracle.rel_l2(dw, ref)
    print(f"headline layer wgrad prec {prec}: {e:.3e}", flush=True)
    assert dw.shape == HEAD["wd"] and e <= bound


def _bf16_value(a):
    v = a.to_numpy()
    return (v.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def test_headline_layer_bf16_arrays_step(af):
    """The benchmarked step: bfloat16 arrays in, FP32 results, three passes,
    against FP64 of the bf16-rounded operands (only FP32 accumulation and
    the ordering of sums remain) and against FP64 of the FP32 operands (the
    precision claim, 1e-2)."""
    st, pd, dl = HEAD["st"], HEAD["pd"], HEAD["dl"]
    rng, x, w = _layer(SEED + 2, HEAD["xd"], HEAD["wd"])
    X, W = af.Array(x).astype(af.bf16), af.Array(w).astype(af.bf16)
    y = af.convolve2NN(X, W, st, pd, dl, out_type=af.f32)
    dy = rng.standard_normal(y.dims).astype(np.float32)
    DY = af.Array(dy).astype(af.bf16)
    dx = af.convolve2GradientNN(DY, X, W, y, st, pd, dl, af.AF_CONV_GRADIENT_DATA,
                                out_type=af.f32).to_numpy()
    dw = af.convolve2GradientNN(DY, X, W, y, st, pd, dl, af.AF_CONV_GRADIENT_FILTER,
                                out_type=af.f32).to_numpy()
    xr, wr, dyr = _bf16_value(X), _bf16_value(W), _bf16_value(DY)
    n = [0, 17, 31]
    assert oracle.rel_l2(y.to_numpy()[..., n], oracle.conv2_nn_direct_f64(xr[..., n], wr, st, pd, dl)) <= 1e-5
    assert oracle.rel_l2(dx[..., n], oracle.conv2_data_grad_direct_f64(dyr[..., n], xr[..., n], wr,
                                                                      st, pd, dl)) <= 1e-5
    assert oracle.rel_l2(dw, oracle.conv2_filter_grad_direct_f64(dyr, xr, wr, st, pd, dl)) <= 1e-5
    assert oracle.rel_l2(y.to_numpy()[..., n], oracle.conv2_nn_direct_f64(x[..., n], w, st, pd, dl)) <= 1e-2


S2 = dict(xd=(28, 28, 128, 16), wd=(3, 3, 128, 128), st=(2, 2), pd=(1, 1), dl=(1, 1))


@pytest.mark.parametrize("prec,bound", [(4, 1e-5), (1, 1e-2)], ids=["f32", "bf16"])
def test_stride2_layer_all_passes(af, prec, bound):
    st, pd, dl = S2["st"], S2["pd"], S2["dl"]
    rng, x, w = _layer(SEED + 3, S2["xd"], S2["wd"])
    y = af.convolve2NN(x, w, st, pd, dl, precision=prec)
    assert y.dims == (14, 14, 128, 16)
    dy = rng.standard_normal(y.dims).astype(np.float32)
    dx = af.convolve2GradientNN(dy, x, w, y, st, pd, dl, af.AF_CONV_GRADIENT_DATA,
                                precision=prec).to_numpy()
    dw = af.convolve2GradientNN(dy, x, w, y, st, pd, dl, af.AF_CONV_GRADIENT_FILTER,
                                precision=prec).to_numpy()
    e = [oracle.rel_l2(y.to_numpy(), oracle.conv2_nn_direct_f64(x, w, st, pd, dl)),
         oracle.rel_l2(dx, oracle.conv2_data_grad_direct_f64(dy, x, w, st, pd, dl)),
         oracle.rel_l2(dw, oracle.conv2_filter_grad_direct_f64(dy, x, w, st, pd, dl))]
    print(f"stride-2 layer prec {prec}: {['%.3e' % v for v in e]}", flush=True)
    assert max(e) <= bound, e
    if prec == 4:   # and against the reference's own im2col/col2im composition
        assert oracle.rel_l2(dx, oracle.conv2_data_grad(dy, x, w, st, pd, dl)) <= 1e-5
        assert oracle.rel_l2(dw, oracle.conv2_filter_grad(dy, x, w, st, pd, dl)) <= 1e-5
