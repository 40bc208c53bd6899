"""af_harris on the GPU against the oracle's restatement of the reference
(src/api/c/harris.cpp:52-95, src/backend/cpu/harris.cpp:26-148,
src/backend/cpu/kernel/harris.hpp), and the af_features handle functions.

The GPU path rounds every product and sum as the reference does and smooths
through the separable kernel that keeps the reference's T-typed temp and tap
order, so the corner list must equal the restatement's exactly: the same
count, the same (x, y) in the same order (scan order, or the stable
descending sort of the strongest max_corners), the same float scores.

Cases: f32 and f64 images; Gaussian windows (sigma 0.5, 1, 2.5: 3, 5 and 15
taps) and box windows (block_size 3, 8); max_corners 0 with a minimum
response, max_corners below and above the number found; a rectangle image
with known corners; an image too flat to have corners (n = 0); the
reference's argument checks."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _smooth_random(rng, shape, dt):
    import scipy.ndimage as ndi

    return ndi.gaussian_filter(rng.random(shape), 1.5).astype(dt)


CASES = [
    # (shape, dtype, max_corners, min_response, sigma, block_size)
    ((96, 128), np.float32, 0, 1e-30, 1.0, 0),
    ((96, 128), np.float32, 40, 0.0, 1.0, 0),
    ((96, 128), np.float32, 100000, 0.0, 1.0, 0),
    ((130, 75), np.float64, 25, 0.0, 2.5, 0),
    ((130, 75), np.float64, 0, 1e-30, 0.5, 0),
    ((64, 64), np.float32, 30, 0.0, 1.0, 3),
    ((200, 150), np.float32, 60, 0.0, 1.0, 8),
]


def _id(c):
    return "{}x{}_{}_mc{}_s{}_b{}".format(c[0][0], c[0][1], np.dtype(c[1]).name, c[2], c[4], c[5])


@pytest.mark.parametrize("case", CASES, ids=_id)
def test_harris_matches_restatement(af, case):
    shape, dt, mc, mr, sigma, bs = case
    rng = np.random.default_rng(20261202 + len(shape) + mc)
    img = _smooth_random(rng, shape, dt)
    f = af.harris(img, mc, mr, sigma, bs, 0.04)
    x, y, s, o, z = f.arrays()
    rx, ry, rs = oracle.harris(img, mc, mr, sigma, bs, 0.04)
    assert f.n == len(rx) > 0
    np.testing.assert_array_equal(x, rx)
    np.testing.assert_array_equal(y, ry)
    np.testing.assert_array_equal(s, rs)
    assert np.all(o == 0) and np.all(z == 1)


def test_harris_rectangle_corners(af):
    img = np.zeros((64, 80), np.float32)
    img[20:40, 30:60] = 1.0
    f = af.harris(img, 500, 1e5, 1.0, 0, 0.04)
    x, y, s, _, _ = f.arrays()
    assert sorted(zip(x.tolist(), y.tolist())) == [(30.0, 20.0), (30.0, 39.0), (59.0, 20.0), (59.0, 39.0)]


def test_harris_no_corners(af):
    img = np.ones((40, 40), np.float32)
    f = af.harris(img, 0, 1.0, 1.0, 0, 0.04)
    assert f.n == 0
    x, y, s, o, z = f.arrays()
    assert x.size == y.size == s.size == o.size == z.size == 0


@pytest.mark.parametrize("args,code", [
    ((np.zeros((8, 8, 2), np.float32), 10, 0.0, 1.0, 0, 0.04), 203),
    ((np.zeros((32, 32), np.float32), 0, 0.0, 1.0, 0, 0.04), 202),
    ((np.zeros((32, 32), np.float32), 10, 0.0, 1.0, 0, 0.001), 202),
    ((np.zeros((32, 32), np.float32), 10, 0.0, 7.0, 0, 0.04), 202),
    ((np.zeros((32, 32), np.float32), 10, 0.0, 1.0, 40, 0.04), 202),
    ((np.zeros((32, 32), np.int32), 10, 0.0, 1.0, 0, 0.04), 204),
])
def test_harris_argument_errors(af, args, code):
    with pytest.raises(af.AfError) as e:
        af.harris(*args)
    assert e.value.code == code


def test_features_handles(af):
    import ctypes

    L = af.lib()
    h = ctypes.c_void_p()
    af.check(L.af_create_features(ctypes.byref(h), 7))
    n = ctypes.c_longlong()
    af.check(L.af_get_features_num(ctypes.byref(n), h))
    assert n.value == 7
    a = ctypes.c_void_p()
    af.check(L.af_get_features_score(ctypes.byref(a), h))
    d = [ctypes.c_longlong() for _ in range(4)]
    af.check(L.af_get_dims(*[ctypes.byref(v) for v in d], a))
    assert [v.value for v in d] == [7, 1, 1, 1]
    h2 = ctypes.c_void_p()
    af.check(L.af_retain_features(ctypes.byref(h2), h))
    af.check(L.af_release_features(h))
    af.check(L.af_get_features_num(ctypes.byref(n), h2))
    assert n.value == 7
    af.check(L.af_get_features_xpos(ctypes.byref(a), h2))
    af.check(L.af_get_dims(*[ctypes.byref(v) for v in d], a))
    assert d[0].value == 7
    af.check(L.af_release_features(h2))
