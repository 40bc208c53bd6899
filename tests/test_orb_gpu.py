"""af_orb on the GPU against the oracle's restatement of the reference
(src/api/c/orb.cpp:58-89, src/backend/cpu/orb.cpp:44-283,
src/backend/cpu/kernel/orb.hpp, fast.hpp, resize.hpp).

Bars: the reference's own ORB test (test/orb.cpp:196-225) compares x and y
exactly, score, orientation and size within 1e-3, and each descriptor word
within one differing bit. The same bars hold here, on top of which the
feature count, x, y, size and score must match exactly (every step that
produces them rounds as the reference does) and at least 99% of the
descriptor words must be identical: the angle and its sine and cosine come
from the device's math library, so a sample whose rotated offset lands
within an ulp of a half-integer may round the other way.

Cases: f32 and f64 images, with and without the Gaussian blur, scale
factors 1.2 / 1.5 / 2 and level counts that stop at the patch size, a
768x1024 image over five levels, a single-level image, an image too small for a level (no features), and the
reference's argument checks."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _image(shape, dt, seed):
    import scipy.ndimage as ndi

    rng = np.random.default_rng(seed)
    x = ndi.gaussian_filter(rng.random(shape), 2.0)
    x = (x - x.min()) / (x.max() - x.min()) * 255.0
    x += (rng.random(shape) > 0.997) * 60.0
    return x.astype(dt)


CASES = [
    # (shape, dtype, fast_thr, max_feat, scl_fctr, levels, blur)
    ((200, 256), np.float32, 20.0, 400, 1.5, 4, True),
    ((200, 256), np.float32, 20.0, 400, 1.5, 4, False),
    ((240, 180), np.float32, 10.0, 250, 1.2, 8, True),
    ((160, 200), np.float64, 15.0, 300, 2.0, 3, True),
    ((160, 200), np.float64, 15.0, 150, 1.5, 8, False),
    ((60, 70), np.float32, 10.0, 100, 1.5, 4, True),
    # a camera-sized image: five levels, the full feature budget
    ((768, 1024), np.float32, 20.0, 500, 1.5, 5, True),
]


def _id(c):
    return "{}x{}_{}_t{}_m{}_s{}_l{}_{}".format(c[0][0], c[0][1], np.dtype(c[1]).name, c[2], c[3], c[4], c[5],
                                                "blur" if c[6] else "raw")


def _popcount(v):
    v = v.astype(np.uint64)
    c = np.zeros(v.shape, np.int64)
    for b in range(32):
        c += ((v >> np.uint64(b)) & np.uint64(1)).astype(np.int64)
    return c


@pytest.mark.parametrize("case", CASES, ids=_id)
def test_orb_matches_restatement(af, case):
    shape, dt, thr, mf, sf, lv, blur = case
    img = _image(shape, dt, 20261203 + shape[0] + lv)
    f, d = af.orb(img, thr, mf, sf, lv, blur)
    x, y, s, o, z = f.arrays()
    rx, ry, rs, ro, rz, rd = oracle.orb(img, thr, mf, sf, lv, blur)
    assert rx.size > 0
    assert f.n == rx.size <= mf
    assert d.dims[:2] == (8, f.n)
    desc = d.to_numpy()[:, :, 0, 0].T  # word w of feature f at [f, w]
    np.testing.assert_array_equal(x, rx)
    np.testing.assert_array_equal(y, ry)
    np.testing.assert_array_equal(z, rz)
    np.testing.assert_array_equal(s, rs)
    assert np.max(np.abs(o - ro)) <= 1e-3
    ham = _popcount(desc ^ rd.astype(np.uint32))
    assert ham.max() <= 1, f"descriptor words differ by {ham.max()} bits"
    assert np.mean(ham == 0) >= 0.99


def test_orb_too_small_for_a_level(af):
    # min side 40 / 1.5 < 31: no level, no features
    f, d = af.orb(_image((40, 40), np.float32, 5), 20.0, 100, 1.5, 4, False)
    assert f.n == 0
    assert d.elements() == 0
    assert all(a.size == 0 for a in f.arrays())


@pytest.mark.parametrize("args,code", [
    ((np.zeros((64, 64, 2), np.float32), 20.0, 100, 1.5, 4, False), 202),
    ((np.zeros((6, 64), np.float32), 20.0, 100, 1.5, 4, False), 202),
    ((np.zeros((64, 64), np.float32), 0.0, 100, 1.5, 4, False), 202),
    ((np.zeros((64, 64), np.float32), 20.0, 0, 1.5, 4, False), 202),
    ((np.zeros((64, 64), np.float32), 20.0, 100, 1.0, 4, False), 202),
    ((np.zeros((64, 64), np.float32), 20.0, 100, 1.5, 0, False), 202),
    ((np.zeros((64, 64), np.int32), 20.0, 100, 1.5, 4, False), 204),
])
def test_orb_argument_errors(af, args, code):
    with pytest.raises(af.AfError) as e:
        af.orb(*args)
    assert e.value.code == code
