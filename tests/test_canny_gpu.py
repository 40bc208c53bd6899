"""af_canny on the GPU against the oracle's restatement of the reference
(src/api/c/canny.cpp:198-285, src/backend/cpu/kernel/sobel.hpp, canny.hpp,
histogram.hpp). Every step rounds as the reference does (the smoothing runs
the separable kernel with the reference's temp and tap order), so the edge
map must equal the restatement's exactly.

Cases: MANUAL thresholds (several low/high pairs) and AUTO_OTSU; the L2 and
the L1 (is_fast) magnitude; f32, f64, u8 and s16 images; a batch of images
(the Otsu threshold per slice, the MANUAL normalisation over the batch); a
rectangle with known edges; the reference's argument checks."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _image(shape, dt, seed):
    import scipy.ndimage as ndi

    rng = np.random.default_rng(seed)
    x = ndi.gaussian_filter(rng.random(shape), (2.0, 2.0) + (0,) * (len(shape) - 2)) * 255.0
    x += (rng.random(shape) > 0.995) * 80.0  # a few bright specks
    return x.astype(dt)


CASES = [
    # (shape, dtype, threshold type, low, high, is_fast)
    ((96, 80), np.float32, 0, 0.1, 0.3, False),
    ((96, 80), np.float32, 0, 0.05, 0.5, True),
    ((96, 80), np.float32, 1, 0.4, 0.0, False),
    ((64, 120), np.float64, 0, 0.2, 0.6, False),
    ((120, 64), np.uint8, 1, 0.5, 0.0, True),
    ((70, 70), np.int16, 0, 0.1, 0.3, False),
    ((60, 50, 3), np.float32, 1, 0.4, 0.0, False),
    ((60, 50, 3), np.float32, 0, 0.1, 0.3, False),
]


def _id(c):
    return "{}_{}_t{}_l{}_h{}_{}".format("x".join(map(str, c[0])), np.dtype(c[1]).name, c[2], c[3], c[4],
                                         "l1" if c[5] else "l2")


@pytest.mark.parametrize("case", CASES, ids=_id)
def test_canny_matches_restatement(af, case):
    shape, dt, ct, lo, hi, fast = case
    img = _image(shape, dt, 20261202 + len(shape) + ct)
    got = af.canny(img, ct, lo, hi, 3, fast)
    assert got.type == af.b8
    g = got.to_numpy().astype(np.uint8)
    ref = oracle.canny(img, ct, lo, hi, 3, fast)
    assert g.shape == ref.shape
    assert ref.sum() > 0
    np.testing.assert_array_equal(g, ref)


def test_canny_rectangle(af):
    img = np.zeros((48, 64), np.float32)
    img[12:36, 16:48] = 255.0
    g = af.canny(img, 0, 0.1, 0.3).to_numpy().astype(np.uint8)
    np.testing.assert_array_equal(g, oracle.canny(img, 0, 0.1, 0.3))


@pytest.mark.parametrize("args,code", [
    ((np.zeros((4, 8), np.float32), 0, 0.1, 0.3, 3), 203),
    ((np.zeros((16, 16), np.float32), 0, 0.1, 0.3, 5), 202),
    ((np.zeros((16, 16), np.complex64), 0, 0.1, 0.3, 3), 204),
])
def test_canny_argument_errors(af, args, code):
    with pytest.raises(af.AfError) as e:
        af.canny(*args)
    assert e.value.code == code
