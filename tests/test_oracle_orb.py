"""Pins the oracle's ORB restatement (src/backend/cpu/orb.cpp,
src/backend/cpu/kernel/orb.hpp, fast.hpp, resize.hpp) before it is trusted.

  * the BRIEF pair table against a digest of the reference's table
    (tests/golden/orb_pattern_digest.json, made by make_orb_pattern_digest.py);
  * bilinear resize: identity at the same size, exact decimation by 2, exact
    reproduction of a linear function when upsampling;
  * FAST: the circle is the radius-3 Bresenham circle; an isolated bright or
    dark dot is the only corner, with the closed-form score 16*(|d| - thr);
    the vectorised windows equal a scalar loop written as the reference's
    incremental window sum with its early exits; the feature cap keeps the
    first corners in scan order;
  * Harris: a linear ramp has det = 0 only if the cross term is right, and
    the response -k*tr^2 with ix along dim1 and iy along dim0;
  * the centroid angle of ramps along dim1, dim0 and their negatives;
  * the descriptor's bit order and rotation on an image whose values order
    pixels by position;
  * the pyramid bookkeeping (level count, per-level budget, sizes).
"""
import hashlib
import json
import math
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))


def test_pattern_matches_reference_digest():
    with open(os.path.join(HERE, "golden", "orb_pattern_digest.json")) as f:
        g = json.load(f)
    p = oracle.orb_pattern().ravel().tolist()
    assert len(p) == g["count"] == 1024
    assert sum(p) == g["sum"] and sum(v * v for v in p) == g["sum_sq"]
    assert hashlib.sha256(",".join(map(str, p)).encode()).hexdigest() == g["sha256"]
    assert p[:4] == g["first_pairs"][0] and p[4:8] == g["first_pairs"][1] and p[-4:] == g["last_pair"]
    # every sample of a 31x31 patch rotated by any angle stays within 19
    assert max(abs(v) for v in p) <= 13


def test_resize_identity_and_decimation():
    rng = np.random.default_rng(1)
    for dt in (np.float32, np.float64):
        x = rng.random((23, 17)).astype(dt)
        np.testing.assert_array_equal(oracle.resize_bilinear(x, 23, 17), x)
        y = rng.random((40, 30)).astype(dt)
        np.testing.assert_array_equal(oracle.resize_bilinear(y, 20, 15), y[::2, ::2])


def test_resize_reproduces_linear_functions():
    i = np.arange(20, dtype=np.float64)[:, None]
    j = np.arange(16, dtype=np.float64)[None, :]
    x = 3.0 * i - 2.0 * j + 1.0
    out = oracle.resize_bilinear(x, 40, 32)
    fx = np.arange(40, dtype=np.float32) / (np.float32(40) / np.float32(20))
    fy = np.arange(32, dtype=np.float32) / (np.float32(32) / np.float32(16))
    want = 3.0 * fx.astype(np.float64)[:, None] - 2.0 * fy.astype(np.float64)[None, :] + 1.0
    # the last source row/column repeats (clamped neighbour), so compare inside
    np.testing.assert_allclose(out[:38, :30], want[:38, :30], rtol=0, atol=1e-6)


def test_fast_circle_is_bresenham_radius_3():
    c = oracle._fast_circle()
    assert len(set(c)) == 16
    assert all(dy * dy + dx * dx in (8, 9, 10) for dy, dx in c)
    for a, b in zip(c, c[1:] + c[:1]):
        assert max(abs(a[0] - b[0]), abs(a[1] - b[1])) == 1


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_fast_isolated_dot(sign):
    img = np.full((40, 50), 100.0, np.float32)
    img[17, 23] = 100.0 + sign * 155.0
    x, y, s = oracle.fast(img, 20.0, 9, True, 0.15, 3)
    assert x.tolist() == [23.0] and y.tolist() == [17.0]
    assert s[0] == np.float32(16 * (155.0 - 20.0))


def _fast_scalar(img, thr, arc, edge, max_feat):
    """The reference's locate_features loop (fast.hpp:51-157) as written:
    early exits on opposite pairs, the incremental window sums."""
    f = np.float32
    thr = f(thr)
    d0, d1 = img.shape
    im = img.astype(np.float32)
    circ = oracle._fast_circle()

    def t(y, x, p):
        v = im[y, x]
        return int(v > p + thr) - int(v < p - thr)

    out = []
    for y in range(edge, d0 - edge):
        for x in range(edge, d1 - edge):
            p = im[y, x]
            d = t(y - 3, x, p) | t(y + 3, x, p)
            if d == 0:
                continue
            for (a, b) in (((-2, 2), (2, -2)), ((0, 3), (0, -3)), ((2, 2), (-2, -2))):
                d &= t(y + a[0], x + a[1], p) | t(y + b[0], x + b[1], p)
            if d == 0:
                continue
            for (a, b) in (((-3, 1), (3, -1)), ((-1, 3), (1, -3)), ((1, 3), (-1, -3)), ((3, 1), (-3, -1))):
                d &= t(y + a[0], x + a[1], p) | t(y + b[0], x + b[1], p)
            if d == 0:
                continue
            vals = [t(y + dy, x + dx, p) for dy, dx in circ]
            s = sum(vals[:arc])
            mx, mn = max(0, s), min(0, s)
            for i in range(arc, 16):
                s += vals[i] - vals[i - arc]
                mx, mn = max(mx, s), min(mn, s)
            for i in range(arc - 1):
                s += vals[i] - vals[16 - arc + i]
                mx, mn = max(mx, s), min(mn, s)
            if mx == arc or mn == -arc:
                sb = f(0)
                sd = f(0)
                for dy, dx in circ:
                    px = im[y + dy, x + dx]
                    sb = f(sb + f(int(px > p + thr)) * f(abs(px - p) - thr))
                    sd = f(sd + f(int(px < p - thr)) * f(abs(p - px) - thr))
                out.append((x, y, max(sb, sd)))
    return out[:max_feat]


def test_fast_matches_reference_loop():
    import scipy.ndimage as ndi

    rng = np.random.default_rng(7)
    img = (ndi.gaussian_filter(rng.random((48, 56)), 1.0) * 600).astype(np.float32)
    ref = _fast_scalar(img, 8.0, 9, 3, 10 ** 9)
    x, y, s = oracle.fast(img, 8.0, 9, False, 1.0, 3)
    assert len(ref) > 20
    assert list(zip(x.astype(int).tolist(), y.astype(int).tolist(), s.tolist())) == \
        [(a, b, float(c)) for a, b, c in ref]


def test_fast_cap_keeps_first_in_scan_order():
    import scipy.ndimage as ndi

    rng = np.random.default_rng(8)
    img = (ndi.gaussian_filter(rng.random((40, 40)), 1.0) * 600).astype(np.float32)
    full = oracle.fast(img, 8.0, 9, False, 1.0, 3)
    cap = int(math.ceil(np.float32(img.size) * np.float32(0.01)))
    part = oracle.fast(img, 8.0, 9, False, 0.01, 3)
    assert full[0].size > cap == part[0].size
    for a, b in zip(full, part):
        np.testing.assert_array_equal(a[:cap], b)


def test_harris_ramp_closed_form():
    # value = 1*column + 2*row: ix = 2 (along dim1), iy = 4 (along dim0)
    d0, d1 = 60, 90
    img = (np.arange(d1)[None, :] * 1.0 + np.arange(d0)[:, None] * 2.0).astype(np.float32)
    xs = np.array([30, 10, 50], np.float32)
    ys = np.array([25, 30, 37], np.float32)
    x, y, s = oracle._orb_harris(img, xs, ys)
    # x = 10 is within 22 of the border and is dropped
    assert x.tolist() == [30.0, 50.0] and y.tolist() == [25.0, 37.0]
    tr = np.float32(49 * 4 + 49 * 16)
    rs = np.float32(0.001)
    rs = rs * rs * rs * rs
    want = (np.float32(0.0) - np.float32(0.04) * (tr * tr)) * rs
    assert np.all(s == want)


def test_centroid_angle_of_ramps():
    d0, d1 = 64, 64
    cols = np.arange(d1, dtype=np.float32)[None, :].repeat(d0, 0)
    rows = np.arange(d0, dtype=np.float32)[:, None].repeat(d1, 1)
    x = np.array([32.0], np.float32)
    y = np.array([30.0], np.float32)
    assert oracle._orb_angle(cols, x, y)[0] == np.float32(0.0)
    assert oracle._orb_angle(rows, x, y)[0] == np.float32(math.pi / 2)
    assert oracle._orb_angle(-cols, x, y)[0] == np.float32(math.pi)
    assert oracle._orb_angle(-rows, x, y)[0] == np.float32(-math.pi / 2)


def test_descriptor_bits_and_rotation():
    d0, d1 = 80, 80
    img = (np.arange(d1)[None, :] * 1000 + np.arange(d0)[:, None]).astype(np.float64)
    pat = oracle.orb_pattern()
    x = np.array([40.0], np.float32)
    y = np.array([38.0], np.float32)

    def bits(desc):
        return np.array([(int(desc[0, k // 32]) >> (k % 32)) & 1 for k in range(256)])

    d, ox, oy, sz = oracle._orb_describe(img, x, y, np.array([0.0], np.float32), np.float32(1.5))
    want = (pat[:, 0] < pat[:, 2]) | ((pat[:, 0] == pat[:, 2]) & (pat[:, 1] < pat[:, 3]))
    np.testing.assert_array_equal(bits(d), want.astype(int))
    assert ox[0] == 60.0 and oy[0] == 57.0 and sz[0] == np.float32(31 * 1.5)
    # a quarter turn maps offset (dx, dy) to (-dy, dx)
    d, _, _, _ = oracle._orb_describe(img, x, y, np.array([np.float32(math.pi / 2)], np.float32), np.float32(1))
    a_x, a_y, b_x, b_y = -pat[:, 1], pat[:, 0], -pat[:, 3], pat[:, 2]
    want = (a_x < b_x) | ((a_x == b_x) & (a_y < b_y))
    np.testing.assert_array_equal(bits(d), want.astype(int))


def test_pyramid_bookkeeping():
    import scipy.ndimage as ndi

    rng = np.random.default_rng(9)
    img = (ndi.gaussian_filter(rng.random((200, 256)), 2.0) * 1000).astype(np.float32)
    x, y, s, o, z, d = oracle.orb(img, 20.0, 300, 1.5, 8, False)
    # min side 200: 133.3, 88.9, 59.3, 39.5 >= 31, then 26.3: four levels
    assert set(z.tolist()) <= {np.float32(31) * oracle._powf(1.5, i)
                               for i in range(4)}
    assert 0 < x.size <= 300 and d.shape == (x.size, 8)
    assert np.all(x < 256) and np.all(y < 200) and np.all(x >= 0) and np.all(y >= 0)
    # the scale is the C library's powf (correctly rounded where numpy's
    # float32 power is not)
    assert oracle._powf(1.2, 3) == np.float32(float(np.float32(1.2)) ** 3)
    # too small for one level: no features
    e = oracle.orb(img[:40, :40], 20.0, 300, 1.5, 8, False)
    assert all(a.size == 0 for a in e)
