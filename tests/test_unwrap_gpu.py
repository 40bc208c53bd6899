"""af_unwrap / af_wrap / af_wrap_v2 on the GPU against the oracle.

Bit-exact for all 13 dtypes the reference instantiates (unwrap is a gather;
wrap adds in the reference's patch order in the element type), both patch
orientations, over a spread of test/wrap.cpp's WRAP_INIT geometries; the
documentation example (test/wrap.cpp:200-246); the MaxDim case
(test/wrap.cpp:183-198: 65536 slices along dim 3); wrap_v2 adding into an
existing array; and the argument checks of src/api/c/unwrap.cpp:44-49 and
src/api/c/wrap.cpp:46-63.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

DTYPES = [np.float32, np.float64, np.complex64, np.complex128, np.int32, np.uint32, np.int64,
          np.uint64, np.int16, np.uint16, np.int8, np.uint8, np.bool_]

GEOMS = [(300, 100, 3, 3, 1, 1, 0, 0), (300, 100, 3, 3, 2, 2, 1, 1), (300, 100, 4, 4, 2, 2, 3, 3),
         (300, 100, 8, 8, 8, 8, 7, 7), (300, 100, 31, 31, 8, 8, 15, 15),
         (300, 100, 8, 12, 8, 12, 7, 11), (300, 100, 15, 10, 15, 10, 14, 9), (37, 29, 5, 3, 2, 3, 2, 1)]


def _data(rng, shape, dt, lim):
    dt = np.dtype(dt)
    if dt == np.bool_:
        return rng.integers(0, 2, shape).astype(np.bool_)
    if dt.kind == "c":
        return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dt)
    if dt.kind == "f":
        return rng.standard_normal(shape).astype(dt)
    lo = 0 if dt.kind == "u" else -lim
    return rng.integers(lo, lim, shape).astype(dt)


def _bits(a):
    """Compare one-byte types (bool storage included) as raw bytes."""
    a = np.ascontiguousarray(a)
    return a.view(np.uint8) if a.dtype.itemsize == 1 else a


@pytest.mark.parametrize("dt", DTYPES, ids=lambda d: np.dtype(d).name)
@pytest.mark.parametrize("g", GEOMS, ids=lambda g: "x".join(map(str, g)))
def test_unwrap_wrap_bit_exact(af, dt, g):
    ix, iy, wx, wy, sx, sy, px, py = g
    rng = np.random.default_rng(hash((ix, iy, wx, wy, sx, sy, px, py)) % 2**32)
    x = _data(rng, (ix, iy, 2), dt, max(2, 250 // (wx * wy)))
    for col in (True, False):
        u = af.unwrap(x, wx, wy, sx, sy, px, py, col)
        ur = oracle.unwrap_any(x, wx, wy, sx, sy, px, py, col)
        ug = u.to_numpy()
        assert ug.shape == ur.shape
        np.testing.assert_array_equal(_bits(ug), _bits(ur))
        # wrap a fresh patch array (not just the unwrap output) to exercise sums
        c = _data(rng, ur.shape, dt, 100)
        w = af.wrap(c, ix, iy, wx, wy, sx, sy, px, py, col).to_numpy()
        wr = oracle.wrap_any(c, ix, iy, wx, wy, sx, sy, px, py, col)
        assert w.shape == wr.shape == (ix, iy, 2, 1)
        np.testing.assert_array_equal(_bits(w), _bits(wr))


def test_doc_snippet(af):
    a = np.arange(1, 10, dtype=np.float32).reshape(3, 3, order="F")
    u = af.unwrap(a, 2, 2, 2, 2, 1, 1)
    np.testing.assert_array_equal(u.to_numpy()[:, :, 0, 0],
                                  [[0, 0, 0, 5], [0, 0, 4, 6], [0, 2, 0, 8], [1, 3, 7, 9]])
    np.testing.assert_array_equal(af.wrap(u, 3, 3, 2, 2, 2, 2, 1, 1).to_numpy()[:, :, 0, 0], a)
    b = np.ones((3, 3), np.float32)
    wb = af.wrap(af.unwrap(b, 2, 2, 1, 1), 3, 3, 2, 2, 1, 1).to_numpy()[:, :, 0, 0]
    np.testing.assert_array_equal(wb, [[1, 2, 1], [2, 4, 2], [1, 2, 1]])


def test_max_dim(af):
    """test/wrap.cpp:183-198: range(5, 5, 1, 65536) survives unwrap + wrap."""
    x = np.broadcast_to(np.arange(5, dtype=np.float32)[:, None, None, None], (5, 5, 1, 65536)).copy()
    u = af.unwrap(x, 5, 5, 5, 5, 0, 0)
    assert u.dims == (25, 1, 1, 65536)
    np.testing.assert_array_equal(af.wrap(u, 5, 5, 5, 5, 5, 5, 0, 0).to_numpy(), x)


@pytest.mark.parametrize("dt", [np.float32, np.int32, np.complex64], ids=lambda d: np.dtype(d).name)
def test_wrap_v2_adds_into_existing(af, dt):
    rng = np.random.default_rng(5)
    c = _data(rng, (9, 30, 1), dt, 50)                 # 3x3 windows, 6x5 patch grid on 12x10 (s2 p1)
    base = _data(rng, (12, 10), dt, 50)
    out = af.Array(base)
    r = af.wrap(c, 12, 10, 3, 3, 2, 2, 1, 1, True, out=out)
    assert r is out
    ref = oracle.wrap_any(c, 12, 10, 3, 3, 2, 2, 1, 1, True, out=base)
    np.testing.assert_array_equal(out.to_numpy(), ref)
    # NULL output: a fresh array, as af_wrap
    import ctypes
    h = ctypes.c_void_p()
    ca = af.Array(c)
    af.check(af.lib().af_wrap_v2(ctypes.byref(h), ca.handle, 12, 10, 3, 3, 2, 2, 1, 1, True))
    np.testing.assert_array_equal(af.Array(handle=h.value).to_numpy(),
                                  oracle.wrap_any(c, 12, 10, 3, 3, 2, 2, 1, 1, True))


def test_argument_errors(af):
    x = np.zeros((10, 8), np.float32)
    for args, code in [((0, 3, 1, 1, 0, 0), 202), ((13, 3, 1, 1, 1, 0), 202), ((3, 3, 0, 1, 0, 0), 202),
                       ((3, 3, 1, 1, 3, 0), 202), ((3, 3, 1, 1, -1, 0), 202)]:
        with pytest.raises(af.AfError) as e:
            af.unwrap(x, *args)
        assert e.value.code == code, args
    with pytest.raises(af.AfError) as e:
        af.unwrap(np.zeros((10, 8), np.float16), 3, 3, 1, 1)
    assert e.value.code == 204
    u = af.unwrap(x, 3, 3, 1, 1)             # [9, 48]
    with pytest.raises(af.AfError) as e:
        af.wrap(u, 10, 8, 3, 2, 1, 1)          # patch size mismatch
    assert e.value.code == 203
    with pytest.raises(af.AfError) as e:
        af.wrap(u, 11, 8, 3, 3, 1, 1)          # patch count mismatch
    assert e.value.code == 203
