"""Spatial af_convolve1/2/3 on the GPU against the oracle.

Bar: output dims exact; FP32 within 1e-5 relative L2 of the reference-CPU
restatement; integer-valued data (the reference's own ASSERT_EQ fixtures use
such data, test/convolve.cpp:43-96) bit-exact; integer dtypes exact.
"""
import numpy as np
import pytest

import oracle
from slab import slab_check

pytestmark = pytest.mark.gpu

SPATIAL = 1  # AF_CONV_SPATIAL

# (signal dims, filter dims, rank) covering every batch kind, odd and even
# filters, the tiled FP32 2-D sizes and the generic path.
CASES = [
    ((32,), (4,), 1),                 # vector_same-like, even filter
    ((100,), (7,), 1),
    ((257,), (129,), 1),              # long filter
    ((32, 3), (5,), 1),               # LHS
    ((17,), (4, 3), 1),               # RHS
    ((20, 3), (5, 3), 1),             # SAME
    ((15, 17), (5, 5), 2),            # rectangle
    ((15, 17), (5, 5, 2), 2),         # rectangle_same_one2many (RHS)
    ((64, 48, 3), (3, 3), 2),         # LHS
    ((33, 31, 2, 2), (4, 2, 2, 2), 2),  # SAME with even filter
    ((300, 200), (7, 7), 2),
    ((129, 65, 2), (9, 9), 2),
    ((33, 31), (2, 4), 2),
    ((40, 30), (11, 3), 2),           # generic 2-D
    ((130, 70), (1, 1), 2),
    # rows of whole 16-byte vectors with odd square filters take the strip
    # kernel: partial last strip, several 256-row segments, every batch kind
    ((260, 530, 2), (7, 7), 2),
    ((132, 300), (9, 9), 2),
    ((68, 40), (5, 5, 3), 2),         # RHS
    ((36, 20, 2, 2), (3, 3, 2, 2), 2),  # SAME
    ((1024, 24), (7, 7), 2),
    ((64, 600), (5, 5), 2),
    ((128, 300), (7, 3), 2),          # odd non-square filters: strip kernel too
    ((64, 70, 2), (3, 9), 2),         # LHS
    ((100, 40), (9, 5, 2), 2),        # RHS
    ((68, 300), (5, 7), 2),
    ((10, 11, 2, 2), (4, 2, 2, 2), 3),  # cuboid_same_many2many shape
    ((16, 16, 16), (3, 3, 3), 3),
    ((12, 10, 8, 2), (3, 3, 3), 3),   # LHS
    ((9, 8, 7), (3, 2, 4, 3), 3),     # RHS
    # 3x3x3 takes the FP32 3-D tile kernel: partial blocks in every dimension,
    # rows that are and are not whole 16-byte vectors, every batch kind
    ((70, 20, 19), (3, 3, 3), 3),
    ((45, 33, 20), (3, 3, 3), 3),
    ((9, 8, 7), (3, 3, 3, 2), 3),     # RHS
    ((10, 11, 6, 2), (3, 3, 3, 2), 3),  # SAME
]


def _ids(c):
    return f"s{'x'.join(map(str, c[0]))}_f{'x'.join(map(str, c[1]))}_r{c[2]}"


def _call(af, rank, x, w, mode):
    fn = {1: af.convolve1, 2: af.convolve2, 3: af.convolve3}[rank]
    return fn(x, w, mode, SPATIAL)


def _ref(x, w, rank, expand):
    return oracle.convolve_dispatch(x, w, rank, expand)


@pytest.mark.parametrize("case", CASES, ids=_ids)
@pytest.mark.parametrize("mode", [0, 1], ids=["default", "expand"])
def test_f32_random(af, rng, case, mode):
    sd, fd, rank = case
    x = rng.random(sd).astype(np.float32)
    w = rng.uniform(-1, 1, fd).astype(np.float32)
    out = _call(af, rank, x, w, mode)
    ref = _ref(x, w, rank, mode == 1)
    got = out.to_numpy()
    assert got.shape == ref.shape
    assert oracle.rel_l2(got, ref) <= 1e-5


@pytest.mark.parametrize("case", CASES, ids=_ids)
@pytest.mark.parametrize("mode", [0, 1], ids=["default", "expand"])
def test_f32_integer_valued_bit_exact(af, rng, case, mode):
    sd, fd, rank = case
    x = rng.integers(0, 10, sd).astype(np.float32)
    w = rng.integers(-3, 4, fd).astype(np.float32)
    got = _call(af, rank, x, w, mode).to_numpy()
    ref = _ref(x, w, rank, mode == 1)
    assert got.shape == ref.shape
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("case", CASES[::3], ids=_ids)
def test_f64(af, rng, case):
    sd, fd, rank = case
    x = rng.random(sd)
    w = rng.uniform(-1, 1, fd)
    got = _call(af, rank, x, w, 0).to_numpy()
    ref = _ref(x, w, rank, False)
    assert got.shape == ref.shape
    assert oracle.rel_l2(got, ref) <= 1e-13


@pytest.mark.parametrize("case", CASES[::3], ids=_ids)
def test_int32_exact(af, rng, case):
    # int32 signal, fractional float filter: the float sum truncated to int is
    # the reference's value bit for bit (each product rounded before the add)
    sd, fd, rank = case
    x = rng.integers(0, 50, sd).astype(np.int32)
    w = rng.uniform(-1, 1, fd).astype(np.float32)
    got = _call(af, rank, x, w, 1).to_numpy()
    ref = _ref(x, w, rank, True)
    assert got.dtype == np.int32 and got.shape == ref.shape
    np.testing.assert_array_equal(got, ref)


def test_uint8_filter_cast(af, rng):
    # int filter is cast to accT = float (src/api/c/convolve.cpp:46)
    # sums stay below 256 so the float -> uint8 conversion is in range
    x = rng.integers(0, 10, (40, 30)).astype(np.uint8)
    w = rng.integers(0, 2, (3, 3)).astype(np.int32)
    got = af.convolve2(x, w, 0, SPATIAL).to_numpy()
    ref = oracle.convolve(x, w.astype(np.float32), 2, False)
    np.testing.assert_array_equal(got, ref)


def test_complex_linearity(af, rng):
    # complex conv == combination of four real convolutions (no complex
    # oracle: pinned by bilinearity against the FP64 real oracle)
    xr, xi = rng.random((30, 20)), rng.random((30, 20))
    wr, wi = rng.uniform(-1, 1, (5, 3)), rng.uniform(-1, 1, (5, 3))
    x = (xr + 1j * xi).astype(np.complex64)
    w = (wr + 1j * wi).astype(np.complex64)
    got = af.convolve2(x, w, 1, SPATIAL).to_numpy()
    c = lambda a, b: oracle.convolve(a.astype(np.float64), b.astype(np.float64), 2, True)  # noqa: E731
    xr32, xi32 = x.real.astype(np.float64), x.imag.astype(np.float64)
    wr32, wi32 = w.real.astype(np.float64), w.imag.astype(np.float64)
    ref = (c(xr32, wr32) - c(xi32, wi32)) + 1j * (c(xr32, wi32) + c(xi32, wr32))
    assert got.shape == ref.shape
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= 1e-5


# ---------------------------------------------------------- known answers
def test_known_answer_1d(af):
    # test/convolve.cpp:522-535 (values printed with 4 decimals)
    a = np.array([0.0000, 0.1315, 0.7556, 0.4587, 0.5328, 0.2190, 0.0470, 0.6789,
                  0.6793, 0.9347], np.float32)
    b = np.array([0.3835, 0.5194, 0.8310, 0.0346], np.float32)
    c = np.array([0.3581, 0.6777, 1.0750, 0.7679, 0.5903, 0.4851, 0.6598, 1.2770,
                  1.0734, 0.8002])
    got = af.convolve(a, b).to_numpy().ravel()
    np.testing.assert_allclose(got, c, atol=1.5e-4)


def test_known_answer_2d(af):
    # test/convolve.cpp:537-562: constant(0.5,5,5) conv constant(1,2,2)
    got = af.convolve(np.full((5, 5), 0.5, np.float32), np.ones((2, 2), np.float32)).to_numpy()
    exp = np.full((5, 5), 2.0)
    exp[4, :] = 1.0
    exp[:, 4] = 1.0
    exp[4, 4] = 0.5
    np.testing.assert_array_equal(got[:, :, 0, 0], exp)


def test_known_answer_3d(af):
    # test/convolve.cpp:564-620: constant(1,4,4,4) conv constant(0.5,2,2,2)
    got = af.convolve(np.ones((4, 4, 4), np.float32), np.full((2, 2, 2), 0.5, np.float32))
    got = got.to_numpy()[:, :, :, 0]
    # in-range taps per axis: 2 everywhere except the last index
    per = np.where(np.arange(4) < 3, 2, 1)
    exp = 0.5 * per[:, None, None] * per[None, :, None] * per[None, None, :]
    np.testing.assert_array_equal(got, exp)


# ------------------------------------------------------ dispatch semantics
def test_batch_equivalence_lhs_rhs_same(af, rng):
    x = rng.random((21, 19, 3)).astype(np.float32)
    w = rng.uniform(-1, 1, (5, 5, 3)).astype(np.float32)
    same = af.convolve2(x, w, 0, SPATIAL).to_numpy()
    for k in range(3):
        one = af.convolve2(np.ascontiguousarray(x[:, :, k]), np.ascontiguousarray(w[:, :, k]),
                           0, SPATIAL).to_numpy()
        np.testing.assert_array_equal(same[:, :, k, 0], one[:, :, 0, 0])
    lhs = af.convolve2(x, w[:, :, 0], 1, SPATIAL).to_numpy()
    rhs = af.convolve2(x[:, :, 0], w, 1, SPATIAL).to_numpy()
    assert lhs.shape == (25, 23, 3, 1) and rhs.shape == (25, 23, 3, 1)
    for k in range(3):
        one = af.convolve2(np.ascontiguousarray(x[:, :, k]), w[:, :, 0], 1, SPATIAL).to_numpy()
        np.testing.assert_array_equal(lhs[:, :, k, 0], one[:, :, 0, 0])
        one = af.convolve2(x[:, :, 0], np.ascontiguousarray(w[:, :, k]), 1, SPATIAL).to_numpy()
        np.testing.assert_array_equal(rhs[:, :, k, 0], one[:, :, 0, 0])


def test_rank_fallback(af, rng):
    # af_convolve2 on a 1-D filter falls back to af_convolve1 (convolve.cpp:231-234)
    x = rng.random((30, 4)).astype(np.float32)
    w = rng.uniform(-1, 1, (5,)).astype(np.float32)
    got = af.convolve2(x, w, 0, SPATIAL).to_numpy()
    ref = oracle.convolve(x, w, 1, False)
    assert got.shape == ref.shape
    assert oracle.rel_l2(got, ref) <= 1e-6


def test_empty_returns_signal(af):
    x = af.Array(np.ones((4, 5), np.float32))
    w = af.Array(np.zeros((0,), np.float32), dims=(0,))
    out = af.convolve2(x, w, 0, SPATIAL)
    assert out.dims == (4, 5, 1, 1)
    np.testing.assert_array_equal(out.to_numpy()[:, :, 0, 0], np.ones((4, 5)))


def test_diff_batch_is_arg_error(af, rng):
    x = rng.random((8, 8, 2, 1)).astype(np.float32)
    w = rng.random((3, 3, 1, 2)).astype(np.float32)
    with pytest.raises(af.AfError) as e:
        af.convolve2(x, w, 0, SPATIAL)
    assert e.value.code == af.AF_ERR_ARG


def test_auto_routes_to_freq(af, rng):
    # a non-square filter wider than 5 goes to the frequency domain under
    # AUTO (src/api/c/convolve.cpp:99-130): same result as the spatial call
    x = rng.random((40, 40)).astype(np.float32)
    w = rng.random((7, 3)).astype(np.float32)
    assert oracle.is_freq_domain(2, (40, 40, 1, 1), (7, 3, 1, 1), 0)
    auto = af.convolve2(x, w, 0, 0).to_numpy()
    spatial = af.convolve2(x, w, 0, SPATIAL).to_numpy()
    assert auto.shape == spatial.shape
    assert oracle.rel_l2(auto, spatial) <= 1e-5
    assert oracle.rel_l2(auto, oracle.convolve(x, w, 2, False)) <= 1e-5


def test_f16_signal_is_type_error(af, rng):
    x = rng.random((8, 8)).astype(np.float16)
    w = rng.random((3, 3)).astype(np.float16)
    with pytest.raises(af.AfError) as e:
        af.convolve2(x, w, 0, SPATIAL)
    assert e.value.code == af.AF_ERR_TYPE


def test_large_batch_grid(af):
    # many slices (test/convolve.cpp:750-806 uses 65536 slices, sum-checked)
    x = np.ones((4, 4, 65536), np.float32)
    w = np.ones((3, 3), np.float32)
    got = af.convolve2(x, w, 0, SPATIAL).to_numpy()
    one = oracle.convolve(np.ones((4, 4), np.float32), w, 2, False)
    assert got.shape == (4, 4, 65536, 1)
    assert float(got.sum()) == float(one.sum()) * 65536



def test_strip_long_segments(af, rng):
    """32 slices of 128 x 4096 give the strip kernel 64-row segments, so its
    four-block input ring wraps within a workgroup; checked on row slabs at
    segment edges and at both image borders."""
    x = rng.random((128, 4096, 32), dtype=np.float32)
    w = rng.uniform(-1, 1, (7, 7)).astype(np.float32)
    for mode in (0, 1):
        got = af.convolve2(x, w, mode, SPATIAL).to_numpy()
        h = 0 if mode else 3
        for k in (0, 13, 31):
            for y0 in (0, 48, 64, 120, 2040, got.shape[1] - 16):
                slab_check(got, x, k, lambda xs: oracle.convolve(xs, w, 2, True), 7, h, h, y0)
