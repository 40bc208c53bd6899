"""Spatial and separable convolution for every af_dtype the reference accepts,
bit for bit against the oracle's instantiation of the reference loops for that
dtype.

The reference's typed tests (test/convolve.cpp:35-37, 44-96) run cdouble,
cfloat, float, double, int, uint, char, schar, uchar, short, ushort, intl and
uintl through af_convolve1/2/3 with a filter of the signal's type and compare
with ASSERT_EQ. The same bar holds here:

  * integer-valued data with a filter of the signal's type (cast to the
    accumulator type, src/api/c/convolve.cpp:46): every dtype exact;
  * integer signals with fractional float filters: exact as well, because the
    GPU rounds each product before adding it and truncates the sum as the
    reference CPU loop does (kernel/convolve.hpp:31-35);
  * separable (src/backend/cpu/kernel/convolve.hpp:213-289): the tap is cast
    to the signal's type before the product (:233) and the column pass is
    rounded to that type (the T-typed temp): exact for integer dtypes with
    integer and fractional taps beyond +-1.

Sums stay inside each type's range (unsigned types get non-negative filters),
where float-to-integer conversion is defined on both sides. b8 runs as the
reference's `char`: results are compared as int8 bytes.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

SPATIAL = 1

DTYPES = [np.complex128, np.complex64, np.float32, np.float64, np.int32, np.uint32,
          np.bool_, np.int8, np.uint8, np.int16, np.uint16, np.int64, np.uint64]
INT_DTYPES = [np.int32, np.uint32, np.bool_, np.int8, np.uint8, np.int16, np.uint16,
              np.int64, np.uint64]

# (signal dims, filter dims, rank): the reference fixtures' shapes and batch
# kinds (vector, rectangle, cuboid; one2many, many2one, many2many).
SHAPES = [
    ((32,), (4,), 1),
    ((32, 3), (5,), 1),            # LHS
    ((17,), (4, 3), 1),            # RHS
    ((20, 3), (5, 3), 1),          # SAME
    ((15, 17), (5, 5), 2),
    ((15, 17, 2), (3, 2), 2),      # LHS, even filter
    ((12, 11), (3, 3, 2), 2),      # RHS
    ((10, 9, 2, 2), (4, 2, 2, 2), 2),  # SAME
    ((9, 8, 7), (3, 2, 2), 3),
    ((10, 11, 2, 2), (2, 2, 2), 3),   # LHS
]


def _id(c):
    return f"s{'x'.join(map(str, c[0]))}_f{'x'.join(map(str, c[1]))}_r{c[2]}"


def _name(dt):
    return np.dtype(dt).name


def _call(af, rank, x, w, mode):
    fn = {1: af.convolve1, 2: af.convolve2, 3: af.convolve3}[rank]
    return fn(x, w, mode, SPATIAL)


def _bytes(a):
    """Raw values: b8 results are compared as int8 (sums of taps exceed 1)."""
    a = np.asarray(a)
    return a.view(np.int8) if a.dtype == np.bool_ else a


def _ints(rng, shape, dt, hi):
    if np.dtype(dt) == np.bool_:
        return rng.integers(0, 2, shape).astype(np.bool_)
    v = rng.integers(0, hi, shape)
    if np.issubdtype(np.dtype(dt), np.complexfloating):
        return (v + 1j * rng.integers(0, hi, shape)).astype(dt)
    return v.astype(dt)


@pytest.mark.parametrize("case", SHAPES, ids=_id)
@pytest.mark.parametrize("dt", DTYPES, ids=_name)
@pytest.mark.parametrize("mode", [0, 1], ids=["default", "expand"])
def test_typed_filter_exact(af, rng, case, dt, mode):
    """Signal and filter of one dtype, small integer values: ASSERT_EQ."""
    sd, fd, rank = case
    x = _ints(rng, sd, dt, 5)
    w = _ints(rng, fd, dt, 2)
    out = _call(af, rank, x, w, mode)
    assert out.type == af.Array(np.zeros(1, dt)).type
    got = out.to_numpy()
    ref = oracle.convolve_dispatch(x, w, rank, mode == 1)  # filter cast to accT there
    assert got.shape == ref.shape
    np.testing.assert_array_equal(_bytes(got), _bytes(ref))


@pytest.mark.parametrize("case", SHAPES, ids=_id)
@pytest.mark.parametrize("dt", INT_DTYPES, ids=_name)
@pytest.mark.parametrize("mode", [0, 1], ids=["default", "expand"])
def test_integer_signal_fractional_filter_exact(af, rng, case, dt, mode):
    """Integer signals with fractional float32 filters: the float sum is
    truncated to the signal's type; exact when products are rounded before
    they are added, as in the reference loop."""
    sd, fd, rank = case
    x = _ints(rng, sd, dt, 10 if np.dtype(dt).itemsize > 1 else 4)
    unsigned = np.dtype(dt).kind == "u"
    lo = 0.0 if unsigned else -1.0
    w = rng.uniform(lo, 1.0, fd).astype(np.float32)
    got = _call(af, rank, x, w, mode).to_numpy()
    ref = oracle.convolve_dispatch(x, w, rank, mode == 1)
    assert got.shape == ref.shape
    np.testing.assert_array_equal(_bytes(got), _bytes(ref))


def test_complex_random_vs_oracle(af, rng):
    """Random complex values against the oracle's complex instantiation."""
    for dt, bar in ((np.complex64, 1e-6), (np.complex128, 1e-14)):
        x = (rng.random((37, 23, 2)) + 1j * rng.random((37, 23, 2))).astype(dt)
        w = (rng.uniform(-1, 1, (5, 4)) + 1j * rng.uniform(-1, 1, (5, 4))).astype(dt)
        for mode in (0, 1):
            got = af.convolve2(x, w, mode, SPATIAL).to_numpy()
            ref = oracle.convolve(x, w, 2, mode == 1)
            assert got.shape == ref.shape
            assert oracle.rel_l2(got, ref) <= bar


# ---------------------------------------------------------------- separable
SEP_SHAPES = [((40, 33, 2), 5, 4), ((31, 20), 3, 7), ((19, 26, 2, 2), 4, 3)]


@pytest.mark.parametrize("case", SEP_SHAPES, ids=lambda c: f"s{'x'.join(map(str, c[0]))}_c{c[1]}_r{c[2]}")
@pytest.mark.parametrize("dt", INT_DTYPES, ids=_name)
@pytest.mark.parametrize("mode", [0, 1], ids=["default", "expand"])
def test_separable_integer_exact(af, rng, case, dt, mode):
    """Integer signals with taps beyond +-1, integer and fractional: the
    reference truncates each tap to the signal's type (kernel/convolve.hpp:233)
    and rounds the column pass to it, so taps of 2.7 act as 2."""
    sd, cl, rl = case
    unsigned = np.dtype(dt).kind == "u" or np.dtype(dt) == np.bool_
    x = _ints(rng, sd, dt, 3 if np.dtype(dt).itemsize > 1 else 2)
    lo = 0.0 if unsigned else -2.9
    cf = rng.uniform(lo, 2.9, cl).astype(np.float32)
    rf = rng.uniform(lo, 2.9, rl).astype(np.float32)
    cf[0], rf[-1] = 2.0, (1.0 if unsigned else -2.0)   # integer taps as well
    got = af.convolve_sep(cf, rf, x, mode).to_numpy()
    ref = oracle.convolve2_sep(cf, rf, x, mode == 1)
    assert got.shape == ref.shape
    assert np.any(_bytes(ref) != 0)       # taps beyond +-1 survive truncation
    np.testing.assert_array_equal(_bytes(got), _bytes(ref))


@pytest.mark.parametrize("dt", [np.complex64, np.complex128], ids=_name)
@pytest.mark.parametrize("mode", [0, 1], ids=["default", "expand"])
def test_separable_complex(af, rng, dt, mode):
    """Complex separable convolution (convolve2<cfloat,cfloat>,
    src/api/c/convolve.cpp:276-283): against the oracle's complex
    instantiation, and against the 2-D convolution with the complex
    outer-product filter in FP64."""
    x = (rng.random((30, 21, 2)) + 1j * rng.random((30, 21, 2))).astype(dt)
    cf = (rng.uniform(-1, 1, 5) + 1j * rng.uniform(-1, 1, 5)).astype(dt)
    rf = (rng.uniform(-1, 1, 4) + 1j * rng.uniform(-1, 1, 4)).astype(dt)
    got = af.convolve_sep(cf, rf, x, mode).to_numpy()
    ref = oracle.convolve2_sep(cf, rf, x, mode == 1)
    assert got.shape == ref.shape
    bar = 1e-6 if dt == np.complex64 else 1e-14
    assert oracle.rel_l2(got, ref) <= bar
    # bilinearity: (x * (c (x) r)) as a 2-D convolution with the complex outer product
    full = oracle.convolve(x.astype(np.complex128), np.outer(cf, rf).astype(np.complex128), 2,
                           mode == 1)
    assert oracle.rel_l2(got, full) <= (1e-5 if dt == np.complex64 else 1e-12)
    # integer-valued complex data: exact
    xi = (rng.integers(0, 4, (18, 11)) + 1j * rng.integers(0, 4, (18, 11))).astype(dt)
    ci = (rng.integers(-2, 3, 3) + 1j * rng.integers(-2, 3, 3)).astype(dt)
    ri = (rng.integers(-2, 3, 5) + 1j * rng.integers(-2, 3, 5)).astype(dt)
    np.testing.assert_array_equal(af.convolve_sep(ci, ri, xi, mode).to_numpy(),
                                  oracle.convolve2_sep(ci, ri, xi, mode == 1))


@pytest.mark.parametrize("dt", [np.complex64, np.complex128, np.int16, np.float64], ids=_name)
def test_separable_scalar_filters_all_types(af, rng, dt):
    """Scalar column and row filters: y = T(accT(x) * (c * r)) elementwise
    (src/api/c/convolve.cpp:57-68), complex signals included."""
    if np.issubdtype(np.dtype(dt), np.complexfloating):
        x = (rng.random((9, 7)) + 1j * rng.random((9, 7))).astype(dt)
        c, r = np.array([0.5 - 1.25j], dt), np.array([3.0 + 0.5j], dt)
    elif np.issubdtype(np.dtype(dt), np.integer):
        x = rng.integers(0, 20, (9, 7)).astype(dt)
        c, r = np.float32([0.5]), np.float32([2.5])
    else:
        x = rng.random((9, 7)).astype(dt)
        c, r = np.array([0.5], dt), np.array([3.0], dt)
    got = af.convolve_sep(c, r, x, 0).to_numpy()
    ref = oracle.convolve2_sep(c, r, x, False)
    assert got.shape == ref.shape
    if np.issubdtype(np.dtype(dt), np.complexfloating):
        # numpy's complex product may use fused multiply-adds: within an ulp
        assert oracle.rel_l2(got, ref) <= (1e-6 if dt == np.complex64 else 1e-15)
    else:
        np.testing.assert_array_equal(got, ref)
