"""Pins the CPU oracle (test infrastructure) before it is trusted.

Each check is chosen so that a plausible restatement mistake (a dropped tap,
a sign or index error, a transposed operand, a wrong centring offset) fails:
  * known answers printed in the reference's tests (test/convolve.cpp:522-620,
    1170-1185), stored in tests/golden/known_answers.json;
  * scipy.signal.convolve (textbook full convolution) for every rank/mode;
  * batch-kind equivalences (LHS/RHS/SAME == loops of NONE);
  * separable == outer-product 2-D (examples/getting_started/convolve.cpp:29-52);
  * im2col+GEMM NN path == independent FP64 closed form, == spatial DEFAULT
    (test/convolve.cpp:828-894), adjoint identities, central differences;
  * the Gaussian builder's closed form.
"""
import json
import os

import numpy as np
import pytest
import scipy.signal as ss

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def golden():
    with open(os.path.join(HERE, "golden", "known_answers.json")) as f:
        return json.load(f)


def test_known_answer_1d(golden):
    g = golden["convolve_1d"]
    got = oracle.convolve(np.float32(g["a"]), np.float32(g["b"]), 1, False).ravel()
    np.testing.assert_allclose(got, g["c"], atol=1.5e-4)


def test_known_answer_2d(golden):
    g = golden["convolve_2d"]
    got = oracle.convolve(np.full(g["signal_dims"], g["signal_value"], np.float32),
                          np.full(g["filter_dims"], g["filter_value"], np.float32), 2, False)
    np.testing.assert_array_equal(got[:, :, 0, 0], np.array(g["output"]))


def test_known_answer_3d(golden):
    g = golden["convolve_3d"]
    got = oracle.convolve(np.full(g["signal_dims"], g["signal_value"], np.float32),
                          np.full(g["filter_dims"], g["filter_value"], np.float32), 3, False)
    # output slices as printed, indexed [k][i][j]
    exp = np.array(g["output_slices"])
    for k in range(4):
        np.testing.assert_array_equal(got[:, :, k, 0], exp[k])


def _scipy_conv(x, w, expand):
    dt = np.complex128 if (np.iscomplexobj(x) or np.iscomplexobj(w)) else np.float64
    full = ss.convolve(x.astype(dt), w.astype(dt), mode="full", method="direct")
    if expand:
        return full
    sl = tuple(slice(f // 2, f // 2 + s) for s, f in zip(x.shape, w.shape))
    return full[sl]


@pytest.mark.parametrize("sd,fd", [((31,), (4,)), ((31,), (7,)), ((13, 9), (4, 3)),
                                   ((13, 9), (5, 5)), ((7, 8, 9), (2, 3, 4)),
                                   ((7, 8, 9), (3, 3, 3))])
@pytest.mark.parametrize("expand", [False, True])
def test_spatial_vs_scipy(rng, sd, fd, expand):
    x = rng.random(sd)
    w = rng.uniform(-1, 1, fd)
    got = oracle.convolve(x, w, len(sd), expand)
    ref = _scipy_conv(x, w, expand)
    assert got.shape[:len(sd)] == ref.shape
    np.testing.assert_allclose(got.reshape(ref.shape, order="F"), ref, rtol=1e-12, atol=1e-12)


def test_batch_kinds_are_loops_of_none(rng):
    x = rng.random((9, 8, 3))
    w = rng.uniform(-1, 1, (3, 2, 3))
    assert oracle.identify_batch_kind(2, (9, 8, 3, 1), (3, 2, 3, 1)) == oracle.AF_BATCH_SAME
    assert oracle.identify_batch_kind(2, (9, 8, 3, 1), (3, 2, 1, 1)) == oracle.AF_BATCH_LHS
    assert oracle.identify_batch_kind(2, (9, 8, 1, 1), (3, 2, 3, 1)) == oracle.AF_BATCH_RHS
    assert oracle.identify_batch_kind(2, (9, 8, 3, 1), (3, 2, 1, 2)) == oracle.AF_BATCH_DIFF
    assert oracle.identify_batch_kind(2, (9, 8, 3, 1), (3, 2, 2, 1)) == oracle.AF_BATCH_UNSUPPORTED
    for expand in (False, True):
        same = oracle.convolve(x, w, 2, expand)
        lhs = oracle.convolve(x, w[:, :, 0], 2, expand)
        rhs = oracle.convolve(x[:, :, 0], w, 2, expand)
        for k in range(3):
            np.testing.assert_array_equal(same[:, :, k, 0],
                                          oracle.convolve(x[:, :, k], w[:, :, k], 2, expand)[:, :, 0, 0])
            np.testing.assert_array_equal(lhs[:, :, k, 0],
                                          oracle.convolve(x[:, :, k], w[:, :, 0], 2, expand)[:, :, 0, 0])
            np.testing.assert_array_equal(rhs[:, :, k, 0],
                                          oracle.convolve(x[:, :, 0], w[:, :, k], 2, expand)[:, :, 0, 0])


def test_out_dims_rules():
    N, L, R, S = (oracle.AF_BATCH_NONE, oracle.AF_BATCH_LHS, oracle.AF_BATCH_RHS,
                  oracle.AF_BATCH_SAME)
    assert oracle.conv_out_dims((10, 11, 1, 1), (3, 4, 1, 1), N, 2, True) == (12, 14, 1, 1)
    assert oracle.conv_out_dims((10, 11, 5, 1), (3, 4, 1, 1), L, 2, True) == (12, 14, 5, 1)
    assert oracle.conv_out_dims((10, 11, 1, 1), (3, 4, 5, 1), R, 2, True) == (12, 14, 5, 1)
    assert oracle.conv_out_dims((10, 11, 1, 1), (3, 4, 5, 1), R, 2, False) == (10, 11, 5, 1)
    assert oracle.conv_out_dims((10, 11, 5, 2), (3, 4, 5, 2), S, 2, True) == (12, 14, 5, 2)
    assert oracle.ndims((1, 1, 1, 1)) == 1 and oracle.ndims((0, 3, 1, 1)) == 0
    assert oracle.ndims((1, 3, 1, 1)) == 2 and oracle.ndims((4, 1, 1, 2)) == 4


def test_is_freq_domain_rules():
    assert oracle.is_freq_domain(2, (40, 40, 1, 1), (7, 3, 1, 1), 0)
    assert not oracle.is_freq_domain(2, (40, 40, 1, 1), (5, 3, 1, 1), 0)
    assert oracle.is_freq_domain(2, (40, 40, 1, 1), (19, 19, 1, 1), 0)
    assert oracle.is_freq_domain(2, (40, 40, 1, 1), (3, 3, 10, 1), 0)
    assert not oracle.is_freq_domain(2, (40, 40, 1, 1), (3, 3, 9, 1), 0)
    assert oracle.is_freq_domain(1, (400, 1, 1, 1), (129, 1, 1, 1), 0)
    assert not oracle.is_freq_domain(2, (40, 40, 1, 1), (19, 19, 1, 1), 1)


@pytest.mark.parametrize("expand", [False, True])
def test_separable_is_outer_product(rng, expand):
    x = rng.random((23, 19, 2)).astype(np.float32)
    cf = rng.uniform(-1, 1, 5).astype(np.float32)
    rf = rng.uniform(-1, 1, 4).astype(np.float32)
    got = oracle.convolve2_sep(cf, rf, x, expand)
    full = oracle.convolve(x.astype(np.float64), np.outer(cf, rf).astype(np.float64), 2, expand)
    assert got.shape == full.shape
    assert oracle.rel_l2(got, full) <= 1e-6


# (x dims, w dims, stride, pad, dilation)
NN = [((10, 10, 1, 1), (3, 3, 1, 1), (1, 1), (1, 1), (1, 1)),
      ((10, 10, 1, 1), (3, 5, 1, 1), (3, 1), (1, 1), (2, 1)),
      ((8, 10, 3, 2), (3, 3, 3, 4), (2, 2), (1, 1), (1, 1)),
      ((9, 7, 5, 2), (2, 4, 5, 3), (1, 2), (0, 2), (2, 1))]


@pytest.mark.parametrize("g", NN)
def test_nn_restatement_vs_closed_form(rng, g):
    x = rng.standard_normal(g[0])
    w = rng.standard_normal(g[1])
    y = oracle.conv2_nn(x, w, *g[2:])
    y64 = oracle.conv2_nn_direct_f64(x, w, *g[2:])
    assert y.shape == y64.shape
    assert oracle.rel_l2(y, y64) <= 1e-13
    dy = rng.standard_normal(y.shape)
    assert oracle.rel_l2(oracle.conv2_data_grad(dy, x, w, *g[2:]),
                         oracle.conv2_data_grad_direct_f64(dy, x, w, *g[2:])) <= 1e-13
    assert oracle.rel_l2(oracle.conv2_filter_grad(dy, x, w, *g[2:]),
                         oracle.conv2_filter_grad_direct_f64(dy, x, w, *g[2:])) <= 1e-13
    # adjoint identities <conv(x,w), dy> = <x, dX(dy)> = <w, dW(dy)>
    lhs = float(np.sum(y64 * dy))
    assert abs(lhs - float(np.sum(x * oracle.conv2_data_grad_direct_f64(dy, x, w, *g[2:])))) \
        <= 1e-10 * max(1, abs(lhs))
    assert abs(lhs - float(np.sum(w * oracle.conv2_filter_grad_direct_f64(dy, x, w, *g[2:])))) \
        <= 1e-10 * max(1, abs(lhs))


def test_nn_matches_spatial_default(rng):
    # Conv2ConsistencyTest invariant, test/convolve.cpp:828-894
    for fs in (3, 5, 7, 9):
        x = rng.standard_normal((19, 17))
        w = rng.standard_normal((fs, fs))
        nn = oracle.conv2_nn(x, w, (1, 1), (fs // 2, fs // 2), (1, 1))
        sp = oracle.convolve(x, w, 2, False)
        np.testing.assert_allclose(nn, sp, atol=1e-12)


def test_zero_padding_issue2817():
    s = np.ones((5, 5), np.float32)
    f = np.full((3, 3), 1 / 9, np.float32)
    y = oracle.conv2_nn(s, f, (1, 1), (0, 0), (1, 1))
    assert y.shape == (3, 3, 1, 1) and np.sum(np.abs(y - 1)) < 1e-5
    g = np.full((3, 3), 1 / 9, np.float32)
    dw = oracle.conv2_filter_grad(g, s, f, (1, 1), (0, 0), (1, 1))
    assert np.sum(np.abs(dw - y)) < 1e-5


def test_nn_finite_differences(rng):
    st, pd, dl = (2, 2), (1, 1), (1, 1)
    x = rng.standard_normal((6, 6, 2, 2))
    w = rng.standard_normal((3, 3, 2, 3))
    g = rng.standard_normal(oracle.conv2_nn_direct_f64(x, w, st, pd, dl).shape)
    h = 1e-6
    L = lambda xx, ww: float(np.sum(oracle.conv2_nn_direct_f64(xx, ww, st, pd, dl) * g))  # noqa
    dx = oracle.conv2_data_grad(g, x, w, st, pd, dl)
    for idx in [(0, 0, 0, 0), (3, 2, 1, 1), (5, 5, 0, 1), (1, 4, 1, 0)]:
        xp, xm = x.copy(), x.copy()
        xp[idx] += h
        xm[idx] -= h
        assert abs((L(xp, w) - L(xm, w)) / (2 * h) - dx[idx]) < 1e-6
    dw = oracle.conv2_filter_grad(g, x, w, st, pd, dl)
    for idx in [(0, 0, 0, 0), (2, 1, 1, 2), (1, 2, 0, 1)]:
        wp, wm = w.copy(), w.copy()
        wp[idx] += h
        wm[idx] -= h
        assert abs((L(x, wp) - L(x, wm)) / (2 * h) - dw[idx]) < 1e-6


def test_nn_out_size_truncates():
    assert oracle.nn_out_size(10, 3, 1, 1, 1) == 10
    assert oracle.nn_out_size(10, 3, 3, 1, 1) == 4
    assert oracle.nn_out_size(10, 5, 5, 1, 1) == 2
    # (2 - 5) / 2 truncates toward zero to -1, so the size is 0
    assert oracle.nn_out_size(2, 5, 2, 0, 1) == 0


def test_gaussian_kernel_closed_form():
    # src/api/c/gaussian_kernel.cpp:35-88
    g = oracle.gaussian_kernel(17, 1, 3.0, dtype=np.float64).ravel()
    assert abs(g.sum() - 1) < 1e-15
    np.testing.assert_allclose(g, g[::-1], rtol=0, atol=0)
    i = np.arange(17)
    np.testing.assert_allclose(g / g[8], np.exp(-0.5 * ((i - 8) / 3.0) ** 2), rtol=1e-14)
    d = oracle.gaussian_kernel(5, 5, dtype=np.float64)  # default sigma 0.25 * n
    r = np.arange(5) - 2.0
    e = np.exp(-0.5 * ((r[:, None] / 1.25) ** 2 + (r[None, :] / 1.25) ** 2))
    np.testing.assert_allclose(d, e / e.sum(), rtol=1e-14)


# ------------------------------------------- integer and complex instantiations
ALL_INT = [np.int32, np.uint32, np.int16, np.uint16, np.int64, np.uint64, np.uint8, np.int8,
           np.bool_]


def _raw(a):
    a = np.asarray(a)
    return a.view(np.int8) if a.dtype == np.bool_ else a


@pytest.mark.parametrize("dt", ALL_INT + [np.complex64, np.complex128],
                         ids=lambda d: np.dtype(d).name)
def test_every_dtype_matches_scipy_on_integer_data(dt):
    """Every instantiation computes the textbook convolution: integer-valued
    data and filters give exact sums, so scipy's FP64 result is the answer."""
    rng = np.random.default_rng(3)
    for sd, fd, rank in (((23,), (4,), 1), ((13, 11), (3, 4), 2), ((7, 6, 5), (2, 3, 2), 3)):
        if np.dtype(dt) == np.bool_:
            x = rng.integers(0, 2, sd).astype(np.bool_)
        elif np.issubdtype(np.dtype(dt), np.complexfloating):
            x = (rng.integers(0, 4, sd) + 1j * rng.integers(0, 4, sd)).astype(dt)
        else:
            x = rng.integers(0, 4, sd).astype(dt)
        w = rng.integers(0, 3, fd).astype(np.float32)
        for expand in (False, True):
            got = oracle.convolve(x, w, rank, expand)
            ref = _scipy_conv(x.astype(np.complex128), w, expand)
            assert got.dtype == (np.int8 if np.dtype(dt) == np.bool_ else np.dtype(dt))
            np.testing.assert_array_equal(got.reshape(ref.shape).astype(np.complex128), ref)


@pytest.mark.parametrize("dt", [np.int32, np.int16, np.int64, np.int8], ids=lambda d: np.dtype(d).name)
def test_integer_output_truncates_toward_zero(dt):
    """InT(accum) truncates (kernel/convolve.hpp:35): x * 0.5 on odd negative
    and positive values rounds toward zero, not to the nearest integer nor
    down."""
    x = np.array([-7, -5, -3, -1, 1, 3, 5, 7], dtype=dt)
    got = oracle.convolve(x, np.float32([0.5]), 1, False).ravel()
    np.testing.assert_array_equal(got, np.trunc(x.astype(np.float64) * 0.5).astype(dt))


@pytest.mark.parametrize("dt", [np.int32, np.uint8, np.int16, np.uint64, np.int8, np.bool_],
                         ids=lambda d: np.dtype(d).name)
def test_separable_taps_truncate_to_signal_type(dt):
    """convolve2_separable casts each tap to InT (kernel/convolve.hpp:233):
    fractional taps act as their truncations, so the result equals the exact
    integer 2-D convolution with the outer product of the truncated taps."""
    rng = np.random.default_rng(11)
    unsigned = np.dtype(dt).kind in "ub"
    x = (rng.integers(0, 2, (17, 14)) if np.dtype(dt).itemsize == 1
         else rng.integers(0, 9, (17, 14))).astype(dt)
    cf = np.float32([2.7, 1.2, 0.6, 2.2]) if unsigned else np.float32([2.7, -1.5, 0.6, -2.2])
    rf = np.float32([1.9, 2.5, 1.1]) if unsigned else np.float32([-1.9, 2.5, 1.1])
    ct, rt = np.trunc(cf), np.trunc(rf)
    for expand in (False, True):
        got = oracle.convolve2_sep(cf, rf, x, expand)
        ref = _scipy_conv(x.astype(np.float64), np.outer(ct, rt), expand)
        np.testing.assert_array_equal(_raw(got)[:, :, 0, 0], ref)
        # the untruncated taps would give a different answer (the pin bites)
        assert not np.array_equal(_scipy_conv(x.astype(np.float64), np.outer(cf, rf), expand), ref)


def _brute_2d_float_filter(x, w, expand, dt):
    """Pure-Python loop in the reference's order (kernel/convolve.hpp:39-75)
    with float32 products rounded before each float32 add, then truncation."""
    S0, S1 = x.shape
    F0, F1 = w.shape
    O0, O1 = (S0 + F0 - 1, S1 + F1 - 1) if expand else (S0, S1)
    i0, j0 = (0, 0) if expand else (F0 // 2, F1 // 2)
    out = np.zeros((O0, O1), dtype=np.int64)
    for oj in range(O1):
        for oi in range(O0):
            acc = np.float32(0)
            for wj in range(F1):
                for wi in range(F0):
                    a, b = oi + i0 - wi, oj + j0 - wj
                    s = x[a, b] if (0 <= a < S0 and 0 <= b < S1) else 0
                    acc = np.float32(acc + np.float32(np.float32(int(s)) * w[wi, wj]))
            out[oi, oj] = int(acc)   # truncation toward zero
    return out


@pytest.mark.parametrize("dt", ALL_INT, ids=lambda d: np.dtype(d).name)
def test_integer_instantiations_brute_force(dt):
    """Fractional float filters on integer signals: every instantiation
    against a pure-Python brute force of the same rounding points."""
    rng = np.random.default_rng(7)
    unsigned = np.dtype(dt).kind in "ub"
    x = (rng.integers(0, 2, (9, 8)) if np.dtype(dt) == np.bool_ else rng.integers(0, 7, (9, 8)))
    x = x.astype(dt)
    w = rng.uniform(0.0 if unsigned else -1.0, 1.0, (3, 4)).astype(np.float32)
    for expand in (False, True):
        got = _raw(oracle.convolve(x, w, 2, expand))[:, :, 0, 0]
        np.testing.assert_array_equal(got.astype(np.int64), _brute_2d_float_filter(x, w, expand, dt))


def test_complex_instantiation_matches_numpy_fp64():
    """The complex instantiations against an FP64 evaluation of the complex
    sum (scipy on complex128)."""
    rng = np.random.default_rng(5)
    x = rng.random((12, 9)) + 1j * rng.random((12, 9))
    w = rng.uniform(-1, 1, (4, 3)) + 1j * rng.uniform(-1, 1, (4, 3))
    for expand in (False, True):
        ref = _scipy_conv(x, w, expand)
        g64 = oracle.convolve(x, w, 2, expand)[:, :, 0, 0]
        g32 = oracle.convolve(x.astype(np.complex64), w.astype(np.complex64), 2, expand)[:, :, 0, 0]
        assert oracle.rel_l2(g64, ref) <= 1e-14
        assert oracle.rel_l2(g32, ref) <= 1e-6


# ------------------------------------------------------ frequency domain
def _np_fft_linear(x, w, expand):
    """numpy FFT convolution with nextpow2(s + f - 1) padding per axis, then
    the crop at f // 2 (DEFAULT) or 0 (EXPAND)."""
    shp = [1 << int(np.ceil(np.log2(a + b - 1))) for a, b in zip(x.shape, w.shape)]
    y = np.fft.ifftn(np.fft.fftn(x, shp) * np.fft.fftn(w, shp))
    full = tuple(slice(0, a + b - 1) for a, b in zip(x.shape, w.shape))
    y = y[full]
    if not expand:
        y = y[tuple(slice(b // 2, b // 2 + a) for a, b in zip(x.shape, w.shape))]
    return y


@pytest.mark.parametrize("sd,fd", [((37,), (9,)), ((20, 13), (7, 4)), ((9, 8, 7), (3, 4, 2))])
def test_fft_convolve_is_padded_fft_product(sd, fd):
    """The restated frequency-domain result equals numpy's FFT convolution
    with the reference's nextpow2 padding and crop offsets (no wrap-around,
    offset f // 2 in DEFAULT mode)."""
    rng = np.random.default_rng(1)
    x, w = rng.standard_normal(sd), rng.standard_normal(fd)
    for expand in (False, True):
        got = oracle.fft_convolve(x, w, len(sd), expand).reshape(
            tuple(a + b - 1 for a, b in zip(sd, fd)) if expand else sd)
        np.testing.assert_allclose(got, _np_fft_linear(x, w, expand).real, atol=1e-12)


def test_fft_convolve_output_rules():
    """Integer outputs: rounded half away from zero on the batched path,
    truncated on the DIFF path; filter taps truncated to the signal type on
    the batched path only."""
    x = np.array([3, -3, 5], np.int32)
    # batched: taps [0.5, 1.5] -> int [0, 1]: y = x shifted, exact integers
    got = oracle.fft_convolve(x, np.float32([0.5, 1.5]), 1, True).ravel()
    np.testing.assert_array_equal(got, ss.convolve(x, [0, 1]))
    # DIFF (signal [3,1,2], filter [2,1,1,2]): taps stay fractional, truncation
    xs = np.array([[[1], [2]], [[-3], [4]], [[5], [-6]]], np.int32).reshape(3, 1, 2, 1)
    fs = np.array([0.5, 0.25], np.float32).reshape(2, 1, 1, 1) * np.ones((1, 1, 1, 2), np.float32)
    assert oracle.identify_batch_kind(1, xs.shape, fs.shape) == oracle.AF_BATCH_DIFF
    got = oracle.fft_convolve(xs, fs, 1, True)
    assert got.shape == (4, 1, 2, 2)
    exact = ss.convolve(xs[:, 0, 1, 0].astype(np.float64), [0.5, 0.25])
    np.testing.assert_array_equal(got[:, 0, 1, 1], np.trunc(exact).astype(np.int32))
    # roundf ties away from zero
    np.testing.assert_array_equal(oracle._round_half_away(np.array([2.5, -2.5, 0.5, -0.49])),
                                  [3.0, -3.0, 1.0, -0.0])


# ---------------------------------------------------------------- unwrap / wrap
def _wrap_factor(ix, iy, wx, wy, sx, sy, px, py):
    """How many patches cover each pixel: the count test/wrap.cpp:88-108
    builds for its wrap(unwrap(x)) == x * factor check."""
    f = np.zeros((ix, iy), dtype=np.int64)
    nx = (ix + 2 * px - wx) // sx + 1
    ny = (iy + 2 * py - wy) // sy + 1
    for wyi in range(ny):
        for wxi in range(nx):
            x0, y0 = wxi * sx - px, wyi * sy - py
            xs = [x for x in range(x0, x0 + wx) if 0 <= x < ix]
            ys = [y for y in range(y0, y0 + wy) if 0 <= y < iy]
            for y in ys:
                f[xs, y] += 1
    return f


def test_unwrap_wrap_doc_snippet():
    """The known answers of test/wrap.cpp:200-246 (af::wrap documentation)."""
    a = np.arange(1, 10, dtype=np.float32).reshape(3, 3, order="F")
    u = oracle.unwrap_any(a, 2, 2, 2, 2, 1, 1, True)[:, :, 0, 0]
    np.testing.assert_array_equal(u, [[0, 0, 0, 5], [0, 0, 4, 6], [0, 2, 0, 8], [1, 3, 7, 9]])
    np.testing.assert_array_equal(oracle.wrap_any(u, 3, 3, 2, 2, 2, 2, 1, 1, True)[:, :, 0, 0], a)
    b = np.ones((3, 3), np.float32)
    ub = oracle.unwrap_any(b, 2, 2, 1, 1, 0, 0, True)
    np.testing.assert_array_equal(ub[:, :, 0, 0], np.ones((4, 4)))
    np.testing.assert_array_equal(oracle.wrap_any(ub, 3, 3, 2, 2, 1, 1, 0, 0, True)[:, :, 0, 0],
                                  [[1, 2, 1], [2, 4, 2], [1, 2, 1]])


# (ix, iy, wx, wy, sx, sy, px, py): a spread of test/wrap.cpp:159-181's WRAP_INIT cases
WRAP_GEOMS = [(30, 10, 3, 3, 1, 1, 0, 0), (30, 10, 3, 3, 2, 2, 1, 1), (30, 10, 4, 4, 2, 2, 3, 3),
              (30, 10, 8, 8, 8, 8, 7, 7), (33, 21, 8, 12, 8, 12, 7, 11), (30, 20, 15, 10, 15, 10, 14, 9)]


@pytest.mark.parametrize("g", WRAP_GEOMS)
@pytest.mark.parametrize("is_column", [True, False])
def test_wrap_of_unwrap_is_factor_times_input(g, is_column):
    ix, iy, wx, wy, sx, sy, px, py = g
    rng = np.random.default_rng(7)
    x = rng.integers(-50, 50, (ix, iy, 2)).astype(np.int32)
    u = oracle.unwrap_any(x, wx, wy, sx, sy, px, py, is_column)
    r = oracle.wrap_any(u, ix, iy, wx, wy, sx, sy, px, py, is_column)
    f = _wrap_factor(ix, iy, wx, wy, sx, sy, px, py)
    np.testing.assert_array_equal(r[:, :, :, 0], x * f[:, :, None])
    rf = oracle.wrap_any(u.astype(np.float32), ix, iy, wx, wy, sx, sy, px, py, is_column)
    np.testing.assert_array_equal(rf[:, :, :, 0], (x * f[:, :, None]).astype(np.float32))


def test_unwrap_any_matches_c_restatement():
    rng = np.random.default_rng(8)
    x = rng.random((17, 13, 3), dtype=np.float32)
    for g in [(3, 3, 1, 1, 1, 1), (5, 4, 2, 3, 2, 1), (4, 4, 4, 4, 0, 0)]:
        a = oracle.unwrap_any(x, *g, is_column=False)
        b = oracle.unwrap(x, *g, 1, 1)
        np.testing.assert_array_equal(a, b)


def test_wrap_serial_from_zero_matches_c_restatement():
    rng = np.random.default_rng(9)
    cols = rng.standard_normal((20, 9, 2, 1)).astype(np.float32)   # 5x4 patches of 3x3 on 11x9
    a = oracle.wrap_any(cols, 11, 9, 3, 3, 2, 2, 0, 0, False)
    b = oracle._wrap_serial(cols, np.zeros((11, 9, 2, 1), np.float32), 3, 3, 2, 2, 0, 0)
    np.testing.assert_array_equal(a, b)


def test_unwrap_wrap_adjoint():
    """<unwrap(x), c> == <x, wrap(c)> in FP64."""
    rng = np.random.default_rng(10)
    x = rng.standard_normal((23, 17, 2))
    for g in [(3, 3, 1, 1, 1, 1), (5, 3, 2, 3, 4, 2), (6, 6, 6, 6, 0, 0)]:
        u = oracle.unwrap_any(x, *g, is_column=True)
        c = rng.standard_normal(u.shape)
        w = oracle.wrap_any(c, 23, 17, *g, is_column=True)
        assert abs(np.vdot(u, c) - np.vdot(oracle.as4(x), w)) <= 1e-9 * np.abs(u).sum() * 10


def test_wrap_integer_types_wrap_around():
    """Integer wrap adds modulo 2^bits, as C++ += on the reference's types."""
    u = np.full((4, 4, 1, 1), 200, dtype=np.uint8)      # 2x2 windows, stride 1 on 3x3
    r = oracle.wrap_any(u, 3, 3, 2, 2, 1, 1, 0, 0, True)
    np.testing.assert_array_equal(r[:, :, 0, 0], np.array([[200, 144, 200], [144, 32, 144],
                                                           [200, 144, 200]], dtype=np.uint8))


# ---------------------------------------------------------------- FIR / IIR / DoG
def test_iir_oracle_is_lfilter():
    """The reference's recurrence (kernel/iir.hpp) is the direct-form
    difference equation a0 y[n] = sum b x - sum_{k>=1} a_k y[n-k]: scipy's
    lfilter computes the same filter."""
    rng = np.random.default_rng(11)
    x = rng.standard_normal((200, 3))
    b = rng.standard_normal(5)
    a = np.array([1.7, -0.4, 0.2, 0.05])
    y = oracle.iir(b, a, x)[:, :, 0, 0]
    ref = np.stack([ss.lfilter(b, a, x[:, j]) for j in range(3)], axis=1)
    np.testing.assert_allclose(y, ref, rtol=1e-10, atol=1e-12)
    # per-column feedforward and feedback taps (the SAME batch)
    B = rng.standard_normal((4, 3))
    A = np.stack([np.array([2.0, 0.3, -0.1]), np.array([1.0, -0.5, 0.25]),
                  np.array([-1.5, 0.2, 0.1])], axis=1)
    y = oracle.iir(B, A, x)[:, :, 0, 0]
    ref = np.stack([ss.lfilter(B[:, j], A[:, j], x[:, j]) for j in range(3)], axis=1)
    np.testing.assert_allclose(y, ref, rtol=1e-10, atol=1e-12)


def test_iir_single_tap_is_normalised_fir():
    rng = np.random.default_rng(12)
    x = rng.standard_normal(64)
    b = rng.standard_normal(7)
    y = oracle.iir(b, np.array([2.5]), x)[:, 0, 0, 0]
    np.testing.assert_allclose(y, np.convolve(x, b / 2.5)[:64], rtol=1e-12, atol=1e-12)


def test_fir_is_truncated_full_convolution():
    rng = np.random.default_rng(13)
    x = rng.standard_normal((50, 2))
    b = rng.standard_normal((6, 2))
    y = oracle.fir(b, x)[:, :, 0, 0]
    ref = np.stack([np.convolve(x[:, j], b[:, j])[:50] for j in range(2)], axis=1)
    np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12)


def test_dog_constant_image_and_scipy():
    """DoG of a constant image is zero away from the borders (both windows
    sum to 1), and equals the difference of two zero-padded scipy
    convolutions everywhere."""
    x = np.ones((40, 30), np.float32)
    d = oracle.dog(x, 3, 2)[:, :, 0, 0]
    assert np.abs(d[7:-7, 7:-7]).max() < 1e-6
    rng = np.random.default_rng(14)
    x = rng.random((40, 30))
    d = oracle.dog(x, 2, 1)[:, :, 0, 0]
    # the windows are af_gaussian_kernel's f32 values, cast to the f64 input type
    g1 = oracle.gaussian_kernel(5, 5).astype(np.float64)
    g2 = oracle.gaussian_kernel(3, 3).astype(np.float64)
    ref = ss.convolve2d(x, g1, mode="same") - ss.convolve2d(x, g2, mode="same")
    np.testing.assert_allclose(d, ref, rtol=1e-12, atol=1e-12)


# ------------------------------------------------------------------ Harris
def test_harris_square_corners():
    """A bright rectangle on a dark field: exactly its four corners, with
    equal responses by symmetry (x = dim1 index, y = dim0 index)."""
    img = np.zeros((64, 80), np.float32)
    img[20:40, 30:60] = 1.0
    x, y, s = oracle.harris(img, 500, 1e5, 1.0, 0, 0.04)
    assert sorted(zip(x.tolist(), y.tolist())) == [(30.0, 20.0), (30.0, 39.0), (59.0, 20.0), (59.0, 39.0)]
    assert np.ptp(s) <= 1e-6 * s.max()


def test_harris_window_closed_forms():
    g = oracle.harris_filter(7, 1.3, np.float64)
    xs = np.arange(7) - 3
    ref = np.exp(-xs ** 2 / (2 * 1.3 ** 2))
    np.testing.assert_allclose(g, ref / ref.sum(), rtol=1e-14)
    b = oracle.harris_filter(5, 0.2, np.float32)
    assert np.all(b == np.float32(1) / np.float32(5))


def test_harris_responses_match_fp64_formulation():
    """The response map of the restatement (gradient, products, separable
    smoothing, det - k tr^2) against scipy in FP64: every corner it reports
    is a strict local maximum of the FP64 response, and the scores agree."""
    import scipy.ndimage as ndi

    rng = np.random.default_rng(3)
    img = ndi.gaussian_filter(rng.random((90, 70)), 2.0)
    h = oracle.harris_filter(5, 1.0, np.float64)
    iy, ix = np.gradient(img)
    sm = lambda a: ndi.convolve1d(ndi.convolve1d(a, h, axis=0, mode="constant"), h, axis=1, mode="constant")
    ixx, ixy, iyy = sm(ix * ix), sm(ix * iy), sm(iy * iy)
    r = ixx * iyy - ixy * ixy - 0.04 * (ixx + iyy) ** 2
    x, y, s = oracle.harris(img, 0, 1e-30, 1.0, 0, 0.04)
    assert len(x) > 5
    for xi, yi, si in zip(x.astype(int), y.astype(int), s):
        assert abs(r[yi, xi] - si) <= 1e-6 * np.abs(r).max()
        nb = r[yi - 1:yi + 2, xi - 1:xi + 2].copy()
        nb[1, 1] = -np.inf
        assert r[yi, xi] > nb.max()


def test_harris_top_k_is_stable_sort_of_detections():
    rng = np.random.default_rng(4)
    img = rng.random((60, 50)).astype(np.float32)
    x0, y0, s0 = oracle.harris(img, 0, 1e-30, 1.0, 0, 0.04)   # min_r = 0, scan order
    k = len(x0) // 3
    x, y, s = oracle.harris(img, k, 0.0, 1.0, 0, 0.04)
    order = sorted(range(len(s0)), key=lambda i: -float(s0[i]))[:k]
    assert np.array_equal(x, x0[order]) and np.array_equal(y, y0[order]) and np.array_equal(s, s0[order])
    xa, ya, sa = oracle.harris(img, 10 * len(x0), 0.0, 1.0, 0, 0.04)  # more asked than found: scan order
    assert np.array_equal(xa, x0) and np.array_equal(sa, s0)


@pytest.mark.parametrize("args,code", [
    (((8, 8, 2), 10, 0.0, 1.0, 0, 0.04), 203),
    (((32, 32), 0, 0.0, 1.0, 0, 0.04), 202),
    (((32, 32), 10, 0.0, 1.0, 0, 0.001), 202),
    (((32, 32), 10, 0.0, 7.0, 0, 0.04), 202),
    (((32, 32), 10, 0.0, 1.0, 40, 0.04), 202),
])
def test_harris_argument_errors(args, code):
    shape, *rest = args
    with pytest.raises(ValueError) as e:
        oracle.harris(np.zeros(shape, np.float32), *rest)
    assert e.value.args[0] == code


# ------------------------------------------------------------------ Canny
def test_canny_sobel_is_scipy_sobel_mirror():
    import scipy.ndimage as ndi

    x = np.random.default_rng(5).random((23, 31)).astype(np.float32)
    dx, dy = oracle.sobel3(x)
    np.testing.assert_allclose(dx, ndi.sobel(x.astype(np.float64), axis=0, mode="mirror"), atol=1e-5)
    np.testing.assert_allclose(dy, ndi.sobel(x.astype(np.float64), axis=1, mode="mirror"), atol=1e-5)


def test_canny_rectangle_edges_on_boundary():
    img = np.zeros((48, 64), np.float32)
    img[12:36, 16:48] = 255.0
    e = oracle.canny(img, 0, 0.1, 0.3)[..., 0, 0]
    ys, xs = np.nonzero(e)
    assert len(ys) > 60
    # every edge pixel lies within one pixel of the rectangle's boundary
    dist = np.minimum.reduce([np.abs(ys - 11.5), np.abs(ys - 35.5), np.abs(xs - 15.5), np.abs(xs - 47.5)])
    inside = (ys >= 10) & (ys <= 37) & (xs >= 14) & (xs <= 49)
    assert np.all(inside) and np.all(dist <= 1.5)


def test_canny_hysteresis_is_seeded_components():
    """Edge tracking marks exactly the 8-connected components of
    (weak | seeds) that hold a seed (pixels off the image border)."""
    import scipy.ndimage as ndi

    rng = np.random.default_rng(6)
    strong = rng.random((30, 25)) > 0.97
    weak = (rng.random((30, 25)) > 0.55) & ~strong
    strong[0, :] = strong[-1, :] = strong[:, 0] = strong[:, -1] = False
    weak[0, :] = weak[-1, :] = weak[:, 0] = weak[:, -1] = False
    out = oracle._hysteresis(strong, weak)
    lab, _ = ndi.label(weak | strong, structure=np.ones((3, 3)))
    seeded = np.unique(lab[strong])
    ref = np.isin(lab, seeded[seeded > 0])
    np.testing.assert_array_equal(out.astype(bool), ref)


def test_canny_otsu_matches_between_class_variance():
    """The threshold is the argmax of the between-class variance
    q0 q1 (mu0 - mu1)^2 over the histogram (FP64 evaluation)."""
    rng = np.random.default_rng(7)
    sup = np.concatenate([rng.normal(20, 3, 400), rng.normal(70, 5, 300), np.zeros(300)]).clip(0).astype(np.float32)
    sup = sup.reshape(25, 40)
    maxv = np.float32(sup.max())
    nb = int(maxv)
    t = oracle._otsu_threshold(sup, nb, maxv)
    step = float(maxv) / nb
    hist = np.bincount(np.clip((sup.ravel() / step).astype(int), 0, nb - 1), minlength=nb).astype(np.float64)
    p = hist / hist.sum()
    w = np.arange(nb)
    best, bi = -1, 0
    for b in range(nb - 1):
        q0 = p[:b + 1].sum()
        q1 = 1 - q0
        if q0 <= 0 or q1 <= 0:
            continue
        m0 = (w[:b + 1] * p[:b + 1]).sum() / q0
        m1 = (w[b + 1:] * p[b + 1:]).sum() / q1
        v = q0 * q1 * (m0 - m1) ** 2
        if v > best + 1e-9 * abs(best):
            best, bi = v, b
    assert abs(t - bi) <= 1


@pytest.mark.parametrize("args,code", [
    (((4, 8), 0, 0.1, 0.3, 3), 203),
    (((16, 16), 0, 0.1, 0.3, 5), 202),
    (((16,), 0, 0.1, 0.3, 3), 203),
])
def test_canny_argument_errors(args, code):
    shape, *rest = args
    with pytest.raises(ValueError) as e:
        oracle.canny(np.zeros(shape, np.float32), *rest)
    assert e.value.args[0] == code
