"""Frequency-domain convolution (af_fft_convolve1/2/3 and the AUTO route of
af_convolve1/2/3) on the GPU against the oracle's restatement of the
reference's semantics (src/api/c/fftconvolve.cpp, src/backend/cpu/fftconvolve.cpp).

Bars: output dims exact; FP32 (and integer signals computed in FP32) within
1e-5 relative L2, FP64 within 1e-12; integer outputs of the batched kinds
(rounded to nearest) exact on integer-valued data; DIFF integer outputs are a
truncation of an FFT result, so a value whose exact sum is an integer may land
one below it (the reference computes them the same way): within 1 there.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

DEFAULT, EXPAND = 0, 1
AUTO, SPATIAL, FREQ = 0, 1, 2

# (signal dims, filter dims, rank): every batch kind including DIFF
CASES = [
    ((37,), (9,), 1),
    ((40, 3), (7,), 1),              # LHS
    ((25,), (6, 4), 1),              # RHS
    ((30, 2), (5, 2), 1),            # SAME
    ((33, 21), (7, 4), 2),
    ((30, 20, 3), (5, 5), 2),        # LHS
    ((17, 19), (3, 9, 2), 2),        # RHS
    ((16, 12, 2, 2), (4, 6, 2, 2), 2),  # SAME
    ((20, 14, 3, 1), (5, 3, 1, 2), 2),  # DIFF: 3 x 2 pairs
    ((12, 9, 8), (3, 4, 5), 3),
    ((10, 9, 6, 2), (3, 3, 3), 3),   # LHS
    ((9, 8, 7, 2), (2, 3, 4, 2), 3),  # SAME
]


def _id(c):
    return f"s{'x'.join(map(str, c[0]))}_f{'x'.join(map(str, c[1]))}_r{c[2]}"


def _fft(af, rank, x, w, mode):
    fn = {1: af.lib().af_fft_convolve1, 2: af.lib().af_fft_convolve2, 3: af.lib().af_fft_convolve3}[rank]
    X, W = af.Array(x), af.Array(w)
    return af._out(fn, X.handle, W.handle, mode)


@pytest.mark.parametrize("case", CASES, ids=_id)
@pytest.mark.parametrize("dt", [np.float32, np.float64, np.complex64, np.complex128],
                         ids=lambda d: np.dtype(d).name)
@pytest.mark.parametrize("mode", [DEFAULT, EXPAND], ids=["default", "expand"])
def test_float_types(af, rng, case, dt, mode):
    sd, fd, rank = case
    x = rng.standard_normal(sd)
    w = rng.standard_normal(fd)
    if np.issubdtype(np.dtype(dt), np.complexfloating):
        x = x + 1j * rng.standard_normal(sd)
        w = w + 1j * rng.standard_normal(fd)
    x, w = x.astype(dt), w.astype(dt)
    got = _fft(af, rank, x, w, mode).to_numpy()
    ref = oracle.fft_convolve(x, w, rank, mode == EXPAND)
    assert got.shape == ref.shape and got.dtype == ref.dtype
    bar = 1e-12 if dt in (np.float64, np.complex128) else 1e-5
    assert oracle.rel_l2(got, ref) <= bar


@pytest.mark.parametrize("case", CASES, ids=_id)
@pytest.mark.parametrize("dt", [np.int32, np.uint8, np.int16, np.uint32, np.int64, np.int8, np.bool_],
                         ids=lambda d: np.dtype(d).name)
@pytest.mark.parametrize("mode", [DEFAULT, EXPAND], ids=["default", "expand"])
def test_integer_types(af, rng, case, dt, mode):
    sd, fd, rank = case
    small = np.dtype(dt).itemsize == 1
    x = (rng.integers(0, 2, sd) if small else rng.integers(0, 20, sd)).astype(dt)
    # float filter with fractional taps: cast to the signal type first on the
    # batched path (truncation), kept fractional on the DIFF path
    unsigned = np.dtype(dt).kind in "ub"
    w = rng.uniform(0.0 if unsigned else -2.5, 2.5, fd).astype(np.float32)
    got = _fft(af, rank, x, w, mode).to_numpy()
    ref = oracle.fft_convolve(x, w, rank, mode == EXPAND)
    assert got.shape == ref.shape
    g, r = got.view(np.int8) if got.dtype == np.bool_ else got, ref
    kind = oracle.identify_batch_kind(rank, oracle.dims4(x), oracle.dims4(w))
    if kind == oracle.AF_BATCH_DIFF:
        assert np.max(np.abs(g.astype(np.int64) - r.astype(np.int64))) <= 1
    else:
        np.testing.assert_array_equal(g, r)


def test_auto_domain_routes(af, rng):
    """AUTO sends these to the frequency domain (src/api/c/convolve.cpp:99-130):
    a 2-D filter wider than 17, a non-square one wider than 5, ten or more
    filters, a 1-D filter longer than 128 and a 3-D filter wider than 5; the
    results equal the spatial convolution."""
    cases = [((64, 50), (19, 19), 2), ((40, 40), (7, 3), 2), ((30, 30), (3, 3, 10), 2),
             ((600,), (130,), 1), ((14, 13, 12), (6, 3, 3), 3)]
    for sd, fd, rank in cases:
        x = rng.random(sd).astype(np.float32)
        w = rng.uniform(-1, 1, fd).astype(np.float32)
        assert oracle.is_freq_domain(rank, oracle.dims4(x), oracle.dims4(w), AUTO)
        fn = {1: af.convolve1, 2: af.convolve2, 3: af.convolve3}[rank]
        for mode in (DEFAULT, EXPAND):
            auto = fn(x, w, mode, AUTO).to_numpy()
            ref = oracle.convolve_dispatch(x, w, rank, mode == EXPAND)
            assert auto.shape == ref.shape
            assert oracle.rel_l2(auto, ref) <= 1e-5, (sd, fd)


def test_diff_batch_through_auto_and_freq(af, rng):
    """AF_BATCH_DIFF is rejected in the spatial domain (convolve.cpp:149) and
    served by the frequency domain: AUTO picks it, FREQ forces it."""
    x = rng.random((18, 15, 2, 1)).astype(np.float32)
    w = rng.uniform(-1, 1, (5, 5, 1, 3)).astype(np.float32)
    with pytest.raises(af.AfError) as e:
        af.convolve2(x, w, DEFAULT, SPATIAL)
    assert e.value.code == af.AF_ERR_ARG
    ref = oracle.fft_convolve(x, w, 2, False)
    assert ref.shape == (18, 15, 2, 3)
    for dom in (AUTO, FREQ):
        got = af.convolve2(x, w, DEFAULT, dom).to_numpy()
        assert got.shape == ref.shape
        assert oracle.rel_l2(got, ref) <= 1e-5
        for i in range(2):
            for j in range(3):
                one = oracle.convolve(np.ascontiguousarray(x[:, :, i, 0]), np.ascontiguousarray(w[:, :, 0, j]),
                                      2, False)
                assert oracle.rel_l2(got[:, :, i, j], one[:, :, 0, 0]) <= 1e-5


@pytest.mark.parametrize("sd,fd,rank", [((5000,), (200,), 1), ((4600, 3), (31, 2), 2),
                                        ((3, 5000), (2, 200), 2)], ids=["1d", "2d_axis0", "2d_axis1"])
def test_axes_longer_than_4096(af, rng, sd, fd, rank):
    """Padded axes of 8192 run as two four-step passes (in-LDS transforms of
    4096 and 2, twiddles fused), along the contiguous and a strided axis."""
    x = rng.standard_normal(sd).astype(np.float32)
    w = rng.standard_normal(fd).astype(np.float32)
    for mode in (DEFAULT, EXPAND):
        got = _fft(af, rank, x, w, mode).to_numpy()
        ref = oracle.fft_convolve(x, w, rank, mode == EXPAND)
        assert got.shape == ref.shape
        assert oracle.rel_l2(got, ref) <= 1e-5


def test_large_image_31x31(af, rng):
    """1024 x 1024 with a 31 x 31 filter (AUTO: frequency domain; padded to
    2048 x 2048) against the spatial restatement."""
    x = rng.random((1024, 1024), dtype=np.float32)
    w = rng.uniform(-1, 1, (31, 31)).astype(np.float32)
    got = af.convolve2(x, w, DEFAULT, AUTO).to_numpy()
    ref = oracle.convolve(x, w, 2, False)
    assert got.shape == ref.shape
    assert oracle.rel_l2(got, ref) <= 1e-5


def test_fft_errors(af, rng):
    # unsupported batch pairing -> AF_ERR_ARG; f16 -> AF_ERR_TYPE
    x = rng.random((8, 8, 2, 3)).astype(np.float32)
    w = rng.random((3, 3, 3, 2)).astype(np.float32)
    with pytest.raises(af.AfError) as e:
        _fft(af, 2, x, w, DEFAULT)
    assert e.value.code == af.AF_ERR_ARG
    with pytest.raises(af.AfError) as e:
        _fft(af, 2, x[:, :, 0, 0].astype(np.float16), w[:, :, 0, 0].astype(np.float16), DEFAULT)
    assert e.value.code == af.AF_ERR_TYPE
