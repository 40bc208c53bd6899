"""The internal callers of the convolution path on the GPU: af_fir, af_iir,
af_dog and af_gaussian_kernel, against the oracle.

Shapes follow the reference's own tests: test/iir.cpp:73-120 (FIR and the
single-feedback-tap IIR at 10000x1 / 5000x10 signals with 1000x1, 50x10,
100x1 filters; the long and many-filter cases run af_convolve1's AUTO route
into the frequency domain), test/dog.cpp:43-83 (512x512 and 512x512x3,
InvalidArray). Bars: FP32 1e-5 and FP64 1e-12 relative L2 (the frequency
route: 1e-5 / 1e-11), complex alike.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

FT = {"f32": (np.float32, 1e-5), "f64": (np.float64, 1e-11), "c32": (np.complex64, 1e-5),
      "c64": (np.complex128, 1e-11)}


def _rand(rng, shape, dt):
    v = rng.random(shape)
    if np.dtype(dt).kind == "c":
        v = v + 1j * rng.random(shape)
    return v.astype(dt)


def test_gaussian_kernel(af):
    for r, c, sr, sc in [(3, 3, 0, 0), (5, 9, 0, 0), (17, 1, 3.0, 0), (1, 17, 0, 3.0), (7, 4, 1.5, 2.5)]:
        g = af.gaussian_kernel(r, c, sr, sc).to_numpy()[:, :, 0, 0]
        ref = oracle.gaussian_kernel(r, c, sr, sc)
        assert g.shape == ref.shape and g.dtype == np.float32
        assert oracle.rel_l2(g, ref) <= 1e-6


@pytest.mark.parametrize("t", list(FT))
@pytest.mark.parametrize("shape", [(10000, 1, 1000, 1), (10000, 1, 50, 10), (5000, 10, 100, 1),
                                   (5000, 10, 50, 10), (300, 3, 7, 3)],
                         ids=lambda s: "x%dx%d_b%dx%d" % s)
def test_fir(af, t, shape):
    dt, bound = FT[t]
    xr, xc, br, bc = shape
    rng = np.random.default_rng(21)
    x = _rand(rng, (xr, xc) if xc > 1 else (xr,), dt)
    b = _rand(rng, (br, bc) if bc > 1 else (br,), dt)
    y = af.fir(b, x).to_numpy()
    ref = oracle.fir(b, x)
    assert y.shape == ref.shape
    assert oracle.rel_l2(y, ref) <= bound


@pytest.mark.parametrize("t", list(FT))
@pytest.mark.parametrize("shape", [(10000, 1, 50, 10), (5000, 10, 100, 1), (5000, 10, 50, 10)],
                         ids=lambda s: "x%dx%d_b%dx%d" % s)
def test_iir_single_feedback_tap(af, t, shape):
    """test/iir.cpp:82-120: y == convolve1(x, b / a0, EXPAND)[:x0]."""
    dt, bound = FT[t]
    xr, xc, br, bc = shape
    rng = np.random.default_rng(22)
    x = _rand(rng, (xr, xc) if xc > 1 else (xr,), dt)
    b = _rand(rng, (br, bc) if bc > 1 else (br,), dt)
    a = (_rand(rng, (1, bc), dt) + 0.5).astype(dt)
    y = af.iir(b, a, x).to_numpy()
    ref = oracle.iir(b, a, x)
    assert y.shape == ref.shape
    assert oracle.rel_l2(y, ref) <= bound


def _stable_a(rng, na, dt):
    """Feedback taps 2 * poly(poles) with every pole inside radius 0.8 (real
    poles for real types), so the recurrence stays bounded."""
    if np.dtype(dt).kind == "c":
        poles = 0.8 * rng.random(na - 1) * np.exp(2j * np.pi * rng.random(na - 1))
    else:
        poles = rng.uniform(-0.8, 0.8, na - 1)
    return (2.0 * np.poly(poles)).astype(dt)


@pytest.mark.parametrize("t", list(FT))
@pytest.mark.parametrize("na", [3, 4, 9, 16, 21])
def test_iir_recurrence(af, t, na):
    dt, bound = FT[t]
    rng = np.random.default_rng(23 + na)
    x = _rand(rng, (400, 3), dt)
    b = _rand(rng, (6,), dt)
    a = _stable_a(rng, na, dt)
    assert a.shape == (na,)
    y = af.iir(b, a, x).to_numpy()
    ref = oracle.iir(b, a, x)
    assert y.shape == ref.shape == (400, 3, 1, 1)
    e = oracle.rel_l2(y, ref)
    print(f"iir {t} na={na}: rel-L2 {e:.3e}, max |diff| {np.abs(y - ref).max():.3e}", flush=True)
    assert e <= bound * 10


@pytest.mark.parametrize("t", ["f32", "f64"])
def test_iir_per_column_taps(af, t):
    """b and a with one column per signal column (the SAME batch)."""
    dt, bound = FT[t]
    rng = np.random.default_rng(31)
    x = _rand(rng, (257, 4), dt)
    b = _rand(rng, (5, 4), dt)
    a = np.stack([_stable_a(rng, 4, dt) for _ in range(4)], axis=1)
    y = af.iir(b, a, x).to_numpy()
    ref = oracle.iir(b, a, x)
    assert y.shape == ref.shape and oracle.rel_l2(y, ref) <= bound * 10


def test_iir_errors(af):
    x = np.zeros(10, np.float32)
    with pytest.raises(af.AfError) as e:
        af.iir(np.ones(3, np.float64), np.ones(2, np.float32), x)
    assert e.value.code == 202
    with pytest.raises(af.AfError) as e:
        af.iir(np.ones(3, np.int32), np.ones(2, np.int32), np.zeros(10, np.int32))
    assert e.value.code == 204


@pytest.mark.parametrize("dt", [np.float32, np.float64, np.uint8, np.int32, np.uint16],
                         ids=lambda d: np.dtype(d).name)
@pytest.mark.parametrize("shape", [(512, 512), (512, 512, 3), (97, 61)], ids=lambda s: "x".join(map(str, s)))
def test_dog(af, dt, shape):
    rng = np.random.default_rng(41)
    x = (rng.random(shape) * (200 if np.dtype(dt).kind in "iu" else 1)).astype(dt)
    for r1, r2 in [(3, 2), (1, 5)]:
        d = af.dog(x, r1, r2)
        assert d.type == (af.f64 if dt == np.float64 else af.f32)
        # the oracle smooths with the library's own f32 windows (pinned to
        # the closed form in test_gaussian_kernel): a last-bit difference in
        # a window is amplified by the difference of two close smoothings
        win = tuple(af.gaussian_kernel(2 * r + 1, 2 * r + 1).to_numpy()[:, :, 0, 0] for r in (r1, r2))
        ref = oracle.dog(x, r1, r2, windows=win)
        got = d.to_numpy()
        assert got.shape == ref.shape
        bound = 1e-12 if dt == np.float64 else 1e-5
        assert oracle.rel_l2(got, ref) <= bound, (r1, r2)


def test_dog_reference_checks(af):
    """test/dog.cpp: a constant 512x512 (and x3) image; 1-D input is an error."""
    for shape in [(512, 512), (512, 512, 3)]:
        x = np.ones(shape, np.float32)
        k1 = af.gaussian_kernel(3, 3).to_numpy()[:, :, 0, 0]
        k2 = af.gaussian_kernel(2, 2).to_numpy()[:, :, 0, 0]
        diff = (oracle.convolve(x, k1, 2, False) - oracle.convolve(x, k2, 2, False))
        out = af.dog(x, 3, 2).to_numpy()
        assert float(np.sum(out - diff)) < 1e-2
    with pytest.raises(af.AfError) as e:
        af.dog(np.zeros(512, np.float32), 3, 2)
    assert e.value.code == 202
