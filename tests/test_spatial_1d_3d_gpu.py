"""The FP32 1-D tile kernel (conv1_tile: af_convolve1 for filters of up to
2048 taps, every batch kind) and the 5x5x5 variant of the 3-D tile kernel
(conv3_tile<5,5,5>), against the oracle's restatement of the reference's
loops (src/backend/cpu/kernel/convolve.hpp:19-211) and scipy in FP64.

Cases (1-D): signal lengths below, at and across the 1024-output tile (1, 5,
1023, 1024, 1025, 100003), filters of 1, 2 (even), 3, 31, 128, 129 and 2048
taps, the filter longer than the signal, DEFAULT and EXPAND, and the batch
kinds NONE, LHS (many signals, one filter), RHS (one signal, many filters)
and SAME. Cases (3-D): 5x5x5 filters on volumes that are and are not whole
32x8x8 blocks, DEFAULT and EXPAND, a batch of volumes (LHS over dim 3).

Bars: output dims exact, 1e-5 relative L2 against the restatement and FP64
(3-D); the 1-D kernel rounds each product before adding it in the
reference's tap order, so it equals the restatement bit for bit."""
import numpy as np
import pytest
import scipy.signal

import oracle

pytestmark = pytest.mark.gpu

DEFAULT, EXPAND = 0, 1
SPATIAL = 1

CASES_1D = [
    # (signal dims, filter dims)
    ((1,), (3,)), ((5,), (1,)), ((1023,), (2,)), ((1024,), (31,)), ((1025,), (128,)),
    ((100003,), (129,)), ((4000,), (2048,)), ((7,), (31,)),
    ((3000, 7), (17,)),            # LHS: 7 signals, one filter
    ((3000,), (9, 5)),             # RHS: one signal, 5 filters
    ((2500, 3, 2), (33, 3, 2)),    # SAME: signal and filter batches paired
]


def _id(c):
    return "s{}_f{}".format("x".join(map(str, c[0])), "x".join(map(str, c[1])))


@pytest.mark.parametrize("mode", [DEFAULT, EXPAND], ids=["default", "expand"])
@pytest.mark.parametrize("case", CASES_1D, ids=_id)
def test_conv1_tile_f32(af, rng, case, mode):
    sd, fd = case
    x = rng.random(sd, dtype=np.float64).astype(np.float32)
    w = rng.uniform(-1, 1, fd).astype(np.float32)
    got = af.convolve1(x, w, mode, SPATIAL).to_numpy()
    ref = oracle.convolve(x, w, 1, mode == EXPAND)
    assert got.shape == ref.shape
    np.testing.assert_array_equal(got, ref)
    if len(sd) == 1 and len(fd) == 1:
        full = scipy.signal.convolve(x.astype(np.float64), w.astype(np.float64), mode="full")
        want = full if mode == EXPAND else full[fd[0] // 2: fd[0] // 2 + sd[0]]
        assert oracle.rel_l2(got.ravel(), want) <= 1e-5


@pytest.mark.parametrize("case", [((1025,), (7,)), ((3000, 7), (31,)), ((4000,), (2048,))], ids=_id)
def test_conv1_tile_integer_valued_exact(af, rng, case):
    sd, fd = case
    x = rng.integers(-8, 9, sd).astype(np.float32)
    w = rng.integers(-3, 4, fd).astype(np.float32)
    got = af.convolve1(x, w, DEFAULT, SPATIAL).to_numpy()
    ref = oracle.convolve(x, w, 1, False)
    np.testing.assert_array_equal(got, ref)


CASES_3D = [((32, 8, 8), (5, 5, 5)), ((37, 11, 9), (5, 5, 5)), ((20, 20, 20, 3), (5, 5, 5))]


@pytest.mark.parametrize("mode", [DEFAULT, EXPAND], ids=["default", "expand"])
@pytest.mark.parametrize("case", CASES_3D, ids=_id)
def test_conv3_tile_5cubed_f32(af, rng, case, mode):
    sd, fd = case
    x = rng.random(sd, dtype=np.float64).astype(np.float32)
    w = rng.uniform(-1, 1, fd).astype(np.float32)
    got = af.convolve3(x, w, mode, SPATIAL).to_numpy()
    ref = oracle.convolve(x, w, 3, mode == EXPAND)
    assert got.shape == ref.shape
    assert oracle.rel_l2(got, ref) <= 1e-5
    if len(sd) == 3:
        full = scipy.signal.convolve(x.astype(np.float64), w.astype(np.float64), mode="full")
        want = full if mode == EXPAND else full[2:2 + sd[0], 2:2 + sd[1], 2:2 + sd[2]]
        assert oracle.rel_l2(got[..., 0], want) <= 1e-5
