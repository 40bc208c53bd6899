"""Ragged batches (afhip_ragged_plan_*): every problem must equal its own
af_convolve2(signal, filter, mode, AF_CONV_SPATIAL) call (oracle within 1e-5
relative L2, dims exact), every graph replay must give the same result, and
the autotuner's choice must be cached."""
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _problems(rng, n, lo=2, hi=300):
    sig, flt, modes = [], [], []
    for _ in range(n):
        w, h = (int(v) for v in np.exp(rng.uniform(np.log(lo), np.log(hi), 2)))
        fw, fh = (int(v) for v in rng.integers(1, 16, 2))
        sig.append(rng.random((w, h), dtype=np.float32))
        flt.append(rng.uniform(-1, 1, (fw, fh)).astype(np.float32))
        modes.append(int(rng.integers(0, 2)))
    return sig, flt, modes


def test_ragged_matches_per_call_oracle(af, rng, tmp_path, monkeypatch):
    monkeypatch.setenv("AFHIP_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    sig, flt, modes = _problems(rng, 60)
    plan = af.RaggedPlan(sig, flt, modes)
    plan.run()
    first = [o.to_numpy() for o in plan.outputs]
    for x, w, m, got in zip(sig, flt, modes, first):
        ref = oracle.convolve_dispatch(x, w, 2, m == 1)
        assert got.shape == ref.shape
        assert oracle.rel_l2(got, ref) <= 1e-5
    for _ in range(3):  # graph replays are deterministic
        plan.run()
        for a, b in zip(first, plan.outputs):
            np.testing.assert_array_equal(a, b.to_numpy())
    assert plan.variant in (0, 1, 2)
    assert (tmp_path / "tune.json").exists()
    # a second plan over the same problems reuses the cached choice
    plan2 = af.RaggedPlan(sig, flt, modes)
    assert plan2.variant == plan.variant


def test_ragged_one_shot_and_cast_filters(af, rng, monkeypatch, tmp_path):
    monkeypatch.setenv("AFHIP_AUTOTUNE_CACHE", "off")
    sig, flt, modes = _problems(rng, 12, 2, 80)
    flt_i = [np.round(f * 3).astype(np.int32) for f in flt]  # cast to f32 like castArray
    outs = af.convolve2_ragged(sig, flt_i, modes)
    for x, w, m, o in zip(sig, flt_i, modes, outs):
        ref = oracle.convolve_dispatch(x, w.astype(np.float32), 2, m == 1)
        np.testing.assert_allclose(o.to_numpy(), ref, rtol=1e-5, atol=1e-4)


def test_ragged_rejects_large_filters(af, rng, monkeypatch):
    monkeypatch.setenv("AFHIP_AUTOTUNE_CACHE", "off")
    with pytest.raises(af.AfError) as e:
        af.convolve2_ragged([np.ones((20, 20), np.float32)], [np.ones((17, 3), np.float32)], [0])
    assert e.value.code == af.AF_ERR_ARG


def test_ragged_rejects_1d_signal(af, monkeypatch):
    # a (W, 1) signal makes af_convolve2 fall back to a 1-D convolution
    monkeypatch.setenv("AFHIP_AUTOTUNE_CACHE", "off")
    with pytest.raises(af.AfError) as e:
        af.convolve2_ragged([np.ones((20, 1), np.float32)], [np.ones((3, 3), np.float32)], [0])
    assert e.value.code == af.AF_ERR_ARG


def test_ragged_empty(af, monkeypatch):
    monkeypatch.setenv("AFHIP_AUTOTUNE_CACHE", "off")
    assert af.convolve2_ragged([], [], []) == []
