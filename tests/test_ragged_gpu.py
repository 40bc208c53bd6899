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


def _check_all(sig, flt, modes, outs, workers=16):
    """Every output against its own af_convolve2 call restated by the oracle
    (the oracle is C behind ctypes, which releases the GIL: a thread pool)."""
    import concurrent.futures as cf

    def one(i):
        ref = oracle.convolve_dispatch(sig[i], flt[i], 2, modes[i] == 1)[:, :, 0, 0]
        got = outs[i]
        assert got.shape == ref.shape, (i, got.shape, ref.shape)
        return oracle.rel_l2(got, ref)

    with cf.ThreadPoolExecutor(workers) as ex:
        errs = list(ex.map(one, range(len(sig))))
    worst = int(np.argmax(errs))
    assert errs[worst] <= 1e-5, (worst, errs[worst])


def test_ragged_full_problem_list(af, monkeypatch):
    """The committed 2000-problem list (image sides log-uniform in [64, 1024],
    odd filter sides 3..15, mixed DEFAULT/EXPAND): tens of thousands of
    (problem, tile) items on a grid of 2048 persistent workgroups, so every
    workgroup walks many items and reuses its LDS tile and filter across
    problems. Every output is checked, and a replay gives the same bits."""
    import json

    monkeypatch.setenv("AFHIP_AUTOTUNE_CACHE", "off")
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "golden", "ragged_problems.json")) as f:
        probs = json.load(f)["problems"]
    rng = np.random.default_rng(20261202 + 4)
    sig = [rng.random((w, h), dtype=np.float32) for w, h, _, _, _ in probs]
    flt = [rng.uniform(-1, 1, (fw, fh)).astype(np.float32) for _, _, fw, fh, _ in probs]
    modes = [m for *_, m in probs]
    plan = af.RaggedPlan(sig, flt, modes)
    plan.run()
    outs = [o.to_numpy()[:, :, 0, 0] for o in plan.outputs]
    print(f"ragged full list: {len(probs)} problems, variant {plan.variant}", flush=True)
    _check_all(sig, flt, modes, outs)
    plan.run()
    for i in (0, 777, 1999):
        np.testing.assert_array_equal(plan.outputs[i].to_numpy()[:, :, 0, 0], outs[i])


@pytest.mark.parametrize("grid", [1, 7, 64])
def test_ragged_small_grid_loops_over_items(af, rng, monkeypatch, grid):
    """AFHIP_RAGGED_GRID caps the persistent grid: with 1, 7 or 64 workgroups
    for hundreds of items, each workgroup runs many items back to back
    (problems of different sizes, filters and modes in one workgroup)."""
    monkeypatch.setenv("AFHIP_AUTOTUNE_CACHE", "off")
    monkeypatch.setenv("AFHIP_RAGGED_GRID", str(grid))
    sig, flt, modes = _problems(rng, 80, 2, 200)
    plan = af.RaggedPlan(sig, flt, modes)
    plan.run()
    _check_all(sig, flt, modes, [o.to_numpy()[:, :, 0, 0] for o in plan.outputs])
