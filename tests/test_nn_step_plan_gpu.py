"""The captured convolve2_nn training step (afhip_nn_step_plan): forward, data
gradient and filter gradient replayed as one hipGraph.

Every replay must equal the eager af_* calls bit for bit (the same kernels
and fixed-order sums run), on several replays in a row with eager library
work and host synchronisation between them; a replay after x, w and dy are
overwritten in place (device-to-device copies into their storage) must
equal the eager calls on the new values; the outputs match the FP64 closed
forms to the precision's bar. Geometries: a reduced benchmark layer in
BF16, FP32 and split-BF16, the stride-2 gradient layer in BF16 (split GEMM,
phase-ordered data gradient), and an unaligned FP16 layer."""
import ctypes

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

CASES = [
    # (x dims, w dims, stride, pad, dilation, dtype, precision, bar vs FP64)
    ((56, 56, 64, 4), (3, 3, 64, 128), (1, 1), (1, 1), (1, 1), "bf16", None, 1e-2),
    ((56, 56, 64, 2), (3, 3, 64, 128), (1, 1), (1, 1), (1, 1), "f32", None, 1e-5),
    ((24, 20, 32, 2), (3, 3, 32, 48), (1, 1), (1, 1), (1, 1), "f32", 3, 1e-5),
    ((28, 28, 128, 4), (3, 3, 128, 128), (2, 2), (1, 1), (1, 1), "bf16", None, 1e-2),
    ((15, 13, 24, 3), (3, 3, 24, 40), (1, 1), (1, 1), (1, 1), "f16", None, 2e-3),
]


def _id(c):
    return "x{}_w{}_s{}_{}_{}".format("x".join(map(str, c[0])), "x".join(map(str, c[1])),
                                      "x".join(map(str, c[2])), c[5], c[6])


def _np(a):
    v = a.to_numpy()
    if v.dtype == np.uint16:
        return (v.astype(np.uint32) << 16).view(np.float32)
    return v


def _copy_into(af, dst, src):
    """Overwrite dst's storage with src's values (same dtype and size)."""
    t = dst.type
    rc = af.lib().afhip_k_cast(t, ctypes.c_void_p(dst.device_ptr()), t, ctypes.c_void_p(src.device_ptr()),
                               int(np.prod(dst.dims)), ctypes.c_void_p(af.stream()))
    assert rc == 0


@pytest.mark.parametrize("case", CASES, ids=_id)
def test_step_plan_replays_match_eager(af, case):
    xd, wd, st, pd, dl, dt, prec, bar = case
    rng = np.random.default_rng(20261202)
    t = {"bf16": af.bf16, "f16": af.f16, "f32": af.f32}[dt]

    def arrays():
        x = rng.standard_normal(xd).astype(np.float32)
        w = (rng.standard_normal(wd) * np.sqrt(2.0 / (wd[0] * wd[1] * wd[2]))).astype(np.float32)
        yd = oracle.conv2_nn(x[..., :1], w, st, pd, dl).shape[:3] + (xd[3],)
        dy = rng.standard_normal(yd).astype(np.float32)
        A = [af.Array(v) for v in (x, w, dy)]
        if t != af.f32:
            A = [a.astype(t) for a in A]
        return A

    X, W, DY = arrays()

    def eager():
        kw = {} if prec is None else {"precision": prec}
        y = af.convolve2NN(X, W, st, pd, dl, **kw)
        dx = af.convolve2GradientNN(DY, X, W, y, st, pd, dl, af.AF_CONV_GRADIENT_DATA, **kw)
        dw = af.convolve2GradientNN(DY, X, W, y, st, pd, dl, af.AF_CONV_GRADIENT_FILTER, **kw)
        return [_np(a) for a in (y, dx, dw)]

    plan = af.NNStepPlan(X, W, DY, st, pd, dl, precision=prec)
    ref = eager()
    for rep in range(3):
        plan.run()
        got = [_np(a) for a in (plan.y, plan.dx, plan.dw)]
        for g, r in zip(got, ref):
            assert g.shape == r.shape
            assert np.array_equal(g, r), f"replay {rep}"
        # eager library work and a host sync between replays
        eager()
        af.sync()

    # new values in the same storage: the replay reads them
    X2, W2, DY2 = arrays()
    _copy_into(af, X, X2)
    _copy_into(af, W, W2)
    _copy_into(af, DY, DY2)
    plan.run()
    got = [_np(a) for a in (plan.y, plan.dx, plan.dw)]
    ref = eager()
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)

    x, w, dy = (_np(a).astype(np.float64) for a in (X, W, DY))
    assert oracle.rel_l2(got[0], oracle.conv2_nn_direct_f64(x, w, st, pd, dl)) <= bar
    assert oracle.rel_l2(got[1], oracle.conv2_data_grad_direct_f64(dy, x, w, st, pd, dl)) <= bar
    assert oracle.rel_l2(got[2], oracle.conv2_filter_grad_direct_f64(dy, x, w, st, pd, dl)) <= bar
    del plan


def test_step_plan_argument_errors(af):
    x = np.ones((8, 8, 4, 1), np.float32)
    w = np.ones((3, 3, 5, 2), np.float32)   # channel mismatch
    dy = np.ones((6, 6, 2, 1), np.float32)
    with pytest.raises(af.AfError) as e:
        af.NNStepPlan(x, w, dy, (1, 1), (0, 0), (1, 1))
    assert e.value.code == af.AF_ERR_SIZE
