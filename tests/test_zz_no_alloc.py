"""afhip_k_conv2_nn_patch_noalloc against afhip_k_conv2_nn_patch (the same
kernels with a caller workspace): bit-exact outputs on two calls in a row,
the second with a smaller layer (so the static scratch still holds the first
call's packed weights), for the forward and the stride-1 data gradient; the
size limit answers 203 without touching any buffer."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F32, BF16 = 0, 64


def _dims(v):
    return (ctypes.c_longlong * 4)(*v)


def _layer(af, rng, n, c, h, w, k, dt):
    x = rng.standard_normal((w, h, c, n)).astype(np.float32)
    f = (rng.standard_normal((3, 3, c, k)) * np.sqrt(2.0 / (9 * c))).astype(np.float32)
    X, Wt = af.Array(x), af.Array(f)
    if dt == BF16:
        X, Wt = X.astype(af.bf16), Wt.astype(af.bf16)
    return X, Wt


def _run_both(af, pass_, X, Wt, xd, wd, yd, dt, prec):
    L = af.lib()
    st, pd, dl = _dims((1, 1, 0, 0)), _dims((1, 1, 0, 0)), _dims((1, 1, 0, 0))
    xdm, wdm, ydm = _dims(xd), _dims(wd), _dims(yd)
    out_dims = yd if pass_ == 0 else xd
    need = L.afhip_k_conv2_nn_patch_workspace(pass_, dt, prec, xdm, wdm, ydm, st, pd, dl)
    assert 0 < need <= L.afhip_k_conv2_nn_patch_noalloc_limit()
    ws = af.constant(0, (need,), af.u8)
    o1 = af.constant(0, out_dims, af.f32)
    o2 = af.constant(7, out_dims, af.f32)  # stale values must all be overwritten
    s = ctypes.c_void_p(af.stream())
    inp = X.device_ptr()
    r1 = L.afhip_k_conv2_nn_patch(pass_, dt, F32, prec, ctypes.c_void_p(o1.device_ptr()),
                                  ctypes.c_void_p(inp), ctypes.c_void_p(Wt.device_ptr()),
                                  xdm, wdm, ydm, st, pd, dl, ctypes.c_void_p(ws.device_ptr()), need, s)
    r2 = L.afhip_k_conv2_nn_patch_noalloc(pass_, dt, F32, prec, ctypes.c_void_p(o2.device_ptr()),
                                          ctypes.c_void_p(inp), ctypes.c_void_p(Wt.device_ptr()),
                                          xdm, wdm, ydm, st, pd, dl, s)
    assert r1 == 0 and r2 == 0, (r1, r2)
    af.sync()
    a, b = o1.to_numpy(), o2.to_numpy()
    assert np.all(np.isfinite(a))
    np.testing.assert_array_equal(a, b)
    return a


@pytest.mark.parametrize("dt,prec", [(F32, 4), (BF16, 1)], ids=["f32", "bf16"])
def test_noalloc_matches_workspace_entry(af, dt, prec):
    rng = np.random.default_rng(77)
    # first call: N=4, C=64, 56x56, K=128; second: N=2, C=16, 20x18, K=24
    for n, c, h, w, k in [(4, 64, 56, 56, 128), (2, 16, 18, 20, 24)]:
        X, Wt = _layer(af, rng, n, c, h, w, k, dt)
        xd, wd, yd = (w, h, c, n), (3, 3, c, k), (w, h, k, n)
        y = _run_both(af, 0, X, Wt, xd, wd, yd, dt, prec)
        assert y.shape == (w, h, k, n)
        # data gradient: the input is dy [W,H,K,N]
        dy = af.Array(rng.standard_normal(yd).astype(np.float32))
        if dt == BF16:
            dy = dy.astype(af.bf16)
        dx = _run_both(af, 1, dy, Wt, xd, wd, yd, dt, prec)
        assert dx.shape == (w, h, c, n)


def test_noalloc_limit_is_size_error(af):
    L = af.lib()
    lim = L.afhip_k_conv2_nn_patch_noalloc_limit()
    assert lim == 16 << 20
    # FP32 packed weights of a 3x3, C=1024, K=1024 layer exceed 16 MiB
    xd, wd, yd = _dims((16, 16, 1024, 1)), _dims((3, 3, 1024, 1024)), _dims((16, 16, 1024, 1))
    st, pd, dl = _dims((1, 1, 0, 0)), _dims((1, 1, 0, 0)), _dims((1, 1, 0, 0))
    need = L.afhip_k_conv2_nn_patch_workspace(0, F32, 4, xd, wd, yd, st, pd, dl)
    if need == 0:
        pytest.skip("layer not served by the patch path")
    assert need > lim
    rc = L.afhip_k_conv2_nn_patch_noalloc(0, F32, F32, 4, None, None, None, xd, wd, yd, st, pd, dl,
                                          ctypes.c_void_p(af.stream()))
    assert rc == 203
