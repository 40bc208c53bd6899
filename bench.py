"""Headline benchmark: convolve2_nn forward + backward on one MI355X.

One step = af_convolve2_nn forward, then af_convolve2_gradient_nn for the data
gradient and for the filter gradient, on inputs already resident in HBM, for
N=32 images of 64 channels at 56x56 with 128 filters of 3x3, stride 1, pad 1
(the reference's convolve2_nn configuration). The timed arithmetic is BF16
operands on the matrix cores with FP32 accumulation; the same step is also
timed in FP32 (the plain path) for the BF16/FP32 ratio.

value = algorithmic FLOPs of (fwd + dgrad + wgrad) / step time, in TFLOP/s,
where each pass counts 2*N*K*OH*OW*C*fh*fw FLOPs.

Prints ONE JSON line. --gpus N with N > 1 exits: this build runs one GPU.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Workload: x [W,H,C,N], w [fw,fh,C,K]
XD = (56, 56, 64, 32)
WD = (3, 3, 64, 128)
ST, PD, DL = (1, 1), (1, 1), (1, 1)
SEED = 20261202 + 2
PEAK = {"hbm_GBs": 8000.0, "bf16_TFs": 2500.0, "fp32_TFs": 157.3,
        # FP32 FMAs on the vector ALUs without packed FP32 instructions (not
        # used here): half of the 157.3 TFLOP/s
        "fp32_valu_TFs": 78.6}
# Mangled names of the step's kernels (as rocprofv3 reports them).
KERNEL_SYMBOL = {
    "conv_fwd": "afhip::conv_wres<1, unsigned short, unsigned short, 2, 4, 2, 0>",
    "conv_dgrad": "afhip::conv_wres<1, unsigned short, unsigned short, 4, 4, 1, 1>",
    "wgrad_partial": "afhip::wgrad_pc<3, 3, 1, 0, true>",
    "wgrad_reduce": "afhip::wgrad_sum<unsigned short>",
}
# HBM bytes per launch from rocprofv3 FETCH_SIZE + WRITE_SIZE (profiles/),
# filled in from the committed counter summary when present.
TRAFFIC = {}
_tp = os.path.join(ROOT, "profiles", "nn_step_traffic.json")
if os.path.exists(_tp):
    with open(_tp) as _f:
        TRAFFIC = json.load(_f).get("bytes_per_launch", {})


def nn_flops(xd, wd, yd):
    return 2.0 * xd[3] * wd[3] * yd[0] * yd[1] * xd[2] * wd[0] * wd[1]


class HipEvents:
    """hipEvent pairs on an explicit stream (the library's stream)."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")
        self.h.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.h.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.h.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                               ctypes.c_void_p]
        self.h.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def create(self, n):
        evs = []
        for _ in range(n):
            e = ctypes.c_void_p()
            assert self.h.hipEventCreate(ctypes.byref(e)) == 0
            evs.append(e)
        return evs

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, ctypes.c_void_p(stream)) == 0

    def ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        t = ctypes.c_float()
        assert self.h.hipEventElapsedTime(ctypes.byref(t), a, b) == 0
        return float(t.value)


def _dims(d):
    return (ctypes.c_longlong * 4)(*d)


def _two(v):
    return (ctypes.c_longlong * 2)(*v)


class NNStep:
    """One training step of the layer through the kernel C ABI
    (include/afhip_kernels.h): forward, data gradient, filter gradient, each
    as its individual kernel launches on the library stream, with outputs and
    workspaces allocated once. Weights are re-packed every step, as they would
    be after an optimizer update."""

    def __init__(self, af, X, W, DY, in_type, out_type, precision):
        self.af, self.L = af, af.lib()
        self.it, self.ot, self.pr = in_type, out_type, precision
        self.X, self.W, self.DY = X, W, DY
        yd = (XD[0], XD[1], WD[3], XD[3])
        self.xd, self.wd, self.yd = _dims(XD), _dims(WD), _dims(yd)
        self.st, self.pd, self.dl = _two(ST), _two(PD), _two(DL)
        L = self.L
        esz = 4 if out_type == af.f32 else 2
        self.Y = af.constant(0, yd, out_type)
        self.DX = af.constant(0, XD, out_type)
        self.DW = af.constant(0, WD, out_type)
        g = (self.xd, self.wd, self.yd, self.st, self.pd, self.dl)
        self.ws_f = L.afhip_k_conv2_nn_patch_workspace(0, in_type, precision, *g)
        self.ws_d = L.afhip_k_conv2_nn_patch_workspace(1, in_type, precision, *g)
        self.ws_w = L.afhip_k_conv2_nn_wgrad_rows_workspace(in_type, precision, *g)
        self.rows = self.ws_w > 0
        if not self.rows:
            self.ws_w = L.afhip_k_conv2_nn_wgrad_workspace(*g)
        assert self.ws_f > 0 and self.ws_d > 0, "patch path expected for this layer"
        self.WSF = af.constant(0, (max(self.ws_f, 1),), af.u8)
        self.WSD = af.constant(0, (max(self.ws_d, 1),), af.u8)
        self.WSW = af.constant(0, (max(self.ws_w, 1),), af.u8)
        self.g = g
        self.stream = ctypes.c_void_p(af.stream())
        del esz

    def kernels(self):
        """(name, launch) for each kernel of the step, in order."""
        L, g, s = self.L, self.g, self.stream
        p = lambda a: ctypes.c_void_p(a.device_ptr())  # noqa: E731
        it, ot, pr = self.it, self.ot, self.pr
        ks = [
            ("pack_fwd", lambda: L.afhip_k_conv2_nn_patch_pack(0, it, pr, p(self.W), *g, p(self.WSF),
                                                               self.ws_f, s)),
            ("conv_fwd", lambda: L.afhip_k_conv2_nn_patch_run(0, it, ot, pr, p(self.Y), p(self.X), *g,
                                                              p(self.WSF), self.ws_f, s)),
            ("pack_dgrad", lambda: L.afhip_k_conv2_nn_patch_pack(1, it, pr, p(self.W), *g, p(self.WSD),
                                                                 self.ws_d, s)),
            ("conv_dgrad", lambda: L.afhip_k_conv2_nn_patch_run(1, it, ot, pr, p(self.DX), p(self.DY), *g,
                                                                p(self.WSD), self.ws_d, s)),
        ]
        if self.rows:
            ks += [("wgrad_partial", lambda: L.afhip_k_conv2_nn_wgrad_rows_partial(
                       it, pr, p(self.X), p(self.DY), *g, p(self.WSW), self.ws_w, s)),
                   ("wgrad_reduce", lambda: L.afhip_k_conv2_nn_wgrad_rows_reduce(
                       it, ot, pr, p(self.DW), *g, p(self.WSW), self.ws_w, s))]
        else:
            ks += [("wgrad", lambda: L.afhip_k_conv2_nn_wgrad(it, ot, pr, p(self.DW), p(self.X), p(self.DY),
                                                               *g, p(self.WSW), self.ws_w, s))]
        return ks

    def outputs(self):
        return [self.Y, self.DX, self.DW]


PASS_OF = {"pack_fwd": "fwd", "conv_fwd": "fwd", "pack_dgrad": "dgrad", "conv_dgrad": "dgrad",
           "wgrad_partial": "wgrad", "wgrad_reduce": "wgrad", "wgrad": "wgrad"}


def time_step(af, ev, stream, step, steps, warmup):
    ks = step.kernels()
    for _ in range(warmup):
        for _, k in ks:
            assert k() == 0
    af.sync()
    marks = ev.create(len(ks) * steps + 1)
    t0 = time.perf_counter()
    ev.record(marks[0], stream)
    i = 1
    for _ in range(steps):
        for _, k in ks:
            rc = k()
            if rc != 0:
                raise RuntimeError(f"kernel launch failed with {rc}")
            ev.record(marks[i], stream)
            i += 1
    af.sync()
    wall = time.perf_counter() - t0
    per_kernel = {n: 0.0 for n, _ in ks}
    i = 0
    for _ in range(steps):
        for n, _ in ks:
            per_kernel[n] += ev.ms(marks[i], marks[i + 1])
            i += 1
    per_kernel = {n: v / steps for n, v in per_kernel.items()}
    total = ev.ms(marks[0], marks[-1]) / steps
    return total, per_kernel, wall


def af_step(af, X, W, DY, out_type, precision=None, geom=(ST, PD, DL)):
    """One step through the af_* C API: (y, dx, dw)."""
    st, pd, dl = geom
    y = af.convolve2NN(X, W, st, pd, dl, precision=precision, out_type=out_type)
    dx = af.convolve2GradientNN(DY, X, W, y, st, pd, dl, af.AF_CONV_GRADIENT_DATA, precision=precision,
                                out_type=out_type)
    dw = af.convolve2GradientNN(DY, X, W, y, st, pd, dl, af.AF_CONV_GRADIENT_FILTER, precision=precision,
                                out_type=out_type)
    return [y, dx, dw]


def time_af_api(af, ev, stream, X, W, DY, out_type, steps, warmup=2, precision=None):
    """The step through the af_* C API (pool allocation and dispatch on every
    call), timed with one HIP event pair around all steps."""
    hold = []
    for _ in range(warmup):
        hold = af_step(af, X, W, DY, out_type, precision)
    af.sync()
    e0, e1 = ev.create(2)
    ev.record(e0, stream)
    for _ in range(steps):
        hold = af_step(af, X, W, DY, out_type, precision)
    ev.record(e1, stream)
    ms = ev.ms(e0, e1) / steps
    del hold
    return ms


def time_graph(af, ev, stream, X, W, DY, steps, warmup):
    """K replays of the step captured by afhip_nn_step_plan between one event
    pair on the library stream (ms per step)."""
    plan = af.NNStepPlan(X, W, DY, ST, PD, DL)
    for _ in range(warmup):
        plan.run()
    af.sync()
    e0, e1 = ev.create(2)
    ev.record(e0, stream)
    for _ in range(steps):
        plan.run()
    ev.record(e1, stream)
    ms = ev.ms(e0, e1) / steps
    del plan
    return ms


def time_cold(af, ev, stream, X, W, DY, out_type, reps=5):
    """The step with the 256 MB MALL and the L2s flushed before it: a 1 GiB
    fill runs first (outside the timed region), then one step between an
    event pair. Median of reps."""
    flush = af.constant(0, (256 * 1024 * 1024,), af.f32)
    L = af.lib()
    hold = af_step(af, X, W, DY, out_type)
    af.sync()
    ts = []
    for i in range(reps):
        assert L.afhip_k_fill(af.f32, ctypes.c_void_p(flush.device_ptr()), float(i), 256 * 1024 * 1024,
                              ctypes.c_void_p(stream)) == 0
        e0, e1 = ev.create(2)
        ev.record(e0, stream)
        hold = af_step(af, X, W, DY, out_type)
        ev.record(e1, stream)
        ts.append(ev.ms(e0, e1))
    del hold, flush
    return float(np.median(ts))


def host_dispatch_us(af):
    """Host time to issue one call (no synchronisation): af_convolve2 on a
    64x64 image with a 5x5 filter and af_convolve2_nn of the headline layer,
    first call in the process and the mean of 200 (resp. 20) later calls."""
    import time as _t
    x = af.Array(np.ones((64, 64), np.float32))
    w = af.Array(np.ones((5, 5), np.float32))
    af.sync()
    t0 = _t.perf_counter()
    o = af.convolve2(x, w, af.AF_CONV_DEFAULT, af.AF_CONV_SPATIAL)
    first = (_t.perf_counter() - t0) * 1e6
    af.sync()
    hold = []
    t0 = _t.perf_counter()
    for _ in range(200):
        hold.append(af.convolve2(x, w, af.AF_CONV_DEFAULT, af.AF_CONV_SPATIAL))
    per = (_t.perf_counter() - t0) / 200 * 1e6
    af.sync()
    del hold, o
    return {"convolve2_64sq_5x5_first_call": round(first, 1), "convolve2_64sq_5x5": round(per, 2)}


def fp64_errors(x, w, dy, outs):
    """rel-L2 of (y, dx, dw) against FP64 evaluations of the closed forms
    (stride 1, pad 1, dilation 1), written with tensordot for speed."""
    import oracle

    y, dx, dw = [_as_f64(o) for o in outs]
    y64 = _fwd64(x, w)
    dx64 = _dgrad64(dy, w, x.shape)
    dw64 = _wgrad64(x, dy, w.shape)
    return (oracle.rel_l2(y, y64), oracle.rel_l2(dx, dx64), oracle.rel_l2(dw, dw64))


def _as_f64(a):
    v = a.to_numpy()
    if v.dtype == np.uint16:  # bfloat16 bits
        v = (v.astype(np.uint32) << 16).view(np.float32)
    return v.astype(np.float64)


def _fwd64(x, w):
    x = x.astype(np.float64)
    w = w.astype(np.float64)
    W, H, C, N = x.shape
    fw, fh, _, K = w.shape
    xp = np.zeros((W + 2, H + 2, C, N))
    xp[1:-1, 1:-1] = x
    y = np.zeros((W, H, K, N))
    for i in range(fw):
        for j in range(fh):
            xs = xp[i:i + W, j:j + H]                       # [W,H,C,N]
            y += np.tensordot(xs, w[fw - 1 - i, fh - 1 - j], axes=([2], [0])).transpose(0, 1, 3, 2)
    return y


def _dgrad64(dy, w, xshape):
    dy = dy.astype(np.float64)
    w = w.astype(np.float64)
    W, H, C, N = xshape
    fw, fh, _, K = w.shape
    dxp = np.zeros((W + 2, H + 2, C, N))
    for i in range(fw):
        for j in range(fh):
            t = np.tensordot(dy, w[fw - 1 - i, fh - 1 - j], axes=([2], [1]))  # [W,H,N,C]
            dxp[i:i + W, j:j + H] += t.transpose(0, 1, 3, 2)
    return dxp[1:-1, 1:-1]


def _wgrad64(x, dy, wshape):
    x = x.astype(np.float64)
    dy = dy.astype(np.float64)
    W, H, C, N = x.shape
    fw, fh, _, K = wshape
    xp = np.zeros((W + 2, H + 2, C, N))
    xp[1:-1, 1:-1] = x
    dw = np.zeros(wshape)
    for i in range(fw):
        for j in range(fh):
            xs = xp[i:i + W, j:j + H]
            dw[fw - 1 - i, fh - 1 - j] = np.tensordot(xs, dy, axes=([0, 1, 3], [0, 1, 3]))
    return dw


def cpu_baseline(budget_s=15.0):
    """The oracle's restatement of the reference CPU backend (im2col + numpy
    SGEMM + col2im) on a bounded sample: the same layer with fewer images."""
    import oracle

    threads = _blas_threads()
    rng = np.random.default_rng(SEED)
    n = 4
    xd = XD[:3] + (n,)
    x = rng.standard_normal(xd).astype(np.float32)
    w = (rng.standard_normal(WD) * np.sqrt(2.0 / (WD[2] * 9))).astype(np.float32)
    y = oracle.conv2_nn(x, w, ST, PD, DL)
    dy = rng.standard_normal(y.shape).astype(np.float32)
    flops = 3 * nn_flops(xd, WD, y.shape)
    reps, t = 0, 0.0
    t0 = time.perf_counter()
    while t < budget_s and reps < 200:
        oracle.conv2_nn(x, w, ST, PD, DL)
        oracle.conv2_data_grad(dy, x, w, ST, PD, DL)
        oracle.conv2_filter_grad(dy, x, w, ST, PD, DL)
        reps += 1
        t = time.perf_counter() - t0
    return {"value": flops * reps / t / 1e12, "unit": "TFLOP/s", "cores": int(threads),
            "kind": "port",
            "sample": f"FP32 fwd+dgrad+wgrad on N={n} images of the same layer "
                      f"(x {list(xd)}, w {list(WD)}), {reps} reps in {t:.1f} s; "
                      "im2col/col2im single-threaded C, SGEMM on numpy's OpenBLAS"}


def _cpu_threads():
    return max(1, min(16, os.cpu_count() or 1))


def cpu_rate(fn, items, units_per_item, budget_s=1.5, threads=1):
    """Runs fn over items (a thread pool of `threads` workers; the oracle's C
    loops release the GIL) until budget_s has passed; returns units / s."""
    import concurrent.futures as cf

    done, t = 0, 0.0
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as ex:
        while t < budget_s:
            list(ex.map(fn, items))
            done += len(items)
            t = time.perf_counter() - t0
    return units_per_item * done / t


def cpu_line(fn, items, units_per_item, unit, sample, budget_s=1.5):
    """The reference-faithful single-thread rate (the CPU backend runs each
    kernel on one worker thread, src/backend/cpu/queue.hpp:98-114) and the
    same loops over a thread pool on the host's cores."""
    one = cpu_rate(fn, items[:1], units_per_item, budget_s, 1)
    nt = _cpu_threads()
    many = cpu_rate(fn, items, units_per_item, budget_s, nt)
    return {"value": round(one, 4), "unit": unit, "cores": 1, "kind": "port", "sample": sample,
            "threaded": {"value": round(many, 4), "cores": nt}}


def spatial_secondary(af, ev, stream, reps=10, cpu=True):
    """FP32 spatial configurations (reported beside the headline): the
    reference CPU-backend semantics, timed per call with HIP events on the
    library stream; HBM roofline from algorithmic bytes (each input and output
    element once); the oracle's C loops on the host beside each."""
    import oracle

    rng = np.random.default_rng(20261202 + 1)
    res = {}

    def timeit(fn):
        fn()
        af.sync()
        e0, e1 = ev.create(2)
        ev.record(e0, stream)
        hold = None
        for _ in range(reps):
            hold = fn()
        ev.record(e1, stream)
        ms = ev.ms(e0, e1) / reps
        del hold
        return ms

    def entry(name, ms, nbytes, nflops, npix, extra=None):
        gbs = nbytes / (ms * 1e-3) / 1e9
        tfs = nflops / (ms * 1e-3) / 1e12
        d = {"ms": round(ms, 4), "GB_s": round(gbs, 1), "hbm_frac": round(gbs / PEAK["hbm_GBs"], 3),
             "TFLOP_s": round(tfs, 2), "valu_frac": round(tfs / PEAK["fp32_valu_TFs"], 3),
             # the binding roofline: the larger of the HBM and VALU times at peak
             "roofline_frac": round(max(gbs / PEAK["hbm_GBs"], tfs / PEAK["fp32_valu_TFs"]), 3),
             "GPix_s": round(npix / (ms * 1e-3) / 1e9, 2)}
        if extra:
            d.update(extra)
        res[name] = d

    nt = _cpu_threads()
    # 2a: 256 x 1024^2, 7x7, AF_BATCH_LHS
    xh = rng.random((1024, 1024, 256), dtype=np.float32)
    x = af.Array(xh)
    w7 = rng.standard_normal((7, 7)).astype(np.float32)
    W7 = af.Array(w7)
    ms = timeit(lambda: af.convolve2(x, W7, af.AF_CONV_DEFAULT, af.AF_CONV_SPATIAL))
    n = 1024 * 1024 * 256
    imgs = [np.ascontiguousarray(xh[:, :, k]) for k in range(nt)]
    entry("2a_conv2_lhs_256x1024sq_7x7", ms, 8.0 * n + 4 * 49, 2.0 * 49 * n, n, None if not cpu else {
        "cpu_baseline": cpu_line(lambda im: oracle.convolve(im, w7, 2, False), imgs, 1024 * 1024 / 1e9, "GPix/s",
                                 "1024x1024 slices of the same batch, 7x7 DEFAULT, the oracle's C loops")})
    # 2b: separable 17+17 Gaussian sigma 3 on the same images
    g = oracle.gaussian_kernel(17, 1, 3.0).ravel()
    G = af.Array(g)
    ms = timeit(lambda: af.convolve_sep(G, G, x, af.AF_CONV_DEFAULT))
    entry("2b_sep_17x17_256x1024sq", ms, 8.0 * n + 8 * 17, 2.0 * 34 * n, n, None if not cpu else {
        "cpu_baseline": cpu_line(lambda im: oracle.convolve2_sep(g, g, im, False), imgs, 1024 * 1024 / 1e9,
                                 "GPix/s", "1024x1024 slices of the same batch, 17+17 taps, the oracle's "
                                           "two passes with the T-typed temp")})
    del x, xh, imgs
    # 2c: 128^3 volume, 3x3x3
    vh = rng.random((128, 128, 128), dtype=np.float32)
    v = af.Array(vh)
    w3h = rng.uniform(-1, 1, (3, 3, 3)).astype(np.float32)
    w3 = af.Array(w3h)
    ms = timeit(lambda: af.convolve3(v, w3, af.AF_CONV_DEFAULT, af.AF_CONV_SPATIAL))
    n3 = 128 ** 3
    extra = {"note": "16 MiB working set: L2/MALL resident"}
    if cpu:
        extra["cpu_baseline"] = cpu_line(lambda im: oracle.convolve(im, w3h, 3, False), [vh] * nt, n3 / 1e9,
                                         "GPix/s", "the whole 128^3 volume, 3x3x3 DEFAULT")
    entry("2c_conv3_128cube_3x3x3", ms, 8.0 * n3, 2.0 * 27 * n3, n3, extra)
    # 1: 512^2, 5x5
    s1h = rng.random((512, 512), dtype=np.float32)
    s1 = af.Array(s1h)
    w5h = rng.uniform(-1, 1, (5, 5)).astype(np.float32)
    w5 = af.Array(w5h)
    ms = timeit(lambda: af.convolve2(s1, w5, af.AF_CONV_DEFAULT, af.AF_CONV_SPATIAL))
    n1 = 512 * 512
    extra = {"note": "2 MiB working set: cache resident"}
    if cpu:
        extra["cpu_baseline"] = cpu_line(lambda im: oracle.convolve(im, w5h, 2, False), [s1h] * nt, n1 / 1e9,
                                         "GPix/s", "the whole 512x512 image, 5x5 DEFAULT")
    entry("1_conv2_512sq_5x5", ms, 8.0 * n1, 2.0 * 25 * n1, n1, extra)
    return res


def callers_secondary(af, reps=5, cpu=True):
    """The frequency-domain route and the callers of the convolution path
    (DoG, IIR, Harris, Canny, ORB), each on one 1024x1024 f32 image through
    the af_* API; wall clock around reps calls (Harris, Canny and ORB read
    counts back to the host inside the call). GB/s counts the image read once
    and the result written once, a lower bound on each pipeline's traffic, so
    hbm_frac is an upper bound on how close it runs to HBM. The oracle's
    restatement on the host beside each, on a crop (one thread, and a
    thread pool)."""
    import oracle
    import scipy.ndimage as ndi

    rng = np.random.default_rng(20261202 + 7)
    base = ndi.gaussian_filter(rng.random((1024, 1024)), 2.0)
    img = ((base - base.min()) / (base.max() - base.min()) * 255.0).astype(np.float32)
    img += (rng.random(img.shape) > 0.997).astype(np.float32) * 60.0
    X = af.Array(img)
    nt = _cpu_threads()
    res = {}

    def wall(fn):
        hold = fn()
        af.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            hold = fn()
        af.sync()
        del hold
        return (time.perf_counter() - t0) / reps * 1e3

    def entry(name, fn, out_bytes, cpu_fn, side, what):
        ms = wall(fn)
        nbytes = img.nbytes + out_bytes
        gbs = nbytes / (ms * 1e-3) / 1e9
        d = {"ms": round(ms, 4), "GPix_s": round(img.size / (ms * 1e-3) / 1e9, 3), "GB_s": round(gbs, 1),
             "hbm_frac": round(gbs / PEAK["hbm_GBs"], 4)}
        if cpu:
            crop = np.ascontiguousarray(img[:side, :side])
            d["cpu_baseline"] = cpu_line(lambda c: cpu_fn(c), [crop] * nt, side * side / 1e9, "GPix/s",
                                         f"{side}x{side} crop of the image, {what}")
        res[name] = d

    f31 = rng.standard_normal((31, 31)).astype(np.float32)
    F31 = af.Array(f31)
    entry("fft_convolve2_1024sq_31x31", lambda: af.convolve2(X, F31, af.AF_CONV_DEFAULT, af.AF_CONV_FREQ),
          img.nbytes, lambda c: oracle.fft_convolve(c, f31, 2, False), 256, "the oracle's padded FFT product")
    entry("dog_1024sq_r1_r2", lambda: af.dog(X, 1, 2), img.nbytes, lambda c: oracle.dog(c, 1, 2), 256,
          "the oracle's two smoothings and difference")
    b = np.array([0.2, 0.3, 0.2], np.float32)
    a = np.array([1.0, -0.5, 0.1], np.float32)
    B, A = af.Array(b), af.Array(a)
    entry("iir_1024_columns_x1024", lambda: af.iir(B, A, X), img.nbytes, lambda c: oracle.iir(b, a, c), 256,
          "the oracle's FIR part and recurrence per column")
    entry("harris_1024sq_s1", lambda: af.harris(X, 500, 1e5, 1.0, 0, 0.04), 5 * 4 * 500,
          lambda c: oracle.harris(c, 500, 1e5, 1.0, 0, 0.04), 128, "the oracle's restatement")
    entry("canny_1024sq_manual", lambda: af.canny(X, 0, 0.1, 0.3), img.size, lambda c: oracle.canny(c, 0, 0.1, 0.3),
          256, "the oracle's restatement")
    entry("orb_1024sq_l4_blur", lambda: af.orb(X, 20.0, 400, 1.5, 4, True), 400 * (5 * 4 + 32),
          lambda c: oracle.orb(c, 20.0, 400, 1.5, 4, True), 128, "the oracle's restatement")
    return res


def ragged_secondary(af, ev, stream, cpu=True):
    """cfg 5: 2000 independent spatial 2-D convolutions (committed list in
    tests/golden/ragged_problems.json) as one graph-captured persistent
    launch, next to looping af_convolve2 over the same problems."""
    with open(os.path.join(ROOT, "tests", "golden", "ragged_problems.json")) as f:
        probs = json.load(f)["problems"]
    rng = np.random.default_rng(20261202 + 4)
    sig, flt, modes = [], [], []
    flops = bytes_ = 0.0
    for w_, h_, fw, fh, m in probs:
        sig.append(af.Array(rng.random((w_, h_), dtype=np.float32)))
        flt.append(af.Array(rng.random((fw, fh), dtype=np.float32)))
        modes.append(m)
        ow, oh = (w_ + fw - 1, h_ + fh - 1) if m else (w_, h_)
        flops += 2.0 * fw * fh * ow * oh
        bytes_ += 4.0 * (w_ * h_ + ow * oh)
    t0 = time.perf_counter()
    plan = af.RaggedPlan(sig, flt, modes)
    setup_s = time.perf_counter() - t0
    plan.run()
    af.sync()
    e0, e1 = ev.create(2)
    reps = 5
    ev.record(e0, stream)
    for _ in range(reps):
        plan.run()
    ev.record(e1, stream)
    ms_graph = ev.ms(e0, e1) / reps
    # per-call loop (host dispatch + one launch per problem)
    outs = [af.convolve2(s_, f_, m, af.AF_CONV_SPATIAL) for s_, f_, m in zip(sig, flt, modes)]
    af.sync()
    del outs
    t0 = time.perf_counter()
    ev.record(e0, stream)
    outs = [af.convolve2(s_, f_, m, af.AF_CONV_SPATIAL) for s_, f_, m in zip(sig, flt, modes)]
    ev.record(e1, stream)
    ms_loop = ev.ms(e0, e1)
    loop_wall = (time.perf_counter() - t0) * 1e3
    del outs, plan
    cpu_b = None
    if cpu:
        import oracle

        host = [(rng.random((w_, h_), dtype=np.float32), rng.random((fw, fh), dtype=np.float32), m)
                for w_, h_, fw, fh, m in probs[:64]]
        fl = sum(2.0 * p_[1].size * (p_[0].shape[0] + (p_[1].shape[0] - 1 if p_[2] else 0))
                 * (p_[0].shape[1] + (p_[1].shape[1] - 1 if p_[2] else 0)) for p_ in host) / len(host)
        cpu_b = cpu_line(lambda t: oracle.convolve(t[0], t[1], 2, bool(t[2])), host, fl / 1e9, "GFLOP/s",
                         "the first 64 problems of the committed list, one af_convolve2 each, the oracle's "
                         "C loops")
    return {"problems": len(probs), "ms_graph": round(ms_graph, 3), "ms_per_call_loop": round(ms_loop, 3),
            "cpu_baseline": cpu_b,
            "ms_per_call_loop_host_wall": round(loop_wall, 3), "speedup": round(ms_loop / ms_graph, 2),
            "TFLOP_s": round(flops / (ms_graph * 1e-3) / 1e12, 2),
            "GB_s": round(bytes_ / (ms_graph * 1e-3) / 1e9, 1),
            "hbm_frac": round(bytes_ / (ms_graph * 1e-3) / 1e9 / PEAK["hbm_GBs"], 3),
            "valu_frac": round(flops / (ms_graph * 1e-3) / 1e12 / PEAK["fp32_valu_TFs"], 3),
            "algorithmic_flops": flops, "algorithmic_bytes": bytes_,
            "plan_setup_s_incl_autotune": round(setup_s, 2)}


def precision_lines(af, ev, stream, x, w, dy, X32, W32, DY32, steps, warmup):
    """The headline step with FP16 operands (f16 arrays, FP32 accumulation,
    f16 results) and in split-BF16 mode (f32 arrays, hi + lo bfloat16
    operands, three products, f32 results): TFLOP/s (algorithmic FLOPs, as
    the headline) and rel-L2 against FP64 of the FP32 operands."""
    yd = (XD[0], XD[1], WD[3], XD[3])
    flops = 3 * nn_flops(XD, WD, yd)
    res = {}
    Xh, Wh, DYh = X32.astype(af.f16), W32.astype(af.f16), DY32.astype(af.f16)
    for name, args, kw in [("fp16", (Xh, Wh, DYh, af.f16), {}),
                           ("split_bf16", (X32, W32, DY32, af.f32), {"precision": af.PRECISION_SPLIT_BF16})]:
        ms = time_af_api(af, ev, stream, *args, steps, warmup, **kw)
        outs = af_step(af, *args, **kw)
        errs = fp64_errors(x, w, dy, outs)
        res[name] = {"tflops": round(flops / (ms * 1e-3) / 1e12, 3), "ms_per_step": round(ms, 4),
                     "rel_l2_vs_fp64": {n: float(f"{e:.3e}") for n, e in zip(["fwd", "dgrad", "wgrad"], errs)}}
    return res


# the stride-2 gradient layer: x [28,28,128,16], w [3,3,128,128], s2 p1 d1
S2_XD = (28, 28, 128, 16)
S2_WD = (3, 3, 128, 128)
S2_GEOM = ((2, 2), (1, 1), (1, 1))


def stride2_line(af, ev, stream, reps=20, cpu=True):
    """Forward, data gradient and filter gradient of the stride-2 layer, FP32
    and BF16 (bf16 arrays, FP32 accumulation), each pass timed alone through
    the af_* API; rel-L2 of the BF16 passes against FP64."""
    import oracle

    rng = np.random.default_rng(20261202 + 3)
    x = rng.standard_normal(S2_XD).astype(np.float32)
    w = (rng.standard_normal(S2_WD) * np.sqrt(2.0 / (S2_WD[2] * 9))).astype(np.float32)
    st, pd, dl = S2_GEOM
    yd = (14, 14, S2_WD[3], S2_XD[3])
    dy = rng.standard_normal(yd).astype(np.float32)
    fl = nn_flops(S2_XD, S2_WD, yd)
    X32, W32, DY32 = af.Array(x), af.Array(w), af.Array(dy)
    out = {"workload": "x [28,28,128,16], w [3,3,128,128], stride 2, pad 1; dy [14,14,128,16]",
           "flops_per_pass": fl}
    for tag, (X, W, DY, ot) in {"fp32": (X32, W32, DY32, af.f32),
                                "bf16": (X32.astype(af.bf16), W32.astype(af.bf16), DY32.astype(af.bf16),
                                         af.bf16)}.items():
        y = af.convolve2NN(X, W, st, pd, dl)
        calls = {
            "fwd": lambda: af.convolve2NN(X, W, st, pd, dl),
            "dgrad": lambda: af.convolve2GradientNN(DY, X, W, y, st, pd, dl, af.AF_CONV_GRADIENT_DATA),
            "wgrad": lambda: af.convolve2GradientNN(DY, X, W, y, st, pd, dl, af.AF_CONV_GRADIENT_FILTER),
        }
        peak = PEAK["bf16_TFs"] if tag == "bf16" else PEAK["fp32_TFs"]
        line = {}
        for pname, fn in calls.items():
            hold = fn()
            af.sync()
            e0, e1 = ev.create(2)
            ev.record(e0, stream)
            for _ in range(reps):
                hold = fn()
            ev.record(e1, stream)
            ms = ev.ms(e0, e1) / reps
            tf = fl / (ms * 1e-3) / 1e12
            line[pname] = {"us": round(ms * 1e3, 2), "TFLOP_s": round(tf, 2), "roofline_frac": round(tf / peak, 4)}
            if tag == "bf16":
                got = _as_f64(hold)
                ref = (oracle.conv2_nn_direct_f64(x, w, st, pd, dl) if pname == "fwd" else
                       oracle.conv2_data_grad_direct_f64(dy, x, w, st, pd, dl) if pname == "dgrad" else
                       oracle.conv2_filter_grad_direct_f64(dy, x, w, st, pd, dl))
                line[pname]["rel_l2_vs_fp64"] = float(f"{oracle.rel_l2(got, ref):.3e}")
            del hold
        tot = sum(v["us"] for v in line.values())
        line["step_TFLOP_s"] = round(3 * fl / (tot * 1e-6) / 1e12, 2)
        out[tag] = line
    if cpu:
        n = 2
        xs, dys = x[..., :n], dy[..., :n]
        sample = (xs, dys)

        def one(t):
            oracle.conv2_nn(t[0], w, st, pd, dl)
            oracle.conv2_data_grad(t[1], t[0], w, st, pd, dl)
            oracle.conv2_filter_grad(t[1], t[0], w, st, pd, dl)
        r = cpu_rate(one, [sample], 3 * fl * n / S2_XD[3] / 1e12, 3.0, 1)
        out["cpu_baseline"] = {"value": round(r, 4), "unit": "TFLOP/s", "cores": _blas_threads(), "kind": "port",
                               "sample": f"FP32 fwd+dgrad+wgrad on N={n} images of the layer; im2col/col2im "
                                         "single-threaded C, SGEMM on numpy's OpenBLAS"}
    return out


def _blas_threads():
    try:
        from threadpoolctl import threadpool_info
        return int(max([i.get("num_threads", 1) for i in threadpool_info() if i.get("user_api") == "blas"] or [1]))
    except Exception:  # pragma: no cover
        return os.cpu_count() or 1


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 error check")
    ap.add_argument("--no-spatial", action="store_true", help="skip the spatial secondary lines")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy (cast_kernel<float, float>) 5 times, "
                         "to scale rocprofv3 FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--callers-only", action="store_true",
                    help="only run the frequency-domain route and the callers (secondary_callers) and print "
                         "that line, for kernel traces of them")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU sharding is a design note, not a runnable path.", file=sys.stderr)
        sys.exit(2)

    import arrayfire_amd as af

    af.lib()
    stream = af.stream()
    if args.copy_calibration:
        n = 64 * 1024 * 1024  # floats: 256 MiB
        A = af.constant(1.0, (n,), af.f32)
        B = af.constant(0.0, (n,), af.f32)
        for _ in range(5):
            rc = af.lib().afhip_k_cast(af.f32, ctypes.c_void_p(B.device_ptr()), af.f32,
                                       ctypes.c_void_p(A.device_ptr()), n, ctypes.c_void_p(stream))
            assert rc == 0
        af.sync()
        print(json.dumps({"copy_calibration_bytes_each_way": 4 * n, "launches": 5}))
        return
    if args.callers_only:
        print(json.dumps({"secondary_callers": callers_secondary(af, cpu=not args.no_cpu)}))
        return
    ev = HipEvents()
    rng = np.random.default_rng(SEED)
    x = rng.standard_normal(XD).astype(np.float32)
    w = (rng.standard_normal(WD) * np.sqrt(2.0 / (WD[2] * WD[0] * WD[1]))).astype(np.float32)
    yd = (XD[0], XD[1], WD[3], XD[3])
    dy = rng.standard_normal(yd).astype(np.float32)
    X32, W32, DY32 = af.Array(x), af.Array(w), af.Array(dy)
    Xb, Wb, DYb = X32.astype(af.bf16), W32.astype(af.bf16), DY32.astype(af.bf16)
    af.sync()

    flops_pass = nn_flops(XD, WD, yd)
    # BF16 step: bf16 operands and bf16 results (the af convention: the output
    # dtype is the input dtype), FP32 accumulation inside every kernel
    # whole-step throughput: the af_* API, K steps between one event pair
    af.sync()
    t0 = time.perf_counter()
    ms_bf16 = time_af_api(af, ev, stream, Xb, Wb, DYb, af.bf16, args.steps, args.warmup)
    wall_bf16 = time.perf_counter() - t0
    ms_f32 = time_af_api(af, ev, stream, X32, W32, DY32, af.f32, max(3, args.steps // 2),
                         max(1, args.warmup // 2))
    # the same BF16 step as one captured hipGraph (afhip_nn_step_plan)
    ms_graph = time_graph(af, ev, stream, Xb, Wb, DYb, args.steps, args.warmup)
    # per-kernel durations: the same step through the kernel ABI with an event
    # after every launch (the events add small gaps; the af_* timing above has
    # none)
    stepb = NNStep(af, Xb, Wb, DYb, af.bf16, af.bf16, af.PRECISION_DEFAULT)
    ms_kabi, kern_bf16, _ = time_step(af, ev, stream, stepb, args.steps, args.warmup)
    step32 = NNStep(af, X32, W32, DY32, af.f32, af.f32, af.PRECISION_DEFAULT)
    _, kern_f32, _ = time_step(af, ev, stream, step32, max(3, args.steps // 2), max(1, args.warmup // 2))
    tf_bf16 = 3 * flops_pass / (ms_bf16 * 1e-3) / 1e12
    tf_f32 = 3 * flops_pass / (ms_f32 * 1e-3) / 1e12
    names = ["fwd", "dgrad", "wgrad"]
    per_bf16 = [sum(v for k, v in kern_bf16.items() if PASS_OF[k] == n) for n in names]
    per_f32 = [sum(v for k, v in kern_f32.items() if PASS_OF[k] == n) for n in names]

    errs = None
    if not args.no_check:
        # the BF16 operands are the RNE roundings of x, w, dy: the error
        # against FP64 includes that input rounding
        errs = fp64_errors(x, w, dy, stepb.outputs())

    # roofline of the dominant kernel (largest average duration)
    kname = max(kern_bf16, key=kern_bf16.get)
    nx, nw, ny = np.prod(XD), np.prod(WD), np.prod(yd)
    kbytes = {"conv_fwd": 2 * (nx + ny) + 2 * nw, "conv_dgrad": 2 * (ny + nx) + 2 * nw,
              "wgrad_partial": 2 * (nx + ny), "wgrad_reduce": 0, "pack_fwd": 0, "pack_dgrad": 0}
    bytes_k = float(kbytes.get(kname, 2 * (nx + ny + nw)))
    flops_k = flops_pass if kname in ("conv_fwd", "conv_dgrad", "wgrad_partial", "wgrad") else 0.0
    t_s = kern_bf16[kname] * 1e-3
    t_mem = bytes_k / (PEAK["hbm_GBs"] * 1e9)
    t_cmp = flops_k / (PEAK["bf16_TFs"] * 1e12)
    if t_mem >= t_cmp:
        roof = {"bound": "hbm", "achieved": bytes_k / t_s / 1e9, "peak": PEAK["hbm_GBs"], "unit": "GB/s"}
    else:
        roof = {"bound": "mfma", "achieved": flops_k / t_s / 1e12, "peak": PEAK["bf16_TFs"],
                "unit": "TFLOP/s"}
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["traffic"] = TRAFFIC.get(kname)
    roof["kernel"] = KERNEL_SYMBOL.get(kname, kname)
    roof["kernel_avg_us"] = round(kern_bf16[kname] * 1e3, 2)
    roof["algorithmic_bytes"] = int(bytes_k)
    roof["algorithmic_flops"] = flops_k

    out = {
        "metric": "convolve2_nn fwd+bwd TFLOP/s, BF16-MFMA vs FP32 plain, at rel-L2 <=1e-2",
        "value": round(tf_bf16, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_bf16, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic (seeded normal inputs, He-scaled filters)",
        "config": {"workload": "convolve2_nn fwd+dgrad+wgrad, x [56,56,64,32] (N=32, C=64, "
                               "56x56), w [3,3,64,128], stride 1, pad 1, dilation 1; bf16 "
                               "operands and results, fp32 accumulation",
                   "global_batch": XD[3], "parallelism": "none (1 GPU)"},
        "ms_per_step_hipgraph": round(ms_graph, 4),
        "tflops_hipgraph": round(3 * flops_pass / (ms_graph * 1e-3) / 1e12, 3),
        "fp32_plain_tflops": round(tf_f32, 3),
        "bf16_over_fp32": round(tf_bf16 / tf_f32, 3),
        "ms_per_pass_bf16": {n: round(float(v), 4) for n, v in zip(names, per_bf16)},
        "ms_per_pass_fp32": {n: round(float(v), 4) for n, v in zip(names, per_f32)},
        "us_per_kernel_bf16": {n: round(v * 1e3, 2) for n, v in kern_bf16.items()},
        "us_per_kernel_fp32": {n: round(v * 1e3, 2) for n, v in kern_f32.items()},
        "ms_per_step_bf16_kernel_abi_with_events": round(ms_kabi, 4),
        "host_wall_ms_per_step_bf16": round(wall_bf16 / args.steps * 1e3, 4),
        "rel_l2_vs_fp64_bf16": None if errs is None else {n: float(f"{e:.3e}") for n, e in
                                                         zip(names, errs)},
        "roofline": roof,
    }
    out["cache_state"] = ("warm: the step's ~90 MB working set stays in the 256 MB MALL between steps; "
                          "ms_per_step_cold_mall flushes it with a 1 GiB fill before each timed step")
    out["ms_per_step_cold_mall"] = round(time_cold(af, ev, stream, Xb, Wb, DYb, af.bf16), 4)
    out["host_dispatch_us_per_call"] = host_dispatch_us(af)
    if not args.no_check:
        out["precision_modes"] = precision_lines(af, ev, stream, x, w, dy, X32, W32, DY32,
                                                 max(3, args.steps // 2), max(1, args.warmup // 2))
    out["stride2_layer"] = stride2_line(af, ev, stream, cpu=not args.no_cpu)
    if not args.no_spatial:
        out["secondary_fp32_spatial"] = spatial_secondary(af, ev, stream, cpu=not args.no_cpu)
        out["secondary_fp32_ragged"] = ragged_secondary(af, ev, stream, cpu=not args.no_cpu)
        out["secondary_callers"] = callers_secondary(af, cpu=not args.no_cpu)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
