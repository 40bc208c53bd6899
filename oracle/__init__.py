"""CPU oracle for ArrayFire's convolution path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this package, and only as the checker (or as the timed CPU
baseline).  The product library (``arrayfire_amd``) never imports it.

Two families of functions live here:

* **Restatements of the reference CPU backend** (same loops, index rules,
  accumulation types and rounding points).  The spatial/separable/im2col/col2im
  loops are in ``conv_oracle.c`` (built by ``oracle/Makefile``); the host-side
  composition (output dims, batch kind, im2col + SGEMM + col2im) is here:

  - ``ndims``                ... src/backend/common/dim4.cpp:48-58
  - ``identify_batch_kind``  ... src/api/c/convolve.cpp:77-97
  - ``is_freq_domain``       ... src/api/c/convolve.cpp:99-130
  - ``conv_out_dims``        ... src/backend/cpu/convolve.cpp:44-58
  - ``convolve``             ... src/backend/cpu/convolve.cpp:38-66 + kernel/convolve.hpp:132-211
  - ``convolve2_sep``        ... src/backend/cpu/convolve.cpp:68-99 + kernel/convolve.hpp:213-289
  - ``conv2_nn``             ... src/backend/cpu/convolve.cpp:123-167
  - ``conv2_data_grad``      ... src/backend/cpu/convolve.cpp:179-211
  - ``conv2_filter_grad``    ... src/backend/cpu/convolve.cpp:213-243
  - ``gaussian_kernel``      ... src/api/c/gaussian_kernel.cpp:35-88
  - ``harris``               ... src/backend/cpu/harris.cpp:26-148
  - ``canny``                ... src/api/c/canny.cpp:198-285
  - ``orb`` (``fast``, ``resize_bilinear``)
                             ... src/backend/cpu/orb.cpp:44-283, kernel/orb.hpp,
                                 kernel/fast.hpp, kernel/resize.hpp:86-130

  The SGEMM the reference delegates to CBLAS (src/backend/cpu/blas.cpp:223-331,
  a third-party library with no pinned version) is numpy's ``matmul`` on
  float32 operands, which calls the OpenBLAS sgemm bundled with numpy.

* **FP64 brute-force definitions** (``*_direct_f64``), written from the
  closed forms of the operations (true convolution with zero borders, and its
  two adjoints), independent of the im2col composition above.

Arrays are numpy arrays of shape ``(d0, d1, d2, d3)`` whose element
``a[i0, i1, i2, i3]`` is ArrayFire's ``a(i0, i1, i2, i3)``; flat buffers are
taken in Fortran order, which is ArrayFire's column-major memory order.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

AF_BATCH_NONE, AF_BATCH_LHS, AF_BATCH_RHS, AF_BATCH_SAME, AF_BATCH_DIFF, \
    AF_BATCH_UNSUPPORTED = 1, 2, 3, 4, 5, 6
AF_CONV_AUTO, AF_CONV_SPATIAL, AF_CONV_FREQ = 0, 1, 2

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build() -> None:
    """Compile the C restatement (liboracle_conv.so) in place."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle_conv.so")
        if not os.path.exists(path):
            build()
        lib = ctypes.CDLL(path)
        p = ctypes.c_void_p
        i64 = ctypes.c_int64
        for suf in _ALL_SUFFIXES:
            f = getattr(lib, f"orc_convolve_nd_{suf}")
            f.argtypes = [p, p, p, p, p, p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
            f.restype = None
            g = getattr(lib, f"orc_convolve2_sep_{suf}")
            g.argtypes = [p, p, p, p, p, i64, p, i64, ctypes.c_int]
            g.restype = ctypes.c_int
        for suf in ("f32", "f64"):
            u = getattr(lib, f"orc_unwrap_{suf}")
            u.argtypes = [p, p, p] + [i64] * 8
            u.restype = None
            w = getattr(lib, f"orc_wrap_dilated_{suf}")
            w.argtypes = [p, p, p, p] + [i64] * 8
            w.restype = None
        _LIB = lib
    return _LIB


# ----------------------------------------------------------------- helpers
# signal dtype -> (C instantiation, accumulator dtype), src/api/c/convolve.cpp:155-209.
# b8 (numpy bool) runs as the reference's `char`: outputs are returned as int8
# because a sum of several taps need not be 0 or 1.
_SUFFIX = {np.dtype(np.float32): ("f32", np.float32),
           np.dtype(np.float64): ("f64", np.float64),
           np.dtype(np.complex64): ("c32", np.complex64),
           np.dtype(np.complex128): ("c64", np.complex128),
           np.dtype(np.int32): ("s32", np.float32),
           np.dtype(np.uint32): ("u32", np.float32),
           np.dtype(np.int16): ("s16", np.float32),
           np.dtype(np.uint16): ("u16", np.float32),
           np.dtype(np.int64): ("s64", np.float32),
           np.dtype(np.uint64): ("u64", np.float32),
           np.dtype(np.uint8): ("u8", np.float32),
           np.dtype(np.int8): ("s8", np.float32),
           np.dtype(np.bool_): ("b8", np.float32)}
_ALL_SUFFIXES = tuple(v[0] for v in _SUFFIX.values())


def _storage(dt):
    """numpy dtype the C instantiation reads and writes (bool -> int8)."""
    return np.dtype(np.int8) if np.dtype(dt) == np.bool_ else np.dtype(dt)


def as4(a) -> np.ndarray:
    """View ``a`` as a 4-D ArrayFire-shaped array (trailing dims of 1)."""
    a = np.asarray(a)
    return a.reshape(tuple(a.shape) + (1,) * (4 - a.ndim), order="F")


def dims4(a) -> tuple:
    return tuple(int(v) for v in as4(a).shape)


def _flat(a: np.ndarray, dtype) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=dtype).ravel(order="F"))


def _dimbuf(d) -> np.ndarray:
    return np.asarray(d, dtype=np.int64)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def rel_l2(a, b) -> float:
    """||a - b||_2 / ||b||_2 computed in float64 (complex128 for complex
    operands, so both parts count)."""
    a, b = np.asarray(a), np.asarray(b)
    dt = np.complex128 if (np.iscomplexobj(a) or np.iscomplexobj(b)) else np.float64
    a = a.astype(dt).ravel()
    b = b.astype(dt).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


# ----------------------------------------------------------- host logic
def ndims(d) -> int:
    """dim4::ndims, src/backend/common/dim4.cpp:48-58."""
    d = tuple(d) + (1,) * (4 - len(d))
    n = d[0] * d[1] * d[2] * d[3]
    if n == 0:
        return 0
    if n == 1:
        return 1
    for k in (3, 2, 1):
        if d[k] != 1:
            return k + 1
    return 1


def identify_batch_kind(rank: int, sd, fd) -> int:
    """src/api/c/convolve.cpp:77-97."""
    sn, fn = ndims(sd), ndims(fd)
    if sn == rank and fn == rank:
        return AF_BATCH_NONE
    if sn == rank and rank < fn <= 4:
        return AF_BATCH_RHS
    if rank < sn <= 4 and fn == rank:
        return AF_BATCH_LHS
    if rank < sn <= 4 and rank < fn <= 4:
        match, inter = True, True
        for i in range(rank, 4):
            match &= sd[i] == fd[i]
            inter &= sd[i] == 1 or fd[i] == 1 or sd[i] == fd[i]
        if match:
            return AF_BATCH_SAME
        return AF_BATCH_DIFF if inter else AF_BATCH_UNSUPPORTED
    return AF_BATCH_UNSUPPORTED


def is_freq_domain(rank: int, sd, fd, domain: int) -> bool:
    """src/api/c/convolve.cpp:99-130."""
    if domain == AF_CONV_FREQ:
        return True
    if domain != AF_CONV_AUTO:
        return False
    if identify_batch_kind(rank, sd, fd) == AF_BATCH_DIFF:
        return True
    kb = 1
    for i in range(3, rank - 1, -1):
        kb *= fd[i]
    if kb >= 10:
        return True
    if rank == 1 and fd[0] > 128:
        return True
    if rank == 2:
        if fd[0] > 17 or fd[1] > 17:
            return True
        if fd[0] != fd[1] and fd[0] > 5:
            return True
    if rank == 3 and (fd[0] > 5 or fd[1] > 5 or fd[2] > 5):
        return True
    return False


def conv_out_dims(sd, fd, kind: int, rank: int, expand: bool) -> tuple:
    """Output dims, src/backend/cpu/convolve.cpp:44-58."""
    od = [1, 1, 1, 1]
    if expand:
        for d in range(4):
            if kind in (AF_BATCH_NONE, AF_BATCH_RHS):
                od[d] = sd[d] + fd[d] - 1
            else:
                od[d] = sd[d] + fd[d] - 1 if d < rank else sd[d]
    else:
        od = list(sd)
        if kind == AF_BATCH_RHS:
            for i in range(rank, 4):
                od[i] = fd[i]
    return tuple(int(v) for v in od)


# ------------------------------------------------------- spatial (restated)
def convolve(signal, filt, rank: int, expand: bool) -> np.ndarray:
    """Spatial convolution of rank 1/2/3 with the reference's batch rules.

    Restates cpu::convolve<T, accT> (src/backend/cpu/convolve.cpp:38-66) and
    kernel::convolve_nd (src/backend/cpu/kernel/convolve.hpp:132-211).  The
    filter is converted to accT as src/api/c/convolve.cpp:46 (castArray) does.
    Raises ValueError for the DIFF / UNSUPPORTED batch kinds.
    """
    x = as4(signal)
    suf, acct = _SUFFIX[x.dtype]
    sd, fd = dims4(x), dims4(filt)
    kind = identify_batch_kind(rank, sd, fd)
    if kind in (AF_BATCH_DIFF, AF_BATCH_UNSUPPORTED):
        raise ValueError("unsupported batch kind")
    od = conv_out_dims(sd, fd, kind, rank, expand)
    st = _storage(x.dtype)
    xs = _flat(x, st)
    fs = _flat(as4(filt), acct)
    out = np.zeros(int(np.prod(od)), dtype=st)
    getattr(_lib(), f"orc_convolve_nd_{suf}")(
        _ptr(out), _ptr(_dimbuf(od)), _ptr(xs), _ptr(_dimbuf(sd)), _ptr(fs),
        _ptr(_dimbuf(fd)), kind, rank, int(bool(expand)))
    return out.reshape(od, order="F")


def convolve_dispatch(signal, filt, rank: int, expand: bool) -> np.ndarray:
    """af_convolve{rank} with AF_CONV_SPATIAL: drops to a lower rank while
    either operand has fewer dims (src/api/c/convolve.cpp:231-234, 246-249),
    then convolve()."""
    while rank > 1 and (ndims(dims4(signal)) < rank or ndims(dims4(filt)) < rank):
        rank -= 1
    return convolve(signal, filt, rank, expand)


def _round_half_away(v):
    """roundf: halves away from zero."""
    return np.sign(v) * np.floor(np.abs(v) + 0.5)


def fft_convolve(signal, filt, rank: int, expand: bool) -> np.ndarray:
    """af_fft_convolve{rank} semantics, src/api/c/fftconvolve.cpp:123-219 with
    src/backend/cpu/fftconvolve.cpp:42-203 and the DIFF / complex fallback
    (src/api/c/fftconvolve.cpp:52-121), evaluated exactly (FP64 / complex128
    direct sums through this module's spatial restatement) and then converted
    as the reference converts:

    * batched kinds (NONE, LHS, RHS, SAME) of a real signal: the filter is
      first cast to the signal's type (castArray<T>, so fractional taps on an
      integer signal truncate); integer outputs are rounded half away from
      zero (roundf, kernel/fftconvolve.hpp:186-218), float outputs cast;
    * DIFF batches and complex signals: the filter is cast to the complex
      compute type; the output is a plain cast of the real part (integers
      truncate) or the complex value.

    Output dims: the spatial rules for batched kinds; for DIFF, the base dims
    of those rules and max(s, f) past the rank (broadcast pairing).  The FFT
    itself only adds rounding of the order of the compute type's epsilon,
    which the tests bound."""
    x = as4(signal)
    f = as4(filt)
    sd, fd = dims4(x), dims4(f)
    kind = identify_batch_kind(rank, sd, fd)
    if kind == AF_BATCH_UNSUPPORTED:
        raise ValueError("unsupported batch kind")
    cplx = np.iscomplexobj(x)
    fallback = kind == AF_BATCH_DIFF or cplx
    st = _storage(x.dtype)
    if not fallback:
        f = f.astype(st)                       # castArray<T>(filter)
    wide = np.complex128 if (cplx or np.iscomplexobj(f)) else np.float64
    if kind == AF_BATCH_DIFF:
        od = tuple((sd[d] + fd[d] - 1 if expand else sd[d]) if d < rank else max(sd[d], fd[d])
                   for d in range(4))
    else:
        od = conv_out_dims(sd, fd, kind, rank, expand)
    out = np.zeros(od, dtype=wide)
    xs, fs = x.astype(st).astype(wide), f.astype(wide)
    for ob in np.ndindex(*od[rank:]):
        si = tuple(0 if sd[rank + i] == 1 else ob[i] for i in range(len(ob)))
        fi = tuple(0 if fd[rank + i] == 1 else ob[i] for i in range(len(ob)))
        sl = (slice(None),) * rank
        xs1 = np.ascontiguousarray(xs[sl + si].reshape(sd[:rank] + (1,) * (4 - rank)))
        fs1 = np.ascontiguousarray(fs[sl + fi].reshape(fd[:rank] + (1,) * (4 - rank)))
        r = convolve(xs1, fs1, rank, expand)
        out[sl + ob] = r.reshape(r.shape[:rank])
    if cplx:
        return out.astype(x.dtype)
    v = out.real
    if np.issubdtype(st, np.integer):
        v = np.trunc(v) if kind == AF_BATCH_DIFF else _round_half_away(v)
    return v.astype(st)


def convolve2_sep(col_filter, row_filter, signal, expand: bool) -> np.ndarray:
    """Separable 2-D convolution, src/backend/cpu/convolve.cpp:68-99 and
    src/backend/cpu/kernel/convolve.hpp:213-289 (InT-typed temp).

    The scalar/scalar case of src/api/c/convolve.cpp:57-68 is included."""
    x = as4(signal)
    suf, acct = _SUFFIX[x.dtype]
    cf = np.asarray(col_filter, dtype=acct).ravel(order="F")
    rf = np.asarray(row_filter, dtype=acct).ravel(order="F")
    if cf.size == 1 and rf.size == 1:
        prod = acct(cf[0]) * acct(rf[0])
        return (x.astype(acct) * prod).astype(_storage(x.dtype))
    sd = dims4(x)
    od = list(sd)
    if expand:
        od[0] += cf.size - 1
        od[1] += rf.size - 1
    st = _storage(x.dtype)
    out = np.zeros(int(np.prod(od)), dtype=st)
    xs = _flat(x, st)
    rc = getattr(_lib(), f"orc_convolve2_sep_{suf}")(
        _ptr(out), _ptr(_dimbuf(od)), _ptr(xs), _ptr(_dimbuf(sd)),
        _ptr(np.ascontiguousarray(cf)), cf.size, _ptr(np.ascontiguousarray(rf)),
        rf.size, int(bool(expand)))
    if rc != 0:
        raise MemoryError("oracle temp allocation failed")
    return out.reshape(od, order="F")


# ---------------------------------------------------------- NN (restated)
def nn_out_size(s: int, f: int, stride: int, pad: int, dil: int) -> int:
    """1 + (s + 2p - ((f-1)d + 1)) / stride with C++ truncating division,
    src/backend/cpu/convolve.cpp:130-135."""
    num = s + 2 * pad - ((f - 1) * dil + 1)
    q = abs(num) // stride
    return 1 + (q if num >= 0 else -q)


def unwrap(x, wx, wy, sx, sy, px, py, dx, dy) -> np.ndarray:
    """im2col, src/backend/cpu/unwrap.cpp:23-42 with retCols = false:
    output dims [nx*ny, wx*wy, C, N]."""
    x = as4(x)
    suf = "f64" if x.dtype == np.float64 else "f32"
    dt = np.float64 if suf == "f64" else np.float32
    sd = dims4(x)
    nx = nn_out_size(sd[0], wx, sx, px, dx)
    ny = nn_out_size(sd[1], wy, sy, py, dy)
    od = (nx * ny, wx * wy, sd[2], sd[3])
    out = np.zeros(int(np.prod(od)), dtype=dt)
    getattr(_lib(), f"orc_unwrap_{suf}")(
        _ptr(out), _ptr(_flat(x, dt)), _ptr(_dimbuf(sd)),
        wx, wy, sx, sy, px, py, dx, dy)
    return out.reshape(od, order="F")


def wrap_dilated(cols, ox, oy, wx, wy, sx, sy, px, py, dx, dy) -> np.ndarray:
    """col2im, src/backend/cpu/wrap.cpp:60-75 (zero-initialised output,
    serial scatter-add)."""
    cols = as4(cols)
    suf = "f64" if cols.dtype == np.float64 else "f32"
    dt = np.float64 if suf == "f64" else np.float32
    idm = dims4(cols)
    od = (ox, oy, idm[2], idm[3])
    out = np.zeros(int(np.prod(od)), dtype=dt)
    getattr(_lib(), f"orc_wrap_dilated_{suf}")(
        _ptr(out), _ptr(_dimbuf(od)), _ptr(_flat(cols, dt)), _ptr(_dimbuf(idm)),
        wx, wy, sx, sy, px, py, dx, dy)
    return out.reshape(od, order="F")


def _patch_index(sd, wx, wy, sx, sy, px, py):
    """For every (col, loc) of af_unwrap's patch grid: the flat (x, y) index
    into one [d0, d1] plane and whether it lies inside the image
    (src/backend/cpu/kernel/unwrap.hpp:40-72 with dilation 1)."""
    nx = nn_out_size(sd[0], wx, sx, px, 1)
    ny = nn_out_size(sd[1], wy, sy, py, 1)
    col = np.arange(nx * ny)
    winy, winx = col // nx, col % nx
    loc = np.arange(wx * wy)
    y, x = loc // wx, loc % wx
    xpad = (winx * sx - px)[:, None] + x[None, :]
    ypad = (winy * sy - py)[:, None] + y[None, :]
    inside = (xpad >= 0) & (xpad < sd[0]) & (ypad >= 0) & (ypad < sd[1])
    flat = np.where(inside, ypad * sd[0] + xpad, 0)
    return nx, ny, flat, inside        # [col, loc]


def unwrap_any(x, wx, wy, sx, sy, px, py, is_column=True) -> np.ndarray:
    """af_unwrap for any dtype (src/api/c/unwrap.cpp:37-99): a gather, so
    exact. Output [wx*wy, nx*ny, C, N] (is_column) or [nx*ny, wx*wy, C, N]."""
    x = as4(x)
    sd = dims4(x)
    nx, ny, flat, inside = _patch_index(sd, wx, wy, sx, sy, px, py)
    planes = x.reshape(sd[0] * sd[1], sd[2], sd[3], order="F")
    g = planes[flat.ravel()].reshape(flat.shape + (sd[2], sd[3]))     # [col, loc, C, N]
    g = np.where(inside[:, :, None, None], g, np.zeros((), dtype=x.dtype))
    return np.ascontiguousarray(g.transpose(1, 0, 2, 3) if is_column else g)


def wrap_any(cols, ox, oy, wx, wy, sx, sy, px, py, is_column=True, out=None) -> np.ndarray:
    """af_wrap for any dtype (src/api/c/wrap.cpp:26-115, kernel/wrap.hpp:20-80):
    patch elements added into `out` (zeros by default) in the reference's
    serial order. Floating types run the C restatement (the order fixes the
    rounding; complex parts add independently); integer types add modulo
    2^bits, where order does not matter."""
    cols = as4(cols)
    if is_column:
        cols = np.ascontiguousarray(cols.transpose(1, 0, 2, 3))     # -> [nx*ny, wx*wy, C, N]
    idm = dims4(cols)
    od = (ox, oy, idm[2], idm[3])
    dt = cols.dtype
    base = np.zeros(od, dtype=dt) if out is None else as4(out).astype(dt, copy=True)
    if dt in (np.float32, np.float64):
        if out is None:
            return wrap_dilated(cols, ox, oy, wx, wy, sx, sy, px, py, 1, 1).astype(dt)
        return _wrap_serial(cols, base, wx, wy, sx, sy, px, py)   # the serial sum from `out`
    if dt in (np.complex64, np.complex128):
        re = wrap_any(np.ascontiguousarray(cols.real), ox, oy, wx, wy, sx, sy, px, py, False,
                      None if out is None else base.real.copy())
        im = wrap_any(np.ascontiguousarray(cols.imag), ox, oy, wx, wy, sx, sy, px, py, False,
                      None if out is None else base.imag.copy())
        return (re + 1j * im).astype(dt)
    u = np.dtype(f"u{dt.itemsize}") if dt != np.bool_ else np.dtype(np.uint8)
    nx, ny, flat, inside = _patch_index((ox, oy), wx, wy, sx, sy, px, py)
    acc = base.view(np.int8).view(u) if dt == np.bool_ else base.view(u)
    acc = acc.reshape(ox * oy, idm[2], idm[3], order="F").copy()
    src = cols.view(np.int8).view(u) if dt == np.bool_ else cols.view(u)
    for c in range(idm[2]):
        for n in range(idm[3]):
            v = src[:, :, c, n][inside]
            np.add.at(acc[:, c, n], flat[inside], v)
    res = acc.reshape(od, order="F")
    return res.view(np.int8) if dt == np.bool_ else res.view(dt)


def _wrap_serial(cols, base, wx, wy, sx, sy, px, py):
    """The serial scatter-add of kernel/wrap.hpp:40-72 from a given start
    (pure Python over patches; small cases only)."""
    idm = dims4(cols)
    ox, oy = base.shape[0], base.shape[1]
    out = base.copy()
    nx = nn_out_size(ox, wx, sx, px, 1)
    for n in range(idm[3]):
        for c in range(idm[2]):
            for col in range(idm[0]):
                winy, winx = divmod(col, nx)
                for y in range(wy):
                    yp = winy * sy - py + y
                    for x in range(wx):
                        xp = winx * sx - px + x
                        if 0 <= xp < ox and 0 <= yp < oy:
                            out[xp, yp, c, n] = out[xp, yp, c, n] + cols[col, y * wx + x, c, n]
    return out


def _fmod(a, shape):
    """modDims: reinterpret in column-major order."""
    return np.reshape(a, shape, order="F")


def _flipped_collapsed_filter(w):
    """flip(filter, {1,1,0,0}) then modDims [fw*fh*C, K]
    (src/backend/cpu/convolve.cpp:147-149)."""
    fd = dims4(w)
    wf = as4(w)[::-1, ::-1, :, :]
    return _fmod(wf, (fd[0] * fd[1] * fd[2], fd[3]))


def _gemm_dtype(a):
    # f16 runs its GEMM in f32 and rounds once (src/backend/cpu/blas.cpp:334-343)
    return np.float32 if a.dtype in (np.float16, np.float32) else np.float64


def conv2_nn(signal, filt, stride=(1, 1), padding=(0, 0), dilation=(1, 1)):
    """convolve2_unwrap, src/backend/cpu/convolve.cpp:123-158.

    signal [W,H,C,N], filter [fw,fh,C,K] -> [OW,OH,K,N]."""
    x, w = as4(signal), as4(filt)
    sd, fd = dims4(x), dims4(w)
    gt = _gemm_dtype(x)
    ow = nn_out_size(sd[0], fd[0], stride[0], padding[0], dilation[0])
    oh = nn_out_size(sd[1], fd[1], stride[1], padding[1], dilation[1])
    u = unwrap(x.astype(gt), fd[0], fd[1], stride[0], stride[1], padding[0],
               padding[1], dilation[0], dilation[1])
    u = np.transpose(u, (1, 2, 0, 3))                    # reorder(1,2,0,3)
    ud = u.shape
    u = _fmod(u, (ud[0] * ud[1], ud[2] * ud[3]))
    wf = _flipped_collapsed_filter(w.astype(gt))
    res = np.matmul(u.T, wf)                             # AF_MAT_TRANS, NONE
    res = _fmod(res, (ow, oh, sd[3], fd[3]))
    out = np.transpose(res, (0, 1, 3, 2))                # reorder(0,1,3,2)
    return np.asfortranarray(out).astype(x.dtype)


def _collapsed_gradient(dy):
    cd = dims4(dy)
    g = np.transpose(as4(dy), (0, 1, 3, 2))              # reorder(0,1,3,2)
    return _fmod(g, (cd[0] * cd[1] * cd[3], cd[2]))


def conv2_data_grad(dy, signal, filt, stride=(1, 1), padding=(0, 0),
                    dilation=(1, 1)):
    """conv2DataGradient, src/backend/cpu/convolve.cpp:179-211."""
    x, w = as4(signal), as4(filt)
    sd, fd = dims4(x), dims4(w)
    gt = _gemm_dtype(x)
    wf = _flipped_collapsed_filter(w.astype(gt))
    g = _collapsed_gradient(as4(dy).astype(gt))
    res = np.matmul(g, wf.T)                             # AF_MAT_NONE, TRANS
    res = _fmod(res, (res.shape[0] // sd[3], sd[3], fd[0] * fd[1], sd[2]))
    res = np.transpose(res, (0, 2, 3, 1))                # reorder(0,2,3,1)
    if x.dtype == np.float16:
        res = res.astype(np.float16).astype(np.float32)
    out = wrap_dilated(np.asfortranarray(res), sd[0], sd[1], fd[0], fd[1],
                       stride[0], stride[1], padding[0], padding[1],
                       dilation[0], dilation[1])
    return out.astype(x.dtype)


def conv2_filter_grad(dy, signal, filt, stride=(1, 1), padding=(0, 0),
                      dilation=(1, 1)):
    """conv2FilterGradient, src/backend/cpu/convolve.cpp:213-243."""
    x, w = as4(signal), as4(filt)
    fd = dims4(w)
    gt = _gemm_dtype(x)
    u = unwrap(x.astype(gt), fd[0], fd[1], stride[0], stride[1], padding[0],
               padding[1], dilation[0], dilation[1])
    u = np.transpose(u, (1, 2, 0, 3))
    ud = u.shape
    u = _fmod(u, (ud[0] * ud[1], ud[2] * ud[3]))
    g = _collapsed_gradient(as4(dy).astype(gt))
    res = np.matmul(u, g)                                # AF_MAT_NONE, NONE
    res = _fmod(res, fd)
    return np.asfortranarray(res[::-1, ::-1, :, :]).astype(x.dtype)


# ------------------------------------------------ FP64 brute-force forms
def _tap_slice(xp, i, j, d, s, ow, oh):
    """Strided view xp[i*d0 + ow*s0, j*d1 + oh*s1, :, :] for all (ow, oh)."""
    a0 = i * d[0]
    a1 = j * d[1]
    return xp[a0:a0 + (ow - 1) * s[0] + 1:s[0], a1:a1 + (oh - 1) * s[1] + 1:s[1]]


def _pad(x, p, extra=(0, 0)):
    sd = x.shape
    xp = np.zeros((sd[0] + 2 * p[0] + extra[0], sd[1] + 2 * p[1] + extra[1],
                   sd[2], sd[3]), dtype=np.float64)
    xp[p[0]:p[0] + sd[0], p[1]:p[1] + sd[1]] = x
    return xp


def conv2_nn_direct_f64(signal, filt, stride=(1, 1), padding=(0, 0),
                        dilation=(1, 1)) -> np.ndarray:
    """y[ow,oh,k,n] = sum_{c,i,j} x[ow*s0-p0+i*d0, oh*s1-p1+j*d1, c, n]
    * w[fw-1-i, fh-1-j, c, k], zero outside x; evaluated in float64."""
    x = as4(signal).astype(np.float64)
    w = as4(filt).astype(np.float64)
    sd, fd = x.shape, w.shape
    ow = nn_out_size(sd[0], fd[0], stride[0], padding[0], dilation[0])
    oh = nn_out_size(sd[1], fd[1], stride[1], padding[1], dilation[1])
    xp = _pad(x, padding, extra=(max(0, stride[0]), max(0, stride[1])))
    y = np.zeros((ow, oh, fd[3], sd[3]), dtype=np.float64)
    for i in range(fd[0]):
        for j in range(fd[1]):
            xs = _tap_slice(xp, i, j, dilation, stride, ow, oh)  # [ow,oh,C,N]
            y += np.einsum("abcn,ck->abkn", xs, w[fd[0] - 1 - i, fd[1] - 1 - j], optimize=True)
    return y


def conv2_data_grad_direct_f64(dy, signal, filt, stride=(1, 1),
                               padding=(0, 0), dilation=(1, 1)) -> np.ndarray:
    """Adjoint of conv2_nn_direct_f64 with respect to x, in float64."""
    x = as4(signal)
    w = as4(filt).astype(np.float64)
    g = as4(dy).astype(np.float64)
    sd, fd = x.shape, w.shape
    ow, oh = g.shape[0], g.shape[1]
    xp = np.zeros((sd[0] + 2 * padding[0] + stride[0],
                   sd[1] + 2 * padding[1] + stride[1], sd[2], sd[3]))
    for i in range(fd[0]):
        for j in range(fd[1]):
            t = np.einsum("abkn,ck->abcn", g, w[fd[0] - 1 - i, fd[1] - 1 - j], optimize=True)
            a0, a1 = i * dilation[0], j * dilation[1]
            xp[a0:a0 + (ow - 1) * stride[0] + 1:stride[0],
               a1:a1 + (oh - 1) * stride[1] + 1:stride[1]] += t
    return xp[padding[0]:padding[0] + sd[0], padding[1]:padding[1] + sd[1]]


def conv2_filter_grad_direct_f64(dy, signal, filt, stride=(1, 1),
                                 padding=(0, 0), dilation=(1, 1)) -> np.ndarray:
    """Adjoint of conv2_nn_direct_f64 with respect to w, in float64."""
    x = as4(signal).astype(np.float64)
    g = as4(dy).astype(np.float64)
    fd = dims4(filt)
    ow, oh = g.shape[0], g.shape[1]
    xp = _pad(x, padding, extra=(max(0, stride[0]), max(0, stride[1])))
    dw = np.zeros(fd, dtype=np.float64)
    for i in range(fd[0]):
        for j in range(fd[1]):
            xs = _tap_slice(xp, i, j, dilation, stride, ow, oh)
            dw[fd[0] - 1 - i, fd[1] - 1 - j] = np.einsum("abcn,abkn->ck", xs, g, optimize=True)
    return dw


# ------------------------------------------------------------ filters
def gaussian_kernel(rows: int, cols: int, sigma_r: float = 0.0,
                    sigma_c: float = 0.0, dtype=np.float32) -> np.ndarray:
    """src/api/c/gaussian_kernel.cpp:35-88: exp(-0.5*((i-(n-1)/2)/sigma)^2)
    summed over the two axes, normalised to sum 1; sigma <= 0 means
    0.25 * n.  Evaluated in ``dtype`` as the reference evaluates in T."""
    t = np.zeros((rows, cols), dtype=dtype)
    if cols > 1:
        sig = sigma_c if sigma_c > 0 else 0.25 * cols
        w = (np.arange(cols, dtype=dtype)[None, :] - dtype((cols - 1) / 2.0))
        w = np.broadcast_to(w / dtype(sig), (rows, cols))
        t = t + w * w
    if rows > 1:
        sig = sigma_r if sigma_r > 0 else 0.25 * rows
        w = (np.arange(rows, dtype=dtype)[:, None] - dtype((rows - 1) / 2.0))
        w = np.broadcast_to(w / dtype(sig), (rows, cols))
        t = w * w + t
    t = np.exp(-(dtype(0.5) * t)).astype(dtype)
    return (t / t.sum(dtype=dtype)).astype(dtype)


# ------------------------------------------------ internal callers
def fir(b, x) -> np.ndarray:
    """af_fir (src/api/c/iir.cpp:24-41): the first x0 rows of the EXPAND
    convolution of x with b (spatial restatement; the library may route long
    filters or many filters through the frequency domain, as AUTO does)."""
    x4, b4 = as4(x), as4(b)
    r = convolve_dispatch(x4, b4, 1, True)
    return np.ascontiguousarray(r[:dims4(x4)[0]])


def iir(b, a, x) -> np.ndarray:
    """af_iir (src/api/c/iir.cpp:49-103, src/backend/cpu/iir.cpp:23-45,
    kernel/iir.hpp:20-55): c = first x0 rows of conv(x, b, EXPAND) with the
    batch kind of iir.cpp:25-28, then per column
        y[i] = (c[i] + z[0]) / a[0];  z[k-1] = z[k] - a[k] * y[i]  (k >= 1)
    evaluated in the element type. Pure Python over rows: small cases."""
    x4, b4, a4 = as4(x), as4(b), as4(a)
    if dims4(a4)[0] == 1:
        bn = b4 / (a4 if a4.size == 1 else a4.reshape((1,) + a4.shape[1:]))
        return fir(bn.astype(b4.dtype), x4)
    # iir.cpp:25-28's batch kind (NONE / SAME / RHS / LHS by ndims) equals
    # identify_batch_kind's for every shape af_iir accepts
    c = convolve(x4, b4, 1, True)
    x0 = dims4(x4)[0]
    c = c[:x0]
    y = np.empty_like(c)
    na = dims4(a4)[0]
    per_col = a4.size > na
    for l in range(c.shape[3]):
        for k in range(c.shape[2]):
            for j in range(c.shape[1]):
                av = a4[:, j, k, l] if per_col else a4[:, 0, 0, 0]
                z = np.zeros(na, dtype=c.dtype)
                col = c[:, j, k, l]
                out = y[:, j, k, l]
                for i in range(x0):
                    v = (col[i] + z[0]) / av[0]
                    out[i] = v
                    for q in range(1, na):
                        z[q - 1] = z[q] - av[q] * v
    return y


def dog(x, r1: int, r2: int, windows=None) -> np.ndarray:
    """af_dog (src/api/c/dog.cpp:30-80): accT smoothings with the (2r+1)^2
    f32 Gaussian windows (sigma 0.25 * side) and their difference.
    ``windows`` replaces the two f32 windows (to take the library's own
    af_gaussian_kernel values, which may differ from gaussian_kernel()'s in
    the last bit of an exp)."""
    x4 = as4(x)
    at = np.float64 if x4.dtype == np.float64 else np.float32
    xa = x4.astype(at)
    if windows is None:
        windows = (gaussian_kernel(2 * r1 + 1, 2 * r1 + 1), gaussian_kernel(2 * r2 + 1, 2 * r2 + 1))
    g1, g2 = (np.asarray(w, dtype=np.float32).astype(at) for w in windows)
    return convolve(xa, g1, 2, False) - convolve(xa, g2, 2, False)


# ------------------------------------------------------------------ Harris
def harris_filter(filter_len: int, sigma: float, dtype=np.float32) -> np.ndarray:
    """The window of cpu::harris (src/backend/cpu/harris.cpp:34-44): a box
    1/filter_len for sigma < 0.5, else gaussian1D (src/backend/cpu/utility.hpp:35-48)
    evaluated per tap in double, stored in T, summed and normalised in T."""
    dt = np.dtype(dtype).type
    if sigma < 0.5:
        return np.full(filter_len, dt(1) / dt(filter_len), dtype=dtype)
    s = sigma if sigma > 0 else 0.25 * filter_len
    out = np.zeros(filter_len, dtype=dtype)
    tot = dt(0)
    for i in range(filter_len):
        x = i - (filter_len - 1) // 2
        el = dt(1.0 / math.sqrt(2 * math.pi * s * s) * math.exp(-((x * x) / (2 * (s * s)))))
        out[i] = el
        tot = dt(tot + el)
    for k in range(filter_len):
        out[k] = dt(out[k] / tot)
    return out


def gradient2(x: np.ndarray):
    """cpu::kernel::gradient (src/backend/cpu/kernel/gradient.hpp:18-80) for a
    2-D array: (grad along dim0, grad along dim1), central differences times
    0.5 inside, one-sided differences at the two ends, in T."""
    dt = x.dtype.type
    g0 = np.empty_like(x)
    g1 = np.empty_like(x)
    g0[1:-1, :] = dt(0.5) * (x[2:, :] - x[:-2, :])
    g0[0, :] = x[1, :] - x[0, :]
    g0[-1, :] = x[-1, :] - x[-2, :]
    g1[:, 1:-1] = dt(0.5) * (x[:, 2:] - x[:, :-2])
    g1[:, 0] = x[:, 1] - x[:, 0]
    g1[:, -1] = x[:, -1] - x[:, -2]
    return g0, g1


def harris(img: np.ndarray, max_corners: int, min_response: float, sigma: float,
           block_size: int, k_thr: float):
    """af_harris (src/api/c/harris.cpp:52-95) with cpu::harris
    (src/backend/cpu/harris.cpp:26-148) and its kernels
    (src/backend/cpu/kernel/harris.hpp:18-118). Returns (x, y, score) as
    float32 vectors: x is the dim1 index, y the dim0 index, in detection
    order (dim1 outer, dim0 inner) or, when max_corners > 0 and more corners
    were found, the max_corners strongest after a stable descending sort.
    Raises ValueError(af_err) where the reference's asserts fail."""
    x = np.asarray(img)
    if x.ndim != 2:
        raise ValueError(203)
    if x.dtype not in (np.float32, np.float64):
        raise ValueError(204)
    T = x.dtype.type
    filter_len = block_size if block_size != 0 else int(math.floor(np.float32(6.0) * np.float32(sigma)))
    if block_size == 0 and filter_len % 2 == 0:
        filter_len -= 1
    edge = block_size // 2 if block_size > 0 else filter_len // 2
    if not (x.shape[0] >= 2 * edge + 1 or x.shape[1] >= 2 * edge + 1):
        raise ValueError(202)
    if not (max_corners > 0 or min_response > 0.0):
        raise ValueError(202)
    if not (np.float32(k_thr) >= np.float32(0.01)):
        raise ValueError(202)
    if not (block_size > 2 or (sigma >= 0.5 and sigma <= 5.0)):
        raise ValueError(202)
    if not (block_size <= 32):
        raise ValueError(202)
    h = harris_filter(filter_len, sigma, x.dtype)
    border_len = filter_len // 2 + 1
    iy, ix = gradient2(x)
    ixx = ix * ix
    ixy = ix * iy
    iyy = iy * iy
    ixx = convolve2_sep(h, h, ixx, False)[:, :, 0, 0]
    ixy = convolve2_sep(h, h, ixy, False)[:, :, 0, 0]
    iyy = convolve2_sep(h, h, iyy, False)[:, :, 0, 0]
    d0, d1 = x.shape
    corner_lim = int(np.float32(x.size) * np.float32(0.2))
    r = border_len
    resp = np.zeros_like(x)
    sl = (slice(r, d0 - r), slice(r, d1 - r))
    tr = ixx[sl] + iyy[sl]
    det = ixx[sl] * iyy[sl] - ixy[sl] * ixy[sl]
    resp[sl] = det - T(np.float32(k_thr)) * (tr * tr)
    min_r = 0 if max_corners > 0 else int(min_response)
    mr = T(np.float32(min_r))
    xs, ys, rs = [], [], []
    count = 0
    r1 = border_len + 1
    for xi in range(r1, d1 - r1):
        for yi in range(r1, d0 - r1):
            v = resp[yi, xi]
            nb = resp[yi - 1:yi + 2, xi - 1:xi + 2].copy()
            nb[1, 1] = -np.inf
            if v > nb.max() and v >= mr:
                if count < corner_lim:
                    xs.append(np.float32(xi))
                    ys.append(np.float32(yi))
                    rs.append(np.float32(v))
                count += 1
    xs, ys, rs = (np.array(a, dtype=np.float32) for a in (xs, ys, rs))
    corners_out = min(count, max_corners if max_corners > 0 else corner_lim)
    if corners_out == 0:
        e = np.zeros(0, np.float32)
        return e, e, e
    if max_corners > 0 and count > corners_out:
        order = np.argsort(-rs, kind="stable")
        order = order[:corners_out]
        return xs[order], ys[order], rs[order]
    return xs[:corners_out], ys[:corners_out], rs[:corners_out]


# ------------------------------------------------------------------ Canny
CANNY_SMOOTH = np.array([-0.11021, -0.23691, -0.30576, -0.23691, -0.11021], dtype=np.float32)


def sobel3(x: np.ndarray):
    """cpu::kernel::derivative (src/backend/cpu/kernel/sobel.hpp:20-70) for
    one float32 slice, reflect101 borders: (dx along dim0, dy along dim1)."""
    d0, d1 = x.shape
    r = lambda idx, end: abs(end - abs(end - idx))
    i = np.arange(d0)
    j = np.arange(d1)
    im = np.array([r(v - 1, d0 - 1) for v in i]); ip = np.array([r(v + 1, d0 - 1) for v in i])
    jm = np.array([r(v - 1, d1 - 1) for v in j]); jp = np.array([r(v + 1, d1 - 1) for v in j])
    NW = x[np.ix_(im, jm)]; SW = x[np.ix_(ip, jm)]; NE = x[np.ix_(im, jp)]; SE = x[np.ix_(ip, jp)]
    N = x[np.ix_(im, j)]; S = x[np.ix_(ip, j)]; W = x[np.ix_(i, jm)]; E = x[np.ix_(i, jp)]
    two = np.float32(2)
    dx = ((SW + SE) - (NW + NE)) + two * (S - N)
    dy = ((NE + SE) - (NW + SW)) + two * (E - W)
    return dx, dy


def canny_nms(mag: np.ndarray, gx: np.ndarray, gy: np.ndarray) -> np.ndarray:
    """cpu::kernel::nonMaxSuppression (src/backend/cpu/kernel/canny.hpp:18-110)
    for one slice: interior pixels whose magnitude exceeds both values
    interpolated along the gradient keep it, others (and the border) 0."""
    d0, d1 = mag.shape
    out = np.zeros_like(mag)
    one = np.float32(1)
    for j in range(1, d1 - 1):
        for i in range(1, d0 - 1):
            curr = mag[i, j]
            if curr == 0:
                continue
            se, nw = mag[i + 1, j + 1], mag[i - 1, j - 1]
            ea, we = mag[i + 1, j], mag[i - 1, j]
            ne, sw = mag[i + 1, j - 1], mag[i - 1, j + 1]
            no, so = mag[i, j - 1], mag[i, j + 1]
            dx, dy = gx[i, j], gy[i, j]
            if dx >= 0:
                if dy >= 0:
                    g = (dx - dy) >= 0
                    a1, a2, b1, b2 = (ea if g else so), (we if g else no), se, nw
                    alpha = dy / dx if g else dx / dy
                else:
                    g = (dx + dy) >= 0
                    a1, a2, b1, b2 = (ea if g else no), (we if g else so), ne, sw
                    alpha = -dy / dx if g else dx / -dy
            else:
                if dy >= 0:
                    g = (dx + dy) >= 0
                    a1, a2, b1, b2 = (so if g else we), (no if g else ea), sw, ne
                    alpha = -dx / dy if g else dy / -dx
                else:
                    g = (-dx + dy) >= 0
                    a1, a2, b1, b2 = (we if g else no), (ea if g else so), nw, se
                    alpha = dy / dx if g else dx / dy
            with np.errstate(all="ignore"):
                m1 = (one - alpha) * a1 + alpha * b1
                m2 = (one - alpha) * a2 + alpha * b2
            if curr > m1 and curr > m2:
                out[i, j] = curr
    return out


def _otsu_threshold(sup: np.ndarray, nbins: int, maxval: np.float32) -> np.float32:
    """otsuThreshold (src/api/c/canny.cpp:82-147) for one slice: histogram
    (src/backend/cpu/kernel/histogram.hpp:19-47) over [0, maxval], scans and
    sums in float in bin order, the index of the first largest between-class
    variance (ireduce max: NaN never wins) as the threshold."""
    f32 = np.float32
    step = f32((float(maxval) - 0.0) / float(f32(nbins)))
    hist = np.zeros(nbins, np.int64)
    for v in sup.ravel(order="F"):
        b = int((f32(v) - f32(0.0)) / step) if np.isfinite((f32(v) - f32(0.0)) / step) else 0
        b = min(max(b, 0), nbins - 1)
        hist[b] += 1
    total = f32(sup.shape[0] * sup.shape[1])
    freqs = hist.astype(f32) / total
    cum = np.zeros(nbins, f32)
    cprod = np.zeros(nbins, f32)
    s = f32(0); sp = f32(0); ws = f32(0)
    for b in range(nbins):
        s = f32(s + freqs[b]); cum[b] = s
        p = f32(f32(b) * freqs[b])
        sp = f32(sp + p); cprod[b] = sp
        ws = f32(p + ws)
    sig = np.zeros(max(nbins - 1, 0), f32)
    with np.errstate(all="ignore"):
        for b in range(nbins - 1):
            qL = cum[b]
            qH = f32(f32(1) - qL)
            muL = f32(cprod[b] / qL)
            muH = f32(f32(ws - cprod[b]) / qH)
            diff = f32(muL - muH)
            sig[b] = f32(f32(diff * diff) * f32(qL * qH))
    # MinMaxOp<af_max_t> (src/backend/cpu/kernel/ireduce.hpp:47-63): starts
    # from element 0 (a NaN there becomes the lowest float, whose magnitude
    # only an infinity beats), takes larger magnitudes, keeps the first of
    # equal ones; NaN never compares larger
    if sig.size == 0:
        return f32(0)
    best_v = sig[0] if not np.isnan(sig[0]) else f32(np.finfo(np.float32).min)
    best_i = 0
    for b in range(1, sig.size):
        if abs(sig[b]) > abs(best_v):
            best_v, best_i = sig[b], b
    return f32(best_i)


def _hysteresis(strong: np.ndarray, weak: np.ndarray) -> np.ndarray:
    """cpu::kernel::edgeTrackingHysteresis (src/backend/cpu/kernel/canny.hpp:
    112-186) for one slice: seeds are the strong pixels at linear indices
    d0+1 .. d0+1+(d0-1)*(d1-1)-1 (the reference's scan), each traced through
    8-neighbours (linear offsets) that are weak."""
    d0, d1 = strong.shape
    s = strong.ravel(order="F"); w = weak.ravel(order="F")
    out = np.zeros(s.size, np.uint8)
    n = s.size
    t = d0 + 1
    for _ in range((d0 - 1) * (d1 - 1)):
        if s[t] > 0 and out[t] != 1:
            out[t] = 1
            q = [t]
            while q:
                c = q.pop()
                for it in (c - d0 - 1, c - d0, c - d0 + 1, c - 1, c + 1, c + d0 - 1, c + d0, c + d0 + 1):
                    if 0 <= it < n and w[it] > 0 and out[it] != 1:
                        out[it] = 1
                        q.append(it)
        t += 1
    return out.reshape((d0, d1), order="F")


def canny(img: np.ndarray, threshold_type: int, t1: float, t2: float, sw: int = 3,
          is_fast: bool = False) -> np.ndarray:
    """af_canny (src/api/c/canny.cpp:212-285): 5-tap separable smoothing in
    float, 3x3 Sobel, L1 or L2 magnitude, non-maximum suppression, weak and
    strong candidates (MANUAL: normalised by the suppressed map's min and
    max; AUTO_OTSU: Otsu's threshold per slice), edge tracking by hysteresis.
    Returns uint8 0/1 of the input's dims. Raises ValueError(af_err)."""
    x = np.asarray(img)
    if x.ndim < 2 or (x.ndim >= 2 and x.shape[1] == 1):
        raise ValueError(203)
    if not (x.shape[0] >= 5 and x.shape[1] >= 5):
        raise ValueError(203)
    if sw != 3:
        raise ValueError(202)
    if x.dtype not in (np.float32, np.float64, np.int32, np.uint32, np.int16, np.uint16, np.int8, np.uint8):
        raise ValueError(204)
    x4 = as4(x.astype(np.float32))
    smt = convolve2_sep(CANNY_SMOOTH, CANNY_SMOOTH, x4, False)
    d0, d1, d2, d3 = smt.shape
    sup = np.zeros_like(smt)
    for b3 in range(d3):
        for b2 in range(d2):
            gx, gy = sobel3(smt[:, :, b2, b3])
            if is_fast:
                mag = np.abs(gx) + np.abs(gy)
            else:
                mag = np.sqrt(gx * gx + gy * gy)
            with np.errstate(all="ignore"):
                sup[:, :, b2, b3] = canny_nms(mag, gx, gy)
    maxv = np.float32(sup.max())
    out = np.zeros(sup.shape, np.uint8)
    if threshold_type == 1:
        nbins = int(maxv)
        for b3 in range(d3):
            for b2 in range(d2):
                s = sup[:, :, b2, b3]
                T2 = _otsu_threshold(s, nbins, maxv)
                T1 = np.float32(T2 * np.float32(t1))
                weak = (s >= T1) & (s < T2)
                strong = s >= T2
                out[:, :, b2, b3] = _hysteresis(strong, weak)
    else:
        minv = np.float32(sup.min())
        with np.errstate(all="ignore"):
            norm = (sup - minv) / np.float32(maxv - minv)
        weak = (norm >= np.float32(t1)) & (norm < np.float32(t2))
        strong = norm >= np.float32(t2)
        for b3 in range(d3):
            for b2 in range(d2):
                out[:, :, b2, b3] = _hysteresis(strong[:, :, b2, b3], weak[:, :, b2, b3])
    return out


# -------------------------------------------------------------------- ORB
ORB_PATCH = 31  # REF_PAT_SIZE, src/backend/cpu/kernel/orb.hpp:17


def orb_pattern() -> np.ndarray:
    """The 256 (x1, y1, x2, y2) pairs of the rotated-BRIEF pattern
    (src/backend/cpu/kernel/orb.hpp:24-94), read from the product's data table
    (a table, not code: its content is pinned by a digest of the reference's
    table in tests/golden)."""
    import re

    path = os.path.join(_HERE, "..", "arrayfire_amd", "csrc", "kernels", "orb_pattern.inc")
    body = "\n".join(l for l in open(path) if not l.lstrip().startswith("//"))
    v = [int(t) for t in re.findall(r"-?\d+", body)]
    return np.array(v, dtype=np.int32).reshape(256, 4)


def resize_bilinear(img: np.ndarray, od0: int, od1: int) -> np.ndarray:
    """resize_op<T, AF_INTERP_BILINEAR> (src/backend/cpu/kernel/resize.hpp:86-130)
    for one 2-D slice: source coordinates and weights in float, the four
    weighted taps summed left to right in the working type (float for f32,
    double for f64, the weights rounded to float first)."""
    id0, id1 = img.shape
    f32 = np.float32
    xs = np.arange(od0, dtype=np.float32) / (f32(od0) / f32(id0))
    ys = np.arange(od1, dtype=np.float32) / (f32(od1) / f32(id1))
    i1x = np.minimum(np.floor(xs).astype(np.int64), id0 - 1)
    i1y = np.minimum(np.floor(ys).astype(np.int64), id1 - 1)
    b = (xs - i1x.astype(np.float32)).astype(np.float32)
    a = (ys - i1y.astype(np.float32)).astype(np.float32)
    i2x = np.where(i1x + 1 >= id0, id0 - 1, i1x + 1)
    i2y = np.where(i1y + 1 >= id1, id1 - 1, i1y + 1)
    B, A = np.meshgrid(b, a, indexing="ij")
    one = f32(1.0)
    wt = img.dtype.type
    w1 = ((one - A) * (one - B)).astype(wt)
    w2 = (A * (one - B)).astype(wt)
    w3 = ((one - A) * B).astype(wt)
    w4 = (A * B).astype(wt)
    p1 = img[np.ix_(i1x, i1y)]
    p2 = img[np.ix_(i1x, i2y)]
    p3 = img[np.ix_(i2x, i1y)]
    p4 = img[np.ix_(i2x, i2y)]
    return (((w1 * p1 + w2 * p2) + w3 * p3) + w4 * p4).astype(wt)


# The 16 offsets (dy, dx) of FAST's circle, idx_y/idx_x of
# src/backend/cpu/kernel/fast.hpp:20-31.
def _fast_circle():
    def iy(i):
        return max(-3, min(3, -(i - 12) if i >= 8 else i - 4))

    def ix(i):
        return iy(i + 4) if i < 12 else iy(i - 12)
    return [(iy(i), ix(i)) for i in range(16)]


def fast(img: np.ndarray, thr: float, arc_length: int, nonmax: bool, feature_ratio: float, edge: int):
    """cpu::fast (src/backend/cpu/fast.cpp:27-108) with locate_features and
    non_maximal (src/backend/cpu/kernel/fast.hpp:51-214): a pixel is a corner
    when arc_length contiguous circle pixels are all brighter than p + thr or
    all darker than p - thr (all in float); its score is the larger of the
    bright and dark sums of |p_x - p| - thr, summed in circle order. Corners
    are found in scan order (dim0 index y outer, dim1 index x inner) and the
    first ceil(elements * feature_ratio) are kept; suppression keeps those
    above their 8 neighbours on the score map, with the reference's edge test
    (which compares y with dims[1] and x with dims[0]). Returns float32
    (x, y, score)."""
    f32 = np.float32
    d0, d1 = img.shape
    max_feat = int(math.ceil(f32(img.size) * f32(feature_ratio)))
    thr = f32(thr)
    im = img.astype(np.float32)
    ys = np.arange(edge, d0 - edge)
    xs = np.arange(edge, d1 - edge)
    e = np.zeros(0, np.float32)
    if ys.size == 0 or xs.size == 0:
        return e, e, e
    Y, X = np.meshgrid(ys, xs, indexing="ij")  # scan order: y outer, x inner
    p = im[Y, X]
    circ = _fast_circle()
    tests = []
    for dy, dx in circ:
        px = im[Y + dy, X + dx]
        tests.append((px > p + thr).astype(np.int32) - (px < p - thr).astype(np.int32))
    t = np.stack(tests)  # [16, ny, nx]
    t2 = np.concatenate([t, t[:arc_length - 1]])
    cs = np.concatenate([np.zeros((1,) + t.shape[1:], np.int64), np.cumsum(t2, axis=0)])
    wins = cs[arc_length:] - cs[:-arc_length]  # every cyclic window of arc_length
    corner = (wins.max(axis=0) == arc_length) | (wins.min(axis=0) == -arc_length)
    sb = np.zeros(p.shape, np.float32)
    sd = np.zeros(p.shape, np.float32)
    for dy, dx in circ:
        px = im[Y + dy, X + dx]
        gt = (px > p + thr).astype(np.float32)
        lt = (px < p - thr).astype(np.float32)
        sb = sb + gt * (np.abs(px - p) - thr)
        sd = sd + lt * (np.abs(p - px) - thr)
    score = np.maximum(sb, sd)
    fy, fx = Y[corner], X[corner]
    fs = score[corner]
    n = min(max_feat, fy.size)
    fy, fx, fs = fy[:n], fx[:n], fs[:n]
    if not nonmax:
        return fx.astype(np.float32), fy.astype(np.float32), fs.astype(np.float32)
    V = np.zeros((d0, d1), np.float32)
    V[fy, fx] = fs
    v = V[fy, fx]
    mx = np.full(v.shape, -np.inf, np.float32)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            if dy or dx:
                mx = np.maximum(mx, V[fy + dy, fx + dx])
    skip = (fy >= d1 - edge - 1) | (fy <= edge + 1) | (fx >= d0 - edge - 1) | (fx <= edge + 1)
    keep = (~skip) & (v > mx)
    return fx[keep].astype(np.float32), fy[keep].astype(np.float32), v[keep].astype(np.float32)


def _orb_harris(img, xs, ys, block_size=7, k_thr=0.04, patch_size=ORB_PATCH):
    """harris_response<T, false> (src/backend/cpu/kernel/orb.hpp:110-162):
    keeps the features whose rounded position is patch_r =
    ceil(size*sqrt(2)/2) inside the image; gradients are T differences
    stored in float, the three products summed in float over the block in
    row-major block order; score = (det - k*tr^2) * 0.001^4."""
    f32 = np.float32
    d0, d1 = img.shape
    x = _round_half_away(xs.astype(np.float64)).astype(np.int64)
    y = _round_half_away(ys.astype(np.float64)).astype(np.int64)
    size = f32(2.0) * f32(math.floor(f32(patch_size) / f32(2.0))) + f32(1.0)
    pr = int(math.ceil(f32(size) * f32(math.sqrt(2.0)) / f32(2.0)))
    ok = ~((x < pr) | (y < pr) | (x >= d1 - pr) | (y >= d0 - pr))
    x, y = x[ok], y[ok]
    r = block_size // 2
    ixx = np.zeros(x.shape, np.float32)
    iyy = np.zeros(x.shape, np.float32)
    ixy = np.zeros(x.shape, np.float32)
    for k in range(block_size * block_size):
        i = k // block_size - r
        j = k % block_size - r
        ix = (img[y + j, x + i + 1] - img[y + j, x + i - 1]).astype(np.float32)
        iy = (img[y + j + 1, x + i] - img[y + j - 1, x + i]).astype(np.float32)
        ixx = ixx + ix * ix
        iyy = iyy + iy * iy
        ixy = ixy + ix * iy
    tr = ixx + iyy
    det = ixx * iyy - ixy * ixy
    resp = det - f32(k_thr) * (tr * tr)
    rs = f32(0.001)
    rs = rs * rs * rs * rs
    return x.astype(np.float32), y.astype(np.float32), (resp * rs).astype(np.float32)


def _orb_angle(img, x, y, patch_size=ORB_PATCH):
    """centroid_angle (src/backend/cpu/kernel/orb.hpp:164-193): the patch's
    first moments m01 = sum j*p, m10 = sum i*p accumulated in T in row-major
    patch order, angle = atan2(m01, m10) (evaluated in double, stored as
    float)."""
    T = img.dtype.type
    xi = x.astype(np.int64)
    yi = y.astype(np.int64)
    r = patch_size // 2
    m01 = np.zeros(xi.shape, img.dtype)
    m10 = np.zeros(xi.shape, img.dtype)
    for k in range(patch_size * patch_size):
        i = k // patch_size - r
        j = k % patch_size - r
        p = img[yi + j, xi + i]
        m01 = m01 + T(j) * p
        m10 = m10 + T(i) * p
    return np.array([np.float32(math.atan2(float(a), float(b))) for a, b in zip(m01, m10)], np.float32)


def _orb_describe(img, x, y, ori, scl, patch_size=ORB_PATCH):
    """extract_orb and get_pixel (src/backend/cpu/kernel/orb.hpp:195-264):
    each of the 256 pairs rotated by the feature's angle (sin and cos of the
    float angle, rounded to float; offsets dist*scale*cos - dist*scale*sin in
    float, rounded half away from zero), bit j of word i set when the first
    sample is below the second. Returns (desc[n, 8] uint32, x, y, size) with
    the positions scaled to the level's scale."""
    f32 = np.float32
    pat = orb_pattern()
    d0, d1 = img.shape
    n = x.size
    desc = np.zeros((n, 8), np.uint32)
    xi = x.astype(np.int64)
    yi = y.astype(np.int64)
    pscl = f32(patch_size) / f32(patch_size)
    for f in range(n):
        s = f32(math.sin(float(ori[f])))
        c = f32(math.cos(float(ori[f])))

        def px(dx, dy):
            ox = _round_half_away((dx.astype(np.float32) * pscl * c - dy.astype(np.float32) * pscl * s)
                                  .astype(np.float64))
            oy = _round_half_away((dx.astype(np.float32) * pscl * s + dy.astype(np.float32) * pscl * c)
                                  .astype(np.float64))
            return img[(yi[f] + oy).astype(np.int64), (xi[f] + ox).astype(np.int64)]
        p1 = px(pat[:, 0], pat[:, 1])
        p2 = px(pat[:, 2], pat[:, 3])
        bits = (p1 < p2).astype(np.uint64).reshape(8, 32)
        desc[f] = (bits << np.arange(32, dtype=np.uint64)).sum(axis=1).astype(np.uint32)
    sc = f32(scl)
    ox = _round_half_away((x.astype(np.float32) * sc).astype(np.float64)).astype(np.float32)
    oy = _round_half_away((y.astype(np.float32) * sc).astype(np.float64)).astype(np.float32)
    size = np.full(n, f32(patch_size) * sc, np.float32)
    return desc, ox, oy, size


_LIBM = None


def _powf(b, e) -> np.float32:
    """std::pow(float, float), i.e. the C library's powf (numpy's float32
    power rounds differently, e.g. 1.2f**3)."""
    global _LIBM
    if _LIBM is None:
        _LIBM = ctypes.CDLL("libm.so.6")
        _LIBM.powf.restype = ctypes.c_float
        _LIBM.powf.argtypes = [ctypes.c_float, ctypes.c_float]
    return np.float32(_LIBM.powf(float(np.float32(b)), float(np.float32(e))))


def orb(img: np.ndarray, fast_thr: float, max_feat: int, scl_fctr: float, levels: int, blur_img: bool):
    """af_orb (src/api/c/orb.cpp:58-89) with cpu::orb
    (src/backend/cpu/orb.cpp:44-283): per pyramid level (each level resized
    bilinearly from the previous one), FAST corners (threshold fast_thr, arc
    9, suppression, feature ratio 0.15, edge 22), their Harris responses
    (block 7, k 0.04), the lvl_best strongest by a stable descending sort,
    the centroid angle on the level image, then the rotated-BRIEF descriptor
    on the level image smoothed by the 9-tap sigma-2 Gaussian (when
    blur_img). Levels are concatenated in order. Returns (x, y, score, ori,
    size, desc[n, 8])."""
    f32 = np.float32
    x = np.asarray(img)
    if x.ndim != 2 or x.shape[0] < 7 or x.shape[1] < 7:
        raise ValueError(202)
    if not (f32(fast_thr) > 0 and max_feat > 0 and f32(scl_fctr) > 1 and levels > 0):
        raise ValueError(202)
    if x.dtype not in (np.float32, np.float64):
        raise ValueError(204)
    sf = f32(scl_fctr)
    patch = ORB_PATCH
    min_side = f32(min(x.shape))
    max_levels = 0
    scl_sum = f32(0)
    for i in range(levels):
        min_side = f32(min_side / sf)
        if min_side < patch or max_levels == levels:
            break
        max_levels += 1
        scl_sum = f32(scl_sum + f32(1) / f32(_powf(sf, i)))
    e = np.zeros(0, np.float32)
    if max_levels == 0:
        return e, e, e, e, e, np.zeros((0, 8), np.uint32)
    lvl_best = []
    feat_sum = 0
    for i in range(max_levels - 1):
        ls = f32(_powf(sf, i))
        lvl_best.append(int(math.ceil(f32(f32(max_feat) / scl_sum) / ls)))
        feat_sum += lvl_best[-1]
    lvl_best.append((max_feat - feat_sum) % (1 << 32))
    gauss = harris_filter(9, 2.0, x.dtype) if blur_img else None
    size = f32(2) * f32(math.floor(f32(patch) / f32(2))) + f32(1)
    edge = int(math.ceil(f32(size) * f32(math.sqrt(2.0)) / f32(2)))
    prev = x
    out = []
    for i in range(max_levels):
        ls = f32(_powf(sf, i))
        if i == 0:
            lvl = x
        else:
            ld0 = int(_round_half_away(np.float64(f32(f32(x.shape[0]) / ls))))
            ld1 = int(_round_half_away(np.float64(f32(f32(x.shape[1]) / ls))))
            lvl = resize_bilinear(prev, ld0, ld1)
            prev = lvl
        fx, fy, _ = fast(lvl, fast_thr, 9, True, 0.15, edge)
        if fx.size == 0:
            continue
        hx, hy, hs = _orb_harris(lvl, fx, fy)
        if hx.size == 0:
            continue
        order = np.argsort(-hs, kind="stable")
        usable = min(hx.size, lvl_best[i])
        if usable == 0:
            continue
        order = order[:usable]
        lx, ly, lsc = hx[order], hy[order], hs[order]
        lori = _orb_angle(lvl, lx, ly)
        src = convolve2_sep(gauss, gauss, lvl, False)[:, :, 0, 0] if blur_img else lvl
        desc, ox, oy, osz = _orb_describe(src, lx, ly, lori, ls)
        out.append((ox, oy, lsc, lori, osz, desc))
    if not out:
        return e, e, e, e, e, np.zeros((0, 8), np.uint32)
    return tuple(np.concatenate([o[k] for o in out]) for k in range(6))
