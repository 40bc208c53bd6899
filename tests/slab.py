"""Row-slab check for outputs too large to run through the oracle whole."""
import numpy as np

import oracle


def slab_check(got, x, k, conv_slab, fy, sx, sy, y0):
    """Output rows y0 .. y0 + 15 of slice k of got against the oracle run in
    EXPAND mode on the input row slab they read. Output (x, y) reads input
    rows y + sy - (fy - 1) .. y + sy, which is expand-mode slab output
    (x + sx, y + sy - a); (sx, sy) = 0 for EXPAND, the filter half-widths
    otherwise."""
    H, ow = x.shape[1], got.shape[0]
    a, b = max(0, y0 + sy - (fy - 1)), min(H, y0 + sy + 16)
    ref = conv_slab(np.ascontiguousarray(x[:, a:b, k]))[:, :, 0, 0]
    for y in range(y0, min(y0 + 16, got.shape[1])):
        assert oracle.rel_l2(got[:, y, k, 0], ref[sx:sx + ow, y + sy - a]) <= 1e-5
