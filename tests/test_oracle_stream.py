"""Pins the oracle's luma_mad (src/stream.rs:2654-2666): the u64 sum of
|prev - cur| over the visible luma, divided as f64 by w * h.  Checked
against numpy's integer sum on random strided frames (odd sizes, padded
strides, the 1x1 frame) and against closed forms: identical frames give
0, a frame against its 255-complement gives the mean of |2p - 255|, a
constant offset d gives d."""
import ctypes as C

import numpy as np


def mad(oracle, w, h, a, sa, b, sb):
    from oracle_ctypes import I420
    z = np.zeros(1, np.uint8)
    fa = I420(w, h, a.ctypes.data, z.ctypes.data, z.ctypes.data, sa, 1)
    fb = I420(w, h, b.ctypes.data, z.ctypes.data, z.ctypes.data, sb, 1)
    oracle.lib.orc_luma_mad.restype = C.c_double
    return oracle.lib.orc_luma_mad(C.byref(fa), C.byref(fb))


def test_luma_mad_equals_the_integer_sum_over_the_visible_luma(oracle):
    rng = np.random.default_rng(5)
    for (w, h, pa, pb) in ((1, 1, 0, 0), (3, 2, 1, 0), (37, 29, 5, 11), (64, 64, 0, 0), (353, 17, 0, 9)):
        a = rng.integers(0, 256, (h, w + pa)).astype(np.uint8)
        b = rng.integers(0, 256, (h, w + pb)).astype(np.uint8)
        want = float(np.abs(a[:, :w].astype(np.int64) - b[:, :w]).sum()) / float(w * h)
        got = mad(oracle, w, h, a.reshape(-1), w + pa, b.reshape(-1), w + pb)
        assert got == want, (w, h)


def test_luma_mad_closed_forms(oracle):
    rng = np.random.default_rng(6)
    w, h = 41, 23
    a = rng.integers(0, 256, (h, w)).astype(np.uint8)
    assert mad(oracle, w, h, a.reshape(-1), w, a.reshape(-1).copy(), w) == 0.0
    c = (255 - a).astype(np.uint8)
    want = float(np.abs(2 * a.astype(np.int64) - 255).sum()) / (w * h)
    assert mad(oracle, w, h, a.reshape(-1), w, c.reshape(-1), w) == want
    flat = np.full((h, w), 100, np.uint8)
    for d in (0, 1, 17, 155):
        assert mad(oracle, w, h, flat.reshape(-1), w, (flat + d).astype(np.uint8).reshape(-1), w) == float(d)
