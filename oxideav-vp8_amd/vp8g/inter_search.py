"""Host side of the encoder's inter motion search on the GPU
(src/encoder/inter_search.rs:166-552, src/motion_search.rs:381-442).

* ``MvCostTable.new(contexts)`` tables the f32 bit cost of every MV
  differential component (src/encoder/inter_search.rs:166-178) from the
  frame's two 19-byte MV contexts with the reference's table-backed
  ``mv_component_bits`` (src/motion_vector.rs:345-398) and ``bool_bits``
  (src/encoder.rs:2229-2260).
* ``whole_pixel_search_batch``, ``sub_pixel_refine_batch`` and
  ``small_diamond_search_luma_batch`` run one macroblock per wave through
  libvp8g.so (vp8g_*_batch in include/vp8g.h).

Job and result records are numpy structured arrays laid out as the C ABI's
``vp8g_mb_search`` and ``vp8g_found``.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import Context

# RFC 6386 section 17.2 default MV contexts (src/coded_header.rs:957-967):
# row then column, 19 probabilities each.
DEFAULT_MV_CONTEXT = np.array(
    [[162, 128, 225, 146, 172, 147, 214, 39, 156, 128, 129, 132, 75, 145, 178, 206, 239, 254, 254],
     [164, 128, 204, 170, 119, 235, 140, 230, 228, 128, 130, 130, 74, 148, 180, 203, 236, 254, 254]], np.uint8)

MVP_IS_SHORT, MVP_SIGN, MVP_SHORT, MVP_BITS = 0, 1, 2, 9  # src/motion_vector.rs:83-90
SMALL_MVTREE = (2, 8, 4, 6, -0, -1, -2, -3, 10, 12, -4, -5, -6, -7)  # src/motion_vector.rs:108-116

# BIT_COST_BY_FALSE_PROB[p] = -log2(p / 256), p = 0 floored at 1/256
# (src/encoder.rs:2251-2260).
_BIT_COST = [-math.log2(1.0 / 256.0)] + [-math.log2(p / 256.0) for p in range(1, 256)]


def bool_bits(prob: int, value: bool) -> float:
    """src/encoder.rs:2229-2240."""
    return _BIT_COST[min(256 - prob, 255)] if value else _BIT_COST[prob]


def _small_mv_bits(p, leaf: int) -> float:
    bits, i = 0.0, 0
    for k in (2, 1, 0):
        bit = (leaf >> k) & 1
        bits += bool_bits(int(p[MVP_SHORT + (i >> 1)]), bool(bit))
        nxt = SMALL_MVTREE[i + bit]
        if nxt <= 0:
            break
        i = nxt
    return bits


def mv_component_bits(p, value: int) -> float:
    """src/motion_vector.rs:345-398: fractional bits of one MV component."""
    a = abs(int(value))
    bits = 0.0
    if a > 7:
        bits += bool_bits(int(p[MVP_IS_SHORT]), True)
        for i in range(3):
            bits += bool_bits(int(p[MVP_BITS + i]), bool((a >> i) & 1))
        i = 9
        while True:
            bits += bool_bits(int(p[MVP_BITS + i]), bool((a >> i) & 1))
            i -= 1
            if i <= 3:
                break
        if a & 0xFFF0:
            bits += bool_bits(int(p[MVP_BITS + 3]), bool((a >> 3) & 1))
    else:
        bits += bool_bits(int(p[MVP_IS_SHORT]), False)
        bits += _small_mv_bits(p, a)
    if a != 0:
        bits += bool_bits(int(p[MVP_SIGN]), value < 0)
    return bits


@dataclass
class MvCostTable:
    """src/encoder/inter_search.rs:166-194: bits[component][diff + 1023]
    (component 0 = row) as f32."""
    bits: np.ndarray  # float32 [2, 2047]

    @staticmethod
    def new(contexts=DEFAULT_MV_CONTEXT) -> "MvCostTable":
        t = np.empty((2, 2047), np.float32)
        for comp in range(2):
            p = contexts[comp]
            t[comp] = [np.float32(mv_component_bits(p, d)) for d in range(-1023, 1024)]
        return MvCostTable(t)

    def diff_bits(self, mv_row: int, mv_col: int, base_row: int, base_col: int):
        """MvCostTable::bits: the f32 sum widened, or None out of range."""
        dr, dc = mv_row - base_row, mv_col - base_col
        if not (-1023 <= dr <= 1023 and -1023 <= dc <= 1023):
            return None
        return float(np.float32(self.bits[0, dr + 1023]) + np.float32(self.bits[1, dc + 1023]))


MB_SEARCH = np.dtype([("mb_col", "<u4"), ("mb_row", "<u4"), ("best", "<i2", (2,)), ("range_x", "<i4", (2,)),
                      ("range_y", "<i4", (2,)), ("n_seeds", "<u4"), ("seeds", "<i2", (8, 2)),
                      ("lambda_sad", "<f8"), ("good_enough", "<u4"), ("descend", "<i4")], align=True)
FOUND = np.dtype([("mv", "<i2", (2,)), ("sad", "<u4"), ("cost", "<f8")], align=True)
assert MB_SEARCH.itemsize == 80 and FOUND.itemsize == 16


@dataclass(frozen=True)
class SearchShape:
    """The SearchShape fields the searches read (src/encoder/inter_search.rs:66-93)."""
    descent_rounds: int = 8
    wide_pattern: bool = True
    half_diagonals: bool = True
    quarter: bool = True

    def c(self):
        return (C.c_uint32 * 4)(self.descent_rounds, int(self.wide_pattern), int(self.half_diagonals),
                                int(self.quarter))


# EncodeEffort::inter_shape (src/encoder/inter_search.rs:95-163).
SHAPES = {"fastest": SearchShape(4, False, True, False), "fast": SearchShape(8, True, True, True),
          "balanced": SearchShape(16, True, True, True), "thorough": SearchShape(32, True, True, True)}


def _jobs(jobs) -> np.ndarray:
    j = np.ascontiguousarray(jobs, MB_SEARCH)
    return j.reshape(-1)


def whole_pixel_search_batch(ctx: Context, reference, costs: MvCostTable, jobs, src, shape: SearchShape):
    """whole_pixel_search for every job (structured MB_SEARCH array) with
    sources src[n, 256]; returns a FOUND array."""
    j = _jobs(jobs)
    n = j.size
    s = np.ascontiguousarray(src, np.uint8).reshape(n, 256)
    p, lr = ctx._luma_ref(reference)
    b = np.ascontiguousarray(costs.bits, np.float32)
    out = np.zeros(max(n, 1), FOUND)
    st = ctx.lib.vp8g_whole_pixel_search_batch(ctx.h, C.byref(lr), b.ctypes.data, j.ctypes.data, s.ctypes.data, n,
                                               shape.c(), out.ctypes.data)
    ctx._raise(st, "vp8g_whole_pixel_search_batch")
    return out[:n]


def sub_pixel_refine_batch(ctx: Context, reference, costs: MvCostTable, jobs, src, start, filters,
                           shape: SearchShape):
    """sub_pixel_refine from each start (FOUND array); returns a FOUND array."""
    j = _jobs(jobs)
    n = j.size
    s = np.ascontiguousarray(src, np.uint8).reshape(n, 256)
    st0 = np.ascontiguousarray(start, FOUND).reshape(-1)
    p, lr = ctx._luma_ref(reference)
    b = np.ascontiguousarray(costs.bits, np.float32)
    f = np.ascontiguousarray(filters, np.int32).reshape(8, 6)
    out = np.zeros(max(n, 1), FOUND)
    st = ctx.lib.vp8g_sub_pixel_refine_batch(ctx.h, C.byref(lr), b.ctypes.data, j.ctypes.data, s.ctypes.data,
                                             st0.ctypes.data, n, f.ctypes.data, shape.c(), out.ctypes.data)
    ctx._raise(st, "vp8g_sub_pixel_refine_batch")
    return out[:n]


def small_diamond_search_luma_batch(ctx: Context, reference, mb_xy, src, centers, max_iters: int) -> np.ndarray:
    """small_diamond_search_luma per macroblock: rows of (mv.row, mv.col, sad)."""
    mb_xy = np.ascontiguousarray(mb_xy, np.uint32).reshape(-1, 2)
    n = mb_xy.shape[0]
    s = np.ascontiguousarray(src, np.uint8).reshape(n, 256)
    c = np.ascontiguousarray(centers, np.int16).reshape(n, 2)
    p, lr = ctx._luma_ref(reference)
    out = np.empty((max(n, 1), 4), np.uint16)
    st = ctx.lib.vp8g_small_diamond_search_luma_batch(ctx.h, C.byref(lr), mb_xy.ctypes.data, s.ctypes.data,
                                                      c.ctypes.data, n, int(max_iters), out.ctypes.data)
    ctx._raise(st, "vp8g_small_diamond_search_luma_batch")
    res = np.empty((n, 3), np.int64)
    res[:, 0] = out[:n, 0].view(np.int16)
    res[:, 1] = out[:n, 1].view(np.int16)
    res[:, 2] = out[:n, 2].astype(np.int64) | (out[:n, 3].astype(np.int64) << 16)
    return res
