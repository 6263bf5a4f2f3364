"""Pins the loop-filter oracle (oracle/loop_filter.c).

* The reference's own known answers (src/loop_filter.rs:2700-3830): the
  section 15.4 parameter ladders, the hand-derived common_adjust / MBfilter
  values, the per-MB level cases, the hand-derived frame edge (columns
  13..18 of a 100 | 110 step), the frame-geometry skips and the RD selector
  cases.
* The reference's second formulation of the normal filter: the branchless
  whole-run form (src/loop_filter.rs:407-483, both outcomes computed and
  selected per position in 16-bit arithmetic), restated here in numpy and
  checked against the oracle's scalar segment kernels on random and extreme
  windows across the parameter range, as
  normal_filter_run_matches_scalar (src/loop_filter.rs:2640-2698) does.
* Frame invariants: a flat picture is a fixed point at every level and
  sharpness; level 0 everywhere is the identity; the simple filter leaves
  chroma alone; one filtered MB among level-0 MBs changes only pixels
  within three of its edges (one, for the simple filter).
* The selector: level 0 on a lossless flat picture (base sharpness
  returned), a non-zero level on a hard block edge against a smooth source
  whose SSD beats the unfiltered one; the threaded scorer returns the
  serial scores in candidate order; the sampled selector on one band that
  covers the picture is the full selector's coarse-to-fine search.
"""
import ctypes as C

import numpy as np
import pytest

from oracle_ctypes import LfMb, lf_config

BPRED, COEFFS = 1, 2


def base_cfg(**kw):
    """base_config(), src/loop_filter.rs:2987-3007."""
    d = dict(simple=0, key_frame=1, loop_filter_level=20, sharpness_level=0)
    d.update(kw)
    return lf_config(**d)


def flat(value, mb_cols, mb_rows):
    """flat_planes(), src/loop_filter.rs:3009-3021."""
    return (np.full((mb_rows * 16, mb_cols * 16), value, np.uint8),
            np.full((mb_rows * 8, mb_cols * 8), value, np.uint8),
            np.full((mb_rows * 8, mb_cols * 8), value, np.uint8))


def mbs_of(n, flags=0, ref=0, mode=0, seg=255):
    m = np.zeros((n, 4), np.uint8)
    m[:, 0], m[:, 1], m[:, 2], m[:, 3] = seg, ref, mode, flags
    return m


def seg_call(oracle, fn, args, window):
    buf = (C.c_uint8 * len(window))(*window)
    ret = getattr(oracle.lib, fn)(*args, C.addressof(buf))
    return list(buf), ret


# ---------------------------------------------------------------- params

@pytest.mark.parametrize("level,sharp,key,want", [
    (30, 0, 1, (30, 1, 94, 90)), (0, 0, 1, (1, 0, 5, 1)), (20, 2, 1, (7, 1, 51, 47)), (10, 1, 1, (5, 0, 29, 25)),
    (40, 7, 1, (2, 2, 86, 82)), (8, 5, 1, (2, 0, 22, 18)), (63, 0, 1, (63, 2, 193, 189)),
])
def test_derive_known_answers(oracle, level, sharp, key, want):
    """src/loop_filter.rs:2739-2808."""
    p = oracle.lib.orc_lf_derive(level, sharp, key)
    assert (p.interior_limit, p.hev_threshold, p.mbedge_limit, p.sub_bedge_limit) == want


def test_derive_hev_ladders(oracle):
    """src/loop_filter.rs:2770-2788."""
    for lv, want in ((45, 2), (40, 2), (39, 1), (15, 1), (14, 0)):
        assert oracle.lib.orc_lf_derive(lv, 0, 1).hev_threshold == want
    for lv, want in ((45, 3), (40, 3), (39, 2), (20, 2), (19, 1), (15, 1), (14, 0)):
        assert oracle.lib.orc_lf_derive(lv, 0, 0).hev_threshold == want


# ---------------------------------------------------------------- segments

def test_simple_segment_known_answers(oracle):
    """src/loop_filter.rs:2811-2833."""
    got, _ = seg_call(oracle, "orc_lf_simple_segment", (1,), [10, 10, 200, 200])
    assert got == [10, 10, 200, 200]
    got, _ = seg_call(oracle, "orc_lf_simple_segment", (255,), [100, 100, 116, 116])
    assert got[1] >= 100 and got[2] <= 116 and got[2] - got[1] < 16 and got[0] == 100 and got[3] == 116


def test_common_adjust_hand_derived(oracle):
    """src/loop_filter.rs:2835-2869."""
    got, a = seg_call(oracle, "orc_lf_common_adjust", (0,), [100, 100, 108, 108])
    assert a == 3 and got == [100, 103, 105, 108]
    got, a = seg_call(oracle, "orc_lf_common_adjust", (1,), [90, 100, 108, 120])
    assert a == -1 and got[1] == 99 and got[2] == 109


def test_subblock_and_mb_filter_known_answers(oracle):
    """src/loop_filter.rs:2871-2971."""
    w = [10, 200, 100, 100, 108, 108, 108, 108]
    assert seg_call(oracle, "orc_lf_subblock_filter", (0, 5, 255), w)[0] == w
    assert seg_call(oracle, "orc_lf_mb_filter", (0, 5, 255), w)[0] == w
    b = [100, 100, 100, 100, 110, 110, 110, 110]
    g = seg_call(oracle, "orc_lf_subblock_filter", (50, 100, 255), b)[0]
    assert g[3] > b[3] and g[4] < b[4] and g[2] >= b[2] and g[5] <= b[5]
    b = [100, 100, 100, 110, 120, 120, 120, 120]
    g = seg_call(oracle, "orc_lf_subblock_filter", (0, 100, 255), b)[0]
    assert g[2] == b[2] and g[5] == b[5] and g[3] > b[3] and g[4] < b[4]
    b = [100, 100, 100, 100, 112, 112, 112, 112]
    g = seg_call(oracle, "orc_lf_mb_filter", (50, 100, 255), b)[0]
    assert g[3] > b[3] and g[4] < b[4] and g[2] >= b[2] and g[5] <= b[5] and g[1] >= b[1] and g[6] <= b[6]
    assert g[0] == b[0] and g[7] == b[7]
    b = [100, 100, 100, 110, 120, 120, 120, 120]
    g = seg_call(oracle, "orc_lf_mb_filter", (0, 100, 255), b)[0]
    assert g[1] == b[1] and g[6] == b[6] and g[3] > b[3] and g[4] < b[4]
    g = seg_call(oracle, "orc_lf_mb_filter", (50, 100, 255), [100, 100, 100, 100, 116, 116, 116, 116])[0]
    assert g == [100, 102, 104, 107, 109, 112, 114, 116]


def run_form(taps, thresh, interior, edge, mb_edge):
    """normal_filter_run (src/loop_filter.rs:407-483) over an (8, N) int16
    array of signed taps, in numpy."""
    c = lambda v: np.clip(v, -128, 127).astype(np.int16)  # noqa: E731
    p3, p2, p1, p0, q0, q1, q2, q3 = (taps[k].astype(np.int16) for k in range(8))
    yes = ((np.abs(p0 - q0) * 2 + np.abs(p1 - q1) // 2) <= edge) & (np.abs(p3 - p2) <= interior) & \
        (np.abs(p2 - p1) <= interior) & (np.abs(p1 - p0) <= interior) & (np.abs(q3 - q2) <= interior) & \
        (np.abs(q2 - q1) <= interior) & (np.abs(q1 - q0) <= interior)
    hev = (np.abs(p1 - p0) > thresh) | (np.abs(q1 - q0) > thresh)
    outer = c(p1 - q1)
    step = (3 * (q0 - p0)).astype(np.int16)
    with_outer = c(outer + step)

    def adjust(a):
        return c(a + 4) >> 3, c(a + 3) >> 3
    if mb_edge:
        w = with_outer
        a0, a1, a2 = c((27 * w + 63) >> 7), c((18 * w + 63) >> 7), c((9 * w + 63) >> 7)
        a, b = adjust(with_outer)
        n = [np.where(hev, p2, c(p2 + a2)), np.where(hev, p1, c(p1 + a1)), np.where(hev, c(p0 + b), c(p0 + a0)),
             np.where(hev, c(q0 - a), c(q0 - a0)), np.where(hev, q1, c(q1 - a1)), np.where(hev, q2, c(q2 - a2))]
    else:
        a, b = adjust(np.where(hev, with_outer, c(step)))
        half = (a + 1) >> 1
        n = [p2, np.where(hev, p1, c(p1 + half)), c(p0 + b), c(q0 - a), np.where(hev, q1, c(q1 - half)), q2]
    out = taps.astype(np.int16).copy()
    for k in range(6):
        out[k + 1] = np.where(yes, n[k], taps[k + 1])
    return out


def test_segment_kernels_match_the_whole_run_form(oracle):
    """normal_filter_run_matches_scalar, src/loop_filter.rs:2640-2698."""
    rng = np.random.default_rng(15)
    filtered = 0
    for case in range(1500):
        thresh = int(rng.choice([0, 1, 2, 3, 9, 40]))
        interior = int(rng.choice([1, 2, 5, 9, 20, 63]))
        edge = int(rng.choice([0, 3, 12, 40, 90, 135, 255]))
        base = rng.integers(0, 256, 16)
        spread = rng.choice([0, 2, 6, 20, 256], 16)
        if case % 7 == 0:
            win = rng.choice([0, 255], (16, 8))
        else:
            win = np.where(spread[:, None] == 256, rng.integers(0, 256, (16, 8)),
                           np.clip(base[:, None] + rng.integers(-20, 21, (16, 8)) * spread[:, None] // 20, 0, 255))
        win = win.astype(np.uint8)
        for mb_edge in (False, True):
            got = (run_form(win.T.astype(np.int16) - 128, thresh, interior, edge, mb_edge) + 128).T.astype(np.uint8)
            fn = "orc_lf_mb_filter" if mb_edge else "orc_lf_subblock_filter"
            for n in range(16):
                want = seg_call(oracle, fn, (thresh, interior, edge), list(win[n]))[0]
                assert list(got[n]) == want, (case, mb_edge, list(win[n]), thresh, interior, edge)
                filtered += want != list(win[n])
    assert filtered > 4000


# ---------------------------------------------------------------- levels

def level(oracle, cfg, seg=255, ref=0, mode=0, flags=0):
    return oracle.lib.orc_lf_mb_level(C.byref(cfg), LfMb(seg, ref, mode, flags))


def test_mb_level_known_answers(oracle):
    """src/loop_filter.rs:3037-3117."""
    c = base_cfg()
    assert level(oracle, c) == 20 and level(oracle, c, seg=3) == 20
    c = base_cfg(segmentation_enabled=1, segment_lf_level=[5, -8, 0, 0])
    assert level(oracle, c, seg=0) == 25 and level(oracle, c, seg=1) == 12
    c = base_cfg(segmentation_enabled=1, segment_abs=1, segment_lf_level=[40, 0, 63, 7])
    assert level(oracle, c, seg=0) == 40 and level(oracle, c, seg=2) == 63 and level(oracle, c, seg=1) == 0
    c = base_cfg(segmentation_enabled=1, segment_lf_level=[60, -40, 0, 0])
    assert level(oracle, c, seg=0) == 63 and level(oracle, c, seg=1) == 0
    c = base_cfg(delta_enabled=1, ref_delta=[4, 0, 0, 0], mode_delta=[-3, 0, 0, 0])
    assert level(oracle, c) == 24 and level(oracle, c, flags=BPRED) == 21
    c = base_cfg(loop_filter_level=62, delta_enabled=1, ref_delta=[10, 0, 0, 0])
    assert level(oracle, c) == 63
    c = base_cfg(loop_filter_level=2, delta_enabled=1, ref_delta=[-10, 0, 0, 0])
    assert level(oracle, c) == 0
    c = base_cfg(ref_delta=[30, 0, 0, 0], mode_delta=[30, 0, 0, 0])
    assert level(oracle, c, flags=BPRED) == 20


def test_mb_level_inter_branches(oracle):
    """calculate_mb_filter_level_inter, src/loop_filter.rs:1280-1334: the
    ref delta by reference frame, the mode delta by (ref, inter mode); a
    segment with no id is segment 0."""
    c = base_cfg(key_frame=0, delta_enabled=1, ref_delta=[1, 2, 4, 8], mode_delta=[16, -3, -5, -7],
                 segmentation_enabled=1, segment_lf_level=[10, 0, 0, 0])
    assert level(oracle, c, seg=255) == 31  # no id is segment 0: 20 + 10, then the CURRENT delta
    assert level(oracle, c, seg=255, flags=BPRED) == 47
    assert level(oracle, c, seg=1, ref=1, mode=1) == 20 + 2 - 3
    assert level(oracle, c, seg=1, ref=2, mode=2) == 20 + 4 - 7
    assert level(oracle, c, seg=1, ref=3, mode=3) == 20 + 8 - 5
    assert level(oracle, c, seg=1, ref=3, mode=0) == 20 + 8 - 5
    assert level(oracle, c, seg=1, ref=1, mode=1, flags=BPRED) == 19  # B_PRED flag ignored for inter


# ---------------------------------------------------------------- frames

def test_level_zero_is_noop(oracle):
    """src/loop_filter.rs:3120-3134."""
    y, u, v = flat(100, 2, 1)
    y[:, 16:] = 200
    out = oracle.lf_filter_frame(y, u, v, mbs_of(2), base_cfg(loop_filter_level=0))
    assert all((a == b).all() for a, b in zip(out, (y, u, v)))


def test_normal_mb_edge_hand_derived(oracle):
    """src/loop_filter.rs:3136-3184: columns 13..18 become 101 103 104 106 107 109."""
    y, u, v = flat(100, 2, 1)
    y[:, 16:] = 110
    oy, ou, ov = oracle.lf_filter_frame(y, u, v, mbs_of(2), base_cfg())
    assert (oy[:, 13:19] == [101, 103, 104, 106, 107, 109]).all()
    assert (oy[:, :13] == y[:, :13]).all() and (oy[:, 19:] == y[:, 19:]).all()
    assert (ou == u).all() and (ov == v).all()


def test_leftmost_mb_skips_left_edge(oracle):
    """src/loop_filter.rs:3186-3241."""
    y, u, v = flat(100, 1, 2)
    y[16:] = 120
    u[8:] = 120
    v[8:] = 120
    oy, ou, ov = oracle.lf_filter_frame(y, u, v, mbs_of(2), base_cfg())
    assert (oy[:13] == y[:13]).all() and (oy[19:] == y[19:]).all()
    assert oy[15, 0] != y[15, 0] and oy[16, 0] != y[16, 0]
    assert ou[7, 3] != u[7, 3] and ov[8, 3] != v[8, 3]


def test_simple_touches_only_luma(oracle):
    """src/loop_filter.rs:3243-3272."""
    y, u, v = flat(100, 2, 1)
    y[:, 16:] = 110
    u[:, 8:] = 130
    oy, ou, ov = oracle.lf_filter_frame(y, u, v, mbs_of(2), base_cfg(simple=1))
    assert oy[0, 15] != y[0, 15] and oy[0, 16] != y[0, 16]
    assert (ou == u).all() and (ov == v).all()


def test_subblock_steps_gated_by_coeffs_and_bpred(oracle):
    """src/loop_filter.rs:3274-3356."""
    y, u, v = flat(100, 1, 1)
    y[:, 8:] = 112
    out = oracle.lf_filter_frame(y, u, v, mbs_of(1), base_cfg())
    assert (out[0] == y).all()
    out = oracle.lf_filter_frame(y, u, v, mbs_of(1, flags=COEFFS), base_cfg())
    assert out[0][0, 7] != y[0, 7] and out[0][0, 8] != y[0, 8]
    out = oracle.lf_filter_frame(y, u, v, mbs_of(1, flags=BPRED), base_cfg())
    assert (out[0] != y).any()


def lcg_frame(mb_cols, mb_rows, seed):
    """A textured picture with smooth regions so the filters engage (the luma
    recipe of banded_filter_matches_serial, src/loop_filter.rs:3407-3430)."""
    rng = np.random.default_rng(seed)
    n = mb_cols * mb_rows
    y = rng.integers(0, 256, n * 256).astype(np.int64)
    y = ((y % 24) + (np.arange(n * 256) // 7) % 200).astype(np.uint8).reshape(mb_rows * 16, mb_cols * 16)
    # smooth chroma too, so the chroma edges engage
    k = np.arange(n * 64)
    u = ((rng.integers(0, 256, n * 64) % 12) + (k // 5) % 180).astype(np.uint8).reshape(mb_rows * 8, mb_cols * 8)
    v = ((rng.integers(0, 256, n * 64) % 12) + 40 + (k // 3) % 150).astype(np.uint8).reshape(mb_rows * 8, mb_cols * 8)
    return y, u, v


def test_flat_picture_is_a_fixed_point(oracle):
    for simple in (0, 1):
        for lv in (1, 9, 30, 63):
            for sh in (0, 3, 7):
                y, u, v = flat(77, 3, 2)
                out = oracle.lf_filter_frame(y, u, v, mbs_of(6, flags=COEFFS), base_cfg(
                    simple=simple, loop_filter_level=lv, sharpness_level=sh))
                assert all((a == b).all() for a, b in zip(out, (y, u, v)))


def test_one_filtered_mb_changes_only_pixels_near_its_edges(oracle):
    """Segment overrides give MB (1, 1) of a 3x3 frame level 40 and every
    other MB level 0; the normal filter moves at most three pixels either
    side of the MB's left/top edges and its inner edges (p2..q2 of an MB
    edge, p1..q1 of a sub-block edge), the simple filter one."""
    y, u, v = lcg_frame(3, 3, 4)
    mbs = mbs_of(9, flags=COEFFS, seg=0)
    mbs[4, 0] = 1
    for simple in (0, 1):
        cfg = base_cfg(simple=simple, segmentation_enabled=1, segment_abs=1, segment_lf_level=[0, 40, 0, 0])
        oy, ou, ov = oracle.lf_filter_frame(y, u, v, mbs, cfg)
        reach = 1 if simple else 3
        allowed = np.zeros_like(y, bool)
        allowed[16:32, 16 - reach:32] = True
        allowed[16 - reach:32, 16:32] = True
        assert ((oy != y) <= allowed).all()
        assert (oy != y).any()
        if simple:
            assert (ou == u).all() and (ov == v).all()
        else:
            ca = np.zeros_like(u, bool)
            ca[8:16, 8 - reach:16] = True
            ca[8 - reach:16, 8:16] = True
            assert ((ou != u) <= ca).all() and ((ov != v) <= ca).all()


def run_frame(y, u, v, mbs, cfg_fields):
    """The frame pass of filter_inter_frame_flags (src/loop_filter.rs:1364-1496)
    on top of the whole-run kernel, for the normal filter."""
    y, u, v = (a.astype(np.int16) - 128 for a in (y, u, v))
    rows, cols = y.shape[0] // 16, y.shape[1] // 16

    def edge(pl, x0, y0, n, vertical, p, mb_edge):
        el = p.mbedge_limit if mb_edge else p.sub_bedge_limit
        if vertical:
            t = pl[y0:y0 + n, x0 - 4:x0 + 4].T
            pl[y0:y0 + n, x0 - 4:x0 + 4] = run_form(t, p.hev_threshold, p.interior_limit, el, mb_edge).T
        else:
            t = pl[y0 - 4:y0 + 4, x0:x0 + n]
            pl[y0 - 4:y0 + 4, x0:x0 + n] = run_form(t, p.hev_threshold, p.interior_limit, el, mb_edge)
    for r in range(rows):
        for c in range(cols):
            lv, p, sub = cfg_fields(r * cols + c)
            if lv == 0:
                continue
            if c > 0:
                edge(y, 16 * c, 16 * r, 16, True, p, True)
                edge(u, 8 * c, 8 * r, 8, True, p, True)
                edge(v, 8 * c, 8 * r, 8, True, p, True)
            if sub:
                for q in (1, 2, 3):
                    edge(y, 16 * c + 4 * q, 16 * r, 16, True, p, False)
                edge(u, 8 * c + 4, 8 * r, 8, True, p, False)
                edge(v, 8 * c + 4, 8 * r, 8, True, p, False)
            if r > 0:
                edge(y, 16 * c, 16 * r, 16, False, p, True)
                edge(u, 8 * c, 8 * r, 8, False, p, True)
                edge(v, 8 * c, 8 * r, 8, False, p, True)
            if sub:
                for q in (1, 2, 3):
                    edge(y, 16 * c, 16 * r + 4 * q, 16, False, p, False)
                edge(u, 8 * c, 8 * r + 4, 8, False, p, False)
                edge(v, 8 * c, 8 * r + 4, 8, False, p, False)
    return tuple((a + 128).astype(np.uint8) for a in (y, u, v))


@pytest.mark.parametrize("shape", [(1, 2), (2, 2), (5, 3), (3, 7)])
def test_frame_pass_matches_the_whole_run_form(oracle, shape):
    """The raster pass of the oracle equals the same pass built on the
    whole-run kernel, on the shapes and level mixes of
    banded_filter_matches_serial (src/loop_filter.rs:3407-3500)."""
    mb_cols, mb_rows = shape
    n = mb_cols * mb_rows
    changed = 0
    for case in range(6):
        y, u, v = lcg_frame(mb_cols, mb_rows, case + 10 * mb_cols)
        rng = np.random.default_rng(case)
        mbs = mbs_of(n)
        mbs[:, 1] = rng.choice([0, 1, 2], n)
        mbs[:, 2] = np.where(mbs[:, 1] > 0, rng.choice([1, 3, 2], n), 0)
        mbs[:, 3] = rng.choice([0, BPRED], n) | np.where(rng.integers(0, 3, n) == 0, COEFFS, 0)
        cfg = base_cfg(key_frame=0, loop_filter_level=[9, 30, 63][case // 2], sharpness_level=case % 4,
                       delta_enabled=int(case >= 2), ref_delta=[-40, 0, 12, 0], mode_delta=[0, -6, 0, 0])
        want = oracle.lf_filter_frame(y, u, v, mbs, cfg)
        changed += int((want[0] != y).any())

        def fields(i):
            lv = level(oracle, cfg, *(int(x) for x in mbs[i]))
            return lv, oracle.lib.orc_lf_derive(lv, cfg.sharpness_level, cfg.key_frame), bool(mbs[i, 3])
        got = run_frame(y, u, v, mbs, fields)
        for g, w in zip(got, want):
            assert (g == w).all(), (shape, case)
    assert changed >= 3


# ---------------------------------------------------------------- selection

def src_of(y, u, v, w=None, h=None):
    return (w or y.shape[1], h or y.shape[0], y, u, v)


def test_select_flat_source_picks_zero(oracle):
    """src/loop_filter.rs:3667-3694 and 3774-3831."""
    y, u, v = flat(120, 2, 2)
    mbs = mbs_of(4, flags=COEFFS)
    assert oracle.lf_select(y, u, v, mbs, base_cfg(), src_of(y, u, v)) == (0, 0)
    assert oracle.lf_select(y, u, v, mbs, base_cfg(sharpness_level=3), src_of(y, u, v)) == (0, 3)
    assert oracle.lf_ssd(y, u, v, src_of(y, u, v)) == 0


def test_select_engages_on_block_edge(oracle):
    """src/loop_filter.rs:3696-3772."""
    y, u, v = flat(100, 2, 1)
    y[:, 16:] = 160
    sy = np.tile((100 + np.arange(32) * 60 // 31).astype(np.uint8), (16, 1))
    src = (32, 16, sy, u, v)
    mbs = mbs_of(2, flags=COEFFS)
    unfiltered = oracle.lf_ssd(y, u, v, src)
    lv, sh = oracle.lf_select(y, u, v, mbs, base_cfg(), src)
    assert lv > 0 and sh <= 7
    out = oracle.lf_filter_frame(y, u, v, mbs, base_cfg(loop_filter_level=lv, sharpness_level=sh))
    assert oracle.lf_ssd(*out, src) < unfiltered


def test_visible_window_ssd(oracle):
    """reconstruction_ssd (src/loop_filter.rs:1941-1979) over an odd visible
    window equals numpy's sum over the window (chroma ceil(w/2) x ceil(h/2))."""
    rng = np.random.default_rng(3)
    y, u, v = lcg_frame(3, 2, 1)
    sy = rng.integers(0, 256, (40, 50)).astype(np.uint8)
    su = rng.integers(0, 256, (20, 27)).astype(np.uint8)
    sv = rng.integers(0, 256, (20, 27)).astype(np.uint8)
    w, h = 45, 29
    want = sum(int(((a[:hh, :ww].astype(np.int64) - b[:hh, :ww]) ** 2).sum())
               for a, b, ww, hh in ((y, sy, w, h), (u, su, 23, 15), (v, sv, 23, 15)))
    assert oracle.lf_ssd(y, u, v, (w, h, sy, su, sv)) == want


def test_scores_in_candidate_order_and_level_zero_is_unfiltered(oracle):
    """score_candidates returns the serial scores in candidate order at any
    thread count (src/loop_filter.rs:2097-2131); level 0 is the unfiltered
    SSD even when absolute segment levels would filter
    (src/loop_filter.rs:2062-2074)."""
    y, u, v = lcg_frame(4, 3, 7)
    sy = np.clip(y.astype(np.int16) + 3, 0, 255).astype(np.uint8)
    src = src_of(sy, u, v, 60, 40)
    mbs = mbs_of(12, flags=COEFFS)
    cfg = base_cfg(segmentation_enabled=1, segment_abs=1, segment_lf_level=[30, 30, 30, 30])
    lv = np.array([0, 5, 17, 63, 40, 0, 1], np.uint8)
    sh = np.array([0, 1, 7, 3, 0, 5, 2], np.uint8)
    one = oracle.lf_score_candidates(y, u, v, mbs, cfg, src, lv, sh, threads=1)
    for t in (2, 5, 16):
        assert (oracle.lf_score_candidates(y, u, v, mbs, cfg, src, lv, sh, threads=t) == one).all()
    assert one[0] == one[5] == oracle.lf_ssd(y, u, v, src)


def test_sampled_selector_on_one_band_is_coarse_to_fine(oracle):
    """select_filter_params_sampled (src/loop_filter.rs:2288-2412) on a
    two-row picture has one band covering it; its pick is the first strict
    minimum of the every-eighth grid refined by +-4, +-2, +-1, recomputed
    here from the oracle's candidate scores."""
    y, u, v = lcg_frame(6, 2, 9)
    sy = np.clip(y.astype(np.int16) - 2, 0, 255).astype(np.uint8)
    src = src_of(sy, u, v)
    mbs = mbs_of(12, flags=COEFFS)
    cfg = base_cfg(sharpness_level=2)
    scores = oracle.lf_score_candidates(y, u, v, mbs, cfg, src, np.arange(64), np.full(64, 2), threads=8)
    grid = list(range(0, 57, 8)) + [63]
    best = grid[0]
    for lv in grid[1:]:
        if scores[lv] < scores[best]:
            best = lv
    for st in (4, 2, 1):
        cands = [c for c in (best - st, best + st) if 0 <= c <= 63]
        for c in cands:
            if scores[c] < scores[best]:
                best = c
    got = oracle.lf_select_sampled(y, u, v, mbs, cfg, src, every_rows=8, sweep_sharpness=False)
    assert got == (best, 2)
