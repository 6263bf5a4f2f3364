"""Pins the ARNR half of the CPU oracle (oracle/vp8_oracle.c) before it is
trusted as the GPU checker.

Each function is held by checks that a plausible mistake would fail:
* weight_table: closed-form values of 16*S/(S+d^2) (src/arnr.rs:307-323);
* block_sad / block_sad_generic: fast == generic on every probe MV
  (src/arnr.rs:863-885) and == an independent numpy clamped gather;
* refine_search: == the descent walked over a brute-force numpy SAD table
  of all 31x31 candidates (src/arnr.rs:891-963);
* accumulate_luma_block / accumulate_chroma_block: == the per-pixel
  reference loops (src/arnr.rs:968-1135), incl. the odd 25x23 overhang;
* build_arnr_altref: == an independent per-pixel gather formulation of the
  blend, plus the reference's invariants (identity, denoise, exclusion,
  dimension errors, fixed point, re-entry, thread invariance).
"""
import json
from pathlib import Path

import numpy as np
import pytest

from conftest import H, W, clean_scene, mse, noisy, packed, translated

GOLDEN = Path(__file__).resolve().parent / "golden"


# --------------------------------------------------------------- helpers
def np_sad_table(src, ref, w, h, bx, by, bw, bh, r=15):
    """SAD of the (bx,by,bw,bh) block of `src` against `ref` displaced by
    every (mvx, mvy) in [-r, r]^2 with visible-dim clamping."""
    s = src.reshape(h, w)[by:by + bh, bx:bx + bw].astype(np.int64)
    R = ref.reshape(h, w).astype(np.int64)
    t = np.zeros((2 * r + 1, 2 * r + 1), np.int64)
    ys0, xs0 = np.arange(by, by + bh), np.arange(bx, bx + bw)
    for my in range(-r, r + 1):
        ys = np.clip(ys0 + my, 0, h - 1)
        for mx in range(-r, r + 1):
            xs = np.clip(xs0 + mx, 0, w - 1)
            t[my + r, mx + r] = np.abs(s - R[ys][:, xs]).sum()
    return t


def walk_descent(t, r=15):
    best, bs = (0, 0), t[r, r]
    step = 8
    while step >= 1:
        cx, cy = best
        for dx, dy in [(-step, -step), (0, -step), (step, -step), (-step, 0), (step, 0),
                       (-step, step), (0, step), (step, step)]:
            px, py = cx + dx, cy + dy
            if abs(px) > 15 or abs(py) > 15:
                continue
            if t[py + r, px + r] < bs:
                bs, best = t[py + r, px + r], (px, py)
        step //= 2
    return best[0], best[1], int(bs)


def descent_ties(t, r=15):
    """Probes whose SAD equals the incumbent's during the descent (where
    the strict '<' and the probe order decide the outcome)."""
    best, bs, ties = (0, 0), t[r, r], 0
    step = 8
    while step >= 1:
        cx, cy = best
        for dx, dy in [(-step, -step), (0, -step), (step, -step), (-step, 0), (step, 0),
                       (-step, step), (0, step), (step, step)]:
            v = t[cy + dy + r, cx + dx + r]
            ties += int(v == bs)
            if v < bs:
                bs, best = v, (cx + dx, cy + dy)
        step //= 2
    return ties


def np_blend(frames, center, strength, mvs):
    """Per-pixel gather formulation of the blend + resolve given the search
    table: every output pixel sums its own contributions directly."""
    w, h = frames[center][0], frames[center][1]
    cw, ch = (w + 1) // 2, (h + 1) // 2

    def plane(f, k):
        stride = f[5] if k == 2 else f[6]
        pw, ph = (w, h) if k == 2 else (cw, ch)
        a = np.asarray(f[k], np.int64)
        return np.stack([a[r * stride:r * stride + pw] for r in range(ph)])

    S = 8 << min(strength, 6)
    out = []
    for k, (pw, ph, sub) in zip((2, 3, 4), ((w, h, 1), (cw, ch, 2), (cw, ch, 2))):
        ctr = plane(frames[center], k)
        acc = 16 * ctr
        ws = np.full_like(acc, 16)
        if strength != 0 and len(frames) >= 2:
            yy, xx = np.meshgrid(np.arange(ph), np.arange(pw), indexing="ij")
            by_ = (yy * sub) // 16
            bx_ = (xx * sub) // 16
            for fi, f in enumerate(frames):
                if fi == center:
                    continue
                ref = plane(f, k)
                mvx = mvs[fi, by_, bx_, 0]
                mvy = mvs[fi, by_, bx_, 1]
                sad = mvs[fi, by_, bx_, 2]
                bw = np.minimum(16, w - bx_ * 16)
                bh = np.minimum(16, h - by_ * 16)
                keep = (sad // (bw * bh)) <= 12
                if sub == 2:
                    mvx = np.trunc(mvx / 2).astype(np.int64)
                    mvy = np.trunc(mvy / 2).astype(np.int64)
                rp = ref[np.clip(yy + mvy, 0, ph - 1), np.clip(xx + mvx, 0, pw - 1)]
                d = rp - ctr
                wt = (16 * S) // (S + d * d)
                acc = acc + np.where(keep, wt * rp, 0)
                ws = ws + np.where(keep, wt, 0)
        out.append(((acc + ws // 2) // ws).astype(np.uint8).reshape(-1))
    return out


def stress_pair():
    cy, cu, _ = clean_scene()
    return cy, noisy(cy, 4242, 20), cu, noisy(cu, 2424, 12)


# --------------------------------------------------------------- weights
def test_weight_table_closed_form(oracle):
    import ctypes as C
    for strength in range(7):
        S = 8 << strength
        t = (C.c_uint32 * 256)()
        oracle.lib.orc_weight_table(S, t)
        for d in range(-255, 256):
            assert oracle.lib.orc_pixel_weight(d, S) == (16 * S) // (S + d * d)
            assert t[abs(d)] == (16 * S) // (S + d * d)
    # Hand values at strength 3 (S = 64): 16, 1024/65 -> 15, 1024/128 -> 8, 0.
    assert [oracle.lib.orc_pixel_weight(d, 64) for d in (0, 1, 8, 255)] == [16, 15, 8, 0]


# --------------------------------------------------------------- SAD + search
def test_block_sad_fast_matches_generic_and_numpy(oracle):
    cy, ry, cu, _ = stress_pair()
    keep = []
    src = oracle.i420(*packed(W, H, cy, cu, cu), keep=keep)
    ref = oracle.i420(*packed(W, H, ry, cu, cu), keep=keep)
    for blk in [(16, 16, 16, 16), (0, 0, 16, 16), (32, 32, 16, 16), (32, 0, 16, 16), (0, 32, 16, 16)]:
        table = np_sad_table(cy, ry, W, H, *blk, r=16)
        for mvy in range(-16, 17):
            for mvx in range(-16, 17):
                f = oracle.lib.orc_block_sad(src, ref, *blk, mvx, mvy, 0xFFFFFFFF)
                g = oracle.lib.orc_block_sad_generic(src, ref, *blk, mvx, mvy, 0xFFFFFFFF)
                assert f == g == table[mvy + 16, mvx + 16], (blk, mvx, mvy)


def test_block_sad_early_exit_is_monotone_bound(oracle):
    cy, ry, cu, _ = stress_pair()
    keep = []
    src = oracle.i420(*packed(W, H, cy, cu, cu), keep=keep)
    ref = oracle.i420(*packed(W, H, ry, cu, cu), keep=keep)
    full = oracle.lib.orc_block_sad(src, ref, 16, 16, 16, 16, 3, -2, 0xFFFFFFFF)
    for bound in (1, full // 3, full, full + 1):
        part = oracle.lib.orc_block_sad(src, ref, 16, 16, 16, 16, 3, -2, bound)
        assert part == full if full < bound else part >= bound


@pytest.mark.parametrize("scene", ["noise", "translated", "alien"])
def test_refine_search_matches_bruteforce_descent(oracle, scene):
    import ctypes as C
    cy, cu, cv = clean_scene()
    ref_y = {"noise": lambda: noisy(cy, 99, 8),
             "translated": lambda: noisy(translated(cy, 4), 7331, 5),
             "alien": lambda: (255 - cy).astype(np.uint8)}[scene]()
    keep = []
    src = oracle.i420(*packed(W, H, cy, cu, cv), keep=keep)
    ref = oracle.i420(*packed(W, H, ref_y, cu, cv), keep=keep)
    out = (C.c_int32 * 3)()
    for by in range(0, H, 16):
        for bx in range(0, W, 16):
            bw, bh = min(16, W - bx), min(16, H - by)
            oracle.lib.orc_refine_search(src, ref, bx, by, bw, bh, out)
            want = walk_descent(np_sad_table(cy, ref_y, W, H, bx, by, bw, bh))
            assert tuple(out) == want, (bx, by)


@pytest.mark.parametrize("seed", range(6))
def test_refine_search_tie_order_bruteforce(oracle, seed):
    """Low-entropy planes make equal SADs common, so the probe order and the
    strict '<' decide the winner."""
    import ctypes as C
    w, h = 64, 48
    rng = np.random.default_rng(100 + seed)
    cell = (2, 4, 8)[seed % 3]
    a = np.kron(rng.integers(0, 2, (h // cell + 1, w // cell + 1)), np.ones((cell, cell)))[:h, :w]
    b = np.kron(rng.integers(0, 2, (h // cell + 1, w // cell + 1)), np.ones((cell, cell)))[:h, :w]
    a = (a * 9).astype(np.uint8).reshape(-1)
    b = (b * 9).astype(np.uint8).reshape(-1)
    c = np.zeros((w // 2) * (h // 2), np.uint8)
    keep = []
    src = oracle.i420(*packed(w, h, a, c, c), keep=keep)
    ref = oracle.i420(*packed(w, h, b, c, c), keep=keep)
    out = (C.c_int32 * 3)()
    ties = 0
    for by in range(0, h, 16):
        for bx in range(0, w, 16):
            t = np_sad_table(a, b, w, h, bx, by, 16, 16)
            oracle.lib.orc_refine_search(src, ref, bx, by, 16, 16, out)
            assert tuple(out) == walk_descent(t), (bx, by)
            ties += descent_ties(t)
    assert ties > 0


PROBES8 = [(-8, -8), (0, -8), (8, -8), (-8, 0), (8, 0), (-8, 8), (0, 8), (8, 8)]


def test_refine_search_planted_tie_pairs(oracle):
    """For every pair of step-8 probes, plant an exact copy of the source
    block at both displacements: both reach SAD 0, so the one earlier in
    the probe order must win and stay (every later probe overlaps a copy
    only partially)."""
    import ctypes as C
    w = h = 64
    bx = by = 24
    c = np.zeros((w // 2) * (h // 2), np.uint8)
    src_y = np.zeros((h, w), np.uint8)
    src_y[by:by + 16, bx:bx + 16] = 100
    out = (C.c_int32 * 3)()
    for i in range(8):
        for j in range(i + 1, 8):
            ref_y = np.zeros((h, w), np.uint8)
            for dx, dy in (PROBES8[i], PROBES8[j]):
                ref_y[by + dy:by + dy + 16, bx + dx:bx + dx + 16] = 100
            keep = []
            src = oracle.i420(*packed(w, h, src_y.reshape(-1), c, c), keep=keep)
            ref = oracle.i420(*packed(w, h, ref_y.reshape(-1), c, c), keep=keep)
            oracle.lib.orc_refine_search(src, ref, bx, by, 16, 16, out)
            # opposite probes together cover the zero-MV block, which then
            # keeps its SAD-0 incumbency
            zero_covered = bool((ref_y[by:by + 16, bx:bx + 16] == 100).all())
            want = (0, 0, 0) if zero_covered else (*PROBES8[i], 0)
            assert tuple(out) == want, (i, j, tuple(out))


def test_refine_search_partial_blocks_bruteforce(oracle):
    """Odd geometry: partial right/bottom blocks clamp against visible dims."""
    import ctypes as C
    w, h = 37, 29
    rng = np.random.default_rng(5)
    a = rng.integers(0, 256, w * h, dtype=np.uint8)
    b = np.roll(a.reshape(h, w), (2, -3), (0, 1)).reshape(-1)
    c = np.zeros(((w + 1) // 2) * ((h + 1) // 2), np.uint8)
    keep = []
    src = oracle.i420(*packed(w, h, a, c, c), keep=keep)
    ref = oracle.i420(*packed(w, h, b, c, c), keep=keep)
    out = (C.c_int32 * 3)()
    for by in range(0, h, 16):
        for bx in range(0, w, 16):
            bw, bh = min(16, w - bx), min(16, h - by)
            oracle.lib.orc_refine_search(src, ref, bx, by, bw, bh, out)
            assert tuple(out) == walk_descent(np_sad_table(a, b, w, h, bx, by, bw, bh))


# --------------------------------------------------------------- accumulation
def _ref_accumulate(ctr, f, f_stride, pw, ph, blk, mv, S, acc, wsum, clamp_out):
    bx, by, bw, bh = blk
    for r in range(bh):
        yy = min(by + r, ph - 1) if clamp_out else by + r
        for c in range(bw):
            xx = min(bx + c, pw - 1) if clamp_out else bx + c
            cp = int(ctr[yy * pw + xx])
            rp = int(f[min(max(yy + mv[1], 0), ph - 1) * f_stride + min(max(xx + mv[0], 0), pw - 1)])
            wt = (16 * S) // (S + (rp - cp) ** 2)
            acc[yy * pw + xx] += wt * rp
            wsum[yy * pw + xx] += wt


def test_accumulate_luma_block_matches_reference_loop(oracle):
    import ctypes as C
    cy, ry, _, _ = stress_pair()
    S = 8 << 3
    t = (C.c_uint32 * 256)()
    oracle.lib.orc_weight_table(S, t)
    for blk in [(16, 16, 16, 16), (0, 0, 16, 16), (32, 32, 16, 16)]:
        for mvy in (-15, -7, -1, 0, 1, 7, 15):
            for mvx in (-15, -7, -1, 0, 1, 7, 15):
                a = np.full(W * H, 3, np.uint32)
                ws = np.full(W * H, 5, np.uint32)
                oracle.lib.orc_accumulate_luma_block(cy.ctypes.data, ry.ctypes.data, W, W, H, *blk,
                                                     mvx, mvy, t, a.ctypes.data, ws.ctypes.data)
                b = [3] * (W * H)
                wb = [5] * (W * H)
                _ref_accumulate(cy, ry, W, W, H, blk, (mvx, mvy), S, b, wb, False)
                assert a.tolist() == b and ws.tolist() == wb, (blk, mvx, mvy)


def test_accumulate_chroma_block_matches_reference_loop_odd_plane(oracle):
    import ctypes as C
    cw, ch = 25, 23
    r, c = np.meshgrid(np.arange(ch), np.arange(cw), indexing="ij")
    ctr = (100 + ((r * 5 + c * 3) & 0x3F)).astype(np.uint8).reshape(-1)
    refp = noisy(ctr, 555, 10)
    S = 8 << 3
    t = (C.c_uint32 * 256)()
    oracle.lib.orc_weight_table(S, t)
    for blk in [(0, 0, 8, 8), (8, 8, 8, 8), (24, 16, 8, 8), (16, 16, 8, 8)]:
        for my in (-7, -3, 0, 3, 7):
            for mx in (-7, -3, 0, 3, 7):
                a = np.full(cw * ch, 1, np.uint32)
                ws = np.full(cw * ch, 2, np.uint32)
                oracle.lib.orc_accumulate_chroma_block(ctr.ctypes.data, refp.ctypes.data, cw, cw, ch,
                                                       *blk, mx, my, t, a.ctypes.data, ws.ctypes.data)
                b = [1] * (cw * ch)
                wb = [2] * (cw * ch)
                _ref_accumulate(ctr, refp, cw, cw, ch, blk, (mx, my), S, b, wb, True)
                assert a.tolist() == b and ws.tolist() == wb, (blk, mx, my)


# --------------------------------------------------------------- whole build
def _frames_from_group(g):
    return g.frames()


@pytest.mark.parametrize("geom", [(48, 48, 5, 2, 3), (37, 29, 3, 0, 6), (17, 9, 4, 3, 1),
                                  (64, 40, 2, 1, 4), (1, 1, 3, 1, 3), (15, 33, 3, 2, 2)])
def test_build_matches_gather_formulation(oracle, geom):
    w, h, n, center, strength = geom
    rng = np.random.default_rng(w * 1000 + h)
    base = rng.integers(0, 256, (h, w)).astype(np.int64)
    cw, ch = (w + 1) // 2, (h + 1) // 2
    frames = []
    for i in range(n):
        y = np.clip(np.roll(base, (i - center, 2 * (i - center)), (0, 1)) + rng.integers(-6, 7, (h, w)), 0, 255)
        u = rng.integers(100, 140, (ch, cw))
        v = rng.integers(90, 150, (ch, cw))
        frames.append(packed(w, h, y.astype(np.uint8).reshape(-1), u.astype(np.uint8).reshape(-1),
                             v.astype(np.uint8).reshape(-1)))
    st, out, _, mvs = oracle.build_arnr_altref(frames, center, strength, want_mvs=True)
    assert st == 0
    want = np_blend(frames, center, strength, mvs)
    for got, exp in zip(out, want):
        np.testing.assert_array_equal(got, exp)
    # every searched entry is consistent with a brute-force SAD table
    for fi in range(n):
        if fi == center:
            continue
        for brow in range(mvs.shape[1]):
            for bcol in range(mvs.shape[2]):
                bx, by = bcol * 16, brow * 16
                bw, bh = min(16, w - bx), min(16, h - by)
                t = np_sad_table(frames[center][2], frames[fi][2], w, h, bx, by, bw, bh)
                assert tuple(mvs[fi, brow, bcol]) == walk_descent(t)


def test_strength_zero_and_single_frame_are_identity(oracle):
    y, u, v = clean_scene()
    st, out, _, _ = oracle.build_arnr_altref([packed(W, H, y, u, v)], 0, 3)
    assert st == 0 and all((a == b).all() for a, b in zip(out, (y, u, v)))
    ny = noisy(y, 7, 6)
    st, out, _, _ = oracle.build_arnr_altref([packed(W, H, ny, u, v), packed(W, H, y, u, v)], 0, 0)
    assert (out[0] == ny).all()


def test_static_noisy_window_denoises(oracle):
    cy, cu, cv = clean_scene()
    planes = [noisy(cy, 1000 + i, 6) for i in range(5)]
    st, out, _, _ = oracle.build_arnr_altref([packed(W, H, p, cu, cv) for p in planes], 2, 3)
    assert mse(out[0], cy) < mse(planes[2], cy) / 2


def test_translating_scene_denoises(oracle):
    cy, cu, cv = clean_scene()
    planes = [noisy(translated(cy, (i - 1) * 4), 2000 + i, 6) for i in range(3)]
    st, out, _, _ = oracle.build_arnr_altref([packed(W, H, p, cu, cv) for p in planes], 1, 3)
    assert mse(out[0], cy) < mse(planes[1], cy)


def test_scene_change_neighbour_is_excluded(oracle):
    cy, cu, cv = clean_scene()
    alien = (255 - cy).astype(np.uint8)
    st, out, _, _ = oracle.build_arnr_altref([packed(W, H, cy, cu, cv), packed(W, H, alien, cu, cv)], 0, 6)
    assert (out[0] == cy).all()


def test_dimension_errors_in_reference_order(oracle):
    cy, cu, cv = clean_scene()
    f0 = packed(W, H, cy, cu, cv)
    small = packed(32, 32, np.zeros(1024, np.uint8), np.full(256, 128, np.uint8), np.full(256, 128, np.uint8))
    st, _, err, _ = oracle.build_arnr_altref([f0, small], 0, 3)
    assert st == 2 and (err.src_w, err.src_h, err.ref_w, err.ref_h) == (48, 48, 32, 32)
    st, _, err, _ = oracle.build_arnr_altref([f0, small], 0, 0)  # checked before the strength-0 copy
    assert st == 2
    zero = (0, H, np.zeros(0, np.uint8), np.zeros(0, np.uint8), np.zeros(0, np.uint8), 0, 0)
    st, _, err, _ = oracle.build_arnr_altref([zero], 0, 3)
    assert st == 1 and (err.src_w, err.src_h) == (0, H)
    big = (0x4000, 1, np.zeros(0x4000, np.uint8), np.zeros(0x2000, np.uint8), np.zeros(0x2000, np.uint8), 0x4000, 0x2000)
    st, _, _, _ = oracle.build_arnr_altref([big, f0], 0, 3)
    assert st == 1  # centre dims are checked before any mismatch
    st, _, _, _ = oracle.build_arnr_altref([f0], 1, 3)
    assert st == 3


def test_fixed_point_every_strength(oracle):
    rng = np.random.default_rng(9)
    w, h = 41, 35
    y = rng.integers(0, 256, w * h).astype(np.uint8)
    u = rng.integers(0, 256, 21 * 18).astype(np.uint8)
    v = rng.integers(0, 256, 21 * 18).astype(np.uint8)
    for s in range(8):
        st, out, _, _ = oracle.build_arnr_altref([packed(w, h, y, u, v)] * 4, 1, s)
        assert all((a == b).all() for a, b in zip(out, (y, u, v)))


def test_reentry_and_thread_invariance(oracle):
    import vp8g.synth as synth
    g = synth.arnr_group_cif(96, 64, 5)
    st, out1, _, m1 = oracle.build_arnr_altref(g.frames(), g.center, 3, threads=1, want_mvs=True)
    st, out4, _, m4 = oracle.build_arnr_altref(g.frames(), g.center, 3, threads=4, want_mvs=True)
    assert all((a == b).all() for a, b in zip(out1, out4)) and (m1 == m4).all()
    st, echo, _, _ = oracle.build_arnr_altref([packed(96, 64, *out1)], 0, 0)
    assert all((a == b).all() for a, b in zip(out1, echo))


def test_strided_input_equals_packed(oracle):
    import vp8g.synth as synth
    g = synth.arnr_group_cif(50, 34, 3)
    frames = []
    for f in g.frames():
        w, h, y, u, v, _, _ = f
        cw, ch = (w + 1) // 2, (h + 1) // 2
        ys = np.zeros((h, w + 7), np.uint8); ys[:, :w] = y.reshape(h, w)
        us = np.zeros((ch, cw + 3), np.uint8); us[:, :cw] = u.reshape(ch, cw)
        vs = np.zeros((ch, cw + 3), np.uint8); vs[:, :cw] = v.reshape(ch, cw)
        frames.append((w, h, ys.reshape(-1), us.reshape(-1), vs.reshape(-1), w + 7, cw + 3))
    _, a, _, _ = oracle.build_arnr_altref(g.frames(), g.center, 4)
    _, b, _, _ = oracle.build_arnr_altref(frames, g.center, 4)
    assert all((x == y).all() for x, y in zip(a, b))


def test_cif_group_golden(oracle):
    """Regression anchor: the 352x288 five-frame group's output hashes,
    written by tests/golden/make_golden.py from this oracle."""
    import hashlib
    import vp8g.synth as synth
    gold = json.loads((GOLDEN / "arnr_cif.json").read_text())
    g = synth.arnr_group_cif()
    inp = hashlib.sha256(b"".join(np.concatenate([y.reshape(-1), u.reshape(-1), v.reshape(-1)]).tobytes()
                                  for y, u, v in zip(g.y, g.u, g.v))).hexdigest()
    assert inp == gold["input_sha256"]
    st, out, _, mvs = oracle.build_arnr_altref(g.frames(), g.center, 3, threads=4, want_mvs=True)
    assert hashlib.sha256(b"".join(o.tobytes() for o in out)).hexdigest() == gold["output_sha256"]
    assert hashlib.sha256(mvs.tobytes()).hexdigest() == gold["mvs_sha256"]
