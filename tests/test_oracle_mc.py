"""Pins the motion-compensation / SAD half of the CPU oracle.

* tables: RFC 6386 section 18.3 values typed from the RFC, taps sum to 128
  (src/motion_comp.rs:3025-3057);
* interp: the known answers 30 / 35 / 200 / 0 (src/motion_comp.rs:3077-3098);
* sixtap_2d: == a literal transcription of the spec's Hinterp/Vinterp over
  a strided buffer, all 64 fractions x both sets (src/motion_comp.rs:3100-3186),
  DC pass, whole-fraction copy, accumulator extremes;
* sixtap_mb_luma / sixtap_mb_chroma: == assembly of per-4x4 sixtap_2d
  results over sub-halos cut from the MB halo (src/motion_comp.rs:3330-3381, 3637);
* fetches: clamped corners, MB halo contains every sub-block halo,
  whole-pixel fetch == halo core (src/motion_comp.rs:3420-3475, 4187-4286);
* MV adjust: stored_luma_mv, avg, the closed form (c+1+(c>>31)*2)/2,
  full-pixel truncation (src/motion_comp.rs:2538-2616);
* predict_inter_mb: == per-4x4-sub-block assembly (src/motion_comp.rs:3950);
* SAD: known sums, whole-MV SAD == numpy clamped fetch SAD, half-pel shift
  found exactly, never worsens, clamps at +-1023
  (src/motion_search.rs:918-965, 1200-1340);
* HalfPlanes: each of its 8 SADs == mb_luma_sad_at_mv at that half-pel MV.
"""
import ctypes as C

import numpy as np
import pytest

SPEC_SIXTAP = [[0, 0, 128, 0, 0, 0], [0, -6, 123, 12, -1, 0], [2, -11, 108, 36, -8, 1],
               [0, -9, 93, 50, -6, 0], [3, -16, 77, 77, -16, 3], [0, -6, 50, 93, -9, 0],
               [1, -8, 36, 108, -11, 2], [0, -1, 12, 123, -6, 0]]
SPEC_BILINEAR = [[0, 0, 128 - 16 * k, 16 * k, 0, 0] for k in range(8)]


def mv(row, col):
    from oracle_ctypes import MvS
    return MvS(row, col)


def rand_halo(seed, n):
    """The reference tests' halo LCG (src/motion_comp.rs:3157-3167)."""
    s = seed & 0xFFFFFFFFFFFFFFFF
    out = np.empty(n, np.uint8)
    for i in range(n):
        s = (s * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        out[i] = (s >> 33) & 0xFF
    return out


def ref_interp(fil, p, origin, s):
    """Literal transcription of the section 18.3 interp prose."""
    a, idx = 0, origin - 2 * s
    for tap in fil:
        a += int(p[idx]) * tap
        idx += s
    return min(255, max(0, (a + 64) >> 7))


def ref_two_pass(halo, ow, mx, my, filters):
    hw = ow + 5
    temp = [ref_interp(filters[mx], halo, r * hw + 2 + c, 1) for r in range(hw) for c in range(ow)]
    return np.array([ref_interp(filters[my], temp, (2 + r) * ow + c, ow)
                     for r in range(ow) for c in range(ow)], np.uint8)


def sixtap(oracle, fn, halo, ow, mx, my, filt):
    out = np.zeros(ow * ow, np.uint8)
    f = np.ascontiguousarray(filt, np.int32)
    getattr(oracle.lib, fn)(halo.ctypes.data, mx, my, f.ctypes.data, out.ctypes.data)
    return out


def test_tables_match_spec(oracle):
    assert oracle.SIXTAP.tolist() == SPEC_SIXTAP
    assert oracle.BILINEAR.tolist() == SPEC_BILINEAR
    assert all(sum(r) == 128 for r in SPEC_SIXTAP + SPEC_BILINEAR)


def test_interp_known_answers(oracle):
    def interp(f, s):
        f = np.array(f, np.int32)
        s = np.array(s, np.uint8)
        return oracle.lib.orc_interp(f.ctypes.data, s.ctypes.data)
    sup = [10, 20, 30, 40, 50, 60]
    assert interp(SPEC_SIXTAP[0], sup) == 30
    assert interp(SPEC_SIXTAP[4], sup) == 35  # (4480 + 64) >> 7
    for f in SPEC_SIXTAP:
        assert interp(f, [200] * 6) == 200
    assert interp(SPEC_SIXTAP[2], [0, 255, 0, 0, 255, 0]) == 0  # -4781 >> 7 clamps


@pytest.mark.parametrize("table", ["sixtap", "bilinear"])
def test_sixtap_2d_matches_spec_transcription(oracle, table):
    filt = SPEC_SIXTAP if table == "sixtap" else SPEC_BILINEAR
    for seed in range(3):
        halo = rand_halo(seed * 977 + 13, 81)
        for mx in range(8):
            for my in range(8):
                got = sixtap(oracle, "orc_sixtap_2d", halo, 4, mx, my, filt)
                assert (got == ref_two_pass(halo, 4, mx, my, filt)).all(), (seed, mx, my)


def test_sixtap_flat_whole_and_extremes(oracle):
    flat = np.full(81, 137, np.uint8)
    for mx in range(8):
        for my in range(8):
            for f in (SPEC_SIXTAP, SPEC_BILINEAR):
                assert (sixtap(oracle, "orc_sixtap_2d", flat, 4, mx, my, f) == 137).all()
    halo = rand_halo(42, 81)
    out = sixtap(oracle, "orc_sixtap_2d", halo, 4, 0, 0, SPEC_SIXTAP)
    assert (out.reshape(4, 4) == halo.reshape(9, 9)[2:6, 2:6]).all()
    pos = np.array([[255 if c % 6 in (0, 2, 3, 5) else 0 for c in range(9)] for _ in range(9)], np.uint8)
    assert sixtap(oracle, "orc_sixtap_2d", pos.reshape(-1), 4, 4, 4, SPEC_SIXTAP)[0] == 255
    assert sixtap(oracle, "orc_sixtap_2d", (255 - pos).reshape(-1), 4, 4, 4, SPEC_SIXTAP)[0] == 0


@pytest.mark.parametrize("ow,fn", [(16, "orc_sixtap_mb_luma"), (8, "orc_sixtap_mb_chroma")])
def test_mb_scale_equals_subblock_assembly(oracle, ow, fn):
    hw = ow + 5
    for seed in range(2):
        halo = rand_halo(seed * 31 + 5, hw * hw).reshape(hw, hw)
        for f in (SPEC_SIXTAP, SPEC_BILINEAR):
            for mx in range(8):
                for my in range(8):
                    got = sixtap(oracle, fn, np.ascontiguousarray(halo).reshape(-1), ow, mx, my, f).reshape(ow, ow)
                    for sr in range(ow // 4):
                        for sc in range(ow // 4):
                            sub = np.ascontiguousarray(halo[4 * sr:4 * sr + 9, 4 * sc:4 * sc + 9]).reshape(-1)
                            want = sixtap(oracle, "orc_sixtap_2d", sub, 4, mx, my, f).reshape(4, 4)
                            assert (got[4 * sr:4 * sr + 4, 4 * sc:4 * sc + 4] == want).all()


def test_fetches_clamp_and_nest(oracle):
    w, h = 40, 24
    plane = (np.arange(w * h) * 7 % 251).astype(np.uint8)
    P = plane.reshape(h, w)

    def clamped(x0, y0, rows, cols):
        ys = np.clip(np.arange(y0, y0 + rows), 0, h - 1)
        xs = np.clip(np.arange(x0, x0 + cols), 0, w - 1)
        return P[ys][:, xs]

    for (mbx, mby, r8, c8) in [(0, 0, -24, -40), (16, 0, 5, 3), (32, 16, 40, 17), (16, 8, -3, 100)]:
        m = mv(r8, c8)
        halo = np.zeros(441, np.uint8)
        oracle.lib.orc_fetch_luma_mb_halo(plane.ctypes.data, w, w, h, mbx, mby, m, halo.ctypes.data)
        H = halo.reshape(21, 21)
        assert (H == clamped(mbx + (c8 >> 3) - 2, mby + (r8 >> 3) - 2, 21, 21)).all()
        whole = np.zeros(256, np.uint8)
        oracle.lib.orc_fetch_luma_mb_whole_pixel(plane.ctypes.data, w, w, h, mbx, mby, m, whole.ctypes.data)
        assert (whole.reshape(16, 16) == H[2:18, 2:18]).all()
        for sr in range(4):
            for sc in range(4):
                sub = np.zeros(81, np.uint8)
                oracle.lib.orc_fetch_block_halo(plane.ctypes.data, w, w, h, mbx + 4 * sc, mby + 4 * sr, m, sub.ctypes.data)
                assert (sub.reshape(9, 9) == H[4 * sr:4 * sr + 9, 4 * sc:4 * sc + 9]).all()
        ch = np.zeros(169, np.uint8)
        oracle.lib.orc_fetch_chroma_mb_halo(plane.ctypes.data, w, w, h, mbx // 2, mby // 2, m, ch.ctypes.data)
        assert (ch.reshape(13, 13) == clamped(mbx // 2 + (c8 >> 3) - 2, mby // 2 + (r8 >> 3) - 2, 13, 13)).all()
        cw8 = np.zeros(64, np.uint8)
        oracle.lib.orc_fetch_chroma_mb_whole_pixel(plane.ctypes.data, w, w, h, mbx // 2, mby // 2, m, cw8.ctypes.data)
        assert (cw8.reshape(8, 8) == ch.reshape(13, 13)[2:10, 2:10]).all()


def test_mv_adjust_closed_forms(oracle):
    s = oracle.lib.orc_stored_luma_mv(mv(3, -5))
    assert (s.row, s.col) == (6, -10)
    s = oracle.lib.orc_stored_luma_mv(mv(1023, -1023))
    assert (s.row, s.col) == (2046, -2046)
    s = oracle.lib.orc_stored_luma_mv(mv(20000, -20000))  # wrapping i16
    assert (s.row, s.col) == (40000 - 65536, -40000 + 65536)
    assert oracle.lib.orc_avg(8, 8, 8, 8) == 4 and oracle.lib.orc_avg(-1, -1, -1, -1) == -1
    assert oracle.lib.orc_avg(5, -5, 3, -3) == 0
    for c in [0, 2, 6, 10, 14, 100, -2, -6, -10, -14, -100, 2046, -2046, 7, -7, 1, -1]:
        m = oracle.lib.orc_chroma_mv(mv(c, c))
        closed = int((c + 1 + (-1 if c < 0 else 0) * 2) / 2)  # (c + 1 + (c>>31)*2) / 2, C division
        assert m.row == closed and m.col == closed, c
    f = oracle.lib.orc_apply_full_pixel(mv(13, 9))
    assert (f.row, f.col) == (8, 8)
    f = oracle.lib.orc_apply_full_pixel(mv(-1, -9))
    assert (f.row, f.col) == (-8, -16)


def _ref_planes(oracle, mb_cols, mb_rows, seed):
    from oracle_ctypes import RefPlanes
    rng = np.random.default_rng(seed)
    y = rng.integers(0, 256, mb_cols * 16 * mb_rows * 16).astype(np.uint8)
    u = rng.integers(0, 256, mb_cols * 8 * mb_rows * 8).astype(np.uint8)
    v = rng.integers(0, 256, mb_cols * 8 * mb_rows * 8).astype(np.uint8)
    rp = RefPlanes(y.ctypes.data, u.ctypes.data, v.ctypes.data, mb_cols * 16, mb_cols * 8, mb_cols, mb_rows)
    return rp, (y, u, v)


def test_predict_inter_mb_equals_subblock_assembly(oracle):
    mb_cols, mb_rows = 4, 3
    rp, (y, u, v) = _ref_planes(oracle, mb_cols, mb_rows, 1)
    lw, lh, cw, ch = 64, 48, 32, 24
    rng = np.random.default_rng(2)
    for f in (SPEC_SIXTAP, SPEC_BILINEAR):
        fa = np.ascontiguousarray(f, np.int32)
        for trial in range(40):
            mc, mr = int(rng.integers(0, mb_cols)), int(rng.integers(0, mb_rows))
            qr, qc = (int(x) for x in rng.integers(-90, 91, 2))
            full = trial % 5 == 4
            out = np.zeros(384, np.uint8)
            oracle.lib.orc_predict_inter_mb(C.byref(rp), mc, mr, mv(qr, qc), int(full), fa.ctypes.data, out.ctypes.data)
            ymv = (2 * qr, 2 * qc)
            uv = tuple(int((c + 1 + (-2 if c < 0 else 0)) / 2) for c in ymv)
            if full:
                ymv = tuple(c & ~7 for c in ymv)
                uv = tuple(c & ~7 for c in uv)
            for plane, size, w_, h_, base, ox, m8 in ((y, 16, lw, lh, 0, mc * 16, ymv),
                                                      (u, 8, cw, ch, 256, mc * 8, uv),
                                                      (v, 8, cw, ch, 320, mc * 8, uv)):
                oy = mr * size
                blk = out[base:base + size * size].reshape(size, size)
                for sr in range(size // 4):
                    for sc in range(size // 4):
                        halo = np.zeros(81, np.uint8)
                        oracle.lib.orc_fetch_block_halo(plane.ctypes.data, w_, w_, h_, ox + 4 * sc, oy + 4 * sr,
                                                        mv(m8[0], m8[1]), halo.ctypes.data)
                        want = ref_two_pass(halo, 4, m8[1] & 7, m8[0] & 7, f).reshape(4, 4)
                        assert (blk[4 * sr:4 * sr + 4, 4 * sc:4 * sc + 4] == want).all(), (trial, base)


def test_block_sad_known_sums(oracle):
    a = (np.arange(256) % 251).astype(np.uint8)
    sad = lambda x, y: oracle.lib.orc_block_sad_16x16(x.ctypes.data, y.ctypes.data)
    assert sad(a, a) == 0
    b = np.full(256, 10, np.uint8); c = b.copy(); c[42] = 11
    assert sad(b, c) == 1
    assert sad(np.zeros(256, np.uint8), np.full(256, 255, np.uint8)) == 65280
    x = np.zeros(256, np.uint8); y = np.zeros(256, np.uint8)
    x[:3] = [10, 20, 30]; y[:3] = [3, 5, 2]
    assert sad(x, y) == 50


def _luma_ref(plane, w, h, stride=None):
    from oracle_ctypes import LumaRef
    return LumaRef(plane.ctypes.data, stride or w, w, h)


def test_sad_at_mv_matches_numpy_whole_and_subpel(oracle):
    w, h = 80, 48
    rng = np.random.default_rng(11)
    plane = rng.integers(0, 256, w * h).astype(np.uint8)
    P = plane.reshape(h, w)
    lr = _luma_ref(plane, w, h)
    src = rng.integers(0, 256, 256).astype(np.uint8)
    for mb in [(0, 0), (4, 2), (2, 1)]:
        for q in [(0, 0), (-40, 12), (64, -64), (4, 8), (-1023, 1023)]:
            ys = np.clip(np.arange(16) + mb[1] * 16 + (2 * q[0] >> 3), 0, h - 1)
            xs = np.clip(np.arange(16) + mb[0] * 16 + (2 * q[1] >> 3), 0, w - 1)
            if (2 * q[0]) & 7 == 0 and (2 * q[1]) & 7 == 0:
                want = int(np.abs(src.reshape(16, 16).astype(int) - P[ys][:, xs]).sum())
                assert oracle.lib.orc_mb_luma_sad_at_whole_mv(C.byref(lr), mb[0], mb[1], src.ctypes.data, mv(*q)) == want
                assert oracle.lib.orc_mb_luma_sad_at_mv(C.byref(lr), mb[0], mb[1], src.ctypes.data, mv(*q)) == want
            # sub-pel: the SAD of the six-tap prediction (spec transcription)
            r8, c8 = 2 * q[0], 2 * q[1]
            hy = np.clip(np.arange(21) + mb[1] * 16 + (r8 >> 3) - 2, 0, h - 1)
            hx = np.clip(np.arange(21) + mb[0] * 16 + (c8 >> 3) - 2, 0, w - 1)
            pred = ref_two_pass(np.ascontiguousarray(P[hy][:, hx]).reshape(-1), 16, c8 & 7, r8 & 7, SPEC_SIXTAP)
            want = int(np.abs(src.astype(int) - pred.astype(int)).sum())
            assert oracle.lib.orc_mb_luma_sad_at_mv(C.byref(lr), mb[0], mb[1], src.ctypes.data, mv(*q)) == want


def two_d_ramp(w, h):
    r, c = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    return np.clip(c + 3 * r, 0, 240).astype(np.uint8).reshape(-1)


def test_half_pixel_refine_finds_half_shift_and_keeps_flat_center(oracle):
    w = h = 64
    plane = two_d_ramp(w, h)
    lr = _luma_ref(plane, w, h)
    # source = six-tap prediction at (0, +half pel) -> SAD 0 only there
    P = plane.reshape(h, w)
    hy = np.clip(np.arange(21) + 16 - 2, 0, h - 1)
    hx = np.clip(np.arange(21) + 16 - 2, 0, w - 1)
    src = ref_two_pass(np.ascontiguousarray(P[hy][:, hx]).reshape(-1), 16, 4, 0, SPEC_SIXTAP)
    r = oracle.lib.orc_half_pixel_refine_luma(C.byref(lr), 1, 1, src.ctypes.data, mv(0, 0))
    assert (r.mv.row, r.mv.col, r.sad) == (0, 2, 0)
    flat = np.full(32 * 32, 200, np.uint8)
    r = oracle.lib.orc_half_pixel_refine_luma(C.byref(_luma_ref(flat, 32, 32)), 0, 0,
                                              np.full(256, 200, np.uint8).ctypes.data, mv(0, 0))
    assert (r.mv.row, r.mv.col, r.sad) == (0, 0, 0)


@pytest.mark.parametrize("ladder,step", [("orc_half_pixel_refine_luma", 2),
                                         ("orc_quarter_pixel_refine_luma", 1)])
def test_refine_ladder_tie_order(oracle, ladder, step):
    """Rows constant along x: every column offset predicts the same block,
    so the three candidates of the winning row tie and the first in the
    (row, col) probe order must win."""
    w, h = 48, 48
    rows = (np.arange(h) * 37 % 200).astype(np.uint8)
    plane = np.repeat(rows[:, None], w, 1).reshape(-1)
    lr = _luma_ref(plane, w, h)
    for dr in (-step, step):
        P = plane.reshape(h, w)
        r8 = 2 * dr
        hy = np.clip(np.arange(21) + 16 + (r8 >> 3) - 2, 0, h - 1)
        hx = np.clip(np.arange(21) + 16 - 2, 0, w - 1)
        src = ref_two_pass(np.ascontiguousarray(P[hy][:, hx]).reshape(-1), 16, 0, r8 & 7, SPEC_SIXTAP)
        r = getattr(oracle.lib, ladder)(C.byref(lr), 1, 1, src.ctypes.data, mv(0, 0))
        assert (r.mv.row, r.mv.col, r.sad) == (dr, -step, 0)


def test_refine_ladders_never_worsen_and_clamp(oracle):
    w, h = 48, 48
    rng = np.random.default_rng(3)
    plane = rng.integers(0, 256, w * h).astype(np.uint8)
    lr = _luma_ref(plane, w, h)
    for trial in range(30):
        src = rng.integers(0, 256, 256).astype(np.uint8)
        c = mv(int(rng.integers(-20, 21)) * 4, int(rng.integers(-20, 21)) * 4)
        base = oracle.lib.orc_mb_luma_sad_at_mv(C.byref(lr), 1, 1, src.ctypes.data, c)
        hres = oracle.lib.orc_half_pixel_refine_luma(C.byref(lr), 1, 1, src.ctypes.data, c)
        assert hres.sad <= base and abs(hres.mv.row - c.row) <= 2 and abs(hres.mv.col - c.col) <= 2
        q = oracle.lib.orc_quarter_pixel_refine_luma(C.byref(lr), 1, 1, src.ctypes.data, hres.mv)
        assert q.sad <= hres.sad
        assert q.sad == oracle.lib.orc_mb_luma_sad_at_mv(C.byref(lr), 1, 1, src.ctypes.data, q.mv)
    src = rng.integers(0, 256, 256).astype(np.uint8)
    r = oracle.lib.orc_half_pixel_refine_luma(C.byref(lr), 1, 1, src.ctypes.data, mv(1020, 1020))
    assert -1023 <= r.mv.row <= 1023 and -1023 <= r.mv.col <= 1023


def test_half_planes_equal_sad_at_half_pel(oracle):
    from oracle_ctypes import HalfPlanes
    w, h = 64, 48
    rng = np.random.default_rng(21)
    plane = rng.integers(0, 256, w * h).astype(np.uint8)
    lr = _luma_ref(plane, w, h)
    taps = np.ascontiguousarray(SPEC_SIXTAP[4], np.int32)
    for (mbx, mby, wr, wc) in [(1, 1, 0, 0), (0, 0, -3, 5), (3, 2, 2, -1), (3, 2, 40, 40)]:
        src = rng.integers(0, 256, 256).astype(np.uint8)
        x, y = mbx * 16 + wc, mby * 16 + wr
        win = np.zeros(22 * 22, np.uint8)
        oracle.lib.orc_fetch_window(C.byref(lr), x, y, win.ctypes.data)
        hp = HalfPlanes()
        oracle.lib.orc_half_planes_new(win.ctypes.data, taps.ctypes.data, 1, C.byref(hp))
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if dx == 0 and dy == 0:
                    continue
                got = oracle.lib.orc_half_planes_sad(C.byref(hp), src.ctypes.data, dx, dy)
                want = oracle.lib.orc_mb_luma_sad_at_mv(C.byref(lr), mbx, mby, src.ctypes.data,
                                                        mv(4 * wr + 2 * dy, 4 * wc + 2 * dx))
                assert got == want, (mbx, mby, wr, wc, dx, dy)
