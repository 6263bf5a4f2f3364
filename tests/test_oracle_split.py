"""Pins the oracle's SPLITMV prediction and inter-macroblock
reconstruction (src/motion_comp.rs:1693-1712, 2098-2150, 2162-2385;
src/inverse_transform.rs:170-240, 381-480, 609-678).

* inverse_wht_4x4 / inverse_dct_4x4: the reference's known answers (zero
  in, zero out; the hand-derived two-value WHT; DC 8 -> 1 and DC 64 -> 8;
  src/inverse_transform.rs:990-1089), and against the real-valued 2-D
  inverse DCT the listing approximates, x = T X T^t / 8 with
  T[n][0] = 1, T[n][k] = sqrt(2) cos((2n+1) k pi / 8), within 1.5
  everywhere (rounding plus the listing's truncated products; swapping the
  two constants gives errors above 100 on the same inputs);
  hostile full-range blocks wrap as 32-bit two's complement (checked
  against the same transform in Python integers reduced mod 2^32).
* The DC-only add path equals the general add (:417-480).
* split_chroma_mvs: hand values and the avg() closed form.
* predict_split_mv with 16 equal vectors equals predict_inter_mb (9x9
  sub-block halos against the 21x21 / 13x13 macroblock halos), both tap
  tables, both full_pixel settings, edge macroblocks.
* reconstruct_*: skip returns the prediction; a pure Y2 DC lifts luma
  uniformly and leaves chroma (src/motion_comp.rs:4061-4135)."""
import ctypes as C
import math

import numpy as np
import pytest


def lib(oracle):
    L = oracle.lib
    vp = C.c_void_p
    L.orc_inverse_wht_4x4.argtypes = [vp, vp]
    L.orc_inverse_dct_4x4.argtypes = [vp, vp]
    L.orc_inverse_dct_4x4_add_into.argtypes = [vp, vp, C.c_size_t, C.c_int, C.c_int]
    L.orc_split_chroma_mvs.argtypes = [vp, vp]
    return L


def wht(oracle, x):
    x = np.ascontiguousarray(x, np.int16)
    o = np.zeros(16, np.int16)
    lib(oracle).orc_inverse_wht_4x4(x.ctypes.data, o.ctypes.data)
    return o


def idct(oracle, x):
    x = np.ascontiguousarray(x, np.int16)
    o = np.zeros(16, np.int16)
    lib(oracle).orc_inverse_dct_4x4(x.ctypes.data, o.ctypes.data)
    return o


def test_wht_known_answers(oracle):
    assert (wht(oracle, np.zeros(16)) == 0).all()
    x = np.zeros(16, np.int16)
    x[0] = x[4] = 8
    assert wht(oracle, x).tolist() == [2, 2, 2, 2, 2, 2, 2, 2, 0, 0, 0, 0, 0, 0, 0, 0]


def test_idct_known_answers(oracle):
    assert (idct(oracle, np.zeros(16)) == 0).all()
    for dc, v in ((8, 1), (64, 8)):
        x = np.zeros(16, np.int16)
        x[0] = dc
        assert (idct(oracle, x) == v).all()


def float_idct(x):
    k = np.arange(4)
    T = np.sqrt(2.0) * np.cos((2 * k[:, None] + 1) * k[None, :] * np.pi / 8)
    T[:, 0] = 1.0
    X = np.asarray(x, np.float64).reshape(4, 4)
    return (T @ X @ T.T / 8.0).reshape(16)


def test_idct_tracks_the_real_transform(oracle):
    rng = np.random.default_rng(4)
    worst = 0.0
    for _ in range(3000):
        x = rng.integers(-600, 601, 16) * (rng.random(16) < 0.6)
        worst = max(worst, float(np.abs(idct(oracle, x) - float_idct(x)).max()))
    assert worst <= 1.5


def test_idct_wraps_on_hostile_blocks(oracle):
    def wrap(v):
        v &= 0xFFFFFFFF
        return v - (1 << 32) if v >= (1 << 31) else v

    def one(i0, i4, i8, i12):
        a1, b1 = wrap(i0 + i8), wrap(i0 - i8)
        c1 = wrap((wrap(i4 * 35468) >> 16) - wrap(i12 + (wrap(i12 * 20091) >> 16)))
        d1 = wrap(wrap(i4 + (wrap(i4 * 20091) >> 16)) + (wrap(i12 * 35468) >> 16))
        return [wrap(a1 + d1), wrap(b1 + c1), wrap(b1 - c1), wrap(a1 - d1)]

    def ref(x):
        t = [0] * 16
        for c in range(4):
            o = one(int(x[c]), int(x[4 + c]), int(x[8 + c]), int(x[12 + c]))
            for k in range(4):
                t[4 * k + c] = o[k]
        out = []
        for r in range(4):
            o = one(*t[4 * r:4 * r + 4])
            out += [np.int16(np.int32(wrap(v + 4) >> 3).astype(np.int16)) for v in o]
        return np.array(out, np.int16)

    hostile = [np.full(16, 32767), np.full(16, -32768), np.array([32767, -32768] * 8),
               np.array([-32768, 32767, -1, 12345, -30000, 32767, -32768, 9999, 32767, -32768, 32767, -32768,
                         1, -2, 3, -4])]
    rng = np.random.default_rng(8)
    hostile += [rng.integers(-32768, 32768, 16) for _ in range(40)]
    for x in hostile:
        assert np.array_equal(idct(oracle, x), ref(x)), x


def test_dc_only_add_equals_general_add(oracle):
    L = lib(oracle)
    rng = np.random.default_rng(2)
    for dc in (-4000, -37, -4, -3, 0, 3, 4, 5, 11, 100, 2047, 4000):
        plane = rng.integers(0, 256, (8, 8)).astype(np.uint8)
        a = plane.copy()
        x = np.zeros(16, np.int16)
        x[0] = dc
        L.orc_inverse_dct_4x4_add_into(x.ctypes.data, a.ctypes.data, 8, 1, 1)
        want = plane.copy()
        want[4:8, 4:8] = np.clip(want[4:8, 4:8].astype(np.int32) + idct(oracle, x).reshape(4, 4), 0, 255)
        assert np.array_equal(a, want), dc


def chroma_mvs(oracle, mvs):
    m = np.ascontiguousarray(mvs, np.int16).reshape(16, 2)
    o = np.zeros((4, 2), np.int16)
    lib(oracle).orc_split_chroma_mvs(m.ctypes.data, o.ctypes.data)
    return o


def test_split_chroma_mvs(oracle):
    def avg_closed(s):
        return (s + 4) >> 3 if s >= 0 else -((-s + 4) >> 3)

    m = np.zeros((16, 2), np.int16)
    for k, b in enumerate((0, 1, 4, 5)):          # chroma block 0's luma blocks
        m[b] = (k + 1, -(k + 1))
    got = chroma_mvs(oracle, m)
    assert tuple(got[0]) == (3, -3) and (got[1:] == 0).all()  # sum 2 * 10 = 20 -> (20 + 4) >> 3
    rng = np.random.default_rng(6)
    for _ in range(200):
        m = rng.integers(-1023, 1024, (16, 2)).astype(np.int16)
        got = chroma_mvs(oracle, m)
        for c in range(4):
            bs = [b for b in range(16) if ((b >> 1) & 1) + ((b >> 2) & 2) == c]
            assert tuple(got[c]) == tuple(avg_closed(int(sum(2 * int(m[b, i]) for b in bs))) for i in (0, 1))
    for v in range(-40, 41):   # four equal vectors: the chroma vector is the luma value in eighth pels
        assert (chroma_mvs(oracle, np.full((16, 2), v)) == v).all()


def build_reference(mb_cols, mb_rows):
    """src/motion_comp.rs:2716-2725."""
    lw, lh, cw, ch = 16 * mb_cols, 16 * mb_rows, 8 * mb_cols, 8 * mb_rows
    y = (np.arange(lw * lh) % 251).astype(np.uint8)
    u = ((np.arange(cw * ch) * 3 + 1) % 251).astype(np.uint8)
    v = ((np.arange(cw * ch) * 7 + 2) % 251).astype(np.uint8)
    return y, u, v


def ref_planes(y, u, v, mb_cols, mb_rows):
    from oracle_ctypes import RefPlanes
    return RefPlanes(y.ctypes.data, u.ctypes.data, v.ctypes.data, 16 * mb_cols, 8 * mb_cols, mb_cols, mb_rows)


def test_split_with_equal_vectors_equals_predict_inter_mb(oracle):
    from oracle_ctypes import MvS
    rng = np.random.default_rng(11)
    y = rng.integers(0, 256, 64 * 48).astype(np.uint8)
    u = rng.integers(0, 256, 32 * 24).astype(np.uint8)
    v = rng.integers(0, 256, 32 * 24).astype(np.uint8)
    rp = ref_planes(y, u, v, 4, 3)
    for filt in (oracle.SIXTAP, oracle.BILINEAR):
        f = np.ascontiguousarray(filt)
        for fp in (0, 1):
            for _ in range(60):
                mv = rng.integers(-90, 91, 2).astype(np.int16)
                mbc, mbr = int(rng.integers(0, 4)), int(rng.integers(0, 3))
                a = np.zeros(384, np.uint8)
                b = np.zeros(384, np.uint8)
                mvs = np.ascontiguousarray(np.tile(mv, (16, 1)))
                oracle.lib.orc_predict_split_mv(C.byref(rp), mbc, mbr, mvs.ctypes.data, fp, f.ctypes.data,
                                                a.ctypes.data)
                oracle.lib.orc_predict_inter_mb(C.byref(rp), mbc, mbr, MvS(int(mv[0]), int(mv[1])), fp,
                                                f.ctypes.data, b.ctypes.data)
                assert np.array_equal(a, b), (mv, mbc, mbr, fp)


def test_reconstruct_skip_and_dc_lift(oracle):
    from oracle_ctypes import MvS
    y, u, v = build_reference(3, 3)
    rp = ref_planes(y, u, v, 3, 3)
    f = np.ascontiguousarray(oracle.SIXTAP)
    pred = np.zeros(384, np.uint8)
    oracle.lib.orc_predict_inter_mb(C.byref(rp), 1, 1, MvS(3, 5), 0, f.ctypes.data, pred.ctypes.data)
    coeffs = np.zeros((25, 16), np.int16)
    coeffs[0, 0] = 64
    out = np.zeros(384, np.uint8)
    oracle.lib.orc_reconstruct_inter_mb(C.byref(rp), 1, 1, MvS(3, 5), 0, f.ctypes.data, 1, coeffs.ctypes.data,
                                        out.ctypes.data)
    assert np.array_equal(out, pred)
    oracle.lib.orc_reconstruct_inter_mb(C.byref(rp), 1, 1, MvS(3, 5), 0, f.ctypes.data, 0, coeffs.ctypes.data,
                                        out.ctypes.data)
    dc = np.zeros(16, np.int16)
    dc[0] = wht(oracle, coeffs[0])[0]
    delta = int(idct(oracle, dc)[0])
    assert delta != 0
    assert np.array_equal(out[256:], pred[256:])
    assert np.array_equal(out[:256], np.clip(pred[:256].astype(np.int32) + delta, 0, 255).astype(np.uint8))
    # SPLITMV: every Y block carries its own DC (no Y2)
    mvs = np.ascontiguousarray(np.tile(np.array([3, 5], np.int16), (16, 1)))
    sc = np.zeros((24, 16), np.int16)
    sc[5, 0] = 80
    oracle.lib.orc_reconstruct_split_mv_mb(C.byref(rp), 1, 1, mvs.ctypes.data, 0, f.ctypes.data, 0, sc.ctypes.data,
                                           out.ctypes.data)
    want = pred.copy().reshape(-1)
    blk = want[:256].reshape(16, 16)
    blk[4:8, 4:8] = np.clip(blk[4:8, 4:8].astype(np.int32) + (80 + 4) // 8, 0, 255)
    assert np.array_equal(out, want)
