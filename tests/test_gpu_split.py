"""GPU parity for SPLITMV prediction and the inter-macroblock
reconstructions (predict_split_mv, reconstruct_inter_mb,
reconstruct_split_mv_mb; src/motion_comp.rs:2098-2385) against the oracle,
bit-exact, through the C ABI: every macroblock of CIF and 1080p reference
frames (edge macroblocks included), both tap tables and a caller table,
full_pixel on and off, sparse realistic and full-range hostile
coefficients (the inverse DCT's wrapping arithmetic), skip flags."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def planes(w, h, seed):
    import vp8g.synth as synth
    g = synth.arnr_group_cif(w, h, 3, clip=seed)
    y = synth.pad_to_mb(g.y[1], 16)
    u = synth.pad_to_mb(g.u[1], 8)
    v = synth.pad_to_mb(g.v[1], 8)
    return np.ascontiguousarray(y), np.ascontiguousarray(u), np.ascontiguousarray(v)


def split_vectors(rng, n):
    """Sixteen vectors per macroblock mixing whole-pel and sub-pel fractions
    (the pattern of src/motion_comp.rs:4485-4493) plus random ones."""
    m = np.zeros((n, 16, 2), np.int16)
    for i in range(16):
        fr = 0 if i % 2 == 0 else (i % 7) + 1
        fc = 0 if i % 3 == 0 else (i % 5) + 1
        m[:, i] = (((i % 3) - 1) * 8 + fr, ((i % 4) - 2) * 8 + fc)
    rnd = rng.random(n) < 0.6
    m[rnd] = rng.integers(-64, 65, (int(rnd.sum()), 16, 2))
    return m


def coeffs(rng, n, blocks, hostile_frac=0.1):
    c = (rng.integers(-60, 61, (n, blocks, 16)) * (rng.random((n, blocks, 16)) < 0.15)).astype(np.int16)
    c[:, :, 0] = rng.integers(-400, 401, (n, blocks))
    hostile = rng.random(n) < hostile_frac
    c[hostile] = rng.integers(-32768, 32768, (int(hostile.sum()), blocks, 16))
    return c


def oracle_rp(y, u, v, mb_cols, mb_rows):
    from oracle_ctypes import RefPlanes
    return RefPlanes(y.ctypes.data, u.ctypes.data, v.ctypes.data, 16 * mb_cols, 8 * mb_cols, mb_cols, mb_rows)


def mb_grid(mb_cols, mb_rows):
    return np.stack(np.meshgrid(np.arange(mb_cols), np.arange(mb_rows)), -1).reshape(-1, 2).astype(np.uint32)


def assert_mbs(got, want, what):
    if not np.array_equal(got, want):
        bad = np.nonzero((got != want).any(1))[0]
        raise AssertionError(f"{what}: {len(bad)} of {len(got)} macroblocks differ, first {bad[:6]}")


CUSTOM = np.array([[0, 0, 128, 0, 0, 0], [5, -20, 120, 30, -9, 2], [1, -12, 101, 44, -6, 0],
                   [-3, 7, 90, 40, -6, 0], [2, -16, 78, 78, -16, 2], [0, -6, 50, 93, -9, 0],
                   [9, -30, 60, 100, -15, 4], [0, -1, 12, 123, -6, 0]], np.int32)


@pytest.mark.parametrize("size", [(352, 288), (1920, 1080)], ids=["cif", "1080p"])
def test_predict_split_mv_matches_oracle(gpu, oracle, size):
    import vp8g
    w, h = size
    y, u, v = planes(w, h, 3)
    mbc, mbr = y.shape[1] // 16, y.shape[0] // 16
    rng = np.random.default_rng(w)
    xy = mb_grid(mbc, mbr)
    n = len(xy)
    mvs = split_vectors(rng, n)
    rp = vp8g.ReferencePlanes(y.reshape(-1), u.reshape(-1), v.reshape(-1), 16 * mbc, 8 * mbc, mbc, mbr)
    orp = oracle_rp(y, u, v, mbc, mbr)
    tables = [vp8g.SIXTAP_FILTERS, vp8g.BILINEAR_FILTERS] + ([CUSTOM] if w < 1000 else [])
    for ti, filt in enumerate(tables):
        for fp in (False, True):
            got = gpu.predict_split_mv_batch(rp, xy, mvs, fp, filt)
            want = np.zeros((n, 384), np.uint8)
            f = np.ascontiguousarray(filt, np.int32)
            oracle.lib.orc_predict_split_mv_batch(C.byref(orp), xy.ctypes.data, mvs.ctypes.data, n, int(fp),
                                                  f.ctypes.data, want.ctypes.data)
            assert_mbs(got, want, f"predict_split_mv table {ti} full_pixel={fp}")


@pytest.mark.parametrize("kind", ["inter", "split"])
def test_reconstructions_match_oracle(gpu, oracle, kind):
    import vp8g
    y, u, v = planes(176, 144, 5)
    mbc, mbr = 11, 9
    rng = np.random.default_rng(len(kind))
    xy = mb_grid(mbc, mbr)
    xy = np.concatenate([xy, xy])  # every macroblock twice: coded and skipped variants
    n = len(xy)
    skip = (rng.random(n) < 0.25).astype(np.uint8)
    rp = vp8g.ReferencePlanes(y.reshape(-1), u.reshape(-1), v.reshape(-1), 16 * mbc, 8 * mbc, mbc, mbr)
    orp = oracle_rp(y, u, v, mbc, mbr)
    for filt in (vp8g.SIXTAP_FILTERS, vp8g.BILINEAR_FILTERS):
        f = np.ascontiguousarray(filt, np.int32)
        for fp in (False, True):
            want = np.zeros((n, 384), np.uint8)
            if kind == "inter":
                mv = rng.integers(-64, 65, (n, 2)).astype(np.int16)
                cf = coeffs(rng, n, 25)
                got = gpu.reconstruct_inter_mb_batch(rp, xy, mv, fp, filt, skip, cf)
                oracle.lib.orc_reconstruct_inter_mb_batch(C.byref(orp), xy.ctypes.data, mv.ctypes.data, n, int(fp),
                                                          f.ctypes.data, skip.ctypes.data, cf.ctypes.data,
                                                          want.ctypes.data)
            else:
                mvs = split_vectors(rng, n)
                cf = coeffs(rng, n, 24)
                got = gpu.reconstruct_split_mv_mb_batch(rp, xy, mvs, fp, filt, skip, cf)
                oracle.lib.orc_reconstruct_split_mv_mb_batch(C.byref(orp), xy.ctypes.data, mvs.ctypes.data, n,
                                                             int(fp), f.ctypes.data, skip.ctypes.data,
                                                             cf.ctypes.data, want.ctypes.data)
            assert_mbs(got, want, f"reconstruct {kind} full_pixel={fp}")


def test_split_equal_vectors_equal_predict_inter_mb_and_dev_entries(gpu):
    import torch
    import vp8g
    y, u, v = planes(352, 288, 7)
    mbc, mbr = 22, 18
    rng = np.random.default_rng(3)
    xy = mb_grid(mbc, mbr)
    n = len(xy)
    mv = rng.integers(-64, 65, (n, 2)).astype(np.int16)
    rp = vp8g.ReferencePlanes(y.reshape(-1), u.reshape(-1), v.reshape(-1), 16 * mbc, 8 * mbc, mbc, mbr)
    split = gpu.predict_split_mv_batch(rp, xy, np.repeat(mv[:, None, :], 16, 1), False, vp8g.SIXTAP_FILTERS)
    whole = gpu.predict_inter_mb_batch(rp, xy, mv, False, vp8g.SIXTAP_FILTERS)
    assert_mbs(split, whole, "16 equal split vectors vs predict_inter_mb")
    # device entries equal the host entries
    mvs = split_vectors(rng, n)
    skip = (rng.random(n) < 0.3).astype(np.uint8)
    cf = coeffs(rng, n, 24)
    host = gpu.reconstruct_split_mv_mb_batch(rp, xy, mvs, True, vp8g.SIXTAP_FILTERS, skip, cf)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a).reshape(-1).view(np.uint8).copy()).cuda()
    dy, du, dv, dxy, dmv, dsk, dcf = t(y), t(u), t(v), t(xy), t(mvs), t(skip), t(cf)
    out = torch.zeros(n * 384, dtype=torch.uint8, device="cuda")
    drp = vp8g._RefPlanes(dy.data_ptr(), du.data_ptr(), dv.data_ptr(), 16 * mbc, 8 * mbc, mbc, mbr)
    f = np.ascontiguousarray(vp8g.SIXTAP_FILTERS, np.int32)
    torch.cuda.synchronize()
    assert gpu.lib.vp8g_reconstruct_split_mv_mb_batch_dev(gpu.h, C.byref(drp), dxy.data_ptr(), dmv.data_ptr(), n, 1,
                                                          f.ctypes.data, dsk.data_ptr(), dcf.data_ptr(),
                                                          out.data_ptr()) == 0
    gpu.synchronize()
    assert_mbs(out.cpu().numpy().reshape(n, 384), host, "reconstruct_split_mv_mb _dev")
    # and the module-level helper restates split_chroma_mvs
    cm = vp8g.split_chroma_mvs([vp8g.Mv(int(r), int(c)) for r, c in mvs[0]])
    assert len(cm) == 4
