"""GPU parity for motion compensation and sub-pel SAD scoring through the
C ABI against the CPU oracle, bit-exact (src/motion_comp.rs:1986-2078,
src/motion_search.rs:470-728)."""
import ctypes as C
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from conftest import assert_same

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def oracle_predict(oracle, mc, mb_xy, mvs, full, filt):
    from oracle_ctypes import RefPlanes
    rp = RefPlanes(mc.ref_y.ctypes.data, mc.ref_u.ctypes.data, mc.ref_v.ctypes.data,
                   mc.mb_cols * 16, mc.mb_cols * 8, mc.mb_cols, mc.mb_rows)
    n = len(mb_xy)
    out = np.zeros((n, 384), np.uint8)
    f = np.ascontiguousarray(filt, np.int32)
    mb_xy = np.ascontiguousarray(mb_xy, np.uint32)
    mvs = np.ascontiguousarray(mvs, np.int16)
    oracle.lib.orc_predict_inter_mb_batch(rp, mb_xy.ctypes.data, mvs.ctypes.data, n, int(full),
                                          f.ctypes.data, out.ctypes.data)
    return out


def ref_planes(vp8g, mc):
    return vp8g.ReferencePlanes(mc.ref_y.reshape(-1), mc.ref_u.reshape(-1), mc.ref_v.reshape(-1),
                                mc.mb_cols * 16, mc.mb_cols * 8, mc.mb_cols, mc.mb_rows)


@pytest.fixture(scope="module")
def cif_mc():
    import vp8g.synth as synth
    return synth.mc_workload(synth.arnr_group_cif(), config=0)


@pytest.mark.parametrize("table", ["sixtap", "bilinear"])
@pytest.mark.parametrize("full", [False, True])
def test_predict_cif_golden(gpu, oracle, cif_mc, table, full):
    import vp8g
    filt = vp8g.SIXTAP_FILTERS if table == "sixtap" else vp8g.BILINEAR_FILTERS
    got = gpu.predict_inter_mb_batch(ref_planes(vp8g, cif_mc), cif_mc.mb_xy, cif_mc.mvs, full, filt)
    want = oracle_predict(oracle, cif_mc, cif_mc.mb_xy, cif_mc.mvs, full, filt)
    assert_same(got, want, f"{table} full_pixel={full}")
    gold = json.loads((GOLDEN / "mc_cif.json").read_text())
    assert hashlib.sha256(got.tobytes()).hexdigest() == gold[f"predict_{table}_full{int(full)}_sha256"]


def test_predict_extreme_vectors_and_positions(gpu, oracle, cif_mc):
    import vp8g
    rng = np.random.default_rng(4)
    n = 3000
    mb_xy = np.stack([rng.integers(0, cif_mc.mb_cols + 3, n), rng.integers(0, cif_mc.mb_rows + 3, n)], 1)
    choices = np.array([-32768, 32767, -1023, 1023, -16384, 16383, 0, 1, -1, 7, -7, 400, -400], np.int64)
    mvs = np.where(rng.random((n, 2)) < 0.5, rng.choice(choices, (n, 2)), rng.integers(-2000, 2001, (n, 2)))
    mvs = mvs.astype(np.int16)
    for filt in (vp8g.SIXTAP_FILTERS, vp8g.BILINEAR_FILTERS):
        for full in (False, True):
            got = gpu.predict_inter_mb_batch(ref_planes(vp8g, cif_mc), mb_xy, mvs, full, filt)
            assert_same(got, oracle_predict(oracle, cif_mc, mb_xy, mvs, full, filt), f"extreme full={full}",
                        lambda i: (tuple(mb_xy[i[0]]), tuple(mvs[i[0]])))


def test_predict_custom_tap_table(gpu, oracle, cif_mc):
    """Any caller-supplied 8x6 table (the reference takes &[[i32; 6]; 8])."""
    import vp8g
    rng = np.random.default_rng(8)
    filt = rng.integers(-40, 160, (8, 6)).astype(np.int32)
    filt[3] = [0, 0, 128, 0, 0, 0]
    # row 0 is not the identity here: whole-pixel vectors must still copy
    mvs = cif_mc.mvs.copy()
    mvs[::3] = mvs[::3] // 4 * 4
    y8 = (mvs.astype(np.int64) * 2) & 7
    for full in (False, True):
        got = gpu.predict_inter_mb_batch(ref_planes(vp8g, cif_mc), cif_mc.mb_xy, mvs, full, filt)
        want = oracle_predict(oracle, cif_mc, cif_mc.mb_xy, mvs, full, filt)
        assert_same(got, want, f"custom taps full_pixel={full}",
                    lambda i: (tuple(mvs[i[0]]), "luma(mx,my)", (int(y8[i[0], 1]), int(y8[i[0], 0]))))


def test_predict_single_mb_api(gpu, oracle, cif_mc):
    import vp8g
    r = gpu.predict_inter_mb(ref_planes(vp8g, cif_mc), 3, 2, vp8g.Mv(-13, 22), False, vp8g.SIXTAP_FILTERS)
    want = oracle_predict(oracle, cif_mc, [[3, 2]], [[-13, 22]], False, vp8g.SIXTAP_FILTERS)[0]
    assert (r.y == want[:256]).all() and (r.u == want[256:320]).all() and (r.v == want[320:]).all()


def test_predict_1080p_all_mbs(gpu, oracle):
    import vp8g
    import vp8g.synth as synth
    g = synth.arnr_group_regions(n=5)
    mc = synth.mc_workload(g)
    for filt in (vp8g.SIXTAP_FILTERS, vp8g.BILINEAR_FILTERS):
        got = gpu.predict_inter_mb_batch(ref_planes(vp8g, mc), mc.mb_xy, mc.mvs, False, filt)
        assert_same(got, oracle_predict(oracle, mc, mc.mb_xy, mc.mvs, False, filt), "1080p predict")


def oracle_sad(oracle, plane, stride, w, h, mb_xy, src, mvs):
    from oracle_ctypes import LumaRef
    lr = LumaRef(plane.ctypes.data, stride, w, h)
    n = len(mb_xy)
    out = np.zeros(n, np.uint32)
    mb_xy = np.ascontiguousarray(mb_xy, np.uint32)
    src = np.ascontiguousarray(src, np.uint8)
    mvs = np.ascontiguousarray(mvs, np.int16)
    oracle.lib.orc_mb_luma_sad_at_mv_batch(lr, mb_xy.ctypes.data, src.ctypes.data, mvs.ctypes.data, n,
                                           out.ctypes.data)
    return out


def test_sad_1080p_and_visible_dims(gpu, oracle):
    import vp8g
    import vp8g.synth as synth
    g = synth.arnr_group_regions(n=5)
    mc = synth.mc_workload(g)
    plane = mc.ref_y.reshape(-1)
    lw, lh = mc.mb_cols * 16, mc.mb_rows * 16
    got = gpu.mb_luma_sad_at_mv_batch(vp8g.LumaRef(plane, lw, lw, lh), mc.mb_xy, mc.src, mc.mvs)
    assert_same(got, oracle_sad(oracle, plane, lw, lw, lh, mc.mb_xy, mc.src, mc.mvs), "sad")
    # visible-dim clamp bounds narrower than the stride (encoder tests' LumaRef)
    got = gpu.mb_luma_sad_at_mv_batch(vp8g.LumaRef(plane, lw, 1917, 1079), mc.mb_xy, mc.src, mc.mvs)
    assert_same(got, oracle_sad(oracle, plane, lw, 1917, 1079, mc.mb_xy, mc.src, mc.mvs), "sad visible")
    # whole-pel vectors: mb_luma_sad_at_whole_mv
    whole = (mc.mvs // 4 * 4).astype(np.int16)
    got = gpu.mb_luma_sad_at_mv_batch(vp8g.LumaRef(plane, lw, lw, lh), mc.mb_xy, mc.src, whole)
    assert_same(got, oracle_sad(oracle, plane, lw, lw, lh, mc.mb_xy, mc.src, whole), "sad whole")


def test_sad_cif_golden(gpu, cif_mc):
    import vp8g
    gold = json.loads((GOLDEN / "mc_cif.json").read_text())
    lw, lh = cif_mc.mb_cols * 16, cif_mc.mb_rows * 16
    got = gpu.mb_luma_sad_at_mv_batch(vp8g.LumaRef(cif_mc.ref_y.reshape(-1), lw, lw, lh), cif_mc.mb_xy,
                                      cif_mc.src, cif_mc.mvs)
    assert hashlib.sha256(got.tobytes()).hexdigest() == gold["sad_sha256"]


@pytest.mark.parametrize("ladder", ["half", "quarter"])
def test_refine_ladders(gpu, oracle, cif_mc, ladder):
    import vp8g
    from oracle_ctypes import LumaRef, MvS
    rng = np.random.default_rng(12 if ladder == "half" else 13)
    lw, lh = cif_mc.mb_cols * 16, cif_mc.mb_rows * 16
    plane = cif_mc.ref_y.reshape(-1)
    n = len(cif_mc.mb_xy)
    unit = 4 if ladder == "half" else 2
    centers = (rng.integers(-255, 256, (n, 2)) * unit).clip(-1023, 1023).astype(np.int16)
    centers[:20] = [[1023 // unit * unit, -(1023 // unit * unit)]] * 20
    centers[20:40] = [[-1020, 1020]] * 20
    fn = gpu.half_pixel_refine_luma_batch if ladder == "half" else gpu.quarter_pixel_refine_luma_batch
    got = fn(vp8g.LumaRef(plane, lw, lw, lh), cif_mc.mb_xy, cif_mc.src, centers)
    lr = LumaRef(plane.ctypes.data, lw, lw, lh)
    ofn = oracle.lib.orc_half_pixel_refine_luma if ladder == "half" else oracle.lib.orc_quarter_pixel_refine_luma
    for i in range(n):
        src = np.ascontiguousarray(cif_mc.src[i])
        r = ofn(C.byref(lr), int(cif_mc.mb_xy[i, 0]), int(cif_mc.mb_xy[i, 1]), src.ctypes.data,
                MvS(int(centers[i, 0]), int(centers[i, 1])))
        assert tuple(got[i]) == (r.mv.row, r.mv.col, r.sad), i


def test_refine_rejects_out_of_range_center(gpu, cif_mc):
    import vp8g
    lw, lh = cif_mc.mb_cols * 16, cif_mc.mb_rows * 16
    with pytest.raises(vp8g.Vp8gError):
        gpu.half_pixel_refine_luma_batch(vp8g.LumaRef(cif_mc.ref_y.reshape(-1), lw, lw, lh),
                                         [[0, 0]], cif_mc.src[:1], [[1024, 0]])


@pytest.mark.parametrize("diagonals", [True, False])
def test_half_plane_sads_match_oracle(gpu, oracle, diagonals):
    """HalfPlanes over the 1080p reference at random whole-pel positions,
    including positions far outside the plane (all reads clamp)."""
    import vp8g
    import vp8g.synth as synth
    from oracle_ctypes import HalfPlanes, LumaRef
    g = synth.arnr_group_cif()
    mc = synth.mc_workload(g, config=0)
    lw, lh = mc.mb_cols * 16, mc.mb_rows * 16
    plane = mc.ref_y.reshape(-1)
    rng = np.random.default_rng(33 + diagonals)
    n = 600
    xy = np.stack([rng.integers(-40, lw + 40, n), rng.integers(-40, lh + 40, n)], 1).astype(np.int32)
    src = rng.integers(0, 256, (n, 256)).astype(np.uint8)
    for taps in (vp8g.SIXTAP_FILTERS[4], vp8g.BILINEAR_FILTERS[4], np.array([5, -30, 90, 70, -12, 5], np.int32)):
        got = gpu.half_plane_sads_batch(vp8g.LumaRef(plane, lw, lw, lh), xy, src, taps, diagonals)
        lr = LumaRef(plane.ctypes.data, lw, lw, lh)
        t = np.ascontiguousarray(taps, np.int32)
        want = np.zeros((n, 8), np.uint32)
        for i in range(n):
            win = np.zeros(22 * 22, np.uint8)
            oracle.lib.orc_fetch_window(C.byref(lr), int(xy[i, 0]), int(xy[i, 1]), win.ctypes.data)
            hp = HalfPlanes()
            oracle.lib.orc_half_planes_new(win.ctypes.data, t.ctypes.data, int(diagonals), C.byref(hp))
            s = np.ascontiguousarray(src[i])
            k = 0
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    if dx == 0 and dy == 0:
                        continue
                    diag = dx != 0 and dy != 0
                    want[i, k] = 0xFFFFFFFF if (diag and not diagonals) else \
                        oracle.lib.orc_half_planes_sad(C.byref(hp), s.ctypes.data, dx, dy)
                    k += 1
        assert_same(got, want, f"half planes taps={list(taps)} diagonals={diagonals}")


def test_device_refine_and_mc_entries_match_host_entries(gpu):
    """The _dev refinement and MC entries give the host entries' results
    (those are pinned against the oracle above) on the same inputs."""
    import torch
    import vp8g
    import vp8g.synth as synth
    g = synth.arnr_group_cif(96, 64, 3, clip=6)
    w, h = g.width, g.height
    bw, bh = w // 16, h // 16
    rng = np.random.default_rng(12)
    n = bw * bh
    mb_xy = np.stack(np.meshgrid(np.arange(bw), np.arange(bh)), -1).reshape(-1, 2).astype(np.uint32)
    centers = rng.integers(-40, 41, (n, 2)).astype(np.int16)
    src = np.ascontiguousarray(g.y[1].reshape(bh, 16, bw, 16).transpose(0, 2, 1, 3).reshape(n, 256))
    ref = vp8g.LumaRef(g.y[0].reshape(-1), w, w, h)
    want_h = gpu.half_pixel_refine_luma_batch(ref, mb_xy, src, centers)
    want_q = gpu.quarter_pixel_refine_luma_batch(ref, mb_xy, src, centers)
    dy = torch.from_numpy(g.y[0].reshape(-1).copy()).cuda()
    dxy = torch.from_numpy(mb_xy.astype(np.int32)).cuda()
    dsrc = torch.from_numpy(src).cuda()
    dc = torch.from_numpy(centers).cuda()
    oh = torch.zeros((n, 2), dtype=torch.int32, device="cuda")
    oq = torch.zeros_like(oh)
    torch.cuda.synchronize()
    gpu.refine_batch_dev(dy.data_ptr(), w, w, h, dxy.data_ptr(), dsrc.data_ptr(), dc.data_ptr(), n, oh.data_ptr(), False)
    gpu.refine_batch_dev(dy.data_ptr(), w, w, h, dxy.data_ptr(), dsrc.data_ptr(), dc.data_ptr(), n, oq.data_ptr(), True)
    gpu.synchronize()
    for got, want in ((oh, want_h), (oq, want_q)):
        raw = got.cpu().numpy()
        rows = (raw[:, 0] & 0xFFFF).astype(np.uint16).view(np.int16)
        cols = ((raw[:, 0] >> 16) & 0xFFFF).astype(np.uint16).view(np.int16)
        assert np.array_equal(np.stack([rows, cols], 1), want[:, :2]) and np.array_equal(raw[:, 1].astype(np.uint32), want[:, 2])
