"""GPU parity of the section 15 loop filter (include/vp8g.h loop-filter
entry points) against the oracle, bit-exact:

* whole-frame filtering (filter_inter_frame_flags / filter_frame_flags,
  src/loop_filter.rs:1364-1496, 2506-2636) over the frame shapes of
  banded_filter_matches_serial (src/loop_filter.rs:3407-3500) plus shapes
  that recycle waves over more than 32 MB rows, odd row counts, one-row and
  one-column frames, both filter types, key and inter frames, segment
  overrides and ref / mode deltas; through the host entry point and in
  place on device planes at 1920x1088 and 3840x2176;
* candidate scores (score_filter_candidate, src/loop_filter.rs:2062-2095)
  on an odd visible window with an unaligned source, whole-frame and banded
  (the bands of select_filter_params_sampled, src/loop_filter.rs:2303-2365);
* the level / sharpness search (select_filter_params_threaded and
  select_filter_params_sampled, src/loop_filter.rs:2133-2200, 2288-2412);
* the reference's frame-level known answers (src/loop_filter.rs:3136-3184)
  and argument errors.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BPRED, COEFFS = 1, 2


def lf():
    from vp8g import loop_filter
    return loop_filter


def cfg_pair(**kw):
    """The same configuration as the GPU's FrameFilterConfig and the
    oracle's orc_lf_config."""
    from oracle_ctypes import lf_config
    L = lf()
    g = L.FrameFilterConfig(**kw)
    o = lf_config(simple=int(g.simple), key_frame=int(g.key_frame), loop_filter_level=g.loop_filter_level,
                  sharpness_level=g.sharpness_level, segmentation_enabled=int(g.segmentation_enabled),
                  segment_abs=int(g.segment_abs), delta_enabled=int(g.delta_enabled),
                  segment_lf_level=list(g.segment_lf_level), ref_delta=g.ref_deltas(), mode_delta=g.mode_deltas())
    return g, o


def picture(mb_cols, mb_rows, seed, stride_pad=0):
    """A reconstruction with 4x4 block steps over a smooth picture (so the
    filters engage at every level) and the smooth source it came from."""
    rng = np.random.default_rng(seed)
    H, W = mb_rows * 16, mb_cols * 16
    yy, xx = np.mgrid[0:H, 0:W]
    src_y = (96 + 60 * np.sin(xx / 23.0 + seed) * np.cos(yy / 31.0) + rng.normal(0, 3, (H, W)))
    src_y = np.clip(src_y, 0, 255).astype(np.uint8)
    blocks = rng.integers(-9, 10, (H // 4, W // 4)).repeat(4, 0).repeat(4, 1)
    rec_y = np.clip(src_y.astype(np.int32) + blocks, 0, 255).astype(np.uint8)
    ch, cw = mb_rows * 8, mb_cols * 8
    cy, cx = np.mgrid[0:ch, 0:cw]
    src_u = np.clip(128 + 40 * np.sin(cx / 11.0) + rng.normal(0, 2, (ch, cw)), 0, 255).astype(np.uint8)
    src_v = np.clip(120 + 30 * np.cos(cy / 9.0) + rng.normal(0, 2, (ch, cw)), 0, 255).astype(np.uint8)
    cb = rng.integers(-7, 8, (2, ch // 4, cw // 4)).repeat(4, 1).repeat(4, 2)
    rec_u = np.clip(src_u.astype(np.int32) + cb[0], 0, 255).astype(np.uint8)
    rec_v = np.clip(src_v.astype(np.int32) + cb[1], 0, 255).astype(np.uint8)
    if stride_pad:
        rec_y = np.pad(rec_y, ((0, 0), (0, stride_pad)), constant_values=7)
        rec_u = np.pad(rec_u, ((0, 0), (0, stride_pad)), constant_values=7)
        rec_v = np.pad(rec_v, ((0, 0), (0, stride_pad)), constant_values=7)
    return (rec_y, rec_u, rec_v), (src_y, src_u, src_v)


def random_mbs(n, seed, inter=True, segments=True):
    rng = np.random.default_rng(seed)
    m = np.zeros((n, 4), np.uint8)
    m[:, 0] = rng.choice([255, 0, 1, 2, 3], n) if segments else 255
    if inter:
        m[:, 1] = rng.choice([0, 1, 2, 3], n)
        m[:, 2] = np.where(m[:, 1] > 0, rng.choice([0, 1, 2, 3], n), 0)
    m[:, 3] = rng.choice([0, BPRED], n) | np.where(rng.integers(0, 3, n) == 0, COEFFS, 0)
    return m


CONFIGS = [
    dict(key_frame=True, loop_filter_level=30),
    dict(key_frame=True, loop_filter_level=63, sharpness_level=5, simple=True),
    dict(key_frame=False, loop_filter_level=9, sharpness_level=3, delta_enabled=True, ref_delta_current=-4,
         ref_delta_golden=12, zero_mv_mode_delta=-6, bpred_mode_delta=7, split_mv_mode_delta=3,
         other_mv_mode_delta=-2, ref_delta_last=5, ref_delta_altref=-9),
    dict(key_frame=False, loop_filter_level=40, sharpness_level=7, segmentation_enabled=True, segment_abs=False,
         segment_lf_level=(5, -30, 20, 0), delta_enabled=True, ref_delta_last=10),
    dict(key_frame=False, loop_filter_level=20, simple=True, segmentation_enabled=True, segment_abs=True,
         segment_lf_level=(0, 63, 17, 40)),
]

SHAPES = [(1, 1), (1, 2), (2, 1), (2, 2), (5, 3), (3, 7), (9, 6), (4, 37), (7, 33), (1, 40)]


def oracle_filter(oracle, planes, mbs, ocfg, mb_cols):
    return oracle.lf_filter_frame(*planes, mbs, ocfg, mb_cols=mb_cols)


@pytest.mark.parametrize("shape", SHAPES)
def test_frame_filter_matches_oracle(gpu, oracle, shape):
    L = lf()
    mb_cols, mb_rows = shape
    for i, kw in enumerate(CONFIGS):
        g, o = cfg_pair(**kw)
        pad = 16 * (i % 2)
        planes, _ = picture(mb_cols, mb_rows, seed=i + 7 * mb_rows, stride_pad=pad)
        mbs = random_mbs(mb_cols * mb_rows, seed=i + 100 * mb_cols, inter=not g.key_frame)
        want = oracle_filter(oracle, planes, mbs, o, mb_cols)
        kp = L.KeyframePlanes(*(a.copy() for a in planes), mb_cols, mb_rows)
        L.filter_mbs(gpu, kp, mbs, g)
        for name, got_p, want_p in zip("yuv", (kp.y, kp.u, kp.v), want):
            assert (got_p == want_p).all(), (shape, i, name, np.argwhere(got_p != want_p)[:4])


@pytest.mark.parametrize("mb_cols,mb_rows", [(120, 68), (240, 136)])
def test_frame_filter_in_place_on_device(gpu, oracle, mb_cols, mb_rows):
    import torch
    L = lf()
    g, o = cfg_pair(key_frame=False, loop_filter_level=36, sharpness_level=2, delta_enabled=True,
                    ref_delta_last=-5, zero_mv_mode_delta=4)
    planes, _ = picture(mb_cols, mb_rows, seed=mb_cols)
    mbs = random_mbs(mb_cols * mb_rows, seed=3)
    want = oracle_filter(oracle, planes, mbs, o, mb_cols)
    dev = [torch.from_numpy(a.copy()).cuda() for a in planes]
    dm = torch.from_numpy(mbs.copy()).cuda()
    pl = L._Planes(dev[0].data_ptr(), dev[1].data_ptr(), dev[2].data_ptr(), mb_cols * 16, mb_cols * 8, mb_cols,
                   mb_rows)
    torch.cuda.synchronize()
    st = gpu.lib.vp8g_loop_filter_frame_dev(gpu.h, C.byref(pl), dm.data_ptr(), C.byref(g.c()))
    assert st == 0
    assert gpu.lib.vp8g_loop_filter_status(gpu.h) == 0
    for name, d, w in zip("yuv", dev, want):
        got = d.cpu().numpy()
        assert (got == w).all(), (name, np.argwhere(got != w)[:4])


def test_frame_known_answer(gpu):
    """filter_frame_normal_mb_edge_hand_derived, src/loop_filter.rs:3136-3184."""
    L = lf()
    y = np.full((16, 32), 100, np.uint8)
    y[:, 16:] = 110
    u = np.full((8, 16), 100, np.uint8)
    v = u.copy()
    kp = L.KeyframePlanes(y, u, v, 2, 1)
    L.filter_frame_flags(gpu, kp, [False, False], [False, False],
                         L.FrameFilterConfig(key_frame=True, loop_filter_level=20))
    assert (kp.y[:, 13:19] == [101, 103, 104, 106, 107, 109]).all()
    assert (kp.y[:, :13] == 100).all() and (kp.y[:, 19:] == 110).all()
    assert (kp.u == 100).all() and (kp.v == 100).all()


def test_scores_match_oracle(gpu, oracle):
    L = lf()
    mb_cols, mb_rows = 23, 9
    planes, src = picture(mb_cols, mb_rows, seed=5)
    w, h = 361, 139
    # unaligned source planes: views one byte into a wider buffer
    srcs = []
    for a in src:
        b = np.zeros((a.shape[0], a.shape[1] + 3), np.uint8)
        b[:, 1:1 + a.shape[1]] = a
        srcs.append(b[:, 1:])
    mbs = random_mbs(mb_cols * mb_rows, seed=9)
    g, o = cfg_pair(key_frame=False, loop_filter_level=25, sharpness_level=1, segmentation_enabled=True,
                    segment_abs=True, segment_lf_level=(30, 12, 0, 50))
    cands = [(0, 0), (1, 0), (4, 3), (17, 7), (33, 2), (63, 0), (63, 7), (40, 5), (0, 6), (9, 4)]
    lv = np.array([c[0] for c in cands], np.uint8)
    sh = np.array([c[1] for c in cands], np.uint8)
    osrc = (w, h, *(np.ascontiguousarray(a) for a in srcs))
    want = oracle.lf_score_candidates(*planes, mbs, o, osrc, lv, sh, threads=8)
    sp = L.SourcePlanes(w, h, *srcs)
    kp = L.KeyframePlanes(*planes, mb_cols, mb_rows)
    got = L.score_candidates(gpu, kp, mbs, g, sp, cands)
    assert (got == want).all(), (got, want)
    assert got[0] == oracle.lf_ssd(*planes, osrc)


@pytest.mark.parametrize("every", [2, 3, 8])
def test_banded_scores_match_oracle(gpu, oracle, every):
    """The sampled search's candidate score: the bands' SSDs summed."""
    L = lf()
    mb_cols, mb_rows = 17, 13
    planes, src = picture(mb_cols, mb_rows, seed=every)
    mbs = random_mbs(mb_cols * mb_rows, seed=every + 1)
    g, o = cfg_pair(key_frame=False, loop_filter_level=25, delta_enabled=True, ref_delta_golden=-8)
    w, h = mb_cols * 16 - 5, mb_rows * 16 - 9
    kp = L.KeyframePlanes(*planes, mb_cols, mb_rows)
    sp = L.SourcePlanes(w, h, *src)
    cands = [(lv_, s_) for lv_ in (0, 8, 21, 63) for s_ in (0, 4)]
    got = L.score_candidates(gpu, kp, mbs, g, sp, cands, every_rows=every)
    bands = list(range(0, mb_rows, every))
    for (lv_, s_), gv in zip(cands, got):
        want = 0
        for b in bands:
            rows = min(2, mb_rows - b)
            y0 = b * 16
            if y0 >= h:
                continue
            bp = (planes[0][y0:y0 + rows * 16], planes[1][y0 // 2:y0 // 2 + rows * 8],
                  planes[2][y0 // 2:y0 // 2 + rows * 8])
            bs = (w, min(rows * 16, h - y0), src[0][y0:], src[1][y0 // 2:], src[2][y0 // 2:])
            bs = (bs[0], bs[1], *(np.ascontiguousarray(a) for a in bs[2:]))
            want += int(oracle.lf_score_candidates(*(np.ascontiguousarray(a) for a in bp),
                                                   mbs[b * mb_cols:(b + rows) * mb_cols], o, bs,
                                                   np.array([lv_], np.uint8), np.array([s_], np.uint8),
                                                   threads=1)[0])
        assert int(gv) == want, (lv_, s_)


@pytest.mark.parametrize("mb_cols,mb_rows,base_sharp", [(40, 23, 0), (120, 68, 3)])
def test_select_matches_oracle(gpu, oracle, mb_cols, mb_rows, base_sharp):
    L = lf()
    planes, src = picture(mb_cols, mb_rows, seed=mb_rows)
    mbs = random_mbs(mb_cols * mb_rows, seed=mb_cols, segments=False)
    g, o = cfg_pair(key_frame=True, loop_filter_level=20, sharpness_level=base_sharp)
    w, h = mb_cols * 16 - 3, mb_rows * 16 - 8
    osrc = (w, h, *src)
    kp = L.KeyframePlanes(*planes, mb_cols, mb_rows)
    sp = L.SourcePlanes(w, h, *src)
    want = oracle.lf_select(*planes, mbs, o, osrc, threads=8)
    assert want[0] > 0  # the block steps make filtering pay
    assert L.select_filter_params(gpu, kp, mbs, g, sp) == want
    every = L.sampling_for_key_frame(mb_rows)
    for sweep in (True, False):
        want = oracle.lf_select_sampled(*planes, mbs, o, osrc, every, sweep, threads=8)
        assert L.select_filter_params_sampled(gpu, kp, mbs, g, sp, every, sweep) == want


def test_select_flat_source_picks_zero(gpu):
    """src/loop_filter.rs:3667-3694, 3774-3831."""
    L = lf()
    y = np.full((32, 32), 120, np.uint8)
    u = np.full((16, 16), 120, np.uint8)
    kp = L.KeyframePlanes(y, u, u.copy(), 2, 2)
    mbs = L.mb_records(None, [False] * 4, [True] * 4)
    sp = L.SourcePlanes(32, 32, y, u, u.copy())
    assert L.select_filter_params(gpu, kp, mbs, L.FrameFilterConfig(key_frame=True, loop_filter_level=20), sp) \
        == (0, 0)
    assert L.select_filter_params(gpu, kp, mbs, L.FrameFilterConfig(key_frame=True, loop_filter_level=20,
                                                                     sharpness_level=3), sp) == (0, 3)


def test_argument_errors(gpu):
    import torch
    L = lf()
    lib, h = gpu.lib, gpu.h
    g = L.FrameFilterConfig(loop_filter_level=20).c()
    t = torch.zeros(64 * 64, dtype=torch.uint8, device="cuda")
    mbs = torch.zeros(16, dtype=torch.int32, device="cuda")
    base = t.data_ptr()
    good = L._Planes(base, base + 2048, base + 2560, 64, 32, 4, 2)
    assert lib.vp8g_loop_filter_frame_dev(h, C.byref(good), mbs.data_ptr(), C.byref(g)) == 0
    bad = L._Planes(base + 8, base + 2048, base + 2560, 64, 32, 4, 2)       # luma not 16-aligned
    assert lib.vp8g_loop_filter_frame_dev(h, C.byref(bad), mbs.data_ptr(), C.byref(g)) == 4
    bad = L._Planes(base, base + 2048, base + 2560, 48, 32, 4, 2)           # stride < 64
    assert lib.vp8g_loop_filter_frame_dev(h, C.byref(bad), mbs.data_ptr(), C.byref(g)) == 4
    big = L.FrameFilterConfig(loop_filter_level=256).c()
    assert lib.vp8g_loop_filter_frame_dev(h, C.byref(good), mbs.data_ptr(), C.byref(big)) == 4
    empty = L._Planes(None, None, None, 0, 0, 0, 0)
    assert lib.vp8g_loop_filter_frame_dev(h, C.byref(empty), None, C.byref(g)) == 0
    src = L._Source(64, 32, base, base + 2048, base + 2560, 64, 32)
    lv = (C.c_uint8 * 300)()
    out = torch.zeros(300, dtype=torch.int64, device="cuda")
    assert lib.vp8g_loop_filter_score_dev(h, C.byref(good), mbs.data_ptr(), C.byref(g), C.byref(src), lv, lv,
                                          300, 0, out.data_ptr()) == 4
    wide = L._Source(65, 32, base, base + 2048, base + 2560, 65, 33)        # wider than the padded frame
    assert lib.vp8g_loop_filter_score_dev(h, C.byref(good), mbs.data_ptr(), C.byref(g), C.byref(wide), lv, lv,
                                          1, 0, out.data_ptr()) == 4
    assert lib.vp8g_loop_filter_status(h) == 0
