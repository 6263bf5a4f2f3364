"""The stream driver's GPU pieces: luma_mad (src/stream.rs:2654-2666)
against the oracle, and the lookahead grouping of vp8g.stream against the
reference stream tests' known answers (tests/encoder_altref_stream.rs:
98-260, 350-424: three groups for window 4; keys at 0, 5, 10 for a
5-frame cadence; a scene cut keys its frame and starts a group; resizing
is rejected) with every anchor equal to the oracle's
build_arnr_altref(group, len - 1) on the same frames."""
import numpy as np
import pytest

from conftest import Lcg

pytestmark = pytest.mark.gpu

W = H = 64


def clean_frame(t):
    """tests/encoder_altref_stream.rs:55-68."""
    r, c = np.mgrid[0:H, 0:W]
    sc = np.clip(c - t, 0, W - 1)
    y = (48 + ((r * 5 + sc * 3) & 0x9F)).astype(np.uint8).reshape(-1)
    return y, np.full(32 * 32, 118 + t, np.uint8), np.full(32 * 32, 134 - t, np.uint8)


def noisy_frame(t):
    """tests/encoder_altref_stream.rs:70-78."""
    y, u, v = clean_frame(t)
    rng = Lcg(0x5EED0000 + t)
    ny = np.array([min(255, max(0, int(p) + rng.noise(5))) for p in y], np.uint8)
    return ny, u, v


def frame(vp8g, y, u, v, w=W, h=H):
    return vp8g.I420Frame.packed(w, h, y, u, v)


def run(gpu, frames, chunked=False, **cfg):
    import vp8g.stream as vs
    b = vs.AltrefStreamBatcher(gpu, vs.AltrefStreamConfig(**cfg))
    groups = []
    if chunked:
        groups += b.push_frames(frames)
    else:
        for f in frames:
            groups += b.push_frame(f)
    groups += b.finish()
    assert b.finish() == []
    return groups


def check_anchors(oracle, frames, groups, strength=3):
    for g in groups:
        fs = frames[g.start:g.start + g.length]
        if g.length < 2:
            assert g.anchor is None
            continue
        tuples = [(f.width, f.height, f.y, f.u, f.v, f.y_stride, f.uv_stride) for f in fs]
        st, want, _, _ = oracle.build_arnr_altref(tuples, g.length - 1, strength)
        assert st == 0
        for name, got, exp in zip("yuv", (g.anchor.y, g.anchor.u, g.anchor.v), want):
            assert np.array_equal(got, exp), f"group {g.start}+{g.length} plane {name}"


def test_luma_mad_matches_oracle(gpu, oracle):
    import vp8g
    import vp8g.stream as vs
    rng = np.random.default_rng(1)
    keep = []
    from oracle_ctypes import I420
    import ctypes as C
    for (w, h, pad) in ((64, 64, 0), (37, 29, 5), (1920, 1080, 0), (3, 2, 1)):
        prev, cur = [], []
        for _ in range(5):
            for lst in (prev, cur):
                y = rng.integers(0, 256, h * (w + pad)).astype(np.uint8)
                c = np.zeros(((h + 1) // 2) * ((w + 1) // 2 + pad), np.uint8)
                lst.append(vp8g.I420Frame(w, h, y, c, c, w + pad, (w + 1) // 2 + pad))
        got = vs.luma_mad_batch(gpu, prev, cur)
        for i in range(5):
            a, b = (I420(f.width, f.height, f.y.ctypes.data, f.u.ctypes.data, f.v.ctypes.data, f.y_stride,
                         f.uv_stride) for f in (prev[i], cur[i]))
            assert got[i] == oracle.lib.orc_luma_mad(C.byref(a), C.byref(b)), (w, h, i)


def test_window_groups_and_anchors(gpu, oracle):
    import vp8g
    frames = [frame(vp8g, *noisy_frame(t)) for t in range(12)]
    for chunked in (False, True):
        groups = run(gpu, frames, chunked, altref_window=4)
        assert [(g.start, g.length) for g in groups] == [(0, 4), (4, 4), (8, 4)]
        assert [g.start for g in groups if g.key] == [0]
        check_anchors(oracle, frames, groups)


def test_keyframe_cadence_closes_groups_early(gpu, oracle):
    import vp8g
    frames = [frame(vp8g, *noisy_frame(t)) for t in range(11)]
    groups = run(gpu, frames, keyframe_interval=5, altref_window=4)
    assert [(g.start, g.length) for g in groups] == [(0, 4), (4, 1), (5, 4), (9, 1), (10, 1)]
    assert [g.start for g in groups if g.key] == [0, 5, 10]
    check_anchors(oracle, frames, groups)


def test_scene_cut_closes_the_group_and_keys(gpu, oracle):
    import vp8g
    frames = []
    for t in range(10):
        y, u, v = noisy_frame(t)
        if t >= 6:
            y = (255 - y).astype(np.uint8)
        frames.append(frame(vp8g, y, u, v))
    groups = run(gpu, frames, altref_window=4, scene_cut_mad_threshold=40.0)
    assert [g.start for g in groups if g.key] == [0, 6]
    assert [(g.start, g.length) for g in groups] == [(0, 4), (4, 2), (6, 4)]
    assert [g.forced_key for g in groups] == [False, False, True]
    check_anchors(oracle, frames, groups)
    groups = run(gpu, frames, chunked=True, altref_window=4, scene_cut_mad_threshold=0.0)
    assert [g.start for g in groups if g.key] == [0]
    assert [(g.start, g.length) for g in groups] == [(0, 4), (4, 4), (8, 2)]
    check_anchors(oracle, frames, groups)


def test_zero_window_and_dimension_lock(gpu):
    import vp8g
    import vp8g.stream as vs
    with pytest.raises(ValueError):
        vs.AltrefStreamBatcher(gpu, vs.AltrefStreamConfig(altref_window=0))
    b = vs.AltrefStreamBatcher(gpu, vs.AltrefStreamConfig())
    b.push_frame(frame(vp8g, *noisy_frame(0)))
    small = vp8g.I420Frame.packed(32, 32, np.zeros(1024, np.uint8), np.full(256, 128, np.uint8),
                                  np.full(256, 128, np.uint8))
    with pytest.raises(vs.DimensionsChanged):
        b.push_frame(small)


def test_long_uneven_stream_matches_per_group_oracle(gpu, oracle):
    """Three scenes of a region-motion clip (cuts at 14 and 27), window 7,
    keys every 16 frames, pushed in uneven chunks: every group's anchor
    equals the oracle's and the grouping equals one-frame-at-a-time push."""
    import vp8g
    import vp8g.synth as synth
    frames = []
    for scene, (n, seed, lift) in enumerate(((14, 1, 0), (13, 2, 110), (13, 3, -110))):
        g = synth.arnr_group_regions(176, 144, n, vmax=4, occluders=2, clip=seed, config=6)
        for t in range(n):   # brightness offsets keep every cut's MAD far above 40
            y = np.clip(g.y[t].astype(np.int32) + lift, 0, 255).astype(np.uint8)
            frames.append(frame(vp8g, y, g.u[t], g.v[t], 176, 144))
    import vp8g.stream as vs
    b = vs.AltrefStreamBatcher(gpu, vs.AltrefStreamConfig(keyframe_interval=16, altref_window=7))
    groups, i = [], 0
    for k in (5, 1, 9, 13, 2, 10):
        groups += b.push_frames(frames[i:i + k])
        i += k
    groups += b.finish()
    single = run(gpu, frames, keyframe_interval=16, altref_window=7)
    assert [(g.start, g.length, g.key) for g in groups] == [(g.start, g.length, g.key) for g in single]
    starts = [g.start for g in groups if g.key]
    assert 0 in starts and 14 in starts and 27 in starts and 16 in starts and 32 in starts
    check_anchors(oracle, frames, groups)
