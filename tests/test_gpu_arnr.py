"""GPU parity for ARNR altref synthesis: the gfx950 path through the C ABI
against the CPU oracle, bit-exact, plus the reference's invariants and
error behaviour (src/arnr.rs:476-650, 732-829)."""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from conftest import H, W, clean_scene, mse, noisy, packed, translated

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def as_frames(vp8g, tuples):
    return [vp8g.I420Frame(*t) for t in tuples]


def check(gpu, oracle, tuples, center, strength):
    import vp8g
    st, want, _, _ = oracle.build_arnr_altref(tuples, center, strength, threads=8)
    assert st == 0
    pic = gpu.build_arnr_altref(as_frames(vp8g, tuples), center, vp8g.ArnrConfig(strength))
    for name, got, exp in zip("yuv", (pic.y, pic.u, pic.v), want):
        if not np.array_equal(got, exp):
            bad = np.nonzero(got != exp)[0]
            raise AssertionError(f"plane {name}: {bad.size} bytes differ, first at {bad[:8]}")
    return pic


def test_reference_scenes(gpu, oracle):
    cy, cu, cv = clean_scene()
    # identity (strength 0, single frame)
    pic = check(gpu, oracle, [packed(W, H, cy, cu, cv)], 0, 3)
    assert (pic.y == cy).all()
    ny = noisy(cy, 7, 6)
    pic = check(gpu, oracle, [packed(W, H, ny, cu, cv), packed(W, H, cy, cu, cv)], 0, 0)
    assert (pic.y == ny).all()
    # static noisy window denoises
    planes = [noisy(cy, 1000 + i, 6) for i in range(5)]
    pic = check(gpu, oracle, [packed(W, H, p, cu, cv) for p in planes], 2, 3)
    assert mse(pic.y, cy) < mse(planes[2], cy) / 2
    # translating window denoises
    planes = [noisy(translated(cy, (i - 1) * 4), 2000 + i, 6) for i in range(3)]
    pic = check(gpu, oracle, [packed(W, H, p, cu, cv) for p in planes], 1, 3)
    assert mse(pic.y, cy) < mse(planes[1], cy)
    # alien neighbour excluded
    alien = (255 - cy).astype(np.uint8)
    pic = check(gpu, oracle, [packed(W, H, cy, cu, cv), packed(W, H, alien, cu, cv)], 0, 6)
    assert (pic.y == cy).all()


GEOMS = [(1, 1, 3, 1), (2, 3, 2, 0), (15, 33, 3, 2), (16, 16, 5, 2), (17, 9, 4, 3),
         (37, 29, 3, 0), (64, 40, 2, 1), (100, 70, 7, 3), (33, 47, 11, 5), (176, 144, 9, 4)]


@pytest.mark.parametrize("geom", GEOMS, ids=lambda g: "x".join(map(str, g)))
@pytest.mark.parametrize("strength", [0, 1, 3, 6, 9])
def test_random_geometries(gpu, oracle, geom, strength):
    w, h, n, center = geom
    rng = np.random.default_rng(w * 7919 + h * 31 + n)
    cw, ch = (w + 1) // 2, (h + 1) // 2
    base = rng.integers(0, 256, (h, w)).astype(np.int64)
    tuples = []
    for i in range(n):
        y = np.clip(np.roll(base, (i - center, 2 * (i - center)), (0, 1)) + rng.integers(-6, 7, (h, w)), 0, 255)
        tuples.append(packed(w, h, y.astype(np.uint8).reshape(-1),
                             rng.integers(0, 256, cw * ch).astype(np.uint8),
                             rng.integers(0, 256, cw * ch).astype(np.uint8)))
    check(gpu, oracle, tuples, center, strength)


def test_strided_unaligned_planes(gpu, oracle):
    import vp8g.synth as synth
    g = synth.arnr_group_cif(98, 66, 5)
    tuples = []
    for (w, h, y, u, v, _, _) in g.frames():
        cw, ch = (w + 1) // 2, (h + 1) // 2
        ys = np.full((h, w + 13), 7, np.uint8); ys[:, :w] = y.reshape(h, w)
        us = np.full((ch, cw + 5), 9, np.uint8); us[:, :cw] = u.reshape(ch, cw)
        vs = np.full((ch, cw + 5), 11, np.uint8); vs[:, :cw] = v.reshape(ch, cw)
        tuples.append((w, h, ys.reshape(-1), us.reshape(-1), vs.reshape(-1), w + 13, cw + 5))
    check(gpu, oracle, tuples, g.center, 3)


def test_cif_group_golden(gpu, oracle):
    import vp8g
    import vp8g.synth as synth
    gold = json.loads((GOLDEN / "arnr_cif.json").read_text())
    g = synth.arnr_group_cif()
    pic = check(gpu, oracle, g.frames(), g.center, 3)
    assert hashlib.sha256(pic.y.tobytes() + pic.u.tobytes() + pic.v.tobytes()).hexdigest() == gold["output_sha256"]
    # launch shape never changes bytes
    for waves, remap, mbs in ((1, 1, 1), (2, 0, 3), (3, 1, 8), (4, 0, 2), (16, 1, 5)):
        gpu.set_arnr_waves(waves)
        gpu.set_xcd_remap(remap)
        gpu.set_arnr_mbs(mbs)
        try:
            p2 = gpu.build_arnr_altref(as_frames(vp8g, g.frames()), g.center, vp8g.ArnrConfig(3))
        finally:
            gpu.set_arnr_waves(0)
            gpu.set_xcd_remap(1)
            gpu.set_arnr_mbs(0)
        assert (p2.y == pic.y).all() and (p2.u == pic.u).all() and (p2.v == pic.v).all()


def test_device_entry_and_mv_table(gpu, oracle):
    """The device-resident entry point and its per-block alignment table."""
    import ctypes as C
    import torch
    import vp8g
    import vp8g.synth as synth
    g = synth.arnr_group_cif(160, 96, 7, clip=3)
    st, want, _, mvs = oracle.build_arnr_altref(g.frames(), g.center, 4, want_mvs=True)
    dev = [(torch.from_numpy(y.copy()).cuda(), torch.from_numpy(u.copy()).cuda(), torch.from_numpy(v.copy()).cuda())
           for y, u, v in zip(g.y, g.u, g.v)]
    descs = (vp8g._I420 * g.n)()
    for i, (y, u, v) in enumerate(dev):
        descs[i] = vp8g._I420(g.width, g.height, y.data_ptr(), u.data_ptr(), v.data_ptr(), g.width, (g.width + 1) // 2)
    cw, ch = (g.width + 1) // 2, (g.height + 1) // 2
    oy = torch.empty(g.width * g.height, dtype=torch.uint8, device="cuda")
    ou = torch.empty(cw * ch, dtype=torch.uint8, device="cuda")
    ov = torch.empty(cw * ch, dtype=torch.uint8, device="cuda")
    mv_dev = torch.zeros(mvs.size, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    gpu.build_arnr_altref_dev(descs, g.n, g.center, 4, oy.data_ptr(), ou.data_ptr(), ov.data_ptr(),
                              mv_dev.data_ptr())
    gpu.synchronize()
    assert (oy.cpu().numpy() == want[0]).all()
    assert (ou.cpu().numpy() == want[1]).all() and (ov.cpu().numpy() == want[2]).all()
    got = mv_dev.cpu().numpy().reshape(mvs.shape)
    assert (got == mvs).all()


@pytest.mark.parametrize("layout", ["aligned", "odd_offsets", "mixed_strides", "odd_width"])
def test_device_entry_plane_layouts(gpu, oracle, layout):
    """Device-resident frames in every layout the kernel stages differently:
    16-byte aligned planes of one stride (windows by DMA, pieces outside the
    left and right edges filled from the edge pixel), planes at odd byte
    offsets and frames of different strides (clamped gathers), and a width
    that is not a multiple of 16 (gathers at the right edge only)."""
    import torch
    import vp8g
    import vp8g.synth as synth
    w, h = (136, 72) if layout == "odd_width" else (144, 80)
    g = synth.arnr_group_cif(w, h, 5, clip=6)
    cw, ch = (w + 1) // 2, (h + 1) // 2
    keep, descs, tuples = [], (vp8g._I420 * g.n)(), []
    for i, (y, u, v) in enumerate(zip(g.y, g.u, g.v)):
        pad = {"mixed_strides": 16 * (i % 3), "odd_width": 8}.get(layout, 0)
        off = (1 + 2 * i) if layout == "odd_offsets" else 0
        ys, cs = w + pad, cw + pad

        def plane(p, pw, ph, stride):
            buf = np.full(off + stride * ph, 77, np.uint8)
            buf[off:].reshape(ph, stride)[:, :pw] = np.asarray(p).reshape(ph, pw)
            t = torch.from_numpy(buf).cuda()
            keep.append(t)
            return t.data_ptr() + off, buf[off:]
        py, hy = plane(y, w, h, ys)
        pu, hu = plane(u, cw, ch, cs)
        pv, hv = plane(v, cw, ch, cs)
        descs[i] = vp8g._I420(w, h, py, pu, pv, ys, cs)
        tuples.append((w, h, hy, hu, hv, ys, cs))
    st, want, _, mvs = oracle.build_arnr_altref(tuples, g.center, 3, want_mvs=True)
    assert st == 0
    oy = torch.empty(w * h, dtype=torch.uint8, device="cuda")
    ou = torch.empty(cw * ch, dtype=torch.uint8, device="cuda")
    ov = torch.empty(cw * ch, dtype=torch.uint8, device="cuda")
    mv_dev = torch.zeros(mvs.size, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    gpu.build_arnr_altref_dev(descs, g.n, g.center, 3, oy.data_ptr(), ou.data_ptr(), ov.data_ptr(),
                              mv_dev.data_ptr())
    gpu.synchronize()
    assert (mv_dev.cpu().numpy().reshape(mvs.shape) == mvs).all()
    assert (oy.cpu().numpy() == want[0]).all()
    assert (ou.cpu().numpy() == want[1]).all() and (ov.cpu().numpy() == want[2]).all()


def test_1080p_region_motion_group(gpu, oracle):
    """The 1920x1080 seven-frame region-motion group with occluders."""
    import vp8g.synth as synth
    g = synth.arnr_group_regions()
    check(gpu, oracle, g.frames(), g.center, 3)


def test_errors_follow_reference_order(gpu):
    import vp8g
    cy, cu, cv = clean_scene()
    f0 = vp8g.I420Frame.packed(W, H, cy, cu, cv)
    small = vp8g.I420Frame.packed(32, 32, np.zeros(1024, np.uint8), np.full(256, 128, np.uint8),
                                  np.full(256, 128, np.uint8))
    with pytest.raises(vp8g.ReferenceDimensionsMismatch) as e:
        gpu.build_arnr_altref([f0, small], 0, vp8g.ArnrConfig(3))
    assert e.value.source == (48, 48) and e.value.reference == (32, 32)
    with pytest.raises(vp8g.ReferenceDimensionsMismatch):
        gpu.build_arnr_altref([f0, small], 0, vp8g.ArnrConfig(0))
    zero = vp8g.I420Frame(0, H, np.zeros(0, np.uint8), np.zeros(0, np.uint8), np.zeros(0, np.uint8), 0, 0)
    with pytest.raises(vp8g.InvalidDimensions):
        gpu.build_arnr_altref([zero], 0, vp8g.ArnrConfig(3))
    with pytest.raises(IndexError):
        gpu.build_arnr_altref([f0], 1, vp8g.ArnrConfig(3))
    with pytest.raises(vp8g.Vp8gError):
        gpu.build_arnr_altref([f0] * 65, 0, vp8g.ArnrConfig(3))
    assert vp8g.ArnrConfig.new(99).strength == 6 and vp8g.ArnrConfig.new(2).strength == 2


def test_fixed_point_and_reentry(gpu):
    import vp8g
    rng = np.random.default_rng(9)
    w, h = 41, 35
    y = rng.integers(0, 256, w * h).astype(np.uint8)
    u = rng.integers(0, 256, 21 * 18).astype(np.uint8)
    v = rng.integers(0, 256, 21 * 18).astype(np.uint8)
    f = vp8g.I420Frame.packed(w, h, y, u, v)
    for s in range(8):
        pic = gpu.build_arnr_altref([f] * 4, 1, vp8g.ArnrConfig(s))
        assert (pic.y == y).all() and (pic.u == u).all() and (pic.v == v).all()
    echo = gpu.build_arnr_altref([pic.as_i420()], 0, vp8g.ArnrConfig(0))
    assert (echo.y == pic.y).all()


def test_max_group_length(gpu, oracle):
    import vp8g.synth as synth
    g = synth.arnr_group_cif(48, 32, 64, clip=5)
    check(gpu, oracle, g.frames(), 40, 3)


@pytest.mark.parametrize("threads,streams,bands", [(1, 1, 1), (3, 2, 2), (8, 4, 8), (0, 3, 0), (2, 2, 5)])
def test_host_copy_shapes_leave_bytes_unchanged(gpu, oracle, threads, streams, bands):
    """The host threads and row bands of the host-pointer entry (and the
    copy-stream option, accepted without effect) change no output byte:
    strided planes of odd sizes (partial last MB rows, bands of uneven
    MB-row counts, more bands than MB rows, rows whose chroma is odd-sized
    and interleaved U|V on the device), and a call after a smaller one that
    reuses the staging."""
    import vp8g
    import vp8g.synth as synth
    try:
        gpu.set_host_threads(threads)
        gpu.set_copy_streams(streams)
        gpu.set_host_bands(bands)
        for (w, h, n, c) in ((98, 66, 5, 2), (353, 289, 7, 3), (37, 21, 2, 0)):
            g = synth.arnr_group_cif(w, h, n, clip=w)
            tuples = []
            for (fw, fh, y, u, v, _, _) in g.frames():
                cw, ch = (fw + 1) // 2, (fh + 1) // 2
                ys = np.full((fh, fw + 7), 3, np.uint8); ys[:, :fw] = y.reshape(fh, fw)
                us = np.full((ch, cw + 3), 5, np.uint8); us[:, :cw] = u.reshape(ch, cw)
                vs = np.full((ch, cw + 3), 6, np.uint8); vs[:, :cw] = v.reshape(ch, cw)
                tuples.append((fw, fh, ys.reshape(-1), us.reshape(-1), vs.reshape(-1), fw + 7, cw + 3))
            check(gpu, oracle, tuples, c, 3)
        with pytest.raises(vp8g.Vp8gError):
            gpu.set_copy_streams(5)
        with pytest.raises(vp8g.Vp8gError):
            gpu.set_host_threads(65)
        with pytest.raises(vp8g.Vp8gError):
            gpu.set_host_bands(9)
    finally:
        gpu.set_host_threads(0)
        gpu.set_copy_streams(0)
        gpu.set_host_bands(0)
