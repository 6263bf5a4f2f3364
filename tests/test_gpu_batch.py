"""GPU parity for batched ARNR: many groups of uneven resolution and length
in one persistent, graph-captured launch (vp8g_arnr_batch_*), each group's
bytes and alignment table equal to its own build_arnr_altref result
(src/arnr.rs:488-650) as the oracle computes it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class DeviceGroup:
    """One group uploaded to the GPU with its output buffers."""

    def __init__(self, torch, tuples, center, strength, want_mvs=True):
        self.tuples, self.center, self.strength = tuples, center, strength
        w, h = tuples[center][0], tuples[center][1]
        self.w, self.h = w, h
        cw, ch = (w + 1) // 2, (h + 1) // 2
        self.planes = []
        self.frames = []
        for (fw, fh, y, u, v, ys, cs) in tuples:
            ty, tu, tv = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (y, u, v))
            self.planes.append((ty, tu, tv))
            self.frames.append((fw, fh, ty.data_ptr(), tu.data_ptr(), tv.data_ptr(), ys, cs))
        self.oy = torch.full((w * h,), 0xA5, dtype=torch.uint8, device="cuda")
        self.ou = torch.full((cw * ch,), 0xA5, dtype=torch.uint8, device="cuda")
        self.ov = torch.full((cw * ch,), 0xA5, dtype=torch.uint8, device="cuda")
        self.mv = None
        if want_mvs:
            n = len(tuples)
            self.mv = torch.zeros(n * ((h + 15) // 16) * ((w + 15) // 16) * 3, dtype=torch.int32,
                                  device="cuda")

    def job(self, vp8g):
        return vp8g.ArnrJob(self.frames, self.center, self.strength, self.oy.data_ptr(),
                            self.ou.data_ptr(), self.ov.data_ptr(),
                            self.mv.data_ptr() if self.mv is not None else None)

    def outputs(self):
        return self.oy.cpu().numpy(), self.ou.cpu().numpy(), self.ov.cpu().numpy()


def expect(oracle, tuples, center, strength):
    st, want, _, mvs = oracle.build_arnr_altref(tuples, center, strength, threads=8, want_mvs=True)
    assert st == 0
    return want, mvs


def compare(tag, dg, want, mvs):
    for name, got, exp in zip("yuv", dg.outputs(), want):
        if not np.array_equal(got, exp):
            bad = np.nonzero(got != exp)[0]
            raise AssertionError(f"{tag} plane {name}: {bad.size} bytes differ, first at {bad[:8]}")
    if dg.mv is not None:
        got = dg.mv.cpu().numpy().reshape(mvs.shape)
        if not np.array_equal(got, mvs):
            raise AssertionError(f"{tag}: alignment table differs at {np.argwhere(got != mvs)[:4]}")


@pytest.fixture(scope="module")
def clips64():
    import vp8g.synth as synth
    specs = synth.batch_specs(64)
    return specs, synth.batch_clips(64, specs=specs)


@pytest.mark.timeout(900)
def test_uneven_batch_64_clips_bit_exact(gpu, oracle, clips64):
    """64 clips, resolutions in {176x144 .. 1920x1080}, lengths in
    {3, 5, 7, 9, 11}, region motion with occluders; one persistent launch,
    replayed three times with host work between replays."""
    import torch
    import vp8g
    specs, clips = clips64
    print(f"batch of {len(clips)} clips: {specs[:4]} ...", flush=True)
    groups = [DeviceGroup(torch, g.frames(), g.center, 3) for g in clips]
    torch.cuda.synchronize()
    batch = gpu.create_arnr_batch([dg.job(vp8g) for dg in groups])
    shape = batch.shape()
    assert shape["units"] > 0 and shape["workgroups"] > 0
    wants = [expect(oracle, g.frames(), g.center, 3) for g in clips]
    for replay in range(3):
        for dg in groups:
            dg.oy.fill_(0x5A)
            if dg.mv is not None:
                dg.mv.zero_()
        torch.cuda.synchronize()
        batch.run()
        gpu.synchronize()
        for i, (dg, (want, mvs)) in enumerate(zip(groups, wants)):
            compare(f"replay {replay} clip {i} {specs[i]}", dg, want, mvs)
    batch.close()


def test_batch_reads_frames_at_run_time(gpu, oracle):
    """A replay reads the frame buffers as they are when it runs."""
    import torch
    import vp8g
    import vp8g.synth as synth
    a = synth.arnr_group_regions(208, 120, 5, vmax=6, occluders=3, clip=1, config=2)
    b = synth.arnr_group_regions(208, 120, 5, vmax=6, occluders=3, clip=2, config=2)
    dg = DeviceGroup(torch, a.frames(), a.center, 4)
    torch.cuda.synchronize()
    batch = gpu.create_arnr_batch([dg.job(vp8g)])
    batch.run()
    gpu.synchronize()
    compare("first content", dg, *expect(oracle, a.frames(), a.center, 4))
    for (ty, tu, tv), (_, _, y, u, v, _, _) in zip(dg.planes, b.frames()):
        ty.copy_(torch.from_numpy(y))
        tu.copy_(torch.from_numpy(u))
        tv.copy_(torch.from_numpy(v))
    torch.cuda.synchronize()
    batch.run()
    gpu.synchronize()
    compare("second content", dg, *expect(oracle, b.frames(), b.center, 4))
    batch.close()


def edge_tuples(w, h, n, seed, pad_y=0, pad_c=0):
    rng = np.random.default_rng(seed)
    cw, ch = (w + 1) // 2, (h + 1) // 2
    base = rng.integers(0, 256, (h, w)).astype(np.int64)
    out = []
    for i in range(n):
        y = np.clip(np.roll(base, (i - n // 2, 2 * (i - n // 2)), (0, 1)) + rng.integers(-6, 7, (h, w)), 0, 255)
        ys = np.full((h, w + pad_y), 3, np.uint8)
        ys[:, :w] = y
        us = np.full((ch, cw + pad_c), 5, np.uint8)
        us[:, :cw] = rng.integers(0, 256, (ch, cw))
        vs = np.full((ch, cw + pad_c), 7, np.uint8)
        vs[:, :cw] = rng.integers(0, 256, (ch, cw))
        out.append((w, h, ys.reshape(-1), us.reshape(-1), vs.reshape(-1), w + pad_y, cw + pad_c))
    return out


EDGE_JOBS = [  # (w, h, n, center, strength, pad_y, pad_c)
    (1, 1, 3, 1, 3, 0, 0), (2, 3, 2, 0, 6, 0, 0), (15, 33, 3, 2, 1, 5, 3), (17, 9, 4, 3, 9, 0, 0),
    (64, 40, 1, 0, 3, 0, 0), (37, 29, 5, 2, 0, 0, 0), (100, 70, 7, 3, 3, 13, 5),
    (33, 47, 11, 5, 2, 0, 1), (176, 144, 9, 4, 5, 16, 8), (48, 32, 64, 40, 3, 0, 0),
]


def test_batch_edge_geometries_and_shapes(gpu, oracle):
    """Tiny, ragged, unaligned, single-frame, strength-0 and 64-frame groups
    in one batch, under every launch shape."""
    import torch
    import vp8g
    tuples = [edge_tuples(w, h, n, 31 * i + 7, py, pc) for i, (w, h, n, _, _, py, pc) in enumerate(EDGE_JOBS)]
    wants = [expect(oracle, t, j[3], j[4]) for t, j in zip(tuples, EDGE_JOBS)]
    groups = [DeviceGroup(torch, t, j[3], j[4]) for t, j in zip(tuples, EDGE_JOBS)]
    torch.cuda.synchronize()
    for mbs, waves, wgs in ((0, 0, 0), (1, 1, 1), (2, 3, 2), (8, 4, 0), (4, 16, 1), (8, 2, 32)):
        gpu.apply_tuning({"mbs": mbs, "waves": waves, "wgs_per_cu": wgs})
        try:
            batch = gpu.create_arnr_batch([dg.job(vp8g) for dg in groups])
        finally:
            gpu.apply_tuning({})
        for dg in groups:
            dg.oy.fill_(0x5A)
            dg.mv.zero_()
        torch.cuda.synchronize()
        batch.run()
        gpu.synchronize()
        for i, (dg, (want, mvs)) in enumerate(zip(groups, wants)):
            compare(f"shape {(mbs, waves, wgs)} {batch.shape()} job {i} {EDGE_JOBS[i]}", dg, want, mvs)
        batch.close()


def test_batch_validation_names_the_failing_job(gpu):
    import torch
    import vp8g
    good = DeviceGroup(torch, edge_tuples(32, 32, 3, 1), 1, 3)
    other = DeviceGroup(torch, edge_tuples(16, 16, 3, 2), 1, 3)
    mixed = vp8g.ArnrJob(good.frames[:2] + other.frames[:1], 0, 3, good.oy.data_ptr(),
                         good.ou.data_ptr(), good.ov.data_ptr())
    with pytest.raises(vp8g.ReferenceDimensionsMismatch) as e:
        gpu.create_arnr_batch([good.job(vp8g), mixed])
    assert e.value.job == 1 and e.value.reference == (16, 16)
    bad_center = vp8g.ArnrJob(good.frames, 3, 3, good.oy.data_ptr(), good.ou.data_ptr(), good.ov.data_ptr())
    with pytest.raises(IndexError) as e:
        gpu.create_arnr_batch([good.job(vp8g), good.job(vp8g), bad_center])
    assert e.value.job == 2
    zero = vp8g.ArnrJob([(0, 8) + good.frames[0][2:]], 0, 3, good.oy.data_ptr(), good.ou.data_ptr(),
                        good.ov.data_ptr())
    with pytest.raises(vp8g.InvalidDimensions) as e:
        gpu.create_arnr_batch([zero])
    assert e.value.job == 0
    with pytest.raises(vp8g.Vp8gError):
        gpu.create_arnr_batch([])
