"""The 4K (c4) pipeline on the GPU against the oracle, stage by stage.

Content model of the c4 workload: region motion, occluders, sigma-4
noise, 11-frame groups.  At an oracle-sized frame (640x368, velocities
[-16, 16] px/frame, 8 occluders) every stage is compared with the oracle
bit-exactly:

* the two pyramid levels (2x2 rounded means, clamped), against numpy;
* the exact +-2 search of every level around that level's seeds
  (proposals at the top, the coarser result doubled below) against
  orc_block_search_group with the same seeds, and the seed upsampling rule;
* the exhaustive +-64 search (on a 3-frame subgroup: the first, centre and
  last frames, 5 frames apart) against the oracle;
* half_pixel_refine_luma then quarter_pixel_refine_luma of every macroblock
  (src/motion_search.rs:564-728) and the six-tap predict_inter_mb at the
  refined vectors (src/motion_comp.rs:1986-2078) against the oracle, and
  the reported residual against the sum of the oracle predictions' SADs.

At the full 3840x2160, N = 11, velocities [-48, 48], 40 occluders, the
size-independent properties: every refinement never raises the SAD, the
reported residual equals the sum of the quarter-pel SADs (two different
kernels), and the +-64 and +-2 results of sampled blocks equal the
oracle's."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def np_down2(p, w, h):
    """2x2 rounded mean with clamped coordinates: one pyramid level."""
    p = p.reshape(h, w).astype(np.int64)
    ow, oh = (w + 1) // 2, (h + 1) // 2
    ys, xs = np.arange(oh), np.arange(ow)
    y0, y1 = 2 * ys, np.minimum(2 * ys + 1, h - 1)
    x0, x1 = 2 * xs, np.minimum(2 * xs + 1, w - 1)
    s = p[y0][:, x0] + p[y0][:, x1] + p[y1][:, x0] + p[y1][:, x1]
    return ((s + 2) >> 2).astype(np.uint8).reshape(-1)


def np_upsample(coarse, fbh, fbw):
    n, cbh, cbw, _ = coarse.shape
    r = np.minimum(np.arange(fbh) // 2, cbh - 1)
    c = np.minimum(np.arange(fbw) // 2, cbw - 1)
    out = np.zeros((n, fbh, fbw, 3), np.int32)
    out[..., :2] = 2 * coarse[:, r][:, :, c, :2]
    return out


def level_tuples(lv, n):
    w, h = lv["w"], lv["h"]
    cw = (w + 1) // 2
    if "planes" not in lv:
        return None
    return [(w, h, y.cpu().numpy(), u.cpu().numpy(), v.cpu().numpy(), w, cw) for y, u, v in lv["planes"]]


def run_pipeline(ctx, g, record=True):
    import torch
    from vp8g.motion import Pyramid, subpel_and_mc
    from vp8g.proposer import WEIGHTS_DIR, DeviceGroup, Proposer
    d = DeviceGroup(torch, g)
    pyr = Pyramid(torch, ctx, d, levels=2)
    prop = Proposer(ctx, params=Proposer.load_params(WEIGHTS_DIR / "proposer_4k_l2.safetensors"))
    pyr.build()
    pyr.search(prop, g.center, 2)
    ctx.synchronize()
    prop.close()
    stream = torch.cuda.ExternalStream(ctx.stream())
    sub = subpel_and_mc(ctx, d.planes, g.width, g.height, g.center, pyr.result, stream, record=record)
    ctx.synchronize()
    return d, pyr, sub


def centre_mbs_np(y, w, h):
    bw, bh = w // 16, h // 16
    return np.ascontiguousarray(y.reshape(h, w).reshape(bh, 16, bw, 16).transpose(0, 2, 1, 3).reshape(bh * bw, 256))


def check_subpel_and_mc(oracle, g, table, sub, mbs=None):
    """Per non-centre frame: half then quarter ladders and the six-tap MC of
    the macroblocks `mbs` (all if None) against the oracle."""
    from oracle_ctypes import LumaRef, MvS, RefPlanes
    w, h = g.width, g.height
    bw, bh = w // 16, h // 16
    src = centre_mbs_np(g.y[g.center], w, h)
    idx = np.arange(bw * bh) if mbs is None else np.asarray(mbs)
    total = 0
    for rec in sub["frames"]:
        f = rec["frame"]
        y = np.ascontiguousarray(g.y[f]).reshape(-1)
        u = np.ascontiguousarray(g.u[f]).reshape(-1)
        v = np.ascontiguousarray(g.v[f]).reshape(-1)
        lr = LumaRef(y.ctypes.data, w, w, h)
        t = table[f].reshape(-1, 3)
        for i in idx:
            mbc, mbr = int(i % bw), int(i // bw)
            s = np.ascontiguousarray(src[i])
            hr = oracle.lib.orc_half_pixel_refine_luma(C.byref(lr), mbc, mbr, s.ctypes.data,
                                                       MvS(int(4 * t[i, 1]), int(4 * t[i, 0])))
            assert (rec["half"][i, 0], rec["half"][i, 1], rec["half"][i, 2]) == (hr.mv.row, hr.mv.col, hr.sad), \
                f"half-pel ladder frame {f} MB {i}"
            qr = oracle.lib.orc_quarter_pixel_refine_luma(C.byref(lr), mbc, mbr, s.ctypes.data, hr.mv)
            assert (rec["quarter"][i, 0], rec["quarter"][i, 1], rec["quarter"][i, 2]) == (qr.mv.row, qr.mv.col,
                                                                                       qr.sad), \
                f"quarter-pel ladder frame {f} MB {i}"
        rp = RefPlanes(y.ctypes.data, u.ctypes.data, v.ctypes.data, w, w // 2, bw, bh)
        xy = np.stack([idx % bw, idx // bw], 1).astype(np.uint32)
        mv = np.ascontiguousarray(rec["quarter"][idx, :2].astype(np.int16))
        want = np.zeros((len(idx), 384), np.uint8)
        f6 = np.ascontiguousarray(oracle.SIXTAP)
        oracle.lib.orc_predict_inter_mb_batch(C.byref(rp), xy.ctypes.data, mv.ctypes.data, len(idx), 0,
                                              f6.ctypes.data, want.ctypes.data)
        assert np.array_equal(rec["pred"][idx], want), f"predict_inter_mb frame {f}"
        total += int(np.abs(want[:, :256].astype(np.int64) - src[idx].astype(np.int64)).sum())
    return total


def test_pipeline_stages_match_oracle_at_640x368(gpu, oracle):
    import torch
    import vp8g.synth as synth
    g = synth.arnr_group_regions(640, 368, 11, vmax=16, occluders=8, clip=5, config=4)
    n, c = g.n, g.center
    d, pyr, sub = run_pipeline(gpu, g)
    # pyramid levels
    prev = [(np.ascontiguousarray(y).reshape(-1), np.ascontiguousarray(u).reshape(-1),
             np.ascontiguousarray(v).reshape(-1)) for y, u, v in zip(g.y, g.u, g.v)]
    pw, ph = g.width, g.height
    for k in (1, 2):
        lv = pyr.levels[k]
        tup = level_tuples(lv, n)
        for f in range(n):
            cwp, chp = (pw + 1) // 2, (ph + 1) // 2
            want = (np_down2(prev[f][0], pw, ph), np_down2(prev[f][1], cwp, chp), np_down2(prev[f][2], cwp, chp))
            for name, got, exp in zip("yuv", tup[f][2:5], want):
                assert np.array_equal(got, exp), f"level {k} frame {f} plane {name}"
        prev = [t[2:5] for t in tup]
        pw, ph = lv["w"], lv["h"]
    # +-2 at every level around that level's seeds; seed upsampling below the top
    for k in (2, 1, 0):
        lv = pyr.levels[k]
        tup = level_tuples(lv, n) if k else g.frames()
        seeds = lv["seeds"].cpu().numpy()
        got = lv["out"].cpu().numpy()
        want = oracle.block_search_group(tup, c, 2, seeds)
        others = [i for i in range(n) if i != c]
        assert np.array_equal(got[others], want[others]), f"+-2 search at level {k}"
        if k < 2:
            up = np_upsample(pyr.levels[k + 1]["out"].cpu().numpy(), lv["bh"], lv["bw"])
            assert np.array_equal(seeds[others, ..., :2], up[others, ..., :2]), f"seed upsampling to level {k}"
    # exhaustive +-64 on the first, centre and last frames
    fr = g.frames()
    sub3 = [fr[0], fr[c], fr[-1]]
    import vp8g
    descs = (vp8g._I420 * 3)()
    for i, f in enumerate((0, c, n - 1)):
        y, u, v = d.planes[f]
        descs[i] = vp8g._I420(g.width, g.height, y.data_ptr(), u.data_ptr(), v.data_ptr(), g.width,
                              (g.width + 1) // 2)
    ex = torch.full((3, d.bh, d.bw, 3), -7, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    gpu.block_search_dev(descs, 3, 1, 64, ex.data_ptr())
    gpu.synchronize()
    want = oracle.block_search_group(sub3, 1, 64)
    got = ex.cpu().numpy()
    assert np.array_equal(got[[0, 2]], want[[0, 2]]), "exhaustive +-64"
    # sub-pel ladders and MC at the pyramid's vectors, and the residual
    total = check_subpel_and_mc(oracle, g, pyr.result.cpu().numpy(), sub)
    assert total == sub["residual_sad"]


@pytest.fixture(scope="module")
def group_4k():
    import vp8g.synth as synth
    return synth.arnr_group_regions(3840, 2160, 11, vmax=48, occluders=40, clip=0, config=4)


def test_pipeline_properties_at_3840x2160(gpu, oracle, group_4k):
    import torch
    g = group_4k
    n, c = g.n, g.center
    d, pyr, sub = run_pipeline(gpu, g)
    table = pyr.result.cpu().numpy()
    others = [i for i in range(n) if i != c]
    # each refinement keeps or lowers the SAD: whole-pel (table) >= half >= quarter
    qsum = 0
    for rec in sub["frames"]:
        whole = table[rec["frame"]].reshape(-1, 3)[:, 2]
        assert (rec["half"][:, 2] <= whole).all(), f"half-pel SAD above whole-pel, frame {rec['frame']}"
        assert (rec["quarter"][:, 2] <= rec["half"][:, 2]).all(), f"quarter above half, frame {rec['frame']}"
        qsum += int(rec["quarter"][:, 2].sum())
    # the residual of the MC kernel equals the sum of the ladder kernel's SADs
    assert qsum == sub["residual_sad"]
    # sampled blocks: the exhaustive +-64 and the level-0 +-2 search against the oracle
    ex = d.table(torch)
    torch.cuda.synchronize()
    gpu.block_search_dev(d.descs, n, c, 64, ex.data_ptr())
    gpu.synchronize()
    ex_np = ex.cpu().numpy()
    seeds = pyr.levels[0]["seeds"].cpu().numpy()
    keep = []
    ctr = oracle.i420(*g.frames()[c], keep=keep)
    rng = np.random.default_rng(44)
    for f in (0, 4, 6, 10):
        fr = oracle.i420(*g.frames()[f], keep=keep)
        for _ in range(12):
            r, col = int(rng.integers(0, d.bh)), int(rng.integers(0, d.bw))
            out = (C.c_int32 * 3)()
            oracle.lib.orc_block_search(C.byref(ctr), C.byref(fr), 16 * col, 16 * r, 16, 16, 0, 0, 64, out)
            assert tuple(out) == tuple(ex_np[f, r, col]), f"+-64 frame {f} block ({r}, {col})"
            sx, sy = int(seeds[f, r, col, 0]), int(seeds[f, r, col, 1])
            oracle.lib.orc_block_search(C.byref(ctr), C.byref(fr), 16 * col, 16 * r, 16, 16, sx, sy, 2, out)
            assert tuple(out) == tuple(table[f, r, col]), f"+-2 frame {f} block ({r}, {col})"
    # the learned pyramid finds the exhaustive minimum on most matched blocks
    from vp8g.motion import block_area
    e, p = ex_np[others], table[others]
    matched = e[..., 2] < 13 * block_area(g.width, g.height)[None]
    agree = float((p[..., 2] == e[..., 2])[matched].mean())
    print({"agreement_matched": agree, "matched_fraction": float(matched.mean())})
    assert agree >= 0.7
