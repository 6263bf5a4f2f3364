"""Whole-pel search -> half/quarter-pel refinement -> six-tap motion
compensation for every block of every non-centre frame of a group, on the
device (the 4K pipeline: exhaustive +-64 or learned-seeded search, then
half_pixel_refine_luma / quarter_pixel_refine_luma, src/motion_search.rs:
564-728, then predict_inter_mb, src/motion_comp.rs:1986-2078).

The residual-energy proxy of the motion field is the SAD and SSE of the
motion-compensated luma against the centre frame's macroblocks.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import SIXTAP_FILTERS, Context


def centre_mbs(torch, y, width: int, height: int):
    """[mb_rows * mb_cols, 256] source macroblocks of a luma plane (edge
    replicated to whole macroblocks)."""
    bw, bh = (width + 15) // 16, (height + 15) // 16
    p = y.reshape(height, width)
    if bw * 16 != width or bh * 16 != height:
        p = torch.nn.functional.pad(p[None, None].float(), (0, bw * 16 - width, 0, bh * 16 - height),
                                    mode="replicate")[0, 0].to(torch.uint8)
    return p.reshape(bh, 16, bw, 16).permute(0, 2, 1, 3).reshape(bh * bw, 256).contiguous()


def subpel_and_mc(ctx: Context, planes, width: int, height: int, center: int, table, stream,
                  record: bool = False) -> dict:
    """For every non-centre frame f: the table's whole-pel vectors refined to
    half then quarter pel against frame f, then the six-tap prediction of
    every macroblock from frame f at the refined vector.  Frames must be
    whole macroblocks (width, height multiples of 16), as ReferencePlanes
    requires.  Returns timings and the residual proxy; with `record`, also
    "frames": per non-centre frame the host copies of the half- and
    quarter-pel results ([n_mb] x {row, col, sad}) and the predictions
    ([n_mb, 384])."""
    import torch
    assert width % 16 == 0 and height % 16 == 0
    dev = torch.device("cuda:0")
    bw, bh = width // 16, height // 16
    n_mb = bw * bh
    cw = width // 2
    with torch.cuda.stream(stream):
        src = centre_mbs(torch, planes[center][0], width, height)
        rr, cc = torch.meshgrid(torch.arange(bh, device=dev), torch.arange(bw, device=dev), indexing="ij")
        mb_xy = torch.stack([cc.reshape(-1), rr.reshape(-1)], 1).to(torch.int32).contiguous()
        half = torch.empty((n_mb, 2), dtype=torch.int32, device=dev)
        quarter = torch.empty((n_mb, 2), dtype=torch.int32, device=dev)
        pred = torch.empty((n_mb, 384), dtype=torch.uint8, device=dev)
        centers = torch.empty((n_mb, 2), dtype=torch.int16, device=dev)
    sad_total = sse_total = 0
    t_refine = t_mc = 0.0
    recorded = []

    def result_rows(t):
        v = t.view(torch.int16).reshape(n_mb, 4).cpu().numpy()
        sad = v[:, 2].astype(np.uint16).astype(np.int64) | (v[:, 3].astype(np.uint16).astype(np.int64) << 16)
        return np.stack([v[:, 0].astype(np.int64), v[:, 1].astype(np.int64), sad], 1)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    for f in range(len(planes)):
        if f == center:
            continue
        y, u, v = planes[f]
        with torch.cuda.stream(stream):
            t = table[f].reshape(n_mb, 3)
            centers[:, 0] = (t[:, 1] * 4).to(torch.int16)   # row = 4 * mvy (quarter pel)
            centers[:, 1] = (t[:, 0] * 4).to(torch.int16)   # col = 4 * mvx
        ev[0].record(stream)
        ctx.refine_batch_dev(y.data_ptr(), width, width, height, mb_xy.data_ptr(), src.data_ptr(),
                             centers.data_ptr(), n_mb, half.data_ptr(), quarter=False)
        with torch.cuda.stream(stream):
            # vp8g_search_result = {int16 row, int16 col, uint32 sad}
            centers.copy_(half.view(torch.int16).reshape(n_mb, 4)[:, :2])
        if record:
            recorded.append({"frame": f, "half": None})
            torch.cuda.synchronize()
            recorded[-1]["half"] = result_rows(half)
        ctx.refine_batch_dev(y.data_ptr(), width, width, height, mb_xy.data_ptr(), src.data_ptr(),
                             centers.data_ptr(), n_mb, quarter.data_ptr(), quarter=True)
        with torch.cuda.stream(stream):
            centers.copy_(quarter.view(torch.int16).reshape(n_mb, 4)[:, :2])
        ev[1].record(stream)
        ctx.predict_inter_mb_batch_dev(y.data_ptr(), u.data_ptr(), v.data_ptr(), width, cw, bw, bh,
                                       mb_xy.data_ptr(), centers.data_ptr(), n_mb, False, SIXTAP_FILTERS,
                                       pred.data_ptr())
        ev[2].record(stream)
        with torch.cuda.stream(stream):
            d = pred[:, :256].to(torch.int32) - src.to(torch.int32)
            sad_total += int(d.abs().sum().item())
            sse_total += int((d * d).sum().item())
        t_refine += ev[0].elapsed_time(ev[1])
        t_mc += ev[1].elapsed_time(ev[2])
        if record:
            recorded[-1]["quarter"] = result_rows(quarter)
            recorded[-1]["pred"] = pred.cpu().numpy()
    pairs = len(planes) - 1
    out = {"ms_subpel": t_refine, "ms_mc": t_mc, "residual_sad": sad_total, "residual_sse": sse_total,
           "residual_sad_per_px": sad_total / (pairs * n_mb * 256.0),
           "mc_bytes": pairs * n_mb * (384 + 384 + 4)}
    if record:
        out["frames"] = recorded
    return out


def block_area(width: int, height: int) -> np.ndarray:
    """[brows, bcols] pixel count of every 16x16 block (partial edge blocks
    smaller), for the occlusion cut-off sad / (bw * bh) > 12
    (src/arnr.rs:603-608)."""
    bw, bh = (width + 15) // 16, (height + 15) // 16
    return (np.minimum(16, height - 16 * np.arange(bh))[:, None] *
            np.minimum(16, width - 16 * np.arange(bw))[None, :])


class Pyramid:
    """Coarser levels of a device group for coarse-to-fine search: level
    k + 1 is vp8g_downsample2_i420_dev of level k (level 0 = the group).
    Each level holds its frames' planes, descriptors, block grid and two
    alignment tables (seeds and search result)."""

    def __init__(self, torch, ctx: Context, d, levels: int = 2):
        from . import _I420
        self.ctx, self.n = ctx, d.g.n
        dev = torch.device("cuda:0")
        w, h = d.g.width, d.g.height
        self.levels = [{"w": w, "h": h, "descs": d.descs, "bw": d.bw, "bh": d.bh}]
        for _ in range(levels):
            w, h = (w + 1) // 2, (h + 1) // 2
            cw, ch = (w + 1) // 2, (h + 1) // 2
            planes = [(torch.empty(w * h, dtype=torch.uint8, device=dev),
                       torch.empty(cw * ch, dtype=torch.uint8, device=dev),
                       torch.empty(cw * ch, dtype=torch.uint8, device=dev)) for _ in range(self.n)]
            descs = (_I420 * self.n)()
            for i, (y, u, v) in enumerate(planes):
                descs[i] = _I420(w, h, y.data_ptr(), u.data_ptr(), v.data_ptr(), w, cw)
            self.levels.append({"w": w, "h": h, "descs": descs, "planes": planes,
                                "bw": (w + 15) // 16, "bh": (h + 15) // 16})
        for lv in self.levels:
            shape = (self.n, lv["bh"], lv["bw"], 3)
            lv["seeds"] = torch.zeros(shape, dtype=torch.int32, device=dev)
            lv["out"] = torch.zeros(shape, dtype=torch.int32, device=dev)

    def build(self) -> None:
        """Enqueue the downsampling of every frame to every level."""
        for k in range(1, len(self.levels)):
            self.ctx.downsample2_dev(self.levels[k - 1]["descs"], self.levels[k]["descs"], self.n)

    def search(self, prop, center: int, refine: int = 2) -> None:
        """Proposals at the coarsest level, then at every level from the
        coarsest to level 0 an exact +-refine search around seeds (the
        proposals, then the coarser level's result with the vector doubled).
        The result is level 0's "out" table."""
        top = self.levels[-1]
        prop.propose(top["descs"], self.n, center, top["seeds"].data_ptr(), 16 - refine)
        for k in range(len(self.levels) - 1, -1, -1):
            lv = self.levels[k]
            if k < len(self.levels) - 1:
                up = self.levels[k + 1]
                self.ctx.upsample_table_dev(up["out"].data_ptr(), self.n, up["bw"], up["bh"], lv["seeds"].data_ptr(),
                                            lv["bw"], lv["bh"])
            self.ctx.block_search_dev(lv["descs"], self.n, center, refine, lv["out"].data_ptr(),
                                      lv["seeds"].data_ptr())

    @property
    def result(self):
        return self.levels[0]["out"]


def evaluate_group(ctx: Context, prop, g, radius: int = 64, refine: int = 2, reps: int = 3, pyramid_prop=None,
                   pyramid_levels: int = 2) -> dict:
    """Exhaustive +-radius vs learned-seeded (+-refine) whole-pel search on a
    host group, each followed by half/quarter-pel refinement and six-tap MC
    of every macroblock of every non-centre frame.  `prop` (may be None)
    proposes at full resolution; with `pyramid_prop` (a proposer trained at
    that pyramid level) the learned search runs coarse to fine (Pyramid:
    building the levels is included in its time)."""
    import torch
    from .proposer import DeviceGroup, psnr
    stream = torch.cuda.ExternalStream(ctx.stream())
    d = DeviceGroup(torch, g)
    ex, seeds, ref = d.table(torch), d.table(torch), d.table(torch)
    torch.cuda.synchronize()
    n, c = g.n, g.center

    def timed(fn):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    t_ex = timed(lambda: ctx.block_search_dev(d.descs, n, c, radius, ex.data_ptr()))
    ctx.synchronize()
    others = [i for i in range(n) if i != c]
    e = ex.cpu().numpy()[others]
    matched = e[..., 2] < 13 * block_area(g.width, g.height)[None]
    out = {"width": g.width, "height": g.height, "frames": n, "radius": radius,
           "matched_fraction": float(matched.mean()), "ms_search_exhaustive": t_ex}
    legs = [("exhaustive", ex, t_ex)]
    if prop is not None:
        t_pr = timed(lambda: prop.propose(d.descs, n, c, seeds.data_ptr(), radius - refine))
        t_rf = timed(lambda: ctx.block_search_dev(d.descs, n, c, refine, ref.data_ptr(), seeds.data_ptr()))
        ctx.synchronize()
        same = ref.cpu().numpy()[others][..., 2] == e[..., 2]
        out["single_level"] = {"agreement": float(same.mean()), "agreement_matched": float(same[matched].mean()),
                               "ms_search": t_pr + t_rf, "ms_propose": t_pr, "ms_refine": t_rf,
                               "search_speedup": t_ex / (t_pr + t_rf)}
        legs.append(("learned", ref, t_pr + t_rf))
    if pyramid_prop is not None:
        pyr = Pyramid(torch, ctx, d, levels=pyramid_levels)
        t_build = timed(pyr.build)
        t_psearch = timed(lambda: pyr.search(pyramid_prop, c, refine))
        ctx.synchronize()
        pr = pyr.result.cpu().numpy()[others]
        psame = pr[..., 2] == e[..., 2]
        out["pyramid"] = {"levels": pyramid_levels, "ms_build": t_build, "ms_search": t_psearch,
                          "agreement": float(psame.mean()), "agreement_matched": float(psame[matched].mean()),
                          "search_speedup": t_ex / (t_build + t_psearch)}
        legs.append(("learned_pyramid", pyr.result, t_build + t_psearch))
    anchors = d.bw * d.bh
    for name, tab, t_search in legs:
        r = subpel_and_mc(ctx, d.planes, g.width, g.height, c, tab, stream)
        total = t_search + r["ms_subpel"] + r["ms_mc"]
        r.update({"ms_total": total, "anchor_mb_per_s": anchors / (total * 1e-3),
                  "mc_gb_per_s": r["mc_bytes"] / (r["ms_mc"] * 1e-3) / 1e9})
        out[name] = r
    return out
