"""Host mirror of the lagged altref stream driver's lookahead grouping on
the GPU (Vp8AltrefStreamEncoder::push_frame / finish / flush_group,
src/stream.rs:2740-2957), without the bitstream encoder around it.

The grouping decisions are the reference's, made on the host in push
order: frame dimensions lock at the first frame (DimensionsChanged, as
StreamEncodeError::DimensionsChanged); a frame whose luma MAD against the
previous pushed frame exceeds ``scene_cut_mad_threshold`` (luma_mad,
src/stream.rs:2654-2666) or whose index is a scheduled key closes the
running group before it enters, and a scene cut keys the next group; a
group closes when it holds ``altref_window`` frames; ``finish`` drains
the tail.  A group of at least two frames gets an invisible anchor,
build_arnr_altref_threaded(group, len - 1, arnr) (src/stream.rs:2888-2896).

What the GPU does: the MADs of a pushed chunk are one launch
(vp8g_luma_sad_batch_dev), and the anchors of every group a chunk closes
are one persistent, graph-captured launch (vp8g_arnr_batch_*), so a
stream's groups reach the device together instead of one call per group.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import (_I420, ArnrConfig, ArnrJob, ArnrPicture, Context, I420Frame, VP8G_OK, _ptr)


@dataclass(frozen=True)
class AltrefStreamConfig:
    """The grouping fields of AltrefStreamConfig (src/stream.rs:2111-2150,
    defaults :2463-2482)."""
    keyframe_interval: int = 0
    altref_window: int = 8
    arnr: ArnrConfig = field(default_factory=ArnrConfig)
    scene_cut_mad_threshold: float = 40.0


class DimensionsChanged(Exception):
    """StreamEncodeError::DimensionsChanged { first, got }."""

    def __init__(self, first, got):
        super().__init__(f"DimensionsChanged {{ first: {first}, got: {got} }}")
        self.first, self.got = first, got


@dataclass
class AltrefGroup:
    """One flushed lookahead group: source index of its first frame, its
    frame count, whether it opens with a key frame (scheduled, forced by a
    scene cut, or the stream's first), and the anchor's altref (None for a
    single-frame group)."""
    start: int
    length: int
    key: bool
    forced_key: bool
    anchor: ArnrPicture | None = None


class AltrefStreamBatcher:
    """Lookahead grouping of a stream with the anchors of every closed group
    built on the GPU in one batched launch per pushed chunk."""

    def __init__(self, ctx: Context, config: AltrefStreamConfig = AltrefStreamConfig()):
        if config.altref_window == 0:
            raise ValueError("altref_window must be at least 1 (Vp8AltrefStreamEncoder::new returns None)")
        import torch
        self.torch = torch
        self.ctx, self.config = ctx, config
        self.dims = None
        self.pending: list = []          # (host frame, device planes) of the running group
        self.pending_start = 0
        self.pending_forced_key = False
        self.last_pushed = None          # device planes of the previous pushed frame
        self.input_count = 0
        self.has_reference = False       # the reference's `last` slot is filled after the first group

    # -- device copies
    def _upload(self, f: I420Frame):
        torch = self.torch
        w, h = f.width, f.height
        cw, ch = (w + 1) // 2, (h + 1) // 2

        def plane(a, pw, ph, stride):
            a = np.ascontiguousarray(a, np.uint8).reshape(-1)
            rows = np.lib.stride_tricks.as_strided(a, (ph, pw), (stride, 1)) if ph else a[:0]
            return torch.from_numpy(np.ascontiguousarray(rows).reshape(-1)).cuda()

        y = plane(f.y, w, h, f.y_stride)
        u = plane(f.u, cw, ch, f.uv_stride)
        v = plane(f.v, cw, ch, f.uv_stride)
        return (w, h, y, u, v)

    @staticmethod
    def _desc(d):
        w, h, y, u, v = d
        return _I420(w, h, y.data_ptr(), u.data_ptr(), v.data_ptr(), w, (w + 1) // 2)

    def _mads(self, devs) -> list:
        """luma_mad of each uploaded frame against the frame pushed before it
        (None for the stream's first)."""
        torch = self.torch
        prevs = ([self.last_pushed] if self.last_pushed is not None else []) + devs[:-1]
        cur = devs if self.last_pushed is not None else devs[1:]
        out = [] if self.last_pushed is not None else [None]
        if cur:
            n = len(cur)
            pa, ca = (_I420 * n)(), (_I420 * n)()
            for i in range(n):
                pa[i], ca[i] = self._desc(prevs[i]), self._desc(cur[i])
            sad = torch.zeros(n, dtype=torch.int64, device="cuda")
            torch.cuda.synchronize()
            st = self.ctx.lib.vp8g_luma_sad_batch_dev(self.ctx.h, pa, ca, n, sad.data_ptr())
            self.ctx._raise(st, "vp8g_luma_sad_batch_dev")
            self.ctx.synchronize()
            w, h = cur[0][0], cur[0][1]
            out += [float(s) / float(w * h) for s in sad.cpu().numpy().astype(np.uint64)]
        return out

    def _is_key_index(self, index: int) -> bool:
        k = self.config.keyframe_interval
        return index == 0 or (k > 0 and index % k == 0)

    def _flush_group(self, closed: list) -> None:
        group = self.pending
        start, forced = self.pending_start, self.pending_forced_key
        self.pending, self.pending_forced_key = [], False
        key = self._is_key_index(start) or forced or not self.has_reference
        self.has_reference = True
        closed.append((AltrefGroup(start, len(group), key, forced), [d for _, d in group]))

    # -- the reference's entry points
    def push_frames(self, frames) -> list:
        """Push frames in order; returns the groups they close, with anchors.
        Dimensions are checked for the whole chunk before any frame enters."""
        frames = list(frames)
        for f in frames:
            dims = (f.width, f.height)
            if self.dims is None:
                self.dims = dims
            elif dims != self.dims:
                raise DimensionsChanged(self.dims, dims)
        devs = [self._upload(f) for f in frames]
        mads = self._mads(devs)
        closed: list = []
        thr = self.config.scene_cut_mad_threshold
        for f, d, mad in zip(frames, devs, mads):
            scene_cut = thr > 0.0 and mad is not None and mad > thr
            if (scene_cut or self._is_key_index(self.input_count)) and self.pending:
                self._flush_group(closed)
            if not self.pending:
                self.pending_start = self.input_count
                self.pending_forced_key = scene_cut
            self.pending.append((f, d))
            self.last_pushed = d
            self.input_count += 1
            if len(self.pending) >= self.config.altref_window:
                self._flush_group(closed)
        return self._anchors(closed)

    def push_frame(self, frame: I420Frame) -> list:
        return self.push_frames([frame])

    def finish(self) -> list:
        """Drain the running group (idempotent)."""
        closed: list = []
        if self.pending:
            self._flush_group(closed)
        return self._anchors(closed)

    def _anchors(self, closed) -> list:
        """Build every anchored group's altref in one batched launch."""
        torch = self.torch
        jobs, outs = [], []
        for g, devs in closed:
            if g.length < 2:
                continue
            w, h = devs[0][0], devs[0][1]
            cw, ch = (w + 1) // 2, (h + 1) // 2
            o = (torch.empty(w * h, dtype=torch.uint8, device="cuda"),
                 torch.empty(cw * ch, dtype=torch.uint8, device="cuda"),
                 torch.empty(cw * ch, dtype=torch.uint8, device="cuda"))
            frames = [(dw, dh, y.data_ptr(), u.data_ptr(), v.data_ptr(), dw, (dw + 1) // 2)
                      for dw, dh, y, u, v in devs]
            jobs.append(ArnrJob(frames, len(devs) - 1, int(self.config.arnr.strength), o[0].data_ptr(),
                                o[1].data_ptr(), o[2].data_ptr()))
            outs.append((g, o, w, h))
        if jobs:
            torch.cuda.synchronize()
            batch = self.ctx.create_arnr_batch(jobs)
            batch.run()
            self.ctx.synchronize()
            batch.close()
            for g, (oy, ou, ov), w, h in outs:
                g.anchor = ArnrPicture(w, h, oy.cpu().numpy(), ou.cpu().numpy(), ov.cpu().numpy())
        return [g for g, _ in closed]


def luma_mad_batch(ctx: Context, prev_frames, cur_frames) -> np.ndarray:
    """luma_mad (src/stream.rs:2654-2666) of each (prev, cur) pair of host
    frames through vp8g_luma_mad_batch."""
    import ctypes as C
    prev_frames, cur_frames = list(prev_frames), list(cur_frames)
    n = len(prev_frames)
    pa, ca = (_I420 * max(n, 1))(), (_I420 * max(n, 1))()
    keep = []
    for arr, fs in ((pa, prev_frames), (ca, cur_frames)):
        for i, f in enumerate(fs):
            y = np.ascontiguousarray(f.y, np.uint8).reshape(-1)
            keep.append(y)
            arr[i] = _I420(f.width, f.height, _ptr(y), _ptr(y), _ptr(y), f.y_stride, f.uv_stride)
    out = np.zeros(max(n, 1), np.float64)
    st = ctx.lib.vp8g_luma_mad_batch(ctx.h, pa, ca, n, out.ctypes.data_as(C.c_void_p))
    ctx._raise(st, "vp8g_luma_mad_batch")
    return out[:n]
