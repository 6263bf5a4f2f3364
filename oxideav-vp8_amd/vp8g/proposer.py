"""Learned motion proposer: host side of the gfx950 CNN (vp8g_proposer_* in
include/vp8g.h), its training data, training loop and evaluation.

The proposer has no counterpart in the reference crate.  It proposes one
whole-pel vector per 16x16 block of the centre frame for every other frame
of a group, from the 4x-downsampled (anchor, neighbour) luma pair; a +-2
exact search (vp8g_block_search_dev, radius 2) around each proposal then
replaces the +-16 exhaustive search, and the table-driven altref
(vp8g_build_arnr_altref_table_dev) blends with the refined vectors.

Training data: region-motion groups synthesised on the GPU
(vp8g_synth_luma_dev) with the test workloads' content model, labelled by
the exhaustive +-16 GPU search (ties to the shortest vector).
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import (_I420, Context, Vp8gError, VP8G_OK, _SynthRect, _SynthSpec)
from . import synth as _synth

FEAT = 128  # correlation-layer channels per block
LAYERS = ((1, FEAT, 32), (9, 32, 32), (9, 32, 64), (9, 64, 32), (9, 32, 32))  # (taps, cin, cout)
PARAM_COUNT = sum(t * ci * co + co for t, ci, co in LAYERS) + 66
WEIGHTS_DIR = Path(__file__).resolve().parent.parent / "weights"


def he_init(seed: int) -> np.ndarray:
    """He-normal conv and head weights (fp32), zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for t, ci, co in LAYERS:
        parts.append(rng.normal(0.0, math.sqrt(2.0 / (t * ci)), t * co * ci))
        parts.append(np.zeros(co))
    parts.append(rng.normal(0.0, math.sqrt(1.0 / 32), 64))
    parts.append(np.zeros(2))
    return np.concatenate(parts).astype(np.float32)


class Proposer:
    """One proposer (parameters, optimizer state, work arena) on a context."""

    def __init__(self, ctx: Context, params: np.ndarray | None = None, seed: int = 0):
        self.ctx, self.lib = ctx, ctx.lib
        h = C.c_void_p()
        ctx._raise(self.lib.vp8g_proposer_create(ctx.h, C.byref(h)), "vp8g_proposer_create")
        self.h = h
        self.set_params(he_init(seed) if params is None else params)

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.vp8g_proposer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- parameters
    def set_params(self, params: np.ndarray) -> None:
        p = np.ascontiguousarray(params, np.float32)
        if p.size != PARAM_COUNT:
            raise ValueError(f"expected {PARAM_COUNT} parameters, got {p.size}")
        self.ctx._raise(self.lib.vp8g_proposer_set_params(self.h, p.ctypes.data, p.size),
                        "vp8g_proposer_set_params")

    def params(self) -> np.ndarray:
        p = np.empty(PARAM_COUNT, np.float32)
        self.ctx._raise(self.lib.vp8g_proposer_get_params(self.h, p.ctypes.data, p.size),
                        "vp8g_proposer_get_params")
        return p

    def grads(self):
        g = np.empty(PARAM_COUNT, np.float32)
        loss = C.c_float()
        self.ctx._raise(self.lib.vp8g_proposer_get_grads(self.h, g.ctypes.data, g.size, C.byref(loss)),
                        "vp8g_proposer_get_grads")
        return g, float(loss.value)

    def save(self, path, meta: dict | None = None) -> None:
        from safetensors.numpy import save_file
        md = {"param_count": str(PARAM_COUNT),
              "layout": "conv [taps][cout][cin]+bias: 1x1 128-32, 3x3 32-32, 32-64, 64-32, 32-32; head [2][32]+[2]"}
        md.update({k: str(v) for k, v in (meta or {}).items()})
        save_file({"params": self.params()}, str(path), metadata=md)

    @staticmethod
    def load_params(path) -> np.ndarray:
        from safetensors.numpy import load_file
        return load_file(str(path))["params"].astype(np.float32)

    # ---- compute (device pointers; enqueue on the context's stream)
    def prepare(self, descs, n: int, center: int, planes_ptr: int) -> None:
        """Downsampled (anchor, neighbour) planes of the group's pairs."""
        self.ctx._raise(self.lib.vp8g_proposer_prepare_dev(self.h, descs, n, center, planes_ptr),
                        "vp8g_proposer_prepare_dev")

    def forward(self, planes_ptr: int, images: int, bw: int, bh: int, a4_ptr: int | None = None) -> None:
        self.ctx._raise(self.lib.vp8g_proposer_forward_dev(self.h, planes_ptr, images, bw, bh, a4_ptr),
                        "vp8g_proposer_forward_dev")

    def propose(self, descs, n: int, center: int, table_ptr: int, clamp: int = 14) -> None:
        self.ctx._raise(self.lib.vp8g_proposer_propose_dev(self.h, descs, n, center, clamp, table_ptr),
                        "vp8g_proposer_propose_dev")

    def backward(self, planes_ptr: int, labels_ptr: int, images: int, bw: int, bh: int) -> None:
        self.ctx._raise(self.lib.vp8g_proposer_backward_dev(self.h, planes_ptr, labels_ptr, images, bw, bh),
                        "vp8g_proposer_backward_dev")

    def train_step(self, planes_ptr: int, labels_ptr: int, images: int, bw: int, bh: int, lr: float,
                   loss_ptr: int | None = None) -> None:
        self.ctx._raise(self.lib.vp8g_proposer_train_step_dev(self.h, planes_ptr, labels_ptr, images, bw, bh,
                                                              lr, loss_ptr),
                        "vp8g_proposer_train_step_dev")


# ------------------------------------------------------------ GPU content
def regions_spec(width: int, height: int, n: int, vmax: int, occluders: int, clip: int, config: int):
    """The parameters synth.arnr_group_regions draws for a group's luma."""
    base = _synth._seed(config, clip, 0)
    vel = _synth.uniform_int(base + 4, 128, -vmax, vmax).reshape(64, 2).astype(np.int32)
    rects = []
    for t in range(n):
        seed = base + 500 + 7 * t
        r = _synth.uniform_int(seed, 4 * occluders, 0, 1 << 30).reshape(occluders, 4) if occluders else []
        for k in range(occluders):
            rw = min(max(1, 16 + int(r[k, 0]) % 81), width)
            rh = min(max(1, 16 + int(r[k, 1]) % 81), height)
            x0 = int(r[k, 2]) % (width - rw + 1)
            y0 = int(r[k, 3]) % (height - rh + 1)
            rects.append(((seed * 131 + k) & 0xFFFFFFFFFFFFFFFF, x0, y0, rw, rh))
    noise = np.array([(base + 100 + 3 * t) & 0xFFFFFFFFFFFFFFFF for t in range(n)], np.uint64)
    return base + 1, vel, rects, noise


def synth_group_luma(ctx: Context, out_ptr: int, clean_ptr: int | None, width: int, height: int, n: int,
                     vmax: int = 12, occluders: int = 20, clip: int = 0, config: int = 1,
                     sigma: float = 4.0) -> None:
    """Luma planes of synth.arnr_group_regions(...) generated on the device
    into [n][h][w] at out_ptr (and the noise-free planes at clean_ptr)."""
    tex, vel, rects, noise = regions_spec(width, height, n, vmax, occluders, clip, config)
    r = (_SynthRect * max(len(rects), 1))()
    for i, (s, x0, y0, w, h) in enumerate(rects):
        r[i] = _SynthRect(s, x0, y0, w, h)
    vel = np.ascontiguousarray(vel)
    spec = _SynthSpec(width, height, n, occluders, tex & 0xFFFFFFFFFFFFFFFF, vel.ctypes.data,
                      C.cast(r, C.c_void_p), noise.ctypes.data, sigma)
    ctx._raise(ctx.lib.vp8g_synth_luma_dev(ctx.h, C.byref(spec), out_ptr, clean_ptr), "vp8g_synth_luma_dev")


def luma_descs(torch, frames, width: int, height: int):
    """vp8g_i420 descriptors over a device [n][h][w] luma tensor (the chroma
    pointers reference a shared dummy plane; no luma-only entry reads them)."""
    cw, ch = (width + 1) // 2, (height + 1) // 2
    dummy = torch.zeros(cw * ch, dtype=torch.uint8, device=frames.device)
    n = frames.shape[0]
    d = (_I420 * n)()
    for i in range(n):
        d[i] = _I420(width, height, frames[i].data_ptr(), dummy.data_ptr(), dummy.data_ptr(), width, cw)
    return d, dummy


@dataclass
class Dataset:
    inputs: object   # torch uint8 [pairs, 2, 4 bh, 4 bw] downsampled (anchor, neighbour) planes
    labels: object   # torch int32 [pairs, cells, 2]
    sads: object     # torch int32 [pairs, cells] exhaustive minimum SAD
    bw: int
    bh: int


def build_dataset(ctx: Context, prop: Proposer, groups: int, width: int = 1920, height: int = 1080,
                  n: int = 7, vmax: int = 12, occluders: int = 20, config: int = 30, clip0: int = 0,
                  radius: int = 16, log=None, level: int = 0) -> Dataset:
    """`groups` synthetic groups -> (n - 1) * groups labelled pairs.  With
    level > 0 the groups are generated at width x height and the pairs and
    labels are taken `level` pyramid levels down (vp8g_downsample2_i420_dev
    applied `level` times), for a proposer that runs at that level."""
    import torch
    dev = torch.device("cuda:0")
    full_w, full_h = width, height
    for _ in range(level):
        width, height = (width + 1) // 2, (height + 1) // 2
    bw, bh = (width + 15) // 16, (height + 15) // 16
    cells = bw * bh
    pairs = groups * (n - 1)
    inputs = torch.empty((pairs, 2, 4 * bh, 4 * bw), dtype=torch.uint8, device=dev)
    labels = torch.empty((pairs, cells, 2), dtype=torch.int32, device=dev)
    sads = torch.empty((pairs, cells), dtype=torch.int32, device=dev)
    chain = []  # (frames, descs, dummy) from full resolution down to the level
    w, h = full_w, full_h
    for k in range(level + 1):
        f = torch.empty((n, h, w), dtype=torch.uint8, device=dev)
        chain.append((f,) + luma_descs(torch, f, w, h))
        w, h = (w + 1) // 2, (h + 1) // 2
    full = chain[0][0]
    frames, descs, dummy = chain[-1]
    table = torch.zeros((n, bh, bw, 3), dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    stream = torch.cuda.ExternalStream(ctx.stream())
    center = n // 2
    others = torch.tensor([i for i in range(n) if i != center], device=dev)
    t0 = time.perf_counter()
    for g in range(groups):
        synth_group_luma(ctx, full.data_ptr(), None, full_w, full_h, n, vmax, occluders, clip0 + g, config)
        for k in range(1, level + 1):
            ctx.downsample2_dev(chain[k - 1][1], chain[k][1], n)
        ctx.block_search_dev(descs, n, center, radius, table.data_ptr())
        k0 = g * (n - 1)
        prop.prepare(descs, n, center, inputs[k0].data_ptr())
        with torch.cuda.stream(stream):  # ordered after the search on the library's stream
            sel = table.index_select(0, others)
            labels[k0:k0 + n - 1] = sel[..., :2].reshape(n - 1, cells, 2)
            sads[k0:k0 + n - 1] = sel[..., 2].reshape(n - 1, cells)
        if log and (g + 1) % max(1, groups // 10) == 0:
            ctx.synchronize()
            log(f"[dataset] {g + 1}/{groups} groups ({(g + 1) * (n - 1)} pairs) in {time.perf_counter() - t0:.1f}s")
    ctx.synchronize()
    del chain, dummy
    return Dataset(inputs, labels, sads, bw, bh)


IGNORE = -32768  # label mvx marking a block left out of the loss


def training_labels(data: Dataset, max_sad_per_pixel: int | None = 12):
    """Labels with the blocks whose exhaustive minimum SAD exceeds the
    occlusion cut-off (sad / 256 > max_sad_per_pixel: no real match within
    the search range) marked IGNORE; the altref excludes those blocks too."""
    if max_sad_per_pixel is None:
        return data.labels
    lab = data.labels.clone()
    bad = data.sads >= (max_sad_per_pixel + 1) * 256
    lab[..., 0][bad] = IGNORE
    return lab


def train(prop: Proposer, data: Dataset, steps: int, batch: int = 32, lr: float = 2e-3, seed: int = 0,
          log=None, log_every: int = 100, max_sad_per_pixel: int | None = 12, warmup: int = 500):
    """Adam with a linear warm-up and a cosine learning-rate decay over
    `steps` minibatches drawn without replacement per epoch.  Returns the
    per-step losses."""
    import torch
    ctx = prop.ctx
    pairs = data.inputs.shape[0]
    labels = training_labels(data, max_sad_per_pixel)
    g = torch.Generator(device="cpu").manual_seed(seed)
    loss_dev = torch.zeros(steps, dtype=torch.float32, device=data.inputs.device)
    xb = torch.empty((batch,) + tuple(data.inputs.shape[1:]), dtype=data.inputs.dtype, device=data.inputs.device)
    yb = torch.empty((batch,) + tuple(data.labels.shape[1:]), dtype=data.labels.dtype, device=data.labels.device)
    stream = torch.cuda.ExternalStream(ctx.stream())
    order = torch.randperm(pairs, generator=g)
    pos = 0
    t0 = time.perf_counter()
    for s in range(steps):
        if pos + batch > pairs:
            order = torch.randperm(pairs, generator=g)
            pos = 0
        pos += batch
        with torch.cuda.stream(stream):
            idx = order[pos - batch:pos].to(data.inputs.device)
            torch.index_select(data.inputs, 0, idx, out=xb)
            torch.index_select(labels, 0, idx, out=yb)
        warm = min(1.0, (s + 1) / max(1, warmup))
        lr_s = warm * lr * 0.5 * (1.0 + math.cos(math.pi * s / max(1, steps)))
        prop.train_step(xb.data_ptr(), yb.data_ptr(), batch, data.bw, data.bh, lr_s,
                        loss_dev[s:].data_ptr())
        if log and ((s + 1) % log_every == 0 or s + 1 == steps):
            ctx.synchronize()
            recent = loss_dev[max(0, s + 1 - log_every):s + 1].mean().item()
            log(f"[train] step {s + 1}/{steps} loss {recent:.5f} lr {lr_s:.2e} ({time.perf_counter() - t0:.1f}s)")
    ctx.synchronize()
    return loss_dev.cpu().numpy()


# ------------------------------------------------------------ evaluation
def psnr(a, b) -> float:
    mse = float(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2))
    return float("inf") if mse == 0 else 10.0 * math.log10(255.0 ** 2 / mse)


class DeviceGroup:
    """A host group (synth.Group) uploaded with descriptors and tables."""

    def __init__(self, torch, g):
        dev = torch.device("cuda:0")
        self.g = g
        w, h = g.width, g.height
        cw = (w + 1) // 2
        self.planes = [tuple(torch.from_numpy(np.ascontiguousarray(p).reshape(-1)).to(dev) for p in t)
                       for t in zip(g.y, g.u, g.v)]
        self.descs = (_I420 * g.n)()
        for i, (y, u, v) in enumerate(self.planes):
            self.descs[i] = _I420(w, h, y.data_ptr(), u.data_ptr(), v.data_ptr(), w, cw)
        self.bw, self.bh = (w + 15) // 16, (h + 15) // 16
        self.shape = (g.n, self.bh, self.bw, 3)

    def table(self, torch):
        return torch.zeros(self.shape, dtype=torch.int32, device="cuda")

    def outputs(self, torch):
        w, h = self.g.width, self.g.height
        cw, ch = (w + 1) // 2, (h + 1) // 2
        return (torch.empty(w * h, dtype=torch.uint8, device="cuda"),
                torch.empty(cw * ch, dtype=torch.uint8, device="cuda"),
                torch.empty(cw * ch, dtype=torch.uint8, device="cuda"))


def evaluate(ctx: Context, prop: Proposer, groups, strength: int = 3, radius: int = 16, refine: int = 2,
             reps: int = 5) -> dict:
    """Learned-seeded search (propose + exact +-refine) against the exact
    +-radius search on host-generated groups: min-SAD agreement (all blocks
    and blocks with a real match, sad < 13 per pixel), altref PSNR of each
    against the noise-free centre, and the time of each search."""
    import torch
    stream = torch.cuda.ExternalStream(ctx.stream())
    res = {"groups": len(groups), "agree": [], "agree_matched": [], "matched_frac": [], "psnr_exhaustive": [],
           "psnr_learned": [], "psnr_descent": [], "psnr_noisy_centre": [], "ms_exhaustive": [],
           "ms_learned": [], "ms_propose": [], "ms_refine": []}

    def timed(fn):
        for _ in range(2):
            fn()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record(stream)
        for _ in range(reps):
            fn()
        ev[1].record(stream)
        ev[1].synchronize()
        return ev[0].elapsed_time(ev[1]) / reps

    for g in groups:
        d = DeviceGroup(torch, g)
        ex, seeds, ref = d.table(torch), d.table(torch), d.table(torch)
        oy, ou, ov = d.outputs(torch)
        torch.cuda.synchronize()
        n, c = g.n, g.center
        t_ex = timed(lambda: ctx.block_search_dev(d.descs, n, c, radius, ex.data_ptr()))
        t_pr = timed(lambda: prop.propose(d.descs, n, c, seeds.data_ptr(), radius - refine))
        t_rf = timed(lambda: ctx.block_search_dev(d.descs, n, c, refine, ref.data_ptr(), seeds.data_ptr()))
        ctx.synchronize()
        others = [i for i in range(n) if i != c]
        e, r = ex.cpu().numpy()[others], ref.cpu().numpy()[others]
        bw_px = np.minimum(16, g.width - 16 * np.arange(d.bw))
        bh_px = np.minimum(16, g.height - 16 * np.arange(d.bh))
        area = bh_px[:, None] * bw_px[None, :]
        matched = e[..., 2] < 13 * area[None]
        same = r[..., 2] == e[..., 2]
        res["agree"].append(float(same.mean()))
        res["agree_matched"].append(float(same[matched].mean()))
        res["matched_frac"].append(float(matched.mean()))
        clean_y = g.clean[c][0]
        for name, tab in (("exhaustive", ex), ("learned", ref)):
            ctx.build_arnr_altref_table_dev(d.descs, n, c, strength, tab.data_ptr(), oy.data_ptr(), ou.data_ptr(),
                                            ov.data_ptr())
            ctx.synchronize()
            res[f"psnr_{name}"].append(psnr(oy.cpu().numpy().reshape(g.height, g.width), clean_y))
        ctx.build_arnr_altref_dev(d.descs, n, c, strength, oy.data_ptr(), ou.data_ptr(), ov.data_ptr())
        ctx.synchronize()
        res["psnr_descent"].append(psnr(oy.cpu().numpy().reshape(g.height, g.width), clean_y))
        res["psnr_noisy_centre"].append(psnr(g.y[c], clean_y))
        res["ms_exhaustive"].append(t_ex)
        res["ms_propose"].append(t_pr)
        res["ms_refine"].append(t_rf)
        res["ms_learned"].append(t_pr + t_rf)
    out = {k: (float(np.mean(v)) if isinstance(v, list) else v) for k, v in res.items()}
    out["speedup"] = out["ms_exhaustive"] / out["ms_learned"]
    out["psnr_loss_db"] = out["psnr_exhaustive"] - out["psnr_learned"]
    return out


def main(argv=None) -> int:
    """Train the proposer on GPU-synthesised, exhaustively labelled pairs and
    report the learned-seeded search against the exhaustive one."""
    import argparse
    import json
    import sys

    import torch
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--out", required=True, help="safetensors file for the trained parameters")
    ap.add_argument("--init", help="start from these parameters (fine-tuning)")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--frames", type=int, default=7)
    ap.add_argument("--vmax", type=int, default=12)
    ap.add_argument("--occluders", type=int, default=20)
    ap.add_argument("--radius", type=int, default=16)
    ap.add_argument("--pairs", type=int, default=20000)
    ap.add_argument("--steps", type=int, default=60000)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--config", type=int, default=30, help="seed family of the training content")
    ap.add_argument("--eval-groups", type=int, default=2)
    ap.add_argument("--level", type=int, default=0,
                    help="pyramid level of the training pairs (groups are generated at width x height)")
    args = ap.parse_args(argv)

    def log(m):
        print(m, file=sys.stderr, flush=True)

    torch.cuda.init()
    ctx = Context(0)
    init = Proposer.load_params(args.init) if args.init else None
    prop = Proposer(ctx, params=init, seed=args.seed)
    groups = (args.pairs + args.frames - 2) // (args.frames - 1)
    t0 = time.perf_counter()
    data = build_dataset(ctx, prop, groups, args.width, args.height, args.frames, args.vmax, args.occluders,
                         args.config, radius=args.radius, log=log, level=args.level)
    t_data = time.perf_counter() - t0
    t0 = time.perf_counter()
    losses = train(prop, data, args.steps, args.batch, args.lr, args.seed, log=log,
                   log_every=max(1, args.steps // 20))
    t_train = time.perf_counter() - t0
    del data
    meta = {"pairs": groups * (args.frames - 1), "steps": args.steps, "batch": args.batch, "lr": args.lr,
            "width": args.width, "height": args.height, "frames": args.frames, "vmax": args.vmax,
            "occluders": args.occluders, "radius": args.radius, "config": args.config, "level": args.level,
            "init": args.init or "he", "final_loss": float(losses[-100:].mean())}
    prop.save(args.out, meta)
    log(f"[proposer] saved {args.out}")
    if args.level == 0:
        evals = [_synth.arnr_group_regions(args.width, args.height, args.frames, args.vmax, args.occluders,
                                           clip=i, config=3) for i in range(args.eval_groups)]
        report = evaluate(ctx, prop, evals, radius=args.radius)
    else:  # a pyramid-level proposer is evaluated through the pyramid search (vp8g.motion)
        report = {}
    report.update({"train": meta, "dataset_s": t_data, "train_s": t_train})
    print(json.dumps(report), flush=True)
    prop.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(main())
