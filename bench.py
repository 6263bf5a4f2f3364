"""Throughput of the ARNR + MC + sub-pel SAD path on one MI355X.

One step = one pass of the hot path over one 1920x1080 I420 lookahead group
of 7 frames (8x8 region motion, velocities in [-12, 12] px/frame, 20
occluding rectangles per frame, sigma-4 noise):

  1. build_arnr_altref of the whole group (centre 3, strength 3),
  2. predict_inter_mb for all 8160 MBs under random quarter-pel MVs in
     [-64, 64] (six-tap) against frame 4, and
  3. mb_luma_sad_at_mv for all 8160 MBs at the same MVs.

Inputs are resident in HBM before timing starts.  `value` = anchor MBs
(16x16 luma + 8x8 U + 8x8 V of the altref) per second over the whole step.
The `roofline` object describes the dominant kernel (the fused ARNR kernel)
timed with HIP events on the stream it runs on; `cpu_baseline` times the CPU
oracle on the same workload on this host's cores.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "oxideav-vp8_amd"))

HBM_PEAK_GBS = 8000.0
VALU_PEAK_OPS = 256 * 64 * 2.4e9   # 256 CUs x 64 lanes x 2.4 GHz, 32-bit integer VALU
METRIC = "ARNR altref MB/s @1080p (exact vs learned-seeded) + % HBM/VALU roofline"


def arnr_algorithmic(w: int, h: int, n: int):
    """Algorithmic HBM bytes and integer VALU ops of one group:
    bytes: every frame's planes read once + the altref written once;
    ops: (n-1)*33 probes of bw*bh abs-diffs counted 4 per v_sad_u8, blend
    4 ops/pixel/neighbour, resolve 4 ops/pixel."""
    cw, ch = (w + 1) // 2, (h + 1) // 2
    frame_bytes = w * h + 2 * cw * ch
    ops = 0
    for by in range(0, h, 16):
        bh = min(16, h - by)
        for bx in range(0, w, 16):
            bw = min(16, w - bx)
            px = bw * bh + 2 * ((bw + 1) // 2) * ((bh + 1) // 2)
            ops += (n - 1) * 33 * bw * bh // 4 + (n - 1) * px * 4 + px * 4
    return (n + 1) * frame_bytes, ops


def cpu_baseline(group, mc, seconds: float, cores: int):
    sys.path.insert(0, str(ROOT / "oracle"))
    from concurrent.futures import ThreadPoolExecutor

    from oracle_ctypes import LumaRef, Oracle, RefPlanes

    orc = Oracle()
    frames = group.frames()
    n_mb = len(mc.mb_xy)
    rp = RefPlanes(mc.ref_y.ctypes.data, mc.ref_u.ctypes.data, mc.ref_v.ctypes.data,
                   mc.mb_cols * 16, mc.mb_cols * 8, mc.mb_cols, mc.mb_rows)
    lr = LumaRef(mc.ref_y.ctypes.data, mc.mb_cols * 16, mc.mb_cols * 16, mc.mb_rows * 16)
    pred = np.zeros((n_mb, 384), np.uint8)
    sad = np.zeros(n_mb, np.uint32)
    filt = np.ascontiguousarray(orc.SIXTAP)
    pool = ThreadPoolExecutor(cores)
    chunks = np.array_split(np.arange(n_mb), cores)

    def mc_chunk(idx):
        if len(idx) == 0:
            return
        s = int(idx[0])
        orc.lib.orc_predict_inter_mb_batch(rp, mc.mb_xy[s:].ctypes.data, mc.mvs[s:].ctypes.data,
                                           len(idx), 0, filt.ctypes.data, pred[s:].ctypes.data)
        orc.lib.orc_mb_luma_sad_at_mv_batch(lr, mc.mb_xy[s:].ctypes.data, mc.src[s:].ctypes.data,
                                            mc.mvs[s:].ctypes.data, len(idx), sad[s:].ctypes.data)

    def one(threads):
        st, _, _, _ = orc.build_arnr_altref(frames, group.center, 3, threads=threads)
        assert st == 0
        if threads == 1:
            mc_chunk(np.arange(n_mb))
        else:
            list(pool.map(mc_chunk, chunks))

    res = {}
    for threads in (cores, 1):
        reps, t0 = 0, time.perf_counter()
        while True:
            one(threads)
            reps += 1
            el = time.perf_counter() - t0
            if el >= seconds / 2:
                break
        res[threads] = (reps, el)
    pool.shutdown()
    anchors = ((group.width + 15) // 16) * ((group.height + 15) // 16)
    reps, el = res[cores]
    r1, e1 = res[1]
    return {
        "value": anchors * reps / el,
        "unit": "anchor-MB/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"{reps} full steps (1080p 7-frame ARNR build + 8160-MB six-tap MC + SAD) "
                   f"in {el:.1f} s on {cores} threads; 1 thread: {anchors * r1 / e1:.0f} anchor-MB/s "
                   f"({r1} steps in {e1:.1f} s)"),
        "value_1thread": anchors * r1 / e1,
        "host": host_info(),
    }


def end_to_end(ctx, vp8g, frames_host, group, anchors: int, ref_out) -> dict:
    """vp8g_build_arnr_altref on host buffers (H2D copy, kernel, D2H copy),
    timed on the wall clock over at least 5 calls and 0.5 s."""
    ctx.build_arnr_altref(frames_host, group.center, vp8g.ArnrConfig(3))
    reps, t0 = 0, time.perf_counter()
    while reps < 5 or time.perf_counter() - t0 < 0.5:
        pic = ctx.build_arnr_altref(frames_host, group.center, vp8g.ArnrConfig(3))
        reps += 1
    secs = (time.perf_counter() - t0) / reps
    if not np.array_equal(pic.y, ref_out.cpu().numpy()):
        raise RuntimeError("host entry point output differs from the device entry point")
    # the same C call as a compiled caller makes it: descriptors built once,
    # output buffers reused
    import ctypes as C
    n = len(frames_host)
    descs, keep = (vp8g._I420 * n)(), []
    for i, f in enumerate(frames_host):
        planes = [np.ascontiguousarray(p, np.uint8).reshape(-1) for p in (f.y, f.u, f.v)]
        keep += planes
        descs[i] = vp8g._I420(f.width, f.height, *(p.ctypes.data for p in planes), f.y_stride, f.uv_stride)
    oy, ou, ov = (np.empty_like(p) for p in (pic.y, pic.u, pic.v))
    err = vp8g._Err()
    call = lambda: ctx.lib.vp8g_build_arnr_altref(ctx.h, descs, n, group.center, 3, oy.ctypes.data, ou.ctypes.data,
                                                  ov.ctypes.data, C.byref(err))
    call()
    reps_c, t0 = 0, time.perf_counter()
    while reps_c < 5 or time.perf_counter() - t0 < 0.5:
        if call() != 0:
            raise RuntimeError("vp8g_build_arnr_altref failed")
        reps_c += 1
    secs_c = (time.perf_counter() - t0) / reps_c
    if not np.array_equal(oy, pic.y):
        raise RuntimeError("direct C call output differs from the mirror's")
    in_bytes = sum(p.size for t in group.frames() for p in t[2:5])
    return {"ms_per_group": secs * 1e3, "anchor_mb_per_s": anchors / secs,
            "ms_per_group_c": secs_c * 1e3, "anchor_mb_per_s_c": anchors / secs_c, "bytes_in": int(in_bytes),
            "bytes_out": int(pic.y.size + pic.u.size + pic.v.size),
            "note": "vp8g_build_arnr_altref on host buffers (H2D copy, kernel, D2H copy): ms_per_group through "
                    "the Python mirror, ms_per_group_c as a direct C call with prebuilt descriptors"}


def host_info() -> dict:
    """CPU model and thread counts of the machine the CPU leg ran on."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count()}


def bench_learned(ctx, group, descs, stream, reps: int = 10) -> dict:
    """The same group's altref three ways, each search + table-driven blend
    timed with HIP events on the library's stream: the reference's 4-step
    descent (the fused kernel), the exhaustive +-16 search, and the
    learned-seeded search (proposer and exact +-2 refinement; trained
    weights oxideav-vp8_amd/weights/proposer_1080p.safetensors), and the
    same with the optional second stage (exact +-2 around each block's
    border vector and its neighbours' vectors).  Reports anchor-MB/s,
    min-SAD agreement with the exhaustive search and luma PSNR against the
    noise-free centre."""
    import torch

    import vp8g
    from vp8g.proposer import WEIGHTS_DIR, Proposer, psnr
    wpath = WEIGHTS_DIR / "proposer_1080p.safetensors"
    if not wpath.exists():
        return {"skipped": f"no weights at {wpath}"}
    prop = Proposer(ctx, params=Proposer.load_params(wpath))
    w, h, n, c = group.width, group.height, group.n, group.center
    bw, bh = (w + 15) // 16, (h + 15) // 16
    cw, ch = (w + 1) // 2, (h + 1) // 2
    dev = torch.device("cuda:0")
    ex = torch.zeros((n, bh, bw, 3), dtype=torch.int32, device=dev)
    seeds, first, ref, fb = torch.zeros_like(ex), torch.zeros_like(ex), torch.zeros_like(ex), torch.zeros_like(ex)
    queued = torch.zeros(1, dtype=torch.int32, device=dev)
    oy = torch.empty(w * h, dtype=torch.uint8, device=dev)
    ou = torch.empty(cw * ch, dtype=torch.uint8, device=dev)
    ov = torch.empty(cw * ch, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()

    def timed(fn):
        for _ in range(2):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    def altref_exhaustive():
        ctx.block_search_dev(descs, n, c, 16, ex.data_ptr())
        ctx.build_arnr_altref_table_dev(descs, n, c, 3, ex.data_ptr(), oy.data_ptr(), ou.data_ptr(), ov.data_ptr())

    def search_learned():
        prop.propose(descs, n, c, seeds.data_ptr(), 14)
        ctx.block_search_dev(descs, n, c, 2, first.data_ptr(), seeds.data_ptr())

    def search_learned_two_stage():
        search_learned()
        ctx.refine_candidates_dev(descs, n, c, seeds.data_ptr(), first.data_ptr(), ref.data_ptr())

    def search_learned_fallback():
        search_learned()
        ctx.search_fallback_dev(descs, n, c, first.data_ptr(), fb.data_ptr(), 16, True, queued.data_ptr())

    def altref_learned_fallback():
        search_learned_fallback()
        ctx.build_arnr_altref_table_dev(descs, n, c, 3, fb.data_ptr(), oy.data_ptr(), ou.data_ptr(), ov.data_ptr())

    def altref_learned():
        search_learned()
        ctx.build_arnr_altref_table_dev(descs, n, c, 3, first.data_ptr(), oy.data_ptr(), ou.data_ptr(), ov.data_ptr())

    def altref_learned_two_stage():
        search_learned_two_stage()
        ctx.build_arnr_altref_table_dev(descs, n, c, 3, ref.data_ptr(), oy.data_ptr(), ou.data_ptr(), ov.data_ptr())

    def altref_descent():
        ctx.build_arnr_altref_dev(descs, n, c, 3, oy.data_ptr(), ou.data_ptr(), ov.data_ptr())

    anchors = bw * bh
    clean = group.clean[c][0]
    out = {}
    for name, fn in (("descent", altref_descent), ("exhaustive16", altref_exhaustive), ("learned", altref_learned),
                     ("learned_two_stage", altref_learned_two_stage), ("learned_fallback", altref_learned_fallback)):
        ms = timed(fn)
        ctx.synchronize()
        out[name] = {"ms": ms, "anchor_mb_per_s": anchors / (ms * 1e-3),
                     "psnr_y_vs_clean": psnr(oy.cpu().numpy().reshape(h, w), clean)}
    t_search_ex = timed(lambda: ctx.block_search_dev(descs, n, c, 16, ex.data_ptr()))
    t_search_le = timed(search_learned)
    t_search_le2 = timed(search_learned_two_stage)
    t_search_fb = timed(search_learned_fallback)
    ctx.synchronize()
    others = [i for i in range(n) if i != c]
    e, r1, r2 = ex.cpu().numpy()[others], first.cpu().numpy()[others], ref.cpu().numpy()[others]
    r3 = fb.cpu().numpy()[others]
    area = (np.minimum(16, h - 16 * np.arange(bh))[:, None] * np.minimum(16, w - 16 * np.arange(bw))[None, :])
    matched = e[..., 2] < 13 * area[None]
    same = r1[..., 2] == e[..., 2]
    same2 = r2[..., 2] == e[..., 2]
    same3 = r3[..., 2] == e[..., 2]
    out.update({
        "agreement": float(same.mean()),
        "agreement_matched": float(same[matched].mean()),
        "matched_fraction": float(matched.mean()),
        "psnr_loss_db": out["exhaustive16"]["psnr_y_vs_clean"] - out["learned"]["psnr_y_vs_clean"],
        "search_ms": {"exhaustive16": t_search_ex, "learned": t_search_le, "learned_two_stage": t_search_le2,
                      "learned_fallback": t_search_fb},
        "search_speedup": t_search_ex / t_search_le,
        "two_stage": {"agreement_matched": float(same2[matched].mean()),
                      "psnr_loss_db": out["exhaustive16"]["psnr_y_vs_clean"] -
                      out["learned_two_stage"]["psnr_y_vs_clean"],
                      "search_speedup": t_search_ex / t_search_le2},
        "fallback": {"agreement": float(same3.mean()), "agreement_matched": float(same3[matched].mean()),
                     "queued_fraction": int(queued.item()) / float(e[..., 2].size),
                     "psnr_loss_db": out["exhaustive16"]["psnr_y_vs_clean"] -
                     out["learned_fallback"]["psnr_y_vs_clean"],
                     "search_speedup": t_search_ex / t_search_fb,
                     "rule": "exhaustive +-16 for entries failing the occlusion cut-off with a 4-neighbour under it"},
        "weights": str(wpath.name),
    })
    prop.close()
    return out


def measure_next_rows(ctx, group, cores: int, cpu_seconds: float = 2.0, reps: int = 10) -> dict:
    """The rows beside the hot path, each over every macroblock of a 1080p
    frame (the bench group's centre against its next frame, edge-padded to
    1920x1088), GPU time from HIP events on the library stream (device
    inputs), the oracle on `cores` host threads beside it:

    * encoder inter search: whole_pixel_search then sub_pixel_refine
      (Balanced shape, zero and census seeds, lambda_sad 8, good_enough 896,
      +-32 px range, default MV contexts);
    * SPLITMV: predict_split_mv (16 vectors per MB in [-64, 64] quarter
      pel), reconstruct_split_mv_mb and reconstruct_inter_mb with sparse
      dequantised coefficients (a quarter of the MBs skipped);
    * luma_mad of 64 consecutive 1080p frame pairs in one call."""
    import ctypes as C
    from concurrent.futures import ThreadPoolExecutor

    import torch

    import vp8g
    import vp8g.inter_search as IS
    import vp8g.synth as synth

    sys.path.insert(0, str(ROOT / "oracle"))
    from oracle_ctypes import I420, LumaRef, Oracle, RefPlanes

    orc = Oracle()
    dev = torch.device("cuda:0")
    stream = torch.cuda.ExternalStream(ctx.stream(), device=dev)
    c = group.center
    cur = np.ascontiguousarray(synth.pad_to_mb(group.y[c], 16))
    ref = np.ascontiguousarray(synth.pad_to_mb(group.y[c + 1], 16))
    ru = np.ascontiguousarray(synth.pad_to_mb(group.u[c + 1], 8))
    rv = np.ascontiguousarray(synth.pad_to_mb(group.v[c + 1], 8))
    h, w = ref.shape
    bw, bh = w // 16, h // 16
    n = bw * bh
    xy = np.stack(np.meshgrid(np.arange(bw), np.arange(bh)), -1).reshape(-1, 2).astype(np.uint32)
    src = np.ascontiguousarray(cur.reshape(bh, 16, bw, 16).transpose(0, 2, 1, 3).reshape(n, 256))
    rng = np.random.default_rng(8160)
    costs = IS.MvCostTable.new()
    jobs = np.zeros(n, IS.MB_SEARCH)
    jobs["mb_col"], jobs["mb_row"] = xy[:, 0], xy[:, 1]
    jobs["range_x"], jobs["range_y"] = (-32, 32), (-32, 32)
    jobs["n_seeds"] = 2
    jobs["seeds"][:, 1] = rng.integers(-24, 25, (n, 2)) * 4
    jobs["best"] = jobs["seeds"][:, 1]
    jobs["lambda_sad"] = 8.0
    jobs["good_enough"] = 896
    jobs["descend"] = 1
    shape = IS.SHAPES["balanced"]
    split_mvs = rng.integers(-64, 65, (n, 16, 2)).astype(np.int16)
    mvs = rng.integers(-64, 65, (n, 2)).astype(np.int16)
    skip = (rng.random(n) < 0.25).astype(np.uint8)
    cf25 = (rng.integers(-40, 41, (n, 25, 16)) * (rng.random((n, 25, 16)) < 0.1)).astype(np.int16)
    cf24 = np.ascontiguousarray(cf25[:, 1:])
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a).reshape(-1).view(np.uint8).copy()).to(dev)
    d = {k: t(v) for k, v in (("ref", ref), ("u", ru), ("v", rv), ("xy", xy), ("src", src), ("costs", costs.bits),
                              ("jobs", jobs), ("smv", split_mvs), ("mv", mvs), ("skip", skip), ("cf25", cf25),
                              ("cf24", cf24))}
    found = torch.zeros(n * 16, dtype=torch.uint8, device=dev)
    found2 = torch.zeros(n * 16, dtype=torch.uint8, device=dev)
    out = torch.zeros(n * 384, dtype=torch.uint8, device=dev)
    lr = vp8g._LumaRef(d["ref"].data_ptr(), w, w, h)
    rp = vp8g._RefPlanes(d["ref"].data_ptr(), d["u"].data_ptr(), d["v"].data_ptr(), w, w // 2, bw, bh)
    f6 = np.ascontiguousarray(vp8g.SIXTAP_FILTERS, np.int32)
    lib = ctx.lib

    def ok(st):
        if st:
            raise RuntimeError(f"vp8g launch failed: {st}")

    def timed(fn):
        for _ in range(2):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    wps = lambda: ok(lib.vp8g_whole_pixel_search_batch_dev(ctx.h, C.byref(lr), d["costs"].data_ptr(),
                                                           d["jobs"].data_ptr(), d["src"].data_ptr(), n, shape.c(),
                                                           found.data_ptr()))
    spr = lambda: ok(lib.vp8g_sub_pixel_refine_batch_dev(ctx.h, C.byref(lr), d["costs"].data_ptr(),
                                                         d["jobs"].data_ptr(), d["src"].data_ptr(), found.data_ptr(),
                                                         n, f6.ctypes.data, shape.c(), found2.data_ptr()))
    ms_wps, ms_spr = timed(wps), timed(spr)
    split = lambda: ok(lib.vp8g_predict_split_mv_batch_dev(ctx.h, C.byref(rp), d["xy"].data_ptr(),
                                                           d["smv"].data_ptr(), n, 0, f6.ctypes.data, out.data_ptr()))
    rsplit = lambda: ok(lib.vp8g_reconstruct_split_mv_mb_batch_dev(ctx.h, C.byref(rp), d["xy"].data_ptr(),
                                                                   d["smv"].data_ptr(), n, 0, f6.ctypes.data,
                                                                   d["skip"].data_ptr(), d["cf24"].data_ptr(),
                                                                   out.data_ptr()))
    rinter = lambda: ok(lib.vp8g_reconstruct_inter_mb_batch_dev(ctx.h, C.byref(rp), d["xy"].data_ptr(),
                                                                d["mv"].data_ptr(), n, 0, f6.ctypes.data,
                                                                d["skip"].data_ptr(), d["cf25"].data_ptr(),
                                                                out.data_ptr()))
    ms_split, ms_rsplit, ms_rinter = timed(split), timed(rsplit), timed(rinter)
    # luma_mad of a 256-frame 1080p stream (255 consecutive pairs, 530 MB of
    # luma: more than the last-level cache holds), one call
    gw, gh = group.width, group.height
    n_mad = 255
    stream_y = torch.randint(0, 256, (n_mad + 1, gh * gw), dtype=torch.uint8, device=dev)
    pa, ca = (vp8g._I420 * n_mad)(), (vp8g._I420 * n_mad)()
    for i in range(n_mad):
        a, b = stream_y[i].data_ptr(), stream_y[i + 1].data_ptr()
        pa[i] = vp8g._I420(gw, gh, a, a, a, gw, (gw + 1) // 2)
        ca[i] = vp8g._I420(gw, gh, b, b, b, gw, (gw + 1) // 2)
    sads = torch.zeros(n_mad, dtype=torch.int64, device=dev)
    mad = lambda: ok(lib.vp8g_luma_sad_batch_dev(ctx.h, pa, ca, n_mad, sads.data_ptr()))
    ms_mad = timed(mad)
    torch.cuda.synchronize()

    # CPU: the oracle over the same inputs on `cores` threads
    pool = ThreadPoolExecutor(cores)
    chunks = [np.arange(k, n, cores) for k in range(cores)]
    olr = LumaRef(ref.ctypes.data, w, w, h)
    orp = RefPlanes(ref.ctypes.data, ru.ctypes.data, rv.ctypes.data, w, w // 2, bw, bh)
    oj, keep = orc._orc_jobs(jobs, src)
    ofound = np.zeros(n, IS.FOUND)
    ofound2 = np.zeros_like(ofound)
    oout = np.zeros((n, 384), np.uint8)
    bits = np.ascontiguousarray(costs.bits)
    sh = (C.c_uint32 * 4)(shape.descent_rounds, 1, 1, 1)
    contig = [np.ascontiguousarray(ix) for ix in np.array_split(np.arange(n), cores)]

    def cpu_time(fn):
        reps_c, t0 = 0, time.perf_counter()
        while True:
            list(pool.map(fn, contig))
            reps_c += 1
            if time.perf_counter() - t0 >= cpu_seconds:
                break
        return (time.perf_counter() - t0) / reps_c * 1e3

    def o_search(ix):
        s0, k = int(ix[0]), len(ix)
        orc.lib.orc_whole_pixel_search_batch(C.byref(olr), bits.ctypes.data, oj[s0:].ctypes.data, k, sh,
                                             ofound[s0:].ctypes.data)
        orc.lib.orc_sub_pixel_refine_batch(C.byref(olr), bits.ctypes.data, oj[s0:].ctypes.data,
                                           ofound[s0:].ctypes.data, k, f6.ctypes.data, sh, ofound2[s0:].ctypes.data)

    def o_split(ix):
        s0, k = int(ix[0]), len(ix)
        orc.lib.orc_reconstruct_split_mv_mb_batch(C.byref(orp), xy[s0:].ctypes.data, split_mvs[s0:].ctypes.data, k,
                                                  0, f6.ctypes.data, skip[s0:].ctypes.data, cf24[s0:].ctypes.data,
                                                  oout[s0:].ctypes.data)

    def o_inter(ix):
        s0, k = int(ix[0]), len(ix)
        orc.lib.orc_reconstruct_inter_mb_batch(C.byref(orp), xy[s0:].ctypes.data, mvs[s0:].ctypes.data, k, 0,
                                               f6.ctypes.data, skip[s0:].ctypes.data, cf25[s0:].ctypes.data,
                                               oout[s0:].ctypes.data)

    cpu_search, cpu_split, cpu_inter = cpu_time(o_search), cpu_time(o_split), cpu_time(o_inter)
    # the oracle over the first 33 frames of the same stream, scaled to 255 pairs
    ys = stream_y[:33].cpu().numpy()
    i420 = [I420(gw, gh, y.ctypes.data, y.ctypes.data, y.ctypes.data, gw, (gw + 1) // 2) for y in ys]
    wmad = np.array([orc.lib.orc_luma_mad(C.byref(i420[i]), C.byref(i420[i + 1])) for i in range(32)])
    got_mad = sads[:32].cpu().numpy().astype(np.float64) / float(gw * gh)

    def o_mad(ix):
        for i in ix:
            orc.lib.orc_luma_mad(C.byref(i420[i]), C.byref(i420[i + 1]))

    reps_c, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < cpu_seconds:
        list(pool.map(o_mad, np.array_split(np.arange(32), cores)))
        reps_c += 1
    cpu_mad = (time.perf_counter() - t0) / reps_c * 1e3 * n_mad / 32
    pool.shutdown()
    # the GPU results equal the oracle's on these inputs
    got2 = found2.cpu().numpy().view(ofound.dtype)
    match = {"inter_search": bool(np.array_equal(got2["mv"], ofound2["mv"]) and
                                  np.array_equal(got2["cost"], ofound2["cost"]))}
    rinter()
    torch.cuda.synchronize()
    match["reconstruct_inter_mb"] = bool(np.array_equal(out.cpu().numpy().reshape(n, 384), oout))
    match["luma_mad"] = bool(np.array_equal(got_mad, wmad))

    def rec(ms, cpu_ms, units, bytes_per_unit, what):
        gbs = bytes_per_unit * units / (ms * 1e-3) / 1e9
        return {"ms": ms, "units": units, "unit": what, "per_s": units / (ms * 1e-3),
                "cpu_ms": cpu_ms, "cpu_cores": cores, "gpu_over_cpu": cpu_ms / ms if cpu_ms else None,
                "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": gbs / HBM_PEAK_GBS, "algorithmic_bytes_per_unit": bytes_per_unit}}

    return {
        "workload": "1920x1088 frame (8160 MBs) of the bench group against its next frame; luma_mad over a "
                    "256-frame 1080p stream (255 pairs, one call)",
        # per MB: the 256-byte source, its 20x20 reference neighbourhood and
        # the 16-byte result (inter search); 384 B reference + 384 B
        # output + MVs + coefficients (reconstructions)
        "inter_search": dict(rec(ms_wps + ms_spr, cpu_search, n, 256 + 400 + 16 + 80, "MBs"),
                             ms_whole_pixel=ms_wps, ms_sub_pixel=ms_spr, shape="balanced"),
        "predict_split_mv": rec(ms_split, None, n, 384 + 384 + 64, "MBs"),
        "reconstruct_split_mv_mb": rec(ms_rsplit, cpu_split, n, 384 + 384 + 64 + 769, "MBs"),
        "reconstruct_inter_mb": rec(ms_rinter, cpu_inter, n, 384 + 384 + 4 + 801, "MBs"),
        # each frame of the stream read once
        "luma_mad": rec(ms_mad, cpu_mad, n_mad, gw * gh, "frame pairs"),
        "bit_exact_vs_oracle": match,
    }


def measure_loop_filter(ctx, group, cores: int, reps: int = 10, cpu_seconds: float = 2.0) -> dict:
    """The section 15 loop filter over a 1920x1088 frame (8160 MBs): the
    bench group's centre frame is the source, the reconstruction is that
    frame plus a +-9 step per 4x4 block (chroma +-7), and every MB carries
    a seeded inter-frame record (segment, ref frame, inter mode, B_PRED /
    coefficient flags).  GPU legs on the library stream with device inputs:

    * frame: filter_inter_frame_flags over the frame, in place (level 36,
      sharpness 2, ref / mode deltas on);
    * score64: 64 candidates (every level at the base sharpness, the level
      stages of the search in one launch);
    * select / select_sampled: select_filter_params and
      select_filter_params_sampled (key-frame sampling, sweep on), wall
      clock of the synchronous calls.

    CPU: the oracle's raster pass on one thread (frame) and its searches
    scoring candidates on `cores` threads, as the reference's threaded
    search does."""
    import ctypes as C

    import torch

    import vp8g
    from vp8g import loop_filter as L
    import vp8g.synth as synth

    sys.path.insert(0, str(ROOT / "oracle"))
    from oracle_ctypes import Oracle, lf_config

    orc = Oracle()
    dev = torch.device("cuda:0")
    stream = torch.cuda.ExternalStream(ctx.stream(), device=dev)
    c = group.center
    sy = np.ascontiguousarray(synth.pad_to_mb(group.y[c], 16))
    su = np.ascontiguousarray(synth.pad_to_mb(group.u[c], 8))
    sv = np.ascontiguousarray(synth.pad_to_mb(group.v[c], 8))
    H, W = sy.shape
    bw, bh = W // 16, H // 16
    n = bw * bh
    rng = np.random.default_rng(1536)
    ry = np.clip(sy.astype(np.int32) + rng.integers(-9, 10, (H // 4, W // 4)).repeat(4, 0).repeat(4, 1), 0,
                 255).astype(np.uint8)
    cb = rng.integers(-7, 8, (2, H // 8, W // 8)).repeat(4, 1).repeat(4, 2)
    ru = np.clip(su.astype(np.int32) + cb[0], 0, 255).astype(np.uint8)
    rv = np.clip(sv.astype(np.int32) + cb[1], 0, 255).astype(np.uint8)
    mbs = np.zeros((n, 4), np.uint8)
    mbs[:, 0] = rng.choice([255, 0, 1, 2, 3], n)
    mbs[:, 1] = rng.choice([0, 1, 2, 3], n)
    mbs[:, 2] = np.where(mbs[:, 1] > 0, rng.choice([1, 2, 3], n), 0)
    mbs[:, 3] = rng.choice([0, 1], n) | np.where(rng.random(n) < 0.6, 2, 0)
    fcfg = dict(key_frame=False, loop_filter_level=36, sharpness_level=2, delta_enabled=True, ref_delta_last=-4,
                ref_delta_golden=6, zero_mv_mode_delta=3, split_mv_mode_delta=-2, segmentation_enabled=True,
                segment_lf_level=(0, 4, -6, 10))
    g = L.FrameFilterConfig(**fcfg)
    og = lf_config(simple=0, key_frame=0, loop_filter_level=36, sharpness_level=2, delta_enabled=1,
                   segmentation_enabled=1, segment_abs=0, segment_lf_level=[0, 4, -6, 10], ref_delta=g.ref_deltas(),
                   mode_delta=g.mode_deltas())
    sel_g = L.FrameFilterConfig(key_frame=True, loop_filter_level=20, sharpness_level=0)
    sel_og = lf_config(simple=0, key_frame=1, loop_filter_level=20, sharpness_level=0)
    vis_w, vis_h = group.width, group.height
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a).reshape(-1).copy()).to(dev)
    d = {k: t(v) for k, v in (("ry", ry), ("ru", ru), ("rv", rv), ("sy", sy), ("su", su), ("sv", sv),
                              ("mbs", mbs))}
    wy, wu, wv = d["ry"].clone(), d["ru"].clone(), d["rv"].clone()
    work = L._Planes(wy.data_ptr(), wu.data_ptr(), wv.data_ptr(), W, W // 2, bw, bh)
    recon = L._Planes(d["ry"].data_ptr(), d["ru"].data_ptr(), d["rv"].data_ptr(), W, W // 2, bw, bh)
    src = L._Source(vis_w, vis_h, d["sy"].data_ptr(), d["su"].data_ptr(), d["sv"].data_ptr(), W, W // 2)
    lib = ctx.lib
    torch.cuda.synchronize()

    def ok(st):
        if st:
            raise RuntimeError(f"vp8g loop filter call failed: {st}")

    def timed(fn):
        for _ in range(2):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    gc = g.c()
    frame = lambda: ok(lib.vp8g_loop_filter_frame_dev(ctx.h, C.byref(work), d["mbs"].data_ptr(), C.byref(gc)))
    # correctness first: one pass from the reconstruction against the oracle
    frame()
    ok(lib.vp8g_loop_filter_status(ctx.h))
    want = orc.lf_filter_frame(ry, ru, rv, mbs, og)
    exact = {"frame": bool(all(np.array_equal(a.cpu().numpy().reshape(b.shape), b)
                               for a, b in zip((wy, wu, wv), want)))}
    ms_frame = timed(frame)
    lv = np.arange(64, dtype=np.uint8)
    sh = np.zeros(64, np.uint8)
    ssd = torch.zeros(64, dtype=torch.int64, device=dev)
    sgc = sel_g.c()
    score = lambda: ok(lib.vp8g_loop_filter_score_dev(ctx.h, C.byref(recon), d["mbs"].data_ptr(), C.byref(sgc),
                                                      C.byref(src), lv.ctypes.data, sh.ctypes.data, 64, 0,
                                                      ssd.data_ptr()))
    ms_score = timed(score)
    every = L.sampling_for_key_frame(bh)
    lo, so = C.c_uint8(), C.c_uint8()

    def select(every_rows, sweep):
        ok(lib.vp8g_loop_filter_select_dev(ctx.h, C.byref(recon), d["mbs"].data_ptr(), C.byref(sgc),
                                           C.byref(src), every_rows, sweep, C.byref(lo), C.byref(so)))
        return lo.value, so.value

    def wall(fn, k=reps):
        fn()
        t0 = time.perf_counter()
        for _ in range(k):
            r = fn()
        return (time.perf_counter() - t0) / k * 1e3, r

    ms_select, pick = wall(lambda: select(0, 1))
    ms_sampled, pick_s = wall(lambda: select(every, 1))
    torch.cuda.synchronize()

    # CPU: the oracle
    osrc = (vis_w, vis_h, sy, su, sv)

    def cpu_wall(fn):
        k, t0 = 0, time.perf_counter()
        while True:
            r = fn()
            k += 1
            if time.perf_counter() - t0 >= cpu_seconds:
                break
        return (time.perf_counter() - t0) / k * 1e3, r

    cpu_frame, _ = cpu_wall(lambda: orc.lf_filter_frame(ry, ru, rv, mbs, og))
    cpu_select, want_pick = cpu_wall(lambda: orc.lf_select(ry, ru, rv, mbs, sel_og, osrc, threads=cores))
    cpu_sampled, want_pick_s = cpu_wall(
        lambda: orc.lf_select_sampled(ry, ru, rv, mbs, sel_og, osrc, every, True, threads=cores))
    exact["select"] = pick == want_pick
    exact["select_sampled"] = pick_s == want_pick_s
    wscore = orc.lf_score_candidates(ry, ru, rv, mbs, sel_og, osrc, lv, sh, threads=cores)
    exact["score64"] = bool(np.array_equal(ssd.cpu().numpy().astype(np.uint64), wscore))
    # bytes a pass must move: the frame read once and written once (384 B
    # per MB each way) plus the 4-byte MB record; scoring reads the source
    # once instead of writing
    frame_bytes = n * (384 + 384 + 4)
    gbs = frame_bytes / (ms_frame * 1e-3) / 1e9
    return {
        "workload": f"1920x1088 reconstruction (8160 MBs) of the bench group's centre frame with 4x4 block steps, "
                    f"seeded inter MB records; visible {vis_w}x{vis_h}; sampled search every {every} MB rows",
        "frame": {"ms": ms_frame, "frames_per_s": 1e3 / ms_frame, "mbs_per_s": n / (ms_frame * 1e-3),
                  "cpu_ms": cpu_frame, "cpu_cores": 1, "gpu_over_cpu": cpu_frame / ms_frame,
                  "roofline": {"bound": "row wavefront (latency)", "achieved": gbs, "peak": HBM_PEAK_GBS,
                               "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                               "algorithmic_bytes_per_unit": 384 + 384 + 4, "units": n,
                               "critical_path_mb_steps": bw + 2 * (bh - 1),
                               "us_per_critical_step": ms_frame * 1e3 / (bw + 2 * (bh - 1))}},
        "score64": {"ms": ms_score, "candidates": 64, "frames_per_s": 64e3 / ms_score},
        "select": {"ms": ms_select, "pick": list(pick), "cpu_ms": cpu_select, "cpu_cores": cores,
                   "gpu_over_cpu": cpu_select / ms_select},
        "select_sampled": {"ms": ms_sampled, "every_rows": every, "pick": list(pick_s), "cpu_ms": cpu_sampled,
                           "cpu_cores": cores, "gpu_over_cpu": cpu_sampled / ms_sampled},
        "bit_exact_vs_oracle": exact,
    }


def bench_loop_filter(args) -> int:
    import torch

    import vp8g
    import vp8g.synth as synth

    group = synth.arnr_group_regions()
    torch.cuda.init()
    ctx = vp8g.Context(0)
    cores = max(1, min(os.cpu_count() or 1, 16))
    r = measure_loop_filter(ctx, group, cores, reps=max(3, args.steps // 5),
                            cpu_seconds=0.0 if args.no_cpu else 2.0)
    ctx.close()
    print(json.dumps({"loop_filter": r}), flush=True)
    return 0


def measure_4k(ctx, reps: int = 3) -> dict:
    """3840x2160 I420, 11-frame group (centre 5), region velocities in
    [-48, 48] px/frame, 40 occluders per frame: exhaustive +-64 and the
    coarse-to-fine learned-seeded whole-pel search (proposals at quarter
    resolution with the weights fine-tuned there, exact +-2 at every level),
    then half- and quarter-pel refinement and six-tap MC of all 32400
    macroblocks of each of the 10 neighbours.  Throughput = anchor MBs per
    second of the learned pipeline; the residual proxy (SAD / SSE of the MC
    luma vs the centre) stands in for encoded stream bytes."""
    import vp8g.synth as synth
    from vp8g.motion import evaluate_group
    from vp8g.proposer import WEIGHTS_DIR, Proposer
    t0 = time.perf_counter()
    g = synth.arnr_group_regions(3840, 2160, 11, vmax=48, occluders=40, clip=0, config=4)
    print(f"[bench] 4K group generated in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    ppath = WEIGHTS_DIR / "proposer_4k_l2.safetensors"
    pprop = Proposer(ctx, params=Proposer.load_params(ppath))
    r = evaluate_group(ctx, None, g, radius=64, refine=2, reps=reps, pyramid_prop=pprop, pyramid_levels=2)
    pprop.close()
    anchors = 240 * 135
    le, ex = r["learned_pyramid"], r["exhaustive"]
    mc_gbs = le["mc_gb_per_s"]
    return {
        "metric": "4K learned-seeded search + sub-pel refine + six-tap MC, anchor-MB/s (vs exhaustive +-64)",
        "value": le["anchor_mb_per_s"],
        "unit": "anchor-MB/s",
        "ms_per_group": le["ms_total"],
        "exhaustive_ms_per_group": ex["ms_total"],
        "search_speedup": r["pyramid"]["search_speedup"],
        "agreement": r["pyramid"]["agreement"],
        "agreement_matched": r["pyramid"]["agreement_matched"],
        "matched_fraction": r["matched_fraction"],
        "residual_sad_per_px": {"learned": le["residual_sad_per_px"], "exhaustive": ex["residual_sad_per_px"]},
        "roofline": {"bound": "hbm", "kernel": "predict_inter_mb_kernel", "achieved": mc_gbs, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": mc_gbs / HBM_PEAK_GBS,
                     "algorithmic_bytes_per_launch": 32400 * (384 + 384 + 4)},
        "config": {"workload": "3840x2160 I420, 11-frame group (centre 5), 8x8 region motion [-48,48] px/frame, "
                               "40 occluders/frame, sigma 4 noise", "anchor_mbs": anchors,
                   "search": "coarse to fine: proposals at quarter resolution, exact +-2 at quarter, half and full",
                   "weights": ppath.name},
        "report": r,
        "stream_bytes_qi44": "not measurable here (reference encoder not buildable); residual SAD/SSE reported",
    }


def bench_4k(args) -> int:
    import torch

    import vp8g
    torch.cuda.init()
    ctx = vp8g.Context(0)
    r = measure_4k(ctx, reps=max(1, min(args.steps, 3)))
    ctx.close()
    result = {"metric": r["metric"], "value": r["value"], "unit": r["unit"], "n_gpus": 1, "steps": 1, "warmup": 1,
              "ms_per_step": r["ms_per_group"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
              "dtype": "u8", "data": "synthetic", "config": dict(r["config"], parallelism="single GPU")}
    result.update({k: v for k, v in r.items() if k not in result})
    print(json.dumps(result), flush=True)
    return 0


def measure_batch(ctx, steps: int, warmup: int) -> dict:
    """64 clips (resolutions uniform over 176x144 .. 1920x1080, lengths over
    {3, 5, 7, 9, 11}, region motion + occluders) as one persistent,
    graph-captured launch, against the same clips as back-to-back
    single-group launches.  Throughput = anchor MBs of all clips per second."""
    import torch

    import vp8g
    import vp8g.synth as synth
    from vp8g.autotune import DeviceClip, time_per_clip_loop

    t0 = time.perf_counter()
    specs = synth.batch_specs(64)
    groups = synth.batch_clips(64, specs=specs)
    print(f"[bench] 64 clips generated in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    clips = [DeviceClip(torch, g) for g in groups]
    torch.cuda.synchronize()
    batch = ctx.create_arnr_batch([c.job(vp8g) for c in clips])
    shape = batch.shape()
    stream = torch.cuda.ExternalStream(ctx.stream(), device=torch.device("cuda:0"))
    for _ in range(warmup):
        batch.run()
    ctx.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    t_wall = time.perf_counter()
    ev[0].record(stream)
    for k in range(steps):
        batch.run()
        ev[k + 1].record(stream)
    ev[-1].synchronize()
    wall = time.perf_counter() - t_wall
    per = [ev[k].elapsed_time(ev[k + 1]) for k in range(steps)]
    ms_step = ev[0].elapsed_time(ev[-1]) / steps
    print(f"[bench] batch: {steps} replays, {ms_step:.4f} ms/replay (wall {wall:.3f}s)", file=sys.stderr, flush=True)
    loop_ms = time_per_clip_loop(torch, vp8g, ctx, clips, reps=max(3, steps // 5))
    print(f"[bench] per-clip launches: {loop_ms:.4f} ms per pass", file=sys.stderr, flush=True)
    anchors = sum(c.anchor_mbs for c in clips)
    nb = sum(c.anchor_mbs * (len(c.frames) - 1) for c in clips)
    bytes_alg = sum(arnr_algorithmic(g.width, g.height, g.n)[0] for g in groups)
    ms_kernel = float(np.mean(per))
    achieved = bytes_alg / (ms_kernel * 1e-3) / 1e9
    batch.close()
    return {
        "metric": "ARNR batched aggregate anchor-MB/s (64 uneven clips, one persistent launch)",
        "value": anchors / (ms_step * 1e-3),
        "unit": "anchor-MB/s",
        "steps": steps,
        "ms_per_step": ms_step,
        "config": {
            "workload": "64 clips, resolution uniform over {176x144, 352x288, 640x360, 1280x720, "
                        "1920x1080}, group length uniform over {3,5,7,9,11}, centre n/2, strength 3, "
                        "8x8 region motion [-12,12], 20 occluders/frame, sigma 4 noise",
            "anchor_mbs_per_step": anchors,
            "mb_neighbour_pairs_per_step": nb,
            "shape": shape,
        },
        "per_clip_launches": {"ms_per_pass": loop_ms, "anchor_mb_per_s": anchors / (loop_ms * 1e-3),
                              "batch_speedup": loop_ms / ms_step},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "arnr_batch_kernel", "algorithmic_bytes_per_launch": bytes_alg},
    }


def bench_batch(args) -> int:
    import torch

    import vp8g
    torch.cuda.init()
    ctx = vp8g.Context(0)
    r = measure_batch(ctx, args.steps, args.warmup)
    ctx.close()
    result = {"metric": r["metric"], "value": r["value"], "unit": r["unit"], "n_gpus": 1, "steps": args.steps,
              "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
              "vs_baseline": None, "dtype": "u8", "data": "synthetic",
              "config": dict(r["config"], parallelism="single GPU"), "cpu_baseline": None}
    result.update({k: v for k, v in r.items() if k not in result})
    print(json.dumps(result), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=16.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--arnr-waves", type=int, default=0)
    ap.add_argument("--no-xcd-remap", action="store_true")
    ap.add_argument("--arnr-mbs", type=int, default=0)
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-seeded altref leg")
    ap.add_argument("--no-e2e", action="store_true", help="skip the end-to-end (host buffers) leg")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the batched-clips (c2_batch) and 4K pipeline (c4_pipeline) sub-records")
    ap.add_argument("--workload", choices=("step", "batch", "4k", "loop_filter"), default="step",
                    help="step: the 1080p group step (default); batch: 64 uneven clips in one "
                         "persistent graph-captured launch against per-clip launches; 4k: 3840x2160 "
                         "11-frame group, exhaustive +-64 vs learned-seeded search, then half/quarter-pel "
                         "refinement and six-tap MC of every macroblock; loop_filter: the section 15 "
                         "loop filter legs alone")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB (5 times), the known-size "
                         "kernel that scales FETCH_SIZE / WRITE_SIZE counter readings")
    args = ap.parse_args()
    if args.calibrate_copy_mib:
        import torch
        n = args.calibrate_copy_mib << 20
        src = torch.randint(0, 255, (n,), dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": n, "launches": 5}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the path shards per clip/group with no exchange; see the design notes)",
              file=sys.stderr, flush=True)
        return 2

    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "4k":
        return bench_4k(args)
    if args.workload == "loop_filter":
        return bench_loop_filter(args)

    import torch

    import vp8g
    import vp8g.synth as synth

    t0 = time.perf_counter()
    group = synth.arnr_group_regions()
    mc = synth.mc_workload(group)
    print(f"[bench] workload generated in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)

    dev = torch.device("cuda:0")
    torch.cuda.init()
    # torch is loaded first, so libvp8g.so binds to the same HIP runtime and
    # the library's own stream can be wrapped for event timing.
    ctx = vp8g.Context(0)
    stream = torch.cuda.ExternalStream(ctx.stream(), device=dev)
    if args.arnr_waves:
        ctx.set_arnr_waves(args.arnr_waves)
    if args.no_xcd_remap:
        ctx.set_xcd_remap(False)
    if args.arnr_mbs:
        ctx.set_arnr_mbs(args.arnr_mbs)

    w, h, n = group.width, group.height, group.n
    cw, ch = (w + 1) // 2, (h + 1) // 2
    planes = [(torch.from_numpy(y.reshape(-1).copy()).to(dev), torch.from_numpy(u.reshape(-1).copy()).to(dev),
               torch.from_numpy(v.reshape(-1).copy()).to(dev)) for y, u, v in zip(group.y, group.u, group.v)]
    descs = (vp8g._I420 * n)()
    for i, (y, u, v) in enumerate(planes):
        descs[i] = vp8g._I420(w, h, y.data_ptr(), u.data_ptr(), v.data_ptr(), w, cw)
    oy = torch.empty(w * h, dtype=torch.uint8, device=dev)
    ou = torch.empty(cw * ch, dtype=torch.uint8, device=dev)
    ov = torch.empty(cw * ch, dtype=torch.uint8, device=dev)

    n_mb = len(mc.mb_xy)
    ry = torch.from_numpy(mc.ref_y.reshape(-1).copy()).to(dev)
    ru = torch.from_numpy(mc.ref_u.reshape(-1).copy()).to(dev)
    rv = torch.from_numpy(mc.ref_v.reshape(-1).copy()).to(dev)
    xy = torch.from_numpy(mc.mb_xy.reshape(-1).astype(np.int32)).to(dev)
    mvs = torch.from_numpy(mc.mvs.reshape(-1).copy()).to(dev)
    src = torch.from_numpy(mc.src.reshape(-1).copy()).to(dev)
    pred = torch.empty(n_mb * 384, dtype=torch.uint8, device=dev)
    sad = torch.empty(n_mb, dtype=torch.int32, device=dev)
    rp = vp8g._RefPlanes(ry.data_ptr(), ru.data_ptr(), rv.data_ptr(), mc.mb_cols * 16, mc.mb_cols * 8,
                         mc.mb_cols, mc.mb_rows)
    lr = vp8g._LumaRef(ry.data_ptr(), mc.mb_cols * 16, mc.mb_cols * 16, mc.mb_rows * 16)
    filt = np.ascontiguousarray(vp8g.SIXTAP_FILTERS, np.int32)
    lib = ctx.lib
    torch.cuda.synchronize()

    def step():
        st = lib.vp8g_build_arnr_altref_dev(ctx.h, descs, n, group.center, 3, oy.data_ptr(),
                                            ou.data_ptr(), ov.data_ptr(), None, None)
        st |= lib.vp8g_predict_inter_mb_batch_dev(ctx.h, rp, xy.data_ptr(), mvs.data_ptr(), n_mb, 0,
                                                  filt.ctypes.data, pred.data_ptr())
        st |= lib.vp8g_mb_luma_sad_at_mv_batch_dev(ctx.h, lr, xy.data_ptr(), src.data_ptr(),
                                                   mvs.data_ptr(), n_mb, sad.data_ptr())
        if st:
            raise RuntimeError(f"vp8g launch failed: {st}")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ref_out, ref_pred, ref_sad = oy.clone(), pred.clone(), sad.clone()
    print(f"[bench] warmup done ({args.warmup} steps)", file=sys.stderr, flush=True)

    # The step: three C-ABI calls from Python on the library's stream, host
    # launch cost included (no events between the kernels).
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    start.record(stream)
    for _ in range(args.steps):
        step()
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    ms_step = start.elapsed_time(end) / args.steps
    if not (torch.equal(ref_out, oy) and torch.equal(ref_pred, pred) and torch.equal(ref_sad, sad)):
        raise RuntimeError("step output changed between steps")

    def capture(fn, times=1):
        """fn's launches, `times` over, as one HIP graph captured on a stream
        the library is pointed at."""
        g = torch.cuda.CUDAGraph()
        cap = torch.cuda.Stream(device=dev)
        torch.cuda.synchronize()
        ctx.set_stream(cap.cuda_stream)
        try:
            with torch.cuda.graph(g, stream=cap):
                for _ in range(times):
                    fn()
        finally:
            ctx.set_stream(None)
        torch.cuda.synchronize()
        return g, cap

    def replay_ms(g, cap, reps):
        with torch.cuda.stream(cap):
            g.replay()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(cap)
            for _ in range(reps):
                g.replay()
            b.record(cap)
        b.synchronize()
        return a.elapsed_time(b) / reps

    # The same step as one HIP graph replay (reported; not the headline).
    g_step, cap = capture(step)
    graph_ms_step = replay_ms(g_step, cap, args.steps)
    if not (torch.equal(ref_out, oy) and torch.equal(ref_pred, pred) and torch.equal(ref_sad, sad)):
        raise RuntimeError("graph replay output differs from the eager step")

    # Per-kernel durations for the roofline: each kernel alone, 10 launches
    # per graph replay, so the host never paces the GPU.
    def only(i):
        def fn():
            if i == 0:
                st = lib.vp8g_build_arnr_altref_dev(ctx.h, descs, n, group.center, 3, oy.data_ptr(),
                                                    ou.data_ptr(), ov.data_ptr(), None, None)
            elif i == 1:
                st = lib.vp8g_predict_inter_mb_batch_dev(ctx.h, rp, xy.data_ptr(), mvs.data_ptr(), n_mb, 0,
                                                         filt.ctypes.data, pred.data_ptr())
            else:
                st = lib.vp8g_mb_luma_sad_at_mv_batch_dev(ctx.h, lr, xy.data_ptr(), src.data_ptr(),
                                                          mvs.data_ptr(), n_mb, sad.data_ptr())
            if st:
                raise RuntimeError(f"vp8g launch failed: {st}")
        return fn
    kreps = max(2, args.steps // 10)
    arnr_ms, mc_ms, sad_ms = (replay_ms(*capture(only(i), 10), kreps) / 10 for i in range(3))
    if not torch.equal(ref_out, oy):
        raise RuntimeError("altref output changed between steps")
    print(f"[bench] timed {args.steps} steps: {ms_step:.4f} ms/step (one-graph replay {graph_ms_step:.4f}; "
          f"wall {wall:.3f}s)", file=sys.stderr, flush=True)

    anchors = ((w + 15) // 16) * ((h + 15) // 16)
    bytes_alg, ops_alg = arnr_algorithmic(w, h, n)
    achieved = bytes_alg / (arnr_ms * 1e-3) / 1e9
    traffic = None
    tpath = ROOT / "profiles" / "arnr_fused_1080p_traffic.json"
    if tpath.exists():
        traffic = json.loads(tpath.read_text()).get("hbm_bytes_per_launch")

    result = {
        "metric": METRIC,
        "value": anchors / (ms_step * 1e-3),
        "unit": "anchor-MB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {
            "workload": "1920x1080 I420, 7-frame group (centre 3, strength 3), 8x8 region motion "
                        "[-12,12] px/frame, 20 occluders/frame, sigma 4 noise; + predict_inter_mb "
                        "(six-tap) and mb_luma_sad_at_mv for all 8160 MBs at quarter-pel MVs in [-64,64]",
            "anchor_mbs_per_step": anchors,
            "group_frames": n,
            "parallelism": "single GPU",
            "launch": "three C-ABI calls per step from Python on the library stream",
        },
        "graph_ms_per_step": graph_ms_step,
        "kernel_timing": "each kernel alone, 10 launches per HIP graph replay",
        "kernels_ms": {"arnr_fused": arnr_ms, "predict_inter_mb": mc_ms, "mb_luma_sad_at_mv": sad_ms},
        "arnr_only_anchor_mb_per_s": anchors / (arnr_ms * 1e-3),
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "arnr_fused_kernel",
            "algorithmic_bytes_per_launch": bytes_alg,
        },
        "valu": {
            "ops_per_launch": ops_alg,
            "achieved_tops": ops_alg / (arnr_ms * 1e-3) / 1e12,
            "peak_tops": VALU_PEAK_OPS / 1e12,
            "frac": ops_alg / (arnr_ms * 1e-3) / VALU_PEAK_OPS,
        },
        "cpu_baseline": None,
    }
    # End to end through the host-pointer entry point: frames copied in,
    # the altref copied out (PCIe included); not the headline.
    frames_host = [vp8g.I420Frame(*t) for t in group.frames()] if not args.no_e2e else None
    if frames_host is not None:
        result["end_to_end"] = end_to_end(ctx, vp8g, frames_host, group, anchors, ref_out)
    h = hashlib.sha256()
    for t in group.frames():
        for p in t[2:5]:
            h.update(np.ascontiguousarray(p).tobytes())
    result["inputs"] = {"prng": "splitmix64", "seed_base": synth._seed(1, 0, 0), "group_sha256": h.hexdigest()}
    if not args.no_learned:
        result["learned_seeded"] = bench_learned(ctx, group, descs, stream, reps=max(5, args.steps // 5))
    if not args.no_extra:
        cores = max(1, min(os.cpu_count() or 1, 16))
        result["next_rows"] = measure_next_rows(ctx, group, cores)
        result["next_rows"]["loop_filter"] = measure_loop_filter(ctx, group, cores)
        result["c2_batch"] = measure_batch(ctx, max(5, min(args.steps, 20)), 2)
        c4 = measure_4k(ctx, reps=2)
        c4.pop("report")
        result["c4_pipeline"] = c4
    if not args.no_cpu:
        cores = max(1, min(os.cpu_count() or 1, 16))
        print(f"[bench] timing CPU oracle on {cores} threads", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(group, mc, args.cpu_seconds, cores)
        result["gpu_over_cpu"] = result["value"] / result["cpu_baseline"]["value"]
    ctx.close()
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
