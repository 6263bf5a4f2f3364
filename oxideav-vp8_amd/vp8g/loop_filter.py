"""The section 15 loop filter and the encoder's filter-level search over
the gfx950 C ABI (include/vp8g.h), under the reference's names:

* ``filter_frame_flags(ctx, planes, modes, has_coeffs, config)`` —
  src/loop_filter.rs:2506-2636 (key frames);
* ``filter_inter_frame_flags(ctx, planes, modes, has_coeffs, ref_frames,
  inter_modes, config)`` — src/loop_filter.rs:1364-1496;
* ``select_filter_params(ctx, planes, modes, has_coeffs, base, src,
  inter=None)`` / ``select_filter_level`` — src/loop_filter.rs:2017-2060,
  2133-2200;
* ``select_filter_params_sampled`` / ``select_filter_level_sampled`` —
  src/loop_filter.rs:2224-2412;
* ``score_candidates`` — the candidate scores of score_filter_candidate
  (src/loop_filter.rs:2062-2131).

Planes are filtered in place like the reference's ``&mut KeyframePlanes``.
Every call runs on the GPU; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

NO_SEGMENT = 255
ZEROMV, SPLITMV = 1, 2
BPRED, HAS_COEFFS = 1, 2
# RefFrame / InterMode as the C ABI numbers them
REF_INTRA, REF_LAST, REF_GOLDEN, REF_ALTREF = 0, 1, 2, 3
INTER_NEAREST, INTER_NEAR, INTER_ZERO, INTER_NEW, INTER_SPLIT = "nearest", "near", "zero", "new", "split"


class _Config(C.Structure):
    _fields_ = [("simple", C.c_int32), ("key_frame", C.c_int32), ("loop_filter_level", C.c_int32),
                ("sharpness_level", C.c_int32), ("segmentation_enabled", C.c_int32), ("segment_abs", C.c_int32),
                ("delta_enabled", C.c_int32), ("segment_lf_level", C.c_int16 * 4), ("ref_delta", C.c_int16 * 4),
                ("mode_delta", C.c_int16 * 4)]


class _Planes(C.Structure):
    _fields_ = [("y", C.c_void_p), ("u", C.c_void_p), ("v", C.c_void_p), ("y_stride", C.c_size_t),
                ("uv_stride", C.c_size_t), ("mb_cols", C.c_size_t), ("mb_rows", C.c_size_t)]


class _Source(C.Structure):
    _fields_ = [("width", C.c_size_t), ("height", C.c_size_t), ("y", C.c_void_p), ("u", C.c_void_p),
                ("v", C.c_void_p), ("y_stride", C.c_size_t), ("uv_stride", C.c_size_t)]


EXPORTS = ("vp8g_loop_filter_frame", "vp8g_loop_filter_frame_dev", "vp8g_loop_filter_score_dev",
           "vp8g_loop_filter_select", "vp8g_loop_filter_select_dev", "vp8g_loop_filter_status")


def bind(lib: C.CDLL) -> None:
    vp, sz, i32 = C.c_void_p, C.c_size_t, C.c_int
    P, S, G = C.POINTER(_Planes), C.POINTER(_Source), C.POINTER(_Config)
    sig = {
        "vp8g_loop_filter_frame": [vp, P, vp, G],
        "vp8g_loop_filter_frame_dev": [vp, P, vp, G],
        "vp8g_loop_filter_score_dev": [vp, P, vp, G, S, vp, vp, sz, sz, vp],
        "vp8g_loop_filter_select": [vp, P, vp, G, S, sz, i32, vp, vp],
        "vp8g_loop_filter_select_dev": [vp, P, vp, G, S, sz, i32, vp, vp],
        "vp8g_loop_filter_status": [vp],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.restype = i32
        fn.argtypes = args


@dataclass
class FrameFilterConfig:
    """src/loop_filter.rs:891-944."""
    simple: bool = False
    key_frame: bool = False
    loop_filter_level: int = 0
    sharpness_level: int = 0
    segmentation_enabled: bool = False
    segment_abs: bool = False
    segment_lf_level: tuple = (0, 0, 0, 0)
    delta_enabled: bool = False
    ref_delta_current: int = 0
    bpred_mode_delta: int = 0
    ref_delta_last: int = 0
    ref_delta_golden: int = 0
    ref_delta_altref: int = 0
    zero_mv_mode_delta: int = 0
    other_mv_mode_delta: int = 0
    split_mv_mode_delta: int = 0

    def ref_deltas(self) -> list:
        """src/loop_filter.rs:1193-1200."""
        return [self.ref_delta_current, self.ref_delta_last, self.ref_delta_golden, self.ref_delta_altref]

    def mode_deltas(self) -> list:
        """src/loop_filter.rs:1202-1209."""
        return [self.bpred_mode_delta, self.zero_mv_mode_delta, self.other_mv_mode_delta, self.split_mv_mode_delta]

    def c(self) -> _Config:
        g = _Config(int(bool(self.simple)), int(bool(self.key_frame)), int(self.loop_filter_level),
                    int(self.sharpness_level), int(bool(self.segmentation_enabled)), int(bool(self.segment_abs)),
                    int(bool(self.delta_enabled)))
        g.segment_lf_level[:] = [int(x) for x in self.segment_lf_level]
        g.ref_delta[:] = [int(x) for x in self.ref_deltas()]
        g.mode_delta[:] = [int(x) for x in self.mode_deltas()]
        return g


@dataclass
class KeyframePlanes:
    """The MB-padded reconstruction (src/frame.rs): y is (mb_rows*16,
    y_stride), u / v (mb_rows*8, uv_stride), uint8, C-contiguous."""
    y: np.ndarray
    u: np.ndarray
    v: np.ndarray
    mb_cols: int
    mb_rows: int

    @staticmethod
    def from_arrays(y, u, v, mb_cols: int | None = None) -> "KeyframePlanes":
        y, u, v = (np.ascontiguousarray(a, np.uint8) for a in (y, u, v))
        return KeyframePlanes(y, u, v, mb_cols if mb_cols is not None else y.shape[1] // 16, y.shape[0] // 16)

    def c(self) -> _Planes:
        return _Planes(self.y.ctypes.data, self.u.ctypes.data, self.v.ctypes.data, self.y.shape[1],
                       self.u.shape[1], self.mb_cols, self.mb_rows)


@dataclass
class SourcePlanes:
    """src/loop_filter.rs:1916-1932 (visible window and source planes)."""
    width: int
    height: int
    y: np.ndarray
    u: np.ndarray
    v: np.ndarray

    def c(self) -> _Source:
        return _Source(self.width, self.height, self.y.ctypes.data, self.u.ctypes.data, self.v.ctypes.data,
                       self.y.strides[0], self.u.strides[0])


@dataclass
class InterFilterInfo:
    """src/loop_filter.rs:1984-1990: per-MB reference frame (None = intra, or
    REF_LAST / REF_GOLDEN / REF_ALTREF) and inter mode (None for intra)."""
    ref_frames: list = field(default_factory=list)
    inter_modes: list = field(default_factory=list)


def mb_records(segment_ids, bpred, has_coeffs, inter: InterFilterInfo | None = None) -> np.ndarray:
    """The (n, 4) uint8 vp8g_lf_mb rows: segment id (None -> NO_SEGMENT),
    ref frame, inter mode, flags."""
    n = len(bpred)
    m = np.zeros((n, 4), np.uint8)
    for i in range(n):
        s = segment_ids[i] if segment_ids is not None else None
        m[i, 0] = NO_SEGMENT if s is None else s
        m[i, 3] = (BPRED if bpred[i] else 0) | (HAS_COEFFS if has_coeffs[i] else 0)
        if inter is not None and inter.ref_frames[i] is not None:
            m[i, 1] = inter.ref_frames[i]
            im = inter.inter_modes[i]
            m[i, 2] = ZEROMV if im == INTER_ZERO else (SPLITMV if im == INTER_SPLIT else 3)
    return m


class LoopFilterError(RuntimeError):
    def __init__(self, status: int, what: str):
        super().__init__(f"{what}: vp8g status {status}")
        self.status = status


def _lib(ctx):
    return ctx.lib


def _check(st: int, what: str) -> None:
    if st != 0:
        raise LoopFilterError(st, what)


def _mbs(mbs) -> np.ndarray:
    return np.ascontiguousarray(mbs, np.uint8).reshape(-1, 4)


def filter_mbs(ctx, planes: KeyframePlanes, mbs, config: FrameFilterConfig) -> None:
    """Filter `planes` in place from vp8g_lf_mb records."""
    m = _mbs(mbs)
    assert m.shape[0] == planes.mb_cols * planes.mb_rows
    _check(_lib(ctx).vp8g_loop_filter_frame(ctx.h, C.byref(planes.c()), m.ctypes.data, C.byref(config.c())),
           "vp8g_loop_filter_frame")


def filter_frame_flags(ctx, planes: KeyframePlanes, bpred, has_coeffs, config: FrameFilterConfig,
                       segment_ids=None) -> None:
    """src/loop_filter.rs:2506-2636; `bpred[i]` is y_mode == B_PRED."""
    filter_mbs(ctx, planes, mb_records(segment_ids, bpred, has_coeffs), config)


def filter_inter_frame_flags(ctx, planes: KeyframePlanes, bpred, has_coeffs, ref_frames, inter_modes,
                             config: FrameFilterConfig, segment_ids=None) -> None:
    """src/loop_filter.rs:1364-1496; SPLITMV MBs carry bpred (their recorded
    y_mode is B_PRED)."""
    filter_mbs(ctx, planes, mb_records(segment_ids, bpred, has_coeffs, InterFilterInfo(ref_frames, inter_modes)),
               config)


def score_candidates(ctx, planes: KeyframePlanes, mbs, base: FrameFilterConfig, src: SourcePlanes, candidates,
                     every_rows: int = 0) -> np.ndarray:
    """SSD of each (level, sharpness) candidate (score_filter_candidate,
    src/loop_filter.rs:2062-2095) through the device entry point; inputs
    are copied to the device first."""
    import torch
    dev = torch.device("cuda")
    cand = np.asarray(candidates, np.uint8).reshape(-1, 2)
    lv, sh = np.ascontiguousarray(cand[:, 0]), np.ascontiguousarray(cand[:, 1])
    ty, tu, tv = (torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (planes.y, planes.u, planes.v))
    keep = []

    def upload(a):
        # the plane's bytes at the same address offset mod 16 as on the host
        rows, width, stride = a.shape[0], a.shape[1], a.strides[0]
        n = (rows - 1) * stride + width if rows else 0
        off = a.ctypes.data % 16
        t = torch.zeros(off + n + 16, dtype=torch.uint8, device=dev)
        if n:
            host = np.ctypeslib.as_array((C.c_uint8 * n).from_address(a.ctypes.data))
            t[off:off + n] = torch.from_numpy(host.copy()).to(dev)
        keep.append(t)
        return t.data_ptr() + off
    sy, su, sv = (upload(a) for a in (src.y, src.u, src.v))
    tm = torch.from_numpy(_mbs(mbs).copy()).to(dev)
    out = torch.zeros(len(lv), dtype=torch.int64, device=dev)
    pl = _Planes(ty.data_ptr(), tu.data_ptr(), tv.data_ptr(), planes.y.shape[1], planes.u.shape[1],
                 planes.mb_cols, planes.mb_rows)
    so = _Source(src.width, src.height, sy, su, sv, src.y.strides[0], src.u.strides[0])
    lib = _lib(ctx)
    torch.cuda.synchronize()
    _check(lib.vp8g_loop_filter_score_dev(ctx.h, C.byref(pl), tm.data_ptr(), C.byref(base.c()), C.byref(so),
                                          lv.ctypes.data, sh.ctypes.data, len(lv), every_rows, out.data_ptr()),
           "vp8g_loop_filter_score_dev")
    _check(lib.vp8g_loop_filter_status(ctx.h), "vp8g_loop_filter_status")
    return out.cpu().numpy().astype(np.uint64)


def _select(ctx, planes, mbs, base, src, every_rows, sweep):
    lv, sh = C.c_uint8(), C.c_uint8()
    m = _mbs(mbs)
    _check(_lib(ctx).vp8g_loop_filter_select(ctx.h, C.byref(planes.c()), m.ctypes.data, C.byref(base.c()),
                                             C.byref(src.c()), every_rows, int(bool(sweep)), C.byref(lv),
                                             C.byref(sh)), "vp8g_loop_filter_select")
    return lv.value, sh.value


def select_filter_params(ctx, planes: KeyframePlanes, mbs, base: FrameFilterConfig, src: SourcePlanes) -> tuple:
    """src/loop_filter.rs:2047-2060 / 2133-2200: (level, sharpness)."""
    return _select(ctx, planes, mbs, base, src, 0, True)


def select_filter_level(ctx, planes: KeyframePlanes, mbs, base: FrameFilterConfig, src: SourcePlanes) -> int:
    """src/loop_filter.rs:2017-2026."""
    return select_filter_params(ctx, planes, mbs, base, src)[0]


def select_filter_params_sampled(ctx, planes: KeyframePlanes, mbs, base: FrameFilterConfig, src: SourcePlanes,
                                 every_rows: int, sweep_sharpness: bool) -> tuple:
    """src/loop_filter.rs:2288-2412."""
    return _select(ctx, planes, mbs, base, src, max(int(every_rows), 2), sweep_sharpness)


def select_filter_level_sampled(ctx, planes: KeyframePlanes, mbs, base: FrameFilterConfig, src: SourcePlanes,
                                every_rows: int) -> int:
    """src/loop_filter.rs:2224-2258."""
    return select_filter_params_sampled(ctx, planes, mbs, base, src, every_rows, False)[0]


def sampling_for_key_frame(mb_rows: int) -> int:
    """FilterSampling::for_key_frame's every_rows (src/loop_filter.rs:2260-2286)."""
    bands = min(max(mb_rows // 32, 3), 6)
    return max(mb_rows // bands, 2)
