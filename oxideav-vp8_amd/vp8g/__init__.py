"""Python host-side mirror of oxideav-vp8's ARNR / MC / SAD interface over
the gfx950 C ABI (include/vp8g.h).

Names, argument meaning and error behaviour follow the reference crate:

* ``build_arnr_altref(frames, center, cfg)`` — src/arnr.rs:476-482; raises
  ``InvalidDimensions`` / ``ReferenceDimensionsMismatch`` (the two
  ``EncodeError`` variants of src/encoder.rs:69-104 it can return) and
  ``IndexError`` where the reference panics (src/arnr.rs:494-498).
* ``predict_inter_mb(reference, mb_col, mb_row, luma_mv, full_pixel,
  filters)`` — src/motion_comp.rs:1986-2078, returns ``ReconstructedMb``.
* ``mb_luma_sad_at_mv(reference, mb_col, mb_row, src_y, mv)`` —
  src/motion_search.rs:470-529.
* ``half_pixel_refine_luma`` / ``quarter_pixel_refine_luma`` —
  src/motion_search.rs:564-728.

Every call runs on the GPU through libvp8g.so; there is no CPU path.  If
the library is missing or no gfx950 device is present the calls raise.
"""
from __future__ import annotations

import ctypes as C
import threading
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import loop_filter as _lf

_LIB_PATH = Path(__file__).resolve().parent.parent / "libvp8g.so"

VP8G_OK = 0
VP8G_E_INVALID_DIMENSIONS = 1
VP8G_E_REF_DIMS_MISMATCH = 2
VP8G_E_CENTER_OUT_OF_RANGE = 3
VP8G_E_ARG = 4
VP8G_E_DEVICE = 5
VP8G_E_NO_DEVICE = 6
MAX_FRAMES = 64
OPT_ARNR_WAVES = 1
OPT_XCD_REMAP = 2
OPT_ARNR_MBS = 3
OPT_BATCH_WGS_PER_CU = 4
OPT_FAULT = 5
OPT_HOST_THREADS = 6
OPT_COPY_STREAMS = 7
OPT_HOST_BANDS = 8
FAULT_NONE, FAULT_ALLOC, FAULT_COPY, FAULT_LAUNCH = 0, 1, 2, 3

# RFC 6386 section 18.3 tap tables, src/motion_comp.rs:294-303, 328-337.
SIXTAP_FILTERS = np.array(
    [[0, 0, 128, 0, 0, 0], [0, -6, 123, 12, -1, 0], [2, -11, 108, 36, -8, 1],
     [0, -9, 93, 50, -6, 0], [3, -16, 77, 77, -16, 3], [0, -6, 50, 93, -9, 0],
     [1, -8, 36, 108, -11, 2], [0, -1, 12, 123, -6, 0]], dtype=np.int32)
BILINEAR_FILTERS = np.array(
    [[0, 0, 128, 0, 0, 0], [0, 0, 112, 16, 0, 0], [0, 0, 96, 32, 0, 0], [0, 0, 80, 48, 0, 0],
     [0, 0, 64, 64, 0, 0], [0, 0, 48, 80, 0, 0], [0, 0, 32, 96, 0, 0], [0, 0, 16, 112, 0, 0]],
    dtype=np.int32)


def filter_set_for_version(version: int) -> np.ndarray:
    """src/motion_comp.rs:370-376: version 0 -> six-tap, else bilinear."""
    return SIXTAP_FILTERS if version == 0 else BILINEAR_FILTERS


class _I420(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("y", C.c_void_p),
                ("u", C.c_void_p), ("v", C.c_void_p), ("y_stride", C.c_size_t),
                ("uv_stride", C.c_size_t)]


class _Err(C.Structure):
    _fields_ = [("src_w", C.c_uint32), ("src_h", C.c_uint32), ("ref_w", C.c_uint32),
                ("ref_h", C.c_uint32)]


class _RefPlanes(C.Structure):
    _fields_ = [("y", C.c_void_p), ("u", C.c_void_p), ("v", C.c_void_p),
                ("y_stride", C.c_size_t), ("uv_stride", C.c_size_t),
                ("mb_cols", C.c_size_t), ("mb_rows", C.c_size_t)]


class _LumaRef(C.Structure):
    _fields_ = [("plane", C.c_void_p), ("stride", C.c_size_t), ("width", C.c_size_t),
                ("height", C.c_size_t)]


class _ArnrJob(C.Structure):
    _fields_ = [("frames", C.POINTER(_I420)), ("n_frames", C.c_size_t), ("center", C.c_size_t),
                ("strength", C.c_uint8), ("out_y", C.c_void_p), ("out_u", C.c_void_p),
                ("out_v", C.c_void_p), ("mv_out", C.c_void_p)]


class _SynthRect(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("x0", C.c_uint32), ("y0", C.c_uint32), ("w", C.c_uint32),
                ("h", C.c_uint32)]


class _SynthSpec(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("n", C.c_size_t),
                ("rects_per_frame", C.c_size_t), ("texture_seed", C.c_uint64),
                ("vel", C.c_void_p), ("rects", C.c_void_p), ("noise_seeds", C.c_void_p),
                ("sigma", C.c_double)]


EXPORTS = (
    "vp8g_ctx_create", "vp8g_ctx_destroy", "vp8g_ctx_set_stream", "vp8g_ctx_stream",
    "vp8g_ctx_synchronize", "vp8g_ctx_set_option", "vp8g_status_str",
    "vp8g_build_arnr_altref", "vp8g_build_arnr_altref_dev",
    "vp8g_predict_inter_mb_batch", "vp8g_predict_inter_mb_batch_dev",
    "vp8g_mb_luma_sad_at_mv_batch", "vp8g_mb_luma_sad_at_mv_batch_dev",
    "vp8g_half_pixel_refine_luma_batch", "vp8g_quarter_pixel_refine_luma_batch",
    "vp8g_half_plane_sads_batch",
    "vp8g_arnr_batch_create", "vp8g_arnr_batch_run", "vp8g_arnr_batch_shape",
    "vp8g_arnr_batch_destroy", "vp8g_ctx_load_tuning", "vp8g_tuning_lookup", "vp8g_arnr_shape",
    "vp8g_block_search_dev", "vp8g_build_arnr_altref_table_dev",
    "vp8g_downsample2_i420_dev", "vp8g_upsample_table_dev", "vp8g_refine_candidates_dev",
    "vp8g_proposer_param_count", "vp8g_proposer_create", "vp8g_proposer_destroy",
    "vp8g_proposer_set_params", "vp8g_proposer_get_params", "vp8g_proposer_get_grads",
    "vp8g_proposer_prepare_dev", "vp8g_proposer_forward_dev", "vp8g_proposer_propose_dev",
    "vp8g_proposer_backward_dev", "vp8g_proposer_train_step_dev", "vp8g_synth_luma_dev",
    "vp8g_half_pixel_refine_luma_batch_dev", "vp8g_quarter_pixel_refine_luma_batch_dev",
    "vp8g_search_fallback_dev", "vp8g_search_fallback_noalloc_dev",
    "vp8g_whole_pixel_search_batch", "vp8g_whole_pixel_search_batch_dev", "vp8g_sub_pixel_refine_batch",
    "vp8g_sub_pixel_refine_batch_dev", "vp8g_small_diamond_search_luma_batch",
    "vp8g_small_diamond_search_luma_batch_dev",
    "vp8g_predict_split_mv_batch", "vp8g_predict_split_mv_batch_dev", "vp8g_reconstruct_inter_mb_batch",
    "vp8g_reconstruct_inter_mb_batch_dev", "vp8g_reconstruct_split_mv_mb_batch",
    "vp8g_reconstruct_split_mv_mb_batch_dev", "vp8g_luma_mad_batch", "vp8g_luma_sad_batch_dev",
) + _lf.EXPORTS
FALLBACK_GATE_NEIGHBOUR = 1


def load_library(path: Path | None = None) -> C.CDLL:
    p = Path(path) if path else _LIB_PATH
    if not p.exists():
        raise RuntimeError(f"libvp8g.so not built at {p}; run oxideav-vp8_amd/build.py")
    lib = C.CDLL(str(p))
    vp, sz, i32, u8 = C.c_void_p, C.c_size_t, C.c_int, C.c_uint8
    sig = {
        "vp8g_ctx_create": (i32, [i32, C.POINTER(vp)]),
        "vp8g_ctx_destroy": (None, [vp]),
        "vp8g_ctx_set_stream": (i32, [vp, vp]),
        "vp8g_ctx_stream": (vp, [vp]),
        "vp8g_ctx_synchronize": (i32, [vp]),
        "vp8g_ctx_set_option": (i32, [vp, i32, i32]),
        "vp8g_status_str": (C.c_char_p, [i32]),
        "vp8g_build_arnr_altref": (i32, [vp, C.POINTER(_I420), sz, sz, u8, vp, vp, vp, C.POINTER(_Err)]),
        "vp8g_build_arnr_altref_dev": (i32, [vp, C.POINTER(_I420), sz, sz, u8, vp, vp, vp,
                                             C.POINTER(_Err), vp]),
        "vp8g_predict_inter_mb_batch": (i32, [vp, C.POINTER(_RefPlanes), vp, vp, sz, i32, vp, vp]),
        "vp8g_predict_inter_mb_batch_dev": (i32, [vp, C.POINTER(_RefPlanes), vp, vp, sz, i32, vp, vp]),
        "vp8g_mb_luma_sad_at_mv_batch": (i32, [vp, C.POINTER(_LumaRef), vp, vp, vp, sz, vp]),
        "vp8g_mb_luma_sad_at_mv_batch_dev": (i32, [vp, C.POINTER(_LumaRef), vp, vp, vp, sz, vp]),
        "vp8g_half_pixel_refine_luma_batch": (i32, [vp, C.POINTER(_LumaRef), vp, vp, vp, sz, vp]),
        "vp8g_quarter_pixel_refine_luma_batch": (i32, [vp, C.POINTER(_LumaRef), vp, vp, vp, sz, vp]),
        "vp8g_half_plane_sads_batch": (i32, [vp, C.POINTER(_LumaRef), vp, vp, sz, vp, i32, vp]),
        "vp8g_arnr_batch_create": (i32, [vp, C.POINTER(_ArnrJob), sz, C.POINTER(vp),
                                         C.POINTER(sz), C.POINTER(_Err)]),
        "vp8g_arnr_batch_run": (i32, [vp]),
        "vp8g_arnr_batch_shape": (i32, [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32),
                                        C.POINTER(sz)]),
        "vp8g_arnr_batch_destroy": (None, [vp]),
        "vp8g_ctx_load_tuning": (i32, [vp, C.c_char_p]),
        "vp8g_tuning_lookup": (i32, [C.c_char_p, C.c_uint32, C.c_uint32, sz, i32, C.POINTER(i32),
                                     C.POINTER(i32), C.POINTER(i32)]),
        "vp8g_arnr_shape": (i32, [vp, C.c_uint32, C.c_uint32, sz, u8, C.POINTER(i32), C.POINTER(i32)]),
        "vp8g_block_search_dev": (i32, [vp, C.POINTER(_I420), sz, sz, i32, vp, vp, C.POINTER(_Err)]),
        "vp8g_downsample2_i420_dev": (i32, [vp, C.POINTER(_I420), C.POINTER(_I420), sz]),
        "vp8g_refine_candidates_dev": (i32, [vp, C.POINTER(_I420), sz, sz, vp, vp, vp, C.POINTER(_Err)]),
        "vp8g_search_fallback_dev": (i32, [vp, C.POINTER(_I420), sz, sz, vp, i32, i32, vp, vp, C.POINTER(_Err)]),
        "vp8g_search_fallback_noalloc_dev": (i32, [vp, C.POINTER(_I420), sz, sz, vp, i32, i32, vp, vp,
                                                   C.POINTER(_Err)]),
        "vp8g_luma_mad_batch": (i32, [vp, C.POINTER(_I420), C.POINTER(_I420), sz, vp]),
        "vp8g_luma_sad_batch_dev": (i32, [vp, C.POINTER(_I420), C.POINTER(_I420), sz, vp]),
        "vp8g_predict_split_mv_batch": (i32, [vp, C.POINTER(_RefPlanes), vp, vp, sz, i32, vp, vp]),
        "vp8g_predict_split_mv_batch_dev": (i32, [vp, C.POINTER(_RefPlanes), vp, vp, sz, i32, vp, vp]),
        "vp8g_reconstruct_inter_mb_batch": (i32, [vp, C.POINTER(_RefPlanes), vp, vp, sz, i32, vp, vp, vp, vp]),
        "vp8g_reconstruct_inter_mb_batch_dev": (i32, [vp, C.POINTER(_RefPlanes), vp, vp, sz, i32, vp, vp, vp, vp]),
        "vp8g_reconstruct_split_mv_mb_batch": (i32, [vp, C.POINTER(_RefPlanes), vp, vp, sz, i32, vp, vp, vp, vp]),
        "vp8g_reconstruct_split_mv_mb_batch_dev": (i32, [vp, C.POINTER(_RefPlanes), vp, vp, sz, i32, vp, vp, vp,
                                                         vp]),
        "vp8g_whole_pixel_search_batch": (i32, [vp, C.POINTER(_LumaRef), vp, vp, vp, sz, vp, vp]),
        "vp8g_whole_pixel_search_batch_dev": (i32, [vp, C.POINTER(_LumaRef), vp, vp, vp, sz, vp, vp]),
        "vp8g_sub_pixel_refine_batch": (i32, [vp, C.POINTER(_LumaRef), vp, vp, vp, vp, sz, vp, vp, vp]),
        "vp8g_sub_pixel_refine_batch_dev": (i32, [vp, C.POINTER(_LumaRef), vp, vp, vp, vp, sz, vp, vp, vp]),
        "vp8g_small_diamond_search_luma_batch": (i32, [vp, C.POINTER(_LumaRef), vp, vp, vp, sz, C.c_uint32, vp]),
        "vp8g_small_diamond_search_luma_batch_dev": (i32, [vp, C.POINTER(_LumaRef), vp, vp, vp, sz, C.c_uint32,
                                                           vp]),
        "vp8g_upsample_table_dev": (i32, [vp, vp, sz, i32, i32, vp, i32, i32]),
        "vp8g_build_arnr_altref_table_dev": (i32, [vp, C.POINTER(_I420), sz, sz, u8, vp, vp, vp, vp,
                                                   C.POINTER(_Err)]),
        "vp8g_proposer_param_count": (sz, []),
        "vp8g_proposer_create": (i32, [vp, C.POINTER(vp)]),
        "vp8g_proposer_destroy": (None, [vp]),
        "vp8g_proposer_set_params": (i32, [vp, vp, sz]),
        "vp8g_proposer_get_params": (i32, [vp, vp, sz]),
        "vp8g_proposer_get_grads": (i32, [vp, vp, sz, C.POINTER(C.c_float)]),
        "vp8g_proposer_prepare_dev": (i32, [vp, C.POINTER(_I420), sz, sz, vp]),
        "vp8g_proposer_forward_dev": (i32, [vp, vp, i32, i32, i32, vp]),
        "vp8g_proposer_propose_dev": (i32, [vp, C.POINTER(_I420), sz, sz, i32, vp]),
        "vp8g_proposer_backward_dev": (i32, [vp, vp, vp, i32, i32, i32]),
        "vp8g_proposer_train_step_dev": (i32, [vp, vp, vp, i32, i32, i32, C.c_float, vp]),
        "vp8g_synth_luma_dev": (i32, [vp, C.POINTER(_SynthSpec), vp, vp]),
        "vp8g_half_pixel_refine_luma_batch_dev": (i32, [vp, C.POINTER(_LumaRef), vp, vp, vp, sz, vp]),
        "vp8g_quarter_pixel_refine_luma_batch_dev": (i32, [vp, C.POINTER(_LumaRef), vp, vp, vp, sz, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lf.bind(lib)
    return lib


# ---------------------------------------------------------------- errors
class EncodeError(Exception):
    """src/encoder.rs:69-135 (the variants build_arnr_altref can return)."""


class InvalidDimensions(EncodeError):
    def __init__(self, width: int, height: int):
        super().__init__(f"InvalidDimensions {{ width: {width}, height: {height} }}")
        self.width, self.height = width, height


class ReferenceDimensionsMismatch(EncodeError):
    def __init__(self, source: tuple, reference: tuple):
        super().__init__(f"ReferenceDimensionsMismatch {{ source: {source}, reference: {reference} }}")
        self.source, self.reference = source, reference


class Vp8gError(RuntimeError):
    def __init__(self, status: int, what: str):
        super().__init__(f"{what}: status {status}")
        self.status = status


# ---------------------------------------------------------------- types
@dataclass
class I420Frame:
    """src/encoder.rs:5419-5434.  Planes are 1-D uint8 arrays."""
    width: int
    height: int
    y: np.ndarray
    u: np.ndarray
    v: np.ndarray
    y_stride: int
    uv_stride: int

    @staticmethod
    def packed(width: int, height: int, y, u, v) -> "I420Frame":
        """I420Frame::packed, src/encoder.rs:5510-5520."""
        return I420Frame(width, height, np.ascontiguousarray(y, np.uint8).reshape(-1),
                         np.ascontiguousarray(u, np.uint8).reshape(-1),
                         np.ascontiguousarray(v, np.uint8).reshape(-1),
                         width, (width + 1) // 2)


@dataclass(frozen=True)
class ArnrConfig:
    """src/arnr.rs:52-83."""
    strength: int = 3
    MAX_STRENGTH = 6

    @staticmethod
    def new(strength: int) -> "ArnrConfig":
        return ArnrConfig(min(int(strength), ArnrConfig.MAX_STRENGTH))

    def with_strength(self, v: int) -> "ArnrConfig":
        return ArnrConfig(int(v))


@dataclass
class ArnrPicture:
    """src/arnr.rs:90-145: tightly packed output planes."""
    width: int
    height: int
    y: np.ndarray
    u: np.ndarray
    v: np.ndarray

    def as_i420(self) -> I420Frame:
        return I420Frame.packed(self.width, self.height, self.y, self.u, self.v)


@dataclass(frozen=True)
class Mv:
    """src/motion_vector.rs:64-70 (quarter-pel)."""
    row: int = 0
    col: int = 0


@dataclass
class ReferencePlanes:
    """src/motion_comp.rs:131-146 (MB-aligned planes)."""
    y: np.ndarray
    u: np.ndarray
    v: np.ndarray
    y_stride: int
    uv_stride: int
    mb_cols: int
    mb_rows: int


@dataclass
class LumaRef:
    """src/motion_search.rs:230-239."""
    plane: np.ndarray
    stride: int
    width: int
    height: int


@dataclass
class ReconstructedMb:
    """src/reconstruct.rs:176-195."""
    y: np.ndarray = field(default_factory=lambda: np.zeros(256, np.uint8))
    u: np.ndarray = field(default_factory=lambda: np.zeros(64, np.uint8))
    v: np.ndarray = field(default_factory=lambda: np.zeros(64, np.uint8))


@dataclass(frozen=True)
class SearchResult:
    """src/motion_search.rs:248-255."""
    mv: Mv
    sad: int


# ---------------------------------------------------------------- context
def _ptr(a: np.ndarray) -> int:
    # the array interface's address: a third cheaper than a.ctypes.data,
    # which builds a ctypes helper per call (24 per ARNR group call)
    return a.__array_interface__["data"][0]


def _check_plane(name: str, a: np.ndarray, need: int) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint8).reshape(-1)
    if a.size < need:
        raise ValueError(f"{name} plane holds {a.size} bytes, needs {need}")
    return a


class Context:
    """One device context (a HIP stream plus a staging arena).  Not locked:
    use one per thread."""

    def __init__(self, device: int = 0, lib_path: Path | None = None):
        self.lib = load_library(lib_path)
        h = C.c_void_p()
        st = self.lib.vp8g_ctx_create(device, C.byref(h))
        if st != VP8G_OK:
            raise Vp8gError(st, f"vp8g_ctx_create(device={device}) failed "
                                f"({self.lib.vp8g_status_str(st).decode()})")
        self.h = h

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.vp8g_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _raise(self, st: int, what: str):
        if st != VP8G_OK:
            raise Vp8gError(st, f"{what} ({self.lib.vp8g_status_str(st).decode()})")

    def set_stream(self, stream: int | None) -> None:
        self._raise(self.lib.vp8g_ctx_set_stream(self.h, stream), "set_stream")

    def stream(self) -> int:
        return self.lib.vp8g_ctx_stream(self.h) or 0

    def synchronize(self) -> None:
        self._raise(self.lib.vp8g_ctx_synchronize(self.h), "synchronize")

    def set_arnr_waves(self, waves: int) -> None:
        self._raise(self.lib.vp8g_ctx_set_option(self.h, OPT_ARNR_WAVES, waves), "set_arnr_waves")

    def set_arnr_mbs(self, mbs: int) -> None:
        self._raise(self.lib.vp8g_ctx_set_option(self.h, OPT_ARNR_MBS, mbs), "set_arnr_mbs")

    def set_xcd_remap(self, on: bool) -> None:
        self._raise(self.lib.vp8g_ctx_set_option(self.h, OPT_XCD_REMAP, int(on)), "set_xcd_remap")

    def set_batch_wgs_per_cu(self, wgs: int) -> None:
        self._raise(self.lib.vp8g_ctx_set_option(self.h, OPT_BATCH_WGS_PER_CU, wgs),
                    "set_batch_wgs_per_cu")

    def set_host_threads(self, threads: int) -> None:
        """Host threads of build_arnr_altref's packing and copy-out (0 = default)."""
        self._raise(self.lib.vp8g_ctx_set_option(self.h, OPT_HOST_THREADS, threads), "set_host_threads")

    def set_copy_streams(self, streams: int) -> None:
        """Accepted (0..4) and without effect: build_arnr_altref uses one copy stream."""
        self._raise(self.lib.vp8g_ctx_set_option(self.h, OPT_COPY_STREAMS, streams), "set_copy_streams")

    def set_host_bands(self, bands: int) -> None:
        """Row bands of the host-buffer ARNR entry point (1..8; 0 = default 3)."""
        self._raise(self.lib.vp8g_ctx_set_option(self.h, OPT_HOST_BANDS, int(bands)), "set_host_bands")

    def set_fault(self, mode: int) -> None:
        """Host-side fault injection (FAULT_ALLOC / FAULT_COPY / FAULT_LAUNCH,
        FAULT_NONE to clear): the chosen step of later calls reports
        VP8G_E_DEVICE without running and writes no output."""
        self._raise(self.lib.vp8g_ctx_set_option(self.h, OPT_FAULT, int(mode)), "set_fault")

    def load_tuning(self, path: str | Path | None) -> None:
        """Load a tuning table (None clears it); see vp8g.autotune."""
        self._raise(self.lib.vp8g_ctx_load_tuning(self.h, None if path is None else str(path).encode()),
                    "vp8g_ctx_load_tuning")

    def arnr_shape(self, width: int, height: int, n_frames: int, strength: int = 3) -> dict:
        m, w = C.c_int(), C.c_int()
        self._raise(self.lib.vp8g_arnr_shape(self.h, width, height, n_frames, strength, C.byref(m),
                                             C.byref(w)), "vp8g_arnr_shape")
        return {"mbs": m.value, "waves": w.value}

    def apply_tuning(self, entry: dict) -> None:
        """Set the launch-shape options of one tuning-table entry
        (keys ``mbs``, ``waves``, ``wgs_per_cu``; missing keys keep 0 =
        library default)."""
        self.set_arnr_mbs(int(entry.get("mbs", 0)))
        self.set_arnr_waves(int(entry.get("waves", 0)))
        self.set_batch_wgs_per_cu(int(entry.get("wgs_per_cu", 0)))

    # ---- ARNR
    def build_arnr_altref(self, frames, center: int, cfg: ArnrConfig | None = None) -> ArnrPicture:
        cfg = cfg or ArnrConfig()
        frames = list(frames)
        n = len(frames)
        if center >= n:
            raise IndexError(f"build_arnr_altref: center index {center} out of range ({n} frames)")
        keep, descs = [], (_I420 * max(n, 1))()
        ctr = frames[center]
        cw0 = (ctr.width + 1) // 2
        for i, f in enumerate(frames):
            y = np.ascontiguousarray(f.y, np.uint8).reshape(-1)
            u = np.ascontiguousarray(f.u, np.uint8).reshape(-1)
            v = np.ascontiguousarray(f.v, np.uint8).reshape(-1)
            keep += [y, u, v]
            # Only well-formed planes are dereferenced; the dimension checks
            # below run first, as in the reference.
            if f.width == ctr.width and f.height == ctr.height and 0 < f.width <= 0x3FFF \
                    and 0 < f.height <= 0x3FFF:
                need_y = (f.height - 1) * f.y_stride + f.width
                need_c = ((f.height + 1) // 2 - 1) * f.uv_stride + cw0
                if y.size < need_y or u.size < need_c or v.size < need_c:
                    raise ValueError(f"frame {i}: plane buffers shorter than stride x height")
            descs[i] = _I420(f.width, f.height, _ptr(y) if y.size else None,
                             _ptr(u) if u.size else None, _ptr(v) if v.size else None,
                             f.y_stride, f.uv_stride)
        w, h = ctr.width, ctr.height
        cw, ch = (w + 1) // 2, (h + 1) // 2
        oy = np.empty(max(w * h, 1), np.uint8)
        ou = np.empty(max(cw * ch, 1), np.uint8)
        ov = np.empty(max(cw * ch, 1), np.uint8)
        err = _Err()
        st = self.lib.vp8g_build_arnr_altref(self.h, descs, n, center, int(cfg.strength) & 0xFF,
                                             _ptr(oy), _ptr(ou), _ptr(ov), C.byref(err))
        if st == VP8G_E_INVALID_DIMENSIONS:
            raise InvalidDimensions(err.src_w, err.src_h)
        if st == VP8G_E_REF_DIMS_MISMATCH:
            raise ReferenceDimensionsMismatch((err.src_w, err.src_h), (err.ref_w, err.ref_h))
        if st == VP8G_E_CENTER_OUT_OF_RANGE:
            raise IndexError(f"build_arnr_altref: center index {center} out of range ({n} frames)")
        self._raise(st, "vp8g_build_arnr_altref")
        return ArnrPicture(w, h, oy[:w * h], ou[:cw * ch], ov[:cw * ch])

    def build_arnr_altref_dev(self, descs, n: int, center: int, strength: int,
                              oy: int, ou: int, ov: int, mv_out: int | None = None) -> None:
        """Device-resident form: ``descs`` is a ctypes ``_I420`` array holding
        device pointers; outputs are device pointers.  Enqueues only."""
        err = _Err()
        st = self.lib.vp8g_build_arnr_altref_dev(self.h, descs, n, center, strength, oy, ou, ov,
                                                 C.byref(err), mv_out)
        self._raise(st, "vp8g_build_arnr_altref_dev")

    def block_search_dev(self, descs, n: int, center: int, radius: int, out: int,
                         seeds: int | None = None) -> None:
        """Exhaustive +-radius block search over device-resident frames into
        the alignment-table layout at device pointer ``out``.  Enqueues only."""
        err = _Err()
        st = self.lib.vp8g_block_search_dev(self.h, descs, n, center, radius, seeds, out, C.byref(err))
        self._raise(st, "vp8g_block_search_dev")

    def refine_candidates_dev(self, descs, n: int, center: int, seeds: int | None, first: int, out: int) -> None:
        """Second stage of the seeded search: exact +-2 around each block's own
        first-stage vector (when it lies on the first square's border) and its
        4-neighbours' first-stage vectors, keeping the best.  Enqueues only."""
        err = _Err()
        st = self.lib.vp8g_refine_candidates_dev(self.h, descs, n, center, seeds, first, out, C.byref(err))
        self._raise(st, "vp8g_refine_candidates_dev")

    def search_fallback_dev(self, descs, n: int, center: int, first: int, out: int, radius: int = 16,
                            gate: bool = True, queued: int | None = None) -> None:
        """Exhaustive fallback of a seeded search: entries of the device table
        ``first`` whose SAD fails the occlusion cut-off (with ``gate``, only
        those with a 4-neighbour under it) are replaced in ``out`` by the
        better of themselves and the exhaustive +-radius search; ``queued``
        (device uint32) receives the count.  Enqueues only."""
        err = _Err()
        flags = FALLBACK_GATE_NEIGHBOUR if gate else 0
        st = self.lib.vp8g_search_fallback_dev(self.h, descs, n, center, first, radius, flags, out, queued,
                                               C.byref(err))
        self._raise(st, "vp8g_search_fallback_dev")

    def downsample2_dev(self, src, dst, n: int = 1) -> None:
        """One pyramid level of ``n`` frames: ``dst`` (``_I420`` descriptors of
        ceil(w/2) x ceil(h/2), device planes; one descriptor or an array)
        receives the 2x2 rounded means of ``src``'s Y, U and V.  Enqueues
        only."""
        sp = src if isinstance(src, C.Array) else C.pointer(src)
        dp = dst if isinstance(dst, C.Array) else C.pointer(dst)
        self._raise(self.lib.vp8g_downsample2_i420_dev(self.h, sp, dp, n), "vp8g_downsample2_i420_dev")

    def upsample_table_dev(self, coarse: int, n: int, coarse_w: int, coarse_h: int, fine: int, fine_w: int,
                           fine_h: int) -> None:
        """Seeds one pyramid level finer: fine block (r, c) takes coarse block
        (r // 2, c // 2) (clamped) with its vector doubled.  Enqueues only."""
        self._raise(self.lib.vp8g_upsample_table_dev(self.h, coarse, n, coarse_w, coarse_h, fine, fine_w, fine_h),
                    "vp8g_upsample_table_dev")

    def build_arnr_altref_table_dev(self, descs, n: int, center: int, strength: int, table: int,
                                    oy: int, ou: int, ov: int) -> None:
        """The altref from a given device alignment table.  Enqueues only."""
        err = _Err()
        st = self.lib.vp8g_build_arnr_altref_table_dev(self.h, descs, n, center, strength, table, oy,
                                                       ou, ov, C.byref(err))
        self._raise(st, "vp8g_build_arnr_altref_table_dev")

    def create_arnr_batch(self, jobs) -> "ArnrBatch":
        """Batched ARNR over device-resident groups.  ``jobs`` is a sequence
        of ``ArnrJob``; raises as ``build_arnr_altref`` does for the first
        job that fails validation (the exception carries ``job``)."""
        jobs = list(jobs)
        arr = (_ArnrJob * max(len(jobs), 1))()
        keep = []
        for i, j in enumerate(jobs):
            descs = (_I420 * max(len(j.frames), 1))()
            for k, f in enumerate(j.frames):
                descs[k] = _I420(*f)
            keep.append(descs)
            arr[i] = _ArnrJob(descs, len(j.frames), j.center, int(j.strength) & 0xFF,
                              j.out_y, j.out_u, j.out_v, j.mv_out)
        h = C.c_void_p()
        failed = C.c_size_t(0)
        err = _Err()
        st = self.lib.vp8g_arnr_batch_create(self.h, arr, len(jobs), C.byref(h), C.byref(failed),
                                             C.byref(err))
        exc = None
        if st == VP8G_E_INVALID_DIMENSIONS:
            exc = InvalidDimensions(err.src_w, err.src_h)
        elif st == VP8G_E_REF_DIMS_MISMATCH:
            exc = ReferenceDimensionsMismatch((err.src_w, err.src_h), (err.ref_w, err.ref_h))
        elif st == VP8G_E_CENTER_OUT_OF_RANGE:
            exc = IndexError(f"arnr batch job {failed.value}: center index out of range")
        if exc is not None:
            exc.job = failed.value
            raise exc
        self._raise(st, "vp8g_arnr_batch_create")
        return ArnrBatch(self, h)

    # ---- motion compensation
    def predict_inter_mb_batch(self, reference: ReferencePlanes, mb_xy, mvs, full_pixel: bool,
                               filters) -> np.ndarray:
        mb_xy = np.ascontiguousarray(mb_xy, np.uint32).reshape(-1, 2)
        mvs = np.ascontiguousarray(mvs, np.int16).reshape(-1, 2)
        n = mb_xy.shape[0]
        if mvs.shape[0] != n:
            raise ValueError("one MV per macroblock")
        filt = np.ascontiguousarray(filters, np.int32).reshape(8, 6)
        ys = reference.y_stride
        y = _check_plane("y", reference.y, (reference.mb_rows * 16 - 1) * ys + reference.mb_cols * 16)
        cneed = (reference.mb_rows * 8 - 1) * reference.uv_stride + reference.mb_cols * 8
        u = _check_plane("u", reference.u, cneed)
        v = _check_plane("v", reference.v, cneed)
        rp = _RefPlanes(_ptr(y), _ptr(u), _ptr(v), ys, reference.uv_stride, reference.mb_cols,
                        reference.mb_rows)
        out = np.empty((max(n, 1), 384), np.uint8)
        st = self.lib.vp8g_predict_inter_mb_batch(self.h, C.byref(rp), _ptr(mb_xy), _ptr(mvs), n,
                                                  1 if full_pixel else 0, _ptr(filt), _ptr(out))
        self._raise(st, "vp8g_predict_inter_mb_batch")
        return out[:n]

    def _ref_planes(self, reference: ReferencePlanes):
        ys = reference.y_stride
        y = _check_plane("y", reference.y, (reference.mb_rows * 16 - 1) * ys + reference.mb_cols * 16)
        cneed = (reference.mb_rows * 8 - 1) * reference.uv_stride + reference.mb_cols * 8
        u = _check_plane("u", reference.u, cneed)
        v = _check_plane("v", reference.v, cneed)
        return (y, u, v), _RefPlanes(_ptr(y), _ptr(u), _ptr(v), ys, reference.uv_stride, reference.mb_cols,
                                     reference.mb_rows)

    def _recon(self, fn, what, reference, mb_xy, mvs, per_mb, full_pixel, filters, skip=None, coeffs=None,
               blocks=0):
        mb_xy = np.ascontiguousarray(mb_xy, np.uint32).reshape(-1, 2)
        n = mb_xy.shape[0]
        mvs = np.ascontiguousarray(mvs, np.int16).reshape(n, per_mb, 2)
        filt = np.ascontiguousarray(filters, np.int32).reshape(8, 6)
        keep, rp = self._ref_planes(reference)
        out = np.empty((max(n, 1), 384), np.uint8)
        args = [self.h, C.byref(rp), _ptr(mb_xy), _ptr(mvs), n, 1 if full_pixel else 0, _ptr(filt)]
        if blocks:
            sk = np.ascontiguousarray(skip, np.uint8).reshape(n)
            cf = np.ascontiguousarray(coeffs, np.int16).reshape(n, blocks, 16)
            args += [_ptr(sk), _ptr(cf)]
        self._raise(fn(*args, _ptr(out)), what)
        return out[:n]

    def predict_split_mv_batch(self, reference: ReferencePlanes, mb_xy, split_mvs, full_pixel: bool,
                               filters) -> np.ndarray:
        """predict_split_mv per macroblock: split_mvs is (n, 16, 2) of (row,
        col) quarter-pel sub-block vectors; returns (n, 384)."""
        return self._recon(self.lib.vp8g_predict_split_mv_batch, "vp8g_predict_split_mv_batch", reference, mb_xy,
                           split_mvs, 16, full_pixel, filters)

    def reconstruct_inter_mb_batch(self, reference: ReferencePlanes, mb_xy, mvs, full_pixel: bool, filters, skip,
                                   coeffs) -> np.ndarray:
        """reconstruct_inter_mb per macroblock; coeffs (n, 25, 16) = y2, y[16],
        u[4], v[4] (dequantised)."""
        return self._recon(self.lib.vp8g_reconstruct_inter_mb_batch, "vp8g_reconstruct_inter_mb_batch", reference,
                           mb_xy, mvs, 1, full_pixel, filters, skip, coeffs, 25)

    def reconstruct_split_mv_mb_batch(self, reference: ReferencePlanes, mb_xy, split_mvs, full_pixel: bool, filters,
                                      skip, coeffs) -> np.ndarray:
        """reconstruct_split_mv_mb per macroblock; coeffs (n, 24, 16) = y[16],
        u[4], v[4] (no Y2)."""
        return self._recon(self.lib.vp8g_reconstruct_split_mv_mb_batch, "vp8g_reconstruct_split_mv_mb_batch",
                           reference, mb_xy, split_mvs, 16, full_pixel, filters, skip, coeffs, 24)

    def predict_split_mv(self, reference: ReferencePlanes, mb_col: int, mb_row: int, split_luma_mvs,
                         full_pixel: bool, filters) -> ReconstructedMb:
        """src/motion_comp.rs:2227-2320."""
        mvs = [[m.row, m.col] if isinstance(m, Mv) else list(m) for m in split_luma_mvs]
        o = self.predict_split_mv_batch(reference, [[mb_col, mb_row]], [mvs], full_pixel, filters)[0]
        return ReconstructedMb(o[:256].copy(), o[256:320].copy(), o[320:].copy())

    def predict_inter_mb(self, reference: ReferencePlanes, mb_col: int, mb_row: int, luma_mv: Mv,
                         full_pixel: bool, filters) -> ReconstructedMb:
        o = self.predict_inter_mb_batch(reference, [[mb_col, mb_row]], [[luma_mv.row, luma_mv.col]],
                                        full_pixel, filters)[0]
        return ReconstructedMb(o[:256].copy(), o[256:320].copy(), o[320:].copy())

    # ---- SAD scoring
    def _luma_ref(self, reference: LumaRef):
        p = _check_plane("luma", reference.plane,
                         (reference.height - 1) * reference.stride + reference.width)
        return p, _LumaRef(_ptr(p), reference.stride, reference.width, reference.height)

    def mb_luma_sad_at_mv_batch(self, reference: LumaRef, mb_xy, src, mvs) -> np.ndarray:
        mb_xy = np.ascontiguousarray(mb_xy, np.uint32).reshape(-1, 2)
        n = mb_xy.shape[0]
        src = np.ascontiguousarray(src, np.uint8).reshape(n, 256)
        mvs = np.ascontiguousarray(mvs, np.int16).reshape(n, 2)
        p, lr = self._luma_ref(reference)
        out = np.empty(max(n, 1), np.uint32)
        st = self.lib.vp8g_mb_luma_sad_at_mv_batch(self.h, C.byref(lr), _ptr(mb_xy), _ptr(src),
                                                   _ptr(mvs), n, _ptr(out))
        self._raise(st, "vp8g_mb_luma_sad_at_mv_batch")
        return out[:n]

    def mb_luma_sad_at_mv(self, reference: LumaRef, mb_col: int, mb_row: int, src_y, mv: Mv) -> int:
        return int(self.mb_luma_sad_at_mv_batch(reference, [[mb_col, mb_row]], src_y,
                                                [[mv.row, mv.col]])[0])

    def _refine(self, fn, reference: LumaRef, mb_xy, src, centers) -> np.ndarray:
        mb_xy = np.ascontiguousarray(mb_xy, np.uint32).reshape(-1, 2)
        n = mb_xy.shape[0]
        src = np.ascontiguousarray(src, np.uint8).reshape(n, 256)
        centers = np.ascontiguousarray(centers, np.int16).reshape(n, 2)
        p, lr = self._luma_ref(reference)
        out = np.empty((max(n, 1), 4), np.uint16)  # row, col, sad lo, sad hi
        st = fn(self.h, C.byref(lr), _ptr(mb_xy), _ptr(src), _ptr(centers), n, _ptr(out))
        self._raise(st, "refine ladder")
        res = np.empty((n, 3), np.int64)
        res[:, 0] = out[:n, 0].view(np.int16)
        res[:, 1] = out[:n, 1].view(np.int16)
        res[:, 2] = out[:n, 2].astype(np.int64) | (out[:n, 3].astype(np.int64) << 16)
        return res

    def refine_batch_dev(self, plane: int, stride: int, width: int, height: int, mb_xy: int, src: int,
                         centers: int, n: int, out: int, quarter: bool) -> None:
        """Half- (quarter=False) or quarter-pel refinement over device arrays
        (vp8g_search_result out[n]); enqueues only."""
        ref = _LumaRef(plane, stride, width, height)
        fn = (self.lib.vp8g_quarter_pixel_refine_luma_batch_dev if quarter
              else self.lib.vp8g_half_pixel_refine_luma_batch_dev)
        self._raise(fn(self.h, C.byref(ref), mb_xy, src, centers, n, out), "refine_batch_dev")

    def predict_inter_mb_batch_dev(self, y: int, u: int, v: int, y_stride: int, uv_stride: int, mb_cols: int,
                                   mb_rows: int, mb_xy: int, mvs: int, n: int, full_pixel: bool, filters,
                                   out: int) -> None:
        """predict_inter_mb over device arrays into out (n x 384); enqueues only."""
        rp = _RefPlanes(y, u, v, y_stride, uv_stride, mb_cols, mb_rows)
        f = np.ascontiguousarray(filters, np.int32)
        self._raise(self.lib.vp8g_predict_inter_mb_batch_dev(self.h, C.byref(rp), mb_xy, mvs, n, int(full_pixel),
                                                             f.ctypes.data, out),
                    "vp8g_predict_inter_mb_batch_dev")

    def half_plane_sads_batch(self, reference: LumaRef, xy, src, taps, diagonals: bool = True) -> np.ndarray:
        """fetch_window + HalfPlanes (src/encoder/inter_search.rs:266-346):
        (n, 8) SADs in (dy, dx) raster order without (0, 0)."""
        xy = np.ascontiguousarray(xy, np.int32).reshape(-1, 2)
        n = xy.shape[0]
        src = np.ascontiguousarray(src, np.uint8).reshape(n, 256)
        taps = np.ascontiguousarray(taps, np.int32).reshape(6)
        p, lr = self._luma_ref(reference)
        out = np.empty((max(n, 1), 8), np.uint32)
        st = self.lib.vp8g_half_plane_sads_batch(self.h, C.byref(lr), _ptr(xy), _ptr(src), n, _ptr(taps),
                                                 1 if diagonals else 0, _ptr(out))
        self._raise(st, "vp8g_half_plane_sads_batch")
        return out[:n]

    def half_pixel_refine_luma_batch(self, reference, mb_xy, src, centers) -> np.ndarray:
        """Rows of (mv.row, mv.col, sad)."""
        return self._refine(self.lib.vp8g_half_pixel_refine_luma_batch, reference, mb_xy, src, centers)

    def quarter_pixel_refine_luma_batch(self, reference, mb_xy, src, centers) -> np.ndarray:
        return self._refine(self.lib.vp8g_quarter_pixel_refine_luma_batch, reference, mb_xy, src,
                            centers)

    def half_pixel_refine_luma(self, reference, mb_col, mb_row, src_y, center: Mv) -> SearchResult:
        r = self.half_pixel_refine_luma_batch(reference, [[mb_col, mb_row]], src_y,
                                              [[center.row, center.col]])[0]
        return SearchResult(Mv(int(r[0]), int(r[1])), int(r[2]))

    def quarter_pixel_refine_luma(self, reference, mb_col, mb_row, src_y, center: Mv) -> SearchResult:
        r = self.quarter_pixel_refine_luma_batch(reference, [[mb_col, mb_row]], src_y,
                                                 [[center.row, center.col]])[0]
        return SearchResult(Mv(int(r[0]), int(r[1])), int(r[2]))


def split_chroma_mvs(split_luma_mvs) -> list:
    """src/motion_comp.rs:2162-2203: the four chroma sub-block vectors
    (eighth-pel) of sixteen quarter-pel luma sub-block vectors, each the
    avg() of its four doubled luma vectors (host arithmetic)."""
    mvs = [(m.row, m.col) if isinstance(m, Mv) else tuple(int(v) for v in m) for m in split_luma_mvs]
    out = []
    for c in range(4):
        s = [0, 0]
        for b in range(16):
            if ((b >> 1) & 1) + ((b >> 2) & 2) == c:
                s[0] += 2 * mvs[b][0]
                s[1] += 2 * mvs[b][1]
        r = [(v + 4) >> 3 if v >= 0 else -((-v + 4) >> 3) for v in s]
        out.append(Mv(int(np.int16(r[0])), int(np.int16(r[1]))))
    return out


# ------------------------------------------------ module-level reference API
_default = threading.local()


@dataclass
class ArnrJob:
    """One group of a batch: ``frames`` holds per-frame tuples
    (width, height, y, u, v, y_stride, uv_stride) of device pointers;
    outputs are device pointers (packed planes), ``mv_out`` optional."""
    frames: list
    center: int
    strength: int
    out_y: int
    out_u: int
    out_v: int
    mv_out: int | None = None


class ArnrBatch:
    """A created batch (one persistent launch captured in a HIP graph)."""

    def __init__(self, ctx: Context, h: C.c_void_p):
        self.ctx, self.h = ctx, h

    def run(self) -> None:
        """Enqueue one replay on the context's stream."""
        self.ctx._raise(self.ctx.lib.vp8g_arnr_batch_run(self.h), "vp8g_arnr_batch_run")

    def shape(self) -> dict:
        m, w, g, u = C.c_int(), C.c_int(), C.c_int(), C.c_size_t()
        self.ctx._raise(self.ctx.lib.vp8g_arnr_batch_shape(self.h, C.byref(m), C.byref(w),
                                                           C.byref(g), C.byref(u)),
                        "vp8g_arnr_batch_shape")
        return {"mbs": m.value, "waves": w.value, "workgroups": g.value, "units": u.value}

    def close(self) -> None:
        if getattr(self, "h", None):
            self.ctx.lib.vp8g_arnr_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tuning_lookup(path, width: int, height: int, n_frames: int, batch: bool = False,
                  lib: C.CDLL | None = None) -> dict:
    """The entry of a tuning-table file for a group or batch, as the library
    reads it (no device needed); zeros when nothing matches."""
    lib = lib or load_library()
    m, w, g = C.c_int(), C.c_int(), C.c_int()
    st = lib.vp8g_tuning_lookup(str(path).encode(), width, height, n_frames, int(batch), C.byref(m),
                                C.byref(w), C.byref(g))
    if st != VP8G_OK:
        raise Vp8gError(st, f"vp8g_tuning_lookup({path}) failed")
    return {"mbs": m.value, "waves": w.value, "wgs_per_cu": g.value}


TUNING_PATH = Path(__file__).resolve().parent.parent / "tuning" / "gfx950.json"


def default_context() -> Context:
    ctx = getattr(_default, "ctx", None)
    if ctx is None:
        ctx = Context(0)
        _default.ctx = ctx
    return ctx


def build_arnr_altref(frames, center: int, cfg: ArnrConfig | None = None) -> ArnrPicture:
    return default_context().build_arnr_altref(frames, center, cfg)


def predict_inter_mb(reference, mb_col, mb_row, luma_mv, full_pixel, filters) -> ReconstructedMb:
    return default_context().predict_inter_mb(reference, mb_col, mb_row, luma_mv, full_pixel, filters)


def mb_luma_sad_at_mv(reference, mb_col, mb_row, src_y, mv) -> int:
    return default_context().mb_luma_sad_at_mv(reference, mb_col, mb_row, src_y, mv)


def half_pixel_refine_luma(reference, mb_col, mb_row, src_y, center) -> SearchResult:
    return default_context().half_pixel_refine_luma(reference, mb_col, mb_row, src_y, center)


def quarter_pixel_refine_luma(reference, mb_col, mb_row, src_y, center) -> SearchResult:
    return default_context().quarter_pixel_refine_luma(reference, mb_col, mb_row, src_y, center)
