"""ctypes loader for the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker / the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"


class I420(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("y", C.c_void_p),
                ("u", C.c_void_p), ("v", C.c_void_p), ("y_stride", C.c_size_t),
                ("uv_stride", C.c_size_t)]


class Err(C.Structure):
    _fields_ = [("src_w", C.c_uint32), ("src_h", C.c_uint32), ("ref_w", C.c_uint32),
                ("ref_h", C.c_uint32)]


class MvS(C.Structure):
    _fields_ = [("row", C.c_int16), ("col", C.c_int16)]


class RefPlanes(C.Structure):
    _fields_ = [("y", C.c_void_p), ("u", C.c_void_p), ("v", C.c_void_p),
                ("y_stride", C.c_size_t), ("uv_stride", C.c_size_t),
                ("mb_cols", C.c_size_t), ("mb_rows", C.c_size_t)]


class LumaRef(C.Structure):
    _fields_ = [("plane", C.c_void_p), ("stride", C.c_size_t), ("width", C.c_size_t),
                ("height", C.c_size_t)]


class SearchResult(C.Structure):
    _fields_ = [("mv", MvS), ("sad", C.c_uint32)]


# orc_mb_search: vp8g_mb_search's fields plus the source pointer.
ORC_MB_SEARCH = np.dtype([("mb_col", "<u4"), ("mb_row", "<u4"), ("best", "<i2", (2,)), ("range_x", "<i4", (2,)),
                          ("range_y", "<i4", (2,)), ("n_seeds", "<u4"), ("seeds", "<i2", (8, 2)),
                          ("lambda_sad", "<f8"), ("good_enough", "<u4"), ("descend", "<i4"), ("src", "<u8")],
                         align=True)
ORC_FOUND = np.dtype([("mv", "<i2", (2,)), ("sad", "<u4"), ("cost", "<f8")], align=True)


class LfParams(C.Structure):
    _fields_ = [("interior_limit", C.c_uint8), ("hev_threshold", C.c_uint8), ("mbedge_limit", C.c_uint8),
                ("sub_bedge_limit", C.c_uint8)]


class LfConfig(C.Structure):
    """orc_lf_config: FrameFilterConfig with ref_delta [CURRENT, LAST, GOLDEN,
    ALTREF] and mode_delta [B_PRED, ZEROMV, NEAREST/NEAR/NEW, SPLITMV]."""
    _fields_ = [("simple", C.c_int32), ("key_frame", C.c_int32), ("loop_filter_level", C.c_int32),
                ("sharpness_level", C.c_int32), ("segmentation_enabled", C.c_int32), ("segment_abs", C.c_int32),
                ("delta_enabled", C.c_int32), ("segment_lf_level", C.c_int16 * 4), ("ref_delta", C.c_int16 * 4),
                ("mode_delta", C.c_int16 * 4)]


class LfMb(C.Structure):
    _fields_ = [("segment_id", C.c_uint8), ("ref_frame", C.c_uint8), ("inter_mode", C.c_uint8),
                ("flags", C.c_uint8)]


class LfPlanes(C.Structure):
    _fields_ = [("y", C.c_void_p), ("u", C.c_void_p), ("v", C.c_void_p), ("y_stride", C.c_size_t),
                ("uv_stride", C.c_size_t), ("mb_cols", C.c_size_t), ("mb_rows", C.c_size_t)]


class LfSource(C.Structure):
    _fields_ = [("width", C.c_size_t), ("height", C.c_size_t), ("y", C.c_void_p), ("u", C.c_void_p),
                ("v", C.c_void_p), ("y_stride", C.c_size_t), ("uv_stride", C.c_size_t)]


def lf_config(**kw) -> LfConfig:
    """An LfConfig from keyword fields; list values fill the array fields."""
    c = LfConfig()
    for k, v in kw.items():
        if k in ("segment_lf_level", "ref_delta", "mode_delta"):
            getattr(c, k)[:] = list(v)
        else:
            setattr(c, k, int(v))
    return c


class HalfPlanes(C.Structure):
    _fields_ = [("h", C.c_uint8 * (22 * 17)), ("v", C.c_uint8 * (17 * 16)),
                ("d", C.c_uint8 * (17 * 17))]


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


class Oracle:
    def __init__(self, path: Path | None = None):
        p = Path(path) if path else LIB
        if not p.exists():
            build()
        self.lib = L = C.CDLL(str(p))
        vp, sz, i32, u32, u8 = C.c_void_p, C.c_size_t, C.c_int, C.c_uint32, C.c_uint8
        L.orc_build_arnr_altref.argtypes = [C.POINTER(I420), sz, sz, u8, sz, vp, vp, vp,
                                            C.POINTER(Err), vp]
        L.orc_build_arnr_altref.restype = i32
        L.orc_block_sad.argtypes = [C.POINTER(I420), C.POINTER(I420)] + [i32] * 6 + [u32]
        L.orc_block_sad.restype = u32
        L.orc_block_sad_generic.argtypes = L.orc_block_sad.argtypes
        L.orc_block_sad_generic.restype = u32
        L.orc_refine_search.argtypes = [C.POINTER(I420), C.POINTER(I420)] + [i32] * 4 + [vp]
        L.orc_refine_search.restype = None
        L.orc_block_search.argtypes = [C.POINTER(I420), C.POINTER(I420)] + [i32] * 7 + [vp]
        L.orc_block_search.restype = None
        L.orc_block_search_group.argtypes = [C.POINTER(I420), sz, sz, i32, vp, sz, vp]
        L.orc_block_search_group.restype = i32
        L.orc_search_fallback.argtypes = [C.POINTER(I420), sz, sz, vp, i32, i32, vp, vp]
        L.orc_search_fallback.restype = i32
        L.orc_build_arnr_altref_mv.argtypes = [C.POINTER(I420), sz, sz, u8, vp, vp, vp, vp,
                                               C.POINTER(Err)]
        L.orc_build_arnr_altref_mv.restype = i32
        L.orc_pixel_weight.argtypes = [C.c_int32, u32]
        L.orc_pixel_weight.restype = u32
        L.orc_weight_table.argtypes = [u32, vp]
        L.orc_accumulate_luma_block.argtypes = [vp, vp, sz] + [i32] * 8 + [vp, vp, vp]
        L.orc_accumulate_chroma_block.argtypes = [vp, vp, sz] + [i32] * 8 + [vp, vp, vp]
        L.orc_interp.argtypes = [vp, vp]
        L.orc_interp.restype = u8
        L.orc_avg.argtypes = [C.c_int32] * 4
        L.orc_avg.restype = C.c_int32
        L.orc_stored_luma_mv.argtypes = [MvS]
        L.orc_stored_luma_mv.restype = MvS
        L.orc_chroma_mv.argtypes = [MvS]
        L.orc_chroma_mv.restype = MvS
        L.orc_apply_full_pixel.argtypes = [MvS]
        L.orc_apply_full_pixel.restype = MvS
        L.orc_fetch_block_halo.argtypes = [vp, sz, sz, sz, sz, sz, MvS, vp]
        L.orc_sixtap_2d.argtypes = [vp, i32, i32, vp, vp]
        L.orc_fetch_luma_mb_halo.argtypes = [vp, sz, sz, sz, sz, sz, MvS, vp]
        L.orc_sixtap_mb_luma.argtypes = [vp, i32, i32, vp, vp]
        L.orc_fetch_chroma_mb_halo.argtypes = [vp, sz, sz, sz, sz, sz, MvS, vp]
        L.orc_sixtap_mb_chroma.argtypes = [vp, i32, i32, vp, vp]
        L.orc_fetch_luma_mb_whole_pixel.argtypes = [vp, sz, sz, sz, sz, sz, MvS, vp]
        L.orc_fetch_chroma_mb_whole_pixel.argtypes = [vp, sz, sz, sz, sz, sz, MvS, vp]
        L.orc_predict_inter_mb.argtypes = [C.POINTER(RefPlanes), sz, sz, MvS, i32, vp, vp]
        L.orc_predict_inter_mb_batch.argtypes = [C.POINTER(RefPlanes), vp, vp, sz, i32, vp, vp]
        L.orc_block_sad_16x16.argtypes = [vp, vp]
        L.orc_block_sad_16x16.restype = u32
        L.orc_mb_luma_sad_at_whole_mv.argtypes = [C.POINTER(LumaRef), sz, sz, vp, MvS]
        L.orc_mb_luma_sad_at_whole_mv.restype = u32
        L.orc_mb_luma_sad_at_mv.argtypes = [C.POINTER(LumaRef), sz, sz, vp, MvS]
        L.orc_mb_luma_sad_at_mv.restype = u32
        L.orc_mb_luma_sad_at_mv_batch.argtypes = [C.POINTER(LumaRef), vp, vp, vp, sz, vp]
        L.orc_half_pixel_refine_luma.argtypes = [C.POINTER(LumaRef), sz, sz, vp, MvS]
        L.orc_half_pixel_refine_luma.restype = SearchResult
        L.orc_quarter_pixel_refine_luma.argtypes = [C.POINTER(LumaRef), sz, sz, vp, MvS]
        L.orc_quarter_pixel_refine_luma.restype = SearchResult
        L.orc_fetch_window.argtypes = [C.POINTER(LumaRef), i32, i32, vp]
        L.orc_half_planes_new.argtypes = [vp, vp, i32, C.POINTER(HalfPlanes)]
        L.orc_half_planes_sad.argtypes = [C.POINTER(HalfPlanes), vp, i32, i32]
        L.orc_half_planes_sad.restype = u32
        L.orc_luma_mad.argtypes = [C.POINTER(I420), C.POINTER(I420)]
        L.orc_luma_mad.restype = C.c_double
        L.orc_filter_block_4x4.argtypes = [vp, sz, sz, sz, sz, sz, MvS, vp, vp]
        L.orc_predict_split_mv.argtypes = [C.POINTER(RefPlanes), sz, sz, vp, i32, vp, vp]
        L.orc_reconstruct_inter_mb.argtypes = [C.POINTER(RefPlanes), sz, sz, MvS, i32, vp, i32, vp, vp]
        L.orc_reconstruct_split_mv_mb.argtypes = [C.POINTER(RefPlanes), sz, sz, vp, i32, vp, i32, vp, vp]
        L.orc_predict_split_mv_batch.argtypes = [C.POINTER(RefPlanes), vp, vp, sz, i32, vp, vp]
        L.orc_reconstruct_inter_mb_batch.argtypes = [C.POINTER(RefPlanes), vp, vp, sz, i32, vp, vp, vp, vp]
        L.orc_reconstruct_split_mv_mb_batch.argtypes = [C.POINTER(RefPlanes), vp, vp, sz, i32, vp, vp, vp, vp]
        L.orc_small_diamond_search_luma.argtypes = [C.POINTER(LumaRef), sz, sz, vp, MvS, u32]
        L.orc_small_diamond_search_luma.restype = SearchResult
        L.orc_whole_pixel_search_batch.argtypes = [C.POINTER(LumaRef), vp, vp, sz, vp, vp]
        L.orc_sub_pixel_refine_batch.argtypes = [C.POINTER(LumaRef), vp, vp, vp, sz, vp, vp, vp]
        L.orc_lf_derive.argtypes = [u8, u8, i32]
        L.orc_lf_derive.restype = LfParams
        L.orc_lf_common_adjust.argtypes = [i32, vp]
        L.orc_lf_common_adjust.restype = i32
        L.orc_lf_simple_segment.argtypes = [u8, vp]
        L.orc_lf_subblock_filter.argtypes = [u8, u8, u8, vp]
        L.orc_lf_mb_filter.argtypes = [u8, u8, u8, vp]
        L.orc_lf_mb_level.argtypes = [C.POINTER(LfConfig), LfMb]
        L.orc_lf_mb_level.restype = u8
        L.orc_lf_filter_frame.argtypes = [C.POINTER(LfPlanes), vp, C.POINTER(LfConfig)]
        L.orc_lf_ssd.argtypes = [C.POINTER(LfPlanes), C.POINTER(LfSource)]
        L.orc_lf_ssd.restype = C.c_uint64
        L.orc_lf_score.argtypes = [C.POINTER(LfPlanes), vp, C.POINTER(LfConfig), C.POINTER(LfSource), u8, u8]
        L.orc_lf_score.restype = C.c_uint64
        L.orc_lf_score_candidates.argtypes = [C.POINTER(LfPlanes), vp, C.POINTER(LfConfig), C.POINTER(LfSource),
                                              vp, vp, sz, sz, vp]
        L.orc_lf_select_params.argtypes = [C.POINTER(LfPlanes), vp, C.POINTER(LfConfig), C.POINTER(LfSource), sz,
                                           vp, vp]
        L.orc_lf_select_params_sampled.argtypes = [C.POINTER(LfPlanes), vp, C.POINTER(LfConfig),
                                                   C.POINTER(LfSource), sz, i32, sz, vp, vp]
        self.SIXTAP = np.ctypeslib.as_array((C.c_int32 * 48).in_dll(L, "ORC_SIXTAP_FILTERS")).reshape(8, 6).copy()
        self.BILINEAR = np.ctypeslib.as_array((C.c_int32 * 48).in_dll(L, "ORC_BILINEAR_FILTERS")).reshape(8, 6).copy()

    # ---- ARNR
    @staticmethod
    def i420(width, height, y, u, v, y_stride, uv_stride, keep: list) -> I420:
        y = np.ascontiguousarray(y, np.uint8).reshape(-1)
        u = np.ascontiguousarray(u, np.uint8).reshape(-1)
        v = np.ascontiguousarray(v, np.uint8).reshape(-1)
        keep += [y, u, v]
        return I420(width, height, _p(y) if y.size else None, _p(u) if u.size else None,
                    _p(v) if v.size else None, y_stride, uv_stride)

    def _arr(self, frames, keep):
        arr = (I420 * max(len(frames), 1))()
        for i, f in enumerate(frames):
            arr[i] = self.i420(*f, keep=keep)
        return arr

    def block_search_group(self, frames, center, radius, seeds=None, threads=8):
        """Exhaustive search table (n, brows, bcols, 3) for every non-centre
        frame against the centre; centre rows are zero."""
        keep: list = []
        arr = self._arr(frames, keep)
        w, h = frames[center][0], frames[center][1]
        out = np.zeros((len(frames), (h + 15) // 16, (w + 15) // 16, 3), np.int32)
        sd = None
        if seeds is not None:
            sd = np.ascontiguousarray(seeds, np.int32)
            assert sd.shape == out.shape
        st = self.lib.orc_block_search_group(arr, len(frames), center, radius,
                                             _p(sd) if sd is not None else None, threads, _p(out))
        assert st == 0, st
        return out

    def search_fallback(self, frames, center, first, radius=16, gate=True):
        """Exhaustive fallback over the table `first` (n, brows, bcols, 3):
        returns (out table, queued count)."""
        keep: list = []
        arr = self._arr(frames, keep)
        f = np.ascontiguousarray(first, np.int32)
        out = f.copy()
        q = C.c_uint32(0)
        st = self.lib.orc_search_fallback(arr, len(frames), center, _p(f), radius, int(gate), _p(out), C.byref(q))
        assert st == 0, st
        return out, int(q.value)

    # ---- encoder inter search
    def _orc_jobs(self, jobs, src):
        src = np.ascontiguousarray(src, np.uint8).reshape(-1, 256)
        out = np.zeros(len(jobs), ORC_MB_SEARCH)
        for name in jobs.dtype.names:
            out[name] = jobs[name]
        out["src"] = src.ctypes.data + 256 * np.arange(len(jobs), dtype=np.uint64)
        return out, src

    def whole_pixel_search_batch(self, plane, stride, width, height, bits, jobs, src, shape):
        """whole_pixel_search per job (vp8g_mb_search structured array);
        shape = (descent_rounds, wide_pattern, half_diagonals, quarter)."""
        oj, keep = self._orc_jobs(jobs, src)
        p = np.ascontiguousarray(plane, np.uint8)
        lr = LumaRef(p.ctypes.data, stride, width, height)
        b = np.ascontiguousarray(bits, np.float32)
        sh = (C.c_uint32 * 4)(*[int(v) for v in shape])
        out = np.zeros(len(jobs), ORC_FOUND)
        self.lib.orc_whole_pixel_search_batch(C.byref(lr), b.ctypes.data, oj.ctypes.data, len(jobs), sh,
                                              out.ctypes.data)
        return out

    def sub_pixel_refine_batch(self, plane, stride, width, height, bits, jobs, src, start, filters, shape):
        oj, keep = self._orc_jobs(jobs, src)
        p = np.ascontiguousarray(plane, np.uint8)
        lr = LumaRef(p.ctypes.data, stride, width, height)
        b = np.ascontiguousarray(bits, np.float32)
        st = np.ascontiguousarray(start, ORC_FOUND)
        f = np.ascontiguousarray(filters, np.int32)
        sh = (C.c_uint32 * 4)(*[int(v) for v in shape])
        out = np.zeros(len(jobs), ORC_FOUND)
        self.lib.orc_sub_pixel_refine_batch(C.byref(lr), b.ctypes.data, oj.ctypes.data, st.ctypes.data, len(jobs),
                                            f.ctypes.data, sh, out.ctypes.data)
        return out

    def build_arnr_altref_mv(self, frames, center, strength, table):
        """Altref from a given alignment table (n, brows, bcols, 3)."""
        keep: list = []
        arr = self._arr(frames, keep)
        w, h = frames[center][0], frames[center][1]
        cw, ch = (w + 1) // 2, (h + 1) // 2
        t = np.ascontiguousarray(table, np.int32)
        oy = np.zeros(max(w * h, 1), np.uint8)
        ou = np.zeros(max(cw * ch, 1), np.uint8)
        ov = np.zeros(max(cw * ch, 1), np.uint8)
        err = Err()
        st = self.lib.orc_build_arnr_altref_mv(arr, len(frames), center, strength, _p(t), _p(oy),
                                               _p(ou), _p(ov), C.byref(err))
        if st != 0:
            return st, None
        return st, (oy[:w * h], ou[:cw * ch], ov[:cw * ch])

    def build_arnr_altref(self, frames, center, strength=3, threads=1, want_mvs=False):
        """frames: list of (w, h, y, u, v, y_stride, uv_stride).  Returns
        (status, (y, u, v) or None, err, mvs or None)."""
        keep: list = []
        n = len(frames)
        arr = (I420 * max(n, 1))()
        for i, f in enumerate(frames):
            arr[i] = self.i420(*f, keep=keep)
        err = Err()
        if center >= n:
            st = self.lib.orc_build_arnr_altref(arr, n, center, strength, threads, None, None,
                                                None, C.byref(err), None)
            return st, None, err, None
        w, h = frames[center][0], frames[center][1]
        cw, ch = (w + 1) // 2, (h + 1) // 2
        oy = np.zeros(max(w * h, 1), np.uint8)
        ou = np.zeros(max(cw * ch, 1), np.uint8)
        ov = np.zeros(max(cw * ch, 1), np.uint8)
        mvs = None
        if want_mvs:
            mvs = np.zeros((n, (h + 15) // 16, (w + 15) // 16, 3), np.int32)
        st = self.lib.orc_build_arnr_altref(arr, n, center, strength, threads, _p(oy), _p(ou),
                                            _p(ov), C.byref(err), _p(mvs) if want_mvs else None)
        if st != 0:
            return st, None, err, None
        return st, (oy[:w * h], ou[:cw * ch], ov[:cw * ch]), err, mvs

    # ---- loop filter.  Planes are (y, u, v) 2-D uint8 arrays of the
    # MB-padded reconstruction (mb_rows*16 x y_stride, mb_rows*8 x uv_stride);
    # mbs is an (mb_rows*mb_cols, 4) uint8 array of orc_lf_mb rows.
    @staticmethod
    def _lf_planes(y, u, v, keep) -> LfPlanes:
        for a in (y, u, v):
            assert a.dtype == np.uint8 and a.flags.c_contiguous
        keep += [y, u, v]
        return LfPlanes(_p(y), _p(u), _p(v), y.shape[1], u.shape[1], y.shape[1] // 16, y.shape[0] // 16)

    @staticmethod
    def _lf_source(src, keep) -> LfSource:
        w, h, y, u, v = src
        keep += [y, u, v]
        return LfSource(w, h, _p(y), _p(u), _p(v), y.shape[1], u.shape[1])

    @staticmethod
    def _lf_mbs(mbs, keep):
        m = np.ascontiguousarray(mbs, np.uint8).reshape(-1, 4)
        keep.append(m)
        return _p(m)

    def lf_filter_frame(self, y, u, v, mbs, cfg: LfConfig, mb_cols=None):
        """Filtered copies of (y, u, v); mb_cols defaults to y_stride // 16."""
        keep = []
        y, u, v = (np.array(a, np.uint8, order="C", copy=True) for a in (y, u, v))
        pl = self._lf_planes(y, u, v, keep)
        if mb_cols is not None:
            pl.mb_cols = mb_cols
        self.lib.orc_lf_filter_frame(C.byref(pl), self._lf_mbs(mbs, keep), C.byref(cfg))
        return y, u, v

    def lf_ssd(self, y, u, v, src) -> int:
        keep = []
        return int(self.lib.orc_lf_ssd(C.byref(self._lf_planes(y, u, v, keep)), C.byref(self._lf_source(src, keep))))

    def lf_score_candidates(self, y, u, v, mbs, cfg, src, levels, sharps, threads=8):
        keep = []
        lv = np.ascontiguousarray(levels, np.uint8)
        sh = np.ascontiguousarray(sharps, np.uint8)
        out = np.zeros(lv.size, np.uint64)
        self.lib.orc_lf_score_candidates(C.byref(self._lf_planes(y, u, v, keep)), self._lf_mbs(mbs, keep),
                                         C.byref(cfg), C.byref(self._lf_source(src, keep)), _p(lv), _p(sh),
                                         lv.size, threads, _p(out))
        return out

    def lf_select(self, y, u, v, mbs, cfg, src, threads=8):
        keep = []
        lv = np.zeros(1, np.uint8)
        sh = np.zeros(1, np.uint8)
        self.lib.orc_lf_select_params(C.byref(self._lf_planes(y, u, v, keep)), self._lf_mbs(mbs, keep),
                                      C.byref(cfg), C.byref(self._lf_source(src, keep)), threads, _p(lv), _p(sh))
        return int(lv[0]), int(sh[0])

    def lf_select_sampled(self, y, u, v, mbs, cfg, src, every_rows, sweep_sharpness, threads=8):
        keep = []
        lv = np.zeros(1, np.uint8)
        sh = np.zeros(1, np.uint8)
        self.lib.orc_lf_select_params_sampled(C.byref(self._lf_planes(y, u, v, keep)), self._lf_mbs(mbs, keep),
                                              C.byref(cfg), C.byref(self._lf_source(src, keep)), every_rows,
                                              int(bool(sweep_sharpness)), threads, _p(lv), _p(sh))
        return int(lv[0]), int(sh[0])
