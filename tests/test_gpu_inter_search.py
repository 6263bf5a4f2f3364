"""GPU parity for the encoder inter search (whole_pixel_search,
sub_pixel_refine, small_diamond_search_luma; src/encoder/inter_search.rs:
380-552, src/motion_search.rs:381-442) against the oracle through the C
ABI: the vector, the SAD and the f64 cost of every macroblock bit-exact.
The oracle keeps the reference's early SAD exit and the GPU computes full
SADs, so equal results also confirm that the exit never changes a
decision."""
import ctypes as C

import numpy as np
import pytest

from test_oracle_inter_search import blurred, feature_plane, mb

pytestmark = pytest.mark.gpu


def frame_pair(w, h, seed, shift=(2, -3)):
    cur = blurred(seed, h, w, 2)
    ref = (np.roll(cur, shift, (0, 1)).astype(np.int32) + np.random.default_rng(seed).integers(-6, 7, (h, w)))
    return cur, np.clip(ref, 0, 255).astype(np.uint8)


def random_jobs(rng, bw, bh, n=None):
    import vp8g.inter_search as IS
    cols, rows = np.meshgrid(np.arange(bw), np.arange(bh))
    cols, rows = cols.reshape(-1), rows.reshape(-1)
    if n is not None:
        pick = rng.choice(cols.size, n, replace=False)
        cols, rows = cols[pick], rows[pick]
    n = cols.size
    j = np.zeros(n, IS.MB_SEARCH)
    j["mb_col"], j["mb_row"] = cols, rows
    j["best"] = rng.integers(-20, 21, (n, 2)) * rng.choice([1, 4], (n, 1))
    lo = rng.integers(-40, 1, (n, 2))
    hi = rng.integers(0, 41, (n, 2))
    j["range_x"] = np.stack([lo[:, 0], hi[:, 0]], 1)
    j["range_y"] = np.stack([lo[:, 1], hi[:, 1]], 1)
    j["n_seeds"] = rng.integers(0, 9, n)
    j["seeds"] = rng.integers(-120, 121, (n, 8, 2))
    j["lambda_sad"] = rng.choice([0.0, 0.7, 3.5, 17.25, 90.0], n)
    j["good_enough"] = rng.choice([0, 64, 640, 5000], n)
    j["descend"] = rng.random(n) < 0.85
    return j


def oracle_wps(oracle, ref, jobs, src, costs, shape):
    return oracle.whole_pixel_search_batch(ref, ref.shape[1], ref.shape[1], ref.shape[0], costs.bits, jobs, src,
                                           (shape.descent_rounds, shape.wide_pattern, shape.half_diagonals,
                                            shape.quarter))


def oracle_spr(oracle, ref, jobs, src, start, costs, filters, shape):
    return oracle.sub_pixel_refine_batch(ref, ref.shape[1], ref.shape[1], ref.shape[0], costs.bits, jobs, src,
                                         start, filters, (shape.descent_rounds, shape.wide_pattern,
                                                          shape.half_diagonals, shape.quarter))


def assert_found(got, want, what):
    for name in ("mv", "sad", "cost"):
        g, w = got[name], want[name]
        if not np.array_equal(g, w):
            bad = np.nonzero((g != w).reshape(len(g), -1).any(1))[0]
            i = int(bad[0])
            raise AssertionError(f"{what}: {len(bad)} of {len(g)} differ in {name}; first {i}: "
                                 f"got {got[i]} want {want[i]}")


@pytest.mark.parametrize("effort", ["fastest", "fast", "balanced", "thorough"])
def test_whole_pixel_search_and_sub_pixel_refine_match_oracle(gpu, oracle, effort):
    import vp8g
    import vp8g.inter_search as IS
    shape = IS.SHAPES[effort]
    rng = np.random.default_rng(len(effort))
    w, h = 352, 288
    cur, ref = frame_pair(w, h, 40 + len(effort))
    costs = IS.MvCostTable.new()
    jobs = random_jobs(rng, w // 16, h // 16)
    src = np.stack([mb(cur, int(c), int(r)) for c, r in zip(jobs["mb_col"], jobs["mb_row"])])
    lref = vp8g.LumaRef(ref.reshape(-1), w, w, h)
    got = IS.whole_pixel_search_batch(gpu, lref, costs, jobs, src, shape)
    want = oracle_wps(oracle, ref, jobs, src, costs, shape)
    assert_found(got, want, f"whole_pixel_search {effort}")
    # refine every finite start, six-tap and bilinear half taps
    ok = np.isfinite(want["cost"])
    for name, filt in (("sixtap", vp8g.SIXTAP_FILTERS), ("bilinear", vp8g.BILINEAR_FILTERS)):
        g2 = IS.sub_pixel_refine_batch(gpu, lref, costs, jobs[ok], src[ok], want[ok], filt, shape)
        w2 = oracle_spr(oracle, ref, jobs[ok], src[ok], want[ok], costs, filt, shape)
        assert_found(g2, w2, f"sub_pixel_refine {effort} {name}")


def test_sub_pixel_refine_any_start_and_shapes(gpu, oracle):
    """Starts off the whole-pel grid and near the +-1020 / +-1022 limits,
    diagonals and quarter steps on and off."""
    import vp8g
    import vp8g.inter_search as IS
    rng = np.random.default_rng(5)
    w, h = 160, 96
    cur, ref = frame_pair(w, h, 77, (1, 1))
    costs = IS.MvCostTable.new()
    jobs = random_jobs(rng, w // 16, h // 16)
    n = len(jobs)
    src = np.stack([mb(cur, int(c), int(r)) for c, r in zip(jobs["mb_col"], jobs["mb_row"])])
    start = np.zeros(n, IS.FOUND)
    start["mv"] = rng.integers(-30, 31, (n, 2))
    start["mv"][:6] = [(1020, -1019), (-1022, 1021), (1018, 1022), (-1023, -1023), (2, 1023), (32767, -32768)]
    start["sad"] = rng.integers(0, 20000, n)
    start["cost"] = rng.uniform(0, 30000, n)
    start["cost"][:3] = np.inf
    lref = vp8g.LumaRef(ref.reshape(-1), w, w, h)
    for shape in (IS.SearchShape(8, True, True, True), IS.SearchShape(8, True, False, True),
                  IS.SearchShape(8, True, True, False), IS.SearchShape(8, False, False, False)):
        g = IS.sub_pixel_refine_batch(gpu, lref, costs, jobs, src, start, vp8g.SIXTAP_FILTERS, shape)
        o = oracle_spr(oracle, ref, jobs, src, start, costs, vp8g.SIXTAP_FILTERS, shape)
        assert_found(g, o, f"sub_pixel_refine {shape}")


def test_1080p_frame_search_matches_oracle(gpu, oracle):
    """Every macroblock of a 1080p frame pair (MB-aligned 1920x1088) at the
    Balanced shape, then the sub-pel refinement of every result."""
    import vp8g
    import vp8g.inter_search as IS
    import vp8g.synth as synth
    g = synth.arnr_group_regions(1920, 1080, 3, vmax=6, occluders=10, clip=1, config=5)
    cur = synth.pad_to_mb(g.y[1], 16)
    ref = synth.pad_to_mb(g.y[2], 16)
    h, w = ref.shape
    rng = np.random.default_rng(1080)
    costs = IS.MvCostTable.new()
    jobs = random_jobs(rng, w // 16, h // 16)
    jobs["lambda_sad"] = 6.0
    jobs["good_enough"] = 256 * 8 // 2
    jobs["descend"] = 1
    src = np.stack([mb(cur, int(c), int(r)) for c, r in zip(jobs["mb_col"], jobs["mb_row"])])
    lref = vp8g.LumaRef(np.ascontiguousarray(ref).reshape(-1), w, w, h)
    shape = IS.SHAPES["balanced"]
    got = IS.whole_pixel_search_batch(gpu, lref, costs, jobs, src, shape)
    want = oracle_wps(oracle, ref, jobs, src, costs, shape)
    assert_found(got, want, "1080p whole_pixel_search")
    ok = np.isfinite(want["cost"])
    g2 = IS.sub_pixel_refine_batch(gpu, lref, costs, jobs[ok], src[ok], want[ok], vp8g.SIXTAP_FILTERS, shape)
    w2 = oracle_spr(oracle, ref, jobs[ok], src[ok], want[ok], costs, vp8g.SIXTAP_FILTERS, shape)
    assert_found(g2, w2, "1080p sub_pixel_refine")


def test_small_diamond_matches_oracle(gpu, oracle):
    import vp8g
    import vp8g.inter_search as IS
    from oracle_ctypes import LumaRef, MvS
    rng = np.random.default_rng(9)
    w, h = 256, 160
    cur, ref = frame_pair(w, h, 91, (-2, 3))
    bw, bh = w // 16, h // 16
    xy = np.stack(np.meshgrid(np.arange(bw), np.arange(bh)), -1).reshape(-1, 2).astype(np.uint32)
    n = len(xy)
    src = np.stack([mb(cur, int(c), int(r)) for c, r in xy])
    centers = rng.integers(-60, 61, (n, 2)).astype(np.int16)
    centers[:4] = [(32767, -32768), (1023, -1023), (5, -7), (-1, 1)]
    p = np.ascontiguousarray(ref)
    lr = LumaRef(p.ctypes.data, w, w, h)
    for iters in (0, 1, 4, 32):
        got = IS.small_diamond_search_luma_batch(gpu, vp8g.LumaRef(p.reshape(-1), w, w, h), xy, src, centers, iters)
        for i in range(n):
            s = np.ascontiguousarray(src[i])
            r = oracle.lib.orc_small_diamond_search_luma(C.byref(lr), int(xy[i, 0]), int(xy[i, 1]), s.ctypes.data,
                                                         MvS(int(centers[i, 0]), int(centers[i, 1])), iters)
            assert tuple(got[i]) == (r.mv.row, r.mv.col, r.sad), (iters, i)
    # the reference's planted translation (src/motion_search.rs:1028-1055)
    ref2 = feature_plane(64, 64, 100, 220, 8, 8, 5)
    src2 = mb(feature_plane(64, 64, 100, 220, 5, 6, 5), 0, 0)
    got = IS.small_diamond_search_luma_batch(gpu, vp8g.LumaRef(ref2.reshape(-1), 64, 64, 64), [[0, 0]], src2,
                                             [[0, 0]], 64)
    assert tuple(got[0]) == (8, 12, 0)


def test_dev_entries_match_host_entries(gpu):
    import torch
    import vp8g
    import vp8g.inter_search as IS
    rng = np.random.default_rng(17)
    w, h = 192, 128
    cur, ref = frame_pair(w, h, 5)
    costs = IS.MvCostTable.new()
    jobs = random_jobs(rng, w // 16, h // 16)
    n = len(jobs)
    src = np.stack([mb(cur, int(c), int(r)) for c, r in zip(jobs["mb_col"], jobs["mb_row"])])
    lref = vp8g.LumaRef(ref.reshape(-1), w, w, h)
    shape = IS.SHAPES["thorough"]
    host_w = IS.whole_pixel_search_batch(gpu, lref, costs, jobs, src, shape)
    host_s = IS.sub_pixel_refine_batch(gpu, lref, costs, jobs, src, host_w, vp8g.SIXTAP_FILTERS, shape)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).reshape(-1).copy()).cuda()
    dplane, dcosts, djobs, dsrc, dw = dev(ref), dev(costs.bits), dev(jobs), dev(src), dev(np.zeros(n, IS.FOUND))
    ds = dev(np.zeros(n, IS.FOUND))
    lr = vp8g._LumaRef(dplane.data_ptr(), w, w, h)
    torch.cuda.synchronize()
    assert gpu.lib.vp8g_whole_pixel_search_batch_dev(gpu.h, C.byref(lr), dcosts.data_ptr(), djobs.data_ptr(),
                                                     dsrc.data_ptr(), n, shape.c(), dw.data_ptr()) == 0
    f = np.ascontiguousarray(vp8g.SIXTAP_FILTERS, np.int32)
    assert gpu.lib.vp8g_sub_pixel_refine_batch_dev(gpu.h, C.byref(lr), dcosts.data_ptr(), djobs.data_ptr(),
                                                   dsrc.data_ptr(), dw.data_ptr(), n, f.ctypes.data, shape.c(),
                                                   ds.data_ptr()) == 0
    gpu.synchronize()
    got_w = dw.cpu().numpy().view(IS.FOUND)
    got_s = ds.cpu().numpy().view(IS.FOUND)
    assert_found(got_w, host_w, "whole_pixel_search _dev")
    assert_found(got_s, host_s, "sub_pixel_refine _dev")


def test_host_entries_reject_bad_jobs(gpu):
    import vp8g
    import vp8g.inter_search as IS
    w = h = 64
    ref = np.zeros((h, w), np.uint8)
    lref = vp8g.LumaRef(ref.reshape(-1), w, w, h)
    costs = IS.MvCostTable.new()
    src = np.zeros((1, 256), np.uint8)
    base = random_jobs(np.random.default_rng(0), 2, 2, 1)
    for field, value in (("n_seeds", 9), ("range_x", (5, -5)), ("range_y", (-9000, 0)), ("lambda_sad", np.nan),
                         ("lambda_sad", -1.0), ("good_enough", 1 << 30), ("mb_col", 1 << 20)):
        j = base.copy()
        j[field] = value
        with pytest.raises(vp8g.Vp8gError):
            IS.whole_pixel_search_batch(gpu, lref, costs, j, src, IS.SHAPES["fast"])
    bad = np.array(vp8g.SIXTAP_FILTERS)
    bad[4, 0] = (1 << 20) + 1
    with pytest.raises(vp8g.Vp8gError):
        IS.sub_pixel_refine_batch(gpu, lref, costs, base, src, np.zeros(1, IS.FOUND), bad, IS.SHAPES["fast"])
