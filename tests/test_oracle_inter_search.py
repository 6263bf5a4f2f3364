"""Pins the encoder inter-search restatements: the host MvCostTable builder
(vp8g.inter_search) and the oracle's small_diamond_search_luma,
whole_pixel_search and sub_pixel_refine.

* mv_component_bits is checked bit for bit against the reference test's
  own literal restatement (a small-tree DFS with an inline -log2,
  src/motion_vector.rs:825-910) over every component and three contexts,
  plus the zero-cost and growth tests (:803-821, :911-928).
* small_diamond_search_luma re-runs the reference's tests
  (src/motion_search.rs:995-1206) on the same planes.
* whole_pixel_search and sub_pixel_refine have no tests in the reference;
  they are pinned by planted exact whole-, half- and quarter-pel
  translations, the cost identity (the returned cost is the returned SAD,
  recomputed by a numpy gather, plus lambda times the f32 table sum), local
  optimality over the one-pixel pattern, hand-rounded seed positions, the
  two-pixel pattern reaching what the one-pixel pattern cannot in the same
  rounds, and sub-pel SADs equal to mb_luma_sad_at_mv at the same vector."""
import ctypes as C
import math

import numpy as np
import pytest


# ------------------------------------------------------------ MvCostTable
def tree_walk_bits(p, value):
    """The reference test's mv_component_bits_reference
    (src/motion_vector.rs:825-888): per-bool -log2 and a DFS path."""
    def bb(prob, bit):
        p0 = max(int(prob), 1) / 256.0
        q = 1.0 - p0 if bit else p0
        return -math.log2(max(q, 1.0 / 256.0))

    tree = (2, 8, 4, 6, -0, -1, -2, -3, 10, 12, -4, -5, -6, -7)

    def find(i, target, path):
        for bit in range(2):
            nxt = tree[i + bit]
            path.append(bit == 1)
            if nxt <= 0:
                if nxt == -target:
                    return True
            elif find(nxt, target, path):
                return True
            path.pop()
        return False

    a = abs(value)
    bits = 0.0
    if a > 7:
        bits += bb(p[0], True)
        for i in range(3):
            bits += bb(p[9 + i], (a >> i) & 1)
        i = 9
        while True:
            bits += bb(p[9 + i], (a >> i) & 1)
            i -= 1
            if i <= 3:
                break
        if a & 0xFFF0:
            bits += bb(p[9 + 3], (a >> 3) & 1)
    else:
        bits += bb(p[0], False)
        path = []
        assert find(0, a, path)
        tb, i = 0.0, 0
        for bit in path:
            tb += bb(p[2 + (i >> 1)], bit)
            nxt = tree[i + int(bit)]
            if nxt <= 0:
                break
            i = nxt
        bits += tb
    if a != 0:
        bits += bb(p[1], value < 0)
    return bits


def test_mv_component_bits_matches_tree_walk_over_full_range():
    import vp8g.inter_search as IS
    synthetic = [[0, 1, 7, 128, 200, 254, 255][i % 7] for i in range(19)]
    for ctx in (IS.DEFAULT_MV_CONTEXT[0], IS.DEFAULT_MV_CONTEXT[1], synthetic):
        for v in range(-1023, 1024):
            assert IS.mv_component_bits(ctx, v) == tree_walk_bits(ctx, v), v


def test_mv_component_bits_zero_cost_and_growth():
    import vp8g.inter_search as IS
    ctx = IS.DEFAULT_MV_CONTEXT[0]
    b0 = IS.mv_component_bits(ctx, 0)
    assert 0.0 < b0 < 8.0
    b4, b32, b256 = (IS.mv_component_bits(ctx, v) for v in (4, 32, 256))
    assert b0 < b4 < b32 < b256


def test_cost_table_is_the_f32_of_component_bits():
    import vp8g.inter_search as IS
    t = IS.MvCostTable.new()
    for comp in (0, 1):
        for d in (-1023, -9, -1, 0, 1, 7, 8, 15, 16, 1023):
            assert t.bits[comp, d + 1023] == np.float32(IS.mv_component_bits(IS.DEFAULT_MV_CONTEXT[comp], d))
    assert t.diff_bits(4, 0, 4, 0) == float(np.float32(t.bits[0, 1023]) + np.float32(t.bits[1, 1023]))
    assert t.diff_bits(1024, 0, 0, 0) is None


# ------------------------------------------------------------ planes
def ramp_plane(w, h):
    """src/motion_search.rs:741-749."""
    y, x = np.mgrid[0:h, 0:w]
    return ((x + y * 3) % 251).astype(np.uint8)


def feature_plane(w, h, bg, fg, fx, fy, feat):
    """src/motion_search.rs:758-780."""
    p = np.full((h, w), bg, np.uint8)
    p[fy:fy + feat, fx:fx + feat] = fg
    return p


def mb(plane, mb_col, mb_row):
    return np.ascontiguousarray(plane[16 * mb_row:16 * mb_row + 16, 16 * mb_col:16 * mb_col + 16]).reshape(256)


def luma_ref(oracle, plane, keep):
    from oracle_ctypes import LumaRef
    p = np.ascontiguousarray(plane)
    keep.append(p)
    return LumaRef(p.ctypes.data, p.shape[1], p.shape[1], p.shape[0])


def diamond(oracle, plane, src, center, iters, mb_col=0, mb_row=0):
    from oracle_ctypes import MvS
    keep = []
    lr = luma_ref(oracle, plane, keep)
    s = np.ascontiguousarray(src, np.uint8)
    r = oracle.lib.orc_small_diamond_search_luma(C.byref(lr), mb_col, mb_row, s.ctypes.data, MvS(*center), iters)
    return (r.mv.row, r.mv.col), r.sad


def test_small_diamond_reference_cases(oracle):
    # exact translations (src/motion_search.rs:995-1055)
    ref = feature_plane(64, 32, 128, 240, 6, 4, 4)
    src = mb(feature_plane(64, 32, 128, 240, 4, 4, 4), 0, 0)
    assert diamond(oracle, ref, src, (0, 0), 32) == ((0, 8), 0)
    ref = feature_plane(64, 64, 100, 220, 8, 8, 5)
    src = mb(feature_plane(64, 64, 100, 220, 5, 6, 5), 0, 0)
    assert diamond(oracle, ref, src, (0, 0), 64) == ((8, 12), 0)
    # zero iterations and identical source (:1057-1096)
    ramp = ramp_plane(32, 32)
    assert diamond(oracle, ramp, mb(ramp, 0, 0), (0, 0), 0) == ((0, 0), 0)
    assert diamond(oracle, ramp, mb(ramp, 0, 0), (0, 0), 8) == ((0, 0), 0)
    # snapping toward zero and clamping (:1098-1149)
    flat = np.full((32, 32), 200, np.uint8)
    assert diamond(oracle, flat, np.full(256, 200, np.uint8), (1, 5), 0) == ((0, 4), 0)
    flat = np.full((32, 32), 128, np.uint8)
    (r, c), s = diamond(oracle, flat, np.full(256, 128, np.uint8), (32767, -32768), 4)
    assert -1023 <= r <= 1023 and -1023 <= c <= 1023 and r % 4 == 0 and c % 4 == 0 and s == 0
    # off-picture start and never-worse (:1151-1206)
    (r, c), s = diamond(oracle, ramp, mb(ramp, 0, 0), (-64, -64), 8)
    assert r % 4 == 0 and c % 4 == 0
    ref = feature_plane(48, 48, 90, 230, 6, 6, 5)
    src = mb(feature_plane(48, 48, 90, 230, 4, 6, 5), 0, 0)
    assert diamond(oracle, ref, src, (0, 0), 8) == ((0, 8), 0)


# ------------------------------------------------------------ whole-pixel search
def np_whole_sad(plane, x, y, src):
    h, w = plane.shape
    ys = np.clip(np.arange(y, y + 16), 0, h - 1)
    xs = np.clip(np.arange(x, x + 16), 0, w - 1)
    return int(np.abs(plane[ys][:, xs].astype(np.int64) - src.reshape(16, 16).astype(np.int64)).sum())


def make_jobs(n):
    import vp8g.inter_search as IS
    j = np.zeros(n, IS.MB_SEARCH)
    j["range_x"] = (-16, 16)
    j["range_y"] = (-16, 16)
    j["n_seeds"] = 1
    j["lambda_sad"] = 2.0
    j["descend"] = 1
    return j


SHAPE_WIDE = (16, 1, 1, 1)


def run_wps(oracle, plane, jobs, src, costs, shape=SHAPE_WIDE):
    return oracle.whole_pixel_search_batch(plane, plane.shape[1], plane.shape[1], plane.shape[0], costs.bits, jobs,
                                           src, shape)


def blurred(seed, h, w, passes=3):
    rng = np.random.default_rng(seed)
    p = rng.integers(0, 256, (h, w)).astype(np.float64)
    for _ in range(passes):
        p = (np.roll(p, 1, 0) + p + np.roll(p, -1, 0)) / 3
        p = (np.roll(p, 1, 1) + p + np.roll(p, -1, 1)) / 3
    return np.clip(np.round(p), 0, 255).astype(np.uint8)


def test_whole_pixel_search_finds_planted_translations(oracle):
    import vp8g.inter_search as IS
    costs = IS.MvCostTable.new()
    for dx, dy in ((3, -2), (-5, 1), (0, 4), (2, 3)):
        ref = feature_plane(96, 96, 100, 230, 40 + dx, 40 + dy, 7)
        src = mb(feature_plane(96, 96, 100, 230, 40, 40, 7), 2, 2)
        jobs = make_jobs(1)
        jobs["mb_col"], jobs["mb_row"] = 2, 2
        jobs["lambda_sad"] = 0.25
        out = run_wps(oracle, ref, jobs, src, costs)
        assert (int(out["mv"][0, 0]), int(out["mv"][0, 1])) == (4 * dy, 4 * dx) and int(out["sad"][0]) == 0


def test_whole_pixel_search_cost_identity_and_local_optimum(oracle):
    import vp8g.inter_search as IS
    costs = IS.MvCostTable.new()
    rng = np.random.default_rng(3)
    cur, ref = blurred(1, 96, 128), blurred(2, 96, 128)
    ref = np.roll(cur, (2, -3), (0, 1)) // 2 + ref // 2
    n = 24
    jobs = make_jobs(n)
    jobs["mb_col"] = rng.integers(0, 8, n)
    jobs["mb_row"] = rng.integers(0, 6, n)
    jobs["best"] = rng.integers(-12, 13, (n, 2)) * 4
    jobs["n_seeds"] = 3
    jobs["seeds"][:, :3] = rng.integers(-40, 41, (n, 3, 2))
    jobs["lambda_sad"] = rng.uniform(0.5, 12.0, n)
    jobs["good_enough"] = 0
    src = np.stack([mb(cur, int(c), int(r)) for c, r in zip(jobs["mb_col"], jobs["mb_row"])])
    out = run_wps(oracle, ref, jobs, src, costs, (64, 1, 1, 1))
    for i in range(n):
        row, col = int(out["mv"][i, 0]), int(out["mv"][i, 1])
        x0, y0 = 16 * int(jobs["mb_col"][i]), 16 * int(jobs["mb_row"][i])
        br, bc = int(jobs["best"][i, 0]), int(jobs["best"][i, 1])
        lam = float(jobs["lambda_sad"][i])
        sad = np_whole_sad(ref, x0 + col // 4, y0 + row // 4, src[i])
        assert int(out["sad"][i]) == sad
        assert float(out["cost"][i]) == sad + lam * costs.diff_bits(row, col, br, bc)
        if sad == 0:
            continue
        for ddx, ddy in ((-1, 0), (1, 0), (0, -1), (0, 1)):
            px, py = col // 4 + ddx, row // 4 + ddy
            if not (-16 <= px <= 16 and -16 <= py <= 16):
                continue
            b = costs.diff_bits(4 * py, 4 * px, br, bc)
            if b is None:
                continue
            assert not (np_whole_sad(ref, x0 + px, y0 + py, src[i]) + lam * b < float(out["cost"][i]))


def test_whole_pixel_search_seed_rounding_and_clamping(oracle):
    """Seeds (row, col) in quarter pixels round as (v + 2) >> 2 and clamp into
    the range: (5, -6) -> (px, py) = (-1, 1); (-3, 3) -> (1, -1);
    (2, 1) -> (0, 1); (100, 100) -> (16, 16) under a +-16 range.  With the
    only exact match at the hand-rounded position and no descent, the
    search returns it."""
    import vp8g.inter_search as IS
    costs = IS.MvCostTable.new()
    for seed, (px, py) in (((5, -6), (-1, 1)), ((-3, 3), (1, -1)), ((2, 1), (0, 1)), ((100, 100), (16, 16))):
        ref = feature_plane(128, 128, 60, 200, 48 + px, 48 + py, 6)
        src = mb(feature_plane(128, 128, 60, 200, 48, 48, 6), 3, 3)
        jobs = make_jobs(1)
        jobs["mb_col"], jobs["mb_row"] = 3, 3
        jobs["n_seeds"] = 2
        jobs["seeds"][0, 0] = (40, -40)   # rounds to (-10, 10): no match there
        jobs["seeds"][0, 1] = seed
        jobs["descend"] = 0
        jobs["lambda_sad"] = 0.1
        out = run_wps(oracle, ref, jobs, src, costs)
        assert (int(out["mv"][0, 0]), int(out["mv"][0, 1])) == (4 * py, 4 * px) and int(out["sad"][0]) == 0


def test_two_pixel_pattern_reaches_what_one_pixel_cannot(oracle):
    import vp8g.inter_search as IS
    costs = IS.MvCostTable.new()
    ref = feature_plane(96, 96, 100, 230, 44, 40, 6)
    src = mb(feature_plane(96, 96, 100, 230, 40, 40, 6), 2, 2)
    jobs = make_jobs(1)
    jobs["mb_col"], jobs["mb_row"] = 2, 2
    jobs["lambda_sad"] = 0.1
    jobs["good_enough"] = 10
    wide = run_wps(oracle, ref, jobs, src, costs, (2, 1, 1, 1))
    narrow = run_wps(oracle, ref, jobs, src, costs, (2, 0, 1, 1))
    assert (int(wide["mv"][0, 1]), int(wide["sad"][0])) == (16, 0)
    assert int(narrow["mv"][0, 1]) == 8 and int(narrow["sad"][0]) > 0


# ------------------------------------------------------------ sub-pixel refine
def predict_luma(oracle, plane, mb_col, mb_row, mv):
    from oracle_ctypes import MvS, RefPlanes
    p = np.ascontiguousarray(plane)
    h, w = p.shape
    cu = np.full((h // 2) * (w // 2), 128, np.uint8)
    rp = RefPlanes(p.ctypes.data, cu.ctypes.data, cu.ctypes.data, w, w // 2, w // 16, h // 16)
    out = np.zeros(384, np.uint8)
    f = np.ascontiguousarray(oracle.SIXTAP)
    oracle.lib.orc_predict_inter_mb(C.byref(rp), mb_col, mb_row, MvS(*mv), 0, f.ctypes.data, out.ctypes.data)
    return out[:256].copy()


def run_spr(oracle, plane, jobs, src, start, costs, shape=(16, 1, 1, 1), filters=None):
    f = oracle.SIXTAP if filters is None else filters
    return oracle.sub_pixel_refine_batch(plane, plane.shape[1], plane.shape[1], plane.shape[0], costs.bits, jobs,
                                         src, start, f, shape)


def start_at(costs, plane, jobs, src, i, row, col):
    import vp8g.inter_search as IS
    st = np.zeros(1, IS.FOUND)
    x0, y0 = 16 * int(jobs["mb_col"][i]), 16 * int(jobs["mb_row"][i])
    sad = np_whole_sad(plane, x0 + col // 4, y0 + row // 4, src)
    st["mv"] = (row, col)
    st["sad"] = sad
    st["cost"] = sad + float(jobs["lambda_sad"][i]) * costs.diff_bits(row, col, int(jobs["best"][i, 0]),
                                                                      int(jobs["best"][i, 1]))
    return st


def test_sub_pixel_refine_finds_planted_half_and_quarter_pel(oracle):
    import vp8g.inter_search as IS
    costs = IS.MvCostTable.new()
    ref = blurred(9, 96, 96, passes=2)
    for mv, needs_quarter in (((2, -2), False), ((0, 2), False), ((1, 0), True), ((-2, 1), True)):
        src = predict_luma(oracle, ref, 2, 2, mv)
        jobs = make_jobs(1)
        jobs["mb_col"], jobs["mb_row"] = 2, 2
        jobs["lambda_sad"] = 0.01
        st = start_at(costs, ref, jobs, src, 0, 0, 0)
        got = run_spr(oracle, ref, jobs, src, st, costs)
        assert (int(got["mv"][0, 0]), int(got["mv"][0, 1])) == mv and int(got["sad"][0]) == 0
        if needs_quarter:
            no_q = run_spr(oracle, ref, jobs, src, st, costs, shape=(16, 1, 1, 0))
            assert int(no_q["sad"][0]) > 0


def test_sub_pixel_refine_sads_and_costs(oracle):
    """Every chosen vector's SAD is mb_luma_sad_at_mv's (six-tap: HalfPlanes
    with the six-tap half taps is the two-pass filter at half-pel vectors),
    the cost is SAD + lambda * bits, and never above the start's."""
    import vp8g.inter_search as IS
    from oracle_ctypes import LumaRef, MvS
    costs = IS.MvCostTable.new()
    rng = np.random.default_rng(12)
    cur, ref = blurred(21, 96, 128, 2), blurred(22, 96, 128, 2)
    ref = np.roll(cur, (1, 2), (0, 1)) // 2 + ref // 2
    n = 30
    jobs = make_jobs(n)
    jobs["mb_col"] = rng.integers(0, 8, n)
    jobs["mb_row"] = rng.integers(0, 6, n)
    jobs["best"] = rng.integers(-8, 9, (n, 2)) * 2
    jobs["lambda_sad"] = rng.uniform(0.5, 10.0, n)
    src = np.stack([mb(cur, int(c), int(r)) for c, r in zip(jobs["mb_col"], jobs["mb_row"])])
    starts = np.concatenate([start_at(costs, ref, jobs, src[i], i, *(4 * rng.integers(-4, 5, 2)))
                             for i in range(n)])
    for shape in ((16, 1, 1, 1), (16, 1, 0, 1), (16, 1, 1, 0)):
        out = run_spr(oracle, ref, jobs, src, starts, costs, shape)
        keep = []
        lr = LumaRef(np.ascontiguousarray(ref).ctypes.data, 128, 128, 96)
        keep.append(ref)
        for i in range(n):
            row, col = int(out["mv"][i, 0]), int(out["mv"][i, 1])
            s = np.ascontiguousarray(src[i])
            want = oracle.lib.orc_mb_luma_sad_at_mv(C.byref(lr), int(jobs["mb_col"][i]), int(jobs["mb_row"][i]),
                                                    s.ctypes.data, MvS(row, col))
            assert int(out["sad"][i]) == want
            b = costs.diff_bits(row, col, int(jobs["best"][i, 0]), int(jobs["best"][i, 1]))
            assert float(out["cost"][i]) == want + float(jobs["lambda_sad"][i]) * b
            assert float(out["cost"][i]) <= float(starts["cost"][i])
            if not shape[2]:
                d = (row - int(starts["mv"][i, 0]), col - int(starts["mv"][i, 1]))
                assert not (abs(d[0]) == 2 and abs(d[1]) == 2), "diagonal probed with half_diagonals off"
