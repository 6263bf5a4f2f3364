"""Pins the oracle's exhaustive block search and its altref-from-a-table
entry.  The search is a project-defined judge (the reference's ARNR search
is the 4-step descent of src/arnr.rs:264-300); it is pinned here by an
independent numpy brute force, planted ties, exact translations and the
invariant that it never loses to the descent, and the table-driven altref
by reproducing build_arnr_altref from the descent's own table."""
import numpy as np
import pytest


def clamp_gather(plane, h, w, ys, xs):
    return plane[np.clip(ys, 0, h - 1)[:, None], np.clip(xs, 0, w - 1)[None, :]]


def np_block_search(src, ref, bx, by, bw, bh, sx, sy, radius):
    """All (2R+1)^2 SADs by array gathers, then one lexsort on
    (sad, |mvx|+|mvy|, mvy, mvx)."""
    h, w = src.shape
    a = src[by:by + bh, bx:bx + bw].astype(np.int64)
    rows = []
    for dy in range(-radius, radius + 1):
        for dx in range(-radius, radius + 1):
            mx, my = sx + dx, sy + dy
            b = clamp_gather(ref, h, w, np.arange(by, by + bh) + my, np.arange(bx, bx + bw) + mx)
            rows.append((int(np.abs(a - b).sum()), abs(mx) + abs(my), my, mx))
    t = np.array(rows)
    i = np.lexsort((t[:, 3], t[:, 2], t[:, 1], t[:, 0]))[0]
    return int(t[i, 3]), int(t[i, 2]), int(t[i, 0])


def frame_tuple(y):
    h, w = y.shape
    cw, ch = (w + 1) // 2, (h + 1) // 2
    z = np.full(cw * ch, 128, np.uint8)
    return (w, h, np.ascontiguousarray(y).reshape(-1), z, z, w, cw)


@pytest.mark.parametrize("radius", [0, 1, 2, 5])
def test_block_search_matches_brute_force(oracle, radius):
    import ctypes as C
    from oracle_ctypes import I420
    rng = np.random.default_rng(100 + radius)
    h, w = 37, 45
    src = rng.integers(0, 256, (h, w)).astype(np.uint8)
    ref = np.clip(np.roll(src, (2, -3), (0, 1)).astype(int) + rng.integers(-9, 10, (h, w)), 0, 255).astype(np.uint8)
    ref[5:9, 10:30] = 7  # flat patch -> many equal SADs
    keep = []
    a = oracle.i420(*frame_tuple(src), keep=keep)
    b = oracle.i420(*frame_tuple(ref), keep=keep)
    for (bx, by) in [(0, 0), (16, 0), (32, 16), (16, 32), (32, 32)]:
        bw, bh = min(16, w - bx), min(16, h - by)
        for (sx, sy) in [(0, 0), (-3, 2), (7, -6)]:
            out = (C.c_int32 * 3)()
            oracle.lib.orc_block_search(C.byref(a), C.byref(b), bx, by, bw, bh, sx, sy, radius, out)
            assert tuple(out) == np_block_search(src, ref, bx, by, bw, bh, sx, sy, radius)


def test_ties_go_to_the_shortest_vector(oracle):
    import ctypes as C
    flat = np.full((32, 32), 90, np.uint8)
    keep = []
    a = oracle.i420(*frame_tuple(flat), keep=keep)
    out = (C.c_int32 * 3)()
    oracle.lib.orc_block_search(C.byref(a), C.byref(a), 16, 16, 16, 16, 0, 0, 4, out)
    assert tuple(out) == (0, 0, 0)
    # every candidate around (3, -2) ties: the smallest |mvx|+|mvy| wins,
    # then the smaller mvy, then the smaller mvx
    oracle.lib.orc_block_search(C.byref(a), C.byref(a), 0, 0, 16, 16, 3, -2, 2, out)
    assert tuple(out) == (1, 0, 0)
    oracle.lib.orc_block_search(C.byref(a), C.byref(a), 0, 0, 16, 16, -2, -2, 1, out)
    assert tuple(out) == (-1, -1, 0)


def diagonal_pair(n=64, seed=5):
    """src(x, y) = g[x + y], ref(x, y) = g[x + y - 1]: every vector with
    mvx + mvy = 1 matches exactly, so (1, 0) and (0, 1) tie at |mv| = 1 and
    the smaller mvy must win."""
    g = np.random.default_rng(seed).integers(0, 256, 2 * n + 2).astype(np.uint8)
    yy, xx = np.mgrid[0:n, 0:n]
    return g[xx + yy + 1], g[xx + yy]


def test_equal_length_ties_go_to_the_smaller_mvy(oracle):
    import ctypes as C
    src, ref = diagonal_pair()
    keep = []
    a = oracle.i420(*frame_tuple(src), keep=keep)
    b = oracle.i420(*frame_tuple(ref), keep=keep)
    out = (C.c_int32 * 3)()
    oracle.lib.orc_block_search(C.byref(a), C.byref(b), 16, 16, 16, 16, 0, 0, 2, out)
    assert tuple(out) == (1, 0, 0)
    oracle.lib.orc_block_search(C.byref(a), C.byref(b), 16, 16, 16, 16, -3, 5, 2, out)
    assert tuple(out) == (-2, 3, 0)  # mvx + mvy = 1 with the smallest |mvx|+|mvy| in the box


@pytest.mark.parametrize("shift", [(0, 0), (3, -5), (-16, 16), (11, 2)])
def test_exact_translation_is_found(oracle, shift):
    dx, dy = shift
    rng = np.random.default_rng(7)
    big = rng.integers(0, 256, (140, 140)).astype(np.uint8)
    src = big[40:104, 40:104]
    ref = big[40 + dy:104 + dy, 40 + dx:104 + dx]  # ref(x + dx, y + dy) = src(x, y)... shifted window
    n_src, n_ref = frame_tuple(src), frame_tuple(ref)
    tab = oracle.block_search_group([n_ref, n_src], 1, 16)
    # interior blocks of frame 0 match exactly at mv = (-dx, -dy)
    for by in (1, 2):
        for bx in (1, 2):
            if abs(dx) <= 16 and abs(dy) <= 16:
                assert tuple(tab[0, by, bx]) == (-dx, -dy, 0)
    assert (tab[1] == 0).all()  # centre rows untouched


def test_exhaustive_never_loses_to_the_descent(oracle):
    import vp8g.synth as synth
    g = synth.arnr_group_regions(96, 80, 5, vmax=6, occluders=2, clip=4, config=3)
    st, _, _, desc = oracle.build_arnr_altref(g.frames(), g.center, 3, want_mvs=True)
    assert st == 0
    ex = oracle.block_search_group(g.frames(), g.center, 16)
    for f in range(g.n):
        if f == g.center:
            continue
        assert (ex[f, :, :, 2] <= desc[f, :, :, 2]).all()
        # where the descent found the same SAD the exhaustive search found,
        # the exhaustive winner is never longer
        same = ex[f, :, :, 2] == desc[f, :, :, 2]
        l1e = np.abs(ex[f, :, :, 0]) + np.abs(ex[f, :, :, 1])
        l1d = np.abs(desc[f, :, :, 0]) + np.abs(desc[f, :, :, 1])
        assert (l1e[same] <= l1d[same]).all()


@pytest.mark.parametrize("strength", [0, 1, 3, 6])
def test_altref_from_the_descent_table_is_the_altref(oracle, strength):
    import vp8g.synth as synth
    g = synth.arnr_group_regions(80, 56, 5, vmax=5, occluders=3, clip=2, config=3)
    st, want, _, tab = oracle.build_arnr_altref(g.frames(), g.center, strength, want_mvs=True)
    assert st == 0
    st, got = oracle.build_arnr_altref_mv(g.frames(), g.center, strength, tab)
    assert st == 0
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


def test_altref_from_a_rejecting_table_is_the_centre(oracle):
    import vp8g.synth as synth
    g = synth.arnr_group_cif(48, 40, 3, clip=1)
    tab = np.zeros((3, 3, 3, 3), np.int32)
    tab[..., 2] = 13 * 256  # every block over the occlusion cut-off
    st, got = oracle.build_arnr_altref_mv(g.frames(), g.center, 4, tab)
    assert st == 0
    assert np.array_equal(got[0], g.y[g.center].reshape(-1))
    assert np.array_equal(got[1], g.u[g.center].reshape(-1))


# ---------------------------------------------------------------- fallback
FAILS = 65535  # above any 16x16 SAD (<= 65280): always fails the cut-off


def _fallback_group(seed, w=64, h=48, n=3):
    rng = np.random.default_rng(seed)
    return [frame_tuple(rng.integers(0, 256, (h, w)).astype(np.uint8)) for _ in range(n)]


def test_fallback_queue_rule_known_answer(oracle):
    """A hand-set failure pattern on a 4 x 3 block grid: the gated queue is
    the 7 failed blocks with a passing 4-neighbour, the ungated one all 9
    failed blocks; queued entries take the exhaustive result (every first
    SAD is above any real SAD), the others are copied."""
    frames = _fallback_group(7)
    n, center = len(frames), 1
    pattern = ["FPFF", "FFFF", "PFFP"]
    first = np.zeros((n, 3, 4, 3), np.int32)
    for f in (0, 2):
        for r, row in enumerate(pattern):
            for c, ch in enumerate(row):
                first[f, r, c] = (5, -3, FAILS if ch == "F" else 0)
    exhaustive = oracle.block_search_group(frames, center, 4)
    gated = {(0, 0), (0, 2), (1, 0), (1, 1), (1, 3), (2, 1), (2, 2)}
    for gate, want_set in ((True, gated), (False, {(r, c) for r in range(3) for c in range(4)
                                                     if pattern[r][c] == "F"})):
        out, q = oracle.search_fallback(frames, center, first, radius=4, gate=gate)
        assert q == 2 * len(want_set)
        for f in (0, 2):
            for r in range(3):
                for c in range(4):
                    want = exhaustive[f, r, c] if (r, c) in want_set else first[f, r, c]
                    assert tuple(out[f, r, c]) == tuple(want), (gate, f, r, c)
        assert (out[center] == first[center]).all()


def test_fallback_keeps_the_better_entry_under_the_key_order(oracle):
    """Random frames: every block fails, so with the gate off every entry
    is queued.  A first entry equal to the exhaustive minimum in SAD wins
    only by the (|mv|, mvy, mvx) order."""
    frames = _fallback_group(11)
    n, center = len(frames), 0
    ex = oracle.block_search_group(frames, center, 3)
    first = ex.copy()
    cases = {}
    for f in (1, 2):
        for r in range(3):
            for c in range(4):
                mx, my, sad = (int(v) for v in ex[f, r, c])
                k = (3 * f + r * 4 + c) % 4
                if k == 0:    # same SAD, longer vector: the exhaustive entry wins
                    first[f, r, c] = (mx + 40, my, sad)
                    cases[(f, r, c)] = (mx, my, sad)
                elif k == 1:  # same SAD and length, mvy = -length: the smaller (mvy, mvx) wins
                    l1 = abs(mx) + abs(my)
                    first[f, r, c] = (0, -l1, sad)
                    cases[(f, r, c)] = (0, -l1, sad) if (-l1, 0) < (my, mx) else (mx, my, sad)
                elif k == 2:  # smaller SAD than any real one: the first entry wins
                    first[f, r, c] = (99, 99, max(0, sad - 1)) if sad >= 13 * 256 else (mx, my, sad)
                    cases[(f, r, c)] = tuple(first[f, r, c])
                else:         # larger SAD: the exhaustive entry wins
                    first[f, r, c] = (1, 1, FAILS)
                    cases[(f, r, c)] = (mx, my, sad)
    out, q = oracle.search_fallback(frames, center, first, radius=3, gate=False)
    assert q == sum(1 for (f, r, c) in cases if int(first[f, r, c, 2]) >= 13 * 256)
    for (f, r, c), want in cases.items():
        assert tuple(out[f, r, c]) == tuple(want), (f, r, c)
