"""GPU parity for the exhaustive block search (vp8g_block_search_dev) and
the table-driven altref (vp8g_build_arnr_altref_table_dev) against the
oracle, bit-exact, through the C ABI."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class Dev:
    """A group's planes in device memory plus the ctypes descriptor array."""

    def __init__(self, tuples):
        import torch
        import vp8g
        self.tuples = tuples
        self.t = []
        self.descs = (vp8g._I420 * len(tuples))()
        for i, (w, h, y, u, v, ys, cs) in enumerate(tuples):
            ty, tu, tv = (torch.from_numpy(np.ascontiguousarray(p)).cuda() for p in (y, u, v))
            self.t.append((ty, tu, tv))
            self.descs[i] = vp8g._I420(w, h, ty.data_ptr(), tu.data_ptr(), tv.data_ptr(), ys, cs)
        w, h = tuples[0][0], tuples[0][1]
        self.shape = (len(tuples), (h + 15) // 16, (w + 15) // 16, 3)

    def table(self, fill=0):
        import torch
        return torch.full(self.shape, fill, dtype=torch.int32, device="cuda")


def search(gpu, dev, center, radius, seeds=None):
    import torch
    out = dev.table(-7)
    sd = None
    if seeds is not None:
        sd = torch.from_numpy(np.ascontiguousarray(seeds, np.int32)).cuda()
    torch.cuda.synchronize()
    gpu.block_search_dev(dev.descs, len(dev.tuples), center, radius, out.data_ptr(),
                         sd.data_ptr() if sd is not None else None)
    gpu.synchronize()
    got = out.cpu().numpy()
    return got


def expect_table(oracle, tuples, center, radius, seeds=None):
    want = oracle.block_search_group(tuples, center, radius, seeds)
    want[center] = -7  # centre rows untouched
    return want


def assert_tables(got, want, tag):
    if not np.array_equal(got, want):
        bad = np.argwhere((got != want).any(-1))
        f, r, c = bad[0]
        raise AssertionError(f"{tag}: {len(bad)} blocks differ; first (frame {f}, row {r}, col {c}): "
                             f"got {got[f, r, c]} want {want[f, r, c]}")


def tuples_of(w, h, n, seed, pad=0, noise=6):
    rng = np.random.default_rng(seed)
    cw, ch = (w + 1) // 2, (h + 1) // 2
    base = rng.integers(0, 256, (h + 40, w + 40)).astype(np.int64)
    out = []
    for i in range(n):
        dx, dy = int(rng.integers(-9, 10)), int(rng.integers(-9, 10))
        y = np.clip(base[20 + dy:20 + dy + h, 20 + dx:20 + dx + w] + rng.integers(-noise, noise + 1, (h, w)), 0, 255)
        ys = np.full((h, w + pad), 1, np.uint8)
        ys[:, :w] = y
        c = rng.integers(0, 256, (ch, cw + pad)).astype(np.uint8)
        out.append((w, h, ys.reshape(-1), c.reshape(-1), c.reshape(-1).copy(), w + pad, cw + pad))
    return out


GEOMS = [(16, 16, 2, 0, 0), (48, 32, 3, 1, 0), (37, 29, 3, 2, 0), (45, 37, 4, 1, 3), (13, 7, 2, 1, 0),
         (100, 70, 5, 2, 5), (64, 48, 3, 0, 0)]


@pytest.mark.parametrize("radius", [2, 16, 64])
@pytest.mark.parametrize("geom", GEOMS, ids=lambda g: "x".join(map(str, g[:3])) + f"c{g[3]}p{g[4]}")
def test_block_search_matches_oracle(gpu, oracle, radius, geom):
    w, h, n, center, pad = geom
    tuples = tuples_of(w, h, n, w * 131 + h * 7 + n + radius, pad)
    dev = Dev(tuples)
    got = search(gpu, dev, center, radius)
    assert_tables(got, expect_table(oracle, tuples, center, radius), f"R={radius} {geom}")


@pytest.mark.parametrize("radius", [2, 16, 64])
def test_seeded_search_matches_oracle(gpu, oracle, radius):
    tuples = tuples_of(72, 56, 4, 99 + radius, 3)
    dev = Dev(tuples)
    rng = np.random.default_rng(radius)
    seeds = np.zeros(dev.shape, np.int32)
    seeds[..., :2] = rng.integers(-40, 41, dev.shape[:3] + (2,))
    seeds[0, 0, 0, :2] = (20000, -30000)  # clamped to +-16384
    got = search(gpu, dev, 1, radius, seeds)
    assert_tables(got, expect_table(oracle, tuples, 1, radius, seeds), f"seeded R={radius}")


def test_search_ties_and_translations(gpu):
    def ft(y):
        h, w = y.shape
        cw, ch = (w + 1) // 2, (h + 1) // 2
        z = np.full(cw * ch, 128, np.uint8)
        return (w, h, np.ascontiguousarray(y).reshape(-1), z, z, w, cw)
    g = np.random.default_rng(5).integers(0, 256, 130).astype(np.uint8)
    yy, xx = np.mgrid[0:64, 0:64]
    src, ref = g[xx + yy + 1], g[xx + yy]
    dev = Dev([ft(ref), ft(src)])
    got = search(gpu, dev, 1, 2)
    assert tuple(got[0, 1, 1]) == (1, 0, 0)   # (1,0) and (0,1) tie; smaller mvy wins
    flat = np.full((48, 48), 77, np.uint8)
    got = search(gpu, Dev([ft(flat), ft(flat)]), 1, 16)
    assert (got[0] == 0).all()                 # every candidate ties: (0, 0)
    big = np.random.default_rng(8).integers(0, 256, (400, 400)).astype(np.uint8)
    for (dx, dy) in ((13, -7), (-60, 41)):
        a = big[100:292, 100:292]
        b = big[100 + dy:292 + dy, 100 + dx:292 + dx]
        got = search(gpu, Dev([ft(b), ft(a)]), 1, 64)
        # block (5, 5) matches inside the frame at (80 - dx, 80 - dy)
        assert tuple(got[0, 5, 5]) == (-dx, -dy, 0)


def test_1080p_exhaustive_16_matches_oracle(gpu, oracle):
    import vp8g.synth as synth
    g = synth.arnr_group_regions()
    dev = Dev(g.frames())
    got = search(gpu, dev, g.center, 16)
    assert_tables(got, expect_table(oracle, g.frames(), g.center, 16), "1080p R=16")


def table_altref(gpu, dev, center, strength, table):
    import torch
    w, h = dev.tuples[0][0], dev.tuples[0][1]
    cw, ch = (w + 1) // 2, (h + 1) // 2
    oy = torch.full((w * h,), 3, dtype=torch.uint8, device="cuda")
    ou = torch.full((cw * ch,), 3, dtype=torch.uint8, device="cuda")
    ov = torch.full((cw * ch,), 3, dtype=torch.uint8, device="cuda")
    t = torch.from_numpy(np.ascontiguousarray(table, np.int32)).cuda()
    torch.cuda.synchronize()
    gpu.build_arnr_altref_table_dev(dev.descs, len(dev.tuples), center, strength, t.data_ptr(),
                                    oy.data_ptr(), ou.data_ptr(), ov.data_ptr())
    gpu.synchronize()
    return oy.cpu().numpy(), ou.cpu().numpy(), ov.cpu().numpy()


@pytest.mark.parametrize("strength", [0, 3, 6])
def test_table_altref_from_the_descent_table_equals_the_altref(gpu, oracle, strength):
    import vp8g.synth as synth
    g = synth.arnr_group_regions(200, 120, 7, vmax=8, occluders=4, clip=3, config=3)
    st, want, _, tab = oracle.build_arnr_altref(g.frames(), g.center, strength, want_mvs=True)
    assert st == 0
    got = table_altref(gpu, Dev(g.frames()), g.center, strength, tab)
    for name, a, b in zip("yuv", got, want):
        assert np.array_equal(a, b), name


@pytest.mark.parametrize("radius", [16, 64])
def test_table_altref_from_exhaustive_vectors(gpu, oracle, radius):
    """Vectors beyond the descent's +-15 window (clamped reads)."""
    import vp8g.synth as synth
    g = synth.arnr_group_regions(176, 144, 5, vmax=24, occluders=3, clip=4, config=4)
    dev = Dev(g.frames())
    tab = search(gpu, dev, g.center, radius)
    tab[g.center] = 0
    st, want = oracle.build_arnr_altref_mv(g.frames(), g.center, 3, tab)
    assert st == 0
    got = table_altref(gpu, dev, g.center, 3, tab)
    for name, a, b in zip("yuv", got, want):
        assert np.array_equal(a, b), name


# (w, h, pad): 4-byte strides with a ragged right edge (dword reads), odd
# strides and rows under 8 pixels (byte reads)
@pytest.mark.parametrize("geom", [(45, 37, 3), (13, 7, 0), (6, 20, 2), (37, 29, 0), (8, 9, 0), (64, 33, 0)],
                         ids=lambda g: "x".join(map(str, g)))
def test_table_altref_large_vectors_ragged(gpu, oracle, geom):
    w, h, pad = geom
    tuples = tuples_of(w, h, 5, 17, pad)
    rng = np.random.default_rng(w * h)
    bh, bw = (h + 15) // 16, (w + 15) // 16
    t2 = np.zeros((5, bh, bw, 3), np.int32)
    t2[..., :2] = rng.integers(-80, 81, (5, bh, bw, 2))
    t2[..., :2][rng.random((5, bh, bw)) < 0.5] = rng.integers(-3, 4, 2)
    t2[..., 2] = rng.integers(0, 20 * min(w, 16) * min(h, 16), (5, bh, bw))
    st, want = oracle.build_arnr_altref_mv(tuples, 2, 5, t2)
    assert st == 0
    got = table_altref(gpu, Dev(tuples), 2, 5, t2)
    for name, a, b in zip("yuv", got, want):
        assert np.array_equal(a, b), name


def abs_key(t):
    """(sad, |mvx| + |mvy|, mvy, mvx) as one ordered integer key."""
    t = t.astype(np.int64)
    return (t[..., 2] << 48) + ((np.abs(t[..., 0]) + np.abs(t[..., 1])) << 32) + ((t[..., 1] + 32768) << 16) + \
        (t[..., 0] + 32768)


def expect_candidates(oracle, tuples, center, seeds, first):
    """The second stage restated with the oracle's +-2 block search: slot 0
    holds the own first-stage vector, searched again only when it lies on
    the border of the first square around the seed; slots 1..4 hold the
    left, right, upper and lower blocks' vectors (an absent neighbour repeats
    the own vector), each searched when it differs from the seed and from
    every earlier slot."""
    n, bh, bw, _ = first.shape
    own = first[..., :2]
    slots = [own]
    for dr, dc in ((0, -1), (0, 1), (-1, 0), (1, 0)):
        nb = own.copy()
        rs, cs = slice(max(0, -dr), bh - max(0, dr)), slice(max(0, -dc), bw - max(0, dc))
        rd, cd = slice(max(0, dr), bh - max(0, -dr)), slice(max(0, dc), bw - max(0, -dc))
        nb[:, rs, cs] = own[:, rd, cd]
        slots.append(nb)
    sd = seeds[..., :2]
    best = first.copy()
    for t, cand in enumerate(slots):
        use = (cand != sd).any(-1)
        if t == 0:
            use &= (np.abs(cand - sd) == 2).any(-1)
        for u in range(t):
            use &= (slots[u] != cand).any(-1)
        tab = np.zeros_like(first)
        tab[..., :2] = cand
        res = oracle.block_search_group(tuples, center, 2, tab)
        better = use & (abs_key(res) < abs_key(best))
        best[better] = res[better]
    return best


@pytest.mark.parametrize("geom", [(176, 144, 5, 0), (45, 37, 3, 3), (100, 70, 5, 1)], ids=lambda g: "x".join(map(str, g)))
def test_refine_candidates_matches_oracle(gpu, oracle, geom):
    import torch
    w, h, n, pad = geom
    tuples = tuples_of(w, h, n, 23, pad)
    center = n // 2
    dev = Dev(tuples)
    rng = np.random.default_rng(w + h)
    exact = expect_table(oracle, tuples, center, 16)
    seeds = exact.copy()
    seeds[..., :2] += rng.integers(-5, 6, seeds[..., :2].shape) * (rng.random(seeds.shape[:3]) < 0.5)[..., None]
    seeds[..., 2] = 0
    seeds[center] = 0
    first = search(gpu, dev, center, 2, seeds)
    assert_tables(first, expect_table(oracle, tuples, center, 2, seeds), "first stage")
    ts = torch.from_numpy(np.ascontiguousarray(seeds, np.int32)).cuda()
    tf = torch.from_numpy(np.ascontiguousarray(first, np.int32)).cuda()
    out = dev.table(-7)
    torch.cuda.synchronize()
    gpu.refine_candidates_dev(dev.descs, n, center, ts.data_ptr(), tf.data_ptr(), out.data_ptr())
    gpu.synchronize()
    got = out.cpu().numpy()
    others = [i for i in range(n) if i != center]
    want = expect_candidates(oracle, tuples, center, seeds, first)
    assert_tables(got[others], want[others], "candidates")
    assert (got[center] == -7).all()
    # the second stage never loses against the first
    assert (abs_key(got[others]) <= abs_key(first[others])).all()


@pytest.mark.parametrize("geom", [(64, 48, 3, 0), (40, 24, 2, 1)], ids=lambda g: "x".join(map(str, g)))
def test_refine_candidates_centre_first_without_seeds(gpu, oracle, geom):
    """Centre frame 0 (every neighbour frame lies after it) and no seed
    table: the seeds are then the zero vector, as a zero table gives."""
    import torch
    w, h, n, pad = geom
    tuples = tuples_of(w, h, n, 29, pad)
    center = 0
    dev = Dev(tuples)
    zero = np.zeros((n, (h + 15) // 16, (w + 15) // 16, 3), np.int32)
    first = search(gpu, dev, center, 2)
    assert_tables(first, expect_table(oracle, tuples, center, 2, zero), "first stage")
    tf = torch.from_numpy(np.ascontiguousarray(first, np.int32)).cuda()
    out = dev.table(-7)
    torch.cuda.synchronize()
    gpu.refine_candidates_dev(dev.descs, n, center, None, tf.data_ptr(), out.data_ptr())
    gpu.synchronize()
    got = out.cpu().numpy()
    want = expect_candidates(oracle, tuples, center, zero, first)
    assert_tables(got[1:], want[1:], "candidates")
    assert (got[0] == -7).all()


def fallback(gpu, dev, center, first, radius=16, gate=True):
    import torch
    tf = torch.from_numpy(np.ascontiguousarray(first, np.int32)).cuda()
    out = dev.table(-7)
    q = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    gpu.search_fallback_dev(dev.descs, len(dev.tuples), center, tf.data_ptr(), out.data_ptr(), radius, gate,
                            q.data_ptr())
    gpu.synchronize()
    return out.cpu().numpy(), int(q.item())


@pytest.mark.parametrize("radius", [16, 64])
@pytest.mark.parametrize("gate", [True, False], ids=["gated", "all"])
@pytest.mark.parametrize("geom", [(176, 144, 5, 0), (45, 37, 3, 3), (100, 70, 5, 1), (64, 48, 3, 0)],
                         ids=lambda g: "x".join(map(str, g)))
def test_search_fallback_matches_oracle(gpu, oracle, radius, gate, geom):
    """First stage: +-2 around seeds that are exact on half of the blocks and
    up to 12 px off elsewhere, so both passing and failing entries occur."""
    w, h, n, pad = geom
    tuples = tuples_of(w, h, n, 31 + w + radius, pad)
    center = n // 2
    dev = Dev(tuples)
    rng = np.random.default_rng(w * h + radius)
    seeds = expect_table(oracle, tuples, center, 16)
    off = rng.integers(-12, 13, seeds[..., :2].shape) * (rng.random(seeds.shape[:3]) < 0.5)[..., None]
    seeds[..., :2] += off
    seeds[..., 2] = 0
    seeds[center] = 0
    first = search(gpu, dev, center, 2, seeds)
    assert_tables(first, expect_table(oracle, tuples, center, 2, seeds), "first stage")
    got, q = fallback(gpu, dev, center, first, radius, gate)
    want, wq = oracle.search_fallback(tuples, center, first, radius, gate)
    want[center] = -7
    assert_tables(got, want, f"fallback R={radius} gate={gate}")
    assert q == wq
    others = [i for i in range(n) if i != center]
    assert (abs_key(got[others]) <= abs_key(first[others])).all()


def test_search_fallback_rejects_bad_arguments(gpu):
    import torch
    import vp8g
    tuples = tuples_of(32, 32, 3, 5)
    dev = Dev(tuples)
    t = dev.table()
    with pytest.raises(vp8g.Vp8gError):   # radius 2 is not a fallback radius
        gpu.search_fallback_dev(dev.descs, 3, 1, t.data_ptr(), dev.table().data_ptr(), radius=2)
    with pytest.raises(vp8g.Vp8gError):   # out aliases first
        gpu.search_fallback_dev(dev.descs, 3, 1, t.data_ptr(), t.data_ptr())
    with pytest.raises(vp8g.Vp8gError) as ei:   # centre index out of range, checked first
        gpu.search_fallback_dev(dev.descs, 3, 5, t.data_ptr(), t.data_ptr(), radius=2)
    assert ei.value.status == vp8g.VP8G_E_CENTER_OUT_OF_RANGE
    torch.cuda.synchronize()
