"""vp8g_search_fallback_noalloc_dev (the fallback with its queue in the
library's static device scratch; no device allocation) against
vp8g_search_fallback_dev (queue in the context's workspace): identical
tables and queue counts, bit-exact, on two calls in a row with the second
group smaller than the first, and rejection of groups beyond the scratch
limit before any work is enqueued.  Every call goes through the C ABI."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCRATCH_ITEMS = 1 << 20


def _group(w, h, n, seed):
    rng = np.random.default_rng(seed)
    cw, ch = (w + 1) // 2, (h + 1) // 2
    base = rng.integers(0, 256, (h + 40, w + 40)).astype(np.int64)
    out = []
    for _ in range(n):
        dx, dy = int(rng.integers(-9, 10)), int(rng.integers(-9, 10))
        y = np.clip(base[20 + dy:20 + dy + h, 20 + dx:20 + dx + w] + rng.integers(-6, 7, (h, w)), 0, 255)
        c = rng.integers(0, 256, (ch, cw)).astype(np.uint8)
        out.append((w, h, y.astype(np.uint8).reshape(-1), c.reshape(-1), c.reshape(-1).copy(), w, cw))
    return out


class _Dev:
    def __init__(self, tuples):
        import torch
        import vp8g
        self.t = []
        self.descs = (vp8g._I420 * len(tuples))()
        for i, (w, h, y, u, v, ys, cs) in enumerate(tuples):
            ty, tu, tv = (torch.from_numpy(np.ascontiguousarray(p)).cuda() for p in (y, u, v))
            self.t.append((ty, tu, tv))
            self.descs[i] = vp8g._I420(w, h, ty.data_ptr(), tu.data_ptr(), tv.data_ptr(), ys, cs)
        w, h = tuples[0][0], tuples[0][1]
        self.n = len(tuples)
        self.shape = (self.n, (h + 15) // 16, (w + 15) // 16, 3)


def _first_stage(gpu, oracle, tuples, dev, center, seed):
    """A +-2 table around seeds that are exact on half of the blocks and up
    to 12 px off on the rest (passing and failing entries both occur)."""
    import torch
    rng = np.random.default_rng(seed)
    seeds = oracle.block_search_group(tuples, center, 16)
    seeds[..., :2] += rng.integers(-12, 13, seeds[..., :2].shape) * (rng.random(seeds.shape[:3]) < 0.5)[..., None]
    seeds[..., 2] = 0
    seeds[center] = 0
    ts = torch.from_numpy(np.ascontiguousarray(seeds, np.int32)).cuda()
    first = torch.zeros(dev.shape, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    assert gpu.lib.vp8g_block_search_dev(gpu.h, dev.descs, dev.n, center, 2, ts.data_ptr(), first.data_ptr(),
                                         None) == 0
    gpu.synchronize()
    return first


def _call(gpu, fn, dev, center, first, radius, flags, fill):
    import torch
    out = torch.full(dev.shape, fill, dtype=torch.int32, device="cuda")
    q = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    st = fn(gpu.h, dev.descs, dev.n, center, first.data_ptr(), radius, flags, out.data_ptr(), q.data_ptr(), None)
    assert st == 0, st
    gpu.synchronize()
    return out.cpu().numpy(), int(q.item())


def test_noalloc_matches_the_allocating_entry_on_two_calls(gpu, oracle):
    import vp8g
    lib = gpu.lib
    cases = [((352, 288, 5, 2), 16, vp8g.FALLBACK_GATE_NEIGHBOUR), ((100, 70, 3, 0), 64, 0)]
    for k, ((w, h, n, center), radius, flags) in enumerate(cases):
        tuples = _group(w, h, n, 1000 + k)
        dev = _Dev(tuples)
        first = _first_stage(gpu, oracle, tuples, dev, center, 2000 + k)
        want, wq = _call(gpu, lib.vp8g_search_fallback_dev, dev, center, first, radius, flags, -7)
        got, q = _call(gpu, lib.vp8g_search_fallback_noalloc_dev, dev, center, first, radius, flags, -7)
        assert np.array_equal(got, want), f"call {k}: tables differ"
        assert q == wq and q > 0, (q, wq)
        # and the oracle agrees with both
        ref, rq = oracle.search_fallback(tuples, center, first.cpu().numpy(), radius, bool(flags))
        others = [i for i in range(n) if i != center]
        assert np.array_equal(got[others], ref[others]) and rq == q


def test_noalloc_rejects_groups_beyond_its_scratch(gpu):
    import torch
    import vp8g
    # 64 frames of 2080x2080 (130 x 130 blocks): 63 * 16900 entries > 2^20.
    w = h = 2080
    assert 63 * ((w + 15) // 16) * ((h + 15) // 16) > SCRATCH_ITEMS
    y = torch.zeros(w * h, dtype=torch.uint8, device="cuda")
    uv = torch.zeros(((w + 1) // 2) * ((h + 1) // 2), dtype=torch.uint8, device="cuda")
    descs = (vp8g._I420 * 64)()
    for i in range(64):
        descs[i] = vp8g._I420(w, h, y.data_ptr(), uv.data_ptr(), uv.data_ptr(), w, (w + 1) // 2)
    first = torch.zeros((64, 130, 130, 3), dtype=torch.int32, device="cuda")
    out = torch.full_like(first, -7)
    torch.cuda.synchronize()
    st = gpu.lib.vp8g_search_fallback_noalloc_dev(gpu.h, descs, 64, 0, first.data_ptr(), 16, 0, out.data_ptr(),
                                                  None, None)
    gpu.synchronize()
    assert st == vp8g.VP8G_E_ARG
    assert bool((out == -7).all())
    # within the limit (17 frames: 16 * 16900 entries) it runs
    st = gpu.lib.vp8g_search_fallback_noalloc_dev(gpu.h, descs, 17, 0, first.data_ptr(), 16, 0, out.data_ptr(),
                                                  None, None)
    gpu.synchronize()
    assert st == 0
    assert bool((out[1:17] == 0).all())  # every entry (sad 0) passes the cut-off: copied
    assert bool((out[0] == -7).all()) and bool((out[17:] == -7).all())
