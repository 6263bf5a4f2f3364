"""GPU parity for the coarse-to-fine pyramid steps (vp8g_downsample2_i420_dev,
vp8g_upsample_table_dev) against numpy restatements, bit-exact, through
the C ABI.  These entry points have no counterpart in the reference."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def down2(p: np.ndarray) -> np.ndarray:
    """2x2 rounded mean with coordinates clamped to the plane."""
    h, w = p.shape
    oh, ow = (h + 1) // 2, (w + 1) // 2
    y0, x0 = 2 * np.arange(oh), 2 * np.arange(ow)
    y1, x1 = np.minimum(y0 + 1, h - 1), np.minimum(x0 + 1, w - 1)
    q = p.astype(np.int32)
    s = q[y0][:, x0] + q[y0][:, x1] + q[y1][:, x0] + q[y1][:, x1]
    return ((s + 2) >> 2).astype(np.uint8)


def test_down2_restatement_known_answers():
    # the closed form on tiny planes: a 1x1 plane is itself, a 2x2 plane its
    # rounded mean (ties round up), an odd column repeats the edge
    assert down2(np.array([[7]], np.uint8)).tolist() == [[7]]
    assert down2(np.array([[0, 1], [1, 0]], np.uint8)).tolist() == [[1]]  # (2 + 2) >> 2
    assert down2(np.array([[0, 0], [1, 0]], np.uint8)).tolist() == [[0]]  # (1 + 2) >> 2
    assert down2(np.array([[10, 20, 30]], np.uint8)).tolist() == [[15, 30]]


@pytest.mark.parametrize("geom", [(3840, 2160, 0), (37, 29, 3), (1, 1, 0), (2, 5, 2), (64, 48, 16), (33, 2, 1)],
                         ids=lambda g: "x".join(map(str, g)))
def test_downsample2_matches_numpy(gpu, geom):
    import torch
    import vp8g
    w, h, pad = geom
    cw, ch = (w + 1) // 2, (h + 1) // 2
    ow, oh = (w + 1) // 2, (h + 1) // 2
    ocw, och = (ow + 1) // 2, (oh + 1) // 2
    rng = np.random.default_rng(w * 131 + h)
    ys, cs = w + pad, cw + pad
    y = rng.integers(0, 256, (h, ys), dtype=np.uint8)
    u = rng.integers(0, 256, (ch, cs), dtype=np.uint8)
    v = rng.integers(0, 256, (ch, cs), dtype=np.uint8)
    ty, tu, tv = (torch.from_numpy(a.reshape(-1).copy()).cuda() for a in (y, u, v))
    oys, ocs = ow + pad, ocw + pad
    dy = torch.full((oh * oys,), 7, dtype=torch.uint8, device="cuda")
    du = torch.full((och * ocs,), 7, dtype=torch.uint8, device="cuda")
    dv = torch.full((och * ocs,), 7, dtype=torch.uint8, device="cuda")
    src = vp8g._I420(w, h, ty.data_ptr(), tu.data_ptr(), tv.data_ptr(), ys, cs)
    dst = vp8g._I420(ow, oh, dy.data_ptr(), du.data_ptr(), dv.data_ptr(), oys, ocs)
    torch.cuda.synchronize()
    gpu.downsample2_dev(src, dst)
    gpu.synchronize()
    for name, plane, out, pw, ph, ps, opw, oph, ops in (
            ("y", y, dy, w, h, ys, ow, oh, oys), ("u", u, du, cw, ch, cs, ocw, och, ocs),
            ("v", v, dv, cw, ch, cs, ocw, och, ocs)):
        got = out.cpu().numpy().reshape(oph, ops)
        assert np.array_equal(got[:, :opw], down2(plane[:, :pw])), name
        assert (got[:, opw:] == 7).all(), f"{name}: padding written"


def test_downsample2_many_frames(gpu):
    """20 frames in one call (two launches of at most 16), each frame with
    its own planes and strides."""
    import torch
    import vp8g
    w, h, n = 45, 23, 20
    cw, ch, ow, oh = (w + 1) // 2, (h + 1) // 2, (w + 1) // 2, (h + 1) // 2
    ocw, och = (ow + 1) // 2, (oh + 1) // 2
    rng = np.random.default_rng(5)
    src, dst = (vp8g._I420 * n)(), (vp8g._I420 * n)()
    keep, want = [], []
    for i in range(n):
        pad = i % 4
        planes = [rng.integers(0, 256, (h, w + pad), dtype=np.uint8)] + \
                 [rng.integers(0, 256, (ch, cw + pad), dtype=np.uint8) for _ in range(2)]
        t = [torch.from_numpy(a.reshape(-1).copy()).cuda() for a in planes]
        o = [torch.zeros(oh * (ow + pad), dtype=torch.uint8, device="cuda")] + \
            [torch.zeros(och * (ocw + pad), dtype=torch.uint8, device="cuda") for _ in range(2)]
        src[i] = vp8g._I420(w, h, t[0].data_ptr(), t[1].data_ptr(), t[2].data_ptr(), w + pad, cw + pad)
        dst[i] = vp8g._I420(ow, oh, o[0].data_ptr(), o[1].data_ptr(), o[2].data_ptr(), ow + pad, ocw + pad)
        keep.append((t, o, pad))
        want.append([down2(planes[0][:, :w]), down2(planes[1][:, :cw]), down2(planes[2][:, :cw])])
    torch.cuda.synchronize()
    gpu.downsample2_dev(src, dst, n)
    gpu.synchronize()
    for i, ((t, o, pad), wnt) in enumerate(zip(keep, want)):
        got = [o[0].cpu().numpy().reshape(oh, ow + pad)[:, :ow]] + \
              [o[k].cpu().numpy().reshape(och, ocw + pad)[:, :ocw] for k in (1, 2)]
        for k in range(3):
            assert np.array_equal(got[k], wnt[k]), (i, "yuv"[k])


def test_downsample2_rejects_wrong_sizes(gpu):
    import torch
    import vp8g
    t = torch.zeros(4096, dtype=torch.uint8, device="cuda")
    src = vp8g._I420(16, 16, t.data_ptr(), t.data_ptr(), t.data_ptr(), 16, 8)
    bad = vp8g._I420(9, 8, t.data_ptr(), t.data_ptr(), t.data_ptr(), 9, 5)
    with pytest.raises(vp8g.Vp8gError):
        gpu.downsample2_dev(src, bad)


@pytest.mark.parametrize("dims", [(120, 68, 240, 135), (3, 2, 5, 4), (1, 1, 2, 2), (60, 34, 120, 68), (4, 4, 5, 9)],
                         ids=lambda d: "x".join(map(str, d)))
def test_upsample_table_matches_numpy(gpu, dims):
    import torch
    cw, ch, fw, fh = dims
    n = 3
    rng = np.random.default_rng(cw * fw)
    coarse = rng.integers(-300, 301, (n, ch, cw, 3)).astype(np.int32)
    tc = torch.from_numpy(coarse).cuda()
    tf = torch.full((n, fh, fw, 3), 99, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    gpu.upsample_table_dev(tc.data_ptr(), n, cw, ch, tf.data_ptr(), fw, fh)
    gpu.synchronize()
    r = np.minimum(np.arange(fh) // 2, ch - 1)
    c = np.minimum(np.arange(fw) // 2, cw - 1)
    want = np.zeros((n, fh, fw, 3), np.int32)
    want[..., :2] = 2 * coarse[:, r][:, :, c, :2]
    assert np.array_equal(tf.cpu().numpy(), want)


def test_upsample_table_rejects_grids_beyond_1024_blocks(gpu):
    """Grid sides are block counts of frames up to 0x3FFF pixels: 1..1024;
    anything else is refused before a launch (the fine table is untouched)."""
    import torch
    import vp8g
    tc = torch.zeros(3 * 4, dtype=torch.int32, device="cuda")
    tf = torch.full((3 * 4,), 7, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    for cw, ch, fw, fh in ((1025, 1, 2, 2), (1, 1025, 2, 2), (1, 1, 1025, 2), (1, 1, 2, 1025), (1, 1, 0, 2),
                           (0, 1, 2, 2), (1, 1, 16383, 16383)):
        with pytest.raises(vp8g.Vp8gError):
            gpu.upsample_table_dev(tc.data_ptr(), 64, cw, ch, tf.data_ptr(), fw, fh)
    gpu.synchronize()
    assert (tf.cpu().numpy() == 7).all()
