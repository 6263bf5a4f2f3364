"""GPU checks of the learned proposer kernels (no reference-crate
counterpart): the synthetic-frame generator against the host generator,
input preparation, the MFMA forward pass against the bf16-rounded
reference model, gradients against the reference model (bf16 and fp64),
training, checkpoints, and the propose -> +-2 refine pipeline."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def ref():
    import proposer_ref
    return proposer_ref


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def bits_to_f64(t):
    u = t.cpu().numpy().astype(np.uint16).astype(np.uint32) << 16
    return u.view(np.float32).astype(np.float64)


@pytest.fixture(scope="module")
def prop(gpu):
    from vp8g.proposer import Proposer
    p = Proposer(gpu, seed=11)
    yield p
    p.close()


def test_gpu_synth_matches_host_generator(gpu):
    import torch
    import vp8g.synth as synth
    from vp8g.proposer import synth_group_luma
    w, h, n = 200, 120, 3
    g = synth.arnr_group_regions(w, h, n, vmax=7, occluders=4, clip=5, config=7)
    out = torch.empty((n, h, w), dtype=torch.uint8, device="cuda")
    clean = torch.empty((n, h, w), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    synth_group_luma(gpu, out.data_ptr(), clean.data_ptr(), w, h, n, 7, 4, 5, 7)
    got, gclean = out.cpu().numpy(), clean.cpu().numpy()
    base = synth._seed(7, 5, 0)
    y0 = synth.blurred_noise(base + 1, h, w)
    vel = synth.uniform_int(base + 4, 128, -7, 7).reshape(64, 2)
    for t in range(n):
        c = synth._occlude(synth._region_warp(y0, t, vel, 8, 1), base + 500 + 7 * t, 4, 1)
        assert np.array_equal(gclean[t], c), t
        d = np.abs(got[t].astype(int) - g.y[t].astype(int))
        assert d.max() <= 1 and (d > 0).mean() < 1e-3, (t, d.max(), (d > 0).mean())


def test_prepare_matches_downsampling(gpu, prop):
    import torch
    import vp8g.synth as synth
    from vp8g.proposer import luma_descs
    w, h, n = 70, 45, 3   # ragged: clamped edge blocks
    g = synth.arnr_group_cif(w, h, n, clip=2)
    frames = torch.from_numpy(np.stack(g.y)).cuda()
    descs, dummy = luma_descs(torch, frames, w, h)
    bw, bh = (w + 15) // 16, (h + 15) // 16
    out = torch.zeros((n - 1, 2, 4 * bh, 4 * bw), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    prop.prepare(descs, n, 1, out.data_ptr())
    gpu.synchronize()
    got = out.cpu().numpy()
    for k, f in enumerate([0, 2]):
        for fi, plane in ((0, g.y[1]), (1, g.y[f])):
            yy = np.minimum(np.arange(bh * 16), h - 1)
            xx = np.minimum(np.arange(bw * 16), w - 1)
            p = plane[yy][:, xx].astype(np.int64)
            s = p.reshape(bh * 4, 4, bw * 4, 4).sum((1, 3))
            assert np.array_equal(got[k, fi], (s + 8) >> 4), (k, fi)


def rand_planes(torch, rng, images, bh, bw):
    """Correlated pairs: the neighbour is the anchor shifted plus noise."""
    a = rng.integers(0, 256, (images, 4 * bh + 8, 4 * bw + 8))
    pl = np.empty((images, 2, 4 * bh, 4 * bw), np.uint8)
    for i in range(images):
        pl[i, 0] = a[i, 4:4 + 4 * bh, 4:4 + 4 * bw]
        dy, dx = rng.integers(-3, 4, 2)
        pl[i, 1] = np.clip(a[i, 4 + dy:4 + dy + 4 * bh, 4 + dx:4 + dx + 4 * bw] + rng.integers(-5, 6, (4 * bh, 4 * bw)), 0, 255)
    return pl, torch.from_numpy(pl).cuda()


@pytest.mark.parametrize("shape", [(2, 5, 7), (1, 68, 120), (3, 9, 130)])
def test_forward_matches_bf16_reference(gpu, prop, shape):
    import torch
    images, bh, bw = shape
    rng = np.random.default_rng(images * 100 + bw)
    params = ref().he_init(5 + bw).astype(np.float32)
    params[4096:4128] = rng.normal(0, 0.05, 32)
    prop.set_params(params)
    pl, plt = rand_planes(torch, rng, images, bh, bw)
    a4 = torch.zeros((images, bh * bw, 32), dtype=torch.int16, device="cuda")
    torch.cuda.synchronize()
    prop.forward(plt.data_ptr(), images, bw, bh, a4.data_ptr())
    gpu.synchronize()
    got = bits_to_f64(a4).reshape(images, bh, bw, 32)
    _, acts = ref().forward(params.astype(np.float64), pl, bf16=True)
    want = acts[-1]
    err = np.abs(got - want)
    tol = 2.0 ** -7 * np.abs(want) + 1e-3
    assert rel_l2(got, want) < 1e-2
    assert (err > tol).mean() < 2e-3, (err > tol).mean()


@pytest.mark.parametrize("w,h,n", [(200, 120, 5), (600, 344, 4), (1040, 280, 3), (40, 24, 9), (16, 16, 2)])
def test_propose_fused_first_layer_matches_forward(gpu, prop, w, h, n):
    """propose applies the first layer inside the features kernel and runs
    the four 3x3 layers and the head as one launch per 16 x 16-cell tile
    (recomputing a 4-cell halo, zeros outside the image); its proposals
    must equal the per-layer forward pass followed by the head, computed
    here on the host in fp32 (a value at a rounding tie may land one unit
    off).  Shapes: one partial tile per image; 3 x 2 tiles with partial
    last ones; a single tile row with 5 columns; 3 x 2 cells in 8 images;
    a single cell."""
    import torch
    import vp8g.synth as synth
    from test_gpu_search import Dev
    rng = np.random.default_rng(3)
    params = ref().he_init(17).astype(np.float32)
    params[4096:4128] = rng.normal(0, 0.05, 32)
    prop.set_params(params)
    g = synth.arnr_group_regions(w, h, n, vmax=8, occluders=4, clip=2, config=3)
    dev = Dev(g.frames())
    n, c = len(dev.tuples), g.center
    bh, bw = dev.shape[1], dev.shape[2]
    table = dev.table(-5)
    planes = torch.zeros((n - 1, 2, 4 * bh, 4 * bw), dtype=torch.uint8, device="cuda")
    a4 = torch.zeros((n - 1, bh * bw, 32), dtype=torch.int16, device="cuda")
    torch.cuda.synchronize()
    prop.propose(dev.descs, n, c, table.data_ptr(), 14)
    prop.prepare(dev.descs, n, c, planes.data_ptr())
    prop.forward(planes.data_ptr(), n - 1, bw, bh, a4.data_ptr())
    gpu.synchronize()
    x = bits_to_f64(a4).astype(np.float32).reshape(n - 1, bh, bw, 32)
    head = params[-66:].astype(np.float32)  # [2][32] weights then [2] bias
    pred = np.empty((n - 1, bh, bw, 2), np.float32)
    for o in range(2):
        s = np.full((n - 1, bh, bw), head[64 + o], np.float32)
        for k in range(32):
            s = (s.astype(np.float64) + head[32 * o + k].astype(np.float64) * x[..., k]).astype(np.float32)
        pred[..., o] = s
    want = np.clip(np.rint(pred * 8), -14, 14).astype(np.int32)
    assert want.size < 64 or (np.abs(want) < 14).mean() > 0.2  # most vectors are not clamped
    got = table.cpu().numpy()[[i for i in range(n) if i != c]]
    assert (got[..., 2] == 0).all()
    diff = np.abs(got[..., :2] - want)
    assert diff.max() <= 1
    assert (diff > 0).mean() < 1e-3, (diff > 0).mean()


def test_gradients_match_reference(gpu, prop):
    import torch
    rng = np.random.default_rng(4)
    images, bh, bw = 2, 6, 9
    params = ref().he_init(21).astype(np.float32)
    prop.set_params(params)
    pl, plt = rand_planes(torch, rng, images, bh, bw)
    labels = rng.integers(-16, 17, (images, bh, bw, 2)).astype(np.int32)
    labels[0, 1, :4, 0] = -32768  # left out of the loss
    lt = torch.from_numpy(labels.reshape(images, bh * bw, 2)).cuda()
    torch.cuda.synchronize()
    prop.backward(plt.data_ptr(), lt.data_ptr(), images, bw, bh)
    g, loss = prop.grads()
    p64 = params.astype(np.float64)
    loss_b, want_b = ref().loss_and_grads(p64, pl, labels, bf16=True)
    loss_f, want_f = ref().loss_and_grads(p64, pl, labels, bf16=False)
    assert abs(loss - loss_b) < 1e-3 * abs(loss_b) + 1e-6
    o = 0
    for t, ci, co in ref().LAYERS + ((1, 32, 2),):
        n = t * ci * co + co if co != 2 else 66
        for name, want, tol in (("bf16", want_b, 3e-2), ("fp64", want_f, 8e-2)):
            e = rel_l2(g[o:o + n], want[o:o + n])
            assert e < tol, (name, ci, co, e)
        o += n


def test_training_reduces_the_loss_and_checkpoints_reload(gpu, tmp_path):
    import torch
    from vp8g.proposer import Proposer, build_dataset, train
    p = Proposer(gpu, seed=3)
    data = build_dataset(gpu, p, groups=2, width=320, height=192, n=5, vmax=4, occluders=2, config=40)
    losses = train(p, data, steps=200, batch=8, lr=3e-3, seed=1)  # one batch: fits it
    assert np.isfinite(losses).all()
    assert losses[-20:].mean() < 0.8 * losses[:20].mean(), (losses[:20].mean(), losses[-20:].mean())
    path = tmp_path / "proposer.safetensors"
    p.save(path, {"steps": 150})
    q = Proposer(gpu, params=Proposer.load_params(path))
    assert np.array_equal(q.params(), p.params())
    frames = torch.from_numpy(np.random.default_rng(0).integers(0, 256, (5, 192, 320), dtype=np.uint8)).cuda()
    from vp8g.proposer import luma_descs
    descs, dummy = luma_descs(torch, frames, 320, 192)
    t1 = torch.zeros((5, 12, 20, 3), dtype=torch.int32, device="cuda")
    t2 = torch.zeros_like(t1)
    torch.cuda.synchronize()
    p.propose(descs, 5, 2, t1.data_ptr())
    q.propose(descs, 5, 2, t2.data_ptr())
    gpu.synchronize()
    assert torch.equal(t1, t2)
    assert int(t1[..., :2].abs().max()) <= 14
    q.close()
    p.close()


def test_propose_then_refine_never_worse_than_the_seed(gpu, oracle):
    import torch
    import vp8g.synth as synth
    from vp8g.proposer import Proposer
    g = synth.arnr_group_regions(256, 160, 5, vmax=6, occluders=3, clip=8, config=41)
    from test_gpu_search import Dev
    dev = Dev(g.frames())
    p = Proposer(gpu, seed=7)
    seeds = dev.table(0)
    refined = dev.table(0)
    torch.cuda.synchronize()
    p.propose(dev.descs, g.n, g.center, seeds.data_ptr())
    gpu.block_search_dev(dev.descs, g.n, g.center, 2, refined.data_ptr(), seeds.data_ptr())
    gpu.synchronize()
    s, r = seeds.cpu().numpy(), refined.cpu().numpy()
    want = oracle.block_search_group(g.frames(), g.center, 2, s)
    others = [i for i in range(g.n) if i != g.center]
    assert np.array_equal(r[others], want[others])
    p.close()
