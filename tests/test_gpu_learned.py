"""What the trained proposer gains, and the exhaustive fallback on its
output, on a held-out 1080p group (region motion [-12, 12] px/frame, 20
occluders per frame, N = 7; seed family 3, clip 2, which no training run
used).

The same exact +-2 search (vp8g_block_search_dev, radius 2) is seeded
three ways: by the trained weights, by the zero vector (the search without
the network) and by an untrained, He-initialised network.  The exhaustive
+-16 search is the judge: min-SAD agreement on every block and on the
blocks with a real match (exhaustive SAD under the occlusion cut-off), and
the luma PSNR of each table-driven altref against the noise-free centre.
The fallback (vp8g_search_fallback_dev) then re-searches the gated failed
entries of the trained table exhaustively; it matches the oracle and never
loses agreement."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def held_out():
    import vp8g.synth as synth
    return synth.arnr_group_regions(1920, 1080, 7, vmax=12, occluders=20, clip=2, config=3)


def _area(g):
    bw = (g.width + 15) // 16
    bh = (g.height + 15) // 16
    return (np.minimum(16, g.height - 16 * np.arange(bh))[:, None] *
            np.minimum(16, g.width - 16 * np.arange(bw))[None, :])


def _score(g, ex, tab):
    others = [i for i in range(g.n) if i != g.center]
    e, r = ex[others], tab[others]
    matched = e[..., 2] < 13 * _area(g)[None]
    same = r[..., 2] == e[..., 2]
    return float(same.mean()), float(same[matched].mean())


def test_trained_network_gain_over_zero_seeds_and_untrained(gpu, held_out):
    import torch
    from vp8g.proposer import WEIGHTS_DIR, DeviceGroup, Proposer, he_init, psnr
    g = held_out
    wpath = WEIGHTS_DIR / "proposer_1080p.safetensors"
    d = DeviceGroup(torch, g)
    n, c = g.n, g.center
    ex, seeds, ref = d.table(torch), d.table(torch), d.table(torch)
    oy, ou, ov = d.outputs(torch)
    torch.cuda.synchronize()
    gpu.block_search_dev(d.descs, n, c, 16, ex.data_ptr())
    gpu.synchronize()
    ex_np = ex.cpu().numpy()
    clean = g.clean[c][0]

    def altref_psnr(table):
        gpu.build_arnr_altref_table_dev(d.descs, n, c, 3, table.data_ptr(), oy.data_ptr(), ou.data_ptr(),
                                        ov.data_ptr())
        gpu.synchronize()
        return psnr(oy.cpu().numpy().reshape(g.height, g.width), clean)

    res = {}
    for name in ("trained", "zero", "untrained"):
        if name == "zero":
            seeds.zero_()
        else:
            prop = Proposer(gpu, params=Proposer.load_params(wpath) if name == "trained" else he_init(7))
            prop.propose(d.descs, n, c, seeds.data_ptr(), 14)
            gpu.synchronize()
            prop.close()
        gpu.block_search_dev(d.descs, n, c, 2, ref.data_ptr(), seeds.data_ptr())
        gpu.synchronize()
        agree, agree_m = _score(g, ex_np, ref.cpu().numpy())
        res[name] = {"agreement": agree, "agreement_matched": agree_m, "psnr": altref_psnr(ref)}
    res["exhaustive_psnr"] = altref_psnr(ex)
    print(json.dumps(res))
    t, z, u = res["trained"], res["zero"], res["untrained"]
    assert t["agreement_matched"] >= 0.9
    assert t["agreement_matched"] >= z["agreement_matched"] + 0.4
    assert t["agreement_matched"] >= u["agreement_matched"] + 0.4
    assert t["psnr"] >= z["psnr"] + 0.3 and t["psnr"] >= u["psnr"] + 0.3
    assert res["exhaustive_psnr"] - t["psnr"] <= 0.1


def test_fallback_on_the_trained_table_matches_oracle(gpu, oracle, held_out):
    import torch
    from vp8g.proposer import WEIGHTS_DIR, DeviceGroup, Proposer
    g = held_out
    d = DeviceGroup(torch, g)
    n, c = g.n, g.center
    ex, seeds, first, out = d.table(torch), d.table(torch), d.table(torch), d.table(torch)
    q = torch.zeros(1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    prop = Proposer(gpu, params=Proposer.load_params(WEIGHTS_DIR / "proposer_1080p.safetensors"))
    prop.propose(d.descs, n, c, seeds.data_ptr(), 14)
    gpu.block_search_dev(d.descs, n, c, 2, first.data_ptr(), seeds.data_ptr())
    gpu.search_fallback_dev(d.descs, n, c, first.data_ptr(), out.data_ptr(), 16, True, q.data_ptr())
    gpu.block_search_dev(d.descs, n, c, 16, ex.data_ptr())
    gpu.synchronize()
    prop.close()
    f_np, o_np, ex_np = first.cpu().numpy(), out.cpu().numpy(), ex.cpu().numpy()
    want, wq = oracle.search_fallback(g.frames(), c, f_np, 16, True)
    others = [i for i in range(n) if i != c]
    assert np.array_equal(o_np[others], want[others])
    assert int(q.item()) == wq
    a0, m0 = _score(g, ex_np, f_np)
    a1, m1 = _score(g, ex_np, o_np)
    print(json.dumps({"queued": wq, "queued_fraction": wq / ((n - 1) * d.bw * d.bh), "agreement": [a0, a1],
                      "agreement_matched": [m0, m1]}))
    assert m1 >= m0 and a1 >= a0
