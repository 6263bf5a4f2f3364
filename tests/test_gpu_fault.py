"""Host-side fault injection (VP8G_OPT_FAULT, $VP8G_FAULT): an entry point
whose allocation, copy or launch step fails reports VP8G_E_DEVICE, writes
none of the caller's outputs, and leaves the context usable.  The failure
is forced on the host, so no faulty work reaches the device."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SENTINEL = 0xA5


def _frames(vp8g, g):
    descs = (vp8g._I420 * g.n)()
    keep = []
    for i, (w, h, y, u, v, ys, cs) in enumerate(g.frames()):
        planes = [np.ascontiguousarray(p, np.uint8).reshape(-1) for p in (y, u, v)]
        keep.append(planes)
        descs[i] = vp8g._I420(w, h, *(p.ctypes.data for p in planes), ys, cs)
    return descs, keep


@pytest.mark.parametrize("mode", ["alloc", "copy", "launch"])
def test_host_arnr_fault_leaves_outputs_untouched(gpu, oracle, mode):
    import vp8g
    import vp8g.synth as synth
    g = synth.arnr_group_cif(96, 64, 5)
    descs, keep = _frames(vp8g, g)
    w, h = g.width, g.height
    cw, ch = (w + 1) // 2, (h + 1) // 2
    outs = [np.full(w * h, SENTINEL, np.uint8), np.full(cw * ch, SENTINEL, np.uint8),
            np.full(cw * ch, SENTINEL, np.uint8)]
    mode_id = {"alloc": vp8g.FAULT_ALLOC, "copy": vp8g.FAULT_COPY, "launch": vp8g.FAULT_LAUNCH}[mode]
    gpu.set_fault(mode_id)
    try:
        st = gpu.lib.vp8g_build_arnr_altref(gpu.h, descs, g.n, g.center, 3, *(o.ctypes.data for o in outs), None)
    finally:
        gpu.set_fault(vp8g.FAULT_NONE)
    assert st == vp8g.VP8G_E_DEVICE
    for o in outs:
        assert (o == SENTINEL).all(), "an output was written by a failed call"
    # The context still works after the injected failure.
    st = gpu.lib.vp8g_build_arnr_altref(gpu.h, descs, g.n, g.center, 3, *(o.ctypes.data for o in outs), None)
    assert st == vp8g.VP8G_OK
    rc, want, _, _ = oracle.build_arnr_altref(g.frames(), g.center, 3)
    assert rc == 0
    for o, e in zip(outs, want):
        assert np.array_equal(o, e)


@pytest.mark.parametrize("mode", ["alloc", "copy", "launch"])
def test_host_mc_and_sad_faults(gpu, mode):
    import vp8g
    import vp8g.synth as synth
    g = synth.arnr_group_cif(176, 144, 3)
    mc = synth.mc_workload(g, config=0)
    n = len(mc.mb_xy)
    rp = vp8g._RefPlanes(mc.ref_y.ctypes.data, mc.ref_u.ctypes.data, mc.ref_v.ctypes.data, mc.mb_cols * 16,
                         mc.mb_cols * 8, mc.mb_cols, mc.mb_rows)
    lr = vp8g._LumaRef(mc.ref_y.ctypes.data, mc.mb_cols * 16, mc.mb_cols * 16, mc.mb_rows * 16)
    filt = np.ascontiguousarray(vp8g.SIXTAP_FILTERS, np.int32)
    pred = np.full((n, 384), SENTINEL, np.uint8)
    sad = np.full(n, 0xA5A5A5A5, np.uint32)
    res = np.full((n, 2), 0xA5A5A5A5, np.uint32)
    centers = np.zeros((n, 2), np.int16)
    mode_id = {"alloc": vp8g.FAULT_ALLOC, "copy": vp8g.FAULT_COPY, "launch": vp8g.FAULT_LAUNCH}[mode]
    gpu.set_fault(mode_id)
    try:
        sts = [
            gpu.lib.vp8g_predict_inter_mb_batch(gpu.h, C.byref(rp), mc.mb_xy.ctypes.data, mc.mvs.ctypes.data, n, 0,
                                                filt.ctypes.data, pred.ctypes.data),
            gpu.lib.vp8g_mb_luma_sad_at_mv_batch(gpu.h, C.byref(lr), mc.mb_xy.ctypes.data, mc.src.ctypes.data,
                                                 mc.mvs.ctypes.data, n, sad.ctypes.data),
            gpu.lib.vp8g_half_pixel_refine_luma_batch(gpu.h, C.byref(lr), mc.mb_xy.ctypes.data, mc.src.ctypes.data,
                                                      centers.ctypes.data, n, res.ctypes.data),
        ]
    finally:
        gpu.set_fault(vp8g.FAULT_NONE)
    assert sts == [vp8g.VP8G_E_DEVICE] * 3
    assert (pred == SENTINEL).all() and (sad == 0xA5A5A5A5).all() and (res == 0xA5A5A5A5).all()


def test_dev_entry_launch_fault_writes_nothing(gpu):
    import torch
    import vp8g
    import vp8g.synth as synth
    g = synth.arnr_group_cif(64, 48, 3)
    dev = [(torch.from_numpy(np.ascontiguousarray(p).reshape(-1)).cuda() for p in t[2:5]) for t in g.frames()]
    dev = [tuple(d) for d in dev]
    descs = (vp8g._I420 * g.n)()
    for i, (y, u, v) in enumerate(dev):
        descs[i] = vp8g._I420(g.width, g.height, y.data_ptr(), u.data_ptr(), v.data_ptr(), g.width,
                              (g.width + 1) // 2)
    cw, ch = (g.width + 1) // 2, (g.height + 1) // 2
    oy = torch.full((g.width * g.height,), SENTINEL, dtype=torch.uint8, device="cuda")
    ou = torch.full((cw * ch,), SENTINEL, dtype=torch.uint8, device="cuda")
    ov = torch.full((cw * ch,), SENTINEL, dtype=torch.uint8, device="cuda")
    table = torch.full((g.n, (g.height + 15) // 16, (g.width + 15) // 16, 3), -7, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    gpu.set_fault(vp8g.FAULT_LAUNCH)
    try:
        st1 = gpu.lib.vp8g_build_arnr_altref_dev(gpu.h, descs, g.n, g.center, 3, oy.data_ptr(), ou.data_ptr(),
                                                 ov.data_ptr(), None, None)
        st2 = gpu.lib.vp8g_block_search_dev(gpu.h, descs, g.n, g.center, 16, None, table.data_ptr(), None)
    finally:
        gpu.set_fault(vp8g.FAULT_NONE)
    gpu.synchronize()
    assert st1 == st2 == vp8g.VP8G_E_DEVICE
    assert bool((oy == SENTINEL).all()) and bool((ou == SENTINEL).all()) and bool((ov == SENTINEL).all())
    assert bool((table == -7).all())


def test_batch_create_faults_and_env_variable(gpu):
    import vp8g
    import vp8g.synth as synth
    from vp8g.autotune import DeviceClip
    import torch
    g = synth.arnr_group_cif(64, 48, 3)
    clip = DeviceClip(torch, g)
    torch.cuda.synchronize()
    for mode in (vp8g.FAULT_ALLOC, vp8g.FAULT_COPY, vp8g.FAULT_LAUNCH):
        gpu.set_fault(mode)
        try:
            with pytest.raises(vp8g.Vp8gError) as ei:
                gpu.create_arnr_batch([clip.job(vp8g)])
        finally:
            gpu.set_fault(vp8g.FAULT_NONE)
        assert ei.value.status == vp8g.VP8G_E_DEVICE
    # $VP8G_FAULT sets a new context's initial mode.
    os.environ["VP8G_FAULT"] = "launch"
    try:
        ctx = vp8g.Context(0)
    finally:
        del os.environ["VP8G_FAULT"]
    try:
        with pytest.raises(vp8g.Vp8gError) as ei:
            ctx.build_arnr_altref([vp8g.I420Frame(*t) for t in g.frames()], g.center, vp8g.ArnrConfig(3))
        assert ei.value.status == vp8g.VP8G_E_DEVICE
        ctx.set_fault(vp8g.FAULT_NONE)
        ctx.build_arnr_altref([vp8g.I420Frame(*t) for t in g.frames()], g.center, vp8g.ArnrConfig(3))
    finally:
        ctx.close()
