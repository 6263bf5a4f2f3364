"""Pins the fp64 reference model of the learned proposer (no reference-crate
counterpart): central finite differences of the loss against its analytic
gradient for every layer, the transposed-conv identity, and the bf16
rounding helper against exact bfloat16 values."""
import numpy as np
import pytest


def ref():
    import proposer_ref
    return proposer_ref


def test_param_count_matches_the_library():
    import vp8g
    assert ref().param_count() == vp8g.load_library().vp8g_proposer_param_count() == 59650


def test_bf16_round_ties_to_even():
    r = ref()
    assert r.bf16_round(1.0) == 1.0
    assert r.bf16_round(1.0 + 2.0 ** -8) == 1.0            # tie -> even mantissa
    assert r.bf16_round(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert r.bf16_round(1.0 + 2.0 ** -7) == 1.0 + 2.0 ** -7  # representable
    assert r.bf16_round(-3.3) == -3.296875


def test_input_grad_is_the_adjoint_of_the_conv():
    r = ref()
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 3, 5, 4))
    w = rng.normal(size=(9, 6, 4))
    g = rng.normal(size=(2, 3, 5, 6))
    # <conv(x), g> == <x, conv^T(g)>
    assert np.isclose((r.conv3x3(x, w) * g).sum(), (x * r.conv3x3_input_grad(g, w)).sum())
    dw = r.conv3x3_weight_grad(g, x)
    eps = 1e-6
    k = (4, 2, 3)
    w2 = w.copy()
    w2[k] += eps
    fd = ((r.conv3x3(x, w2) - r.conv3x3(x, w)) * g).sum() / eps
    assert np.isclose(fd, dw[k], rtol=1e-5, atol=1e-6)


def test_features_are_block_sads_of_the_downsampled_pair():
    r = ref()
    rng = np.random.default_rng(9)
    planes = rng.integers(0, 256, (1, 2, 12, 16)).astype(np.uint8)
    x = r.features(planes)
    A, B = planes[0, 0].astype(int), planes[0, 1].astype(int)
    for (by, bx) in ((0, 0), (1, 2), (2, 3)):
        for (ddx, ddy) in ((0, 0), (-4, 3), (4, -4), (1, 1)):
            s = 0
            for j in range(4):
                for i in range(4):
                    y, xx = 4 * by + j, 4 * bx + i
                    s += abs(A[y, xx] - B[min(max(y + ddy, 0), 11), min(max(xx + ddx, 0), 15)])
            assert x[0, by, bx, 9 * (ddy + 4) + (ddx + 4)] == s / 1024.0
        assert np.allclose(x[0, by, bx, 81:97], A[4 * by:4 * by + 4, 4 * bx:4 * bx + 4].reshape(-1) / 255.0)
        assert np.allclose(x[0, by, bx, 97:113], B[4 * by:4 * by + 4, 4 * bx:4 * bx + 4].reshape(-1) / 255.0)
        assert (x[0, by, bx, 113:] == 0).all()


def test_gradients_match_finite_differences():
    r = ref()
    rng = np.random.default_rng(2)
    params = r.he_init(3)
    params[4096:4128] = rng.normal(0, 0.1, 32)  # non-zero biases
    planes = rng.integers(0, 256, (2, 2, 12, 16)).astype(np.uint8)
    labels = rng.integers(-16, 17, (2, 3, 4, 2))
    loss, grads = r.loss_and_grads(params, planes, labels)
    x0 = planes
    # probe parameters in every tensor
    offsets, o = [], 0
    for t, ci, co in r.LAYERS:
        offsets += [o + 17, o + t * co * ci - 5, o + t * co * ci + 3]
        o += t * co * ci + co
    offsets += [o + 7, o + 40, o + 64, o + 65]
    eps = 1e-6
    for i in offsets:
        p1, p2 = params.copy(), params.copy()
        p1[i] += eps
        p2[i] -= eps
        fd = (r.loss_and_grads(p1, x0, labels)[0] - r.loss_and_grads(p2, x0, labels)[0]) / (2 * eps)
        assert np.isclose(fd, grads[i], rtol=1e-4, atol=1e-9), (i, fd, grads[i])
