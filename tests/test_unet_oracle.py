"""Pins the CPU oracle of the U-Net (oracle/unet_oracle.py): the convolution
against scipy's correlation, the adjoint pairs the backward pass is built
from, and every weight gradient against central finite differences (FP64)."""
import numpy as np
import pytest
from scipy.signal import correlate

from oracle import unet_oracle as U


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.0e-3, 65504.0], dtype=np.float32)
    r = U.bf16_round(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to even (1.0)
    assert r[0] == 1.0 and r[1] == 1.0
    assert r[2] == np.float32(1.0078125)
    assert abs(r[3] - x[3]) <= abs(x[3]) * 2 ** -8
    assert r[4] == 65536.0


def test_conv3x3_is_scipy_correlation():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(9, 11, 5))
    w = rng.normal(size=(9, 4, 5))
    b = rng.normal(size=4)
    out = U.conv3x3(x, w, b)
    for co in range(4):
        ref = np.full((9, 11), b[co])
        for ci in range(5):
            k = w[:, co, ci].reshape(3, 3)  # k[dy+1, dx+1]
            ref += correlate(x[:, :, ci], k, mode="same")
        np.testing.assert_allclose(out[:, :, co], ref, rtol=1e-12, atol=1e-12)


def test_adjoint_pairs():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(8, 12, 6))
    w = rng.normal(size=(9, 7, 6))
    y = rng.normal(size=(8, 12, 7))
    assert np.isclose(np.vdot(U.conv3x3(x, w), y), np.vdot(x, U.conv3x3_dx(y, w)), rtol=1e-12)
    d = rng.normal(size=(8, 12, 7))
    yy = rng.normal(size=(4, 6, 7))
    assert np.isclose(np.vdot(U.pool2(d), yy), np.vdot(d, 0.25 * U.up2(yy)), rtol=1e-12)
    assert np.isclose(np.vdot(U.up2(yy), d), np.vdot(yy, U.up2_bwd(d)), rtol=1e-12)
    # dW is the adjoint in w: <conv(x, w), y> is linear in w with gradient conv3x3_dw(y, x)
    g = U.conv3x3_dw(y, x)
    assert np.isclose(np.vdot(U.conv3x3(x, w), y), np.vdot(w, g), rtol=1e-12)


def test_weight_count():
    assert U.num_weights() == sum(9 * a * b + b for a, b in U.CFG) + 33
    assert len(U.init_weights(0)) == U.num_weights()


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(2)
    H, W = 8, 8
    wv = U.init_weights(3).astype(np.float64)
    # non-zero biases so that ReLUs are exercised on both sides
    x3 = rng.normal(size=(3, H, W))
    target = rng.normal(size=(H, W))
    mask = (rng.random((H, W)) > 0.2).astype(np.float64)
    _, loss, g = U.forward_backward(wv, x3, target, mask, bf16=False)
    # one index inside every layer's weights and biases, plus the head
    offsets, o = [], 0
    for ci, co in U.CFG:
        offsets.append(o + rng.integers(0, 9 * ci * co))
        offsets.append(o + 9 * ci * co + rng.integers(0, co))
        o += 9 * ci * co + co
    offsets += [o + 5, o + 32]
    eps = 1e-6
    for i in offsets:
        wp, wm = wv.copy(), wv.copy()
        wp[i] += eps
        wm[i] -= eps
        lp = U.forward_backward(wp, x3, target, mask, bf16=False)[1]
        lm = U.forward_backward(wm, x3, target, mask, bf16=False)[1]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(g[i])) + 1e-7, (i, fd, g[i])


def test_bf16_mode_tracks_fp64():
    rng = np.random.default_rng(4)
    H, W = 16, 8
    wv = U.init_weights(5)
    x3 = rng.normal(size=(3, H, W)).astype(np.float32)
    target = rng.normal(size=(H, W)).astype(np.float32)
    mask = np.ones((H, W), dtype=np.float32)
    o32, l32, g32 = U.forward_backward(wv, x3, target, mask, bf16=True)
    o64, l64, g64 = U.forward_backward(wv.astype(np.float64), x3.astype(np.float64), target, mask, bf16=False)
    assert np.linalg.norm(o32 - o64) <= 5e-2 * np.linalg.norm(o64)
    assert np.linalg.norm(g32 - g64) <= 1e-1 * np.linalg.norm(g64)
