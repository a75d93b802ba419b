"""CPU oracle of the learned free-surface operator (test infrastructure only).

A numpy restatement of the U-Net of si3d_amd/csrc/s3g_unet.hip: forward pass,
masked mean-square loss and the gradient of every weight.  With
``bf16=True`` every tensor the GPU stores in bfloat16 (input image,
activations, weights used by the convolutions, activation gradients) is
rounded to bfloat16 at the same points, so the GPU and the oracle differ only
in FP32 summation order; with ``bf16=False`` everything is FP64 (used to pin
the backward pass against finite differences).

The network has no counterpart in the reference (its free-surface solve is
NSPCG, si3d_procedures.f90:1870-2065); the parity of this file is pinned by
the tests in tests/test_unet_oracle.py: 3x3 convolution against
scipy.signal.correlate, every weight gradient against central finite
differences in FP64, pooling/upsampling adjointness.

Weight vector layout: for each of the 14 convolutions (Cin, Cout) in CFG,
w[9][Cout][Cin] then b[Cout]; then the 1x1 head w[32], b[1].
"""
from __future__ import annotations

import numpy as np

CFG = [(16, 32), (32, 32), (32, 64), (64, 64), (64, 128), (128, 128), (128, 256),
       (256, 256), (384, 128), (128, 128), (192, 64), (64, 64), (96, 32), (32, 32)]
LEVEL = [0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0]


def num_weights():
    return sum(9 * ci * co + co for ci, co in CFG) + 33


def bf16_round(a):
    """Round float32 values to the nearest bfloat16 (ties to even), as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack(wvec):
    layers, o = [], 0
    for ci, co in CFG:
        w = wvec[o:o + 9 * co * ci].reshape(9, co, ci)
        o += 9 * co * ci
        b = wvec[o:o + co]
        o += co
        layers.append((w, b))
    hw, hb = wvec[o:o + 32], wvec[o + 32]
    return layers, hw, hb


def shift(x, dy, dx):
    """y[h, w] = x[h + dy, w + dx] with zeros outside (x: H x W x C)."""
    H, W = x.shape[:2]
    y = np.zeros_like(x)
    hs, he = max(0, -dy), min(H, H - dy)
    ws, we = max(0, -dx), min(W, W - dx)
    y[hs:he, ws:we] = x[hs + dy:he + dy, ws + dx:we + dx]
    return y


def conv3x3(x, w, b=None):
    """out[p, co] = b[co] + sum_{t, ci} x[p + d(t), ci] w[t, co, ci], t = 3(dy+1)+(dx+1)."""
    H, W, _ = x.shape
    out = np.zeros((H, W, w.shape[1]), dtype=x.dtype)
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        out += shift(x, dy, dx) @ w[t].T
    if b is not None:
        out += b
    return out


def conv3x3_dx(dy_, w):
    """Adjoint of conv3x3 in x: dx[q, ci] = sum_{t, co} dy[q - d(t), co] w[t, co, ci]."""
    H, W, _ = dy_.shape
    dx_ = np.zeros((H, W, w.shape[2]), dtype=dy_.dtype)
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        dx_ += shift(dy_, -dy, -dx) @ w[t]
    return dx_


def conv3x3_dw(dy_, x):
    """dW[t, co, ci] = sum_p dy[p, co] x[p + d(t), ci]."""
    out = np.zeros((9, dy_.shape[2], x.shape[2]), dtype=x.dtype)
    C = dy_.shape[2]
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        out[t] = dy_.reshape(-1, C).T @ shift(x, dy, dx).reshape(-1, x.shape[2])
    return out


def pool2(x):
    H, W, C = x.shape
    return 0.25 * x.reshape(H // 2, 2, W // 2, 2, C).sum(axis=(1, 3))


def up2(x):
    return np.repeat(np.repeat(x, 2, axis=0), 2, axis=1)


def up2_bwd(d):
    H, W, C = d.shape
    return d.reshape(H // 2, 2, W // 2, 2, C).sum(axis=(1, 3))


def forward_backward(wvec, x3, target, mask, bf16=True):
    """Returns (out[H, W], loss, grad vector).  x3: (3, H, W)."""
    dt = np.float32 if bf16 else np.float64
    R = bf16_round if bf16 else (lambda a: a)
    wvec = np.asarray(wvec, dtype=dt)
    layers, hw, hb = unpack(wvec)
    wq = [(R(w), b) for w, b in layers]  # weights as the convolutions see them
    _, H, W = x3.shape
    X0 = np.zeros((H, W, 16), dtype=dt)
    X0[:, :, :3] = np.moveaxis(x3, 0, -1)
    X0 = R(X0)

    def cf(i, x):
        return R(np.maximum(conv3x3(x, wq[i][0], wq[i][1]), 0))

    A0 = cf(0, X0)
    S0 = cf(1, A0)
    X1 = R(pool2(S0))
    A1 = cf(2, X1)
    S1 = cf(3, A1)
    X2 = R(pool2(S1))
    A2 = cf(4, X2)
    S2 = cf(5, A2)
    X3 = R(pool2(S2))
    A3 = cf(6, X3)
    B3 = cf(7, A3)
    C2 = np.concatenate([up2(B3), S2], axis=2)
    A4 = cf(8, C2)
    B2 = cf(9, A4)
    C1 = np.concatenate([up2(B2), S1], axis=2)
    A5 = cf(10, C1)
    B1 = cf(11, A5)
    C0 = np.concatenate([up2(B1), S0], axis=2)
    A6 = cf(12, C0)
    B0 = cf(13, A6)
    out = B0 @ hw + hb
    nm = max(1.0, float(mask.sum()))
    e = (out - target) * mask
    loss = float((e.astype(np.float64) ** 2).sum() / nm)
    dout = (2.0 * e / nm).astype(dt)

    grads = [None] * 14
    dhw = (B0 * dout[:, :, None]).sum(axis=(0, 1))
    dhb = dout.sum()
    dB0 = R(dout[:, :, None] * hw * (B0 > 0))

    def cb(i, dy_, x, mask_t):
        w = wq[i][0]
        grads[i] = (conv3x3_dw(dy_, x), dy_.sum(axis=(0, 1)))
        if mask_t is None:
            return conv3x3_dx(dy_, w)
        return R(conv3x3_dx(dy_, w) * (mask_t > 0))

    dA6 = cb(13, dB0, A6, A6)
    dC0 = cb(12, dA6, C0, C0)
    dB1 = R(up2_bwd(dC0[:, :, :64]))
    dA5 = cb(11, dB1, A5, A5)
    dC1 = cb(10, dA5, C1, C1)
    dB2 = R(up2_bwd(dC1[:, :, :128]))
    dA4 = cb(9, dB2, A4, A4)
    dC2 = cb(8, dA4, C2, C2)
    dB3 = R(up2_bwd(dC2[:, :, :256]))
    dA3 = cb(7, dB3, A3, A3)
    dX3 = R(cb(6, dA3, X3, None))
    dS2 = R((dC2[:, :, 256:] + 0.25 * up2(dX3)) * (S2 > 0))
    dA2 = cb(5, dS2, A2, A2)
    dX2 = R(cb(4, dA2, X2, None))
    dS1 = R((dC1[:, :, 128:] + 0.25 * up2(dX2)) * (S1 > 0))
    dA1 = cb(3, dS1, A1, A1)
    dX1 = R(cb(2, dA1, X1, None))
    dS0 = R((dC0[:, :, 64:] + 0.25 * up2(dX1)) * (S0 > 0))
    dA0 = cb(1, dS0, A0, A0)
    grads[0] = (conv3x3_dw(dA0, X0), dA0.sum(axis=(0, 1)))

    g = []
    for gw, gb in grads:
        g.append(gw.ravel())
        g.append(gb.ravel())
    g.append(dhw.ravel())
    g.append(np.array([dhb]))
    return out, loss, np.concatenate(g).astype(np.float64)


def init_weights(seed=0, scale=1.0):
    """He-initialised weights (numpy RNG; the GPU library has its own)."""
    rng = np.random.default_rng(seed)
    parts = []
    for ci, co in CFG:
        fan = 9 * (3 if ci == 16 else ci)
        w = rng.normal(0, np.sqrt(2.0 / fan), size=(9, co, ci)) * scale
        if ci == 16:
            w[:, :, 3:] = 0
        parts += [w.ravel(), rng.normal(0, 0.05, size=co)]
    parts += [rng.normal(0, 0.1 / np.sqrt(32), size=32), np.array([0.01])]
    return np.concatenate(parts).astype(np.float32)
