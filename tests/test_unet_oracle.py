"""Pins of the U-Net CPU oracle (oracle/unet_oracle.py) that need no GPU:
the 3x3x3 convolution against a brute-force loop, the input- and
weight-gradient convolutions and pooling/upsampling against their adjoint
identities, the whole backward pass against central finite differences
(float64, no BF16 rounding), and the BF16 rounding against torch's."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), ".."))
U = pytest.importorskip("oracle.unet_oracle")


def brute_conv(x, w):
    B, D, H, W, Ci = x.shape
    Co = w.shape[1]
    y = np.zeros((B, D, H, W, Co))
    for b in range(B):
        for z in range(D):
            for yy in range(H):
                for xx in range(W):
                    for tap in range(27):
                        dz, dy, dx = tap // 9 - 1, (tap // 3) % 3 - 1, tap % 3 - 1
                        zz, y2, x2 = z + dz, yy + dy, xx + dx
                        if 0 <= zz < D and 0 <= y2 < H and 0 <= x2 < W:
                            y[b, z, yy, xx] += w[tap] @ x[b, zz, y2, x2]
    return y


def test_conv_matches_brute_force():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 3, 4, 5, 2))
    w = rng.standard_normal((27, 3, 2))
    assert np.allclose(U.conv3(x, w), brute_conv(x, w), atol=1e-12)


def test_conv_gradients_are_adjoints():
    """<conv(x, w), g> = <x, conv_dx(g, w)> = <w, conv_dw(x, g)>."""
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 4, 6, 5, 3))
    w = rng.standard_normal((27, 4, 3))
    g = rng.standard_normal((2, 4, 6, 5, 4))
    lhs = np.sum(U.conv3(x, w) * g)
    assert np.isclose(lhs, np.sum(x * U.conv3_dx(g, w)), rtol=1e-12)
    assert np.isclose(lhs, np.sum(w * U.conv3_dw(x, g)), rtol=1e-12)


def test_pool_and_upsample_adjoints():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((1, 4, 6, 8, 3))
    y = rng.standard_normal((1, 2, 3, 4, 3))
    assert np.isclose(np.sum(U.pool(x) * y), np.sum(x * U.pool_bwd(y, x.shape)), rtol=1e-12)
    assert np.isclose(np.sum(U.up(y) * x), np.sum(y * U.up_bwd(x)), rtol=1e-12)
    # pooling a constant keeps it, upsampling then pooling is the identity
    assert np.allclose(U.pool(U.up(y)), y)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    _, n = U.offsets()
    p = rng.standard_normal(n) * 0.3
    x = rng.standard_normal((1, 2, 4, 4, 4))
    t = rng.standard_normal((1, 4, 4, 4))
    loss, g, _ = U.backward(p, x, t, quant=False)
    off, _ = U.offsets()
    # coordinates in every layer (input channels 0/1 of the first layer) and the head
    idx = [off[0] + 0, off[0] + 27 * 16 * 16 - 15, off[3] + 123, off[5] + 4567, off[6] + 999, off[9] + 77,
           off[-1] + 3]
    for i in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp = U.backward(pp, x, t, quant=False)[0]
        lm = U.backward(pm, x, t, quant=False)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(4)
    a = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10.0 ** rng.integers(-30, 30, 10000),
                        np.array([0.0, -0.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, np.inf, -np.inf], np.float32)])
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(U.bf16(a).astype(np.float32).view(np.uint32), ref.view(np.uint32))


def test_homogeneity_of_the_network():
    """Bias-free ReLU network: f(a x) = a f(x) for a > 0 (the warm start
    normalises its input by the residual's maximum and rescales)."""
    rng = np.random.default_rng(5)
    _, n = U.offsets()
    p = rng.standard_normal(n) * 0.2
    x = rng.standard_normal((1, 2, 4, 8, 4))
    x[:, 1] = (rng.random((1, 4, 8, 4)) > 0.5)
    y1, _ = U.forward(p, x, quant=False)
    xs = x.copy()
    xs[:, 0] *= 3.0
    xs[:, 1] *= 3.0
    y3, _ = U.forward(p, xs, quant=False)
    assert np.allclose(y3, 3.0 * y1, rtol=1e-10, atol=1e-12)
