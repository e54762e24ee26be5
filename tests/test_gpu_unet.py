"""GPU tests of the warm-start U-Net (blaze_amd/csrc/k_unet.hip) against the
CPU oracle (oracle/unet_oracle.py, BF16 storage emulated at the same points):
forward output and every layer's weight gradient, with tolerances stated
below (the GPU accumulates in FP32 on the MFMA units, the oracle in
float64, so a BF16 rounding of an intermediate can differ by one unit)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(__file__), ".."))

pytestmark = pytest.mark.gpu

B = pytest.importorskip("blaze_amd")
U = pytest.importorskip("oracle.unet_oracle")

FWD_RTOL = 2e-2   # relative L2 of the output
GRAD_RTOL = 5e-2  # relative L2 of each layer's weight gradient


def rel_l2(a, b):
    return float(np.linalg.norm((np.asarray(a, np.float64) - b).ravel()) / (np.linalg.norm(np.ravel(b)) + 1e-30))


def sample(seed, shape):
    rng = np.random.default_rng(seed)
    Bn, D, H, W = shape
    x = np.zeros((Bn, 2, D, H, W), np.float32)
    x[:, 0] = rng.standard_normal((Bn, D, H, W))
    x[:, 1] = rng.random((Bn, D, H, W)) > 0.7
    t = rng.standard_normal((Bn, D, H, W)).astype(np.float32) * 0.1
    return x, t


# (1, 32, 32, 32): the network grid of a 256^3 scene, whose levels run the
# convolution with the reduction split over 2 and 4 waves; (4, 32, 32, 32)
# also without the split (the first level's 8192 tiles)
@pytest.mark.parametrize("shape", [(1, 8, 8, 8), (2, 4, 8, 12), (1, 32, 32, 32), (4, 32, 32, 32)])
def test_forward_matches_oracle(shape):
    print(f"[unet] forward {shape}", flush=True)
    net = B.Net(seed=7)
    x, _ = sample(1, shape)
    got = net.forward(x)
    want, _ = U.forward(net.params().astype(np.float64), x, quant=True)
    e = rel_l2(got, want)
    print(f"[unet] forward rel L2 {e:.2e}", flush=True)
    assert e <= FWD_RTOL


def test_gradients_match_oracle():
    print("[unet] gradients", flush=True)
    net = B.Net(seed=11)
    x, t = sample(2, (1, 8, 8, 8))
    loss = net.train_step(x, t, update=False)
    g = net.grads()
    p = net.params().astype(np.float64)
    lo, go, _ = U.backward(p, x, t, quant=True)
    assert abs(loss - lo) <= 1e-2 * abs(lo)
    off, _ = U.offsets()
    for l, name in enumerate(U.NAMES):
        a, b = g[off[l]:off[l + 1]], go[off[l]:off[l + 1]]
        e = rel_l2(a, b)
        print(f"[unet] grad {name} rel L2 {e:.2e}", flush=True)
        assert e <= GRAD_RTOL, name
    e = rel_l2(g[off[-1]:], go[off[-1]:])
    assert e <= GRAD_RTOL, "head"


def test_training_reduces_loss():
    """Adam on one fixed batch: the loss falls by 10x within 300 steps."""
    print("[unet] training", flush=True)
    net = B.Net(seed=3)
    x, t = sample(5, (2, 8, 8, 8))
    # a smooth target the network can represent: a local average of the input
    from scipy.ndimage import uniform_filter

    t = np.stack([uniform_filter(x[b, 0].astype(np.float64), size=3, mode="constant") for b in range(2)])
    t = t.astype(np.float32)
    first = net.train_step(x, t, lr=2e-3)
    last = first
    for _ in range(300):
        last = net.train_step(x, t, lr=2e-3)
    print(f"[unet] loss {first:.3e} -> {last:.3e}", flush=True)
    assert last <= 0.1 * first


def test_params_roundtrip_and_seed():
    a, b = B.Net(seed=1), B.Net(seed=1)
    assert np.array_equal(a.params(), b.params())
    c = B.Net(seed=2)
    assert not np.array_equal(a.params(), c.params())
    p = a.params() * 0.5
    a.set_params(p)
    assert np.array_equal(a.params(), p)
