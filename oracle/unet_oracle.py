"""CPU oracle of the warm-start U-Net (test infrastructure only: the product
never imports it). Forward and backward of the network that
blaze_amd/csrc/k_unet.hip runs on the GPU, in float64, optionally rounding
every stored activation and gradient to BF16 (round to nearest even) at the
same points as the GPU kernels (quant=True), so the two agree up to the
FP32 accumulation order of the MFMA units.

There is no reference counterpart (the reference has no learned
component); the oracle is pinned by its own tests: the convolution against
a brute-force loop, the backward pass against central finite differences
(quant=False), pooling/upsampling adjointness.

Network (channels-last, params vector layout identical to the library):
  x0(16: res, mask, 0..) -e1a-> a1(16) -e1b-> s1(16) -pool-> p1 -e2a-> a2(32)
  -e2b-> s2(32) -pool-> p2 -e3a-> a3(64) -e3b-> b3(64) -up-> [u2, s2](96)
  -d2a-> a4(32) -d2b-> b4(32) -up-> [u1, s1](48) -d1a-> a5(16) -d1b-> b5(16)
  -1x1-> out; leaky ReLU (slope LEAK below 0) after every 3x3x3
  convolution; no biases.
"""
import numpy as np

CONV = [(16, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (96, 32), (32, 32), (48, 16), (16, 16)]
NAMES = ["e1a", "e1b", "e2a", "e2b", "e3a", "e3b", "d2a", "d2b", "d1a", "d1b"]
HEAD = 16
LEAK = 0.1


def offsets():
    off, o = [], 0
    for cin, cout in CONV:
        off.append(o)
        o += 27 * cout * cin
    off.append(o)
    return off, o + HEAD


def bf16(x):
    """Round to BF16 (nearest even), returned as float64."""
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)
    nan = np.isnan(a)
    if nan.any():
        out = out.copy()
        out[nan] = np.nan
    return out.astype(np.float64)


def unpack(params):
    off, n = offsets()
    p = np.asarray(params, np.float64)
    assert p.size == n
    ws = [p[off[l]:off[l + 1]].reshape(27, CONV[l][1], CONV[l][0]) for l in range(len(CONV))]
    return ws, p[off[-1]:off[-1] + HEAD]


def _shift(x, dz, dy, dx):
    """y[b, z, y, x] = x[b, z+dz, y+dy, x+dx] with zeros outside."""
    B, D, H, W, C = x.shape
    out = np.zeros_like(x)
    zs, ze = max(0, -dz), min(D, D - dz)
    ys, ye = max(0, -dy), min(H, H - dy)
    xs, xe = max(0, -dx), min(W, W - dx)
    out[:, zs:ze, ys:ye, xs:xe] = x[:, zs + dz:ze + dz, ys + dy:ye + dy, xs + dx:xe + dx]
    return out


def conv3(x, w):
    """3x3x3 zero-padded convolution: y[v, co] = sum_tap sum_ci x[v+tap, ci] w[tap, co, ci]."""
    B, D, H, W, _ = x.shape
    y = np.zeros((B, D, H, W, w.shape[1]))
    for tap in range(27):
        dz, dy, dx = tap // 9 - 1, (tap // 3) % 3 - 1, tap % 3 - 1
        y += _shift(x, dz, dy, dx) @ w[tap].T
    return y


def conv3_dw(x, gy):
    """dW[tap, co, ci] = sum_v gy[v, co] x[v+tap, ci]."""
    C_in, C_out = x.shape[-1], gy.shape[-1]
    g2 = gy.reshape(-1, C_out)
    dw = np.zeros((27, C_out, C_in))
    for tap in range(27):
        dz, dy, dx = tap // 9 - 1, (tap // 3) % 3 - 1, tap % 3 - 1
        dw[tap] = g2.T @ _shift(x, dz, dy, dx).reshape(-1, C_in)
    return dw


def conv3_dx(gy, w):
    """Input gradient: dx[v, ci] = sum_tap sum_co gy[v - tap, co] w[tap, co, ci]."""
    B, D, H, W, _ = gy.shape
    dx_ = np.zeros((B, D, H, W, w.shape[2]))
    for tap in range(27):
        dz, dy, dx = tap // 9 - 1, (tap // 3) % 3 - 1, tap % 3 - 1
        dx_ += _shift(gy, -dz, -dy, -dx) @ w[tap]
    return dx_


def pool(x):
    B, D, H, W, C = x.shape
    return x.reshape(B, D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4, 6))


def pool_bwd(g, shape):
    return np.repeat(np.repeat(np.repeat(g, 2, 1), 2, 2), 2, 3) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 1), 2, 2), 2, 3)


def up_bwd(g):
    B, D, H, W, C = g.shape
    return g.reshape(B, D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4, 6))


def forward(params, x_in, quant=True):
    """x_in: (B, 2, D, H, W). Returns (out (B, D, H, W), cache)."""
    q = bf16 if quant else (lambda a: np.asarray(a, np.float64))
    ws, head = unpack(params)
    wq = [q(w) for w in ws]
    B, _, D, H, W = x_in.shape
    x0 = np.zeros((B, D, H, W, 16))
    x0[..., 0] = x_in[:, 0]
    x0[..., 1] = x_in[:, 1]
    x0 = q(x0)
    relu = lambda a: np.where(a > 0, a, LEAK * a)
    c = {"x0": x0, "w": wq}
    c["a1"] = q(relu(conv3(x0, wq[0])))
    c["s1"] = q(relu(conv3(c["a1"], wq[1])))
    c["p1"] = q(pool(c["s1"]))
    c["a2"] = q(relu(conv3(c["p1"], wq[2])))
    c["s2"] = q(relu(conv3(c["a2"], wq[3])))
    c["p2"] = q(pool(c["s2"]))
    c["a3"] = q(relu(conv3(c["p2"], wq[4])))
    c["b3"] = q(relu(conv3(c["a3"], wq[5])))
    c["cat2"] = np.concatenate([up(c["b3"]), c["s2"]], axis=-1)
    c["a4"] = q(relu(conv3(c["cat2"], wq[6])))
    c["b4"] = q(relu(conv3(c["a4"], wq[7])))
    c["cat1"] = np.concatenate([up(c["b4"]), c["s1"]], axis=-1)
    c["a5"] = q(relu(conv3(c["cat1"], wq[8])))
    c["b5"] = q(relu(conv3(c["a5"], wq[9])))
    hq = q(head)
    c["head"] = hq
    out = c["b5"] @ hq
    if quant:
        out = out.astype(np.float32).astype(np.float64)
    return out, c


def backward(params, x_in, target, quant=True):
    """Mean-squared-error loss and its gradient with respect to the params
    vector (same layout as the library's)."""
    q = bf16 if quant else (lambda a: np.asarray(a, np.float64))
    out, c = forward(params, x_in, quant)
    n = out.size
    d = out - np.asarray(target, np.float64)
    loss = float(np.sum(d * d) / n)
    gout = 2.0 * d / n
    if quant:
        gout = gout.astype(np.float32).astype(np.float64)
    wq = c["w"]
    grads = [None] * len(CONV)
    mask = lambda g, a: g * np.where(a > 0, 1.0, LEAK)
    # head
    g_head = np.einsum("bzyx,bzyxc->c", gout, c["b5"])
    g_b5 = q(mask(gout[..., None] * c["head"], c["b5"]))
    grads[9] = conv3_dw(c["a5"], g_b5)
    g_a5 = q(mask(conv3_dx(g_b5, wq[9]), c["a5"]))
    grads[8] = conv3_dw(c["cat1"], g_a5)
    g_cat1 = q(conv3_dx(g_a5, wq[8]))
    g_b4 = q(mask(up_bwd(g_cat1[..., :32]), c["b4"]))
    grads[7] = conv3_dw(c["a4"], g_b4)
    g_a4 = q(mask(conv3_dx(g_b4, wq[7]), c["a4"]))
    grads[6] = conv3_dw(c["cat2"], g_a4)
    g_cat2 = q(conv3_dx(g_a4, wq[6]))
    g_b3 = q(mask(up_bwd(g_cat2[..., :64]), c["b3"]))
    grads[5] = conv3_dw(c["a3"], g_b3)
    g_a3 = q(mask(conv3_dx(g_b3, wq[5]), c["a3"]))
    grads[4] = conv3_dw(c["p2"], g_a3)
    g_p2 = q(conv3_dx(g_a3, wq[4]))
    g_s2 = q(mask(g_cat2[..., 64:] + g_p2.repeat(2, 1).repeat(2, 2).repeat(2, 3) / 8.0, c["s2"]))
    grads[3] = conv3_dw(c["a2"], g_s2)
    g_a2 = q(mask(conv3_dx(g_s2, wq[3]), c["a2"]))
    grads[2] = conv3_dw(c["p1"], g_a2)
    g_p1 = q(conv3_dx(g_a2, wq[2]))
    g_s1 = q(mask(g_cat1[..., 32:] + g_p1.repeat(2, 1).repeat(2, 2).repeat(2, 3) / 8.0, c["s1"]))
    grads[1] = conv3_dw(c["a1"], g_s1)
    g_a1 = q(mask(conv3_dx(g_s1, wq[1]), c["a1"]))
    grads[0] = conv3_dw(c["x0"], g_a1)
    off, ntot = offsets()
    g = np.zeros(ntot)
    for l in range(len(CONV)):
        g[off[l]:off[l + 1]] = grads[l].ravel()
    g[off[-1]:] = g_head
    return loss, g, out
