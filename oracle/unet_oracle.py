"""CPU reference of the flux-form surrogate network (TEST INFRASTRUCTURE ONLY).

Only tests/ and bench.py's baselines may import this module; it checks the GPU
kernels of csrc/amb_unet.hip, it is never the product.  The surrogate has no
counterpart in the reference; this file restates the network the GPU code
implements, in float64 NumPy, with an option to round to bfloat16 at exactly the
points where the GPU stores BF16 tensors (so GPU and CPU gradients agree to the
accumulation order).  Its own backward pass is pinned by finite differences of its
forward in tests/test_unet_oracle.py.

Network (NHWC, periodic padding, 3x3 convolutions unless noted):
  x0  = [phi, lamb/2, zeta/5, phi0/0.6, 256 (phi - bf16(phi)), 0...]   16 channels
  e1  = gelu(conv(x0; W1) + b1)                            32 ch, H x W
  p1  = avgpool2(e1)                                        32 ch, H/2
  e2  = gelu(conv(p1; W2) + b2)                            64 ch, H/2
  p2  = avgpool2(e2)                                        64 ch, H/4
  bt  = gelu(conv(p2; W3) + b3)                            128 ch, H/4
  u2  = conv1x1(bt; W4) + b4                                64 ch, H/4
  s2  = up2(u2) + e2                                        64 ch, H/2
  d2  = gelu(conv(s2; W5) + b5)                            64 ch, H/2
  u1  = conv1x1(d2; W6) + b6                                32 ch, H/2
  s1  = up2(u1) + e1                                        32 ch, H
  d1  = gelu(conv(s1; W7) + b7)                            32 ch, H
  J   = 1e-3 (conv(d1; Wh) + bh)                            2 ch (y, x), FP32
  phi' = phi - div J,  (div J)(i,j) = Jy(i,j) - Jy(i-1,j) + Jx(i,j) - Jx(i,j-1)
Loss (per batch): mean over samples of sum_p (phi' - target)^2 / (sum_p (target - phi)^2
+ LOSS_FLOOR * sites).
Weights W[tap][co][ci] with tap = 3*(dy+1) + (dx+1) reading in[y+dy][x+dx].
"""
from __future__ import annotations

import numpy as np

# Per-site floor of the per-sample loss normaliser |target - phi|^2 + floor * sites.
LOSS_FLOOR = 1e-8

LAYERS = [  # name, cin, cout, taps, gelu
    ("c1", 16, 32, 9, True),
    ("c2", 32, 64, 9, True),
    ("c3", 64, 128, 9, True),
    ("c4", 128, 64, 1, False),
    ("c5", 64, 64, 9, True),
    ("c6", 64, 32, 1, False),
    ("c7", 32, 32, 9, True),
    ("head", 32, 2, 9, False),
]
PARAM_SCALE = (2.0, 5.0, 0.6)  # lamb, zeta, phi0 are divided by these in the input
LO_SCALE = 256.0  # input channel 4: LO_SCALE (phi - bf16(phi))
J_SCALE = 1e-3  # J = J_SCALE (head convolution)
CIN_REAL = 5  # input channels carrying data; the rest of c1's 16 are zero
C0, C1 = 0.7978845608028654, 0.044715


def param_layout():
    """(name, offset of W, offset of b, cin, cout, taps) in the flat FP32 parameter vector."""
    out, off = [], 0
    for name, cin, cout, taps, _ in LAYERS:
        w_off = off
        off += taps * cout * cin
        b_off = off
        off += cout
        out.append((name, w_off, b_off, cin, cout, taps))
    return out, off


def bf16(a):
    """Round float64/float32 values to bfloat16 (round to nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1 + np.tanh(C0 * (x + C1 * x**3)))


def gelu_grad(x):
    t = np.tanh(C0 * (x + C1 * x**3))
    return 0.5 * (1 + t) + 0.5 * x * (1 - t * t) * C0 * (1 + 3 * C1 * x * x)


def shift(x, dy, dx):
    """out[n, y, x] = x[n, y+dy, x+dx] (periodic)."""
    return np.roll(np.roll(x, -dy, axis=1), -dx, axis=2)


def taps_of(taps):
    return [(0, 0)] if taps == 1 else [(t // 3 - 1, t % 3 - 1) for t in range(9)]


def conv(x, w, b):
    """x (N,H,W,ci), w (taps, co, ci), b (co,) -> (N,H,W,co)."""
    out = np.zeros(x.shape[:3] + (w.shape[1],))
    for t, (dy, dx) in enumerate(taps_of(w.shape[0])):
        out += shift(x, dy, dx) @ w[t].T
    return out + b


def conv_grad(x, w, g):
    """Gradients of sum(g * conv(x, w, b)) -> (dx, dw, db)."""
    dx = np.zeros_like(x)
    dw = np.zeros_like(w)
    for t, (dy, ddx) in enumerate(taps_of(w.shape[0])):
        xs = shift(x, dy, ddx)
        dw[t] = np.einsum("nhwo,nhwi->oi", g, xs)
        dx += shift(g @ w[t], -dy, -ddx)
    return dx, dw, g.sum(axis=(0, 1, 2))


def avgpool(x):
    n, h, w, c = x.shape
    return x.reshape(n, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4))


def avgpool_grad(g):
    return np.repeat(np.repeat(g, 2, axis=1), 2, axis=2) / 4


def up2(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up2_grad(g):
    n, h, w, c = g.shape
    return g.reshape(n, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4))


def div_back(J):
    """J (N,H,W,2) [y, x] -> backward-difference divergence (N,H,W)."""
    return (J[..., 0] - np.roll(J[..., 0], 1, axis=1)) + (J[..., 1] - np.roll(J[..., 1], 1, axis=2))


def div_back_T(g):
    """Adjoint of div_back: returns dJ (N,H,W,2)."""
    return np.stack([g - np.roll(g, -1, axis=1), g - np.roll(g, -1, axis=2)], axis=-1)


def unpack(theta):
    layout, total = param_layout()
    assert theta.size == total
    P = {}
    for name, w_off, b_off, cin, cout, taps in layout:
        P[name] = (theta[w_off:b_off].reshape(taps, cout, cin), theta[b_off:b_off + cout])
    return P


def make_input(phi, params):
    """Input channels from the field as the device holds it (float32)."""
    n, h, w = phi.shape
    p32 = np.asarray(phi, dtype=np.float32)
    x = np.zeros((n, h, w, 16))
    x[..., 0] = p32
    x[..., 4] = (p32 - bf16(p32).astype(np.float32)) * np.float32(LO_SCALE)
    for c in range(3):
        x[..., 1 + c] = (np.asarray(params)[:, c] / PARAM_SCALE[c])[:, None, None]
    return x


def forward(theta, phi, params, emulate_bf16=True):
    """Returns (J, phi_next, cache)."""
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    P = unpack(theta)
    Wq = {k: (r(w) if k != "head" else w, b) for k, (w, b) in P.items()}
    c = {}
    c["x0"] = r(make_input(phi, params))

    def act(name, x):
        pre = conv(x, *Wq[name])
        c[name + "_pre"] = r(pre)
        return r(gelu(pre))

    c["e1"] = act("c1", c["x0"])
    c["p1"] = r(avgpool(c["e1"]))
    c["e2"] = act("c2", c["p1"])
    c["p2"] = r(avgpool(c["e2"]))
    c["bt"] = act("c3", c["p2"])
    c["u2"] = r(conv(c["bt"], *Wq["c4"]))
    c["s2"] = r(up2(c["u2"]) + c["e2"])
    c["d2"] = act("c5", c["s2"])
    c["u1"] = r(conv(c["d2"], *Wq["c6"]))
    c["s1"] = r(up2(c["u1"]) + c["e1"])
    c["d1"] = act("c7", c["s1"])
    J = J_SCALE * conv(c["d1"], *Wq["head"])
    return J, phi - div_back(J), c


def loss_and_grad(theta, phi, params, target, emulate_bf16=True):
    """Loss and gradient w.r.t. the flat parameter vector (gradients are taken with
    respect to the rounded weights the forward uses, as the GPU does)."""
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    P = unpack(theta)
    Wq = {k: (r(w) if k != "head" else w, b) for k, (w, b) in P.items()}
    J, pred, c = forward(theta, phi, params, emulate_bf16)
    n = phi.shape[0]
    norm = ((target - phi) ** 2).sum(axis=(1, 2)) + LOSS_FLOOR * phi.shape[1] * phi.shape[2]
    res = pred - target
    loss = float(np.mean((res**2).sum(axis=(1, 2)) / norm))
    g = 2 * res / (norm[:, None, None] * n)
    dJ = -J_SCALE * div_back_T(g)
    grads = {}
    dd1, dwh, dbh = conv_grad(c["d1"], Wq["head"][0], dJ)
    grads["head"] = (dwh, dbh)

    def back_act(name, x_in, dout):
        dpre = r(dout * gelu_grad(c[name + "_pre"]))
        dx, dw, db = conv_grad(x_in, Wq[name][0], dpre)
        grads[name] = (dw, db)
        return dx

    ds1 = back_act("c7", c["s1"], dd1)
    du1 = r(up2_grad(ds1))
    dd2, dw6, db6 = conv_grad(c["d2"], Wq["c6"][0], du1)
    grads["c6"] = (dw6, db6)
    ds2 = back_act("c5", c["s2"], dd2)
    du2 = r(up2_grad(ds2))
    dbt, dw4, db4 = conv_grad(c["bt"], Wq["c4"][0], du2)
    grads["c4"] = (dw4, db4)
    dp2 = back_act("c3", c["p2"], dbt)
    de2 = ds2 + avgpool_grad(dp2)
    dp1 = back_act("c2", c["p1"], de2)
    de1 = ds1 + avgpool_grad(dp1)
    back_act("c1", c["x0"], de1)

    layout, total = param_layout()
    flat = np.zeros(total)
    for name, w_off, b_off, cin, cout, taps in layout:
        dw, db = grads[name]
        flat[w_off:b_off] = dw.reshape(-1)
        flat[b_off:b_off + cout] = db
    return loss, flat


def init_params(seed):
    """He-uniform weights (zero on the 11 padding input channels of c1), zero biases."""
    rng = np.random.default_rng(seed)
    layout, total = param_layout()
    theta = np.zeros(total)
    for name, w_off, b_off, cin, cout, taps in layout:
        real_cin = CIN_REAL if name == "c1" else cin
        bound = np.sqrt(6.0 / (taps * real_cin))
        if name == "head":
            bound *= 0.1
        w = rng.uniform(-bound, bound, size=(taps, cout, cin))
        if name == "c1":
            w[:, :, CIN_REAL:] = 0
        theta[w_off:b_off] = w.reshape(-1)
    return theta
