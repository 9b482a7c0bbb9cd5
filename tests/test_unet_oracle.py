"""Pins the CPU reference of the surrogate network (oracle/unet_oracle.py): its
analytic backward pass against central finite differences of its own forward, in
float64 without bfloat16 rounding, and the exact mass conservation of the
flux-form update."""
import numpy as np
import pytest

from oracle import unet_oracle as U


def sample(n=2, h=8, w=8, seed=0):
    rng = np.random.default_rng(seed)
    phi = rng.normal(0, 0.3, size=(n, h, w))
    params = np.stack([rng.uniform(-2, 2, n), rng.uniform(-5, 0, n), rng.uniform(-0.6, 0.6, n)], axis=1)
    target = phi + 0.05 * rng.normal(size=phi.shape)
    return phi, params, target


def test_param_layout():
    layout, total = U.param_layout()
    assert layout[0][0] == "c1" and layout[-1][0] == "head"
    assert total == sum(t * co * ci + co for _, ci, co, t, _ in U.LAYERS)


def test_backward_matches_finite_differences():
    phi, params, target = sample()
    theta = U.init_params(1)
    theta += 0.05 * np.random.default_rng(2).normal(size=theta.size)  # move off zero biases
    layout, _ = U.param_layout()
    theta[layout[0][1]:layout[0][2]].reshape(9, 32, 16)[:, :, 5:] = 0
    loss, g = U.loss_and_grad(theta, phi, params, target, emulate_bf16=False)
    rng = np.random.default_rng(3)
    idx = []
    for name, w_off, b_off, cin, cout, taps in layout:
        real = [i for i in range(w_off, b_off) if name != "c1" or (i - w_off) % cin < 5]
        idx += list(rng.choice(real, 3, replace=False)) + [b_off + int(rng.integers(cout))]
    eps = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += eps
        tm[i] -= eps
        lp = U.loss_and_grad(tp, phi, params, target, emulate_bf16=False)[0]
        lm = U.loss_and_grad(tm, phi, params, target, emulate_bf16=False)[0]
        fd = (lp - lm) / (2 * eps)
        assert fd == pytest.approx(g[i], rel=2e-5, abs=1e-9), i


def test_flux_form_conserves_mass_exactly():
    phi, params, _ = sample(n=3, h=16, w=16, seed=4)
    theta = U.init_params(5)
    J, nxt, _ = U.forward(theta, phi, params, emulate_bf16=False)
    assert np.allclose(nxt.sum(axis=(1, 2)), phi.sum(axis=(1, 2)), atol=1e-12)
    # the divergence and its adjoint are consistent
    rng = np.random.default_rng(6)
    a, b = rng.normal(size=(1, 16, 16, 2)), rng.normal(size=(1, 16, 16))
    assert np.sum(U.div_back(a) * b) == pytest.approx(np.sum(a * U.div_back_T(b)), rel=1e-12)


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.0e-3, 65504.0])
    r = U.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125  # ties to even, then up
    assert np.all(np.abs(r - x) <= np.abs(x) * 2**-8)
