"""Pins of the learned-closure oracle (oracle/mlp.py) and of the closure's
host-side data pipeline (dualsphysics_amd/closure.py), on the CPU.

- forward and gradient in FP64 against torch autograd (an independent
  implementation), and the gradient against central finite differences;
- BF16 rounding against torch's float32 -> bfloat16 conversion;
- the training-data filter against closed forms of linear velocity fields
  u = S x (filtered gradient = S, residual stress = sigma^2 S S^T with
  sigma^2 = (1/3) int W r^2 dV, least-squares eddy viscosity sigma^2 g / 2 for
  S = g diag(1, 1, -2));
- the reference's Smagorinsky viscosity (JSph.cpp:1528-1532,
  JSphGpu_ker.cu:1387-1436) from the strain convention of the interaction.
"""
import math

import numpy as np
import pytest
import torch

from dualsphysics_amd import closure
from oracle import mlp


def _data(n=77, seed=3):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, 11)) * np.array([30.0] * 9 + [20.0, 0.0]) + \
        np.array([0.0] * 9 + [1000.0, 0.02])
    mean = x.mean(0)
    scale = np.where(x.std(0) > 0, 1 / np.where(x.std(0) > 0, x.std(0), 1), 0)
    t = np.abs(rng.standard_normal(n)) * 1e-4
    return x, t, mean, scale, 1e-4


def _torch_loss(p, x, t, mean, scale, out_scale):
    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    w1 = pt[mlp.OFF_W1:mlp.OFF_B1].reshape(64, 16)[:, :11]
    b1 = pt[mlp.OFF_B1:mlp.OFF_W2]
    w2 = pt[mlp.OFF_W2:mlp.OFF_B2].reshape(64, 64)
    b2 = pt[mlp.OFF_B2:mlp.OFF_W3]
    w3 = pt[mlp.OFF_W3:mlp.OFF_B3]
    b3 = pt[mlp.OFF_B3]
    xn = (torch.tensor(x) - torch.tensor(mean)) * torch.tensor(scale)
    h1 = torch.nn.functional.silu(xn @ w1.T + b1)
    h2 = torch.nn.functional.silu(h1 @ w2.T + b2)
    o = h2 @ w3 + b3
    y = torch.nn.functional.softplus(o)
    loss = torch.mean((y - torch.tensor(t) / out_scale) ** 2)
    loss.backward()
    return out_scale * y.detach().numpy(), float(loss.detach()), pt.grad.numpy()


def test_forward_and_gradient_match_torch_autograd():
    x, t, mean, scale, os_ = _data()
    p = mlp.init_params(5).astype(np.float64)
    p[mlp.OFF_B1:mlp.OFF_W2] = np.linspace(-0.5, 0.5, 64)
    p[mlp.OFF_B3] = -1.0
    nu_t, loss_t, g_t = _torch_loss(p, x, t, mean, scale, os_)
    nu, _ = mlp.forward(p, x, mean, scale, os_, emulate_bf16=False)
    loss, g = mlp.loss_and_grad(p, x, t, mean, scale, os_, emulate_bf16=False)
    np.testing.assert_allclose(nu, nu_t, rtol=1e-12)
    assert abs(loss - loss_t) <= 1e-12 * loss_t
    g_t = g_t.copy()
    g_t[:mlp.OFF_B1].reshape(64, 16)[:, 11:] = 0  # padding columns carry no gradient
    np.testing.assert_allclose(g, g_t, rtol=1e-9, atol=1e-12 * np.abs(g_t).max())


def test_gradient_matches_finite_differences():
    x, t, mean, scale, os_ = _data(n=40, seed=9)
    p = mlp.init_params(2).astype(np.float64)
    _, g = mlp.loss_and_grad(p, x, t, mean, scale, os_, emulate_bf16=False)
    rng = np.random.default_rng(0)
    blocks = [(mlp.OFF_W1, mlp.OFF_B1), (mlp.OFF_B1, mlp.OFF_W2), (mlp.OFF_W2, mlp.OFF_B2),
              (mlp.OFF_B2, mlp.OFF_W3), (mlp.OFF_W3, mlp.OFF_B3), (mlp.OFF_B3, mlp.N_PARAM)]
    for lo, hi in blocks:
        for i in rng.choice(np.arange(lo, hi), min(6, hi - lo), replace=False):
            if i < mlp.OFF_B1 and i % 16 >= 11:
                continue
            e = 1e-6
            pp, pm = p.copy(), p.copy()
            pp[i] += e
            pm[i] -= e
            fd = (mlp.loss_and_grad(pp, x, t, mean, scale, os_, False)[0] -
                  mlp.loss_and_grad(pm, x, t, mean, scale, os_, False)[0]) / (2 * e)
            assert abs(fd - g[i]) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(g).max()), (i, fd, g[i])


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10,
                        # exact ties at the 16th bit: round to even both ways
                        (np.arange(1, 200, dtype=np.uint32) << 16 | 0x8000).view(np.float32),
                        np.array([0.0, -0.0, 1e-30, -3.3e38], np.float32)])
    ref = torch.tensor(x).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)
    np.testing.assert_array_equal(mlp.bf16(x), ref)


def test_bf16_emulated_gradient_close_to_exact():
    x, t, mean, scale, os_ = _data(n=200)
    p = mlp.init_params(4)
    _, g0 = mlp.loss_and_grad(p, x, t, mean, scale, os_, emulate_bf16=False)
    _, g1 = mlp.loss_and_grad(p, x, t, mean, scale, os_, emulate_bf16=True)
    assert np.linalg.norm(g1 - g0) < 0.03 * np.linalg.norm(g0)


def _wendland_sigma2(h):
    from scipy.integrate import quad
    awen = 21 / (16 * math.pi * h ** 3)
    w = lambda r: awen * (2 * r / h + 1) * (1 - 0.5 * r / h) ** 4
    m0 = quad(lambda r: w(r) * 4 * math.pi * r * r, 0, 2 * h)[0]
    m2 = quad(lambda r: w(r) * r * r * 4 * math.pi * r * r, 0, 2 * h)[0]
    assert abs(m0 - 1) < 1e-10  # the kernel is normalised
    return m2 / 3


@pytest.mark.parametrize("offset", [(0, 0, 0), (0.31, -0.17, 0.42)])
def test_filter_of_linear_field_closed_forms(offset):
    dpf, dpc = 0.01, 0.02
    hc = 1.2 * math.sqrt(3) * dpc
    g = 7.0
    s = g * np.diag([1.0, 1.0, -2.0]) + np.array([[0, 2.0, 0], [2.0, 0, 0], [0, 0, 0]])
    k = np.arange(-12, 13) * dpf
    pos = np.stack(np.meshgrid(k, k, k, indexing="ij"), -1).reshape(-1, 3)
    u0 = np.array([0.3, -0.1, 0.2])
    vel = u0 + pos @ s.T
    rho = np.full(len(pos), 1000.0)
    probe = np.array(offset, float)[None] * dpf
    mom = mlp.filter_moments(pos, vel, rho, 1000.0 * dpf ** 3, probe, hc)
    assert abs(mom[0, 0] - 1) < 2e-3          # partition of unity on the lattice
    feats, nu = closure.targets_from_moments(mom, dpc)
    np.testing.assert_allclose(feats[0, :9].reshape(3, 3), s, atol=5e-3 * g)
    assert feats[0, 9] == pytest.approx(1000.0) and feats[0, 10] == pytest.approx(dpc)
    # residual stress of a linear field: tau = S Cov S^T, Cov -> sigma^2 I
    sig2 = _wendland_sigma2(hc)
    w = mom[0, 0]
    tau = np.empty((3, 3))
    comp = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    ub = mom[0, 1:4] / w
    for j, (a, b) in enumerate(comp):
        tau[a, b] = tau[b, a] = mom[0, 4 + j] / w - ub[a] * ub[b]
    np.testing.assert_allclose(tau, sig2 * s @ s.T, rtol=0.02, atol=0.02 * sig2 * g * g)
    # least-squares eddy viscosity: -tau_d:S_d / (2 S_d:S_d) with S_d = s (traceless)
    nu_cf = -sig2 * np.trace(s @ s.T @ s) / (2 * np.sum(s * s))
    assert nu[0] == pytest.approx(max(nu_cf, 0.0), rel=0.03)
    assert nu_cf > 0


def test_smagorinsky_viscosity_matches_reference_form():
    rng = np.random.default_rng(4)
    gr = rng.standard_normal((50, 3, 3)) * 20
    dp = 0.02
    smag = (2.0 / 3.0) * (0.12 * dp) ** 2
    # interaction strain convention: xx, xy (= dvx/dy + dvy/dx), xz, yy, yz, zz
    xx, yy, zz = gr[:, 0, 0], gr[:, 1, 1], gr[:, 2, 2]
    xy = gr[:, 0, 1] + gr[:, 1, 0]
    xz = gr[:, 0, 2] + gr[:, 2, 0]
    yz = gr[:, 1, 2] + gr[:, 2, 1]
    pow1 = xx * xx + yy * yy + zz * zz
    prr = xy * xy + xz * xz + yz * yz + pow1 + pow1
    np.testing.assert_allclose(closure.smagorinsky_nu(gr.reshape(-1, 9), dp), smag * np.sqrt(prr),
                               rtol=1e-12)


def test_planted_error_in_oracle_gradient_is_caught():
    """A transposed W2 in the backward pass (a plausible layout mistake)
    is detected by the autograd comparison."""
    x, t, mean, scale, os_ = _data()
    p = mlp.init_params(5).astype(np.float64)
    _, _, g_t = _torch_loss(p, x, t, mean, scale, os_)
    w2 = p[mlp.OFF_W2:mlp.OFF_B2].reshape(64, 64)
    q = p.copy()
    q[mlp.OFF_W2:mlp.OFF_B2] = w2.T.ravel()
    _, g_bad = mlp.loss_and_grad(q, x, t, mean, scale, os_, emulate_bf16=False)
    # the W1 gradient flows through W2 in the backward pass: it must differ
    sl = slice(mlp.OFF_W1, mlp.OFF_B1)
    assert np.linalg.norm(g_bad[sl] - g_t[sl]) > 0.05 * np.linalg.norm(g_t[sl])
