"""CPU oracle of the learned SPS closure network (numpy, FP64 arithmetic on
the same BF16-rounded operands as the kernels).

TEST INFRASTRUCTURE ONLY: the checker of dualsphysics_amd/csrc/learned.hip
(forward, gradient, training-data filter) in tests/; the product never
imports it. The closure has no counterpart in the reference: the network
replaces the Smagorinsky eddy viscosity of the reference's SPS model
(JSphGpu_ker.cu:1387-1436); the pins are torch autograd (forward and
gradient in FP64), torch's BF16 conversion, central finite differences and,
for the filter, closed forms of linear velocity fields.

Network: x' = (x - mean) * scale (11 features), z1 = W1 x' + b1, h1 = silu(z1),
z2 = W2 h1 + b2, h2 = silu(z2), o = w3 . h2 + b3, nu = out_scale * softplus(o).
Loss L = mean_i (softplus(o_i) - t_i / out_scale)^2. The kernels round to
BF16: x', W1, W2, w3, h1 (as operand of layer 2 and of dW2), dZ2 (operand of
dW2 and dH1), dZ1 and x' (operands of dW1); this oracle rounds at the same
places.
"""
from __future__ import annotations

import numpy as np

N_IN, N_PAD, N_HID = 11, 16, 64
OFF_W1, OFF_B1 = 0, 64 * 16
OFF_W2 = OFF_B1 + 64
OFF_B2 = OFF_W2 + 64 * 64
OFF_W3 = OFF_B2 + 64
OFF_B3 = OFF_W3 + 64
N_PARAM = OFF_B3 + 1


def bf16(x) -> np.ndarray:
    """Round FP32 values to BF16 (nearest even), returned as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p: np.ndarray):
    p = np.asarray(p, np.float64)
    w1 = p[OFF_W1:OFF_B1].reshape(64, 16)[:, :N_IN]
    return (w1, p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(64, 64), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3], p[OFF_B3])


def silu(z):
    return z / (1 + np.exp(-z))


def dsilu(z):
    s = 1 / (1 + np.exp(-z))
    return s * (1 + z * (1 - s))


def softplus(z):
    return np.where(z > 20, z, np.log1p(np.exp(np.minimum(z, 20))))


def sigmoid(z):
    return 1 / (1 + np.exp(-z))


def forward(p, x, mean, scale, out_scale, emulate_bf16=True):
    """Returns nu and the intermediates."""
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    w1, b1, w2, b2, w3, b3 = unpack(p)
    xn = r((np.asarray(x, np.float64) - mean) * scale)
    z1 = xn @ r(w1).T + b1
    h1 = r(silu(z1))
    z2 = h1 @ r(w2).T + b2
    o = silu(z2) @ r(w3) + b3
    return out_scale * softplus(o), dict(xn=xn, z1=z1, h1=h1, z2=z2, o=o)


def loss_and_grad(p, x, t, mean, scale, out_scale, emulate_bf16=True):
    """Mean squared error of softplus(o) against t/out_scale and its
    gradient with respect to the parameter block (padding columns 0)."""
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    w1, b1, w2, b2, w3, b3 = unpack(p)
    n = len(x)
    nu, f = forward(p, x, mean, scale, out_scale, emulate_bf16)
    y = softplus(f["o"])
    e = y - np.asarray(t, np.float64) / out_scale
    loss = float(np.mean(e * e))
    dout = 2 * e / n * sigmoid(f["o"])
    h2 = silu(f["z2"])
    gw3 = dout @ h2
    gb3 = dout.sum()
    dz2 = dout[:, None] * r(w3)[None, :] * dsilu(f["z2"])
    gb2 = dz2.sum(0)
    rdz2 = r(dz2)
    gw2 = rdz2.T @ f["h1"]
    dz1 = (rdz2 @ r(w2)) * dsilu(f["z1"])
    gb1 = dz1.sum(0)
    gw1 = r(dz1).T @ f["xn"]
    g = np.zeros(N_PARAM)
    g1 = np.zeros((64, 16))
    g1[:, :N_IN] = gw1
    g[OFF_W1:OFF_B1] = g1.ravel()
    g[OFF_B1:OFF_W2] = gb1
    g[OFF_W2:OFF_B2] = gw2.ravel()
    g[OFF_B2:OFF_W3] = gb2
    g[OFF_W3:OFF_B3] = gw3
    g[OFF_B3] = gb3
    return loss, g


def init_params(seed: int = 0) -> np.ndarray:
    """Glorot-uniform weights, zero biases (FP32)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAM, np.float32)
    lim1 = np.sqrt(6 / (N_IN + N_HID))
    w1 = np.zeros((64, 16), np.float32)
    w1[:, :N_IN] = rng.uniform(-lim1, lim1, (64, N_IN))
    p[OFF_W1:OFF_B1] = w1.ravel()
    lim2 = np.sqrt(6 / (2 * N_HID))
    p[OFF_W2:OFF_B2] = rng.uniform(-lim2, lim2, 64 * 64)
    lim3 = np.sqrt(6 / (N_HID + 1))
    p[OFF_W3:OFF_B3] = rng.uniform(-lim3, lim3, 64)
    return p


def filter_moments(pos, vel, rho, mass, probes, hf):
    """Moments of the coarse-scale filter of a fine particle state at probe
    points (brute force over all given particles, FP64), in the order of
    dsph_filter_probes: [sum VW, sum VW u (3), sum VW u_a u_b (xx,xy,xz,yy,yz,zz),
    sum VW rho, sum V u_a dW/dx_b (9, row-major), sum V dW/dx_b (3)] with the
    3-D Wendland C2 kernel of smoothing length hf (support 2 hf), V = m/rho,
    dW/dx taken at the probe."""
    pos = np.asarray(pos, np.float64)
    vel = np.asarray(vel, np.float64)
    rho = np.asarray(rho, np.float64)
    awen = 21.0 / (16.0 * np.pi * hf ** 3)
    out = np.zeros((len(probes), 23))
    comp = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for i, p in enumerate(np.asarray(probes, np.float64)):
        d = p[None, :] - pos
        r2 = np.einsum("nk,nk->n", d, d)
        m = r2 < 4 * hf * hf
        d, u, rh = d[m], vel[m], rho[m]
        q = np.sqrt(r2[m]) / hf
        w1 = 1 - 0.5 * q
        w = awen * (2 * q + 1) * w1 ** 4
        fac = -5 * awen / hf ** 2 * w1 ** 3
        vol = mass / rh
        gw = fac[:, None] * d
        out[i, 0] = np.sum(vol * w)
        out[i, 1:4] = (vol * w) @ u
        for k, (a, b) in enumerate(comp):
            out[i, 4 + k] = np.sum(vol * w * u[:, a] * u[:, b])
        out[i, 10] = np.sum(vol * w * rh)
        out[i, 11:20] = np.einsum("n,na,nb->ab", vol, u, gw).ravel()
        out[i, 20:23] = vol @ gw
    return out
