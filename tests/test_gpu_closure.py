"""Learned SPS closure on the GPU (csrc/learned.hip) against the oracle
(oracle/mlp.py), and its a-priori gain over the reference's Smagorinsky model.

Tolerances: the kernels use BF16 operands with FP32 accumulation and the
hardware exp; the oracle rounds the same operands to BF16 and computes in
FP64. Forward: relative error of nu <= 1e-2 per sample, median <= 1e-3.
Gradient: norm-wise relative error <= 2e-2 per parameter block. A layout or
transposition mistake gives O(1) errors.
"""
import ctypes as C
import math

import numpy as np
import pytest

from dualsphysics_amd import _native as N
from dualsphysics_amd import case as casemod
from dualsphysics_amd import closure
from tests.helpers import Dev

pytestmark = pytest.mark.gpu

BLOCKS = {"W1": (0, 1024), "b1": (1024, 1088), "W2": (1088, 5184), "b2": (5184, 5248),
          "w3": (5248, 5312), "b3": (5312, 5313)}


def _features(n, seed=0, dp=0.02):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, 9)) * rng.lognormal(1.0, 1.0, (n, 1))
    rho = 1000 + rng.standard_normal((n, 1)) * 15
    return np.concatenate([g, rho, np.full((n, 1), dp)], 1).astype(np.float32)


def _net(lib, ctx, feats, seed=1, out_scale=1e-5):
    from oracle import mlp
    p = mlp.init_params(seed)
    rng = np.random.default_rng(seed)
    p[mlp.OFF_B1:mlp.OFF_W2] = rng.uniform(-0.5, 0.5, 64)
    p[mlp.OFF_B2:mlp.OFF_W3] = rng.uniform(-0.5, 0.5, 64)
    p[mlp.OFF_B3] = 0.3
    mean, scale, _ = closure.normalisation(feats, np.ones(len(feats)))
    return closure.Mlp(lib, ctx, p, mean, scale, out_scale), p, mean, scale


@pytest.mark.parametrize("n", [1, 16, 1003])
def test_forward_matches_oracle(lib, ctx, n):
    from oracle import mlp
    feats = _features(max(n, 64), seed=n)[:n]
    net, p, mean, scale = _net(lib, ctx, _features(4096))
    try:
        got = net.forward(feats)
        exp, _ = mlp.forward(p, feats, mean, scale, net.m.out_scale)
        rel = np.abs(got - exp) / np.abs(exp)
        assert rel.max() <= 1e-2 and np.median(rel) <= 1e-3, (rel.max(), np.median(rel))
    finally:
        net.free()


@pytest.mark.parametrize("n,scratch_blocks", [(1003, 1024), (5000, 7)])
def test_gradient_matches_oracle(lib, ctx, n, scratch_blocks):
    """scratch_blocks < tiles makes each block loop over several tiles."""
    from oracle import mlp
    feats = _features(n, seed=3)
    rng = np.random.default_rng(4)
    t = (closure.smagorinsky_nu(feats[:, :9], 0.02) * rng.uniform(0.5, 2.0, n)).astype(np.float32)
    net, p, mean, scale = _net(lib, ctx, feats, out_scale=float(t.mean()))
    try:
        sb = scratch_blocks * 4 * (net.nparam + 1)
        df, dt = Dev(lib, ctx, feats), Dev(lib, ctx, t)
        ds = Dev(lib, ctx, nbytes=sb, dtype=np.uint8, shape=(sb,))
        dg = Dev.zeros(lib, ctx, net.nparam, np.float32)
        dl = Dev.zeros(lib, ctx, 1, np.float32)
        rc = lib.dsph_mlp_grad(ctx, C.byref(net.m), n, df.ptr, dt.ptr, ds.ptr, sb, dg.ptr, dl.ptr)
        assert rc == 0, lib.dsph_last_error(ctx)
        g, loss = dg.get(), float(dl.get()[0])
        lo, go = mlp.loss_and_grad(p, feats, t, mean, scale, net.m.out_scale)
        assert abs(loss - lo) <= 2e-3 * lo, (loss, lo)
        for name, (a, b) in BLOCKS.items():
            err = np.linalg.norm(g[a:b] - go[a:b]) / np.linalg.norm(go[a:b])
            assert err <= 2e-2, (name, err)
        assert np.all(g[:1024].reshape(64, 16)[:, 11:] == 0)
    finally:
        net.free()


def test_scratch_too_small_is_an_error(lib, ctx):
    feats = _features(64)
    net, *_ = _net(lib, ctx, feats)
    try:
        df, dt = Dev(lib, ctx, feats), Dev.zeros(lib, ctx, 64, np.float32)
        ds = Dev.zeros(lib, ctx, 100, np.uint8)
        dg = Dev.zeros(lib, ctx, net.nparam, np.float32)
        rc = lib.dsph_mlp_grad(ctx, C.byref(net.m), 64, df.ptr, dt.ptr, ds.ptr, 100, dg.ptr, None)
        assert rc != 0 and b"scratch" in lib.dsph_last_error(ctx)
    finally:
        net.free()


def test_adam_step_matches_formula(lib, ctx):
    from oracle import mlp
    feats = _features(256)
    net, p, mean, scale = _net(lib, ctx, feats)
    try:
        rng = np.random.default_rng(8)
        g1 = rng.standard_normal(net.nparam).astype(np.float32) * 1e-2
        g2 = rng.standard_normal(net.nparam).astype(np.float32) * 1e-2
        lr, b1, b2, eps = 1e-3, 0.9, 0.999, 1e-8
        m = np.zeros(net.nparam)
        v = np.zeros(net.nparam)
        pe = p.astype(np.float64)
        for step, g in enumerate((g1, g2), start=1):
            dgr = Dev(lib, ctx, g)
            assert lib.dsph_mlp_adam(ctx, C.byref(net.m), dgr.ptr, lr, b1, b2, eps, step) == 0
            gg = g.astype(np.float64).copy()
            gg[:1024].reshape(64, 16)[:, 11:] = 0
            m = b1 * m + (1 - b1) * gg
            v = b2 * v + (1 - b2) * gg * gg
            pe -= lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
        got = net.params()
        np.testing.assert_allclose(got, pe, rtol=0, atol=1e-6)
        # the BF16 copies follow the new parameters
        nu = net.forward(feats)
        exp, _ = mlp.forward(got, feats, mean, scale, net.m.out_scale)
        assert np.max(np.abs(nu - exp) / exp) <= 1e-2
    finally:
        net.free()


def test_training_reduces_loss(lib, ctx):
    feats = _features(8192, seed=11)
    t = closure.smagorinsky_nu(feats[:, :9], 0.02).astype(np.float32)
    mean, scale, os_ = closure.normalisation(feats, t)
    net = closure.Mlp(lib, ctx, closure.init_params(0), mean, scale, os_)
    try:
        hist = net.train(feats, t, steps=400, batch=2048, lr=3e-3)
        assert np.isfinite(hist).all()
        assert np.mean(hist[-20:]) < 0.05 * np.mean(hist[:5]), (hist[:5], hist[-20:])
        pred = net.forward(feats)
        rel = np.sqrt(np.mean((pred - t) ** 2)) / np.sqrt(np.mean(t ** 2))
        assert rel < 0.15, rel
    finally:
        net.free()


def test_sps_tau_learned_matches_formula(lib, ctx):
    """tau of KerComputeSpsTau (JSphGpu_ker.cu:1387-1436) with the network's
    eddy viscosity in place of smag*sqrt(prr); boundary rows untouched."""
    from oracle import mlp
    npb, n = 37, 1000
    rng = np.random.default_rng(5)
    feats = _features(n, seed=6)
    velrho = np.zeros((npb + n, 4), np.float32)
    velrho[:, :3] = rng.standard_normal((npb + n, 3))
    velrho[:, 3] = 1000.0
    velrho[npb:, 3] = feats[:, 9]
    velgrad = np.zeros((npb + n, 9), np.float32)
    velgrad[npb:] = feats[:, :9]
    strain = (rng.standard_normal((npb + n, 6)) * 10).astype(np.float32)
    net, p, mean, scale = _net(lib, ctx, feats)
    try:
        blin = float((2 / 3) * 0.0066 * 0.02 ** 2)
        dvr, dvg, dst = Dev(lib, ctx, velrho), Dev(lib, ctx, velgrad), Dev(lib, ctx, strain)
        tau0 = np.full((npb + n, 6), 7.0, np.float32)
        dtau = Dev(lib, ctx, tau0)
        dnu = Dev.zeros(lib, ctx, npb + n, np.float32)
        rc = lib.dsph_sps_tau_learned(ctx, C.byref(net.m), npb + n, npb, blin, 0.02, dvr.ptr,
                                      dvg.ptr, dst.ptr, dtau.ptr, dnu.ptr)
        assert rc == 0, lib.dsph_last_error(ctx)
        tau, nu = dtau.get(), dnu.get()
        nu_o, _ = mlp.forward(p, feats, mean, scale, net.m.out_scale)
        np.testing.assert_allclose(nu[npb:], nu_o, rtol=1e-2)
        assert np.all(nu[:npb] == 0) and np.all(tau[:npb] == 7.0)
        s = strain[npb:].astype(np.float64)
        xx, xy, xz, yy, yz, zz = s.T
        pow1 = xx * xx + yy * yy + zz * zz
        prr = xy * xy + xz * xz + yz * yz + pow1 + pow1
        vs = nu[npb:].astype(np.float64)
        sumsps = -(vs * (xx + yy + zz) + blin * prr)
        r = 1 / velrho[npb:, 3].astype(np.float64)
        exp = np.stack([r * (2 * vs * xx + sumsps), r * vs * xy, r * vs * xz,
                        r * (2 * vs * yy + sumsps), r * vs * yz, r * (2 * vs * zz + sumsps)], 1)
        np.testing.assert_allclose(tau[npb:], exp, rtol=1e-5, atol=1e-6 * np.abs(exp).max())
    finally:
        net.free()


def test_filter_probes_match_oracle():
    """dsph_filter_probes on a running dam break vs the oracle's brute force."""
    from dualsphysics_amd.sim import SphGpuSingle
    from oracle import mlp
    c = casemod.dam_break(dp=0.02)
    sim = SphGpuSingle(c)
    try:
        sim.run(300)
        d = sim.download()
        fl = d["idp"] >= c.nbound
        pos, vr = d["pos"][fl], d["velrho"][fl]
        rng = np.random.default_rng(0)
        probes = np.concatenate([pos[rng.choice(len(pos), 150, replace=False)],
                                 rng.uniform([-0.1, -0.1, -0.1], [1.8, 0.8, 0.5], (50, 3))])
        hf = 1.2 * math.sqrt(3) * 0.04
        got = sim.filter_probes(probes, hf)
        # the solver's fluid mass is FP32 (JSph.cpp:1481-1546)
        exp = mlp.filter_moments(pos, vr[:, :3], vr[:, 3], float(np.float32(c.massfluid)),
                                 probes, hf)
        scale = np.abs(exp).max(0)
        scale[scale == 0] = 1
        err = np.abs(got - exp) / scale
        assert err.max() <= 1e-9, err.max(0)
        assert (got[:150, 0] > 0).all()
    finally:
        sim.close()


def test_closure_in_simulation_matches_oracle(lib, ctx):
    """Coarse dam break with Laminar+SPS and the learned closure: the eddy
    viscosity of the interaction equals the oracle network on the
    interaction's own velocity gradient, density and dp."""
    from dualsphysics_amd.sim import SphGpuSingle
    from oracle import mlp
    prm = casemod.Parameters(tvisco=N.VISCO_LAMINARSPS, visco=1e-6)
    c = casemod.dam_break(dp=0.03, params=prm)
    feats = _features(4096, dp=0.03)
    net, p, mean, scale = _net(lib, ctx, feats, out_scale=1e-6)
    sim = SphGpuSingle(c)
    try:
        sim.set_closure(net)
        sim.run(50)
        sim.interact()
        nu, g = sim.closure_nu(with_velgrad=True)
        d = sim.download()
        fl = d["idp"] >= c.nbound
        f = np.concatenate([g[fl], d["velrho"][fl, 3:4], np.full((fl.sum(), 1), 0.03)], 1)
        exp, _ = mlp.forward(p, f, mean, scale, net.m.out_scale)
        rel = np.abs(nu[fl] - exp) / exp
        assert rel.max() <= 1e-2 and np.median(rel) <= 1e-3, (rel.max(), np.median(rel))
        assert np.all(nu[~fl] == 0)
        st = sim.stats()
        assert st["np"] == c.np
    finally:
        sim.close()
        net.free()


def test_learned_closure_a_priori_gain(lib, ctx):
    """Filtered fine run (dp 0.01) at the coarse scale (dp 0.02): train on
    t = 0.2..0.5 s, evaluate on held-out probes at t = 0.6 s. Metric: mean
    squared error of the eddy viscosity against the filtered optimum nu*.
    The learned closure must beat the reference's Smagorinsky model and the
    best single-constant Smagorinsky fit by a clear margin (MSE below 0.6x
    the better of the two)."""
    fine = casemod.dam_break(dp=0.01)
    dpc = 0.02
    times = [0.2, 0.3, 0.4, 0.5, 0.6]
    feats, tgt, ti = closure.training_data(fine, dpc, times, nprobe=6000, seed=1, log=print)
    test = ti == len(times) - 1
    xtr, ytr = feats[~test], tgt[~test]
    xte, yte = feats[test], tgt[test]
    mean, scale, os_ = closure.normalisation(xtr, ytr)
    net = closure.Mlp(lib, ctx, closure.init_params(0), mean, scale, os_)
    try:
        hist = net.train(xtr, ytr, steps=1500, batch=4096, lr=2e-3)
        pred = net.forward(xte)
    finally:
        net.free()
    smag = closure.smagorinsky_nu(xte[:, :9], dpc)
    s_tr = closure.smagorinsky_nu(xtr[:, :9], dpc)
    cfit = float(np.dot(s_tr, ytr) / np.dot(s_tr, s_tr))
    mse = lambda a: float(np.mean((a - yte) ** 2))
    m_l, m_s, m_f = mse(pred), mse(smag), mse(cfit * smag)
    print(f"a-priori MSE (held out, {len(yte)} probes): learned {m_l:.4e}  "
          f"Smagorinsky {m_s:.4e}  fitted-constant Smagorinsky (x{cfit:.3f}) {m_f:.4e}  "
          f"zero {mse(0 * yte):.4e}; train loss {hist[0]:.4f} -> {np.mean(hist[-50:]):.4f}")
    # observed: learned 7.5e-9, Smagorinsky 2.4e-8, fitted constant 2.2e-8
    assert m_l < 0.6 * min(m_s, m_f)
