"""Learned SPS closure: training data from fine-resolution runs, training on
the GPU, and use in coarse runs.

The network (csrc/learned.hip) predicts the eddy viscosity nu_t of the
sub-particle-scale stress from 11 features: the velocity gradient
dv_a/dx_b (9, row-major), the density and dp. In a Laminar+SPS run it
replaces the Smagorinsky viscosity smag*|S| of the reference's SPS model
(JSphGpu_ker.cu:1387-1436, smag = (2/3)(0.12 dp)^2).

Training data (a-priori, from this solver): a fine run (dp_f) is filtered at
the coarse scale (the Wendland kernel of the coarse h) at probe points inside
the fluid (dsph_filter_probes): filtered velocity u~, velocity products
(u u)~, density rho~ and the SPH velocity gradient G~ of the filtered field.
The target is the least-squares eddy viscosity of the residual stress
tau = (u u)~ - u~ u~:  nu* = max(0, -tau_dev : S~_dev / (2 S~_dev : S~_dev)),
S~ = (G~ + G~^T)/2. Features: G~, rho~, dp_coarse.

Training runs entirely on the GPU (dsph_mlp_grad: fused forward + backward
on the matrix cores; dsph_mlp_adam).
"""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np

from . import _native as N

N_FEAT = 11


def _alloc(lib, ctx, nbytes):
    p = C.c_void_p()
    rc = lib.dsph_malloc(ctx, max(int(nbytes), 4), C.byref(p))
    if rc:
        raise RuntimeError(lib.dsph_last_error(ctx).decode())
    return p


class Mlp:
    """The network's device buffers (FP32 parameters, BF16 copies, Adam
    moments) and feature normalisation on one context."""

    def __init__(self, lib, ctx, params: np.ndarray, feat_mean, feat_scale, out_scale: float):
        self.lib, self.ctx = lib, ctx
        self.nparam = lib.dsph_mlp_param_count()
        nbf = lib.dsph_mlp_bf16_count()
        params = np.ascontiguousarray(params, np.float32)
        if params.shape != (self.nparam,):
            raise ValueError(f"expected {self.nparam} parameters")
        self._bufs = [_alloc(lib, ctx, 4 * self.nparam), _alloc(lib, ctx, 2 * nbf),
                      _alloc(lib, ctx, 4 * self.nparam), _alloc(lib, ctx, 4 * self.nparam)]
        m = N.DsphMlp()
        m.d_params, m.d_wbf16, m.d_adam_m, m.d_adam_v = [b.value for b in self._bufs]
        for k in range(N_FEAT):
            m.feat_mean[k] = float(feat_mean[k])
            m.feat_scale[k] = float(feat_scale[k])
        m.out_scale = float(out_scale)
        self.m = m
        self.set_params(params)
        z = np.zeros(self.nparam, np.float32)
        self._h2d(m.d_adam_m, z)
        self._h2d(m.d_adam_v, z)
        self.step = 0

    def _h2d(self, dst, a):
        a = np.ascontiguousarray(a)
        rc = self.lib.dsph_memcpy_h2d(self.ctx, C.c_void_p(dst), a.ctypes.data_as(C.c_void_p),
                                      a.nbytes)
        if rc:
            raise RuntimeError(self.lib.dsph_last_error(self.ctx).decode())

    def _d2h(self, src, a):
        rc = self.lib.dsph_memcpy_d2h(self.ctx, a.ctypes.data_as(C.c_void_p), C.c_void_p(src),
                                      a.nbytes)
        if rc:
            raise RuntimeError(self.lib.dsph_last_error(self.ctx).decode())
        return a

    def _check(self, rc):
        if rc:
            raise RuntimeError(self.lib.dsph_last_error(self.ctx).decode())

    def set_params(self, p):
        self._h2d(self.m.d_params, np.ascontiguousarray(p, np.float32))
        self._check(self.lib.dsph_mlp_sync(self.ctx, C.byref(self.m)))

    def synchronize(self):
        """Waits for the network's pending work (training) on its context."""
        self._check(self.lib.dsph_synchronize(self.ctx))

    def params(self) -> np.ndarray:
        return self._d2h(self.m.d_params, np.empty(self.nparam, np.float32))

    def forward(self, feats: np.ndarray) -> np.ndarray:
        feats = np.ascontiguousarray(feats, np.float32)
        n = len(feats)
        df, do = _alloc(self.lib, self.ctx, feats.nbytes), _alloc(self.lib, self.ctx, 4 * n)
        try:
            self._h2d(df.value, feats)
            self._check(self.lib.dsph_mlp_forward(self.ctx, C.byref(self.m), n, df, do))
            return self._d2h(do.value, np.empty(n, np.float32))
        finally:
            self.lib.dsph_free(self.ctx, df)
            self.lib.dsph_free(self.ctx, do)

    def grad(self, feats, targets):
        """(loss, gradient) of the mean squared error on one batch."""
        feats = np.ascontiguousarray(feats, np.float32)
        targets = np.ascontiguousarray(targets, np.float32)
        n = len(feats)
        sb = 1024 * 4 * (self.nparam + 1)
        bufs = [_alloc(self.lib, self.ctx, x) for x in (feats.nbytes, targets.nbytes, sb,
                                                        4 * self.nparam, 4)]
        try:
            self._h2d(bufs[0].value, feats)
            self._h2d(bufs[1].value, targets)
            self._check(self.lib.dsph_mlp_grad(self.ctx, C.byref(self.m), n, bufs[0], bufs[1],
                                               bufs[2], sb, bufs[3], bufs[4]))
            g = self._d2h(bufs[3].value, np.empty(self.nparam, np.float32))
            l = self._d2h(bufs[4].value, np.empty(1, np.float32))
            return float(l[0]), g
        finally:
            for b in bufs:
                self.lib.dsph_free(self.ctx, b)

    def train(self, feats, targets, steps: int = 2000, batch: int = 4096, lr: float = 2e-3,
              seed: int = 0, log_every: int = 0):
        """Adam on mini-batches; data resident on the device. Returns the
        loss history (every step, from the last batch)."""
        feats = np.ascontiguousarray(feats, np.float32)
        targets = np.ascontiguousarray(targets, np.float32)
        n = len(feats)
        batch = min(batch, n)
        lib, ctx = self.lib, self.ctx
        rng = np.random.default_rng(seed)
        sb = 1024 * 4 * (self.nparam + 1)
        dfe = _alloc(lib, ctx, 4 * N_FEAT * batch)
        dta = _alloc(lib, ctx, 4 * batch)
        dsc = _alloc(lib, ctx, sb)
        dgr = _alloc(lib, ctx, 4 * self.nparam)
        dlo = _alloc(lib, ctx, 4)
        hist = []
        try:
            for it in range(steps):
                idx = rng.choice(n, batch, replace=False) if batch < n else np.arange(n)
                self._h2d(dfe.value, feats[idx])
                self._h2d(dta.value, targets[idx])
                self._check(lib.dsph_mlp_grad(ctx, C.byref(self.m), batch, dfe, dta, dsc, sb,
                                              dgr, dlo))
                self.step += 1
                self._check(lib.dsph_mlp_adam(ctx, C.byref(self.m), dgr, lr, 0.9, 0.999, 1e-8,
                                              self.step))
                l = self._d2h(dlo.value, np.empty(1, np.float32))
                hist.append(float(l[0]))
                if log_every and it % log_every == 0:
                    print(f"  step {it} loss {hist[-1]:.5f}", flush=True)
        finally:
            for b in (dfe, dta, dsc, dgr, dlo):
                lib.dsph_free(ctx, b)
        return hist

    def save(self, path):
        np.savez(path, params=self.params(), feat_mean=np.array(self.m.feat_mean),
                 feat_scale=np.array(self.m.feat_scale), out_scale=self.m.out_scale)

    @classmethod
    def load(cls, lib, ctx, path):
        d = np.load(path, allow_pickle=False)
        return cls(lib, ctx, d["params"], d["feat_mean"], d["feat_scale"], float(d["out_scale"]))

    def free(self):
        for b in self._bufs:
            self.lib.dsph_free(self.ctx, b)
        self._bufs = []


def init_params(seed: int = 0) -> np.ndarray:
    """Glorot-uniform weights, zero biases; block layout of dsph_mlp."""
    rng = np.random.default_rng(seed)
    w1 = np.zeros((64, 16), np.float32)
    w1[:, :N_FEAT] = rng.uniform(-1, 1, (64, N_FEAT)) * math.sqrt(6 / (N_FEAT + 64))
    w2 = rng.uniform(-1, 1, (64, 64)).astype(np.float32) * math.sqrt(6 / 128)
    w3 = rng.uniform(-1, 1, 64).astype(np.float32) * math.sqrt(6 / 65)
    return np.concatenate([w1.ravel(), np.zeros(64, np.float32), w2.ravel(),
                           np.zeros(64, np.float32), w3, np.zeros(1, np.float32)]).astype(np.float32)


def smagorinsky_nu(grad: np.ndarray, dp: float) -> np.ndarray:
    """The reference's SPS eddy viscosity smag*|S| with smag = (2/3)(0.12 dp)^2
    (JSph.cpp:1528-1532, JSphGpu_ker.cu:1387-1436), |S| = sqrt(2 S:S)."""
    g = grad.reshape(-1, 3, 3)
    s = 0.5 * (g + g.transpose(0, 2, 1))
    return (2.0 / 3.0) * (0.12 * dp) ** 2 * np.sqrt(2.0 * np.einsum("nij,nij->n", s, s))


def targets_from_moments(mom: np.ndarray, dp_coarse: float):
    """Features [n,11] and nu* [n] from the 23 filter moments per probe;
    probes with little kernel support (free surface) are dropped."""
    w = mom[:, 0]
    ok = w > 0.9
    mom = mom[ok]
    w = w[ok]
    u = mom[:, 1:4] / w[:, None]
    uu = mom[:, 4:10] / w[:, None]
    rho = mom[:, 10] / w
    gu = mom[:, 11:20].reshape(-1, 3, 3)
    gw = mom[:, 20:23]
    g = gu - u[:, :, None] * gw[:, None, :]      # dv_a/dx_b
    tau = np.empty((len(u), 3, 3))
    comp = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for k, (a, b) in enumerate(comp):
        tau[:, a, b] = tau[:, b, a] = uu[:, k] - u[:, a] * u[:, b]
    s = 0.5 * (g + g.transpose(0, 2, 1))
    eye = np.eye(3)[None]
    sd = s - np.trace(s, axis1=1, axis2=2)[:, None, None] / 3 * eye
    td = tau - np.trace(tau, axis1=1, axis2=2)[:, None, None] / 3 * eye
    ss = np.einsum("nij,nij->n", sd, sd)
    nu = -np.einsum("nij,nij->n", td, sd) / np.maximum(2 * ss, 1e-30)
    nu = np.where(ss > 1e-12, np.maximum(nu, 0.0), 0.0)
    feats = np.concatenate([g.reshape(-1, 9), rho[:, None],
                            np.full((len(rho), 1), dp_coarse)], 1)
    return feats.astype(np.float32), nu.astype(np.float32)


def training_data(fine_case, dp_coarse: float, times, nprobe: int = 4000, seed: int = 0,
                  device: int = 0, log=print):
    """Runs the fine case on the GPU and filters it at the coarse scale at
    ``times``; returns (features, targets, index of the time of each row)."""
    from .sim import SphGpuSingle
    rng = np.random.default_rng(seed)
    hc = 1.2 * math.sqrt(3.0) * dp_coarse
    sim = SphGpuSingle(fine_case, device=device)
    feats, tgts, tidx = [], [], []
    try:
        for it, t in enumerate(times):
            sim.run_until(t)
            d = sim.download()
            fluid = d["idp"] >= fine_case.nbound
            pos = d["pos"][fluid]
            sel = rng.choice(len(pos), min(nprobe, len(pos)), replace=False)
            mom = sim.filter_probes(pos[sel], hc)
            f, y = targets_from_moments(mom, dp_coarse)
            feats.append(f)
            tgts.append(y)
            tidx.append(np.full(len(y), it))
            log(f"t={sim.stats()['timestep']:.3f}: {len(y)} samples, mean nu* {y.mean():.3e}")
    finally:
        sim.close()
    return np.concatenate(feats), np.concatenate(tgts), np.concatenate(tidx)


def normalisation(feats: np.ndarray, targets: np.ndarray):
    mean = feats.mean(0)
    std = feats.std(0)
    scale = np.where(std > 0, 1.0 / np.where(std > 0, std, 1.0), 0.0)
    out_scale = float(max(targets.mean(), 1e-12))
    return mean.astype(np.float32), scale.astype(np.float32), out_scale


# Algorithmic FLOP per sample of the network (11 -> 64 -> 64 -> 1): the
# forward products, and forward + backward for the gradient (dW1, dW2, dH1,
# dW3/dH2); element-wise activations not counted.
FLOP_FORWARD = 2 * (N_FEAT * 64 + 64 * 64 + 64)
FLOP_GRAD = FLOP_FORWARD + 2 * (64 + 64 + 64 * 64 + 64 * 64 + N_FEAT * 64)
PEAK_BF16_TFLOPS = 2500.0   # MI355X dense BF16 matrix peak


def bench(n: int = 1 << 22, reps: int = 5, device: int = 0) -> dict:
    """Times the network's kernels on n random samples: inference
    (dsph_mlp_forward) and one training gradient (dsph_mlp_grad over n
    samples in 1024 workgroups); returns samples/s and algorithmic TFLOP/s
    against the BF16 matrix peak."""
    import time
    lib = N.load()
    ctx = C.c_void_p()
    if lib.dsph_ctx_create(device, C.byref(ctx)):
        raise RuntimeError("dsph_ctx_create failed")
    rng = np.random.default_rng(0)
    feats = rng.standard_normal((n, N_FEAT)).astype(np.float32)
    targets = np.abs(rng.standard_normal(n)).astype(np.float32) * 1e-6
    net = Mlp(lib, ctx, init_params(0), np.zeros(N_FEAT), np.ones(N_FEAT), 1e-6)
    sb = 1024 * 4 * (net.nparam + 1)
    bufs = [_alloc(lib, ctx, x) for x in (feats.nbytes, targets.nbytes, 4 * n, sb,
                                          4 * net.nparam, 4)]
    df, dt_, do, dsc, dgr, dlo = bufs
    out = {}
    try:
        net._h2d(df.value, feats)
        net._h2d(dt_.value, targets)
        for name, call, flop in (
                ("forward", lambda: lib.dsph_mlp_forward(ctx, C.byref(net.m), n, df, do),
                 FLOP_FORWARD),
                ("grad", lambda: lib.dsph_mlp_grad(ctx, C.byref(net.m), n, df, dt_, dsc, sb, dgr,
                                                   dlo), FLOP_GRAD)):
            net._check(call())
            net.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                net._check(call())
            net.synchronize()
            s = (time.perf_counter() - t0) / reps
            tf = n * flop / s / 1e12
            out[name] = {"ms": 1e3 * s, "samples_per_s": n / s, "tflops": tf,
                         "frac_bf16_peak": tf / PEAK_BF16_TFLOPS, "flop_per_sample": flop}
    finally:
        for b in bufs:
            lib.dsph_free(ctx, b)
        net.free()
        lib.dsph_ctx_destroy(ctx)
    return out


if __name__ == "__main__":
    import json
    import sys
    nn = int(sys.argv[sys.argv.index("--bench") + 1]) if "--bench" in sys.argv else 1 << 22
    print(json.dumps({"samples": nn, **bench(nn)}))
