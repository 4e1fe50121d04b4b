"""A-posteriori judgement of the learned SPS closure on the dam break.

A fine run (dp_fine, Laminar+SPS with the Smagorinsky model) provides both
the truth for the observables and the training data (filtered at the coarse
scale). A network is trained on the GPU from those data; then two coarse runs
(dp_coarse) of the same dam break, one with the reference's Smagorinsky
eddy viscosity and one with the network, are compared with the fine run on
the surge-front position and the probe pressure on the downstream wall.

Observables use fixed physical definitions for all resolutions: front = the
largest x of fluid within 0.06 m of the floor; pressure probes at
(1.58, 0.335, 0.02 | 0.05 | 0.10) m (Shepard interpolation with each run's
own kernel).

  python -m dualsphysics_amd.closure_eval [--fine-dp 0.004] [--coarse-dp 0.02]
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import _native as N
from . import case as casemod
from . import closure
from . import observables as obs
from .sim import SphGpuSingle

PROBES = np.array([[1.58, 0.335, 0.02], [1.58, 0.335, 0.05], [1.58, 0.335, 0.10]])
FRONT_ZMAX = 0.06


def _sample(c, sim):
    d = sim.download()
    fl = d["idp"] >= c.nbound
    front = obs.front_position(d["pos"], fl, zmax=FRONT_ZMAX)
    p = obs.probe_pressure(PROBES, d["pos"], d["velrho"][:, 3], fl, c.massfluid, c.h, c.b,
                           c.rhop0, c.gamma)
    return front, p


def _params():
    return casemod.Parameters(tvisco=N.VISCO_LAMINARSPS, visco=1e-6)


def run_observed(c, times, closure_net=None, filter_at=(), dp_coarse=None, nprobe=6000,
                 seed=0, log=print):
    """Runs case c sampling the observables at `times`; optionally filters the
    state at `filter_at` (training data at scale dp_coarse)."""
    rng = np.random.default_rng(seed)
    sim = SphGpuSingle(c)
    fronts, press, feats, tgts = [], [], [], []
    t0 = time.perf_counter()
    try:
        if closure_net is not None:
            sim.set_closure(closure_net)
        times = [round(float(t), 9) for t in times]
        filter_at = [round(float(t), 9) for t in filter_at]
        want = sorted(set(times) | set(filter_at))
        for i, t in enumerate(want):
            sim.run_until(t)
            if i % 10 == 0:
                log(f"    {c.name}: t={t:.2f} s, {sim.stats()['nstep']} steps, "
                    f"{time.perf_counter() - t0:.1f}s")
            if t in times:
                f, p = _sample(c, sim)
                fronts.append(f)
                press.append(p)
            if t in filter_at:
                d = sim.download()
                pos = d["pos"][d["idp"] >= c.nbound]
                sel = rng.choice(len(pos), min(nprobe, len(pos)), replace=False)
                hc = 1.2 * math.sqrt(3.0) * dp_coarse
                mom = sim.filter_probes(pos[sel], hc)
                x, y = closure.targets_from_moments(mom, dp_coarse)
                feats.append(x)
                tgts.append(y)
        nstep = sim.stats()["nstep"]
    finally:
        sim.close()
    log(f"  {c.name}: {c.np} particles, {nstep} steps, {time.perf_counter() - t0:.1f}s")
    out = {"front": np.array(fronts), "pressure": np.array(press), "nstep": nstep,
           "seconds": time.perf_counter() - t0}
    if feats:
        out["features"] = np.concatenate(feats)
        out["targets"] = np.concatenate(tgts)
    return out


def evaluate(fine_dp=0.004, coarse_dp=0.02, tmax=1.0, tout=0.01,
             train_times=(0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9), steps=2000, log=print,
             fine_file=None):
    """fine_file: npz of a previous fine run (observables and training data)
    to reuse; written after the fine run when it does not exist."""
    import os
    import dualsphysics_amd._native as Nn
    times = [round(k * tout, 9) for k in range(int(round(tmax / tout)) + 1)]
    train_times = [t for t in train_times if t < tmax]
    fine = casemod.dam_break(dp=fine_dp, params=_params())
    if fine_file and os.path.exists(fine_file):
        rf = dict(np.load(fine_file, allow_pickle=False))
        log(f"fine run dp={fine_dp} from {fine_file}")
    else:
        log(f"fine run dp={fine_dp}")
        rf = run_observed(fine, times, filter_at=train_times, dp_coarse=coarse_dp, log=log)
        if fine_file:
            np.savez(fine_file, **rf)
    x, y = rf["features"], rf["targets"]
    mean, scale, os_ = closure.normalisation(x, y)
    lib = Nn.load()
    import ctypes as C
    ctx = C.c_void_p()
    assert lib.dsph_ctx_create(0, C.byref(ctx)) == 0
    net = closure.Mlp(lib, ctx, closure.init_params(0), mean, scale, os_)
    try:
        hist = net.train(x, y, steps=steps, batch=4096, lr=2e-3)
        log(f"trained on {len(y)} samples: loss {hist[0]:.3f} -> {np.mean(hist[-50:]):.3f}")
        coarse = casemod.dam_break(dp=coarse_dp, params=_params())
        log(f"coarse runs dp={coarse_dp}")
        rs = run_observed(coarse, times, log=log)
        rl = run_observed(coarse, times, closure_net=net, log=log)
    finally:
        net.free()
        lib.dsph_ctx_destroy(ctx)
    t = np.array(times)
    scale_p = 1000.0 * 9.81 * 0.3
    win = t >= 0.5
    before = t <= 0.5

    def metrics(r):
        fe = np.abs(r["front"] - rf["front"]) / rf["front"]
        pm = np.abs(r["pressure"][win].mean(0) - rf["pressure"][win].mean(0)) / scale_p
        # Impact: the peak of each probe's pressure over the run.
        pk = np.abs(r["pressure"].max(0) - rf["pressure"].max(0)) / rf["pressure"].max(0)
        return {"front_max_rel": float(fe.max()), "front_mean_rel": float(fe.mean()),
                "front_mean_rel_before_impact": float(fe[before].mean()),
                "pressure_mean_diff_over_rho_g_H": [float(v) for v in pm],
                "pressure_peak_rel": [float(v) for v in pk],
                "seconds": r["seconds"], "steps": int(r["nstep"])}

    res = {"fine_dp": fine_dp, "coarse_dp": coarse_dp, "tmax": tmax,
           "fine": {"seconds": float(rf["seconds"]), "steps": int(rf["nstep"]),
                    "particles": fine.np},
           "train_samples": int(len(y)),
           "smagorinsky": metrics(rs), "learned": metrics(rl)}
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--fine-dp", type=float, default=0.004)
    ap.add_argument("--coarse-dp", type=float, default=0.02)
    ap.add_argument("--tmax", type=float, default=1.0)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--fine-file", default=None,
                    help="npz cache of the fine run (read if present, else written)")
    a = ap.parse_args(argv)
    log = lambda *m: print(*m, file=sys.stderr, flush=True)
    res = evaluate(a.fine_dp, a.coarse_dp, a.tmax, steps=a.steps, log=log,
                   fine_file=a.fine_file)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
