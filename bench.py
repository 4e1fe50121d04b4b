#!/usr/bin/env python3
"""Benchmark of the SPH step on one MI355X.

Workload: the 3-D dam break at dp=0.0025 m (5,145,600 fluid + 934,016 DBC wall
particles; Wendland C2, Tait EOS, artificial viscosity 0.01, Fourtakas
delta-SPH 0.1 (DDT2Full), Symplectic, CFL 0.2), synthetic particles from the
case generator. A step is one Symplectic time step (two interactions, two
neighbour-list divides, predictor and corrector), dt kept on the device.

Metric: pair interactions/s, counted with the reference's rule (JDsPips: real
pairs r^2 <= (2h)^2, r^2 >= 1e-18; fluid particles against fluid and boundary,
boundary particles against fluid; two interactions per Symplectic step),
averaged over the first and last state of the timed window. steps/s reported
alongside.

Prints one JSON line on stdout; progress on stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# Algorithmic FP32 work per real pair (FMA = 2, div = 1, add/mul/cmp/max = 1;
# transcendentals not counted): fluid-fluid, fluid-boundary, boundary-fluid.
F_FF, F_FB, F_BF = 79, 61, 39
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_times(lib, ctx):
    import ctypes as C
    lib.dsph_ctx_timing_collect(ctx)
    out = {}
    i = 0
    name = C.create_string_buffer(128)
    cnt = C.c_uint64()
    ms = C.c_double()
    while lib.dsph_ctx_timing_get(ctx, i, name, 128, C.byref(cnt), C.byref(ms)) == 0:
        out[name.value.decode()] = (int(cnt.value), float(ms.value))
        i += 1
    return out


def cpu_baseline(case, steps: int):
    """The oracle (C++/OpenMP restatement of the reference CPU path) timed on
    a bounded sample of the same workload: `steps` Symplectic steps of the same
    particles, pair rule as above."""
    from oracle.oracle import OracleSim
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    o = OracleSim(case)
    tset = time.perf_counter() - t0
    p0 = o.pips()
    t0 = time.perf_counter()
    o.run(steps)
    dt = time.perf_counter() - t0
    p1 = o.pips()
    pairs = 2 * steps * ((p0[0] + p0[1]) + (p1[0] + p1[1])) / 2
    return {"value": pairs / dt, "unit": "pair interactions/s", "cores": threads, "kind": "port",
            "sample": f"{steps} Symplectic steps (+divides) of the same {case.np}-particle dam "
                      f"break on the host cores, {dt:.1f} s (setup {tset:.1f} s not timed)",
            "steps_per_s": steps / dt}


REF_FAST = ROOT / "oracle" / "_ref" / "DualSPHysics5.4CPU_fast"


def cpu_baseline_reference(case, steps: int, pairs_step: float):
    """The reference CPU solver itself (built from its own sources by
    oracle/ref.mk with its release flags: -O3 -fopenmp -ffast-math) on the
    same particles (written as its case XML + .bi4 by dualsphysics_amd/bi4.py), run for
    `steps` Symplectic steps on the host cores; time = its own "Simulation
    Runtime" (initialisation excluded). Pairs per step are those of the same
    initial state counted with the same rule on the GPU."""
    import re
    import subprocess
    import tempfile
    from dualsphysics_amd import bi4
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    with tempfile.TemporaryDirectory() as td:
        td = Path(td)
        bi4.write_case(case, td / "in", "case", timemax=100.0, timeout=100.0)
        cmd = [str(REF_FAST), str(td / "in" / "case"), str(td / "out"), "-cpu",
               f"-ompthreads:{threads}", f"-nsteps:{steps}", "-sv:none", "-svres:0"]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=1800)
    twall = time.perf_counter() - t0
    if r.returncode != 0:
        raise RuntimeError("reference CPU run failed: " + r.stdout[-2000:] + r.stderr[-1000:])
    m = re.search(r"Simulation Runtime\.*:\s*([0-9.eE+-]+)\s*sec", r.stdout)
    ms = re.search(r"Steps of simulation\.*:\s*([0-9,]+)", r.stdout)
    tsim = float(m.group(1))
    nstep = int(ms.group(1).replace(",", "")) if ms else steps
    per = tsim / nstep
    return {"value": pairs_step / per, "unit": "pair interactions/s", "cores": threads,
            "kind": "reference",
            "sample": f"{nstep} Symplectic steps (+divides) of the same {case.np}-particle dam "
                      f"break by the reference CPU solver (DualSPHysics v5.4 built from its "
                      f"sources, -O3 -fopenmp -ffast-math -march=x86-64-v3), {threads} OpenMP "
                      f"threads: Simulation Runtime {tsim:.1f} s ({twall:.0f} s with case "
                      f"writing and initialisation)",
            "steps_per_s": nstep / tsim}


def ensemble_bench(n: int, tend: float, sample: int):
    """Batched ensemble of ``n`` uneven dam breaks (seeded draws), every case
    run from t=0 to ``tend``: cases/s of the ensemble (one graph launch per 24
    steps for all cases, fluid-kernel variant and launch shape (plain grid or
    persistent) per size bucket and boundary block size from the autotuner) against the
    plain single-case path (one SphGpuSingle per case, default parameters)
    timed on a size-stratified sample of ``sample`` cases. Set-up (case
    generation, upload, first divide, graph capture) is not timed for either."""
    import numpy as np
    import torch
    from dualsphysics_amd.ensemble import Ensemble, dam_break_ensemble
    from dualsphysics_amd.sim import SphGpuSingle
    t0 = time.perf_counter()
    cases = dam_break_ensemble(n)
    sizes = np.array([c.np for c in cases])
    log(f"ensemble: {n} cases, {sizes.sum()} particles ({sizes.min()}-{sizes.max()}), "
        f"generated in {time.perf_counter() - t0:.1f}s")
    ens = Ensemble(cases, tend=tend)
    chosen, tuned = ens.autotune(reps=2)
    shapes = {}
    for b, name in enumerate(("fluid<16384", "fluid<131072", "fluid>=131072")):
        sh = ens.shape(b)
        if sh["variant"] >= 0:
            shapes[name] = {k: (round(v, 4) if isinstance(v, float) else v)
                            for k, v in sh.items()}
    log(f"  autotuner: {shapes}")
    ens.run(max_steps=0)  # graph capture
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    steps = ens.run(check_every=4)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    states = [ens.state(k) for k in range(n)]
    info = ens.info()
    ens.close()
    stopped = sum(1 for s in states if not s["active"] and s["time"] >= tend)
    part_steps = sum(float(s["np0"]) * s["nstep"] for s in states)
    log(f"  ensemble: {steps} steps in {el:.2f}s, {stopped}/{n} cases reached t={tend}")
    order = np.argsort(sizes, kind="stable")
    pick = sorted(set(order[np.linspace(0, n - 1, min(sample, n)).round().astype(int)].tolist()))
    times = []
    for k in pick:
        sim = SphGpuSingle(cases[k])
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        sim.run_until(tend)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t1)
        sim.close()
    if len(pick) == n:
        plain_total = float(np.sum(times))
        plain_se = 0.0
        how = f"all {n} cases timed"
    else:
        # Size-stratified sample: mean time x cases, with the standard error
        # of that estimate.
        plain_total = float(np.mean(times)) * n
        plain_se = float(np.std(times, ddof=1) / np.sqrt(len(times))) * n if len(times) > 1 else None
        how = f"{len(pick)} sampled cases, estimated for all {n}"
    log(f"  plain: {how}: {plain_total:.1f}s, mean {np.mean(times)*1e3:.1f} ms/case")
    return {"cases": n, "particles": int(sizes.sum()), "np_min": int(sizes.min()),
            "np_max": int(sizes.max()), "tend_s": tend, "steps": steps, "seconds": el,
            "cases_per_s": n / el, "plain_cases_per_s": n / plain_total,
            "speedup": plain_total / el, "particle_steps_per_s": part_steps / el,
            "cases_completed": stopped, "plain_sample": len(pick),
            "plain_seconds": plain_total, "plain_seconds_stderr": plain_se,
            "autotune": shapes, "steps_per_graph": info["steps_per_graph"],
            "note": "plain = one SphGpuSingle run per case to the same end time, set-up "
                    "untimed (" + how + ")"}


def flume_bench(steps: int, warmup: int, dp: float = 0.005):
    """The full-size wave-flume configuration: 12 x 1 x 1.2 m flume, water
    depth 0.6 m, piston wavemaker (amplitude 0.1 m, period 1.6 s), dp = dp,
    mDBC + laminar+SPS; ``steps`` timed Symplectic steps after ``warmup``."""
    import torch
    from dualsphysics_amd import case as casemod
    from dualsphysics_amd import _native as N
    from dualsphysics_amd.sim import SphGpuSingle
    t0 = time.perf_counter()
    log(f"flume: generating the dp={dp:g} case ...")
    prm = casemod.Parameters(tboundary=N.BC_MDBC, tvisco=N.VISCO_LAMINARSPS, visco=1e-6)
    c = casemod.flume(dp=dp, length=12.0, width=1.0, height=1.2, depth=0.6, ampl=0.1,
                      period=1.6, params=prm)
    log(f"flume: {c.nfluid} fluid + {c.nbound} boundary particles "
        f"({time.perf_counter() - t0:.1f}s to generate)")
    sim = SphGpuSingle(c)
    del c
    sim.run(warmup)
    torch.cuda.synchronize()
    p0 = sim.pips()
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        n = min(10, steps - done)
        sim.run(n)
        done += n
        log(f"  flume {done}/{steps} steps")
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    st = sim.stats()
    sim.close()
    pairs = 2.0 * (p0[0] + p0[1])
    return {"workload": f"wave_flume_12x1x1.2m_dp{dp:g}_mdbc_laminar_sps_piston",
            "particles": st["np"], "steps": steps, "ms_per_step": 1e3 * el / steps,
            "pairs_per_step": pairs, "pair_interactions_per_s": pairs * steps / el,
            "sim_time_s": st["timestep"], "dt_last": st["dt_last"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--dp", type=float, default=None,
                    help="particle spacing (default 0.0025 m in 3-D, 0.0002 m in 2-D)")
    ap.add_argument("--workload", choices=["dam_break_3d", "dam_break_2d"],
                    default="dam_break_3d",
                    help="dam_break_2d: the x-z section of the same tank and column (data2d)")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-timing", action="store_true",
                    help="skip the instrumented window (per-kernel HIP events)")
    ap.add_argument("--timed-steps", type=int, default=200,
                    help="steps of the instrumented window after the headline window")
    ap.add_argument("--forces-kernel", choices=["tile", "rows"], default="tile")
    ap.add_argument("--cellmode", choices=["full", "half"], default="full",
                    help="cell size 2h (the reference's default) or h (-cellmode:half)")
    ap.add_argument("--ensemble", type=int, default=512, metavar="N",
                    help="cases of the batched-ensemble measurement (0: skip)")
    ap.add_argument("--ens-tend", type=float, default=0.01,
                    help="physical end time of every ensemble case [s]")
    ap.add_argument("--ens-plain-sample", type=int, default=512,
                    help="cases timed on the plain single-case path (all by default)")
    ap.add_argument("--flume-steps", type=int, default=0, metavar="N",
                    help="also time N steps of the full-size wave flume (~58M fluid particles)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES (3 times): the known-size "
                         "kernel that calibrates the HBM byte counters of a profiler")
    a = ap.parse_args()
    if a.copy_probe:
        import torch
        n = a.copy_probe // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n,
                          "launches": 3}), flush=True)
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs the SPH step on one GPU only; "
                          "multi-GPU sharding is a design note (x-slab halo exchange), "
                          "not implemented", "n_gpus": a.gpus}))
        sys.exit(2)

    import torch
    from dualsphysics_amd import case as casemod
    from dualsphysics_amd import _native as N
    from dualsphysics_amd.sim import SphGpuSingle

    t0 = time.perf_counter()
    if a.dp is None:
        a.dp = 0.0025 if a.workload == "dam_break_3d" else 0.0002
    gen = casemod.dam_break if a.workload == "dam_break_3d" else casemod.dam_break_2d
    c = gen(dp=a.dp, params=casemod.Parameters(
        cellmode=2 if a.cellmode == "half" else 1))
    log(f"case: {c.nfluid} fluid + {c.nbound} boundary particles ({time.perf_counter()-t0:.1f}s)")
    sim = SphGpuSingle(c)
    lib = N.load()
    ctx = lib.dsph_sim_ctx(sim._h)
    lib.dsph_ctx_set_option(ctx, N.OPT_FORCES_KERNEL,
                            N.FORCES_TILE if a.forces_kernel == "tile" else N.FORCES_ROWS)
    p_init = sim.pips()
    log("warmup ...")
    sim.run(a.warmup)
    torch.cuda.synchronize()
    p0 = sim.pips()
    # Headline window: kernel timing off (no HIP events between launches).
    log(f"timing {a.steps} steps ...")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done = 0
    chunk = max(1, min(a.steps, 250))
    while done < a.steps:
        n = min(chunk, a.steps - done)
        sim.run(n)
        done += n
        log(f"  {done}/{a.steps} steps")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    p1 = sim.pips()
    st = sim.stats()
    pairs_step = 2 * ((p0[0] + p0[1]) + (p1[0] + p1[1])) / 2
    value = pairs_step * a.steps / elapsed
    ms_step = 1e3 * elapsed / a.steps

    # Second window with per-kernel HIP events on the context stream: kernel
    # averages and the roofline of the fluid interaction kernel.
    kt, kt_steps, ms_step_timed = {}, 0, None
    roof = None
    if not a.no_timing:
        kt_steps = max(1, min(a.steps, a.timed_steps))
        lib.dsph_ctx_timing_reset(ctx)
        lib.dsph_ctx_set_timing(ctx, 1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim.run(kt_steps)
        torch.cuda.synchronize()
        ms_step_timed = 1e3 * (time.perf_counter() - t0) / kt_steps
        kt = kernel_times(lib, ctx)
        lib.dsph_ctx_set_timing(ctx, 0)
        p2 = sim.pips()
        log(f"instrumented window: {kt_steps} steps, {ms_step_timed:.3f} ms/step")
        if "k_forces_fluid" in kt:
            cnt, ms = kt["k_forces_fluid"]
            ff = ((p1[0] - p1[4]) + (p2[0] - p2[4])) / 2
            fb = (p1[4] + p2[4]) / 2
            flops = ff * F_FF + fb * F_FB
            avg_s = ms / cnt / 1e3
            ach = flops / avg_s / 1e12
            traffic = None
            tfile = ROOT / "profiles" / (f"k_forces_fluid_traffic_dam_break_dp{a.dp:g}.json"
                                         if a.workload == "dam_break_3d" else
                                         f"k_forces_fluid_traffic_{a.workload}_dp{a.dp:g}.json")
            if tfile.exists() and a.forces_kernel == "tile":
                traffic = json.loads(tfile.read_text()).get("traffic_bytes_per_launch")
            roof = {"bound": "valu", "kernel": "k_forces_fluid", "achieved": round(ach, 3),
                    "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(ach / PEAK_FP32_TFLOPS, 4), "traffic": traffic,
                    "avg_launch_ms": round(ms / cnt, 4), "launches": cnt,
                    "flop_per_launch": flops,
                    "note": "FP32 kernel bound by the latency of its neighbour gathers "
                            "(rocprof counters in profiles/); peak = MI355X FP32 vector peak "
                            "(packed FP32); kernel time from HIP events on the context stream over a "
                            "second window of the same run (the headline window has no events); "
                            "traffic = rocprofv3 FETCH_SIZE/WRITE_SIZE per launch scaled by a "
                            "1 GiB copy calibration (profiles/)"}
    sim.close()
    del sim
    flume = flume_bench(a.flume_steps, 3) if a.flume_steps > 0 else None
    ens = None
    if a.ensemble > 0:
        ens = ensemble_bench(a.ensemble, a.ens_tend, a.ens_plain_sample)
    cpu = None
    if not a.no_cpu_baseline:
        if REF_FAST.exists():
            log(f"cpu baseline: {a.cpu_steps} steps of the reference CPU solver ...")
            cpu = cpu_baseline_reference(c, a.cpu_steps, 2.0 * (p_init[0] + p_init[1]))
        else:
            log(f"cpu baseline: {a.cpu_steps} steps of the oracle ...")
            cpu = cpu_baseline(c, a.cpu_steps)
    out = {
        "metric": "pair interactions/s + steps/s at 5M particles; ensemble cases/s vs plain",
        "value": value, "unit": "pair interactions/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms_step, "steps_per_s": 1e3 / ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{a.workload}_dp{a.dp:g}", "n_fluid": c.nfluid,
                   "n_bound": c.nbound, "dp": a.dp, "kernel": "wendland", "forces_kernel": a.forces_kernel,
                   "cellmode": a.cellmode,
                   "visco": "artificial 0.01", "ddt": "fourtakas_full 0.1",
                   "step": "symplectic", "cfl": 0.2, "pairs_per_step": pairs_step,
                   "sim_time_s": st["timestep"]},
        "roofline": roof,
        "cpu_baseline": cpu,
        "ensemble": ens,
        "flume": flume,
        "kernels_ms": {k: {"launches": v[0], "total_ms": round(v[1], 3),
                           "avg_ms": round(v[1] / max(v[0], 1), 5)} for k, v in kt.items()},
        "instrumented_window": {"steps": kt_steps, "ms_per_step": ms_step_timed},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
