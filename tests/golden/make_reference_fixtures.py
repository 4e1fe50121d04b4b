#!/usr/bin/env python3
"""Writes the reference-run fixtures of tests/golden/ by running the
reference DualSPHysics CPU build (oracle/_ref, compiled from the reference's
own sources by oracle/ref.mk, IEEE arithmetic) on the case this package
generates, saving a PART file every step (-svsteps:1, FP64 positions).

Fixture ref_dambreak_dp0.03.npz (data only, no code of the reference):
  case      dp, seed and parameters to regenerate the identical initial
            particles with dualsphysics_amd.case.dam_break;
  steps     the saved steps;
  time      simulated time after each saved step (FP64);
  disp      fluid displacement from the initial position, FP32 [k, nf, 3];
  vel       fluid velocity [k, nf, 3]; rho: density of all particles [k, np];
  order     the ids in the reference's own particle order (sorted by cell)
            after each saved step [k, np].
Other arrays are in id order. A case with a floating body
(ref_dambreak_dp0.03_floating) also holds disp_float and vel_float, the
displacement (FP64) and velocity of the body's particles [k, nfloat, 3].

ref_observables_dp0.01.npz (--observables, ~1 h on 8 cores): front position
and probe pressure every 0.01 s over 1 s of the dp=0.01 dam break.

ref_forces_dambreak_dp0.03.npz (--forces): the reference solver's own
interaction results over the first 10 steps of the dp=0.03 dam break
(DensityDT=3). The reference runs 10 steps saving a PART every step; for
each saved state, oracle/_ref/ref_forces (a driver of this repository
linked with the reference's objects) resumes the reference from that PART
and writes its per-particle ace and ar after Interaction_Forces (the
predictor interaction of the next step). Arrays per step k = 1..10:
  order     ids in the reference's particle order [k, np];
  posf      fluid positions, FP64 [k, nf, 3]; vel [k, nf, 3]; rho [k, np]
            (boundaries are fixed at the case's positions);
  ace       [k, nf, 3] (fluid) and ar [k, np], in id order.

Run in the development container (the reference is not available on the
GPU hosts):  python tests/golden/make_reference_fixtures.py [--observables | --gauges |
             --forces] [names]
"""
from __future__ import annotations

import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from dualsphysics_amd import case as cm  # noqa: E402
from dualsphysics_amd import bi4  # noqa: E402

REF_EXE = ROOT / "oracle" / "_ref" / "DualSPHysics5.4CPU_ieee"
OUT = Path(__file__).resolve().parent
CASES = {
    # name: (generator, dp, nsteps, saved steps, Parameters overrides)
    "ref_dambreak_dp0.03": ("dam_break", 0.03, 10, (1, 2, 5, 10), {}),
    "ref_dambreak_dp0.03_verlet_laminar": ("dam_break", 0.03, 10, (1, 5, 10),
                                           dict(tstep=1, tvisco=3, visco=1e-4, tdensity=2)),
    # Flume with a piston wavemaker (moving boundary): DBC, and mDBC with
    # laminar+SPS viscosity (the flume configuration's options).
    "ref_flume_dp0.01_piston": ("flume", 0.01, 10, (1, 5, 10), {}),
    "ref_flume_dp0.01_piston_mdbc_sps": ("flume", 0.01, 10, (1, 5, 10),
                                         dict(tboundary=2, tvisco=2, visco=1e-6)),
    # Flap wavemakers (matrix motions, MoveMatBound): rotation sinusoid with
    # DBC; constant angular velocity with mDBC (the normals turn with the flap).
    "ref_flume_dp0.01_flap": ("flume", 0.01, 10, (1, 5, 10), dict(wavemaker="flap")),
    "ref_flume_dp0.01_flap_rot_mdbc": ("flume", 0.01, 10, (1, 5, 10),
                                       dict(wavemaker="flap_rot", tboundary=2, tvisco=2,
                                            visco=1e-6)),
    # More option combinations: Molteni-Colagrossi DDT; no DDT with
    # laminar+SPS; Verlet with mDBC and the piston.
    "ref_dambreak_dp0.03_ddt1": ("dam_break", 0.03, 10, (1, 5, 10), dict(tdensity=1)),
    "ref_dambreak_dp0.03_noddt_sps": ("dam_break", 0.03, 10, (1, 5, 10),
                                      dict(tdensity=0, tvisco=2, visco=1e-6)),
    "ref_flume_dp0.01_piston_mdbc_verlet": ("flume", 0.01, 10, (1, 5, 10),
                                            dict(tboundary=2, tstep=1, tvisco=3, visco=1e-4)),
    # Standard shifting: full, and without boundaries with the free-surface
    # threshold.
    "ref_dambreak_dp0.03_shift_full": ("dam_break", 0.03, 10, (1, 5, 10), dict(shifting=3)),
    "ref_dambreak_dp0.03_shift_nobound_tfs": ("dam_break", 0.03, 10, (1, 5, 10),
                                              dict(shifting=1, shifttfs=1.5)),
    # Molteni-Colagrossi DDT with mDBC (boundaries inside the term).
    "ref_flume_dp0.01_piston_mdbc_ddt1": ("flume", 0.01, 10, (1, 5, 10),
                                          dict(tboundary=2, tdensity=1)),
    # More combinations: mDBC2 with Verlet and the Molteni-Colagrossi DDT;
    # shifting that spares the moving piston (NoFixed) with mDBC; a boundary
    # viscosity factor.
    "ref_dambreak_dp0.03_mdbc2_verlet_ddt1": ("dam_break", 0.03, 10, (1, 5, 10),
                                              dict(tboundary=2, slipmode=2, tstep=1, tdensity=1,
                                                   tvisco=3, visco=1e-4)),
    "ref_flume_dp0.01_piston_mdbc_shift_nofixed": ("flume", 0.01, 10, (1, 5, 10),
                                                   dict(tboundary=2, shifting=2, shifttfs=1.5)),
    "ref_dambreak_dp0.03_viscbound2": ("dam_break", 0.03, 10, (1, 5, 10),
                                       dict(viscoboundfactor=2.0)),
    # Half cells (cells of size h, -cellmode:half).
    "ref_dambreak_dp0.03_halfcells": ("dam_break", 0.03, 10, (1, 5, 10), dict(cellmode=2)),
    # mDBC2 (SlipMode no-slip) on the dam break's fixed walls.
    "ref_dambreak_dp0.03_mdbc2": ("dam_break", 0.03, 10, (1, 5, 10),
                                  dict(tboundary=2, slipmode=2, tvisco=3, visco=1e-4)),
    # mDBC2 free slip (SlipMode 3); no-slip with no penetration
    # (NoPenetration, MDBC2_NoPen) under the Symplectic corrector; free slip
    # with no penetration under Verlet. The column starts moving at 1 m/s
    # towards the floor and the back wall, so that fluid particles come
    # within 0.75 of a normal length of the first wall layer (where the
    # no-penetration correction acts) within the saved steps.
    "ref_dambreak_dp0.03_mdbc2_freeslip": ("dam_break", 0.03, 40, (1, 10, 20, 40),
                                           dict(tboundary=2, slipmode=3, tvisco=3, visco=1e-4,
                                                vel0=(-1.0, 0.0, -1.0))),
    "ref_dambreak_dp0.03_mdbc2_nopen": ("dam_break", 0.03, 40, (1, 10, 20, 40),
                                        dict(tboundary=2, slipmode=2, nopenetration=1, tvisco=3,
                                             visco=1e-4, vel0=(-1.0, 0.0, -1.0))),
    "ref_dambreak_dp0.03_mdbc2_freeslip_nopen_verlet": (
        "dam_break", 0.03, 40, (1, 10, 20, 40),
        dict(tboundary=2, slipmode=3, nopenetration=1, tstep=1, tvisco=3, visco=1e-4,
             vel0=(-1.0, 0.0, -1.0))),
    # The same start without no penetration (the correction's effect).
    "ref_dambreak_dp0.03_mdbc2_moving": ("dam_break", 0.03, 40, (1, 10, 20, 40),
                                         dict(tboundary=2, slipmode=2, tvisco=3, visco=1e-4,
                                              vel0=(-1.0, 0.0, -1.0))),
    # 2-D dam break (<data2d>, particles in the plane y = 0): DBC with the
    # benchmark options; mDBC with laminar+SPS; mDBC2 no-slip under Verlet;
    # mDBC2 free slip with no penetration, the column thrown at 1 m/s
    # against the floor and the back wall.
    "ref_dambreak2d_dp0.005": ("dam_break_2d", 0.005, 40, (1, 10, 20, 40), {}),
    "ref_dambreak2d_dp0.005_mdbc_sps": ("dam_break_2d", 0.005, 40, (1, 10, 20, 40),
                                        dict(tboundary=2, tvisco=2, visco=1e-6)),
    "ref_dambreak2d_dp0.005_mdbc2_verlet": ("dam_break_2d", 0.005, 40, (1, 10, 20, 40),
                                            dict(tboundary=2, slipmode=2, tstep=1, tvisco=3,
                                                 visco=1e-4)),
    "ref_dambreak2d_dp0.005_mdbc2_freeslip_nopen": (
        "dam_break_2d", 0.005, 40, (1, 10, 20, 40),
        dict(tboundary=2, slipmode=3, nopenetration=1, tvisco=3, visco=1e-4,
             vel0=(-2.0, -2.0))),
    # 2-D shifting (free-surface factor (w - tfs)/(2 - tfs)): no boundaries
    # with the threshold, and full; 2-D channel periodic in x (DBC, and mDBC
    # with Verlet and laminar viscosity).
    "ref_dambreak2d_dp0.005_shift_nobound_tfs": ("dam_break_2d", 0.005, 40, (1, 10, 20, 40),
                                                 dict(shifting=1, shifttfs=1.5)),
    "ref_dambreak2d_dp0.005_shift_full": ("dam_break_2d", 0.005, 40, (1, 10, 20, 40),
                                          dict(shifting=3)),
    "ref_channel2d_dp0.005_periodic": ("periodic_channel_2d", 0.005, 60, (1, 30, 60),
                                       dict(periodic=1)),
    "ref_channel2d_dp0.005_periodic_mdbc_verlet": ("periodic_channel_2d", 0.005, 60, (1, 30, 60),
                                                   dict(periodic=1, tboundary=2, tstep=1,
                                                        tvisco=3, visco=1e-4)),
    # 2-D dam break with a floating box of half the water density at the
    # column's surface (in-plane motion: x, z and rotation about y), and the
    # same box thrown at 1 m/s spinning at 2 rad/s about y.
    "ref_dambreak2d_dp0.005_floating": ("dam_break_floating_2d", 0.005, 40, (1, 10, 20, 40), {}),
    "ref_dambreak2d_dp0.005_floating_spin": ("dam_break_floating_2d", 0.005, 40, (1, 10, 20, 40),
                                             dict(velini=(1.0, 0.0, -0.5),
                                                  omegaini=(0.0, 2.0, 0.0))),
    # 2-D flume: piston with mDBC + laminar+SPS; flap (rotation sinusoid,
    # DBC); constant rotation with mDBC (normals turned with the flap). The
    # mDBC cases save from step 10: after one step the fluid has moved
    # 1.6e-7 m (the residual of the hydrostatic start), and the few-ulp
    # spread of the mDBC boundary densities between FP32 builds is 2e-3 of
    # that, above the bound; from step 10 on it is below 1e-3.
    "ref_flume2d_dp0.005_piston_mdbc_sps": ("flume_2d", 0.005, 40, (10, 20, 40),
                                            dict(tboundary=2, tvisco=2, visco=1e-6)),
    "ref_flume2d_dp0.005_flap": ("flume_2d", 0.005, 40, (1, 10, 20, 40),
                                 dict(wavemaker="flap")),
    "ref_flume2d_dp0.005_flap_rot_mdbc": ("flume_2d", 0.005, 40, (10, 20, 40),
                                          dict(wavemaker="flap_rot", tboundary=2, tvisco=3,
                                               visco=1e-4)),
    # mDBC2 with moving boundaries (motion velocity and acceleration of the
    # wavemaker in the pressure clone and the no-penetration sums): the
    # piston with no-slip; a piston of 0.1 m stroke (0.79 m/s at the start,
    # so that it closes on the fluid within 0.75 of a normal length after
    # about 20 steps) with no-slip + no penetration; the flap at a constant
    # rotation with free slip.
    "ref_flume_dp0.01_piston_mdbc2": ("flume", 0.01, 40, (1, 10, 20, 40),
                                      dict(tboundary=2, slipmode=2, tvisco=3, visco=1e-4)),
    "ref_flume_dp0.01_piston_mdbc2_nopen": ("flume", 0.01, 40, (1, 10, 20, 40),
                                            dict(tboundary=2, slipmode=2, nopenetration=1,
                                                 tvisco=3, visco=1e-4, ampl=0.1)),
    "ref_flume_dp0.01_flap_rot_mdbc2_freeslip": ("flume", 0.01, 40, (1, 10, 20, 40),
                                                 dict(wavemaker="flap_rot", tboundary=2,
                                                      slipmode=3, tvisco=3, visco=1e-4)),
    # Shifting with mDBC2 (boundaries switched off have mass 0 in the sums):
    # full shifting on the dam break with no-slip; shifting that spares
    # fixed walls, with the moving piston and free slip.
    "ref_dambreak_dp0.03_mdbc2_shift_full": ("dam_break", 0.03, 10, (1, 5, 10),
                                             dict(tboundary=2, slipmode=2, tvisco=3, visco=1e-4,
                                                  shifting=3)),
    "ref_flume_dp0.01_piston_mdbc2_freeslip_shift_nofixed": (
        "flume", 0.01, 40, (1, 10, 20, 40),
        dict(tboundary=2, slipmode=3, tvisco=3, visco=1e-4, shifting=2, shifttfs=1.5)),
    # Floating box under the Verlet integrator (floating particles keep their
    # velocity in the step, RunFloating moves them): at rest, and thrown
    # spinning in 2-D; 10 steps, as the other Verlet cases (the FP32 spread
    # of these Verlet runs grows past the bounds by step 20;
    # the oracle in CPU semantics reproduces the reference bit for bit over
    # 40 steps).
    "ref_dambreak_dp0.03_floating_verlet": ("dam_break_floating", 0.03, 10, (1, 5, 10),
                                            dict(tstep=1)),
    "ref_dambreak2d_dp0.005_floating_spin_verlet": (
        "dam_break_floating_2d", 0.005, 10, (1, 5, 10),
        dict(tstep=1, velini=(1.0, 0.0, -0.5), omegaini=(0.0, 2.0, 0.0))),
    # Periodic channel with mDBC2 no-slip (the periodic copies of the walls
    # get their own modes and tangential velocities).
    "ref_channel_dp0.02_periodic_mdbc2": ("periodic_channel", 0.02, 60, (1, 30, 60),
                                          dict(periodic=1, tboundary=2, slipmode=2, tvisco=3,
                                               visco=1e-4)),
    # Channel periodic in x with the fluid moving at 2 m/s: particles leave
    # through one face and re-enter through the other (Symplectic + DBC, and
    # Verlet + mDBC + laminar viscosity).
    "ref_dambreak_dp0.03_floating": ("dam_break_floating", 0.03, 40, (1, 10, 20, 40), {}),
    "ref_channel_dp0.02_periodic": ("periodic_channel", 0.02, 60, (1, 30, 60), dict(periodic=1)),
    "ref_channel_dp0.02_periodic_mdbc_verlet": ("periodic_channel", 0.02, 60, (1, 30, 60),
                                                dict(periodic=1, tboundary=2, tstep=1, tvisco=3,
                                                     visco=1e-4)),
}


# Keys of a CASES entry that are arguments of the case generator, not Parameters.
GENERATOR_ARGS = ("wavemaker", "vel0", "velini", "omegaini", "ampl")


def run_reference(case, nsteps: int, workdir: Path) -> list[dict]:
    bi4.write_case(case, workdir / "in", "case", timemax=10.0, timeout=10.0)
    cmd = [str(REF_EXE), str(workdir / "in" / "case"), str(workdir / "out"), "-cpu",
           "-ompthreads:8", f"-nsteps:{nsteps}", "-svsteps:1", "-saveposdouble:1", "-stable:1",
           "-svres:0"] + (["-cellmode:half"] if case.params.cellmode == 2 else [])
    r = subprocess.run(cmd, cwd=workdir, capture_output=True, text=True, timeout=3600)
    if r.returncode != 0:
        raise RuntimeError(r.stdout[-4000:] + r.stderr[-2000:])
    return [bi4.read_part(workdir / "out" / "data" / f"Part_{k:04d}.bi4")
            for k in range(nsteps + 1)]


def observables_fixture(dp: float = 0.01, tmax: float = 1.0, tout: float = 0.01):
    """ref_observables_dp<dp>.npz: front position and probe pressure every
    ``tout`` s over ``tmax`` s of the dam break (cfg0 at dp=0.01), from the
    reference's PART files with dualsphysics_amd.observables."""
    from dualsphysics_amd import observables as obs
    case = cm.dam_break(dp=dp)
    with tempfile.TemporaryDirectory() as td:
        w = Path(td)
        bi4.write_case(case, w / "in", "case", timemax=tmax, timeout=tout)
        cmd = [str(REF_EXE), str(w / "in" / "case"), str(w / "out"), "-cpu", "-ompthreads:8",
               "-saveposdouble:1", "-stable:1", "-svres:0"]
        r = subprocess.run(cmd, cwd=w, capture_output=True, text=True, timeout=6 * 3600)
        if r.returncode != 0:
            raise RuntimeError(r.stdout[-4000:] + r.stderr[-2000:])
        files = sorted((w / "out" / "data").glob("Part_*.bi4"))
        times, front, press = [], [], []
        for f in files:
            pt = bi4.read_part(f)
            fluid = pt["idp"] >= case.nbound
            o = obs.sample(case, pt["pos"], pt["rho"], fluid)
            times.append(pt["time"]); front.append(o["front"]); press.append(o["pressure"])
    name = f"ref_observables_dp{dp:g}"
    np.savez_compressed(OUT / f"{name}.npz", dp=dp, seed=42, tmax=tmax, tout=tout,
                        time=np.array(times), front=np.array(front), pressure=np.array(press),
                        probes=obs.dam_break_probes(dp=dp))
    print(f"wrote {name}.npz ({len(times)} samples)")


GAUGE_DP = 0.03
GAUGE_TMAX = 0.02


def gauge_system(case):
    """The gauges of ref_gauges_dambreak_dp0.03.npz: two velocity points, two
    SWL lines, a maxz line and the force on the tank walls."""
    from dualsphysics_amd import gauges as G
    return G.GaugeSystem(case, [
        G.Velocity("v1", (0.3, 0.335, 0.1)), G.Velocity("v2", (0.05, 0.2, 0.25)),
        G.Swl("s1", (0.2, 0.335, 0.0), (0.2, 0.335, 0.4), 0.015),
        G.Swl("s2", (0.38, 0.335, 0.0), (0.38, 0.335, 0.4), 0.015),
        G.MaxZ("m1", (0.35, 0.335, 0.0), 0.4, 0.06), G.Force("f1", 0)], computedt=0.001)


def gauges_fixture():
    """ref_gauges_dambreak_dp0.03.npz: the reference's gauge CSV series
    (GaugesVel/SWL/MaxZ) over 0.02 s of the dp=0.03 dam break."""
    from dualsphysics_amd import gauges as G
    case = cm.dam_break(dp=GAUGE_DP)
    gs = gauge_system(case)
    with tempfile.TemporaryDirectory() as td:
        w = Path(td)
        bi4.write_case(case, w / "in", "case", timemax=GAUGE_TMAX, timeout=0.005, gauges=gs)
        cmd = [str(REF_EXE), str(w / "in" / "case"), str(w / "out"), "-cpu", "-ompthreads:8",
               "-saveposdouble:1", "-stable:1", "-svres:0"]
        r = subprocess.run(cmd, cwd=w, capture_output=True, text=True, timeout=3600)
        if r.returncode != 0:
            raise RuntimeError(r.stdout[-4000:] + r.stderr[-2000:])
        kinds = {G.Velocity: "Vel", G.Swl: "SWL", G.MaxZ: "MaxZ", G.Force: "Force"}
        arrays = {g.name: G.read_csv(w / "out" / f"Gauges{kinds[type(g)]}_{g.name}.csv")
                  for g in gs.gauges}
    np.savez_compressed(OUT / "ref_gauges_dambreak_dp0.03.npz", dp=GAUGE_DP, seed=42,
                        tmax=GAUGE_TMAX, **arrays)
    print("wrote ref_gauges_dambreak_dp0.03.npz", {k: v.shape for k, v in arrays.items()})


REF_FORCES = ROOT / "oracle" / "_ref" / "ref_forces"
FORCES_DP, FORCES_STEPS = 0.03, 10


def read_forces(path: Path):
    """(idp, ace [np,3], ar [np]) in the solver's order from a ref_forces file."""
    b = np.fromfile(path, dtype=np.uint8)
    n = int(b[:4].view(np.uint32)[0])
    rec = b[4:].reshape(n, 20)
    idp = rec[:, :4].copy().view(np.uint32).ravel()
    f = rec[:, 4:].copy().view(np.float32).reshape(n, 4)
    return idp, f[:, :3].copy(), f[:, 3].copy()


def forces_fixture():
    """ref_forces_dambreak_dp0.03.npz (see the module doc)."""
    if not REF_FORCES.exists():
        sys.exit(f"{REF_FORCES} missing: make -f oracle/ref.mk forces")
    case = cm.dam_break(dp=FORCES_DP)
    nb = case.nbound
    with tempfile.TemporaryDirectory() as td:
        w = Path(td)
        parts = run_reference(case, FORCES_STEPS, w)
        assert np.array_equal(parts[0]["pos"], case.pos)
        steps = list(range(1, FORCES_STEPS + 1))
        out = {k: [] for k in ("order", "posf", "vel", "rho", "ace", "ar", "time")}
        for k in steps:
            fout = w / f"forces_{k}.bin"
            cmd = [str(REF_FORCES), str(w / "in" / "case"), str(w / f"f{k}"), "-cpu",
                   "-ompthreads:8", f"-partbegin:{k}", str(w / "out" / "data"), "-stable:1",
                   "-svres:0", str(fout)]
            r = subprocess.run(cmd, cwd=w, capture_output=True, text=True, timeout=3600)
            if r.returncode != 0:
                raise RuntimeError(r.stdout[-4000:] + r.stderr[-2000:])
            idp, ace, ar = read_forces(fout)
            pt = parts[k]
            # The resumed solver keeps the PART's order (already sorted by cell).
            assert np.array_equal(idp, pt["idp_saved"]), f"step {k}: order changed on resume"
            assert np.array_equal(pt["pos"][:nb], case.pos[:nb]), "fixed boundaries moved"
            o = np.argsort(idp, kind="stable")
            out["order"].append(idp)
            out["posf"].append(pt["pos"][nb:])
            out["vel"].append(pt["vel"][nb:])
            out["rho"].append(pt["rho"])
            out["ace"].append(ace[o][nb:])
            out["ar"].append(ar[o])
            out["time"].append(pt["time"])
    np.savez_compressed(OUT / "ref_forces_dambreak_dp0.03.npz", generator=np.array("dam_break"),
                        dp=FORCES_DP, seed=42, params=np.array(repr([])), nbound=nb,
                        nfluid=case.nfluid, steps=np.array(steps),
                        **{k: np.stack(v) for k, v in out.items()})
    print(f"wrote ref_forces_dambreak_dp0.03.npz ({case.np} particles, steps {steps})")


def main():
    if "--forces" in sys.argv:
        forces_fixture()
        return
    if not REF_EXE.exists():
        sys.exit(f"{REF_EXE} missing: make -f oracle/ref.mk (needs /root/reference)")
    if "--observables" in sys.argv:
        observables_fixture()
        return
    if "--gauges" in sys.argv:
        gauges_fixture()
        return
    only = [a for a in sys.argv[1:] if not a.startswith("--")]
    for name, (gen, dp, nsteps, saved, over) in CASES.items():
        if only and name not in only:
            continue
        over = dict(over)
        genargs = {k: over.pop(k) for k in list(over) if k in GENERATOR_ARGS}
        params = cm.Parameters(**over)
        case = getattr(cm, gen)(dp=dp, params=params, **genargs)
        with tempfile.TemporaryDirectory() as td:
            parts = run_reference(case, nsteps, Path(td))
        p0 = parts[0]
        assert np.array_equal(p0["idp"], case.idp)
        assert np.array_equal(p0["pos"], case.pos), "reference did not load the same particles"
        assert np.array_equal(p0["rho"], case.velrho[:, 3])
        fl = slice(case.nbound, None)
        mv = slice(case.npb - case.nmoving, case.npb)
        ft = slice(case.npb, case.nbound)
        extra = {}
        if case.nfloat:
            # Floating particles (ids [npb, nbound)): displacement and velocity.
            extra = dict(
                disp_float=np.stack([(parts[k]["pos"] - p0["pos"])[ft] for k in saved]),
                vel_float=np.stack([parts[k]["vel"][ft] for k in saved]))
        np.savez_compressed(
            OUT / f"{name}.npz", **extra,
            generator=np.array(gen), dp=dp, seed=42, params=np.array(repr(sorted(over.items()))),
            genargs=np.array(repr(sorted(genargs.items()))),
            nmoving=case.nmoving,
            nbound=case.nbound, nfluid=case.nfluid, steps=np.array(saved),
            time=np.array([parts[k]["time"] for k in saved]),
            disp=np.stack([(parts[k]["pos"] - p0["pos"])[fl].astype(np.float32) for k in saved]),
            disp_moving=np.stack([(parts[k]["pos"] - p0["pos"])[mv] for k in saved]),
            vel_moving=np.stack([parts[k]["vel"][mv] for k in saved]),
            vel=np.stack([parts[k]["vel"][fl] for k in saved]),
            order=np.stack([parts[k]["idp_saved"] for k in saved]),
            rho=np.stack([parts[k]["rho"] for k in saved]))
        print(f"wrote {name}.npz ({case.np} particles, steps {saved})")


if __name__ == "__main__":
    main()
