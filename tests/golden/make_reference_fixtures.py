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
  vel       fluid velocity [k, nf, 3]; rho: density of all particles [k, np].
Particles are in id order.

Run in the development container (the reference is not available on the
GPU hosts):  python tests/golden/make_reference_fixtures.py
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
from oracle import bi4  # noqa: E402

REF_EXE = ROOT / "oracle" / "_ref" / "DualSPHysics5.4CPU_ieee"
OUT = Path(__file__).resolve().parent
CASES = {
    # name: (dp, nsteps, saved steps, Parameters overrides)
    "ref_dambreak_dp0.03": (0.03, 10, (1, 2, 5, 10), {}),
    "ref_dambreak_dp0.03_verlet_laminar": (0.03, 10, (1, 5, 10),
                                           dict(tstep=1, tvisco=3, visco=1e-4, tdensity=2)),
}


def run_reference(case, nsteps: int, workdir: Path) -> list[dict]:
    bi4.write_case(case, workdir / "in", "case", timemax=10.0, timeout=10.0)
    cmd = [str(REF_EXE), str(workdir / "in" / "case"), str(workdir / "out"), "-cpu",
           "-ompthreads:8", f"-nsteps:{nsteps}", "-svsteps:1", "-saveposdouble:1", "-stable:1",
           "-svres:0"]
    r = subprocess.run(cmd, cwd=workdir, capture_output=True, text=True, timeout=3600)
    if r.returncode != 0:
        raise RuntimeError(r.stdout[-4000:] + r.stderr[-2000:])
    return [bi4.read_part(workdir / "out" / "data" / f"Part_{k:04d}.bi4")
            for k in range(nsteps + 1)]


def main():
    if not REF_EXE.exists():
        sys.exit(f"{REF_EXE} missing: make -f oracle/ref.mk (needs /root/reference)")
    for name, (dp, nsteps, saved, over) in CASES.items():
        params = cm.Parameters(**over)
        case = cm.dam_break(dp=dp, params=params)
        with tempfile.TemporaryDirectory() as td:
            parts = run_reference(case, nsteps, Path(td))
        p0 = parts[0]
        assert np.array_equal(p0["idp"], case.idp)
        assert np.array_equal(p0["pos"], case.pos), "reference did not load the same particles"
        assert np.array_equal(p0["rho"], case.velrho[:, 3])
        fl = slice(case.nbound, None)
        np.savez_compressed(
            OUT / f"{name}.npz",
            dp=dp, seed=42, params=np.array(repr(sorted(over.items()))),
            nbound=case.nbound, nfluid=case.nfluid, steps=np.array(saved),
            time=np.array([parts[k]["time"] for k in saved]),
            disp=np.stack([(parts[k]["pos"] - p0["pos"])[fl].astype(np.float32) for k in saved]),
            vel=np.stack([parts[k]["vel"][fl] for k in saved]),
            rho=np.stack([parts[k]["rho"] for k in saved]))
        print(f"wrote {name}.npz ({case.np} particles, steps {saved})")


if __name__ == "__main__":
    main()
