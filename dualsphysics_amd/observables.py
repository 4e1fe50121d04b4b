"""Integrated observables of the dam break: surge-front position and probe
pressure, computed from particle states (FP64 positions, density).

The reference computes such quantities with its gauge system
(JDsGaugeItem.cpp, JDsGauge_ker.cu) and the closed MeasureTool; both
definitions here are this package's, applied identically to the states of
every solver that is compared (the HIP path, the oracle, the reference CPU
build's PART files):

- front position: the largest x of the fluid particles within ``zmax`` of the
  tank floor (the leading edge of the surge along the bottom);
- probe pressure: Shepard-normalised SPH interpolation of the Tait pressure
  of the fluid particles, p(x) = sum_j V_j p_j W(x - x_j) / sum_j V_j W(x - x_j)
  with V_j = m/rho_j and the Wendland C2 kernel of the case; 0 when no fluid
  particle is within 2h.
"""
from __future__ import annotations

import math

import numpy as np


def front_position(pos: np.ndarray, fluid: np.ndarray, floor_z: float = 0.0,
                   zmax: float | None = None, dp: float = 0.01) -> float:
    """Leading-edge x of the fluid within ``zmax`` (default 3 dp) of the floor."""
    zmax = 3 * dp if zmax is None else zmax
    sel = fluid & (pos[:, 2] < floor_z + zmax)
    return float(pos[sel, 0].max()) if np.any(sel) else float("nan")


def tait_pressure(rho: np.ndarray, b: float, rho0: float, gamma: float) -> np.ndarray:
    return b * ((rho.astype(np.float64) / rho0) ** gamma - 1.0)


def wendland(r: np.ndarray, h: float) -> np.ndarray:
    q = r / h
    w = 21.0 / (16.0 * math.pi * h ** 3) * (1 - 0.5 * q) ** 4 * (2 * q + 1)
    return np.where(q < 2.0, w, 0.0)


def probe_pressure(points: np.ndarray, pos: np.ndarray, rho: np.ndarray, fluid: np.ndarray,
                   mass: float, h: float, b: float, rho0: float, gamma: float) -> np.ndarray:
    """Shepard-interpolated pressure at ``points`` [k,3]."""
    sel = fluid
    x, r, m = pos[sel], rho[sel].astype(np.float64), mass
    p = tait_pressure(r, b, rho0, gamma)
    out = np.zeros(len(points))
    for k, pt in enumerate(points):
        d = np.linalg.norm(x - pt[None, :], axis=1)
        near = d < 2 * h
        if not np.any(near):
            continue
        w = wendland(d[near], h) * (m / r[near])
        s = w.sum()
        out[k] = float((w * p[near]).sum() / s) if s > 0 else 0.0
    return out


def dam_break_probes(tank=(1.6, 0.67, 0.4), dp: float = 0.01) -> np.ndarray:
    """Pressure probes on the downstream wall (the wall the surge hits),
    centred in y at heights 0.02, 0.05 and 0.10 m, one particle spacing
    inside the fluid domain."""
    x = tank[0] - dp
    y = 0.5 * tank[1]
    return np.array([[x, y, 0.02], [x, y, 0.05], [x, y, 0.10]])


def sample(case, pos: np.ndarray, rho: np.ndarray, fluid: np.ndarray,
           probes: np.ndarray | None = None) -> dict:
    """Observables of one state of a dam-break case."""
    probes = dam_break_probes(dp=case.dp) if probes is None else probes
    return {"front": front_position(pos, fluid, dp=case.dp),
            "pressure": probe_pressure(probes, pos, rho, fluid, case.massfluid, case.h, case.b,
                                       case.rhop0, case.gamma)}
