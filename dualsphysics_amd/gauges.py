"""Measurement gauges (JDsGaugeSystem / JDsGaugeItem): velocity at a point,
surface water level (SWL) along a line and the maximum fluid height near a
vertical line, evaluated on the GPU (``dsph_sim_gauge_points`` /
``dsph_sim_gauge_maxz``) for all gauges of a sampling time in one launch
each, and written as the reference's CSV files (GaugesVel_<name>.csv,
GaugesSWL_<name>.csv, GaugesMaxZ_<name>.csv; JDsGaugeItem.cpp:380-400,
675-695, 965-980).

Sampling follows JGaugeItem::Update/SetTimeStep (JDsGaugeItem.h:274,
JDsGaugeItem.cpp:239-246): a gauge is computed after the first step whose
time reaches its next compute time, which then moves to the next multiple
of computedt after that time. The reference evaluates its gauges after the
step and before the step's final cell division; here they are evaluated on
the state after the division (the same particles and positions, cell lists
up to date).

Example::

    from dualsphysics_amd import gauges, sim
    gs = gauges.GaugeSystem(case, [gauges.Velocity("v1", (0.5, 0.3, 0.1))], computedt=0.01)
    s = sim.SphGpuSingle(case)
    gauges.run(s, gs, tmax=1.0)
    gs.save_csv("out")
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
from pathlib import Path

import numpy as np


@dataclasses.dataclass
class Velocity:
    """Velocity at a point: sum of W*V*v over the fluid within 2h
    (JGaugeVelocity, JDsGaugeItem.cpp:427-478; no Shepard normalisation)."""
    name: str
    point: tuple
    computedt: float | None = None


@dataclasses.dataclass
class Swl:
    """Free-surface position along point0 -> point2 (JGaugeSwl,
    JDsGaugeItem.cpp:600-790): the line is sampled every ~pointdp; the
    surface is where the SPH mass sum first drops below masslimit
    (default 0.5 massfluid in 3-D), interpolated linearly."""
    name: str
    point0: tuple
    point2: tuple
    pointdp: float
    masslimit: float | None = None
    computedt: float | None = None

    def line(self):
        """(PointNp, PointDir) as JGaugeSwl::ConfigPoints (JDsGaugeItem.cpp:615-629)."""
        p0, p2 = np.array(self.point0, float), np.array(self.point2, float)
        dis = float(np.linalg.norm(p2 - p0))
        if dis <= 0:
            return 0, np.zeros(3)
        n = int(dis / self.pointdp)
        if dis - self.pointdp * n >= self.pointdp * 0.1:
            n += 1
        if n < 1:
            n += 1
        return n, (p2 - p0) / dis * (dis / n)


@dataclasses.dataclass
class MaxZ:
    """Highest fluid particle within distlimit (horizontally) of the vertical
    line [point0, point0 + height] (JGaugeMaxZ, JDsGaugeItem.cpp:1023-1080)."""
    name: str
    point0: tuple
    height: float
    distlimit: float
    computedt: float | None = None


@dataclasses.dataclass
class Force:
    """Pressure force of the fluid on the boundary block ``mkbound`` (0 = the
    fixed walls, 1 = the moving object here; JGaugeForce,
    JDsGaugeItem.cpp:1847-1905): massbound times the sum over its particles
    of -(P1+P2)/(rho1 rho2) massf gradW."""
    name: str
    mkbound: int
    computedt: float | None = None

    def ids(self, case):
        nfix = case.nbound - case.nmoving
        if self.mkbound == 0:
            return 0, nfix
        if self.mkbound == 1 and case.nmoving:
            return nfix, case.nmoving
        raise ValueError(f"no boundary block mkbound={self.mkbound}")


def swl_surface(masses: np.ndarray, point0, pointdir, masslimit: float) -> np.ndarray:
    """Surface point from the masses at point0 + k*pointdir, k = 0..np
    (JGaugeSwl::CalculeCpuT, JDsGaugeItem.cpp:770-784; FP32 interpolation)."""
    f32 = np.float32
    ml = f32(masslimit)
    mpre = f32(0)
    p0, d = np.array(point0, float), np.array(pointdir, float)
    ptpos = p0.copy()
    for cp, m in enumerate(np.asarray(masses, np.float32)):
        if m > ml:
            mpre = m
        if m < ml and mpre:
            fxm1 = f32(f32(ml - mpre) / f32(m - mpre)) - f32(1)
            return ptpos + d * float(fxm1)
        ptpos = ptpos + d
    npnt = len(masses) - 1
    return p0 + d * (npnt if mpre else 0)


class GaugeSystem:
    """Gauges of one case; ``computedt`` is the default sampling interval
    (the reference's default is the PART interval)."""

    def __init__(self, case, gauges, computedt: float = 0.0, computeend: float = math.inf):
        self.case = case
        self.computeend = computeend
        self.gauges = list(gauges)
        names = [g.name for g in self.gauges]
        if len(set(names)) != len(names):
            raise ValueError("gauge names must be unique")
        self.computedt = computedt
        for g in self.gauges:
            if isinstance(g, Swl) and g.masslimit is None:
                g.masslimit = 0.5 * case.massfluid
        self.results = {g.name: [] for g in self.gauges}
        self._next = {g.name: 0.0 for g in self.gauges}

    def _dt(self, g):
        return self.computedt if g.computedt is None else g.computedt

    def due(self, t: float):
        """Gauges to compute at time t: t reached the next compute time and
        lies in the compute window (JGaugeItem::Update, JDsGaugeItem.h:274)."""
        return [g for g in self.gauges if t >= self._next[g.name] and t <= self.computeend]

    def compute(self, s, t: float | None = None, force: bool = False):
        """Evaluates the gauges due at the simulation's current time (all of
        them with ``force``) on the GPU and records the results."""
        from . import _native as N
        t = s.stats()["timestep"] if t is None else t
        gs = self.gauges if force else self.due(t)
        if not gs:
            return
        pts, slices = [], {}
        for g in gs:
            if isinstance(g, Velocity):
                slices[g.name] = (len(pts), 1)
                pts.append(np.array(g.point, float))
            elif isinstance(g, Swl):
                n, d = g.line()
                p0 = np.array(g.point0, float)
                slices[g.name] = (len(pts), n + 1)
                pts.extend(p0 + d * k for k in range(n + 1))
        vel = mass = None
        if pts:
            P = np.ascontiguousarray(np.array(pts), np.float64)
            vel = np.empty((len(pts), 3), np.float32)
            mass = np.empty(len(pts), np.float32)
            s._check(s.lib.dsph_sim_gauge_points(s._h, len(pts), P.ctypes.data_as(C.c_void_p),
                                                 vel.ctypes.data_as(C.c_void_p),
                                                 mass.ctypes.data_as(C.c_void_p)))
        mz = [g for g in gs if isinstance(g, MaxZ)]
        zmax = None
        if mz:
            P = np.ascontiguousarray([g.point0 for g in mz], np.float64)
            HD = np.ascontiguousarray([(g.height, g.distlimit) for g in mz], np.float32)
            zmax = np.empty(len(mz), np.float32)
            s._check(s.lib.dsph_sim_gauge_maxz(s._h, len(mz), P.ctypes.data_as(C.c_void_p),
                                               HD.ctypes.data_as(C.c_void_p),
                                               zmax.ctypes.data_as(C.c_void_p)))
        forces = {}
        for g in gs:
            if isinstance(g, Force):
                i0, cnt = g.ids(self.case)
                ace = np.empty((cnt, 3), np.float32)
                s._check(s.lib.dsph_sim_gauge_force(s._h, i0, cnt, ace.ctypes.data_as(C.c_void_p)))
                acc = np.zeros(3, np.float32)
                for a in ace:            # FP32 sum in id order, as the reference
                    acc = acc + a
                forces[g.name] = acc * np.float32(self.case.massbound)
        for g in gs:
            if isinstance(g, Force):
                self.results[g.name].append((t, forces[g.name]))
            elif isinstance(g, Velocity):
                i, _ = slices[g.name]
                self.results[g.name].append((t, vel[i].copy()))
            elif isinstance(g, Swl):
                i, n = slices[g.name]
                _, d = g.line()
                self.results[g.name].append((t, swl_surface(mass[i:i + n], g.point0, d,
                                                            g.masslimit)))
            else:
                self.results[g.name].append((t, float(zmax[mz.index(g)])))
            dt = self._dt(g)
            if dt > 0:   # next compute time: the first multiple of dt after t
                k = math.floor(t / dt)
                nxt = dt * k
                self._next[g.name] = nxt if nxt > t else dt * (k + 1)
            else:
                self._next[g.name] = t

    def next_time(self) -> float:
        return min(self._next.values()) if self._next else math.inf

    def save_csv(self, directory):
        """GaugesVel_/GaugesSWL_/GaugesMaxZ_<name>.csv in the reference's
        columns (';' separated, %g)."""
        d = Path(directory)
        d.mkdir(parents=True, exist_ok=True)
        g3 = lambda v: ";".join(f"{float(np.float32(x)):g}" for x in v)
        for g in self.gauges:
            rows = self.results[g.name]
            if isinstance(g, Velocity):
                head = "time [s];velx [m/s];vely [m/s];velz [m/s];posx [m];posy [m];posz [m]"
                lines = [f"{t:g};{g3(v)};{g3(g.point)}" for t, v in rows]
                f = d / f"GaugesVel_{g.name}.csv"
            elif isinstance(g, Swl):
                head = ("time [s];swlx [m];swly [m];swlz [m];pos0x [m];pos0y [m];pos0z [m];"
                        "pos2x [m];pos2y [m];pos2z [m]")
                lines = [f"{t:g};{g3(v)};{g3(g.point0)};{g3(g.point2)}" for t, v in rows]
                f = d / f"GaugesSWL_{g.name}.csv"
            elif isinstance(g, Force):
                head = "time [s];force [N];forcex [N];forcey [N];forcez [N]"
                lines = [f"{t:g};{float(np.float32(np.sqrt(np.float32(np.dot(v, v))))):g};{g3(v)}"
                         for t, v in rows]
                f = d / f"GaugesForce_{g.name}.csv"
            else:
                head = "time [s];zmax [m];posx [m];posy [m];posz [m]"
                lines = [f"{t:g};{float(np.float32(z)):g};{g3(g.point0)}" for t, z in rows]
                f = d / f"GaugesMaxZ_{g.name}.csv"
            f.write_text("\n".join([head] + lines) + "\n")


def read_csv(path) -> np.ndarray:
    """Numeric rows of a gauge CSV (ours or the reference's)."""
    rows = []
    for line in Path(path).read_text().splitlines()[1:]:
        if line.strip():
            rows.append([float(x) for x in line.split(";") if x.strip() != ""])
    return np.array(rows)


def read_xml(case, xml_path) -> GaugeSystem:
    """Gauges of <case><execution><special><gauges> (JGaugeSystem::ReadXml,
    JDsGaugeSystem.cpp:244-310): velocity, swl and maxz items with their
    computedt (default: the case's TimeOut)."""
    import xml.etree.ElementTree as ET
    root = ET.parse(xml_path).getroot()
    gl = root.find("execution/special/gauges")
    if gl is None:
        return GaugeSystem(case, [])
    v3 = lambda e: (float(e.get("x")), float(e.get("y")), float(e.get("z")))
    fval = lambda e, tag, a="value": (float(e.find(tag).get(a))
                                      if e.find(tag) is not None else None)
    dflt = gl.find("default")
    cdt = (fval(dflt, "computedt") if dflt is not None else None)
    if cdt is None:
        cdt = case.params.timeout or 0.0
    out = []
    for e in gl:
        if e.tag in ("default",) or e.tag.startswith("_"):
            continue
        name = e.get("name")
        cd = fval(e, "computedt")
        if e.tag == "velocity":
            out.append(Velocity(name, v3(e.find("point")), cd))
        elif e.tag == "swl":
            pe = e.find("pointdp")
            pdp = (float(pe.get("value")) if pe.get("value") is not None
                   else case.dp * float(pe.get("coefdp")))
            ml = e.find("masslimit")
            mlv = None
            if ml is not None:
                mlv = (float(ml.get("value")) if ml.get("value") is not None
                       else case.massfluid * float(ml.get("coef")))
            out.append(Swl(name, v3(e.find("point0")), v3(e.find("point2")), pdp, mlv, cd))
        elif e.tag == "force":
            out.append(Force(name, int(e.find("target").get("mkbound")), cd))
        elif e.tag == "maxz":
            dl = e.find("distlimit")
            if dl.get("value") is not None:
                dist = float(dl.get("value"))
            elif dl.get("coefdp") is not None:
                dist = case.dp * float(dl.get("coefdp"))
            else:
                dist = case.h * float(dl.get("coefh"))
            out.append(MaxZ(name, v3(e.find("point0")), fval(e, "height"), dist, cd))
        else:
            raise ValueError(f"gauge <{e.tag}> is not built")
    return GaugeSystem(case, out, computedt=cdt)


def run(s, gs: GaugeSystem, tmax: float, log=None):
    """Advances simulation ``s`` to ``tmax`` (the step that crosses it
    included), computing every gauge on the initial state
    (JSphGpuSingle::RunFirstGaugeSystem, JSphGpuSingle.cpp:913-916) and then
    each gauge after the step that reaches its next compute time, within the
    compute window that ends at tmax (the reference's default end, TimeMax)."""
    if gs.computeend == math.inf:
        gs.computeend = tmax
    gs.compute(s, s.stats()["timestep"], force=True)
    while True:
        t = s.stats()["timestep"]
        if t >= tmax:
            break
        target = min(gs.next_time(), tmax)
        if target <= t:
            s.run(1)
        else:
            s.run_until(target)
        gs.compute(s)
        if log:
            log(f"t={s.stats()['timestep']:.6f}")
