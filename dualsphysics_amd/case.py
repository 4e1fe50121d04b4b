"""Case front-end: box dam-break lattice generator and execution parameters.

GenCase (the reference's closed case generator) is not available, so the
initial particles are defined here and fed identically to every solver that
is compared (this package, the oracle, the reference CPU build):

- fluid lattice ``x=(i+1/2)dp`` filling ``[0,Lx_col]x[0,Ly]x[0,H]``;
- fixed DBC walls: ``layers`` lattice layers outside the tank interior
  ``[0,Lx]x[0,Ly]x[0,Lz]`` on the bottom and the four sides (open top);
- every coordinate jittered by ``U(-1e-4 dp, +1e-4 dp)`` from SplitMix64
  (seed 42, one draw per component in particle-id order, boundary first);
- hydrostatic fluid density ``rho0 (1 + rho0 g (H - z)/B)^(1/gamma)``,
  boundary ``rho0``, zero velocity.

Constants follow the reference's rules (JCaseCtes.cpp:33-44, 298-302):
``h = coefh sqrt(3 dp^2)``, ``c0 = coefsound sqrt(g hswl)``,
``B = c0^2 rho0 / gamma``, ``m = rho0 dp^3``; the simulation domain is the
particle bounding box widened by ``0.05 h`` (JSph.cpp:2174-2182,
JPartsLoad4.cpp:378-399) with ``posmax.z = default + 50%``
(ResizeMapLimits, JSph.cpp:1445-1476).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math

import numpy as np

from . import _native as N

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, count: int) -> np.ndarray:
    """``count`` successive SplitMix64 outputs (Steele, Lea & Flood 2014)."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, count: int) -> np.ndarray:
    """Doubles in [0,1) from the top 53 bits of SplitMix64."""
    return (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


@dataclasses.dataclass
class Parameters:
    """The ``<execution><parameters>`` subset of the solver (JSph.cpp:663-854)."""
    tstep: int = 2                 # StepAlgorithm 1 Verlet, 2 Symplectic
    verletsteps: int = 40
    tkernel: int = N.KERNEL_WENDLAND
    tvisco: int = N.VISCO_ARTIFICIAL
    visco: float = 0.01
    viscoboundfactor: float = 1.0
    tdensity: int = N.DDT_DDT2FULL
    ddtvalue: float = 0.1
    tboundary: int = N.BC_DBC
    slipmode: int = N.SLIP_VEL0     # mDBC SlipMode 1 vel=0, 2 no-slip, 3 free slip (mDBC2)
    nopenetration: int = 0          # mDBC2 NoPenetration (SlipMode 2 or 3 only)
    shifting: int = N.SHIFT_NONE    # Shifting 0 none, 1 no boundaries, 2 no fixed, 3 full
    shiftcoef: float = -2.0
    shifttfs: float = 0.0
    periodic: int = 0               # periodic axes, bits 1 X, 2 Y, 4 Z (XPeriodicIncY ...)
    # Increments of the periodic axes [axis][component] (XPeriodicIncY/Z,
    # YPeriodicIncX/Z, ZPeriodicIncX/Y; JSph.cpp:813-819): an inclined
    # periodic axis. Each axis' own component is ignored (the solver sets it
    # to -MapRealSize, JSph.cpp:2186-2188).
    perinc: tuple = ((0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0))
    cflnumber: float = 0.2
    dtini: float = 0.0
    dtmin: float = 0.0
    rhopoutmin: float = 700.0
    rhopoutmax: float = 1300.0
    cellmode: int = 1              # 1 Full, 2 Half
    celldomfixed: int = 0
    timemax: float = 1.0
    timeout: float = 0.0           # TimeOut: time between PART files (0: none)


@dataclasses.dataclass
class Floating:
    """A floating body: the <floating> block of the case particles
    (JCaseParts.cpp:274-297). Its particles are the ids [begin, begin+count),
    inside the boundary id block after the fixed and moving ones."""
    begin: int
    count: int
    massbody: float
    masspart: float
    center: tuple
    inertia: tuple                 # 3x3 about the centre, row-major
    velini: tuple = (0.0, 0.0, 0.0)
    omegaini: tuple = (0.0, 0.0, 0.0)


@dataclasses.dataclass
class Case:
    name: str
    dp: float
    h: float
    rhop0: float
    gamma: float
    cs0: float
    b: float
    massfluid: float
    massbound: float
    gravity: tuple
    hswl: float
    params: Parameters
    idp: np.ndarray        # uint32 [np], boundary ids first
    pos: np.ndarray        # float64 [np,3]
    velrho: np.ndarray     # float32 [np,4]
    nbound: int
    nfluid: int
    maprealposmin: tuple
    maprealposmax: tuple
    boundnor: np.ndarray | None = None   # float32 [nbound,3] (mDBC), not doubled
    nmoving: int = 0                     # moving boundary ids [npb-nmoving, npb)
    motion: dict | None = None           # rectilinear sinusoid: freq, ampl, phase (3 each), start, finish
    floatings: list | None = None        # Floating bodies: ids [npb, nbound)
    data2d: bool = False                 # 2-D case: every particle at y = data2dposy
    data2dposy: float = 0.0

    @property
    def np(self) -> int:
        return int(self.idp.shape[0])

    @property
    def nfloat(self) -> int:
        """Floating particles (CaseNfloat)."""
        return sum(f.count for f in self.floatings) if self.floatings else 0

    @property
    def npb(self) -> int:
        """Fixed and moving boundary particles (CaseNpb); the floating ones
        follow them in the boundary id block."""
        return self.nbound - self.nfloat

    def to_native(self) -> N.DsphCase:
        c = N.DsphCase()
        p = self.params
        c.dp, c.kernelh, c.cteb, c.gamma = self.dp, self.h, self.b, self.gamma
        c.rhopzero, c.cs0 = self.rhop0, self.cs0
        c.massfluid, c.massbound = self.massfluid, self.massbound
        for k in range(3):
            c.gravity[k] = self.gravity[k]
            c.maprealposmin[k] = self.maprealposmin[k]
            c.maprealposmax[k] = self.maprealposmax[k]
        c.cflnumber, c.dtini, c.dtmin = p.cflnumber, p.dtini, p.dtmin
        c.tstep, c.verletsteps, c.tkernel = p.tstep, p.verletsteps, p.tkernel
        c.tvisco, c.tdensity, c.tboundary = p.tvisco, p.tdensity, p.tboundary
        c.visco, c.viscoboundfactor, c.ddtvalue = p.visco, p.viscoboundfactor, p.ddtvalue
        c.rhopoutmin, c.rhopoutmax = p.rhopoutmin, p.rhopoutmax
        c.cellmode, c.celldomfixed = p.cellmode, p.celldomfixed
        c.slipmode = p.slipmode
        c.nopenetration = int(getattr(p, "nopenetration", 0))
        c.shiftmode, c.shiftcoef, c.shifttfs = p.shifting, p.shiftcoef, p.shifttfs
        c.periactive = int(p.periodic)
        for a in range(3):
            for k in range(3):
                c.perinc[a][k] = float(p.perinc[a][k]) if a != k else 0.0
        c.simulate2d, c.simulate2dposy = int(bool(self.data2d)), float(self.data2dposy)
        c.casenp, c.casenbound = self.np, self.nbound
        c.casenfixed = self.npb - self.nmoving
        c.casenmoving, c.casenfloat, c.casenfluid = self.nmoving, self.nfloat, self.nfluid
        if self.floatings:
            arr = (N.DsphFloating * len(self.floatings))()
            for k, f in enumerate(self.floatings):
                a = arr[k]
                a.begin, a.count = int(f.begin), int(f.count)
                a.massbody, a.masspart = float(f.massbody), float(f.masspart)
                for i in range(3):
                    a.center[i] = float(f.center[i])
                    a.velini[i] = float(f.velini[i])
                    a.omegaini[i] = float(f.omegaini[i])
                for i in range(9):
                    a.inertia[i] = float(f.inertia[i])
            c.ftcount = len(self.floatings)
            c.floatings = C.cast(arr, C.POINTER(N.DsphFloating))
            c._floatings_keep = arr   # the array lives as long as the struct
        if self.nmoving and self.motion:
            m = self.motion
            kind = m.get("type", "rectsinu")
            c.motion.type = {"rectsinu": 1, "rotsinu": 2, "rot": 3}[kind]
            c.motion.idbegin, c.motion.count = self.npb - self.nmoving, self.nmoving
            if kind == "rectsinu":
                for k in range(3):
                    c.motion.freq[k] = m["freq"][k]
                    c.motion.ampl[k] = m["ampl"][k]
                    c.motion.phase[k] = m["phase"][k]
            else:
                for k in range(3):
                    c.motion.axisp1[k] = m["axisp1"][k]
                    c.motion.axisp2[k] = m["axisp2"][k]
                if kind == "rotsinu":
                    c.motion.angfreq, c.motion.angampl = m["angfreq"], m["angampl"]
                    c.motion.angphase = m["angphase"]
                else:
                    c.motion.angvel = m["angvel"]
            c.motion.start, c.motion.finish = m.get("start", 0.0), m.get("finish", 1e30)
        return c


def _lattice(nx0, nx1, ny0, ny1, nz0, nz1, dp):
    """Lattice points ((i+1/2)dp, (j+1/2)dp, (k+1/2)dp), x fastest then y, z."""
    k, j, i = np.meshgrid(np.arange(nz0, nz1), np.arange(ny0, ny1), np.arange(nx0, nx1),
                          indexing="ij")
    pts = np.stack([(i.ravel() + 0.5) * dp, (j.ravel() + 0.5) * dp, (k.ravel() + 0.5) * dp], 1)
    return pts, np.stack([i.ravel(), j.ravel(), k.ravel()], 1)


def _box_walls(nxt: int, nyt: int, nzt: int, layers: int, dp: float) -> np.ndarray:
    """Wall lattice of an open-top box with interior [0,nxt)x[0,nyt)x[0,inf):
    every point of [-layers, nxt+layers) x [-layers, nyt+layers) x
    [-layers, nzt) outside the interior, in lattice order (x fastest, then y,
    then z); built per z-plane (full plane below the floor, a ring above)."""
    j, i = np.meshgrid(np.arange(-layers, nyt + layers), np.arange(-layers, nxt + layers),
                       indexing="ij")
    i, j = i.ravel(), j.ravel()
    ring = ~((i >= 0) & (i < nxt) & (j >= 0) & (j < nyt))
    planes_i = [i] * layers + [i[ring]] * nzt
    planes_j = [j] * layers + [j[ring]] * nzt
    planes_k = [np.full(len(i), k) for k in range(-layers, 0)] + \
               [np.full(int(ring.sum()), k) for k in range(nzt)]
    ii = np.concatenate(planes_i)
    jj = np.concatenate(planes_j)
    kk = np.concatenate(planes_k)
    return np.stack([(ii + 0.5) * dp, (jj + 0.5) * dp, (kk + 0.5) * dp], 1)


def _n(length, dp):
    return int(math.floor(length / dp + 0.5))


def map_limits(pos: np.ndarray, h: float, zextra: float = 0.5, periodic: int = 0,
               dp: float = 0.0):
    """Map limits of the simulation domain from the particles (see module doc);
    a periodic axis takes a border of dp/2 (JPartsLoad4::CalculeLimitsPos,
    JPartsLoad4.cpp:389-399)."""
    border = np.full(3, float(np.float32(h)) * 0.05)
    for a in range(3):
        if periodic & (1 << a):
            border[a] = dp / 2.0
    pmin = pos.min(0) - border
    pmax = pos.max(0) + border
    pmax = pmax.copy()
    pmax[2] = pmax[2] + (pmax[2] - pmin[2]) * zextra
    return tuple(float(v) for v in pmin), tuple(float(v) for v in pmax)


def dam_break(dp: float = 0.01, tank=(1.6, 0.67, 0.4), column=(0.4, 0.67, 0.3), layers: int = 2,
              coefh: float = 1.2, coefsound: float = 20.0, rhop0: float = 1000.0,
              gamma: float = 7.0, g: float = 9.81, seed: int = 42, jitter: float = 1e-4,
              params: Parameters | None = None, name: str | None = None,
              vel0=(0.0, 0.0, 0.0)) -> Case:
    """3-D dam break in a box tank (the reference's CaseDambreak geometry without
    the obstacle): column ``column`` at the x=0 end of tank ``tank``; the fluid
    starts with velocity ``vel0`` (at rest by default).
    ``boundnor`` (mDBC normals) is the vector from each wall particle to the
    nearest point of the tank interior [0,L]x[0,W]x[0,inf)."""
    params = params or Parameters()
    nxt, nyt, nzt = _n(tank[0], dp), _n(tank[1], dp), _n(tank[2], dp)
    nxf, nyf, nzf = _n(column[0], dp), _n(column[1], dp), _n(column[2], dp)
    # Walls: every lattice point of the widened box outside the interior.
    wpts = _box_walls(nxt, nyt, nzt, layers, dp)
    fpts, _ = _lattice(0, nxf, 0, nyf, 0, nzf, dp)
    nb, nf = wpts.shape[0], fpts.shape[0]
    pos = np.concatenate([wpts, fpts], 0)
    u = uniform01(seed, 3 * pos.shape[0]).reshape(-1, 3)
    pos = pos + (2.0 * u - 1.0) * (jitter * dp)
    h = coefh * math.sqrt(3.0 * dp * dp)
    hswl = column[2]
    cs0 = coefsound * math.sqrt(g * hswl)
    b = cs0 * cs0 * rhop0 / gamma
    mass = rhop0 * dp ** 3
    velrho = np.zeros((pos.shape[0], 4), np.float32)
    velrho[:nb, 3] = rhop0
    zf = pos[nb:, 2]
    velrho[nb:, 3] = (rhop0 * (1.0 + rhop0 * g * (hswl - zf) / b) ** (1.0 / gamma)).astype(np.float32)
    velrho[nb:, :3] = np.asarray(vel0, np.float32)
    bp = pos[:nb]
    q = np.stack([np.clip(bp[:, 0], 0.0, nxt * dp), np.clip(bp[:, 1], 0.0, nyt * dp),
                  np.maximum(bp[:, 2], 0.0)], 1)
    boundnor = (q - bp).astype(np.float32)
    idp = np.arange(pos.shape[0], dtype=np.uint32)
    mn, mx = map_limits(pos, h)
    return Case(name=name or f"dambreak_dp{dp:g}", dp=dp, h=h, rhop0=rhop0, gamma=gamma,
                cs0=cs0, b=b, massfluid=mass, massbound=mass, gravity=(0.0, 0.0, -g),
                hswl=hswl, params=params, idp=idp, pos=pos, velrho=velrho, nbound=nb,
                nfluid=nf, maprealposmin=mn, maprealposmax=mx, boundnor=boundnor)


def dam_break_2d(dp: float = 0.01, tank=(1.6, 0.4), column=(0.4, 0.3), layers: int = 2,
                 coefh: float = 1.2, coefsound: float = 20.0, rhop0: float = 1000.0,
                 gamma: float = 7.0, g: float = 9.81, seed: int = 42, jitter: float = 1e-4,
                 posy: float = 0.0, params: Parameters | None = None, name: str | None = None,
                 vel0=(0.0, 0.0)) -> Case:
    """2-D dam break (the x-z section of :func:`dam_break`): every particle
    in the plane y = ``posy``; walls of ``layers`` lattice rows around the
    tank interior [0,L]x[0,inf), the column ``column`` at the x=0 end,
    starting with velocity ``vel0`` (x, z). 2-D constants as GenCase sets
    them for <data2d>: h = coefh*sqrt(2 dp^2) (JSph.cpp:1667), particle mass
    rho0*dp^2. The jitter moves x and z only."""
    params = params or Parameters()
    nxt, nzt = _n(tank[0], dp), _n(tank[1], dp)
    nxf, nzf = _n(column[0], dp), _n(column[1], dp)
    k, i = np.meshgrid(np.arange(-layers, nzt), np.arange(-layers, nxt + layers), indexing="ij")
    i, k = i.ravel(), k.ravel()
    wall = ~((i >= 0) & (i < nxt) & (k >= 0))
    wpts = np.stack([(i[wall] + 0.5) * dp, np.full(int(wall.sum()), posy),
                     (k[wall] + 0.5) * dp], 1)
    kf, jf = np.meshgrid(np.arange(nzf), np.arange(nxf), indexing="ij")
    fpts = np.stack([(jf.ravel() + 0.5) * dp, np.full(kf.size, posy), (kf.ravel() + 0.5) * dp], 1)
    nb, nf = wpts.shape[0], fpts.shape[0]
    pos = np.concatenate([wpts, fpts], 0)
    u = uniform01(seed, 2 * pos.shape[0]).reshape(-1, 2)
    pos[:, 0] += (2.0 * u[:, 0] - 1.0) * (jitter * dp)
    pos[:, 2] += (2.0 * u[:, 1] - 1.0) * (jitter * dp)
    h = coefh * math.sqrt(2.0 * dp * dp)
    hswl = column[1]
    cs0 = coefsound * math.sqrt(g * hswl)
    b = cs0 * cs0 * rhop0 / gamma
    mass = rhop0 * dp ** 2
    velrho = np.zeros((pos.shape[0], 4), np.float32)
    velrho[:nb, 3] = rhop0
    zf = pos[nb:, 2]
    velrho[nb:, 3] = (rhop0 * (1.0 + rhop0 * g * (hswl - zf) / b) ** (1.0 / gamma)).astype(np.float32)
    velrho[nb:, 0] = np.float32(vel0[0])
    velrho[nb:, 2] = np.float32(vel0[1])
    bp = pos[:nb]
    q = np.stack([np.clip(bp[:, 0], 0.0, nxt * dp), bp[:, 1], np.maximum(bp[:, 2], 0.0)], 1)
    boundnor = (q - bp).astype(np.float32)
    idp = np.arange(pos.shape[0], dtype=np.uint32)
    mn, mx = map_limits(pos, h)
    return Case(name=name or f"dambreak2d_dp{dp:g}", dp=dp, h=h, rhop0=rhop0, gamma=gamma,
                cs0=cs0, b=b, massfluid=mass, massbound=mass, gravity=(0.0, 0.0, -g),
                hswl=hswl, params=params, idp=idp, pos=pos, velrho=velrho, nbound=nb,
                nfluid=nf, maprealposmin=mn, maprealposmax=mx, boundnor=boundnor,
                data2d=True, data2dposy=posy)


def rigid_body(pos: np.ndarray, masspart: float):
    """Centre of mass and inertia tensor about it (row-major) of equal point
    masses at ``pos``: the body data GenCase derives from a floating body's
    particles."""
    p = np.asarray(pos, np.float64)
    cen = p.mean(0)
    r = p - cen
    r2 = (r * r).sum(1)
    inertia = masspart * (np.eye(3) * r2.sum() - r.T @ r)
    return tuple(float(v) for v in cen), tuple(float(v) for v in inertia.ravel())


def dam_break_floating(dp: float = 0.03, box=(0.12, 0.12, 0.12), box_center=(0.16, 0.335, 0.3),
                       rhobody: float = 500.0, velini=(0.0, 0.0, 0.0),
                       omegaini=(0.0, 0.0, 0.0), params: Parameters | None = None,
                       name: str | None = None, **kw) -> Case:
    """The dam break (``dam_break``) with a floating box of density
    ``rhobody``: the lattice points inside the box ``box`` centred at
    ``box_center`` become the body's particles (ids after the walls) and
    the fluid points there are removed. The default box sits half-immersed
    at the column's free surface, as a box of half the water density floats
    at rest. Mass per particle rhobody*dp^3; centre and inertia from the
    (jittered) particles."""
    base = dam_break(dp=dp, params=params, name=name or f"dambreak_float_dp{dp:g}",
                     jitter=0.0, **{k: v for k, v in kw.items() if k != "jitter"})
    nb = base.nbound
    lo = np.array(box_center) - 0.5 * np.array(box)
    hi = np.array(box_center) + 0.5 * np.array(box)
    # Lattice points of the tank interior inside the box.
    i0 = np.ceil(lo / dp - 0.5).astype(int)
    i1 = np.floor(hi / dp - 0.5).astype(int) + 1
    bpts, _ = _lattice(i0[0], i1[0], i0[1], i1[1], i0[2], i1[2], dp)
    fl = base.pos[nb:]
    inside = np.all((fl > lo - 0.5 * dp) & (fl < hi + 0.5 * dp), axis=1)
    fpts = fl[~inside]
    pos = np.concatenate([base.pos[:nb], bpts, fpts], 0)
    seed, jitter = kw.get("seed", 42), kw.get("jitter", 1e-4)
    u = uniform01(seed, 3 * pos.shape[0]).reshape(-1, 3)
    pos = pos + (2.0 * u - 1.0) * (jitter * dp)
    nfl, nf = bpts.shape[0], fpts.shape[0]
    velrho = np.zeros((pos.shape[0], 4), np.float32)
    velrho[:nb + nfl, 3] = base.rhop0
    zf = pos[nb + nfl:, 2]
    g = -base.gravity[2]
    velrho[nb + nfl:, 3] = (base.rhop0 * (1.0 + base.rhop0 * g * (base.hswl - zf) / base.b)
                            ** (1.0 / base.gamma)).astype(np.float32)
    masspart = rhobody * dp ** 3
    cen, inertia = rigid_body(pos[nb:nb + nfl], masspart)
    ft = Floating(begin=nb, count=nfl, massbody=masspart * nfl, masspart=masspart, center=cen,
                  inertia=inertia, velini=tuple(velini), omegaini=tuple(omegaini))
    mn, mx = map_limits(pos, base.h)
    return dataclasses.replace(base, idp=np.arange(pos.shape[0], dtype=np.uint32), pos=pos,
                               velrho=velrho, nbound=nb + nfl, nfluid=nf, maprealposmin=mn,
                               maprealposmax=mx, boundnor=None, floatings=[ft])


def dam_break_floating_2d(dp: float = 0.005, box=(0.12, 0.12), box_center=(0.16, 0.3),
                          rhobody: float = 500.0, velini=(0.0, 0.0, 0.0),
                          omegaini=(0.0, 0.0, 0.0), params: Parameters | None = None,
                          name: str | None = None, **kw) -> Case:
    """The 2-D dam break (``dam_break_2d``) with a floating box (x-z section
    ``box`` centred at ``box_center``) of density ``rhobody``, half-immersed
    at the column's surface by default; mass per particle rhobody*dp^2,
    centre and inertia from the (jittered) particles."""
    base = dam_break_2d(dp=dp, params=params, name=name or f"dambreak2d_float_dp{dp:g}",
                        jitter=0.0, **{k: v for k, v in kw.items() if k != "jitter"})
    nb = base.nbound
    posy = base.data2dposy
    lo = np.array(box_center) - 0.5 * np.array(box)
    hi = np.array(box_center) + 0.5 * np.array(box)
    i0 = np.ceil(lo / dp - 0.5).astype(int)
    i1 = np.floor(hi / dp - 0.5).astype(int) + 1
    k, i = np.meshgrid(np.arange(i0[1], i1[1]), np.arange(i0[0], i1[0]), indexing="ij")
    bpts = np.stack([(i.ravel() + 0.5) * dp, np.full(i.size, posy), (k.ravel() + 0.5) * dp], 1)
    fl = base.pos[nb:]
    xz = fl[:, [0, 2]]
    inside = np.all((xz > lo - 0.5 * dp) & (xz < hi + 0.5 * dp), axis=1)
    fpts = fl[~inside]
    pos = np.concatenate([base.pos[:nb], bpts, fpts], 0)
    seed, jitter = kw.get("seed", 42), kw.get("jitter", 1e-4)
    u = uniform01(seed, 2 * pos.shape[0]).reshape(-1, 2)
    pos[:, 0] += (2.0 * u[:, 0] - 1.0) * (jitter * dp)
    pos[:, 2] += (2.0 * u[:, 1] - 1.0) * (jitter * dp)
    nfl, nf = bpts.shape[0], fpts.shape[0]
    velrho = np.zeros((pos.shape[0], 4), np.float32)
    velrho[:nb + nfl, 3] = base.rhop0
    zf = pos[nb + nfl:, 2]
    g = -base.gravity[2]
    velrho[nb + nfl:, 3] = (base.rhop0 * (1.0 + base.rhop0 * g * (base.hswl - zf) / base.b)
                            ** (1.0 / base.gamma)).astype(np.float32)
    masspart = rhobody * dp ** 2
    cen, inertia = rigid_body(pos[nb:nb + nfl], masspart)
    ft = Floating(begin=nb, count=nfl, massbody=masspart * nfl, masspart=masspart, center=cen,
                  inertia=inertia, velini=tuple(velini), omegaini=tuple(omegaini))
    mn, mx = map_limits(pos, base.h)
    return dataclasses.replace(base, idp=np.arange(pos.shape[0], dtype=np.uint32), pos=pos,
                               velrho=velrho, nbound=nb + nfl, nfluid=nf, maprealposmin=mn,
                               maprealposmax=mx, boundnor=None, floatings=[ft])


def periodic_channel(dp: float = 0.02, length: float = 0.4, width: float = 0.2,
                     depth: float = 0.2, layers: int = 2, u0: float = 2.0, coefh: float = 1.2,
                     coefsound: float = 20.0, rhop0: float = 1000.0, gamma: float = 7.0,
                     g: float = 9.81, seed: int = 42, jitter: float = 1e-4,
                     params: Parameters | None = None, name: str | None = None) -> Case:
    """Channel periodic in x: a floor and two side walls (``layers`` thick)
    under a fluid layer of ``depth`` moving at ``u0`` along x, hydrostatic
    density. The lattice spans [0, length) in x, so that it continues across
    the periodic faces (map limits dp/2 beyond the outer particles)."""
    params = params or Parameters(periodic=1)
    if not params.periodic & 1:
        raise ValueError("periodic_channel needs periodic=1 (x)")
    nx, ny, nz = _n(length, dp), _n(width, dp), _n(depth, dp)
    k, j, i = np.meshgrid(np.arange(-layers, nz), np.arange(-layers, ny + layers),
                          np.arange(0, nx), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    wall = (k < 0) | (j < 0) | (j >= ny)
    wpts = np.stack([(i[wall] + 0.5) * dp, (j[wall] + 0.5) * dp, (k[wall] + 0.5) * dp], 1)
    fpts, _ = _lattice(0, nx, 0, ny, 0, nz, dp)
    nb, nf = wpts.shape[0], fpts.shape[0]
    pos = np.concatenate([wpts, fpts], 0)
    u = uniform01(seed, 3 * pos.shape[0]).reshape(-1, 3)
    pos = pos + (2.0 * u - 1.0) * (jitter * dp)
    h = coefh * math.sqrt(3.0 * dp * dp)
    cs0 = coefsound * math.sqrt(g * depth)
    b = cs0 * cs0 * rhop0 / gamma
    mass = rhop0 * dp ** 3
    velrho = np.zeros((pos.shape[0], 4), np.float32)
    velrho[:nb, 3] = rhop0
    velrho[nb:, 0] = u0
    zf = pos[nb:, 2]
    velrho[nb:, 3] = (rhop0 * (1.0 + rhop0 * g * (depth - zf) / b) ** (1.0 / gamma)).astype(np.float32)
    bp = pos[:nb]
    q = np.stack([bp[:, 0], np.clip(bp[:, 1], 0.0, ny * dp), np.maximum(bp[:, 2], 0.0)], 1)
    boundnor = (q - bp).astype(np.float32)
    idp = np.arange(pos.shape[0], dtype=np.uint32)
    mn, mx = map_limits(pos, h, periodic=params.periodic, dp=dp)
    return Case(name=name or f"channel_dp{dp:g}", dp=dp, h=h, rhop0=rhop0, gamma=gamma,
                cs0=cs0, b=b, massfluid=mass, massbound=mass, gravity=(0.0, 0.0, -g),
                hswl=depth, params=params, idp=idp, pos=pos, velrho=velrho, nbound=nb,
                nfluid=nf, maprealposmin=mn, maprealposmax=mx, boundnor=boundnor)


def periodic_channel_2d(dp: float = 0.01, length: float = 0.4, depth: float = 0.2,
                        layers: int = 2, u0: float = 2.0, coefh: float = 1.2,
                        coefsound: float = 20.0, rhop0: float = 1000.0, gamma: float = 7.0,
                        g: float = 9.81, seed: int = 42, jitter: float = 1e-4, posy: float = 0.0,
                        params: Parameters | None = None, name: str | None = None) -> Case:
    """2-D section of :func:`periodic_channel` (particles in the plane
    y = ``posy``): a floor ``layers`` thick under a fluid layer of ``depth``
    moving at ``u0`` along x, periodic in x, 2-D constants (h from sqrt(2 dp^2),
    mass rho0 dp^2); the jitter moves x and z only."""
    params = params or Parameters(periodic=1)
    if not params.periodic & 1:
        raise ValueError("periodic_channel_2d needs periodic=1 (x)")
    nx, nz = _n(length, dp), _n(depth, dp)
    k, i = np.meshgrid(np.arange(-layers, nz), np.arange(0, nx), indexing="ij")
    i, k = i.ravel(), k.ravel()
    wall = k < 0
    pts = np.stack([(i + 0.5) * dp, np.full(i.size, posy), (k + 0.5) * dp], 1)
    pos = np.concatenate([pts[wall], pts[~wall]], 0)
    nb, nf = int(wall.sum()), int((~wall).sum())
    u = uniform01(seed, 2 * pos.shape[0]).reshape(-1, 2)
    pos[:, 0] += (2.0 * u[:, 0] - 1.0) * (jitter * dp)
    pos[:, 2] += (2.0 * u[:, 1] - 1.0) * (jitter * dp)
    h = coefh * math.sqrt(2.0 * dp * dp)
    cs0 = coefsound * math.sqrt(g * depth)
    b = cs0 * cs0 * rhop0 / gamma
    mass = rhop0 * dp ** 2
    velrho = np.zeros((pos.shape[0], 4), np.float32)
    velrho[:nb, 3] = rhop0
    velrho[nb:, 0] = u0
    zf = pos[nb:, 2]
    velrho[nb:, 3] = (rhop0 * (1.0 + rhop0 * g * (depth - zf) / b) ** (1.0 / gamma)).astype(np.float32)
    bp = pos[:nb]
    q = np.stack([bp[:, 0], bp[:, 1], np.maximum(bp[:, 2], 0.0)], 1)
    boundnor = (q - bp).astype(np.float32)
    idp = np.arange(pos.shape[0], dtype=np.uint32)
    mn, mx = map_limits(pos, h, periodic=params.periodic, dp=dp)
    return Case(name=name or f"channel2d_dp{dp:g}", dp=dp, h=h, rhop0=rhop0, gamma=gamma,
                cs0=cs0, b=b, massfluid=mass, massbound=mass, gravity=(0.0, 0.0, -g),
                hswl=depth, params=params, idp=idp, pos=pos, velrho=velrho, nbound=nb,
                nfluid=nf, maprealposmin=mn, maprealposmax=mx, boundnor=boundnor,
                data2d=True, data2dposy=posy)


def dam_break_cfg0(**kw) -> Case:
    """3-D dam break at dp=0.01 m: 80,400 fluid + 60,248 wall particles."""
    return dam_break(dp=0.01, **kw)


def dam_break_cfg1(**kw) -> Case:
    """3-D dam break at dp=0.0025 m: 5,145,600 fluid + 934,016 wall particles."""
    return dam_break(dp=0.0025, **kw)


def flume(dp: float = 0.01, length: float = 0.6, width: float = 0.1, height: float = 0.3,
          depth: float = 0.15, ampl: float = 0.02, period: float = 0.8, phase: float = 0.0,
          layers: int = 2, coefh: float = 1.2, coefsound: float = 20.0, rhop0: float = 1000.0,
          gamma: float = 7.0, g: float = 9.81, seed: int = 42, jitter: float = 1e-4,
          params: Parameters | None = None, name: str | None = None,
          wavemaker: str = "piston", flap_deg: float = 4.0, angvel: float = 10.0) -> Case:
    """Wave flume with a piston wavemaker (the flume configuration scaled by
    its arguments): fluid ``[0,length]x[0,width]x[0,depth]``; fixed walls
    (bottom, two sides, far end, ``layers`` lattice layers, open top) extended
    behind the piston's stroke; a moving wall of ``layers`` layers at x<0 over
    the tank height driven by x(t) = ampl sin(2 pi t/period + phase).
    Ids: fixed walls, then the piston, then the fluid. ``boundnor`` (mDBC
    normals) is the vector from each wall particle to the nearest point of
    the fluid region (the piston face for piston particles).
    ``wavemaker="flap"``: the same wall instead rotates about the hinge line
    x=0, z=0 (axis along +y) by flap_deg*sin(2 pi t/period + phase) degrees
    (rotation sinusoid); ``"flap_rot"``: at a constant angular velocity of
    ``angvel`` degrees/s."""
    params = params or Parameters()
    nx, ny, nz = _n(length, dp), _n(width, dp), _n(height, dp)
    nback = int(math.ceil(ampl / dp)) + layers      # stroke behind x=0 plus the piston
    # Fixed walls: lattice of the widened box minus the interior (x from -nback).
    wpts, widx = _lattice(-nback - layers, nx + layers, -layers, ny + layers, -layers, nz, dp)
    inside = ((widx[:, 0] >= -nback - layers) & (widx[:, 0] < nx) & (widx[:, 1] >= 0) &
              (widx[:, 1] < ny) & (widx[:, 2] >= 0))
    inside &= widx[:, 0] >= -nback                      # x-end wall behind the stroke stays
    fixed = wpts[~inside]
    ppts, _ = _lattice(-layers, 0, 0, ny, 0, nz, dp)    # piston layers
    fpts, _ = _lattice(0, nx, 0, ny, 0, _n(depth, dp), dp)
    nfix, nmov, nf = fixed.shape[0], ppts.shape[0], fpts.shape[0]
    pos = np.concatenate([fixed, ppts, fpts], 0)
    u = uniform01(seed, 3 * pos.shape[0]).reshape(-1, 3)
    pos = pos + (2.0 * u - 1.0) * (jitter * dp)
    h = coefh * math.sqrt(3.0 * dp * dp)
    cs0 = coefsound * math.sqrt(g * depth)
    b = cs0 * cs0 * rhop0 / gamma
    mass = rhop0 * dp ** 3
    velrho = np.zeros((pos.shape[0], 4), np.float32)
    nb = nfix + nmov
    velrho[:nb, 3] = rhop0
    zf = pos[nb:, 2]
    velrho[nb:, 3] = (rhop0 * (1.0 + rhop0 * g * (depth - zf) / b) ** (1.0 / gamma)).astype(np.float32)
    # Normals: to the nearest point of the fluid region [xlo,length]x[0,width]x[0,inf).
    bp = pos[:nb]
    xlo = np.where(np.arange(nb) >= nfix, 0.0, -nback * dp)
    q = np.stack([np.clip(bp[:, 0], xlo, length), np.clip(bp[:, 1], 0.0, width),
                  np.maximum(bp[:, 2], 0.0)], 1)
    boundnor = (q - bp).astype(np.float32)
    idp = np.arange(pos.shape[0], dtype=np.uint32)
    mn, mx = map_limits(pos, h)
    motion = {"freq": (1.0 / period, 0.0, 0.0), "ampl": (ampl, 0.0, 0.0),
              "phase": (phase, 0.0, 0.0), "start": 0.0, "finish": 1e30}
    if wavemaker == "flap":
        motion = {"type": "rotsinu", "axisp1": (0.0, 0.0, 0.0), "axisp2": (0.0, 1.0, 0.0),
                  "angfreq": 1.0 / period, "angampl": flap_deg, "angphase": phase,
                  "start": 0.0, "finish": 1e30}
    elif wavemaker == "flap_rot":
        motion = {"type": "rot", "axisp1": (0.0, 0.0, 0.0), "axisp2": (0.0, 1.0, 0.0),
                  "angvel": angvel, "start": 0.0, "finish": 1e30}
    elif wavemaker != "piston":
        raise ValueError(f"unknown wavemaker {wavemaker!r}")
    return Case(name=name or f"flume_dp{dp:g}", dp=dp, h=h, rhop0=rhop0, gamma=gamma, cs0=cs0,
                b=b, massfluid=mass, massbound=mass, gravity=(0.0, 0.0, -g), hswl=depth,
                params=params, idp=idp, pos=pos, velrho=velrho, nbound=nb, nfluid=nf,
                maprealposmin=mn, maprealposmax=mx, boundnor=boundnor, nmoving=nmov,
                motion=motion)


def _lattice_xz(nx0, nx1, nz0, nz1, dp, posy):
    """2-D lattice points ((i+1/2)dp, posy, (k+1/2)dp), x fastest then z."""
    k, i = np.meshgrid(np.arange(nz0, nz1), np.arange(nx0, nx1), indexing="ij")
    i, k = i.ravel(), k.ravel()
    return np.stack([(i + 0.5) * dp, np.full(i.size, posy), (k + 0.5) * dp], 1), np.stack([i, k], 1)


def flume_2d(dp: float = 0.005, length: float = 0.6, height: float = 0.3, depth: float = 0.15,
             ampl: float = 0.02, period: float = 0.8, phase: float = 0.0, layers: int = 2,
             coefh: float = 1.2, coefsound: float = 20.0, rhop0: float = 1000.0,
             gamma: float = 7.0, g: float = 9.81, seed: int = 42, jitter: float = 1e-4,
             posy: float = 0.0, params: Parameters | None = None, name: str | None = None,
             wavemaker: str = "piston", flap_deg: float = 4.0, angvel: float = 10.0) -> Case:
    """2-D section of :func:`flume` (particles in the plane y = ``posy``):
    fluid [0,length]x[0,depth], fixed floor and far-end wall extended behind
    the wavemaker's stroke, a piston (rectilinear sinusoid in x) or a flap
    (rotation about the hinge line x=0, z=0, axis +y: sinusoid or constant
    angular velocity) of ``layers`` layers. Ids: fixed walls, the wavemaker,
    the fluid; 2-D constants (h from sqrt(2 dp^2), mass rho0 dp^2), jitter in
    x and z only."""
    params = params or Parameters()
    nx, nz = _n(length, dp), _n(height, dp)
    nback = int(math.ceil(ampl / dp)) + layers
    wpts, widx = _lattice_xz(-nback - layers, nx + layers, -layers, nz, dp, posy)
    inside = (widx[:, 0] >= -nback) & (widx[:, 0] < nx) & (widx[:, 1] >= 0)
    fixed = wpts[~inside]
    ppts, _ = _lattice_xz(-layers, 0, 0, nz, dp, posy)
    fpts, _ = _lattice_xz(0, nx, 0, _n(depth, dp), dp, posy)
    nfix, nmov, nf = fixed.shape[0], ppts.shape[0], fpts.shape[0]
    pos = np.concatenate([fixed, ppts, fpts], 0)
    u = uniform01(seed, 2 * pos.shape[0]).reshape(-1, 2)
    pos[:, 0] += (2.0 * u[:, 0] - 1.0) * (jitter * dp)
    pos[:, 2] += (2.0 * u[:, 1] - 1.0) * (jitter * dp)
    h = coefh * math.sqrt(2.0 * dp * dp)
    cs0 = coefsound * math.sqrt(g * depth)
    b = cs0 * cs0 * rhop0 / gamma
    mass = rhop0 * dp ** 2
    velrho = np.zeros((pos.shape[0], 4), np.float32)
    nb = nfix + nmov
    velrho[:nb, 3] = rhop0
    zf = pos[nb:, 2]
    velrho[nb:, 3] = (rhop0 * (1.0 + rhop0 * g * (depth - zf) / b) ** (1.0 / gamma)).astype(np.float32)
    bp = pos[:nb]
    xlo = np.where(np.arange(nb) >= nfix, 0.0, -nback * dp)
    q = np.stack([np.clip(bp[:, 0], xlo, length), bp[:, 1], np.maximum(bp[:, 2], 0.0)], 1)
    boundnor = (q - bp).astype(np.float32)
    idp = np.arange(pos.shape[0], dtype=np.uint32)
    mn, mx = map_limits(pos, h)
    motion = {"freq": (1.0 / period, 0.0, 0.0), "ampl": (ampl, 0.0, 0.0),
              "phase": (phase, 0.0, 0.0), "start": 0.0, "finish": 1e30}
    if wavemaker == "flap":
        motion = {"type": "rotsinu", "axisp1": (0.0, posy, 0.0), "axisp2": (0.0, posy + 1.0, 0.0),
                  "angfreq": 1.0 / period, "angampl": flap_deg, "angphase": phase,
                  "start": 0.0, "finish": 1e30}
    elif wavemaker == "flap_rot":
        motion = {"type": "rot", "axisp1": (0.0, posy, 0.0), "axisp2": (0.0, posy + 1.0, 0.0),
                  "angvel": angvel, "start": 0.0, "finish": 1e30}
    elif wavemaker != "piston":
        raise ValueError(f"unknown wavemaker {wavemaker!r}")
    return Case(name=name or f"flume2d_dp{dp:g}", dp=dp, h=h, rhop0=rhop0, gamma=gamma, cs0=cs0,
                b=b, massfluid=mass, massbound=mass, gravity=(0.0, 0.0, -g), hswl=depth,
                params=params, idp=idp, pos=pos, velrho=velrho, nbound=nb, nfluid=nf,
                maprealposmin=mn, maprealposmax=mx, boundnor=boundnor, nmoving=nmov,
                motion=motion, data2d=True, data2dposy=posy)
