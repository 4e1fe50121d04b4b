"""Python face of the C++ host driver (``dsph_sim_*``, mirror of JSphGpuSingle).

Example::

    from dualsphysics_amd import case, sim
    c = case.dam_break_cfg0()
    s = sim.SphGpuSingle(c)
    s.run(10)                 # 10 Symplectic steps + divides, dt kept on device
    st = s.stats()
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N


class DsphError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[dsph error {code}] {msg}")
        self.code = code


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class SphGpuSingle:
    """Single-GPU SPH simulation of a :class:`dualsphysics_amd.case.Case`."""

    def __init__(self, case, device: int = 0):
        self.lib = N.load()
        self.case = case
        self._h = C.c_void_p()
        idp = np.ascontiguousarray(case.idp, np.uint32)
        pos = np.ascontiguousarray(case.pos, np.float64)
        vr = np.ascontiguousarray(case.velrho, np.float32)
        bn = None if case.boundnor is None else np.ascontiguousarray(case.boundnor, np.float32)
        self._native_case = case.to_native()
        rc = self.lib.dsph_sim_create(device, C.byref(self._native_case), case.np, _ptr(idp),
                                      _ptr(pos), _ptr(vr), None if bn is None else _ptr(bn),
                                      C.byref(self._h))
        if rc:
            msg = self._err()
            self.close()
            raise DsphError(rc, msg)

    def _err(self) -> str:
        if not self._h:
            return "simulation not created"
        return self.lib.dsph_sim_last_error(self._h).decode()

    def _check(self, rc):
        if rc:
            raise DsphError(rc, self._err())

    def close(self):
        if self._h:
            self.lib.dsph_sim_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, nsteps: int, sync_dt: bool = False):
        """Runs ``nsteps`` time steps. ``sync_dt=True`` follows the reference's
        host-side dt (one readback per interaction); otherwise dt stays on the
        device and the step has no host round trip besides the divide's counts."""
        self._check(self.lib.dsph_sim_run(self._h, nsteps, 1 if sync_dt else 0))

    def run_until(self, t: float, max_steps: int = 10_000_000) -> int:
        """Steps until the simulated time reaches ``t`` (the step that crosses
        ``t`` included, as the reference's output times); returns the number
        of steps run. Runs batches of 90 % of the steps the last dt predicts
        (dt may grow within a batch), then single steps near ``t``."""
        n = 0
        while n < max_steps:
            st = self.stats()
            if st["timestep"] >= t:
                break
            if st["dt_last"] > 0:
                k = int(0.9 * (t - st["timestep"]) / st["dt_last"]) - 1
                k = max(1, min(k, max_steps - n))
            else:
                k = 1  # dt not known before the first step
            self.run(k)
            n += k
        return n

    def save_part(self, directory, cpart: int, posdouble: bool = True):
        """Saves the current state as <directory>/Part_<cpart>.bi4 in the
        reference's format (particles in the solver's order, time, step,
        SymplecticDtPre) and Part_Head.ibi4 when missing: a restart point the
        reference's own -partbegin and :meth:`resume` both load."""
        from pathlib import Path
        from . import bi4
        d = Path(directory)
        d.mkdir(parents=True, exist_ok=True)
        if not (d / "Part_Head.ibi4").exists():
            bi4.write_part_head(d / "Part_Head.ibi4", self.case)
        st = self.stats()
        g = self.download()
        bi4.write_part(d / f"Part_{cpart:04d}.bi4", self.case, g["idp"], g["pos"], g["velrho"],
                       st["timestep"], st["nstep"], cpart,
                       symplecticdtpre=st["symplecticdtpre"], posdouble=posdouble)
        return d / f"Part_{cpart:04d}.bi4"

    @classmethod
    def resume(cls, case, part_file, device: int = 0) -> "SphGpuSingle":
        """A simulation of ``case`` continued from a PART file (the reference's
        -partbegin; JPartsLoad4.cpp:150-290, JSph.cpp:2214-2219): particles in
        the saved order, simulated time, step count and SymplecticDtPre. The
        continued run equals the uninterrupted one for states that live in
        the PART (positions, velocity, density); Laminar+SPS stresses, the
        Verlet previous state and turned mDBC normals restart as in the
        reference (zero stresses, VelrhoM1 = Velrho, the case's normals)."""
        import dataclasses
        from . import bi4
        pt = bi4.read_part(part_file, sort_by_id=False)
        vr = np.concatenate([pt["vel"], pt["rho"][:, None]], 1).astype(np.float32)
        nf = int(np.count_nonzero(pt["idp"] >= case.nbound))
        c2 = dataclasses.replace(case, idp=pt["idp"].astype(np.uint32), pos=pt["pos"],
                                 velrho=vr, nfluid=nf)
        s = cls(c2, device)
        dtpre = pt["symplecticdtpre"] if pt["symplecticdtpre"] is not None else 0.0
        s._check(s.lib.dsph_sim_set_time(s._h, float(pt["time"]), float(dtpre),
                                         int(pt["step"] or 0)))
        return s

    def stats(self) -> dict:
        st = N.DsphStats()
        self._check(self.lib.dsph_sim_stats(self._h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in N.DsphStats._fields_}

    def download(self):
        """Current particles in sorted order: dict of idp, pos, velrho, code
        (the periodic copies left out, as the reference's PART files)."""
        n = self.stats()["np"]
        idp = np.empty(n, np.uint32)
        pos = np.empty((n, 3), np.float64)
        vr = np.empty((n, 4), np.float32)
        code = np.empty(n, np.uint16)
        cnt = C.c_uint32()
        self._check(self.lib.dsph_sim_download(self._h, C.byref(cnt), idp.ctypes.data_as(N.u32p),
                                               _ptr(pos), _ptr(vr), _ptr(code)))
        keep = (code & 0xe000) != 0x2000
        if not keep.all():
            idp, pos, vr, code = idp[keep], pos[keep], vr[keep], code[keep]
        return {"idp": idp, "pos": pos, "velrho": vr, "code": code}

    def _keep(self, n: int):
        """Mask of the rows download() returns (periodic copies left out), or
        None when every row is kept."""
        code = np.empty(n, np.uint16)
        cnt = C.c_uint32()
        self._check(self.lib.dsph_sim_download(self._h, C.byref(cnt), None, None, None,
                                               _ptr(code)))
        keep = (code & 0xe000) != 0x2000
        return None if keep.all() else keep

    def forces(self, with_copies: bool = False):
        """ace (float3) and ar of the last interaction, sorted order, row for
        row as download() (periodic copies left out unless ``with_copies``)."""
        n = self.stats()["np"]
        ace = np.empty((n, 3), np.float32)
        ar = np.empty(n, np.float32)
        self._check(self.lib.dsph_sim_download_forces(self._h, _ptr(ace), _ptr(ar)))
        keep = None if with_copies else self._keep(n)
        if keep is not None:
            ace, ar = ace[keep], ar[keep]
        return ace, ar

    def interact(self):
        """Forces of the current state (no time advance); read with forces()."""
        self._check(self.lib.dsph_sim_interact(self._h))

    def pips(self):
        """Pair counts {real fluid, real bound, checked fluid, checked bound,
        real fluid-boundary} of the current state (JDsPips rule)."""
        res = (C.c_uint64 * 5)()
        self._check(self.lib.dsph_sim_pips(self._h, res))
        return [int(v) for v in res]

    def floating(self, cf: int = 0) -> dict:
        """State of floating body cf after the last step: centre, angles,
        linear and angular velocity and acceleration (StFloatingData)."""
        st = N.DsphFloatingState()
        self._check(self.lib.dsph_sim_floating(self._h, cf, C.byref(st)))
        return {"center": np.array(st.center[:]), "angles": np.array(st.angles[:], np.float32),
                "fvel": np.array(st.fvel[:], np.float32),
                "fomega": np.array(st.fomega[:], np.float32),
                "facelin": np.array(st.facelin[:], np.float32),
                "faceang": np.array(st.faceang[:], np.float32)}

    def set_closure(self, mlp):
        """Uses the learned eddy viscosity of ``mlp`` (a
        :class:`dualsphysics_amd.closure.Mlp` on this simulation's device) in
        place of the Smagorinsky viscosity (Laminar+SPS runs); None restores
        the Smagorinsky model. The network buffers must outlive the run."""
        self._closure = mlp
        if mlp is not None:
            mlp.synchronize()
        self._check(self.lib.dsph_sim_set_closure(self._h, None if mlp is None else C.byref(mlp.m)))

    def closure_nu(self, with_velgrad: bool = False):
        """Eddy viscosity of the last interaction (sorted order, row for row
        as download(); 0 for boundary) and, with ``with_velgrad``, the
        velocity gradient [np,9] it was evaluated on."""
        n = self.stats()["np"]
        nu = np.empty(n, np.float32)
        g = np.empty((n, 9), np.float32) if with_velgrad else None
        self._check(self.lib.dsph_sim_closure_nu(self._h, _ptr(nu), None if g is None else _ptr(g)))
        keep = self._keep(n)
        if keep is not None:
            nu = nu[keep]
            g = None if g is None else g[keep]
        return (nu, g) if with_velgrad else nu

    def filter_probes(self, probes: np.ndarray, hf: float) -> np.ndarray:
        """Filter moments [n,23] of the current fluid state at ``probes``
        (Wendland kernel of smoothing length ``hf``), see dsph_filter_probes."""
        probes = np.ascontiguousarray(probes, np.float64)
        out = np.empty((len(probes), 23), np.float64)
        self._check(self.lib.dsph_sim_filter_probes(self._h, len(probes), _ptr(probes), hf,
                                                    _ptr(out)))
        return out


def run_case(directory, name: str, out_dir, device: int = 0, timemax: float | None = None,
             timeout: float | None = None, sync_dt: bool = False, log=None) -> list:
    """Runs the case <directory>/<name>.xml + .bi4 (GenCase output) to its
    TimeMax, saving Part_0000 and then a PART every TimeOut into
    <out_dir>/data, as JSphGpuSingle::Run's main loop (JSphGpuSingle.cpp:
    1000-1040: a PART after the step that reaches the next output time,
    which is then the next multiple of TimeOut after the current time).
    Returns the paths of the PART files."""
    import math
    from pathlib import Path
    from . import bi4
    c = bi4.read_case(directory, name)
    tmax = c.params.timemax if timemax is None else timemax
    tout = c.params.timeout if timeout is None else timeout
    data = Path(out_dir) / "data"
    s = SphGpuSingle(c, device)
    try:
        parts = [s.save_part(data, 0)]
        if tout <= 0:
            tout = tmax
        tnext = tout
        while True:
            s.run_until(tnext)
            t = s.stats()["timestep"]
            parts.append(s.save_part(data, len(parts)))
            if log:
                log(f"{name}: t={t:.6f} s, {s.stats()['nstep']} steps, part {len(parts) - 1}")
            if t >= tmax:
                break
            tnext = (math.floor(t / tout + 1e-9) + 1) * tout
    finally:
        s.close()
    return parts
