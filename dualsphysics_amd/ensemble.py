"""Batched ensembles of independent cases on one GPU (the dsph_ens_* C ABI).

Every case runs the Symplectic step with DBC boundaries, its own constants,
cell grid (fixed to its map) and adaptive dt; all cases advance together,
each stage of the step one launch for all of them, 24 steps per HIP graph
launch; after :meth:`Ensemble.autotune` the fluid interaction of a size
bucket may be one persistent launch whose resident waves pull work from
counters.
A case stops by itself at its end time. Results of a case equal a single
:class:`dualsphysics_amd.sim.SphGpuSingle` run of that case with
``celldomfixed=1`` and device-side dt, bit for bit.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _native as N
from . import case as casemod

VARIANTS = ("tile-ring8", "tile-ring12", "tile-ring16", "rows64")
BUCKETS = ("fluid<16384", "fluid<131072", "fluid>=131072")


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class EnsembleError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"dsph error {code}: {msg}")
        self.code = code


class Ensemble:
    """Many cases on one GPU. ``tend``: end time per case (or one value)."""

    def __init__(self, cases, tend, device: int = 0):
        self.lib = N.load()
        self.cases = list(cases)
        n = len(self.cases)
        tend = np.broadcast_to(np.asarray(tend, np.float64), (n,)).copy()
        natives = (N.DsphCase * n)()
        for k, c in enumerate(self.cases):
            natives[k] = c.to_native()
        self._natives = natives
        npk = np.array([c.np for c in self.cases], np.uint32)
        idp = np.ascontiguousarray(np.concatenate([c.idp for c in self.cases]), np.uint32)
        pos = np.ascontiguousarray(np.concatenate([c.pos for c in self.cases]), np.float64)
        vr = np.ascontiguousarray(np.concatenate([c.velrho for c in self.cases]), np.float32)
        self._h = C.c_void_p()
        rc = self.lib.dsph_ens_create(device, n, natives, _ptr(npk), _ptr(idp), _ptr(pos), _ptr(vr),
                                      _ptr(tend), C.byref(self._h))
        if rc:
            msg = self._err()
            self.close()
            raise EnsembleError(rc, msg)
        self.tend = tend

    def _err(self) -> str:
        return self.lib.dsph_ens_last_error(self._h).decode() if self._h else "not created"

    def _check(self, rc):
        if rc:
            raise EnsembleError(rc, self._err())

    def close(self):
        if self._h:
            self.lib.dsph_ens_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, max_steps: int = 1 << 30, check_every: int = 8, graph: bool = True) -> int:
        """Runs until every case has reached its end time (or max_steps)."""
        self._check(self.lib.dsph_ens_set_graph(self._h, 1 if graph else 0))
        steps = C.c_uint32()
        self._check(self.lib.dsph_ens_run(self._h, min(max_steps, 0xffffffff), check_every,
                                          C.byref(steps)))
        return steps.value

    def state(self, i: int) -> dict:
        st = N.DsphEnsState()
        self._check(self.lib.dsph_ens_case_state(self._h, i, C.byref(st)))
        return {f: getattr(st, f) for f, _ in N.DsphEnsState._fields_}

    def download(self, i: int) -> dict:
        """Current particles of case i in sorted order (idp, pos, velrho, code)."""
        n0 = self.cases[i].np
        idp = np.empty(n0, np.uint32)
        pos = np.empty((n0, 3), np.float64)
        vr = np.empty((n0, 4), np.float32)
        code = np.empty(n0, np.uint16)
        cnt = C.c_uint32()
        self._check(self.lib.dsph_ens_download(self._h, i, C.byref(cnt), _ptr(idp), _ptr(pos),
                                               _ptr(vr), _ptr(code)))
        n = cnt.value
        return {"idp": idp[:n], "pos": pos[:n], "velrho": vr[:n], "code": code[:n]}

    def autotune(self, reps: int = 3):
        """Chooses the fluid-kernel variant per size bucket; returns
        ({bucket: variant}, ms[bucket][variant])."""
        var = np.empty(3, np.int32)
        ms = np.empty(12, np.float32)
        self._check(self.lib.dsph_ens_autotune(self._h, reps, _ptr(var), _ptr(ms)))
        chosen = {BUCKETS[b]: VARIANTS[v] for b, v in enumerate(var) if v >= 0}
        return chosen, ms.reshape(3, 4)

    def set_variant(self, bucket: int, variant: int):
        """Plain grid (one workgroup per wave of particles) with ``variant``."""
        self._check(self.lib.dsph_ens_set_variant(self._h, bucket, variant))

    def shape(self, bucket: int) -> dict:
        """Launch shape of a bucket's fluid kernel and of the boundary kernel
        (with the autotuner's times)."""
        sh = N.DsphEnsShape()
        self._check(self.lib.dsph_ens_get_shape(self._h, bucket, C.byref(sh)))
        d = {f: getattr(sh, f) for f, _ in N.DsphEnsShape._fields_}
        d["variant_name"] = VARIANTS[sh.variant] if sh.variant >= 0 else None
        return d

    def set_shape(self, bucket: int, variant: int, persistent: bool = False, chunk: int = 1,
                  waves_per_cu: int = 20, bound_block: int = 128, bound_tile: bool = False):
        sh = N.DsphEnsShape(variant=variant, persistent=int(persistent), chunk=chunk,
                            waves_per_cu=waves_per_cu, bound_block=bound_block,
                            bound_tile=int(bound_tile))
        self._check(self.lib.dsph_ens_set_shape(self._h, bucket, C.byref(sh)))

    def info(self) -> dict:
        a = np.empty(8, np.uint32)
        self._check(self.lib.dsph_ens_info(self._h, _ptr(a)))
        return {"cases": int(a[0]), "particles": int(a[1]), "keys": int(a[2]),
                "cases_per_bucket": [int(v) for v in a[3:6]], "steps_per_graph": int(a[6]),
                "steps_run": int(a[7])}


def dam_break_ensemble(n: int = 512, seed: int = 2024, min_np: int = 3000,
                       max_np: int = 250_000, params: casemod.Parameters | None = None):
    """``n`` uneven 3-D dam breaks drawn from ``seed``: tank length L in
    [0.5, 2.0] m, column height H in [0.1, 0.5] m, dp in [0.008, 0.02] m
    (uniform, SplitMix64); tank width 0.67 m, column 0.25 L x 0.67 m x H, tank
    height max(1.33 H, 0.2) m. dp is then scaled so that the case has between
    ``min_np`` and ``max_np`` particles (the draws alone span ~1k-500k)."""
    u = casemod.uniform01(seed, 3 * n).reshape(n, 3)
    out = []
    for k in range(n):
        length = 0.5 + 1.5 * u[k, 0]
        hcol = 0.1 + 0.4 * u[k, 1]
        dp = 0.008 + 0.012 * u[k, 2]
        tank = (length, 0.67, max(1.33 * hcol, 0.2))
        column = (0.25 * length, 0.67, hcol)

        def count(d):
            nf = (casemod._n(column[0], d) * casemod._n(column[1], d) * casemod._n(column[2], d))
            nx, ny, nz = casemod._n(tank[0], d) + 4, casemod._n(tank[1], d) + 4, casemod._n(tank[2], d) + 2
            nw = nx * ny * nz - (nx - 4) * (ny - 4) * (nz - 2)
            return nf + nw

        for _ in range(40):
            m = count(dp)
            if m < min_np:
                dp *= 0.97
            elif m > max_np:
                dp *= 1.03
            else:
                break
        out.append(casemod.dam_break(dp=dp, tank=tank, column=column, seed=seed + 1 + k,
                                     params=params, name=f"ens{k:03d}_dp{dp:.4f}"))
    return out
