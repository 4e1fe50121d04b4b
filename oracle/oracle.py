"""ctypes bindings of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline. The
product (dualsphysics_amd) never imports this module.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboracle.so"
SEM_GPU, SEM_CPU, SEM_F64 = 0, 1, 2


class OrcCte(C.Structure):
    _fields_ = [(n, C.c_float) for n in (
        "massb", "massf", "kernelh", "kernelsize2", "poscellsize", "awen", "bwenh", "cs0",
        "eta2", "ddtkh", "ddtgz", "scell", "kernelsize", "dp", "cteb", "gamma", "rhopzero",
        "ovrhopzero", "movlimit")] + [("cellcode", C.c_uint32)] + [
        (n, C.c_double * 3) for n in ("domposmin", "maprealposmin", "maprealsize", "mapposmin")] + [
        ("tboundary", C.c_uint32), ("periactive", C.c_uint32), ("perinc", (C.c_double * 3) * 3),
        ("sim2d", C.c_uint32)]


class OrcDiv(C.Structure):
    _fields_ = [("scelldiv", C.c_int), ("ncx", C.c_int), ("ncy", C.c_int), ("ncz", C.c_int),
                ("ncw", C.c_int), ("cellfluid", C.c_uint32), ("czx", C.c_int), ("czy", C.c_int),
                ("czz", C.c_int), ("begincell", C.c_void_p)]


_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing: python -m dualsphysics_amd.build")
        lib = C.CDLL(str(LIB_PATH))
        vp, u, i, f, d = C.c_void_p, C.c_uint32, C.c_int, C.c_float, C.c_double
        sig = {
            "orc_cell_code": ([u, u, u], u),
            "orc_dcell": ([u, vp, vp, f, u, vp], None),
            "orc_limits": ([u, u, u, vp, vp, vp, vp], None),
            "orc_presort": ([i, u, u, u, vp, vp, vp, vp, vp], None),
            "orc_sort_cells": ([u, u, u, u, vp, vp, vp], None),
            "orc_poscell": ([u, vp, vp, f, vp], None),
            "orc_interaction": ([i, i, i, i, f, f, C.POINTER(OrcCte), C.POINTER(OrcDiv), u, u, u,
                                 vp, vp, vp, vp, vp, vp, vp, vp, vp, vp], None),
            "orc_sps_tau": ([u, u, f, f, vp, vp, vp], None),
            "orc_sym_pre": ([u, u, vp, vp, vp, d, f, f, f, vp, vp, vp, vp], None),
            "orc_sym_cor": ([u, u, vp, vp, vp, d, d, f, f, f, vp, vp, vp, vp], None),
            "orc_verlet": ([u, u, vp, vp, vp, vp, d, d, f, f, f, vp, vp, vp, vp], None),
            "orc_step_pos": ([u, u, C.POINTER(OrcCte), vp, vp, vp, vp, vp], None),
            "orc_dt_variable": ([d, d, f, f, d, d, d], d),
            "orc_pips": ([C.POINTER(OrcCte), C.POINTER(OrcDiv), u, u, u, vp, vp, vp], None),
            "orc_bruteforce_fp64": ([C.POINTER(OrcCte), i, i, f, f, u, u, vp, vp, vp, vp, vp],
                                    None),
            "orc_sim_create": ([vp, i, u, vp, vp, vp, vp], vp),
            "orc_mdbc": ([i, i, C.POINTER(OrcCte), C.POINTER(OrcDiv), u, vp, vp, vp, vp, vp], None),
            "orc_mdbc2": ([i, i, C.POINTER(OrcCte), C.POINTER(OrcDiv), u, vp, vp, vp, vp, vp, vp,
                           vp, d, d, vp, vp, vp], None),
            "orc_mdbc2_keep": ([u, vp, vp, vp], None),
            "orc_mdbc2_slip": ([i, i, i, C.POINTER(OrcCte), C.POINTER(OrcDiv), u, vp, vp, vp,
                                vp, vp, vp, vp, d, d, vp, vp, vp], None),
            "orc_nopen_step": ([u, u, vp, vp, d, vp, vp], None),
            "orc_interaction_np": ([i, i, i, i, f, f, C.POINTER(OrcCte), C.POINTER(OrcDiv), u,
                                    u, u] + [vp] * 12 + [i, vp, vp, vp, vp, vp], None),
            "orc_interaction_range": ([i, i, i, i, f, f, C.POINTER(OrcCte), C.POINTER(OrcDiv),
                                       u, u, u] + [vp] * 12 + [i, vp, vp, vp, vp, vp]
                                      + [C.c_long] * 4, None),
            "orc_interaction_range_mv": ([i, i, i, i, f, f, C.POINTER(OrcCte),
                                          C.POINTER(OrcDiv), u, u, u] + [vp] * 12
                                         + [i, vp, vp, vp, vp, vp, vp] + [C.c_long] * 4, None),
            "orc_motion_vel_ace": ([u, u, u, d, vp, vp, vp, vp], None),
            "orc_matrix_rot": ([d, vp, vp, vp], None),
            "orc_interaction_ext": ([i, i, i, i, f, f, C.POINTER(OrcCte), C.POINTER(OrcDiv), u,
                                     u, u] + [vp] * 12 + [i, vp, vp], None),
            "orc_run_shifting": ([u, u, d, f, f, f, d, vp, vp, u], None),
            "orc_add_shift": ([u, u, vp, vp], None),
            "orc_interaction_mdbc2": ([i, i, i, i, f, f, C.POINTER(OrcCte), C.POINTER(OrcDiv), u,
                                       u, u, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp],
                                      None),
            "orc_sim_destroy": ([vp], None),
            "orc_sim_error": ([vp], C.c_char_p),
            "orc_sim_run": ([vp, u], i),
            "orc_sim_download": ([vp, vp, vp, vp, vp], u),
            "orc_sim_forces": ([vp, vp, vp], None),
            "orc_sim_stats": ([vp, vp, vp], None),
            "orc_sim_celldiv": ([vp, vp, vp], u),
            "orc_sim_pips": ([vp, vp], None),
            "orc_sim_interact": ([vp], None),
            "orc_sim_arrays": ([vp, vp, vp, C.POINTER(OrcCte)], None),
            "orc_sim_floating": ([vp, u, vp], i),
            "orc_interaction_ft": ([i, i, i, i, f, f, C.POINTER(OrcCte), C.POINTER(OrcDiv), u,
                                    u, u] + [vp] * 12 + [i, vp, vp, vp], None),
            "orc_kernel_eval": ([i, f, f, f, f, C.POINTER(d), C.POINTER(d)], None),
            "orc_eos": ([i, i, C.POINTER(OrcCte), f], d),
        }
        for name, (args, res) in sig.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class OracleSim:
    """Whole-step CPU oracle (JSphCpuSingle flow), prec 0 = FP32, 1 = FP64."""

    def __init__(self, case, prec: int = 0):
        self.lib = load()
        self.case = case
        self._nc = case.to_native()
        idp = np.ascontiguousarray(case.idp, np.uint32)
        pos = np.ascontiguousarray(case.pos, np.float64)
        vr = np.ascontiguousarray(case.velrho, np.float32)
        bn = (None if getattr(case, "boundnor", None) is None
              else np.ascontiguousarray(case.boundnor, np.float32))
        self._bn = bn
        self._h = self.lib.orc_sim_create(C.cast(C.pointer(self._nc), C.c_void_p), prec,
                                          case.np, _p(idp), _p(pos), _p(vr), _p(bn))
        err = self.lib.orc_sim_error(self._h).decode()
        if err:
            raise RuntimeError("oracle: " + err)

    def close(self):
        if self._h:
            self.lib.orc_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, n: int):
        if self.lib.orc_sim_run(self._h, n):
            raise RuntimeError("oracle: " + self.lib.orc_sim_error(self._h).decode())

    def stats(self) -> dict:
        u = (C.c_uint32 * 6)()
        d = (C.c_double * 6)()
        self.lib.orc_sim_stats(self._h, u, d)
        return {"np": u[0], "npb": u[1], "npbok": u[2], "nct": u[3], "nstep": u[4],
                "npf_out_total": u[5], "timestep": d[0], "dt_last": d[1], "dtpre": d[2],
                "velmax": d[3], "acemax": d[4], "viscdtmax": d[5]}

    def download(self):
        n = self.stats()["np"]
        idp = np.empty(n, np.uint32)
        pos = np.empty((n, 3), np.float64)
        vr = np.empty((n, 4), np.float32)
        code = np.empty(n, np.uint16)
        self.lib.orc_sim_download(self._h, _p(idp), _p(pos), _p(vr), _p(code))
        keep = (code & 0xe000) != 0x2000   # periodic copies left out, order kept
        if not keep.all():
            idp, pos, vr, code = idp[keep], pos[keep], vr[keep], code[keep]
        return {"idp": idp, "pos": pos, "velrho": vr, "code": code}

    def forces(self):
        n = self.stats()["np"]
        ace = np.empty((n, 3), np.float32)
        ar = np.empty(n, np.float32)
        self.lib.orc_sim_forces(self._h, _p(ace), _p(ar))
        return ace, ar

    def interact(self):
        self.lib.orc_sim_interact(self._h)

    def celldiv(self):
        c6 = (C.c_uint32 * 6)()
        nctt = self.lib.orc_sim_celldiv(self._h, c6, None)
        be = np.empty(2 * nctt, np.int32)
        self.lib.orc_sim_celldiv(self._h, c6, _p(be))
        return list(c6), be.reshape(-1, 2)

    def arrays(self):
        """dcell, poscell and the constants of the current (sorted) state."""
        n = self.stats()["np"]
        dcell = np.empty(n, np.uint32)
        poscell = np.empty((n, 4), np.float32)
        cte = OrcCte()
        self.lib.orc_sim_arrays(self._h, _p(dcell), _p(poscell), C.byref(cte))
        return dcell, poscell, cte

    def floating(self, cf: int = 0) -> dict:
        """State of floating body cf: centre, angles, linear and angular
        velocity and acceleration (StFloatingData)."""
        from dualsphysics_amd import _native as N
        st = N.DsphFloatingState()
        if self.lib.orc_sim_floating(self._h, cf, C.cast(C.pointer(st), C.c_void_p)):
            raise IndexError(f"no floating body {cf}")
        return {"center": np.array(st.center[:]), "angles": np.array(st.angles[:], np.float32),
                "fvel": np.array(st.fvel[:], np.float32),
                "fomega": np.array(st.fomega[:], np.float32),
                "facelin": np.array(st.facelin[:], np.float32),
                "faceang": np.array(st.faceang[:], np.float32)}

    def pips(self):
        res = (C.c_uint64 * 4)()
        self.lib.orc_sim_pips(self._h, res)
        return [int(v) for v in res]
