"""The benchmark dam break at its own size on the GPU against the oracle.

3-D dam break at dp = 0.0025 m: 5,145,600 fluid + 934,016 wall particles
(6,079,616 in all), 1.2M cells of 2h, 21-bit sort keys (three radix passes,
the odd pass count included). Two Symplectic steps with host dt; after each
step's final divide:

- the particle order is bit for bit the order the oracle's own divide gives
  to the GPU's state (the oracle re-divides the downloaded state and must
  leave it in place);
- the force interaction on that state (C ABI on the oracle-prepared arrays,
  and the simulation's own interaction) is within 1e-5 norm-wise of the FP32
  oracle (reference GPU semantics) for ace, ar and viscdt, with the
  per-particle error and the error against the FP64 yardstick printed;
- the PIPS pair counts are exact.

Reference: JSphGpu_ker.cu:789-990 (interaction), JCellDivGpu_ker.cu:120-127
(sort), JDsPips_ker.cu:54-124 (pair count).
"""
import ctypes as C
import dataclasses
import sys
import time

import numpy as np
import pytest

from dualsphysics_amd import _native as N
from dualsphysics_amd import case as casemod
from tests.helpers import Dev, divdata, per_particle_rel, rel_err, to_cte

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _state_case(c, g):
    return dataclasses.replace(c, idp=g["idp"].copy(), pos=g["pos"].copy(),
                               velrho=g["velrho"].copy())


@pytest.mark.timeout(1500)
def test_dam_break_dp0_0025_order_forces_pips():
    import oracle.oracle as orc
    from dualsphysics_amd.sim import SphGpuSingle
    from oracle.oracle import OracleSim
    t0 = time.time()
    c = casemod.dam_break_cfg1()
    assert c.nfluid == 5_145_600 and c.nbound == 934_016
    _say(f"case built: {c.np} particles ({time.time() - t0:.1f} s)")
    lib = N.load()
    s = SphGpuSingle(c)
    ctx = C.c_void_p()
    assert lib.dsph_ctx_create(0, C.byref(ctx)) == 0
    try:
        assert lib.dsph_ctx_reserve(ctx, c.np + 1024, 0) == 0, lib.dsph_last_error(ctx)
        for step in (1, 2):
            s.run(1, sync_dt=True)
            st = s.stats()
            g = s.download()
            assert st["np"] == c.np
            _say(f"step {step}: t={st['timestep']:.4e} s, {st['nct']} cells")
            o = OracleSim(_state_case(c, g))
            od = o.download()
            nd = int(np.count_nonzero(od["idp"] != g["idp"]))
            assert nd == 0, f"step {step}: {nd} particles out of the oracle's order"
            dcell, poscell, cte = o.arrays()
            celldiv, be = o.celldiv()
            ost = o.stats()
            n, npb, npbok = ost["np"], ost["npb"], ost["npbok"]
            # GPU interaction on the oracle-prepared arrays (viscdt included).
            bc = Dev(lib, ctx, be)
            dv = divdata(celldiv, bc, cte)
            acte = to_cte(cte)
            assert lib.dsph_set_constants(ctx, C.byref(acte)) == 0
            gd = {k: Dev(lib, ctx, a) for k, a in (("dcell", dcell), ("poscell", poscell),
                                                  ("velrho", od["velrho"]), ("code", od["code"]))}
            ace = Dev.zeros(lib, ctx, (n, 3), np.float32)
            ar = Dev.zeros(lib, ctx, n, np.float32)
            vdt = Dev.zeros(lib, ctx, n, np.float32)
            t = N.DsphInterParms()
            t.tkernel, t.tvisco, t.tdensity = N.KERNEL_WENDLAND, N.VISCO_ARTIFICIAL, N.DDT_DDT2FULL
            t.viscob, t.viscof = 0.01, 0.01
            t.np, t.npb, t.npbok = n, npb, npbok
            t.divdata = dv
            t.d_dcell, t.d_poscell = gd["dcell"].ptr, gd["poscell"].ptr
            t.d_velrho, t.d_code = gd["velrho"].ptr, gd["code"].ptr
            t.d_viscdt, t.d_ar, t.d_ace = vdt.ptr, ar.ptr, ace.ptr
            assert lib.dsph_interaction_forces(ctx, C.byref(t)) == 0, lib.dsph_last_error(ctx)
            gace, gar, gvdt = ace.get(), ar.get(), vdt.get()
            # The simulation's own interaction on its own arrays gives the same bits.
            s.interact()
            sace, sar = s.forces()
            np.testing.assert_array_equal(sace.view(np.uint32), gace.view(np.uint32))
            np.testing.assert_array_equal(sar.view(np.uint32), gar.view(np.uint32))
            odiv = orc.OrcDiv(1, int(dv.ncx), int(dv.ncy), int(dv.ncz), int(dv.ncw),
                              int(dv.cellfluid), int(dv.cellzerox), int(dv.cellzeroy),
                              int(dv.cellzeroz), be.ctypes.data)
            pos = np.ascontiguousarray(od["pos"])
            res = {}
            for prec, sem in ((0, orc.SEM_GPU), (1, orc.SEM_F64)):
                t1 = time.time()
                oa = np.zeros((n, 3), np.float32)
                oar = np.zeros(n, np.float32)
                ov = np.zeros(n, np.float32)
                os_ = np.zeros((n, 6), np.float32)
                orc.load().orc_interaction(prec, sem, N.VISCO_ARTIFICIAL, N.DDT_DDT2FULL, 0.01,
                                           0.01, C.byref(cte), C.byref(odiv), n, npb, npbok,
                                           dcell.ctypes.data, poscell.ctypes.data,
                                           pos.ctypes.data, od["velrho"].ctypes.data, None,
                                           oa.ctypes.data, oar.ctypes.data, ov.ctypes.data,
                                           os_.ctypes.data, None)
                res[prec] = (oa, oar, ov)
                _say(f"step {step}: oracle prec {prec} interaction {time.time() - t1:.1f} s")
            (oa, oar, ov), (a64, r64, v64) = res[0], res[1]
            fl = slice(npb, None)
            e = {"ace": rel_err(gace[fl], oa[fl]), "ar": rel_err(gar, oar),
                 "viscdt": rel_err(gvdt, ov)}
            pp = per_particle_rel(gace[fl], oa[fl])
            ppr = per_particle_rel(gar[fl], oar[fl])
            e64 = {"gpu ace": rel_err(gace[fl], a64[fl]), "oracle32 ace": rel_err(oa[fl], a64[fl]),
                   "gpu ar": rel_err(gar, r64), "oracle32 ar": rel_err(oar, r64)}
            _say(f"step {step}: gpu vs fp32 oracle " +
                 " ".join(f"{k} {v:.2e}" for k, v in e.items()) +
                 f" | per particle ace max {pp.max():.2e} p99.99 {np.quantile(pp, 0.9999):.2e}"
                 f" ar max {ppr.max():.2e} | vs fp64 " +
                 " ".join(f"{k} {v:.2e}" for k, v in e64.items()))
            for k, v in e.items():
                assert v < TOL, (step, k, e)
            assert e64["gpu ace"] < 2 * e64["oracle32 ace"] + 1e-6
            assert e64["gpu ar"] < 2 * e64["oracle32 ar"] + 1e-6
            # PIPS counts of the state: GPU sim = GPU on the oracle arrays = oracle.
            opips = o.pips()
            res4 = (C.c_uint64 * 4)()
            assert lib.dsph_pips_count_sync(ctx, npbok, 0, n - npb, npb, C.byref(dv),
                                            gd["dcell"].ptr, gd["poscell"].ptr, res4) == 0
            assert list(res4) == opips
            assert s.pips()[:4] == opips
            _say(f"step {step}: pips {opips}")
            o.close()
            from tests import helpers
            helpers.release_all()
    finally:
        s.close()
        lib.dsph_ctx_destroy(ctx)


@pytest.mark.timeout(1500)
def test_wave_flume_dp0_005_order_and_sampled_forces():
    """The wave-flume configuration at its own size: 12 x 1 x 1.2 m, depth
    0.6 m, dp = 0.005 m, 57,600,000 fluid + 3,606,768 wall particles
    (61,206,768 in all), mDBC, laminar+SPS, piston. One Symplectic step on
    the GPU (host dt, DDT2Full as the configuration's default); then

    - the particle order is bit for bit the order the oracle's own divide
      gives to the GPU's state;
    - the force interaction on that state (C ABI on the oracle-prepared
      arrays; mDBC constants, laminar+SPS with a random tau) is within 1e-5
      norm-wise of the FP32 oracle (reference GPU semantics) for ace, ar,
      viscdt and the strain on a sample: the fluid particles of one
      contiguous twelfth of the sorted fluid range (4.8M) and a tenth of the
      boundary range, with the error against the FP64 yardstick printed.
      The oracle computes only the sample (orc_interaction_range); the GPU
      computes all 61.2M particles, so 32-bit offsets, the sort and the
      neighbour ranges run at this size."""
    import oracle.oracle as orc
    from dualsphysics_amd.sim import SphGpuSingle
    from oracle.oracle import OracleSim
    t0 = time.time()
    prm = casemod.Parameters(tboundary=N.BC_MDBC, tvisco=N.VISCO_LAMINARSPS, visco=1e-6)
    c = casemod.flume(dp=0.005, length=12.0, width=1.0, height=1.2, depth=0.6, ampl=0.1,
                      period=1.6, params=prm)
    assert c.nfluid == 57_600_000 and c.nbound == 3_606_768
    _say(f"flume built: {c.np} particles ({time.time() - t0:.1f} s)")
    s = SphGpuSingle(c)
    try:
        s.run(1, sync_dt=True)
        st = s.stats()
        g = s.download()
    finally:
        s.close()
    assert st["np"] == c.np
    _say(f"step 1: t={st['timestep']:.4e} s, {st['nct']} cells ({time.time() - t0:.1f} s)")
    nb = c.nbound
    assert np.all(g["idp"][:nb] < nb)
    sc = dataclasses.replace(_state_case(c, g), boundnor=c.boundnor[g["idp"][:nb]].copy())
    del c
    o = OracleSim(sc)
    lib = N.load()
    ctx = C.c_void_p()
    assert lib.dsph_ctx_create(0, C.byref(ctx)) == 0
    try:
        od = o.download()
        nd = int(np.count_nonzero(od["idp"] != g["idp"]))
        assert nd == 0, f"{nd} particles out of the oracle's order"
        dcell, poscell, cte = o.arrays()
        celldiv, be = o.celldiv()
        ost = o.stats()
        n, npb, npbok = ost["np"], ost["npb"], ost["npbok"]
        _say(f"oracle order equal; np {n} npb {npb} npbok {npbok} ({time.time() - t0:.1f} s)")
        del g
        assert lib.dsph_ctx_reserve(ctx, n + 1024, 0) == 0, lib.dsph_last_error(ctx)
        bc = Dev(lib, ctx, be)
        dv = divdata(celldiv, bc, cte)
        acte = to_cte(cte, tboundary=N.BC_MDBC)
        assert lib.dsph_set_constants(ctx, C.byref(acte)) == 0
        rng = np.random.default_rng(11)
        tau = (rng.standard_normal((n, 6)) * 1e-3).astype(np.float32)
        gd = {k: Dev(lib, ctx, a) for k, a in (("dcell", dcell), ("poscell", poscell),
                                              ("velrho", od["velrho"]), ("code", od["code"]),
                                              ("tau", tau))}
        ace = Dev.zeros(lib, ctx, (n, 3), np.float32)
        ar = Dev.zeros(lib, ctx, n, np.float32)
        vdt = Dev.zeros(lib, ctx, n, np.float32)
        strain = Dev.zeros(lib, ctx, (n, 6), np.float32)
        t = N.DsphInterParms()
        t.tkernel, t.tvisco, t.tdensity = N.KERNEL_WENDLAND, N.VISCO_LAMINARSPS, N.DDT_DDT2FULL
        t.viscob = t.viscof = 1e-6
        t.np, t.npb, t.npbok = n, npb, npbok
        t.divdata = dv
        t.d_dcell, t.d_poscell = gd["dcell"].ptr, gd["poscell"].ptr
        t.d_velrho, t.d_code = gd["velrho"].ptr, gd["code"].ptr
        t.d_spstaurho2 = gd["tau"].ptr
        t.d_viscdt, t.d_ar, t.d_ace, t.d_sps2strain = vdt.ptr, ar.ptr, ace.ptr, strain.ptr
        assert lib.dsph_interaction_forces(ctx, C.byref(t)) == 0, lib.dsph_last_error(ctx)
        gace, gar, gvdt, gst = ace.get(), ar.get(), vdt.get(), strain.get()
        _say(f"gpu interaction done ({time.time() - t0:.1f} s)")
        from tests import helpers
        helpers.release_all()
        nf = n - npb
        fb, fe = npb + 6 * (nf // 12), npb + 7 * (nf // 12)
        bb, be_ = 4 * (npbok // 10), 5 * (npbok // 10)
        odiv = orc.OrcDiv(1, int(dv.ncx), int(dv.ncy), int(dv.ncz), int(dv.ncw),
                          int(dv.cellfluid), int(dv.cellzerox), int(dv.cellzeroy),
                          int(dv.cellzeroz), be.ctypes.data)
        pos = np.ascontiguousarray(od["pos"])
        res = {}
        for prec, sem in ((0, orc.SEM_GPU), (1, orc.SEM_F64)):
            t1 = time.time()
            oa = np.zeros((n, 3), np.float32)
            oar = np.zeros(n, np.float32)
            ov = np.zeros(n, np.float32)
            os_ = np.zeros((n, 6), np.float32)
            orc.load().orc_interaction_range(
                prec, sem, N.VISCO_LAMINARSPS, N.DDT_DDT2FULL, 1e-6, 1e-6, C.byref(cte),
                C.byref(odiv), n, npb, npbok, dcell.ctypes.data, poscell.ctypes.data,
                pos.ctypes.data, od["velrho"].ctypes.data, tau.ctypes.data, oa.ctypes.data,
                oar.ctypes.data, ov.ctypes.data, os_.ctypes.data, None, None, None, 0, None,
                None, None, None, None, fb, fe, bb, be_)
            res[prec] = (oa, oar, ov, os_)
            _say(f"oracle prec {prec} on the sample: {time.time() - t1:.1f} s")
        (oa, oar, ov, os_), (a64, r64, v64, s64) = res[0], res[1]
        fs, bs = slice(fb, fe), slice(bb, be_)
        e = {"ace": rel_err(gace[fs], oa[fs]), "ar": rel_err(gar[fs], oar[fs]),
             "viscdt": rel_err(gvdt[fs], ov[fs]), "strain": rel_err(gst[fs], os_[fs]),
             "bound ar": rel_err(gar[bs], oar[bs]), "bound viscdt": rel_err(gvdt[bs], ov[bs])}
        e64 = {"gpu ace": rel_err(gace[fs], a64[fs]), "oracle32 ace": rel_err(oa[fs], a64[fs]),
               "gpu ar": rel_err(gar[fs], r64[fs]), "oracle32 ar": rel_err(oar[fs], r64[fs])}
        pp = per_particle_rel(gace[fs], oa[fs])
        _say("sample: gpu vs fp32 oracle " + " ".join(f"{k} {v:.2e}" for k, v in e.items()) +
             f" | per particle ace max {pp.max():.2e} p99.99 {np.quantile(pp, 0.9999):.2e}"
             " | vs fp64 " + " ".join(f"{k} {v:.2e}" for k, v in e64.items()))
        assert np.abs(oa[fs]).max() > 0 and np.abs(oar[bs]).max() > 0
        for k, v in e.items():
            assert v < TOL, (k, e)
        assert e64["gpu ace"] < 2 * e64["oracle32 ace"] + 1e-6
        assert e64["gpu ar"] < 2 * e64["oracle32 ar"] + 1e-6
    finally:
        o.close()
        lib.dsph_ctx_destroy(ctx)
