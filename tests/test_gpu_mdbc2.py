"""mDBC2 (SlipMode no-slip and free slip) on the GPU against the oracle: the
correction kernel (density, boundary mode, tangential velocity), the fluid
interaction with boundary modes and with the no-penetration sums, the
no-penetration step correction, and the integrators' rule for mDBC2
boundaries. The whole step against the reference CPU build is in
test_gpu_vs_reference.py (fixtures ref_dambreak_dp0.03_mdbc2*)."""
import dataclasses
import ctypes as C

import numpy as np
import pytest

import oracle.oracle as orc
from dualsphysics_amd import _native as N
from dualsphysics_amd import case as casemod
from tests.helpers import (Dev, divdata, hydrostatic_mdbc2_case, mdbc2_state, oracle_mdbc2,
                           perturbed_case, rel_err, to_cte)

pytestmark = pytest.mark.gpu


def _gpu_mdbc2(lib, ctx, c, s, velrho, slipmode=N.SLIP_NOSLIP, mv=None, ma=None):
    acte = to_cte(s["cte"], tboundary=N.BC_MDBC)
    assert lib.dsph_set_constants(ctx, C.byref(acte)) == 0
    bc = Dev(lib, ctx, s["be"])
    dv = divdata(s["celldiv"], bc, s["cte"])
    pos = s["pos"]
    npb = s["npb"]
    d_xy = Dev(lib, ctx, np.ascontiguousarray(pos[:, :2]))
    d_z = Dev(lib, ctx, np.ascontiguousarray(pos[:, 2]))
    d_ps = Dev(lib, ctx, s["poscell"])
    d_code = Dev(lib, ctx, s["d"]["code"])
    d_bn = Dev(lib, ctx, s["bn2"])
    d_vr = Dev(lib, ctx, velrho)
    d_bm = Dev.zeros(lib, ctx, npb, np.uint8)
    d_tv = Dev.zeros(lib, ctx, (npb, 3), np.float32)
    g = (C.c_float * 3)(*c.gravity)
    mp = (C.c_double * 3)(*s["cte"].mapposmin)
    d_mv = Dev(lib, ctx, mv) if mv is not None else None
    d_ma = Dev(lib, ctx, ma) if ma is not None else None
    pmv = d_mv.ptr if d_mv else None
    pma = d_ma.ptr if d_ma else None
    if slipmode == N.SLIP_NOSLIP:
        rc = lib.dsph_mdbc2_correction(ctx, npb, g, C.byref(dv), mp, d_xy.ptr, d_z.ptr, d_ps.ptr,
                                       d_code.ptr, d_bn.ptr, pmv, pma, d_vr.ptr, d_bm.ptr,
                                       d_tv.ptr)
    else:
        rc = lib.dsph_mdbc2_correction_slip(ctx, slipmode, npb, g, C.byref(dv), mp, d_xy.ptr,
                                            d_z.ptr, d_ps.ptr, d_code.ptr, d_bn.ptr, pmv, pma,
                                            d_vr.ptr, d_bm.ptr, d_tv.ptr)
    assert rc == 0, lib.dsph_last_error(ctx)
    return d_vr.get(), d_bm.get(), d_tv.get(), dv, bc


def _cases():
    lay = hydrostatic_mdbc2_case()
    p = casemod.Parameters(tboundary=N.BC_MDBC, slipmode=N.SLIP_NOSLIP, tvisco=N.VISCO_LAMINAR,
                           visco=1e-4)
    db = perturbed_case(casemod.dam_break(dp=0.02, params=p))
    db2 = perturbed_case(casemod.dam_break_2d(dp=0.005, params=p))
    return {"hydrostatic_layer": lay, "perturbed_dam_break": db, "perturbed_dam_break_2d": db2}


@pytest.mark.parametrize("name", ["hydrostatic_layer", "perturbed_dam_break",
                                  "perturbed_dam_break_2d"])
def test_mdbc2_correction_matches_oracle(lib, ctx, name):
    c = _cases()[name]
    s = mdbc2_state(c)
    npb = s["npb"]
    print(f"case {name}: {s['n']} particles, {npb} boundary", flush=True)
    g_vr, g_bm, g_tv, _, _ = _gpu_mdbc2(lib, ctx, c, s, s["d"]["velrho"])
    o_vr, o_bm, o_tv = oracle_mdbc2(c, s, 0, orc.SEM_GPU)
    f_vr, f_bm, f_tv = oracle_mdbc2(c, s, 1, orc.SEM_F64)
    # Modes: the submerged sign test, exact except at its rounding edge.
    same = g_bm == o_bm
    assert same.mean() > 0.999, (np.count_nonzero(~same), npb)
    e_rho = np.abs(g_vr[:npb, 3] - o_vr[:npb, 3])[same].max()
    e_rho64 = np.abs(g_vr[:npb, 3] - f_vr[:npb, 3])[same].max()
    e_tv = rel_err(g_tv[same], o_tv[same])
    print(f"{name}: modes {np.bincount(g_bm, minlength=3)} rho |gpu-oracle32| {e_rho:.2e} "
          f"|gpu-fp64| {e_rho64:.2e} tangenvel {e_tv:.2e}")
    # FP32 densities of ~1000: the 4x4 solve amplifies rounding to ~1e-3.
    assert e_rho < 5e-3 and e_rho64 < 5e-3
    assert e_tv < 1e-4
    np.testing.assert_array_equal(g_vr[npb:], s["d"]["velrho"][npb:])
    np.testing.assert_array_equal(g_vr[:npb, :3], s["d"]["velrho"][:npb, :3])


@pytest.mark.parametrize("kernel", [N.FORCES_TILE, N.FORCES_ROWS], ids=["tile", "rows"])
@pytest.mark.parametrize("tvisco,tddt", [(N.VISCO_LAMINAR, N.DDT_DDT2FULL),
                                         (N.VISCO_ARTIFICIAL, N.DDT_DDT2FULL),
                                         (N.VISCO_LAMINARSPS, N.DDT_DDT2)])
def test_interaction_with_boundary_modes_matches_oracle(lib, ctx, tvisco, tddt, kernel):
    assert lib.dsph_ctx_set_option(ctx, N.OPT_FORCES_KERNEL, kernel) == 0
    c = _cases()["perturbed_dam_break"]
    s = mdbc2_state(c)
    n, npb, npbok = s["n"], s["npb"], s["npbok"]
    o_vr, o_bm, o_tv = oracle_mdbc2(c, s, 0, orc.SEM_GPU)
    assert (o_bm == 2).any() and (o_bm == 1).any()
    rng = np.random.default_rng(5)
    tau = (rng.standard_normal((n, 6)) * 1e-3).astype(np.float32)
    visco = 1e-3 if tvisco != N.VISCO_ARTIFICIAL else 0.01
    acte = to_cte(s["cte"], tboundary=N.BC_MDBC)
    assert lib.dsph_set_constants(ctx, C.byref(acte)) == 0
    bc = Dev(lib, ctx, s["be"])
    dv = divdata(s["celldiv"], bc, s["cte"])
    gd = {k: Dev(lib, ctx, a) for k, a in (("dcell", s["dcell"]), ("poscell", s["poscell"]),
                                          ("velrho", o_vr), ("code", s["d"]["code"]),
                                          ("bm", o_bm), ("tv", o_tv), ("tau", tau))}
    ace = Dev.zeros(lib, ctx, (n, 3), np.float32)
    ar = Dev.zeros(lib, ctx, n, np.float32)
    vdt = Dev.zeros(lib, ctx, n, np.float32)
    strain = Dev.zeros(lib, ctx, (n, 6), np.float32)
    t = N.DsphInterParms()
    t.tkernel, t.tvisco, t.tdensity = N.KERNEL_WENDLAND, tvisco, tddt
    t.viscob, t.viscof = visco, visco
    t.np, t.npb, t.npbok = n, npb, npbok
    t.divdata = dv
    t.d_dcell, t.d_poscell, t.d_velrho, t.d_code = (gd["dcell"].ptr, gd["poscell"].ptr,
                                                     gd["velrho"].ptr, gd["code"].ptr)
    t.d_spstaurho2 = gd["tau"].ptr if tvisco == N.VISCO_LAMINARSPS else None
    t.d_viscdt, t.d_ar, t.d_ace, t.d_sps2strain = vdt.ptr, ar.ptr, ace.ptr, strain.ptr
    t.d_boundmode, t.d_tangenvel = gd["bm"].ptr, gd["tv"].ptr
    assert lib.dsph_interaction_forces(ctx, C.byref(t)) == 0, lib.dsph_last_error(ctx)
    g = {"ace": ace.get(), "ar": ar.get(), "viscdt": vdt.get(), "strain": strain.get()}
    res = {}
    olib = orc.load()
    for with_modes in (True, False):
        oa = np.zeros((n, 3), np.float32)
        oar = np.zeros(n, np.float32)
        ov = np.zeros(n, np.float32)
        os_ = np.zeros((n, 6), np.float32)
        olib.orc_interaction_mdbc2(
            0, orc.SEM_GPU, tvisco, tddt, visco, visco, C.byref(s["cte"]), C.byref(s["odiv"]), n,
            npb, npbok, s["dcell"].ctypes.data, s["poscell"].ctypes.data, s["pos"].ctypes.data,
            o_vr.ctypes.data, tau.ctypes.data if tvisco == N.VISCO_LAMINARSPS else None,
            oa.ctypes.data, oar.ctypes.data, ov.ctypes.data, os_.ctypes.data, None,
            o_bm.ctypes.data if with_modes else None, o_tv.ctypes.data if with_modes else None)
        res[with_modes] = {"ace": oa, "ar": oar, "viscdt": ov, "strain": os_}
    fl = slice(npb, None)
    o = res[True]
    e = {k: rel_err(g[k][fl] if k != "viscdt" else g[k], o[k][fl] if k != "viscdt" else o[k])
         for k in ("ace", "ar", "viscdt")}
    if tvisco == N.VISCO_LAMINARSPS:
        e["strain"] = rel_err(g["strain"][fl], o["strain"][fl])
    # The modes change the result well beyond the tolerance (the test sees them).
    d_modes = rel_err(res[False]["ace"][fl], o["ace"][fl])
    print(f"tvisco {tvisco} tddt {tddt}: {e}, without modes the oracle differs by {d_modes:.2e}")
    assert lib.dsph_ctx_set_option(ctx, N.OPT_FORCES_KERNEL, N.FORCES_TILE) == 0
    for k, v in e.items():
        assert v < 1e-5, (k, v)
    assert d_modes > 20 * 1e-5


def test_integrators_keep_mdbc2_boundaries(lib, ctx):
    """Boundary particles with mode >= 1 keep velrhopre (Symplectic) and
    velrho2 (Verlet); mode 0 follows the DBC update; fluid is unaffected."""
    n, npb = 600, 200
    rng = np.random.default_rng(3)
    vpre = np.zeros((n, 4), np.float32)
    vpre[:, :3] = rng.standard_normal((n, 3)).astype(np.float32)
    vpre[:, 3] = (1000 + 5 * rng.random(n)).astype(np.float32)
    vnow = vpre.copy()
    vnow[:, 3] += 1.0
    ar = (rng.standard_normal(n) * 50).astype(np.float32)
    ace = (rng.standard_normal((n, 3))).astype(np.float32)
    bm = rng.integers(0, 3, npb).astype(np.uint8)
    code = np.zeros(n, np.uint16)
    code[npb:] = 0x1800
    G = (C.c_float * 3)(0.0, 0.0, -9.81)
    dt = 1e-4
    out = {}
    for mode in ("pre", "cor", "verlet"):
        for with_bm in (True, False):
            d_code = Dev(lib, ctx, code)
            d_mxy = Dev.zeros(lib, ctx, (n, 2), np.float64)
            d_mz = Dev.zeros(lib, ctx, n, np.float64)
            d_bm = Dev(lib, ctx, bm).ptr if with_bm else None
            if mode == "verlet":
                d_new = Dev.zeros(lib, ctx, (n, 4), np.float32)
                rc = lib.dsph_step_verlet(ctx, n, npb, Dev(lib, ctx, vnow).ptr,
                                          Dev(lib, ctx, vpre).ptr, d_bm, Dev(lib, ctx, ar).ptr,
                                          Dev(lib, ctx, ace).ptr, None, dt, 2 * dt, 1000., 700., 1300.,
                                          G, d_code.ptr, d_mxy.ptr, d_mz.ptr, d_new.ptr)
            else:
                d_new = Dev(lib, ctx, vnow)
                fn = lib.dsph_step_symplectic_pre if mode == "pre" else lib.dsph_step_symplectic_cor
                extra = (dt * .5,) if mode == "pre" else (dt * .5, dt)
                rc = fn(ctx, n, npb, Dev(lib, ctx, vpre).ptr, d_bm, Dev(lib, ctx, ar).ptr,
                        Dev(lib, ctx, ace).ptr, None, *extra, None, 0, 1000., 700., 1300., G,
                        d_code.ptr, d_mxy.ptr, d_mz.ptr, d_new.ptr)
            assert rc == 0, lib.dsph_last_error(ctx)
            out[(mode, with_bm)] = d_new.get()
        a, b = out[(mode, True)], out[(mode, False)]
        keep = bm >= 1
        np.testing.assert_array_equal(a[:npb][keep], vpre[:npb][keep])
        np.testing.assert_array_equal(a[:npb][~keep], b[:npb][~keep])
        np.testing.assert_array_equal(a[npb:], b[npb:])
        # The same rule in the oracle.
        ref = b.copy()
        orc.load().orc_mdbc2_keep(npb, bm.ctypes.data, vpre.ctypes.data, ref.ctypes.data)
        np.testing.assert_array_equal(a, ref)


def test_mdbc2_freeslip_correction_matches_oracle(lib, ctx):
    """Free slip: the same density and modes as no-slip, the tangential
    velocity of the ghost node's Shepard velocity instead of its mirror."""
    c = _cases()["perturbed_dam_break"]
    s = mdbc2_state(c)
    npb = s["npb"]
    g_vr, g_bm, g_tv, _, _ = _gpu_mdbc2(lib, ctx, c, s, s["d"]["velrho"], N.SLIP_FREESLIP)
    o_vr, o_bm, o_tv = oracle_mdbc2(c, s, 0, orc.SEM_GPU, slipmode=N.SLIP_FREESLIP)
    n_vr, n_bm, n_tv = oracle_mdbc2(c, s, 0, orc.SEM_GPU, slipmode=N.SLIP_NOSLIP)
    same = g_bm == o_bm
    assert same.mean() > 0.999, (np.count_nonzero(~same), npb)
    e_rho = np.abs(g_vr[:npb, 3] - o_vr[:npb, 3])[same].max()
    on = same & (g_bm == 1)
    e_tv = rel_err(g_tv[on], o_tv[on])
    d_ns = rel_err(n_tv[on], o_tv[on])
    print(f"free slip: rho |gpu-oracle32| {e_rho:.2e} tangenvel {e_tv:.2e}; no-slip's "
          f"tangential velocity differs by {d_ns:.2e}")
    assert e_rho < 5e-3 and e_tv < 1e-4
    assert d_ns > 0.5
    assert np.abs(o_vr[:npb, 3] - n_vr[:npb, 3]).max() == 0.0   # density: slip mode plays no part


@pytest.mark.parametrize("slipmode", [N.SLIP_NOSLIP, N.SLIP_FREESLIP], ids=["noslip", "freeslip"])
def test_mdbc2_correction_with_moving_walls_matches_oracle(lib, ctx, slipmode):
    """Moving walls: the motion acceleration enters the pressure clone and
    the motion velocity the no-slip mirror (JSphGpu_mdbc_iker.cu:411-470);
    random motion velocity (0.3 m/s) and acceleration (5 m/s^2) per wall
    particle against the oracle, and the result differs from fixed walls."""
    c = _cases()["perturbed_dam_break"]
    s = mdbc2_state(c)
    npb, n = s["npb"], s["n"]
    rng = np.random.default_rng(5)
    mv = np.zeros((n, 3), np.float32)
    ma = np.zeros((n, 3), np.float32)
    mv[:npb] = (rng.standard_normal((npb, 3)) * 0.3).astype(np.float32)
    ma[:npb] = (rng.standard_normal((npb, 3)) * 5.0).astype(np.float32)
    g_vr, g_bm, g_tv, _, _ = _gpu_mdbc2(lib, ctx, c, s, s["d"]["velrho"], slipmode, mv, ma)
    o_vr, o_bm, o_tv = oracle_mdbc2(c, s, 0, orc.SEM_GPU, slipmode=slipmode, motionvel=mv,
                                    motionace=ma)
    f_vr, _, f_tv = oracle_mdbc2(c, s, 0, orc.SEM_GPU, slipmode=slipmode)
    same = g_bm == o_bm
    assert same.mean() > 0.999, (np.count_nonzero(~same), npb)
    on = same & (g_bm == 1)
    e_rho = np.abs(g_vr[:npb, 3] - o_vr[:npb, 3])[same].max()
    e_tv = rel_err(g_tv[same], o_tv[same])
    d_rho = np.abs(o_vr[:npb, 3] - f_vr[:npb, 3])[on].max()
    d_tv = rel_err(f_tv[same], o_tv[same])
    print(f"slip {slipmode}: rho |gpu-oracle32| {e_rho:.2e} tangenvel {e_tv:.2e}; moving vs "
          f"fixed walls: rho {d_rho:.2e} tangenvel {d_tv:.2e}")
    assert e_rho < 5e-3 and e_tv < 1e-4
    assert d_rho > 0.05
    if slipmode == N.SLIP_NOSLIP:
        assert d_tv > 0.1


@pytest.mark.parametrize("device_dt", [False, True], ids=["host_dt", "device_dt"])
def test_motion_vel_ace_matches_oracle(lib, ctx, device_dt):
    """CopyMotionVelAce: the moving particles (ids in [idbegin, idbegin+count)
    among the boundary particles, any order) get float((v - motionvel)/dt)
    and motionvel = v, bit for bit; the others are untouched."""
    rng = np.random.default_rng(3)
    npb, idbegin, count = 5000, 1200, 1700
    idp = rng.permutation(npb).astype(np.uint32)
    vr = (rng.standard_normal((npb, 4))).astype(np.float32)
    mv = (rng.standard_normal((npb, 3))).astype(np.float32)
    ma = np.full((npb, 3), 7.0, np.float32)
    dt = 1.7e-4
    d_idp, d_vr = Dev(lib, ctx, idp), Dev(lib, ctx, vr)
    d_mv, d_ma = Dev(lib, ctx, mv), Dev(lib, ctx, ma)
    d_dt = Dev(lib, ctx, np.array([0.0, 0.0, dt, 0.0], np.float64)) if device_dt else None
    rc = lib.dsph_mdbc2_motion_vel_ace(ctx, npb, idbegin, count, 0.0 if device_dt else dt,
                                       d_dt.ptr if device_dt else None, 2, d_idp.ptr, d_vr.ptr,
                                       d_mv.ptr, d_ma.ptr)
    assert rc == 0, lib.dsph_last_error(ctx)
    omv, oma = mv.copy(), ma.copy()
    orc.load().orc_motion_vel_ace(npb, idbegin, count, dt, idp.ctypes.data, vr.ctypes.data,
                                  omv.ctypes.data, oma.ctypes.data)
    np.testing.assert_array_equal(d_mv.get(), omv)
    np.testing.assert_array_equal(d_ma.get(), oma)
    moving = (idp >= idbegin) & (idp < idbegin + count)
    assert moving.sum() == count and np.all(oma[~moving] == 7.0)


def _nopen_case():
    """The perturbed mDBC2 dam break with the fluid moved 0.3 dp towards the
    floor and the back wall, so that the first fluid layers sit closer than
    0.75 of a normal length to the first wall layer (where the
    no-penetration correction acts), about half of them moving towards it."""
    c = _cases()["perturbed_dam_break"]
    pos = c.pos.copy()
    pos[c.nbound:, 0] -= 0.3 * c.dp
    pos[c.nbound:, 2] -= 0.3 * c.dp
    return dataclasses.replace(c, pos=pos)


@pytest.mark.parametrize("moving", [False, True], ids=["fixed", "moving"])
@pytest.mark.parametrize("tvisco,tddt", [(N.VISCO_LAMINAR, N.DDT_DDT2FULL),
                                         (N.VISCO_LAMINARSPS, N.DDT_DDT2)])
def test_interaction_no_penetration_matches_oracle(lib, ctx, tvisco, tddt, moving):
    """The no-penetration sums against the oracle, with fixed walls and with
    walls moving at random velocities (0.5 m/s, the test velocity relative to
    the wall's motion velocity)."""
    c = _nopen_case()
    s = mdbc2_state(c)
    n, npb, npbok = s["n"], s["npb"], s["npbok"]
    o_vr, o_bm, o_tv = oracle_mdbc2(c, s, 0, orc.SEM_GPU)
    rng = np.random.default_rng(9)
    tau = (rng.standard_normal((n, 6)) * 1e-3).astype(np.float32)
    visco = 1e-3
    acte = to_cte(s["cte"], tboundary=N.BC_MDBC)
    assert lib.dsph_set_constants(ctx, C.byref(acte)) == 0
    bc = Dev(lib, ctx, s["be"])
    dv = divdata(s["celldiv"], bc, s["cte"])
    bn2 = np.ascontiguousarray(s["bn2"], np.float32)
    gd = {k: Dev(lib, ctx, a) for k, a in (("dcell", s["dcell"]), ("poscell", s["poscell"]),
                                          ("velrho", o_vr), ("code", s["d"]["code"]),
                                          ("bm", o_bm), ("tv", o_tv), ("tau", tau),
                                          ("bn", bn2))}
    mv = np.zeros((n, 3), np.float32)
    if moving:
        mv[:npb] = (np.random.default_rng(4).standard_normal((npb, 3)) * 0.5).astype(np.float32)
    d_mv = Dev(lib, ctx, mv)
    ace = Dev.zeros(lib, ctx, (n, 3), np.float32)
    ar = Dev.zeros(lib, ctx, n, np.float32)
    vdt = Dev.zeros(lib, ctx, n, np.float32)
    strain = Dev.zeros(lib, ctx, (n, 6), np.float32)
    ns = Dev(lib, ctx, np.full((n, 4), 7.0, np.float32))   # fluid rows are all rewritten
    t = N.DsphInterParms()
    t.tkernel, t.tvisco, t.tdensity = N.KERNEL_WENDLAND, tvisco, tddt
    t.viscob, t.viscof = visco, visco
    t.np, t.npb, t.npbok = n, npb, npbok
    t.divdata = dv
    t.d_dcell, t.d_poscell, t.d_velrho, t.d_code = (gd["dcell"].ptr, gd["poscell"].ptr,
                                                     gd["velrho"].ptr, gd["code"].ptr)
    t.d_spstaurho2 = gd["tau"].ptr if tvisco == N.VISCO_LAMINARSPS else None
    t.d_viscdt, t.d_ar, t.d_ace, t.d_sps2strain = vdt.ptr, ar.ptr, ace.ptr, strain.ptr
    t.d_boundmode, t.d_tangenvel = gd["bm"].ptr, gd["tv"].ptr
    t.d_boundnor, t.d_nopenshift = gd["bn"].ptr, ns.ptr
    t.d_motionvel = d_mv.ptr if moving else None
    assert lib.dsph_interaction_forces(ctx, C.byref(t)) == 0, lib.dsph_last_error(ctx)
    g = {"ace": ace.get(), "ar": ar.get(), "viscdt": vdt.get(), "ns": ns.get()}
    oa = np.zeros((n, 3), np.float32)
    oar = np.zeros(n, np.float32)
    ov = np.zeros(n, np.float32)
    os_ = np.zeros((n, 6), np.float32)
    ons = np.zeros((n, 4), np.float32)
    orc.load().orc_interaction_range_mv(
        0, orc.SEM_GPU, tvisco, tddt, visco, visco, C.byref(s["cte"]), C.byref(s["odiv"]), n,
        npb, npbok, s["dcell"].ctypes.data, s["poscell"].ctypes.data, s["pos"].ctypes.data,
        o_vr.ctypes.data, tau.ctypes.data if tvisco == N.VISCO_LAMINARSPS else None,
        oa.ctypes.data, oar.ctypes.data, ov.ctypes.data, os_.ctypes.data, None,
        o_bm.ctypes.data, o_tv.ctypes.data, 0, None, None, None, bn2.ctypes.data,
        ons.ctypes.data, mv.ctypes.data if moving else None, 0, 2**62, 0, 2**62)
    fl = slice(npb, None)
    e = {"ace": rel_err(g["ace"][fl], oa[fl]), "ar": rel_err(g["ar"][fl], oar[fl]),
         "viscdt": rel_err(g["viscdt"], ov)}
    gw, ow = g["ns"][fl, 3], ons[fl, 3]
    same = gw == ow
    flagged = same & (ow > 5)
    e["nopenshift"] = rel_err(g["ns"][fl][flagged, :3], ons[fl][flagged, :3])
    print(f"tvisco {tvisco} tddt {tddt}: {int(flagged.sum())} particles corrected, "
          f"{int((~same).sum())} flags differ; {e}", flush=True)
    assert flagged.sum() > 100
    assert same.mean() > 0.999
    assert np.all(g["ns"][fl][~(gw > 5)] == 0)      # unflagged: cleared, w = 0
    for k, v in e.items():
        assert v < 1e-5, (k, v)


def test_no_penetration_step_matches_oracle(lib, ctx):
    """dsph_mdbc2_nopen_step on random flags and corrections (some axes 0)
    equals the oracle's restatement bit for bit."""
    n, npb, dt = 5000, 700, 1.3e-4
    rng = np.random.default_rng(3)
    ns = (rng.standard_normal((n, 4)) * 0.2).astype(np.float32)
    ns[:, 3] = np.where(rng.random(n) < 0.6, 10.0, 0.0)
    ns[rng.random((n, 4)) < 0.3] = 0.0
    base = rng.standard_normal((n, 4)).astype(np.float32)
    vr = rng.standard_normal((n, 4)).astype(np.float32)
    mov = rng.standard_normal((n, 3))
    ovr, omov = vr.copy(), mov.copy()
    orc.load().orc_nopen_step(n, npb, ns.ctypes.data, base.ctypes.data, dt, ovr.ctypes.data,
                              omov.ctypes.data)
    d_vr = Dev(lib, ctx, vr)
    d_xy = Dev(lib, ctx, np.ascontiguousarray(mov[:, :2]))
    d_z = Dev(lib, ctx, np.ascontiguousarray(mov[:, 2]))
    rc = lib.dsph_mdbc2_nopen_step(ctx, n, npb, Dev(lib, ctx, ns).ptr, Dev(lib, ctx, base).ptr,
                                   dt, None, 0, d_vr.ptr, d_xy.ptr, d_z.ptr)
    assert rc == 0, lib.dsph_last_error(ctx)
    np.testing.assert_array_equal(d_vr.get(), ovr)
    np.testing.assert_array_equal(d_xy.get(), omov[:, :2])
    np.testing.assert_array_equal(d_z.get(), omov[:, 2])
    assert not np.array_equal(ovr, vr)
