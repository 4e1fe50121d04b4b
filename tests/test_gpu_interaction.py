"""Force interaction on the GPU vs the oracle (reference rows a12-a17, a25).

Parity metric: norm-wise relative error max|x_gpu - x_oracle| / max|x_oracle|
(per component set) with the oracle in the reference's FP32 semantics; each
result is also stated against the FP64 yardstick (FP64 pair arithmetic on
FP64 positions). Tolerances are written in the tests.
"""
import ctypes as C

import numpy as np
import pytest

from dualsphysics_amd import _native as N
from dualsphysics_amd import case as casemod
from tests.helpers import Dev, divdata, perturbed_case, rel_err, to_cte, per_particle_rel

pytestmark = pytest.mark.gpu

TOL_FP32 = 1e-5      # GPU vs FP32 oracle, norm-wise relative
TOL_PP = 1e-4        # GPU vs FP32 oracle, per particle (floor 1e-3 RMS)


def _prepared(c):
    from oracle.oracle import OracleSim
    o = OracleSim(c)
    d = o.download()
    dcell, poscell, cte = o.arrays()
    celldiv, be = o.celldiv()
    return o, d, dcell, poscell, cte, celldiv, be


def _run_both(c, tvisco, tddt, visco, viscob, lib, ctx, with_tau=False, kernel=N.FORCES_TILE):
    import oracle.oracle as orc
    assert lib.dsph_ctx_set_option(ctx, N.OPT_FORCES_KERNEL, kernel) == 0
    o, d, dcell, poscell, cte, celldiv, be = _prepared(c)
    st = o.stats()
    n, npb, npbok = st["np"], st["npb"], st["npbok"]
    rng = np.random.default_rng(2)
    tau = (rng.standard_normal((n, 6)) * 1e-3).astype(np.float32) if with_tau else None
    # GPU
    bc = Dev(lib, ctx, be)
    dv = divdata(celldiv, bc, cte)
    acte = to_cte(cte)
    assert lib.dsph_set_constants(ctx, C.byref(acte)) == 0
    gd = {k: Dev(lib, ctx, a) for k, a in (("dcell", dcell), ("poscell", poscell),
                                          ("velrho", d["velrho"]), ("code", d["code"]))}
    ace = Dev.zeros(lib, ctx, (n, 3), np.float32)
    ar = Dev.zeros(lib, ctx, n, np.float32)
    vdt = Dev.zeros(lib, ctx, n, np.float32)
    strain = Dev.zeros(lib, ctx, (n, 6), np.float32)
    taud = Dev(lib, ctx, tau) if with_tau else None
    t = N.DsphInterParms()
    t.tkernel, t.tvisco, t.tdensity = N.KERNEL_WENDLAND, tvisco, tddt
    t.viscob, t.viscof = viscob, visco
    t.np, t.npb, t.npbok = n, npb, npbok
    t.divdata = dv
    t.d_dcell, t.d_poscell, t.d_velrho, t.d_code = (gd["dcell"].ptr, gd["poscell"].ptr,
                                                     gd["velrho"].ptr, gd["code"].ptr)
    t.d_spstaurho2 = taud.ptr if with_tau else None
    t.d_viscdt, t.d_ar, t.d_ace, t.d_sps2strain = vdt.ptr, ar.ptr, ace.ptr, strain.ptr
    rc = lib.dsph_interaction_forces(ctx, C.byref(t))
    assert rc == 0, lib.dsph_last_error(ctx)
    g = {"ace": ace.get(), "ar": ar.get(), "viscdt": vdt.get(), "strain": strain.get()}
    # Oracle FP32 (GPU semantics) and FP64 yardstick.
    olib = orc.load()
    odiv = orc.OrcDiv(1, int(dv.ncx), int(dv.ncy), int(dv.ncz), int(dv.ncw), int(dv.cellfluid),
                      int(dv.cellzerox), int(dv.cellzeroy), int(dv.cellzeroz),
                      be.ctypes.data)
    res = {}
    for prec, sem in ((0, orc.SEM_GPU), (1, orc.SEM_F64)):
        oa = np.zeros((n, 3), np.float32)
        oar = np.zeros(n, np.float32)
        ov = np.zeros(n, np.float32)
        os_ = np.zeros((n, 6), np.float32)
        olib.orc_interaction(prec, sem, tvisco, tddt, viscob, visco, C.byref(cte), C.byref(odiv),
                             n, npb, npbok, dcell.ctypes.data, poscell.ctypes.data,
                             np.ascontiguousarray(d["pos"]).ctypes.data, d["velrho"].ctypes.data,
                             tau.ctypes.data if with_tau else None, oa.ctypes.data,
                             oar.ctypes.data, ov.ctypes.data, os_.ctypes.data, None)
        res[prec] = {"ace": oa, "ar": oar, "viscdt": ov, "strain": os_}
    return g, res, npb


CASES = [
    ("hydrostatic_art_ddt2full", False, 1, 3, 0.01, 0.01),
    ("perturbed_art_ddt2full", True, 1, 3, 0.01, 0.01),
    ("perturbed_art_ddt2", True, 1, 2, 0.01, 0.01),
    ("perturbed_art_ddt1", True, 1, 1, 0.01, 0.01),
    ("perturbed_art_noddt", True, 1, 0, 0.05, 0.02),
    ("perturbed_laminar", True, 3, 3, 1e-3, 1e-3),
    ("perturbed_laminarsps", True, 2, 3, 1e-3, 1e-3),
]


@pytest.mark.parametrize("kernel", [N.FORCES_TILE, N.FORCES_ROWS], ids=["tile", "rows"])
@pytest.mark.parametrize("name,pert,tvisco,tddt,visco,viscob", CASES, ids=[x[0] for x in CASES])
def test_interaction_parity(lib, ctx, name, pert, tvisco, tddt, visco, viscob, kernel):
    c = casemod.dam_break(dp=0.02)
    if pert:
        c = perturbed_case(c)
    g, res, npb = _run_both(c, tvisco, tddt, visco, viscob, lib, ctx, with_tau=(tvisco == 2),
                            kernel=kernel)
    o32, o64 = res[0], res[1]
    fl = slice(npb, None)
    e_ace = rel_err(g["ace"][fl], o32["ace"][fl])
    e_ar = rel_err(g["ar"], o32["ar"])
    e_vdt = rel_err(g["viscdt"], o32["viscdt"])
    e64_gpu = rel_err(g["ace"][fl], o64["ace"][fl])
    e64_orc = rel_err(o32["ace"][fl], o64["ace"][fl])
    e64_ar_gpu = rel_err(g["ar"], o64["ar"])
    e64_ar_orc = rel_err(o32["ar"], o64["ar"])
    print(f"{name}: gpu-vs-fp32 ace {e_ace:.2e} ar {e_ar:.2e} viscdt {e_vdt:.2e} | "
          f"vs-fp64 ace gpu {e64_gpu:.2e} oracle32 {e64_orc:.2e} ar gpu {e64_ar_gpu:.2e} "
          f"oracle32 {e64_ar_orc:.2e}")
    assert e_ace < TOL_FP32
    # In the hydrostatic state ar is pure rounding noise around 0 (norm-wise
    # relative error is then meaningless): there the GPU must stay inside the
    # FP32 noise band of the oracle itself.
    assert e_ar < max(TOL_FP32, 3 * e64_ar_orc if not pert else 0.0)
    assert e_vdt < TOL_FP32
    # GPU FP32 is as close to FP64 as the FP32 oracle is (within 2x + 1e-6).
    assert e64_gpu < 2 * e64_orc + 1e-6
    if pert:
        assert e64_ar_gpu < 2 * e64_ar_orc + 1e-6
    pp = per_particle_rel(g["ace"][fl], o32["ace"][fl])
    assert np.max(pp) < TOL_PP, np.max(pp)
    if tvisco == 2:
        assert rel_err(g["strain"][fl], o32["strain"][fl]) < TOL_FP32


def test_tile_kernel_equals_rows_kernel_bitwise(lib, ctx):
    """Both fluid kernels sum the same pairs in the same order: identical bits."""
    c = perturbed_case(casemod.dam_break(dp=0.01), seed=3)
    gq, _, npb = _run_both(c, 1, 3, 0.01, 0.01, lib, ctx, kernel=N.FORCES_TILE)
    gr, _, _ = _run_both(c, 1, 3, 0.01, 0.01, lib, ctx, kernel=N.FORCES_ROWS)
    lib.dsph_ctx_set_option(ctx, N.OPT_FORCES_KERNEL, N.FORCES_TILE)
    for k in ("ace", "ar", "viscdt"):
        np.testing.assert_array_equal(gq[k].view(np.uint32), gr[k].view(np.uint32))


def test_pips_counts_exact(lib, ctx):
    import oracle.oracle as orc
    c = perturbed_case(casemod.dam_break(dp=0.02))
    o, d, dcell, poscell, cte, celldiv, be = _prepared(c)
    st = o.stats()
    bc = Dev(lib, ctx, be)
    dv = divdata(celldiv, bc, cte)
    res = (C.c_uint64 * 4)()
    rc = lib.dsph_pips_count_sync(ctx, st["npbok"], 0, st["np"] - st["npb"], st["npb"],
                                  C.byref(dv), Dev(lib, ctx, dcell).ptr, Dev(lib, ctx, poscell).ptr,
                                  res)
    assert rc == 0
    assert list(res) == o.pips()


def test_sps_tau(lib, ctx):
    import oracle.oracle as orc
    n, npb = 5000, 1000
    rng = np.random.default_rng(9)
    s2 = (rng.standard_normal((n, 6)) * 10).astype(np.float32)
    vr = np.ones((n, 4), np.float32)
    vr[:, 3] = (1000 + rng.standard_normal(n)).astype(np.float32)
    tau = Dev.zeros(lib, ctx, (n, 6), np.float32)
    smag, blin = 1.9e-6, 2.2e-7
    rc = lib.dsph_sps_tau(ctx, n, npb, smag, blin, Dev(lib, ctx, vr).ptr, Dev(lib, ctx, s2).ptr,
                          tau.ptr)
    assert rc == 0
    ot = np.zeros((n, 6), np.float32)
    orc.load().orc_sps_tau(n, npb, smag, blin, vr.ctypes.data, s2.ctypes.data, ot.ctypes.data)
    assert rel_err(tau.get()[npb:], ot[npb:]) < 1e-6


def test_boundary_kernel_forms_equal_bitwise(lib, ctx):
    """The boundary interaction's tile form and its one-lane-per-particle
    form give the same bits, and so does the automatic choice between them
    over the interactions in which it times both (DSPH_OPT_BOUND_KERNEL)."""
    c = perturbed_case(casemod.dam_break(dp=0.01), seed=4)
    o, d, dcell, poscell, cte, celldiv, be = _prepared(c)
    st = o.stats()
    n, npb, npbok = st["np"], st["npb"], st["npbok"]
    bc = Dev(lib, ctx, be)
    dv = divdata(celldiv, bc, cte)
    assert lib.dsph_set_constants(ctx, C.byref(to_cte(cte))) == 0
    gd = {k: Dev(lib, ctx, a) for k, a in (("dcell", dcell), ("poscell", poscell),
                                          ("velrho", d["velrho"]), ("code", d["code"]))}
    outs = []
    try:
        for form, reps in ((N.BOUND_TILE, 1), (N.BOUND_ROWS, 1), (N.BOUND_AUTO, 7)):
            assert lib.dsph_ctx_set_option(ctx, N.OPT_BOUND_KERNEL, form) == 0
            for _ in range(reps):
                ace = Dev.zeros(lib, ctx, (n, 3), np.float32)
                ar = Dev.zeros(lib, ctx, n, np.float32)
                vdt = Dev.zeros(lib, ctx, n, np.float32)
                t = N.DsphInterParms()
                t.tkernel, t.tvisco, t.tdensity = N.KERNEL_WENDLAND, 1, 3
                t.viscob = t.viscof = 0.01
                t.np, t.npb, t.npbok = n, npb, npbok
                t.divdata = dv
                t.d_dcell, t.d_poscell = gd["dcell"].ptr, gd["poscell"].ptr
                t.d_velrho, t.d_code = gd["velrho"].ptr, gd["code"].ptr
                t.d_viscdt, t.d_ar, t.d_ace = vdt.ptr, ar.ptr, ace.ptr
                assert lib.dsph_interaction_forces(ctx, C.byref(t)) == 0, lib.dsph_last_error(ctx)
                outs.append((ar.get(), vdt.get()))
    finally:
        lib.dsph_ctx_set_option(ctx, N.OPT_BOUND_KERNEL, N.BOUND_AUTO)
    assert np.abs(outs[0][0][:npbok]).max() > 0
    for a, v in outs[1:]:
        np.testing.assert_array_equal(a.view(np.uint32), outs[0][0].view(np.uint32))
        np.testing.assert_array_equal(v.view(np.uint32), outs[0][1].view(np.uint32))
    assert lib.dsph_ctx_set_option(ctx, N.OPT_BOUND_KERNEL, 3) != 0
