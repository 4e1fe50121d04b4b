"""Standard shifting on the GPU against the oracle: the pair sums of the
fluid interaction (tile and rows kernels, which also agree with each other
bit for bit), and the displacement kernel. The whole step against the
reference CPU build is in test_gpu_vs_reference.py (shifting fixtures)."""
import ctypes as C

import numpy as np
import pytest

import oracle.oracle as orc
from dualsphysics_amd import _native as N
from dualsphysics_amd import case as casemod
from tests.helpers import Dev, divdata, perturbed_case, rel_err, to_cte

pytestmark = pytest.mark.gpu


def _state():
    from oracle.oracle import OracleSim
    c = perturbed_case(casemod.flume(dp=0.012, width=0.3))
    o = OracleSim(c)
    d = o.download()
    dcell, poscell, cte = o.arrays()
    c6, be = o.celldiv()
    st = o.stats()
    return c, d, dcell, poscell, cte, c6, be, st


@pytest.mark.parametrize("shiftmode", [1, 2, 3])
def test_shifting_sums_match_oracle_and_kernels_agree(lib, ctx, shiftmode):
    c, d, dcell, poscell, cte, c6, be, st = _state()
    n, npb, npbok = st["np"], st["npb"], st["npbok"]
    acte = to_cte(cte)
    assert lib.dsph_set_constants(ctx, C.byref(acte)) == 0
    bc = Dev(lib, ctx, be)
    dv = divdata(c6, bc, cte)
    out = {}
    for kernel in (N.FORCES_TILE, N.FORCES_ROWS):
        assert lib.dsph_ctx_set_option(ctx, N.OPT_FORCES_KERNEL, kernel) == 0
        gd = {k: Dev(lib, ctx, a) for k, a in (("dcell", dcell), ("poscell", poscell),
                                              ("velrho", d["velrho"]), ("code", d["code"]))}
        sh = Dev.zeros(lib, ctx, (n, 4), np.float32)
        ace = Dev.zeros(lib, ctx, (n, 3), np.float32)
        ar = Dev.zeros(lib, ctx, n, np.float32)
        vdt = Dev.zeros(lib, ctx, n, np.float32)
        t = N.DsphInterParms()
        t.tkernel, t.tvisco, t.tdensity = N.KERNEL_WENDLAND, N.VISCO_ARTIFICIAL, N.DDT_DDT2FULL
        t.viscob = t.viscof = 0.01
        t.np, t.npb, t.npbok = n, npb, npbok
        t.divdata = dv
        t.d_dcell, t.d_poscell, t.d_velrho, t.d_code = (gd["dcell"].ptr, gd["poscell"].ptr,
                                                         gd["velrho"].ptr, gd["code"].ptr)
        t.d_viscdt, t.d_ar, t.d_ace = vdt.ptr, ar.ptr, ace.ptr
        t.shiftmode, t.d_shiftposfs = shiftmode, sh.ptr
        assert lib.dsph_interaction_forces(ctx, C.byref(t)) == 0, lib.dsph_last_error(ctx)
        out[kernel] = (sh.get(), ace.get())
    assert lib.dsph_ctx_set_option(ctx, N.OPT_FORCES_KERNEL, N.FORCES_TILE) == 0
    np.testing.assert_array_equal(out[N.FORCES_TILE][0], out[N.FORCES_ROWS][0])
    np.testing.assert_array_equal(out[N.FORCES_TILE][1], out[N.FORCES_ROWS][1])
    # Oracle, GPU semantics.
    odiv = orc.OrcDiv(1, int(dv.ncx), int(dv.ncy), int(dv.ncz), int(dv.ncw), int(dv.cellfluid),
                      int(dv.cellzerox), int(dv.cellzeroy), int(dv.cellzeroz), be.ctypes.data)
    osh = np.zeros((n, 4), np.float32)
    z = lambda *s: np.zeros(s, np.float32)
    oa, oar, ov = z(n, 3), z(n), z(n)
    orc.load().orc_interaction_ext(
        0, orc.SEM_GPU, N.VISCO_ARTIFICIAL, N.DDT_DDT2FULL, 0.01, 0.01, C.byref(cte),
        C.byref(odiv), n, npb, npbok, dcell.ctypes.data, poscell.ctypes.data,
        np.ascontiguousarray(d["pos"]).ctypes.data, d["velrho"].ctypes.data, None, oa.ctypes.data,
        oar.ctypes.data, ov.ctypes.data, None, None, None, None, shiftmode, d["code"].ctypes.data,
        osh.ctypes.data)
    g = out[N.FORCES_TILE][0][npb:]
    o = osh[npb:]
    cancelled = o[:, 0] == np.float32(np.finfo(np.float32).max)
    assert np.array_equal(g[:, 0] == np.finfo(np.float32).max, cancelled)
    if shiftmode == 3:
        assert not cancelled.any()
    else:
        assert cancelled.any()
    ok = ~cancelled
    e = rel_err(g[ok], o[ok])
    ew = rel_err(g[:, 3], o[:, 3])
    print(f"shiftmode {shiftmode}: {cancelled.sum()} cancelled, sums {e:.2e}, divergence {ew:.2e}")
    assert e < 1e-5 and ew < 1e-5


@pytest.mark.parametrize("sim2d", [0, 1], ids=["3d", "2d"])
def test_shifting_displacement_matches_oracle(lib, ctx, sim2d):
    """The displacement pass: free-surface factor (w - tfs)/(3 - tfs), or
    (w - tfs)/(2 - tfs) in 2-D (JSphShifting.cpp:467), from the context's
    constants."""
    cte = N.DsphCte()
    cte.simulate2d = sim2d
    assert lib.dsph_set_constants(ctx, C.byref(cte)) == 0
    rng = np.random.default_rng(1)
    n, pini = 5000, 100
    vr = np.zeros((n + pini, 4), np.float32)
    vr[:, :3] = rng.standard_normal((n + pini, 3)).astype(np.float32)
    sh = (rng.standard_normal((n + pini, 4)) * 50).astype(np.float32)
    sh[:, 3] = rng.uniform(0, 3, n + pini).astype(np.float32)
    sh[rng.random(n + pini) < 0.1, 0] = np.finfo(np.float32).max
    for tfs in (0.0, 1.5):
        d_vr = Dev(lib, ctx, vr)
        d_sh = Dev(lib, ctx, sh)
        assert lib.dsph_run_shifting(ctx, n, pini, 1e-4, None, 0, 1.0, -2.0, float(np.float32(0.0249)),
                                     tfs, 0.01, d_vr.ptr, d_sh.ptr) == 0
        got = d_sh.get()
        ref = sh.copy()
        orc.load().orc_run_shifting(n, pini, 1e-4, -2.0, float(np.float32(0.0249)), tfs, 0.01,
                                    vr.ctypes.data, ref.ctypes.data, sim2d)
        np.testing.assert_array_equal(got[:pini], sh[:pini])
        np.testing.assert_allclose(got[pini:, :3], ref[pini:, :3], rtol=2e-7, atol=1e-12)
        assert (np.abs(got[pini:, :3]) <= np.float32(0.001)).all()


@pytest.mark.parametrize("shiftmode", [2, 3])
def test_shifting_sums_with_mdbc2_modes_match_oracle(lib, ctx, shiftmode):
    """Shifting with mDBC2 boundaries (the one-lane-per-particle kernel with
    both pair rules): random boundary modes, a wall particle switched off
    (mode 2) having mass 0 in the shifting sums too; sums, divergence, ace and
    ar against the oracle in GPU semantics."""
    c, d, dcell, poscell, cte, c6, be, st = _state()
    n, npb, npbok = st["np"], st["npb"], st["npbok"]
    acte = to_cte(cte, tboundary=N.BC_MDBC)
    assert lib.dsph_set_constants(ctx, C.byref(acte)) == 0
    bc = Dev(lib, ctx, be)
    dv = divdata(c6, bc, cte)
    rng = np.random.default_rng(shiftmode)
    bm = rng.integers(1, 3, npb).astype(np.uint8)
    tv = (rng.standard_normal((npb, 3)) * 0.1).astype(np.float32)
    gd = {k: Dev(lib, ctx, a) for k, a in (("dcell", dcell), ("poscell", poscell),
                                          ("velrho", d["velrho"]), ("code", d["code"]),
                                          ("bm", bm), ("tv", tv))}
    sh = Dev.zeros(lib, ctx, (n, 4), np.float32)
    ace = Dev.zeros(lib, ctx, (n, 3), np.float32)
    ar = Dev.zeros(lib, ctx, n, np.float32)
    vdt = Dev.zeros(lib, ctx, n, np.float32)
    t = N.DsphInterParms()
    t.tkernel, t.tvisco, t.tdensity = N.KERNEL_WENDLAND, N.VISCO_LAMINAR, N.DDT_DDT2FULL
    t.viscob = t.viscof = 1e-4
    t.np, t.npb, t.npbok = n, npb, npbok
    t.divdata = dv
    t.d_dcell, t.d_poscell, t.d_velrho, t.d_code = (gd["dcell"].ptr, gd["poscell"].ptr,
                                                     gd["velrho"].ptr, gd["code"].ptr)
    t.d_viscdt, t.d_ar, t.d_ace = vdt.ptr, ar.ptr, ace.ptr
    t.d_boundmode, t.d_tangenvel = gd["bm"].ptr, gd["tv"].ptr
    t.shiftmode, t.d_shiftposfs = shiftmode, sh.ptr
    assert lib.dsph_interaction_forces(ctx, C.byref(t)) == 0, lib.dsph_last_error(ctx)
    odiv = orc.OrcDiv(1, int(dv.ncx), int(dv.ncy), int(dv.ncz), int(dv.ncw), int(dv.cellfluid),
                      int(dv.cellzerox), int(dv.cellzeroy), int(dv.cellzeroz), be.ctypes.data)
    osh = np.zeros((n, 4), np.float32)
    z = lambda *s: np.zeros(s, np.float32)
    oa, oar, ov = z(n, 3), z(n), z(n)
    cte.tboundary = N.BC_MDBC
    orc.load().orc_interaction_ext(
        0, orc.SEM_GPU, N.VISCO_LAMINAR, N.DDT_DDT2FULL, 1e-4, 1e-4, C.byref(cte),
        C.byref(odiv), n, npb, npbok, dcell.ctypes.data, poscell.ctypes.data,
        np.ascontiguousarray(d["pos"]).ctypes.data, d["velrho"].ctypes.data, None, oa.ctypes.data,
        oar.ctypes.data, ov.ctypes.data, None, None, bm.ctypes.data, tv.ctypes.data, shiftmode,
        d["code"].ctypes.data, osh.ctypes.data)
    g, o = sh.get()[npb:], osh[npb:]
    cancelled = o[:, 0] == np.float32(np.finfo(np.float32).max)
    assert np.array_equal(g[:, 0] == np.finfo(np.float32).max, cancelled)
    ok = ~cancelled
    e = {"sums": rel_err(g[ok, :3], o[ok, :3]), "divergence": rel_err(g[:, 3], o[:, 3]),
         "ace": rel_err(ace.get()[npb:], oa[npb:]), "ar": rel_err(ar.get()[npb:], oar[npb:])}
    print(f"shiftmode {shiftmode} with mDBC2 modes ({int((bm == 2).sum())} off): "
          f"{int(cancelled.sum())} cancelled, {e}")
    for k, v in e.items():
        assert v < 1e-5, (k, v)
