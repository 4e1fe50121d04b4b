"""dsph_interaction_forces_noalloc (the interaction without any device
allocation, its scratch in the library's static device array) against
dsph_interaction_forces: the same kernels, so the same bits, on two calls in
a row on a context that never reserved scratch, the second with fewer
particles; np above DSPH_NOALLOC_NP_MAX is refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from dualsphysics_amd import _native as N
from dualsphysics_amd import case as casemod
from tests.helpers import Dev, divdata, perturbed_case, release_all, to_cte

pytestmark = pytest.mark.gpu


def _inputs(lib, ctx, c, tvisco):
    """Oracle-prepared (sorted) arrays of case c on the device of ctx and
    the interaction parameters over them (outputs zeroed)."""
    from oracle.oracle import OracleSim
    o = OracleSim(c)
    d = o.download()
    dcell, poscell, cte = o.arrays()
    celldiv, be = o.celldiv()
    st = o.stats()
    o.close()
    n, npb, npbok = st["np"], st["npb"], st["npbok"]
    acte = to_cte(cte)
    assert lib.dsph_set_constants(ctx, C.byref(acte)) == 0, lib.dsph_last_error(ctx)
    bc = Dev(lib, ctx, be)
    t = N.DsphInterParms()
    t.tkernel, t.tvisco, t.tdensity = N.KERNEL_WENDLAND, tvisco, N.DDT_DDT2FULL
    t.viscob, t.viscof = (0.01, 0.01) if tvisco == N.VISCO_ARTIFICIAL else (1e-3, 1e-3)
    t.np, t.npb, t.npbok = n, npb, npbok
    t.divdata = divdata(celldiv, bc, cte)
    keep = [bc]
    for name, a in (("d_dcell", dcell), ("d_poscell", poscell), ("d_velrho", d["velrho"]),
                    ("d_code", d["code"])):
        dv = Dev(lib, ctx, a)
        keep.append(dv)
        setattr(t, name, dv.ptr)
    if tvisco == N.VISCO_LAMINARSPS:
        tau = (np.random.default_rng(5).standard_normal((n, 6)) * 1e-3).astype(np.float32)
        dv = Dev(lib, ctx, tau)
        keep.append(dv)
        t.d_spstaurho2 = dv.ptr
    out = {"ace": Dev.zeros(lib, ctx, (n, 3), np.float32),
           "ar": Dev.zeros(lib, ctx, n, np.float32),
           "viscdt": Dev.zeros(lib, ctx, n, np.float32),
           "strain": Dev.zeros(lib, ctx, (n, 6), np.float32)}
    t.d_ace, t.d_ar, t.d_viscdt = out["ace"].ptr, out["ar"].ptr, out["viscdt"].ptr
    t.d_sps2strain = out["strain"].ptr
    return t, out, keep


def _run(lib, ctx, c, tvisco, entry):
    t, out, _keep = _inputs(lib, ctx, c, tvisco)
    rc = getattr(lib, entry)(ctx, C.byref(t))
    assert rc == 0, (entry, lib.dsph_last_error(ctx))
    assert lib.dsph_synchronize(ctx) == 0
    return t.np, {k: v.get() for k, v in out.items()}


@pytest.mark.parametrize("tvisco", [N.VISCO_ARTIFICIAL, N.VISCO_LAMINARSPS],
                         ids=["artificial", "laminarsps"])
def test_noalloc_equals_interaction_forces(lib, ctx, tvisco):
    fresh = C.c_void_p()
    assert lib.dsph_ctx_create(0, C.byref(fresh)) == 0
    try:
        sizes = []
        for dp, seed in ((0.02, 11), (0.03, 12)):   # second call: fewer particles
            c = perturbed_case(casemod.dam_break(dp=dp), seed=seed)
            n, ref = _run(lib, ctx, c, tvisco, "dsph_interaction_forces")
            n2, got = _run(lib, fresh, c, tvisco, "dsph_interaction_forces_noalloc")
            assert n2 == n
            sizes.append(n)
            print(f"dp={dp}: np={n}", flush=True)
            for k in ("ace", "ar", "viscdt") + (("strain",) if tvisco == N.VISCO_LAMINARSPS
                                                else ()):
                np.testing.assert_array_equal(got[k].view(np.uint32), ref[k].view(np.uint32),
                                              err_msg=f"dp={dp} {k}")
            assert np.abs(ref["ace"]).max() > 0
        assert sizes[1] < sizes[0]
    finally:
        release_all()   # buffers of the fresh context go before it
        lib.dsph_ctx_destroy(fresh)


def test_noalloc_refuses_more_particles_than_its_scratch(lib):
    fresh = C.c_void_p()
    assert lib.dsph_ctx_create(0, C.byref(fresh)) == 0
    try:
        c = casemod.dam_break(dp=0.05)
        t, _out, _keep = _inputs(lib, fresh, c, N.VISCO_ARTIFICIAL)
        t.np = N.NOALLOC_NP_MAX + 1   # refused before any launch
        rc = lib.dsph_interaction_forces_noalloc(fresh, C.byref(t))
        assert rc == N.DSPH_E_ARG
        assert b"DSPH_NOALLOC_NP_MAX" in lib.dsph_last_error(fresh)
    finally:
        release_all()
        lib.dsph_ctx_destroy(fresh)
