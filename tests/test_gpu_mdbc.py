"""mDBC with SlipMode vel=0 on the GPU (k_mdbc, the boundary of the wave
flume) against the oracle, kernel by kernel: the density of every boundary
particle extrapolated from its ghost node (JSphGpu_mdbc_iker.cu:34-150), in
3-D (4x4 systems) and 2-D (3x3 systems over (1, x, z)).

Tolerance: densities of ~1000 kg/m^3 from the 4x4 first-order system, whose
inverse is rounded to FP32 (the reference GPU build's tmatrix4f). The
rounding of the sums is amplified by the condition number of the moment
matrix near walls: on these cases the FP32 oracle itself is up to
5.6e-3 kg/m^3 from the FP64 yardstick (and 2.1e-3 from the exact linear
field, test_mdbc.py). Per particle, the GPU must agree with the FP32 oracle
within that spread (1.5x the oracle's FP32-vs-FP64 maximum on the same case,
and 8e-3 kg/m^3 = 8e-6 relative at most), and be as close to the FP64
yardstick as the FP32 oracle is (2x + 1e-4)."""
import ctypes as C

import numpy as np
import pytest

import oracle.oracle as orc
from dualsphysics_amd import _native as N
from dualsphysics_amd import case as casemod
from tests.helpers import Dev, divdata, mdbc2_state, perturbed_case, to_cte
from tests.test_mdbc import DEPTH, linear_layer, oracle_mdbc

pytestmark = pytest.mark.gpu

TOL_RHO = 8e-3


def _gpu_mdbc(lib, ctx, s):
    acte = to_cte(s["cte"], tboundary=N.BC_MDBC)
    assert lib.dsph_set_constants(ctx, C.byref(acte)) == 0
    bc = Dev(lib, ctx, s["be"])
    dv = divdata(s["celldiv"], bc, s["cte"])
    pos = s["pos"]
    d_vr = Dev(lib, ctx, s["d"]["velrho"])
    mp = (C.c_double * 3)(*s["cte"].mapposmin)
    rc = lib.dsph_mdbc_correction(ctx, s["npb"], C.byref(dv), mp,
                                  Dev(lib, ctx, np.ascontiguousarray(pos[:, :2])).ptr,
                                  Dev(lib, ctx, np.ascontiguousarray(pos[:, 2])).ptr,
                                  Dev(lib, ctx, s["poscell"]).ptr, Dev(lib, ctx, s["d"]["code"]).ptr,
                                  Dev(lib, ctx, s["bn2"]).ptr, d_vr.ptr)
    assert rc == 0, lib.dsph_last_error(ctx)
    return d_vr.get()


def _cases():
    p = casemod.Parameters(tboundary=N.BC_MDBC, tvisco=N.VISCO_LAMINARSPS, visco=1e-6)
    return {
        "flume_dp0.01": lambda: casemod.flume(dp=0.01, params=p),
        "perturbed_dam_break_dp0.02": lambda: perturbed_case(casemod.dam_break(dp=0.02, params=p)),
        "linear_layer_dp0.02": linear_layer,
        "perturbed_dam_break_2d_dp0.005": lambda: perturbed_case(
            casemod.dam_break_2d(dp=0.005, params=p)),
    }


@pytest.mark.parametrize("name", list(_cases()))
def test_mdbc_correction_matches_oracle(lib, ctx, name):
    c = _cases()[name]()
    s = mdbc2_state(c)
    npb = s["npb"]
    print(f"case {name}: {s['n']} particles, {npb} boundary", flush=True)
    g = _gpu_mdbc(lib, ctx, s)
    o32 = oracle_mdbc(s, 0, orc.SEM_GPU)
    o64 = oracle_mdbc(s, 1, orc.SEM_F64)
    e32 = np.abs(g[:npb, 3].astype(np.float64) - o32[:npb, 3])
    e64 = np.abs(g[:npb, 3].astype(np.float64) - o64[:npb, 3])
    eo = np.abs(o32[:npb, 3].astype(np.float64) - o64[:npb, 3])
    changed = np.count_nonzero(g[:npb, 3] != s["d"]["velrho"][:npb, 3])
    print(f"{name}: {changed}/{npb} densities corrected; |gpu-oracle32| max {e32.max():.2e} "
          f"(bit-equal {np.mean(e32 == 0):.3f}); |gpu-fp64| max {e64.max():.2e}; "
          f"|oracle32-fp64| max {eo.max():.2e}")
    assert changed > npb // 10
    assert e32.max() < min(TOL_RHO, max(1.5 * eo.max(), 1e-3))
    assert e64.max() < 2 * eo.max() + 1e-4
    # Fluid and the velocities of boundary particles are untouched.
    np.testing.assert_array_equal(g[npb:], s["d"]["velrho"][npb:])
    np.testing.assert_array_equal(g[:npb, :3], s["d"]["velrho"][:npb, :3])
    if name.startswith("linear_layer"):
        pb = s["pos"][:npb]
        bn = s["bn2"][:npb]
        inner = ((bn[:, 0] == 0) & (bn[:, 1] == 0) & (bn[:, 2] > 0) & (pb[:, 0] > 0.15) &
                 (pb[:, 0] < 0.45) & (pb[:, 1] > 0.15) & (pb[:, 1] < 0.45))
        rho_exact = c.rhop0 + c.rhop0 * 9.81 * (DEPTH - pb[:, 2]) / c.cs0 ** 2
        assert np.abs(g[:npb, 3] - rho_exact)[inner].max() < TOL_RHO
