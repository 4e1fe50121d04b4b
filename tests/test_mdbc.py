"""mDBC with SlipMode vel=0 (JSphCpu_mdbc.cpp:36-146, GPU
KerInteractionMdbcCorrection_Fast JSphGpu_mdbc_iker.cu:34-150) in the oracle,
without a GPU: the closed form of the first-order correction.

At the ghost node g = x_b + n (n the doubled normal) the 4x4 system
A o = b with A = sum V2 [W, gradW] x [1, g - x2] and b = sum m2 [W, gradW]
is solved; for a density field that is linear in space, rho(x2) =
rho(g) + grad(rho).(x2 - g), b = A [rho(g), -grad(rho)] exactly (m2 = rho2
V2), so o reproduces the field and the boundary particle gets
rho(g) - grad(rho).n = rho(x_b): the linear field at its own position. A
hydrostatic layer with a linear density profile pins it: FP64 to rounding,
FP32 within the amplification of the 4x4 inverse."""
import ctypes as C

import numpy as np
import pytest

import oracle.oracle as orc
from dualsphysics_amd import _native as N
from tests.helpers import hydrostatic_mdbc2_case, mdbc2_state

DEPTH = 0.2


def linear_layer(dp=0.02, dim2=False):
    """Hydrostatic layer of depth DEPTH in a 0.6 m x 0.6 m tank (or its 2-D
    section) with mDBC SlipMode vel=0, density linear in z (see
    hydrostatic_mdbc2_case), at rest."""
    import dataclasses
    c = hydrostatic_mdbc2_case(dp=dp, depth=DEPTH, u=0.0, dim2=dim2)
    prm = dataclasses.replace(c.params, slipmode=N.SLIP_VEL0)
    return dataclasses.replace(c, params=prm)


def oracle_mdbc(s, prec, sem, velrho=None):
    vr = (s["d"]["velrho"] if velrho is None else velrho).copy()
    orc.load().orc_mdbc(prec, sem, C.byref(s["cte"]), C.byref(s["odiv"]), s["npb"],
                        s["pos"].ctypes.data, s["poscell"].ctypes.data,
                        s["d"]["code"].ctypes.data, s["bn2"].ctypes.data, vr.ctypes.data)
    return vr


@pytest.fixture(scope="module")
def layer():
    c = linear_layer()
    return c, mdbc2_state(c)


# (prec, sem, density tolerance in kg/m^3 around rho ~ 1000-1003)
SEMS = [(1, orc.SEM_F64, 1e-6), (0, orc.SEM_GPU, 4e-3), (0, orc.SEM_CPU, 4e-3)]


@pytest.mark.parametrize("prec,sem,tol", SEMS, ids=["fp64", "fp32_gpu", "fp32_cpu"])
def test_floor_density_is_the_linear_field_at_the_particle(layer, prec, sem, tol):
    c, s = layer
    vr = oracle_mdbc(s, prec, sem)
    npb = s["npb"]
    pb = s["pos"][:npb]
    bn = s["bn2"][:npb]
    floor = (bn[:, 0] == 0) & (bn[:, 1] == 0) & (bn[:, 2] > 0)
    inner = floor & (pb[:, 0] > 0.15) & (pb[:, 0] < 0.45) & (pb[:, 1] > 0.15) & (pb[:, 1] < 0.45)
    assert inner.sum() > 100
    rho_exact = c.rhop0 + c.rhop0 * 9.81 * (DEPTH - pb[:, 2]) / c.cs0 ** 2
    err = np.abs(vr[:npb, 3] - rho_exact)[inner]
    print(f"prec {prec} sem {sem}: max |rho_b - rho(x_b)| {err.max():.2e} over {inner.sum()}")
    assert err.max() < tol
    # The wrong sign of the extrapolation would give rho(x_b + 2n): off by
    # 2 |d rho/dz| n_z >= 2 * 12.5 kg/m^4 * 0.02 m = 0.5 kg/m^3, far above tol.
    wrong = c.rhop0 + c.rhop0 * 9.81 * (DEPTH - (pb[:, 2] + 2 * bn[:, 2])) / c.cs0 ** 2
    assert np.abs(vr[:npb, 3] - wrong)[inner].min() > 10 * tol
    # Only densities of boundary particles change.
    np.testing.assert_array_equal(vr[npb:], s["d"]["velrho"][npb:])
    np.testing.assert_array_equal(vr[:npb, :3], s["d"]["velrho"][:npb, :3])


def test_dry_ghost_nodes_keep_rho0(layer):
    """A ghost node without fluid within 2h (A singular, no kernel sum) gives
    the boundary particle rho0 (JSphCpu_mdbc.cpp: rhofinal stays FLT_MAX)."""
    c, s = layer
    vr = oracle_mdbc(s, 1, orc.SEM_F64)
    npb = s["npb"]
    pb = s["pos"][:npb]
    dry = pb[:, 2] > DEPTH + 2 * c.h
    assert dry.sum() > 0
    assert np.all(vr[:npb, 3][dry] == np.float32(c.rhop0))


@pytest.mark.parametrize("prec,sem,tol", SEMS, ids=["fp64", "fp32_gpu", "fp32_cpu"])
def test_2d_floor_density_is_the_linear_field_at_the_particle(prec, sem, tol):
    """2-D (all dy = 0): the 3x3 system over (1, x, z) reproduces the linear
    field as the 4x4 one does in 3-D; the singular 4x4 system (solved as 3-D)
    would fall back to the Shepard sum, off by far more than tol."""
    c = linear_layer(dp=0.01, dim2=True)
    s = mdbc2_state(c)
    assert s["cte"].sim2d == 1
    vr = oracle_mdbc(s, prec, sem)
    npb = s["npb"]
    pb = s["pos"][:npb]
    bn = s["bn2"][:npb]
    floor = (bn[:, 0] == 0) & (bn[:, 2] > 0)
    inner = floor & (pb[:, 0] > 0.15) & (pb[:, 0] < 0.45)
    assert inner.sum() > 40 and np.all(bn[:, 1] == 0)
    rho_exact = c.rhop0 + c.rhop0 * 9.81 * (DEPTH - pb[:, 2]) / c.cs0 ** 2
    err = np.abs(vr[:npb, 3] - rho_exact)[inner]
    print(f"2-D prec {prec} sem {sem}: max |rho_b - rho(x_b)| {err.max():.2e} over {inner.sum()}")
    assert err.max() < tol
    s["cte"].sim2d = 0
    shep = oracle_mdbc(s, prec, sem)
    assert np.abs(shep[:npb, 3] - rho_exact)[inner].min() > 10 * tol
