"""mDBC2 (SlipMode no-slip) in the oracle, without a GPU: closed forms of the
correction on a hydrostatic layer, and the reference-run fixture
distinguishing no-slip from the vel=0 mDBC (the oracle-vs-reference runs of
the fixture are in test_oracle_vs_reference.py)."""
import numpy as np
import pytest

import oracle.oracle as orc
from tests.helpers import (compare_with_reference, hydrostatic_mdbc2_case,
                           load_reference_fixture, mdbc2_state, oracle_mdbc2)

DEPTH, U = 0.2, 0.3


@pytest.fixture(scope="module")
def layer():
    c = hydrostatic_mdbc2_case(depth=DEPTH, u=U)
    return c, mdbc2_state(c)


# (prec, sem, density tolerance): FP64 reproduces the linear field to
# rounding; FP32 within ~16 ulp of 1000 (the 4x4 inverse amplifies rounding).
SEMS = [(1, orc.SEM_F64, 1e-6), (0, orc.SEM_GPU, 2e-3), (0, orc.SEM_CPU, 2e-3)]


@pytest.mark.parametrize("prec,sem,tol", SEMS, ids=["fp64", "fp32_gpu", "fp32_cpu"])
def test_floor_density_and_tangential_velocity_closed_form(layer, prec, sem, tol):
    c, s = layer
    vr, bm, tv = oracle_mdbc2(c, s, prec, sem)
    npb = s["npb"]
    pb = s["pos"][:npb]
    bn = s["bn2"][:npb]
    floor = (bn[:, 0] == 0) & (bn[:, 1] == 0) & (bn[:, 2] > 0)
    inner = floor & (pb[:, 0] > 0.15) & (pb[:, 0] < 0.45) & (pb[:, 1] > 0.15) & (pb[:, 1] < 0.45)
    assert inner.sum() > 100
    rho_exact = c.rhop0 + c.rhop0 * 9.81 * (DEPTH - pb[:, 2]) / c.cs0 ** 2
    assert np.abs(vr[:npb, 3] - rho_exact)[inner].max() < tol
    assert np.all(bm[inner] == 1)
    np.testing.assert_allclose(tv[inner], np.tile([-U, 0.0, 0.0], (inner.sum(), 1)),
                               atol=2e-6 if prec == 0 else 1e-12)
    # Fluid and the velocities of boundary particles are untouched (MDBC2_KEEPVEL).
    np.testing.assert_array_equal(vr[npb:], s["d"]["velrho"][npb:])
    np.testing.assert_array_equal(vr[:npb, :3], s["d"]["velrho"][:npb, :3])


@pytest.mark.parametrize("prec,sem,tol", SEMS, ids=["fp64", "fp32_gpu", "fp32_cpu"])
def test_2d_floor_density_and_tangential_velocity_closed_form(prec, sem, tol):
    """2-D section of the layer (all dy = 0, 3x3 systems over (1, x, z), the
    2-D Wendland constants): the same closed forms as in 3-D."""
    c = hydrostatic_mdbc2_case(dp=0.01, depth=DEPTH, u=U, dim2=True)
    s = mdbc2_state(c)
    assert s["cte"].sim2d == 1
    vr, bm, tv = oracle_mdbc2(c, s, prec, sem)
    npb = s["npb"]
    pb = s["pos"][:npb]
    bn = s["bn2"][:npb]
    inner = (bn[:, 0] == 0) & (bn[:, 2] > 0) & (pb[:, 0] > 0.15) & (pb[:, 0] < 0.45)
    assert inner.sum() > 40
    rho_exact = c.rhop0 + c.rhop0 * 9.81 * (DEPTH - pb[:, 2]) / c.cs0 ** 2
    err = np.abs(vr[:npb, 3] - rho_exact)[inner].max()
    print(f"2-D prec {prec} sem {sem}: max |rho_b - rho(x_b)| {err:.2e}")
    assert err < tol
    assert np.all(bm[inner] == 1)
    np.testing.assert_allclose(tv[inner], np.tile([-U, 0.0, 0.0], (inner.sum(), 1)),
                               atol=2e-6 if prec == 0 else 1e-12)


def test_dry_ghost_nodes_switch_the_particle_off(layer):
    c, s = layer
    vr, bm, tv = oracle_mdbc2(c, s, 1, orc.SEM_F64)
    npb = s["npb"]
    pb = s["pos"][:npb]
    # Side-wall particles whose ghost node has no fluid within 2h.
    h2 = 2 * c.h
    dry = pb[:, 2] > DEPTH + h2
    assert dry.sum() > 0
    assert np.all(bm[dry] == 2)
    assert np.all(vr[:npb, 3][dry] == np.float32(c.rhop0))
    # Tangential part of the motion velocity (0 for fixed walls).
    assert np.all(tv[dry] == 0)
    assert set(np.unique(bm)) <= {1, 2}


def test_wrong_slip_mode_leaves_the_reference_bounds():
    """The mDBC2 fixture is sensitive to the slip mode: the oracle with the
    vel=0 mDBC on the same particles leaves the reference bounds."""
    import dataclasses
    from oracle.oracle import OracleSim
    fx, case = load_reference_fixture("ref_dambreak_dp0.03_mdbc2")
    bad = dataclasses.replace(case, params=dataclasses.replace(case.params, slipmode=1))
    o = OracleSim(bad, prec=1)
    o.run(int(fx["steps"][-1]))
    d = o.download()
    idx = np.argsort(d["idp"], kind="stable")
    with pytest.raises(AssertionError):
        compare_with_reference(fx, len(fx["steps"]) - 1, o.stats()["timestep"], case.pos,
                               d["pos"][idx], d["velrho"][idx, :3], d["velrho"][idx, 3])
