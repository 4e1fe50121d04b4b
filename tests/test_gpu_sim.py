"""Whole Symplectic steps (ComputeStep_Sym + RunCellDivide) on the GPU vs the
oracle, over the first 10 steps of the coarse 3-D dam break (dp=0.02 m).

- forces of every step's state: GPU vs FP32 oracle on the identical state
  (per-particle ace/ar within 1e-5 norm-wise relative), and both vs FP64;
- neighbour order after every divide: identical particle order (bit-exact);
- trajectories: GPU and oracle run independently from the same initial
  particles; positions, velocities and densities compared by particle id.
"""
import dataclasses

import numpy as np
import pytest

from dualsphysics_amd import case as casemod
from tests.helpers import per_particle_rel, rel_err

pytestmark = pytest.mark.gpu


def _state_case(c, state):
    """The case with particles replaced by a downloaded (sorted) state."""
    return dataclasses.replace(c, idp=state["idp"].copy(), pos=state["pos"].copy(),
                               velrho=state["velrho"].copy())


def test_first_10_steps_forces_and_order():
    from dualsphysics_amd.sim import SphGpuSingle
    from oracle.oracle import OracleSim
    c = casemod.dam_break(dp=0.02)
    s = SphGpuSingle(c)
    worst = {"ace": 0.0, "ar": 0.0, "pp_ace": 0.0}
    for step in range(10):
        s.run(1, sync_dt=True)
        s.interact()
        g = s.download()
        gace, gar = s.forces()
        o = OracleSim(_state_case(c, g))
        od = o.download()
        # The oracle re-divides the GPU's state: the order must already be sorted.
        np.testing.assert_array_equal(od["idp"], g["idp"])
        o.interact()
        oace, oar = o.forces()
        o64 = OracleSim(_state_case(c, g), prec=1)
        o64.interact()
        a64, r64 = o64.forces()
        npb = s.stats()["npb"]
        fl = slice(npb, None)
        e_ace = rel_err(gace[fl], oace[fl])
        e_ar = rel_err(gar, oar)
        pp = float(np.max(per_particle_rel(gace[fl], oace[fl])))
        print(f"step {step + 1}: ace {e_ace:.2e} ar {e_ar:.2e} pp-ace {pp:.2e} | vs fp64: "
              f"gpu ace {rel_err(gace[fl], a64[fl]):.2e} ar {rel_err(gar, r64):.2e}; "
              f"oracle32 ace {rel_err(oace[fl], a64[fl]):.2e} ar {rel_err(oar, r64):.2e}")
        worst["ace"] = max(worst["ace"], e_ace)
        worst["ar"] = max(worst["ar"], e_ar)
        worst["pp_ace"] = max(worst["pp_ace"], pp)
    assert worst["ace"] < 1e-5 and worst["ar"] < 1e-5, worst
    assert worst["pp_ace"] < 1e-4, worst
    s.close()


def test_trajectory_matches_oracle():
    from dualsphysics_amd.sim import SphGpuSingle
    from oracle.oracle import OracleSim
    c = casemod.dam_break(dp=0.02)
    s = SphGpuSingle(c)
    o = OracleSim(c)
    s.run(20, sync_dt=True)
    o.run(20)
    gs, os_ = s.stats(), o.stats()
    assert gs["np"] == os_["np"]
    assert abs(gs["timestep"] - os_["timestep"]) <= 1e-6 * os_["timestep"]
    g, d = s.download(), o.download()
    gi, oi = np.argsort(g["idp"]), np.argsort(d["idp"])
    dpos = np.abs(g["pos"][gi] - d["pos"][oi]).max()
    dvel = np.abs(g["velrho"][gi, :3] - d["velrho"][oi, :3]).max()
    vmax = np.abs(d["velrho"][:, :3]).max()
    drho = np.abs(g["velrho"][gi, 3] - d["velrho"][oi, 3]).max()
    print(f"20 steps: max|dpos| {dpos:.2e} m (dp=0.02), max|dv| {dvel:.2e} (vmax {vmax:.2e}),"
          f" max|drho| {drho:.2e}")
    assert dpos < 1e-6 * c.dp * 100
    assert dvel < 1e-4 * vmax + 1e-6
    assert drho < 1e-3


def test_device_dt_equals_host_dt():
    """The graph-friendly mode (dt kept on the device) gives the same run."""
    from dualsphysics_amd.sim import SphGpuSingle
    c = casemod.dam_break(dp=0.02)
    a, b = SphGpuSingle(c), SphGpuSingle(c)
    a.run(8, sync_dt=True)
    b.run(8, sync_dt=False)
    sa, sb = a.stats(), b.stats()
    assert abs(sa["timestep"] - sb["timestep"]) <= 1e-12 * sa["timestep"]
    ga, gb = a.download(), b.download()
    np.testing.assert_array_equal(ga["idp"], gb["idp"])
    np.testing.assert_array_equal(ga["pos"], gb["pos"])
    np.testing.assert_array_equal(ga["velrho"], gb["velrho"])


@pytest.mark.parametrize("wavemaker,tboundary", [("piston", 1), ("flap", 1), ("flap_rot", 2)])
def test_device_motion_equals_host_motion(wavemaker, tboundary):
    """Moving boundaries with dt and the motion law on the device follow the
    host path (the reference's): the same run up to the last bits of the
    device's and the host's sin/cos."""
    from dualsphysics_amd.sim import SphGpuSingle
    prm = casemod.Parameters(tboundary=tboundary)
    c = casemod.flume(dp=0.015, wavemaker=wavemaker, params=prm)
    a, b = SphGpuSingle(c), SphGpuSingle(c)
    a.run(12, sync_dt=True)
    b.run(12, sync_dt=False)
    sa, sb = a.stats(), b.stats()
    assert abs(sa["timestep"] - sb["timestep"]) <= 1e-12 * sa["timestep"]
    ga, gb = a.download(), b.download()
    ia, ib = np.argsort(ga["idp"]), np.argsort(gb["idp"])
    nb, nm = c.nbound, c.nmoving
    mv = slice(nb - nm, nb)
    dmov = np.abs(ga["pos"][ia][mv] - c.pos[mv]).max()
    assert dmov > 1e-6, "the wavemaker did not move"
    assert np.abs(ga["pos"][ia] - gb["pos"][ib]).max() < 1e-12
    assert np.abs(ga["velrho"][ia] - gb["velrho"][ib]).max() < 1e-3
    a.close()
    b.close()


def test_determinism():
    from dualsphysics_amd.sim import SphGpuSingle
    c = casemod.dam_break(dp=0.02)
    a, b = SphGpuSingle(c), SphGpuSingle(c)
    a.run(5)
    b.run(5)
    ga, gb = a.download(), b.download()
    np.testing.assert_array_equal(ga["pos"], gb["pos"])
    np.testing.assert_array_equal(ga["velrho"], gb["velrho"])


def test_pips_matches_oracle():
    from dualsphysics_amd.sim import SphGpuSingle
    from oracle.oracle import OracleSim
    c = casemod.dam_break(dp=0.02)
    s = SphGpuSingle(c)
    o = OracleSim(c)
    g = s.pips()
    assert g[:4] == o.pips()
    assert 0 < g[4] < g[0]


@pytest.mark.parametrize("nfluid", [1, 0])
def test_tiny_and_empty_fluid(nfluid):
    """Edge cases of the step: a fluid of one particle resting on the floor of
    an empty tank (matches the oracle), and the same tank without fluid (runs,
    nothing moves)."""
    from dualsphysics_amd.sim import SphGpuSingle
    from oracle.oracle import OracleSim
    c = casemod.dam_break(dp=0.02, tank=(0.2, 0.2, 0.3), column=(0.02, 0.02, 0.02))
    assert c.nfluid == 1
    if nfluid == 0:
        n = c.nbound
        c = dataclasses.replace(c, idp=c.idp[:n], pos=c.pos[:n], velrho=c.velrho[:n], nfluid=0)
    s = SphGpuSingle(c)
    try:
        s.run(5, sync_dt=True)
        st = s.stats()
        assert st["np"] == c.np and st["nstep"] == 5 and st["timestep"] > 0
        g = s.download()
        if nfluid:
            o = OracleSim(c)
            o.run(5)
            od = o.download()
            np.testing.assert_array_equal(g["idp"], od["idp"])
            fl = g["idp"] >= c.nbound
            # FP32 velocity rounding integrated into FP64 positions: ~1e-9 relative
            assert rel_err(g["pos"][fl], od["pos"][fl]) < 1e-7
            assert rel_err(g["velrho"][fl], od["velrho"][fl]) < 1e-5
        else:
            order = np.argsort(g["idp"])
            np.testing.assert_array_equal(g["pos"][order], c.pos)
    finally:
        s.close()


def test_half_cells_match_oracle_and_kernels_agree():
    """CELLMODE_Half (cells of size h, 5x5 rows of 5 cells): GPU forces on a
    stepped state equal the FP32 oracle within 1e-5 (norm-wise), the neighbour
    order is identical, and the tile and rows fluid kernels give the same bits."""
    from dualsphysics_amd import _native as N
    from dualsphysics_amd.sim import SphGpuSingle
    from oracle.oracle import OracleSim
    c = casemod.dam_break(dp=0.02, params=casemod.Parameters(cellmode=2))
    s = SphGpuSingle(c)
    try:
        s.run(3, sync_dt=True)
        ctx = s.lib.dsph_sim_ctx(s._h)
        out = {}
        for k in (N.FORCES_TILE, N.FORCES_ROWS):
            assert s.lib.dsph_ctx_set_option(ctx, N.OPT_FORCES_KERNEL, k) == 0
            s.interact()
            out[k] = s.forces()
        s.lib.dsph_ctx_set_option(ctx, N.OPT_FORCES_KERNEL, N.FORCES_TILE)
        np.testing.assert_array_equal(out[N.FORCES_TILE][0], out[N.FORCES_ROWS][0])
        np.testing.assert_array_equal(out[N.FORCES_TILE][1], out[N.FORCES_ROWS][1])
        g = s.download()
        gace, gar = out[N.FORCES_TILE]
        o = OracleSim(_state_case(c, g))
        np.testing.assert_array_equal(o.download()["idp"], g["idp"])
        o.interact()
        oace, oar = o.forces()
        npb = s.stats()["npb"]
        assert rel_err(gace[npb:], oace[npb:]) < 1e-5
        assert rel_err(gar, oar) < 1e-5
    finally:
        s.close()
