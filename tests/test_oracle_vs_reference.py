"""The oracle against the reference CPU build on identical initial particles
(fixtures written by tests/golden/make_reference_fixtures.py from
oracle/_ref, the reference's own sources compiled with IEEE arithmetic).

The oracle runs with the reference CPU build's pair semantics (prec=2: FP32,
float(FP64 difference) distances, two passes), with the GPU semantics
(prec=0) the HIP kernels follow, and in FP64 (prec=1); all three stay within
FP32 rounding of the reference over the first 10 steps (bounds in
tests/helpers.REF_TOL). Cases: dam break with Symplectic + artificial
viscosity + DDT2Full (the benchmark configuration) and Verlet + laminar
viscosity + DDT2; wave flume with a piston wavemaker (moving boundary,
rectilinear sinusoid) with DBC, and with mDBC + laminar+SPS viscosity; dam
break with mDBC2 (SlipMode no-slip) + laminar viscosity; flume with a flap
wavemaker (matrix motion: rotation sinusoid with DBC, constant rotation with
mDBC and the normals turned with the flap); dam break with the
Molteni-Colagrossi DDT and with laminar+SPS without DDT; the piston flume
with Verlet and mDBC; half cells; standard shifting (no boundaries with the
free-surface threshold, and full); a channel periodic in x (Symplectic +
DBC, Verlet + mDBC + laminar viscosity) over 60 steps in which particles
leave through one face and re-enter through the other; the dam break with
a floating box of half the water density at the column's surface over 40
steps (rigid-body motion from the fluid forces); mDBC2 free slip, and
no penetration (Symplectic with no-slip, Verlet with free slip), with the
column thrown at 1 m/s against the floor and the back wall over 40 steps;
the 2-D dam break (<data2d>: 2-D Wendland constants, 3x3 mDBC systems over
(1, x, z), 2-D SPS length) with DBC, mDBC + laminar+SPS, mDBC2 no-slip
under Verlet and mDBC2 free slip with no penetration, over 40 steps; 2-D
shifting (no boundaries with the free-surface threshold, whose 2-D factor
(w - tfs)/(2 - tfs) the fixture pins: the 3-D factor puts the displacement
3 % off after one step; and full) over 40 steps; a 2-D channel periodic in
x (DBC; Verlet + mDBC + laminar) over 60 steps; the 2-D dam break with a
floating box at rest, and thrown at 1 m/s spinning at 2 rad/s about y, over
40 steps (reproduced bit for bit in CPU semantics); the 2-D flume with the
piston (mDBC + laminar+SPS), the flap (rotation sinusoid, DBC) and the flap
at a constant rotation (mDBC) over 40 steps; mDBC2 with moving walls (the
motion acceleration in the pressure clone, the motion velocity in the
no-penetration test): the piston with no-slip, a 0.1 m piston with no-slip
and no penetration (acting from step ~25, when the piston has closed on the
fluid), the flap at constant rotation with free slip, over 40 steps;
shifting with mDBC2 (full shifting with no-slip on the dam break; shifting
that spares fixed walls with the moving piston and free slip); floating
bodies under Verlet (3-D at rest, 2-D thrown spinning), 10 steps.
"""
import numpy as np
import pytest

from oracle.oracle import OracleSim
from tests.helpers import compare_with_reference, load_reference_fixture

FIXTURES = ["ref_dambreak_dp0.03", "ref_dambreak_dp0.03_verlet_laminar",
            "ref_flume_dp0.01_piston", "ref_flume_dp0.01_piston_mdbc_sps",
            "ref_dambreak_dp0.03_mdbc2", "ref_flume_dp0.01_flap",
            "ref_flume_dp0.01_flap_rot_mdbc", "ref_dambreak_dp0.03_ddt1",
            "ref_dambreak_dp0.03_noddt_sps", "ref_flume_dp0.01_piston_mdbc_verlet",
            "ref_dambreak_dp0.03_halfcells", "ref_dambreak_dp0.03_shift_nobound_tfs",
            "ref_dambreak_dp0.03_shift_full", "ref_flume_dp0.01_piston_mdbc_ddt1",
            "ref_dambreak_dp0.03_mdbc2_verlet_ddt1", "ref_flume_dp0.01_piston_mdbc_shift_nofixed",
            "ref_dambreak_dp0.03_viscbound2", "ref_channel_dp0.02_periodic",
            "ref_channel_dp0.02_periodic_mdbc_verlet", "ref_dambreak_dp0.03_floating",
            "ref_dambreak_dp0.03_mdbc2_moving", "ref_dambreak_dp0.03_mdbc2_freeslip",
            "ref_dambreak_dp0.03_mdbc2_nopen", "ref_dambreak_dp0.03_mdbc2_freeslip_nopen_verlet",
            "ref_dambreak2d_dp0.005", "ref_dambreak2d_dp0.005_mdbc_sps",
            "ref_dambreak2d_dp0.005_mdbc2_verlet", "ref_dambreak2d_dp0.005_mdbc2_freeslip_nopen",
            "ref_dambreak2d_dp0.005_shift_nobound_tfs", "ref_dambreak2d_dp0.005_shift_full",
            "ref_channel2d_dp0.005_periodic", "ref_channel2d_dp0.005_periodic_mdbc_verlet",
            "ref_dambreak2d_dp0.005_floating", "ref_dambreak2d_dp0.005_floating_spin",
            "ref_flume2d_dp0.005_piston_mdbc_sps", "ref_flume2d_dp0.005_flap",
            "ref_flume2d_dp0.005_flap_rot_mdbc", "ref_flume_dp0.01_piston_mdbc2",
            "ref_flume_dp0.01_piston_mdbc2_nopen", "ref_flume_dp0.01_flap_rot_mdbc2_freeslip",
            "ref_dambreak_dp0.03_mdbc2_shift_full",
            "ref_flume_dp0.01_piston_mdbc2_freeslip_shift_nofixed",
            "ref_dambreak_dp0.03_floating_verlet", "ref_dambreak2d_dp0.005_floating_spin_verlet",
            "ref_channel_dp0.02_periodic_mdbc2"]


@pytest.mark.parametrize("prec", [2, 0, 1], ids=["fp32_cpu_semantics", "fp32_gpu_semantics",
                                                 "fp64"])
@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_follows_reference_cpu_build(name, prec):
    fx, case = load_reference_fixture(name)
    o = OracleSim(case, prec=prec)
    done = 0
    for k, step in enumerate(fx["steps"]):
        o.run(int(step) - done)
        done = int(step)
        d = o.download()
        idx = np.argsort(d["idp"], kind="stable")
        e = compare_with_reference(fx, k, o.stats()["timestep"], case.pos, d["pos"][idx],
                                   d["velrho"][idx, :3], d["velrho"][idx, 3],
                                   order=d["idp"])
        print(name, prec, int(step), e)


def test_fixture_detects_a_planted_error():
    """A deliberately wrong oracle input (gravity 1% off) leaves the bounds:
    the fixture comparison is sensitive to O(1e-2) physics errors."""
    import dataclasses
    fx, case = load_reference_fixture(FIXTURES[0])
    bad = dataclasses.replace(case, gravity=(0.0, 0.0, -9.81 * 1.01))
    o = OracleSim(bad, prec=1)
    o.run(int(fx["steps"][-1]))
    d = o.download()
    idx = np.argsort(d["idp"], kind="stable")
    with pytest.raises(AssertionError):
        compare_with_reference(fx, len(fx["steps"]) - 1, o.stats()["timestep"], case.pos,
                               d["pos"][idx], d["velrho"][idx, :3], d["velrho"][idx, 3])


@pytest.mark.parametrize("name,wrong", [
    ("ref_dambreak_dp0.03_mdbc2_nopen", dict(nopenetration=0)),
    ("ref_dambreak_dp0.03_mdbc2_freeslip", dict(slipmode=2)),
    ("ref_dambreak_dp0.03_mdbc2_freeslip_nopen_verlet", dict(nopenetration=0)),
    ("ref_dambreak2d_dp0.005_mdbc2_freeslip_nopen", dict(nopenetration=0)),
    ("ref_dambreak2d_dp0.005_mdbc2_freeslip_nopen", dict(slipmode=2)),
    ("ref_flume_dp0.01_piston_mdbc2_nopen", dict(nopenetration=0)),
])
def test_mdbc2_options_are_pinned_by_their_fixtures(name, wrong):
    """The oracle run without the option a fixture was made with (no
    penetration off, no-slip instead of free slip) leaves the bounds by the
    last saved step: each of these fixtures detects its option."""
    import dataclasses
    fx, case = load_reference_fixture(name)
    bad = dataclasses.replace(case, params=dataclasses.replace(case.params, **wrong))
    o = OracleSim(bad, prec=2)
    try:
        o.run(int(fx["steps"][-1]))
        d, t = o.download(), o.stats()["timestep"]
    finally:
        o.close()
    idx = np.argsort(d["idp"], kind="stable")
    with pytest.raises(AssertionError):
        compare_with_reference(fx, len(fx["steps"]) - 1, t, case.pos, d["pos"][idx],
                               d["velrho"][idx, :3], d["velrho"][idx, 3])


@pytest.mark.parametrize("name", ["ref_dambreak2d_dp0.005", "ref_dambreak2d_dp0.005_mdbc_sps"])
def test_2d_mode_is_pinned_by_its_fixtures(name):
    """The 2-D fixtures run as 3-D (3-D Wendland constants and SPS length,
    singular 4x4 mDBC systems falling back to Shepard) leave the bounds at
    the first saved step (particles excluded, or the fields off)."""
    import dataclasses
    fx, case = load_reference_fixture(name)
    bad = dataclasses.replace(case, data2d=False)
    o = OracleSim(bad, prec=2)
    try:
        o.run(int(fx["steps"][0]))
        d, t = o.download(), o.stats()["timestep"]
    finally:
        o.close()
    if d["idp"].shape[0] != case.np:
        return
    idx = np.argsort(d["idp"], kind="stable")
    with pytest.raises(AssertionError):
        compare_with_reference(fx, 0, t, case.pos, d["pos"][idx], d["velrho"][idx, :3],
                               d["velrho"][idx, 3])
