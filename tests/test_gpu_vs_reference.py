"""The HIP path against the reference CPU build on identical initial
particles: the whole step (divide, interaction, integration, dt) for the
first 10 to 60 steps, from the reference-run fixtures of tests/golden/
(bounds in tests/helpers.REF_TOL, the FP32 rounding spread between the
reference and the oracle). The 2-D cases (<data2d>, particles in the plane
y = 0) cover DBC, mDBC with laminar+SPS, mDBC2 no-slip under Verlet, mDBC2
free slip with no penetration, shifting (with the free-surface threshold,
and full), a channel periodic in x, a floating box (at rest; thrown and
spinning) and the flume with a piston or a flap. mDBC2 with moving walls:
the piston flume with no-slip, with no-slip and no penetration, and the flap
at constant rotation with free slip."""
import numpy as np
import pytest

from tests.helpers import compare_with_reference, load_reference_fixture

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["ref_dambreak_dp0.03", "ref_dambreak_dp0.03_verlet_laminar",
                                  "ref_flume_dp0.01_piston",
                                  "ref_flume_dp0.01_piston_mdbc_sps",
                                  "ref_dambreak_dp0.03_mdbc2", "ref_flume_dp0.01_flap",
                                  "ref_flume_dp0.01_flap_rot_mdbc", "ref_dambreak_dp0.03_ddt1",
                                  "ref_dambreak_dp0.03_noddt_sps",
                                  "ref_flume_dp0.01_piston_mdbc_verlet",
                                  "ref_dambreak_dp0.03_halfcells",
                                  "ref_dambreak_dp0.03_shift_nobound_tfs",
                                  "ref_dambreak_dp0.03_shift_full",
                                  "ref_flume_dp0.01_piston_mdbc_ddt1",
                                  "ref_dambreak_dp0.03_mdbc2_verlet_ddt1",
                                  "ref_flume_dp0.01_piston_mdbc_shift_nofixed",
                                  "ref_dambreak_dp0.03_viscbound2",
                                  "ref_channel_dp0.02_periodic",
                                  "ref_channel_dp0.02_periodic_mdbc_verlet",
                                  "ref_dambreak_dp0.03_mdbc2_moving",
                                  "ref_dambreak_dp0.03_mdbc2_freeslip",
                                  "ref_dambreak_dp0.03_mdbc2_nopen",
                                  "ref_dambreak_dp0.03_mdbc2_freeslip_nopen_verlet",
                                  "ref_dambreak2d_dp0.005", "ref_dambreak2d_dp0.005_mdbc_sps",
                                  "ref_dambreak2d_dp0.005_mdbc2_verlet",
                                  "ref_dambreak2d_dp0.005_mdbc2_freeslip_nopen",
                                  "ref_dambreak2d_dp0.005_shift_nobound_tfs",
                                  "ref_dambreak2d_dp0.005_shift_full",
                                  "ref_channel2d_dp0.005_periodic",
                                  "ref_channel2d_dp0.005_periodic_mdbc_verlet",
                                  "ref_dambreak2d_dp0.005_floating",
                                  "ref_dambreak2d_dp0.005_floating_spin",
                                  "ref_flume2d_dp0.005_piston_mdbc_sps",
                                  "ref_flume2d_dp0.005_flap",
                                  "ref_flume2d_dp0.005_flap_rot_mdbc",
                                  "ref_flume_dp0.01_piston_mdbc2",
                                  "ref_flume_dp0.01_piston_mdbc2_nopen",
                                  "ref_flume_dp0.01_flap_rot_mdbc2_freeslip",
                                  "ref_dambreak_dp0.03_mdbc2_shift_full",
                                  "ref_flume_dp0.01_piston_mdbc2_freeslip_shift_nofixed",
                                  "ref_dambreak_dp0.03_floating_verlet",
                                  "ref_dambreak2d_dp0.005_floating_spin_verlet",
                                  "ref_channel_dp0.02_periodic_mdbc2"])
def test_gpu_steps_follow_reference_cpu_build(name):
    from dualsphysics_amd.sim import SphGpuSingle
    fx, case = load_reference_fixture(name)
    sim = SphGpuSingle(case)
    done = 0
    for k, step in enumerate(fx["steps"]):
        sim.run(int(step) - done, sync_dt=True)
        done = int(step)
        d = sim.download()
        idx = np.argsort(d["idp"], kind="stable")
        e = compare_with_reference(fx, k, sim.stats()["timestep"], case.pos, d["pos"][idx],
                                   d["velrho"][idx, :3], d["velrho"][idx, 3],
                                   order=d["idp"])
        print(name, int(step), e)
    sim.close()
