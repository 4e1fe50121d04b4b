"""Checkpoint / resume on the GPU: a run saved to a PART file
(SphGpuSingle.save_part, the reference's format) and continued with
SphGpuSingle.resume equals the uninterrupted run bit for bit (particle
order, positions, velocity, density, time), as the reference's own
-partbegin does (tests/test_case_io.py)."""
import numpy as np
import pytest

from dualsphysics_amd import case as casemod
from dualsphysics_amd.sim import SphGpuSingle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("sync_dt", [True, False], ids=["host_dt", "device_dt"])
@pytest.mark.parametrize("name", ["dam_break", "dam_break_mdbc2", "flume_flap"])
def test_resume_equals_uninterrupted_run(tmp_path, name, sync_dt):
    if name == "dam_break":
        c = casemod.dam_break(dp=0.03)
    elif name == "dam_break_mdbc2":
        c = casemod.dam_break(dp=0.03, params=casemod.Parameters(tboundary=2, slipmode=2,
                                                                 tvisco=3, visco=1e-4))
    else:
        c = casemod.flume(dp=0.012, wavemaker="flap")
    print(f"case {name}: {c.np} particles", flush=True)
    a = SphGpuSingle(c)
    a.run(12, sync_dt=sync_dt)
    b = SphGpuSingle(c)
    b.run(5, sync_dt=sync_dt)
    part = b.save_part(tmp_path / "data", 5)
    b.close()
    r = SphGpuSingle.resume(c, part)
    assert r.stats()["nstep"] == 5
    r.run(7, sync_dt=sync_dt)
    ga, gr = a.download(), r.download()
    sa, sr = a.stats(), r.stats()
    assert sa["timestep"] == sr["timestep"] and sr["nstep"] == 12
    for k in ("idp", "pos", "velrho", "code"):
        np.testing.assert_array_equal(ga[k], gr[k], err_msg=k)
    a.close()
    r.close()
