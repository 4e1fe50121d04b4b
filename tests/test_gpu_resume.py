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


def test_run_case_from_files(tmp_path):
    """A case given as XML + .bi4 runs to TimeMax with a PART every TimeOut;
    the last PART equals a direct run of the same steps."""
    from dualsphysics_amd import bi4
    from dualsphysics_amd.sim import run_case
    c = casemod.dam_break(dp=0.04)
    bi4.write_case(c, tmp_path / "in", "case", timemax=0.006, timeout=0.002)
    parts = run_case(tmp_path / "in", "case", tmp_path / "out")
    assert len(parts) == 4
    times = [bi4.read_part(p)["time"] for p in parts]
    assert times[0] == 0.0
    for k, t in enumerate(times[1:], 1):
        assert t >= 0.002 * k - 1e-12 and t < 0.002 * k + 0.001
    last = bi4.read_part(parts[-1], sort_by_id=False)
    s = SphGpuSingle(c)
    s.run(int(last["step"]))
    g = s.download()
    np.testing.assert_array_equal(g["idp"], last["idp"])
    np.testing.assert_array_equal(g["pos"], last["pos"])
    np.testing.assert_array_equal(g["velrho"][:, 3], last["rho"])
    s.close()
