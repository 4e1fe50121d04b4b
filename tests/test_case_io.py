"""Case and PART files (dualsphysics_amd.bi4) without a GPU: the case XML +
.bi4 round trip, PART round trip, refusal of options outside the built path,
and — where the reference CPU build exists (marker ref) — the reference
resuming from a PART written here exactly as from its own."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from dualsphysics_amd import bi4
from dualsphysics_amd import case as cm

ROOT = Path(__file__).resolve().parent.parent
REF_EXE = ROOT / "oracle" / "_ref" / "DualSPHysics5.4CPU_ieee"


def _cases():
    mdbc = cm.Parameters(tboundary=2, slipmode=1, tvisco=2, visco=1e-6)
    return [cm.dam_break(dp=0.03),
            cm.dam_break(dp=0.03, params=cm.Parameters(tboundary=2, slipmode=2, tvisco=3,
                                                       visco=1e-4, tstep=1, tdensity=2)),
            cm.flume(dp=0.01), cm.flume(dp=0.01, wavemaker="flap", params=mdbc),
            cm.flume(dp=0.01, wavemaker="flap_rot"),
            cm.periodic_channel(dp=0.02),
            cm.periodic_channel(dp=0.02, params=cm.Parameters(periodic=3, tboundary=2)),
            cm.periodic_channel(dp=0.02, params=cm.Parameters(
                periodic=1, perinc=((0.0, 0.0, 0.01), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0)))),
            cm.dam_break_2d(dp=0.02, posy=0.25),
            cm.flume_2d(dp=0.01, wavemaker="flap", params=mdbc)]


@pytest.mark.parametrize("k", range(10))
def test_case_round_trip(tmp_path, k):
    c = _cases()[k]
    bi4.write_case(c, tmp_path, "case")
    r = bi4.read_case(tmp_path, "case")
    for f in ("dp", "h", "rhop0", "gamma", "b", "massfluid", "massbound", "gravity", "nbound",
              "nfluid", "nmoving", "maprealposmin", "maprealposmax", "data2d", "data2dposy"):
        assert getattr(r, f) == getattr(c, f), f
    import dataclasses
    assert r.params.timeout == c.params.timemax   # TimeOut 0 is written as TimeMax
    assert dataclasses.replace(r.params, timeout=c.params.timeout) == c.params
    assert abs(r.cs0 - c.cs0) <= 1e-12 * c.cs0
    for f in ("idp", "pos", "velrho"):
        np.testing.assert_array_equal(getattr(r, f), getattr(c, f), err_msg=f)
    if c.params.tboundary == 2:
        np.testing.assert_array_equal(r.boundnor, c.boundnor)
    if c.motion:
        m = dict(c.motion)
        m["finish"] = min(m["finish"], 1e6)   # written as the motion's duration
        assert r.motion == m


def test_part_round_trip(tmp_path):
    c = cm.dam_break(dp=0.05)
    rng = np.random.default_rng(0)
    order = rng.permutation(c.np).astype(np.uint32)
    vr = c.velrho.copy()
    vr[:, :3] = rng.standard_normal((c.np, 3)).astype(np.float32)
    bi4.write_part(tmp_path / "Part_0003.bi4", c, c.idp[order], c.pos[order], vr[order], 0.125,
                   42, 3, symplecticdtpre=1.5e-4)
    p = bi4.read_part(tmp_path / "Part_0003.bi4", sort_by_id=False)
    np.testing.assert_array_equal(p["idp"], c.idp[order])
    np.testing.assert_array_equal(p["pos"], c.pos[order])
    np.testing.assert_array_equal(p["vel"], vr[order, :3])
    np.testing.assert_array_equal(p["rho"], vr[order, 3])
    assert (p["time"], p["step"], p["cpart"], p["symplecticdtpre"]) == (0.125, 42, 3, 1.5e-4)
    q = bi4.read_part(tmp_path / "Part_0003.bi4")
    np.testing.assert_array_equal(q["idp"], c.idp)
    np.testing.assert_array_equal(q["pos"], c.pos)


def test_options_outside_the_path_are_refused(tmp_path):
    c = cm.dam_break(dp=0.05)
    bi4.write_case(c, tmp_path, "case")
    xml = (tmp_path / "case.xml").read_text()
    (tmp_path / "dtfix.xml").write_text(xml.replace('key="DtFixed" value="0"',
                                                    'key="DtFixed" value="0.0001"'))
    shutil.copy(tmp_path / "case.bi4", tmp_path / "dtfix.bi4")
    with pytest.raises(ValueError, match="fixed dt"):
        bi4.read_case(tmp_path, "dtfix")
    (tmp_path / "cubic.xml").write_text(xml.replace('key="Kernel" value="2"',
                                                    'key="Kernel" value="1"'))
    shutil.copy(tmp_path / "case.bi4", tmp_path / "cubic.bi4")
    with pytest.raises(ValueError, match="Wendland"):
        bi4.read_case(tmp_path, "cubic")


def test_periodic_keys_of_the_case_xml(tmp_path):
    """Periodic parameters as the solver reads them (JSph.cpp:812-826,
    JCaseEParms.cpp:454-466): increments of an inclined axis, the two-axis
    keys, parameters switched off with active="false", unknown keys refused;
    PART and Part_Head headers carry PeriMode and the increments (own
    component -MapRealSize) that the reference's resume checks."""
    c = cm.periodic_channel(dp=0.02)
    bi4.write_case(c, tmp_path, "case")
    xml = (tmp_path / "case.xml").read_text()
    assert 'key="XPeriodicIncY"' in xml

    def variant(name, text):
        (tmp_path / f"{name}.xml").write_text(text)
        shutil.copy(tmp_path / "case.bi4", tmp_path / f"{name}.bi4")
        return bi4.read_case(tmp_path, name)

    r = variant("incl", xml.replace('key="XPeriodicIncZ" value="0.0"',
                                    'key="XPeriodicIncZ" value="0.25"'))
    assert r.params.periodic == 1 and r.params.perinc[0] == (0.0, 0.0, 0.25)
    nat = r.to_native()
    assert list(nat.perinc[0]) == [0.0, 0.0, 0.25]
    r = variant("off", xml.replace('key="XPeriodicIncY" value="0.0"',
                                   'key="XPeriodicIncY" value="0.0" active="false"')
                .replace('key="XPeriodicIncZ" value="0.0"',
                         'key="XPeriodicIncZ" value="0.0" active="false"'))
    assert r.params.periodic == 0
    r = variant("xy", xml.replace('<parameter key="XPeriodicIncY" value="0.0" />',
                                  '<parameter key="XYPeriodic" value="1" />'))
    assert r.params.periodic == 3 and r.params.perinc == ((0.0,) * 3,) * 3
    with pytest.raises(ValueError, match="not recognised"):
        variant("bad", xml.replace('key="XPeriodicIncY"', 'key="XPeriodicIncW"'))
    bi4.write_part(tmp_path / "Part_0001.bi4", c, c.idp, c.pos, c.velrho, 0.0, 0, 1)
    bi4.write_part_head(tmp_path / "Part_Head.ibi4", c)
    size = c.maprealposmax[0] - c.maprealposmin[0]
    for f in ("Part_0001.bi4", "Part_Head.ibi4"):
        root = bi4.read_file(tmp_path / f)
        assert root.get("PeriMode") == 1
        assert tuple(root.get("PeriXinc")) == (-size, 0.0, 0.0)
        assert tuple(root.get("PeriYinc")) == (0.0, 0.0, 0.0)


@pytest.mark.ref
@pytest.mark.parametrize("kind", ["dam_break", "periodic_channel", "flume_2d"])
def test_reference_resumes_from_a_part_written_here(tmp_path, kind):
    """The reference CPU build continued (-partbegin) from a PART + head
    written by write_part / write_part_head (from the data of its own
    Part_0005) gives its own continued run bit for bit, which equals its
    uninterrupted run. The periodic channel checks the periodic header
    (the reference refuses a PART whose PeriMode differs from the case); the
    2-D flume checks the moving block and the 2-D header (the reference
    refuses a PART whose Data2d or Data2dPosY differs from the case)."""
    if not REF_EXE.exists():
        pytest.skip("reference CPU build absent")
    c = {"dam_break": lambda: cm.dam_break(dp=0.04),
         "periodic_channel": lambda: cm.periodic_channel(dp=0.02),
         "flume_2d": lambda: cm.flume_2d(dp=0.01, posy=0.1)}[kind]()
    bi4.write_case(c, tmp_path / "in", "case", timemax=10.0, timeout=10.0)

    def run(out, extra):
        cmd = [str(REF_EXE), str(tmp_path / "in" / "case"), str(out), "-cpu", "-ompthreads:4",
               "-svsteps:1", "-saveposdouble:1", "-stable:1", "-svres:0"] + extra
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stdout[-2000:]
    run(tmp_path / "full", ["-nsteps:8"])
    src = tmp_path / "full" / "data" / "Part_0004.bi4"
    pt = bi4.read_part(src, sort_by_id=False)
    ours = tmp_path / "ours"
    ours.mkdir()
    vr = np.concatenate([pt["vel"], pt["rho"][:, None]], 1)
    bi4.write_part(ours / "Part_0004.bi4", c, pt["idp"], pt["pos"], vr, pt["time"], pt["step"], 4,
                   symplecticdtpre=pt["symplecticdtpre"])
    bi4.write_part_head(ours / "Part_Head.ibi4", c)
    run(tmp_path / "res", ["-nsteps:4", "-partbegin:4", str(ours)])
    a = bi4.read_part(tmp_path / "res" / "data" / "Part_0008.bi4", sort_by_id=False)
    f = bi4.read_part(tmp_path / "full" / "data" / "Part_0008.bi4", sort_by_id=False)
    for k in ("idp", "pos", "vel", "rho"):
        np.testing.assert_array_equal(a[k], f[k], err_msg=k)
    assert a["time"] == f["time"]
