"""Gauges on the GPU: point velocities, masses and maxz against brute-force
FP64 sums over all particles (numpy), and a gauge run of the dam break
against the reference CPU build's own gauge CSV output on the identical
particles (fixture ref_gauges_dambreak_dp0.03.npz, written by
tests/golden/make_reference_fixtures.py --gauges)."""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from dualsphysics_amd import case as casemod
from dualsphysics_amd import gauges as G
from dualsphysics_amd.sim import SphGpuSingle
from tests.helpers import perturbed_case

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def _brute(c, d, pts):
    """FP64 sums of W*V*v and W*V*m at pts over every fluid particle within 2h
    (FP32 kernel value and distance, as the gauges)."""
    fl = d["code"] & 0x1800 == 0x1800
    pos, vr = d["pos"][fl], d["velrho"][fl]
    h = np.float32(c.h)
    awen = np.float32(0.41778 / (float(h) ** 3))
    ks2 = np.float32(2 * h) ** 2
    vel = np.zeros((len(pts), 3))
    mass = np.zeros(len(pts))
    for i, p in enumerate(pts):
        dr = (p[None, :] - pos).astype(np.float32)
        rr2 = (dr * dr).sum(1, dtype=np.float32)
        m = (rr2 <= ks2) & (rr2 >= np.float32(1e-18))
        qq = np.sqrt(rr2[m]) / h
        w = awen * (qq + qq + 1) * (1 - 0.5 * qq) ** 4
        w = (w * (np.float32(c.massfluid) / vr[m, 3])).astype(np.float64)
        vel[i] = (w[:, None] * vr[m, :3]).sum(0)
        mass[i] = (w * np.float32(c.massfluid)).sum()
    return vel, mass


def test_gauge_points_and_maxz_match_brute_force():
    c = perturbed_case(casemod.dam_break(dp=0.03))
    s = SphGpuSingle(c)
    s.run(3)
    d = s.download()
    rng = np.random.default_rng(0)
    pts = np.column_stack([rng.uniform(-0.05, 1.0, 200), rng.uniform(0, 0.67, 200),
                           rng.uniform(0, 0.35, 200)])
    vel = np.empty((200, 3), np.float32)
    mass = np.empty(200, np.float32)
    s._check(s.lib.dsph_sim_gauge_points(s._h, 200, pts.ctypes.data_as(C.c_void_p),
                                         vel.ctypes.data_as(C.c_void_p),
                                         mass.ctypes.data_as(C.c_void_p)))
    bv, bm = _brute(c, d, pts)
    scale_v = np.abs(bv).max()
    assert np.abs(vel - bv).max() < 1e-5 * scale_v
    assert np.abs(mass - bm).max() < 1e-5 * bm.max()
    assert (bm > 0).sum() > 50 and (bm == 0).sum() > 5
    # maxz: highest fluid particle within distlimit horizontally.
    mp = np.column_stack([rng.uniform(0, 0.6, 50), rng.uniform(0.05, 0.6, 50), np.zeros(50)])
    hd = np.tile(np.array([0.4, 0.05], np.float32), (50, 1))
    zmax = np.empty(50, np.float32)
    s._check(s.lib.dsph_sim_gauge_maxz(s._h, 50, mp.ctypes.data_as(C.c_void_p),
                                       hd.ctypes.data_as(C.c_void_p),
                                       zmax.ctypes.data_as(C.c_void_p)))
    fl = d["code"] & 0x1800 == 0x1800
    pos = d["pos"][fl]
    for i in range(50):
        dr = (mp[i, None, :2] - pos[:, :2]).astype(np.float32)
        m = (dr * dr).sum(1) <= np.float32(0.05) ** 2
        m &= pos[:, 2] <= 0.4 + 0.2   # inside the searched cell layers
        exp = np.float32(pos[m, 2].max()) if m.any() else np.float32(c.maprealposmin[2])
        assert zmax[i] == max(exp, np.float32(c.maprealposmin[2])), i
    s.close()


def test_gauges_follow_the_reference_cpu_build(tmp_path):
    """Velocity, SWL and maxz series of the reference's CSV files (6
    significant digits) on the same particles: identical sampling times,
    values within the FP32 spread of the two runs."""
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "mkfx", Path(__file__).resolve().parent / "golden" / "make_reference_fixtures.py")
    fx = np.load(GOLDEN / "ref_gauges_dambreak_dp0.03.npz")
    c = casemod.dam_break(dp=float(fx["dp"]))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    gs = mk.gauge_system(c)
    s = SphGpuSingle(c)
    G.run(s, gs, float(fx["tmax"]))
    s.close()
    gs.save_csv(tmp_path)
    kinds = {G.Velocity: "Vel", G.Swl: "SWL", G.MaxZ: "MaxZ", G.Force: "Force"}
    for g in gs.gauges:
        ref = fx[g.name]
        got = G.read_csv(tmp_path / f"Gauges{kinds[type(g)]}_{g.name}.csv")
        assert got.shape == ref.shape, (g.name, got.shape, ref.shape)
        np.testing.assert_allclose(got[:, 0], ref[:, 0], rtol=0, atol=2e-6, err_msg=g.name)
        if isinstance(g, G.Velocity):
            err = np.abs(got[:, 1:4] - ref[:, 1:4]).max() / np.abs(ref[:, 1:4]).max()
            assert err < 1e-3, (g.name, err)
        elif isinstance(g, G.Swl):
            assert np.abs(got[:, 1:4] - ref[:, 1:4]).max() < 2e-5, g.name
        elif isinstance(g, G.Force):
            err = np.abs(got[:, 1:5] - ref[:, 1:5]).max() / np.abs(ref[:, 1]).max()
            assert err < 1e-4, (g.name, err)
        else:
            assert np.abs(got[:, 1] - ref[:, 1]).max() < 2e-5, g.name
        print(g.name, "ok")
