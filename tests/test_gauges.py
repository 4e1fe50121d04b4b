"""Gauges without a GPU: the SWL surface search and line sampling against
closed forms, and the gauge definitions through the case XML (the keys the
reference reads; the reference's own CSV output for these gauges is the
fixture ref_gauges_dambreak_dp0.03.npz used by test_gpu_gauges.py)."""
import numpy as np
import pytest

from dualsphysics_amd import bi4
from dualsphysics_amd import case as cm
from dualsphysics_amd import gauges as G


def test_swl_surface_interpolates_the_mass_crossing():
    p0, d = np.array([1.0, 2.0, 0.0]), np.array([0.0, 0.0, 0.1])
    # Wet up to k=2 (mass 2), 0.5 at k=3: the crossing of the limit 1 is at k = 2 + 1/1.5.
    s = G.swl_surface([2, 2, 2, 0.5, 0], p0, d, 1.0)
    np.testing.assert_allclose(s, p0 + d * (2 + 1 / 1.5), rtol=0, atol=1e-7)
    # Dry line: point0; wet line: the last point.
    np.testing.assert_array_equal(G.swl_surface([0, 0, 0], p0, d, 1.0), p0)
    np.testing.assert_allclose(G.swl_surface([3, 3, 3], p0, d, 1.0), p0 + 2 * d)
    # A dry start then fluid: the surface is the first wet->dry crossing.
    s = G.swl_surface([0.2, 3, 3, 0, 0], p0, d, 1.0)
    np.testing.assert_allclose(s, p0 + d * 3 * 1.0 + d * ((1 - 3) / (0 - 3) - 1), atol=1e-7)


def test_swl_line_sampling():
    g = G.Swl("s", (0, 0, 0), (0, 0, 0.4), 0.015)
    n, d = g.line()
    assert n == 27 and abs(np.linalg.norm(d) * n - 0.4) < 1e-15
    g2 = G.Swl("s", (0, 0, 0), (0, 0, 0.3), 0.015)
    assert g2.line()[0] == 20


def test_gauges_through_the_case_xml(tmp_path):
    c = cm.dam_break(dp=0.05)
    gs = G.GaugeSystem(c, [G.Velocity("v", (0.3, 0.3, 0.1)),
                           G.Swl("s", (0.2, 0.3, 0.0), (0.2, 0.3, 0.4), 0.02, masslimit=0.06),
                           G.MaxZ("m", (0.3, 0.3, 0.0), 0.4, 0.05, computedt=0.002)],
                       computedt=0.001)
    bi4.write_case(c, tmp_path, "case", gauges=gs)
    r = G.read_xml(c, tmp_path / "case.xml")
    assert r.computedt == gs.computedt
    assert r.gauges == gs.gauges
    with pytest.raises(ValueError, match="unique"):
        G.GaugeSystem(c, [G.Velocity("v", (0, 0, 0)), G.Velocity("v", (1, 0, 0))])


def test_csv_columns(tmp_path):
    c = cm.dam_break(dp=0.05)
    gs = G.GaugeSystem(c, [G.Velocity("v", (0.3, 0.3, 0.1)), G.MaxZ("m", (0.3, 0.3, 0.0), 0.4, 0.05)])
    gs.results["v"] = [(0.0, np.zeros(3, np.float32)), (0.5, np.array([1, 2, 3], np.float32))]
    gs.results["m"] = [(0.5, 0.25)]
    gs.save_csv(tmp_path)
    v = G.read_csv(tmp_path / "GaugesVel_v.csv")
    np.testing.assert_allclose(v[1], [0.5, 1, 2, 3, 0.3, 0.3, 0.1])
    assert (tmp_path / "GaugesVel_v.csv").read_text().startswith(
        "time [s];velx [m/s];vely [m/s];velz [m/s];posx [m];posy [m];posz [m]")
    np.testing.assert_allclose(G.read_csv(tmp_path / "GaugesMaxZ_m.csv")[0], [0.5, 0.25, 0.3, 0.3, 0])
