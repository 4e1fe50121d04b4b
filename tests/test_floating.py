"""Floating bodies in the oracle, pinned by mathematics where the reference
fixture (ref_dambreak_dp0.03_floating, test_oracle_vs_reference.py) leaves
room: a body far from any fluid is a free rigid body, so
- under gravity its velocity is exactly g*t (summed step by step in FP32 as
  the reference does: v += dt*a), its angular velocity stays zero and its
  particles translate together;
- without gravity, spinning at omega, it rotates rigidly: every particle
  keeps its distance to the centre, the angles grow as omega*t, and each
  particle's velocity is omega x (x - centre).
Also: the case files carry the body (write_case/read_case round trip)."""
import dataclasses

import numpy as np
import pytest

from dualsphysics_amd import bi4
from dualsphysics_amd import case as cm
from oracle.oracle import OracleSim


def _lonely_box(omega=(0.0, 0.0, 0.0), g=9.81):
    """A 4x4x4-particle box (density 500) high above the water column of a
    dam break (no fluid within 2h of it)."""
    c = cm.dam_break_floating(dp=0.03, box_center=(1.2, 0.335, 0.33), omegaini=omega)
    if g != 9.81:
        c = dataclasses.replace(c, gravity=(0.0, 0.0, -g))
    return c


def _body(o, c):
    d = o.download()
    idx = np.argsort(d["idp"], kind="stable")
    ft = slice(c.npb, c.nbound)
    return d["pos"][idx][ft], d["velrho"][idx, :3][ft]


def test_free_fall_is_g_t_and_rigid():
    c = _lonely_box()
    o = OracleSim(c, prec=2)
    # Explicit Euler of the body (FtComputeAceVel): centre += dt v (FP64),
    # then v = float(dt a + v), step by step with the run's own dt.
    zc, v, t0 = c.floatings[0].center[2], np.float32(0.0), 0.0
    for _ in range(30):
        o.run(1)
        t = o.stats()["timestep"]
        dt = t - t0
        t0 = t
        zc = zc + dt * float(v)
        v = np.float32(dt * float(np.float32(-9.81)) + float(v))
    f = o.floating(0)
    assert f["center"][2] == zc and f["fvel"][2] == v
    pos, vel = _body(o, c)
    # fvel.z accumulates float(dt*(-g) + v) step by step: within FP32 rounding of -g t.
    assert abs(f["fvel"][2] + 9.81 * t) < 1e-6 * 9.81 * t + 1e-7
    assert np.all(f["fomega"] == 0) and f["fvel"][0] == 0 and f["fvel"][1] == 0
    disp = pos - c.pos[c.npb:c.nbound]
    assert np.abs(disp - disp[0]).max() < 1e-12          # pure translation
    assert np.abs(vel - f["fvel"]).max() == 0            # every particle at the body velocity
    # ... which is the continuous fall z0 - g t^2/2 up to the O(dt) Euler term.
    dz = f["center"][2] - c.floatings[0].center[2]
    assert abs(dz + 0.5 * 9.81 * t * t) < 9.81 * t * 1e-3


def test_spin_without_gravity_is_a_rigid_rotation():
    w = (0.0, 0.0, 2.0)
    c = _lonely_box(omega=w, g=0.0)
    o = OracleSim(c, prec=2)
    o.run(1)
    dt1 = o.stats()["timestep"]
    o.run(29)
    f = o.floating(0)
    t = o.stats()["timestep"]
    pos, vel = _body(o, c)
    cen0 = np.array(c.floatings[0].center)
    r0 = np.linalg.norm(c.pos[c.npb:c.nbound] - cen0, axis=1)
    r = np.linalg.norm(pos - f["center"], axis=1)
    assert np.abs(r - r0).max() < 5e-5 * r0.max()        # distances kept (explicit update)
    # Internal pair forces of the body cancel to FP32 rounding.
    assert abs(f["angles"][2] - 2.0 * t) < 1e-5 and np.abs(f["fvel"]).max() < 1e-12
    assert np.abs(f["fomega"] - np.array(w, np.float32)).max() < 1e-6
    expect = np.cross(np.array(w), pos - f["center"])
    assert np.abs(vel - expect).max() < 1e-6
    # The particles turned by 2 (t - dt1) about z: they start at rest, so the
    # first step (each step moves them by dt times their velocity at its
    # start) leaves them in place while the body's angle counts from t = 0.
    a0 = np.arctan2(*(c.pos[c.npb:c.nbound] - cen0)[:, [1, 0]].T)
    a1 = np.arctan2(*(pos - f["center"])[:, [1, 0]].T)
    da = np.angle(np.exp(1j * (a1 - a0)))
    assert abs(np.mean(da) - 2.0 * (t - dt1)) < 1e-3 * 2.0 * t


def test_body_data_from_particles():
    """Centre and inertia of the generated box: point masses on a cube
    lattice of n^3 points, spacing dp: I = m n^3 * 2 (n^2-1) dp^2 / 12."""
    c = cm.dam_break_floating(dp=0.03, jitter=0.0)
    f = c.floatings[0]
    n = round(f.count ** (1 / 3))
    assert n ** 3 == f.count
    ii = f.masspart * f.count * 2 * (n * n - 1) * 0.03 ** 2 / 12
    inertia = np.array(f.inertia).reshape(3, 3)
    np.testing.assert_allclose(np.diag(inertia), ii, rtol=1e-12)
    assert np.abs(inertia - np.diag(np.diag(inertia))).max() < 1e-12 * ii


def test_case_files_carry_the_floating_body(tmp_path):
    c = cm.dam_break_floating(dp=0.03, velini=(0.1, 0.0, 0.0), omegaini=(0.0, 0.5, 0.0))
    bi4.write_case(c, tmp_path, "flt")
    r = bi4.read_case(tmp_path, "flt")
    assert r.nbound == c.nbound and r.nfloat == c.nfloat and r.npb == c.npb
    assert len(r.floatings) == 1
    a, b = r.floatings[0], c.floatings[0]
    assert (a.begin, a.count) == (b.begin, b.count)
    for k in ("massbody", "masspart"):
        assert getattr(a, k) == getattr(b, k)
    for k in ("center", "inertia", "velini", "omegaini"):
        np.testing.assert_array_equal(np.array(getattr(a, k)), np.array(getattr(b, k)))


@pytest.mark.parametrize("prm", [dict(tboundary=2), dict(shifting=3)], ids=["mdbc", "shifting"])
def test_unsupported_floating_options_are_refused(prm):
    c = cm.dam_break_floating(dp=0.05, params=cm.Parameters(**prm))
    with pytest.raises(RuntimeError, match="floating"):
        OracleSim(c)
