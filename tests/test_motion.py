"""Matrix motions (rotations of a moving boundary, MoveMatBound) without a
GPU: the library's rotation matrix (Rodrigues form, host function of the C
ABI) against the oracle's restatement of JMatrix4::MatrixRot (JMatrix4.h:
465-497, six composed matrices), and closed forms. The sign convention and
the motion laws are pinned by the reference-run fixtures
ref_flume_dp0.01_flap (rotation sinusoid) and ref_flume_dp0.01_flap_rot_mdbc
(constant angular velocity, mDBC normals turned with the flap), run by
test_oracle_vs_reference.py and test_gpu_vs_reference.py."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle.oracle as orc
from dualsphysics_amd import _native as N


def _lib_matrix(kind, p1, p2, t, dt, **kw):
    m = N.DsphMotion()
    m.type = kind
    for k in range(3):
        m.axisp1[k], m.axisp2[k] = p1[k], p2[k]
    for k, v in kw.items():
        setattr(m, k, v)
    out = (C.c_double * 12)()
    assert N.load().dsph_motion_matrix(C.byref(m), t, dt, out) == 0
    return np.array(out).reshape(3, 4)


def _oracle_matrix(ang, p1, p2):
    out = np.zeros(16)
    a1, a2 = np.array(p1, float), np.array(p2, float)
    orc.load().orc_matrix_rot(ang, a1.ctypes.data, a2.ctypes.data, out.ctypes.data)
    return out.reshape(4, 4)


def test_rotation_matrix_equals_the_composed_reference_form():
    rng = np.random.default_rng(4)
    for _ in range(200):
        p1 = rng.uniform(-2, 2, 3)
        p2 = p1 + rng.standard_normal(3)
        if rng.random() < 0.2:   # axes along x (the reference's L1 == 0 branch)
            p2 = p1 + np.array([rng.choice([-1.0, 1.0]) * rng.uniform(0.1, 2), 0, 0])
        ang = rng.uniform(-30, 30)
        got = _lib_matrix(3, p1, p2, 0.0, 1.0, angvel=ang)
        ref = _oracle_matrix(ang, p1, p2)
        np.testing.assert_allclose(got, ref[:3], rtol=0, atol=1e-13)
        np.testing.assert_allclose(ref[3], [0, 0, 0, 1], atol=0)


def test_rotation_closed_forms():
    # About +z through the origin: (1,0,0) -> (cos a, -sin a, 0) (a positive
    # angle turns clockwise seen from +z).
    a = 30.0
    m = _oracle_matrix(a, (0, 0, 0), (0, 0, 1))
    x = m @ np.array([1.0, 0, 0, 1])
    np.testing.assert_allclose(x[:3], [math.cos(math.radians(a)), -math.sin(math.radians(a)), 0],
                               atol=1e-15)
    # Points of the axis are fixed, R is orthonormal with det +1, and the
    # rotation angle is recovered from the trace.
    rng = np.random.default_rng(1)
    p1, p2 = rng.standard_normal(3), rng.standard_normal(3)
    m = _lib_matrix(3, p1, p2, 0.0, 1.0, angvel=a)
    for s in (0.0, 0.3, 1.7):
        q = p1 + s * (p2 - p1)
        np.testing.assert_allclose(m[:, :3] @ q + m[:, 3], q, atol=1e-14)
    r = m[:, :3]
    np.testing.assert_allclose(r @ r.T, np.eye(3), atol=1e-15)
    assert abs(np.linalg.det(r) - 1) < 1e-14
    assert abs(math.degrees(math.acos((np.trace(r) - 1) / 2)) - a) < 1e-10


def test_motion_laws_over_a_step():
    """Rotation sinusoid: the step angle is ampl*(sin(ph+w(t+dt))-sin(ph+wt));
    constant rotation: angvel*dt (recovered from the matrix)."""
    p1, p2 = (0.0, 0.0, 0.0), (0.0, 1.0, 0.0)
    f, A, ph, t, dt = 1.25, 4.0, 0.3, 0.37, 1e-3
    m = _lib_matrix(2, p1, p2, t, dt, angfreq=f, angampl=A, angphase=ph)
    ang = A * (math.sin(ph + 2 * math.pi * f * (t + dt)) - math.sin(ph + 2 * math.pi * f * t))
    np.testing.assert_allclose(m, _oracle_matrix(ang, p1, p2)[:3], atol=1e-15)
    m3 = _lib_matrix(3, p1, p2, t, dt, angvel=10.0)
    np.testing.assert_allclose(m3, _oracle_matrix(10.0 * dt, p1, p2)[:3], atol=1e-15)


def test_invalid_motion_type_is_refused():
    m = N.DsphMotion()
    m.type = 1
    out = (C.c_double * 12)()
    assert N.load().dsph_motion_matrix(C.byref(m), 0.0, 1.0, out) == N.DSPH_E_ARG
