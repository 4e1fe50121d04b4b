"""Periodic boundaries on the GPU (RunPeriodic, JSphGpuSingle.cpp:276-366;
KerPeriodicMakeList / KerPeriodicDuplicate*, JSphGpu_ker.cu:2106-2400;
KerUpdatePos with periodic axes, JSphGpu_ker.cu:334-392) against the oracle,
on a channel periodic in x whose fluid moves at 2 m/s so that particles leave
through one face and re-enter through the other:
- particle order and copy count: after every step the GPU's particles, with
  the periodic copies left out, are in the order the oracle gives the same
  state (the oracle makes its own copies and sorts them), and both hold the
  same number of copies;
- trajectories over 60 steps (host dt and device dt), compared by id, with
  the particles that crossed a face among them;
- a channel periodic in x and y (corner copies from copies) with mDBC;
- the list kernel alone through the C ABI: entries in particle order over
  partial and many blocks, skipped codes, truncation at nmax.
The comparison with the reference CPU build on the same channel is in
test_gpu_vs_reference.py.
"""
import dataclasses

import numpy as np
import pytest

from dualsphysics_amd import case as casemod

pytestmark = pytest.mark.gpu


def _state_case(c, state):
    return dataclasses.replace(c, idp=state["idp"].copy(), pos=state["pos"].copy(),
                               velrho=state["velrho"].copy())


def test_order_and_copies_follow_oracle():
    from dualsphysics_amd.sim import SphGpuSingle
    from oracle.oracle import OracleSim
    c = casemod.periodic_channel(dp=0.02)
    s = SphGpuSingle(c)
    for step in range(12):
        s.run(1, sync_dt=True)
        g = s.download()
        o = OracleSim(_state_case(c, g))
        od = o.download()
        np.testing.assert_array_equal(od["idp"], g["idp"])
        assert o.stats()["np"] == s.stats()["np"] > c.np   # the copies are there
        assert o.stats()["npb"] == s.stats()["npb"]
    s.close()


@pytest.mark.parametrize("sync_dt", [True, False], ids=["host_dt", "device_dt"])
def test_trajectory_matches_oracle(sync_dt):
    from dualsphysics_amd.sim import SphGpuSingle
    from oracle.oracle import OracleSim
    c = casemod.periodic_channel(dp=0.02)
    s = SphGpuSingle(c)
    o = OracleSim(c)
    s.run(60, sync_dt=sync_dt)
    o.run(60)
    gs, os_ = s.stats(), o.stats()
    assert gs["np"] == os_["np"] and gs["npf_out_total"] == os_["npf_out_total"] == 0
    assert abs(gs["timestep"] - os_["timestep"]) <= 1e-6 * os_["timestep"]
    g, d = s.download(), o.download()
    assert len(g["idp"]) == len(d["idp"]) == c.np
    gi, oi = np.argsort(g["idp"]), np.argsort(d["idp"])
    crossed = (d["pos"][oi, 0] - c.pos[:, 0]) < -0.1
    assert crossed.sum() > 100
    dpos = np.abs(g["pos"][gi] - d["pos"][oi]).max()
    dvel = np.abs(g["velrho"][gi, :3] - d["velrho"][oi, :3]).max()
    drho = np.abs(g["velrho"][gi, 3] - d["velrho"][oi, 3]).max()
    print(f"60 steps ({int(crossed.sum())} crossed): max|dpos| {dpos:.2e} max|dv| {dvel:.2e} "
          f"max|drho| {drho:.2e}")
    assert dpos < 1e-4 * c.dp
    assert dvel < 1e-4 * 2.0
    assert drho < 1e-2


@pytest.mark.parametrize("tvisco,sync_dt,cellmode", [(3, True, 1), (2, False, 1), (3, True, 2)],
                         ids=["laminar_host_dt", "laminarsps_device_dt", "laminar_half_cells"])
def test_xy_periodic_mdbc_matches_oracle(tvisco, sync_dt, cellmode):
    """Channel periodic in x and y with mDBC: laminar viscosity (host dt),
    laminar+SPS with dt on the device, and half cells (cells of size h)."""
    from dualsphysics_amd.sim import SphGpuSingle
    from oracle.oracle import OracleSim
    p = casemod.Parameters(periodic=3, tboundary=2, tvisco=tvisco,
                           visco=1e-4 if tvisco == 3 else 1e-6, cellmode=cellmode)
    c = casemod.periodic_channel(dp=0.02, params=p)
    s = SphGpuSingle(c)
    o = OracleSim(c)
    for _ in range(3):
        s.run(10, sync_dt=sync_dt)
        o.run(10)
        g, d = s.download(), o.download()
        np.testing.assert_array_equal(g["idp"], d["idp"])
        assert s.stats()["np"] == o.stats()["np"]
    dpos = np.abs(g["pos"] - d["pos"]).max()
    drho = np.abs(g["velrho"][:, 3] - d["velrho"][:, 3]).max()
    print(f"xy periodic + mDBC (visco {tvisco}, cellmode {cellmode}), 30 steps: "
          f"max|dpos| {dpos:.2e} max|drho| {drho:.2e}")
    assert dpos < 1e-4 * c.dp and drho < 1e-2
    s.close()


@pytest.mark.parametrize("prm", [dict(shifting=3)], ids=["shifting"])
def test_unbuilt_periodic_combinations_are_refused(prm):
    from dualsphysics_amd.sim import DsphError, SphGpuSingle
    c = casemod.periodic_channel(dp=0.04, params=casemod.Parameters(periodic=1, **prm))
    with pytest.raises(DsphError, match="periodic boundaries with"):
        SphGpuSingle(c)


def _list_expected(pos, code, pini, n, mn, mx, inc):
    """Entries of the periodic list in particle order, plus before minus."""
    out = []
    for p in range(pini, pini + n):
        if (int(code[p]) & 0xe000) > 0x2000:
            continue
        for sg, flag in ((1.0, 0), (-1.0, 0x80000000)):
            q = pos[p] + sg * inc
            if np.all(mn <= q) and np.all(q < mx):
                out.append(p | flag)
    return np.array(out, np.uint32)


@pytest.mark.parametrize("n,pini,nmax", [(0, 0, 16), (1, 0, 16), (255, 3, 4096), (256, 0, 4096),
                                         (257, 1, 4096), (100003, 7, 1 << 20), (5000, 0, 37)])
def test_make_list_matches_rule(lib, ctx, n, pini, nmax):
    """dsph_periodic_make_list over partial and many blocks, with codes that
    are copied (normal, periodic) and skipped (ignore, out), both faces, and
    a list longer than nmax (count still complete, entries truncated)."""
    import ctypes as C
    from tests.helpers import Dev
    rng = np.random.default_rng(n + pini)
    tot = pini + n
    mn = np.array([-0.1, 0.0, 0.0]); mx = np.array([1.1, 0.5, 0.5])
    inc = np.array([-1.0, 0.0, 0.0])
    pos = np.stack([rng.uniform(0.0, 1.0, tot), rng.uniform(0.0, 0.5, tot),
                    rng.uniform(0.0, 0.5, tot)], 1)
    code = rng.choice(np.array([0x1800, 0x3800, 0x5800, 0x9800, 0x0000], np.uint16), tot)
    posxy = np.ascontiguousarray(pos[:, :2]).ravel() if tot else np.zeros(2)
    posz = np.ascontiguousarray(pos[:, 2]) if tot else np.zeros(1)
    d_xy, d_z = Dev(lib, ctx, posxy), Dev(lib, ctx, posz)
    d_code = Dev(lib, ctx, code if tot else np.zeros(1, np.uint16))
    d_work = Dev(lib, ctx, np.zeros((n + 255) // 256 + 1, np.uint32))
    d_list = Dev(lib, ctx, np.full(nmax, 0xdeadbeef, np.uint32))
    cnt = C.c_uint32(12345)
    rc = lib.dsph_periodic_make_list(ctx, n, pini, nmax, mn.ctypes.data, mx.ctypes.data,
                                     inc.ctypes.data, d_xy.ptr, d_z.ptr, d_code.ptr, d_work.ptr,
                                     d_list.ptr, C.byref(cnt))
    assert rc == 0, lib.dsph_last_error(ctx)
    exp = _list_expected(pos, code, pini, n, mn, mx, inc)
    assert cnt.value == exp.size
    got = d_list.get()
    k = min(exp.size, nmax)
    np.testing.assert_array_equal(got[:k], exp[:k])
    assert np.all(got[k:] == 0xdeadbeef)


def _free_block_channel(dp=0.02, length=1.0, u0=20.0):
    """A fluid block 0.4 x 0.6 x 0.6 m at x in [0.45, 0.85] of a channel
    periodic in x, without walls or gravity, moving at u0 along x: at load no
    particle is near a periodic face (the arrays get the minimum headroom for
    copies, 4096); when the block crosses the face, about 7000 copies are
    made while the previous ones are still held."""
    import math
    from dualsphysics_amd import _native as N
    pts, _ = casemod._lattice(round(0.45 / dp), round(0.85 / dp), 0, round(0.6 / dp), 0,
                              round(0.6 / dp), dp)
    u = casemod.uniform01(42, 3 * pts.shape[0]).reshape(-1, 3)
    pos = pts + (2.0 * u - 1.0) * (1e-4 * dp)
    h = 1.2 * math.sqrt(3.0) * dp
    cs0 = 20.0 * math.sqrt(9.81 * 0.2)
    rho0, gamma = 1000.0, 7.0
    vr = np.zeros((pos.shape[0], 4), np.float32)
    vr[:, 0], vr[:, 3] = u0, rho0
    prm = casemod.Parameters(periodic=1, tvisco=N.VISCO_ARTIFICIAL, tdensity=N.DDT_NONE)
    mn, mx = casemod.map_limits(pos, h, periodic=1, dp=dp)
    mn = (0.0 - dp / 2, mn[1], mn[2])
    mx = (length - dp / 2, mx[1], mx[2] + 0.2)
    return casemod.Case(name="free_block", dp=dp, h=h, rhop0=rho0, gamma=gamma, cs0=cs0,
                        b=cs0 * cs0 * rho0 / gamma, massfluid=rho0 * dp ** 3,
                        massbound=rho0 * dp ** 3, gravity=(0.0, 0.0, 0.0), hswl=0.2, params=prm,
                        idp=np.arange(pos.shape[0], dtype=np.uint32), pos=pos, velrho=vr,
                        nbound=0, nfluid=pos.shape[0], maprealposmin=mn, maprealposmax=mx)


def test_copies_beyond_the_initial_capacity_grow_the_arrays():
    """The arrays are sized at load for the copies of that moment; a block
    that later crosses a periodic face needs more: the arrays grow (contents
    kept) and the run follows the oracle, order and copies included."""
    from dualsphysics_amd.sim import SphGpuSingle
    from oracle.oracle import OracleSim
    c = _free_block_channel()
    s = SphGpuSingle(c)
    o = OracleSim(c)
    assert s.stats()["np"] == c.np        # nothing near a face at load
    most = 0
    for _ in range(10):
        s.run(20, sync_dt=True)
        o.run(20)
        most = max(most, s.stats()["np"] - c.np)
        assert s.stats()["np"] == o.stats()["np"]
        if most > 4096:
            break
    print(f"{c.np} particles, up to {most} copies after {s.stats()['nstep']} steps")
    # Old and new copies coexist during the copy step: twice the copies
    # exceed the 4096 of headroom, so the arrays must have grown.
    assert 2 * most > 2 * 4096, "the block never needed more copies than the initial headroom"
    g, d = s.download(), o.download()
    np.testing.assert_array_equal(g["idp"], d["idp"])
    assert np.abs(g["pos"] - d["pos"]).max() < 1e-4 * c.dp
    s.close()
