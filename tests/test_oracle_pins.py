"""Pins of the CPU oracle (no GPU): each check is chosen so that a plausible
mistake in the restatement (a dropped term, a wrong sign or index, a
transposed operand, a missed neighbour) fails it. The checks come from
closed forms, conservation laws, consistency of SPH on a lattice, published
test vectors and brute force; the reference CPU build's own outputs are used
in test_oracle_vs_reference.py.
"""
import ctypes as C
import math

import numpy as np
import pytest

from dualsphysics_amd import case as cm
import oracle.oracle as orc
from oracle.oracle import OracleSim


# ------------------------------------------------------------------ case generator
def test_splitmix64_published_vector():
    # SplitMix64 from state 0 (Steele, Lea & Flood 2014; Vigna's reference C).
    out = cm.splitmix64(0, 3)
    assert [int(v) for v in out] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_dam_break_cfg0_counts_and_constants():
    c = cm.dam_break(dp=0.01)
    assert c.nfluid == 40 * 67 * 30 == 80400
    # Two wall layers around a 160x67x40 interior, open top.
    assert c.nbound == (164 * 71 * 42) - (160 * 67 * 40) == 60248
    assert math.isclose(c.h, 1.2 * math.sqrt(3) * 0.01)
    assert math.isclose(c.cs0, 20 * math.sqrt(9.81 * 0.3))
    assert math.isclose(c.b, c.cs0 ** 2 * 1000 / 7)
    jit = c.pos - np.round(c.pos / c.dp - 0.5)[:, :] * c.dp - 0.5 * c.dp
    assert np.abs(jit).max() <= 1e-4 * c.dp * (1 + 1e-9)
    fl = slice(c.nbound, None)
    z = c.pos[fl, 2]
    rho = 1000 * (1 + 1000 * 9.81 * (0.3 - z) / c.b) ** (1 / 7)
    np.testing.assert_allclose(c.velrho[fl, 3], rho, rtol=1e-7)


# ------------------------------------------------------------------ kernel and EOS
def _kernel(prec, r, h):
    lib = orc.load()
    awen = 21.0 / (16.0 * math.pi * h ** 3)
    bwenh = -5.0 * awen / h ** 2
    wab, fac = C.c_double(), C.c_double()
    lib.orc_kernel_eval(prec, r * r, h, awen, bwenh, C.byref(wab), C.byref(fac))
    return wab.value, fac.value


def test_wendland_normalised_and_gradient_consistent():
    h = 0.02
    r = np.linspace(0.0, 2 * h, 4001)
    w = np.array([_kernel(1, x, h)[0] for x in r])
    integral = np.trapezoid(w * 4 * np.pi * r ** 2, r)
    assert abs(integral - 1.0) < 1e-6
    # fac = (dW/dr)/r by central differences (r^2 enters as FP32, hence the
    # step and tolerance).
    for x in (0.3 * h, 0.9 * h, 1.5 * h, 1.7 * h):
        e = 1e-3 * h
        dw = (_kernel(1, x + e, h)[0] - _kernel(1, x - e, h)[0]) / (2 * e)
        assert math.isclose(_kernel(1, x, h)[1], dw / x, rel_tol=1e-4)
    w0, f0 = _kernel(1, 0.0, h)
    w2, f2 = _kernel(1, 2 * h, h)
    assert abs(w2) < 1e-12 * w0 and abs(f2) < 1e-8 * abs(f0)


def test_reference_kernel_constants_are_the_closed_forms():
    # awen = 21/(16 pi) / h^3 and bwenh = -5 awen / h^2, which the reference
    # rounds to 0.41778 and -2.08891 (FunSphKernel.h:198-207).
    assert abs(21 / (16 * math.pi) - 0.41778) < 5e-6
    assert abs(-5 * 21 / (16 * math.pi) + 2.08891) < 5e-6


def test_tait_eos():
    lib = orc.load()
    c = orc.OrcCte()
    c.cteb, c.gamma, c.rhopzero, c.ovrhopzero = 1.6817e5, 7.0, 1000.0, np.float32(1 / 1000.0)
    assert lib.orc_eos(1, orc.SEM_CPU, C.byref(c), 1000.0) == 0.0
    for rho in (990.0, 1003.5, 1010.0):
        exact = 1.6817e5 * ((rho / 1000.0) ** 7 - 1)
        assert math.isclose(lib.orc_eos(1, orc.SEM_CPU, C.byref(c), rho), exact, rel_tol=1e-12)
        # FP32 semantics: the power of the FP32 ratio rounded once.
        got = lib.orc_eos(0, orc.SEM_GPU, C.byref(c), rho)
        assert math.isclose(got, exact, rel_tol=1e-4, abs_tol=0.05)


# ------------------------------------------------------------------ cells and sort
def test_cell_code_round_trip():
    lib = orc.load()
    rng = np.random.default_rng(5)
    for _ in range(50):
        nc = rng.integers(1, 700, 3)
        dcc = lib.orc_cell_code(*map(int, nc))
        sx, sy = (dcc >> 10) & 31, (dcc >> 5) & 31
        for _ in range(20):
            cx, cy, cz = (int(rng.integers(0, n)) for n in nc)
            cel = (cx << ((dcc >> 10) & 31)) | (cy << ((dcc >> 15) & 31)) | cz
            assert cel < 2 ** 31
            assert cel >> sx == cx
            assert ((cel << (dcc >> 25)) & 0xFFFFFFFF) >> sy == cy
            assert ((cel << (dcc & 31)) & 0xFFFFFFFF) >> (dcc & 31) == cz


@pytest.mark.parametrize("n,nkeys", [(1, 1), (1000, 7), (50000, 3000), (20000, 1)])
def test_counting_sort_is_the_stable_sort(n, nkeys):
    lib = orc.load()
    rng = np.random.default_rng(n)
    keys = rng.integers(0, nkeys, n).astype(np.uint32)
    cellpart = keys.copy()
    sortpart = np.empty(n, np.uint32)
    be = np.zeros(2 * nkeys, np.int32)
    lib.orc_sort_cells(n, 0, nkeys, 0, cellpart.ctypes.data, sortpart.ctypes.data, be.ctypes.data)
    ref = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(sortpart, ref)
    sk = keys[ref]
    for k in range(nkeys):
        idx = np.nonzero(sk == k)[0]
        if idx.size:
            assert tuple(be[2 * k:2 * k + 2]) == (idx[0], idx[-1] + 1)
        else:
            assert tuple(be[2 * k:2 * k + 2]) == (0, 0)


# ------------------------------------------------------------------ interaction physics
def _cube(n=14, dp=0.02, vel=None, rho=None, tvisco=1, visco=0.01, tdensity=0, seed=7):
    pts, _ = cm._lattice(0, n, 0, n, 0, n, dp)
    u = cm.uniform01(seed, pts.size).reshape(-1, 3)
    pos = pts + (2 * u - 1) * 1e-4 * dp
    h = 1.2 * math.sqrt(3) * dp
    cs0 = 20 * math.sqrt(9.81 * 0.3)
    vr = np.zeros((len(pos), 4), np.float32)
    vr[:, 3] = 1000.0 if rho is None else rho
    if vel is not None:
        vr[:, :3] = vel(pos)
    mn, mx = cm.map_limits(pos, h)
    p = cm.Parameters(tvisco=tvisco, visco=visco, tdensity=tdensity)
    return cm.Case(name="cube", dp=dp, h=h, rhop0=1000.0, gamma=7.0, cs0=cs0,
                   b=cs0 ** 2 * 1000 / 7, massfluid=1000 * dp ** 3, massbound=1000 * dp ** 3,
                   gravity=(0, 0, -9.81), hswl=0.3, params=p,
                   idp=np.arange(len(pos), dtype=np.uint32), pos=pos, velrho=vr, nbound=0,
                   nfluid=len(pos), maprealposmin=mn, maprealposmax=mx)


def _forces(case, prec=1):
    o = OracleSim(case, prec=prec)
    o.interact()
    ace, ar = o.forces()
    return o, ace.astype(np.float64), ar.astype(np.float64), o.download()


def test_continuity_of_a_linear_velocity_field():
    """d rho/dt = -rho div v: exact for v = A x on a lattice away from the
    free surface (first-order consistency of the SPH gradient)."""
    A = np.array([[1.0, 0.3, 0.0], [0.0, 2.0, 0.0], [0.1, 0.0, 3.0]]) * 1e-2
    c = _cube(vel=lambda x: (x - x.mean(0)) @ A.T)
    o, ace, ar, d = _forces(c)
    h, L = c.h, 14 * c.dp
    inner = np.all((d["pos"] > 2.2 * h) & (d["pos"] < L - 2.2 * h), axis=1)
    assert inner.sum() > 30
    np.testing.assert_allclose(ar[inner], -1000 * np.trace(A), rtol=2e-3)


@pytest.mark.parametrize("tvisco,visco", [(1, 0.05), (3, 1e-3)])
def test_pair_forces_conserve_momentum(tvisco, visco):
    """Pressure and viscous pair forces are antisymmetric: sum m a = 0 for a
    fluid-only cloud with arbitrary velocities and densities."""
    rng = np.random.default_rng(11)
    c = _cube(vel=lambda x: rng.standard_normal(x.shape) * 0.05,
              rho=(1000 + rng.standard_normal(14 ** 3) * 5).astype(np.float32), tvisco=tvisco,
              visco=visco)
    o, ace, ar, d = _forces(c)
    m = c.massfluid
    tot = (m * ace).sum(0)
    scale = np.abs(m * ace).sum()
    assert scale > 1.0
    assert np.abs(tot).max() < 1e-8 * scale


def test_artificial_viscosity_acts_on_approaching_pairs_only():
    """A receding (expanding) velocity field feels no artificial viscosity;
    a converging one does."""
    out = {}
    for sign in (1.0, -1.0):
        for visco in (0.0, 0.1):
            c = _cube(vel=lambda x: sign * 0.05 * (x - x.mean(0)), visco=visco)
            out[sign, visco] = _forces(c)[1]
    np.testing.assert_array_equal(out[1.0, 0.0], out[1.0, 0.1])
    assert np.abs(out[-1.0, 0.1] - out[-1.0, 0.0]).max() > 1e-3


def test_hydrostatic_column_balances_gravity():
    """In the resting column the pressure force per unit mass is +g away from
    walls and surface (sign and size of the pressure term, EOS, kernel)."""
    c = cm.dam_break(dp=0.02, params=cm.Parameters(tdensity=0))
    o, ace, ar, d = _forces(c)
    st = o.stats()
    x, h = d["pos"], c.h
    inner = ((np.arange(len(x)) >= st["npb"]) & (x[:, 0] > 2.5 * h) & (x[:, 0] < 0.4 - 2.5 * h) &
             (x[:, 1] > 2.5 * h) & (x[:, 1] < 0.67 - 2.5 * h) & (x[:, 2] > 2.5 * h) &
             (x[:, 2] < 0.3 - 2.5 * h))
    assert inner.sum() > 500
    assert abs(np.median(ace[inner, 2]) - 9.81) < 0.02 * 9.81
    assert np.abs(ace[inner, :2]).max() < 0.01


def test_fourtakas_term_vanishes_for_the_hydrostatic_state():
    """DDT2/DDT2Full subtract the hydrostatic density difference: for the
    resting column the density diffusion is small, the Molteni term (DDT1)
    is not (wrong sign or missing drho fails this)."""
    res = {}
    for tdd in (1, 3):
        c = cm.dam_break(dp=0.02, params=cm.Parameters(tdensity=tdd))
        o, ace, ar, d = _forces(c)
        st = o.stats()
        x, h = d["pos"], c.h
        inner = ((np.arange(len(x)) >= st["npb"]) & (x[:, 2] > 2.5 * h) & (x[:, 2] < 0.3 - 2.5 * h)
                 & (x[:, 0] > 2.5 * h) & (x[:, 0] < 0.4 - 2.5 * h))
        res[tdd] = np.median(np.abs(ar[inner]))
    assert res[1] > 0.02
    assert res[3] < 0.2 * res[1]


def _sorted_state(case, prec=1):
    o = OracleSim(case, prec=prec)
    d = o.download()
    dcell, poscell, cte = o.arrays()
    c6, be = o.celldiv()
    st = o.stats()
    odiv = orc.OrcDiv(1, c6[3] - c6[0] + 1, c6[4] - c6[1] + 1, c6[5] - c6[2] + 1,
                      (c6[3] - c6[0] + 1) * (c6[4] - c6[1] + 1), 0, c6[0], c6[1], c6[2],
                      be.ctypes.data)
    odiv.cellfluid = odiv.ncw * odiv.ncz + 1
    return o, d, dcell, poscell, cte, odiv, be, st


@pytest.mark.parametrize("tvisco,tddt", [(1, 3), (3, 2)])
def test_cell_list_sums_equal_brute_force(tvisco, tddt):
    """The interaction over cell rows equals the O(N^2) sum over all pairs
    (no missed or extra neighbours) in FP64."""
    c = cm.dam_break(dp=0.04)
    rng = np.random.default_rng(3)
    c.velrho[c.nbound:, :3] = rng.standard_normal((c.nfluid, 3)) * 0.05
    o, d, dcell, poscell, cte, odiv, be, st = _sorted_state(c)
    n, npb, npbok = st["np"], st["npb"], st["npbok"]
    lib = orc.load()
    pos = np.ascontiguousarray(d["pos"])
    vr = np.ascontiguousarray(d["velrho"])
    ace = np.zeros((n, 3), np.float32)
    ar = np.zeros(n, np.float32)
    vdt = np.zeros(n, np.float32)
    strain = np.zeros((n, 6), np.float32)
    lib.orc_interaction(1, orc.SEM_F64, tvisco, tddt, 0.01, 0.01, C.byref(cte), C.byref(odiv), n,
                        npb, npbok, dcell.ctypes.data, poscell.ctypes.data, pos.ctypes.data,
                        vr.ctypes.data, None, ace.ctypes.data, ar.ctypes.data, vdt.ctypes.data,
                        strain.ctypes.data, None)
    bace = np.zeros((n, 3))
    bar = np.zeros(n)
    bvdt = np.zeros(n)
    lib.orc_bruteforce_fp64(C.byref(cte), tvisco, tddt, 0.01, 0.01, n, npb, pos.ctypes.data,
                            vr.ctypes.data, bace.ctypes.data, bar.ctypes.data, bvdt.ctypes.data)
    fl = slice(npb, None)
    np.testing.assert_allclose(ace[fl], bace[fl], rtol=1e-5, atol=1e-6 * np.abs(bace[fl]).max())
    np.testing.assert_allclose(ar[fl], bar[fl], rtol=1e-5, atol=1e-6 * np.abs(bar[fl]).max())
    np.testing.assert_allclose(vdt[fl], bvdt[fl], rtol=1e-5, atol=1e-6 * np.abs(bvdt).max())


def test_pips_counts_equal_brute_force():
    """JDsPips rule: pairs with r^2 <= (2h)^2 and >= 1e-18; fluid p1 over
    fluid+boundary, boundary p1 (NpbOk) over fluid."""
    c = cm.dam_break(dp=0.04)
    o, d, dcell, poscell, cte, odiv, be, st = _sorted_state(c, prec=0)
    npb = st["npb"]
    real = o.pips()
    # Brute force from the FP32 cell-relative positions (the rule's own
    # distance), all pairs.
    ps = poscell
    cell = ps[:, 3].view(np.uint32)
    cxyz = np.stack([cell >> 19, (cell & 0x7FE00) >> 9, cell & 0x1FF], 1).astype(np.int64)
    pcs = np.float32(cte.poscellsize)
    rf = rb = 0
    ks2 = np.float32(cte.kernelsize2)
    for p1 in range(len(ps)):
        dr = (ps[p1, :3] - ps[:, :3]) + pcs * (cxyz[p1] - cxyz).astype(np.float32)
        rr2 = (dr * dr).sum(1, dtype=np.float32)
        hit = (rr2 <= ks2) & (rr2 >= np.float32(1e-18))
        if p1 >= npb:
            rf += int(hit.sum())
        elif p1 < st["npbok"]:
            rb += int(hit[npb:].sum())
    assert real[:2] == [rf, rb]


# ------------------------------------------------------------------ integrators and dt
def test_symplectic_is_exact_for_constant_acceleration():
    """Predictor + corrector with constant ace: v = v0 + (a+g) dt and
    x = x0 + v0 dt + (a+g) dt^2 / 2 (to FP32 velocity rounding)."""
    lib = orc.load()
    n, npb = 64, 0
    rng = np.random.default_rng(2)
    vr0 = np.zeros((n, 4), np.float32)
    vr0[:, :3] = rng.standard_normal((n, 3)).astype(np.float32)
    vr0[:, 3] = 1000
    ace = rng.standard_normal((n, 3)).astype(np.float32)
    ar = np.zeros(n, np.float32)
    g = (C.c_float * 3)(0, 0, -9.81)
    code = np.full(n, 0x1800, np.uint16)
    dt = 1e-3
    mov = np.zeros((n, 3))
    vr_half = np.zeros_like(vr0)
    lib.orc_sym_pre(n, npb, vr0.ctypes.data, ar.ctypes.data, ace.ctypes.data, dt / 2, 1000.0,
                    700.0, 1300.0, g, code.ctypes.data, mov.ctypes.data, vr_half.ctypes.data)
    np.testing.assert_allclose(mov, vr0[:, :3].astype(np.float64) * dt / 2, rtol=1e-15)
    vr1 = vr_half.copy()
    lib.orc_sym_cor(n, npb, vr0.ctypes.data, ar.ctypes.data, ace.ctypes.data, dt / 2, dt, 1000.0,
                    700.0, 1300.0, g, code.ctypes.data, mov.ctypes.data, vr1.ctypes.data)
    acg = ace.astype(np.float64) + np.array([0, 0, -9.81], np.float32).astype(np.float64)
    np.testing.assert_allclose(vr1[:, :3], vr0[:, :3] + acg * dt, rtol=1e-6, atol=1e-7)
    x_exact = vr0[:, :3].astype(np.float64) * dt + 0.5 * acg * dt * dt
    np.testing.assert_allclose(mov, x_exact, rtol=1e-6, atol=1e-12)
    # Density: rho_pre (2 - eps)/(2 + eps), eps = -ar/rho dt: unchanged for ar = 0.
    assert np.all(vr1[:, 3] == 1000)


def test_symplectic_density_is_second_order():
    """With a constant ar the corrector's rational update matches
    rho(t) = rho0 + ar t to O(dt^2)."""
    lib = orc.load()
    n = 4
    vr0 = np.zeros((n, 4), np.float32)
    vr0[:, 3] = 1000
    ace = np.zeros((n, 3), np.float32)
    ar = np.full(n, 50.0, np.float32)
    g = (C.c_float * 3)(0, 0, 0)
    code = np.full(n, 0x1800, np.uint16)
    mov = np.zeros((n, 3))
    errs = []
    for dt in (1e-3, 5e-4):
        vh = np.zeros_like(vr0)
        lib.orc_sym_pre(n, 0, vr0.ctypes.data, ar.ctypes.data, ace.ctypes.data, dt / 2, 1000.0,
                        700.0, 1300.0, g, code.ctypes.data, mov.ctypes.data, vh.ctypes.data)
        lib.orc_sym_cor(n, 0, vr0.ctypes.data, ar.ctypes.data, ace.ctypes.data, dt / 2, dt,
                        1000.0, 700.0, 1300.0, g, code.ctypes.data, mov.ctypes.data,
                        vh.ctypes.data)
        errs.append(abs(float(vh[0, 3]) - (1000 + 50 * dt * 1000 / float(vh[0, 3]) * 0 + 50 * dt)))
    assert errs[0] < 1e-3


def test_dt_variable_closed_form():
    lib = orc.load()
    h, cs0, cfl = 0.0207846, 34.31, 0.2
    for velmax, acemax, vdt in ((0.0, 0.0, 0.0), (1.0, 12.0, 3.0), (9.0, 500.0, 40.0)):
        hf, vf = float(np.float32(h)), float(np.float32(vdt))
        dt1 = math.sqrt(hf / acemax) if acemax else math.inf
        dt2 = hf / (max(cs0, 10 * velmax) + hf * vf)
        exp = cfl * min(dt1, dt2)
        got = lib.orc_dt_variable(velmax, acemax, vdt, h, cs0, cfl, 0.0)
        assert math.isclose(got, exp, rel_tol=1e-12)
    assert lib.orc_dt_variable(1e6, 1e12, 1e6, h, cs0, cfl, 1e-5) == 1e-5


def test_interaction_range_equals_full_interaction_on_its_particles():
    """orc_interaction_range (the oracle on a sample of a large state) gives
    the full interaction's bits on the sampled fluid and boundary particles
    and leaves the others alone."""
    import ctypes as C
    import numpy as np
    import oracle.oracle as orc
    from dualsphysics_amd import case as casemod
    from oracle.oracle import OracleSim
    c = casemod.dam_break(dp=0.04)
    o = OracleSim(c)
    try:
        od = o.download()
        dcell, poscell, cte = o.arrays()
        celldiv, be = o.celldiv()
        st = o.stats()
    finally:
        o.close()
    n, npb, npbok = st["np"], st["npb"], st["npbok"]
    c6 = celldiv
    ncx, ncy, ncz = c6[3] - c6[0] + 1, c6[4] - c6[1] + 1, c6[5] - c6[2] + 1
    odiv = orc.OrcDiv(1, int(ncx), int(ncy), int(ncz), int(ncx * ncy), int(ncx * ncy * ncz + 1),
                      int(c6[0]), int(c6[1]), int(c6[2]), be.ctypes.data)
    pos = np.ascontiguousarray(od["pos"])
    out = {}
    for rng in (None, (npb + 100, npb + 900, 50, 300)):
        a = np.full((n, 3), 7.0, np.float32)
        ar = np.full(n, 7.0, np.float32)
        v = np.zeros(n, np.float32)
        args = (0, orc.SEM_GPU, 1, 3, 0.01, 0.01, C.byref(cte), C.byref(odiv), n, npb, npbok,
                dcell.ctypes.data, poscell.ctypes.data, pos.ctypes.data,
                od["velrho"].ctypes.data, None, a.ctypes.data, ar.ctypes.data, v.ctypes.data,
                None, None, None, None, 0, None, None, None, None, None)
        r = (0, 2 ** 62, 0, 2 ** 62) if rng is None else rng
        orc.load().orc_interaction_range(*args, *r)
        out[rng is None] = (a, ar, v)
    (fa, far, fv), (sa, sar, sv) = out[True], out[False]
    fs, bs = slice(npb + 100, npb + 900), slice(50, 300)
    np.testing.assert_array_equal(sa[fs], fa[fs])
    np.testing.assert_array_equal(sar[fs], far[fs])
    np.testing.assert_array_equal(sar[bs], far[bs])
    np.testing.assert_array_equal(sv[bs], fv[bs])
    assert np.all(sa[npb + 900:] == 7.0) and np.all(sar[300:npb] == 7.0)
    assert np.all(sa[npb:npb + 100] == 7.0)
