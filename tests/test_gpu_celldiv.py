"""Neighbour list on the GPU vs the oracle: bit-exact cell keys, stable sort
permutation, BeginEndCell ranges, gathers and PosCell (reference rows a1-a11)."""
import ctypes as C

import numpy as np
import pytest

from dualsphysics_amd import case as casemod
from tests.helpers import Dev, u3

pytestmark = pytest.mark.gpu


def _sort_gpu(lib, ctx, keys, limit):
    n = keys.size
    k = Dev(lib, ctx, keys.astype(np.uint32))
    v = Dev(lib, ctx, np.arange(n, dtype=np.uint32))
    rc = lib.dsph_sort_by_cell(ctx, k.ptr, v.ptr, n, limit)
    assert rc == 0, lib.dsph_last_error(ctx)
    return k.get(), v.get()


@pytest.mark.parametrize("n", [0, 1, 2, 63, 64, 65, 1000, 4095, 4096, 4097, 65537, 1000003])
@pytest.mark.parametrize("limit", [1, 2, 255, 256, 257, 70001, (1 << 21) + 5, 1 << 24])
def test_sort_random_keys_is_stable(lib, ctx, n, limit):
    rng = np.random.default_rng(n * 7 + limit)
    keys = rng.integers(0, limit, n, dtype=np.uint64).astype(np.uint32)
    ks, vs = _sort_gpu(lib, ctx, keys, limit)
    perm = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(vs, perm)
    np.testing.assert_array_equal(ks, keys[perm])


@pytest.mark.parametrize("pattern", ["equal", "sorted", "reversed", "nearly_sorted", "max_key"])
def test_sort_patterns(lib, ctx, pattern):
    n, limit = 300001, 1 << 20
    rng = np.random.default_rng(3)
    if pattern == "equal":
        keys = np.full(n, 12345, np.uint32)
    elif pattern == "sorted":
        keys = np.sort(rng.integers(0, limit, n)).astype(np.uint32)
    elif pattern == "reversed":
        keys = np.sort(rng.integers(0, limit, n))[::-1].astype(np.uint32)
    elif pattern == "nearly_sorted":
        keys = np.sort(rng.integers(0, limit, n)).astype(np.int64)
        idx = rng.integers(0, n, n // 100)
        keys[idx] += rng.integers(-3, 4, idx.size)
        keys = np.clip(keys, 0, limit - 1).astype(np.uint32)
    else:
        keys = np.full(n, limit - 1, np.uint32)
        keys[::7] = 0
    ks, vs = _sort_gpu(lib, ctx, keys, limit)
    perm = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(vs, perm)
    np.testing.assert_array_equal(ks, keys[perm])


@pytest.fixture(scope="module")
def cstate():
    """A dam break with some excluded particles, prepared by the oracle."""
    from oracle.oracle import OracleSim
    c = casemod.dam_break(dp=0.02)
    o = OracleSim(c)
    d = o.download()
    dcell, poscell, cte = o.arrays()
    celldiv, be = o.celldiv()
    return c, o, d, dcell, poscell, cte, celldiv, be


def test_cell_limits(lib, ctx, cstate):
    import oracle.oracle as orc
    c, o, d, dcell, poscell, cte, celldiv, be = cstate
    code = d["code"].copy()
    code[c.nbound + 5] = 0x1800 | 0x8000          # an excluded fluid particle
    olib = orc.load()
    for n, pini in [(c.nbound, 0), (d["idp"].size - c.nbound, c.nbound), (0, 0)]:
        mn, mx = (C.c_uint32 * 3)(), (C.c_uint32 * 3)()
        rc = lib.dsph_cell_limits_sync(ctx, n, pini, cte.cellcode, Dev(lib, ctx, dcell).ptr,
                                       Dev(lib, ctx, code).ptr, mn, mx)
        assert rc == 0
        omn, omx = np.zeros(3, np.uint32), np.zeros(3, np.uint32)
        olib.orc_limits(n, pini, cte.cellcode, dcell.ctypes.data, code.ctypes.data,
                        omn.ctypes.data, omx.ctypes.data)
        assert list(mn) == list(omn) and list(mx) == list(omx)


@pytest.mark.parametrize("n", [1, 257, 3_000_017])
def test_cell_limits_many_blocks(lib, ctx, cstate, n):
    """Limits over up to 3M particles (the grid-stride loop over 1024 block
    partials and the final block): the cells of the dam break repeated with
    random exclusions, against the oracle; all excluded gives min > max."""
    import oracle.oracle as orc
    c, o, d, dcell0, poscell, cte, celldiv, be = cstate
    rng = np.random.default_rng(n)
    pick = rng.integers(0, dcell0.size, n)
    dcell = np.ascontiguousarray(dcell0[pick])
    code = np.ascontiguousarray(d["code"][pick])
    code[rng.random(n) < 0.3] |= 0x8000
    olib = orc.load()
    for cd in (code, code | np.uint16(0x8000)):
        mn, mx = (C.c_uint32 * 3)(), (C.c_uint32 * 3)()
        rc = lib.dsph_cell_limits_sync(ctx, n, 0, cte.cellcode, Dev(lib, ctx, dcell).ptr,
                                       Dev(lib, ctx, cd).ptr, mn, mx)
        assert rc == 0, lib.dsph_last_error(ctx)
        omn, omx = np.zeros(3, np.uint32), np.zeros(3, np.uint32)
        olib.orc_limits(n, 0, cte.cellcode, dcell.ctypes.data, cd.ctypes.data,
                        omn.ctypes.data, omx.ctypes.data)
        assert list(mn) == list(omn) and list(mx) == list(omx)
    assert all(a > b for a, b in zip(mn, mx))   # every particle excluded


@pytest.mark.parametrize("full", [1, 0])
def test_presort_sort_beginend_match_oracle(lib, ctx, cstate, full):
    import oracle.oracle as orc
    c, o, d, dcell, poscell, cte, celldiv, be = cstate
    olib = orc.load()
    np_ = d["idp"].size
    npb = c.nbound
    code = d["code"].copy()
    dcell = dcell.copy()
    rng = np.random.default_rng(11)
    # Excluded particles of every kind land in the special boxes.
    for k, sp in zip(rng.choice(np.arange(npb, np_), 6, replace=False),
                     [0x8000, 0x6000, 0xa000, 0x4000, 0x8000, 0x6000]):
        code[k] = (code[k] & 0x1fff) | sp
        dcell[k] = 0xffffffff if sp != 0x4000 else dcell[k]
    cmin = celldiv[:3]
    nc = [celldiv[3 + i] - celldiv[i] + 1 for i in range(3)]
    nct = nc[0] * nc[1] * nc[2]
    nctt = 2 * nct + 5
    dd, dc = Dev(lib, ctx, dcell), Dev(lib, ctx, code)
    cp = Dev.zeros(lib, ctx, np_, np.uint32)
    sp = Dev.zeros(lib, ctx, np_, np.uint32)
    if full:
        rc = lib.dsph_presort_full(ctx, np_, cte.cellcode, dd.ptr, dc.ptr, u3(cmin), u3(nc), cp.ptr, sp.ptr)
    else:
        rc = lib.dsph_presort_fluid(ctx, np_ - npb, npb, cte.cellcode, dd.ptr, dc.ptr, u3(cmin),
                                    u3(nc), cp.ptr, sp.ptr)
    assert rc == 0
    ocp = np.zeros(np_, np.uint32)
    acmin, anc = np.array(cmin, np.uint32), np.array(nc, np.uint32)
    olib.orc_presort(full, np_, npb, cte.cellcode, dcell.ctypes.data, code.ctypes.data,
                     acmin.ctypes.data, anc.ctypes.data, ocp.ctypes.data)
    pini = 0 if full else npb
    gcp = cp.get()
    np.testing.assert_array_equal(gcp[pini:], ocp[pini:])
    # Sort (GPU radix vs oracle counting sort) and BeginEndCell.
    if full:
        rc = lib.dsph_sort_by_cell(ctx, cp.ptr, sp.ptr, np_, nctt)
    else:
        rc = lib.dsph_sort_by_cell(ctx, cp.off(4 * npb), sp.off(4 * npb), np_ - npb, nctt)
    assert rc == 0
    osp = np.zeros(np_, np.uint32)
    obe = np.zeros(2 * nctt, np.int32)
    olib.orc_sort_cells(np_, pini, nctt, 0 if full else nct + 1, ocp.ctypes.data,
                        osp.ctypes.data, obe.ctypes.data)
    np.testing.assert_array_equal(sp.get()[pini:], osp[pini:])
    np.testing.assert_array_equal(cp.get()[pini:], ocp[pini:])
    bc = Dev(lib, ctx, np.full(2 * nctt, -7, np.int32))
    rc = lib.dsph_begin_end_cell(ctx, full, np_, npb, nctt, nct + 1, cp.ptr, bc.ptr)
    assert rc == 0
    gbe = bc.get()
    k0 = 0 if full else 2 * (nct + 1)
    np.testing.assert_array_equal(gbe[k0:], obe[k0:])
    # The special-box readback used by the divide.
    out = (C.c_int32 * 8)()
    assert lib.dsph_cell_begin_end_sync(ctx, bc.ptr, 2 * nct + 1, 4, out) == 0
    np.testing.assert_array_equal(np.array(out), obe[2 * (2 * nct + 1): 2 * (2 * nct + 5)])


def test_gather_and_poscell(lib, ctx, cstate):
    import oracle.oracle as orc
    c, o, d, dcell, poscell, cte, celldiv, be = cstate
    np_ = d["idp"].size
    rng = np.random.default_rng(5)
    sortpart = np.arange(np_, dtype=np.uint32)
    pini = c.nbound
    sortpart[pini:] = pini + rng.permutation(np_ - pini).astype(np.uint32)
    idp, code = d["idp"], d["code"]
    posxy = np.ascontiguousarray(d["pos"][:, :2])
    posz = np.ascontiguousarray(d["pos"][:, 2])
    vr = d["velrho"]
    ins = [Dev(lib, ctx, a) for a in (sortpart, idp, code, dcell, posxy, posz, vr)]
    outs = [Dev.zeros(lib, ctx, a.shape, a.dtype) for a in (idp, code, dcell, posxy, posz, vr)]
    pc = Dev.zeros(lib, ctx, (np_, 4), np.float32)
    posmin = (C.c_double * 3)(*cte.mapposmin)
    rc = lib.dsph_sort_data_basic_poscell(ctx, np_, pini, *[b.ptr for b in ins],
                                          *[b.ptr for b in outs], posmin, cte.poscellsize, pc.ptr)
    assert rc == 0
    src = np.where(np.arange(np_) < pini, np.arange(np_), sortpart)
    for a, b in zip((idp, code, dcell, posxy, posz, vr), outs):
        np.testing.assert_array_equal(b.get(), a[src])
    pos2 = d["pos"][src]
    opc = np.zeros((np_, 4), np.float32)
    pos2 = np.ascontiguousarray(pos2)
    pmin = np.array(cte.mapposmin, np.float64)
    orc.load().orc_poscell(np_, pos2.ctypes.data, pmin.ctypes.data, cte.poscellsize,
                           opc.ctypes.data)
    np.testing.assert_array_equal(pc.get().view(np.uint32), opc.view(np.uint32))


def test_sim_divide_matches_oracle(cstate):
    """Whole divide through the host driver: same order, same cell domain."""
    from dualsphysics_amd.sim import SphGpuSingle
    c, o, d, dcell, poscell, cte, celldiv, be = cstate
    s = SphGpuSingle(c)
    g = s.download()
    np.testing.assert_array_equal(g["idp"], d["idp"])
    np.testing.assert_array_equal(g["pos"], d["pos"])
    np.testing.assert_array_equal(g["velrho"], d["velrho"])
    assert s.stats()["nct"] == o.stats()["nct"]
    assert s.stats()["npbok"] == o.stats()["npbok"]
    s.close()
