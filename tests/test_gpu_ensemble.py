"""Batched ensembles on the GPU: every case of an ensemble equals, bit for
bit, a single-case run of the same particles (fixed cell domain, device-side
dt); graph replays equal eager launches; the autotuner changes the speed,
never the results; cases stop at their own end times."""
import numpy as np
import pytest

from dualsphysics_amd import _native as N
from dualsphysics_amd import case as casemod
from dualsphysics_amd.ensemble import Ensemble, EnsembleError
from dualsphysics_amd.sim import SphGpuSingle

pytestmark = pytest.mark.gpu


def _cases(tvisco=N.VISCO_ARTIFICIAL, tddt=N.DDT_DDT2FULL):
    prm = casemod.Parameters(celldomfixed=1, tvisco=tvisco, tdensity=tddt,
                             visco=0.01 if tvisco == N.VISCO_ARTIFICIAL else 1e-4)
    shapes = [  # dp, tank, column
        (0.04, (1.6, 0.67, 0.4), (0.4, 0.67, 0.3)),
        (0.03, (0.8, 0.4, 0.3), (0.2, 0.4, 0.2)),
        (0.05, (2.0, 0.6, 0.5), (0.5, 0.6, 0.4)),
        (0.035, (0.6, 0.5, 0.25), (0.3, 0.5, 0.12)),
        (0.045, (1.2, 0.67, 0.3), (0.3, 0.67, 0.25)),
    ]
    return [casemod.dam_break(dp=dp, tank=t, column=c, seed=11 + k, params=prm)
            for k, (dp, t, c) in enumerate(shapes)]


def _single(c, nsteps):
    s = SphGpuSingle(c)
    try:
        s.run(nsteps)
        return s.download(), s.stats()
    finally:
        s.close()


def _assert_same(a, b):
    for k in ("idp", "pos", "velrho", "code"):
        assert a[k].shape == b[k].shape, k
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


@pytest.mark.parametrize("tvisco,tddt", [(N.VISCO_ARTIFICIAL, N.DDT_DDT2FULL),
                                         (N.VISCO_LAMINAR, N.DDT_NONE)])
def test_ensemble_equals_single_runs_bitwise(tvisco, tddt):
    cases = _cases(tvisco, tddt)
    nsteps = 12
    ens = Ensemble(cases, tend=10.0)
    try:
        assert ens.run(max_steps=nsteps) == nsteps
        for k, c in enumerate(cases):
            ref, st = _single(c, nsteps)
            got = ens.download(k)
            _assert_same(got, ref)
            s = ens.state(k)
            assert s["nstep"] == nsteps and s["active"] == 1
            assert s["time"] == st["timestep"] and s["dt_last"] == st["dt_last"]
            assert s["np"] == st["np"]
    finally:
        ens.close()


def test_graph_replay_equals_eager_launches():
    cases = _cases()
    a = Ensemble(cases, tend=10.0)
    b = Ensemble(cases, tend=10.0)
    try:
        a.run(max_steps=9, graph=True)
        b.run(max_steps=9, graph=False)
        a.run(max_steps=6, graph=True)   # replays of the same graph
        b.run(max_steps=6, graph=False)
        for k in range(len(cases)):
            _assert_same(a.download(k), b.download(k))
            assert a.state(k)["time"] == b.state(k)["time"]
    finally:
        a.close()
        b.close()


def test_autotuner_changes_speed_not_results():
    cases = _cases()
    a = Ensemble(cases, tend=10.0)
    b = Ensemble(cases, tend=10.0)
    try:
        chosen, ms = a.autotune(reps=2)
        assert chosen and np.all(ms[ms > 0] < 1e4)
        for bucket in range(3):
            sh = a.shape(bucket)
            print(f"bucket {bucket}: {sh}", flush=True)
            if sh["variant"] >= 0:
                assert sh["ms_plain"] > 0 and sh["ms_persistent"] > 0 and sh["ms_bound"] > 0
                best = min(sh["ms_plain"], sh["ms_persistent"])
                assert (sh["ms_persistent"] if sh["persistent"] else sh["ms_plain"]) == best
        for bucket in range(3):
            b.set_variant(bucket, 3)   # the one-lane-per-particle kernel everywhere
        a.run(max_steps=6)
        b.run(max_steps=6)
        for k in range(len(cases)):
            _assert_same(a.download(k), b.download(k))
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("variant,chunk,wpc,bound_block,bound_tile", [
    (0, 1, 10, 64, True),     # about one work item per wave; tile boundary kernel
    (1, 4, 40, 256, False),
    (2, 64, 1, 128, False),   # a handful of waves draining every XCD segment
])
def test_persistent_launch_changes_speed_not_results(variant, chunk, wpc, bound_block,
                                                     bound_tile):
    """The fluid interaction as one persistent launch per bucket (resident
    waves pulling chunks of work items from per-XCD counters, which the step
    clears with a fill kernel before each launch) inside the 24-step graph,
    against the plain one-workgroup-per-wave launches run eagerly."""
    cases = _cases()
    a = Ensemble(cases, tend=10.0)
    b = Ensemble(cases, tend=10.0)
    try:
        for bucket in range(3):
            a.set_shape(bucket, variant, persistent=True, chunk=chunk, waves_per_cu=wpc,
                        bound_block=bound_block, bound_tile=bound_tile)
            sh = a.shape(bucket)
            assert sh["persistent"] == 1 and sh["chunk"] == chunk
            assert sh["bound_block"] == bound_block and sh["bound_tile"] == int(bound_tile)
        assert a.run(max_steps=27, graph=True) == 27   # one 24-step launch + one 3-step launch
        assert b.run(max_steps=27, graph=False) == 27
        assert a.info()["steps_per_graph"] == 24
        for k in range(len(cases)):
            _assert_same(a.download(k), b.download(k))
            assert a.state(k)["time"] == b.state(k)["time"]
    finally:
        a.close()
        b.close()


def test_shape_arguments_are_checked():
    ens = Ensemble(_cases()[:1], tend=1.0)
    try:
        with pytest.raises(EnsembleError, match="tile variant"):
            ens.set_shape(0, 3, persistent=True)
        with pytest.raises(EnsembleError, match="bound_block"):
            ens.set_shape(0, 1, bound_block=96)
        with pytest.raises(EnsembleError, match="bound_block"):
            ens.set_shape(0, 1, bound_block=128, bound_tile=True)
        with pytest.raises(EnsembleError, match="tile variant"):
            ens.set_shape(0, 1, persistent=True, chunk=0)
    finally:
        ens.close()


def test_cases_stop_at_their_end_times():
    cases = _cases()
    tend = np.array([0.002, 0.004, 0.001, 0.003, 0.0025])
    ens = Ensemble(cases, tend=tend)
    try:
        steps = ens.run(check_every=1)
        assert 0 < steps < 1000
        for k in range(len(cases)):
            s = ens.state(k)
            assert s["active"] == 0 and s["flags"] & 1 == 0
            assert s["time"] >= tend[k] and s["time"] - s["dt_last"] < tend[k]
        # a stopped case no longer changes
        before = [ens.download(k) for k in range(len(cases))]
        ens.run(max_steps=3)
        for k in range(len(cases)):
            _assert_same(ens.download(k), before[k])
    finally:
        ens.close()


def test_unsupported_options_are_errors():
    c = casemod.dam_break(dp=0.05, params=casemod.Parameters(tstep=1, celldomfixed=1))
    with pytest.raises(EnsembleError, match="Symplectic"):
        Ensemble([c], tend=1.0)
    c2 = casemod.dam_break(dp=0.05, params=casemod.Parameters(tdensity=N.DDT_DDT2))
    with pytest.raises(EnsembleError, match="DDT2Full"):
        Ensemble([c2], tend=1.0)


@pytest.mark.timeout(1500)
def test_512_case_ensemble_at_its_configuration():
    """The benchmark ensemble itself (512 seeded dam breaks, 7.2k-247k
    particles, 32.9M in all): after 12 steps (4 graph replays of 3 steps) a
    size-stratified sample of 16 cases equals single-case runs bit for bit,
    and the smallest and the largest case have forces within 1e-5 (norm-wise)
    of the FP32 oracle on their state, in the oracle's own particle order."""
    import dataclasses
    import sys
    from dualsphysics_amd.ensemble import dam_break_ensemble
    from oracle.oracle import OracleSim
    from tests.helpers import rel_err
    cases = dam_break_ensemble(512)
    sizes = np.array([c.np for c in cases])
    nsteps = 12
    ens = Ensemble(cases, tend=10.0)
    try:
        ens.autotune(reps=1)
        assert ens.run(max_steps=nsteps) == nsteps
        order = np.argsort(sizes, kind="stable")
        pick = sorted(set(order[np.linspace(0, len(cases) - 1, 16).round().astype(int)].tolist()))
        assert len(pick) == 16 and order[0] in pick and order[-1] in pick
        for k in pick:
            c = cases[k]
            c1 = dataclasses.replace(c, params=dataclasses.replace(c.params, celldomfixed=1))
            s = SphGpuSingle(c1)
            try:
                s.run(nsteps)
                ref, st = s.download(), s.stats()
                got = ens.download(k)
                _assert_same(got, ref)
                es = ens.state(k)
                assert es["nstep"] == nsteps and es["time"] == st["timestep"]
                print(f"case {k}: {c.np} particles, t={st['timestep']:.4e} s: bitwise equal",
                      flush=True)
                if k in (order[0], order[-1]):
                    s.interact()
                    gace, gar = s.forces()
                    o = OracleSim(dataclasses.replace(c1, idp=ref["idp"].copy(),
                                                      pos=ref["pos"].copy(),
                                                      velrho=ref["velrho"].copy()))
                    np.testing.assert_array_equal(o.download()["idp"], ref["idp"])
                    o.interact()
                    oace, oar = o.forces()
                    npb = st["npb"]
                    e_ace = rel_err(gace[npb:], oace[npb:])
                    e_ar = rel_err(gar, oar)
                    print(f"case {k}: forces vs oracle ace {e_ace:.2e} ar {e_ar:.2e}", flush=True)
                    assert e_ace < 1e-5 and e_ar < 1e-5
                    o.close()
            finally:
                s.close()
            sys.stdout.flush()
    finally:
        ens.close()
