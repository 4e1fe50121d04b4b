"""The HIP interaction against the reference solver's own per-particle forces
over the first 10 steps of the dp=0.03 dam break (fixture
ref_forces_dambreak_dp0.03.npz, see test_reference_forces.py), on the
reference's own states in its own particle order.

The HIP path follows the reference GPU build's semantics (cell-relative FP32
PosCell distances, one accumulator set), the fixture is the CPU build's: the
two differ by FP32 rounding of the distances, which the pair sums amplify
where they cancel. Per step the test states the HIP path's error against the
reference next to the same error of the oracle in GPU semantics and of the
FP64 yardstick, and requires:
- HIP vs the oracle in GPU semantics within 1e-5 norm-wise (ace, ar);
- HIP vs reference no further than the oracle's GPU-semantics spread from
  the reference plus that 1e-5 (norm-wise, ace and ar);
- HIP vs FP64 no further than the oracle in GPU semantics is, plus 1e-5;
  the reference CPU build's own distance from FP64 is printed beside it.
"""
import numpy as np
import pytest

from tests.helpers import (load_reference_fixture, per_particle_rel, reference_forces_state,
                           rel_err)
from tests.test_reference_forces import oracle_forces

pytestmark = pytest.mark.gpu

TOL = 1e-5


def test_hip_forces_follow_reference_over_10_steps():
    from dualsphysics_amd.sim import SphGpuSingle
    fx, case = load_reference_fixture("ref_forces_dambreak_dp0.03")
    nb = case.nbound
    for k in range(len(fx["steps"])):
        st = reference_forces_state(fx, case, k)
        s = SphGpuSingle(st)
        try:
            g = s.download()
            np.testing.assert_array_equal(g["idp"], st.idp)
            s.interact()
            ga, gr = s.forces()
        finally:
            s.close()
        idx = np.argsort(g["idp"])
        ga, gr = ga[idx], gr[idx]
        ra, rr = fx["ace"][k], fx["ar"][k]
        oa, orr = oracle_forces(st, 0)
        fa, frr = oracle_forces(st, 1)
        e = {"hip-ref": (rel_err(ga[nb:], ra), rel_err(gr, rr)),
             "orc32gpu-ref": (rel_err(oa[nb:], ra), rel_err(orr, rr)),
             "hip-orc32gpu": (rel_err(ga[nb:], oa[nb:]), rel_err(gr, orr)),
             "hip-fp64": (rel_err(ga[nb:], fa[nb:]), rel_err(gr, frr)),
             "orc32gpu-fp64": (rel_err(oa[nb:], fa[nb:]), rel_err(orr, frr)),
             "ref-fp64": (rel_err(ra, fa[nb:]), rel_err(rr, frr))}
        pp = per_particle_rel(ga[nb:], ra)
        ppo = per_particle_rel(oa[nb:], ra)
        print(f"step {int(fx['steps'][k])}: " +
              " ".join(f"{n} ace {a:.2e} ar {b:.2e};" for n, (a, b) in e.items()) +
              f" per particle hip-ref median {np.median(pp):.1e} max {pp.max():.1e} "
              f"(oracle gpu-semantics median {np.median(ppo):.1e} max {ppo.max():.1e})",
              flush=True)
        for j in (0, 1):
            assert e["hip-orc32gpu"][j] < TOL
            assert e["hip-ref"][j] < e["orc32gpu-ref"][j] + TOL
            assert e["hip-fp64"][j] < e["orc32gpu-fp64"][j] + TOL
