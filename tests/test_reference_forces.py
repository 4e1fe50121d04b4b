"""The oracle against the reference solver's own interaction results.

Fixture ref_forces_dambreak_dp0.03.npz (tests/golden/make_reference_fixtures.py
--forces): the reference CPU build (oracle/ref.mk, its own sources, IEEE
flags) runs 10 steps of the dp=0.03 dam break (Symplectic, artificial
viscosity, DDT2Full), and for every saved state a driver linked with its
objects (oracle/ref_forces.cpp) writes the per-particle ace and ar of its
Interaction_Forces (JSphCpuSingle.cpp:685-748) on that state.

- The oracle in the CPU build's semantics (FP32 pairs on float(FP64
  difference) distances, two passes, delta array, FP64 EOS) reproduces them
  bit for bit for every particle at every step.
- The reference GPU build's semantics (distances from cell-relative FP32
  PosCell, one pass) and the FP64 yardstick are FP32-rounding away from the
  CPU build: that spread (printed) bounds any FP32 restatement that is not the
  CPU build itself.
"""
import dataclasses

import numpy as np
import pytest

from oracle.oracle import OracleSim
from tests.helpers import (load_reference_fixture, per_particle_rel, reference_forces_state,
                           rel_err)

FIXTURE = "ref_forces_dambreak_dp0.03"


@pytest.fixture(scope="module")
def fixture():
    return load_reference_fixture(FIXTURE)


def oracle_forces(st, prec):
    """ace [np,3] and ar [np] of the oracle on state st, in id order."""
    o = OracleSim(st, prec=prec)
    d = o.download()
    np.testing.assert_array_equal(d["idp"], st.idp)   # the reference's order is kept
    o.interact()
    a, r = o.forces()
    idx = np.argsort(d["idp"])
    return a[idx], r[idx]


@pytest.mark.parametrize("k", range(10))
def test_oracle_cpu_semantics_equals_reference_bitwise(fixture, k):
    fx, case = fixture
    nb = case.nbound
    a, r = oracle_forces(reference_forces_state(fx, case, k), prec=2)
    np.testing.assert_array_equal(a[nb:].view(np.uint32), fx["ace"][k].view(np.uint32))
    np.testing.assert_array_equal(r.view(np.uint32), fx["ar"][k].view(np.uint32))


@pytest.mark.parametrize("k", [0, 4, 9])
def test_spread_of_other_semantics(fixture, k):
    """GPU semantics and FP64 against the CPU build: within FP32 rounding
    (norm-wise ace 1.5e-4 / 6e-5, ar 1e-5); the per-particle spread is
    printed (it reaches 1e-3 where the pair sum cancels)."""
    fx, case = fixture
    nb = case.nbound
    st = reference_forces_state(fx, case, k)
    ra, rr = fx["ace"][k], fx["ar"][k]
    for prec, name, tol in ((0, "gpu-semantics", 1.5e-4), (1, "fp64", 6e-5)):
        a, r = oracle_forces(st, prec)
        e_ace, e_ar = rel_err(a[nb:], ra), rel_err(r, rr)
        pp = per_particle_rel(a[nb:], ra)
        print(f"step {int(fx['steps'][k])} {name}: ace {e_ace:.2e} (per particle median "
              f"{np.median(pp):.1e} max {pp.max():.1e}) ar {e_ar:.2e}")
        assert e_ace < tol and e_ar < 1e-5


def test_fixture_detects_a_tiny_planted_error(fixture):
    """The oracle with the fluid mass 2^-20 (1e-6) too large no longer
    matches the reference bit for bit: the comparison sees single-ulp
    changes of the physics."""
    fx, case = fixture
    nb = case.nbound
    st = reference_forces_state(fx, case, 0)
    bad = dataclasses.replace(st, massfluid=st.massfluid * (1 + 2.0 ** -20))
    a, r = oracle_forces(bad, prec=2)
    assert np.count_nonzero(a[nb:] != fx["ace"][0]) > 0.5 * case.nfluid
