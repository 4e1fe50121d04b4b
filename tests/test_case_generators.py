"""Case generators on the CPU: the plane-wise wall lattice equals the
filtered full-box lattice, and the ensemble draws are deterministic and
within their stated size range."""
import numpy as np

from dualsphysics_amd import case as cm
from dualsphysics_amd import ensemble as ens


def test_box_walls_equal_filtered_full_lattice():
    for nxt, nyt, nzt, layers in [(10, 7, 5, 2), (3, 4, 2, 1), (17, 9, 11, 3), (1, 1, 1, 2)]:
        pts, idx = cm._lattice(-layers, nxt + layers, -layers, nyt + layers, -layers, nzt, 0.01)
        inside = ((idx[:, 0] >= 0) & (idx[:, 0] < nxt) & (idx[:, 1] >= 0) & (idx[:, 1] < nyt) &
                  (idx[:, 2] >= 0))
        np.testing.assert_array_equal(cm._box_walls(nxt, nyt, nzt, layers, 0.01), pts[~inside])


def test_dam_break_cfg0_counts():
    c = cm.dam_break_cfg0()
    assert c.nfluid == 80_400 and c.nbound == 60_248


def test_ensemble_draws_deterministic_and_in_range():
    a = ens.dam_break_ensemble(12, seed=5)
    b = ens.dam_break_ensemble(12, seed=5)
    for x, y in zip(a, b):
        assert x.dp == y.dp and np.array_equal(x.pos, y.pos)
    sizes = np.array([c.np for c in a])
    assert sizes.min() >= 3000 and sizes.max() <= 250_000
    assert len({c.dp for c in a}) == len(a)   # uneven cases
    for c in a:
        assert 0.5 <= c.pos[:, 0].max() + 0.0 and c.hswl <= 0.5 + 1e-12
