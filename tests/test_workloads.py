"""Host-side checks of the parametric family used by the learned warm start:
every instance is feasible by construction (x0(z) in the box, A x0(z) within
the row bounds), instances are reproducible from their seeds, and the shared
data does not change between instances."""
import numpy as np

from pdhcg_amd import workloads


def test_parametric_family_instances_are_feasible_and_seeded():
    fam = workloads.ParametricFamily(n=300, m=150, nnz=3000, rank=10)
    Z, ps = fam.sample(16, 3)
    assert Z.shape == (16, 16) and len(ps) == 16
    for z, p in zip(Z, ps):
        x0 = np.clip(2.5 + fam.W_x @ z.astype(np.float64), 0.0, 5.0)
        ax = p.A @ x0
        assert np.all(p.con_lb <= ax) and np.all(ax <= p.con_ub)
        assert np.all(p.con_lb < p.con_ub)
        assert p.A is fam.A and p.R is fam.R
    Z2, ps2 = fam.sample(16, 3)
    assert np.array_equal(Z, Z2)
    assert all(np.array_equal(a.c, b.c) and np.array_equal(a.con_lb, b.con_lb) for a, b in zip(ps, ps2))
    assert not np.array_equal(ps[0].c, ps[1].c)
    # the objective moves with z by c_spread per coordinate (plus 0.01 noise)
    dc = np.std(np.stack([p.c for p in ps]) - fam.c_bar)
    assert 0.05 < dc < 0.2
