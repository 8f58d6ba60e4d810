"""Domain_Decomposition with NTarget > 1 (domain.c:67-145) through the C
ABI against the oracle on the same particles and P.Cost: the same bunches
as top nodes (keys, levels, particle ranges, float centres), the same
target for each and the same number of passes; then the tree on those top
nodes and the FP64 walk, bit for bit as in the other parity tests."""
import numpy as np
import pytest

import oracle
from tandav_amd import Engine, TDVError, ics
from test_gpu_parity import _compare

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ewald():
    return oracle.cached_ewald_table()


@pytest.mark.parametrize("case,ntarget", [("clustered", 2), ("clustered", 8),
                                          ("clustered", 16),
                                          ("periodic", 8), ("periodic", 64)])
def test_domain_targets_match_oracle(case, ntarget, ewald):
    from test_oracle_tree import clustered
    if case == "periodic":
        pos, _, m = ics.zeldovich_box(32, seed=7)
        kw = dict(periodic=True, softening=1 / 1280.)
    else:
        pos, m = clustered(30000, 5)
        kw = dict(periodic=False, softening=1e-3)
    table = ewald if kw["periodic"] else None
    print("case", case, "ntarget", ntarget, flush=True)
    s = oracle.Sim(pos, m, alpha_rel=0.005, nthreads=16, ewald_table=table,
                   **kw)
    s.setup_domain()
    s.domain_decomposition()
    s.compute_acceleration(True)            # P.Cost for the decomposition
    tasks_s, it_s = s.domain_decomposition_targets(ntarget)
    s.compute_acceleration(False)
    e = Engine(alpha_rel=0.005, compat=True, **kw)
    if kw["periodic"]:
        e.set_ewald_table(ewald)
    e.set_particles(pos, m)
    e.setup_domain()
    e.domain_decomposition()
    e.compute_acceleration(True)
    tasks_e, it_e = e.domain_decomposition_targets(ntarget)
    assert it_e == it_s
    np.testing.assert_array_equal(tasks_e, tasks_s)
    assert len(tasks_e) == len(s.top)
    e.compute_acceleration(False)
    _compare(s, e)


def test_domain_targets_rejects_too_few_bunches():
    """Every particle in one level-2 cell: one bunch for two targets (the
    reference would index past its bunch list) is an error."""
    g = np.random.Generator(np.random.Philox(4))
    pos = 0.01 * g.random((5000, 3))        # one cell of the unit box
    m = np.full(5000, 1.0 / 5000)
    e = Engine(periodic=True, boxsize=1.0, softening=1e-3)
    e.set_particles(pos, m)
    with pytest.raises(TDVError):
        e.domain_decomposition_targets(2)
