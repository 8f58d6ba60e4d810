"""The dynamic tree update on the device (tree kicks inside
tdv_kick_halfstep, tdv_tree_update_drift) against oracle/tree_update.py,
Tandav's tree_update.c restated: every node's Dp, UPDATED bit and CoM and
every top node's float Dp, Level sign and CoM bit for bit, for an active
subset spread over time bins; and the identity Dp Mass = sum of the
node's kicked particles' m dt a."""
import numpy as np
import pytest

from oracle import tree_update as U
from tandav_amd import Engine, engine, ics

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("periodic", [True, False])
def test_tree_kicks_and_drift_match_oracle(periodic):
    if periodic:
        pos, vel, m = ics.zeldovich_box(24, seed=7)
        soft = 1 / 960.
    else:
        pos, vel, m = ics.hernquist_sphere(20000, seed=2)
        soft = 0.01
    n = len(m)
    e = Engine(periodic=periodic, softening=soft, alpha_rel=0.005)
    e.set_particles(pos, m, vel=vel)
    e.setup_domain()
    e.domain_decomposition()
    e.compute_acceleration(True)
    g = np.random.Generator(np.random.Philox(17))
    bins = g.integers(1, 9, n).astype(np.int32)
    e.make_active_list(31, time_bin=bins)
    e.set_timeline(engine.timeline(False, 0.0, 1.0))
    it = (g.integers(0, 1 << 14, n) << 9).astype(np.uint32)
    e.set_integer_times(it, it)
    act = np.sort(g.choice(n, n // 4, replace=False)).astype(np.int32)
    e.set_active(act)
    tree0, top0 = e.tree(), e.top_nodes()
    st = e.get("mass", "acc", "tree_parent", "time_bin")
    e.kick_halfstep(update_tree=True)
    _, itk = e.integer_times()
    step = (np.left_shift(np.uint32(1), st["time_bin"].astype(np.uint32))
            >> np.uint32(1)).astype(np.uint32)
    t0 = it.astype(np.float64) / 2.0 ** 31
    t1 = (it + step).astype(np.float64) / 2.0 ** 31
    dt = t1 - t0                                   # Particle_Kick_Step
    want_tree, want_top = U.kicks(tree0, top0, st["mass"], st["acc"], dt,
                                  st["tree_parent"], act)
    tree1, top1 = e.tree(), e.top_nodes()
    for f in ("Bitfield", "Dp", "CoM"):
        np.testing.assert_array_equal(tree1[f], want_tree[f], err_msg=f)
    for f in ("Level", "Dp", "CoM"):
        np.testing.assert_array_equal(top1[f], want_top[f], err_msg=f)
    upd = (tree1["Bitfield"] >> 11) & 1 == 1
    assert upd.sum() > 10
    # Dp Mass = the momentum the node's kicked particles took
    mdt_a = (st["mass"] * dt)[:, None] * st["acc"]
    on = np.zeros(n, bool)
    on[act] = True
    for j in np.flatnonzero(upd)[:200]:
        first = _first_particle(tree1, j)
        sel = np.arange(first, first + tree1["Npart"][j])
        want = (mdt_a[sel] * on[sel, None]).sum(0)
        np.testing.assert_allclose(tree1["Dp"][j] * tree1["Mass"][j], want,
                                   rtol=1e-11, atol=1e-13 * np.abs(want).max())
    e.tree_update_drift(0.01)
    want_tree, want_top = U.drift(want_tree, want_top, 0.01)
    tree2, top2 = e.tree(), e.top_nodes()
    for f in ("Bitfield", "Dp", "CoM"):
        np.testing.assert_array_equal(tree2[f], want_tree[f], err_msg=f)
    for f in ("Level", "Dp", "CoM"):
        np.testing.assert_array_equal(top2[f], want_top[f], err_msg=f)
    assert not ((tree2["Bitfield"] >> 11) & 1).any()


def _first_particle(tree, j):
    """A node's first particle: the first bundle at or below it in the
    depth-first array."""
    k = j
    while tree["DNext"][k] >= 0:
        k += 1
    return -int(tree["DNext"][k]) - 1
