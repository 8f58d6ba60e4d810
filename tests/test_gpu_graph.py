"""The persistent work-queue walk and the captured walk graph against the
plain FP32 walk: the same groups run the same code, so the results must be
bit-identical, on every replay, with eager work and synchronisations between
the replays, for the full active set and for subsets (which split each
group's walk over several waves).  The split walk against the whole walk:
the same interactions (Cost and counts equal) summed in another order."""
import numpy as np
import pytest

from tandav_amd import Engine, ics

pytestmark = pytest.mark.gpu


def _engine(pos, m, queue, split=0, periodic=True, softening=1 / 1280.):
    e = Engine(periodic=periodic, softening=softening, alpha_rel=0.01,
               precision="fp32", compat=False, ewald_size=0.125,
               ewald_accuracy=1.0, fp32_order=2, fp32_criterion=3,
               fp32_queue=queue, fp32_split=split)
    e.set_particles(pos, m)
    e.setup_domain()
    e.domain_decomposition()
    e.compute_acceleration(True)
    e.compute_acceleration(False)
    return e


def test_queue_walk_matches_grid_walk():
    pos, _, m = ics.zeldovich_box(32, seed=7)
    a = _engine(pos, m, 0, split=1)
    b = _engine(pos, m, 1, split=1)
    ga, gb = a.get("acc", "cost"), b.get("acc", "cost")
    np.testing.assert_array_equal(ga["acc"], gb["acc"])
    np.testing.assert_array_equal(ga["cost"], gb["cost"])
    assert a.interaction_counts() == b.interaction_counts()


def test_graph_replays_match_eager_walk():
    pos, _, m = ics.zeldovich_box(32, seed=7)
    e = _engine(pos, m, 1)
    e.capture_walk()
    st = e.get("pos", "mass", "acc", "last_acc_mag")
    n = len(m)
    g = np.random.Generator(np.random.Philox(3))
    subsets = [np.arange(n), np.sort(g.choice(n, n // 7, replace=False)),
               np.sort(g.choice(n, 100, replace=False)), np.arange(n)]
    for k, sub in enumerate(subsets):
        e.set_active(sub)
        e.update_particles(st["pos"], acc=st["acc"],
                           last_acc_mag=st["last_acc_mag"])
        e.compute_acceleration(False)          # eager: safe_last_accel + walk
        want = e.get("acc", "cost", "last_acc_mag")
        want_counts = e.interaction_counts()
        for rep in range(3):
            e.update_particles(st["pos"], acc=st["acc"],
                               last_acc_mag=st["last_acc_mag"])
            e.launch_walk()
            got = e.get("acc", "cost", "last_acc_mag")
            for f in ("acc", "cost", "last_acc_mag"):
                np.testing.assert_array_equal(got[f], want[f],
                                              err_msg="%s subset %d rep %d"
                                              % (f, k, rep))
            assert e.interaction_counts() == want_counts
            # eager work between replays
            e.get("pos")
            e.forcetest([0, 1])


@pytest.mark.parametrize("case", ["zeldovich", "clustered", "hernquist"])
def test_split_walk_matches_whole_walk(case):
    if case == "zeldovich":
        pos, _, m = ics.zeldovich_box(32, seed=7)
        kw = {}
    elif case == "clustered":
        pos, _, m = ics.clustered_box(1 << 16, nhalo=32, seed=11)
        kw = dict(softening=1 / 2560.)
    else:
        pos, _, m = ics.hernquist_sphere(30000, seed=4)
        kw = dict(periodic=False, softening=0.01)
    whole = _engine(pos, m, 0, split=1, **kw)
    n = len(m)
    g = np.random.Generator(np.random.Philox(5))
    # the full set (auto: 16384 / groups parts), a forced 3-way split, and
    # a few groups (64 parts)
    ref = whole.get("pos", "last_acc_mag")
    zero = np.zeros((n, 3))
    for split, sub in ((0, np.arange(n)), (3, np.arange(n)),
                       (0, np.sort(g.choice(n, 150, replace=False)))):
        parts = _engine(pos, m, 0, split=split, **kw)
        for e in (whole, parts):         # the same criterion input
            e.set_active(sub)
            e.update_particles(ref["pos"], acc=zero,
                               last_acc_mag=ref["last_acc_mag"])
            e.tree_acceleration(False)
        a, b = whole.get("acc", "cost"), parts.get("acc", "cost")
        np.testing.assert_array_equal(a["cost"][sub], b["cost"][sub])
        assert whole.interaction_counts() == parts.interaction_counts()
        amag = np.linalg.norm(a["acc"][sub], axis=1)
        err = np.linalg.norm(a["acc"][sub] - b["acc"][sub], axis=1) / amag
        # the same interactions summed in another order: FP32 rounding of
        # partial sums, largest where |a| is a small residue of large terms
        # (the near-uniform box), far below the walk's 2e-3 error bar
        assert np.median(err) < 1e-5, (case, split, len(sub), np.median(err))
        assert err.max() < 1e-4, (case, split, len(sub), err.max())
        parts.close()


@pytest.mark.parametrize("case", ["zeldovich", "clustered", "hernquist"])
def test_item_split_walk_matches_queue_split_walk(case):
    """A split walk launched one workgroup per work item (group, part) and
    the work-queue kernel running the same parts: the same interactions in
    the same order, so bit-identical results, for the automatic split of a
    few groups and for forced 3- and 64-way splits of larger sets, eagerly
    and as graph replays (whose grid is sized for the largest active set)."""
    if case == "zeldovich":
        pos, _, m = ics.zeldovich_box(32, seed=7)
        kw = {}
    elif case == "clustered":
        pos, _, m = ics.clustered_box(1 << 16, nhalo=32, seed=11)
        kw = dict(softening=1 / 2560.)
    else:
        pos, _, m = ics.hernquist_sphere(30000, seed=4)
        kw = dict(periodic=False, softening=0.01)
    n = len(m)
    g = np.random.Generator(np.random.Philox(8))
    for split, sub in ((0, np.sort(g.choice(n, 150, replace=False))),
                       (3, np.sort(g.choice(n, n // 5, replace=False))),
                       (64, np.arange(4096, dtype=np.int32))):
        items = _engine(pos, m, 0, split=split, **kw)
        queue = _engine(pos, m, 1, split=split, **kw)
        ref = items.get("pos", "acc", "last_acc_mag")
        for e in (items, queue):
            e.set_active(sub)
            e.update_particles(ref["pos"], acc=ref["acc"],
                               last_acc_mag=ref["last_acc_mag"])
            e.compute_acceleration(False)   # as a replay: safe_last_accel
        a, b = items.get("acc", "cost"), queue.get("acc", "cost")
        np.testing.assert_array_equal(a["acc"], b["acc"], err_msg=case)
        np.testing.assert_array_equal(a["cost"], b["cost"], err_msg=case)
        assert items.interaction_counts() == queue.interaction_counts()
        want = a
        items.capture_walk()
        for rep in range(2):
            items.update_particles(ref["pos"], acc=ref["acc"],
                                   last_acc_mag=ref["last_acc_mag"])
            items.launch_walk()
            got = items.get("acc", "cost")
            np.testing.assert_array_equal(got["acc"], want["acc"])
            np.testing.assert_array_equal(got["cost"], want["cost"])
            items.get("pos")
        items.close()
        queue.close()


def test_cost_ordered_walk_matches_list_order():
    """An eager walk of >= 2048 whole groups launches them by descending
    Cost of the previous walk; a captured walk graph keeps the list order.
    Every group runs the same code either way: bit-identical results, with
    the grid and the work-queue kernel."""
    pos, _, m = ics.zeldovich_box(64, seed=7)        # 4096 groups
    for queue in (0, 1):
        e = _engine(pos, m, queue, split=1)
        st = e.get("pos", "acc", "last_acc_mag")
        e.update_particles(st["pos"], acc=st["acc"],
                           last_acc_mag=st["last_acc_mag"])
        e.compute_acceleration(False)
        want = e.get("acc", "cost")
        assert want["cost"].min() > 0
        e.capture_walk()
        for rep in range(2):
            e.update_particles(st["pos"], acc=st["acc"],
                               last_acc_mag=st["last_acc_mag"])
            e.launch_walk()
            got = e.get("acc", "cost")
            np.testing.assert_array_equal(got["acc"], want["acc"])
            np.testing.assert_array_equal(got["cost"], want["cost"])
        e.close()


def test_device_active_list_matches_host_list():
    """Make_Active_Particle_List on the device (Time_Bin <= max bin, in
    particle order) drives the eager walk and a captured graph exactly as
    the host-made list does; an empty list is an error, as in Tandav."""
    from tandav_amd.engine import TDVError
    pos, _, m = ics.zeldovich_box(32, seed=7)
    n = len(m)
    a = _engine(pos, m, 1)
    b = _engine(pos, m, 1)
    st = a.get("pos", "acc", "last_acc_mag")
    bins = np.random.Generator(np.random.Philox(9)).integers(
        1, 6, n).astype(np.int32)
    b.capture_walk()
    for mb in (1, 3, 5):
        act = np.flatnonzero(bins <= mb).astype(np.int32)
        a.set_active(act)
        for e in (a, b):
            e.update_particles(st["pos"], acc=st["acc"],
                               last_acc_mag=st["last_acc_mag"])
        b.make_active_list(mb, time_bin=bins)
        a.compute_acceleration(False)
        b.launch_walk()
        ga, gb = a.get("acc", "cost"), b.get("acc", "cost")
        np.testing.assert_array_equal(ga["acc"], gb["acc"])
        np.testing.assert_array_equal(ga["cost"], gb["cost"])
    with pytest.raises(TDVError):
        b.make_active_list(0)


@pytest.mark.parametrize("ng", [48, 64])
def test_step_graph_matches_eager_steps(ng):
    """The whole regular step captured as one HIP graph (decomposition,
    build, FP32 walk, kick-drift of one global step) and replayed three
    times, with eager work between the replays, leaves every column, the
    keys, the top nodes and the tree bit for bit as eager steps do.  At
    64^3 (4096 groups) both walk their groups by descending Cost."""
    pos, vel, m = ics.zeldovich_box(ng, seed=7)
    engines = []
    for _ in range(2):
        e = Engine(periodic=True, softening=1 / (40. * ng), alpha_rel=0.01,
                   precision="fp32", compat=False, ewald_size=0.125,
                   ewald_accuracy=1.0, fp32_order=2, fp32_criterion=3)
        e.set_particles(pos, m, vel=vel * 0.3)
        e.setup_domain()
        e.domain_decomposition()
        e.compute_acceleration(True)
        e.domain_decomposition()               # one eager step
        e.compute_acceleration(False)
        e.kick_drift(0.002, 0.004)
        engines.append(e)
    eager, graph = engines
    graph.capture_step(0.002, 0.004)
    cols = ("pos", "vel", "acc", "grav_acc", "cost", "last_acc_mag",
            "tree_parent", "key", "id")
    for step in range(3):
        eager.domain_decomposition()
        eager.compute_acceleration(False)
        eager.kick_drift(0.002, 0.004)
        graph.launch_step()
        a, b = eager.get(*cols), graph.get(*cols)
        for f in cols:
            np.testing.assert_array_equal(a[f], b[f],
                                          err_msg="%s step %d" % (f, step))
        assert eager.interaction_counts() == graph.interaction_counts()
        ta, tb = eager.tree(), graph.tree()
        assert len(ta) == len(tb)
        for f in ta.dtype.names:
            np.testing.assert_array_equal(ta[f], tb[f], err_msg=f)
        np.testing.assert_array_equal(eager.top_nodes(), graph.top_nodes())
        # eager work between replays
        graph.forcetest([0, 5])
        graph.get("pos")


def test_substep_walk_graph():
    """tdv_capture_substep_walk: with nothing moved, its replays equal the
    eager walk bit for bit (subsets of the particles active); after the
    active particles moved (the only ones a substep drifts), a replay takes
    their new positions and differs from an eager walk, which also
    recomputes every node's quadrupole, only by the quadrupoles' change."""
    pos, _, m = ics.clustered_box(1 << 17, nhalo=16, seed=3)
    e = _engine(pos, m, 0)
    st = e.get("pos", "acc", "last_acc_mag")
    n = len(m)
    e.capture_substep_walk()
    g = np.random.Generator(np.random.Philox(8))
    for k, sub in enumerate([np.sort(g.choice(n, n // 50, replace=False)),
                             np.sort(g.choice(n, 64, replace=False))]):
        e.set_active(sub)
        e.update_particles(st["pos"], acc=st["acc"],
                           last_acc_mag=st["last_acc_mag"])
        e.compute_acceleration(False)
        want = e.get("acc", "cost")
        for rep in range(2):
            e.update_particles(st["pos"], acc=st["acc"],
                               last_acc_mag=st["last_acc_mag"])
            e.launch_walk()
            got = e.get("acc", "cost")
            np.testing.assert_array_equal(got["acc"], want["acc"],
                                          err_msg="subset %d rep %d" % (k, rep))
            np.testing.assert_array_equal(got["cost"], want["cost"])
    # the active particles move by 1e-6 of the box
    sub = np.sort(g.choice(n, n // 50, replace=False))
    moved = st["pos"].copy()
    moved[sub] += g.normal(0.0, 1e-6, (len(sub), 3))
    e.set_active(sub)
    e.update_particles(moved, acc=st["acc"], last_acc_mag=st["last_acc_mag"])
    e.launch_walk()
    got = e.get("acc")["acc"][sub]
    e.update_particles(moved, acc=st["acc"], last_acc_mag=st["last_acc_mag"])
    e.compute_acceleration(False)
    want = e.get("acc")["acc"][sub]
    rel = np.linalg.norm(got - want, axis=1) / np.linalg.norm(want, axis=1)
    assert rel.max() < 1e-5
    # and the walk saw the moves: the particles' own records changed
    e.update_particles(st["pos"], acc=st["acc"], last_acc_mag=st["last_acc_mag"])
    e.launch_walk()
    unmoved = e.get("acc")["acc"][sub]
    assert np.abs(unmoved - got).max() > 0
    e.close()
